// Host-code input checks under AddressSanitizer and UndefinedBehaviorSanitizer (built with g++
// by tests/test_host_failures.py; no device code): the profile reader, the dispersion-file
// reader and the option parser of the command lines on every file / argument list given.
// Each input must end in a status code or an exception the programs catch, never in a memory
// error or undefined behaviour (the sanitizers abort the run on the first one).
//
//   host_inputs profile <file>...      ncpa_profile_check; prints "<file>: <code> <message>"
//   host_inputs dispersion <file>...   ncpa_dispersion_read (+ info, arrays, free)
//   host_inputs options <arg>...       the Modess option set: parse_command_line + validate
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "ncpamodes.h"
#include "params.h"

using ncpacli::Options;

static void modess_like(Options &p) {
    p.string("atmosfile");
    p.required("atmosfile");
    p.real("freq");
    p.real_gt("freq", 0.0);
    p.required("freq");
    p.flag("singleprop");
    p.flag("multiprop");
    p.radio("operation_mode", {"singleprop", "multiprop"});
    p.real("azimuth");
    p.real_ge("azimuth", 0.0);
    p.real_lt("azimuth", 360.0);
    p.required_if("azimuth", "singleprop");
    p.integer("Nz_grid", 20000);
    p.int_ge("Nz_grid", 10);
    p.string("ground_impedence_model", "rigid");
    p.one_of("ground_impedence_model", {"rigid"});
    p.string("paramfile", "modess.param");
}

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "profile") {
        for (int i = 2; i < argc; ++i) {
            char msg[512];
            ncpa_profile_info_t info{};
            const int rc = ncpa_profile_check(argv[i], "", &info, msg, (int)sizeof msg);
            // a message cut to the buffer stays terminated
            char tiny[8];
            (void)ncpa_profile_check(argv[i], nullptr, nullptr, tiny, (int)sizeof tiny);
            if (std::strlen(tiny) >= sizeof tiny) return 3;
            for (char *c = msg; *c; ++c)
                if (*c == '\n') *c = '|';   // one output line per file
            std::printf("%s: %d %s\n", argv[i], rc, rc ? msg : "ok");
        }
        return 0;
    }
    if (mode == "dispersion") {
        for (int i = 2; i < argc; ++i) {
            ncpa_dispersion_t *d = nullptr;
            const int rc = ncpa_dispersion_read(argv[i], 2, &d);
            if (rc == NCPA_OK) {
                int nrec = 0;
                size_t nm = 0;
                ncpa_dispersion_info(d, &nrec, &nm);
                std::vector<double> f(nrec), rs(nrec), rr(nrec), a(nm + 1), b(nm + 1), c(nm + 1), e(nm + 1);
                std::vector<int> cnt(nrec);
                ncpa_dispersion_arrays(d, f.data(), cnt.data(), rs.data(), rr.data(), a.data(), b.data(), c.data(),
                                       e.data());
                ncpa_dispersion_free(d);
                std::printf("%s: 0 %d records %zu modes\n", argv[i], nrec, nm);
            } else {
                std::printf("%s: %d\n", argv[i], rc);
            }
        }
        return 0;
    }
    if (mode == "options") {
        Options p;
        modess_like(p);
        std::vector<char *> av;
        av.push_back(argv[0]);
        for (int i = 2; i < argc; ++i) av.push_back(argv[i]);
        try {
            p.parse_command_line((int)av.size(), av.data());
            p.parse_file(p.get_string("paramfile"));
            const bool ok = p.validate();
            std::printf("options: %s\n", ok ? "valid" : "invalid");
        } catch (const std::invalid_argument &e) {
            std::printf("options: rejected: %s\n", e.what());
        }
        return 0;
    }
    return 2;
}
