// Modess / ModBB --dispersion run on the GPU; see ess_run.h.
//
// The reference solves one (frequency, azimuth) at a time (ESSModeSolver::solve,
// ESSModeSolver.cpp:291-510).  Here every problem of the run goes to the device as one batch
// (ncpa_ess_solve), and the outputs are written afterwards in the reference's loop order with
// its file-open modes, so the files on disk end up identical in layout: files the reference
// rewrites for every problem ("w") hold the last problem's data, appended files hold all of them.
#include "ess_run.h"

#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <vector>

#include "common.h"
#include "ncpamodes.h"

namespace ncpacli {

using ncpa::PI;

// Options of Modess (src/modess/modess_parameters.cpp:6-222) or, wide, of WMod
// (src/wmod/wmod_parameters.cpp:6-214: the same set without --modal_starter_file and
// --write_speeds; both default --paramfile to modess.param).
void configure_modes_options(Options &p, bool wide) {
    p.flag("help");
    p.flag("h");
    p.string("paramfile", "modess.param");
    p.flag("printparams");
    p.string("atmosfile");
    p.required("atmosfile");
    p.real("freq");
    p.required("freq");
    p.real_gt("freq", 0.0);
    p.string("atmosheaderfile", "");
    p.real("maxheight_km", 150.0);
    p.real_ge("maxheight_km", 0.1);
    p.real("zground_km", 0.0);
    p.integer("Nz_grid", 20000);
    p.int_ge("Nz_grid", 10);
    p.real("sourceheight_km", 0.0);
    p.real("receiverheight_km", 0.0);
    p.real("maxrange_km", 1000.0);
    p.real_ge("maxrange_km", 0.01);
    p.integer("Nrng_steps", 1000);
    p.int_gt("Nrng_steps", 0);
    p.string("ground_impedence_model", "rigid");
    p.one_of("ground_impedence_model", {"rigid"});
    p.string("use_attn_file", "");
    if (!wide) p.string("modal_starter_file", "");
    p.string("dispersion_file", "");
    p.flag("append_dispersion_file");
    p.string("filetag", "");
    p.flag("singleprop");
    p.flag("multiprop");
    p.radio("operation_mode", {"singleprop", "multiprop"});
    p.real("azimuth");
    p.real_ge("azimuth", 0.0);
    p.real_lt("azimuth", 360.0);
    p.required_if("azimuth", "singleprop");
    for (const char *k : {"azimuth_start", "azimuth_end", "azimuth_step"}) {
        p.real(k);
        p.real_ge(k, 0.0);
        p.real_le(k, 360.0);
        p.required_if(k, "multiprop");
    }
    for (const char *k : {"write_2d_tloss", "write_lossless", "write_phase_speeds", "write_modes",
                          "write_atm_profile", "Lamb_wave_BC", "turnoff_WKB", "wvnum_filter"})
        p.flag(k);
    if (!wide) p.flag("write_speeds");
    p.real("c_min");
    p.real_gt("c_min", 0.0);
    p.required_if("c_min", "wvnum_filter");
    p.real("c_max");
    p.real_gt("c_max", 0.0);
    p.required_if("c_max", "wvnum_filter");
    // dummy parameters the solver reads (modess_parameters.cpp:200-205, wmod_parameters.cpp:181-184)
    p.flag("broadband");
    p.real("f_min", 0.0);
    p.real("f_max", 0.0);
    p.real("f_step", 0.0);
}

EssRun ess_from_options(const Options &p) {
    EssRun r;
    r.atmosfile = p.get_string("atmosfile");
    r.atmosheaderfile = p.get_string("atmosheaderfile");
    r.gnd_imp_model = p.get_string("ground_impedence_model");
    r.usrattfile = p.get_string("use_attn_file");
    if (p.has("modal_starter_file")) r.modstartfile = p.get_string("modal_starter_file");
    r.z_min = p.get_real("zground_km") * 1000.0;
    r.maxrange = p.get_real("maxrange_km") * 1000.0;
    r.maxheight = p.get_real("maxheight_km") * 1000.0;
    r.sourceheight = p.get_real("sourceheight_km") * 1000.0;
    r.receiverheight = p.get_real("receiverheight_km") * 1000.0;
    r.Nz_grid = p.get_int("Nz_grid");
    r.Nrng_steps = p.get_int("Nrng_steps");
    r.Lamb_wave_BC = p.get_bool("Lamb_wave_BC");
    r.write_2D_TLoss = p.get_bool("write_2d_tloss");
    r.write_atmosphere = p.get_bool("write_atm_profile");
    r.write_phase_speeds = p.get_bool("write_phase_speeds");
    if (p.has("write_speeds")) r.write_speeds = p.get_bool("write_speeds");
    r.write_modes = p.get_bool("write_modes");
    r.dispersion_file = p.get_string("dispersion_file");
    r.append_dispersion_file = p.get_bool("append_dispersion_file");
    r.Nby2Dprop = p.get_bool("multiprop");
    r.turnoff_WKB = p.get_bool("turnoff_WKB");
    r.write_lossless = p.get_bool("write_lossless");
    r.user_tag = p.get_string("filetag");
    if (!r.user_tag.empty()) r.user_tag += ".";
    r.wvnum_filter = p.get_bool("wvnum_filter");
    if (r.wvnum_filter) {
        r.c_min = p.get_real("c_min");
        r.c_max = p.get_real("c_max");
    }
    // WKB is not used when any of these are saved (ESSModeSolver.cpp:78-81, WModeSolver.cpp:75-77)
    if (r.write_phase_speeds || r.write_speeds || r.write_2D_TLoss || r.write_modes || !r.dispersion_file.empty())
        r.turnoff_WKB = true;
    if (r.Nby2Dprop) {   // ESSModeSolver.cpp:84-88
        r.azi_min = p.get_real("azimuth_start");
        const double azi_max = p.get_real("azimuth_end");
        r.azi_step = p.get_real("azimuth_step");
        r.Naz = (int)((azi_max - r.azi_min) / r.azi_step) + 1;
    } else {
        r.azi_min = p.get_real("azimuth");
    }
    if (p.get_bool("broadband")) {   // ESSModeSolver.cpp:98-114
        const double f_min = p.get_real("f_min"), f_step = p.get_real("f_step"), f_max = p.get_real("f_max");
        if (f_min >= f_max) throw std::runtime_error("f_min must be less than f_max!");
        const int nf = (int)std::floor((f_max - f_min) / f_step) + 1;
        for (int i = 0; i < nf; ++i) r.freqs.push_back(f_min + (double)i * f_step);
    } else {
        r.freqs.push_back(p.get_real("freq"));
    }
    if (p.has("method")) r.wide = p.get_string("method") == "wmod";   // ModBB_main.cpp:86-90
    if (r.gnd_imp_model != "rigid")
        throw std::invalid_argument("This ground impedance model is not implemented yet: " + r.gnd_imp_model);
    return r;
}

namespace {

struct Fail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Session {
    ncpa_ctx_t *ctx = nullptr;
    ncpa_modes_t *modes = nullptr;
    std::vector<void *> dev;
    ~Session() {
        for (void *d : dev) ncpa_dev_free(ctx, d);
        if (modes) ncpa_modes_free(modes);
        if (ctx) ncpa_ctx_destroy(ctx);
    }
    void check(int rc) const {
        if (rc != NCPA_OK)
            throw Fail(ctx ? ncpa_ctx_error(ctx)
                           : "ncpamodes: cannot create a device context (no HIP device?), code " + std::to_string(rc));
    }
    double *alloc(size_t n) {
        void *d = nullptr;
        check(ncpa_dev_alloc(ctx, n * sizeof(double), &d));
        dev.push_back(d);
        return (double *)d;
    }
    void download(double *host, const double *d, size_t n) { check(ncpa_dev_download(ctx, host, d, n * sizeof(double))); }
};

FILE *open_or_throw(const std::string &path, const char *mode) {
    FILE *f = std::fopen(path.c_str(), mode);
    if (!f) throw Fail("cannot open " + path);
    return f;
}

// --use_attn_file: "altitude[km] attenuation[Np/m]" lines interpolated onto the profile nodes
// with the Steffen spline (Atmosphere1D::read_attenuation_from_file, Atmosphere1D.cpp:307-424).
// A table that starts above the profile's lowest node or ends below its highest is first
// extended by a node at that altitude carrying the table's end value (the reference's intent;
// its copies move bytes instead of doubles, Atmosphere1D.cpp:368-392), so the end values are
// held outside the table and the Steffen slopes at the table's end nodes are those of the
// extended table.
std::vector<double> attenuation_at_nodes(const std::string &path, const std::vector<double> &z_nodes_km) {
    std::ifstream in(path);
    if (!in) throw Fail("cannot open attenuation file " + path);
    std::vector<double> za, aa;
    std::string line;
    while (std::getline(in, line)) {
        const size_t b = line.find_first_not_of(" \t\r");
        if (b == std::string::npos || line[b] == '#') continue;
        for (char &c : line)
            if (c == ',') c = ' ';
        std::istringstream ss(line);
        double z, a;
        std::string extra;
        if (!(ss >> z >> a) || (ss >> extra))
            throw Fail("Atmosphere1D - Error parsing attenuation line:\n" + line +
                       "\nMust be formatted as:\naltitude  attenuation");
        za.push_back(z);
        aa.push_back(a);
    }
    if (za.size() < 2) throw Fail("attenuation file " + path + " needs at least two lines");
    for (size_t i = 1; i < za.size(); ++i)
        if (!(za[i] > za[i - 1])) throw Fail("attenuation file " + path + ": altitudes must increase");
    if (!z_nodes_km.empty() && za.front() > z_nodes_km.front()) {
        za.insert(za.begin(), z_nodes_km.front());
        aa.insert(aa.begin(), aa.front());
    }
    if (!z_nodes_km.empty() && za.back() < z_nodes_km.back()) {
        za.push_back(z_nodes_km.back());
        aa.push_back(aa.back());
    }
    std::vector<double> out(z_nodes_km.size());
    auto yf = [&aa](int k) { return aa[k]; };
    for (size_t i = 0; i < out.size(); ++i) out[i] = ncpa::steffen_eval(za.data(), (int)za.size(), yf, z_nodes_km[i]);
    return out;
}

// Atmosphere1D::print_atmosphere with the keys ESSModeSolver.cpp:468-484 asks for
void write_atmosphere(const std::string &path, const ncpa_profile_info_t &info, const std::vector<double> &z,
                      const std::vector<double> &c0, const std::vector<double> &ws, const std::vector<double> &wd,
                      const std::vector<double> &T, const std::vector<double> &P, const std::vector<double> &rho,
                      double azimuth) {
    std::ofstream os(path);
    if (!os) throw Fail("cannot open " + path);
    if (info.has_z0) os << "#% 0, Z0, m, " << info.z0 << std::endl;
    os << "#% 0, AZ, deg CW from N, " << azimuth << std::endl;
    os << "#% 1, Z, m" << std::endl;
    const char *cols[] = {"U, m/s", "V, m/s", "T, K", "RHO, kg/m3", "P, Pa", "C0, m/s", "CEFF, m/s"};
    for (int c = 0; c < 7; ++c) os << "#% " << c + 2 << ", " << cols[c] << std::endl;
    os.setf(std::ios::scientific, std::ios::floatfield);
    os.setf(std::ios::right, std::ios::adjustfield);
    os.precision(6);
    const double deg = PI / 180.0;
    for (size_t i = 0; i < z.size(); ++i) {
        const double u = ws[i] * std::sin(wd[i] * deg), v = ws[i] * std::cos(wd[i] * deg);
        const double ceff = c0[i] + ws[i] * std::cos((wd[i] - azimuth) * deg);
        os.width(9);
        os << z[i] << " " << u << " " << v << " " << T[i] << " " << rho[i] << " " << P[i] << " " << c0[i] << " "
           << ceff << std::endl;
    }
}

}  // namespace

// A path the run will write: its directory must be writable, and the file too if it exists
// (the reference fopen()s its outputs only when it writes them and does not check the result,
// ModeSolver.cpp:319-323; here an unwritable output stops the run before any device work).
void check_writable(const std::string &path) {
    const size_t slash = path.find_last_of('/');
    const std::string dir = slash == std::string::npos ? "." : (slash == 0 ? "/" : path.substr(0, slash));
    if (::access(dir.c_str(), W_OK) != 0)
        throw std::runtime_error("cannot open " + path + " for writing (directory " + dir + ")");
    if (::access(path.c_str(), F_OK) == 0 && ::access(path.c_str(), W_OK) != 0)
        throw std::runtime_error("cannot open " + path + " for writing");
}

namespace {

// Everything the run reads or writes, checked on the host before a device context exists:
// the profile (the reference's own messages), the attenuation table and the output paths.
void preflight(const EssRun &r) {
    char msg[1024];
    ncpa_profile_info_t info{};
    if (ncpa_profile_check(r.atmosfile.c_str(), r.atmosheaderfile.c_str(), &info, msg, (int)sizeof msg) != NCPA_OK)
        throw Fail(msg);
    if (!r.usrattfile.empty()) (void)attenuation_at_nodes(r.usrattfile, {});
    const char *ext = r.wide ? ".wnm" : ".nm";
    std::vector<std::string> out;
    for (int ll = 0; ll < (r.write_lossless ? 2 : 1); ++ll) {
        if (r.Nby2Dprop) out.push_back(r.user_tag + (ll ? "Nby2D_tloss_1d.lossless" : "Nby2D_tloss_1d") + ext);
        else out.push_back(r.user_tag + (ll ? "tloss_1d.lossless" : "tloss_1d") + ext);
        if (r.write_2D_TLoss && !r.Nby2Dprop) out.push_back(r.user_tag + (ll ? "tloss_2d.lossless" : "tloss_2d") + ext);
    }
    if (!r.Nby2Dprop) {
        if (r.write_phase_speeds) out.push_back(r.user_tag + "phasespeeds" + ext);
        if (r.write_speeds) out.push_back(r.user_tag + "speeds.nm");
        if (r.write_modes) out.push_back(r.user_tag + "mode_0" + ext);
        if (r.write_atmosphere) out.push_back(r.user_tag + "atm_profile" + ext);
        if (!r.modstartfile.empty()) out.push_back(r.user_tag + r.modstartfile);
        if (!r.dispersion_file.empty()) out.push_back(r.dispersion_file);
    }
    for (const std::string &f : out) check_writable(f);
}

}  // namespace

int ess_run(const EssRun &r) {
    Session s;
    try {
        preflight(r);
        s.check(ncpa_ctx_create(0, nullptr, &s.ctx));
        int pid = -1;
        ncpa_profile_info_t info{};
        s.check(ncpa_profile_read(s.ctx, r.atmosfile.c_str(), r.atmosheaderfile.c_str(), &pid, &info));
        const int nn = info.n_nodes;
        std::vector<double> zn(nn), c0n(nn), wsn(nn), wdn(nn), Tn(nn), Pn(nn), rhon(nn);
        s.check(ncpa_profile_nodes(s.ctx, pid, zn.data(), c0n.data(), wsn.data(), wdn.data(), Tn.data(), Pn.data(),
                                   rhon.data()));
        if (!r.usrattfile.empty()) {
            std::vector<double> zkm(nn);
            for (int i = 0; i < nn; ++i) zkm[i] = zn[i] / 1000.0;
            std::vector<double> a = attenuation_at_nodes(r.usrattfile, zkm);
            s.check(ncpa_profile_set_attenuation(s.ctx, pid, a.data()));
        }
        ncpa_grid_t g{};
        s.check(ncpa_grid_make(s.ctx, pid, r.Nz_grid, r.maxheight / 1000.0, r.z_min / 1000.0, &g));
        if (g.maxheight < r.maxheight)
            std::cout << std::endl << "maxheight adjusted to: " << g.maxheight << " m (max available in profile file)"
                      << std::endl;
        if (g.z_min > r.z_min)
            std::cout << "Adjusting minimum altitude to profile ground height of " << g.z_min << " m" << std::endl;
        const int Nz = g.Nz;
        // WModeSolver::solve spaces the matrix rows by (maxheight - z_min)/Nz (WModeSolver.cpp:297)
        const double dz = r.wide ? (g.maxheight - g.z_min) / Nz : g.dz;
        const char *ext = r.wide ? ".wnm" : ".nm";   // WModeSolver.h:4-34 file names
        std::vector<double> rho(Nz);
        s.check(ncpa_profile_sample(s.ctx, pid, &g, rho.data(), nullptr, nullptr, nullptr));
        double adm = 0.0;
        if (r.Lamb_wave_BC) {
            s.check(ncpa_profile_admittance(s.ctx, pid, g.z_min, &adm));
            std::cout << "Admittance = " << adm << std::endl;
        }

        const int nf = (int)r.freqs.size(), naz = r.Naz, P = nf * naz;
        std::vector<ncpa_problem_t> probs;
        std::vector<double> azs;
        for (int fi = 0; fi < nf; ++fi)
            for (int it = 0; it < naz; ++it) {
                const double azi = r.azi_min + it * r.azi_step;
                ncpa_problem_t q{};
                q.profile = pid;
                q.flags = (r.turnoff_WKB ? 0 : NCPA_F_WKB) | (r.wvnum_filter ? NCPA_F_WVNUM_FILTER : 0);
                q.freq = r.freqs[fi];
                q.azimuth = azi;
                q.alpha_freq = r.freqs[fi];
                q.zs = r.sourceheight;
                q.zr = r.receiverheight;
                q.admittance = adm;
                q.c_min = r.c_min;
                q.c_max = r.c_max;
                probs.push_back(q);
                if (fi == 0) azs.push_back(azi);
            }
        if (r.wide)
            s.check(ncpa_wmod_solve(s.ctx, &g, probs.data(), P, &s.modes));
        else
            s.check(ncpa_ess_solve(s.ctx, &g, probs.data(), P, &s.modes));
        std::vector<int> counts(P);
        int total = 0;
        s.check(ncpa_modes_counts(s.modes, counts.data(), &total));
        std::vector<int> off(P + 1, 0);
        for (int p = 0; p < P; ++p) off[p + 1] = off[p] + counts[p];
        const size_t nt = std::max(total, 1);
        std::vector<double> kre(nt), kim(nt), psr(nt), ps(nt), pr(nt), kmin(P), kmax(P);
        s.check(ncpa_modes_download(s.modes, kre.data(), kim.data(), psr.data(), ps.data(), pr.data()));
        s.check(ncpa_modes_interval(s.modes, kmin.data(), kmax.data()));

        const int n_r = r.Nrng_steps;
        const double dr = r.maxrange / r.Nrng_steps;   // ESSModeSolver.cpp:313
        const int n_zsrc = (int)std::ceil(r.sourceheight / dz), n_zrcv = (int)std::ceil(r.receiverheight / dz);
        if (n_zsrc >= Nz || n_zrcv >= Nz) throw Fail("source or receiver height above the grid");

        // 1-D TL of every problem (lossy, and lossless when asked)
        std::vector<double> tl[2][3];
        for (int ll = 0; ll < (r.write_lossless ? 2 : 1); ++ll) {
            double *d[3] = {s.alloc((size_t)P * n_r), s.alloc((size_t)P * n_r), s.alloc((size_t)P * n_r)};
            s.check(ncpa_tl1d(s.ctx, s.modes, dr, n_r, ll, d[0], d[1], d[2]));
            for (int c = 0; c < 3; ++c) {
                tl[ll][c].resize((size_t)P * n_r);
                s.download(tl[ll][c].data(), d[c], (size_t)P * n_r);
            }
        }
        auto tl_name = [&](bool multi, bool lossless) {
            return r.user_tag + (multi ? (lossless ? "Nby2D_tloss_1d.lossless" : "Nby2D_tloss_1d")
                                       : (lossless ? "tloss_1d.lossless" : "tloss_1d")) + ext;
        };
        const int last = P - 1;

        for (int fi = 0; fi < nf; ++fi) {
            const double freq = r.freqs[fi];
            for (int it = 0; it < naz; ++it) {
                const int p = fi * naz + it;
                const double azi = azs[it];
                const int nm = counts[p];
                const int o = off[p];
                std::cout << std::endl << "Now processing azimuth = " << azi << " (" << it + 1 << " of " << naz << ")"
                          << std::endl;
                std::printf("______________________________________________________________________\n\n");
                if (r.wide)
                    std::printf(" -> Solving wide-angle problem at %g Hz and %g deg (%d modes)...\n", freq, azi, nm);
                else
                    std::printf(" -> Normal mode solution at %5.3f Hz and %5.2f deg (%d modes)...\n", freq, azi, nm);
                std::printf(" -> Discrete spectrum: %5.2f m/s to %5.2f m/s\n", 2 * PI * freq / kmax[p],
                            2 * PI * freq / kmin[p]);
                if (r.Nby2Dprop) {
                    // ModeSolver::getTLoss1DNx2 (ModeSolver.cpp:387-489): "w" for the first azimuth
                    for (int ll = 0; ll < (r.write_lossless ? 2 : 1); ++ll) {
                        FILE *f = open_or_throw(tl_name(true, ll), it == 0 ? "w" : "a");
                        for (int i = 0; i < n_r; ++i) {
                            const size_t k = (size_t)p * n_r + i;
                            std::fprintf(f, "%10.3f %8.3f %20.12e %20.12e %20.12e\n", (i + 1) * dr / 1000.0, azi,
                                         tl[ll][0][k], tl[ll][1][k], tl[ll][2][k]);
                        }
                        std::fprintf(f, "\n");
                        std::fclose(f);
                    }
                } else if (p == last) {
                    // ModeSolver::getTLoss1D (ModeSolver.cpp:292-381); earlier problems' files
                    // would be overwritten, so only the last one is written
                    std::cout << "Writing to file: 1D transmission loss" << std::endl;
                    for (int ll = 0; ll < (r.write_lossless ? 2 : 1); ++ll) {
                        FILE *f = open_or_throw(tl_name(false, ll), "w");
                        for (int i = 0; i < n_r; ++i) {
                            const size_t k = (size_t)p * n_r + i;
                            std::fprintf(f, "%f %20.12e %20.12e %20.12e\n", (i + 1) * dr / 1000.0, tl[ll][0][k],
                                         tl[ll][1][k], tl[ll][2][k]);
                        }
                        std::fclose(f);
                        std::printf("           file %s created\n", tl_name(false, ll).c_str());
                    }
                    const bool need_full = r.write_modes || r.write_speeds;
                    if (!r.modstartfile.empty()) {
                        // ESSModeSolver::getModalStarter (ESSModeSolver.cpp:670-723)
                        std::cout << "Writing to file: modal starter" << std::endl;
                        const double k0 = 2 * PI * freq / 340.0;
                        const int z_cnd = (int)((340.0 / freq) / 10 / dz);
                        if (z_cnd < 1) throw Fail("modal starter: depth step (340/f)/10 is below the grid spacing");
                        const int nrows = (Nz + z_cnd - 1) / z_cnd;
                        double *d_rows = s.alloc((size_t)nrows * nt);
                        s.check(ncpa_modes_rows(s.ctx, s.modes, 0, z_cnd, nrows, d_rows));
                        std::vector<double> rows((size_t)nrows * nt);
                        s.download(rows.data(), d_rows, rows.size());
                        FILE *f = open_or_throw(r.user_tag + r.modstartfile, "w");
                        for (int j = 0; j < nrows; ++j) {
                            double sum = 0.0;
                            for (int m = 0; m < nm; ++m) sum += ps[o + m] * rows[(size_t)j * nt + o + m] / std::sqrt(kre[o + m]);
                            sum *= PI * std::sqrt(k0);
                            std::fprintf(f, "%10.3f   %16.12e   %16.12e\n", (double)j * z_cnd * dz / 1000.0, sum, 0.0);
                        }
                        std::fprintf(f, "\n");
                        std::fclose(f);
                        std::printf("           file %s created\n", (r.user_tag + r.modstartfile).c_str());
                    }
                    if (r.write_2D_TLoss) {
                        // ModeSolver::getTLoss2D (ModeSolver.cpp:492-565): rows j = 0, stepj, ...
                        std::cout << "Writing to file: 2D transmission loss" << std::endl;
                        int stepj = Nz / 500;
                        if (stepj == 0) stepj = 10;
                        const int nrows = (Nz + stepj - 1) / stepj;
                        double *d_rows = s.alloc((size_t)nrows * nt);
                        s.check(ncpa_modes_rows(s.ctx, s.modes, 0, stepj, nrows, d_rows));
                        for (int ll = 0; ll < (r.write_lossless ? 2 : 1); ++ll) {
                            const size_t per = (size_t)nrows * 2 * n_r;
                            double *d_out = s.alloc((size_t)P * per);
                            s.check(ncpa_field2d(s.ctx, s.modes, d_rows, nrows, dr, n_r, ll, NCPA_PREC_FP64, d_out));
                            std::vector<double> fld(per);
                            s.download(fld.data(), d_out + (size_t)p * per, per);
                            const std::string name = r.user_tag + (ll ? "tloss_2d.lossless" : "tloss_2d") + ext;
                            FILE *f = open_or_throw(name, "w");
                            for (int i = 0; i < n_r; ++i) {
                                for (int j = 0; j < nrows; ++j) {
                                    const size_t k = (size_t)j * 2 * n_r + 2 * i;
                                    std::fprintf(f, "%f %f %15.8e %15.8e\n", (i + 1) * dr / 1000.0,
                                                 (double)j * stepj * dz / 1000.0, fld[k], fld[k + 1]);
                                }
                                std::fprintf(f, "\n");
                            }
                            std::fclose(f);
                            std::printf("           file %s created\n", name.c_str());
                        }
                    }
                    if (r.write_phase_speeds) {
                        // ModeSolver::writePhaseSpeeds (ModeSolver.cpp:603-614)
                        std::cout << "Writing to file: phase speeds" << std::endl;
                        const std::string name = r.user_tag + "phasespeeds" + ext;
                        FILE *f = open_or_throw(name, "w");
                        for (int j = 0; j < nm; ++j)
                            std::fprintf(f, "%d %f %15.8e\n", j, (2 * PI * freq) / kre[o + j], kim[o + j]);
                        std::fclose(f);
                        std::printf("           file %s created\n", name.c_str());
                    }
                    if (need_full) {
                        double *d_rows = s.alloc((size_t)Nz * nt);
                        s.check(ncpa_modes_rows(s.ctx, s.modes, 0, 1, Nz, d_rows));
                        std::vector<double> v((size_t)Nz * nt);
                        s.download(v.data(), d_rows, v.size());
                        if (r.write_modes) {
                            // ModeSolver::writeEigenFunctions (ModeSolver.cpp:618-641)
                            std::cout << "Writing to file: the modes and the phase and group speeds...\n";
                            for (int j = 0; j < nm; ++j) {
                                char name[64];
                                std::snprintf(name, sizeof(name), "mode_%d%s", j, ext);
                                FILE *f = open_or_throw(r.user_tag + name, "w");
                                double chk = 0.0;
                                for (int n = 0; n < Nz; ++n) {
                                    const double x = v[(size_t)n * nt + o + j];
                                    std::fprintf(f, "%f %15.8e\n", n * (dz / 1000.0), x);
                                    chk += x * x * dz;
                                }
                                if (std::fabs(1. - chk) > 0.1) std::printf("Check if eigenfunction %d is normalized!\n", j);
                                std::fclose(f);
                            }
                            std::printf("           %d mode files created\n", nm);
                        } else {
                            std::cout << "Writing to file: the modal phase speeds and the group speeds...\n";
                        }
                        // ESSModeSolver::writePhaseAndGroupSpeeds (ESSModeSolver.cpp:726-752);
                        // WModeSolver writes the mode files only (WModeSolver.cpp:431-434)
                        if (!r.wide) {
                            std::vector<double> ceff(Nz);
                            s.check(ncpa_modes_ceff(s.modes, p, ceff.data()));
                            const std::string name = r.user_tag + "speeds.nm";
                            FILE *f = open_or_throw(name, "w");
                            const double omega = 2 * PI * freq;
                            for (int j = 0; j < nm; ++j) {
                                const double v_phase = omega / kre[o + j];
                                double vg = 0.0;
                                for (int n = 0; n < Nz; ++n) {
                                    const double x = v[(size_t)n * nt + o + j];
                                    vg += x * x / (ceff[n] * ceff[n]);
                                }
                                vg = 1.0 / (vg * v_phase * dz);
                                std::fprintf(f, "%4d %9.3f %9.3f %15.8e\n", j + 1, v_phase, vg, kim[o + j]);
                            }
                            std::fclose(f);
                            std::printf("           Phase and group speeds saved in file %s.\n", name.c_str());
                        }
                    }
                }
                // the dispersion file and the atmosphere dump belong to the single-azimuth branch
                // (ESSModeSolver.cpp:454-484)
                if (!r.Nby2Dprop && !r.dispersion_file.empty() && (r.append_dispersion_file || p == last))
                    std::printf("Writing to file %s: dispersion at freq = %8.3f Hz...\n", r.dispersion_file.c_str(), freq);
                if (!r.Nby2Dprop && r.write_atmosphere && fi == 0) {
                    const std::string name = r.user_tag + "atm_profile" + ext;
                    std::cout << "Writing atmosphere to " << name << std::endl;
                    write_atmosphere(name, info, zn, c0n, wsn, wdn, Tn, Pn, rhon, azi);
                }
            }
        }
        // ModeSolver::writeDispersion (ModeSolver.cpp:569-589) of every problem, in the reference's
        // order and open modes: appended records, or ("w" per problem) the last one
        if (!r.Nby2Dprop && !r.dispersion_file.empty()) {
            const int r0 = r.append_dispersion_file ? 0 : last;
            const int nrec = P - r0;
            std::vector<double> fr(nrec), rs(nrec, rho[n_zsrc]), rr(nrec, rho[n_zrcv]);
            for (int q = 0; q < nrec; ++q) fr[q] = r.freqs[(r0 + q) / naz];
            const size_t o0 = off[r0];
            if (ncpa_dispersion_write(r.dispersion_file.c_str(), r.append_dispersion_file ? 1 : 0, nrec, fr.data(),
                                      counts.data() + r0, rs.data(), rr.data(), kre.data() + o0, kim.data() + o0,
                                      ps.data() + o0, pr.data() + o0, 0) != NCPA_OK)
                throw Fail("cannot write " + r.dispersion_file);
        }
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}

}  // namespace ncpacli
