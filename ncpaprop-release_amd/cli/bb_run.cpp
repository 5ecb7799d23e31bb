// ModBB --propagation on the GPU: NCPA::ModalBroadbandPropagator (src/libbroadband/
// ModalBroadbandPropagator.cpp:25-277) and BroadbandPropagator::get_source_spectrum /
// fft_pulse_prop (src/libbroadband/BroadbandPropagator.cpp:244-651) through the ncpamodes C ABI.
// The dispersion file is parsed here; transfer functions, FFTs and the time-phase product run
// on the device for all receivers at once.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ess_run.h"
#include "ncpamodes.h"
#include "params.h"

namespace ncpacli {

namespace {

struct Fail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// NCPA::count_rows_arbcol (src/common/util.cpp:269-284): newline count
int count_rows(const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) throw std::invalid_argument("file << " + path + " could not be opened.\n");
    int n = 0, c;
    while ((c = std::getc(f)) != EOF)
        if (c == '\n') ++n;
    std::fclose(f);
    return n;
}

struct Dispersion {
    std::vector<double> f;             // DC first
    std::vector<int> count, begin;     // per frequency bin
    std::vector<double> kre, kim, psr; // flat, file order
    double rho_s = 1.0, rho_r = 1.0;
};

// ModalBroadbandPropagator::read_dispersion_file (ModalBroadbandPropagator.cpp:98-138), the
// parse done by ncpa_dispersion_read (memory-mapped, multi-threaded)
Dispersion read_dispersion(const std::string &path) {
    {
        FILE *fp = std::fopen(path.c_str(), "r");
        if (!fp) throw std::invalid_argument("file << " + path + " could not be opened.\n");
        std::fclose(fp);
    }
    std::printf("--> Opening dispersion file %s ...if the program chokes here the file may be corrupt.\n", path.c_str());
    ncpa_dispersion_t *h = nullptr;
    if (ncpa_dispersion_read(path.c_str(), 0, &h) != NCPA_OK) throw Fail(path + ": read_dispersion_data format error");
    int nrec = 0;
    size_t nm = 0;
    ncpa_dispersion_info(h, &nrec, &nm);
    std::vector<double> f(nrec), rs(nrec), rr(nrec), ps(nm), pr(nm);
    std::vector<int> cnt(nrec);
    Dispersion d;
    d.kre.resize(nm);
    d.kim.resize(nm);
    ncpa_dispersion_arrays(h, f.data(), cnt.data(), rs.data(), rr.data(), d.kre.data(), d.kim.data(), ps.data(),
                           pr.data());
    ncpa_dispersion_free(h);
    d.f.push_back(0.0);   // the DC bin (ModalBroadbandPropagator.cpp:114-115)
    d.count.push_back(0);
    d.begin.push_back(0);
    size_t o = 0;
    for (int r = 0; r < nrec; ++r) {
        d.f.push_back(f[r]);
        d.begin.push_back((int)o);
        d.count.push_back(cnt[r]);
        o += (size_t)cnt[r];
    }
    d.psr.resize(nm);
    for (size_t j = 0; j < nm; ++j) d.psr[j] = ps[j] * pr[j];
    if (nrec > 0) {   // rho of the last record read
        d.rho_s = rs[nrec - 1];
        d.rho_r = rr[nrec - 1];
    }
    std::printf("--> Found modes at %d positive frequencies in file %s\n", (int)d.f.size(), path.c_str());
    if (d.f.size() < 3) throw Fail("the dispersion file needs at least two frequencies");
    return d;
}

std::vector<double> read_rows(const std::string &path, int cols) {
    std::ifstream in(path);
    if (!in) throw Fail("file " + path + " could not be opened.");
    std::vector<double> v;
    double x;
    while (in >> x) v.push_back(x);
    if (v.size() % cols) throw Fail("file " + path + " does not hold rows of " + std::to_string(cols) + " numbers");
    return v;
}

struct Dev {
    ncpa_ctx_t *ctx = nullptr;
    std::vector<void *> bufs;
    ~Dev() {
        for (void *b : bufs) ncpa_dev_free(ctx, b);
        if (ctx) ncpa_ctx_destroy(ctx);
    }
    void check(int rc) const {
        if (rc != NCPA_OK)
            throw Fail(ctx ? ncpa_ctx_error(ctx)
                           : "ncpamodes: cannot create a device context (no HIP device?), code " + std::to_string(rc));
    }
    double *put(const std::vector<double> &h, size_t n) {
        void *d = nullptr;
        check(ncpa_dev_alloc(ctx, std::max<size_t>(n, 1) * sizeof(double), &d));
        bufs.push_back(d);
        if (!h.empty()) check(ncpa_dev_upload(ctx, d, h.data(), h.size() * sizeof(double)));
        return (double *)d;
    }
};

}  // namespace

int bb_run(const Options &p) {
    try {
        int NFFT = p.get_int("nfft");
        const std::string out_file = p.get_string("output_waveform_file");
        const std::string disp_file = p.get_string("input_dispersion_file");
        const std::string source_type = p.get_string("source");
        const std::string source_file = p.get_string("source_file");
        const double f_center = p.get_real("f_center");
        const double max_cel = p.get_real("max_celerity");
        const bool single = p.get_string("receiver") == "single";
        const bool zero_attn = p.get_bool("use_zero_attenuation");
        std::vector<double> r_vec;
        if (single) {
            r_vec.push_back(p.get_real("range_km") * 1000.0);
        } else {
            const double rmin = p.get_real("start_range_km"), rmax = p.get_real("end_range_km"),
                         rstep = p.get_real("range_step_km");
            if (rmax <= rmin) throw std::runtime_error("end_range_km must be greater than start_range_km");
            const int Nr = (int)std::floor((rmax - rmin) / rstep) + 1;
            for (int i = 0; i < Nr; ++i) r_vec.push_back((rmin + (double)i * rstep) * 1000.0);
        }
        Dispersion d = read_dispersion(disp_file);
        const int Nfreq = (int)d.f.size();
        const double f_step = d.f[2] - d.f[1];
        if (NFFT < 4 * Nfreq) {   // ModalBroadbandPropagator.cpp:71-74
            NFFT = 4 * Nfreq;
            std::printf("Minimum NFFT is set to NFFT = %d\n", NFFT);
        }
        std::vector<double> rows;
        int kind = NCPA_SRC_IMPULSE, n_in = 0;
        if (source_type == "pulse1") kind = NCPA_SRC_PULSE1;
        else if (source_type == "pulse2") kind = NCPA_SRC_PULSE2;
        else if (source_type == "spectrum") kind = NCPA_SRC_SPECTRUM;
        else if (source_type == "waveform") kind = NCPA_SRC_WAVEFORM;
        if (kind == NCPA_SRC_WAVEFORM) {
            const int filelines = count_rows(source_file);
            while (filelines > NFFT) {   // ModalBroadbandPropagator.cpp:77-84
                NFFT *= 2;
                std::cout << "File " << source_file << " has " << filelines << " lines, increasing NFFT to " << NFFT
                          << std::endl;
            }
            std::cout << "Loading time domain source pulse from file " << source_file << std::endl;
            rows = read_rows(source_file, 2);
            n_in = (int)rows.size() / 2;
        } else if (kind == NCPA_SRC_SPECTRUM) {
            std::cout << "Loading source spectrum from file " << source_file << std::endl;
            rows = read_rows(source_file, 3);
            n_in = (int)rows.size() / 3;
        }
        const double fmx = (double)NFFT * f_step;
        // the outputs, before any device work (BroadbandPropagator.cpp:377-414 and
        // ModalBroadbandPropagator.cpp:200-230 write them unchecked)
        for (const std::string &f : {std::string("source_waveform_input.dat"), std::string("source_spectrum_input.dat"),
                                     out_file})
            check_writable(f);

        Dev dev;
        dev.check(ncpa_ctx_create(0, nullptr, &dev.ctx));
        std::vector<double> S(2 * (size_t)Nfreq), pulse(2 * (size_t)NFFT);
        dev.check(ncpa_bb_source(dev.ctx, kind, d.f.data(), Nfreq, NFFT, f_center, rows.empty() ? nullptr : rows.data(),
                                 n_in, S.data(), pulse.data()));
        {   // BroadbandPropagator.cpp:377-414
            FILE *f = std::fopen("source_waveform_input.dat", "w");
            if (!f) throw Fail("cannot open source_waveform_input.dat");
            const double factor = kind == NCPA_SRC_WAVEFORM ? 1.0 : 2.0;
            for (int i = 0; i < NFFT / 2; ++i) std::fprintf(f, "%12.6f %15.6e\n", 1.0 * i / fmx, factor * pulse[2 * i]);
            std::fclose(f);
            std::cout << "--> Initial source waveform saved in file 'source_waveform_input.dat' " << std::endl
                      << " with format: | Time (s) | Amplitude |" << std::endl;
            f = std::fopen("source_spectrum_input.dat", "w");
            if (!f) throw Fail("cannot open source_spectrum_input.dat");
            for (int i = 0; i < Nfreq; ++i) std::fprintf(f, "%14.9f %15.6e %15.6e\n", d.f[i], S[2 * i], S[2 * i + 1]);
            std::fclose(f);
            std::cout << "--> The source spectrum is saved in file 'source_spectrum_input.dat' " << std::endl
                      << "with format: | Freq (Hz) | Re(S) | Imag(S) |" << std::endl;
        }
        const size_t nm = d.kre.size();
        double *d_kre = dev.put(d.kre, nm), *d_kim = dev.put(d.kim, nm), *d_psr = dev.put(d.psr, nm);
        const int Nr = (int)r_vec.size();
        double *d_H = dev.put({}, 2 * (size_t)Nr * Nfreq);
        double *d_W = dev.put({}, (size_t)Nr * NFFT);
        const double rho_ratio = d.rho_r / d.rho_s;
        dev.check(ncpa_bb_transfer(dev.ctx, 1, Nfreq, d.begin.data(), d.count.data(), d_kre, d_kim, d_psr,
                                   zero_attn ? 1 : 0, r_vec.data(), Nr, &rho_ratio, d_H));
        dev.check(ncpa_bb_pulse(dev.ctx, 1, Nr, Nfreq, d_H, S.data(), d.f.data(), r_vec.data(), max_cel, NFFT, d_W));
        std::vector<double> W((size_t)Nr * NFFT);
        dev.check(ncpa_dev_download(dev.ctx, W.data(), d_W, W.size() * sizeof(double)));

        std::cout << "--> Propagating pulse from source-to-receivers on grid ..." << std::endl;
        std::printf("----------------------------------------------\n");
        std::printf("max_celerity     t0             R\n");
        std::printf("    m/s          sec            km\n");
        std::printf("----------------------------------------------\n");
        FILE *f = std::fopen(out_file.c_str(), "w");
        if (!f) throw Fail("cannot open " + out_file);
        for (int n = 0; n < Nr; ++n) {   // ModalBroadbandPropagator.cpp:208-228
            const double rr = r_vec[n], t0 = rr / max_cel;
            std::printf("%8.3f     %9.3f      %9.3f\n", max_cel, t0, rr / 1000.0);
            for (int i = 0; i < NFFT; ++i)
                std::fprintf(f, "%10.3f %12.6f %15.6e\n", rr / 1000.0, 1.0 * i / fmx + t0, W[(size_t)n * NFFT + i]);
            std::fprintf(f, "\n");
        }
        std::fclose(f);
        std::printf("f_step = %f   1/f_step = %f\n", f_step, 1.0 / f_step);
        std::printf("Time array length = %d; delta_t = %g s\n", NFFT, 1.0 / fmx);
        std::printf("Propagation results saved in file: %s\n", out_file.c_str());
        std::printf("with columns: R (km) | time (s) | pulse(R,t) |\n");
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}

}  // namespace ncpacli
