// Dispersion files: the wire format between ModBB --dispersion and ModBB --propagation.
//
// Record (ModeSolver::writeDispersion, src/libmodes/ModeSolver.cpp:569-589), one line per
// (frequency, azimuth):
//     "%.7e   %d    %.7e   %.7e" f, n_modes, rho(z_src), rho(z_rcv)
//     then per mode "   %.12e   %.12e" Re k, Im k and "   %.12e   %.12e" psi(z_src), psi(z_rcv)
// read back with fscanf by ModalBroadbandPropagator::read_dispersion_file
// (src/libbroadband/ModalBroadbandPropagator.cpp:98-138).
//
// At the sizes of a broadband job (1024 frequencies x hundreds of modes, tens of MB of text)
// printf/scanf take longer than the GPU solve.  Here records are formatted in parallel with
// std::to_chars (specified as printf's %.*e in the C locale, so the bytes are the same) and
// written with one fwrite; the reader maps the file and parses whitespace-separated numbers
// with std::from_chars on several threads, then assembles the records sequentially.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ncpamodes.h"

struct ncpa_dispersion {
    std::vector<double> f, rho_s, rho_r;
    std::vector<int> count;
    std::vector<double> kre, kim, ps, pr;
};

namespace {

inline bool is_space(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

int clamp_threads(int n) {
    if (n <= 0) n = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(n, 64));
}

template <class F>
void parallel_for(int n_items, int n_threads, F fn) {
    const int T = std::min(n_threads, std::max(n_items, 1));
    if (T <= 1) {
        for (int i = 0; i < n_items; ++i) fn(i);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([=, &fn]() {
            for (int i = t; i < n_items; i += T) fn(i);
        });
    for (auto &x : th) x.join();
}

// numbers of [b, e) (split at whitespace); stops at the first token that is not a whole number
struct Chunk {
    std::vector<double> v;
    bool bad = false;   // a token that does not parse: the stream ends before it
};

void parse_chunk(const char *b, const char *e, Chunk &out) {
    out.v.reserve((size_t)(e - b) / 20 + 4);
    const char *p = b;
    while (p < e) {
        while (p < e && is_space(*p)) ++p;
        if (p >= e) break;
        const char *q = p;
        while (q < e && !is_space(*q)) ++q;
        double x;
        const char *s = (*p == '+') ? p + 1 : p;   // from_chars takes no leading '+'
        const auto r = std::from_chars(s, q, x);
        if (r.ec != std::errc() || r.ptr != q) {
            out.bad = true;
            return;
        }
        out.v.push_back(x);
        p = q;
    }
}

char *put_e(char *p, double v, int prec) { return std::to_chars(p, p + 40, v, std::chars_format::scientific, prec).ptr; }

char *put_sep(char *p, int n_spaces) {
    for (int i = 0; i < n_spaces; ++i) *p++ = ' ';
    return p;
}

}  // namespace

extern "C" {

int ncpa_dispersion_read(const char *path, int n_threads, ncpa_dispersion_t **out) {
    if (!path || !out) return NCPA_ERR_ARG;
    *out = nullptr;
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0) return NCPA_ERR_ARG;
    struct stat st;
    if (::fstat(fd, &st) != 0) {
        ::close(fd);
        return NCPA_ERR_ARG;
    }
    const size_t size = (size_t)st.st_size;
    const char *text = nullptr;
    void *map = nullptr;
    if (size > 0) {
        map = ::mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
        if (map == MAP_FAILED) {
            ::close(fd);
            return NCPA_ERR_ALLOC;
        }
        text = (const char *)map;
    }
    ::close(fd);
    // chunks end at whitespace so that no number is split
    const int T = size < (1u << 20) ? 1 : clamp_threads(n_threads);
    std::vector<size_t> cut(T + 1, size);
    cut[0] = 0;
    for (int t = 1; t < T; ++t) {
        size_t c = std::max(cut[t - 1], size * (size_t)t / (size_t)T);
        while (c < size && !is_space(text[c])) ++c;
        cut[t] = c;
    }
    std::vector<Chunk> ch(T);
    parallel_for(T, T, [&](int t) { parse_chunk(text + cut[t], text + cut[t + 1], ch[t]); });
    if (map) ::munmap(map, size);
    // the number stream, up to the first token that did not parse
    size_t total = 0;
    int last_chunk = T - 1;
    for (int t = 0; t < T; ++t) {
        total += ch[t].v.size();
        if (ch[t].bad) {
            last_chunk = t;
            break;
        }
    }
    std::vector<double> num;
    num.reserve(total);
    for (int t = 0; t <= last_chunk; ++t) num.insert(num.end(), ch[t].v.begin(), ch[t].v.end());
    ch.clear();
    // records: while a frequency reads, the count, rho_s, rho_r and 4 numbers per mode follow
    // (ModalBroadbandPropagator.cpp:117-132)
    ncpa_dispersion *d = new ncpa_dispersion();
    size_t i = 0;
    while (i < num.size()) {
        if (i + 4 > num.size()) { delete d; return NCPA_ERR_ARG; }
        const double f = num[i], c = num[i + 1];
        if (!(c >= 0.0) || c != std::floor(c) || c > (double)NCPA_MAX_MODES) { delete d; return NCPA_ERR_ARG; }
        const size_t n = (size_t)c;
        if (i + 4 + 4 * n > num.size()) { delete d; return NCPA_ERR_ARG; }
        d->f.push_back(f);
        d->count.push_back((int)n);
        d->rho_s.push_back(num[i + 2]);
        d->rho_r.push_back(num[i + 3]);
        const double *m = num.data() + i + 4;
        for (size_t j = 0; j < n; ++j) {
            d->kre.push_back(m[4 * j]);
            d->kim.push_back(m[4 * j + 1]);
            d->ps.push_back(m[4 * j + 2]);
            d->pr.push_back(m[4 * j + 3]);
        }
        i += 4 + 4 * n;
    }
    *out = d;
    return NCPA_OK;
}

int ncpa_dispersion_info(const ncpa_dispersion_t *d, int *n_records, size_t *n_modes) {
    if (!d) return NCPA_ERR_ARG;
    if (n_records) *n_records = (int)d->f.size();
    if (n_modes) *n_modes = d->kre.size();
    return NCPA_OK;
}

int ncpa_dispersion_arrays(const ncpa_dispersion_t *d, double *f, int *count, double *rho_s, double *rho_r,
                           double *k_re, double *k_im, double *psi_s, double *psi_r) {
    if (!d) return NCPA_ERR_ARG;
    auto cp = [](double *dst, const std::vector<double> &v) {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(double));
    };
    cp(f, d->f);
    cp(rho_s, d->rho_s);
    cp(rho_r, d->rho_r);
    cp(k_re, d->kre);
    cp(k_im, d->kim);
    cp(psi_s, d->ps);
    cp(psi_r, d->pr);
    if (count && !d->count.empty()) std::memcpy(count, d->count.data(), d->count.size() * sizeof(int));
    return NCPA_OK;
}

void ncpa_dispersion_free(ncpa_dispersion_t *d) { delete d; }

int ncpa_dispersion_write(const char *path, int append, int n_records, const double *f, const int *count,
                          const double *rho_s, const double *rho_r, const double *k_re, const double *k_im,
                          const double *psi_s, const double *psi_r, int n_threads) {
    if (!path || n_records < 0 || (n_records > 0 && (!f || !count || !rho_s || !rho_r))) return NCPA_ERR_ARG;
    std::vector<size_t> off(n_records + 1, 0);
    for (int r = 0; r < n_records; ++r) {
        if (count[r] < 0) return NCPA_ERR_ARG;
        off[r + 1] = off[r] + (size_t)count[r];
    }
    if (off[n_records] > 0 && (!k_re || !k_im || !psi_s || !psi_r)) return NCPA_ERR_ARG;
    std::vector<std::string> line(n_records);
    parallel_for(n_records, clamp_threads(n_threads), [&](int r) {
        const int n = count[r];
        std::string &s = line[r];
        s.resize(96 + 96 * (size_t)n);
        char *p = &s[0];
        p = put_e(p, f[r], 7);
        p = put_sep(p, 3);
        p = std::to_chars(p, p + 16, n).ptr;
        p = put_sep(p, 4);
        p = put_e(p, rho_s[r], 7);
        p = put_sep(p, 3);
        p = put_e(p, rho_r[r], 7);
        const size_t o = off[r];
        for (int j = 0; j < n; ++j) {
            p = put_sep(p, 3);
            p = put_e(p, k_re[o + j], 12);
            p = put_sep(p, 3);
            p = put_e(p, k_im[o + j], 12);
            p = put_sep(p, 3);
            p = put_e(p, psi_s[o + j], 12);
            p = put_sep(p, 3);
            p = put_e(p, psi_r[o + j], 12);
        }
        *p++ = '\n';
        s.resize((size_t)(p - &s[0]));
    });
    FILE *fp = std::fopen(path, append ? "a" : "w");
    if (!fp) return NCPA_ERR_ARG;
    bool ok = true;
    for (const std::string &s : line) ok = ok && std::fwrite(s.data(), 1, s.size(), fp) == s.size();
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? NCPA_OK : NCPA_ERR_ARG;
}

}  // extern "C"
