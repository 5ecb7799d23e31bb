// Broadband pulse synthesis from dispersion data (ModBB --propagation) on MI355X.
//
//   k_kpert        k_pert = sqrt(k^2 + i A) of device-resident modes (ModeSolver.cpp:263-284)
//   k_bb_transfer  H_s(r, f) = i e^{-i pi/4}/sqrt(8 pi r) sqrt(rho_r/rho_s)
//                              * sum_j psi_S psi_R e^{i k_j r}/sqrt(k_j)
//                  with the half-Hann taper on the top floor(0.1 Nfreq) bins
//                  (ModalBroadbandPropagator::compute_modal_sum, ModalBroadbandPropagator.cpp:245-277;
//                   BroadbandPropagator::half_hann, BroadbandPropagator.cpp:654-662).
//                  One block per (set, frequency); the per-mode factor psi_S psi_R / sqrt(k) is
//                  formed once and reused for a chunk of 16 ranges held in registers, so uneven
//                  mode counts cost one pass over each frequency's modes per 16 ranges.
//   k_bb_arg       arg_i = e^{-2 pi i f_i t0} S_i H_i, zero padded to NFFT
//                  (BroadbandPropagator::fft_pulse_prop, BroadbandPropagator.cpp:588-634)
//   k_fft          batched complex FFT, unnormalised, sign -1 (FFTW_FORWARD) or +1
//                  (FFTW_BACKWARD); Stockham autosort through LDS, radix 4/2/3/5 and a generic
//                  prime radix (NFFT = 4 Nfreq = 4100 for 1024 frequencies + DC has the factor 41)
//   k_bb_out       waveform = 2 Re(df * FFT(arg))  (ModalBroadbandPropagator.cpp:223-226)
// Phases Re(k) r and f t0 are formed in FP64 and reduced (mod 2 pi / mod 1) before sincos.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "ctx.h"
#include "internal.h"

namespace ncpa {

typedef double2 cplx;

__device__ __forceinline__ cplx cmul(cplx a, cplx b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return make_double2(a.x - b.x, a.y - b.y); }

// principal square root of x + i y
__device__ __forceinline__ void bb_csqrt(double x, double y, double &re, double &im) {
    const double m = hypot(x, y);
    if (x >= 0.0) {
        re = sqrt(0.5 * (m + x));
        im = (re == 0.0) ? 0.0 : y / (2.0 * re);
    } else {
        im = copysign(sqrt(0.5 * (m - x)), y);
        re = y / (2.0 * im);
    }
}

// e^{2 pi i x}, x reduced to [-1/2, 1/2] first
__device__ __forceinline__ cplx expi2pi(double x) {
    const double xr = x - rint(x);
    double s, c;
    sincospi(2.0 * xr, &s, &c);
    return make_double2(c, s);
}

// n_dev (optional): the count written on the device; n is then the buffer capacity
__global__ void k_kpert(int n, const int *__restrict__ n_dev, const double *__restrict__ k2,
                        const double *__restrict__ absorption, double *__restrict__ kre, double *__restrict__ kim) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (n_dev && i >= *n_dev)) return;
    double re, im;
    bb_csqrt(k2[i], absorption[i], re, im);
    kre[i] = re;
    kim[i] = im;
}

#define BB_RCHUNK 16

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// grid: one block (256 threads) per (set, frequency) = blockIdx.x; H[(s * nr + ir) * nf + f]
__global__ void __launch_bounds__(256) k_bb_transfer(int nf, const int *__restrict__ mbeg, const int *__restrict__ mcnt,
                                                     const double *__restrict__ kre, const double *__restrict__ kim,
                                                     const double *__restrict__ psr, int zero_attn,
                                                     const double *__restrict__ ranges, int nr,
                                                     const double *__restrict__ rho_fac, int smooth,
                                                     double *__restrict__ H) {
    __shared__ double red[4][BB_RCHUNK][2];
    const int sf = blockIdx.x;
    const int s = sf / nf, f = sf - s * nf;
    const int b0 = mbeg[sf], n = mcnt[sf];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // taper weight half_hann(nf - smooth, nf - 1, f)
    double taper = 1.0;
    if (smooth > 0 && f >= nf - smooth) {
        const int beg = nf - smooth, end = nf - 1;
        taper = (end > beg) ? 0.5 * (cos(PI * (double)(f - beg) / (double)(end - beg)) + 1.0) : 0.0;
    }
    for (int r0 = 0; r0 < nr; r0 += BB_RCHUNK) {
        const int nrc = min(BB_RCHUNK, nr - r0);
        double ar[BB_RCHUNK], ai[BB_RCHUNK];
#pragma unroll
        for (int t = 0; t < BB_RCHUNK; ++t) ar[t] = ai[t] = 0.0;
        for (int j = threadIdx.x; j < n; j += 256) {
            const double kr = kre[b0 + j];
            const double ki = zero_attn ? 0.0 : kim[b0 + j];
            double qr, qi;
            bb_csqrt(kr, ki, qr, qi);
            const double sc = psr[b0 + j] / (qr * qr + qi * qi);
            const double wr = sc * qr, wi = -sc * qi;     // psi_S psi_R / sqrt(k)
#pragma unroll
            for (int t = 0; t < BB_RCHUNK; ++t) {
                if (t < nrc) {
                    const double r = ranges[r0 + t];
                    const double ph = kr * r;
                    const double nq = rint(ph * 0.15915494309189535);
                    const double red2 = fma(-nq, 2.4492935982947064e-16, fma(-nq, 6.283185307179586, ph));
                    double sn, cs;
                    sincos(red2, &sn, &cs);
                    const double amp = exp(-ki * r);
                    const double er = cs * amp, ei = sn * amp;
                    ar[t] += er * wr - ei * wi;
                    ai[t] += er * wi + ei * wr;
                }
            }
        }
#pragma unroll
        for (int t = 0; t < BB_RCHUNK; ++t) {
            const double sr = wave_sum(ar[t]);
            const double si = wave_sum(ai[t]);
            if (lane == 0) {
                red[wv][t][0] = sr;
                red[wv][t][1] = si;
            }
        }
        __syncthreads();
        if (threadIdx.x < nrc) {
            const int t = threadIdx.x;
            const double sr = red[0][t][0] + red[1][t][0] + red[2][t][0] + red[3][t][0];
            const double si = red[0][t][1] + red[1][t][1] + red[2][t][1] + red[3][t][1];
            const double r = ranges[r0 + t];
            // i e^{-i pi/4} = sin(pi/4) + i cos(pi/4)
            const double c = 0.7071067811865476 / sqrt(8.0 * PI * r) * rho_fac[s] * taper;
            const size_t o = ((size_t)s * nr + r0 + t) * nf + f;
            H[2 * o] = c * (sr - si);
            H[2 * o + 1] = c * (sr + si);
        }
        __syncthreads();
    }
}

// arg[b][i] = e^{-2 pi i f_i t0_b} S_i H_b[i] for i < nf, 0 for nf <= i < nfft;
// t0 per range (b = s * nr + ir)
__global__ void k_bb_arg(int n_tr, int nf, int nfft, const double *__restrict__ H, const double *__restrict__ S,
                         const double *__restrict__ fvec, const double *__restrict__ t0, int nr,
                         double *__restrict__ arg) {
    const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (tid >= (size_t)n_tr * nfft) return;
    const int i = (int)(tid % nfft);
    const int b = (int)(tid / nfft);
    cplx v = make_double2(0.0, 0.0);
    if (i < nf) {
        const cplx tph = expi2pi(-fvec[i] * t0[b % nr]);
        const cplx s = make_double2(S[2 * i], S[2 * i + 1]);
        const size_t h = (size_t)b * nf + i;
        v = cmul(cmul(tph, s), make_double2(H[2 * h], H[2 * h + 1]));
    }
    arg[2 * tid] = v.x;
    arg[2 * tid + 1] = v.y;
}

// ------------------------------------------------------------------------------------
// Stockham autosort FFT: one block per transform, ping-pong between two N-point buffers
// (LDS when 2 N complex doubles fit, else a per-transform global work buffer).  Stage with
// radix R and span Ns: butterfly j reads src[j + r N/R], twiddles by e^{sgn 2 pi i k r/(Ns R)}
// (k = j mod Ns), does an R-point DFT, writes dst[(j/Ns) Ns R + k + q Ns].
// ------------------------------------------------------------------------------------
struct FftPlan {
    int n;
    int nstages;
    int radix[24];
};

template <int R>
__device__ __forceinline__ void dft_small(cplx *v, double sgn) {
    if constexpr (R == 2) {
        const cplx a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const cplx a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        const cplx b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
        const cplx b1r = make_double2(-sgn * b1.y, sgn * b1.x);   // (sgn i) b1
        v[0] = cadd(a0, b0);
        v[2] = csub(a0, b0);
        v[1] = cadd(a1, b1r);
        v[3] = csub(a1, b1r);
    } else {
        cplx out[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            cplx acc = v[0];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                double s, c;
                sincospi(sgn * 2.0 * (double)((r * q) % R) / R, &s, &c);
                acc = cadd(acc, cmul(v[r], make_double2(c, s)));
            }
            out[q] = acc;
        }
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = out[q];
    }
}

template <int R>
__device__ __forceinline__ void stockham_stage(const cplx *src, cplx *dst, int N, int Ns, double sgn) {
    const int nb = N / R;
    for (int j = threadIdx.x; j < nb; j += blockDim.x) {
        cplx v[R];
        const int k = j % Ns;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const cplx x = src[j + r * nb];
            if (r == 0 || k == 0) {
                v[r] = x;
            } else {
                double s, c;
                const long long span = (long long)Ns * R;
                sincospi(sgn * 2.0 * (double)(((long long)k * r) % span) / (double)span, &s, &c);
                v[r] = cmul(x, make_double2(c, s));
            }
        }
        dft_small<R>(v, sgn);
        const int idxD = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[idxD + r * Ns] = v[r];
    }
}

// generic radix R (any factor): twiddle and R-point DFT folded into one exponent
// (k r + Ns (r q mod R)) / (Ns R)
__device__ void stockham_stage_generic(const cplx *src, cplx *dst, int N, int R, int Ns, double sgn) {
    const int nb = N / R;
    const long long span = (long long)Ns * R;
    for (int j = threadIdx.x; j < nb; j += blockDim.x) {
        const int k = j % Ns;
        const int idxD = (j / Ns) * Ns * R + k;
        for (int q = 0; q < R; ++q) {
            cplx acc = make_double2(0.0, 0.0);
            for (int r = 0; r < R; ++r) {
                const long long num = ((long long)k * r + (long long)Ns * ((r * q) % R)) % span;
                double s, c;
                sincospi(sgn * 2.0 * (double)num / (double)span, &s, &c);
                acc = cadd(acc, cmul(src[j + r * nb], make_double2(c, s)));
            }
            dst[idxD + q * Ns] = acc;
        }
    }
}

__global__ void __launch_bounds__(256) k_fft(FftPlan plan, const double *__restrict__ in, double *__restrict__ out,
                                             double *__restrict__ work, int use_lds, double sgn) {
    extern __shared__ cplx lds[];
    const int N = plan.n;
    const size_t base = (size_t)blockIdx.x * N;
    cplx *b0 = use_lds ? lds : (cplx *)work + 2 * base;
    cplx *b1 = use_lds ? lds + N : (cplx *)work + 2 * base + N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) b0[i] = make_double2(in[2 * (base + i)], in[2 * (base + i) + 1]);
    __syncthreads();
    int Ns = 1;
    cplx *src = b0, *dst = b1;
    for (int st = 0; st < plan.nstages; ++st) {
        const int R = plan.radix[st];
        if (R == 4) stockham_stage<4>(src, dst, N, Ns, sgn);
        else if (R == 2) stockham_stage<2>(src, dst, N, Ns, sgn);
        else if (R == 3) stockham_stage<3>(src, dst, N, Ns, sgn);
        else if (R == 5) stockham_stage<5>(src, dst, N, Ns, sgn);
        else stockham_stage_generic(src, dst, N, R, Ns, sgn);
        __syncthreads();
        Ns *= R;
        cplx *t = src;
        src = dst;
        dst = t;
    }
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        out[2 * (base + i)] = src[i].x;
        out[2 * (base + i) + 1] = src[i].y;
    }
}

// waveform: wave[i] = 2 Re(df * spec[i])
__global__ void k_bb_out(size_t n, const double *__restrict__ spec, double df, double *__restrict__ wave) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) wave[i] = 2.0 * (df * spec[2 * i]);
}

}  // namespace ncpa

// ------------------------------------------------------------------------------------
// Fixed device memory of ncpa_bb_pulse_noalloc: one array per GPU, held until the process
// ends.  In host code `scratch` names a host-side stand-in, not device memory: host code
// uses only the per-device address hipGetSymbolAddress returns (g_scratch_dev).
// ------------------------------------------------------------------------------------
namespace {
constexpr size_t SCRATCH_DOUBLES = NCPA_PULSE_SCRATCH_BYTES / sizeof(double);
constexpr int SCRATCH_MAX_DEVICES = 64;
}  // namespace
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

namespace {
std::mutex g_scratch_mu;
double *g_scratch_dev[SCRATCH_MAX_DEVICES] = {};
}  // namespace

namespace ncpa {
// Device address of `scratch` on ctx's device, looked up once per device with that device
// current (each GPU holds its own copy of the array at its own address).
int pulse_scratch_setup(ncpa_ctx *ctx, double **out) {
    *out = nullptr;
    if (ctx->device < 0 || ctx->device >= SCRATCH_MAX_DEVICES)
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_bb_pulse_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[ctx->device]) {
        NCPA_HIP(ctx, hipSetDevice(ctx->device));
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return hip_error(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "address of the fixed pulse array");
        g_scratch_dev[ctx->device] = static_cast<double *>(ptr);
    }
    *out = g_scratch_dev[ctx->device];
    return NCPA_OK;
}
}  // namespace ncpa

using namespace ncpa;

namespace {

int make_plan(ncpa_ctx *ctx, int n, FftPlan &p) {
    p.n = n;
    p.nstages = 0;
    int m = n;
    const int small[] = {4, 2, 3, 5};
    for (int r : small)
        while (m % r == 0) {
            if (p.nstages >= 24) return set_error(ctx, NCPA_ERR_ARG, "fft: too many stages");
            p.radix[p.nstages++] = r;
            m /= r;
        }
    for (int r = 7; m > 1; r += 2) {
        if (r > 4096) return set_error(ctx, NCPA_ERR_ARG, "fft: a prime factor above 4096 is not supported");
        while (m % r == 0) {
            if (p.nstages >= 24) return set_error(ctx, NCPA_ERR_ARG, "fft: too many stages");
            p.radix[p.nstages++] = r;
            m /= r;
        }
    }
    return NCPA_OK;
}

}  // namespace

namespace ncpa {
// Sets the FFT kernel up for n-point transforms (its LDS attribute when 2 n complex doubles
// need more than 64 KB) and reports the global work buffer it needs: 0 when the ping-pong
// buffers fit in LDS, else 4 n batch doubles.
int bb_prepare_fft(ncpa_ctx *ctx, int n, int batch, size_t *work_doubles) {
    FftPlan plan;
    int rc = make_plan(ctx, n, plan);
    if (rc) return rc;
    int max_lds = 0;
    NCPA_HIP(ctx, hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
    const size_t lds = 2 * (size_t)n * sizeof(cplx);
    const bool use_lds = lds <= (size_t)max_lds;
    if (use_lds && lds > 65536)
        NCPA_HIP(ctx, hipFuncSetAttribute((const void *)k_fft, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    *work_doubles = use_lds ? 0 : 4 * (size_t)n * batch;
    return NCPA_OK;
}

// batch unnormalised n-point FFTs with the global work buffer from bb_prepare_fft (no
// allocation); timed: under the context's kernel timing
int bb_launch_fft(ncpa_ctx *ctx, const double *in, double *out, int n, int batch, int sign, double *work,
                  bool timed) {
    FftPlan plan;
    int rc = make_plan(ctx, n, plan);
    if (rc) return rc;
    const size_t lds = 2 * (size_t)n * sizeof(cplx);
    const bool use_lds = work == nullptr;
    if (timed) {
        TimedLaunch tl(ctx, "fft");
        hipLaunchKernelGGL(k_fft, dim3(batch), dim3(256), use_lds ? lds : 0, ctx->stream, plan, in, out, work,
                           use_lds ? 1 : 0, (double)(sign < 0 ? -1.0 : 1.0));
    } else {
        hipLaunchKernelGGL(k_fft, dim3(batch), dim3(256), use_lds ? lds : 0, ctx->stream, plan, in, out, work,
                           use_lds ? 1 : 0, (double)(sign < 0 ? -1.0 : 1.0));
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(ctx, e, "k_fft launch");
    return NCPA_OK;
}

void bb_launch_kpert(hipStream_t st, int n_cap, const int *n_dev, const double *k2, const double *absn, double *kre,
                     double *kim) {
    hipLaunchKernelGGL(k_kpert, dim3((n_cap + 255) / 256), dim3(256), 0, st, n_cap, n_dev, k2, absn, kre, kim);
}

void bb_launch_transfer(hipStream_t st, int n_bins, int n_freq, const int *beg, const int *cnt, const double *kre,
                        const double *kim, const double *psr, int zero_attn, const double *ranges, int n_ranges,
                        const double *fac, double *H) {
    const int smooth = (int)std::floor(0.1 * n_freq);   // ModalBroadbandPropagator.cpp:272
    hipLaunchKernelGGL(k_bb_transfer, dim3((unsigned)n_bins), dim3(256), 0, st, n_freq, beg, cnt, kre, kim, psr,
                       zero_attn ? 1 : 0, ranges, n_ranges, fac, smooth, H);
}

void bb_launch_arg(hipStream_t st, int n_tr, int n_freq, int nfft, const double *H, const double *S, const double *f,
                   const double *t0, int n_ranges, double *arg) {
    const size_t n = (size_t)n_tr * nfft;
    hipLaunchKernelGGL(k_bb_arg, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n_tr, n_freq, nfft, H, S, f, t0,
                       n_ranges, arg);
}

void bb_launch_out(hipStream_t st, size_t n, const double *spec, double df, double *wave) {
    hipLaunchKernelGGL(k_bb_out, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, spec, df, wave);
}
}  // namespace ncpa

namespace {

int launch_fft(ncpa_ctx *ctx, const double *in, double *out, int n, int batch, int sign) {
    size_t need = 0;
    int rc = bb_prepare_fft(ctx, n, batch, &need);
    if (rc) return rc;
    double *work = nullptr;
    if (need) {
        work = pool_alloc<double>(ctx, need);
        if (!work) return set_error(ctx, NCPA_ERR_ALLOC, "fft: work buffer allocation failed");
    }
    rc = bb_launch_fft(ctx, in, out, n, batch, sign, work, true);
    ctx->pool.put(work);
    return rc;
}

template <class T>
int upload_tmp(ncpa_ctx *ctx, const T *host, size_t n, T **dev) {
    *dev = pool_alloc<T>(ctx, n);
    if (!*dev) return set_error(ctx, NCPA_ERR_ALLOC, "broadband: allocation failed");
    hipError_t e = hipMemcpyAsync(*dev, host, sizeof(T) * n, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return hip_error(ctx, e, "broadband: upload");
    return NCPA_OK;
}

// one unnormalised FFT of a host vector (2 n interleaved doubles) on the device, in place
int host_fft(ncpa_ctx *ctx, std::vector<double> &data, int n, int sign) {
    double *d_in = nullptr;
    double *d_out = pool_alloc<double>(ctx, 2 * (size_t)n);
    int rc = d_out ? NCPA_OK : set_error(ctx, NCPA_ERR_ALLOC, "fft: allocation failed");
    if (!rc) rc = upload_tmp(ctx, data.data(), 2 * (size_t)n, &d_in);
    if (!rc) rc = launch_fft(ctx, d_in, d_out, n, 1, sign);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(data.data(), d_out, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = hip_error(ctx, e, "fft: download");
    }
    ctx->pool.put(d_in);
    ctx->pool.put(d_out);
    return rc;
}

// BroadbandPropagator::pulse_spec_fit (BroadbandPropagator.cpp:420-438)
void pulse_spec_fit(double scale, double f, double &re, double &im) {
    const double fnorm = 0.780776406404415;
    const double x = scale * fnorm * f;
    const double a = x * std::exp((-x - x * x) / 2) * scale / 33.4336 * 4.0 * PI * 5.0;   // times e^{-i pi/2}
    const double ph = 2.0 * PI * x * 0.01;
    // a e^{-i pi/2} e^{i ph} = a (sin ph - i cos ph)
    re = a * std::sin(ph);
    im = -a * std::cos(ph);
}

// BroadbandPropagator::model_pulse_shape (BroadbandPropagator.cpp:469-476)
double model_pulse_shape(double power, double scale, double t) {
    const double x = 2.094 * (1.0 + power) * t / scale;
    return std::pow(x, power) * (1 - x / (1.0 + power)) * std::exp(-x) / 1.387566e+03;
}

// e^{2 pi i x} on the host with x reduced mod 1 first
void expi2pi_host(double x, double &c, double &s) {
    const double xr = x - std::nearbyint(x);
    c = std::cos(2.0 * PI * xr);
    s = std::sin(2.0 * PI * xr);
}

}  // namespace

extern "C" {

int ncpa_fft_batch(ncpa_ctx_t *ctx, const double *dev_in, double *dev_out, int n, int batch, int sign) {
    if (!ctx) return NCPA_ERR_ARG;
    if (!dev_in || !dev_out || n < 1 || batch < 1 || (sign != 1 && sign != -1))
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_fft_batch: bad arguments");
    return launch_fft(ctx, dev_in, dev_out, n, batch, sign);
}

int ncpa_kpert(ncpa_ctx_t *ctx, const double *dev_k2, const double *dev_absorption, int n, double *dev_kre,
               double *dev_kim) {
    if (!ctx) return NCPA_ERR_ARG;
    if (n < 0 || (n > 0 && (!dev_k2 || !dev_absorption || !dev_kre || !dev_kim)))
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_kpert: bad arguments");
    if (n == 0) return NCPA_OK;
    hipLaunchKernelGGL(k_kpert, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, nullptr, dev_k2, dev_absorption,
                       dev_kre, dev_kim);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(ctx, e, "k_kpert launch");
    return NCPA_OK;
}

int ncpa_bb_transfer(ncpa_ctx_t *ctx, int n_sets, int n_freq, const int *mode_begin, const int *mode_count,
                     const double *dev_kre, const double *dev_kim, const double *dev_psi_sr, int zero_attn,
                     const double *ranges_m, int n_ranges, const double *rho_ratio, double *dev_H) {
    if (!ctx) return NCPA_ERR_ARG;
    if (n_sets < 1 || n_freq < 1 || !mode_begin || !mode_count || !dev_kre || !dev_kim || !dev_psi_sr ||
        !ranges_m || n_ranges < 1 || !rho_ratio || !dev_H)
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_bb_transfer: bad arguments");
    if ((size_t)n_sets * n_freq > 0x7fffffff) return set_error(ctx, NCPA_ERR_ARG, "ncpa_bb_transfer: too many bins");
    for (int i = 0; i < n_ranges; ++i)
        if (!(ranges_m[i] > 0.0)) return set_error(ctx, NCPA_ERR_ARG, "ncpa_bb_transfer: ranges must be > 0");
    const size_t nb = (size_t)n_sets * n_freq;
    for (size_t i = 0; i < nb; ++i)
        if (mode_count[i] < 0 || mode_begin[i] < 0)
            return set_error(ctx, NCPA_ERR_ARG, "ncpa_bb_transfer: negative mode begin/count");
    std::vector<double> fac(n_sets);
    for (int s = 0; s < n_sets; ++s) {
        if (!(rho_ratio[s] >= 0.0)) return set_error(ctx, NCPA_ERR_ARG, "ncpa_bb_transfer: rho ratio must be >= 0");
        fac[s] = std::sqrt(rho_ratio[s]);
    }
    int *d_beg = nullptr, *d_cnt = nullptr;
    double *d_r = nullptr, *d_fac = nullptr;
    int rc = upload_tmp(ctx, mode_begin, nb, &d_beg);
    if (!rc) rc = upload_tmp(ctx, mode_count, nb, &d_cnt);
    if (!rc) rc = upload_tmp(ctx, ranges_m, (size_t)n_ranges, &d_r);
    if (!rc) rc = upload_tmp(ctx, fac.data(), (size_t)n_sets, &d_fac);
    if (!rc) {
        const int smooth = (int)std::floor(0.1 * n_freq);   // ModalBroadbandPropagator.cpp:272
        {
            TimedLaunch tl(ctx, "bb_transfer");
            hipLaunchKernelGGL(k_bb_transfer, dim3((unsigned)nb), dim3(256), 0, ctx->stream, n_freq, d_beg, d_cnt,
                               dev_kre, dev_kim, dev_psi_sr, zero_attn ? 1 : 0, d_r, n_ranges, d_fac, smooth, dev_H);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = hip_error(ctx, e, "k_bb_transfer launch");
    }
    ctx->pool.put(d_beg);
    ctx->pool.put(d_cnt);
    ctx->pool.put(d_r);
    ctx->pool.put(d_fac);
    return rc;
}

int ncpa_bb_pulse(ncpa_ctx_t *ctx, int n_sets, int n_ranges, int n_freq, const double *dev_H,
                  const double *spectrum, const double *f_vec, const double *ranges_m, double max_celerity,
                  int nfft, double *dev_wave) {
    if (!ctx) return NCPA_ERR_ARG;
    if (n_sets < 1 || n_ranges < 1 || n_freq < 2 || !dev_H || !spectrum || !f_vec || !ranges_m || !dev_wave ||
        !(max_celerity > 0.0))
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_bb_pulse: bad arguments");
    if (nfft < n_freq) return set_error(ctx, NCPA_ERR_ARG, "fft too short (i.e. NFFT < Nfreq)");
    const int n_tr = n_sets * n_ranges;
    std::vector<double> t0(n_ranges);
    for (int i = 0; i < n_ranges; ++i) t0[i] = ranges_m[i] / max_celerity;   // ModalBroadbandPropagator.cpp:211
    const double df = f_vec[1] - f_vec[0];                                       // BroadbandPropagator.cpp:581
    double *d_S = nullptr, *d_f = nullptr, *d_t0 = nullptr;
    double *d_arg = pool_alloc<double>(ctx, 2 * (size_t)n_tr * nfft);
    double *d_spec = pool_alloc<double>(ctx, 2 * (size_t)n_tr * nfft);
    int rc = (d_arg && d_spec) ? NCPA_OK : set_error(ctx, NCPA_ERR_ALLOC, "ncpa_bb_pulse: allocation failed");
    if (!rc) rc = upload_tmp(ctx, spectrum, 2 * (size_t)n_freq, &d_S);
    if (!rc) rc = upload_tmp(ctx, f_vec, (size_t)n_freq, &d_f);
    if (!rc) rc = upload_tmp(ctx, t0.data(), (size_t)n_ranges, &d_t0);
    const size_t n = (size_t)n_tr * nfft;
    if (!rc) {
        hipLaunchKernelGGL(k_bb_arg, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, n_tr, n_freq, nfft,
                           dev_H, d_S, d_f, d_t0, n_ranges, d_arg);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = hip_error(ctx, e, "k_bb_arg launch");
    }
    if (!rc) rc = launch_fft(ctx, d_arg, d_spec, nfft, n_tr, -1);   // FFTW_FORWARD
    if (!rc) {
        hipLaunchKernelGGL(k_bb_out, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, n, d_spec, df,
                           dev_wave);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = hip_error(ctx, e, "k_bb_out launch");
    }
    ctx->pool.put(d_S);
    ctx->pool.put(d_f);
    ctx->pool.put(d_t0);
    ctx->pool.put(d_arg);
    ctx->pool.put(d_spec);
    return rc;
}

int ncpa_bb_pulse_noalloc(ncpa_ctx_t *ctx, int n_sets, int n_ranges, int n_freq, const double *dev_H,
                          const double *spectrum, const double *f_vec, const double *ranges_m,
                          double max_celerity, int nfft, double *dev_wave) {
    if (!ctx) return NCPA_ERR_ARG;
    if (n_sets < 1 || n_ranges < 1 || n_freq < 2 || !dev_H || !spectrum || !f_vec || !ranges_m || !dev_wave ||
        !(max_celerity > 0.0))
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_bb_pulse_noalloc: bad arguments");
    if (nfft < n_freq) return set_error(ctx, NCPA_ERR_ARG, "fft too short (i.e. NFFT < Nfreq)");
    // the same kernels and launches as ncpa_bb_pulse; the intermediates are carved out of
    // the fixed array: S [2 n_freq] | f [n_freq] | t0 [n_ranges] | arg | spec | FFT work,
    // each segment 256-byte aligned
    const size_t n_tr = (size_t)n_sets * n_ranges, n = n_tr * (size_t)nfft;
    size_t work = 0;
    int rc = bb_prepare_fft(ctx, nfft, (int)std::min<size_t>(n_tr, INT_MAX), &work);
    if (rc) return rc;
    auto seg = [](size_t doubles) { return (doubles + 31) / 32 * 32; };
    const size_t o_f = seg(2 * (size_t)n_freq), o_t0 = o_f + seg((size_t)n_freq), o_arg = o_t0 + seg((size_t)n_ranges);
    const size_t o_spec = o_arg + seg(2 * n), o_work = o_spec + seg(2 * n), need = o_work + seg(work);
    if (n_tr > (size_t)INT_MAX || need > SCRATCH_DOUBLES)
        return set_error(ctx, NCPA_ERR_ARG,
                         "ncpa_bb_pulse_noalloc: the job needs " + std::to_string(need * sizeof(double)) +
                             " bytes of fixed device memory, above the maximum of " +
                             std::to_string((size_t)NCPA_PULSE_SCRATCH_BYTES));
    double *base = nullptr;
    if ((rc = pulse_scratch_setup(ctx, &base))) return rc;
    NCPA_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<double> t0(n_ranges);
    for (int i = 0; i < n_ranges; ++i) t0[i] = ranges_m[i] / max_celerity;   // ModalBroadbandPropagator.cpp:211
    const double df = f_vec[1] - f_vec[0];                                       // BroadbandPropagator.cpp:581
    double *d_S = base, *d_f = base + o_f, *d_t0 = base + o_t0, *d_arg = base + o_arg, *d_spec = base + o_spec;
    double *d_work = work ? base + o_work : nullptr;
    // stream-ordered after this context's earlier use of the array; every element of arg,
    // spec and the FFT work buffer is written before it is read, so nothing is cleared.  The
    // host arrays are pageable: HIP stages them before hipMemcpyAsync returns (as for
    // ncpa_bb_pulse), so t0 may go out of scope at the end of this call.
    NCPA_HIP(ctx, hipMemcpyAsync(d_S, spectrum, sizeof(double) * 2 * n_freq, hipMemcpyHostToDevice, ctx->stream));
    NCPA_HIP(ctx, hipMemcpyAsync(d_f, f_vec, sizeof(double) * n_freq, hipMemcpyHostToDevice, ctx->stream));
    NCPA_HIP(ctx, hipMemcpyAsync(d_t0, t0.data(), sizeof(double) * n_ranges, hipMemcpyHostToDevice, ctx->stream));
    bb_launch_arg(ctx->stream, (int)n_tr, n_freq, nfft, dev_H, d_S, d_f, d_t0, n_ranges, d_arg);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(ctx, e, "k_bb_arg launch");
    if ((rc = bb_launch_fft(ctx, d_arg, d_spec, nfft, (int)n_tr, -1, d_work, false))) return rc;   // FFTW_FORWARD
    bb_launch_out(ctx->stream, n, d_spec, df, dev_wave);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_error(ctx, e, "k_bb_out launch");
    return NCPA_OK;
}


// BroadbandPropagator::get_source_spectrum (BroadbandPropagator.cpp:244-417)
int ncpa_bb_source(ncpa_ctx_t *ctx, int kind, const double *f_vec, int n_freq, int nfft, double f_center,
                   const double *input, int n_input, double *spectrum, double *pulse) {
    if (!ctx) return NCPA_ERR_ARG;
    if (!f_vec || n_freq < 3 || nfft < n_freq || !spectrum)
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_bb_source: need f_vec[>=3], NFFT >= Nfreq and a spectrum buffer");
    if ((kind == NCPA_SRC_SPECTRUM || kind == NCPA_SRC_WAVEFORM) && (!input || n_input < 2))
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_bb_source: spectrum/waveform sources need input rows");
    const double f_step = f_vec[2] - f_vec[1];                  // ModalBroadbandPropagator.cpp:65
    const double fmx = (double)nfft * f_step;                   // BroadbandPropagator.cpp:257
    const double dt = 1.0 / fmx;
    const double f_max = f_vec[n_freq - 1];
    std::vector<double> S(2 * (size_t)n_freq, 0.0);
    double scale = 0.0;
    if (kind == NCPA_SRC_PULSE1 || kind == NCPA_SRC_PULSE2) {   // BroadbandPropagator.cpp:270-283
        if (f_center < 0.0) f_center = f_max / 5.0;
        if (f_center > 0 && f_center <= f_max / 5.0 + 1.0e-16) {
            scale = 1.0 / f_center;
        } else {
            return set_error(ctx, NCPA_ERR_ARG,
                             "f_center is too large: for the built-in pulse f_center should be set smaller than f_max/5");
        }
    }
    std::vector<double> pv;   // time-domain source pulse, 2 nfft
    switch (kind) {
    case NCPA_SRC_IMPULSE:      // ones, delayed by Nfreq/2 (BroadbandPropagator.cpp:260-264)
        for (int i = 0; i < n_freq; ++i) expi2pi_host(f_vec[i] * ((double)n_freq / 2.0), S[2 * i], S[2 * i + 1]);
        break;
    case NCPA_SRC_PULSE1:       // BroadbandPropagator.cpp:294-296
        for (int i = 0; i < n_freq; ++i) {
            double fr, fi, c, s;
            pulse_spec_fit(scale, f_vec[i], fr, fi);
            expi2pi_host(f_vec[i] * scale, c, s);
            S[2 * i] = 0.5 * (fr * c - fi * s);
            S[2 * i + 1] = 0.5 * (fr * s + fi * c);
        }
        break;
    case NCPA_SRC_PULSE2: {     // model_pulse_fft with power 8 (BroadbandPropagator.cpp:300-310, 440-466)
        std::vector<double> p(2 * (size_t)nfft, 0.0);
        for (int i = 0; i < nfft; ++i) p[2 * i] = model_pulse_shape(8.0, scale, i * dt);
        int rc = host_fft(ctx, p, nfft, +1);
        if (rc) return rc;
        for (int i = 0; i < 2 * n_freq; ++i) S[i] = p[i] * dt;
        break;
    }
    case NCPA_SRC_SPECTRUM:     // BroadbandPropagator.cpp:313-325, 478-509
        if (n_input > nfft) return set_error(ctx, NCPA_ERR_ARG, "source spectrum has more points than NFFT");
        if (std::fabs(input[0] - f_vec[0]) > 1.0e-8)
            return set_error(ctx, NCPA_ERR_ARG,
                             "the frequencies of the source spectrum do not match the dispersion file");
        for (int i = 0; i < std::min(n_input, n_freq); ++i) {
            S[2 * i] = input[3 * i + 1];
            S[2 * i + 1] = input[3 * i + 2];
        }
        break;
    case NCPA_SRC_WAVEFORM: {   // BroadbandPropagator.cpp:327-353, 513-537
        if (n_input > nfft) return set_error(ctx, NCPA_ERR_ARG, "source waveform has more samples than NFFT");
        const double dtw = input[2] - input[0];
        pv.assign(2 * (size_t)nfft, 0.0);
        for (int i = 0; i < n_input; ++i) pv[2 * i] = input[2 * i + 1];
        std::vector<double> a = pv;
        int rc = host_fft(ctx, a, nfft, +1);
        if (rc) return rc;
        for (int i = 0; i < 2 * n_freq; ++i) S[i] = a[i] * dtw;
        break;
    }
    default:
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_bb_source: unknown source type");
    }
    if (kind != NCPA_SRC_WAVEFORM && pulse) {
        // f_step * FFTW_FORWARD of the spectrum zero padded to NFFT (BroadbandPropagator.cpp:354-375);
        // the left pad is empty because f_vec[0] = 0
        pv.assign(2 * (size_t)nfft, 0.0);
        std::copy(S.begin(), S.end(), pv.begin());
        int rc = host_fft(ctx, pv, nfft, -1);
        if (rc) return rc;
        for (auto &v : pv) v *= f_step;
    }
    std::copy(S.begin(), S.end(), spectrum);
    if (pulse) std::copy(pv.begin(), pv.end(), pulse);
    return NCPA_OK;
}

}  // extern "C"
