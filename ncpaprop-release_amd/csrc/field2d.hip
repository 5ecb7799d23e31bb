// 2-D transmission-loss field as a batched real x complex contraction on MFMA (gfx950).
//
// For every problem p, output rows z_k and ranges r_i = (i+1) dr (ModeSolver::getTLoss2D,
// src/libmodes/ModeSolver.cpp:524-544):
//     p(r_i, z_k) = 1000 ev / sqrt(r_i) * sum_m psi_m(z_k) psi_m(zs) exp(i k_m r_i) / sqrt(k_m)
// with ev = 4 pi i exp(-i pi/4)/sqrt(8 pi).  Written as C = A B with
//     A[k][m] = psi_m(z_k)                          (rows x modes, real)
//     B[m][2i], B[m][2i+1] = Re, Im of psi_m(zs)/sqrt(k_m) exp(i k_m r_i) 1000 ev/sqrt(r_i)
// so C[k][2i], C[k][2i+1] is the field itself.  The phase k_m r_i is formed in FP64 and
// reduced mod 2 pi before the sincos; B is then stored in the contraction's precision.
//
// Three contraction precisions, same tiling (block 64 x 64 of C, 4 waves of 32 x 32, each a
// 2 x 2 grid of 16 x 16 MFMA tiles, A and B tiles staged in LDS):
//   FP64    v_mfma_f64_16x16x4f64, FP64 output
//   FP32    v_mfma_f32_16x16x4f32, FP32 output
//   BF16X3  A = Ah + Al, B = Bh + Bl split into bf16 pairs, C = Ah Bh + Ah Bl + Al Bh with
//           v_mfma_f32_16x16x32_bf16 and FP32 accumulation, FP32 output
// MFMA operand layouts (checked on gfx950 with dense operands): lane l holds A row l%16 /
// B column l%16 at k-group l/16 (one k per lane for FP64/FP32 16x16x4, 8 consecutive k for
// bf16 16x16x32).  D lane l holds column l%16; its register v is row 4(l/16) + v for the
// FP32 and BF16 instructions but row (l/16) + 4v for v_mfma_f64_16x16x4f64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <string>
#include <vector>

#include "common.h"
#include "ctx.h"

namespace ncpa {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr double F2_TWO_PI_HI = 6.283185307179586;
constexpr double F2_TWO_PI_LO = 2.4492935982947064e-16;
constexpr double F2_INV_TWO_PI = 0.15915494309189535;

__device__ __forceinline__ double f2_reduce(double x) {
    const double nq = rint(x * F2_INV_TWO_PI);
    return fma(-nq, F2_TWO_PI_LO, fma(-nq, F2_TWO_PI_HI, x));
}

__device__ __forceinline__ void f2_csqrt(double x, double y, double &re, double &im) {
    const double m = hypot(x, y);
    if (x >= 0.0) {
        re = sqrt(0.5 * (m + x));
        im = (re == 0.0) ? 0.0 : y / (2.0 * re);
    } else {
        im = copysign(sqrt(0.5 * (m - x)), y);
        re = y / (2.0 * im);
    }
}

enum { PREC_FP64 = 0, PREC_FP32 = 1, PREC_BF16X3 = 2 };

// B rows for every mode: 1-D grid of total * ceil(nr/256) blocks
template <int PREC>
__global__ void __launch_bounds__(256) k_emat(const double *__restrict__ k2, const double *__restrict__ absn,
                                              const double *__restrict__ psis, double dr, int nr, int lossless,
                                              double ev_re, double ev_im, void *__restrict__ bh,
                                              void *__restrict__ bl) {
    const unsigned nb = (nr + 255) / 256;
    const size_t o = blockIdx.x / nb;
    const int i = (blockIdx.x % nb) * blockDim.x + threadIdx.x;
    if (i >= nr) return;
    double kr, ki;
    f2_csqrt(k2[o], absn[o], kr, ki);              // k_pert = sqrt(k^2 + i A)
    double cr, ci;
    if (lossless) {
        ki = 0.0;
        cr = psis[o] / sqrt(kr);
        ci = 0.0;
    } else {
        double sr, si;
        f2_csqrt(kr, ki, sr, si);
        const double den = sr * sr + si * si;
        cr = psis[o] * sr / den;
        ci = -psis[o] * si / den;
    }
    const double r = (i + 1) * dr;
    double sn, cs;
    sincos(f2_reduce(kr * r), &sn, &cs);
    const double amp = exp(-ki * r) * (1000.0 / sqrt(r));
    // (c e^{i k r}) (ev) amp
    const double er = cs * amp, ei = sn * amp;
    const double tr = cr * er - ci * ei, ti = cr * ei + ci * er;
    const double vr = tr * ev_re - ti * ev_im, vi = tr * ev_im + ti * ev_re;
    const size_t idx = o * (size_t)(2 * nr) + 2 * (size_t)i;
    if (PREC == PREC_FP64) {
        double *b = (double *)bh;
        b[idx] = vr;
        b[idx + 1] = vi;
    } else if (PREC == PREC_FP32) {
        float *b = (float *)bh;
        b[idx] = (float)vr;
        b[idx + 1] = (float)vi;
    } else {
        __bf16 *h = (__bf16 *)bh, *l = (__bf16 *)bl;
        const __bf16 hr = (__bf16)(float)vr, hi = (__bf16)(float)vi;
        h[idx] = hr;
        h[idx + 1] = hi;
        l[idx] = (__bf16)(float)(vr - (double)(float)hr);
        l[idx + 1] = (__bf16)(float)(vi - (double)(float)hi);
    }
}

// A (rows x total) from the FP64 eigenvector rows, in the contraction precision
template <int PREC>
__global__ void k_amat(const double *__restrict__ psi, size_t n, void *__restrict__ ah, void *__restrict__ al) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double v = psi[i];
        if (PREC == PREC_FP32) {
            ((float *)ah)[i] = (float)v;
        } else {
            const __bf16 h = (__bf16)(float)v;
            ((__bf16 *)ah)[i] = h;
            ((__bf16 *)al)[i] = (__bf16)(float)(v - (double)(float)h);
        }
    }
}

struct GemmProb {
    size_t a_off;   // column offset into A (first mode of the problem)
    size_t b_off;   // row offset into B
    size_t c_off;   // element offset of the problem's output block
    int K;          // modes of the problem
};

constexpr int BM = 64, BN = 64;

// FP64 / FP32 16x16x4 MFMA GEMM: C[M x N] = A[:, a_off : a_off+K] B[b_off : b_off+K, :]
template <class T, int KS>
__global__ void __launch_bounds__(256) k_gemm_x4(const GemmProb *__restrict__ gp, const T *__restrict__ A,
                                                 size_t lda, const T *__restrict__ B, size_t ldb,
                                                 T *__restrict__ C, size_t ldc, int M, int N) {
    typedef typename std::conditional<sizeof(T) == 8, d4, f4>::type acc_t;
    __shared__ T As[BM][KS + 1];
    __shared__ T Bs[KS][BN + 1];
    const GemmProb P = gp[blockIdx.z];
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    acc_t acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
    const T *Ap = A + P.a_off;
    const T *Bp = B + P.b_off * ldb;
    for (int k0 = 0; k0 < P.K; k0 += KS) {
        for (int e = tid; e < BM * KS; e += 256) {
            const int r = e / KS, c = e % KS;
            const int gr = m0 + r, gc = k0 + c;
            As[r][c] = (gr < M && gc < P.K) ? Ap[(size_t)gr * lda + gc] : (T)0;
        }
        for (int e = tid; e < KS * BN; e += 256) {
            const int r = e / BN, c = e % BN;
            const int gr = k0 + r, gc = n0 + c;
            Bs[r][c] = (gr < P.K && gc < N) ? Bp[(size_t)gr * ldb + gc] : (T)0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < KS; kk += 4) {
            T af[2], bf[2];
            for (int a = 0; a < 2; ++a) af[a] = As[wm + a * 16 + (lane & 15)][kk + (lane >> 4)];
            for (int b = 0; b < 2; ++b) bf[b] = Bs[kk + (lane >> 4)][wn + b * 16 + (lane & 15)];
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) {
                    if constexpr (sizeof(T) == 8)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
                    else
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a], bf[b], acc[a][b], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    T *Cp = C + P.c_off;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                const int rr = (sizeof(T) == 8) ? (lane >> 4) + 4 * v : 4 * (lane >> 4) + v;
                const int r = m0 + wm + a * 16 + rr;
                const int c = n0 + wn + b * 16 + (lane & 15);
                if (r < M && c < N) Cp[(size_t)r * ldc + c] = acc[a][b][v];
            }
}

// split-BF16 (3 products) MFMA GEMM, FP32 accumulate and output
__global__ void __launch_bounds__(256) k_gemm_bf16x3(const GemmProb *__restrict__ gp, const __bf16 *__restrict__ Ah,
                                                     const __bf16 *__restrict__ Al, size_t lda,
                                                     const __bf16 *__restrict__ Bh, const __bf16 *__restrict__ Bl,
                                                     size_t ldb, float *__restrict__ C, size_t ldc, int M, int N) {
    constexpr int KS = 32;
    // k contiguous for both operands so each lane reads its 8 k-values as one 16-byte vector
    __shared__ __attribute__((aligned(16))) __bf16 Ash[BM][KS + 8], Asl[BM][KS + 8];
    __shared__ __attribute__((aligned(16))) __bf16 Bsh[BN][KS + 8], Bsl[BN][KS + 8];
    const GemmProb P = gp[blockIdx.z];
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = f4{0, 0, 0, 0};
    const __bf16 zero = (__bf16)0.0f;
    for (int k0 = 0; k0 < P.K; k0 += KS) {
        for (int e = tid; e < BM * KS; e += 256) {
            const int r = e / KS, c = e % KS;
            const int gr = m0 + r, gc = k0 + c;
            const bool ok = gr < M && gc < P.K;
            const size_t gi = (size_t)gr * lda + P.a_off + gc;
            Ash[r][c] = ok ? Ah[gi] : zero;
            Asl[r][c] = ok ? Al[gi] : zero;
        }
        for (int e = tid; e < KS * BN; e += 256) {
            const int r = e / BN, c = e % BN;     // r: k, c: column
            const int gr = k0 + r, gc = n0 + c;
            const bool ok = gr < P.K && gc < N;
            const size_t gi = (P.b_off + gr) * ldb + gc;
            Bsh[c][r] = ok ? Bh[gi] : zero;
            Bsl[c][r] = ok ? Bl[gi] : zero;
        }
        __syncthreads();
        const int kq = 8 * (lane >> 4);
        bf8 ah[2], al[2], bh[2], bl[2];
        for (int a = 0; a < 2; ++a) {
            const int r = wm + a * 16 + (lane & 15);
            ah[a] = *(const bf8 *)&Ash[r][kq];
            al[a] = *(const bf8 *)&Asl[r][kq];
        }
        for (int b = 0; b < 2; ++b) {
            const int c = wn + b * 16 + (lane & 15);
            bh[b] = *(const bf8 *)&Bsh[c][kq];
            bl[b] = *(const bf8 *)&Bsl[c][kq];
        }
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
            }
        __syncthreads();
    }
    float *Cp = C + P.c_off;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v) {
                const int r = m0 + wm + a * 16 + 4 * (lane >> 4) + v;
                const int c = n0 + wn + b * 16 + (lane & 15);
                if (r < M && c < N) Cp[(size_t)r * ldc + c] = acc[a][b][v];
            }
}

}  // namespace ncpa

using namespace ncpa;

// accessors into the modes handle (defined in ess.hip)
extern "C" int ncpa_modes_device(ncpa_modes_t *m, const int **off, const double **k2, const double **ab,
                                 const double **psi_sr);
int ncpa_modes_internal(ncpa_modes_t *m, int *P, int *total, const int **h_offsets, const int **h_counts,
                        const double **d_psis);

extern "C" {

int ncpa_field2d(ncpa_ctx_t *ctx, ncpa_modes_t *m, const double *dev_psi_rows, int nrows, double dr, int n_r,
                 int lossless, int precision, void *dev_out) {
    if (!ctx || !m || !dev_psi_rows || nrows <= 0 || n_r <= 0 || !(dr > 0.0) || !dev_out ||
        precision < 0 || precision > 2)
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_field2d: bad arguments");
    int P = 0, total = 0;
    const int *hoff = nullptr, *hcnt = nullptr;
    const double *d_psis = nullptr, *d_k2 = nullptr, *d_abs = nullptr;
    ncpa_modes_internal(m, &P, &total, &hoff, &hcnt, &d_psis);
    ncpa_modes_device(m, nullptr, &d_k2, &d_abs, nullptr);
    if (total > 0 && !d_psis)
        return set_error(ctx, NCPA_ERR_STATE, "ncpa_field2d: call ncpa_modes_rows first (it supplies psi(zs))");
    const size_t ldb = 2 * (size_t)n_r, ldc = ldb, lda = (size_t)std::max(total, 1);
    std::vector<GemmProb> gp(P);
    for (int p = 0; p < P; ++p) gp[p] = {(size_t)hoff[p], (size_t)hoff[p], (size_t)p * nrows * ldc, hcnt[p]};
    GemmProb *d_gp = pool_alloc<GemmProb>(ctx, P);
    const size_t esz = precision == PREC_FP64 ? 8 : (precision == PREC_FP32 ? 4 : 2);
    void *bh = ctx->pool.get(std::max<size_t>(total, 1) * ldb * esz);
    void *bl = precision == PREC_BF16X3 ? ctx->pool.get(std::max<size_t>(total, 1) * ldb * esz) : nullptr;
    void *ah = precision == PREC_FP64 ? nullptr : ctx->pool.get((size_t)nrows * lda * esz);
    void *al = precision == PREC_BF16X3 ? ctx->pool.get((size_t)nrows * lda * esz) : nullptr;
    auto cleanup = [&]() {
        ctx->pool.put(d_gp);
        ctx->pool.put(bh);
        ctx->pool.put(bl);
        ctx->pool.put(ah);
        ctx->pool.put(al);
    };
    if (!d_gp || !bh || (precision == PREC_BF16X3 && (!bl || !al)) || (precision != PREC_FP64 && !ah)) {
        cleanup();
        return set_error(ctx, NCPA_ERR_ALLOC, "ncpa_field2d: allocation failed");
    }
    hipError_t e = hipMemcpyAsync(d_gp, gp.data(), sizeof(GemmProb) * P, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
        cleanup();
        return hip_error(ctx, e, "ncpa_field2d copy");
    }
    const std::complex<double> I(0.0, 1.0);
    const std::complex<double> ev = 4 * PI * I * std::exp(-I * PI * 0.25) / std::sqrt(8.0 * PI);
    if (total > 0) {
        TimedLaunch tl(ctx, "emat");
        dim3 ge((unsigned)(((n_r + 255) / 256) * (size_t)total));
        if (precision == PREC_FP64)
            hipLaunchKernelGGL(k_emat<PREC_FP64>, ge, dim3(256), 0, ctx->stream, d_k2, d_abs, d_psis, dr, n_r,
                               lossless, ev.real(), ev.imag(), bh, bl);
        else if (precision == PREC_FP32)
            hipLaunchKernelGGL(k_emat<PREC_FP32>, ge, dim3(256), 0, ctx->stream, d_k2, d_abs, d_psis, dr, n_r,
                               lossless, ev.real(), ev.imag(), bh, bl);
        else
            hipLaunchKernelGGL(k_emat<PREC_BF16X3>, ge, dim3(256), 0, ctx->stream, d_k2, d_abs, d_psis, dr, n_r,
                               lossless, ev.real(), ev.imag(), bh, bl);
        const size_t na = (size_t)nrows * lda;
        if (precision == PREC_FP32)
            hipLaunchKernelGGL(k_amat<PREC_FP32>, dim3(1024), dim3(256), 0, ctx->stream, dev_psi_rows, na, ah, al);
        else if (precision == PREC_BF16X3)
            hipLaunchKernelGGL(k_amat<PREC_BF16X3>, dim3(1024), dim3(256), 0, ctx->stream, dev_psi_rows, na, ah, al);
    }
    {
        TimedLaunch tl(ctx, "field2d");
        dim3 gg((unsigned)((ldb + BN - 1) / BN), (unsigned)((nrows + BM - 1) / BM), (unsigned)P);
        if (precision == PREC_FP64)
            hipLaunchKernelGGL((k_gemm_x4<double, 16>), gg, dim3(256), 0, ctx->stream, d_gp, dev_psi_rows, lda,
                               (const double *)bh, ldb, (double *)dev_out, ldc, nrows, (int)ldb);
        else if (precision == PREC_FP32)
            hipLaunchKernelGGL((k_gemm_x4<float, 16>), gg, dim3(256), 0, ctx->stream, d_gp, (const float *)ah, lda,
                               (const float *)bh, ldb, (float *)dev_out, ldc, nrows, (int)ldb);
        else
            hipLaunchKernelGGL(k_gemm_bf16x3, gg, dim3(256), 0, ctx->stream, d_gp, (const __bf16 *)ah,
                               (const __bf16 *)al, lda, (const __bf16 *)bh, (const __bf16 *)bl, ldb,
                               (float *)dev_out, ldc, nrows, (int)ldb);
    }
    e = hipGetLastError();
    cleanup();
    if (e != hipSuccess) return hip_error(ctx, e, "ncpa_field2d launch");
    return NCPA_OK;
}

}  // extern "C"
