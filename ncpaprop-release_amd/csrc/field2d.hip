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
typedef unsigned u4 __attribute__((ext_vector_type(4)));

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

enum { PREC_FP64 = 0, PREC_FP32 = 1, PREC_BF16X3 = 2, PREC_BF16X6 = 3 };

// B rows for every mode: 1-D grid of total * ceil(nr / (256 * EMAT_R)) blocks; a thread
// forms EMAT_R consecutive ranges, the phasor exact at the first and advanced by the exact
// factor e^{i k dr} for the others.
#define EMAT_R 8
template <int PREC>
__global__ void __launch_bounds__(256) k_emat(const double *__restrict__ k2, const double *__restrict__ absn,
                                              const double *__restrict__ psis, double dr, int nr, int lossless,
                                              double ev_re, double ev_im, void *__restrict__ bh,
                                              void *__restrict__ bl) {
    const unsigned nb = (nr + 256 * EMAT_R - 1) / (256 * EMAT_R);
    const size_t o = blockIdx.x / nb;
    const int i0 = ((blockIdx.x % nb) * blockDim.x + threadIdx.x) * EMAT_R;
    if (i0 >= nr) return;
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
    // g = ev c
    const double gr = cr * ev_re - ci * ev_im, gi = cr * ev_im + ci * ev_re;
    const double r0 = (i0 + 1) * dr;
    double sn, cs;
    sincos(f2_reduce(kr * r0), &sn, &cs);
    const double a0 = exp(-ki * r0);
    double zr = a0 * cs, zi = a0 * sn;
    sincos(f2_reduce(kr * dr), &sn, &cs);
    const double a1 = exp(-ki * dr);
    const double mr = a1 * cs, mi = a1 * sn;
    for (int j = 0; j < EMAT_R && i0 + j < nr; ++j) {
        const int i = i0 + j;
        const double sc = 1000.0 / sqrt((i + 1) * dr);
        const double vr = (gr * zr - gi * zi) * sc, vi = (gr * zi + gi * zr) * sc;
        const double nz = zr * mr - zi * mi;
        zi = zr * mi + zi * mr;
        zr = nz;
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


// ------------------------------------------------------------------------------------
// split-BF16 contraction, tuned path
// ------------------------------------------------------------------------------------
// Operands are split v = v0 + v1 (+ v2) into bf16 parts (S = 2 or 3).  The products kept are
// those with part-index sum < S: S = 2 -> v0w0 + v0w1 + v1w0 (3 MFMAs, ~16-bit products),
// S = 3 -> adds v0w2 + v1w1 + v2w0 (6 MFMAs, ~24-bit, FP32-class).  Accumulation is FP32.
//
// Layouts (per problem p, written by k_asplit / k_esplit, zero padded), tiled so that one
// k-step of one 128-row tile is a single contiguous 8 KB chunk per part:
//   Ahat[s] : [Mp/128][Kp/32][128][32]  rows x modes    (Mp = M rounded up to 128)
//   Bhat[s] : [Np/128][Kp/32][128][32]  columns x modes (Np = 2 n_r rounded up to 128)
// Both MFMA operands read 8 consecutive modes as one 16-byte vector.  Kp = K rounded up to
// 32.  A block computes a 128 x 128 tile of C with 4 waves of 64 x 64 (4 x 4 MFMA 16x16x32
// tiles); tiles of the next k-step are fetched into registers while the current k-step runs
// from LDS.
struct SplitProb {
    size_t a_off, b_off, c_off;   // element offsets of the problem's Ahat, Bhat blocks and C block
    int K, Kp;
};

constexpr int TBM = 128, TBN = 128, TKS = 32, TLD = TKS + 8;   // LDS row: 40 bf16 = 80 B

// element offset of (row, k) in the tiled [row/128][k/32][row%128][k%32] layout
__host__ __device__ __forceinline__ size_t tiled_index(int row, int k, int Kp) {
    return (((size_t)(row >> 7) * (Kp / TKS) + (k >> 5)) * TBM + (row & 127)) * TKS + (k & 31);
}

__device__ __forceinline__ __bf16 bf_round(double v) { return (__bf16)(float)v; }

// A split: one thread per (row, k) of the padded block of problem p (grid.y = p)
template <int S>
__global__ void k_asplit(const double *__restrict__ psi, size_t ldpsi, const SplitProb *__restrict__ sp,
                         const int *__restrict__ col_off, int M, int Mp, size_t split_stride,
                         __bf16 *__restrict__ Ahat) {
    const int p = blockIdx.y;
    const SplitProb P = sp[p];
    const size_t n = (size_t)Mp * P.Kp;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const int row = (int)(e / P.Kp), k = (int)(e % P.Kp);
        double v = (row < M && k < P.K) ? psi[(size_t)row * ldpsi + col_off[p] + k] : 0.0;
        for (int s = 0; s < S; ++s) {
            const __bf16 h = bf_round(v);
            Ahat[s * split_stride + P.a_off + tiled_index(row, k, P.Kp)] = h;
            v -= (double)(float)h;
        }
    }
}

// E split: block = 32 modes x 64 ranges of one problem; phasors formed in FP64, split to
// bf16 parts and written transposed ([column][mode]) through LDS.
struct ETile {
    int p, k0;
};

template <int S>
__global__ void __launch_bounds__(256) k_esplit(const ETile *__restrict__ tiles, const SplitProb *__restrict__ sp,
                                                const int *__restrict__ mode_off, const double *__restrict__ k2,
                                                const double *__restrict__ absn, const double *__restrict__ psis,
                                                double dr, int nr, int lossless, double ev_re, double ev_im,
                                                size_t split_stride, __bf16 *__restrict__ Bhat) {
    __shared__ __attribute__((aligned(16))) __bf16 T[S][128][TKS + 8];
    const ETile t = tiles[blockIdx.y];
    const SplitProb P = sp[t.p];
    const int i0 = blockIdx.x * 64;
    const int mloc = threadIdx.x & 31;       // mode within the tile
    const int k = t.k0 + mloc;               // mode within the problem
    double kr = 0, ki = 0, cr = 0, ci = 0;
    const bool live = k < P.K;
    if (live) {
        const size_t o = (size_t)mode_off[t.p] + k;
        f2_csqrt(k2[o], absn[o], kr, ki);
        if (lossless) {
            ki = 0.0;
            cr = psis[o] / sqrt(kr);
        } else {
            double sr, si;
            f2_csqrt(kr, ki, sr, si);
            const double den = sr * sr + si * si;
            cr = psis[o] * sr / den;
            ci = -psis[o] * si / den;
        }
    }
    // 1000 / sqrt(r) of the tile's 64 ranges
    __shared__ double s_sc[64];
    if (threadIdx.x < 64) s_sc[threadIdx.x] = 1000.0 / sqrt((i0 + (int)threadIdx.x + 1) * dr);
    // g = ev c; the thread's ranges are i0 + w + 8q: e^{i k r} exact at q = 0, then advanced
    // by the exact factor e^{i k 8 dr} (a complex multiply instead of a sincos and an exp)
    const double gr = cr * ev_re - ci * ev_im, gi = cr * ev_im + ci * ev_re;
    double zr = 0.0, zi = 0.0, mr = 0.0, mi = 0.0;
    if (live) {
        const double r0 = (i0 + (threadIdx.x >> 5) + 1) * dr;
        double sn, cs;
        sincos(f2_reduce(kr * r0), &sn, &cs);
        const double a0 = exp(-ki * r0);
        zr = a0 * cs;
        zi = a0 * sn;
        sincos(f2_reduce(kr * (8.0 * dr)), &sn, &cs);
        const double a8 = exp(-ki * (8.0 * dr));
        mr = a8 * cs;
        mi = a8 * sn;
    }
    __syncthreads();
    for (int q = 0; q < 8; ++q) {
        const int il = (threadIdx.x >> 5) + 8 * q;   // range within the tile, 0..63
        const int i = i0 + il;
        double vr = 0.0, vi = 0.0;
        if (live && i < nr) {
            const double sc = s_sc[il];
            vr = (gr * zr - gi * zi) * sc;
            vi = (gr * zi + gi * zr) * sc;
        }
        {
            const double nz = zr * mr - zi * mi;
            zi = zr * mi + zi * mr;
            zr = nz;
        }
        for (int s = 0; s < S; ++s) {
            const __bf16 hr = bf_round(vr), hi = bf_round(vi);
            T[s][2 * il][mloc] = hr;
            T[s][2 * il + 1][mloc] = hi;
            vr -= (double)(float)hr;
            vi -= (double)(float)hi;
        }
    }
    __syncthreads();
    // write 128 columns x 32 modes per split as 16-byte vectors (4 per column)
    for (int s = 0; s < S; ++s)
        for (int v = threadIdx.x; v < 128 * 4; v += 256) {
            const int col = v >> 2, part = v & 3;
            const size_t n = 2 * (size_t)i0 + col;
            const uint4 val = *(const uint4 *)&T[s][col][part * 8];
            *(uint4 *)&Bhat[s * split_stride + P.b_off + tiled_index((int)n, t.k0 + part * 8, P.Kp)] = val;
        }
}

template <int S>
__global__ void __launch_bounds__(256, 2) k_gemm_split(const SplitProb *__restrict__ sp, const __bf16 *__restrict__ Ahat,
                                                       const __bf16 *__restrict__ Bhat, size_t split_stride_a,
                                                       size_t split_stride_b, float *__restrict__ C, size_t ldc,
                                                       int M, int N, int n_tiles_n) {
    __shared__ __attribute__((aligned(16))) __bf16 As[S][TBM][TLD];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[S][TBN][TLD];
    const SplitProb P = sp[blockIdx.z];
    // Blocks are dealt to the 8 XCDs round-robin (id % 8).  All M tiles of one N tile are
    // kept on one XCD, one after another, so its Bhat tile is fetched into one L2 only.
    const int n_tiles_n8 = (n_tiles_n + 7) & ~7;
    const int tiles_m = gridDim.x / n_tiles_n8;
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int tm = slot % tiles_m;
    const int tn = (slot / tiles_m) * 8 + xcd;
    if (tn >= n_tiles_n) return;
    const int m0 = tm * TBM, n0 = tn * TBN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    f4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = f4{0, 0, 0, 0};
    const __bf16 *Ab = Ahat + P.a_off, *Bb = Bhat + P.b_off;
    const int Kp = P.Kp;
    // each thread moves 2 x 16 B of A and 2 x 16 B of B per split and k-step:
    // tile rows 128 x 4 vectors = 512 vectors per operand and split.  Native vector type:
    // arrays of the HIP uint4 struct were placed in scratch memory (a round trip through
    // HBM every k-step); ext_vector_type arrays stay in registers.
    u4 ra[S][2], rb[S][2];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int v = tid + 256 * h;
                const int row = v >> 2, part = v & 3;
                ra[s][h] = *(const u4 *)&Ab[s * split_stride_a + tiled_index(m0 + row, k0 + part * 8, Kp)];
                rb[s][h] = *(const u4 *)&Bb[s * split_stride_b + tiled_index(n0 + row, k0 + part * 8, Kp)];
            }
    };
    auto stash = [&]() {
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int v = tid + 256 * h;
                const int row = v >> 2, part = v & 3;
                *(u4 *)&As[s][row][part * 8] = ra[s][h];
                *(u4 *)&Bs[s][row][part * 8] = rb[s][h];
            }
    };
    fetch(0);
    for (int k0 = 0; k0 < Kp; k0 += TKS) {
        __syncthreads();
        stash();
        __syncthreads();
        if (k0 + TKS < Kp) fetch(k0 + TKS);
        const int kq = 8 * (lane >> 4);
        bf8 af[S][4], bfr[S][4];
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                af[s][a] = *(const bf8 *)&As[s][wm + a * 16 + (lane & 15)][kq];
                bfr[s][a] = *(const bf8 *)&Bs[s][wn + a * 16 + (lane & 15)][kq];
            }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                // smallest products first
                if constexpr (S == 3) {
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[2][a], bfr[0][b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1][a], bfr[1][b], acc[a][b], 0, 0, 0);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][a], bfr[S - 1][b], acc[a][b], 0, 0, 0);
                }
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1][a], bfr[0][b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][a], bfr[1][b], acc[a][b], 0, 0, 0);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][a], bfr[0][b], acc[a][b], 0, 0, 0);
            }
    }
    float *Cp = C + P.c_off;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
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

namespace {

template <int S>
int field2d_split(ncpa_ctx *ctx, int P, int total, const int *hoff, const int *hcnt, const int *d_off,
                  const double *d_k2, const double *d_abs, const double *d_psis, const double *psi, int M,
                  double dr, int n_r, int lossless, float *out) {
    const int N = 2 * n_r;
    const int Mp = (M + TBM - 1) / TBM * TBM, Np = (N + TBN - 1) / TBN * TBN;
    std::vector<SplitProb> sp(P);
    std::vector<ETile> tiles;
    size_t sa = 0, sb = 0;
    for (int p = 0; p < P; ++p) {
        const int K = hcnt[p], Kp = (K + TKS - 1) / TKS * TKS;
        sp[p] = {sa, sb, (size_t)p * M * N, K, Kp};
        sa += (size_t)Mp * Kp;
        sb += (size_t)Np * Kp;
        for (int k0 = 0; k0 < Kp; k0 += TKS) tiles.push_back({p, k0});
    }
    SplitProb *d_sp = pool_alloc<SplitProb>(ctx, P);
    ETile *d_tiles = pool_alloc<ETile>(ctx, std::max<size_t>(tiles.size(), 1));
    __bf16 *Ahat = pool_alloc<__bf16>(ctx, std::max<size_t>(S * sa, 8));
    __bf16 *Bhat = pool_alloc<__bf16>(ctx, std::max<size_t>(S * sb, 8));
    auto cleanup = [&]() {
        ctx->pool.put(d_sp);
        ctx->pool.put(d_tiles);
        ctx->pool.put(Ahat);
        ctx->pool.put(Bhat);
    };
    if (!d_sp || !d_tiles || !Ahat || !Bhat) {
        cleanup();
        return set_error(ctx, NCPA_ERR_ALLOC, "ncpa_field2d: allocation failed");
    }
    hipError_t e = hipMemcpyAsync(d_sp, sp.data(), sizeof(SplitProb) * P, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && !tiles.empty())
        e = hipMemcpyAsync(d_tiles, tiles.data(), sizeof(ETile) * tiles.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
        cleanup();
        return hip_error(ctx, e, "ncpa_field2d copy");
    }
    const std::complex<double> I(0.0, 1.0);
    const std::complex<double> ev = 4 * PI * I * std::exp(-I * PI * 0.25) / std::sqrt(8.0 * PI);
    if (!tiles.empty()) {
        TimedLaunch tl(ctx, "emat");
        hipLaunchKernelGGL(k_asplit<S>, dim3(64, P), dim3(256), 0, ctx->stream, psi, (size_t)total, d_sp, d_off, M, Mp,
                           sa, Ahat);
        hipLaunchKernelGGL(k_esplit<S>, dim3((n_r + 63) / 64, (unsigned)tiles.size()), dim3(256), 0, ctx->stream,
                           d_tiles, d_sp, d_off, d_k2, d_abs, d_psis, dr, n_r, lossless, ev.real(), ev.imag(), sb,
                           Bhat);
    }
    {
        TimedLaunch tl(ctx, "field2d");
        const int tn = Np / TBN, tm = Mp / TBM;
        dim3 g((unsigned)(tm * ((tn + 7) & ~7)), 1, (unsigned)P);
        hipLaunchKernelGGL(k_gemm_split<S>, g, dim3(256), 0, ctx->stream, d_sp, Ahat, Bhat, sa, sb, out, (size_t)N, M,
                           N, tn);
    }
    e = hipGetLastError();
    cleanup();
    if (e != hipSuccess) return hip_error(ctx, e, "ncpa_field2d split launch");
    return NCPA_OK;
}

}  // namespace

extern "C" {

int ncpa_field2d(ncpa_ctx_t *ctx, ncpa_modes_t *m, const double *dev_psi_rows, int nrows, double dr, int n_r,
                 int lossless, int precision, void *dev_out) {
    if (!ctx || !m || !dev_psi_rows || nrows <= 0 || n_r <= 0 || !(dr > 0.0) || !dev_out ||
        precision < 0 || precision > 3)
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_field2d: bad arguments");
    int P = 0, total = 0;
    const int *hoff = nullptr, *hcnt = nullptr;
    const double *d_psis = nullptr, *d_k2 = nullptr, *d_abs = nullptr;
    ncpa_modes_internal(m, &P, &total, &hoff, &hcnt, &d_psis);
    ncpa_modes_device(m, nullptr, &d_k2, &d_abs, nullptr);
    if (total > 0 && !d_psis)
        return set_error(ctx, NCPA_ERR_STATE, "ncpa_field2d: call ncpa_modes_rows first (it supplies psi(zs))");
    if (precision >= PREC_BF16X3) {
        const int *d_off = nullptr;
        ncpa_modes_device(m, &d_off, nullptr, nullptr, nullptr);
        if (precision == PREC_BF16X3)
            return field2d_split<2>(ctx, P, total, hoff, hcnt, d_off, d_k2, d_abs, d_psis, dev_psi_rows, nrows, dr,
                                    n_r, lossless, (float *)dev_out);
        return field2d_split<3>(ctx, P, total, hoff, hcnt, d_off, d_k2, d_abs, d_psis, dev_psi_rows, nrows, dr, n_r,
                                lossless, (float *)dev_out);
    }
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
        dim3 ge((unsigned)(((n_r + 256 * EMAT_R - 1) / (256 * EMAT_R)) * (size_t)total));
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
