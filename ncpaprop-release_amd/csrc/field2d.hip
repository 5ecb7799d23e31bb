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
// Contraction precisions:
//   FP64    v_mfma_f64_16x16x4f64, FP64 output      (block 64 x 64, operands in LDS)
//   FP32    v_mfma_f32_16x16x4f32, FP32 output      (same tiling)
//   split   BF16x3, BF16x6, FP16x3, FP16x4: operands split into 16-bit parts, products on
//           v_mfma_f32_16x16x32_{bf16,f16} with FP32 accumulation, FP32 output (block
//           128 x 128, tiled operand layout; see "split 16-bit contractions" below)
// MFMA operand layouts (checked on gfx950 with dense operands): lane l holds A row l%16 /
// B column l%16 at k-group l/16 (one k per lane for FP64/FP32 16x16x4, 8 consecutive k for
// the 16-bit 16x16x32 forms).  D lane l holds column l%16; its register v is row 4(l/16) + v for the
// FP32 and BF16 instructions but row (l/16) + 4v for v_mfma_f64_16x16x4f64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"
#include "ctx.h"

namespace ncpa {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
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

enum { PREC_FP64 = 0, PREC_FP32 = 1, PREC_BF16X3 = 2, PREC_BF16X6 = 3, PREC_FP16X3 = 4, PREC_FP16X4 = 5,
       PREC_BF16X6F = 6, PREC_FP16X4F = 7 };

// B rows for every mode: 1-D grid of total * ceil(nr / (256 * EMAT_R)) blocks; a thread
// forms EMAT_R consecutive ranges, the phasor exact at the first and advanced by the exact
// factor e^{i k dr} for the others.
#define EMAT_R 8
template <int PREC>
__global__ void __launch_bounds__(256) k_emat(const double *__restrict__ k2, const double *__restrict__ absn,
                                              const double *__restrict__ psis, double dr, int nr, int lossless,
                                              double ev_re, double ev_im, void *__restrict__ bh) {
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
        } else {
            float *b = (float *)bh;
            b[idx] = (float)vr;
            b[idx + 1] = (float)vi;
        }
    }
}

// A (rows x total) from the FP64 eigenvector rows, in FP32
__global__ void k_amat(const double *__restrict__ psi, size_t n, float *__restrict__ ah) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        ah[i] = (float)psi[i];
}

// word 3 of a raw buffer resource on gfx950 (32-bit data format, no swizzle, no stride)
constexpr int RSRC_W3 = 0x00020000;

typedef __attribute__((address_space(3))) void lds_void;

// Block id -> (M tile, N tile).  Blocks are dealt to the 8 XCDs round-robin (id % 8).  The M
// tiles fall into row_groups groups (1, 2, 4 or 8) and the XCDs into as many classes: XCD x
// takes group x % row_groups and every (8 / row_groups)-th N tile from x / row_groups, its
// slots running over the group's M tiles first, so consecutive blocks on an XCD share a B
// tile and the XCD's L2 holds 1 / row_groups of A plus the B tiles in flight.  The groups
// deal the M tiles out in passes of row_groups, alternating direction (groups 0 1 1 0 0 1 ...
// for two): the top row tiles, whose all-zero k-steps the folded contraction skips, are
// shared out evenly, and so is any work that grows or shrinks along the rows.
// Grid: 8 x ceil(tiles_m / row_groups) x ceil(n_tiles / (8 / row_groups)) blocks.
__device__ __forceinline__ void xcd_tile(int bid, int tiles_m, int row_groups, int &tm, int &tn) {
    const int per = (tiles_m + row_groups - 1) / row_groups;
    const int xcd = bid & 7, slot = bid >> 3;
    const int pass = slot % per, g = xcd % row_groups;
    tm = pass * row_groups + ((pass & 1) ? row_groups - 1 - g : g);
    tn = (slot / per) * (8 / row_groups) + xcd / row_groups;
}

// s_waitcnt immediate (gfx9 encoding) that waits until at most n vector memory operations
// are outstanding and leaves the export and LDS/scalar counters alone
constexpr int waitcnt_vm(int n) { return (n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8); }

// 16 bytes per lane from base + voff + soff (bytes; zero beyond `bytes`) into LDS at
// dst + 16 lane, through a raw buffer resource over [base, base + bytes).
__device__ __forceinline__ void buffer_load_lds16(const void *base, unsigned bytes, lds_void *dst, unsigned voff,
                                                  unsigned soff) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(__builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)bytes, RSRC_W3),
                                             dst, 16, voff, soff, 0, 0);
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

// The FP64 / FP32 contraction in 128 x 128 blocks: 4 waves of 64 x 64 (4 x 4 MFMA 16x16x4
// tiles), KS modes per k-step, operands straight from global memory into two LDS stages
// (buffer loads into LDS, 16 B per lane), the next k-step's loads in flight while the current
// one runs.  Each output element is summed over k in the order of k_gemm_x4 (one MFMA per 4
// modes, ascending), so the two kernels agree to the bit.
//   A (rows x modes, row-major, ld lda): a resource over the whole array.  Columns past the
//     problem's K read the next problem's (finite) values, rows past M read zero.
//   B (modes x columns, ld ldb): a resource over the problem's K rows only, so the rows
//     past K read zero and cancel A's extra columns exactly.  Columns past N only reach
//     outputs that are not stored.
// LDS rows of A hold KS elements in 16-byte groups, the group g of row r stored at
// g ^ f(r), so that 16 rows read at one k fall on distinct banks.
template <class T, int KS>
__global__ void __launch_bounds__(256, 2) k_gemm_x4t(const GemmProb *__restrict__ gp, const T *__restrict__ A,
                                                     size_t lda, unsigned a_bytes, const T *__restrict__ B,
                                                     size_t ldb, T *__restrict__ C, size_t ldc, int M, int N,
                                                     int tiles_m, int tiles_n, int row_groups) {
    typedef typename std::conditional<sizeof(T) == 8, d4, f4>::type acc_t;
    constexpr int G = 16 / (int)sizeof(T);   // elements per 16-byte group
    constexpr int AG = KS / G;               // groups per A row
    constexpr int AR = 64 / AG;              // A rows per load instruction
    constexpr int AI = 128 / AR / 4;         // A load instructions per wave and k-step
    constexpr int BG = 128 / G;              // groups per B row
    constexpr int BR = 64 / BG;              // B rows per load instruction
    constexpr int BI = KS / BR / 4;          // B load instructions per wave and k-step
    constexpr int RB = 256 / (KS * (int)sizeof(T));   // A rows per 256 bytes (one pass of the banks)
    static_assert(AI >= 1 && BI >= 1 && AG * AR == 64 && BG * BR == 64, "tile geometry");
    typedef T ATile[128][KS];
    typedef T BTile[KS][128];
    __shared__ __attribute__((aligned(16))) ATile As0, As1;
    __shared__ __attribute__((aligned(16))) BTile Bs0, Bs1;
    const GemmProb P = gp[blockIdx.z];
    int tm, tn;
    xcd_tile(blockIdx.x, tiles_m, row_groups, tm, tn);
    if (tn >= tiles_n || tm >= tiles_m) return;
    const int m0 = tm * 128, n0 = tn * 128;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    acc_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
    const int K = P.K;
    const T *Bp = B + P.b_off * ldb;
    const unsigned b_bytes = (unsigned)((size_t)K * ldb * sizeof(T));
    unsigned voff_a[AI], voff_b[BI];
#pragma unroll
    for (int j = 0; j < AI; ++j) {
        const int row = (wave * AI + j) * AR + lane / AG, gs = lane % AG;
        const int g = gs ^ ((row / RB) % AG);
        voff_a[j] = (unsigned)(((size_t)row * lda + (size_t)g * G) * sizeof(T));
    }
#pragma unroll
    for (int j = 0; j < BI; ++j) {
        const int kr = (wave * BI + j) * BR + lane / BG, cg = lane % BG;
        voff_b[j] = (unsigned)(((size_t)kr * ldb + (size_t)cg * G) * sizeof(T));
    }
    auto issue = [&](ATile &Ax, BTile &Bx, int k0) {
        const unsigned so_a = (unsigned)(((size_t)m0 * lda + P.a_off + k0) * sizeof(T));
        const unsigned so_b = (unsigned)(((size_t)k0 * ldb + n0) * sizeof(T));
#pragma unroll
        for (int j = 0; j < AI; ++j)
            buffer_load_lds16(A, a_bytes, (lds_void *)&Ax[(wave * AI + j) * AR][0], voff_a[j], so_a);
#pragma unroll
        for (int j = 0; j < BI; ++j)
            buffer_load_lds16(Bp, b_bytes, (lds_void *)&Bx[(wave * BI + j) * BR][0], voff_b[j], so_b);
    };
    auto step = [&](ATile &Ac, BTile &Bc, ATile &An, BTile &Bn, int k0) {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (k0 + KS < K) issue(An, Bn, k0 + KS);
#pragma unroll
        for (int kk = 0; kk < KS; kk += 4) {
            const int k = kk + (lane >> 4);
            T af[4], bf[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int r = wm + a * 16 + (lane & 15);
                af[a] = Ac[r][(((k / G) ^ ((r / RB) % AG)) * G) + k % G];
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) bf[b] = Bc[k][wn + b * 16 + (lane & 15)];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    if constexpr (sizeof(T) == 8)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
                    else
                        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a], bf[b], acc[a][b], 0, 0, 0);
                }
        }
    };
    if (K > 0) issue(As0, Bs0, 0);
    for (int k0 = 0; k0 < K; k0 += 2 * KS) {
        step(As0, Bs0, As1, Bs1, k0);
        if (k0 + KS < K) step(As1, Bs1, As0, Bs0, k0 + KS);
    }
    T *Cp = C + P.c_off;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int rr = (sizeof(T) == 8) ? (lane >> 4) + 4 * v : 4 * (lane >> 4) + v;
                const int r = m0 + wm + a * 16 + rr;
                const int c = n0 + wn + b * 16 + (lane & 15);
                if (r < M && c < N) Cp[(size_t)r * ldc + c] = acc[a][b][v];
            }
}

// ------------------------------------------------------------------------------------
// split 16-bit contractions
// ------------------------------------------------------------------------------------
// Operands are split v = v0 + v1 (+ v2) into 16-bit parts (S = 2 or 3), BF16 or FP16.
//   BF16, S = 2: v0w0 + v0w1 + v1w0 (3 MFMAs, ~2^-16 relative)
//   BF16, S = 3: adds v0w2 + v1w1 + v2w0 (6 MFMAs, ~2^-24, FP32-class)
//   FP16, S = 2: v0w0 + v0w1 + v1w0 (3 MFMAs, ~2^-22) or with v1w1 (4 MFMAs: the products
//                of the parts are exact, only the FP32 accumulation rounds).  FP16 has a
//                narrow range, so A rows and B columns are scaled by powers of two first
//                (largest |value| of the row / range to [2^14, 2^15)) and the scales are
//                undone on the FP32 output: C[z][r] = acc 2^-(SA[z] + SB[r]).
// Accumulation is FP32 in every case.
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

// LDS rows of a k-step are 32 halves (64 B) with the four 16-byte groups of row r stored at
// group g ^ ((r >> 2) & 3): a 16-lane write of 4 rows x 4 groups and a 16-lane fragment read
// of 16 rows x 1 group both touch 16 distinct 4-bank quads of the 64 banks.
constexpr int TBM = 128, TBN = 128, TKS = 32, TLD = TKS;
__device__ __forceinline__ int lds_group(int row, int g) { return 8 * (g ^ ((row >> 2) & 3)); }

// element offset of (row, k) in the tiled [row/128][k/32][row%128][k%32] layout
__host__ __device__ __forceinline__ size_t tiled_index(int row, int k, int Kp) {
    return (((size_t)(row >> 7) * (Kp / TKS) + (k >> 5)) * TBM + (row & 127)) * TKS + (k & 31);
}

// One 16-bit part of a split: the part's bits and its value.  BF16 (8-bit significand,
// FP32 range) or FP16 (11-bit significand; operands are pre-scaled by powers of two so
// the largest magnitudes sit at 2^14).
// The conversion goes through FP32: two hardware conversions, where a direct FP64 -> FP16
// rounding is a long integer sequence on gfx950 (the empty asm keeps the compiler from
// merging the two back into it).  The rare double rounding is harmless: the caller
// subtracts the value actually stored, so the parts still sum to the input.
template <bool F16>
__device__ __forceinline__ unsigned short split_part(double v, double &back) {
    if (F16) {
        float f = (float)v;
        asm volatile("" : "+v"(f));
        const _Float16 h = (_Float16)f;
        back = (double)(float)h;
        return __builtin_bit_cast(unsigned short, h);
    }
    float f = (float)v;
    asm volatile("" : "+v"(f));
    const __bf16 h = (__bf16)f;
    back = (double)(float)h;
    return __builtin_bit_cast(unsigned short, h);
}

// The parts of a pair of values (re, im, or two consecutive modes) as one 32-bit word, the
// first in the low half, and the values they stand for.  BF16: one packed conversion (round to nearest even, as the scalar one)
// for both; FP16: two scalar conversions (the packed FP16 form rounds toward zero).
template <bool F16>
__device__ __forceinline__ unsigned split_pair(double vr, double vi, double &br, double &bi) {
    if (F16) {
        const unsigned lo = split_part<true>(vr, br), hi = split_part<true>(vi, bi);
        return lo | (hi << 16);
    }
    typedef float fl2 __attribute__((ext_vector_type(2)));
    typedef __bf16 bh2 __attribute__((ext_vector_type(2)));
    float fr = (float)vr, fi = (float)vi;
    asm volatile("" : "+v"(fr), "+v"(fi));
    const unsigned w = __builtin_bit_cast(unsigned, __builtin_convertvector(fl2{fr, fi}, bh2));
    br = (double)__builtin_bit_cast(float, w << 16);
    bi = (double)__builtin_bit_cast(float, w & 0xffff0000u);
    return w;
}

#define SPLIT_TOP 14   // scaled operands: max |value| in [2^14, 2^15), FP16 max 65504

// exponent e with max * 2^e in [2^14, 2^15) (0 for an all-zero set)
__device__ __forceinline__ int split_scale(double mx) {
    return (mx > 0.0 && isfinite(mx)) ? SPLIT_TOP - ilogb(mx) : 0;
}

// FP16 only: per (problem, row) scale of A from the largest |psi| of the row.  One wave
// per row; grid (ceil(Mp / 4), P), block 256.
__global__ void __launch_bounds__(256) k_rowscale(const double *__restrict__ psi, size_t ldpsi,
                                                  const SplitProb *__restrict__ sp, const int *__restrict__ col_off,
                                                  int M, int Mp, int *__restrict__ SA) {
    const int p = blockIdx.y;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= Mp) return;
    const SplitProb P = sp[p];
    double mx = 0.0;
    if (row < M)
        for (int k = lane; k < P.K; k += 64) mx = fmax(mx, fabs(psi[(size_t)row * ldpsi + col_off[p] + k]));
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) mx = fmax(mx, __shfl_xor(mx, d));
    if (lane == 0) SA[(size_t)p * Mp + row] = split_scale(mx);
}

// A split, one wave per chunk of 128 rows x 32 modes of problem p (one k-step of one row tile,
// the tiled layout's unit: 64 contiguous bytes per row and part), two rows per step: lanes
// 0-31 take the modes of one row and lanes 32-63 those of the next, so each load reads two
// 256-byte runs of psi and each store writes two 64-byte rows of a part.  With F16 every row
// is first scaled by 2^SA[row].  With flag non-null the wave also writes the chunk's flag: 1
// if any part holds a nonzero, else 0.  Modes above their work item's evanescent cut have
// exactly zero rows there (+0 or -0, after the mode's sign flip), so whole chunks of the top
// row tiles are zero; their products are zeros that sum to +0, which the folded contraction
// skips (sign bits are masked: -0 counts as zero).  Grid (ceil(chunks of the largest problem
// / 4), P), block 256.
template <int S, bool F16>
__global__ void __launch_bounds__(256) k_asplit(const double *__restrict__ psi, size_t ldpsi,
                                                const SplitProb *__restrict__ sp, const int *__restrict__ col_off,
                                                int M, int Mp, size_t split_stride, const int *__restrict__ SA,
                                                unsigned short *__restrict__ Ahat, int *__restrict__ flag) {
    const int p = blockIdx.y;
    const SplitProb P = sp[p];
    const int nkt = P.Kp / TKS;
    const int c = (int)blockIdx.x * 4 + ((int)threadIdx.x >> 6), lane = (int)threadIdx.x & 63;
    if (c >= (Mp / TBM) * nkt) return;
    const int tm = c / nkt, k0 = (c % nkt) * TKS;
    const int k = lane & 31;
    const bool live_k = k0 + k < P.K;
    const double *src = psi + col_off[p] + k0 + k;
    unsigned short *dst = Ahat + P.a_off + (size_t)c * (TBM * TKS) + k;
    unsigned any = 0;
    for (int rl = lane >> 5; rl < TBM; rl += 2) {
        const int row = tm * TBM + rl;
        double v = (row < M && live_k) ? src[(size_t)row * ldpsi] : 0.0;
        if (F16) v = ldexp(v, row < M ? SA[(size_t)p * Mp + row] : 0);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            double back;
            const unsigned short h = split_part<F16>(v, back);
            dst[s * split_stride + (size_t)rl * TKS] = h;
            any |= h & 0x7fffu;
            v -= back;
        }
    }
    if (flag) {
        const bool nz = __ballot(any != 0u) != 0ull;
        if (lane == 0) flag[(P.a_off >> 12) + c] = nz ? 1 : 0;
    }
}

// Per mode: k_pert = kr + i ki and g = ev psi(zs) / sqrt(k_pert), the mode's factor in
// B[m, r] = g e^{i k r} 1000 / sqrt(r).
__device__ __forceinline__ void mode_factor(double k2, double absn, double psis, int lossless, double ev_re,
                                            double ev_im, double &kr, double &ki, double &gr, double &gi) {
    f2_csqrt(k2, absn, kr, ki);
    double cr, ci;
    if (lossless) {
        ki = 0.0;
        cr = psis / sqrt(kr);
        ci = 0.0;
    } else {
        double sr, si;
        f2_csqrt(kr, ki, sr, si);
        const double den = sr * sr + si * si;
        cr = psis * sr / den;
        ci = -psis * si / den;
    }
    gr = cr * ev_re - ci * ev_im;
    gi = cr * ev_im + ci * ev_re;
}

// FP16 only: per mode log2 |g| and ki (the envelope |B[m, r]| = |g| e^{-ki r} 1000/sqrt(r))
__global__ void __launch_bounds__(256) k_mode_env(const double *__restrict__ k2, const double *__restrict__ absn,
                                                  const double *__restrict__ psis, size_t n, int lossless,
                                                  double ev_re, double ev_im, double *__restrict__ lg,
                                                  double *__restrict__ kim) {
    const size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (o >= n) return;
    double kr, ki, gr, gi;
    mode_factor(k2[o], absn[o], psis[o], lossless, ev_re, ev_im, kr, ki, gr, gi);
    const double a = hypot(gr, gi);
    lg[o] = a > 0.0 ? log2(a) : -INFINITY;
    kim[o] = ki;
}

// FP16 only: per (problem, range) scale of B from the envelope max_m |B[m, r]|;
// grid (ceil(nr / 256), P)
__global__ void __launch_bounds__(256) k_colscale(const SplitProb *__restrict__ sp, const int *__restrict__ mode_off,
                                                  const double *__restrict__ lg, const double *__restrict__ kim,
                                                  double dr, int nr, int nr_pad, int *__restrict__ SB) {
    __shared__ double s_lg[256], s_ki[256];
    const int p = blockIdx.y;
    const SplitProb P = sp[p];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const double r = (i + 1) * dr;
    double best = -INFINITY;
    for (int m0 = 0; m0 < P.K; m0 += 256) {
        const int m = m0 + threadIdx.x;
        if (m < P.K) {
            s_lg[threadIdx.x] = lg[mode_off[p] + m];
            s_ki[threadIdx.x] = kim[mode_off[p] + m];
        }
        __syncthreads();
        const int cnt = min(256, P.K - m0);
        for (int q = 0; q < cnt; ++q) best = fmax(best, fma(-s_ki[q], r * 1.4426950408889634, s_lg[q]));
        __syncthreads();
    }
    if (i >= nr_pad) return;
    // log2 max|B| = best + log2(1000 / sqrt(r)); the recurrence may exceed the envelope by
    // a few ulp, far inside the factor 2 of headroom below 65504
    const int e = (i < nr && best > -INFINITY) ? SPLIT_TOP - (int)floor(best + log2(1000.0 / sqrt(r))) : 0;
    SB[(size_t)p * nr_pad + i] = e;
}

// Per mode, for the E split: k_pert = kr + i ki, the factor g = gr + i gi (mode_factor) and the
// exact per-range step e^{i k dr} = fr + i fi (FP64 phase reduced mod 2 pi, sincos, exp), each in
// an array of n (structure of arrays, for coalesced reads by the E tiles of the mode).
__global__ void __launch_bounds__(256) k_mode_fac(const double *__restrict__ k2, const double *__restrict__ absn,
                                                  const double *__restrict__ psis, size_t n, int lossless,
                                                  double ev_re, double ev_im, double dr, double *__restrict__ fac) {
    const size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (o >= n) return;
    double kr, ki, gr, gi, sn, cs;
    mode_factor(k2[o], absn[o], psis[o], lossless, ev_re, ev_im, kr, ki, gr, gi);
    sincos(f2_reduce(kr * dr), &sn, &cs);
    const double a1 = exp(-ki * dr);
    fac[o] = kr;
    fac[n + o] = ki;
    fac[2 * n + o] = gr;
    fac[3 * n + o] = gi;
    fac[4 * n + o] = a1 * cs;
    fac[5 * n + o] = a1 * sn;
}

// x of the lane j ^ d (d = 1, 2 or 4, inside each group of 8 lanes) by DPP lane moves, with no
// LDS traffic: quad permutations [1 0 3 2] and [2 3 0 1] for d = 1 and 2; for d = 4 the row
// shifts by 4 (lane i reads lane i + 4, or lane i - 4), chosen by bit 2 of the lane.
template <int D>
__device__ __forceinline__ unsigned xor_lanes(unsigned x, int lane) {
    const int v = (int)x;
    if constexpr (D == 1) return (unsigned)__builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
    if constexpr (D == 2) return (unsigned)__builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
    const int up = __builtin_amdgcn_update_dpp(0, v, 0x104, 0xF, 0xF, false);   // row_shl:4
    const int dn = __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);   // row_shr:4
    return (unsigned)((lane & 4) ? dn : up);
}

// E split: block = 32 modes x 64 ranges of one problem; phasors formed in FP64, split to
// 16-bit parts and written transposed ([column][mode]).  The transpose runs in registers
// (DPP lane moves inside each octet of modes): the block holds 5 KB of LDS, so blocks fit in
// the 16 KB that two 72 KB contraction blocks leave on a CU.
struct ETile {
    int p, k0;
};

template <int S, bool F16>
__global__ void __launch_bounds__(256) k_esplit(const ETile *__restrict__ tiles, const SplitProb *__restrict__ sp,
                                                const int *__restrict__ mode_off, const double *__restrict__ fac,
                                                size_t n_modes, double dr, int nr, size_t split_stride,
                                                const int *__restrict__ SB, int nr_pad,
                                                unsigned short *__restrict__ Bhat) {
    __shared__ double s_sc[64];
    __shared__ double s_z[8][32][2], s_m8[32][2];
    const ETile t = tiles[blockIdx.y];
    const SplitProb P = sp[t.p];
    const int i0 = blockIdx.x * 64;
    const int mloc = threadIdx.x & 31;       // mode within the tile
    const int k = t.k0 + mloc;               // mode within the problem
    double kr = 0, ki = 0, gr = 0, gi = 0;
    const bool live = k < P.K;
    const size_t o = live ? (size_t)mode_off[t.p] + k : 0;
    if (live) {
        gr = fac[2 * n_modes + o];
        gi = fac[3 * n_modes + o];
    }
    // 1000 / sqrt(r) of the tile's 64 ranges (times 2^SB with F16)
    if (threadIdx.x < 64) {
        const int i = i0 + (int)threadIdx.x;
        double sc = 1000.0 / sqrt((i + 1) * dr);
        if (F16 && i < nr_pad) sc = ldexp(sc, SB[(size_t)t.p * nr_pad + i]);
        s_sc[threadIdx.x] = sc;
    }
    // The thread's ranges are i0 + w + 8q (w = tid / 32).  Threads 0-31 form, for their mode,
    // e^{i k r} exact at the tile's first range (FP64 phase reduced mod 2 pi, sincos, exp);
    // the powers of the exact factor e^{i k dr} (k_mode_fac) give the 8 starting phasors
    // (w = 0..7) and the step e^{i k 8 dr}, shared through LDS: one sincos/exp per mode and tile.
    if (threadIdx.x < 32) {
        double ar = 0.0, ai = 0.0, fr = 0.0, fi = 0.0;
        if (live) {
            kr = fac[o];
            ki = fac[n_modes + o];
            fr = fac[4 * n_modes + o];
            fi = fac[5 * n_modes + o];
            const double r0 = (i0 + 1) * dr;
            double sn, cs;
            sincos(f2_reduce(kr * r0), &sn, &cs);
            const double a0 = exp(-ki * r0);
            ar = a0 * cs;
            ai = a0 * sn;
        }
        double pr = 1.0, pi = 0.0;   // e^{i k v dr}
        for (int v = 0; v < 8; ++v) {
            s_z[v][mloc][0] = ar * pr - ai * pi;
            s_z[v][mloc][1] = ar * pi + ai * pr;
            const double t = pr * fr - pi * fi;
            pi = pr * fi + pi * fr;
            pr = t;
        }
        s_m8[mloc][0] = pr;
        s_m8[mloc][1] = pi;
    }
    __syncthreads();
    const int w = (int)threadIdx.x >> 5;
    double zr = s_z[w][mloc][0], zi = s_z[w][mloc][1];
    const double mr = s_m8[mloc][0], mi = s_m8[mloc][1];
    // wd[s][q]: part s of the mode's value at range w + 8q, real part in the low half
    unsigned wd[S][8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int il = w + 8 * q;   // range within the tile, 0..63
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
#pragma unroll
        for (int s = 0; s < S; ++s) {
            double br, bi;
            wd[s][q] = split_pair<F16>(vr, vi, br, bi);
            vr -= br;
            vi -= bi;
        }
    }
    // 8 x 8 transpose among the 8 lanes j = mloc & 7 of an octet of modes (three butterfly
    // exchanges): afterwards wd[s][m] of lane j holds mode 8 (mloc >> 3) + m at range w + 8j
    const int j = mloc & 7;
    auto exchange = [&](auto dc) {
        constexpr int d = decltype(dc)::value;
        const bool up = (j & d) != 0;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if ((q & d) == 0) {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    // both words read and written unconditionally: a select between two
                    // addresses of the array would keep it out of registers
                    const unsigned a = wd[s][q], c = wd[s][q | d];
                    const unsigned r = xor_lanes<d>(up ? a : c, (int)threadIdx.x);
                    wd[s][q] = up ? r : a;
                    wd[s][q | d] = up ? c : r;
                }
            }
    };
    exchange(std::integral_constant<int, 4>());
    exchange(std::integral_constant<int, 2>());
    exchange(std::integral_constant<int, 1>());
    // columns 2 (i0 + w + 8j) (real) and + 1 (imaginary), modes k0 + 8 (mloc >> 3) .. + 7:
    // one 16-byte vector each
    const int n = 2 * (i0 + w + 8 * j), kk = t.k0 + 8 * (mloc >> 3);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        u4 re, im;
#pragma unroll
        for (int x = 0; x < 4; ++x) {
            re[x] = (wd[s][2 * x] & 0xffffu) | (wd[s][2 * x + 1] << 16);
            im[x] = (wd[s][2 * x] >> 16) | (wd[s][2 * x + 1] & 0xffff0000u);
        }
        unsigned short *b = Bhat + s * split_stride + P.b_off;
        *(u4 *)&b[tiled_index(n, kk, P.Kp)] = re;
        *(u4 *)&b[tiled_index(n + 1, kk, P.Kp)] = im;
    }
}

template <bool F16>
__device__ __forceinline__ f4 mfma16(u4 a, u4 b, f4 c) {
    if (F16)
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, a), __builtin_bit_cast(h8, b), c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0, 0);
}

// Epilogue shared by the contractions: the wave's accumulators (FP32, or FP64 when folded)
// to C, undoing the power-of-two scales of the FP16 forms: C = acc 2^-(SA[row] + SB[range]).
template <bool F16, int NB, typename Acc>
__device__ __forceinline__ void store_tile(float *__restrict__ Cp, size_t ldc, int M, int N, int row0, int col0,
                                           int lane, const Acc &acc, const int *__restrict__ SAp,
                                           const int *__restrict__ SBp) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int c = col0 + b * 16 + (lane & 15);
            const int sb = (F16 && c < N) ? SBp[c >> 1] : 0;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = row0 + a * 16 + 4 * (lane >> 4) + v;
                if (r < M && c < N) {
                    const int e = F16 ? -(SAp[r] + sb) : 0;
                    Cp[(size_t)r * ldc + c] = F16 ? (float)ldexp((double)acc[a][b][v], e) : (float)acc[a][b][v];
                }
            }
        }
}

// The NP part products of one 16 x 16 tile and k-step, smallest first: S = 2 -> 3 (hi*hi +
// hi*lo + lo*hi) or 4 (+ lo*lo); S = 3 -> 6 (all products with i + j <= 2).
template <int S, int NP, bool F16>
__device__ __forceinline__ f4 part_products(const u4 (&af)[S], const u4 (&bf)[S], f4 acc) {
    if constexpr (S == 3) {
        acc = mfma16<F16>(af[2], bf[0], acc);
        acc = mfma16<F16>(af[1], bf[1], acc);
        acc = mfma16<F16>(af[0], bf[2], acc);
    }
    if constexpr (S == 2 && NP == 4) acc = mfma16<F16>(af[1], bf[1], acc);
    acc = mfma16<F16>(af[1], bf[0], acc);
    acc = mfma16<F16>(af[0], bf[1], acc);
    return mfma16<F16>(af[0], bf[0], acc);
}

// The order of part_products as a table: product i multiplies A part a(i) by B part b(i);
// first_a(i) / first_b(i): the A / B parts in the order the products first need them.
template <int S, int NP>
struct PartOrder {
    // S = 3: lo*hi, mid*mid, hi*lo, mid*hi, hi*mid, hi*hi; S = 2, NP = 4: lo*lo, lo*hi, hi*lo,
    // hi*hi; S = 2, NP = 3: lo*hi, hi*lo, hi*hi (part 0 = hi)
    __device__ static constexpr int a(int i) {
        return S == 3 ? (i < 3 ? 2 - i : (i == 3 ? 1 : 0)) : (NP == 4 ? (i < 2 ? 1 : 0) : (i == 0 ? 1 : 0));
    }
    __device__ static constexpr int b(int i) {
        return S == 3 ? (i < 3 ? i : (i == 4 ? 1 : 0)) : (NP == 4 ? (i == 0 || i == 2 ? 1 : 0) : (i == 1 ? 1 : 0));
    }
    __device__ static constexpr int first_a(int i) { return S - 1 - i; }
    __device__ static constexpr int first_b(int i) { return S == 2 && NP == 4 ? 1 - i : i; }
};

// Plain split contraction: FP32 accumulation over all modes.  Block tile 128 x 128, 4 waves
// of 64 x 64; operands staged global -> registers -> LDS, the next k-step fetched into
// registers while the current one runs from LDS.
template <int S, int NP, bool F16>
__global__ void __launch_bounds__(256, 2) k_gemm_split(const SplitProb *__restrict__ sp,
                                                       const unsigned short *__restrict__ Ahat,
                                                       const unsigned short *__restrict__ Bhat,
                                                       size_t split_stride_a, size_t split_stride_b,
                                                       float *__restrict__ C, size_t ldc, int M, int N,
                                                       int tiles_m, int n_tiles_n, const int *__restrict__ SA, int Mp,
                                                       const int *__restrict__ SB, int nr_pad, int row_groups) {
    __shared__ __attribute__((aligned(16))) unsigned short As[S][TBM][TLD];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[S][TBN][TLD];
    const SplitProb P = sp[blockIdx.z];
    int tm, tn;
    xcd_tile(blockIdx.x, tiles_m, row_groups, tm, tn);
    if (tn >= n_tiles_n || tm >= tiles_m) return;
    const int m0 = tm * TBM, n0 = tn * TBN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    f4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = f4{0, 0, 0, 0};
    const unsigned short *Ab = Ahat + P.a_off, *Bb = Bhat + P.b_off;
    const int Kp = P.Kp;
    // per part and k-step a tile row is 4 vectors of 16 B: A and B 128 x 4 = 2 per thread.
    // Native vector types: arrays of the HIP uint4 struct were placed in scratch memory (a
    // round trip through HBM every k-step); ext_vector_type arrays stay in registers.
    u4 ra[S][2], rb[S][2];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int v = tid + 256 * h;
                ra[s][h] = *(const u4 *)&Ab[s * split_stride_a + tiled_index(m0 + (v >> 2), k0 + (v & 3) * 8, Kp)];
                rb[s][h] = *(const u4 *)&Bb[s * split_stride_b + tiled_index(n0 + (v >> 2), k0 + (v & 3) * 8, Kp)];
            }
    };
    auto stash = [&]() {
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int v = tid + 256 * h;
                *(u4 *)&As[s][v >> 2][lds_group(v >> 2, v & 3)] = ra[s][h];
                *(u4 *)&Bs[s][v >> 2][lds_group(v >> 2, v & 3)] = rb[s][h];
            }
    };
    if (Kp > 0) fetch(0);   // a problem without modes has no operand block to read
    for (int k0 = 0; k0 < Kp; k0 += TKS) {
        __syncthreads();
        stash();
        __syncthreads();
        if (k0 + TKS < Kp) fetch(k0 + TKS);
        // rows wm + 16a + (lane & 15): the swizzle of k-group lane >> 4 depends on lane only
        const int kq = lds_group(lane & 15, lane >> 4);
        u4 af[4][S], bfr[4][S];
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                af[a][s] = *(const u4 *)&As[s][wm + a * 16 + (lane & 15)][kq];
                bfr[a][s] = *(const u4 *)&Bs[s][wn + a * 16 + (lane & 15)][kq];
            }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = part_products<S, NP, F16>(af[a], bfr[b], acc[a][b]);
    }
    store_tile<F16, 4>(C + P.c_off, ldc, M, N, m0 + wm, n0 + wn, lane, acc,
                       F16 ? SA + (size_t)blockIdx.z * Mp : nullptr, F16 ? SB + (size_t)blockIdx.z * nr_pad : nullptr);
}


// FP64-folded split contraction: each k-step's part products start from zero in FP32 and
// are added to FP64 accumulators, so FP32 sums run over the 32 modes of one k-step only
// (the FP32 running sum over all modes is what limits the plain forms at interference nulls,
// where the partial sums exceed the result by orders of magnitude).  The FP64 accumulators
// take the registers of half the columns: block tile 128 x 64, 4 waves of 64 x 32.
// Operands go from global memory straight into LDS (global_load_lds_dwordx4: 16 B per lane,
// one 1 KB wave instruction = 16 tile rows, LDS address = base + 16 lane) into two stages:
// one barrier per k-step, the next k-step's loads in flight during the current MFMAs.  The
// XOR swizzle of the 16-byte groups is applied on the global side (lane -> row, stored
// group; the group loaded is the logical one), so the fragment reads are those of the
// register-staged kernel.  The two stages are separate arrays so that the compiler can tell
// the loads into one from the reads of the other (no wait on the loads in flight).
template <int S, int NP, bool F16>
__global__ void __launch_bounds__(256, 2) k_gemm_fold(const SplitProb *__restrict__ sp,
                                                      const unsigned short *__restrict__ Ahat,
                                                      const unsigned short *__restrict__ Bhat,
                                                      size_t split_stride_a, size_t split_stride_b,
                                                      float *__restrict__ C, size_t ldc, int M, int N,
                                                      int tiles_m, int n_tiles_n, const int *__restrict__ SA, int Mp,
                                                      const int *__restrict__ SB, int nr_pad, int row_groups,
                                                      const int *__restrict__ aflag) {
    constexpr int BN = 64;
    typedef unsigned short ATile[TBM][TKS];
    typedef unsigned short BTile[BN][TKS];
    __shared__ __attribute__((aligned(16))) ATile As0[S], As1[S];
    __shared__ __attribute__((aligned(16))) BTile Bs0[S], Bs1[S];
    const SplitProb P = sp[blockIdx.z];
    int tm, tn;
    xcd_tile(blockIdx.x, tiles_m, row_groups, tm, tn);
    if (tn >= n_tiles_n || tm >= tiles_m) return;
    const int m0 = tm * TBM, n0 = tn * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 32;
    f4 acc[4][2];
    double acc64[4][2][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc64[a][b][v] = 0.0;
    const unsigned short *Ab = Ahat + P.a_off, *Bb = Bhat + P.b_off;
    const int Kp = P.Kp;
    const int crow = lane >> 2, gs = lane & 3;
    // Buffer resources over each part of the problem's operand blocks (loads outside them
    // return zero).  Buffer loads rather than global ones: a FLAT-encoded load into LDS makes
    // the compiler wait for every LDS read at once (lgkmcnt(0)), buffer loads let each
    // fragment read be waited for on its own.  Per lane the byte offset inside a k-step's
    // tile chunk is fixed; the chunk's offset is a scalar that advances by 8 KB per k-step.
    const unsigned a_bytes = (unsigned)tiles_m * TBM * Kp * 2u;
    const unsigned b_bytes = (unsigned)((n_tiles_n + 1) / 2) * TBM * Kp * 2u;
    unsigned voff_a[2], voff_b;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int row = (2 * wave + j) * 16 + crow;
        voff_a[j] = (unsigned)(row * TKS + 8 * (gs ^ ((row >> 2) & 3))) * 2u;
    }
    {
        const int row = wave * 16 + crow;
        voff_b = (unsigned)(((tn & 1) * 64 + row) * TKS + 8 * (gs ^ ((row >> 2) & 3))) * 2u;
    }
    // k-step k0 into one stage: A = 8 chunks of 16 rows (2 per wave), B = 4 chunks (1 per wave)
    auto issue = [&](ATile *Ax, BTile *Bx, int k0) {
        const int kt = k0 / TKS, nkt = Kp / TKS;
        const unsigned so_a = (unsigned)(tm * nkt + kt) * (TBM * TKS * 2u);
        const unsigned so_b = (unsigned)((tn >> 1) * nkt + kt) * (TBM * TKS * 2u);
#pragma unroll
        for (int s = 0; s < S; ++s) {
#pragma unroll
            for (int j = 0; j < 2; ++j)
                buffer_load_lds16(Ab + s * split_stride_a, a_bytes, (lds_void *)&Ax[s][(2 * wave + j) * 16][0],
                                  voff_a[j], so_a);
            buffer_load_lds16(Bb + s * split_stride_b, b_bytes, (lds_void *)&Bx[s][wave * 16][0], voff_b, so_b);
        }
    };
    const int kq = lds_group(lane & 15, lane >> 4);
    auto step = [&](ATile *Ac, BTile *Bc, ATile *An, BTile *Bn, int k_next) {
        // this wave's loads of the current stage (the only ones in flight), then every wave's
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        // the other stage was last read in the previous k-step, which every wave has left
        if (k_next < Kp) issue(An, Bn, k_next);
        // fragments in the order the products first use them
        u4 af[4][S], bfr[2][S];
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const int sa = PartOrder<S, NP>::first_a(i), sb = PartOrder<S, NP>::first_b(i);
#pragma unroll
            for (int a = 0; a < 4; ++a) af[a][sa] = *(const u4 *)&Ac[sa][wm + a * 16 + (lane & 15)][kq];
#pragma unroll
            for (int b = 0; b < 2; ++b) bfr[b][sb] = *(const u4 *)&Bc[sb][wn + b * 16 + (lane & 15)][kq];
        }
        // one part product at a time over all 8 tiles (each tile's products smallest first,
        // as part_products sums them), so dependent MFMAs are 8 instructions apart
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[a][b] = f4{0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < NP; ++i) {
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = mfma16<F16>(af[a][PartOrder<S, NP>::a(i)], bfr[b][PartOrder<S, NP>::b(i)], acc[a][b]);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc64[a][b][v] += (double)acc[a][b][v];
    };
    // The k-steps whose Psi chunk (this row tile) is not all zero, one bit each (nkt <= 128;
    // without flags every k-step): a skipped k-step's products are exact zeros, and adding
    // +0 to an FP64 accumulator that started at +0 changes no bit, so the field is the same.
    const int nkt = Kp / TKS;
    unsigned long long km0 = 0, km1 = 0;
    if (aflag) {
        const int *fl = aflag + (P.a_off >> 12) + (size_t)tm * nkt;
        km0 = __ballot(lane < nkt && fl[lane] != 0);
        km1 = __ballot(lane + 64 < nkt && fl[lane + 64] != 0);
    } else {
        km0 = nkt >= 64 ? ~0ull : (1ull << nkt) - 1ull;
        km1 = nkt >= 128 ? ~0ull : (nkt > 64 ? (1ull << (nkt - 64)) - 1ull : 0ull);
    }
    // first flagged k-step at or after kt, as a mode offset (Kp if none)
    auto next_k = [&](int kt) -> int {
        if (kt < 64) {
            const unsigned long long m = km0 & (~0ull << kt);
            if (m) return (__ffsll((long long)m) - 1) * TKS;
            kt = 64;
        }
        if (kt < 128) {
            const unsigned long long m = km1 & (~0ull << (kt - 64));
            if (m) return (64 + __ffsll((long long)m) - 1) * TKS;
        }
        return Kp;
    };
    int k0 = next_k(0);   // a problem without modes has no operand block to read
    if (k0 < Kp) issue(As0, Bs0, k0);
    while (k0 < Kp) {   // two k-steps per pass, one per LDS stage
        const int k1 = next_k(k0 / TKS + 1);
        step(As0, Bs0, As1, Bs1, k1);
        int k2 = Kp;
        if (k1 < Kp) {
            k2 = next_k(k1 / TKS + 1);
            step(As1, Bs1, As0, Bs0, k2);
        }
        k0 = k2;
    }
    store_tile<F16, 2>(C + P.c_off, ldc, M, N, m0 + wm, n0 + wn, lane, acc64,
                       F16 ? SA + (size_t)blockIdx.z * Mp : nullptr, F16 ? SB + (size_t)blockIdx.z * nr_pad : nullptr);
}

// The folded contraction with three LDS stages: 8 waves of 64 x 32 in a 128 x 128 block,
// one block per CU (3 x 48 KB of LDS for three parts), the loads of k-steps t + 1 and t + 2
// in flight while k-step t runs.
template <int S, int NP, bool F16>
__global__ void __launch_bounds__(512, 1) k_gemm_fold3(const SplitProb *__restrict__ sp,
                                                       const unsigned short *__restrict__ Ahat,
                                                       const unsigned short *__restrict__ Bhat,
                                                       size_t split_stride_a, size_t split_stride_b,
                                                       float *__restrict__ C, size_t ldc, int M, int N,
                                                       int tiles_m, int n_tiles_n, const int *__restrict__ SA, int Mp,
                                                       const int *__restrict__ SB, int nr_pad, int row_groups) {
    typedef unsigned short Tile[TBM][TKS];
    __shared__ __attribute__((aligned(16))) Tile As0[S], As1[S], As2[S];
    __shared__ __attribute__((aligned(16))) Tile Bs0[S], Bs1[S], Bs2[S];
    const SplitProb P = sp[blockIdx.z];
    int tm, tn;
    xcd_tile(blockIdx.x, tiles_m, row_groups, tm, tn);
    if (tn >= n_tiles_n || tm >= tiles_m) return;
    const int m0 = tm * TBM, n0 = tn * TBN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave >> 2) * 64, wn = (wave & 3) * 32;
    f4 acc[4][2];
    double acc64[4][2][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc64[a][b][v] = 0.0;
    const unsigned short *Ab = Ahat + P.a_off, *Bb = Bhat + P.b_off;
    const int Kp = P.Kp;
    const int crow = lane >> 2, gs = lane & 3;
    const unsigned a_bytes = (unsigned)tiles_m * TBM * Kp * 2u;
    const unsigned b_bytes = (unsigned)n_tiles_n * TBN * Kp * 2u;
    // this wave's 16 rows of every A and B tile chunk (the same lane -> row / group map)
    const int row = wave * 16 + crow;
    const unsigned voff = (unsigned)(row * TKS + 8 * (gs ^ ((row >> 2) & 3))) * 2u;
    auto issue = [&](Tile *Ax, Tile *Bx, int k0) {
        const int kt = k0 / TKS, nkt = Kp / TKS;
        const unsigned so_a = (unsigned)(tm * nkt + kt) * (TBM * TKS * 2u);
        const unsigned so_b = (unsigned)(tn * nkt + kt) * (TBM * TKS * 2u);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            buffer_load_lds16(Ab + s * split_stride_a, a_bytes, (lds_void *)&Ax[s][wave * 16][0], voff, so_a);
            buffer_load_lds16(Bb + s * split_stride_b, b_bytes, (lds_void *)&Bx[s][wave * 16][0], voff, so_b);
        }
    };
    const int kq = lds_group(lane & 15, lane >> 4);
    auto step = [&](Tile *Ac, Tile *Bc, Tile *An, Tile *Bn, int k0) {
        // this wave's loads of the current stage: the loads of the next k-step (2 S per
        // wave) may stay in flight
        // (the builtin, not inline asm, so that the compiler's own wait insertion sees it)
        if (k0 + TKS < Kp)
            __builtin_amdgcn_s_waitcnt(waitcnt_vm(2 * S));
        else
            __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
        // a barrier without the fence of __syncthreads, which would wait for every load in
        // flight; the LDS reads before it have completed (their MFMAs have consumed them)
        asm volatile("s_barrier" ::: "memory");
        // the stage two k-steps ahead was last read in the previous k-step, which every wave
        // has left
        if (k0 + 2 * TKS < Kp) issue(An, Bn, k0 + 2 * TKS);
        u4 af[4][S], bfr[2][S];
#pragma unroll
        for (int i = 0; i < S; ++i) {
            const int sa = PartOrder<S, NP>::first_a(i), sb = PartOrder<S, NP>::first_b(i);
#pragma unroll
            for (int a = 0; a < 4; ++a) af[a][sa] = *(const u4 *)&Ac[sa][wm + a * 16 + (lane & 15)][kq];
#pragma unroll
            for (int b = 0; b < 2; ++b) bfr[b][sb] = *(const u4 *)&Bc[sb][wn + b * 16 + (lane & 15)][kq];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[a][b] = f4{0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < NP; ++i) {
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = mfma16<F16>(af[a][PartOrder<S, NP>::a(i)], bfr[b][PartOrder<S, NP>::b(i)], acc[a][b]);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc64[a][b][v] += (double)acc[a][b][v];
    };
    if (Kp > 0) issue(As0, Bs0, 0);   // a problem without modes has no operand block to read
    if (TKS < Kp) issue(As1, Bs1, TKS);
    for (int k0 = 0; k0 < Kp; k0 += 3 * TKS) {
        step(As0, Bs0, As2, Bs2, k0);
        if (k0 + TKS < Kp) step(As1, Bs1, As0, Bs0, k0 + TKS);
        if (k0 + 2 * TKS < Kp) step(As2, Bs2, As1, Bs1, k0 + 2 * TKS);
    }
    store_tile<F16, 2>(C + P.c_off, ldc, M, N, m0 + wm, n0 + wn, lane, acc64,
                       F16 ? SA + (size_t)blockIdx.z * Mp : nullptr, F16 ? SB + (size_t)blockIdx.z * nr_pad : nullptr);
}

// |TL_test - TL_ref| in dB over the points with TL_ref above a floor (10 log10 |p|^2 > floor_db;
// |p_test|^2 clamped at 1e-300 as a zero field): per block the maximum, the sum of squares and
// the count, folded into acc by one atomic each.  A non-negative double orders like its bits
// as an unsigned 64-bit integer, so the maximum is an integer atomicMax.
struct CompareAcc {
    unsigned long long max_bits, count;
    double sumsq;
};

__global__ void k_compare_init(CompareAcc *acc) {
    if (threadIdx.x == 0) {
        acc->max_bits = 0;
        acc->count = 0;
        acc->sumsq = 0.0;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_tl_compare(const double *__restrict__ ref, const T *__restrict__ test,
                                                    size_t n, double floor_pow, CompareAcc *acc) {
    double mx = 0.0, ss = 0.0;
    unsigned long long cnt = 0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double r0 = ref[2 * i], r1 = ref[2 * i + 1];
        const double t0 = (double)test[2 * i], t1 = (double)test[2 * i + 1];
        const double pr = r0 * r0 + r1 * r1;
        if (pr > floor_pow) {
            const double pt = fmax(t0 * t0 + t1 * t1, 1e-300);
            const double d = fabs(10.0 * log10(pt / pr));
            mx = fmax(mx, d);
            ss += d * d;
            ++cnt;
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, o));
        ss += __shfl_xor(ss, o);
        cnt += __shfl_xor(cnt, o);
    }
    __shared__ double s_mx[4], s_ss[4];
    __shared__ unsigned long long s_cnt[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_mx[wave] = mx;
        s_ss[wave] = ss;
        s_cnt[wave] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            mx = fmax(mx, s_mx[w]);
            ss += s_ss[w];
            cnt += s_cnt[w];
        }
        if (cnt) {
            atomicMax(&acc->max_bits, (unsigned long long)__double_as_longlong(mx));
            atomicAdd(&acc->sumsq, ss);
            atomicAdd(&acc->count, cnt);
        }
    }
}

}  // namespace ncpa

using namespace ncpa;

// accessors into the modes handle (defined in ess.hip)
extern "C" int ncpa_modes_device(ncpa_modes_t *m, const int **off, const double **k2, const double **ab,
                                 const double **psi_sr);
int ncpa_modes_internal(ncpa_modes_t *m, int *P, int *total, const int **h_offsets, const int **h_counts,
                        const double **d_psis);

// Staged operands of one ncpa_field2d call: everything the contraction reads besides the
// eigenvector rows, in device buffers of the staging context's pool.
struct ncpa_field_stage {
    ncpa_ctx *owner = nullptr;
    int precision = 0, P = 0, M = 0, N = 0, total = 0;
    // split precisions
    int Mp = 0, Np = 0, nr_pad = 0;
    size_t sa = 0, sb = 0;
    SplitProb *d_sp = nullptr;
    unsigned short *Ahat = nullptr, *Bhat = nullptr;
    int *SA = nullptr, *SB = nullptr;
    int *aflag = nullptr;   // per 128 x 32 chunk of Ahat: any nonzero (k_asplit); null: none skipped
    // FP64 / FP32
    const double *psi_rows = nullptr;
    GemmProb *d_gp = nullptr;
    void *bh = nullptr;
    float *ah = nullptr;
    size_t lda = 0, ldb = 0;
    int max_k = 0;   // most modes of one problem
    // the last contraction of this stage on another context's stream: the owner's stream
    // waits for it before the blocks go back to the owner's pool
    hipEvent_t foreign_done = nullptr;
};

namespace {

void stage_release(ncpa_field_stage *st) {
    ncpa_ctx *c = st->owner;
    if (st->foreign_done) {
        // later users of these blocks run on the owner's stream: order them after the
        // contraction that ran on the other context (no host wait)
        (void)hipSetDevice(c->device);
        (void)hipStreamWaitEvent(c->stream, st->foreign_done, 0);
        (void)hipEventDestroy(st->foreign_done);
        st->foreign_done = nullptr;
    }
    c->pool.put(st->d_sp);
    c->pool.put(st->Ahat);
    c->pool.put(st->Bhat);
    c->pool.put(st->SA);
    c->pool.put(st->SB);
    c->pool.put(st->aflag);
    c->pool.put(st->d_gp);
    c->pool.put(st->bh);
    c->pool.put(st->ah);
    delete st;
}

// Split operands: A rows and E columns cut into S 16-bit parts in the tiled layout (with
// F16, the power-of-two row and range scales first).
template <int S, bool F16>
int stage_split(ncpa_ctx *ctx, ncpa_field_stage *st, const int *hoff, const int *hcnt, const int *d_off,
                const double *d_k2, const double *d_abs, const double *d_psis, const double *psi, double dr,
                int n_r, int lossless) {
    const int P = st->P, M = st->M, N = st->N, total = st->total;
    const int Mp = (M + TBM - 1) / TBM * TBM, Np = (N + TBN - 1) / TBN * TBN;
    const int nr_pad = Np / 2;
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
    st->Mp = Mp;
    st->Np = Np;
    st->nr_pad = nr_pad;
    st->sa = sa;
    st->sb = sb;
    st->d_sp = pool_alloc<SplitProb>(ctx, P);
    st->Ahat = pool_alloc<unsigned short>(ctx, std::max<size_t>(S * sa, 8));
    st->Bhat = pool_alloc<unsigned short>(ctx, std::max<size_t>(S * sb, 8));
    st->SA = F16 ? pool_alloc<int>(ctx, (size_t)P * Mp) : nullptr;
    st->SB = F16 ? pool_alloc<int>(ctx, (size_t)P * nr_pad) : nullptr;
    // chunk flags: the folded contraction holds one bit per k-step of a row tile, 128 at most
    int max_kp = 0;
    for (int p = 0; p < P; ++p) max_kp = std::max(max_kp, sp[p].Kp);
    const size_t n_chunks = sa / (TBM * TKS);
    const int max_chunks = (Mp / TBM) * (max_kp / TKS);   // of one problem
    const bool flags = max_kp <= 128 * TKS && n_chunks > 0;
    st->aflag = flags ? pool_alloc<int>(ctx, n_chunks) : nullptr;
    // needed only while staging
    ETile *d_tiles = pool_alloc<ETile>(ctx, std::max<size_t>(tiles.size(), 1));
    double *LG = F16 ? pool_alloc<double>(ctx, std::max(total, 1)) : nullptr;
    double *KI = F16 ? pool_alloc<double>(ctx, std::max(total, 1)) : nullptr;
    double *FAC = pool_alloc<double>(ctx, 6 * (size_t)std::max(total, 1));   // k_mode_fac
    auto put_temps = [&]() {
        ctx->pool.put(d_tiles);
        ctx->pool.put(LG);
        ctx->pool.put(KI);
        ctx->pool.put(FAC);
    };
    if (!st->d_sp || !st->Ahat || !st->Bhat || !d_tiles || !FAC || (F16 && (!st->SA || !st->SB || !LG || !KI)) ||
        (flags && !st->aflag)) {
        put_temps();
        return set_error(ctx, NCPA_ERR_ALLOC, "ncpa_field2d: allocation failed");
    }
    hipError_t e = hipMemcpyAsync(st->d_sp, sp.data(), sizeof(SplitProb) * P, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && !tiles.empty())
        e = hipMemcpyAsync(d_tiles, tiles.data(), sizeof(ETile) * tiles.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
        put_temps();
        return hip_error(ctx, e, "ncpa_field2d copy");
    }
    const std::complex<double> I(0.0, 1.0);
    const std::complex<double> ev = 4 * PI * I * std::exp(-I * PI * 0.25) / std::sqrt(8.0 * PI);
    if (!tiles.empty()) {
        TimedLaunch tl(ctx, "emat");
        if (F16) {
            hipLaunchKernelGGL(k_rowscale, dim3((Mp + 3) / 4, P), dim3(256), 0, ctx->stream, psi, (size_t)total,
                               st->d_sp, d_off, M, Mp, st->SA);
            hipLaunchKernelGGL(k_mode_env, dim3((total + 255) / 256), dim3(256), 0, ctx->stream, d_k2, d_abs, d_psis,
                               (size_t)total, lossless, ev.real(), ev.imag(), LG, KI);
            hipLaunchKernelGGL(k_colscale, dim3((nr_pad + 255) / 256, P), dim3(256), 0, ctx->stream, st->d_sp, d_off,
                               LG, KI, dr, n_r, nr_pad, st->SB);
        }
        hipLaunchKernelGGL((k_asplit<S, F16>), dim3((unsigned)((max_chunks + 3) / 4), P), dim3(256), 0, ctx->stream,
                           psi, (size_t)total, st->d_sp, d_off, M, Mp, sa, st->SA, st->Ahat, st->aflag);
        hipLaunchKernelGGL(k_mode_fac, dim3((total + 255) / 256), dim3(256), 0, ctx->stream, d_k2, d_abs, d_psis,
                           (size_t)total, lossless, ev.real(), ev.imag(), dr, FAC);
        hipLaunchKernelGGL((k_esplit<S, F16>), dim3((n_r + 63) / 64, (unsigned)tiles.size()), dim3(256), 0,
                           ctx->stream, d_tiles, st->d_sp, d_off, FAC, (size_t)total, dr, n_r, sb, st->SB, nr_pad,
                           st->Bhat);
    }
    e = hipGetLastError();
    put_temps();   // stream-ordered: later users of these blocks follow the staging kernels
    if (e != hipSuccess) return hip_error(ctx, e, "ncpa_field2d staging launch");
    return NCPA_OK;
}

// FP64 / FP32: the E matrix in the output precision (and A in FP32 for FP32).
int stage_plain(ncpa_ctx *ctx, ncpa_field_stage *st, const int *hoff, const int *hcnt, const double *d_k2,
                const double *d_abs, const double *d_psis, const double *psi, double dr, int n_r, int lossless) {
    const int P = st->P, nrows = st->M, total = st->total, precision = st->precision;
    st->ldb = 2 * (size_t)n_r;
    st->lda = (size_t)std::max(total, 1);
    st->psi_rows = psi;
    std::vector<GemmProb> gp(P);
    for (int p = 0; p < P; ++p) {
        gp[p] = {(size_t)hoff[p], (size_t)hoff[p], (size_t)p * nrows * st->ldb, hcnt[p]};
        st->max_k = std::max(st->max_k, hcnt[p]);
    }
    st->d_gp = pool_alloc<GemmProb>(ctx, P);
    const size_t esz = precision == PREC_FP64 ? 8 : 4;
    st->bh = ctx->pool.get(std::max<size_t>(total, 1) * st->ldb * esz);
    st->ah = precision == PREC_FP64 ? nullptr : pool_alloc<float>(ctx, (size_t)nrows * st->lda);
    if (!st->d_gp || !st->bh || (precision != PREC_FP64 && !st->ah))
        return set_error(ctx, NCPA_ERR_ALLOC, "ncpa_field2d: allocation failed");
    hipError_t e = hipMemcpyAsync(st->d_gp, gp.data(), sizeof(GemmProb) * P, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return hip_error(ctx, e, "ncpa_field2d copy");
    const std::complex<double> I(0.0, 1.0);
    const std::complex<double> ev = 4 * PI * I * std::exp(-I * PI * 0.25) / std::sqrt(8.0 * PI);
    if (total > 0) {
        TimedLaunch tl(ctx, "emat");
        dim3 ge((unsigned)(((n_r + 256 * EMAT_R - 1) / (256 * EMAT_R)) * (size_t)total));
        if (precision == PREC_FP64)
            hipLaunchKernelGGL(k_emat<PREC_FP64>, ge, dim3(256), 0, ctx->stream, d_k2, d_abs, d_psis, dr, n_r,
                               lossless, ev.real(), ev.imag(), st->bh);
        else
            hipLaunchKernelGGL(k_emat<PREC_FP32>, ge, dim3(256), 0, ctx->stream, d_k2, d_abs, d_psis, dr, n_r,
                               lossless, ev.real(), ev.imag(), st->bh);
        if (precision == PREC_FP32)
            hipLaunchKernelGGL(k_amat, dim3(1024), dim3(256), 0, ctx->stream, psi, (size_t)nrows * st->lda, st->ah);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return hip_error(ctx, e, "ncpa_field2d staging launch");
    return NCPA_OK;
}

// Launch variants of the contractions: a kernel form and an XCD row grouping (xcd_tile).
// Within one precision every variant forms each output element with the same operations in
// the same order, so the variants agree to the bit and differ in speed only; the context
// keeps one per precision (ncpa_field_set_variant, ncpa_field_tune).
//   folded splits (BF16x6F, FP16x4F): form 0 = k_gemm_fold (128 x 64 blocks, 2 per CU),
//                                     form 1 = k_gemm_fold3 (128 x 128 blocks, 3 LDS stages)
//   plain splits: k_gemm_split (128 x 128 blocks)
//   FP64 / FP32:  form 1 = k_gemm_x4t (128 x 128 blocks, loads into LDS), form 0 = k_gemm_x4
//                 (64 x 64 blocks, element loads; also taken when an operand exceeds the
//                 2 GiB that the tiled form's 32-bit offsets address)
struct ContractVariant {
    int form, row_groups;
};
constexpr ContractVariant FOLD_VARIANTS[] = {{0, 2}, {0, 4}, {0, 1}, {1, 2}, {1, 4}};
constexpr ContractVariant SPLIT_VARIANTS[] = {{0, 2}, {0, 4}, {0, 1}};
constexpr ContractVariant X4_VARIANTS[] = {{1, 2}, {1, 4}, {1, 1}, {0, 0}};

int contract_variants(int precision) {
    if (precision == PREC_BF16X6F || precision == PREC_FP16X4F) return (int)(sizeof FOLD_VARIANTS / sizeof FOLD_VARIANTS[0]);
    if (precision == PREC_FP64 || precision == PREC_FP32) return (int)(sizeof X4_VARIANTS / sizeof X4_VARIANTS[0]);
    return (int)(sizeof SPLIT_VARIANTS / sizeof SPLIT_VARIANTS[0]);
}

// blocks of an xcd_tile grid over tiles_m x tiles_n tiles
static unsigned xcd_grid(int tiles_m, int tiles_n, int rg) {
    return (unsigned)(8 * ((tiles_m + rg - 1) / rg) * ((tiles_n + 8 / rg - 1) / (8 / rg)));
}

template <int S, int NP, bool F16, bool FOLD>
void launch_split(ncpa_ctx *ctx, const ncpa_field_stage *st, float *out, int variant) {
    // Np is a multiple of 128 (the E-staging tiles), so of either block width
    const int tm = st->Mp / TBM;
    if constexpr (FOLD) {
        const ContractVariant v = FOLD_VARIANTS[variant];
        const int tn = v.form == 1 ? st->Np / TBN : (st->N + 63) / 64;   // N tiles that hold outputs
        const dim3 g(xcd_grid(tm, tn, v.row_groups), 1, (unsigned)st->P);
        if (v.form == 1)
            hipLaunchKernelGGL((k_gemm_fold3<S, NP, F16>), g, dim3(512), 0, ctx->stream, st->d_sp, st->Ahat,
                               st->Bhat, st->sa, st->sb, out, (size_t)st->N, st->M, st->N, tm, tn, st->SA, st->Mp,
                               st->SB, st->nr_pad, v.row_groups);
        else
            hipLaunchKernelGGL((k_gemm_fold<S, NP, F16>), g, dim3(256), 0, ctx->stream, st->d_sp, st->Ahat,
                               st->Bhat, st->sa, st->sb, out, (size_t)st->N, st->M, st->N, tm, tn, st->SA, st->Mp,
                               st->SB, st->nr_pad, v.row_groups, st->aflag);
    } else {
        const ContractVariant v = SPLIT_VARIANTS[variant];
        const int tn = st->Np / TBN;
        const dim3 g(xcd_grid(tm, tn, v.row_groups), 1, (unsigned)st->P);
        hipLaunchKernelGGL((k_gemm_split<S, NP, F16>), g, dim3(256), 0, ctx->stream, st->d_sp, st->Ahat, st->Bhat,
                           st->sa, st->sb, out, (size_t)st->N, st->M, st->N, tm, tn, st->SA, st->Mp, st->SB,
                           st->nr_pad, v.row_groups);
    }
}

template <class T>
void launch_x4(ncpa_ctx *ctx, const ncpa_field_stage *st, const T *A, T *out, int variant) {
    const size_t ldc = st->ldb;
    ContractVariant v = X4_VARIANTS[variant];
    // the tiled form addresses its operands through 32-bit buffer offsets
    const size_t a_bytes = (size_t)st->M * st->lda * sizeof(T);
    const size_t b_bytes = (size_t)st->max_k * st->ldb * sizeof(T);
    if (v.form == 1 && (a_bytes >= (size_t(1) << 31) || b_bytes >= (size_t(1) << 31))) v = ContractVariant{0, 0};
    if (v.form == 1) {
        const int tm = (st->M + 127) / 128, tn = (int)((st->ldb + 127) / 128);
        const dim3 g(xcd_grid(tm, tn, v.row_groups), 1, (unsigned)st->P);
        hipLaunchKernelGGL((k_gemm_x4t<T, 16>), g, dim3(256), 0, ctx->stream, st->d_gp, A, st->lda,
                           (unsigned)a_bytes, (const T *)st->bh, st->ldb, out, ldc, st->M, (int)st->ldb, tm, tn,
                           v.row_groups);
    } else {
        const dim3 gg((unsigned)((st->ldb + BN - 1) / BN), (unsigned)((st->M + BM - 1) / BM), (unsigned)st->P);
        hipLaunchKernelGGL((k_gemm_x4<T, 16>), gg, dim3(256), 0, ctx->stream, st->d_gp, A, st->lda,
                           (const T *)st->bh, st->ldb, out, ldc, st->M, (int)st->ldb);
    }
}

int contract(ncpa_ctx *ctx, const ncpa_field_stage *st, void *dev_out, int variant) {
    TimedLaunch tl(ctx, "field2d");
    float *out = (float *)dev_out;
    if (variant < 0) variant = ctx->field_variant[st->precision];
    switch (st->precision) {
    case PREC_BF16X3: launch_split<2, 3, false, false>(ctx, st, out, variant); break;
    case PREC_BF16X6: launch_split<3, 6, false, false>(ctx, st, out, variant); break;
    case PREC_FP16X3: launch_split<2, 3, true, false>(ctx, st, out, variant); break;
    case PREC_FP16X4: launch_split<2, 4, true, false>(ctx, st, out, variant); break;
    case PREC_BF16X6F: launch_split<3, 6, false, true>(ctx, st, out, variant); break;
    case PREC_FP16X4F: launch_split<2, 4, true, true>(ctx, st, out, variant); break;
    case PREC_FP64: launch_x4<double>(ctx, st, st->psi_rows, (double *)dev_out, variant); break;
    default: launch_x4<float>(ctx, st, st->ah, out, variant);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(ctx, e, "ncpa_field2d launch");
    return NCPA_OK;
}

}  // namespace

extern "C" {

int ncpa_field2d_stage(ncpa_ctx_t *ctx, ncpa_modes_t *m, const double *dev_psi_rows, int nrows, double dr, int n_r,
                       int lossless, int precision, ncpa_field_stage_t **stage) {
    if (stage) *stage = nullptr;
    if (!ctx || !m || !dev_psi_rows || nrows <= 0 || n_r <= 0 || !(dr > 0.0) || !stage || precision < 0 ||
        precision > PREC_FP16X4F)
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_field2d: bad arguments");
    int P = 0, total = 0;
    const int *hoff = nullptr, *hcnt = nullptr, *d_off = nullptr;
    const double *d_psis = nullptr, *d_k2 = nullptr, *d_abs = nullptr;
    ncpa_modes_internal(m, &P, &total, &hoff, &hcnt, &d_psis);
    ncpa_modes_device(m, &d_off, &d_k2, &d_abs, nullptr);
    if (total > 0 && !d_psis)
        return set_error(ctx, NCPA_ERR_STATE, "ncpa_field2d: call ncpa_modes_rows first (it supplies psi(zs))");
    ncpa_field_stage *st = new (std::nothrow) ncpa_field_stage();
    if (!st) return set_error(ctx, NCPA_ERR_ALLOC, "ncpa_field2d: allocation failed");
    st->owner = ctx;
    st->precision = precision;
    st->P = P;
    st->M = nrows;
    st->N = 2 * n_r;
    st->total = total;
    int rc = NCPA_OK;
    switch (precision) {
    case PREC_BF16X3:
        rc = stage_split<2, false>(ctx, st, hoff, hcnt, d_off, d_k2, d_abs, d_psis, dev_psi_rows, dr, n_r, lossless);
        break;
    case PREC_BF16X6:
    case PREC_BF16X6F:
        rc = stage_split<3, false>(ctx, st, hoff, hcnt, d_off, d_k2, d_abs, d_psis, dev_psi_rows, dr, n_r, lossless);
        break;
    case PREC_FP16X3:
    case PREC_FP16X4:
    case PREC_FP16X4F:
        rc = stage_split<2, true>(ctx, st, hoff, hcnt, d_off, d_k2, d_abs, d_psis, dev_psi_rows, dr, n_r, lossless);
        break;
    default: rc = stage_plain(ctx, st, hoff, hcnt, d_k2, d_abs, d_psis, dev_psi_rows, dr, n_r, lossless);
    }
    if (rc != NCPA_OK) {
        stage_release(st);
        return rc;
    }
    *stage = st;
    return NCPA_OK;
}

int ncpa_field2d_contract(ncpa_ctx_t *ctx, ncpa_field_stage_t *stage, void *dev_out) {
    if (!ctx || !stage || !dev_out) return set_error(ctx, NCPA_ERR_ARG, "ncpa_field2d_contract: bad arguments");
    NCPA_HIP(ctx, hipSetDevice(ctx->device));
    const int rc = contract(ctx, stage, dev_out, -1);
    if (rc != NCPA_OK || ctx == stage->owner) return rc;
    // a contraction on another context's stream reads the owner's blocks: mark its end so
    // that ncpa_field_stage_free returns them only behind it
    if (!stage->foreign_done)
        NCPA_HIP(ctx, hipEventCreateWithFlags(&stage->foreign_done, hipEventDisableTiming));
    NCPA_HIP(ctx, hipEventRecord(stage->foreign_done, ctx->stream));
    return NCPA_OK;
}

void ncpa_field_stage_free(ncpa_field_stage_t *stage) {
    if (stage) stage_release(stage);
}

int ncpa_field_variants(int precision, int *n_variants) {
    if (precision < 0 || precision > PREC_FP16X4F || !n_variants) return NCPA_ERR_ARG;
    *n_variants = contract_variants(precision);
    return NCPA_OK;
}

int ncpa_field_variant_info(int precision, int variant, const char **kernel, int *row_groups) {
    if (precision < 0 || precision > PREC_FP16X4F || variant < 0 || variant >= contract_variants(precision))
        return NCPA_ERR_ARG;
    static const char *const split_names[8] = {"", "", "k_gemm_split<2, 3, false>", "k_gemm_split<3, 6, false>",
                                               "k_gemm_split<2, 3, true>", "k_gemm_split<2, 4, true>", "", ""};
    ContractVariant v;
    const char *name;
    if (precision == PREC_BF16X6F || precision == PREC_FP16X4F) {
        v = FOLD_VARIANTS[variant];
        const bool b = precision == PREC_BF16X6F;
        name = v.form == 1 ? (b ? "k_gemm_fold3<3, 6, false>" : "k_gemm_fold3<2, 4, true>")
                           : (b ? "k_gemm_fold<3, 6, false>" : "k_gemm_fold<2, 4, true>");
    } else if (precision == PREC_FP64 || precision == PREC_FP32) {
        v = X4_VARIANTS[variant];
        const bool d = precision == PREC_FP64;
        name = v.form == 1 ? (d ? "k_gemm_x4t<double, 16>" : "k_gemm_x4t<float, 16>")
                           : (d ? "k_gemm_x4<double, 16>" : "k_gemm_x4<float, 16>");
    } else {
        v = SPLIT_VARIANTS[variant];
        name = split_names[precision];
    }
    if (kernel) *kernel = name;
    if (row_groups) *row_groups = v.row_groups;
    return NCPA_OK;
}

int ncpa_field_set_variant(ncpa_ctx_t *ctx, int precision, int variant) {
    if (!ctx || precision < 0 || precision > PREC_FP16X4F || variant < 0 || variant >= contract_variants(precision))
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_field_set_variant: bad arguments");
    ctx->field_variant[precision] = variant;
    return NCPA_OK;
}

int ncpa_field_get_variant(ncpa_ctx_t *ctx, int precision, int *variant) {
    if (!ctx || precision < 0 || precision > PREC_FP16X4F || !variant)
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_field_get_variant: bad arguments");
    *variant = ctx->field_variant[precision];
    return NCPA_OK;
}

int ncpa_field_tune(ncpa_ctx_t *ctx, ncpa_field_stage_t *stage, void *dev_out, int reps, int *best, double *ms,
                    int ms_len) {
    if (!ctx || !stage || !dev_out || reps < 1) return set_error(ctx, NCPA_ERR_ARG, "ncpa_field_tune: bad arguments");
    NCPA_HIP(ctx, hipSetDevice(ctx->device));
    const int nv = contract_variants(stage->precision);
    std::vector<double> t(nv, 0.0);
    hipEvent_t e0, e1;
    NCPA_HIP(ctx, hipEventCreate(&e0));
    NCPA_HIP(ctx, hipEventCreate(&e1));
    int rc = NCPA_OK;
    for (int v = 0; v < nv && rc == NCPA_OK; ++v) {
        rc = contract(ctx, stage, dev_out, v);   // warm-up
        hipEventRecord(e0, ctx->stream);
        for (int r = 0; r < reps && rc == NCPA_OK; ++r) rc = contract(ctx, stage, dev_out, v);
        hipEventRecord(e1, ctx->stream);
        float el = 0.0f;
        if (rc == NCPA_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&el, e0, e1) != hipSuccess))
            rc = set_error(ctx, NCPA_ERR_HIP, "ncpa_field_tune: timing failed");
        t[v] = (double)el / reps;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc != NCPA_OK) return rc;
    int pick = 0;
    for (int v = 1; v < nv; ++v)
        if (t[v] < t[pick]) pick = v;
    // variant 0 (the lowest HBM traffic of its family) stays unless another is faster by more
    // than the launch-to-launch spread (~2 %)
    if (t[pick] > 0.98 * t[0]) pick = 0;
    ctx->field_variant[stage->precision] = pick;
    if (best) *best = pick;
    for (int v = 0; ms && v < std::min(nv, ms_len); ++v) ms[v] = t[v];
    if (ctx != stage->owner) {
        if (!stage->foreign_done)
            NCPA_HIP(ctx, hipEventCreateWithFlags(&stage->foreign_done, hipEventDisableTiming));
        NCPA_HIP(ctx, hipEventRecord(stage->foreign_done, ctx->stream));
    }
    return NCPA_OK;
}

int ncpa_field2d(ncpa_ctx_t *ctx, ncpa_modes_t *m, const double *dev_psi_rows, int nrows, double dr, int n_r,
                 int lossless, int precision, void *dev_out) {
    if (!dev_out) return set_error(ctx, NCPA_ERR_ARG, "ncpa_field2d: bad arguments");
    ncpa_field_stage *st = nullptr;
    int rc = ncpa_field2d_stage(ctx, m, dev_psi_rows, nrows, dr, n_r, lossless, precision, &st);
    if (rc != NCPA_OK) return rc;
    rc = contract(ctx, st, dev_out, -1);
    stage_release(st);   // stream-ordered on ctx: the blocks are reused only after the contraction
    return rc;
}

int ncpa_field_compare(ncpa_ctx_t *ctx, const double *dev_ref, const void *dev_test, int test_precision,
                       size_t n_points, double floor_db, double *max_db, double *rms_db, size_t *n_counted) {
    if (!ctx || !dev_ref || !dev_test || test_precision < 0 || test_precision > PREC_FP16X4F || !max_db)
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_field_compare: bad arguments");
    *max_db = 0.0;
    if (rms_db) *rms_db = 0.0;
    if (n_counted) *n_counted = 0;
    if (n_points == 0) return NCPA_OK;
    CompareAcc *acc = pool_alloc<CompareAcc>(ctx, 1);
    if (!acc) return set_error(ctx, NCPA_ERR_ALLOC, "ncpa_field_compare: allocation failed");
    const double floor_pow = std::pow(10.0, floor_db / 10.0);
    const unsigned blocks = (unsigned)std::min<size_t>((n_points + 1023) / 1024, 8192);
    hipLaunchKernelGGL(k_compare_init, dim3(1), dim3(64), 0, ctx->stream, acc);
    if (test_precision == PREC_FP64)
        hipLaunchKernelGGL(k_tl_compare<double>, dim3(blocks), dim3(256), 0, ctx->stream, dev_ref,
                           (const double *)dev_test, n_points, floor_pow, acc);
    else
        hipLaunchKernelGGL(k_tl_compare<float>, dim3(blocks), dim3(256), 0, ctx->stream, dev_ref,
                           (const float *)dev_test, n_points, floor_pow, acc);
    CompareAcc h{};
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&h, acc, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    ctx->pool.put(acc);
    if (e != hipSuccess) return hip_error(ctx, e, "ncpa_field_compare");
    double mx;
    std::memcpy(&mx, &h.max_bits, sizeof(mx));
    *max_db = mx;
    if (rms_db) *rms_db = h.count ? std::sqrt(h.sumsq / (double)h.count) : 0.0;
    if (n_counted) *n_counted = (size_t)h.count;
    return NCPA_OK;
}

}  // extern "C"
