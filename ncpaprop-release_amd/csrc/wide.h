// FP64 Sturm sweeps over the wide-angle (high-Mach) modal pencil.
//
// WModeSolver solves the quadratic eigenvalue problem (k^2 M + k C + D) v = 0 with
// D = tridiag(1/h^2, -2/h^2 + kd_i, 1/h^2), M = diag(md_i), C = diag(cd_i)
// (src/libmodes/EigenEngine.cpp:220-250, WModeSolver.cpp:549-557).  For real k the matrix
//     T(k) = D + k C + k^2 M
// is symmetric tridiagonal, and the sweeps work on dz^2 T(k) = tridiag(1, a_i(k), 1) with
//     a_i(k) = A_i + k (B_i + k C_i),  A = -2 + dz^2 kd,  B = dz^2 cd,  C = dz^2 md.
// With md = (wc/c0)^2 - 1 < 0 and cd = -2 omega wc/c0^2, dT/dk = C + 2kM is negative definite
// for k > omega |wc|/(c0^2 - wc^2) on every row, which holds on the trapped interval of any
// subsonic atmosphere: every eigenvalue branch mu_i(k) of T(k) then decreases strictly, so
//     n(k) = number of negative eigenvalues of T(k)
// is non-decreasing in k and the QEP eigenvalues in (k_lo, k_hi] are exactly the
// n(k_hi) - n(k_lo) zero crossings, branch i crossing at the i-th one.  The search of the
// effective-sound-speed kernel therefore runs unchanged with k as the shift: count(k) is
// the sign-change count of the continuants of dz^2 T(k) at shift 0, and Newton works on
// det dz^2 T(k) with the k-derivative continuant
//     q_i = a_i'(k) phi_{i+1} + a_i(k) q_{i+1} - q_{i+2},   a_i' = B_i + 2 k C_i.
// At a converged k the unit null vector u of dz^2 T(k) has the adjugate products of the
// effective-sound-speed case (sturm.h) with the identity-shift derivative continuant
// q^mu; the reference's vector is the first half of a unit 2N vector (v, k v)
// (EigenEngine.cpp:376-381), so psi = u / sqrt(dz (1 + k^2)) and the normalising
// derivative is q^mu (1 + k^2).
//
// Three arrays (four with the absorption weights) are staged per sweep through LDS rings
// of two 128-row tiles each (the pattern of sturm.h at half its tile): 8 KB of rings and
// 8-row reads keep a block at ~10 KB of LDS and a wave under 128 VGPRs, so large batches run
// 4 waves per SIMD.
#pragma once
#include "sturm.h"

namespace ncpa {

// Two staging geometries: tiles of T = 256 rows read back 16 rows at a time (fewest
// syncs per row: small batches, where each wave's own instruction stream sets the time), and
// T = 128 rows read 8 at a time (8 KB of rings, under 128 VGPRs: large batches run 4 waves per
// SIMD).  A ring holds two tiles of one array: 2 T doubles.
#define NCPA_WTILE_BIG 256
#define NCPA_WTILE_SMALL 128

template <int T>
struct WTileStage {
    const double *g;
    double *ring;
    int n;
    int lane;
    double pre[T / 64];
    __device__ __forceinline__ void load(int c) {
#pragma unroll
        for (int u = 0; u < T / 64; ++u) {
            const int r = c * T + lane + 64 * u;
            pre[u] = (c >= 0 && r < n) ? g[r] : 0.0;
        }
    }
    __device__ __forceinline__ void store(int c) {
        double *s = ring + (c & 1) * T;
#pragma unroll
        for (int u = 0; u < T / 64; ++u) s[lane + 64 * u] = pre[u];
    }
};

// G rows of NA staged arrays (rings of 2 T doubles) starting at ring offset off
template <int T, int NA, int G>
__device__ __forceinline__ void lds_rows(double (&v)[NA][G], const double *ring, int off) {
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        const double *s = ring + a * 2 * T + off;
#pragma unroll
        for (int u = 0; u < G; ++u) v[a][u] = s[u];
    }
}

// Rows n-1 .. 0 of NA arrays staged together: step(v, i) with v[a] = arr[a][i], then
// chunk_end(lo, m) after each chunk of rows [lo, lo + m) (first the n % 16 top rows, then
// 16-row chunks downwards).  ring holds NA rings of NCPA_RING doubles.  Full-wave entry.
template <int NA, class Step, class ChunkEnd>
__device__ __forceinline__ void sweep_down_big(const double *const (&arr)[NA], double *ring, int n, Step step,
                                             ChunkEnd chunk_end) {
    if (n <= 0) return;
    const int lane = (int)(threadIdx.x & 63);
    TileStage st[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) st[a] = TileStage{arr[a], ring + a * NCPA_RING, n, lane};
    const int ntile = (n + NCPA_TILE - 1) / NCPA_TILE;
#pragma unroll
    for (int a = 0; a < NA; ++a) { st[a].load(ntile - 1); st[a].store(ntile - 1); st[a].load(ntile - 2); }
    wave_lds_sync();
    const int rem = n % NCPA_CHUNK;
    for (int c = ntile - 1; c >= 0; --c) {
        const int off = (c & 1) * NCPA_TILE;
        const int base = c * NCPA_TILE;
        int top = min(n, base + NCPA_TILE);
        if (c == ntile - 1 && rem) {
            for (int i = n - 1; i >= n - rem; --i) {
                double v[NA];
#pragma unroll
                for (int a = 0; a < NA; ++a) v[a] = ring[a * NCPA_RING + off + i - base];
                step(v, i);
            }
            chunk_end(n - rem, rem);
            top = n - rem;
        }
        for (int cb = top - NCPA_CHUNK; cb >= base; cb -= NCPA_CHUNK) {
            double va[NA][NCPA_CHUNK];
#pragma unroll
            for (int a = 0; a < NA; ++a) lds_chunk(va[a], ring + a * NCPA_RING + off + (cb - base));
#pragma unroll
            for (int u = NCPA_CHUNK - 1; u >= 0; --u) {
                double v[NA];
#pragma unroll
                for (int a = 0; a < NA; ++a) v[a] = va[a][u];
                step(v, cb + u);
            }
            chunk_end(cb, NCPA_CHUNK);
        }
        if (c > 0) {
            wave_lds_sync();
#pragma unroll
            for (int a = 0; a < NA; ++a) { st[a].store(c - 1); st[a].load(c - 2); }
            wave_lds_sync();
        }
    }
}

// Rows 0 .. m-1 (ascending) of NA arrays, chunk_end(lo, m') after every chunk of 16 rows and
// after the last m % 16 rows.  Full-wave entry.
template <int NA, class Step, class ChunkEnd>
__device__ __forceinline__ void sweep_up_big(const double *const (&arr)[NA], double *ring, int m, Step step,
                                           ChunkEnd chunk_end) {
    if (m <= 0) return;
    const int lane = (int)(threadIdx.x & 63);
    TileStage st[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) st[a] = TileStage{arr[a], ring + a * NCPA_RING, m, lane};
    const int ntile = (m + NCPA_TILE - 1) / NCPA_TILE;
#pragma unroll
    for (int a = 0; a < NA; ++a) { st[a].load(0); st[a].store(0); st[a].load(1 < ntile ? 1 : -1); }
    wave_lds_sync();
    for (int c = 0; c < ntile; ++c) {
        const int off = (c & 1) * NCPA_TILE;
        const int base = c * NCPA_TILE;
        const int endt = min(m, base + NCPA_TILE);
        for (int cb = base; cb < endt; cb += NCPA_CHUNK) {
            const int len = min(NCPA_CHUNK, endt - cb);
            if (len == NCPA_CHUNK) {
                double va[NA][NCPA_CHUNK];
#pragma unroll
                for (int a = 0; a < NA; ++a) lds_chunk(va[a], ring + a * NCPA_RING + off + (cb - base));
#pragma unroll
                for (int u = 0; u < NCPA_CHUNK; ++u) {
                    double v[NA];
#pragma unroll
                    for (int a = 0; a < NA; ++a) v[a] = va[a][u];
                    step(v, cb + u);
                }
            } else {
                for (int i = cb; i < cb + len; ++i) {
                    double v[NA];
#pragma unroll
                    for (int a = 0; a < NA; ++a) v[a] = ring[a * NCPA_RING + off + i - base];
                    step(v, i);
                }
            }
            chunk_end(cb, len);
        }
        if (c + 1 < ntile) {
            wave_lds_sync();
#pragma unroll
            for (int a = 0; a < NA; ++a) { st[a].store(c + 1); st[a].load(c + 2 < ntile ? c + 2 : -1); }
            wave_lds_sync();
        }
    }
}

// Rows n-1 .. 0 of NA arrays staged together: step(v, i) with v[a] = arr[a][i], then
// chunk_end(lo, m) after each chunk of rows [lo, lo + m) (first the n % 16 top rows, then
// 16-row chunks downwards).  ring holds NA rings of 2 T doubles.  Full-wave entry.
template <int T, int NA, class Step, class ChunkEnd>
__device__ __forceinline__ void sweep_down_compact(const double *const (&arr)[NA], double *ring, int n, Step step,
                                             ChunkEnd chunk_end) {
    constexpr int G = T >= 256 ? 16 : 8;   // rows per LDS read group
    if (n <= 0) return;
    const int lane = (int)(threadIdx.x & 63);
    WTileStage<T> st[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) st[a] = WTileStage<T>{arr[a], ring + a * 2 * T, n, lane};
    const int ntile = (n + T - 1) / T;
#pragma unroll
    for (int a = 0; a < NA; ++a) { st[a].load(ntile - 1); st[a].store(ntile - 1); st[a].load(ntile - 2); }
    wave_lds_sync();
    const int rem = n % NCPA_CHUNK;
    for (int c = ntile - 1; c >= 0; --c) {
        const int off = (c & 1) * T;
        const int base = c * T;
        int top = min(n, base + T);
        if (c == ntile - 1 && rem) {
            for (int i = n - 1; i >= n - rem; --i) {
                double v[NA];
#pragma unroll
                for (int a = 0; a < NA; ++a) v[a] = ring[a * 2 * T + off + i - base];
                step(v, i);
            }
            chunk_end(n - rem, rem);
            top = n - rem;
        }
        for (int cb = top - NCPA_CHUNK; cb >= base; cb -= NCPA_CHUNK) {
#pragma unroll
            for (int h = NCPA_CHUNK / G - 1; h >= 0; --h) {   // the chunk's read groups, top first
                double va[NA][G];
                lds_rows<T, NA, G>(va, ring, off + (cb + G * h - base));
#pragma unroll
                for (int u = G - 1; u >= 0; --u) {
                    double v[NA];
#pragma unroll
                    for (int a = 0; a < NA; ++a) v[a] = va[a][u];
                    step(v, cb + G * h + u);
                }
            }
            chunk_end(cb, NCPA_CHUNK);
        }
        if (c > 0) {
            wave_lds_sync();
#pragma unroll
            for (int a = 0; a < NA; ++a) { st[a].store(c - 1); st[a].load(c - 2); }
            wave_lds_sync();
        }
    }
}

// Rows 0 .. m-1 (ascending) of NA arrays, chunk_end(lo, m') after every chunk of 16 rows and
// after the last m % 16 rows.  Full-wave entry.
template <int T, int NA, class Step, class ChunkEnd>
__device__ __forceinline__ void sweep_up_compact(const double *const (&arr)[NA], double *ring, int m, Step step,
                                           ChunkEnd chunk_end) {
    constexpr int G = T >= 256 ? 16 : 8;
    if (m <= 0) return;
    const int lane = (int)(threadIdx.x & 63);
    WTileStage<T> st[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) st[a] = WTileStage<T>{arr[a], ring + a * 2 * T, m, lane};
    const int ntile = (m + T - 1) / T;
#pragma unroll
    for (int a = 0; a < NA; ++a) { st[a].load(0); st[a].store(0); st[a].load(1 < ntile ? 1 : -1); }
    wave_lds_sync();
    for (int c = 0; c < ntile; ++c) {
        const int off = (c & 1) * T;
        const int base = c * T;
        const int endt = min(m, base + T);
        for (int cb = base; cb < endt; cb += NCPA_CHUNK) {
            const int len = min(NCPA_CHUNK, endt - cb);
            if (len == NCPA_CHUNK) {
#pragma unroll
                for (int h = 0; h < NCPA_CHUNK / G; ++h) {
                    double va[NA][G];
                    lds_rows<T, NA, G>(va, ring, off + (cb + G * h - base));
#pragma unroll
                    for (int u = 0; u < G; ++u) {
                        double v[NA];
#pragma unroll
                        for (int a = 0; a < NA; ++a) v[a] = va[a][u];
                        step(v, cb + G * h + u);
                    }
                }
            } else {
                for (int i = cb; i < cb + len; ++i) {
                    double v[NA];
#pragma unroll
                    for (int a = 0; a < NA; ++a) v[a] = ring[a * 2 * T + off + i - base];
                    step(v, i);
                }
            }
            chunk_end(cb, len);
        }
        if (c + 1 < ntile) {
            wave_lds_sync();
#pragma unroll
            for (int a = 0; a < NA; ++a) { st[a].store(c + 1); st[a].load(c + 2 < ntile ? c + 2 : -1); }
            wave_lds_sync();
        }
    }
}

// The sweeps of a WideSweeps<T>: T = 256 the 16-row reads of sturm.h's tiles, else compact.
template <int T, int NA, class Step, class ChunkEnd>
__device__ __forceinline__ void sweep_down_n(const double *const (&arr)[NA], double *ring, int n, Step step,
                                             ChunkEnd chunk_end) {
    if constexpr (T == NCPA_TILE) sweep_down_big<NA>(arr, ring, n, step, chunk_end);
    else sweep_down_compact<T, NA>(arr, ring, n, step, chunk_end);
}
template <int T, int NA, class Step, class ChunkEnd>
__device__ __forceinline__ void sweep_up_n(const double *const (&arr)[NA], double *ring, int m, Step step,
                                           ChunkEnd chunk_end) {
    if constexpr (T == NCPA_TILE) sweep_up_big<NA>(arr, ring, m, step, chunk_end);
    else sweep_up_compact<T, NA>(arr, ring, m, step, chunk_end);
}

__device__ __forceinline__ double wide_row(const double (&v)[3], double k) { return fma(fma(v[2], k, v[1]), k, v[0]); }

// The wide-angle pencil of one problem: rows A, B, C of dz^2 T(k) and absorption weights W.
// ring: 4 staged arrays (4 x 2 T doubles).
//
// With PQ the counts and Newton sweeps use the factored rows of the physical operator
// (WModeSolver.cpp:549-553): with P = dz omega/c0, Q = dz wc/c0,
//     A = P^2 - 2 (+ beta on row 0),  B = -2 P Q,  C = Q^2 - dz^2,
//     a_i(k) = (P_i - k Q_i)^2 - 2 - dz^2 k^2 (+ beta = 1/(dz adm + 1) on row 0),
// two staged arrays instead of three (the rows 1 .. n-1 are swept, row 0 is stepped last).
template <int T, bool PQ = false>
struct WideSweeps {
    const double *A, *B, *C, *W;
    int n;
    const double *P = nullptr, *Q = nullptr;
    double h2 = 0.0, beta = 0.0;

    // number of negative eigenvalues of dz^2 T(k)
    __device__ __forceinline__ int count(double *ring, double k) const {
        double p1 = 1.0, p2 = 0.0;
        int cnt = 0;
        unsigned bits = 0u;
        if constexpr (PQ) {
            const double K = fma(-h2 * k, k, -2.0);
            const double *const arr[2] = {P + 1, Q + 1};
            auto chunk_end = [&](int, int m) {
                cnt += sign_changes(bits, m);
                bits &= 1u;
                const int e = pow2_exponent(p1);
                p1 = ldexp(p1, -e);
                p2 = ldexp(p2, -e);
            };
            sweep_down_n<T, 2>(
                arr, ring, n - 1,
                [&](const double (&v)[2], int) {
                    const double t = fma(-k, v[1], v[0]);
                    const double p0 = fma(fma(t, t, K), p1, -p2);
                    bits = push_sign(bits, p0);
                    p2 = p1;
                    p1 = p0;
                },
                chunk_end);
            const double t = fma(-k, Q[0], P[0]);
            const double p0 = fma(fma(t, t, K) + beta, p1, -p2);
            bits = push_sign(bits, p0);
            p2 = p1;
            p1 = p0;
            chunk_end(0, 1);
            return cnt;
        }
        const double *const arr[3] = {A, B, C};
        sweep_down_n<T, 3>(
            arr, ring, n,
            [&](const double (&v)[3], int) {
                const double p0 = fma(wide_row(v, k), p1, -p2);
                bits = push_sign(bits, p0);
                p2 = p1;
                p1 = p0;
            },
            [&](int, int m) {
                cnt += sign_changes(bits, m);
                bits &= 1u;
                const int e = pow2_exponent(p1);
                p1 = ldexp(p1, -e);
                p2 = ldexp(p2, -e);
            });
        return cnt;
    }

    // count plus det dz^2 T(k) and its k-derivative (Newton step -p0/q0 in k); COUNT false
    // leaves out the sign history (the count is then taken from det's sign by the caller)
    template <bool COUNT = true>
    __device__ __forceinline__ NewtonOut newton(double *ring, double k) const {
        double p1 = 1.0, p2 = 0.0, q1 = 0.0, q2 = 0.0;
        int cnt = 0;
        unsigned bits = 0u;
        if constexpr (PQ) {
            // q carried as -q/2: da/dk = -2 (Q t + dz^2 k) with t = P - k Q
            const double K = fma(-h2 * k, k, -2.0), hk = h2 * k;
            const double *const arr[2] = {P + 1, Q + 1};
            auto step = [&](double a, double m) {
                const double p0 = fma(a, p1, -p2);
                const double q0 = fma(a, q1, fma(m, p1, -q2));
                if constexpr (COUNT) bits = push_sign(bits, p0);
                p2 = p1;
                p1 = p0;
                q2 = q1;
                q1 = q0;
            };
            auto chunk_end = [&](int, int m) {
                if constexpr (COUNT) {
                    cnt += sign_changes(bits, m);
                    bits &= 1u;
                }
                const int e = pow2_exponent(p1);
                p1 = ldexp(p1, -e);
                p2 = ldexp(p2, -e);
                q1 = ldexp(q1, -e);
                q2 = ldexp(q2, -e);
            };
            sweep_down_n<T, 2>(
                arr, ring, n - 1,
                [&](const double (&v)[2], int) {
                    const double t = fma(-k, v[1], v[0]);
                    step(fma(t, t, K), fma(v[1], t, hk));
                },
                chunk_end);
            const double t = fma(-k, Q[0], P[0]);
            step(fma(t, t, K) + beta, fma(Q[0], t, hk));
            chunk_end(0, 1);
            NewtonOut o;
            o.cnt = cnt;
            o.p0 = p1;
            o.q0 = -2.0 * q1;
            return o;
        }
        const double k2 = k + k;
        const double *const arr[3] = {A, B, C};
        sweep_down_n<T, 3>(
            arr, ring, n,
            [&](const double (&v)[3], int) {
                const double am = wide_row(v, k);
                const double da = fma(k2, v[2], v[1]);
                const double p0 = fma(am, p1, -p2);
                const double q0 = fma(am, q1, fma(da, p1, -q2));
                if constexpr (COUNT) bits = push_sign(bits, p0);
                p2 = p1;
                p1 = p0;
                q2 = q1;
                q1 = q0;
            },
            [&](int, int m) {
                if constexpr (COUNT) {
                    cnt += sign_changes(bits, m);
                    bits &= 1u;
                }
                const int e = pow2_exponent(p1);
                p1 = ldexp(p1, -e);
                p2 = ldexp(p2, -e);
                q1 = ldexp(q1, -e);
                q2 = ldexp(q2, -e);
            });
        NewtonOut o;
        o.cnt = cnt;
        o.p0 = p1;
        o.q0 = q1;
        return o;
    }

    // backward sweep at a converged k: det, its identity-shift derivative q^mu (q0), the
    // absorption derivative (e0) and the records phi_{s+1}, phi_{r+1}
    __device__ __forceinline__ FinalOut final_sweep(double *ring, double k, int srow, int rrow) const {
        double p1 = 1.0, p2 = 0.0, q1 = 0.0, q2 = 0.0, e1 = 0.0, e2 = 0.0;
        int ex = 0;
        FinalOut o;
        o.phi_s1 = SVal{1.0, 0};
        o.phi_r1 = SVal{1.0, 0};
        if constexpr (PQ) {
            const double K = fma(-h2 * k, k, -2.0);
            auto step = [&](double am, double w) {
                const double p0 = fma(am, p1, -p2);
                const double q0 = fma(am, q1, -(q2 + p1));
                const double e0 = fma(am, e1, fma(w, p1, -e2));
                p2 = p1;
                p1 = p0;
                q2 = q1;
                q1 = q0;
                e2 = e1;
                e1 = e0;
            };
            const double *const arr[3] = {P + 1, Q + 1, W + 1};
            sweep_down_n<T, 3>(
                arr, ring, n - 1,
                [&](const double (&v)[3], int i1) {   // row i1 + 1
                    const double t = fma(-k, v[1], v[0]);
                    step(fma(t, t, K), v[2]);
                    if (i1 == srow) o.phi_s1 = SVal{p1, ex};
                    if (i1 == rrow) o.phi_r1 = SVal{p1, ex};
                },
                [&](int, int) {
                    const int e = pow2_exponent(p1);
                    p1 = ldexp(p1, -e);
                    p2 = ldexp(p2, -e);
                    q1 = ldexp(q1, -e);
                    q2 = ldexp(q2, -e);
                    e1 = ldexp(e1, -e);
                    e2 = ldexp(e2, -e);
                    ex += e;
                });
            const double t = fma(-k, Q[0], P[0]);
            step(fma(t, t, K) + beta, W[0]);   // row 0 (no record: records are rows s + 1 >= 1)
            o.p0 = p1;
            o.q0 = q1;
            o.e0 = e1;
            o.ex = ex;
            return o;
        }
        const double *const arr[4] = {A, B, C, W};
        sweep_down_n<T, 4>(
            arr, ring, n,
            [&](const double (&v)[4], int i) {
                const double am = fma(fma(v[2], k, v[1]), k, v[0]);
                const double p0 = fma(am, p1, -p2);
                const double q0 = fma(am, q1, -(q2 + p1));
                const double e0 = fma(am, e1, fma(v[3], p1, -e2));
                p2 = p1;
                p1 = p0;
                q2 = q1;
                q1 = q0;
                e2 = e1;
                e1 = e0;
                if (i == srow + 1) o.phi_s1 = SVal{p1, ex};
                if (i == rrow + 1) o.phi_r1 = SVal{p1, ex};
            },
            [&](int, int) {
                const int e = pow2_exponent(p1);
                p1 = ldexp(p1, -e);
                p2 = ldexp(p2, -e);
                q1 = ldexp(q1, -e);
                q2 = ldexp(q2, -e);
                e1 = ldexp(e1, -e);
                e2 = ldexp(e2, -e);
                ex += e;
            });
        o.p0 = p1;
        o.q0 = q1;
        o.e0 = e1;
        o.ex = ex;
        return o;
    }

    // forward continuants theta_{s-1}, theta_{r-1} of dz^2 T(k) (theta_{-1} = 1)
    __device__ __forceinline__ void theta(double *ring, double k, int srow, int rrow, SVal &th_s, SVal &th_r) const {
        th_s = SVal{1.0, 0};
        th_r = SVal{1.0, 0};
        const int top = srow > rrow ? srow : rrow;
        double t1 = 1.0, t2 = 0.0;
        int ex = 0;
        if constexpr (PQ) {
            if (top <= 0) return;
            const double K = fma(-h2 * k, k, -2.0);
            const double t = fma(-k, Q[0], P[0]);
            t2 = t1;
            t1 = fma(t, t, K) + beta;   // theta_0 (row 0)
            if (srow == 1) th_s = SVal{t1, ex};
            if (rrow == 1) th_r = SVal{t1, ex};
            const double *const arr[2] = {P + 1, Q + 1};
            sweep_up_n<T, 2>(
                arr, ring, top - 1,
                [&](const double (&v)[2], int i1) {   // row i1 + 1
                    const double tt = fma(-k, v[1], v[0]);
                    const double t0 = fma(fma(tt, tt, K), t1, -t2);
                    t2 = t1;
                    t1 = t0;
                    if (i1 + 2 == srow) th_s = SVal{t1, ex};
                    if (i1 + 2 == rrow) th_r = SVal{t1, ex};
                },
                [&](int, int) {
                    const int e = pow2_exponent(t1);
                    t1 = ldexp(t1, -e);
                    t2 = ldexp(t2, -e);
                    ex += e;
                });
            return;
        }
        const double *const arr[3] = {A, B, C};
        sweep_up_n<T, 3>(
            arr, ring, top,
            [&](const double (&v)[3], int i) {
                const double t0 = fma(wide_row(v, k), t1, -t2);
                t2 = t1;
                t1 = t0;
                if (i == srow - 1) th_s = SVal{t1, ex};
                if (i == rrow - 1) th_r = SVal{t1, ex};
            },
            [&](int, int) {
                const int e = pow2_exponent(t1);
                t1 = ldexp(t1, -e);
                t2 = ldexp(t2, -e);
                ex += e;
            });
    }

    // rows-pass sweeps with the row value a_i(k) handed to step(am, i)
    // (PQ: row 0 stepped first going up and last going down, the chunks of rows 1 .. m-1 between)
    template <class Step, class ChunkEnd>
    __device__ __forceinline__ void up(double *ring, int m, double k, Step step, ChunkEnd ce) const {
        if constexpr (PQ) {
            if (m <= 0) return;
            const double K = fma(-h2 * k, k, -2.0);
            const double t = fma(-k, Q[0], P[0]);
            step(fma(t, t, K) + beta, 0);
            ce(0, 1);
            const double *const arr[2] = {P + 1, Q + 1};
            sweep_up_n<T, 2>(
                arr, ring, m - 1,
                [&](const double (&v)[2], int i1) {
                    const double tt = fma(-k, v[1], v[0]);
                    step(fma(tt, tt, K), i1 + 1);
                },
                [&](int lo1, int len) { ce(lo1 + 1, len); });
            return;
        }
        const double *const arr[3] = {A, B, C};
        sweep_up_n<T, 3>(arr, ring, m, [&](const double (&v)[3], int i) { step(wide_row(v, k), i); }, ce);
    }
    template <class Step, class ChunkEnd>
    __device__ __forceinline__ void down(double *ring, int m, double k, Step step, ChunkEnd ce) const {
        if constexpr (PQ) {
            if (m <= 0) return;
            const double K = fma(-h2 * k, k, -2.0);
            const double *const arr[2] = {P + 1, Q + 1};
            sweep_down_n<T, 2>(
                arr, ring, m - 1,
                [&](const double (&v)[2], int i1) {
                    const double tt = fma(-k, v[1], v[0]);
                    step(fma(tt, tt, K), i1 + 1);
                },
                [&](int lo1, int len) { ce(lo1 + 1, len); });
            const double t = fma(-k, Q[0], P[0]);
            step(fma(t, t, K) + beta, 0);
            ce(0, 1);
            return;
        }
        const double *const arr[3] = {A, B, C};
        sweep_down_n<T, 3>(arr, ring, m, [&](const double (&v)[3], int i) { step(wide_row(v, k), i); }, ce);
    }
};

}  // namespace ncpa
