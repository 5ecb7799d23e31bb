// FP64 Sturm sweeps over the finite-difference Helmholtz tridiagonal.
//
// T = tridiag(b, a_i, b), constant off-diagonal b = 1/dz^2 (src/libmodes/EigenEngine.cpp:57-79).
// The sweeps work on T' = dz^2 T = tridiag(1, a'_i, 1), a'_i = dz^2 a_i, whose eigenvalues
// are dz^2 lambda and whose eigenvectors are those of T.  A sweep runs the continuant
// (three-term) recurrence from the top row down to the ground:
//     phi_n = 1, phi_{n+1} = 0,   phi_i = (a'_i - s) phi_{i+1} - phi_{i+2}
// so phi_i = det(T'[i:n, i:n] - s) and phi_0 = det(T' - s).  The LDL^T pivots are
// D_i = phi_i / phi_{i+1}; negative pivots count the eigenvalues below s (Sylvester).
// With the unit off-diagonal a row is one subtract and one fma (the negation is an operand
// modifier); every 16 rows the state is rescaled by an exact power of two so nothing
// over- or underflows.
//
// Derivative continuants carried with it:
//     q_i = d phi_i / ds             -> Newton step on det:  -phi_0 / q_0
//     e_i = d phi_i / d eps  for T' + eps diag(w)  -> absorption u^T W u = -e_0 / q_0
// and the adjugate of a tridiagonal with unit off-diagonal gives the eigenvector products
// at an eigenvalue:
//     u_s u_r = -(-1)^(r-s) theta_{s-1} phi_{r+1} / q_0      (s <= r)
// with theta_k the forward continuant det(T[0:k+1, 0:k+1] - s) from the ground.  Each
// factor is computed in its stable direction, so tiny components (modes trapped aloft)
// come out tiny instead of at a noise floor.
//
// Memory: every lane of a wave sweeps the same problem, so the diagonal is wave-uniform.
// It is staged through a wave-private LDS ring of two 256-row tiles: the 64 lanes load the
// next tile with coalesced vector loads (lane l takes rows t + l + 64u) one whole tile of
// compute ahead of its use, then every lane reads the tile back with broadcast LDS reads,
// 16 rows at a time, the next 16 requested before the current ones are consumed.  A sweep
// must be entered by the full wave (all 64 lanes active): the lanes share the tile loads.
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>

namespace ncpa {

#define NCPA_TILE 256                 // rows per LDS tile
#define NCPA_RING (2 * NCPA_TILE)     // doubles of LDS per staged array (two tiles)
#define NCPA_CHUNK 16                 // rows between rescales / sign-count flushes

// exponent such that ldexp(x, -e) lies in [0.5, 1)
__device__ __forceinline__ int pow2_exponent(double x) {
    int e;
    (void)frexp(x, &e);
    return x == 0.0 ? 0 : e;
}

// An exact zero continuant needs no guard: counting sign changes of phi_{i+1}, phi_i,
// phi_{i-1} = -b^2 phi_{i+1} gives exactly one change whatever the sign bit of the zero,
// the same count as LAPACK's -pivmin replacement of a vanishing pivot; the recurrence
// itself continues with finite values.  Only phi_0 == 0 (s exactly an eigenvalue) is
// counted either way, which bisection tolerates.

// sign bit of x shifted into the low end of a 32-bit history word (one v_alignbit)
__device__ __forceinline__ unsigned push_sign(unsigned bits, double x) {
    return __builtin_amdgcn_alignbit(bits, (unsigned)__double2hiint(x), 31u);
}

// sign changes between consecutive entries of a history of m+1 sign bits (bit 0 newest)
__device__ __forceinline__ int sign_changes(unsigned bits, int m) {
    const unsigned mask = (m >= 32) ? 0xffffffffu : ((1u << m) - 1u);
    return __builtin_popcount((bits ^ (bits >> 1)) & mask);
}

// LDS writes of one lane are read by the other lanes of the same wave: LDS executes a
// wave's instructions in order, so only the compiler has to be kept from reordering.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One array staged tile by tile: load() issues the coalesced global loads of tile c into
// registers, store() writes them to the ring slot of tile c.
struct TileStage {
    const double *g;
    double *ring;
    int n;
    int lane;
    double pre[NCPA_TILE / 64];
    __device__ __forceinline__ void load(int c) {
#pragma unroll
        for (int u = 0; u < NCPA_TILE / 64; ++u) {
            const int r = c * NCPA_TILE + lane + 64 * u;
            pre[u] = (c >= 0 && r < n) ? g[r] : 0.0;
        }
    }
    __device__ __forceinline__ void store(int c) {
        double *s = ring + (c & 1) * NCPA_TILE;
#pragma unroll
        for (int u = 0; u < NCPA_TILE / 64; ++u) s[lane + 64 * u] = pre[u];
    }
};

__device__ __forceinline__ void lds_chunk(double (&v)[NCPA_CHUNK], const double *s) {
#pragma unroll
    for (int u = 0; u < NCPA_CHUNK; ++u) v[u] = s[u];
}

// The first chunk of a tile passes through an empty asm statement: were it a plain load
// like the prefetches inside the chunk loop, the compiler would merge the loop-carried pair
// of loads into one load at the top of the loop and wait on it there, undoing the
// prefetch.  (The asm only pins registers; it waits for this one chunk, once per tile.)
__device__ __forceinline__ void lds_chunk_first(double (&v)[NCPA_CHUNK], const double *s) {
#pragma unroll
    for (int u = 0; u < NCPA_CHUNK; ++u) {
        v[u] = s[u];
        asm volatile("" : "+v"(v[u]));
    }
}

// The full 16-row chunks of one tile, in sweep order: chunk k starts at row
// cb0 - 16k (Down) or cb0 + 16k (up); s points at the tile's first row `base`.  Two register
// buffers alternate: the next chunk is read from LDS before the current one is stepped
// through (the scheduling barrier keeps the reads ahead of the arithmetic).
template <bool Down, class Step>
__device__ __forceinline__ void step_chunk(const double (&v)[NCPA_CHUNK], int cb, Step &step) {
    if (Down) {
#pragma unroll
        for (int u = NCPA_CHUNK - 1; u >= 0; --u) step(v[u], cb + u);
    } else {
#pragma unroll
        for (int u = 0; u < NCPA_CHUNK; ++u) step(v[u], cb + u);
    }
}

// s_waitcnt lgkmcnt(0) (vmcnt and expcnt left at their maxima): the LDS reads issued a
// whole chunk earlier have returned; waiting here, before the next reads are issued, keeps
// the compiler's own wait from also covering the reads just issued.
__device__ __forceinline__ void wait_lds() { __builtin_amdgcn_s_waitcnt(0xc07f); }

template <bool Down, class Step, class ChunkEnd>
__device__ __forceinline__ void tile_chunks(const double *s, int base, int cb0, int nch, Step &step,
                                            ChunkEnd &chunk_end) {
    if (nch <= 0) return;
    const int dir = Down ? -NCPA_CHUNK : NCPA_CHUNK;
    double va[NCPA_CHUNK], vb[NCPA_CHUNK];
    lds_chunk_first(va, s + (cb0 - base));
    int k = 0;
    for (; k + 1 < nch; k += 2) {
        const int c0 = cb0 + k * dir, c1 = c0 + dir;
        lds_chunk(vb, s + (c1 - base));
        __builtin_amdgcn_sched_barrier(0);
        step_chunk<Down>(va, c0, step);
        chunk_end(c0, NCPA_CHUNK);
        __builtin_amdgcn_sched_barrier(0);
        wait_lds();
        const int c2 = (k + 2 < nch) ? c1 + dir : c1;
        lds_chunk(va, s + (c2 - base));
        __builtin_amdgcn_sched_barrier(0);
        step_chunk<Down>(vb, c1, step);
        chunk_end(c1, NCPA_CHUNK);
        __builtin_amdgcn_sched_barrier(0);
        wait_lds();
    }
    if (k < nch) {
        step_chunk<Down>(va, cb0 + k * dir, step);
        chunk_end(cb0 + k * dir, NCPA_CHUNK);
    }
}

// Drive `step(a_i, i)` over rows i = n-1 .. 0, then `chunk_end(lo, m)` after each chunk of
// rows [lo, lo + m): first the n % 16 top rows, then chunks [16c, 16c+15] downwards.
template <class Step, class ChunkEnd>
__device__ __forceinline__ void sweep_down(const double *a, double *ring, int n, Step step, ChunkEnd chunk_end) {
    if (n <= 0) return;
    TileStage st{a, ring, n, (int)(threadIdx.x & 63)};
    const int ntile = (n + NCPA_TILE - 1) / NCPA_TILE;
    st.load(ntile - 1);
    st.store(ntile - 1);
    st.load(ntile - 2);
    wave_lds_sync();
    const int rem = n % NCPA_CHUNK;
    for (int c = ntile - 1; c >= 0; --c) {
        const double *s = ring + (c & 1) * NCPA_TILE;
        const int base = c * NCPA_TILE;
        int top = min(n, base + NCPA_TILE);
        if (c == ntile - 1 && rem) {
            for (int i = n - 1; i >= n - rem; --i) step(s[i - base], i);
            chunk_end(n - rem, rem);
            top = n - rem;
        }
        tile_chunks<true>(s, base, top - NCPA_CHUNK, (top - base) / NCPA_CHUNK, step, chunk_end);
        if (c > 0) {
            wave_lds_sync();
            st.store(c - 1);
            st.load(c - 2);
            wave_lds_sync();
        }
    }
}

// two-array variant (diagonal and absorption weights), same chunking; ring2 holds two
// rings of NCPA_RING doubles.  Chunks for which `rec(lo)` holds are stepped with
// `step_rec` (a step that may record values at given rows), all others with `step`: the
// per-row row tests stay out of the chunks that need none.
template <class Step, class StepRec, class RecPred, class ChunkEnd>
__device__ __forceinline__ void sweep_down2(const double *a, const double *w, double *ring2, int n, Step step,
                                            StepRec step_rec, RecPred rec, ChunkEnd chunk_end) {
    if (n <= 0) return;
    const int lane = (int)(threadIdx.x & 63);
    TileStage sa{a, ring2, n, lane}, sw{w, ring2 + NCPA_RING, n, lane};
    const int ntile = (n + NCPA_TILE - 1) / NCPA_TILE;
    sa.load(ntile - 1);
    sw.load(ntile - 1);
    sa.store(ntile - 1);
    sw.store(ntile - 1);
    sa.load(ntile - 2);
    sw.load(ntile - 2);
    wave_lds_sync();
    const int rem = n % NCPA_CHUNK;
    for (int c = ntile - 1; c >= 0; --c) {
        const double *s = ring2 + (c & 1) * NCPA_TILE;
        const double *t = s + NCPA_RING;
        const int base = c * NCPA_TILE;
        int top = min(n, base + NCPA_TILE);
        if (c == ntile - 1 && rem) {
            for (int i = n - 1; i >= n - rem; --i) step_rec(s[i - base], t[i - base], i);
            chunk_end(n - rem, rem);
            top = n - rem;
        }
        for (int cb = top - NCPA_CHUNK; cb >= base; cb -= NCPA_CHUNK) {
            const bool r = rec(cb);
            // each array read 8 rows at a time (32 VGPRs for both, not 64)
#pragma unroll
            for (int h = NCPA_CHUNK / 8 - 1; h >= 0; --h) {
                double va[8], vw[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    va[u] = s[cb - base + 8 * h + u];
                    vw[u] = t[cb - base + 8 * h + u];
                }
                if (r) {
#pragma unroll
                    for (int u = 7; u >= 0; --u) step_rec(va[u], vw[u], cb + 8 * h + u);
                } else {
#pragma unroll
                    for (int u = 7; u >= 0; --u) step(va[u], vw[u], cb + 8 * h + u);
                }
            }
            chunk_end(cb, NCPA_CHUNK);
        }
        if (c > 0) {
            wave_lds_sync();
            sa.store(c - 1);
            sw.store(c - 1);
            sa.load(c - 2);
            sw.load(c - 2);
            wave_lds_sync();
        }
    }
}

// Drive `step(a_i, i)` over rows i = 0 .. m-1 (ascending), `chunk_end(lo, m)` after every
// chunk [lo, lo + m) of 16 rows and after the last m % 16 rows; same staging as sweep_down.
template <class Step, class ChunkEnd>
__device__ __forceinline__ void sweep_up(const double *a, double *ring, int m, Step step, ChunkEnd chunk_end) {
    if (m <= 0) return;
    TileStage st{a, ring, m, (int)(threadIdx.x & 63)};
    const int ntile = (m + NCPA_TILE - 1) / NCPA_TILE;
    st.load(0);
    st.store(0);
    st.load(1 < ntile ? 1 : -1);
    wave_lds_sync();
    const int nfull = m / NCPA_CHUNK;
    for (int c = 0; c < ntile; ++c) {
        const double *s = ring + (c & 1) * NCPA_TILE;
        const int base = c * NCPA_TILE;
        const int endc = max(base, min(nfull * NCPA_CHUNK, base + NCPA_TILE));   // end of full chunks
        tile_chunks<false>(s, base, base, (endc - base) / NCPA_CHUNK, step, chunk_end);
        const int endt = min(m, base + NCPA_TILE);
        for (int i = endc; i < endt; ++i) step(s[i - base], i);
        if (endt > endc) chunk_end(endc, endt - endc);
        if (c + 1 < ntile) {
            wave_lds_sync();
            st.store(c + 1);
            st.load(c + 2 < ntile ? c + 2 : -1);
            wave_lds_sync();
        }
    }
}

// ---- the diagonal through the scalar data cache into SGPRs ---------------------------------
// The alternative to the LDS ring for launches that fill the chip (several waves per SIMD):
// the rows are loaded by scalar loads (the address is wave-uniform, the array read-only for
// the kernel's lifetime: the constant address space), so each row's value reaches the VALU as
// an SGPR operand with no LDS traffic.  Scalar loads return out of order, so the only wait
// the compiler can place for a chunk is one for every scalar load in flight; left to itself
// it issues the next chunk's loads and then waits for all of them before the current chunk's
// first row.  Here the order is fixed: wait for the current chunk (explicit lgkmcnt(0), which
// the compiler's wait insertion sees), then issue the next chunk's loads, then step through
// the current chunk while they are in flight, with scheduling barriers between the three.
typedef const __attribute__((address_space(4))) double *cdptr;

__device__ __forceinline__ void smem_chunk(double (&v)[NCPA_CHUNK], cdptr p) {
#pragma unroll
    for (int u = 0; u < NCPA_CHUNK; ++u) v[u] = p[u];
}

// rows n-1 .. 0 as sweep_down (the n % 16 top rows first, then 16-row chunks downwards)
template <class Step, class ChunkEnd>
__device__ __forceinline__ void sweep_down_smem(const double *a_, int n, Step step, ChunkEnd chunk_end) {
    if (n <= 0) return;
    const cdptr a = (cdptr)a_;
    const int rem = n % NCPA_CHUNK;
    if (rem) {
        for (int i = n - 1; i >= n - rem; --i) step(a[i], i);
        chunk_end(n - rem, rem);
    }
    int c = n - rem - NCPA_CHUNK;   // first row of the current chunk
    if (c < 0) return;
    // One loop body with no branch between a chunk's loads and their use (the copy into
    // `cur` at its end): split over two bodies, the loads were sunk past the exit test into
    // the body that uses them, right behind their own wait.
    double cur[NCPA_CHUNK];
    smem_chunk(cur, a + c);
    for (;;) {
        wait_lds();   // lgkmcnt(0): the current chunk has arrived
        __builtin_amdgcn_sched_barrier(0);
        double nxt[NCPA_CHUNK];
        smem_chunk(nxt, a + (c >= NCPA_CHUNK ? c - NCPA_CHUNK : c));
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = NCPA_CHUNK - 1; u >= 0; --u) step(cur[u], c + u);
        chunk_end(c, NCPA_CHUNK);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < NCPA_CHUNK; ++u) cur[u] = nxt[u];
        if (c < NCPA_CHUNK) break;
        c -= NCPA_CHUNK;
    }
}

// rows 0 .. m-1 ascending as sweep_up (16-row chunks, then the m % 16 last rows), the
// diagonal by scalar loads in the order of sweep_down_smem
template <class Step, class ChunkEnd>
__device__ __forceinline__ void sweep_up_smem(const double *a_, int m, Step step, ChunkEnd chunk_end) {
    if (m <= 0) return;
    const cdptr a = (cdptr)a_;
    const int nfull = m / NCPA_CHUNK;
    if (nfull > 0) {
        int c = 0;
        double cur[NCPA_CHUNK];
        smem_chunk(cur, a);
        for (;;) {
            wait_lds();
            __builtin_amdgcn_sched_barrier(0);
            double nxt[NCPA_CHUNK];
            smem_chunk(nxt, a + (c + NCPA_CHUNK < nfull * NCPA_CHUNK ? c + NCPA_CHUNK : c));
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < NCPA_CHUNK; ++u) step(cur[u], c + u);
            chunk_end(c, NCPA_CHUNK);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < NCPA_CHUNK; ++u) cur[u] = nxt[u];
            c += NCPA_CHUNK;
            if (c >= nfull * NCPA_CHUNK) break;
        }
    }
    const int endc = nfull * NCPA_CHUNK;
    for (int i = endc; i < m; ++i) step(a[i], i);
    if (m > endc) chunk_end(endc, m - endc);
}

// two arrays (diagonal and absorption weights) by scalar loads, as sweep_down2: the n % 16
// top rows with step_rec, then 16-row chunks downwards, each read as two 8-row halves (the
// next half's loads in flight while the current one is stepped), step_rec where rec(lo)
template <class Step, class StepRec, class RecPred, class ChunkEnd>
__device__ __forceinline__ void sweep_down2_smem(const double *a_, const double *w_, int n, Step step,
                                                 StepRec step_rec, RecPred rec, ChunkEnd chunk_end) {
    if (n <= 0) return;
    const cdptr a = (cdptr)a_, w = (cdptr)w_;
    const int rem = n % NCPA_CHUNK;
    if (rem) {
        for (int i = n - 1; i >= n - rem; --i) step_rec(a[i], w[i], i);
        chunk_end(n - rem, rem);
    }
    const int c = n - rem - NCPA_CHUNK;   // lowest full chunk start is 0; chunks are 16-aligned
    if (c < 0) return;
    int h = c + 8;                        // first row of the current 8-row half
    double ca[8], cw[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        ca[u] = a[h + u];
        cw[u] = w[h + u];
    }
    for (;;) {
        wait_lds();   // lgkmcnt(0): the current half has arrived
        __builtin_amdgcn_sched_barrier(0);
        const int hn = h >= 8 ? h - 8 : h;
        double na[8], nw[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            na[u] = a[hn + u];
            nw[u] = w[hn + u];
        }
        __builtin_amdgcn_sched_barrier(0);
        if (rec(h & ~(NCPA_CHUNK - 1))) {
#pragma unroll
            for (int u = 7; u >= 0; --u) step_rec(ca[u], cw[u], h + u);
        } else {
#pragma unroll
            for (int u = 7; u >= 0; --u) step(ca[u], cw[u], h + u);
        }
        if ((h & (NCPA_CHUNK - 1)) == 0) chunk_end(h, NCPA_CHUNK);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            ca[u] = na[u];
            cw[u] = nw[u];
        }
        if (h < 8) break;
        h -= 8;
    }
}

// rows n-1 .. 0 through the LDS ring (SMEM false) or scalar loads (SMEM true)
template <bool SMEM, class Step, class ChunkEnd>
__device__ __forceinline__ void sweep_down_via(const double *a, double *ring, int n, Step step, ChunkEnd chunk_end) {
    if constexpr (SMEM)
        sweep_down_smem(a, n, step, chunk_end);
    else
        sweep_down(a, ring, n, step, chunk_end);
}

// count of eigenvalues below s (full-wave entry; s may differ per lane)
template <bool SMEM = false>
__device__ __forceinline__ int sturm_count(const double *a, double *ring, int n, double s) {
    double p1 = 1.0, p2 = 0.0;
    int cnt = 0;
    unsigned bits = 0u;   // bit 0: sign of the newest continuant; starts with phi_n = +1
    sweep_down_via<SMEM>(
        a, ring, n,
        [&](double ai, int) {
            const double p0 = fma(ai - s, p1, -p2);
            bits = push_sign(bits, p0);
            p2 = p1;
            p1 = p0;
        },
        [&](int, int m) {
            cnt += sign_changes(bits, m);
            bits &= 1u;
            const int k = pow2_exponent(p1);
            p1 = ldexp(p1, -k);
            p2 = ldexp(p2, -k);
        });
    return cnt;
}

struct NewtonOut {
    int cnt;
    double p0, q0;   // same scale
};

// count plus det and d det/ds (Newton step -p0/q0)
template <bool SMEM = false>
__device__ __forceinline__ NewtonOut sturm_newton(const double *a, double *ring, int n, double s) {
    double p1 = 1.0, p2 = 0.0, q1 = 0.0, q2 = 0.0;
    int cnt = 0;
    unsigned bits = 0u;
    sweep_down_via<SMEM>(
        a, ring, n,
        [&](double ai, int) {
            const double am = ai - s;
            const double p0 = fma(am, p1, -p2);
            const double q0 = fma(am, q1, -(q2 + p1));
            bits = push_sign(bits, p0);
            p2 = p1;
            p1 = p0;
            q2 = q1;
            q1 = q0;
        },
        [&](int, int m) {
            cnt += sign_changes(bits, m);
            bits &= 1u;
            const int k = pow2_exponent(p1);
            p1 = ldexp(p1, -k);
            p2 = ldexp(p2, -k);
            q1 = ldexp(q1, -k);
            q2 = ldexp(q2, -k);
        });
    NewtonOut o;
    o.cnt = cnt;
    o.p0 = p1;
    o.q0 = q1;
    return o;
}

// det and d det/ds without the count: for the phase in which every shift of the wave lies in
// a bracket holding a single eigenvalue, where the count follows from the sign of det alone
// (the sign changes from phi_n = +1 down to phi_0 are even exactly when phi_0's sign bit is
// clear, so count(s) is cl or cl + 1 by that bit): one instruction less per row.
template <bool SMEM = false>
__device__ __forceinline__ NewtonOut sturm_newton_nocount(const double *a, double *ring, int n, double s) {
    double p1 = 1.0, p2 = 0.0, q1 = 0.0, q2 = 0.0;
    sweep_down_via<SMEM>(
        a, ring, n,
        [&](double ai, int) {
            const double am = ai - s;
            const double p0 = fma(am, p1, -p2);
            const double q0 = fma(am, q1, -(q2 + p1));
            p2 = p1;
            p1 = p0;
            q2 = q1;
            q1 = q0;
        },
        [&](int, int) {
            const int k = pow2_exponent(p1);
            p1 = ldexp(p1, -k);
            p2 = ldexp(p2, -k);
            q1 = ldexp(q1, -k);
            q2 = ldexp(q2, -k);
        });
    NewtonOut o;
    o.cnt = -1;
    o.p0 = p1;
    o.q0 = q1;
    return o;
}

// count of a shift inside a bracket whose lower end has count cl and that holds one
// eigenvalue, from the sign bit of det(T' - s) there
__device__ __forceinline__ int count_from_sign(double det, int cl) {
    const int odd = (int)((unsigned)__double2hiint(det) >> 31);
    return cl + ((cl ^ odd) & 1);
}

// scaled value: v * 2^e
struct SVal {
    double v;
    int e;
};

struct FinalOut {
    double p0, q0, e0;   // same scale
    int ex;              // scale exponent of p0, q0, e0
    SVal phi_s1, phi_r1; // phi_{s+1}, phi_{r+1}
};

// backward sweep with absorption derivative and the records phi_{s+1}, phi_{r+1}
template <bool SMEM = false>
__device__ __forceinline__ FinalOut sturm_final(const double *a, const double *w, double *ring2, int n, double s,
                                                int srow, int rrow) {
    double p1 = 1.0, p2 = 0.0, q1 = 0.0, q2 = 0.0, e1 = 0.0, e2 = 0.0;
    int ex = 0;
    FinalOut o;
    o.phi_s1 = SVal{1.0, 0};
    o.phi_r1 = SVal{1.0, 0};
    auto step = [&](double ai, double wi, int) {
        const double am = ai - s;
        const double p0 = fma(am, p1, -p2);
        const double q0 = fma(am, q1, -(q2 + p1));
        const double e0 = fma(am, e1, fma(wi, p1, -e2));
        p2 = p1;
        p1 = p0;
        q2 = q1;
        q1 = q0;
        e2 = e1;
        e1 = e0;
    };
    auto step_rec = [&](double ai, double wi, int i) {
        step(ai, wi, i);
        if (i == srow + 1) o.phi_s1 = SVal{p1, ex};
        if (i == rrow + 1) o.phi_r1 = SVal{p1, ex};
    };
    auto rec = [&](int lo) {
        return (unsigned)(srow + 1 - lo) < (unsigned)NCPA_CHUNK || (unsigned)(rrow + 1 - lo) < (unsigned)NCPA_CHUNK;
    };
    auto chunk_end = [&](int, int) {
        const int k = pow2_exponent(p1);
        p1 = ldexp(p1, -k);
        p2 = ldexp(p2, -k);
        q1 = ldexp(q1, -k);
        q2 = ldexp(q2, -k);
        e1 = ldexp(e1, -k);
        e2 = ldexp(e2, -k);
        ex += k;
    };
    if constexpr (SMEM)
        sweep_down2_smem(a, w, n, step, step_rec, rec, chunk_end);
    else
        sweep_down2(a, w, ring2, n, step, step_rec, rec, chunk_end);
    o.p0 = p1;
    o.q0 = q1;
    o.e0 = e1;
    o.ex = ex;
    return o;
}

// Rows an eigenvalue sweep may leave out at the top of T' - s for every shift s >= s_min.
// Returns n_eff: sweeps run rows n_eff-1 .. 0 from phi_{n_eff} = 1, phi_{n_eff+1} = 0 and add
// n - n_eff to the count.
//  * Counts stay exact.  Every row i >= t0 has fl(a'_i - s_min) <= -2 (the subtraction the
//    sweep itself performs, monotone in s), so in the full recurrence from the top each pivot
//    D_i = (a'_i - s) - 1/D_{i+1} is below -1 (D_n = a'_{n-1} - s <= -2; D_{i+1} < -1 gives
//    -1/D_{i+1} < 1): exactly one negative pivot per row above t0, the n - n_eff added back.
//  * The truncated start replaces the true pivot entering row n_eff - 1 by infinity.  Down the
//    evanescent rows t0 .. n_eff-1 that difference is multiplied by 1/D_i^2 per row, where
//    |D_i| tracks the growth factor r_i of the decaying solution, r_i + 1/r_i = 2 + x_i,
//    x_i = s - a'_i - 2: ln(1/D_i^2) ~ -2 acosh(1 + x_i/2).  The cut is placed where that sum,
//    counted upwards from t0, first reaches `decay` (80: e^-80, so that even a WKB estimate off
//    by a factor two leaves the pivot at t0 exact to far below rounding).  Eigenvalues then
//    move by less than rounding; the left-out rows carry eigenvector components below
//    e^(-decay/2) of those at t0.
//  * n_eff >= min_keep (rows the caller records), and n when no cut qualifies.
// Whole-wave call (64 lanes), reads a' straight from global memory: ~(n - t0) / 64 loads per
// lane, a fraction of one sweep.
// Generic over the row values: am(i) returns row i's diagonal entry of the shifted matrix at
// the lower bound (a'_i - s_min here; the wide-angle pencil's a_i(k_min)); every later shift
// must leave rows with am(i) <= -2 at or below -2.
// Rows with am(i) <= -2 - margin count as evanescent (margin 0 where am is one rounded
// subtraction, monotone in the shift; a small margin where am is a longer expression).
template <class RowAt>
__device__ __forceinline__ int evanescent_rows_at(RowAt am, int n, double decay, int min_keep, double margin = 0.0) {
    const int lane = (int)(threadIdx.x & 63);
    // t0: one above the highest row that is not evanescent at the bound
    int t0 = 0;
    for (int top = n - 1; top >= 0; top -= 64) {
        const int i = top - lane;
        const bool osc = i >= 0 && !(am(i) <= -2.0 - margin);   // (NaN counts as oscillatory)
        const unsigned long long m = __ballot(osc);
        if (m) {
            t0 = top - (__ffsll((unsigned long long)m) - 1) + 1;
            break;
        }
    }
    int cut = n;   // highest row kept
    double carry = 0.0;
    for (int base = t0; base < n; base += 64) {
        const int i = base + lane;
        double d = 0.0;
        if (i < n) {
            const double y = 0.5 * (-2.0 - am(i));          // x_i / 2 >= 0
            d = 2.0 * log1p(y + sqrt(y * (y + 2.0)));       // 2 acosh(1 + y)
        }
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const double v = __shfl_up(d, o);
            if (lane >= o) d += v;
        }
        const unsigned long long m = __ballot(carry + d >= decay);
        if (m) {
            cut = base + __ffsll((unsigned long long)m) - 1;
            break;
        }
        carry += __shfl(d, 63);
    }
    const int n_eff = cut < n ? cut + 1 : n;
    return min(n, max(n_eff, min_keep));
}

__device__ __forceinline__ int evanescent_rows(const double *a, int n, double s_min, double decay, int min_keep) {
    return evanescent_rows_at([&](int i) { return a[i] - s_min; }, n, decay, min_keep);
}

// forward continuants theta_{s-1}, theta_{r-1} (theta_{-1} = 1); rows 0 .. max(s,r)-1
template <bool SMEM = false>
__device__ __forceinline__ void theta_forward(const double *a, double *ring, double s, int srow, int rrow, SVal &th_s,
                                              SVal &th_r) {
    th_s = SVal{1.0, 0};
    th_r = SVal{1.0, 0};
    const int top = srow > rrow ? srow : rrow;
    double t1 = 1.0, t2 = 0.0;
    int ex = 0;
    auto step = [&](double ak, int k) {
        const double t0 = fma(ak - s, t1, -t2);
        t2 = t1;
        t1 = t0;
        if (k == srow - 1) th_s = SVal{t1, ex};
        if (k == rrow - 1) th_r = SVal{t1, ex};
    };
    auto chunk_end = [&](int, int) {
        const int kk = pow2_exponent(t1);
        t1 = ldexp(t1, -kk);
        t2 = ldexp(t2, -kk);
        ex += kk;
    };
    if constexpr (SMEM)
        sweep_up_smem(a, top, step, chunk_end);
    else
        sweep_up(a, ring, top, step, chunk_end);
}

// u_lo u_hi = -(-1)^(hi-lo) theta_{lo-1} phi_{hi+1} / q_0
__device__ __forceinline__ double adjugate_product(SVal th, SVal ph, double q0, int qex, int dist) {
    if (th.v == 0.0 || ph.v == 0.0 || q0 == 0.0) return 0.0;
    const double sgn = ((dist & 1) ? 1.0 : -1.0) * copysign(1.0, th.v) * copysign(1.0, ph.v) *
                       copysign(1.0, q0);
    const double l2 = log2(fabs(th.v)) + th.e + log2(fabs(ph.v)) + ph.e - log2(fabs(q0)) - qex;
    return sgn * exp2(l2);
}

}  // namespace ncpa
