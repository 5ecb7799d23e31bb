// FP64 Sturm sweeps over the finite-difference Helmholtz tridiagonal.
//
// T = tridiag(b, a_i, b), constant off-diagonal b = 1/dz^2 (src/libmodes/EigenEngine.cpp:57-79).
// A sweep runs the continuant (three-term) recurrence from the top row down to the ground:
//     phi_n = 1, phi_{n+1} = 0,   phi_i = (a_i - s) phi_{i+1} - b^2 phi_{i+2}
// so phi_i = det(T[i:n, i:n] - s) and phi_0 = det(T - s).  The LDL^T pivots are
// D_i = phi_i / phi_{i+1}; negative pivots count the eigenvalues below s (Sylvester).
// No division sits on the dependency chain (one mul + one fma per row); every 16 rows the
// state is rescaled by an exact power of two so nothing over- or underflows.
//
// Derivative continuants carried with it:
//     q_i = d phi_i / ds             -> Newton step on det:  -phi_0 / q_0
//     e_i = d phi_i / d eps  for T + eps diag(w)   -> absorption u^T W u = -e_0 / q_0
// and the adjugate of a tridiagonal gives the eigenvector products at an eigenvalue:
//     u_s u_r = -(-b)^(r-s) theta_{s-1} phi_{r+1} / q_0      (s <= r)
// with theta_k the forward continuant det(T[0:k+1, 0:k+1] - s) from the ground.  Each
// factor is computed in its stable direction, so tiny components (modes trapped aloft)
// come out tiny instead of at a noise floor.
//
// Memory: every lane of a wave sweeps the same problem, so the diagonal is wave-uniform.
// It is streamed through the scalar data cache (constant address space loads, 16 rows per
// chunk, the next chunk requested before the current one is consumed), so the per-row work
// is pure FP64 VALU.
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>

namespace ncpa {

typedef const __attribute__((address_space(4))) double *cdptr;

__device__ __forceinline__ cdptr as_const(const double *p) { return (cdptr)(p); }

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

#define NCPA_CHUNK 16

// Drive `step(a_i, i)` over rows i = n-1 .. 0, then `chunk_end(rows_in_chunk)` after each
// chunk of rows; the diagonal comes through double-buffered scalar loads.
template <class Step, class ChunkEnd>
__device__ __forceinline__ void sweep_down(cdptr a, int n, Step step, ChunkEnd chunk_end) {
    const int nfull = n / NCPA_CHUNK;          // chunks [16c, 16c+15], c < nfull
    const int rem = n - nfull * NCPA_CHUNK;    // top rows n-1 .. 16*nfull
    if (rem) {
        for (int i = n - 1; i >= nfull * NCPA_CHUNK; --i) step(a[i], i);
        chunk_end(rem);
    }
    if (nfull == 0) return;
    double buf[NCPA_CHUNK];
    int c = nfull - 1;
#pragma unroll
    for (int u = 0; u < NCPA_CHUNK; ++u) buf[u] = a[c * NCPA_CHUNK + u];
    for (; c >= 0; --c) {
        double nxt[NCPA_CHUNK];
        const int cn = c > 0 ? c - 1 : 0;
#pragma unroll
        for (int u = 0; u < NCPA_CHUNK; ++u) nxt[u] = a[cn * NCPA_CHUNK + u];
#pragma unroll
        for (int u = NCPA_CHUNK - 1; u >= 0; --u) step(buf[u], c * NCPA_CHUNK + u);
        chunk_end(NCPA_CHUNK);
#pragma unroll
        for (int u = 0; u < NCPA_CHUNK; ++u) buf[u] = nxt[u];
    }
}

// two-array variant (diagonal and absorption weights), 8-row chunks
#define NCPA_CHUNK2 8
template <class Step, class ChunkEnd>
__device__ __forceinline__ void sweep_down2(cdptr a, cdptr w, int n, Step step, ChunkEnd chunk_end) {
    const int nfull = n / NCPA_CHUNK2;
    const int rem = n - nfull * NCPA_CHUNK2;
    if (rem) {
        for (int i = n - 1; i >= nfull * NCPA_CHUNK2; --i) step(a[i], w[i], i);
        chunk_end(rem);
    }
    if (nfull == 0) return;
    double ba[NCPA_CHUNK2], bw[NCPA_CHUNK2];
    int c = nfull - 1;
#pragma unroll
    for (int u = 0; u < NCPA_CHUNK2; ++u) {
        ba[u] = a[c * NCPA_CHUNK2 + u];
        bw[u] = w[c * NCPA_CHUNK2 + u];
    }
    for (; c >= 0; --c) {
        double na[NCPA_CHUNK2], nw[NCPA_CHUNK2];
        const int cn = c > 0 ? c - 1 : 0;
#pragma unroll
        for (int u = 0; u < NCPA_CHUNK2; ++u) {
            na[u] = a[cn * NCPA_CHUNK2 + u];
            nw[u] = w[cn * NCPA_CHUNK2 + u];
        }
#pragma unroll
        for (int u = NCPA_CHUNK2 - 1; u >= 0; --u) step(ba[u], bw[u], c * NCPA_CHUNK2 + u);
        chunk_end(NCPA_CHUNK2);
#pragma unroll
        for (int u = 0; u < NCPA_CHUNK2; ++u) {
            ba[u] = na[u];
            bw[u] = nw[u];
        }
    }
}

// Drive `step(a_i, i)` over rows i = 0 .. m-1 (ascending), `chunk_end()` after every
// 16 rows; same scalar double-buffered streaming as sweep_down.
template <class Step, class ChunkEnd>
__device__ __forceinline__ void sweep_up(cdptr a, int m, Step step, ChunkEnd chunk_end) {
    const int nfull = m / NCPA_CHUNK;
    if (nfull > 0) {
        double buf[NCPA_CHUNK];
#pragma unroll
        for (int u = 0; u < NCPA_CHUNK; ++u) buf[u] = a[u];
        for (int c = 0; c < nfull; ++c) {
            double nxt[NCPA_CHUNK];
            const int cn = (c + 1 < nfull) ? c + 1 : c;
#pragma unroll
            for (int u = 0; u < NCPA_CHUNK; ++u) nxt[u] = a[cn * NCPA_CHUNK + u];
#pragma unroll
            for (int u = 0; u < NCPA_CHUNK; ++u) step(buf[u], c * NCPA_CHUNK + u);
            chunk_end();
#pragma unroll
            for (int u = 0; u < NCPA_CHUNK; ++u) buf[u] = nxt[u];
        }
    }
    for (int i = nfull * NCPA_CHUNK; i < m; ++i) step(a[i], i);
}

// count of eigenvalues below s
__device__ __forceinline__ int sturm_count(cdptr a, int n, double b2, double s) {
    double p1 = 1.0, p2 = 0.0;
    int cnt = 0;
    unsigned bits = 0u;   // bit 0: sign of the newest continuant; starts with phi_n = +1
    sweep_down(
        a, n,
        [&](double ai, int) {
            const double p0 = fma(ai - s, p1, -b2 * p2);
            bits = push_sign(bits, p0);
            p2 = p1;
            p1 = p0;
        },
        [&](int m) {
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
__device__ __forceinline__ NewtonOut sturm_newton(cdptr a, int n, double b2, double s) {
    double p1 = 1.0, p2 = 0.0, q1 = 0.0, q2 = 0.0;
    int cnt = 0;
    unsigned bits = 0u;
    sweep_down(
        a, n,
        [&](double ai, int) {
            const double am = ai - s;
            const double p0 = fma(am, p1, -b2 * p2);
            const double q0 = fma(am, q1, fma(-b2, q2, -p1));
            bits = push_sign(bits, p0);
            p2 = p1;
            p1 = p0;
            q2 = q1;
            q1 = q0;
        },
        [&](int m) {
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
__device__ __forceinline__ FinalOut sturm_final(cdptr a, cdptr w, int n, double b2, double s, int srow,
                                                int rrow) {
    double p1 = 1.0, p2 = 0.0, q1 = 0.0, q2 = 0.0, e1 = 0.0, e2 = 0.0;
    int ex = 0;
    FinalOut o;
    o.phi_s1 = SVal{1.0, 0};
    o.phi_r1 = SVal{1.0, 0};
    sweep_down2(
        a, w, n,
        [&](double ai, double wi, int i) {
            const double am = ai - s;
            const double p0 = fma(am, p1, -b2 * p2);
            const double q0 = fma(am, q1, fma(-b2, q2, -p1));
            const double e0 = fma(am, e1, fma(-b2, e2, wi * p1));
            p2 = p1;
            p1 = p0;
            q2 = q1;
            q1 = q0;
            e2 = e1;
            e1 = e0;
            if (i == srow + 1) o.phi_s1 = SVal{p1, ex};
            if (i == rrow + 1) o.phi_r1 = SVal{p1, ex};
        },
        [&](int) {
            const int k = pow2_exponent(p1);
            p1 = ldexp(p1, -k);
            p2 = ldexp(p2, -k);
            q1 = ldexp(q1, -k);
            q2 = ldexp(q2, -k);
            e1 = ldexp(e1, -k);
            e2 = ldexp(e2, -k);
            ex += k;
        });
    o.p0 = p1;
    o.q0 = q1;
    o.e0 = e1;
    o.ex = ex;
    return o;
}

// forward continuants theta_{s-1}, theta_{r-1} (theta_{-1} = 1); rows 0 .. max(s,r)-1
__device__ __forceinline__ void theta_forward(cdptr a, double b2, double s, int srow, int rrow, SVal &th_s,
                                              SVal &th_r) {
    th_s = SVal{1.0, 0};
    th_r = SVal{1.0, 0};
    const int top = srow > rrow ? srow : rrow;
    double t1 = 1.0, t2 = 0.0;
    int ex = 0;
    for (int k = 0; k < top; ++k) {
        const double t0 = fma(a[k] - s, t1, -b2 * t2);
        t2 = t1;
        t1 = t0;
        if (k == srow - 1) th_s = SVal{t1, ex};
        if (k == rrow - 1) th_r = SVal{t1, ex};
        if ((k & (NCPA_CHUNK - 1)) == NCPA_CHUNK - 1) {
            const int kk = pow2_exponent(t1);
            t1 = ldexp(t1, -kk);
            t2 = ldexp(t2, -kk);
            ex += kk;
        }
    }
}

// u_lo u_hi = -(-1)^(hi-lo) b^(hi-lo) theta_{lo-1} phi_{hi+1} / q_0
__device__ __forceinline__ double adjugate_product(SVal th, SVal ph, double q0, int qex, int dist,
                                                   double log2b) {
    if (th.v == 0.0 || ph.v == 0.0 || q0 == 0.0) return 0.0;
    const double sgn = ((dist & 1) ? 1.0 : -1.0) * copysign(1.0, th.v) * copysign(1.0, ph.v) *
                       copysign(1.0, q0);
    const double l2 = log2(fabs(th.v)) + th.e + log2(fabs(ph.v)) + ph.e - log2(fabs(q0)) - qex +
                      dist * log2b;
    return sgn * exp2(l2);
}

}  // namespace ncpa
