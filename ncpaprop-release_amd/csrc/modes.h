// Shared device code of the eigenvalue kernels: the block reductions, the modal interval,
// the search of one 64-lane block over 64 consecutive eigenvalue indices and the
// eigenvector-rows pass, written for any pencil (ess.hip: dz^2 T - s; wide.hip: dz^2 T(k)).
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "internal.h"
#include "sturm.h"

namespace ncpa {

// ------------------------------------------------------------------------------------
// block reductions (256-thread blocks)
// ------------------------------------------------------------------------------------
__device__ inline double block_min(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = sh[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) r = fmin(r, sh[k]);
    __syncthreads();
    return r;
}
__device__ inline double block_sum(double v, double *sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = sh[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) r += sh[k];
    __syncthreads();
    return r;
}
__device__ inline int block_min_int(int v, int *sh) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    int r = sh[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) r = min(r, sh[k]);
    __syncthreads();
    return r;
}

// k_min / k_max of one problem from c_eff on the grid rows (whole block): the WKB cut-off
// for ground-to-ground propagation (ESSModeSolver.cpp:561-607, WModeSolver.cpp:574-609),
// k_min = omega/c_eff(Hgt[top+1]), top = Nz - Nz/10 (ESSModeSolver.cpp:609-614,
// WModeSolver.cpp:611-620), and the wavenumber filter (ESSModeSolver.cpp:366-369,
// WModeSolver.cpp:356-360).  wkb_fallback: drop the cut-off when it falls below k_min, as
// ESSModeSolver.cpp:616-621 does (WModeSolver has no such rule).  dz: the spacing of the
// WKB integral.
//
// The WKB stop test `integral >= 10` decides k_max to one step dkk, so it must come out as
// the reference's left-to-right sum would: every term is formed with the reference's
// operations (no fused multiply-add anywhere in this function), the block sums them in
// parallel, and when that sum lies within the bound of both summation errors of 10 the
// first lane forms the sequential sum itself and that decides.
__device__ inline void interval_bounds(const ProbDev &pr, int Nz, double dz, const double *ce, double *shd,
                                       int *shi, bool wkb_fallback, double &k_min_o, double &k_max_o,
                                       int &used_wkb_o) {
#pragma clang fp contract(off)
    const double om = pr.omega;
    double m = DBL_MAX;
    for (int i = threadIdx.x; i < Nz; i += blockDim.x) m = fmin(m, ce[i]);
    const double ceffmin = fmin(block_min(m, shd), ce[0]);
    double k_max;
    int used_wkb = 0;
    if ((pr.flags & NCPA_F_WKB) && fabs(pr.zs) < 1.0e-3 && fabs(pr.zr) < 1.0e-3) {
        used_wkb = 1;
        const double kmf = om / ceffmin;
        const double kg = om / ce[0];
        const double kmf2 = kmf * kmf, kg2 = kg * kg;
        const double dkk = (kmf2 - kg2) / 100.0;
        double kk = kg2;
        if (dkk > 1.0e-10) {
            // the reference steps ~100 times; the cap only bounds a degenerate dkk
            int steps = 0;
            for (kk = kg2; kk < kmf2 && steps < 1000; kk = kk + dkk, ++steps) {
                // the reference's `while (wkbTerm > dkk)` starts from wkbTerm = 1: with
                // dkk >= 1 its body never runs and the integral stays 0
                if (dkk >= 1.0) continue;
                // rows 0 .. first (inclusive): first = the lowest row with |kk - k_eff^2| <= dkk
                // (every row when none is, where the reference would read past the grid)
                int first = Nz - 1;
                for (int i = threadIdx.x; i < Nz; i += blockDim.x) {
                    const double ke = om / ce[i];
                    if (fabs(kk - ke * ke) <= dkk) { first = min(first, i); break; }
                }
                first = block_min_int(first, shi);
                double s = 0.0;
                for (int i = threadIdx.x; i <= first; i += blockDim.x) {
                    const double ke = om / ce[i];
                    s += dz * sqrt(fabs(kk - ke * ke));
                }
                const double integral = block_sum(s, shd);
                // both sums of n = first + 1 non-negative terms lie within (n - 1) u of the
                // exact sum (u = 2^-53); margin: twice that, with room
                const double margin = 2.5e-16 * double(first + 2) * fmax(integral, 10.0);
                int stop = integral >= 10.0;
                if (fabs(integral - 10.0) <= margin) {   // same branch for the whole block
                    if (threadIdx.x == 0) {
                        double q = 0.0;
                        for (int i = 0; i <= first; ++i) {
                            const double ke = om / ce[i];
                            q = q + dz * sqrt(fabs(kk - ke * ke));
                        }
                        shi[0] = q >= 10.0;
                    }
                    __syncthreads();
                    stop = shi[0];
                    __syncthreads();
                }
                if (stop) break;
            }
        }
        k_max = sqrt(kk);
    } else {
        k_max = om / ceffmin;
    }
    const int top = Nz - Nz / 10;
    double k_min = om / ce[min(top + 1, Nz - 1)];   // Nz < 20 would read past the grid
    if (wkb_fallback && used_wkb && k_max < k_min) k_max = om / ceffmin;
    if (pr.flags & NCPA_F_WVNUM_FILTER) {
        k_min = om / pr.c_max;   // 2*PI*freq/c_max
        k_max = om / pr.c_min;
    }
    k_min_o = k_min;
    k_max_o = k_max;
    used_wkb_o = used_wkb;
}

// ------------------------------------------------------------------------------------
// eigenvalues + sensitivities: one 64-lane block per (problem, 64 consecutive indices)
// ------------------------------------------------------------------------------------
#define NCPA_MAXIT 200

// One lane per eigenvalue index j (ascending), 64 consecutive indices of one problem per
// block.  Phase 0 narrows every bracket by multisection (see below); phase 1 bisects with count-only sweeps until [lo, hi) holds only eigenvalue j;
// phase 2 runs Newton on det(T - s) (count kept as a safeguard) until the step is below
// the working-precision tolerance; phase 3 is one sweep at the converged eigenvalue that
// yields the absorption and the adjugate eigenvector products.  iters_out: sweeps used,
// negative when the bracket collapsed before Newton converged (cluster at working precision).
//
// The search is written once for any pencil whose count of eigenvalues below the shift
// is non-decreasing in the shift: `pen` supplies count(ring, s), newton(ring, s) (count,
// det and d det/ds), narrow(s_min) (told after each multisection pass that no later shift
// lies below s_min), rows() (rows a sweep covers now) and finish(...) (phase 3 and the
// outputs).  EssPencil is dz^2 T - s;
// the wide-angle pencil (wide.h) is dz^2 T(k) with the shift k inside the diagonal.
template <int RING>
struct ModesSharedN {
    double ms_s[64];
    int ms_c[64];
    double hb_lo[64], hb_hi[64];   // brackets of the unconverged eigenvalues
    int hb_cl[64];                 // the count at each bracket's lower end
    int h_list[64];                // their lanes, in lane order
    double ring[RING];             // the staged arrays (RING doubles)
};
using ModesShared = ModesSharedN<2 * NCPA_RING>;

// Next shift of a lane whose bracket [lo, hi) holds the eigenvalues cl .. ch-1 (j among
// them): the midpoint when j is alone, else the (j - cl + 1)-th of ch - cl evenly spaced
// interior points, so the lanes that share a bracket cut it into ch - cl + 1 parts between
// them instead of all testing its midpoint.
__device__ __forceinline__ double next_shift(double lo, double hi, int cl, int ch, int j) {
    const int k = ch - cl;
    return k > 1 ? lo + (double)(j - cl + 1) * ((hi - lo) / (double)(k + 1)) : 0.5 * (lo + hi);
}

template <int MS, class Pen, class Sh>
__device__ __forceinline__ void modes_item(const ModeProb &P, Pen &pen, const ModeWork wk, Sh &sh_,
                                           const ModesOut &out) {
    double *ms_s = sh_.ms_s, *hb_lo = sh_.hb_lo, *hb_hi = sh_.hb_hi, *s_ring = sh_.ring;
    int *ms_c = sh_.ms_c, *h_list = sh_.h_list, *hb_cl = sh_.hb_cl;
    const int lane = (int)threadIdx.x;
    const int j = wk.j0 + lane;
    const bool active = j < P.n_hi;
    double lo = P.lo, hi = P.hi;
    int cl = P.n_lo, ch = P.n_hi;
    // phase 0: multisection.  The 64 lanes evaluate the count at 64 distinct shifts spread
    // over the bracket of the block's eigenvalues; every lane then takes the tightest
    // bracket of its own eigenvalue from the 64 counts.  Two passes isolate most eigenvalues,
    // work that bisection spends ~log2(n) sweeps on per lane.
    const int last = min(P.n_hi - 1, wk.j0 + 63) - wk.j0;   // last active lane
    double blo = lo, bhi = hi;
    int rows_used = 0;   // rows of the sweeps this lane's search used (until it converged)
    for (int pass = 0; pass < MS; ++pass) {
        const double sh = blo + (double)(lane + 1) * ((bhi - blo) * (1.0 / 65.0));
        const int c = pen.count(s_ring, sh);
        rows_used += pen.rows();
        ms_s[lane] = sh;
        ms_c[lane] = c;
        __syncthreads();
        for (int l = 0; l < 64; ++l) {           // counts are non-decreasing in the shift
            const int cc = ms_c[l];
            const double ss = ms_s[l];
            if (cc <= j) {
                if (ss > lo) { lo = ss; cl = cc; }
            } else if (ss < hi) {
                hi = ss;
                ch = cc;
            }
        }
        __syncthreads();
        blo = __shfl(lo, 0);
        bhi = __shfl(hi, last);
        // every later shift of the block lies above lane 0's lower bound, to within the
        // tolerance of a Newton step accepted just outside a bracket: the pencil may drop
        // what no such shift needs (narrow; after every pass, as the bound rises)
        pen.narrow(blo - 16.0 * P.tol);
    }
    double sig = next_shift(lo, hi, cl, ch, j);
    bool done = !active;
    int my_iters = MS;
    bool collapsed = false;
    // phases 1-2.  Lanes whose eigenvalue has converged (and lanes past the problem's last
    // index) do not idle: each iteration they are dealt round-robin to the unconverged
    // eigenvalues of the block and evaluate extra shifts spread over those brackets; every
    // unconverged lane then tightens its bracket from all 64 (shift, count) pairs, so the
    // stragglers that set the block's duration converge several bits per sweep.
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    double prev_step = INFINITY;   // the last Newton step taken (INFINITY after a bisection)
    for (int it = 0; it < NCPA_MAXIT; ++it) {
        const unsigned long long um = __ballot(!done);
        if (um == 0ull) break;
        const int nu = __popcll(um);
        if (!done) {
            hb_lo[lane] = lo;
            hb_hi[lane] = hi;
            hb_cl[lane] = cl;
            h_list[__popcll(um & below)] = lane;
        }
        __syncthreads();
        double sh = sig;
        int sh_cl = cl;   // count at the lower end of the bracket that holds sh
        if (done) {
            const int h = __popcll(~um & below);        // rank among the helpers
            const int idx = h % nu, kk = h / nu;
            const int nh = (64 - nu - idx + nu - 1) / nu;  // helpers dealt to eigenvalue idx
            const int u = h_list[idx];
            const double a = hb_lo[u], b = hb_hi[u];
            sh = a + (double)(kk + 1) * ((b - a) / (double)(nh + 1));
            sh_cl = hb_cl[u];
        }
        const bool newton = !done && (ch - cl == 1);
        int cnt;
        double p0 = 0.0, q0 = 0.0;
        bool have_det = true;   // (wave-uniform) every lane evaluated det and d det/ds at its shift
        if (Pen::sign_counts && __all(done || ch - cl == 1)) {
            // every shift of the wave lies strictly inside a bracket that holds one eigenvalue
            // (its own, or the one it helps with): the count there is cl or cl + 1, told apart
            // by the sign of det, and the sweep needs no sign history
            const NewtonOut o = pen.newton_nocount(s_ring, sh);
            cnt = count_from_sign(o.p0, sh_cl);
            p0 = o.p0;
            q0 = o.q0;
        } else if (__any(newton)) {
            const NewtonOut o = pen.newton(s_ring, sh);
            cnt = o.cnt;
            p0 = o.p0;
            q0 = o.q0;
        } else {
            cnt = pen.count(s_ring, sh);
            have_det = false;
        }
        if (!done) rows_used += pen.rows();
        ms_s[lane] = sh;
        ms_c[lane] = cnt;
        // det and its derivative at every shift of the sweep (the helpers' bracket arrays are
        // free again: every helper has read its bracket before the sweep)
        double *ms_p = hb_lo, *ms_q = hb_hi;
        if (have_det) {
            ms_p[lane] = p0;
            ms_q[lane] = q0;
        }
        __syncthreads();
        if (!done) {
            int lo_l = -1, hi_l = -1;                // lanes whose shifts are the new bracket ends
            for (int l = 0; l < 64; ++l) {           // counts are non-decreasing in the shift
                const int cc = ms_c[l];
                const double ss = ms_s[l];
                if (cc <= j) {
                    if (ss > lo) { lo = ss; cl = cc; lo_l = l; }
                } else if (ss < hi) {
                    hi = ss;
                    ch = cc;
                    hi_l = l;
                }
            }
            sig = next_shift(lo, hi, cl, ch, j);
            double step = INFINITY;
            // Newton from the best of three points: the lane's own shift and the two shifts (the
            // helpers' or other lanes') that now end its bracket, all with det known; near a
            // close neighbour the own step converges only linearly from far off, while a bracket
            // end the helpers put next to the root starts in the quadratic regime.  The smallest
            // step wins (the point nearest the root).
            double dlt = INFINITY, from = sh;
            bool own = false;
            if (newton && have_det) {
                if (q0 != 0.0 && isfinite(p0) && isfinite(q0)) {
                    dlt = -p0 / q0;
                    own = true;
                }
                if (ch - cl == 1) {
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int l = e ? hi_l : lo_l;
                        if (l < 0 || l == lane) continue;
                        const double pe = ms_p[l], qe = ms_q[l];
                        if (qe == 0.0 || !isfinite(pe) || !isfinite(qe)) continue;
                        const double de = -pe / qe;
                        const double ce = ms_s[l] + de;
                        // strictly inside the bracket, whatever the step: a bracket end may be
                        // a neighbour lane's converged shift, whose tiny step points at the
                        // neighbouring eigenvalue, just outside
                        if (fabs(de) < fabs(dlt) && ce > lo && ce < hi) {
                            dlt = de;
                            from = ms_s[l];
                            own = false;
                        }
                    }
                }
            }
            if (newton && isfinite(dlt)) {
                const double cand = from + dlt;
                const double adl = fabs(dlt);
                // Converged when the step is at the tolerance, or (pencils with
                // Pen::early_accept) when two successive Newton steps show the quadratic
                // regime (ratio below 1e-3) and predict the next step, |dlt|^3 / |dlt_prev|^2,
                // a thousandfold below the tolerance: the sweep that would only confirm it is
                // skipped.  The prediction takes the previous step as asymptotic, which it is
                // not when that step came from far off (the first Newton step from a wide
                // bracket): accepted at the tolerance itself, such a lane stopped 300x the
                // tolerance short; the factor covers it.  det T(k) of the wide-angle pencil has
                // other zero branches close enough to break the prediction, so it waits for the
                // step itself.
                // (the prediction needs the step from the lane's own shift, the point the
                // previous step led to)
                const double rat = adl / prev_step;   // 0 after a bisection: no estimate yet
                if (adl <= P.tol ||
                    (Pen::early_accept && own && prev_step < INFINITY && rat <= 1e-3 && 1e3 * adl * rat * rat <= P.tol)) {
                    sig = cand;
                    done = true;
                    my_iters = MS + it + 1;
                } else if (cand > lo && cand < hi) {
                    sig = cand;
                    step = adl;
                }
            }
            prev_step = step;
            if (!done && (hi - lo) <= P.tol) {
                sig = 0.5 * (lo + hi);
                done = true;
                collapsed = (ch - cl) != 1;
                my_iters = -(MS + it + 1);
            }
        }
        __syncthreads();
    }
    if (!done) my_iters = -NCPA_MAXIT - 1;
    pen.finish(P, s_ring, sig, collapsed, ch - cl, j, active, my_iters, rows_used + pen.rows(), out);
}

// ------------------------------------------------------------------------------------
// eigenvectors at a strided set of rows (plus the source row)
// ------------------------------------------------------------------------------------
// Slots 0..nrec-1 are rows row0 + k*stride, slot nrec the source row.  For each mode (lane)
// k_rows records theta_{i-1} (forward continuant from the ground) and phi_{i+1} (backward
// continuant from the top) at every recorded row i, with their power-of-two exponents, and
// q_0 = d det/ds at the eigenvalue.  k_rows_combine then forms
//     u_i^2 = -theta_{i-1} phi_{i+1} / q_0                 (adjugate diagonal)
// picks the row of largest |u| as reference r and writes
//     u_i = [-(-b)^d theta_{lo-1} phi_{hi+1} / q_0] / u_r,  (lo, hi) = sorted(i, r)
// so every factor comes from the direction in which its recurrence is stable.  Signs are
// fixed so that u(source row) >= 0, the convention psi(zs) >= 0 of the mode products.
struct RowsArgs {
    int row0, stride, nrec;
    size_t ld;      // scratch / output leading dimension (total modes)
};

// (the pencil supplies the shift of a mode, the row values at that shift and the factor
// by which q_0 is scaled: the wide-angle vectors are normalised with (1 + k^2) q_0.)
// The sweeps keep their row loops free of branches: every row's continuant goes to a
// 16-row LDS buffer (one column per lane) and the recorded rows of a chunk are copied out
// once the chunk is done, before its rescale.
// The two sweeps are independent and run in separate blocks (`backward` false: theta, true:
// phi and q): each wave is one dependent chain over the rows, so two blocks per work item
// halve the kernel's length and fill twice as many SIMDs.
template <class Pen>
__device__ __forceinline__ void rows_item(const ModeProb &P, const Pen &pen, const ModeWork wk,
                                          const double *__restrict__ k2, RowsArgs ra, double *__restrict__ TH,
                                          int *__restrict__ THE, double *__restrict__ PH, int *__restrict__ PHE,
                                          double *__restrict__ Q0, int *__restrict__ QEX, double *s_ring,
                                          double (*rec)[64], bool backward) {
    const int lane = (int)threadIdx.x;
    const int j = wk.j0 + lane;
    const bool active = j < P.n_hi;
    const size_t o = (size_t)P.out_off + (size_t)(active ? (j - P.n_lo) : 0);
    const double s = active ? pen.shift(k2[o]) : 0.5 * (P.lo + P.hi);   // the pencil's shift
    // rows n_eff .. n-1 lie in the evanescent top of every mode of the block (lane 0 holds its
    // lowest eigenvalue): their components are below e^-40 of those at the top of the
    // oscillatory part and are written as zeros; the sweeps start at row n_eff - 1
    const int n = pen.rows_kept(__shfl(s, 0));
    const size_t src = (size_t)ra.nrec * ra.ld + o;   // slot of the source row
    const int last_rec = ra.row0 + (ra.nrec - 1) * ra.stride;
    // forward: theta_k for k = 0 .. n-2, recorded at row k+1
    if (!backward) {
        double t1 = 1.0, t2 = 0.0;
        int ex = 0;
        int slot = 0, next = ra.row0;   // next recorded row and its slot
        if (next == 0) {
            if (active) { TH[o] = 1.0; THE[o] = 0; }
            slot = 1;
            next = ra.nrec > 1 ? ra.row0 + ra.stride : INT_MAX;
        }
        if (P.srow == 0 && active) { TH[src] = 1.0; THE[src] = 0; }
        for (int k = slot, row = next; k < ra.nrec; ++k, row += ra.stride)   // rows left out
            if (row >= n && active) {
                TH[(size_t)k * ra.ld + o] = 0.0;
                THE[(size_t)k * ra.ld + o] = 0;
            }
        pen.up(
            s_ring, n - 1, s,
            [&](double am, int k) {
                const double t0 = fma(am, t1, -t2);
                t2 = t1;
                t1 = t0;
                rec[k & (NCPA_CHUNK - 1)][lane] = t1;
            },
            [&](int lo, int m) {
                // rows lo+1 .. lo+m now hold theta_{row-1}
                for (; next <= lo + m; ++slot, next = slot < ra.nrec ? next + ra.stride : INT_MAX)
                    if (active) {
                        TH[(size_t)slot * ra.ld + o] = rec[(next - 1) & (NCPA_CHUNK - 1)][lane];
                        THE[(size_t)slot * ra.ld + o] = ex;
                    }
                if (P.srow >= lo + 1 && P.srow <= lo + m && active) {
                    TH[src] = rec[(P.srow - 1) & (NCPA_CHUNK - 1)][lane];
                    THE[src] = ex;
                }
                const int kk = pow2_exponent(t1);
                t1 = ldexp(t1, -kk);
                t2 = ldexp(t2, -kk);
                ex += kk;
            });
    }
    // backward: phi_i for i = n-1 .. 0, recorded at row i-1; q for the normalisation
    else {
        double p1 = 1.0, p2 = 0.0, q1 = 0.0, q2 = 0.0;
        int ex = 0;
        int slot = ra.nrec - 1, next = last_rec;
        for (; slot >= 0 && next >= n; --slot, next -= ra.stride)   // rows left out
            if (active) {
                PH[(size_t)slot * ra.ld + o] = 0.0;
                PHE[(size_t)slot * ra.ld + o] = 0;
            }
        if (slot < 0) next = INT_MIN;
        if (next == n - 1) {
            if (active) { PH[(size_t)slot * ra.ld + o] = 1.0; PHE[(size_t)slot * ra.ld + o] = 0; }
            --slot;
            next = slot >= 0 ? next - ra.stride : INT_MIN;
        }
        if (P.srow == n - 1 && active) { PH[src] = 1.0; PHE[src] = 0; }
        pen.down(
            s_ring, n, s,
            [&](double am, int i) {
                const double p0 = fma(am, p1, -p2);
                const double q0 = fma(am, q1, -(q2 + p1));
                p2 = p1;
                p1 = p0;
                q2 = q1;
                q1 = q0;
                rec[i & (NCPA_CHUNK - 1)][lane] = p1;
            },
            [&](int lo, int m) {
                // rows lo-1 .. lo+m-2 now hold phi_{row+1}
                for (; next >= lo - 1; --slot, next = slot >= 0 ? next - ra.stride : INT_MIN)
                    if (active) {
                        PH[(size_t)slot * ra.ld + o] = rec[(next + 1) & (NCPA_CHUNK - 1)][lane];
                        PHE[(size_t)slot * ra.ld + o] = ex;
                    }
                if (P.srow >= lo - 1 && P.srow <= lo + m - 2 && active) {
                    PH[src] = rec[(P.srow + 1) & (NCPA_CHUNK - 1)][lane];
                    PHE[src] = ex;
                }
                const int k = pow2_exponent(p1);
                p1 = ldexp(p1, -k);
                p2 = ldexp(p2, -k);
                q1 = ldexp(q1, -k);
                q2 = ldexp(q2, -k);
                ex += k;
            });
        if (active) {
            Q0[o] = q1 * pen.qscale(s);
            QEX[o] = ex;
        }
    }
}

// rows_item with the recorded rows stored as the sweep passes them (no record buffer, no LDS):
// the recorded rows and the source row are wave-uniform, so each row's test is a scalar
// compare and a branch taken only at the rows recorded; with the diagonal by scalar loads the
// kernel touches no LDS at all.  The values and exponents stored are those of rows_item.
template <class Pen>
__device__ __forceinline__ void rows_item_direct(const ModeProb &P, const Pen &pen, const ModeWork wk,
                                                 const double *__restrict__ k2, RowsArgs ra,
                                                 double *__restrict__ TH, int *__restrict__ THE,
                                                 double *__restrict__ PH, int *__restrict__ PHE,
                                                 double *__restrict__ Q0, int *__restrict__ QEX, bool backward) {
    const int lane = (int)threadIdx.x;
    const int j = wk.j0 + lane;
    const bool active = j < P.n_hi;
    const size_t o = (size_t)P.out_off + (size_t)(active ? (j - P.n_lo) : 0);
    const double s = active ? pen.shift(k2[o]) : 0.5 * (P.lo + P.hi);
    const int n = pen.rows_kept(__shfl(s, 0));
    const size_t src = (size_t)ra.nrec * ra.ld + o;
    const int last_rec = ra.row0 + (ra.nrec - 1) * ra.stride;
    if (!backward) {
        // forward: theta_k for k = 0 .. n-2; row k+1 records theta_k
        double t1 = 1.0, t2 = 0.0;
        int ex = 0;
        int slot = 0, next = ra.row0;
        if (next == 0) {
            if (active) { TH[o] = 1.0; THE[o] = 0; }
            slot = 1;
            next = ra.nrec > 1 ? ra.row0 + ra.stride : INT_MAX;
        }
        if (P.srow == 0 && active) { TH[src] = 1.0; THE[src] = 0; }
        for (int k = slot, row = next; k < ra.nrec; ++k, row += ra.stride)   // rows left out
            if (row >= n && active) {
                TH[(size_t)k * ra.ld + o] = 0.0;
                THE[(size_t)k * ra.ld + o] = 0;
            }
        if (next >= n) next = INT_MAX;
        int srow = P.srow >= 1 && P.srow <= n - 1 ? P.srow : INT_MAX;
        int ev = min(next, srow);   // the next row that stores anything: one test per row
        pen.up(
            nullptr, n - 1, s,
            [&](double am, int k) {
                const double t0 = fma(am, t1, -t2);
                t2 = t1;
                t1 = t0;
                if (k + 1 == ev) {
                    if (ev == next) {
                        if (active) {
                            TH[(size_t)slot * ra.ld + o] = t1;
                            THE[(size_t)slot * ra.ld + o] = ex;
                        }
                        ++slot;
                        next = (slot < ra.nrec && next + ra.stride < n) ? next + ra.stride : INT_MAX;
                    }
                    if (ev == srow) {
                        if (active) {
                            TH[src] = t1;
                            THE[src] = ex;
                        }
                        srow = INT_MAX;
                    }
                    ev = min(next, srow);
                }
            },
            [&](int, int) {
                const int kk = pow2_exponent(t1);
                t1 = ldexp(t1, -kk);
                t2 = ldexp(t2, -kk);
                ex += kk;
            });
    } else {
        // backward: phi_i for i = n-1 .. 0; row i-1 records phi_i; q for the normalisation
        double p1 = 1.0, p2 = 0.0, q1 = 0.0, q2 = 0.0;
        int ex = 0;
        int slot = ra.nrec - 1, next = last_rec;
        for (; slot >= 0 && next >= n; --slot, next -= ra.stride)   // rows left out
            if (active) {
                PH[(size_t)slot * ra.ld + o] = 0.0;
                PHE[(size_t)slot * ra.ld + o] = 0;
            }
        if (slot < 0) next = INT_MIN;
        if (next == n - 1) {
            if (active) { PH[(size_t)slot * ra.ld + o] = 1.0; PHE[(size_t)slot * ra.ld + o] = 0; }
            --slot;
            next = slot >= 0 ? next - ra.stride : INT_MIN;
        }
        if (P.srow == n - 1 && active) { PH[src] = 1.0; PHE[src] = 0; }
        int srow = P.srow >= 0 && P.srow <= n - 2 ? P.srow : INT_MIN;
        int ev = max(next, srow);   // the next row (going down) that stores anything
        pen.down(
            nullptr, n, s,
            [&](double am, int i) {
                const double p0 = fma(am, p1, -p2);
                const double q0 = fma(am, q1, -(q2 + p1));
                p2 = p1;
                p1 = p0;
                q2 = q1;
                q1 = q0;
                if (i - 1 == ev) {
                    if (ev == next) {
                        if (active) {
                            PH[(size_t)slot * ra.ld + o] = p1;
                            PHE[(size_t)slot * ra.ld + o] = ex;
                        }
                        --slot;
                        next = slot >= 0 ? next - ra.stride : INT_MIN;
                    }
                    if (ev == srow) {
                        if (active) {
                            PH[src] = p1;
                            PHE[src] = ex;
                        }
                        srow = INT_MIN;
                    }
                    ev = max(next, srow);
                }
            },
            [&](int, int) {
                const int k = pow2_exponent(p1);
                p1 = ldexp(p1, -k);
                p2 = ldexp(p2, -k);
                q1 = ldexp(q1, -k);
                q2 = ldexp(q2, -k);
                ex += k;
            });
        if (active) {
            Q0[o] = q1 * pen.qscale(s);
            QEX[o] = ex;
        }
    }
}

}  // namespace ncpa
