// Batched wide-angle, high-Mach normal modes (ncpaprop WMod / ModBB --method wmod) on
// MI355X (gfx950).
//
// The reference assembles the quadratic eigenvalue problem of WModeSolver::getModalTrace
// (src/libmodes/WModeSolver.cpp:514-629) and solves its 2N x 2N linearisation with SLEPc
// Krylov-Schur shift-invert (EigenEngine::doWideAngleCalculation, EigenEngine.cpp:174-395),
// then keeps k in [k_min, k_max] (WModeSolver::doSelect, :491-506) and perturbs
// (ModeSolver::doPerturb, ModeSolver.cpp:263-284).  Here the same eigenpairs come from
// Sturm counts in k on the symmetric tridiagonal T(k) (wide.h), with the search and the
// eigenvector-rows pass of the effective-sound-speed kernels (modes.h):
//   k_wsetup_grid   rows A, B, C of dz^2 T(k) = A + k (B + k C), absorption weights, c_eff
//   k_winterval     k_min / k_max and the counts n(k_min), n(k_max)
//   k_wmodes        one lane per eigenvalue: multisection, bisection, Newton in k, then the
//                   adjugate products and the absorption at the converged k
//   k_wrows         continuants for eigenvector rows (2-D fields, mode files)
// The results are an ncpa_modes_t like ncpa_ess_solve's: k^2, absorption and psi(zs) psi(zr)
// per mode, so the 1-D / 2-D / dispersion / broadband paths apply unchanged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"
#include "ctx.h"
#include "internal.h"
#include "modes.h"
#include "wide.h"

namespace ncpa {

struct NodeArr {
    const double *y;
    __device__ double operator()(int k) const { return y[k]; }
};
// _WC_ = |w| cos(dir - az) at a node (Atmosphere1D.cpp:546-577)
struct NodeWc {
    const double *ws, *wd;
    double az;
    __device__ double operator()(int k) const { return ws[k] * cos(wd[k] * PI / 180.0 - az); }
};

// Rows of one problem at Hgt[i] = z_min + i dz_h (the setParams grid, WModeSolver.cpp:150,
// 202-209), matrix spacing dz_m (WModeSolver.cpp:297):
//   kd = (omega/c0)^2, md = (wc/c0)^2 - 1, cd = -2 omega wc/c0^2 (:549-553), the ground term
//   1/((dz adm + 1) dz^2) added to kd[0] (:568-570), A = -2 + dz^2 kd, B = dz^2 cd,
//   C = dz^2 md; W = 2 (omega/c_T) alpha (ModeSolver.cpp:276-278); CE = c_eff (:554); the
//   factored rows P = dz omega/c0, Q = dz wc/c0 of the counts (wide.h).
// Block (x, p) takes rows [1024 x, 1024 x + 1024) of problem p; as k_setup_grid (ess.hip), the
// Steffen coefficients of the node intervals those rows fall in are formed once per block in
// LDS (c0, the wind component, c_eff, alpha, P, rho), so a row costs a search over the LDS
// abscissae and three fma per property; windows wider than NCPA_WSETUP_IV intervals evaluate
// every row from global memory.  Both paths form the same coefficients (steffen_coef).
#define NCPA_WSETUP_ROWS 1024
#define NCPA_WSETUP_IV 128
__global__ void __launch_bounds__(256) k_wsetup_grid(const ProbDev *__restrict__ probs, int Nz, double z_min,
                                                     double dz_h, double dz_m,
                                                     const double *__restrict__ alpha_nodes,
                                                     const double *__restrict__ ceff_nodes, int ldn,
                                                     double *__restrict__ A, double *__restrict__ BC, size_t plane,
                                                     double *__restrict__ W, double *__restrict__ CE,
                                                     double *__restrict__ PQ) {
    __shared__ SteffenCoef s_cf[6][NCPA_WSETUP_IV];
    __shared__ double s_z[NCPA_WSETUP_IV + 1];
    __shared__ int s_k[2];
    const int p = blockIdx.y;
    const int r0 = blockIdx.x * NCPA_WSETUP_ROWS, r1 = min(Nz, r0 + NCPA_WSETUP_ROWS);
    const ProbDev pr = probs[p];
    const int n = pr.n_nodes;
    const double *zn = pr.prof + PF_Z * n;
    const NodeArr c0n{pr.prof + PF_C0 * n}, cen{ceff_nodes + size_t(p) * ldn}, aln{alpha_nodes + size_t(p) * ldn},
        pn{pr.prof + PF_P * n}, rn{pr.prof + PF_RHO * n};
    const NodeWc wcn{pr.prof + PF_WS * n, pr.prof + PF_WD * n, pr.az_rad};
    // node value k of property f: c0, wc, c_eff, alpha, P, rho
    auto node = [&](int f, int k) {
        switch (f) {
            case 0: return c0n(k);
            case 1: return wcn(k);
            case 2: return cen(k);
            case 3: return aln(k);
            case 4: return pn(k);
            default: return rn(k);
        }
    };
    const double z0 = zn[0], z1 = zn[n - 1];
    if (threadIdx.x == 0) s_k[0] = bsearch_interval(zn, n, z_min + r0 * dz_h);
    if (threadIdx.x == 64) s_k[1] = bsearch_interval(zn, n, z_min + (r1 - 1) * dz_h);
    __syncthreads();
    const int k_lo = s_k[0], k_hi = s_k[1], ni = k_hi - k_lo + 1;
    const bool lds = ni <= NCPA_WSETUP_IV;   // block-uniform
    if (lds) {
        for (int t = threadIdx.x; t < 6 * ni; t += blockDim.x) {
            const int f = t / ni, k = k_lo + t % ni;
            s_cf[f][k - k_lo] = steffen_coef(PtrNodes{zn}, n, [&](int kk) { return node(f, kk); }, k);
        }
        for (int t = threadIdx.x; t <= ni; t += blockDim.x) s_z[t] = zn[min(k_lo + t, n - 1)];
        __syncthreads();
    }
    const double om = pr.omega;
    const double h2 = dz_m * dz_m;
    for (int i = r0 + threadIdx.x; i < r1; i += blockDim.x) {
        const double z = z_min + i * dz_h;
        int k;
        double dx = 0.0;
        if (lds) {
            // largest interval in [k_lo, k_hi] with z_k <= z (bsearch_interval's answer)
            int lo = 0, hi = ni;
            while (hi > lo + 1) {
                const int mid = (hi + lo) >> 1;
                if (s_z[mid] > z) hi = mid; else lo = mid;
            }
            k = lo;
            dx = z - s_z[lo];
        } else {
            k = bsearch_interval(zn, n, z);
        }
        // spline value: end values outside the nodes (AtmosphericProperty1D.cpp:116-126)
        auto at = [&](int f) {
            if (z > z1) return node(f, n - 1);
            if (z < z0) return node(f, 0);
            if (lds) return steffen_poly(s_cf[f][k], dx);
            return steffen_eval_zf(PtrNodes{zn}, n, [&](int kk) { return node(f, kk); }, z, k);
        };
        const double c0 = at(0);
        const double wc = at(1);
        const double ceff = at(2);
        double kd = (om / c0) * (om / c0);
        if (i == 0) kd = (1.0 / (dz_m * pr.admittance + 1)) / h2 + kd;
        const double md = (wc / c0) * (wc / c0) - 1;
        const double cd = -2 * om * (wc / (c0 * c0));
        const double alpha = at(3);
        const double Pz = at(4);
        const double Rz = at(5);
        const double cT = sqrt(1.4 * Pz / Rz);
        const size_t o = size_t(p) * Nz + i;
        A[o] = -2.0 + h2 * kd;
        BC[o] = h2 * cd;
        BC[plane + o] = h2 * md;
        PQ[o] = dz_m * (om / c0);   // the factored rows: P = dz omega/c0, Q = dz wc/c0
        PQ[plane + o] = dz_m * (wc / c0);
        W[o] = (om / cT) * alpha * 2.0;
        CE[o] = ceff;
    }
}

// The pencil of problem p: A and W rows from the ModeProb, B and C from the BC planes.
template <int T, bool PQ = false>
struct WidePencil {
    static constexpr bool early_accept = false;
    WideSweeps<T, PQ> s;

    int keep = 0;      // rows the final sweep records below (source / receiver + 2)
    int dropped = 0;   // evanescent rows left out at the top, one negative pivot each
    __device__ __forceinline__ int count(double *ring, double k) const { return s.count(ring, k) + dropped; }
    __device__ __forceinline__ NewtonOut newton(double *ring, double k) const {
        NewtonOut o = s.newton(ring, k);
        o.cnt += dropped;
        return o;
    }
    // The evanescent cut of the effective-sound-speed pencil with k as the shift: a_i(k) + 2 =
    // dz^2 (kd + k cd + k^2 md) is a concave parabola in k (md < 0 below Mach 1) whose larger
    // root is positive, so a row with a_i(k_min) <= -2 keeps a_i(k) <= -2 for every k >= k_min
    // and contributes one negative pivot at every later shift.  (The margin keeps rows whose
    // value sits within rounding of -2 in the sweep.)
    __device__ __forceinline__ void narrow(double k) {
        const int n0 = s.n;
        int n1;
        if constexpr (PQ) {
            const double K = fma(-s.h2 * k, k, -2.0);
            n1 = evanescent_rows_at(
                [&](int i) {
                    const double t = fma(-k, s.Q[i], s.P[i]);
                    return fma(t, t, K) + (i == 0 ? s.beta : 0.0);
                },
                n0, 80.0, keep, 1e-12);
        } else {
            n1 = evanescent_rows_at(
                [&](int i) { return fma(fma(s.C[i], k, s.B[i]), k, s.A[i]); }, n0, 80.0, keep, 1e-12);
        }
        dropped += n0 - n1;
        s.n = n1;
    }
    __device__ __forceinline__ int rows() const { return s.n; }
    // Newton sweeps without the sign history once every bracket isolates one branch crossing:
    // the branches of T(k) decrease strictly, so counts are monotone in k and the sign of
    // det T(k) is (-1)^count; each dropped row flips it once
    static constexpr bool sign_counts = true;
    __device__ __forceinline__ NewtonOut newton_nocount(double *ring, double k) const {
        NewtonOut o = s.template newton<false>(ring, k);
        if (dropped & 1) {
            o.p0 = -o.p0;
            o.q0 = -o.q0;
        }
        return o;
    }
    __device__ __forceinline__ int rows_kept(double) const { return s.n; }

    // phase 3 at the converged k: psi = u / sqrt(dz (1 + k^2)), absorption u^T W u / (1 + k^2)
    __device__ __forceinline__ void finish(const ModeProb &P, double *ring, double k, bool collapsed, int nclu,
                                           int j, bool active, int my_iters, int rows_used,
                                           const ModesOut &out) const {
        const FinalOut f = s.final_sweep(ring, k, P.srow, P.rrow);
        SVal th_s, th_r;
        s.theta(ring, k, P.srow, P.rrow, th_s, th_r);
        if (!active) return;
        const double qn = f.q0 * (1.0 + k * k);
        const double mult = collapsed ? (double)nclu : 1.0;
        const double us2 = adjugate_product(th_s, f.phi_s1, qn, f.ex, 0) / mult;
        const double ur2 = adjugate_product(th_r, f.phi_r1, qn, f.ex, 0) / mult;
        double usr;
        if (P.srow <= P.rrow)
            usr = adjugate_product(th_s, f.phi_r1, qn, f.ex, P.rrow - P.srow) / mult;
        else
            usr = adjugate_product(th_r, f.phi_s1, qn, f.ex, P.srow - P.rrow) / mult;
        const int o = P.out_off + (j - P.n_lo);
        out.k2[o] = k * k;
        out.absn[o] = (qn != 0.0) ? -f.e0 / qn : 0.0;
        const double ps = sqrt(fmax(us2, 0.0) * P.inv_dz);
        const double pr = sqrt(fmax(ur2, 0.0) * P.inv_dz);
        out.psis[o] = ps;
        out.psir[o] = (usr < 0.0) ? -pr : pr;
        out.psisr[o] = usr * P.inv_dz;
        out.iters[o] = my_iters;
        if (out.rows) out.rows[o] = rows_used;
    }

    // rows pass
    __device__ __forceinline__ double shift(double k2) const { return sqrt(k2); }
    __device__ __forceinline__ double qscale(double k) const { return 1.0 + k * k; }
    template <class Step, class ChunkEnd>
    __device__ __forceinline__ void up(double *ring, int m, double k, Step step, ChunkEnd ce) const {
        s.up(ring, m, k, step, ce);
    }
    template <class Step, class ChunkEnd>
    __device__ __forceinline__ void down(double *ring, int m, double k, Step step, ChunkEnd ce) const {
        s.down(ring, m, k, step, ce);
    }
};

template <int T, bool PQ = false>
__device__ __forceinline__ WideSweeps<T, PQ> wide_sweeps(const ModeProb &P, int prob, const double *BC,
                                                         size_t plane, const double *PQp = nullptr,
                                                         const double *beta = nullptr) {
    const size_t o = (size_t)prob * P.n;
    WideSweeps<T, PQ> w{P.a, BC + o, BC + plane + o, P.w, P.n};
    if constexpr (PQ) {
        w.P = PQp + o;
        w.Q = PQp + plane + o;
        w.h2 = P.h2;
        w.beta = beta[prob];
    }
    return w;
}

constexpr int WT_ONE = NCPA_WTILE_SMALL;   // geometry of the one-off sweeps (interval, counts, rows)

// k_min, k_max (no WKB fallback, WModeSolver.cpp:572-620) and n(k_min), n(k_max)
__global__ void __launch_bounds__(256) k_winterval(const ProbDev *__restrict__ probs, int Nz, double dz_m,
                                                   const double *__restrict__ A, const double *__restrict__ BC,
                                                   size_t plane, const double *__restrict__ CE,
                                                   const double *__restrict__ PQ, IntervalOut *__restrict__ out) {
    __shared__ double shd[8];
    __shared__ int shi[8];
    __shared__ double s_k[2];
    __shared__ double s_ring[3 * 2 * WT_ONE];
    const int p = blockIdx.x;
    const ProbDev pr = probs[p];
    double k_min, k_max;
    int used_wkb;
    interval_bounds(pr, Nz, dz_m, CE + size_t(p) * Nz, shd, shi, false, k_min, k_max, used_wkb);
    if (threadIdx.x == 0) { s_k[0] = k_min; s_k[1] = k_max; }
    __syncthreads();
    // the counts are exact only while every branch decreases: dz^2 dT/dk = diag(B + 2kC) < 0
    // on every row, checked at both ends of the interval (it is linear in k)
    double worst = -DBL_MAX;
    {
        const double *Bp = BC + size_t(p) * Nz, *Cp = BC + plane + size_t(p) * Nz;
        for (int i = threadIdx.x; i < Nz; i += blockDim.x)
            worst = fmax(worst, fmax(fma(2.0 * k_min, Cp[i], Bp[i]), fma(2.0 * k_max, Cp[i], Bp[i])));
    }
    worst = -block_min(-worst, shd);
    if (threadIdx.x < 64) {   // the whole first wave sweeps (lanes 0 and 1 keep their counts)
        // the factored rows, as the eigenvalue kernel counts with them
        const size_t o = size_t(p) * Nz;
        WideSweeps<WT_ONE, true> ws{A + o, BC + o, BC + plane + o, nullptr, Nz};
        ws.P = PQ + o;
        ws.Q = PQ + plane + o;
        ws.h2 = dz_m * dz_m;
        ws.beta = 1.0 / (dz_m * pr.admittance + 1);
        const int c = ws.count(s_ring, s_k[threadIdx.x ? 1 : 0]);
        if (threadIdx.x == 0) out[p].n_lo = c;
        if (threadIdx.x == 1) out[p].n_hi = c;
    }
    if (threadIdx.x == 0) {
        out[p].k_min = k_min;
        out[p].k_max = k_max;
        out[p].lo = k_min;   // shifts of the wide-angle search are wavenumbers
        out[p].hi = k_max;
        out[p].used_wkb = used_wkb;
        out[p].not_monotone = (worst >= 0.0 && k_max > k_min) ? 1 : 0;
    }
}

// counts at two wavenumbers (lane 0: lo, lane 1: hi) for one pencil
__global__ void __launch_bounds__(64) k_wcounts2(const double *__restrict__ A, const double *__restrict__ B, const double *__restrict__ C,
                           int n, double lo, double hi, int *__restrict__ out) {
    __shared__ double s_ring[3 * 2 * WT_ONE];
    const WideSweeps<WT_ONE> ws{A, B, C, nullptr, n};
    const int c = ws.count(s_ring, threadIdx.x ? hi : lo);   // whole wave
    if (threadIdx.x < 2) out[threadIdx.x] = c;
}

// one block per work item (as k_modes; the wide-angle path has no persistent form), staged in
// tiles of T rows
template <int MS, int T, bool PQ>
__global__ void __launch_bounds__(64, T == NCPA_WTILE_SMALL ? 4 : 1) k_wmodes(const ModeProb *__restrict__ mp, const ModeWork *__restrict__ work,
                                               const double *__restrict__ BC, size_t plane,
                                               const double *__restrict__ PQp, const double *__restrict__ beta,
                                               ModesOut out) {
    __shared__ ModesSharedN<(PQ ? 3 : 4) * 2 * T> sh;   // staged arrays of the final sweep: P, Q, W or A, B, C, W
    const ModeWork wk = work[blockIdx.x];
    if (wk.prob < 0) return;
    const ModeProb &P = mp[wk.prob];
    WidePencil<T, PQ> pen{wide_sweeps<T, PQ>(P, wk.prob, BC, plane, PQp, beta)};
    pen.keep = max(P.srow, P.rrow) + 2;
    modes_item<MS>(P, pen, wk, sh, out);
}

template <bool PQ>
__global__ void __launch_bounds__(64) k_wrows(const ModeProb *__restrict__ mp, const ModeWork *__restrict__ work,
                                              const double *__restrict__ BC, size_t plane,
                                              const double *__restrict__ PQp, const double *__restrict__ beta,
                                              const double *__restrict__ k2, RowsArgs ra, double *__restrict__ TH,
                                              int *__restrict__ THE, double *__restrict__ PH, int *__restrict__ PHE,
                                              double *__restrict__ Q0, int *__restrict__ QEX) {
    __shared__ double s_ring[3 * 2 * WT_ONE];
    __shared__ double rec[NCPA_CHUNK][64];
    const ModeWork wk = work[blockIdx.x >> 1];   // block 2w: forward sweep of item w, 2w + 1: backward
    if (wk.prob < 0) return;
    const ModeProb P = mp[wk.prob];
    const WidePencil<WT_ONE, PQ> pen{wide_sweeps<WT_ONE, PQ>(P, wk.prob, BC, plane, PQp, beta)};
    rows_item(P, pen, wk, k2, ra, TH, THE, PH, PHE, Q0, QEX, s_ring, rec, (blockIdx.x & 1) != 0);
}

// Batches of more work items than the GPU has SIMDs run many waves per SIMD, where the
// compact geometry's occupancy wins (broadband: -13 %); below, each wave's own instruction
// stream sets the time and the 256-row tiles win (1 Hz batch of 64 problems: -10 %).
#define NCPA_WIDE_COMPACT_ITEMS 4096

template <int T, bool PQ>
void launch_wmodes_t(int ms, dim3 grid, hipStream_t st, const ModeProb *mp, const ModeWork *work, const double *BC,
                     size_t plane, const double *PQp, const double *beta, const ModesOut &mo) {
    switch (ms) {
        case 1: hipLaunchKernelGGL((k_wmodes<1, T, PQ>), grid, dim3(64), 0, st, mp, work, BC, plane, PQp, beta, mo); break;
        case 3: hipLaunchKernelGGL((k_wmodes<3, T, PQ>), grid, dim3(64), 0, st, mp, work, BC, plane, PQp, beta, mo); break;
        default: hipLaunchKernelGGL((k_wmodes<2, T, PQ>), grid, dim3(64), 0, st, mp, work, BC, plane, PQp, beta, mo); break;
    }
}

// PQ (the factored rows of ncpa_wmod_solve) or nullptr (any kd, md, cd: ncpa_wmod_eigen)
void launch_k_wmodes(int ms, dim3 grid, hipStream_t st, const ModeProb *mp, const ModeWork *work,
                     const double *BC, size_t plane, const double *PQ, const double *beta, const ModesOut &mo) {
    const bool compact = grid.x >= NCPA_WIDE_COMPACT_ITEMS;
    if (PQ && compact)
        launch_wmodes_t<NCPA_WTILE_SMALL, true>(ms, grid, st, mp, work, BC, plane, PQ, beta, mo);
    else if (PQ)
        launch_wmodes_t<NCPA_WTILE_BIG, true>(ms, grid, st, mp, work, BC, plane, PQ, beta, mo);
    else if (compact)
        launch_wmodes_t<NCPA_WTILE_SMALL, false>(ms, grid, st, mp, work, BC, plane, nullptr, nullptr, mo);
    else
        launch_wmodes_t<NCPA_WTILE_BIG, false>(ms, grid, st, mp, work, BC, plane, nullptr, nullptr, mo);
}

void launch_k_wrows(dim3 grid, hipStream_t st, const ModeProb *mp, const ModeWork *work, const double *BC,
                    size_t plane, const double *PQ, const double *beta, const double *k2, const RowsArgs &ra,
                    double *TH, int *THE, double *PH, int *PHE, double *Q0, int *QEX) {
    if (PQ)
        hipLaunchKernelGGL(k_wrows<true>, grid, dim3(64), 0, st, mp, work, BC, plane, PQ, beta, k2, ra, TH, THE, PH,
                           PHE, Q0, QEX);
    else
        hipLaunchKernelGGL(k_wrows<false>, grid, dim3(64), 0, st, mp, work, BC, plane, nullptr, nullptr, k2, ra, TH,
                           THE, PH, PHE, Q0, QEX);
}

}  // namespace ncpa

using namespace ncpa;

extern "C" {

int ncpa_wmod_solve(ncpa_ctx_t *ctx, const ncpa_grid_t *grid, const ncpa_problem_t *problems, int n_problems,
                    ncpa_modes_t **out) {
    if (!ctx || !problems || n_problems <= 0 || !out)
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_wmod_solve: bad arguments");
    *out = nullptr;
    int rc = check_grid(ctx, grid);
    if (rc) return rc;
    if (!(grid->maxheight > grid->z_min))
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_wmod_solve: grid needs maxheight > z_min");
    NCPA_HIP(ctx, hipSetDevice(ctx->device));
    const int Nz = grid->Nz;
    const int P = n_problems;
    // WModeSolver::solve spaces the matrix rows by (maxheight - z_min)/Nz (WModeSolver.cpp:297)
    // while the profile samples keep the setParams spacing (maxheight - z_min)/(Nz - 1)
    ncpa_grid_t gm = *grid;
    gm.dz = (grid->maxheight - grid->z_min) / Nz;
    std::vector<ProbDev> hp(P);
    std::vector<int> srow(P), rrow(P);
    int max_nodes = 0;
    rc = build_problems(ctx, &gm, problems, P, hp.data(), &max_nodes);   // rows ceil(z / dz_m)
    if (rc) return rc;
    for (int p = 0; p < P; ++p) {
        srow[p] = hp[p].srow;
        rrow[p] = hp[p].rrow;
    }
    ncpa_modes *m = new ncpa_modes();
    m->ctx = ctx;
    m->P = P;
    m->wide = true;
    m->grid = gm;
    const size_t plane = size_t(P) * Nz;
    ProbDev *d_probs = pool_alloc<ProbDev>(ctx, P);
    double *d_nodes = pool_alloc<double>(ctx, 2 * size_t(P) * max_nodes);   // alpha and c_eff nodes
    m->d_A = pool_alloc<double>(ctx, plane);
    m->d_BC = pool_alloc<double>(ctx, 2 * plane);
    m->d_W = pool_alloc<double>(ctx, plane);
    m->d_CE = pool_alloc<double>(ctx, plane);
    m->d_PQ = pool_alloc<double>(ctx, 2 * plane);
    m->d_beta = pool_alloc<double>(ctx, P);
    IntervalOut *d_iv = pool_alloc<IntervalOut>(ctx, P);
    auto fail = [&](int code) {
        ctx->pool.put(d_probs);
        ctx->pool.put(d_nodes);
        ctx->pool.put(d_iv);
        release_modes(m);
        delete m;
        return code;
    };
    if (!d_probs || !d_nodes || !m->d_A || !m->d_BC || !m->d_W || !m->d_CE || !m->d_PQ || !m->d_beta || !d_iv)
        return fail(set_error(ctx, NCPA_ERR_ALLOC, "ncpa_wmod_solve: device allocation failed"));
    hipError_t e = hipMemcpyAsync(d_probs, hp.data(), sizeof(ProbDev) * P, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return fail(hip_error(ctx, e, "copy problems"));
    {
        // row 0's ground term of the factored rows: dz^2 / ((dz adm + 1) dz^2) (WModeSolver.cpp:568-570)
        std::vector<double> beta(P);
        for (int p = 0; p < P; ++p) beta[p] = 1.0 / (gm.dz * hp[p].admittance + 1);
        e = hipMemcpyAsync(m->d_beta, beta.data(), sizeof(double) * P, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);   // beta is a host temporary
        if (e != hipSuccess) return fail(hip_error(ctx, e, "copy ground terms"));
    }
    {
        TimedLaunch tl(ctx, "setup");
        launch_node_tables(ctx->stream, d_probs, P, max_nodes, d_nodes);
        hipLaunchKernelGGL(k_wsetup_grid, dim3((Nz + NCPA_WSETUP_ROWS - 1) / NCPA_WSETUP_ROWS, P), dim3(256), 0,
                           ctx->stream, d_probs, Nz,
                           grid->z_min, grid->dz, gm.dz, d_nodes, d_nodes + size_t(P) * max_nodes, max_nodes,
                           m->d_A, m->d_BC, plane, m->d_W, m->d_CE, m->d_PQ);
    }
    {
        TimedLaunch tl(ctx, "interval");
        hipLaunchKernelGGL(k_winterval, dim3(P), dim3(256), 0, ctx->stream, d_probs, Nz, gm.dz, m->d_A, m->d_BC,
                           plane, m->d_CE, m->d_PQ, d_iv);
    }
    if ((e = hipGetLastError()) != hipSuccess) return fail(hip_error(ctx, e, "wide-angle setup kernels"));
    m->iv.resize(P);
    e = hipMemcpyAsync(m->iv.data(), d_iv, sizeof(IntervalOut) * P, hipMemcpyDeviceToHost, ctx->stream);
    if (e != hipSuccess) return fail(hip_error(ctx, e, "copy intervals"));
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return fail(hip_error(ctx, e, "interval sync"));
    for (int p = 0; p < P; ++p)
        if (m->iv[p].not_monotone)
            return fail(set_error(ctx, NCPA_ERR_ARG,
                                  "ncpa_wmod_solve: problem " + std::to_string(p) +
                                      ": the wind is too strong for the wide-angle counts (omega w + k (c0^2 - w^2) "
                                      "not positive on every row for k in [k_min, k_max])"));
    rc = launch_modes(ctx, m, srow, rrow);
    ctx->pool.put(d_nodes);
    ctx->pool.put(d_iv);
    ctx->pool.put(d_probs);
    d_nodes = nullptr;
    d_iv = nullptr;
    d_probs = nullptr;
    if (rc) return fail(rc);
    *out = m;
    return NCPA_OK;
}

int ncpa_wmod_eigen(ncpa_ctx_t *ctx, const double *kd, const double *md, const double *cd, int Nz, double dz,
                    double tol, double k_min, double k_max, int max_modes, int *nconv, double *kH, double *v,
                    int ldv) {
    (void)tol;
    if (!ctx || !kd || !md || !cd || Nz < 3 || !(dz > 0.0) || !nconv || !kH || !v || max_modes <= 0 || ldv < 1 ||
        !(k_min > 0.0) || !(k_max >= k_min))
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_wmod_eigen: bad arguments");
    NCPA_HIP(ctx, hipSetDevice(ctx->device));
    *nconv = 0;
    // dz^2 (D + k C + k^2 M) of EigenEngine::doWideAngleCalculation (EigenEngine.cpp:257-302)
    const double h2 = dz * dz;
    std::vector<double> a(Nz), bc(2 * (size_t)Nz);
    for (int i = 0; i < Nz; ++i) {
        a[i] = -2.0 + h2 * kd[i];
        bc[i] = h2 * cd[i];
        bc[Nz + i] = h2 * md[i];
    }
    ncpa_modes *m = new ncpa_modes();
    m->ctx = ctx;
    m->P = 1;
    m->wide = true;
    m->grid.Nz = Nz;
    m->grid.dz = dz;
    m->grid.z_min = 0.0;
    m->grid.maxheight = dz * Nz;
    m->d_A = pool_alloc<double>(ctx, Nz);
    m->d_BC = pool_alloc<double>(ctx, 2 * (size_t)Nz);
    m->d_W = pool_alloc<double>(ctx, Nz);
    int *d_cnt = pool_alloc<int>(ctx, 2);
    double *d_psi = nullptr;
    auto done = [&](int code) {
        ctx->pool.put(d_cnt);
        ctx->pool.put(d_psi);
        release_modes(m);
        delete m;
        return code;
    };
    if (!m->d_A || !m->d_BC || !m->d_W || !d_cnt)
        return done(set_error(ctx, NCPA_ERR_ALLOC, "ncpa_wmod_eigen: allocation failed"));
    hipError_t e;
    if ((e = hipMemcpyAsync(m->d_A, a.data(), sizeof(double) * Nz, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(m->d_BC, bc.data(), sizeof(double) * bc.size(), hipMemcpyHostToDevice, ctx->stream)) !=
            hipSuccess ||
        (e = hipMemsetAsync(m->d_W, 0, sizeof(double) * Nz, ctx->stream)) != hipSuccess)
        return done(hip_error(ctx, e, "ncpa_wmod_eigen upload"));
    hipLaunchKernelGGL(k_wcounts2, dim3(1), dim3(64), 0, ctx->stream, m->d_A, m->d_BC, m->d_BC + Nz, Nz, k_min,
                       k_max, d_cnt);
    int cnt[2] = {0, 0};
    if ((e = hipMemcpyAsync(cnt, d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
        return done(hip_error(ctx, e, "ncpa_wmod_eigen counts"));
    const int nm = std::max(0, cnt[1] - cnt[0]);
    if (nm > max_modes || nm > NCPA_MAX_MODES)
        return done(set_error(ctx, NCPA_ERR_MAX_MODES,
                              "ncpa_wmod_eigen: " + std::to_string(nm) + " eigenpairs exceed max_modes"));
    if (ldv < nm) return done(set_error(ctx, NCPA_ERR_ARG, "ncpa_wmod_eigen: ldv smaller than the number of modes"));
    for (int i = 0; i < Nz; ++i)   // dT/dk = diag(cd + 2 k md) must be negative on the interval
        if (!(cd[i] + 2.0 * k_min * md[i] < 0.0 && cd[i] + 2.0 * k_max * md[i] < 0.0))
            return done(set_error(ctx, NCPA_ERR_ARG, "ncpa_wmod_eigen: cd + 2 k md is not negative on row " +
                                                         std::to_string(i) + " for k in [k_min, k_max]"));
    m->iv.resize(1);
    m->iv[0] = IntervalOut{k_min, k_max, k_min, k_max, cnt[0], cnt[1], 0, 0};
    int rc = launch_modes(ctx, m, std::vector<int>{0}, std::vector<int>{0});
    if (rc) return done(rc);
    if (nm > 0) {
        d_psi = pool_alloc<double>(ctx, (size_t)Nz * nm);
        if (!d_psi) return done(set_error(ctx, NCPA_ERR_ALLOC, "ncpa_wmod_eigen: allocation failed"));
        rc = launch_rows(ctx, m, 0, 1, Nz, d_psi);
        if (rc) return done(rc);
        std::vector<double> k2(nm), psi((size_t)Nz * nm);
        if ((e = hipMemcpyAsync(k2.data(), m->d_k2, sizeof(double) * nm, hipMemcpyDeviceToHost, ctx->stream)) !=
                hipSuccess ||
            (e = hipMemcpyAsync(psi.data(), d_psi, sizeof(double) * psi.size(), hipMemcpyDeviceToHost, ctx->stream)) !=
                hipSuccess ||
            (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
            return done(hip_error(ctx, e, "ncpa_wmod_eigen download"));
        for (int i = 0; i < nm; ++i) kH[i] = std::sqrt(k2[i]);
        for (int j = 0; j < Nz; ++j) std::memcpy(v + (size_t)j * ldv, psi.data() + (size_t)j * nm, sizeof(double) * nm);
    }
    *nconv = nm;
    return done(NCPA_OK);
}

}  // extern "C"
