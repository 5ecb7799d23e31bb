// Types and launchers shared by the library's translation units (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include "ncpamodes.h"

struct ncpa_ctx;

namespace ncpa {

#define NCPA_XCDS 8   // MI355X accelerator complex dies (workgroups dispatched round-robin)

// ------------------------------------------------------------------------------------
// device-side problem record (built on the host per batch)
// ------------------------------------------------------------------------------------
struct ProbDev {
    const double *prof;   // packed profile nodes
    int n_nodes;
    int flags;
    double omega;
    double az_rad;
    double alpha_freq;
    double admittance;
    double zs, zr;
    double c_min, c_max;
    int user_alpha;
    int srow, rrow;
};

struct IntervalOut {
    double k_min, k_max, lo, hi;
    int n_lo, n_hi;
    int used_wkb;
    int pad;
};

// Per problem of the eigenvalue kernels.  a: the rows of dz^2 T (unit off-diagonal); lo,
// hi, tol and the shifts are in eigenvalues of dz^2 T; outputs are lambda = lambda' / dz^2.
struct ModeProb {
    const double *a;
    const double *w;
    double h2, inv_h2;   // dz^2, 1/dz^2
    double lo, hi;
    double tol;       // absolute convergence tolerance on lambda' (eigenvalue of dz^2 T)
    double inv_dz;
    int n;
    int srow, rrow;
    int n_lo, n_hi;
    int out_off;
};

struct ModeWork {
    int prob;   // -1: padding block (no work)
    int j0;
};

struct ModesOut {
    double *k2, *absn, *psis, *psir, *psisr;
    int *iters;
};

// Written on the device by k_plan_build for the persistent eigenvalue launch.
struct PlanMeta {
    int n_work;     // work-list length (multiple of NCPA_XCDS, padding items have prob = -1)
    int total;      // modes of the batch (counted also when not computed)
    int overflow;   // bit 0: a problem above NCPA_MAX_MODES; bit 1: batch above the plan's
                    // capacity.  Either way no mode is computed (n_work = 0).
    int max_nm;     // most modes of one problem
    int n_work_need;   // work-list length the batch needs (also when it exceeds the capacity)
    int queue[NCPA_XCDS];   // per-XCD work counters of the persistent launch (reset here)
};

// ess.hip.  d_nodes: 2 P max_nodes doubles of scratch (attenuation and c_eff at the nodes)
void launch_setup(hipStream_t st, const ProbDev *probs, int P, int max_nodes, const ncpa_grid_t &grid,
                  double *d_nodes, double *A, double *W, double *CE);
void launch_solve_front(hipStream_t st, const ProbDev *probs, int P, int max_nodes, const ncpa_grid_t &grid,
                        double *d_nodes, double *A, double *W, double *CE, IntervalOut *iv);
void launch_plan_build(hipStream_t st, const ProbDev *probs, const IntervalOut *iv, int P, int Nz, double dz,
                       const double *A, const double *W, int cap_modes, int cap_work, int n_sets, int n_freq,
                       ModeProb *mp, int *offsets, ModeWork *work, int *mbeg, int *mcnt, PlanMeta *meta);
// queue != nullptr: the persistent form (grid = resident blocks); else one block per item,
// n_work from the host or, when n_work_dev is given, from the device (grid = capacity)
void launch_k_modes(int ms, dim3 grid, hipStream_t st, const ModeProb *mp, const ModeWork *work, int n_work,
                    const int *n_work_dev, int *queue, const ModesOut &mo);
int build_problems(ncpa_ctx *ctx, const ncpa_grid_t *grid, const ncpa_problem_t *problems, int n_problems,
                   ProbDev *hp, int *max_nodes);

// broadband.hip
void bb_launch_kpert(hipStream_t st, int n_cap, const int *n_dev, const double *k2, const double *absn,
                     double *kre, double *kim);
void bb_launch_transfer(hipStream_t st, int n_bins, int n_freq, const int *beg, const int *cnt, const double *kre,
                        const double *kim, const double *psr, int zero_attn, const double *ranges, int n_ranges,
                        const double *fac, double *H);
void bb_launch_arg(hipStream_t st, int n_tr, int n_freq, int nfft, const double *H, const double *S,
                   const double *f, const double *t0, int n_ranges, double *arg);
int bb_prepare_fft(ncpa_ctx *ctx, int n, int batch, size_t *scratch_doubles);
int bb_launch_fft(ncpa_ctx *ctx, const double *in, double *out, int n, int batch, int sign, double *scratch);
void bb_launch_out(hipStream_t st, size_t n, const double *spec, double df, double *wave);

}  // namespace ncpa
