// Types and launchers shared by the library's translation units (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

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
    int not_monotone;   // wide-angle: dT/dk is not negative definite on [k_min, k_max]
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
    int *rows;   // rows the sweeps of each mode's search covered (the evanescent top left out)
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

struct RowsArgs;

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
                    const int *n_work_dev, int *queue, const ModesOut &mo, bool smem = false);
int build_problems(ncpa_ctx *ctx, const ncpa_grid_t *grid, const ncpa_problem_t *problems, int n_problems,
                   ProbDev *hp, int *max_nodes);
// attenuation and c_eff at the profile nodes: d_nodes[0 .. P ldn) and [P ldn .. 2 P ldn)
void launch_node_tables(hipStream_t st, const ProbDev *probs, int P, int ldn, double *d_nodes);

// wide.hip: the wide-angle eigenvalue and eigenvector-rows kernels; BC holds the B rows of
// every problem (plane elements) followed by the C rows
void launch_k_wmodes(int ms, dim3 grid, hipStream_t st, const ModeProb *mp, const ModeWork *work,
                     const double *BC, size_t plane, const double *PQ, const double *beta, const ModesOut &mo);
void launch_k_wrows(dim3 grid, hipStream_t st, const ModeProb *mp, const ModeWork *work, const double *BC,
                    size_t plane, const double *PQ, const double *beta, const double *k2, const RowsArgs &ra,
                    double *TH, int *THE, double *PH, int *PHE, double *Q0, int *QEX);

// broadband.hip
void bb_launch_kpert(hipStream_t st, int n_cap, const int *n_dev, const double *k2, const double *absn,
                     double *kre, double *kim);
void bb_launch_transfer(hipStream_t st, int n_bins, int n_freq, const int *beg, const int *cnt, const double *kre,
                        const double *kim, const double *psr, int zero_attn, const double *ranges, int n_ranges,
                        const double *fac, double *H);
void bb_launch_arg(hipStream_t st, int n_tr, int n_freq, int nfft, const double *H, const double *S,
                   const double *f, const double *t0, int n_ranges, double *arg);
int bb_prepare_fft(ncpa_ctx *ctx, int n, int batch, size_t *work_doubles);
int bb_launch_fft(ncpa_ctx *ctx, const double *in, double *out, int n, int batch, int sign, double *work,
                  bool timed = true);
void bb_launch_out(hipStream_t st, size_t n, const double *spec, double df, double *wave);
// device address of ncpa_bb_pulse_noalloc's fixed array on ctx's device (looked up once per device)
int pulse_scratch_setup(ncpa_ctx *ctx, double **out);

}  // namespace ncpa

// Modes of a solved batch (device-resident), shared by the translation units.
struct ncpa_modes {
    ncpa_ctx *ctx = nullptr;
    int P = 0;
    int total = 0;
    bool wide = false;                         // wide-angle (WModeSolver) modes
    ncpa_grid_t grid{};                        // dz: the matrix spacing
    std::vector<int> counts, offsets;
    std::vector<ncpa::IntervalOut> iv;
    std::vector<ncpa::ModeProb> mp;            // host copy of the per-problem records
    int *d_offsets = nullptr;
    double *d_k2 = nullptr, *d_abs = nullptr, *d_psis = nullptr, *d_psir = nullptr, *d_psisr = nullptr;
    int *d_iters = nullptr;
    int *d_rows = nullptr;
    double *d_A = nullptr, *d_W = nullptr, *d_CE = nullptr;   // per-problem grid samples
    double *d_BC = nullptr;                    // wide-angle B and C rows
    double *d_PQ = nullptr, *d_beta = nullptr; // wide-angle factored rows P, Q and ground terms
    double *d_psis_rows = nullptr;             // psi(zs) from the last eigenvector-rows pass
    ncpa::ModeProb *d_mp = nullptr;
    ncpa::ModeWork *d_work = nullptr;
    int *d_mode_prob = nullptr;
    size_t n_work = 0;
};

namespace ncpa {
int check_grid(ncpa_ctx *ctx, const ncpa_grid_t *g);
void release_modes(ncpa_modes *m);
// eigenvalue kernel over the intervals in m->iv (matrices already in m's buffers)
int launch_modes(ncpa_ctx *ctx, ncpa_modes *m, const std::vector<int> &srow, const std::vector<int> &rrow);
// eigenvectors of every mode at rows row0 + k*stride, k < nrec -> dev_psi[k * total + o]
int launch_rows(ncpa_ctx *ctx, ncpa_modes *m, int row0, int stride, int nrec, double *dev_psi);
}  // namespace ncpa
