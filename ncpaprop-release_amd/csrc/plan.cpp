// Broadband job plans: a batched ModBB dispersion solve (n_sets x (n_freq - 1) problems)
// followed by pulse synthesis at the receiver ranges, as one HIP graph.
//
// The reference runs this as two programs over a text file: ModBB --dispersion solves one
// frequency at a time (ESSModeSolver::solve per frequency, src/modbb/ModBB_main.cpp) and
// ModBB --propagation reads the file back (ModalBroadbandPropagator::calculate_waveform,
// src/libbroadband/ModalBroadbandPropagator.cpp:140-243).  Here every frequency and set of
// the job is one batch:
//   k_alpha_nodes, k_setup_grid, k_interval   grids, Sturm counts of [k_min^2, k_max^2]
//   k_plan_build                              mode offsets + XCD-interleaved work list, on the device
//   k_modes (persistent, per-XCD queues)      every eigenvalue of every problem, uneven counts
//   k_kpert, k_bb_transfer                    k_pert, transfer functions H(f, r)
//   k_bb_arg, k_fft, k_bb_out                 e^{-2 pi i f t0} S H, FFT, 2 Re(df .)
// Nothing in the sequence needs a count on the host, so it is captured once into a graph
// (no memset nodes: the queue counters are reset by k_plan_build) and replayed per job.
// An autotuner times the eigenvalue kernel's multisection depth and persistent grid size on
// the plan's own batch before capture.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.h"
#include "internal.h"

using namespace ncpa;

struct ncpa_bb_plan {
    ncpa_ctx *ctx = nullptr;
    ncpa_grid_t grid{};
    int P = 0, max_nodes = 0;
    int n_sets = 0, n_freq = 0, n_ranges = 0, nfft = 0, zero_attn = 0;
    double df = 0.0;
    int cap_modes = 0, cap_work = 0;
    int ms = 2, blocks = 0;
    double tuned_ms = 0.0;
    std::vector<double> tune_log;   // (ms, blocks, milliseconds) triples
    // device buffers (owned)
    ProbDev *probs = nullptr;
    double *alpha = nullptr, *A = nullptr, *W = nullptr, *CE = nullptr;
    IntervalOut *iv = nullptr;
    ModeProb *mp = nullptr;
    int *offsets = nullptr, *mbeg = nullptr, *mcnt = nullptr;
    ModeWork *work = nullptr;
    PlanMeta *meta = nullptr;
    double *k2 = nullptr, *absn = nullptr, *psis = nullptr, *psir = nullptr, *psisr = nullptr;
    int *iters = nullptr;
    double *kre = nullptr, *kim = nullptr;
    double *ranges = nullptr, *fac = nullptr, *S = nullptr, *f = nullptr, *t0 = nullptr;
    double *H = nullptr, *arg = nullptr, *spec = nullptr, *wave = nullptr, *fft_scratch = nullptr;
    std::vector<void *> owned;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

template <class T>
int dev_alloc(ncpa_bb_plan *pl, T **p, size_t n) {
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(pl->ctx, NCPA_ERR_ALLOC, "ncpa_bb_plan: device allocation failed");
    }
    pl->owned.push_back(q);
    *p = static_cast<T *>(q);
    return NCPA_OK;
}

void release(ncpa_bb_plan *pl) {
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    for (void *q : pl->owned) (void)hipFree(q);
    pl->owned.clear();
}

ModesOut modes_out(const ncpa_bb_plan *pl) {
    return ModesOut{pl->k2, pl->absn, pl->psis, pl->psir, pl->psisr, pl->iters, nullptr};
}

// grids, intervals and the device work list
void enqueue_front(ncpa_bb_plan *pl, int cap_modes, int cap_work) {
    hipStream_t st = pl->ctx->stream;
    launch_solve_front(st, pl->probs, pl->P, pl->max_nodes, pl->grid, pl->alpha, pl->A, pl->W, pl->CE, pl->iv);
    launch_plan_build(st, pl->probs, pl->iv, pl->P, pl->grid.Nz, pl->grid.dz, pl->A, pl->W, cap_modes, cap_work,
                      pl->n_sets, pl->n_freq, pl->mp, pl->offsets, pl->work, pl->mbeg, pl->mcnt, pl->meta);
}

// blocks > 0: persistent form with that grid; 0: one block per work-list slot (grid =
// capacity, blocks past the device count exit), diagonal through the LDS ring; -1: the same
// with the count and Newton sweeps reading the diagonal by scalar loads
void enqueue_modes(ncpa_bb_plan *pl, int ms, int blocks) {
    if (blocks > 0)
        launch_k_modes(ms, dim3((unsigned)blocks), pl->ctx->stream, pl->mp, pl->work, 0, &pl->meta->n_work,
                       pl->meta->queue, modes_out(pl));
    else
        launch_k_modes(ms, dim3((unsigned)pl->cap_work), pl->ctx->stream, pl->mp, pl->work, 0, &pl->meta->n_work,
                       nullptr, modes_out(pl), blocks < 0);
}

int enqueue_all(ncpa_bb_plan *pl) {
    hipStream_t st = pl->ctx->stream;
    enqueue_front(pl, pl->cap_modes, pl->cap_work);
    enqueue_modes(pl, pl->ms, pl->blocks);
    bb_launch_kpert(st, pl->cap_modes, &pl->meta->total, pl->k2, pl->absn, pl->kre, pl->kim);
    bb_launch_transfer(st, pl->n_sets * pl->n_freq, pl->n_freq, pl->mbeg, pl->mcnt, pl->kre, pl->kim, pl->psisr,
                       pl->zero_attn, pl->ranges, pl->n_ranges, pl->fac, pl->H);
    const int n_tr = pl->n_sets * pl->n_ranges;
    bb_launch_arg(st, n_tr, pl->n_freq, pl->nfft, pl->H, pl->S, pl->f, pl->t0, pl->n_ranges, pl->arg);
    int rc = bb_launch_fft(pl->ctx, pl->arg, pl->spec, pl->nfft, n_tr, -1, pl->fft_scratch);   // FFTW_FORWARD
    if (rc) return rc;
    bb_launch_out(st, (size_t)n_tr * pl->nfft, pl->spec, pl->df, pl->wave);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(pl->ctx, e, "ncpa_bb_plan launch");
    return NCPA_OK;
}

int read_meta(ncpa_bb_plan *pl, PlanMeta *m) {
    NCPA_HIP(pl->ctx, hipMemcpyAsync(m, pl->meta, sizeof(PlanMeta), hipMemcpyDeviceToHost, pl->ctx->stream));
    NCPA_HIP(pl->ctx, hipStreamSynchronize(pl->ctx->stream));
    return NCPA_OK;
}

int overflow_error(ncpa_bb_plan *pl, const PlanMeta &m) {
    if (m.overflow & 1)
        return set_error(pl->ctx, NCPA_ERR_MAX_MODES,
                         std::to_string(m.max_nm) + " modes exceed MAX_MODES=" + std::to_string(NCPA_MAX_MODES));
    return set_error(pl->ctx, NCPA_ERR_MAX_MODES,
                     "ncpa_bb_plan: " + std::to_string(m.total) + " modes exceed the plan's capacity of " +
                         std::to_string(pl->cap_modes));
}

// Times every (multisection passes, launch form) candidate on the plan's own batch (front
// stages once, then k_plan_build + k_modes per candidate; k_plan_build resets the queues)
// and keeps the fastest.  Launch forms: persistent grids of 4, 8 and 15 blocks per CU (one
// wave per SIMD, two, the LDS limit) and one block per work-list slot with the diagonal
// through the LDS ring (blocks = 0) or through scalar loads (blocks = -1).
int autotune(ncpa_bb_plan *pl, int n_cu) {
    hipStream_t st = pl->ctx->stream;
    hipEvent_t e0, e1;
    NCPA_HIP(pl->ctx, hipEventCreate(&e0));
    NCPA_HIP(pl->ctx, hipEventCreate(&e1));
    const int ms_cand[] = {1, 2, 3};
    const int per_cu[] = {0, -1, 4, 8, 15};   // (-1: scalar-load sweeps, one block per slot)
    double best = 1e300;
    int rc = NCPA_OK;
    enqueue_front(pl, pl->cap_modes, pl->cap_work);
    enqueue_modes(pl, pl->ms, pl->blocks);   // warm-up
    for (int ms : ms_cand)
        for (int k : per_cu) {
            if (rc) break;
            const int blocks = k < 0 ? -1 : n_cu * k;
            launch_plan_build(st, pl->probs, pl->iv, pl->P, pl->grid.Nz, pl->grid.dz, pl->A, pl->W, pl->cap_modes,
                              pl->cap_work, pl->n_sets, pl->n_freq, pl->mp, pl->offsets, pl->work, pl->mbeg,
                              pl->mcnt, pl->meta);
            hipError_t e = hipEventRecord(e0, st);
            enqueue_modes(pl, ms, blocks);
            if (e == hipSuccess) e = hipEventRecord(e1, st);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float t = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
            if (e != hipSuccess) {
                rc = hip_error(pl->ctx, e, "ncpa_bb_plan autotune");
                break;
            }
            pl->tune_log.insert(pl->tune_log.end(), {(double)ms, (double)blocks, (double)t});
            if (t < best) {
                best = t;
                pl->ms = ms;
                pl->blocks = blocks;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    pl->tuned_ms = best;
    return rc;
}

}  // namespace

extern "C" {

int ncpa_bb_plan_create(ncpa_ctx_t *ctx, const ncpa_grid_t *grid, const ncpa_problem_t *problems,
                        const ncpa_bb_job_t *job, ncpa_bb_plan_t **out) {
    if (!ctx) return NCPA_ERR_ARG;
    if (!grid || !problems || !job || !out || job->n_sets < 1 || job->n_freq < 2 || !job->f_vec ||
        !job->spectrum || !job->ranges_m || job->n_ranges < 1 || !job->rho_ratio || !(job->max_celerity > 0.0) ||
        job->nfft < job->n_freq || job->max_modes_total < 0 || grid->Nz < 10 || !(grid->dz > 0.0))
        return set_error(ctx, NCPA_ERR_ARG, "ncpa_bb_plan_create: bad arguments");
    for (int i = 0; i < job->n_ranges; ++i)
        if (!(job->ranges_m[i] > 0.0)) return set_error(ctx, NCPA_ERR_ARG, "ncpa_bb_plan_create: ranges must be > 0");
    for (int s = 0; s < job->n_sets; ++s)
        if (!(job->rho_ratio[s] >= 0.0))
            return set_error(ctx, NCPA_ERR_ARG, "ncpa_bb_plan_create: rho ratio must be >= 0");
    *out = nullptr;
    NCPA_HIP(ctx, hipSetDevice(ctx->device));
    ncpa_bb_plan *pl = new ncpa_bb_plan();
    pl->ctx = ctx;
    pl->grid = *grid;
    pl->n_sets = job->n_sets;
    pl->n_freq = job->n_freq;
    pl->P = job->n_sets * (job->n_freq - 1);
    pl->n_ranges = job->n_ranges;
    pl->nfft = job->nfft;
    pl->zero_attn = job->zero_attn ? 1 : 0;
    pl->df = job->f_vec[1] - job->f_vec[0];   // BroadbandPropagator.cpp:581
    auto fail = [&](int code) {
        release(pl);
        delete pl;
        return code;
    };
    const int P = pl->P, Nz = grid->Nz;
    std::vector<ProbDev> hp(P);
    int rc = build_problems(ctx, grid, problems, P, hp.data(), &pl->max_nodes);
    if (rc) return fail(rc);
    const bool timing = ctx->timing;
    ctx->timing = false;   // no event records inside the captured sequence
    auto done = [&](int code) {
        ctx->timing = timing;
        return code == NCPA_OK ? code : fail(code);
    };
    // per-problem buffers, then a count-only pass (capacity 0) to size the mode buffers
    if ((rc = dev_alloc(pl, &pl->probs, P)) || (rc = dev_alloc(pl, &pl->alpha, 2 * (size_t)P * pl->max_nodes)) ||
        (rc = dev_alloc(pl, &pl->A, (size_t)P * Nz)) || (rc = dev_alloc(pl, &pl->W, (size_t)P * Nz)) ||
        (rc = dev_alloc(pl, &pl->CE, (size_t)P * Nz)) || (rc = dev_alloc(pl, &pl->iv, P)) ||
        (rc = dev_alloc(pl, &pl->mp, P)) || (rc = dev_alloc(pl, &pl->offsets, (size_t)P + 1)) ||
        (rc = dev_alloc(pl, &pl->mbeg, (size_t)pl->n_sets * pl->n_freq)) ||
        (rc = dev_alloc(pl, &pl->mcnt, (size_t)pl->n_sets * pl->n_freq)) || (rc = dev_alloc(pl, &pl->meta, 1)))
        return done(rc);
    hipError_t e = hipMemcpyAsync(pl->probs, hp.data(), sizeof(ProbDev) * P, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return done(hip_error(ctx, e, "ncpa_bb_plan_create upload"));
    enqueue_front(pl, 0, 0);
    PlanMeta m{};
    if ((rc = read_meta(pl, &m))) return done(rc);
    if (m.overflow & 1) return done(overflow_error(pl, m));
    pl->cap_modes = job->max_modes_total > 0 ? job->max_modes_total : m.total + m.total / 2 + 64;
    if (m.total > pl->cap_modes) return done(overflow_error(pl, m));
    // work-list slots: the batch's need with the same 1.5x headroom (the one-block-per-slot
    // launch has this many blocks; a later batch that needs more is reported as overflow)
    pl->cap_work = job->max_modes_total > 0 ? NCPA_XCDS * (pl->cap_modes / 64 + P) + NCPA_XCDS
                                            : m.n_work_need + m.n_work_need / 2 + 8 * NCPA_XCDS;
    const size_t cm = (size_t)pl->cap_modes;
    const int n_tr = pl->n_sets * pl->n_ranges;
    size_t scratch = 0;
    if ((rc = bb_prepare_fft(ctx, pl->nfft, n_tr, &scratch))) return done(rc);
    if ((rc = dev_alloc(pl, &pl->work, (size_t)pl->cap_work)) || (rc = dev_alloc(pl, &pl->k2, cm)) ||
        (rc = dev_alloc(pl, &pl->absn, cm)) || (rc = dev_alloc(pl, &pl->psis, cm)) ||
        (rc = dev_alloc(pl, &pl->psir, cm)) || (rc = dev_alloc(pl, &pl->psisr, cm)) ||
        (rc = dev_alloc(pl, &pl->iters, cm)) || (rc = dev_alloc(pl, &pl->kre, cm)) ||
        (rc = dev_alloc(pl, &pl->kim, cm)) || (rc = dev_alloc(pl, &pl->ranges, pl->n_ranges)) ||
        (rc = dev_alloc(pl, &pl->fac, pl->n_sets)) || (rc = dev_alloc(pl, &pl->S, 2 * (size_t)pl->n_freq)) ||
        (rc = dev_alloc(pl, &pl->f, pl->n_freq)) || (rc = dev_alloc(pl, &pl->t0, pl->n_ranges)) ||
        (rc = dev_alloc(pl, &pl->H, 2 * (size_t)n_tr * pl->n_freq)) ||
        (rc = dev_alloc(pl, &pl->arg, 2 * (size_t)n_tr * pl->nfft)) ||
        (rc = dev_alloc(pl, &pl->spec, 2 * (size_t)n_tr * pl->nfft)) ||
        (rc = dev_alloc(pl, &pl->wave, (size_t)n_tr * pl->nfft)))
        return done(rc);
    if (scratch && (rc = dev_alloc(pl, &pl->fft_scratch, scratch))) return done(rc);
    std::vector<double> fac(pl->n_sets), t0(pl->n_ranges);
    for (int s = 0; s < pl->n_sets; ++s) fac[s] = std::sqrt(job->rho_ratio[s]);
    for (int i = 0; i < pl->n_ranges; ++i) t0[i] = job->ranges_m[i] / job->max_celerity;   // ModalBroadbandPropagator.cpp:211
    if ((e = hipMemcpyAsync(pl->ranges, job->ranges_m, sizeof(double) * pl->n_ranges, hipMemcpyHostToDevice,
                            ctx->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(pl->fac, fac.data(), sizeof(double) * pl->n_sets, hipMemcpyHostToDevice, ctx->stream)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(pl->S, job->spectrum, sizeof(double) * 2 * pl->n_freq, hipMemcpyHostToDevice,
                            ctx->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(pl->f, job->f_vec, sizeof(double) * pl->n_freq, hipMemcpyHostToDevice, ctx->stream)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(pl->t0, t0.data(), sizeof(double) * pl->n_ranges, hipMemcpyHostToDevice, ctx->stream)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
        return done(hip_error(ctx, e, "ncpa_bb_plan_create upload"));
    int n_cu = 0;
    if ((e = hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device)) != hipSuccess)
        return done(hip_error(ctx, e, "ncpa_bb_plan_create"));
    pl->ms = ctx->ms_passes;
    pl->blocks = 0;
    if (job->tune) {
        if ((rc = autotune(pl, std::max(n_cu, 1)))) return done(rc);
        ctx->ms_passes = pl->ms;   // the eager solves of this context follow the tuned depth
        if (pl->blocks <= 0) ctx->modes_smem = pl->blocks < 0;   // and sweep form
    }
    // capture
    if ((e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return done(hip_error(ctx, e, "ncpa_bb_plan capture"));
    rc = enqueue_all(pl);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return done(rc);
    }
    if (e != hipSuccess) return done(hip_error(ctx, e, "ncpa_bb_plan capture"));
    pl->graph = g;
    if ((e = hipGraphInstantiate(&pl->exec, g, nullptr, nullptr, 0)) != hipSuccess)
        return done(hip_error(ctx, e, "ncpa_bb_plan instantiate"));
    ctx->timing = timing;
    *out = pl;
    return NCPA_OK;
}

int ncpa_bb_plan_set_problems(ncpa_bb_plan_t *pl, const ncpa_problem_t *problems) {
    if (!pl || !problems) return NCPA_ERR_ARG;
    std::vector<ProbDev> hp(pl->P);
    int max_nodes = 0;
    int rc = build_problems(pl->ctx, &pl->grid, problems, pl->P, hp.data(), &max_nodes);
    if (rc) return rc;
    if (max_nodes > pl->max_nodes)
        return set_error(pl->ctx, NCPA_ERR_ARG, "ncpa_bb_plan_set_problems: profiles with more nodes than at creation");
    NCPA_HIP(pl->ctx, hipMemcpyAsync(pl->probs, hp.data(), sizeof(ProbDev) * pl->P, hipMemcpyHostToDevice,
                                     pl->ctx->stream));
    NCPA_HIP(pl->ctx, hipStreamSynchronize(pl->ctx->stream));
    return NCPA_OK;
}

int ncpa_bb_plan_run(ncpa_bb_plan_t *pl) {
    if (!pl) return NCPA_ERR_ARG;
    NCPA_HIP(pl->ctx, hipGraphLaunch(pl->exec, pl->ctx->stream));
    return NCPA_OK;
}

int ncpa_bb_plan_result(ncpa_bb_plan_t *pl, const double **dev_wave, const double **dev_kre, const double **dev_kim,
                        const double **dev_psi_sr, const int **dev_offsets, int *total_modes) {
    if (!pl) return NCPA_ERR_ARG;
    PlanMeta m{};
    int rc = read_meta(pl, &m);
    if (rc) return rc;
    if (total_modes) *total_modes = m.total;
    if (m.overflow) return overflow_error(pl, m);
    if (dev_wave) *dev_wave = pl->wave;
    if (dev_kre) *dev_kre = pl->kre;
    if (dev_kim) *dev_kim = pl->kim;
    if (dev_psi_sr) *dev_psi_sr = pl->psisr;
    if (dev_offsets) *dev_offsets = pl->offsets;
    return NCPA_OK;
}

int ncpa_bb_plan_tuning(ncpa_bb_plan_t *pl, int *ms_passes, int *grid_blocks, double *modes_ms, double *log,
                        int *n_log) {
    if (!pl) return NCPA_ERR_ARG;
    if (ms_passes) *ms_passes = pl->ms;
    if (grid_blocks) *grid_blocks = pl->blocks;
    if (modes_ms) *modes_ms = pl->tuned_ms;
    if (n_log) {
        const int n = (int)(pl->tune_log.size() / 3);
        if (log) std::memcpy(log, pl->tune_log.data(), sizeof(double) * 3 * std::min(n, *n_log));
        *n_log = n;
    }
    return NCPA_OK;
}

void ncpa_bb_plan_destroy(ncpa_bb_plan_t *pl) {
    if (!pl) return;
    (void)hipStreamSynchronize(pl->ctx->stream);
    release(pl);
    delete pl;
}

}  // extern "C"
