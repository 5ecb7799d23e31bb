/*
 * ncpamodes.h -- C ABI of the MI355X effective-sound-speed normal-mode library.
 *
 * Every entry point returns an int status (NCPA_OK == 0, negative on error; the message
 * is available from ncpa_ctx_error()).  No C++ exception crosses this boundary.  Host
 * pointers are owned by the caller; device buffers passed in ("dev_" arguments) are
 * caller-allocated device memory; buffers the library allocates live behind handles.
 * Work is ordered on the context's stream; one context per host thread.
 *
 * Reference interfaces replaced (ncpaprop, chetzer-ncpa/ncpaprop-release):
 *   EigenEngine::doESSCalculation        src/libmodes/EigenEngine.h:13-14, EigenEngine.cpp:10-172
 *   ESSModeSolver::setParams/getModalTrace  src/libmodes/ESSModeSolver.cpp:30-240, 518-645
 *   ESSModeSolver::doSelect              src/libmodes/ESSModeSolver.cpp:647-665
 *   ModeSolver::doPerturb                src/libmodes/ModeSolver.cpp:263-284
 *   ModeSolver::getTLoss1D / getTLoss1DNx2  src/libmodes/ModeSolver.cpp:292-381, 387-489
 *   ModeSolver::getTLoss2D               src/libmodes/ModeSolver.cpp:492-565
 *   ModalBroadbandPropagator::compute_modal_sum  src/libbroadband/ModalBroadbandPropagator.cpp:245-277
 *   BroadbandPropagator::fft_pulse_prop  src/libbroadband/BroadbandPropagator.cpp:573-651
 */
#ifndef NCPAMODES_H
#define NCPAMODES_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NCPA_OK 0
#define NCPA_ERR_ARG (-1)
#define NCPA_ERR_HIP (-2)
#define NCPA_ERR_ALLOC (-3)
#define NCPA_ERR_MAX_MODES (-4)
#define NCPA_ERR_STATE (-5)

#define NCPA_MAX_MODES 4000   /* src/libmodes/ESSModeSolver.cpp:17 */

typedef struct ncpa_ctx ncpa_ctx_t;
typedef struct ncpa_modes ncpa_modes_t;
typedef struct ncpa_event ncpa_event_t;

/* ---------------------------------------------------------------------------------- */
/* context                                                                             */
/* ---------------------------------------------------------------------------------- */

/* Create a context on HIP device `device`.  `hip_stream` may be NULL (the library then
 * creates its own stream, default priority) or an existing hipStream_t to share. */
int ncpa_ctx_create(int device, void *hip_stream, ncpa_ctx_t **out);
void ncpa_ctx_destroy(ncpa_ctx_t *ctx);
const char *ncpa_ctx_error(const ncpa_ctx_t *ctx);
int ncpa_ctx_synchronize(ncpa_ctx_t *ctx);
/* Returns the hipStream_t the context launches on. */
void *ncpa_ctx_stream(ncpa_ctx_t *ctx);

/* Per-kernel timing with HIP events recorded on the context stream around each launch of
 * the named kernel family ("modes", "tl1d", "setup", "interval", "rows", "field2d",
 * "bb_transfer", "fft").  ncpa_ctx_kernel_time synchronises and returns the summed
 * duration (ms) and the number of timed launches since the last reset. */
int ncpa_ctx_set_timing(ncpa_ctx_t *ctx, int enable);
int ncpa_ctx_kernel_time(ncpa_ctx_t *ctx, const char *kernel, double *total_ms, int *launches);
int ncpa_ctx_reset_timing(ncpa_ctx_t *ctx);

/* How the eigenvalue sweeps of this context's eager solves (ncpa_ess_solve) and the sweeps of
 * its eigenvector rows (ncpa_modes_rows) read the matrix diagonal: 1 (default) by scalar loads
 * (no LDS, 6 waves per SIMD; the rows store their recorded values as the sweep passes them);
 * 0 through an LDS ring (the rows keep their records in LDS).  The wide-angle solver
 * (ncpa_wmod_solve) reads its rows through LDS in both forms.  Both forms give the same results
 * to the bit.  A broadband plan created with tuning sets the form (and
 * the multisection depth) its autotuner picked.  No counterpart in the reference, which has
 * one CPU path. */
int ncpa_ctx_set_modes_form(ncpa_ctx_t *ctx, int form);

/* Device memory of the context (no counterpart in the reference): live = bytes its
 * allocator has handed out (requests rounded up to 256 B) plus the buffers of
 * ncpa_dev_alloc still allocated; peak = the largest live since the last call with
 * reset_peak != 0 (which starts a new window at the current live); cached = bytes held in
 * the allocator's free blocks.  Any pointer may be NULL. */
int ncpa_ctx_memory(ncpa_ctx_t *ctx, int reset_peak, size_t *live_bytes, size_t *peak_bytes,
                    size_t *cached_bytes);

/* Ordering between contexts (no counterpart in the reference, whose solver is one
 * sequential loop; these let a caller pipeline the stages of consecutive batches over two
 * contexts, e.g. eigenpairs of batch i+1 on one stream while the 2-D field of batch i is
 * contracted on another).  ncpa_event_record marks the current end of ctx's stream;
 * ncpa_event_wait makes work later queued on ctx wait for it (no host wait);
 * ncpa_event_synchronize waits on the host.  A modes handle may be used on another context
 * (e.g. ncpa_field2d) once that context waits for an event recorded after the work that
 * produced it; its buffers belong to the context that solved it, so the caller keeps the
 * handle until the other context's work on it has completed (an event waited for by the
 * owning context, or synchronised on the host) before ncpa_modes_free. */
int ncpa_event_record(ncpa_ctx_t *ctx, ncpa_event_t **ev);
int ncpa_event_wait(ncpa_ctx_t *ctx, ncpa_event_t *ev);
int ncpa_event_synchronize(ncpa_event_t *ev);
void ncpa_event_destroy(ncpa_event_t *ev);

/* Device memory owned by the caller (allocated on the context's device).  Copies are
 * ordered on the context stream; download and upload return once the host buffer may be
 * reused. */
int ncpa_dev_alloc(ncpa_ctx_t *ctx, size_t bytes, void **dev_ptr);
int ncpa_dev_free(ncpa_ctx_t *ctx, void *dev_ptr);
int ncpa_dev_upload(ncpa_ctx_t *ctx, void *dev_dst, const void *host_src, size_t bytes);
int ncpa_dev_download(ncpa_ctx_t *ctx, void *host_dst, const void *dev_src, size_t bytes);
/* Device-to-device copy by a plain streaming kernel ("copy"); bytes must be a multiple of 32.
 * Used to calibrate hardware byte counters against a kernel of known traffic. */
int ncpa_dev_copy(ncpa_ctx_t *ctx, void *dev_dst, const void *dev_src, size_t bytes);

/* ---------------------------------------------------------------------------------- */
/* atmosphere profiles (node arrays in SI units, as ESSModeSolver::setParams leaves them)*/
/* ---------------------------------------------------------------------------------- */

/* Upload one profile.  z [m] ascending; c0, u, v [m/s]; T [K]; P [Pa]; rho [kg/m3].
 * Wind speed/direction and the node-level quantities of Atmosphere1D.cpp:484-610 are
 * formed on the device.  Returns a profile id (>= 0) in *profile_id. */
int ncpa_profile_upload(ncpa_ctx_t *ctx, int n_nodes, const double *z, const double *c0,
                        const double *u, const double *v, const double *T, const double *P,
                        const double *rho, int *profile_id);
/* Optional user attenuation (--use_attn_file): alpha [Np/m] at the profile nodes. */
int ncpa_profile_set_attenuation(ncpa_ctx_t *ctx, int profile_id, const double *alpha_nodes);
int ncpa_profile_clear(ncpa_ctx_t *ctx);

typedef struct ncpa_profile_info {
    int n_nodes;
    int has_z0;        /* scalar Z0 present */
    double z0;         /* ground height [m] */
    double z_bottom;   /* first altitude [m] */
    double z_top;      /* last altitude [m] */
} ncpa_profile_info_t;

/* Read an Atmosphere1D text profile (Atmosphere1D(file, headerfile),
 * src/atmosphere/Atmosphere1D.cpp:44-60,85-285; units src/common/units.cpp:131-178), convert
 * to SI, choose c0 as ESSModeSolver.cpp:159-183 does, and upload it.  header_path may be
 * NULL or "" (the header is then read from `path`). */
int ncpa_profile_read(ncpa_ctx_t *ctx, const char *path, const char *header_path, int *profile_id,
                      ncpa_profile_info_t *info);
/* The same reader without a context or device: parses and checks the profile and fills
 * *info (may be NULL).  NCPA_ERR_ARG with the reference's message (Atmosphere1D.cpp:85-285,
 * ESSModeSolver.cpp:176-210: the messages thrown there) in msg[msg_len] for a missing file, a
 * malformed header or data line, duplicate keys, unknown units, missing U/V/T/P/RHO,
 * altitudes that do not increase or non-finite values.  The command lines call it before
 * any device work. */
int ncpa_profile_check(const char *path, const char *header_path, ncpa_profile_info_t *info, char *msg,
                       int msg_len);
/* Host copies of a profile's SI node arrays (each n_nodes long; any pointer may be NULL):
 * z, c0, wind speed, wind direction [deg CW from N], T, P, rho. */
int ncpa_profile_nodes(ncpa_ctx_t *ctx, int profile_id, double *z, double *c0, double *ws,
                       double *wd_deg, double *T, double *P, double *rho);
/* Lamb-wave ground admittance -rho'(z)/rho(z)/2 from the Steffen spline of rho
 * (ESSModeSolver.cpp:347-349). */
int ncpa_profile_admittance(ncpa_ctx_t *ctx, int profile_id, double z_min, double *admittance);

/* ---------------------------------------------------------------------------------- */
/* batched ESS solve: grid -> modal trace -> eigenvalues -> selection -> perturbation   */
/* ---------------------------------------------------------------------------------- */

/* Vertical grid shared by a batch (ESSModeSolver.cpp:144-151): rows Hgt[i] = z_min + i dz. */
typedef struct ncpa_grid {
    int Nz;            /* --Nz_grid */
    double z_min;      /* [m], after the Z0 adjustment */
    double dz;         /* (maxheight - z_min)/(Nz - 1), maxheight already clamped */
    double maxheight;  /* [m] */
} ncpa_grid_t;

/* The grid ESSModeSolver::setParams builds for a profile: maxheight clamped to the profile
 * top - 1e-6 m (:144-148), z_min raised to the profile's Z0 (:212-220). */
int ncpa_grid_make(ncpa_ctx_t *ctx, int profile_id, int Nz, double maxheight_km,
                   double zground_km, ncpa_grid_t *grid);
/* Steffen samples of rho, P, T on the grid rows (host arrays of Nz, may be NULL). */
int ncpa_profile_sample(ncpa_ctx_t *ctx, int profile_id, const ncpa_grid_t *grid, double *rho,
                        double *P, double *T, double *unused);

/* One (profile, frequency, azimuth) problem. */
typedef struct ncpa_problem {
    int profile;          /* id from ncpa_profile_upload */
    int flags;            /* NCPA_F_* */
    double freq;          /* [Hz] */
    double azimuth;       /* [deg CW from N] */
    double alpha_freq;    /* frequency at which Sutherland-Bass attenuation is evaluated */
    double zs;            /* source height AGL [m] */
    double zr;            /* receiver height AGL [m] */
    double admittance;    /* ground admittance (0 for rigid ground) */
    double c_min, c_max;  /* phase-speed filter [m/s] (NCPA_F_WVNUM_FILTER) */
} ncpa_problem_t;

#define NCPA_F_WKB          1  /* WKB k_max cut-off (ground-to-ground 1-D runs) */
#define NCPA_F_WVNUM_FILTER 2  /* k_min = w/c_max, k_max = w/c_min */

/* Solve a batch of problems; results stay on the device behind *out.
 * Replaces, per problem, getModalTrace + doESSCalculation + doSelect + doPerturb for the
 * mode data the 1-D TL and dispersion outputs need: k_pert and psi(zs) psi(zr). */
int ncpa_ess_solve(ncpa_ctx_t *ctx, const ncpa_grid_t *grid, const ncpa_problem_t *problems,
                   int n_problems, ncpa_modes_t **out);
void ncpa_modes_free(ncpa_modes_t *modes);

/* The mode count of every problem (host array of n_problems) without the modes: the grid,
 * getModalTrace's interval and the Sturm counts at its ends, i.e. the number of modes
 * ncpa_ess_solve would return (ESSModeSolver.cpp:366-377, whose printed count comes from its
 * own, differently rounded, getNumberOfModes).  No counterpart in the reference: it sizes a
 * batch's device memory before the batch is admitted. */
int ncpa_ess_count(ncpa_ctx_t *ctx, const ncpa_grid_t *grid, const ncpa_problem_t *problems,
                   int n_problems, int *counts);

/* Mode counts per problem (host array of n_problems) and total. */
int ncpa_modes_counts(ncpa_modes_t *modes, int *counts, int *total);
/* Per-problem interval [k_min, k_max] actually used (host arrays, may be NULL). */
int ncpa_modes_interval(ncpa_modes_t *modes, double *k_min, double *k_max);
/* Download modes (host arrays of length total, problem-major, ascending k within a
 * problem): k_pert real/imag, psi(zs)*psi(zr), psi(zs), psi(zr) (sign convention:
 * psi(zs) >= 0); any pointer may be NULL. */
int ncpa_modes_download(ncpa_modes_t *modes, double *k_re, double *k_im, double *psi_sr,
                        double *psi_s, double *psi_r);
/* Sweeps each eigenvalue needed (host int[total]); negative when its bracket collapsed
 * to working precision before Newton converged. */
int ncpa_modes_iterations(ncpa_modes_t *modes, int *iters);
/* Rows of the tridiagonal each mode's eigenvalue search swept in total (host int[total]): the
 * sweeps it used, each over the rows below the evanescent top its block leaves out (the
 * work behind the eigenvalue kernel's throughput figures).  No counterpart in the reference. */
int ncpa_modes_sweep_rows(ncpa_modes_t *modes, int *rows);
/* Device views (valid until ncpa_modes_free): offsets[n_problems+1] (int), and per mode
 * k^2, absorption A (k_pert = sqrt(k^2 + i A)) and psi(zs) psi(zr) (double). */
int ncpa_modes_device(ncpa_modes_t *modes, const int **dev_offsets, const double **dev_k2,
                      const double **dev_absorption, const double **dev_psi_sr);

/* Effective sound speed c_eff(Hgt_i) [m/s] on the grid rows of one problem (host array of Nz),
 * as getModalTrace leaves it (ESSModeSolver.cpp:544-556); used by the group-speed writer. */
int ncpa_modes_ceff(ncpa_modes_t *modes, int problem, double *c_eff);

/* Eigenvectors of every mode at grid rows row0 + k*stride (k < nrows), normalised so that
 * sum psi^2 dz = 1 with psi(zs) >= 0: dev_psi[k * total + mode] (device, nrows*total
 * doubles).  Used for 2-D fields and mode files (ModeSolver.cpp:513-540, 618-641). */
int ncpa_modes_rows(ncpa_ctx_t *ctx, ncpa_modes_t *modes, int row0, int stride, int nrows,
                    double *dev_psi);

/* 1-D transmission loss for every problem of `modes` at ranges r_i = (i+1) dr,
 * i < n_r (ModeSolver.cpp:325-356).  Outputs are device buffers of n_problems*n_r
 * doubles: Re p, Im p and the incoherent sum, scaled by 1000 and 4 pi as the reference
 * writes them.  lossless != 0 uses Re(k_pert). */
int ncpa_tl1d(ncpa_ctx_t *ctx, ncpa_modes_t *modes, double dr, int n_r, int lossless,
              double *dev_p_re, double *dev_p_im, double *dev_incoh);

/* Contraction precisions of the 2-D field */
#define NCPA_PREC_FP64   0   /* FP64 MFMA, complex128 output */
#define NCPA_PREC_FP32   1   /* FP32 MFMA, complex64 output */
#define NCPA_PREC_BF16X3 2   /* split-BF16, 2 parts, 3 products (hi*hi + hi*lo + lo*hi), FP32 accumulate */
#define NCPA_PREC_BF16X6 3   /* split-BF16, 3 parts, 6 products (FP32-class accuracy), FP32 accumulate */
#define NCPA_PREC_FP16X3 4   /* split-FP16, 2 parts, 3 products (~2^-22), power-of-two row/range scaling, FP32 accumulate */
#define NCPA_PREC_FP16X4 5   /* split-FP16, 2 parts, 4 products (exact part products), scaled, FP32 accumulate */
#define NCPA_PREC_BF16X6F 6  /* NCPA_PREC_BF16X6, each 32-mode step's FP32 sum folded into FP64 accumulators */
#define NCPA_PREC_FP16X4F 7  /* NCPA_PREC_FP16X4, each 32-mode step's FP32 sum folded into FP64 accumulators */

/* 2-D field of every problem (ModeSolver::getTLoss2D, ModeSolver.cpp:524-544) at the rows of
 * dev_psi_rows (from ncpa_modes_rows, nrows x total) and ranges r_i = (i+1) dr, i < n_r:
 * dev_out[p][k][2i + {0,1}] = Re, Im of p(r_i, z_k), scaled by 1000 and 4 pi as written by
 * the reference; element type double (FP64) or float (every other precision).  lossless != 0 uses
 * Re(k_pert).  The phase k r is formed in FP64 and reduced mod 2 pi in every precision. */
int ncpa_field2d(ncpa_ctx_t *ctx, ncpa_modes_t *modes, const double *dev_psi_rows, int nrows,
                 double dr, int n_r, int lossless, int precision, void *dev_out);

/* ncpa_field2d in two calls, so that the operand staging (phasors, splits, scales) and the
 * contraction can run on different contexts (no counterpart in the reference, whose
 * getTLoss2D is one loop): ncpa_field2d_stage queues the staging on ctx and returns a handle
 * to the staged operands (buffers of ctx's pool); ncpa_field2d_contract queues the
 * contraction on any context that has waited for an event recorded after the staging
 * (ncpa_event_*), writing dev_out as ncpa_field2d does; ncpa_field_stage_free returns the
 * buffers to the staging context's pool, whose stream first waits (on the device, no host
 * wait) for the last contraction queued on another context, so its later work cannot reuse
 * them early; free a stage from the staging context's host thread, after its last
 * ncpa_field2d_contract call.  The eigenvector rows are read by both calls (FP64
 * contraction: by the contraction only) and must stay valid until the contraction
 * completes. */
typedef struct ncpa_field_stage ncpa_field_stage_t;
int ncpa_field2d_stage(ncpa_ctx_t *ctx, ncpa_modes_t *modes, const double *dev_psi_rows, int nrows,
                       double dr, int n_r, int lossless, int precision, ncpa_field_stage_t **stage);
int ncpa_field2d_contract(ncpa_ctx_t *ctx, ncpa_field_stage_t *stage, void *dev_out);
void ncpa_field_stage_free(ncpa_field_stage_t *stage);

/* Contraction variants and their autotuner (no counterpart in the reference).  A
 * precision's contraction has n_variants launch variants (block tile, LDS pipeline depth,
 * assignment of row tiles to the 8 XCDs); within a precision they compute every output
 * element with the same operations in the same order, so they agree to the bit and differ
 * only in speed.  Each context keeps one variant per precision (initially 0), used by
 * ncpa_field2d and ncpa_field2d_contract on that context.  ncpa_field_tune times every
 * variant of the stage's precision on the staged batch (one warm-up and `reps` contractions
 * each, HIP events on ctx's stream, writing dev_out), keeps the fastest for ctx (variant 0,
 * the one with the least HBM traffic, unless another is more than 2 % faster) and returns
 * its index in *best (may be NULL) and the average milliseconds per contraction of the
 * first ms_len variants in ms (may be NULL).  Like ncpa_field2d_contract it may run on a
 * context other than the stage's; it waits for its own timings on the host. */
int ncpa_field_variants(int precision, int *n_variants);
/* The kernel a variant launches (a static string) and its row groups per XCD class (the
 * XCDs split the row tiles into that many groups; 0: the kernel's own block order). */
int ncpa_field_variant_info(int precision, int variant, const char **kernel, int *row_groups);
int ncpa_field_set_variant(ncpa_ctx_t *ctx, int precision, int variant);
int ncpa_field_get_variant(ncpa_ctx_t *ctx, int precision, int *variant);
int ncpa_field_tune(ncpa_ctx_t *ctx, ncpa_field_stage_t *stage, void *dev_out, int reps, int *best, double *ms,
                    int ms_len);

/* Transmission-loss agreement of two fields of n_points complex values (device buffers laid
 * out as ncpa_field2d writes them): dev_ref complex128, dev_test complex128 if test_precision
 * is NCPA_PREC_FP64 and complex64 otherwise.  Over the points with 10 log10 |p_ref|^2 >
 * floor_db: *max_db = max |TL_test - TL_ref|, *rms_db its root mean square (optional),
 * *n_counted the number of such points (optional).  The reference has no counterpart: it is
 * the check that a reduced-precision contraction stays within its stated dB error (TL as
 * written by ModeSolver::getTLoss2D, ModeSolver.cpp:524-544). */
int ncpa_field_compare(ncpa_ctx_t *ctx, const double *dev_ref, const void *dev_test, int test_precision,
                       size_t n_points, double floor_db, double *max_db, double *rms_db, size_t *n_counted);

/* Drop-in for EigenEngine::doESSCalculation (EigenEngine.h:13-14): host diag[Nz]
 * (the (omega/c_eff)^2 trace with the ground term), all eigenpairs with k^2 in
 * [k_min^2, k_max^2]; k2[nconv] ascending; v row-major v[j*ldv + mode] = x_j/sqrt(dz)
 * (sum psi^2 dz = 1).  tol is accepted for signature compatibility (the solver converges
 * to working precision).  Returns NCPA_ERR_MAX_MODES if more than max_modes modes. */
int ncpa_ess_eigen(ncpa_ctx_t *ctx, const double *diag, int Nz, double dz, double tol,
                   const double *k_min, const double *k_max, int max_modes, int *nconv,
                   double *k2, double *v, int ldv);

/* ---- wide-angle, high-Mach modes (WMod, ModBB --method wmod) ---------------------- */

/* Solve a batch with the wide-angle solver: WModeSolver::getModalTrace +
 * EigenEngine::doWideAngleCalculation + doSelect + doPerturb (src/libmodes/WModeSolver.cpp:
 * 322-380, 491-629; EigenEngine.cpp:174-395; ModeSolver.cpp:263-284).  `grid` is the
 * setParams grid of ncpa_grid_make (profile samples every (maxheight - z_min)/(Nz - 1));
 * the matrix, the ground term, the WKB integral, the perturbation and the source/receiver
 * rows use dz = (maxheight - z_min)/Nz as WModeSolver::solve does (:297), and the result's
 * grid carries that dz.  Every real eigenvalue k of the quadratic problem in [k_min, k_max]
 * is returned (ascending), psi normalised as the reference's first half of a unit-norm
 * linearised eigenvector: sum psi^2 dz = 1/(1 + k^2).  The modes serve ncpa_modes_*,
 * ncpa_tl1d, ncpa_modes_rows and ncpa_field2d exactly as those of ncpa_ess_solve. */
int ncpa_wmod_solve(ncpa_ctx_t *ctx, const ncpa_grid_t *grid, const ncpa_problem_t *problems,
                    int n_problems, ncpa_modes_t **out);

/* Drop-in for EigenEngine::doWideAngleCalculation + WModeSolver::doSelect
 * (EigenEngine.cpp:174-395, WModeSolver.cpp:491-506): kd, md, cd as getModalTrace fills them
 * (kd[0] including the ground term).  Returns every real eigenvalue of
 * (k^2 diag(md) + k diag(cd) + D) v = 0 in [k_min, k_max], kH[nconv] ascending, and
 * v[j*ldv + mode] = x_j/sqrt(dz) for the unit-norm linearised eigenvector x = (v, k v).
 * tol is accepted for signature compatibility.  NCPA_ERR_MAX_MODES above max_modes. */
int ncpa_wmod_eigen(ncpa_ctx_t *ctx, const double *kd, const double *md, const double *cd, int Nz,
                    double dz, double tol, double k_min, double k_max, int max_modes, int *nconv,
                    double *kH, double *v, int ldv);

/* ---- dispersion files (host only: no device, no context) -------------------------- */

/* The records ModeSolver::writeDispersion writes (src/libmodes/ModeSolver.cpp:569-589), one
 * line per record: "%.7e   %d    %.7e   %.7e" f, n, rho(z_src), rho(z_rcv), then per mode
 * "   %.12e   %.12e" Re k, Im k and "   %.12e   %.12e" psi(z_src), psi(z_rcv).  Record r's
 * modes are entries [sum count[0..r), + count[r]) of the mode arrays.  Formatted on
 * n_threads threads (<= 0: the host's cores) with the same bytes as printf, written with
 * one fwrite; append != 0 opens the file "a", else "w". */
int ncpa_dispersion_write(const char *path, int append, int n_records, const double *f, const int *count,
                          const double *rho_s, const double *rho_r, const double *k_re,
                          const double *k_im, const double *psi_s, const double *psi_r,
                          int n_threads);

/* Read a dispersion file as ModalBroadbandPropagator::read_dispersion_file does
 * (src/libbroadband/ModalBroadbandPropagator.cpp:98-138): whitespace-separated numbers,
 * records while a frequency reads.  The file is memory-mapped and its numbers parsed on
 * n_threads threads.  NCPA_ERR_ARG: unreadable file, or a record cut short / a mode count
 * that is not a whole number in [0, NCPA_MAX_MODES] (the reference's format error). */
typedef struct ncpa_dispersion ncpa_dispersion_t;
int ncpa_dispersion_read(const char *path, int n_threads, ncpa_dispersion_t **out);
int ncpa_dispersion_info(const ncpa_dispersion_t *d, int *n_records, size_t *n_modes);
/* Copies of the records (any pointer may be NULL): f, count, rho_s, rho_r [n_records];
 * k_re, k_im, psi_s, psi_r [n_modes], file order. */
int ncpa_dispersion_arrays(const ncpa_dispersion_t *d, double *f, int *count, double *rho_s,
                           double *rho_r, double *k_re, double *k_im, double *psi_s, double *psi_r);
void ncpa_dispersion_free(ncpa_dispersion_t *d);

/* ---- broadband pulse synthesis (ModBB --propagation) ------------------------------ */

/* Batched complex FFT of `batch` contiguous transforms of n points (interleaved re/im
 * doubles, device).  sign -1 = FFTW_FORWARD, +1 = FFTW_BACKWARD, both unnormalised
 * (the plans of BroadbandPropagator.cpp:342-345, 365-368, 453-455, 639-641). */
int ncpa_fft_batch(ncpa_ctx_t *ctx, const double *dev_in, double *dev_out, int n, int batch, int sign);

/* k_pert = sqrt(k^2 + i A) (ModeSolver.cpp:277-282) of n device-resident modes, e.g. the
 * k2 / absorption views of ncpa_modes_device, into device arrays dev_kre, dev_kim. */
int ncpa_kpert(ncpa_ctx_t *ctx, const double *dev_k2, const double *dev_absorption, int n,
               double *dev_kre, double *dev_kim);

/* Transfer functions of n_sets dispersion sets (e.g. azimuths) at n_ranges ranges:
 * ModalBroadbandPropagator::compute_modal_sum (ModalBroadbandPropagator.cpp:245-277), i.e.
 * H = i e^{-i pi/4}/sqrt(8 pi r) sqrt(rho_ratio) sum_j psi_S psi_R e^{i k r}/sqrt(k) with the
 * half-Hann taper over the top floor(0.1 n_freq) bins.  Modes of (set s, bin f) are the
 * device entries [mode_begin[s*n_freq+f], + mode_count[...]) of dev_kre, dev_kim,
 * dev_psi_sr (host index arrays; count 0 for the DC bin).  rho_ratio[s] = rho_r/rho_s.
 * zero_attn != 0 takes Im k = 0.  dev_H: complex (interleaved) [n_sets][n_ranges][n_freq]. */
int ncpa_bb_transfer(ncpa_ctx_t *ctx, int n_sets, int n_freq, const int *mode_begin,
                     const int *mode_count, const double *dev_kre, const double *dev_kim,
                     const double *dev_psi_sr, int zero_attn, const double *ranges_m,
                     int n_ranges, const double *rho_ratio, double *dev_H);

/* Propagated waveforms from transfer functions: BroadbandPropagator::fft_pulse_prop
 * (BroadbandPropagator.cpp:573-651) and the factor 2 of ModalBroadbandPropagator.cpp:223-226.
 * spectrum: host complex S[n_freq] (interleaved); f_vec: host frequencies with f_vec[0] the
 * DC bin; t0 = range / max_celerity; df = f_vec[1] - f_vec[0].  dev_wave:
 * [n_sets][n_ranges][nfft] doubles = 2 Re(df FFTW_FORWARD(e^{-2 pi i f t0} S H)). */
int ncpa_bb_pulse(ncpa_ctx_t *ctx, int n_sets, int n_ranges, int n_freq, const double *dev_H,
                  const double *spectrum, const double *f_vec, const double *ranges_m,
                  double max_celerity, int nfft, double *dev_wave);

/* Allocation-free form of ncpa_bb_pulse: the same arguments, kernels and results, with
 * every intermediate (the uploaded spectrum, frequencies and delays, the FFT input and
 * output, the FFT's global work buffer when its ping-pong buffers do not fit in LDS) in one
 * fixed device array of NCPA_PULSE_SCRATCH_BYTES per GPU, which the library holds from its
 * first use on that GPU until the process ends.  It calls no hipMalloc, hipFree or other
 * allocating routine.  No counterpart in the reference (its synthesis allocates per call):
 * for callers that cannot allocate device memory, such as code under HIP graph capture,
 * tight serving loops and harnesses that forbid hipMalloc.  A job whose intermediates
 * (8 bytes x (2 n_freq + n_freq + n_ranges + 4 n_sets n_ranges nfft [+ 4 nfft n_sets n_ranges
 * when 2 nfft complex doubles exceed the LDS of a block]), each part rounded up to 256 bytes)
 * exceed the array returns NCPA_ERR_ARG with nothing launched.  Every call on one GPU uses
 * that GPU's one array: calls on one context are ordered by its stream, but calls that
 * overlap on one GPU (other contexts, streams or host threads) race. */
#define NCPA_PULSE_SCRATCH_BYTES (32u << 20)
int ncpa_bb_pulse_noalloc(ncpa_ctx_t *ctx, int n_sets, int n_ranges, int n_freq, const double *dev_H,
                          const double *spectrum, const double *f_vec, const double *ranges_m,
                          double max_celerity, int nfft, double *dev_wave);

/* Source types of BroadbandPropagator::get_source_spectrum (BroadbandPropagator.cpp:244-417) */
#define NCPA_SRC_IMPULSE  0
#define NCPA_SRC_PULSE1   1
#define NCPA_SRC_PULSE2   2
#define NCPA_SRC_SPECTRUM 3
#define NCPA_SRC_WAVEFORM 4

/* Source spectrum S[n_freq] at f_vec (f_vec[0] = 0, f_step = f_vec[2] - f_vec[1]) and the
 * time-domain source pulse[nfft] (complex, before the factor the reference applies when it
 * writes source_waveform_input.dat); both host arrays, pulse may be NULL.
 * f_center < 0 selects f_max/5 (pulse1, pulse2).  input: for SPECTRUM, n_input rows of
 * (f, Re S, Im S); for WAVEFORM, n_input rows of (t, amplitude); else NULL. */
int ncpa_bb_source(ncpa_ctx_t *ctx, int kind, const double *f_vec, int n_freq, int nfft,
                   double f_center, const double *input, int n_input, double *spectrum,
                   double *pulse);

/* ------------------------------------------------------------------------------------------
 * Broadband job plans.  The batched form of ModBB --dispersion followed by --propagation
 * (ModBB_main.cpp:33-121: one ESSModeSolver::solve per frequency, then
 * ModalBroadbandPropagator::calculate_waveform, ModalBroadbandPropagator.cpp:140-243, over
 * the dispersion file).  Problem p = s (n_freq - 1) + i - 1 is set s (e.g. an azimuth) at
 * f_vec[i]; every eigenvalue of every problem is found in one persistent launch and the
 * waveforms are synthesised without a count on the host, so the whole job is captured once
 * as a HIP graph and replayed by ncpa_bb_plan_run.  Creation runs the batch's interval
 * counts once to size the mode buffers and, with tune != 0, times the eigenvalue kernel's
 * variants on the batch and keeps the fastest: multisection passes 1-3 x launch form (a
 * persistent grid of 4, 8 or 15 blocks per CU pulling work from per-XCD queues, or one
 * block per work-list slot with the diagonal read through LDS or by scalar loads;
 * grid_blocks = 0 and -1 report these two).
 * ------------------------------------------------------------------------------------------ */
typedef struct ncpa_bb_plan ncpa_bb_plan_t;

typedef struct {
    int n_sets;               /* dispersion sets (e.g. azimuths) */
    int n_freq;               /* frequency bins including DC: f_vec[0] = 0 */
    const double *f_vec;      /* host [n_freq] */
    const double *spectrum;   /* host complex S[n_freq], interleaved (ncpa_bb_source) */
    const double *ranges_m;   /* host [n_ranges], > 0 */
    int n_ranges;
    const double *rho_ratio;  /* host [n_sets]: rho_r / rho_s */
    double max_celerity;      /* t0 = range / max_celerity */
    int nfft;                 /* >= n_freq (the reference's rule: max(nfft, 4 n_freq)) */
    int zero_attn;            /* != 0: Im k = 0 in the transfer functions */
    int max_modes_total;      /* mode capacity; 0: 1.5x the modes the batch has at creation */
    int tune;                 /* != 0: autotune the eigenvalue kernel on this batch */
} ncpa_bb_job_t;

int ncpa_bb_plan_create(ncpa_ctx_t *ctx, const ncpa_grid_t *grid, const ncpa_problem_t *problems,
                        const ncpa_bb_job_t *job, ncpa_bb_plan_t **plan);
/* New problem records (same count, grid and job) for the next runs, e.g. other profiles;
 * a batch with more modes than the capacity makes ncpa_bb_plan_result fail. */
int ncpa_bb_plan_set_problems(ncpa_bb_plan_t *plan, const ncpa_problem_t *problems);
/* Enqueues one replay of the job's graph on the context stream (returns at once). */
int ncpa_bb_plan_run(ncpa_bb_plan_t *plan);
/* Waits for the last run and returns its device results (owned by the plan):
 * dev_wave [n_sets][n_ranges][nfft] = 2 Re(df FFT(e^{-2 pi i f t0} S H)) as ncpa_bb_pulse;
 * k_pert (dev_kre, dev_kim), psi_S psi_R (dev_psi_sr) and mode offsets (dev_offsets,
 * [n_problems + 1]) of every mode, problem-major, ascending k.  Any pointer may be NULL.
 * NCPA_ERR_MAX_MODES if a problem or the batch exceeded its limit (nothing computed). */
int ncpa_bb_plan_result(ncpa_bb_plan_t *plan, const double **dev_wave, const double **dev_kre,
                        const double **dev_kim, const double **dev_psi_sr, const int **dev_offsets,
                        int *total_modes);
/* The configuration in use and the tuner's measurements: log holds (passes, blocks, ms)
 * triples (blocks 0 / -1: one block per work-list slot, LDS / scalar loads); *n_log in:
 * capacity in triples, out:
 * triples measured. */
int ncpa_bb_plan_tuning(ncpa_bb_plan_t *plan, int *ms_passes, int *grid_blocks, double *modes_ms,
                        double *log, int *n_log);
void ncpa_bb_plan_destroy(ncpa_bb_plan_t *plan);

#ifdef __cplusplus
}
#endif
#endif /* NCPAMODES_H */
