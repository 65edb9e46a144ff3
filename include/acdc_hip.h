/*
 * acdc_hip.h -- C ABI of the MI355X (gfx950) PCDM hot path.
 *
 * Parallel randomized coordinate descent with ESO step scaling for
 *   loss 0  (ACDC_LOSS_L1_SQUARE)      min 0.5||Ax-b||^2 + lambda||x||_1
 *   loss 2  (ACDC_LOSS_SVM_DUAL_HINGE) hinge-SVM dual, box-clipped SDCA
 *   loss 10 (ACDC_LOSS_NNLS)           min_{x>=0} 0.5||Bx-f||^2 (TTD)
 * Loss ids follow the reference's settings->lossFunction switch
 * (cpp/src/solver/MulticoreEngineExecutor.h:32-58); 10 is new.
 *
 * Conventions: every entry point returns an acdc_status; nothing throws
 * across this boundary.  Host arrays belong to the caller and are copied at
 * create; device memory belongs to the library.  Handles are independent (no
 * globals besides the thread-local error text), calls on one handle must be
 * serialised by the caller.  The residual sign convention is r = Ax - b
 * (cpp/src/solver/lossFunctions/squareLoss/MulticoreSquareLoss.h:42-53).
 * Sizes are int64_t, indices int32_t, values double.
 */
#ifndef ACDC_HIP_H
#define ACDC_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum acdc_status {
  ACDC_OK = 0,
  ACDC_ERR_INVALID = 1,  /* bad argument or inconsistent sizes */
  ACDC_ERR_HIP = 2,      /* HIP runtime error (text in acdc_last_error) */
  ACDC_ERR_NOMEM = 3,    /* host or device allocation failed */
  ACDC_ERR_STATE = 4,    /* call not valid for this handle (e.g. wrong loss) */
  ACDC_ERR_NO_DEVICE = 5 /* no usable gfx950 device */
} acdc_status;

typedef enum acdc_loss {
  ACDC_LOSS_L1_SQUARE = 0,
  ACDC_LOSS_SVM_DUAL_HINGE = 2,
  ACDC_LOSS_NNLS = 10
} acdc_loss;

typedef enum acdc_precision {
  ACDC_PREC_FP64 = 0,        /* plain FP64 everywhere */
  ACDC_PREC_FP32_REFRESH = 1 /* FP32 values/residual/dots/scatter; FP64 x,
                                Lipschitz, periodic r = Ax-b refresh, gap */
} acdc_precision;

typedef enum acdc_beta_mode {
  ACDC_BETA_OMEGA = 0,  /* beta = 1+(omega-1)(tau-1)/(n-1), omega = max nnz
                           per row: MultiCoreSolver.cpp:32-38, :108-110 */
  ACDC_BETA_FORCED = 1, /* beta = forced_beta (OptimizationSettings
                           ::forcedSigma, settingsAndStatistics.h:53) */
  ACDC_BETA_NONE = 2,   /* beta = 1, the multicore Lasso's L_j = 1/||a_j||^2
                           (MulticoreSquareLoss.h:27-40) */
  ACDC_BETA_SPECTRAL = 3 /* beta = 1+(tau-1)(lmax-1)/(n-1), lmax the largest
                            eigenvalue of D^-1/2 A^T A D^-1/2 (segments as
                            columns, D = diag ||a_c||^2) by 20 power steps
                            from a start seeded by `seed`: mSDCA.cpp:188-265,
                            :286 */
} acdc_beta_mode;

typedef enum acdc_step_mode {
  ACDC_STEP_SYNC = 0, /* tau distinct coordinates per step, all deltas from
                         one state then applied (cdn_common.h:235-282);
                         tau = 1 is the serial method (loss_square.h:72-83) */
  ACDC_STEP_ASYNC = 1 /* tau unsynchronised workers, each drawing coordinates
                         with replacement from its own stream and applying
                         its updates with atomic adds: the multicore executor
                         (MulticoreEngineExecutor.h:83-89 with
                         parallel_essentials_posix.h:62-68).  A "step" is one
                         update per worker; tau = 1 is the serial method. */
} acdc_step_mode;

/* Solver settings; the reference counterpart is OptimizationSettings
 * (cpp/src/solver/settingsAndStatistics.h:11-93) plus ProblemData::lambda
 * and ::sigma (structures.h:30-97). */
typedef struct acdc_settings {
  double lambda;           /* ProblemData::lambda; SVM: lambda of the primal */
  int64_t tau;             /* coordinates per step (totalThreads, minibatch b) */
  uint64_t seed;           /* key of the Philox4x32-10 coordinate stream */
  int32_t beta_mode;       /* acdc_beta_mode */
  double forced_beta;      /* used when beta_mode == ACDC_BETA_FORCED */
  int32_t precision;       /* acdc_precision */
  int32_t step_mode;       /* acdc_step_mode */
  int32_t lanes_per_coord; /* 0 = auto, else 8, 16, 32 or 64 lanes per
                              sampled coordinate, or 256 (a whole block per
                              coordinate: synchronous steps, per-step or
                              persistent, for long segments; auto picks it
                              when segments average over 192 entries) */
  int32_t refresh_period;  /* FP32 path: FP64 residual refresh every this
                              many epochs of ceil(n/tau) steps inside
                              acdc_run_updates (0 = off) */
  int32_t launch_mode;     /* acdc_launch_mode: how the steps execute */
} acdc_settings;

/* Execution of synchronous steps with tau > 1 (modes 0-3), or of
 * asynchronous workers (modes 0, 4, 5); the arithmetic is the same. */
typedef enum acdc_launch_mode {
  ACDC_LAUNCH_AUTO = 0,      /* persistent cooperative grid when all tau slots
                                fit on the device at once, else per step,
                                with the streamed scatter when tau >= n/16
                                and it applies (see below) */
  ACDC_LAUNCH_PER_STEP = 1,  /* one gather + one scatter launch per step */
  ACDC_LAUNCH_PERSISTENT = 2, /* persistent grid; fails if it does not fit */
  ACDC_LAUNCH_PER_STEP_STREAMED = 3 /* per step, the scatter as one coalesced
                                pass over a transposed copy of the data (each
                                block owns a range of residual indices and
                                sums into LDS) instead of one atomic per
                                sampled entry: for steps that cover a large
                                share of the entries (SVM minibatches); needs
                                n <= 65535 and n * sizeof(T) <= 96 KiB (FP32:
                                n <= 24576, FP64: n <= 12288).  Same
                                arithmetic as ACDC_LAUNCH_PER_STEP up to the
                                order of the additions into the residual */,
  /* asynchronous workers (ACDC_STEP_ASYNC), same semantics either way: */
  ACDC_LAUNCH_ASYNC_DIRECT = 4, /* each worker loads its own records and
                                entries (any lanes_per_coord) */
  ACDC_LAUNCH_ASYNC_STAGED = 5  /* one worker per 64-lane wave; a loader wave
                                per 8 workers copies the coming updates'
                                records and entries into LDS ahead of use,
                                so a worker waits only on its residual
                                gathers (lanes_per_coord 0 or 64) */
} acdc_launch_mode;

/* One evaluation, all in FP64.  Counterparts: SquareLoss::getResultLogRow
 * (squareLoss/SquareLoss.h:28-38) and SVMDualLoss::getResultLogRow
 * (SVMDualLoss/SVMDualLoss.h:31-42). */
typedef struct acdc_report {
  double primal;    /* Lasso 0.5||r||^2+lambda||x||_1; NNLS 0.5||r||^2;
                       SVM (lambda/2)||w||^2 + mean hinge */
  double dual;      /* Lasso D(theta), theta = -r min(1,lambda/||A^T r||_inf);
                       SVM -(lambda/2)||w||^2 + mean alpha; NNLS: primal - xg */
  double gap;       /* primal - dual */
  double rel_gap;   /* gap / |primal| */
  double grad_inf;  /* ||A^T r||_inf (Lasso, NNLS); 0 for SVM */
  double x_l1;      /* ||x||_1 (sum alpha for SVM) */
  double res_sq;    /* ||r||^2 (||w||^2 for SVM) */
  double kkt;       /* NNLS max_j |min(x_j, g_j)|; 0 otherwise */
  double accuracy;  /* SVM training 0-1 accuracy; 0 otherwise */
  double nnz_x;     /* number of nonzero coordinates */
  uint64_t updates; /* coordinate updates applied since create */
  double epochs;    /* updates / n */
  double kernel_ms; /* device time inside acdc_run_updates so far (update
                       kernels and the refreshes they trigger) */
  double eval_ms;   /* device time of explicit refreshes and reports so far */
  double step_ms;   /* device time of the update kernels alone so far */
  uint64_t step_launches; /* update-kernel launches so far (two per step with
                             per-step launches) */
} acdc_report;

typedef struct acdc_problem acdc_problem;
typedef struct acdc_solver acdc_solver;

/* Text of the last error on this host thread ("" if none). */
const char *acdc_last_error(void);

/* Number of visible gfx950 devices (0 when none). */
acdc_status acdc_device_count(int32_t *count);

/* Column-major data A (m x n) and b (m) for losses 0 and 10.
 * Replaces ProblemData<L,D>::A_csc_* + b as filled by
 * InputOuputHelper::loadCSCData (cpp/src/utils/file_reader.h:108-137).
 * Row indices must lie in [0,m); duplicates within a column are allowed. */
acdc_status acdc_problem_create_csc(int64_t m, int64_t n, int64_t nnz,
                                    const int32_t *col_ptr,
                                    const int32_t *row_idx, const double *vals,
                                    const double *b, acdc_problem **out);

/* Sample-major data for loss 2: n_samples rows of n_features, labels +-1.
 * Replaces ProblemData::A_csr_* + b as filled by
 * loadDistributedSparseSVMRowData (utils/distributed_instances_loader.h:68-193),
 * where n = samples and m = features. */
acdc_status acdc_problem_create_csr(int64_t n_samples, int64_t n_features,
                                    int64_t nnz, const int32_t *row_ptr,
                                    const int32_t *col_idx, const double *vals,
                                    const double *labels, acdc_problem **out);

acdc_status acdc_problem_destroy(acdc_problem *p);

/* omega = max nnz per row (CSC) or max samples per feature (CSR):
 * MultiCoreSolver.cpp:32-38 and :55-62. */
acdc_status acdc_problem_omega(const acdc_problem *p, int32_t *omega);

/* Builds the device data, sets x = 0, r = Ax - b (MulticoreEngineExecutor
 * constructor, MulticoreEngineExecutor.h:25-66: x zeroed :29-30,
 * recomputeResiduals :59, computeReciprocalLipConstants :60). */
acdc_status acdc_solver_create(const acdc_problem *p, int32_t loss,
                               const acdc_settings *s, acdc_solver **out);
acdc_status acdc_solver_destroy(acdc_solver *s);

/* Runs ceil(n_updates / tau) steps of tau coordinates.
 * Replaces AbstractEngineExecutor::executeBulkOfIterations
 * (MulticoreEngineExecutor.h:83-89) with innerIterations = n_updates.
 * Returns once the work is queued on the solver's stream (persistent
 * launches: once it has run), so consecutive calls run back to back; the
 * report, the getters and acdc_synchronize wait for it.  A device error in
 * queued updates is therefore returned by the next call that waits for the
 * stream (acdc_synchronize, acdc_report_get, acdc_get_x, ...), under that
 * call's name. */
acdc_status acdc_run_updates(acdc_solver *s, uint64_t n_updates);

/* r = Ax - b in FP64 (SVM: w = (1/(lambda n)) sum alpha_i y_i x_i).
 * Replaces AbstractLoss::recomputeResiduals (MulticoreSquareLoss.h:42-53,
 * MulticoreSVMDualLoss.h:31-46). */
acdc_status acdc_recompute_residuals(acdc_solver *s);

/* FP64 primal/dual/gap.  Replaces AbstractLoss::computeObjectiveValue
 * (MulticoreL1RegSquareLoss.h:30-39, MulticoreSVMDualLoss.h:63-117).
 * On the FP32 path the residual is first recomputed in FP64 (and written
 * back to the FP32 residual); on the FP64 path the running residual is used. */
acdc_status acdc_report_get(acdc_solver *s, acdc_report *out);

/* x (alpha for SVM), length n.  acdc_set_x is the warm start of
 * lasso_console_solver.cpp:153-172 and recomputes the residual. */
acdc_status acdc_get_x(acdc_solver *s, double *x, int64_t n);
acdc_status acdc_set_x(acdc_solver *s, const double *x, int64_t n);

/* Running residual r (Lasso/NNLS, length m) or w (SVM, n_features). */
acdc_status acdc_get_residuals(acdc_solver *s, double *r, int64_t len);

/* Reciprocal Lipschitz constants L_j = 1/(beta ||a_j||^2) (ProblemData::Li). */
acdc_status acdc_get_lipschitz(acdc_solver *s, double *Li, int64_t n);

/* lambda for the next updates (lambda path, lasso_console_solver.cpp:149-161);
 * SVM: also rescales w for the new 1/(lambda n). */
acdc_status acdc_set_lambda(acdc_solver *s, double lambda);

/* ESO beta in use (ProblemData::sigma, MultiCoreSolver.cpp:108-110). */
acdc_status acdc_get_beta(const acdc_solver *s, double *beta);

/* Coordinates the device samples for steps [step0, step0+nsteps): out has
 * nsteps*tau entries, row-major (ACDC_STEP_ASYNC: entry [s][w] is update
 * step0+s of worker w).  Used to replay the stream on the host. */
acdc_status acdc_sample_coordinates(acdc_solver *s, uint64_t step0,
                                    uint64_t nsteps, int32_t *out);

/* Runs nsteps (1..4096) synchronous steps (tau >= 2) with HIP events around
 * each kernel and returns the mean device time per launch of the gather
 * (phase A) and scatter (phase B) kernels in milliseconds.  The steps count
 * as regular updates.  Counterpart of the reference's per-launch cudaEvent
 * timing (oldSourceFiles/cuda_hso/hso/solver/parallel/randomL2L1Solver.h:169-183). */
acdc_status acdc_profile_step_kernels(acdc_solver *s, uint64_t nsteps,
                                      double *gather_ms, double *scatter_ms);

/* acdc_sample_coordinates without any allocation: the same coordinates,
 * staged in a fixed device array of ACDC_NOALLOC_MAX_SAMPLES int32 entries
 * (4 MiB, held by each device the library runs on until the process ends)
 * instead of a buffer allocated per call.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  nsteps * tau
 * above the maximum is ACDC_ERR_INVALID.  Every call on one device shares
 * the array, so calls that overlap on one device (other host threads, other
 * solvers' streams) race; calls on one solver are serialised by the caller. */
#define ACDC_NOALLOC_MAX_SAMPLES (1 << 20)
acdc_status acdc_sample_coordinates_noalloc(acdc_solver *s, uint64_t step0, uint64_t nsteps,
                                            int32_t *out);

/* Waits for all work queued on the solver's stream. */
acdc_status acdc_synchronize(acdc_solver *s);

/* ---- batched Lasso / NNLS / SVM: many instances x warm-started lambda paths
 * One persistent launch solves every instance along its lambda path: lambda
 * l+1 starts from the x of lambda l (lasso_console_solver.cpp:149-172), and
 * each (instance, lambda) stops at FP64 relative duality gap <= gap_tol
 * (checked with an FP64 refresh every refresh_period epochs, mSDCA.cpp:123
 * stops on the gap the same way) or after max_epochs.  Instances are handed
 * to workgroups by a device work queue, largest first.  Instance i samples
 * with key seed + i and by default tau = 2 * (1024 / lanes_per_coord)
 * coordinates per synchronous step, or 1024 / lanes_per_coord asynchronous
 * workers (capped at its n); its ESO beta comes from its omega.  No
 * counterpart in the reference beyond the lambda-path driver it generalises. */
typedef struct acdc_batch acdc_batch;

typedef struct acdc_batch_settings {
  uint64_t seed;
  int32_t precision;       /* ACDC_PREC_FP32_REFRESH or ACDC_PREC_FP64 */
  int32_t lanes_per_coord; /* 0 = auto, else 8, 16, 32 or 64 */
  int32_t refresh_period;  /* epochs between FP64 refresh + gap checks (>= 1) */
  int32_t max_epochs;      /* per lambda */
  double gap_tol;          /* relative duality gap target (< 0: always max_epochs) */
  int32_t use_graph;       /* capture the run in a HIP graph at create */
  int32_t step_mode;       /* acdc_step_mode inside each instance: SYNC = tau-nice
                              steps with block barriers (replays the oracle);
                              ASYNC = the block's lane groups as unsynchronised
                              workers with LDS atomics */
  int32_t tau;             /* 0 = the block's capacity (SYNC: 2 slots per lane
                              group; ASYNC: one worker per lane group), else
                              at most that, per instance at most n */
  int32_t queue_chunk;     /* instances a block takes from the work queue per
                              access (0 = 1); consecutive in the queue's
                              largest-first order */
  int32_t loss;            /* of every instance: ACDC_LOSS_L1_SQUARE (0),
                              ACDC_LOSS_NNLS (x >= 0, the lambdas are ignored;
                              primal = 1/2 ||Ax-b||^2, dual = 0, rel_gap = the
                              KKT residual max_j |min(x_j, g_j)| over
                              ||A^T b||_inf) or ACDC_LOSS_SVM_DUAL_HINGE
                              (problems from acdc_problem_create_csr, lambdas
                              > 0; x is alpha; primal and dual as
                              acdc_report's; nnz_x counts alpha_j != 0) */
  int32_t beta_mode;       /* ACDC_BETA_OMEGA (0: 1 + (omega-1)(tau-1)/(n-1)
                              per instance) or ACDC_BETA_SPECTRAL (per
                              instance, from 20 power steps on its normalised
                              Gram matrix at create, as acdc_solver_create
                              does: mSDCA.cpp:188-286) */
  int32_t screening;       /* Lasso instances with asynchronous workers: 1 =
                              gap-safe working sets (at every full FP64
                              refresh, coordinates j with s|<a_j,r>| +
                              ||a_j|| sqrt(2 gap) < lambda, s = min(1,
                              lambda/||A^T r||_inf), are provably zero at
                              the optimum and are set aside; the workers draw
                              uniformly from the rest, with its own ESO beta
                              (omega rule), and cheap refreshes over it check
                              its gap; a full refresh confirms the whole
                              problem's gap before a pair counts as
                              converged).  0 = every coordinate throughout.
                              The optimum is the same; the coordinate
                              sequence is not. */
  int32_t ws_newton_iters; /* screened Lasso instances: at most this many
                              conjugate-gradient iterations for a Newton
                              step on the support of x at each restricted
                              refresh of a working set (the quadratic of the
                              support's sign orthant; the step is projected
                              onto that orthant and kept only when the FP64
                              objective falls); 0 = no such steps, at most
                              ACDC_WS_NEWTON_MAX_ITERS.  It makes the
                              near-interpolating pairs of a lambda path
                              (support near m) converge in a few hundred
                              epochs instead of tens of thousands.  No
                              counterpart in the reference. */
} acdc_batch_settings;

#define ACDC_WS_NEWTON_MAX_ITERS 10000

typedef struct acdc_batch_result {
  double lambda, primal, dual, rel_gap, nnz_x;
  int32_t epochs, converged; /* epochs: updates of this lambda over n, rounded up */
  uint64_t updates; /* coordinate updates of the instance up to this lambda */
  double device_us; /* device time of this lambda (refreshes included) */
  double ws_size;   /* coordinates the workers drew from last (n without screening) */
  int32_t status;   /* 0; ACDC_BATCH_NONFINITE: an evaluation of the instance
                       held a NaN or an infinity (data or iterate), so it was
                       abandoned at this lambda and the rest of its path;
                       the other instances are unaffected */
  int32_t newton_steps; /* support Newton steps kept for this lambda (screened
                           Lasso instances with ws_newton_iters > 0) */
} acdc_batch_result;
#define ACDC_BATCH_NONFINITE 1

acdc_status acdc_batch_create(const acdc_problem *const *problems, int32_t count,
                              const double *lambdas /* count x path_len */,
                              int32_t path_len, const acdc_batch_settings *s,
                              acdc_batch **out);
/* Runs the whole batch from x = 0; *device_ms (may be null) gets the device
 * time of the run. */
acdc_status acdc_batch_run(acdc_batch *b, double *device_ms);
acdc_status acdc_batch_results(acdc_batch *b, acdc_batch_result *out, int64_t len);
acdc_status acdc_batch_get_x(acdc_batch *b, int32_t instance, double *x, int64_t n);
acdc_status acdc_batch_instance_info(const acdc_batch *b, int32_t instance, int64_t *tau,
                                     double *beta, int32_t *lds_resident);
acdc_status acdc_batch_destroy(acdc_batch *b);

/* ---- autotuner (no reference counterpart: the reference fixes tau by the
 * thread count and has no lanes-per-coordinate or queue knobs) ------------
 * acdc_batch_tune times short runs of the batch's largest instance (4
 * epochs at lambda = 0.1 lambda_max, or 1 / n for SVM batches: the loss
 * comes from base->loss) for lanes per coordinate in {8,16,32}
 * and tau in {cap/4, cap/2, cap} (cap: the block's slots or workers for
 * those lanes) and returns *base with the lanes and tau of the best score =
 * log(gap at x = 0 / gap after the trial) per second of device time, and
 * with queue_chunk = 1 when the batch has no more instances than the device
 * has CUs (a block per CU: a larger chunk would only idle blocks), else the
 * chunk in {1, 2, 4, 8} with the least device time over the whole batch at
 * two lambdas of two epochs each.
 * acdc_solver_tune_lanes times two epochs (after a warm-up epoch) at lanes
 * 16, 32 and 64 with the other settings of *base and returns the fastest. */
acdc_status acdc_batch_tune(const acdc_problem *const *problems, int32_t count,
                            const acdc_batch_settings *base, acdc_batch_settings *tuned,
                            double *score);
acdc_status acdc_solver_tune_lanes(const acdc_problem *p, int32_t loss,
                                   const acdc_settings *base, int32_t *lanes,
                                   double *updates_per_s);

#ifdef __cplusplus
}
#endif
#endif /* ACDC_HIP_H */
