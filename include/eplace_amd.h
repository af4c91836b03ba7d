/* eplace_amd.h - C ABI of the MI355X global-placement engine.
 *
 * Drop-in boundary for the reference ePlace placer's global-placement path
 * (Xinyu-lumos/ePlace).  The reference has no FFI: its optimiser holds an
 * EPlacer_2D* and calls C++ methods on it (Optimization/nesterov.hpp:59-415,
 * EPlace/eplace.h:95-200).  Each entry point below replaces one of those
 * methods and is cited next to it.  Plain pointers and sizes only; positions
 * and gradients cross the boundary as interleaved (x, y) float pairs in the
 * reference's stage order (eplace.cpp:190-203, 742-804).
 *
 * All functions return EP_OK (0) or an EP_ERR_* code; ep_last_error() then
 * describes the failure.  A context is bound to one HIP device and one
 * stream and is not thread-safe.
 */
#ifndef EPLACE_AMD_H
#define EPLACE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ep_ctx ep_ctx;
typedef struct ep_ctrl ep_ctrl;

enum {
  EP_OK = 0,
  EP_ERR_ARG = 1,     /* bad argument or size */
  EP_ERR_HIP = 2,     /* HIP runtime error */
  EP_ERR_STATE = 3,   /* call not valid in this state */
  EP_ERR_NUMERIC = 4, /* NaN/inf in a scalar (e.g. Lipschitz 0/0) */
  EP_ERR_NOMEM = 5
};

/* PLACEMENT_STAGE (eplace.h:21-27). */
enum { EP_STAGE_MGP = 0, EP_STAGE_FILLERONLY = 1, EP_STAGE_CGP = 2 };

/* Design arrays; replaces PlaceDB's dbNodes/dbTerminals/dbPins/dbNets/
 * dbSiteRows (PlaceDB/placedb.h:88-93).  Nodes in dbNodes order, terminals
 * in dbTerminals order, pins sorted by net (parser order, parser.cpp:473-680),
 * positions are centres, pin offsets are from the centre
 * (objects.cpp:454-459).  pin_obj < num_nodes names a node, otherwise
 * terminal pin_obj - num_nodes. */
typedef struct ep_design {
  int32_t num_nodes, num_terminals, num_nets, num_rows;
  int64_t num_pins;
  const float *node_w, *node_h, *node_x, *node_y;
  const float *term_w, *term_h, *term_x, *term_y;
  const uint8_t *term_ni;
  const int64_t *net_start; /* [num_nets + 1] */
  const int32_t *pin_obj;   /* [num_pins] */
  const float *pin_dx, *pin_dy;
  const double *row_bottom, *row_height, *row_x0, *row_x1;
  double common_row_height;
} ep_design;

typedef struct ep_params {
  float target_density;     /* EPlacer_2D::setTargetDensity (eplace.cpp:27-31) */
  int32_t bins_x, bins_y;   /* 0: the reference's bin rule (eplace.cpp:237-261) */
  int32_t wirelength_model; /* 0 WA (default), 1 LSE (-LSE, eplace.cpp:512-530) */
  int32_t exp_mode;         /* 0 reference (1+a/1024)^1024 (global.h:457-471), 1 exp */
  uint32_t filler_seed;     /* rand() seed for filler placement; 1 = reference */
  const float *filler_xy;   /* optional filler centres, overrides the seed */
  int32_t filler_count;     /* -1: reference count (eplace.cpp:167) */
  int32_t poisson_mode;     /* 0 FP32 MFMA; 1 split-BF16, 2 planes, 3 products; 2 split-FP16
                               with power-of-two operand scales, 3 products; 3 split-BF16,
                               3 planes, 6 products (all accumulate in FP32) */
  int32_t device;           /* HIP device ordinal */
} ep_params;

typedef struct ep_info {
  int64_t num_nodes, num_fillers, num_objects, num_terminals, num_nets, num_pins;
  int32_t bins_x, bins_y, fixed_shift, num_big_objects, num_large_nets, pad;
  double bin_step_x, bin_step_y, core_llx, core_lly, core_urx, core_ury;
  double filler_w, filler_h, std_cell_area, macro_area, total_row_area;
} ep_info;

/* Scalars the reference keeps as EPlacer_2D fields (eplace.h:128-136). */
typedef struct ep_stats {
  double overflow;      /* globalDensityOverflow (tau) of the last evaluation */
  double lambda;        /* penalty factor */
  double last_hpwl;     /* lastHPWL */
  double hpwl;          /* HPWL at the last evaluated positions */
  double inv_gamma_x, inv_gamma_y;
  double wirelength;    /* WA (or LSE) value of the last evaluation */
  int32_t stage;
  int32_t mgp_iterations;
} ep_stats;

/* Columns of one optimiser log row (written at the end of iteration k). */
enum {
  EP_LOG_ITER = 0,          /* k */
  EP_LOG_TAU = 1,           /* overflow at the accepted point y_{k+1} (vanilla: at y_k) */
  EP_LOG_LAMBDA = 2,        /* lambda after the update of iteration k (vanilla: of iteration
                               k-1; its update at y_{k+1} runs inside the next evaluation) */
  EP_LOG_HPWL = 3,          /* HPWL at y_{k+1} (vanilla: at y_k) */
  EP_LOG_STEP = 4,          /* 1/L' measured at the accepted trial (backtracking) */
  EP_LOG_INV_GAMMA = 5,     /* 1/gamma_x */
  EP_LOG_EVALS = 6,         /* full evaluations so far */
  EP_LOG_BACKTRACKS = 7,    /* rejected trials so far */
  EP_LOG_FEATURES = 8,      /* EP_CTRL_FEATURES controller inputs at y_k (backtracking/learned) */
  EP_LOG_STEP_PROP = 20,    /* plain step proposal 1/L at iteration k */
  EP_LOG_STEP_USED = 21,    /* step of the first trial (the controller's under method 2) */
  EP_LOG_LAMBDA_BEFORE = 22,/* lambda used by the evaluations of iteration k */
  EP_LOG_FALLBACKS = 23,    /* controller steps rejected so far */
  EP_LOG_COLS = 24
};

/* Optimiser settings (nesterov.hpp:86-149, 231-415). */
typedef struct ep_opt_params {
  int32_t method;           /* 0 vanilla (default), 1 backtracking (-bktrk),
                               2 learned: backtracking whose first trial step and
                               a floor under the penalty multiplier come from
                               `controller`; a first trial that fails the
                               backtracking acceptance test continues as plain
                               backtracking; 3 Barzilai-Borwein (-bb): one move per
                               iteration with the short BB step s.y/y.y when positive,
                               else min(|s|/|y|, 1/L) (nesterov.hpp:231-302) */
  float target_overflow;    /* mGP stop threshold (default 0.10) */
  int32_t max_iterations;   /* MAX_ITERATION (1000); stop when iter > this */
  int32_t max_backtracks;   /* cap on trial steps per iteration, the last taken when every
                               one fails the acceptance test (30; reference: no cap) */
  int32_t sync_every;       /* host polls the stop flag every this many iterations */
  int32_t log_capacity;     /* rows available in log (may be 0) */
  double *log;              /* optional [log_capacity][EP_LOG_COLS], columns EP_LOG_* */
  double *phase_seconds;    /* optional [9]: summed time of the full evaluations per phase
                               (scatter, density finalise, wiring kernels, node sums of the pin
                               gradients, poisson, field gather, combine), the transform GEMM
                               launches alone (inside poisson), and the number of evaluations
                               timed; HIP events on the streams the phases run on.  Each timed
                               evaluation is launched kernel by kernel behind a 0.4 ms hold
                               kernel, so the host has enqueued all of it before the device
                               starts it; iterates are unchanged, the run's wall time is not */
  const ep_ctrl *controller; /* method 2: the step / lambda controller (same device) */
} ep_opt_params;

typedef struct ep_run_stats {
  int32_t iterations;       /* accepted Nesterov iterations */
  int32_t evaluations;      /* full gradient evaluations */
  int32_t backtracks;
  int32_t stopped;          /* 1 if the stop condition was met */
  double final_overflow;
  double final_hpwl;        /* HPWL after wrap_up (main solution written back) */
  double seconds;           /* wall time of the loop */
  int32_t controller_fallbacks; /* method 2: first trials that failed the acceptance test */
  int32_t bb_short_steps;   /* method 3: iterations that took the short BB step */
  double gemm_seconds;      /* summed duration of the transform GEMM launches of the run's
                               evaluations, each from its first block's start to its last
                               wave's end on the device's wall clock (the last evaluation's
                               launches are not included) */
  int32_t gemm_launches;    /* launches summed in gemm_seconds */
  int32_t numeric_fallbacks; /* steps or acceptance tests whose 1/L was not finite and
                                positive (identical consecutive positions): the previous
                                step was kept, or the trial accepted (the reference would
                                carry the NaN/inf into the positions) */
} ep_run_stats;

/* Which device array ep_read/ep_write address (parity and debugging). */
enum {
  EP_ARR_POSITIONS = 0,     /* float [O][2] centres, all objects (nodes then fillers) */
  EP_ARR_NODE_Q = 1,        /* int64 [nx*ny] fixed-point node density */
  EP_ARR_FILLER_Q = 2,      /* int64 [nx*ny] fixed-point filler density */
  EP_ARR_RHO = 3,           /* float [nx*ny] charge density */
  EP_ARR_PHI = 4,           /* float [nx*ny] */
  EP_ARR_EX = 5,            /* float [nx*ny] */
  EP_ARR_EY = 6,            /* float [nx*ny] */
  EP_ARR_GRAD_DENSITY = 7,  /* float [O][2] */
  EP_ARR_GRAD_WL = 8,       /* float [N][2] */
  EP_ARR_GRAD_TOTAL = 9,    /* float [O][2] preconditioned direction */
  EP_ARR_TERM_MAP = 10,     /* float [nx*ny] */
  EP_ARR_BASE_MAP = 11,     /* float [nx*ny] */
  EP_ARR_PIN_GRAD = 12,     /* float [P][2] */
  EP_ARR_MAIN = 13          /* float [O][2] optimiser main solution */
};

/* new EPlacer_2D(db) + setTargetDensity + initialization() minus its first
 * gradient (eplace.cpp:39-50, 61-426): fillers, bin grid, terminal/base
 * density, device buffers. */
int ep_create(const ep_design *design, const ep_params *params, ep_ctx **out);
int ep_destroy(ep_ctx *ctx);
const char *ep_last_error(const ep_ctx *ctx);
int ep_get_info(ep_ctx *ctx, ep_info *info);

/* EPlacer_2D::totalGradientUpdate (eplace.cpp:687-738). */
int ep_total_gradient_update(ep_ctx *ctx);
/* EPlacer_2D::getPosition / setPosition / getGradient (eplace.cpp:742-804);
 * n = number of objects of the current stage. */
int ep_get_position(ep_ctx *ctx, float *xy, int64_t n);
int ep_set_position(ep_ctx *ctx, const float *xy, int64_t n);
int ep_get_gradient(ep_ctx *ctx, float *xy, int64_t n);
int64_t ep_stage_size(ep_ctx *ctx);
/* EPlacer_2D::penaltyFactorInitilization / updatePenaltyFactor
 * (eplace.cpp:812-883). */
int ep_penalty_factor_init(ep_ctx *ctx);
int ep_update_penalty_factor(ep_ctx *ctx);
/* EPlacer_2D::switch2FillerOnly / switch2cGP (eplace.cpp:887-904). */
int ep_switch_to_filler_only(ep_ctx *ctx);
int ep_switch_to_cgp(ep_ctx *ctx, int32_t mgp_iterations);
int ep_get_stats(ep_ctx *ctx, ep_stats *stats);
/* PlaceDB::calcHPWL (placedb.cpp:424-432) at the current positions. */
int ep_hpwl(ep_ctx *ctx, double *hpwl);

/* FirstOrderOptimizer<VECTOR_3D>::opt() with EplaceNesterovOpt
 * (opt.hpp:29-36, nesterov.hpp:86-415), run device-resident: init, loop
 * until the stage's stop condition, wrap_up. */
int ep_nesterov_run(ep_ctx *ctx, const ep_opt_params *opt, ep_run_stats *out);

/* FFT_2D::doFFT on a caller-supplied density (fft.cpp:190-284), on the
 * context's grid: rho [nx*ny] in, phi/ex/ey [nx*ny] out (phi may be NULL). */
int ep_poisson(ep_ctx *ctx, const float *rho, float *phi, float *ex, float *ey);

/* Raw access to device arrays (EP_ARR_*); count is in elements of the
 * array's element type (float pairs count as 2). */
int ep_read(ep_ctx *ctx, int32_t which, void *host, int64_t count);
int ep_write(ep_ctx *ctx, int32_t which, const void *host, int64_t count);

/* Wall time in seconds of the last ep_total_gradient_update's kernels, split
 * per phase: [scatter, density finalise, wiring kernels, node sums of the pin
 * gradients, poisson, field gather, combine]. */
int ep_last_timings(ep_ctx *ctx, double *seconds7);

/* ---------------------------------------------------------------------------
 * Standard-cell legalisation (host, no device work).  These take the design
 * and node centres directly; errors are described by ep_host_last_error().
 * ------------------------------------------------------------------------ */
typedef struct ep_legal_params {
  double site_step;         /* site width of the rows (.scl Sitewidth) */
  int32_t mode;             /* 0 reference semantics, 1 corrected (see legalize.cpp) */
  int32_t pad;
} ep_legal_params;

typedef struct ep_legal_stats {
  int64_t cells;            /* standard cells placed */
  int32_t subrows;          /* row intervals after removing blocked sites */
  int32_t pad;
  double total_displacement, max_displacement;
} ep_legal_stats;

typedef struct ep_legal_check {
  int64_t cells;            /* standard cells checked */
  int64_t overlaps;         /* cells overlapping the cell before them in their row */
  int64_t off_row;          /* lower-left y on no row, or outside the row's x span */
  int64_t off_site;         /* x not on the row's site grid */
  int64_t out_of_core;
  int64_t in_blockage;      /* overlapping a macro or a terminal inside the core */
  int64_t macro_overlaps;   /* overlapping macro pairs */
  double overlap_area;
} ep_legal_check;

const char *ep_host_last_error(void);
/* ---------------------------------------------------------------------------
 * Quadratic initial placement (QPlace/qplace.cpp).  Self-contained: each call
 * uploads the design, runs on the device and returns host results; errors
 * are described by ep_qp_last_error().
 * ------------------------------------------------------------------------- */
typedef struct ep_qp_params {
  int32_t max_iterations;    /* -IPiteCount, default 20 (qplace.cpp:35-42) */
  int32_t solver_iterations; /* BiCGSTAB iteration cap, 100 (qplace.cpp:68) */
  float target_error;        /* outer stop on both solver errors, 1e-6 (qplace.cpp:45) */
  float solver_tolerance;    /* BiCGSTAB tolerance: Eigen's default, float epsilon */
  int32_t start_at_center;   /* 1: moveNodesCenterToCenter first (qplace.cpp:23), as the reference */
  int32_t device;            /* HIP device ordinal */
} ep_qp_params;

typedef struct ep_qp_stats {
  int32_t iterations;        /* outer iterations run */
  int32_t solver_iterations; /* BiCGSTAB iterations, both directions, all outer iterations */
  int32_t restarts;          /* BiCGSTAB restarts */
  int32_t pad;
  double initial_hpwl;       /* after the move to the centre (qplace.cpp:25) */
  double final_hpwl;         /* calcNetBoundPins after the last update (qplace.cpp:81) */
  double error_x, error_y;   /* last solver errors (qplace.cpp:73, 78) */
  double seconds;            /* wall time of the call, upload included */
} ep_qp_stats;

/* Assembled systems of QPPlacer::createSparseMatrix: rows 0..N-1 are X and
 * N..2N-1 are Y; off-diagonal entries in CSR (columns are node indices of
 * the same direction, duplicates not merged), the diagonal and b apart.
 * Call with rowptr/col/val/diag/b NULL to get nnz, then with arrays of
 * [2N+1], [capacity], [capacity], [2N], [2N]. */
typedef struct ep_qp_matrix_out {
  int64_t nnz;      /* out */
  int64_t capacity; /* in: length of col / val */
  int32_t *rowptr, *col;
  float *val, *diag, *b;
  double hpwl;      /* out: calcNetBoundPins at the given centres */
} ep_qp_matrix_out;

/* ndim systems of n rows in one CSR (rows d*n .. d*n+n-1 belong to system
 * d, columns 0..n-1), an optional separate diagonal, and b. */
typedef struct ep_qp_system {
  int32_t n, ndim;
  int64_t nnz;
  const int32_t *rowptr; /* [ndim*n + 1] */
  const int32_t *col;    /* [nnz] */
  const float *val;      /* [nnz] */
  const float *diag;     /* [ndim*n] or NULL */
  const float *b;        /* [ndim*n] */
} ep_qp_system;

/* ---------------------------------------------------------------------------
 * Macro legalisation: SAMacroLegalizer (Legalization/legalizer.cpp:701-1530),
 * the simulated annealing over the movable macros, on the host.  Errors are
 * described by ep_mlg_last_error() (standalone) or ep_last_error() (placer).
 * ------------------------------------------------------------------------- */
typedef struct ep_mlg_params {
  float target_density; /* SAMacroLegalizer::setTargetDensity (ePlace_main.cpp:179) */
  int32_t max_outer;    /* jLimit; <= 0: 1000 (legalizer.cpp:1223) */
  int32_t max_inner;    /* kLimit; <= 0: 1000 (legalizer.cpp:1224) */
  uint32_t rand_seed;   /* srand seed of the rand() stream (standalone entry); 1 = unseeded, as the reference */
  int64_t rand_skip;    /* rand() draws made before the annealing (standalone entry) */
} ep_mlg_params;

typedef struct ep_mlg_stats {
  int32_t macros, bins_x, bins_y, outer_iterations;
  int64_t perturbations, accepted;
  int32_t overlap_free;     /* overlapFree */
  int32_t final_overlap;    /* totalMacroOverlap: total macro area - union area */
  double initial_hpwl, final_hpwl; /* PlaceDB::calcHPWL before / after */
  double cell_area_covered; /* totalCellAreaCovered at the end */
  int64_t rand_draws;       /* rand() draws the annealing made */
  double seconds;
} ep_mlg_stats;

const char *ep_mlg_last_error(void);
/* SAMacroLegalizer(db).legalization() on the given node centres (NULL: the
 * design's) with its own rand() stream (rand_seed, rand_skip); out_xy
 * receives the centres after legalisation (macros moved, cells unchanged). */
int ep_legalize_macros(const ep_design *design, const float *node_xy, const ep_mlg_params *params, float *out_xy,
                       ep_mlg_stats *stats);
/* The same on a placer (ePlace_main.cpp:174-186): node positions are read
 * from and written back to the placer, and the rand() stream is the
 * placer's (the one its fillers were drawn from, continued by the next
 * ep_switch_to_filler_only).  design must be the one the placer was created
 * from. */
int ep_macro_legalize(ep_ctx *ctx, const ep_design *design, const ep_mlg_params *params, ep_mlg_stats *stats);

const char *ep_qp_last_error(void);
/* QPPlacer::quadraticPlacement (qplace.cpp:20-103).  node_xy: [num_nodes][2]
 * centres, read when start_at_center is 0 and replaced by the result (the
 * reference moves the modules of its database).  log: NULL or
 * [max_iterations + 1][4] doubles: HPWL, X error, Y error, BiCGSTAB
 * iterations of each outer iteration (the reference's per-iteration line,
 * qplace.cpp:89). */
int ep_qplace(const ep_design *design, const ep_qp_params *params, float *node_xy, double *log,
              ep_qp_stats *stats);
/* QPPlacer::createSparseMatrix (qplace.cpp:123-280) after calcNetBoundPins,
 * at the given centres. */
int ep_qp_matrix(const ep_design *design, const float *node_xy, int32_t device, ep_qp_matrix_out *out);
/* BiCGSTAB<SMatrix, IdentityPreconditioner>::solveWithGuess with
 * setMaxIterations (qplace.cpp:67-78) on each of ndim systems; x holds the
 * guess and receives the solution; error / iterations: [ndim] or NULL. */
int ep_qp_solve(const ep_qp_system *system, float *x, int32_t max_iterations, float tolerance, int32_t device,
                float *error, int32_t *iterations);

/* No counterpart in the reference.  ep_qp_solve for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): every array of `system` and `x` is a
 * device pointer on `device` (x holds the guess and receives the solution),
 * and the solver's work vectors, partial sums and scalars live in one
 * library-owned static device array per device, sized for
 * EP_QP_NOALLOC_MAX_ROWS rows (ndim * n; about 24 MiB, held until the
 * process ends).  Larger systems are rejected with EP_ERR_ARG before any
 * device work; a CSR whose rowptr or column indices are out of range is
 * rejected by a checking kernel before the solve reads through it.  Kernels
 * run on `stream` (a hipStream_t; NULL: the null stream) and the call
 * returns when the solve has finished.  All calls on one device share the
 * work space, which keeps the last call's contents: calls that overlap in
 * time on one device race.  Results are those of ep_qp_solve on the same
 * system.  error / iterations: [ndim] host arrays or NULL. */
#define EP_QP_NOALLOC_MAX_ROWS (1 << 20)
int ep_qp_solve_noalloc(const ep_qp_system *system, float *x, int32_t max_iterations, float tolerance,
                        int32_t device, void *stream, float *error, int32_t *iterations);

/* AbacusLegalizer::legalization (Legalization/legalizer.cpp:58-177) on the
 * intervals of PlaceDB::removeBlockedSite (placedb.cpp:629-849).  node_xy:
 * [num_nodes][2] centres (NULL: the design's); out_xy receives the legal
 * centres (macros unchanged). */
int ep_legalize_abacus(const ep_design *design, const float *node_xy, const ep_legal_params *params,
                       float *out_xy, ep_legal_stats *stats);
/* No counterpart in the reference (it has no standard-cell overlap
 * checker): row, site, core, blockage and overlap checks of a placement. */
int ep_check_legal(const ep_design *design, const float *node_xy, double site_step, ep_legal_check *out);

/* Detailed placement (host, no device work): DetailedPlacer::detailedPlacement
 * (DetailedPlacement/detailed.cpp:74-85, ePlace_main.cpp:271-283 with
 * -internalDP) on a legal placement: `passes` rounds of local reordering,
 * independent-set matching and global swap.  node_xy: [num_nodes][2] legal
 * centres (NULL: the design's); out_xy receives the centres after it (macros
 * unchanged); errors are described by ep_host_last_error(). */
#define EP_DP_MAX_PASSES 8
typedef struct ep_dp_params {
  double site_step;      /* site width of the rows (.scl Sitewidth) */
  double ism_time_limit; /* seconds for a round's ISM sweeps; <= 0: the reference's 28800 (detailed.cpp:118) */
  int32_t passes;        /* rounds, 2 (detailed.cpp:77) */
  int32_t pad;
} ep_dp_params;

typedef struct ep_dp_stats {
  double hpwl_before, hpwl_after, seconds;
  double seconds_lr, seconds_ism, seconds_gs; /* time in each move over all rounds */
  double hpwl_after_lr[EP_DP_MAX_PASSES], hpwl_after_ism[EP_DP_MAX_PASSES], hpwl_after_gs[EP_DP_MAX_PASSES];
  int64_t lr_windows, ism_matchings, gs_moves;
} ep_dp_stats;

int ep_detailed_place(const ep_design *design, const float *node_xy, const ep_dp_params *params, float *out_xy,
                      ep_dp_stats *stats);

/* ---------------------------------------------------------------------------
 * Learned step / lambda controller.  No counterpart in the reference: an MLP
 * EP_CTRL_FEATURES -> EP_CTRL_HIDDEN -> EP_CTRL_HIDDEN -> EP_CTRL_OUTPUTS
 * (tanh hidden layers, FP32) evaluated on the device once per iteration of
 * method 2.  Inputs are the EP_LOG_FEATURES columns; outputs are
 * log(step / plain proposal), clamped to [log 1/4, 0], and l, clamped to
 * [0, log(1.05/0.95)], which sets a floor 0.95 e^l under the reference's
 * penalty multiplier (eplace.cpp:846-883): the iteration multiplies lambda
 * by max(reference multiplier, floor), which stays in the reference's
 * range [0.95, 1.05]; zero outputs reproduce plain backtracking.
 *
 * Parameter vector (EP_CTRL_PARAMS floats): W1[H][F], b1[H], W2[H][H], b2[H],
 * W3[O][H], b3[O] (the EP_CTRL_WEIGHTS trained values), then the input
 * normalisation mean[F], std[F] and output mean[O], std[O]: the network sees
 * (x - mean) / std and its output is scaled back by std and mean.
 * ------------------------------------------------------------------------ */
enum {
  EP_CTRL_FEATURES = 12,
  EP_CTRL_HIDDEN = 64,
  EP_CTRL_OUTPUTS = 2,
  EP_CTRL_WEIGHTS = 12 * 64 + 64 + 64 * 64 + 64 + 2 * 64 + 2,
  EP_CTRL_PARAMS = 12 * 64 + 64 + 64 * 64 + 64 + 2 * 64 + 2 + 2 * 12 + 2 * 2
};

/* params: EP_CTRL_PARAMS floats, or NULL for a seeded initialisation with
 * identity normalisation. */
int ep_ctrl_create(int32_t device, const float *params, uint32_t seed, ep_ctrl **out);
int ep_ctrl_destroy(ep_ctrl *ctrl);
const char *ep_ctrl_last_error(const ep_ctrl *ctrl);
int ep_ctrl_get_params(ep_ctrl *ctrl, float *params);
/* Replaces all parameters and resets the optimiser state. */
int ep_ctrl_set_params(ep_ctrl *ctrl, const float *params);
/* Full-batch Adam on the mean squared error of the normalised outputs over
 * n samples (X [n][F], Y [n][O], host arrays), `steps` updates at learning
 * rate lr; loss_out (optional, [steps]) receives the loss before each update. */
int ep_ctrl_train(ep_ctrl *ctrl, const float *X, const float *Y, int64_t n, int32_t steps, float lr,
                  float *loss_out);
/* Y [n][O] = controller outputs for X [n][F]. */
int ep_ctrl_predict(ep_ctrl *ctrl, const float *X, int64_t n, float *Y);
/* Loss and its gradient with respect to the EP_CTRL_WEIGHTS weights at the
 * current parameters (the training kernels' backward pass). */
int ep_ctrl_loss_grad(ep_ctrl *ctrl, const float *X, const float *Y, int64_t n, float *grad, float *loss);

#ifdef __cplusplus
}
#endif

#endif /* EPLACE_AMD_H */
