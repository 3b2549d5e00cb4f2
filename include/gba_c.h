/*
 * gba_c.h -- C ABI of libgba_amd.so, the MI355X (gfx950) bundle-adjustment
 * LM solver.  Plain C: pointers, sizes and POD structs only.
 *
 * The reference has no FFI; its boundary is the C++ library API of
 * FireHorseAutumn/gpu-bundle-adjustment (gba::solve_bal_gpu and friends).
 * Each entry point below names the reference interface it replaces.  The
 * C++ mirror of that API (include/gba/lm.hpp, include/gba/bal_problem.hpp,
 * include/gba/schur.hpp) is implemented on top of these functions inside the
 * same library.
 *
 * Conventions: every int-returning function returns 0 on success and a
 * non-zero code on failure; gba_last_error() then holds a message (thread
 * local).  GPU runtime failures are errors; numerical failure of the LM
 * solve is not (it is reported through gba_lm_result.termination).
 */
#ifndef GBA_C_H
#define GBA_C_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- problem ---------------------------------------------------------- */

/* Mirrors gba::BALProblem (reference include/gba/bal_problem.hpp:15-31):
 * structure of arrays; cameras are [rodrigues(3), t(3), f, k1, k2].  The
 * solver reads the observation arrays and overwrites cameras/points with the
 * last accepted parameters. */
typedef struct gba_problem_view {
    int num_cameras;
    int num_points;
    int num_observations;
    double* cameras;   /* 9 * num_cameras */
    double* points;    /* 3 * num_points */
    int* obs_camera;   /* num_observations */
    int* obs_point;    /* num_observations */
    double* obs_pixel; /* 2 * num_observations, x/y interleaved */
} gba_problem_view;

/* ---- LM solve ----------------------------------------------------------- */

/* Mirrors gba::LMOptions (reference include/gba/lm.hpp:13-19). */
typedef struct gba_lm_options {
    int max_iterations;      /* default 50 */
    double initial_lambda;   /* default 1e-3 (Marquardt diagonal damping) */
    double rel_decrease_tol; /* default 1e-6 */
    int max_inner_trials;    /* default 16 */
    int verbose;             /* per-trial log to stderr */
} gba_lm_options;

enum gba_termination { GBA_CONVERGENCE = 0, GBA_MAX_ITERATIONS = 1, GBA_FAILURE = 2 };

/* Mirrors gba::LMResult (reference include/gba/lm.hpp:21-29); termination is
 * the enum above instead of the reference's string.  The fields after
 * total_time_s have no reference counterpart. */
typedef struct gba_lm_result {
    double initial_cost;
    double final_cost;
    int iterations;               /* accepted outer iterations */
    int termination;              /* enum gba_termination */
    unsigned long long anomalies; /* cumulative non-finite edge count */
    double total_time_s;          /* includes device allocation and upload */
    int trials;                   /* LM trials evaluated */
    double solve_time_s;          /* device-resident part (after upload); in a batch, from
                                     the start of the batch solve to the round that
                                     finished this problem (device clock) */
    double setup_time_s;          /* allocation, upload, observation sort */
    int refine_iterations;        /* mixed precision: PCG steps, summed over trials */
    double max_refine_residual;   /* mixed precision: worst final |D (b - S x)| / |D b| of the
                                     FP64 reduced camera system (S and b in FP64, evaluated
                                     without W from the Jacobians; D the diagonal scaling),
                                     after the FP64 iterative refinement */
    int refine_factor_trials;     /* mixed precision: trials refined with the FP32 factor
                                     (all of them with GBA_PRECOND_FACTOR; with block Jacobi,
                                     those after a trial missed the tolerance) */
    int refine_outer_steps;       /* mixed precision: residuals of the FP64 reduced camera
                                     system evaluated (one per trial plus one per correction
                                     solve), summed over trials */
    int refine_floor_trials;      /* mixed precision: trials whose refinement stopped above the
                                     tolerance because the residual reached the rounding level
                                     of its FP64 evaluation (or stopped shrinking) */
    double max_refine_floor;      /* mixed precision: worst rounding level of the FP64 residual,
                                     relative to |D b|: 32 x 2^-53 x |sum of the magnitudes of
                                     its terms|.  Near a minimum the residual is a small
                                     difference of large terms and cannot be evaluated, in
                                     FP64, below this */
} gba_lm_result;

/* Arithmetic of the reduced camera solve. */
enum gba_precision {
    GBA_PRECISION_FP64 = 0,  /* plain drop-in: FP64 throughout */
    GBA_PRECISION_MIXED = 1  /* FP32 W and S; the FP32-stored S solved in FP64 by CG
                                preconditioned by its FP32 MFMA factor or by its camera
                                blocks (enum gba_mixed_precond), then FP64 iterative
                                refinement against the FP64 system's residual
                                (refine_corrections); FP64 cost, gain ratio,
                                back-substitution */
};

/* Device storage of the observed pixels. */
enum gba_obs_storage {
    GBA_OBS_FP64 = 0, /* two doubles per observation, as given */
    GBA_OBS_FP16 = 1  /* two IEEE binary16 values per observation times a per-camera FP32
                         scale (see gba_encode_obs_fp16); decoded exactly into FP64 */
};

/* Solver configuration beyond gba_lm_options (no reference counterpart). */
typedef struct gba_solver_config {
    int precision;          /* enum gba_precision */
    double refine_tol;      /* mixed: relative residual target of the FP64 iterative
                               refinement, |D (b - S x)| <= tol |D b| on the FP64 reduced
                               camera system (default 1e-10); with no corrections, the
                               inner solve's target on the FP32-stored system */
    int refine_max_iters;   /* mixed: PCG step cap per trial (default 12) */
    double* trace;          /* optional, GBA_TRACE_STRIDE doubles per trial */
    int trace_capacity;     /* trials that fit in trace */
    int obs_storage;        /* enum gba_obs_storage (default FP64) */
    double huber_delta;     /* > 0: Huber loss on |r|^2 with this threshold, as the
                               reference's Ceres harness applies HuberLoss(delta)
                               (benchmarks/bench_ceres.cpp:120-141): cost = sum rho / 2,
                               rho(s) = s (s <= delta^2) or 2 delta sqrt(s) - delta^2,
                               IRLS-weighted linearisation.  0 (default): plain least
                               squares, as solve_bal_gpu. */
    int factor_layout;      /* enum gba_factor_layout (default AUTO) */
    int mixed_precond;      /* mixed: enum gba_mixed_precond (default AUTO) */
    int jacobi_max_iters;   /* mixed, block-Jacobi refinement: step cap per trial; a trial
                               that misses refine_tol within it is refined with the FP32
                               factor instead, and so is the rest of the solve
                               (0: default 200) */
    int refine_corrections; /* mixed: FP64 iterative-refinement corrections per trial, each
                               solving the residual of the FP64 system (evaluated without W,
                               from the Jacobians) on the FP32-stored one; 0: default (1);
                               negative: none, the FP64 residual is still evaluated once and
                               reported */
    int cg_deflation;       /* mixed, block-Jacobi refinement: 0 (default) deflates the CG
                               iteration by the seven directions of the similarity gauge
                               (the near-null space of the damped S); negative: plain
                               block-Jacobi CG */
} gba_solver_config;

/* Preconditioner of the mixed form's FP64 refinement (no reference
 * counterpart).  FACTOR: the FP32 MFMA Cholesky factor of the scaled S (two or
 * three steps per trial).  JACOBI: the inverses of S's 9x9 camera blocks, no
 * factorisation (tens of steps, each one pass over S).  AUTO: JACOBI from 24
 * tile columns of S up (9C above about 1,500), FACTOR below.
 * GBA_MIXED_PRECOND=factor|jacobi in the environment overrides. */
enum gba_mixed_precond { GBA_PRECOND_AUTO = 0, GBA_PRECOND_FACTOR = 1, GBA_PRECOND_JACOBI = 2 };

/* Layout of the reduced camera system's factor (no reference counterpart;
 * the reference always factors a dense S, src/lm.cu:597-604, and names a
 * sparse factor as its next step, :13-17).  ENVELOPE renumbers the cameras by
 * reverse Cuthill-McKee on the camera graph (cameras sharing a point) and
 * factors only the tiles inside each tile row's profile; results are returned
 * in the caller's camera order.  Plain FP64 only.  AUTO takes ENVELOPE for a
 * plain FP64 solve when the camera graph has block density <= 0.5 and the
 * profile needs less than half the dense factor's tile products, else DENSE. */
enum gba_factor_layout { GBA_LAYOUT_AUTO = 0, GBA_LAYOUT_DENSE = 1, GBA_LAYOUT_ENVELOPE = 2 };

/* Trace record per LM trial: iteration, lambda used, trial cost, rho,
 * status (enum below), seconds since solve entry when the trial finished. */
#define GBA_TRACE_STRIDE 6
enum gba_trial_status {
    GBA_TRIAL_ACCEPT = 0,
    GBA_TRIAL_REJECT = 1,
    GBA_TRIAL_NONSPD_V = 2,
    GBA_TRIAL_CHOL_FAIL = 3,
    GBA_TRIAL_DENOM = 4,
    GBA_TRIAL_NONFINITE = 5
};

void gba_default_options(gba_lm_options* opts);
void gba_default_config(gba_solver_config* cfg);

/* Replaces gba::solve_bal_gpu (reference include/gba/lm.hpp:34,
 * src/lm.cu:371-724).  cfg may be NULL (plain FP64).  Blocking; uses its own
 * HIP stream on the current device. */
int gba_solve_bal(gba_problem_view* prob, const gba_lm_options* opts,
                  const gba_solver_config* cfg, gba_lm_result* res);

/* Device-resident solver handle (no reference counterpart): uploads and
 * sorts the observations once, then solves repeatedly from the uploaded
 * initial parameters.  Used by the bench to separate setup from solve. */
typedef struct gba_solver gba_solver;
int gba_solver_create(const gba_problem_view* prob, const gba_solver_config* cfg,
                      gba_solver** out);
int gba_solver_solve(gba_solver* s, const gba_lm_options* opts, gba_lm_result* res);
int gba_solver_params(gba_solver* s, double* cameras, double* points);
/* Times one kernel family on the solver's stream with HIP events at the
 * solver's current parameters: reps launches, average milliseconds per
 * launch in *avg_ms.  kernel: enum gba_kernel_id. */
enum gba_kernel_id {
    GBA_KERNEL_EDGE_POINT = 0,  /* linearisation, point pass (residual, J, W, V, bp) */
    GBA_KERNEL_EDGE_CAMERA = 1, /* linearisation, camera pass (U, bc) */
    GBA_KERNEL_SCHUR = 2,       /* reduced camera system assembly */
    GBA_KERNEL_FACTOR = 3,      /* dense Cholesky of S (all its launches); mixed: + explicit L^-1 */
    GBA_KERNEL_SOLVE = 4,       /* FP64: both triangular solves; mixed: one preconditioner apply */
    GBA_KERNEL_BACKSUB = 5,     /* point back-substitution + trial cost */
    GBA_KERNEL_COST = 6,        /* residual-only cost pass */
    GBA_KERNEL_CG_STEP = 7,     /* mixed, block Jacobi: one refinement step (matrix product
                                   over S + vector update) */
    GBA_KERNEL_CG_SETUP = 8,    /* mixed, block Jacobi: inverses of the camera blocks and
                                   the gauge deflation's basis, A Z and (Z^T A Z)^-1 */
    GBA_KERNEL_REFINE_RESIDUAL = 9, /* mixed: one residual of the FP64 reduced camera system
                                      (matrix-free: point pass, camera chunks, rows) */
    GBA_KERNEL_HEAVY_SYRK = 10  /* the heavy points' dense product S -= Y Y^T alone (part of
                                   GBA_KERNEL_SCHUR; a handle without heavy points: error) */
};
int gba_solver_time_kernel(gba_solver* s, int kernel, double lambda, int reps, double* avg_ms);
/* Shape facts used for roofline accounting: sum_p k_p(k_p+1)/2, matrix
 * order and padded order of S, and the workgroup counts of the edge passes. */
int gba_solver_stats(gba_solver* s, double* pair_count, int* n, int* n_padded);
/* The factor layout the handle uses (enum gba_factor_layout: DENSE or
 * ENVELOPE), the envelope's share of the dense factor's tile products, and
 * the camera graph's block density (both 1 when the graph was not examined
 * or was found too dense).  Any output may be NULL. */
int gba_solver_layout(gba_solver* s, int* layout, double* factor_work_ratio, double* graph_density);
/* The preconditioner of the handle's mixed refinement (enum gba_mixed_precond:
 * FACTOR or JACOBI, AUTO resolved; 0 for a plain FP64 handle). */
int gba_solver_mixed_precond(gba_solver* s, int* mode);
/* The heavy points of the handle's Schur assembly (no reference counterpart;
 * the reference sums every observation pair, src/lm.cu:515-559): points seen
 * at least heavy_k times (0: none) whose share of S is one dense MFMA product
 * instead of pair sums, their count, and the product's padded inner width
 * (3 per point, rounded up to 16).  GBA_SCHUR_HEAVY_K in the environment sets
 * heavy_k when the handle is created (0: off).  Any output may be NULL. */
int gba_solver_heavy(gba_solver* s, int* heavy_k, int* heavy_points, int* width);
void gba_solver_destroy(gba_solver* s);

/* ---- batched solve (no reference counterpart) --------------------------- */

/* Many independent problems solved together (at most 113 cameras each), each
 * with the LM semantics of gba_solve_bal and the same options; every stage of an LM round is one launch for the whole batch, and a
 * round's launch sequences are replayed from HIP graphs captured on first
 * use.  For callers that solve many small reconstructions (the reference
 * solves them one call at a time).  Handle form: create uploads and sorts
 * once and autotunes the dense factor form on the batch (one timed trial per
 * form); solve restarts from the uploaded initial parameters;
 * results[count]; *rounds (optional) receives the number of LM rounds the
 * batch took. */
typedef struct gba_batch gba_batch;
int gba_batch_create(const gba_problem_view* probs, int count, gba_batch** out);
/* gba_batch_create with a precision (enum gba_precision).  GBA_PRECISION_FP64
 * is gba_batch_create.  GBA_PRECISION_MIXED stores W in FP32 and assembles
 * each problem's damped reduced camera system diagonally scaled, D S D with
 * D = (diag(U + lambda diag'(U)))^-1/2, from 64-bit fixed-point sums
 * (resolution 2^-46), stored and factored in FP64; then one FP64 refinement
 * step: the residual of the FP64 system, evaluated without W (the Jacobians
 * are evaluated again at the linearisation point), is solved with the same
 * factor and the correction added (GBA_BATCH_REFINE_STEPS sets the number
 * of steps).  The cost, gain ratio, point back-substitution and LM rules stay
 * FP64 in both. */
int gba_batch_create_precision(const gba_problem_view* probs, int count, int precision, gba_batch** out);
int gba_batch_solve(gba_batch* b, const gba_lm_options* opts, gba_lm_result* results, int* rounds);
int gba_batch_params(gba_batch* b, int index, double* cameras, double* points);
/* What the handle chose: *tile_factor 1 = tile-parallel factor across the
 * batch, 0 = one workgroup per problem; tune_ms[2] the autotuner's trial
 * times for the two forms (0 when a GBA_BATCH_FACTOR setting fixed the
 * form); *graphs 1 when rounds replay captured HIP graphs.  Any output may be
 * NULL. */
int gba_batch_info(gba_batch* b, int* tile_factor, double* tune_ms, int* graphs);
void gba_batch_destroy(gba_batch* b);
/* One-shot form: writes each problem's last accepted parameters back. */
int gba_solve_bal_batched(gba_problem_view* probs, int count, const gba_lm_options* opts,
                          gba_lm_result* results);

/* ---- pieces of the path, host arrays in / out (parity tests) ----------- */

/* The FP16 observation encoding the solver uses with GBA_OBS_FP16 (no
 * reference counterpart; the reference studied FP16 pixel storage offline,
 * docs/RUNGS.md:44,50).  Per camera c: m_c = the largest finite |x| or |y| of
 * its observations, scale[c] = (float)(m_c / 32768), or 1 when m_c is 0 or
 * there is none.  Per observation e (caller order): px16[2e + i] = binary16
 * of pixel_i / scale[c], rounded to nearest even (a NaN stays NaN, an
 * infinity stays infinite).  The decoded pixel is (double)px16 * scale[c],
 * exact in FP64. */
int gba_encode_obs_fp16(const gba_problem_view* prob, float* cam_scale, unsigned short* px16);

/* Linearisation at the given parameters (reference src/lm.cu:205-252 and
 * accumulate() :398-422).  Outputs in the reference layout: U 81C and V 9P
 * with the lower triangle written and the upper zero, W 27O in observation
 * order (FP32-stored W is widened in mixed precision), bc 9C, bp 3P. */
int gba_linearize(const gba_problem_view* prob, int precision, double* U, double* V, double* W,
                  double* bc, double* bp, double* sse, unsigned long long* anomalies);

/* One LM trial's linear algebra at lambda (reference src/lm.cu:469-643):
 * damped V inverse, Schur assembly, dense factor + solve, back-substitution,
 * gain denominator.  *status is enum gba_trial_status (ACCEPT = solved).
 * S_out (9C x 9C, column-major, lower triangle, before factorisation) and
 * rhs_out (9C) may be NULL; in mixed precision S_out is the FP32 matrix
 * widened, without the internal diagonal scaling. */
int gba_schur_step(const gba_problem_view* prob, int precision, double lambda, double* dc,
                   double* dp, double* denom, int* status, double* S_out, double* rhs_out);

/* Residual-only SSE at the given parameters (reference src/lm.cu:255-281). */
int gba_cost(const gba_problem_view* prob, double* sse, unsigned long long* anomalies);

/* gba_linearize / gba_cost under a solver configuration (precision, pixel
 * storage, loss); with a Huber loss *sse is sum rho(|r|^2) and U, V, W, bc,
 * bp come from the IRLS-weighted residuals and Jacobians.  cfg may be NULL. */
int gba_linearize_cfg(const gba_problem_view* prob, const gba_solver_config* cfg, double* U, double* V,
                      double* W, double* bc, double* bp, double* sse, unsigned long long* anomalies);
int gba_cost_cfg(const gba_problem_view* prob, const gba_solver_config* cfg, double* sse,
                 unsigned long long* anomalies);

/* Dense SPD factor + solve of an n x n column-major matrix on the GPU with
 * the solver's blocked MFMA Cholesky (replaces Eigen::LLT at
 * src/lm.cu:597-604).  b (n) is overwritten with the solution; *ok = 0 when
 * the matrix is not positive definite.  precision FP64 or MIXED (FP32
 * factor + FP64 PCG on the FP32 matrix; *refine_iters receives the step count). */
int gba_dense_spd_solve(int n, int precision, const double* A, double* b, int* ok,
                        int* refine_iters);

/* Allocation-free variant of gba_dense_spd_solve (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops.  Plain FP64 only.  A (n x n,
 * column-major, lower triangle read), b (n; overwritten with the solution
 * when the factor succeeds) and ok (one int: 1 solved, 0 not positive
 * definite, -1 a triangular solve gave up waiting) are device pointers; the
 * work is enqueued on stream (NULL = default) and the call returns without
 * waiting.  The intermediate matrices live in one static device array of
 * the library (about 9.5 MB per GPU, held until the process ends), so
 * n <= GBA_NOALLOC_MAX_N (larger n is an error return), and calls that
 * overlap on one GPU race on it: serialise them.  Calls no hipMalloc,
 * hipFree or allocating library routine. */
#define GBA_NOALLOC_MAX_N 1024
int gba_dense_spd_solve_noalloc(int n, const double* A, double* b, int* ok, void* stream);

/* ---- batched small blocks, device pointers (reference include/gba/schur.cuh) */

/* Replaces the batch_chol9_factor_solve kernel (schur.cuh:104-109): factor
 * each symmetric 9x9 (lower triangle read) and solve; ok[i] = 0 if not SPD.
 * A, b, x, ok are device pointers; stream is a hipStream_t (NULL = default). */
int gba_batch_chol9_factor_solve(const double* A, const double* b, double* x, int* ok, int n,
                                 void* stream);
/* Replaces the batch_inverse3_spd kernel (schur.cuh:112-116). */
int gba_batch_inverse3_spd(const double* A, double* Ainv, int* ok, int n, void* stream);

/* ---- host camera model and I/O ----------------------------------------- */

/* Replaces gba::snavely_residual (reference bal_problem.hpp:45-46,
 * src/bal_reader.cpp:100-128). */
void gba_snavely_residual(const double* camera9, const double* point3, const double* observed2,
                          double* out2);
/* Replaces gba::total_squared_error (bal_problem.hpp:50, bal_reader.cpp:130-139). */
double gba_total_squared_error(const gba_problem_view* prob);
/* Replaces gba::edge_residual_jacobian_host (reference lm.hpp:40-42,
 * src/lm.cu:365-369): the residual + analytic Jacobian the GPU kernels use. */
void gba_edge_residual_jacobian_host(const double* camera9, const double* point3,
                                     const double* observed2, double* out_residual2,
                                     double* out_Jc, double* out_Jp);
/* Replaces gba::load_bal (bal_problem.hpp:36, src/bal_reader.cpp:66-98), in
 * two calls: the header, then the arrays into caller-owned storage.  .bz2
 * files are streamed through bzcat as the reference does. */
int gba_bal_header(const char* path, int* num_cameras, int* num_points, int* num_observations);
int gba_load_bal(const char* path, gba_problem_view* out);
/* BAL text writer (the reference has one only in tools/quantize_bal.cu:99-116). */
int gba_write_bal(const char* path, const gba_problem_view* prob);

/* The envelope layout the solver would plan for a problem (host only, no
 * device work; no reference counterpart): *usable = 1 when the camera graph
 * is sparse enough to plan at all (block density <= max_density), then
 * cam_order[C] (internal position -> caller camera, reverse Cuthill-McKee),
 * first[nblk] (first tile column of each 64-row tile row of the renumbered
 * S, nblk = ceil(9C / 64)), the profile's share of the dense factor's tile
 * products and the graph's block density.  cam_order / first may be NULL. */
int gba_plan_envelope(const gba_problem_view* prob, double max_density, int* usable, int* cam_order, int* first,
                      double* work_ratio, double* density);

/* ---- synthetic problems (no reference counterpart) ---------------------- */

enum gba_visibility_law { GBA_LAW_GEOMETRIC = 0, GBA_LAW_PARETO = 1 };

typedef struct gba_synth_spec {
    int num_cameras, num_points, num_observations;
    int kmin, kmax;      /* observations per point */
    int law;             /* enum gba_visibility_law */
    double alpha;        /* Pareto tail index */
    double ring_radius;  /* camera ring radius (default 30) */
    double pixel_noise;  /* pixel noise sigma (1 px) */
    double perturb;      /* relative parameter perturbation (0.01) */
    int view_window;     /* > 0: a point is seen only by cameras within this ring distance of
                            the camera nearest to it in azimuth (a camera graph of bounded
                            degree, as in real reconstructions); 0: by any camera */
} gba_synth_spec;

/* Preset shapes: 1 (C16/P22106/O83718), 2 (C49/P7776/O31843),
 * 3 (C356/P226730/O1255268), 4 (C1936/P525000/O3400000), 5 (batched: C, P,
 * O drawn from the seed: C~U{8..64}, P~U{2000..20000}, O = round(4.5 P)),
 * and two with a limited view (sparse camera graph): 6 (C400/P60000/O300000,
 * view_window 12) and 7 (C2000/P400000/O2000000, view_window 12). */
int gba_synth_preset(int config, unsigned long long seed, gba_synth_spec* out);
int gba_synth_generate(const gba_synth_spec* spec, unsigned long long seed, gba_problem_view* out);
/* The reference's data-free end-to-end problem (tests/test_launch_bounds.cpp
 * :28-80): C cameras, P points, 2P noise-free observations, std::mt19937_64. */
int gba_synth_launch_bounds(int num_cameras, int num_points, unsigned long long seed,
                            gba_problem_view* out);

/* ---- misc ---------------------------------------------------------------- */
/* Waits for all work this library enqueued on the current device (for callers
 * that launched the batched kernels on the default stream). */
int gba_device_synchronize(void);
const char* gba_last_error(void);
const char* gba_version(void);

#ifdef __cplusplus
}
#endif
#endif /* GBA_C_H */
