/*
 * gba_oracle.h -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU restatement (plain C11 + pthreads) of the reference bundle-adjustment
 * LM solve.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it, and only as the checker / timed CPU baseline.  The product
 * path (libgba_amd.so) never links, loads or calls anything here.
 *
 * Every function cites the reference file:line it restates.  All matrices are
 * row-major unless stated otherwise; U/V blocks follow the reference layout
 * (lower triangle written, upper triangle zero).
 */
#ifndef GBA_ORACLE_H
#define GBA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- camera model --------------------------------------------------- */
/* src/bal_reader.cpp:100-128 (zero-order branch for theta <= 1e-12). */
void orc_snavely_residual(const double* cam9, const double* X3, const double* obs2, double* r2);
/* src/bal_reader.cpp:130-139 (sequential sum in observation order). */
double orc_total_squared_error(int O, const double* cams, const double* pts, const int* obs_cam,
                               const int* obs_pt, const double* obs_px);
/* src/lm.cu:58-82 (first-order branch for theta^2 <= 1e-24). */
void orc_edge_residual(const double* cam9, const double* X3, const double* obs2, double* r2);
/* src/lm.cu:93-191: residual + Jc (2x9 row-major) + Jp (2x3 row-major). */
void orc_edge_residual_jacobian(const double* cam9, const double* X3, const double* obs2,
                                double* r2, double* Jc18, double* Jp6);

/* ---- linearisation (src/lm.cu:205-252, accumulate() :398-422) ------- */
/* U 81C, V 9P (lower triangles, upper zero), W 27O (edge order), bc 9C,
 * bp 3P, *sse, *anomalies = number of skipped non-finite edges. Results are
 * bitwise independent of nthreads (fixed per-camera / per-point order). */
void orc_linearize(int C, int P, int O, const double* cams, const double* pts, const int* obs_cam,
                   const int* obs_pt, const double* obs_px, double* U, double* V, double* W,
                   double* bc, double* bp, double* sse, unsigned long long* anomalies,
                   int nthreads);
/* src/lm.cu:255-281: SSE of the residuals, non-finite residuals skipped and counted. */
double orc_cost(int O, const double* cams, const double* pts, const int* obs_cam, const int* obs_pt,
                const double* obs_px, unsigned long long* anomalies, int nthreads);
/* The same two passes with the Huber loss of the reference's Ceres harness
 * (benchmarks/bench_ceres.cpp:120-141; huber_delta <= 0: none): *sse and the
 * cost are sum rho(|r|^2); the linearisation uses sqrt(rho') r and
 * sqrt(rho') J. */
void orc_linearize_loss(int C, int P, int O, const double* cams, const double* pts, const int* obs_cam,
                        const int* obs_pt, const double* obs_px, double huber_delta, double* U, double* V,
                        double* W, double* bc, double* bp, double* sse, unsigned long long* anomalies,
                        int nthreads);
double orc_cost_loss(int O, const double* cams, const double* pts, const int* obs_cam, const int* obs_pt,
                     const double* obs_px, double huber_delta, unsigned long long* anomalies, int nthreads);

/* ---- batched blocks (include/gba/schur.cuh) ------------------------- */
int orc_chol9_factor(const double* A81, double* L81);                   /* :23-38 */
void orc_chol9_solve(const double* L81, const double* b9, double* x9);  /* :41-53 */
int orc_inverse3_spd(const double* A9, double* Ainv9);                  /* :66-96 */
void orc_batch_chol9_factor_solve(const double* A, const double* b, double* x, int* ok, int n);
void orc_batch_inverse3_spd(const double* A, double* Ainv, int* ok, int n);
/* src/lm.cu:289-302: Vinv_p = inverse3_spd(V_p + lambda*diag'(V_p)); returns
 * the number of non-SPD blocks (their Vinv is left untouched). */
long orc_damp_invert_v(int P, const double* V, double lambda, double* Vinv);

/* ---- dense reduced camera system (src/lm.cu:597-604 uses Eigen LLT) - */
/* In-place lower Cholesky of an n x n column-major matrix (lower triangle read
 * and overwritten). Returns 1 on success, 0 when a pivot is not > 0. */
int orc_dense_cholesky(int n, double* A);
/* Solves L L^T x = b in place (b <- x) with L from orc_dense_cholesky. */
void orc_dense_cholesky_solve(int n, const double* L, double* b);

/* ---- one LM trial's linear solve (src/lm.cu:469-643) ----------------- */
enum {
    ORC_TRIAL_OK = 0,
    ORC_TRIAL_NONSPD_V = 1, /* lm.cu:477-481 */
    ORC_TRIAL_LLT_FAIL = 2, /* lm.cu:599-603 */
    ORC_TRIAL_DENOM = 3     /* lm.cu:639-643 */
};
/* dc 9C, dp 3P, *denom; optional S_out (9C x 9C column-major, lower filled,
 * before factorisation) and rhs_out (9C). Returns an ORC_TRIAL_* code. */
int orc_schur_step(int C, int P, int O, const int* obs_cam, const int* obs_pt, const double* U,
                   const double* V, const double* W, const double* bc, const double* bp,
                   double lambda, double* dc, double* dp, double* denom, double* S_out,
                   double* rhs_out, int nthreads);

/* ---- full solve (src/lm.cu:371-724) ---------------------------------- */
typedef struct {
    int max_iterations;      /* include/gba/lm.hpp:14 */
    double initial_lambda;   /* :15 */
    double rel_decrease_tol; /* :16 */
    int max_inner_trials;    /* :17 */
    int verbose;             /* :18 */
    double huber_delta;      /* > 0: Huber loss (no counterpart in LMOptions; the
                                Ceres harness's --loss huber --delta) */
} orc_options;

enum { ORC_CONVERGENCE = 0, ORC_MAX_ITERATIONS = 1, ORC_FAILURE = 2 };

typedef struct {
    double initial_cost, final_cost;
    int iterations;
    int termination;              /* ORC_CONVERGENCE / ORC_MAX_ITERATIONS / ORC_FAILURE */
    unsigned long long anomalies; /* lm.hpp:26-27 */
    double total_time_s;
    int trials;
    /* phase timers, as the reference's verbose breakdown (lm.cu:718-722) */
    double t_linearize, t_dampinv, t_assembly, t_solve, t_trial_cost;
} orc_result;

/* Trial record of the trace (ORC_TRACE_STRIDE doubles per trial): outer
 * iteration, lambda used, trial cost, rho, status (below), nu when the trial
 * ran, gain-ratio denominator (0 when the trial stopped before it). */
#define ORC_TRACE_STRIDE 7
enum {
    ORC_TR_ACCEPT = 0,
    ORC_TR_REJECT = 1,
    ORC_TR_NONSPD_V = 2,
    ORC_TR_LLT_FAIL = 3,
    ORC_TR_DENOM = 4,
    ORC_TR_NONFINITE = 5
};

/* cams/pts are updated in place to the last accepted state (lm.cu:679-680).
 * nthreads: host threads as the reference uses them (Phase 1 and merge use
 * min(C, nthreads); Phase 2 workers are capped by the 16 MiB rule,
 * lm.cu:449-451; the dense LLT is single-threaded); the linearisation, which
 * the reference runs on its GPU, uses nthreads as well. trace may be NULL. */
int orc_lm_solve(int C, int P, int O, double* cams, double* pts, const int* obs_cam,
                 const int* obs_pt, const double* obs_px, const orc_options* opts,
                 orc_result* res, double* trace, int max_trace, int* n_trace, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
