// launchers.hpp -- host-side launch wrappers for the HIP kernels.  Every
// wrapper enqueues on the given stream and returns immediately.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "device_common.hpp"
#include "gba_c.h"

namespace gba_amd {

// k_linearize.hip
void launch_lin_points(hipStream_t st, const DevProblem& pb, const double* cams, const double* pts,
                       double* W64, float* W32, double* VB, double* unit_sse, unsigned* unit_anom);
void launch_lin_cameras(hipStream_t st, const DevProblem& pb, const double* cams, const double* pts,
                        double* gram_part);
void launch_lin_reduce(hipStream_t st, const DevProblem& pb, const double* gram_part,
                       const double* unit_sse, const unsigned* unit_anom, double* U, double* bc,
                       PassStatus* status);
void launch_cost(hipStream_t st, const DevProblem& pb, const double* cams, const double* pts,
                 double* blk_sse, unsigned* blk_anom, PassStatus* status);
int cost_partials(const DevProblem& pb);

// noalloc.hip: FP64 dense SPD factor + solve on caller device buffers, scratch
// in the library's static device array (n <= kNoAllocMaxN); no allocation.
constexpr int kNoAllocMaxN = GBA_NOALLOC_MAX_N;
void dense_spd_solve_noalloc(int n, const double* A, double* b, int* ok, hipStream_t st);

// k_blocks.hip
void launch_clear_flags(hipStream_t st, TrialFlags* flags);
void launch_fill_int(hipStream_t st, int* p, int n, int v);
void launch_fill_double(hipStream_t st, double* p, size_t n, double v);
void launch_damp_invert_v(hipStream_t st, int P, const double* VB, double lambda, double* Vinv6,
                          TrialFlags* flags);
// Batched problems: lambda and the active switch per problem, flags[problem].
void launch_damp_invert_v_batched(hipStream_t st, int P, const double* VB, const int* pt_prob,
                                  const double* lam, const int* active, double* Vinv6,
                                  TrialFlags* flags);
void launch_batch_chol9(hipStream_t st, const double* A, const double* b, double* x, int* ok, int n);
void launch_batch_inverse3(hipStream_t st, const double* A, double* Ainv, int* ok, int n);
void launch_identity_pad(hipStream_t st, double* S64, float* S32, int ld, int n, int npad);

// k_schur.hip
// Shape of the assembly: camera columns per LDS chunk and workgroup size
// (the task list is built for this chunk width).
struct SchurShape {
    int cb = 1;  // cameras per column chunk
    int threads = 512;
};
SchurShape schur_shape(int C, bool w32);
void launch_schur(hipStream_t st, const SchurShape& sh, const DevProblem& pb,
                  const SchurTask* tasks, int num_tasks, const int2* pairs, const double* W64,
                  const float* W32, const double* Vinv6, const double* VB, const double* U,
                  const double* bc, double lambda, const double* dscale, double* S64, float* S32fac,
                  float* S32low, int ld, double* rhs, TrialFlags* flags);
// Diagonal blocks (both triangles) + rhs, one workgroup per camera.  desc:
// per-problem S layout (batched) or NULL (single problem, leading dimension
// ld).  W32 selects the mixed form (D S D into the FP32 copies, rhs scaled);
// W32 without S32fac writes the mixed form's D S D into S64 (the batched
// mixed solver, whose factor is FP64).  launch_schur follows the same rule.
void launch_schur_diag(hipStream_t st, const DevProblem& pb, const BatchDesc* desc, const double* W64,
                       const float* W32, const double* Vinv6, const double* VB, const double* U,
                       const double* bc, double lambda, const double* dscale, double* S64,
                       float* S32fac, float* S32low, int ld, double* rhs, const TrialFlags* flags);
// Batched problems: cam_prob and lam give each row its problem's lambda.
void launch_diag_scale(hipStream_t st, int n, int npad, const double* U, double lambda,
                       double* dscale, const int* cam_prob = nullptr, const double* lam = nullptr);

// k_pairs.hip -- the Schur task pair lists on the device.  Task t = (a, lo,
// hi) in tq[t] (x, y, z) owns slots qoff[t] .. qoff[t + 1] - 1, one per
// observation of camera a in cam_list order; qoff[ntask] = nslots.  count
// writes cnt[0 .. nslots] (pairs per slot, then 0); scan turns cnt into write
// positions pos (exclusive sum over nslots + 1 entries, pair_scan_work bytes
// of scratch); bounds[t] = pos[qoff[t]] for t <= ntask (task t's pairs are
// [bounds[t], bounds[t + 1])); fill writes the pairs (e, f).
size_t pair_scan_work(long long n);
void launch_pair_count(hipStream_t st, int ntask, const int4* tq, const long long* qoff, long long nslots,
                       const int* cam_off, const int* cam_list, const int* spt, const int* pt_off,
                       const int* scam, int heavy_k, long long* cnt);
void launch_pair_scan(hipStream_t st, const long long* cnt, long long* pos, long long n, void* work,
                      size_t work_bytes);
void launch_pair_bounds(hipStream_t st, int ntask, const long long* qoff, const long long* pos,
                        long long* bounds);
void launch_pair_fill(hipStream_t st, int ntask, const int4* tq, const long long* qoff, long long nslots,
                      const int* cam_off, const int* cam_list, const int* spt, const int* pt_off, const int* scam,
                      int heavy_k, const long long* pos, int2* pairs);

// k_heavy.hip -- the heavy points' share of S (pb.heavy_k): Yt (k-major,
// kp x ldy with ldy = npad rounded up to 128, zero outside the observed
// blocks) receives D W_{a,p} L_p for heavy point hpts[h] in columns
// 3h .. 3h + 2 (FP32 with W32 and Yt32, else FP64); then S -= Y Y^T over the
// lower tiles of S (leading dimension ld; FP32 into S32 and, when given, the
// second copy S32b; FP64 into S64).
void launch_heavy_y(hipStream_t st, const DevProblem& pb, const int* hpts, int nh, const double* W64,
                    const float* W32, const double* Vinv6, const double* dscale, double* Yt64, float* Yt32,
                    int ldy);
void launch_heavy_syrk(hipStream_t st, int nblk, int kp, int ldy, int ld, const double* Yt64, const float* Yt32,
                       double* S64, float* S32, float* S32b, const TrialFlags* flags);

// k_layout.hip: a batch's concatenated observations from problem-local to
// batch camera / point indices (problem q owns observations [o_off[q],
// o_off[q + 1]) and adds c0[q], p0[q]).
void launch_offset_obs(hipStream_t st, int O, int count, const long long* o_off, const int* c0, const int* p0,
                       int* cam, int* pt);

// k_dense.hip
// Factors in place; D holds 2*nblk tiles of 64x64: tiles [0, nblk) receive the
// diagonal tiles of L (row-major, upper part zero), tiles [nblk, 2 nblk) their
// inverses transposed; the off-diagonal tiles of L replace S's.  FP32 with X32
// set: X32 also receives X = L^-1, formed inside the factor's launches (lower
// tile (i, j) X_ij, upper tile (j, i) X_ij^T, diagonal tile X_ii below and
// X_ii^T above its diagonal; R32: n_pad^2 floats of scratch).
void launch_cholesky(hipStream_t st, double* S64, float* S32, int ld, int nblk, double* D64,
                     float* D32, TrialFlags* flags, float* X32 = nullptr, float* R32 = nullptr);
// Batched problems (FP64): solve every active problem's S in place, one
// workgroup per problem (desc[count]); dc = S^-1 rhs per problem.  factor =
// true factors S in the same workgroup first; false expects the factor from
// launch_batch_cholesky.
void launch_batch_chol_solve(hipStream_t st, int count, const BatchDesc* desc, double* S, double* D,
                             const double* rhs, double* dc, TrialFlags* flags, const int* active,
                             bool factor);
// Work tables of the tile-parallel batched factor (device arrays of
// (problem, tile) pairs, host offsets): potrf0[count] factors every problem's
// first diagonal tile; step k runs trsm[trsm_off[k] .. trsm_off[k+1]) (tile
// rows below the diagonal, as k_trsm_tiles numbers them) and
// upd[upd_off[k] .. upd_off[k+1]) (trailing tiles, as k_chol_update numbers
// them, tile 0 first) for every problem with more than k+1 tile columns.
struct BatchFactorTables {
    const int2* potrf0 = nullptr;
    const int2* trsm = nullptr;
    const int2* upd = nullptr;
    const int* trsm_off = nullptr;  // host, nsteps entries
    const int* upd_off = nullptr;   // host, nsteps entries
    int nsteps = 0;                 // largest tile-column count
};
// Factors every active problem's S (FP64, desc layout) in place with the
// single-problem tile kernels, one launch per stage and column for the whole
// batch: D receives each problem's diagonal tiles and their inverses as
// launch_cholesky lays them out; a bad pivot sets flags[problem].chol_fail.
void launch_batch_cholesky(hipStream_t st, int count, const BatchFactorTables& t, const BatchDesc* desc,
                           double* S, double* D, TrialFlags* flags, const int* active);
// One-launch form of launch_cholesky (same outputs): a critical-chain
// workgroup and ticketed workers (persist_tasks.hpp).  persist_plan gives the
// per-column first tickets (nblk + 1, to upload) and the grid; sync holds
// persist_sync_ints(nblk) ints (cleared here by a fill kernel).  X32 / R32
// (FP32 only, both or neither): also form X = L^-1.
size_t persist_sync_ints(int nblk);
void persist_plan(int nblk, bool inverse, bool fp64, std::vector<long long>& col_start, int& grid);
void launch_cholesky_persistent(hipStream_t st, double* S64, float* S32, int ld, int nblk, double* D64, float* D32,
                                TrialFlags* flags, const long long* col_start, long long ntasks, int grid, int* sync,
                                float* X32 = nullptr, float* R32 = nullptr);

// Solves L L^T x = b (b, x length nblk*kTile; b and x may alias).  sync holds
// 2*nblk + 2 ints of scratch (done flags and tickets).  first (nblk ints, optional):
// the envelope factor's first tile column of each tile row.  skip (optional):
// a device flag that turns the two solves into empty launches (a refinement
// step after its stop).
void launch_chol_solve(hipStream_t st, const double* L64, const float* L32, int ld, int nblk,
                       const double* D64, const float* D32, const double* b, double* y, double* x,
                       int* sync, TrialFlags* flags,
                       const int* first = nullptr, const double* skip = nullptr);
// y = A x, A symmetric given by the lower triangle of an FP32 row-major
// matrix (FP64 accumulation); work holds symv_work(npad) doubles.
size_t symv_work(int npad);
// done (optional): a device flag that turns the launch into a no-op (the
// refinement's steps after its stopping test).
void launch_symv(hipStream_t st, const float* A, int ld, int npad, const double* x, double* y,
                 double* work, const double* done = nullptr);
int residual_partials(int npad);
// blk holds 2 residual_partials(n) doubles (launch_pcg_dot).
// Refinement dot products with device-side scalars sc[0..7] (see
// k_pcg_reduce): mode 0 start (r.z, b.b), 1 alpha from p.q, 2 beta from the
// new r.z and r.r, the step count and the stopping test (r.r <= tol2 b.b,
// max_it steps); sc[4] is the done flag of the no-op launches.
void launch_pcg_dot(hipStream_t st, int n, const double* a, const double* b, const double* c,
                    const double* d, double* blk, int mode, double* sc, PassStatus* status,
                    const TrialFlags* flags, double tol2, int max_it);
void launch_pcg_update(hipStream_t st, int n, const double* sc, double* x, const double* p, double* r,
                       const double* q);
void launch_pcg_direction(hipStream_t st, int n, const double* sc, const double* z, double* p);
void launch_vec_copy(hipStream_t st, int n, const double* a, double* b);
void launch_vec_axpy(hipStream_t st, int n, double* x, const double* d);
void launch_vec_scale(hipStream_t st, int n, int npad, const double* in, const double* s,
                      double* out);

// k_precond.hip
// z = X^T (X r) = (L L^T)^-1 r; work holds precond_work(npad) doubles.
size_t precond_work(int npad);
void launch_precond_apply(hipStream_t st, const float* X32, int ld, int npad, const double* r,
                          double* z, double* work, const double* done = nullptr);

// k_cg.hip -- block-Jacobi conjugate gradients on the FP32-stored scaled S
// (the mixed form's factor-free inner solver).  A: lower tiles valid, camera
// blocks on the diagonal in both triangles; vectors have npad entries.
// Matrix-product workgroups: each takes a run of up to tpw tiles of one tile
// row, wgtab[w] = (tile row i, run c), listed row by row; rowstart[i] is the
// index of tile row i's first run (nblk + 1 entries).  Work arrays: rp
// (nwg_m x 64), cp (nblk (nblk + 1) / 2 x 64), dpart (nwg_m), gpart
// (kCgPartials nwg_v x 2), st (16).
// Deflation by the similarity gauge (zd non-null): zd and azd hold the
// normalised gauge basis Z of the scaled system and A Z (npad x 8, the eighth
// column zero), einv the 7x7 (Z^T A Z)^-1 (all zero when the basis is
// unusable, which turns the deflation off), zbpart the start's partial sums
// of Z^T b and b.b (nwg_v x 8).
constexpr int kCgPartials = 10;  // per vector workgroup: gamma, r.r, (A Z)^T u
struct CgArgs {
    const float* A;
    int ld, npad, nblk, ncam, tpw;
    const int2* wgtab;
    const int* rowstart;
    int nwg_m, nwg_v;
    const double* b;
    double *x, *r, *u, *p, *s;
    const double* minv;  // ncam x 81: inverse of each camera block
    double *rp, *cp, *dpart, *gpart, *st;
    const double *zd, *azd, *einv;
    double* zbpart;
    PassStatus* status;
    const TrialFlags* flags;
    double tol2;
    int max_it;
};
// Work of the deflation set-up: unnormalised basis (npad x 8), partial
// column norms (nwg_v x 8), partials of A Z per run of tiles
// (nblk x cg_gauge_runs(nblk) x 64 x 8) and of Z^T A Z (nblk x 49).
struct CgDeflWork {
    double *zraw, *npart, *apart, *epart;
    double *zd, *azd, *einv;
};
int cg_vector_groups(int npad);
int cg_gauge_runs(int nblk);
// minv[c] = A_cc^-1 (FP64 Cholesky of the FP32 block); a block that is not
// SPD raises chol_fail.
void launch_cg_block_inverse(hipStream_t st, const float* A, int ld, int ncam, double* minv, TrialFlags* flags);
// Once per trial, after the Schur assembly: the gauge basis of the scaled
// system at cameras cams (internal order), A Z and (Z^T A Z)^-1.
void launch_cg_deflation_setup(hipStream_t st, const CgArgs& g, const double* cams, const double* dscale,
                               const CgDeflWork& w);
// x = 0, r = b, u = M^-1 r (clears the done state); with the deflation
// x = Z (Z^T A Z)^-1 Z^T b, r = b - A x.
void launch_cg_start(hipStream_t st, const CgArgs& g);
// Step k (0, 1, ... in order): the matrix product and the vector update;
// both do nothing once the iteration has stopped.
void launch_cg_step(hipStream_t st, const CgArgs& g, int k);

// k_batch.hip
// Per-problem status of a pass (trial = 0: linearisation SSE and anomalies;
// trial = 1: trial SSE, gain denominator, anomalies, flags).
void launch_batch_status(hipStream_t st, int count, const BatchDesc* desc, int trial,
                         const double* unit_sse, const double* unit_den, const unsigned* unit_anom,
                         const double* dc, const double* U, const double* bc, const double* lam,
                         const TrialFlags* flags, const int* active, PassStatus* status);
void launch_batch_accept(hipStream_t st, int nc, int np, const int* cam_prob, const int* pt_prob,
                         const int* accept, const double* tcams, const double* tpts, double* cams,
                         double* pts);
// Per-problem LM control of the batched solver on the device (the rules of
// src/lm.cu:455-714, as Solver::solve applies them): one state per problem,
// the per-problem switches of the next pass (lin_active: re-linearise;
// trial_active: evaluate a trial; accept: copy the trial parameters) and
// lambda (lam, what the trial kernels damp with).  counters[0] counts the
// rounds in which some problem evaluated a trial, counters[1] the problems
// still running.
struct BatchLmState {
    double sse, lambda, nu, cost0;
    unsigned long long counter;  // non-finite edges seen (the reference's anomaly counter)
    // device wall clock (wall_clock64, hipDeviceAttributeWallClockRate kHz)
    // when the batch solve started and when this problem finished
    unsigned long long t_start, t_done;
    int need_lin, done, converged, failed, inner, iterations, trials, started;
    int done_round;  // the round in which the problem finished
    int pad_;
};
void launch_batch_lm_init(hipStream_t st, int count, double lambda0, BatchLmState* state, double* lam,
                          int* lin_active, int* trial_active, int* accept, int* counters);
// After a linearisation pass (statuses of the problems in lin_active).
void launch_batch_lm_lin(hipStream_t st, int count, const gba_lm_options& opts, const PassStatus* lin_status,
                         BatchLmState* state, const int* lin_active, int* trial_active, int* counters);
// After a trial pass: accept / reject, lambda and nu, termination; sets
// accept, lin_active and lam for the next pass.
void launch_batch_lm_trial(hipStream_t st, int count, const gba_lm_options& opts, const PassStatus* trial_status,
                           BatchLmState* state, const int* trial_active, int* accept, int* lin_active,
                           double* lam, int* counters);

// k_refine.hip
// out = rhs - S x of the damped reduced camera system in FP64 (D (rhs - S x)
// with dscale), with neither S nor W read: passes that evaluate the
// Jacobians again at (cams, pts) -- z_p = V'^-1 (bp - sum W_f^T x) into z
// (3P), then the camera sums per camera chunk into part (18 doubles per
// chunk), then the rows; rot (4C doubles) receives the cameras' rotation
// terms; mag (optional) receives per row the sum of the magnitudes of the
// row's terms, scaled like out (the rounding of the evaluation is a small
// multiple of 2^-53 times it).  x NULL: x = 0 (out = D rhs).  Problems
// switched off by pb.active or whose trial failed (flags) are skipped.
void launch_refine_residual(hipStream_t st, const DevProblem& pb, const double* cams, const double* pts,
                            const double* Vinv6, const double* VB, const double* U, const double* bc, double lambda,
                            const double* x, double* z, double* rot, double* part, const double* dscale, double* out,
                            double* mag, const TrialFlags* flags);
// x_i += s_i d_i for i < n (s NULL: x += d)
void launch_vec_axpy_scaled(hipStream_t st, int n, double* x, const double* d, const double* s);

// k_backsub.hip
void launch_trial_cams(hipStream_t st, int C, const double* cams, const double* dc, double* tcams);
void launch_backsub(hipStream_t st, const DevProblem& pb, const double* W64, const float* W32,
                    const double* Vinv6, const double* VB, const double* dc, const double* pts,
                    const double* tcams, double* tpts, double lambda, double* unit_sse,
                    double* unit_den, unsigned* unit_anom, const TrialFlags* flags);
void launch_trial_finalize(hipStream_t st, const DevProblem& pb, const double* unit_sse,
                           const double* unit_den, const unsigned* unit_anom, const double* dc,
                           const double* U, const double* bc, double lambda, PassStatus* status,
                           const TrialFlags* flags);
// dp for the first 3P entries of out (test helper): recomputes the point step.
void launch_point_step(hipStream_t st, const DevProblem& pb, const double* W64, const float* W32,
                       const double* Vinv6, const double* VB, const double* dc, double* dp);

}  // namespace gba_amd
