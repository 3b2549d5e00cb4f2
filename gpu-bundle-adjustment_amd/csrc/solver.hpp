// solver.hpp -- host driver of the device-resident LM solve.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "device_common.hpp"
#include "gba_c.h"
#include "gba_internal.hpp"
#include "launchers.hpp"

namespace gba_amd {

void hip_check(hipError_t e, const char* what, const char* file, int line);
#define GBA_HIP(x) ::gba_amd::hip_check((x), #x, __FILE__, __LINE__)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    void alloc(size_t count) {
        release();
        n = count;
        if (count) GBA_HIP(hipMalloc(&p, sizeof(T) * count));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

// Work units over the point CSR (points of up to a wavefront of observations
// grouped, larger points alone) and sum_p k_p (k_p + 1) / 2; camera-pass
// chunks over the camera CSR.
void layout_units(const std::vector<int>& pt_off, std::vector<int2>& units, double& pair_count);
// The same over points [0, P) of a CSR slice, appending to units and adding
// to pair_count (a batch builds each problem's units from its slice).
void layout_units_append(const int* pt_off, int P, std::vector<int2>& units, double& pair_count);
void layout_chunks(const std::vector<int>& cam_off, std::vector<int4>& chunks, std::vector<int>& cam_chunk_off);
// The point-major layout (k_layout.hip): perm (sorted position -> caller's
// index, ordered by (point, camera, index)), scam, spt, spx (when spx is
// set), the point CSR pt_off, and the camera index cam_off / cam_list (sorted
// positions by camera, ascending within a camera), from the caller's arrays
// already on the device; *repeated = 1 when some point is observed more than
// once by one camera.  Allocates the outputs; synchronises st.
// Camera-ordered copies of the observations' points and pixels (cspt[i] =
// spt[cam_list[i]], likewise cspx; spx NULL: no pixel copy), enqueued on st.
void build_cam_order_copies(hipStream_t st, int O, const int* cam_list, const int* spt, const double2* spx,
                            DevBuf<int>& cspt, DevBuf<double2>& cspx);
void build_layout_device(hipStream_t st, int C, int P, int O, const int* d_cam, const int* d_pt,
                         const double* d_px, DevBuf<int>& perm, DevBuf<int>& scam, DevBuf<int>& spt,
                         DevBuf<double2>* spx, DevBuf<int>& pt_off, DevBuf<int>& cam_off, DevBuf<int>& cam_list,
                         int* repeated);

// IEEE binary16 bits of x, rounded to nearest even.
uint16_t double_to_half_rne(double x);
// FP16 pixel encoding with a per-camera FP32 scale, in the caller's
// observation order (gba_encode_obs_fp16): cam_scale[C], px16[2 O].
void encode_obs_fp16(const gba_problem_view& prob, float* cam_scale, uint16_t* px16);

// The pair lists of the tasks tq[t] = (a, lo, hi): pairs (e, f) with
// camera(e) = a, camera(f) in [lo, hi), grouped by e (ascending point) and
// ordered by f, built on the device from the sorted layout (d_*; cam_off on
// the host too); task t's pairs are [bounds[t], bounds[t + 1]) of pairs
// (allocated here); points with at least heavy_k observations (heavy_k > 0)
// contribute no pairs.  Synchronises st.
void schur_pairs_device(hipStream_t st, const std::vector<int4>& tq, const std::vector<int>& cam_off,
                        const int* d_cam_off, const int* d_cam_list, const int* d_spt, const int* d_pt_off,
                        const int* d_scam, std::vector<long long>& bounds, DevBuf<int2>& pairs, int heavy_k = 0);
// Schur task and pair lists of one problem for chunks of cb camera columns:
// task (a, c) holds the pairs (e, f) with camera(e) = a and camera(f) in
// [c cb, min((c + 1) cb, a + with_diag)); every (a, c) gets a task, empty or
// not, because the task also writes its blocks of S.  Pairs from
// schur_pairs_device; tasks returned by decreasing pair count (longest
// first).  Synchronises st.
void build_schur_tasks_device(hipStream_t st, int C, int cb, const std::vector<int>& cam_off,
                              const int* d_cam_off, const int* d_cam_list, const int* d_spt,
                              const int* d_pt_off, const int* d_scam, std::vector<SchurTask>& tasks,
                              DevBuf<int2>& pairs, bool with_diag, int heavy_k = 0);

// FP64 preconditioned conjugate gradients on A x = b, A symmetric given by the
// lower triangle of an FP32 matrix, preconditioned by (L L^T)^-1 (PcgPrecond:
// the explicit FP32 inverse X = L^-1 from launch_cholesky with X32, or the
// factor's triangular solves).  Vectors
// have npad entries.  Stops at |r| <= tol |b| or after max_it steps; returns
// the steps taken and the final |r| / |b| in *rel.  work holds pcg_work(npad)
// doubles.  The stopping test runs on the device; steps are enqueued
// first_batch at a time first, then one at a time, with one synchronisation
// per batch (the steps of a batch past the stop are empty launches).
size_t pcg_work(int npad);
// The refinement's preconditioner: the explicit FP32 inverse X = L^-1
// (z = X^T X r, two tile-parallel matrix-vector passes) or, where forming X
// costs more than its cheaper applications save (large matrices), the FP32
// factor itself (forward and backward block substitution, launch_chol_solve).
// Largest tile-column count at which the mixed refinement applies the explicit
// inverse; above it, the triangular solves (measured: C=356, 51 columns,
// inverse 0.07 ms to form + 22 us per application against 0.3 ms per pair of
// solves; n = 17,424, 272 columns, inverse 29 ms to form against 1.7 ms per
// pair of solves).
constexpr int kInversePrecondBlk = 128;
// Smallest tile-column count of S at which GBA_PRECOND_AUTO refines the mixed
// form with block-Jacobi CG instead of the FP32 factor: below it the factor's
// chain of tile columns is short (measured on MI355X: mixed solve at C=356,
// 51 columns, 17.3 ms with block Jacobi against 17.9 ms with the factor; at
// n = 17,424, 273 columns, 0.10 s against 0.34 s).
constexpr int kJacobiMinBlk = 24;
// Tile-column count up to which a block-Jacobi matrix-product workgroup takes
// two tiles (enough workgroups to fill the chip), runs of eight beyond.
constexpr int kJacobiPairBlk = 96;
struct PcgPrecond {
    const float* X32 = nullptr;
    const float* L32 = nullptr;  // factor (lower tiles), with D32 (diagonal tiles and inverses)
    const float* D32 = nullptr;
    int nblk = 0;
    double* tmp = nullptr;  // npad doubles
    int* sync = nullptr;    // 2 nblk + 2 ints
};
int pcg_solve(hipStream_t st, const float* A, const PcgPrecond& M, int ld, int npad, const double* b,
              double* x, double* work, TrialFlags* flags, PassStatus* status, PassStatus* h_status,
              double tol, int max_it, double* rel, int first_batch = 1);

// Envelope layout of S (camera_graph.cpp): the cameras' internal order
// (reverse Cuthill-McKee on the camera graph), and per tile row of the
// renumbered S the first tile column of its profile.
struct EnvelopeLayout {
    std::vector<int> cam_of_pos;  // internal index -> caller camera
    std::vector<int> pos_of_cam;  // caller camera -> internal index
    std::vector<int> first;       // nblk entries
    double work_ratio = 1.0;      // profile tile products / dense tile products
    double density = 1.0;         // camera-graph block density (off-diagonal)
};
// Builds the camera graph; false (dense) when it is too small, too large, or
// denser than max_density (the scan stops there).  Otherwise fills out.
bool plan_envelope(const gba_problem_view& prob, double max_density, EnvelopeLayout& out);
// Per-column TRSM / update tables of the profile for launch_batch_cholesky
// (one problem, index 0; update tile 0 of every column first).
void envelope_factor_tables(const std::vector<int>& first, std::vector<int2>& trsm, std::vector<int>& trsm_off,
                            std::vector<int2>& upd, std::vector<int>& upd_off);

// Result of one enqueued-and-synchronised LM trial.
struct TrialOut {
    TrialFlags flags;
    double trial_sse = 0.0;
    double denom = 0.0;
    unsigned long long anomalies = 0;
    int refine_iters = 0;     // mixed: inner CG / PCG steps, all solves of the trial
    double refine_rel = 0.0;  // mixed: final |D (rhs - S x)| / |D rhs| of the FP64 system
    int refine_outer = 0;     // mixed: FP64 residual evaluations
    double refine_floor = 0.0;  // mixed: rounding level of that residual, relative to |D rhs|
    bool refined_by_factor = false;  // mixed: the FP32 factor preconditioned this trial
};

class Solver {
public:
    Solver(const gba_problem_view& prob, const gba_solver_config& cfg);
    ~Solver();
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    // Uploads the parameters the solver was created with (or the given ones).
    void reset_params(const double* cams = nullptr, const double* pts = nullptr);
    // Linearisation at the current parameters; returns SSE / anomalies.
    void linearize(double* sse, unsigned long long* anomalies);
    // One LM trial at lambda (enqueue + synchronise).  When hook_S is set, S
    // (column-major lower, unscaled, before factorisation) and rhs are copied
    // out after assembly.
    TrialOut trial(double lambda, std::vector<double>* hook_S = nullptr,
                   std::vector<double>* hook_rhs = nullptr);
    // Residual-only cost at the current parameters.
    void cost(double* sse, unsigned long long* anomalies);
    void accept_trial();  // the trial parameters become current
    void solve(const gba_lm_options& opts, gba_lm_result* res, double t_entry);
    void download_params(double* cams, double* pts);
    // Test helpers (host arrays in the reference layout).
    void download_linearization(double* U, double* V, double* W, double* bc, double* bp);
    void download_step(double* dc, double* dp);
    double time_kernel(int kernel, double lambda, int reps);
    double pair_count() const { return pair_count_; }
    // GBA_LAYOUT_DENSE or GBA_LAYOUT_ENVELOPE, and the plan's figures
    void layout(int* lay, double* work_ratio, double* density) const {
        *lay = env_ ? GBA_LAYOUT_ENVELOPE : GBA_LAYOUT_DENSE;
        *work_ratio = env_work_;
        *density = env_density_;
    }
    int n() const { return pb_.n; }
    // mixed refinement preconditioner of the handle (enum gba_mixed_precond)
    void heavy(int* k, int* np, int* w) const {
        *k = pb_.heavy_k;
        *np = nheavy_;
        *w = heavy_kp_;
    }
    int mixed_precond() const { return !mixed_ ? 0 : (cg_mode_ ? GBA_PRECOND_JACOBI : GBA_PRECOND_FACTOR); }
    int npad() const { return pb_.npad; }

private:
    void enqueue_trial_linear_algebra(double lambda, std::vector<double>* hook_S,
                                      std::vector<double>* hook_rhs, TrialOut& out);
    void sync_status();
    // fac_copy: also write the FP32 factor copy of the scaled S (mixed; the
    // block-Jacobi refinement reads only the lower copy)
    // mixed: the scaled S into the lower copy (the inner solver's operator),
    // and into the factor copy when fac_copy
    void enqueue_schur(double lambda, bool fac_copy = true);
    void enqueue_factor();  // the dense factor of S (and X = L^-1 in mixed precision)
    // Mixed, block-Jacobi CG (k_cg.hip) on the FP32-stored scaled system:
    // solves A x = b (A the scaled S, camera-block inverses from
    // launch_cg_block_inverse of the trial); returns the steps, *rel = |r| /
    // |b|, *converged when |r| <= tol |b| (or the trial failed before the
    // solve, which needs no refinement); *last sizes the first batch of steps
    // and receives the count.
    void setup_cg();
    int cg_solve(const double* b, double* x, double tol, int max_it, int* last, double* rel, bool* converged);
    // Mixed: the inner solve of the FP32-stored system (block-Jacobi CG, or
    // PCG with the FP32 factor), then FP64 iterative refinement: the residual
    // of the FP64 reduced camera system (k_refine.hip) against dc_, each
    // correction solved by the same inner solver, until |D (rhs - S dc)| <=
    // tol |D rhs|, the corrections are spent, or the residual stops
    // shrinking.
    void mixed_solve(double lambda, TrialOut& out);
    bool cg_active() const { return cg_mode_ && !cg_off_; }

    gba_solver_config cfg_;
    bool mixed_;
    bool sep_diag_ = true;  // diagonal Schur blocks + rhs from launch_schur_diag
    DevProblem pb_{};
    DevBuf<int> cspt_;       // camera-ordered copies for the camera passes
    DevBuf<double2> cspx_;
    int nblk_ = 0;
    hipStream_t st_ = nullptr;
    std::vector<double> init_cams_, init_pts_;
    double pair_count_ = 0.0;
    int cur_ = 0;  // index of the current parameter buffers

    DevBuf<int> scam_, spt_, perm_, pt_off_, cam_list_, cam_off_, cam_chunk_off_;
    SchurShape sshape_;
    DevBuf<SchurTask> stasks_;
    DevBuf<int2> spairs_;
    int num_stasks_ = 0;
    // heavy points (pb_.heavy_k > 0): their list, the padded width of Y (3 per
    // point) and Y itself, k-major (FP32 in the mixed form)
    DevBuf<int> hpts_;
    int nheavy_ = 0, heavy_kp_ = 0, heavy_ld_ = 0;
    DevBuf<double> Yt64_;
    DevBuf<float> Yt32_;
    DevBuf<double2> spx_;
    DevBuf<unsigned> spx16_;  // FP16 storage: packed pixel pair per sorted observation
    DevBuf<float> cam_scale_;
    DevBuf<int2> units_;
    DevBuf<int4> chunks_;
    DevBuf<double> cams_[2], pts_[2];
    DevBuf<double> W64_;
    DevBuf<float> W32_;
    DevBuf<double> VB_, Vinv_, U_, bc_;
    DevBuf<double> unit_sse_, unit_den_, gram_, blk_sse_;
    DevBuf<unsigned> unit_anom_, blk_anom_;
    DevBuf<double> S64_, D64_;
    DevBuf<float> S32_, S32low_, D32_, X32_, T32_;
    DevBuf<double> rhs_, y_, x_, dc_, dscale_, pcgw_;
    int last_refine_ = 3;  // refinement steps of the last trial (first batch of the next)
    // block-Jacobi refinement: chosen for the handle (cg_mode_), switched off
    // for the rest of a solve after a trial missed the tolerance (cg_off_)
    bool cg_mode_ = false, cg_off_ = false;
    int cg_last_ = 40;       // steps of the last block-Jacobi first solve
    int cg_last_corr_ = 12;  // ... and of its last correction solve
    int pcg_last_corr_ = 2;
    // FP64 refinement: the residual D (rhs - S dc) and the magnitudes of its
    // terms (2 npad), the correction, the point terms and camera rotation
    // terms of the residual passes; host copy of the residual
    DevBuf<double> ref_res_, ref_dx_, ref_z_, ref_rot_;
    DevBuf<double> neg_one_;  // npad entries of -1 (x -= d)
    double* h_res_ = nullptr;  // two slots of residual + magnitudes (4 npad)
    CgArgs cg_{};
    DevBuf<int2> cg_wgtab_;
    DevBuf<int> cg_rowstart_;
    DevBuf<double> cg_minv_, cg_vec_, cg_rp_, cg_cp_, cg_dpart_, cg_gpart_, cg_st_;
    // gauge deflation of the block-Jacobi CG (cg_.zd set when on)
    DevBuf<double> cg_defl_;
    CgDeflWork cg_dw_{};
    // mixed: precondition by the factor's triangular solves instead of the
    // explicit inverse (large matrices; see kInversePrecondBlk)
    bool precond_trsv_ = false;
    PcgPrecond precond() const {
        PcgPrecond M;
        if (!precond_trsv_) {
            M.X32 = X32_.p;
        } else {
            M.L32 = S32_.p;
            M.D32 = D32_.p;
            M.nblk = nblk_;
            M.tmp = y_.p;
            M.sync = sync_.p;
        }
        return M;
    }
    DevBuf<int> sync_;
    // one-launch factor (GBA_FACTOR=persistent; the default is the launch
    // per tile column): per-column first tickets, flags, grid
    bool persist_ = false;
    DevBuf<long long> pcol_;
    DevBuf<int> psync_;
    long long pntasks_ = 0;
    int pgrid_ = 1;
    // envelope layout: cameras renumbered (internal index = pos_of_cam_[caller
    // camera]), profile per tile row, factor tables for launch_batch_cholesky
    bool env_ = false;
    double env_work_ = 1.0, env_density_ = 1.0;
    std::vector<int> cam_of_pos_, pos_of_cam_;
    DevBuf<int> first_, eactive_;
    DevBuf<int2> epotrf_, etrsm_, eupd_;
    std::vector<int> etrsm_off_, eupd_off_;
    DevBuf<BatchDesc> edesc_;
    BatchFactorTables etab_;
    // caller camera order <-> internal order for 9-value camera records
    void cams_to_internal(const double* in, double* out) const;
    void cams_to_caller(const double* in, double* out) const;
    DevBuf<TrialFlags> flags_;
    DevBuf<PassStatus> status_;
    PassStatus* h_status_ = nullptr;
};

}  // namespace gba_amd
