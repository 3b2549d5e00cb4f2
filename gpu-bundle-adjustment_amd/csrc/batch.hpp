// batch.hpp -- batched LM solver: many independent problems advanced together.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "solver.hpp"

namespace gba_amd {

// Solves a batch of independent bundle-adjustment problems, each with the
// reference LM semantics (src/lm.cu:455-714, the same options for all), in
// plain FP64 or in the mixed form (mixed set): W stored FP32, and the damped
// reduced camera system assembled diagonally scaled, D S D with D =
// (diag(U + lambda diag'(U)))^-1/2, from 64-bit fixed-point LDS sums (2^-46,
// one order-independent result), stored and factored in FP64; the step from
// it carries the FP32 rounding of W (percents where lambda is small: the
// damped system is near singular along the gauge freedom), so refine_steps
// FP64 refinement steps follow, each solving the residual of the FP64
// system (k_refine.hip: the Jacobians evaluated again, W not read) with the
// same factor.  All problems live in one concatenated device layout (cameras,
// points and sorted observations side by side); every stage of an LM round is
// one launch for the whole batch, masked per problem, so the launch cost of a
// round is shared by all problems.  A round re-linearises the problems whose
// last trial was accepted and evaluates one trial for every problem still
// running; each problem's accept / reject / lambda / nu rules run on the
// device (k_batch_lm_*), so a round needs no host decision.  The round's
// launch sequence is captured once into a HIP graph (GBA_BATCH_GRAPH=0:
// launched one by one) and replayed rounds_per_sync times between two reads
// of the running-problem count (GBA_BATCH_ROUNDS_PER_SYNC, default 4).  With
// autotune set, the constructor times one trial of the whole batch with each
// dense factor form (tile-parallel across problems, or one workgroup per
// problem) and keeps the faster; GBA_BATCH_FACTOR=tiles or =workgroup fixes
// the form instead.
class BatchSolver {
public:
    BatchSolver(const gba_problem_view* probs, int count, bool autotune = false, bool mixed = false);
    ~BatchSolver();
    BatchSolver(const BatchSolver&) = delete;
    BatchSolver& operator=(const BatchSolver&) = delete;

    // Solves every problem from its initial parameters; res[count].
    void solve(const gba_lm_options& opts, gba_lm_result* res);
    void download_params(int i, double* cams, double* pts);
    int count() const { return B_; }
    int rounds() const { return rounds_; }
    bool tile_factor() const { return tile_factor_; }
    bool graphs() const { return use_graphs_; }
    bool mixed() const { return mixed_; }
    // trial milliseconds measured by the autotuner: tile-parallel factor,
    // workgroup factor (0 when it did not run)
    double tune_ms(int form) const { return tune_ms_[form & 1]; }

private:
    void upload_mask(const std::vector<int>& h, int* d);
    // the launch sequences of a round, masked by pl.active
    void enqueue_lin(const DevProblem& pl);
    void enqueue_trial(const DevProblem& pl, bool tile_factor);
    // one whole round with the device-side LM rules
    void enqueue_round(const gba_lm_options& opts);
    // runs a sequence through its graph (captured on first use) or eagerly
    template <class F>
    void run_sequence(hipGraphExec_t& exec, F&& enqueue);
    void autotune_factor();

    int B_ = 0;
    int rounds_ = 0;
    DevProblem pb_{};
    hipStream_t st_ = nullptr;
    std::vector<BatchDesc> hdesc_;
    std::vector<long long> pt0_;  // first point of each problem, and the total
    std::vector<double> init_cams_, init_pts_;
    SchurShape sshape_;
    int num_stasks_ = 0;

    DevBuf<int> scam_, spt_, pt_off_, cam_list_, cam_off_, cam_chunk_off_, cam_prob_, pt_prob_;
    DevBuf<double2> spx_, cspx_;
    DevBuf<int> cspt_;  // camera-ordered copies for the camera passes
    DevBuf<int2> units_, spairs_;
    DevBuf<int4> chunks_;
    DevBuf<SchurTask> stasks_;
    DevBuf<BatchDesc> desc_;
    DevBuf<double> cams_, pts_, tcams_, tpts_;
    DevBuf<double> W_, VB_, Vinv_, U_, bc_, rhs_, dc_, S_, D_;
    // mixed form: FP32 W, the diagonal scale D and the scaled step D^-1 dc
    bool mixed_ = false;
    int refine_steps_ = 1;  // FP64 refinement steps per trial (GBA_BATCH_REFINE_STEPS)
    DevBuf<float> W32_;
    DevBuf<double> dscale_, ys_, zp_, rot_;
    DevBuf<double> unit_sse_, unit_den_, gram_;
    DevBuf<unsigned> unit_anom_;
    DevBuf<double> lam_;
    DevBuf<int> active_, accept_, lin_active_, trial_active_, counters_;
    DevBuf<BatchLmState> lm_;
    DevBuf<TrialFlags> flags_;
    DevBuf<PassStatus> lin_status_, trial_status_;
    // tile-parallel factor (launch_batch_cholesky) and its work tables
    bool tile_factor_ = true;
    DevBuf<int2> fpotrf_, ftrsm_, fupd_;
    std::vector<int> ftrsm_off_, fupd_off_;
    BatchFactorTables ftab_;
    double tune_ms_[2] = {0.0, 0.0};
    bool use_graphs_ = true;
    int rounds_per_sync_ = 4;
    hipGraphExec_t round_exec_ = nullptr;
    gba_lm_options round_opts_{};  // the options the round graph was captured with
    int* h_counters_ = nullptr;
    double clock_hz_ = 0.0;  // device wall clock rate
    PassStatus* h_status_ = nullptr;  // 2 B entries: linearisation, trial
    double* h_lam_ = nullptr;
    int* h_mask_ = nullptr;           // staging for the per-problem switches
};

}  // namespace gba_amd
