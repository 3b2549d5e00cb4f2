// batch.cpp -- batched LM solver (see batch.hpp).  The per-problem control
// follows Solver::solve, i.e. the reference's src/lm.cu:455-714, problem by
// problem; only the device work is shared.
#include "batch.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

#include "launchers.hpp"

namespace gba_amd {

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename T>
void upload(DevBuf<T>& b, const std::vector<T>& v, hipStream_t st) {
    b.alloc(std::max<size_t>(1, v.size()));
    if (!v.empty()) GBA_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
}

// Largest padded order the one-workgroup-per-problem factor handles.
constexpr int kBatchMaxNpad = 1024;

}  // namespace

BatchSolver::BatchSolver(const gba_problem_view* probs, int count, bool autotune, bool mixed)
    : B_(count), mixed_(mixed) {
    if (count <= 0) throw Error("batch: no problems");
    int Cmax = 1;
    for (int b = 0; b < count; ++b) Cmax = std::max(Cmax, probs[b].num_cameras);
    sshape_ = schur_shape(Cmax, mixed);
    // ---- per-problem descriptors and offsets (host) ----
    std::vector<long long> o_off(1, 0);
    std::vector<int> c0s, p0s, cam_prob, pt_prob;
    long long C0 = 0, P0 = 0, O0 = 0, soff = 0, doff = 0;
    for (int b = 0; b < count; ++b) {
        const gba_problem_view& pr = probs[b];
        const int C = pr.num_cameras, P = pr.num_points, O = pr.num_observations;
        if (C <= 0 || P <= 0 || O <= 0) throw Error("batch: problem sizes must be positive");
        for (int e = 0; e < O; ++e)
            if (pr.obs_camera[e] < 0 || pr.obs_camera[e] >= C || pr.obs_point[e] < 0 || pr.obs_point[e] >= P)
                throw Error("batch: observation index out of range");
        const int n = 9 * C, npad = (n + kTile - 1) / kTile * kTile;
        if (npad > kBatchMaxNpad) throw Error("batch: a problem has more than 113 cameras");
        if (C0 + C > (1LL << 30) || P0 + P > (1LL << 30) || O0 + O > (1LL << 30))
            throw Error("batch: too large");
        BatchDesc d;
        d.c0 = static_cast<int>(C0);
        d.c1 = static_cast<int>(C0 + C);
        d.u0 = d.u1 = 0;  // below, from the point CSR
        d.soff = soff;
        d.doff = doff;
        d.n = n;
        d.npad = npad;
        hdesc_.push_back(d);
        c0s.push_back(static_cast<int>(C0));
        p0s.push_back(static_cast<int>(P0));
        cam_prob.insert(cam_prob.end(), static_cast<size_t>(C), b);
        pt_prob.insert(pt_prob.end(), static_cast<size_t>(P), b);
        init_cams_.insert(init_cams_.end(), pr.cameras, pr.cameras + 9 * static_cast<size_t>(C));
        init_pts_.insert(init_pts_.end(), pr.points, pr.points + 3 * static_cast<size_t>(P));
        pt0_.push_back(P0);
        C0 += C;
        P0 += P;
        O0 += O;
        o_off.push_back(O0);
        soff += static_cast<long long>(npad) * npad;
        doff += 2LL * (npad / kTile) * kTile * kTile;
    }
    pt0_.push_back(P0);
    const int C = static_cast<int>(C0), P = static_cast<int>(P0), O = static_cast<int>(O0);

    // ---- the whole batch's point-major layout on the device: the problems'
    // observations side by side with batch indices, sorted by (point, camera,
    // index) -- points are numbered problem by problem, so each problem's
    // observations stay contiguous and in its own sorted order ----
    GBA_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    std::vector<int> pt_off, cam_off;
    {
        DevBuf<int> ocam, opt, d_c0, d_p0, perm;
        DevBuf<double> opx;
        DevBuf<long long> d_ooff;
        ocam.alloc(static_cast<size_t>(O));
        opt.alloc(static_cast<size_t>(O));
        opx.alloc(2 * static_cast<size_t>(O));
        for (int b = 0; b < count; ++b) {
            const gba_problem_view& pr = probs[b];
            const size_t e0 = static_cast<size_t>(o_off[b]), no = static_cast<size_t>(pr.num_observations);
            GBA_HIP(hipMemcpyAsync(ocam.p + e0, pr.obs_camera, sizeof(int) * no, hipMemcpyHostToDevice, st_));
            GBA_HIP(hipMemcpyAsync(opt.p + e0, pr.obs_point, sizeof(int) * no, hipMemcpyHostToDevice, st_));
            GBA_HIP(hipMemcpyAsync(opx.p + 2 * e0, pr.obs_pixel, sizeof(double) * 2 * no, hipMemcpyHostToDevice,
                                   st_));
        }
        upload(d_ooff, o_off, st_);
        upload(d_c0, c0s, st_);
        upload(d_p0, p0s, st_);
        launch_offset_obs(st_, O, count, d_ooff.p, d_c0.p, d_p0.p, ocam.p, opt.p);
        build_layout_device(st_, C, P, O, ocam.p, opt.p, opx.p, perm, scam_, spt_, &spx_, pt_off_, cam_off_,
                            cam_list_, &pb_.repeated_obs);
        pt_off.resize(static_cast<size_t>(P) + 1);
        cam_off.resize(static_cast<size_t>(C) + 1);
        GBA_HIP(hipMemcpyAsync(pt_off.data(), pt_off_.p, sizeof(int) * pt_off.size(), hipMemcpyDeviceToHost, st_));
        GBA_HIP(hipMemcpyAsync(cam_off.data(), cam_off_.p, sizeof(int) * cam_off.size(), hipMemcpyDeviceToHost,
                               st_));
        GBA_HIP(hipStreamSynchronize(st_));
    }
    // work units per problem (a unit never spans two problems), camera chunks
    std::vector<int2> units;
    std::vector<int4> chunks;
    std::vector<int> cam_chunk_off;
    for (int b = 0; b < count; ++b) {
        double pc = 0.0;
        hdesc_[b].u0 = static_cast<int>(units.size());
        layout_units_append(pt_off.data() + pt0_[b], static_cast<int>(pt0_[b + 1] - pt0_[b]), units, pc);
        hdesc_[b].u1 = static_cast<int>(units.size());
    }
    layout_chunks(cam_off, chunks, cam_chunk_off);
    // Schur tasks of every problem (column chunks of cb cameras from the
    // problem's first camera, blocks b < a), pair lists on the device
    std::vector<int4> tq;
    std::vector<int> tprob;
    for (int b = 0; b < count; ++b)
        for (int a = hdesc_[b].c0; a < hdesc_[b].c1; ++a)
            for (int lo = hdesc_[b].c0; lo < a; lo += sshape_.cb) {
                tq.push_back(make_int4(a, lo, std::min(lo + sshape_.cb, a), 0));
                tprob.push_back(b);
            }
    std::vector<long long> bounds;
    schur_pairs_device(st_, tq, cam_off, cam_off_.p, cam_list_.p, spt_.p, pt_off_.p, scam_.p, bounds, spairs_);
    std::vector<SchurTask> stasks(tq.size());
    for (size_t t = 0; t < tq.size(); ++t) {
        const BatchDesc& d = hdesc_[tprob[t]];
        stasks[t] = SchurTask{tq[t].x, tq[t].y, tq[t].z, d.c0, bounds[t], bounds[t + 1], d.soff, d.npad, 0};
    }
    // longest Schur tasks first across the whole batch
    std::stable_sort(stasks.begin(), stasks.end(), [](const SchurTask& x, const SchurTask& y) {
        return x.end - x.begin > y.end - y.begin;
    });
    num_stasks_ = static_cast<int>(stasks.size());
    upload(units_, units, st_);
    upload(chunks_, chunks, st_);
    upload(cam_chunk_off_, cam_chunk_off, st_);
    upload(cam_prob_, cam_prob, st_);
    upload(pt_prob_, pt_prob, st_);
    upload(stasks_, stasks, st_);
    upload(desc_, hdesc_, st_);
    {
        // work tables of the tile-parallel factor; at each step the lookahead
        // tiles (update tile 0, which factors the next diagonal tile) of all
        // problems come first
        const char* env = std::getenv("GBA_BATCH_FACTOR");
        tile_factor_ = !(env && std::string(env) == "workgroup");
        if (env && (std::string(env) == "workgroup" || std::string(env) == "tiles")) autotune = false;
        int nsteps = 0;
        for (const BatchDesc& d : hdesc_) nsteps = std::max(nsteps, d.npad / kTile);
        std::vector<int2> potrf, trsm, upd;
        for (int q = 0; q < B_; ++q) potrf.push_back(make_int2(q, 0));
        ftrsm_off_.assign(1, 0);
        fupd_off_.assign(1, 0);
        for (int k = 0; k + 1 < nsteps; ++k) {
            for (int q = 0; q < B_; ++q) {
                const int m = hdesc_[q].npad / kTile - k - 1;
                for (int i = 0; i < m; ++i) trsm.push_back(make_int2(q, i));
                if (m > 0) upd.push_back(make_int2(q, 0));
            }
            for (int q = 0; q < B_; ++q) {
                const int m = hdesc_[q].npad / kTile - k - 1;  // negative once q is factored
                if (m <= 0) continue;
                for (int t = 1; t < m * (m + 1) / 2; ++t) upd.push_back(make_int2(q, t));
            }
            ftrsm_off_.push_back(static_cast<int>(trsm.size()));
            fupd_off_.push_back(static_cast<int>(upd.size()));
        }
        upload(fpotrf_, potrf, st_);
        upload(ftrsm_, trsm, st_);
        upload(fupd_, upd, st_);
        ftab_.potrf0 = fpotrf_.p;
        ftab_.trsm = ftrsm_.p;
        ftab_.upd = fupd_.p;
        ftab_.trsm_off = ftrsm_off_.data();
        ftab_.upd_off = fupd_off_.data();
        ftab_.nsteps = nsteps;
        GBA_HIP(hipStreamSynchronize(st_));  // the tables above are released here
    }
    pb_.C = C;
    pb_.P = P;
    pb_.O = O;
    pb_.n = 9 * C;
    pb_.npad = 0;  // per problem (BatchDesc)
    pb_.num_units = static_cast<int>(units.size());
    pb_.num_chunks = static_cast<int>(chunks.size());
    pb_.scam = scam_.p;
    pb_.spt = spt_.p;
    pb_.spx = spx_.p;
    pb_.perm = nullptr;
    pb_.pt_off = pt_off_.p;
    pb_.units = units_.p;
    pb_.cam_list = cam_list_.p;
    pb_.cam_off = cam_off_.p;
    pb_.chunks = chunks_.p;
    pb_.cam_chunk_off = cam_chunk_off_.p;
    pb_.cam_prob = cam_prob_.p;
    pb_.pt_prob = pt_prob_.p;
    build_cam_order_copies(st_, O, cam_list_.p, spt_.p, spx_.p, cspt_, cspx_);
    pb_.cspt = cspt_.p;
    pb_.cspx = cspx_.p;
    cams_.alloc(9 * static_cast<size_t>(C));
    tcams_.alloc(9 * static_cast<size_t>(C));
    pts_.alloc(3 * static_cast<size_t>(P));
    tpts_.alloc(3 * static_cast<size_t>(P));
    if (mixed_) {
        W32_.alloc(kWStride * static_cast<size_t>(O));
        dscale_.alloc(9 * static_cast<size_t>(C));
        ys_.alloc(9 * static_cast<size_t>(C));
        zp_.alloc(3 * static_cast<size_t>(P));
        rot_.alloc(4 * static_cast<size_t>(C));
        if (const char* env = std::getenv("GBA_BATCH_REFINE_STEPS")) refine_steps_ = std::max(0, std::atoi(env));
    } else {
        W_.alloc(kWStride * static_cast<size_t>(O));
    }
    VB_.alloc(9 * static_cast<size_t>(P));
    Vinv_.alloc(6 * static_cast<size_t>(P));
    U_.alloc(81 * static_cast<size_t>(C));
    bc_.alloc(9 * static_cast<size_t>(C));
    rhs_.alloc(9 * static_cast<size_t>(C));
    dc_.alloc(9 * static_cast<size_t>(C));
    S_.alloc(static_cast<size_t>(soff));
    D_.alloc(static_cast<size_t>(doff));
    const size_t nu = std::max<size_t>(1, units.size());
    unit_sse_.alloc(nu);
    unit_den_.alloc(nu);
    unit_anom_.alloc(nu);
    gram_.alloc(55 * std::max<size_t>(1, chunks.size()));
    lam_.alloc(count);
    active_.alloc(count);
    accept_.alloc(count);
    lin_active_.alloc(count);
    trial_active_.alloc(count);
    lm_.alloc(count);
    counters_.alloc(2);
    flags_.alloc(count);
    lin_status_.alloc(count);
    trial_status_.alloc(count);
    pb_.lam = lam_.p;
    GBA_HIP(hipHostMalloc(&h_status_, 2 * sizeof(PassStatus) * count, hipHostMallocDefault));
    GBA_HIP(hipHostMalloc(&h_lam_, sizeof(double) * count, hipHostMallocDefault));
    GBA_HIP(hipHostMalloc(&h_mask_, sizeof(int) * count, hipHostMallocDefault));
    GBA_HIP(hipHostMalloc(&h_counters_, 2 * sizeof(int), hipHostMallocDefault));
    // a point without observations is in no work unit, so the point pass
    // never writes its V and bp: they stay 0 (its damped V is then
    // lambda 1e-12 I, SPD for lambda > 0, and its step is 0, as in the
    // reference, src/lm.cu:289-301)
    launch_fill_double(st_, VB_.p, VB_.n, 0.0);
    launch_fill_double(st_, Vinv_.p, Vinv_.n, 0.0);
    // S: zero, then the identity on each problem's padding
    launch_fill_double(st_, S_.p, static_cast<size_t>(soff), 0.0);
    for (const BatchDesc& d : hdesc_)
        if (d.npad > d.n) launch_identity_pad(st_, S_.p + d.soff, nullptr, d.npad, d.n, d.npad);
    GBA_HIP(hipGetLastError());
    GBA_HIP(hipStreamSynchronize(st_));
    {
        int dev = 0, khz = 0;
        GBA_HIP(hipGetDevice(&dev));
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
            clock_hz_ = 1e3 * khz;
        (void)hipGetLastError();
    }
    if (const char* env = std::getenv("GBA_BATCH_GRAPH")) use_graphs_ = std::string(env) != "0";
    if (const char* env = std::getenv("GBA_BATCH_ROUNDS_PER_SYNC")) rounds_per_sync_ = std::max(1, std::atoi(env));
    if (autotune) autotune_factor();
}

BatchSolver::~BatchSolver() {
    if (round_exec_) (void)hipGraphExecDestroy(round_exec_);
    if (h_counters_) (void)hipHostFree(h_counters_);
    if (h_status_) (void)hipHostFree(h_status_);
    if (h_lam_) (void)hipHostFree(h_lam_);
    if (h_mask_) (void)hipHostFree(h_mask_);
    if (st_) (void)hipStreamDestroy(st_);
}

void BatchSolver::upload_mask(const std::vector<int>& h, int* d) {
    // pinned staging, then a stream-ordered copy; the caller synchronises
    // before the staging buffer is rewritten
    std::memcpy(h_mask_, h.data(), sizeof(int) * B_);
    GBA_HIP(hipMemcpyAsync(d, h_mask_, sizeof(int) * B_, hipMemcpyHostToDevice, st_));
    GBA_HIP(hipStreamSynchronize(st_));
}

void BatchSolver::enqueue_lin(const DevProblem& pl) {
    launch_lin_points(st_, pl, cams_.p, pts_.p, W_.p, W32_.p, VB_.p, unit_sse_.p, unit_anom_.p);
    launch_lin_cameras(st_, pl, cams_.p, pts_.p, gram_.p);
    launch_lin_reduce(st_, pl, gram_.p, unit_sse_.p, unit_anom_.p, U_.p, bc_.p, nullptr);
    launch_batch_status(st_, B_, desc_.p, 0, unit_sse_.p, unit_den_.p, unit_anom_.p, dc_.p, U_.p, bc_.p,
                        lam_.p, flags_.p, pl.active, lin_status_.p);
}

void BatchSolver::enqueue_trial(const DevProblem& pl, bool tile_factor) {
    launch_fill_int(st_, reinterpret_cast<int*>(flags_.p), 4 * B_, 0);
    launch_damp_invert_v_batched(st_, pb_.P, VB_.p, pt_prob_.p, lam_.p, pl.active, Vinv_.p, flags_.p);
    // mixed: S and rhs come out scaled (D S D, D rhs); the factor solves for
    // y = D^-1 dc, then dc = D y
    const int n = 9 * pb_.C;
    if (mixed_) launch_diag_scale(st_, n, n, U_.p, 0.0, dscale_.p, cam_prob_.p, lam_.p);
    launch_schur_diag(st_, pl, desc_.p, W_.p, W32_.p, Vinv_.p, VB_.p, U_.p, bc_.p, 0.0, dscale_.p, S_.p,
                      nullptr, nullptr, 0, rhs_.p, flags_.p);
    if (num_stasks_ > 0)
        launch_schur(st_, sshape_, pl, stasks_.p, num_stasks_, spairs_.p, W_.p, W32_.p, Vinv_.p, VB_.p,
                     U_.p, bc_.p, 0.0, dscale_.p, S_.p, nullptr, nullptr, 0, rhs_.p, flags_.p);
    double* y = mixed_ ? ys_.p : dc_.p;
    if (tile_factor) launch_batch_cholesky(st_, B_, ftab_, desc_.p, S_.p, D_.p, flags_.p, pl.active);
    launch_batch_chol_solve(st_, B_, desc_.p, S_.p, D_.p, rhs_.p, y, flags_.p, pl.active, !tile_factor);
    if (mixed_) {
        launch_vec_scale(st_, n, n, ys_.p, dscale_.p, dc_.p);
        // FP64 refinement with the same factor: the residual of the FP64
        // system (Jacobians evaluated again, W not read), its scaled
        // correction solved, dc += D y
        for (int it = 0; it < refine_steps_; ++it) {
            launch_refine_residual(st_, pl, cams_.p, pts_.p, Vinv_.p, VB_.p, U_.p, bc_.p, 0.0, dc_.p, zp_.p,
                                   rot_.p, gram_.p, dscale_.p, rhs_.p, nullptr, flags_.p);
            launch_batch_chol_solve(st_, B_, desc_.p, S_.p, D_.p, rhs_.p, ys_.p, flags_.p, pl.active, false);
            launch_vec_axpy_scaled(st_, n, dc_.p, ys_.p, dscale_.p);
        }
    }
    launch_trial_cams(st_, pb_.C, cams_.p, dc_.p, tcams_.p);
    launch_backsub(st_, pl, W_.p, W32_.p, Vinv_.p, VB_.p, dc_.p, pts_.p, tcams_.p, tpts_.p, 0.0,
                   unit_sse_.p, unit_den_.p, unit_anom_.p, flags_.p);
    launch_batch_status(st_, B_, desc_.p, 1, unit_sse_.p, unit_den_.p, unit_anom_.p, dc_.p, U_.p, bc_.p,
                        lam_.p, flags_.p, pl.active, trial_status_.p);
}

void BatchSolver::enqueue_round(const gba_lm_options& opts) {
    DevProblem pl = pb_, pt = pb_;
    pl.active = lin_active_.p;
    pt.active = trial_active_.p;
    enqueue_lin(pl);
    launch_batch_lm_lin(st_, B_, opts, lin_status_.p, lm_.p, lin_active_.p, trial_active_.p, counters_.p);
    enqueue_trial(pt, tile_factor_);
    launch_batch_lm_trial(st_, B_, opts, trial_status_.p, lm_.p, trial_active_.p, accept_.p, lin_active_.p,
                          lam_.p, counters_.p);
    launch_batch_accept(st_, pb_.C, pb_.P, cam_prob_.p, pt_prob_.p, accept_.p, tcams_.p, tpts_.p, cams_.p,
                        pts_.p);
}

template <class F>
void BatchSolver::run_sequence(hipGraphExec_t& exec, F&& enqueue) {
    if (!use_graphs_) {
        enqueue();
        return;
    }
    if (!exec) {
        // every pointer and size in the sequence is fixed for the life of the
        // solver; what changes between rounds (lambda, the masks) lives in
        // device buffers that the host rewrites before each replay
        hipGraph_t g = nullptr;
        GBA_HIP(hipStreamBeginCapture(st_, hipStreamCaptureModeRelaxed));
        try {
            enqueue();
        } catch (...) {
            (void)hipStreamEndCapture(st_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        GBA_HIP(hipStreamEndCapture(st_, &g));
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        GBA_HIP(e);
    }
    GBA_HIP(hipGraphLaunch(exec, st_));
}

void BatchSolver::autotune_factor() {
    // every problem linearised at its initial parameters and damped with
    // lambda = 1e-3; one trial per form, twice, the faster time kept (the
    // trial writes only trial buffers; solve starts over from the uploads)
    GBA_HIP(hipMemcpyAsync(cams_.p, init_cams_.data(), sizeof(double) * init_cams_.size(),
                           hipMemcpyHostToDevice, st_));
    GBA_HIP(hipMemcpyAsync(pts_.p, init_pts_.data(), sizeof(double) * init_pts_.size(),
                           hipMemcpyHostToDevice, st_));
    for (int b = 0; b < B_; ++b) h_lam_[b] = 1e-3;
    GBA_HIP(hipMemcpyAsync(lam_.p, h_lam_, sizeof(double) * B_, hipMemcpyHostToDevice, st_));
    upload_mask(std::vector<int>(B_, 1), active_.p);
    DevProblem pl = pb_;
    pl.active = active_.p;
    enqueue_lin(pl);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    GBA_HIP(hipEventCreate(&e0));
    GBA_HIP(hipEventCreate(&e1));
    double best[2] = {1e300, 1e300};
    try {
        for (int rep = 0; rep < 2; ++rep)
            for (int form = 0; form < 2; ++form) {
                GBA_HIP(hipEventRecord(e0, st_));
                enqueue_trial(pl, form == 0);
                GBA_HIP(hipEventRecord(e1, st_));
                GBA_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                GBA_HIP(hipEventElapsedTime(&ms, e0, e1));
                best[form] = std::min(best[form], static_cast<double>(ms));
            }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tune_ms_[0] = best[0];
    tune_ms_[1] = best[1];
    tile_factor_ = best[0] <= best[1];
}

void BatchSolver::solve(const gba_lm_options& opts, gba_lm_result* res) {
    const double t0 = now_s();
    GBA_HIP(hipMemcpyAsync(cams_.p, init_cams_.data(), sizeof(double) * init_cams_.size(),
                           hipMemcpyHostToDevice, st_));
    // the trial points too: the back-substitution writes only observed points,
    // and an accept copies every point of its problem, so a point without
    // observations must already hold its value there (the reference leaves it
    // unchanged: dp = 0, src/lm.cu:289-301,645-646)
    for (double* dst : {pts_.p, tpts_.p})
        GBA_HIP(hipMemcpyAsync(dst, init_pts_.data(), sizeof(double) * init_pts_.size(), hipMemcpyHostToDevice,
                               st_));
    const int B = B_;
    // The LM rules run on the device (k_batch_lm_*): a round is one captured
    // sequence (re-linearise the problems that accepted, apply the rules,
    // evaluate a trial for every running problem, apply the rules, copy the
    // accepted parameters), replayed rounds_per_sync times between reads of
    // the running-problem count.  A round after every problem has finished
    // runs with every switch off.
    launch_batch_lm_init(st_, B, opts.initial_lambda, lm_.p, lam_.p, lin_active_.p, trial_active_.p, accept_.p,
                         counters_.p);
    if (round_exec_ && std::memcmp(&round_opts_, &opts, sizeof opts) != 0) {
        (void)hipGraphExecDestroy(round_exec_);
        round_exec_ = nullptr;
    }
    round_opts_ = opts;
    // an upper bound on the rounds any problem can need: every outer
    // iteration ends with an accept after at most max_inner_trials trials
    const long long max_rounds =
        static_cast<long long>(std::max(0, opts.max_iterations)) * std::max(1, opts.max_inner_trials) + 2;
    long long enq = 0;
    for (;;) {
        for (int r = 0; r < rounds_per_sync_ && enq < max_rounds; ++r, ++enq)
            run_sequence(round_exec_, [&] { enqueue_round(opts); });
        GBA_HIP(hipMemcpyAsync(h_counters_, counters_.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st_));
        GBA_HIP(hipGetLastError());
        GBA_HIP(hipStreamSynchronize(st_));
        if (h_counters_[1] == 0 || enq >= max_rounds) break;
    }
    rounds_ = h_counters_[0];
    std::vector<BatchLmState> hs(static_cast<size_t>(B));
    GBA_HIP(hipMemcpy(hs.data(), lm_.p, sizeof(BatchLmState) * B, hipMemcpyDeviceToHost));
    const double t1 = now_s();
    for (int b = 0; b < B; ++b) {
        const BatchLmState& s = hs[b];
        res[b] = gba_lm_result{};
        res[b].initial_cost = s.cost0;
        res[b].final_cost = s.sse / 2.0;
        res[b].iterations = s.iterations;
        res[b].trials = s.trials;
        res[b].anomalies = s.counter;
        res[b].termination = s.converged ? GBA_CONVERGENCE : (s.failed ? GBA_FAILURE : GBA_MAX_ITERATIONS);
        // this problem's own solve time: device clock from the start of the
        // batch solve to the round that finished it
        res[b].solve_time_s = (s.done && clock_hz_ > 0.0 && s.t_done >= s.t_start)
                                  ? static_cast<double>(s.t_done - s.t_start) / clock_hz_
                                  : t1 - t0;
        res[b].total_time_s = t1 - t0;
    }
}

void BatchSolver::download_params(int i, double* cams, double* pts) {
    if (i < 0 || i >= B_) throw Error("batch: problem index out of range");
    const BatchDesc& d = hdesc_[i];
    GBA_HIP(hipStreamSynchronize(st_));
    GBA_HIP(hipMemcpy(cams, cams_.p + 9 * static_cast<size_t>(d.c0), sizeof(double) * 9 * (d.c1 - d.c0),
                      hipMemcpyDeviceToHost));
    GBA_HIP(hipMemcpy(pts, pts_.p + 3 * static_cast<size_t>(pt0_[i]),
                      sizeof(double) * 3 * static_cast<size_t>(pt0_[i + 1] - pt0_[i]), hipMemcpyDeviceToHost));
}

}  // namespace gba_amd
