// solver.cpp -- host driver of the device-resident LM solve (replaces
// gba::solve_bal_gpu, reference src/lm.cu:371-724).
//
// Everything per iteration stays in HBM: linearisation, damped V inverse,
// Schur assembly, dense factor, solves, back-substitution and trial cost run
// as kernels on one stream; kernels after a failed stage exit early on device
// flags, so a trial is enqueued without host round trips and the host reads one
// small status block per trial (plus one per refinement step in mixed
// precision).  The accept / reject logic, lambda and nu follow the reference
// exactly (src/lm.cu:455-463, :672-714).
#include "solver.hpp"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "launchers.hpp"

namespace gba_amd {

void hip_check(hipError_t e, const char* what, const char* file, int line) {
    if (e != hipSuccess)
        throw Error(std::string("HIP error at ") + file + ":" + std::to_string(line) + ": " +
                    hipGetErrorString(e) + " (" + what + ")");
}

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename T>
void upload(DevBuf<T>& b, const std::vector<T>& v, hipStream_t st) {
    b.alloc(v.size());
    if (!v.empty()) GBA_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st));
}

}  // namespace

void schur_pairs_device(hipStream_t st, const std::vector<int4>& tq, const std::vector<int>& cam_off,
                        const int* d_cam_off, const int* d_cam_list, const int* d_spt, const int* d_pt_off,
                        const int* d_scam, std::vector<long long>& bounds, DevBuf<int2>& pairs, int heavy_k) {
    const int ntask = static_cast<int>(tq.size());
    bounds.assign(static_cast<size_t>(ntask) + 1, 0);
    pairs.release();
    if (ntask == 0) return;
    std::vector<long long> qoff(static_cast<size_t>(ntask) + 1, 0);
    for (int t = 0; t < ntask; ++t) qoff[t + 1] = qoff[t] + (cam_off[tq[t].x + 1] - cam_off[tq[t].x]);
    const long long nslots = qoff.back();
    if (nslots + 1 > INT_MAX) throw Error("too many observation slots for the pair scan");
    DevBuf<int4> d_tq;
    DevBuf<long long> d_qoff, cnt, pos, d_bounds;
    DevBuf<unsigned char> work;
    upload(d_tq, tq, st);
    upload(d_qoff, qoff, st);
    cnt.alloc(static_cast<size_t>(nslots) + 1);
    pos.alloc(static_cast<size_t>(nslots) + 1);
    d_bounds.alloc(static_cast<size_t>(ntask) + 1);
    const size_t wb = pair_scan_work(nslots + 1);
    work.alloc(std::max<size_t>(wb, 1));
    launch_pair_count(st, ntask, d_tq.p, d_qoff.p, nslots, d_cam_off, d_cam_list, d_spt, d_pt_off, d_scam, heavy_k,
                      cnt.p);
    launch_pair_scan(st, cnt.p, pos.p, nslots + 1, work.p, wb);
    launch_pair_bounds(st, ntask, d_qoff.p, pos.p, d_bounds.p);
    GBA_HIP(hipMemcpyAsync(bounds.data(), d_bounds.p, sizeof(long long) * bounds.size(), hipMemcpyDeviceToHost,
                           st));
    GBA_HIP(hipGetLastError());
    GBA_HIP(hipStreamSynchronize(st));
    if (bounds[ntask] > 0) {
        pairs.alloc(static_cast<size_t>(bounds[ntask]));
        launch_pair_fill(st, ntask, d_tq.p, d_qoff.p, nslots, d_cam_off, d_cam_list, d_spt, d_pt_off, d_scam,
                         heavy_k, pos.p, pairs.p);
    }
    GBA_HIP(hipGetLastError());
    GBA_HIP(hipStreamSynchronize(st));  // the temporaries are released on return
}

void build_schur_tasks_device(hipStream_t st, int C, int cb, const std::vector<int>& cam_off,
                              const int* d_cam_off, const int* d_cam_list, const int* d_spt,
                              const int* d_pt_off, const int* d_scam, std::vector<SchurTask>& tasks,
                              DevBuf<int2>& pairs, bool with_diag, int heavy_k) {
    // tasks ordered by camera a, then column chunk
    const int dg = with_diag ? 1 : 0;
    std::vector<int4> tq;
    for (int a = 0; a < C; ++a)
        for (int lo = 0; lo < a + dg; lo += cb) tq.push_back(make_int4(a, lo, std::min(lo + cb, a + dg), 0));
    std::vector<long long> hb;
    schur_pairs_device(st, tq, cam_off, d_cam_off, d_cam_list, d_spt, d_pt_off, d_scam, hb, pairs, heavy_k);
    const int ntask = static_cast<int>(tq.size());
    tasks.resize(static_cast<size_t>(ntask));
    for (int t = 0; t < ntask; ++t)
        tasks[t] = SchurTask{tq[t].x, tq[t].y, tq[t].z, 0, hb[t], hb[t + 1], 0, 0, 0};
    std::stable_sort(tasks.begin(), tasks.end(), [](const SchurTask& x, const SchurTask& y) {
        return x.end - x.begin > y.end - y.begin;
    });
}

size_t pcg_work(int npad) {
    return 5 * static_cast<size_t>(npad) + std::max(symv_work(npad), precond_work(npad)) +
           2 * static_cast<size_t>(residual_partials(npad));
}

int pcg_solve(hipStream_t st, const float* A, const PcgPrecond& M, int ld, int npad, const double* b,
              double* x, double* work, TrialFlags* flags, PassStatus* status, PassStatus* h_status,
              double tol, int max_it, double* rel, int first_batch) {
    const size_t np = static_cast<size_t>(npad);
    double* r = work;
    double* z = work + np;
    double* p = work + 2 * np;
    double* q = work + 3 * np;
    double* sc = work + 4 * np;  // device scalars (k_pcg_reduce)
    double* sw = work + 5 * np;
    double* blk = sw + std::max(symv_work(npad), precond_work(npad));
    auto sync_status = [&]() {
        GBA_HIP(hipGetLastError());
        GBA_HIP(hipMemcpyAsync(h_status, status, sizeof(PassStatus), hipMemcpyDeviceToHost, st));
        GBA_HIP(hipStreamSynchronize(st));
    };
    const double* done = sc + 4;
    auto precond = [&](const double* skip) {
        if (M.X32)
            launch_precond_apply(st, M.X32, ld, npad, r, z, sw, skip);
        else
            launch_chol_solve(st, nullptr, M.L32, ld, M.nblk, nullptr, M.D32, r, M.tmp, z, M.sync, flags, nullptr,
                              skip);
    };
    launch_fill_double(st, x, np, 0.0);
    launch_vec_copy(st, npad, b, r);
    precond(nullptr);
    launch_vec_copy(st, npad, z, p);
    launch_pcg_dot(st, npad, r, z, b, b, blk, 0, sc, status, flags, tol * tol, max_it);
    // Steps are enqueued in batches with one synchronisation per batch: the
    // stopping test runs on the device (k_pcg_reduce), and the steps of a
    // batch after it has fired are empty launches.  alpha = 0 on a step whose
    // p.q or r.z is not > 0 (x and r unchanged), which ends the loop uncounted.
    int enq = 0, batch = std::max(1, first_batch);
    for (;;) {
        const int nb = std::max(0, std::min(batch, max_it - enq));
        for (int k = 0; k < nb; ++k) {
            launch_symv(st, A, ld, npad, p, q, sw, done);
            launch_pcg_dot(st, npad, p, q, p, q, blk, 1, sc, status, flags, tol * tol, max_it);
            launch_pcg_update(st, npad, sc, x, p, r, q);
            precond(done);
            launch_pcg_dot(st, npad, r, z, r, r, blk, 2, sc, status, flags, tol * tol, max_it);
            launch_pcg_direction(st, npad, sc, z, p);
        }
        enq += nb;
        sync_status();
        if (h_status->pcg_done || enq >= max_it) break;
        batch = 1;
    }
    // a trial that failed before the solve runs no refinement: no residual
    const TrialFlags f = h_status->flags;
    *rel = 0.0;
    if (f.nonspd || f.chol_fail || f.solve_err) return 0;
    const double bb = h_status->pcg_bb;
    if (!(bb > 0.0)) return 0;
    *rel = std::sqrt(h_status->pcg_rr / bb);
    return h_status->pcg_it;
}

void layout_units(const std::vector<int>& pt_off, std::vector<int2>& units, double& pair_count) {
    units.clear();
    pair_count = 0.0;
    layout_units_append(pt_off.data(), static_cast<int>(pt_off.size()) - 1, units, pair_count);
}

void layout_units_append(const int* pt_off, int P, std::vector<int2>& units, double& pair_count) {
    int b = -1, e_end = -1;
    for (int p = 0; p < P; ++p) {
        const int k0 = pt_off[p], k1 = pt_off[p + 1], k = k1 - k0;
        if (k == 0) continue;
        pair_count += 0.5 * static_cast<double>(k) * (k + 1);
        if (k > kWave) {
            if (b >= 0) units.push_back(make_int2(b, e_end));
            units.push_back(make_int2(k0, k1));
            b = -1;
        } else if (b < 0) {
            b = k0;
            e_end = k1;
        } else if (k1 - b > kWave) {
            units.push_back(make_int2(b, e_end));
            b = k0;
            e_end = k1;
        } else {
            e_end = k1;
        }
    }
    if (b >= 0) units.push_back(make_int2(b, e_end));
}

void layout_chunks(const std::vector<int>& cam_off, std::vector<int4>& chunks, std::vector<int>& cam_chunk_off) {
    const int C = static_cast<int>(cam_off.size()) - 1;
    chunks.clear();
    cam_chunk_off.assign(static_cast<size_t>(C) + 1, 0);
    for (int c = 0; c < C; ++c) {
        cam_chunk_off[c] = static_cast<int>(chunks.size());
        for (int b = cam_off[c]; b < cam_off[c + 1]; b += kCamChunk)
            chunks.push_back(make_int4(c, b, std::min(b + kCamChunk, cam_off[c + 1]), 0));
    }
    cam_chunk_off[C] = static_cast<int>(chunks.size());
}

Solver::Solver(const gba_problem_view& prob, const gba_solver_config& cfg) : cfg_(cfg) {
    mixed_ = cfg.precision == GBA_PRECISION_MIXED;
    const int C = prob.num_cameras, P = prob.num_points, O = prob.num_observations;
    if (C <= 0 || P <= 0 || O <= 0) throw Error("problem sizes must be positive");
    if (static_cast<long long>(9) * C > (1 << 20)) throw Error("too many cameras for a dense S");
    for (int e = 0; e < O; ++e)
        if (prob.obs_camera[e] < 0 || prob.obs_camera[e] >= C || prob.obs_point[e] < 0 ||
            prob.obs_point[e] >= P)
            throw Error("observation index out of range");
    if (!(cfg.huber_delta >= 0.0) || !std::isfinite(cfg.huber_delta))
        throw Error("huber_delta must be finite and >= 0");
    pb_.huber = cfg.huber_delta;
    pb_.C = C;
    pb_.P = P;
    pb_.O = O;
    pb_.n = 9 * C;
    pb_.npad = (pb_.n + kTile - 1) / kTile * kTile;
    nblk_ = pb_.npad / kTile;
    // mixed refinement preconditioner: the explicit inverse up to
    // kInversePrecondBlk tile columns, the factor's triangular solves beyond
    // (GBA_PRECOND=inverse / =trsv in the environment overrides)
    if (mixed_) {
        precond_trsv_ = nblk_ > kInversePrecondBlk;
        if (const char* env = std::getenv("GBA_PRECOND")) {
            if (std::string(env) == "inverse") precond_trsv_ = false;
            else if (std::string(env) == "trsv") precond_trsv_ = true;
        }
    }

    sshape_ = schur_shape(C, mixed_);
    // diagonal blocks and rhs in their own per-camera kernel (GBA_SCHUR_DIAG=0:
    // inside the task lists, as pairs e == f)
    {
        const char* env = std::getenv("GBA_SCHUR_DIAG");
        sep_diag_ = !(env && env[0] == '0');
    }
    GBA_HIP(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    // factor layout: the envelope (profile) factor for plain FP64 solves of
    // sparse camera graphs, with the cameras renumbered; GBA_LAYOUT in the
    // environment (dense / envelope) overrides the configuration
    {
        int want = cfg.factor_layout;
        if (const char* env = std::getenv("GBA_LAYOUT")) {
            if (std::string(env) == "dense") want = GBA_LAYOUT_DENSE;
            else if (std::string(env) == "envelope") want = GBA_LAYOUT_ENVELOPE;
        }
        if (want == GBA_LAYOUT_ENVELOPE && mixed_) throw Error("the envelope factor layout is plain FP64 only");
        if (want != GBA_LAYOUT_DENSE && want != GBA_LAYOUT_ENVELOPE && want != GBA_LAYOUT_AUTO)
            throw Error("unknown factor layout");
        if (!mixed_ && want != GBA_LAYOUT_DENSE) {
            EnvelopeLayout lay;
            if (plan_envelope(prob, want == GBA_LAYOUT_ENVELOPE ? 1.0 : 0.5, lay)) {
                env_work_ = lay.work_ratio;
                env_density_ = lay.density;
                env_ = want == GBA_LAYOUT_ENVELOPE || lay.work_ratio < 0.5;
            }
            if (env_) {
                cam_of_pos_ = std::move(lay.cam_of_pos);
                pos_of_cam_ = std::move(lay.pos_of_cam);
                upload(first_, lay.first, st_);
                std::vector<int2> trsm, upd;
                envelope_factor_tables(lay.first, trsm, etrsm_off_, upd, eupd_off_);
                const int nblk = static_cast<int>(lay.first.size());
                upload(epotrf_, std::vector<int2>{make_int2(0, 0)}, st_);
                upload(etrsm_, trsm, st_);
                upload(eupd_, upd, st_);
                upload(eactive_, std::vector<int>{1}, st_);
                etab_.potrf0 = epotrf_.p;
                etab_.trsm = etrsm_.p;
                etab_.upd = eupd_.p;
                etab_.trsm_off = etrsm_off_.data();
                etab_.upd_off = eupd_off_.data();
                etab_.nsteps = nblk;
                BatchDesc d{};
                d.c0 = 0;
                d.c1 = C;
                d.n = 9 * C;
                d.npad = nblk * kTile;
                upload(edesc_, std::vector<BatchDesc>{d}, st_);
            }
        }
    }
    init_cams_.assign(9 * static_cast<size_t>(C), 0.0);
    cams_to_internal(prob.cameras, init_cams_.data());
    init_pts_.assign(prob.points, prob.points + 3 * static_cast<size_t>(P));

    // ---- point-major layout on the device: sort, CSR, camera index ----
    const bool fp16 = cfg.obs_storage == GBA_OBS_FP16;
    if (!fp16 && cfg.obs_storage != GBA_OBS_FP64) throw Error("unknown observation storage");
    std::vector<int> oc_internal;  // alive until the synchronisation below
    {
        DevBuf<int> ocam, opt;
        DevBuf<double> opx;
        ocam.alloc(static_cast<size_t>(O));
        opt.alloc(static_cast<size_t>(O));
        const int* oc = prob.obs_camera;
        if (env_) {
            oc_internal.resize(static_cast<size_t>(O));
            for (int e = 0; e < O; ++e) oc_internal[e] = pos_of_cam_[prob.obs_camera[e]];
            oc = oc_internal.data();
        }
        GBA_HIP(hipMemcpyAsync(ocam.p, oc, sizeof(int) * O, hipMemcpyHostToDevice, st_));
        GBA_HIP(hipMemcpyAsync(opt.p, prob.obs_point, sizeof(int) * O, hipMemcpyHostToDevice, st_));
        if (!fp16) {
            opx.alloc(2 * static_cast<size_t>(O));
            GBA_HIP(hipMemcpyAsync(opx.p, prob.obs_pixel, sizeof(double) * 2 * O, hipMemcpyHostToDevice, st_));
        }
        build_layout_device(st_, C, P, O, ocam.p, opt.p, opx.p, perm_, scam_, spt_, fp16 ? nullptr : &spx_,
                            pt_off_, cam_off_, cam_list_, &pb_.repeated_obs);
    }
    std::vector<int> pt_off(static_cast<size_t>(P) + 1), cam_off(static_cast<size_t>(C) + 1);
    GBA_HIP(hipMemcpyAsync(pt_off.data(), pt_off_.p, sizeof(int) * pt_off.size(), hipMemcpyDeviceToHost, st_));
    GBA_HIP(hipMemcpyAsync(cam_off.data(), cam_off_.p, sizeof(int) * cam_off.size(), hipMemcpyDeviceToHost, st_));
    GBA_HIP(hipStreamSynchronize(st_));
    std::vector<int2> units;
    layout_units(pt_off, units, pair_count_);
    std::vector<int4> chunks;
    std::vector<int> cam_chunk_off;
    layout_chunks(cam_off, chunks, cam_chunk_off);
    pb_.num_units = static_cast<int>(units.size());
    pb_.num_chunks = static_cast<int>(chunks.size());
    if (fp16) {
        // only the binary16 pixels and the camera scales reach the device,
        // packed in the sorted order
        std::vector<int> perm(static_cast<size_t>(O));
        GBA_HIP(hipMemcpyAsync(perm.data(), perm_.p, sizeof(int) * O, hipMemcpyDeviceToHost, st_));
        std::vector<float> scale(static_cast<size_t>(C));
        std::vector<uint16_t> h(2 * static_cast<size_t>(O));
        encode_obs_fp16(prob, scale.data(), h.data());
        GBA_HIP(hipStreamSynchronize(st_));
        std::vector<unsigned> packed(static_cast<size_t>(O));
        for (int k = 0; k < O; ++k) {
            const size_t e = static_cast<size_t>(perm[k]);
            packed[k] = static_cast<unsigned>(h[2 * e]) | (static_cast<unsigned>(h[2 * e + 1]) << 16);
        }
        upload(spx16_, packed, st_);
        if (env_) {
            std::vector<float> internal(scale.size());
            for (int c = 0; c < C; ++c) internal[c] = scale[cam_of_pos_[c]];
            scale.swap(internal);
        }
        upload(cam_scale_, scale, st_);
        pb_.spx16 = spx16_.p;
        pb_.cam_scale = cam_scale_.p;
    }
    upload(units_, units, st_);
    // Heavy points: those seen at least heavy_k times, when they hold a fifth
    // or more of the observation pairs, are assembled by one dense product on
    // MFMA (k_heavy.hip) instead of the pair lists.  The dense product costs
    // (9C)^2 x 3 / 2 multiply-adds per point at about half the MFMA peak,
    // whatever the point's k; the pair kernels cost k(k-1)/2 pairs.  Measured
    // at C=356 (assembly per trial, mixed / FP64): none 0.93 / 1.29 ms,
    // k >= 64 1.03 / 1.74, k >= 96 0.83 / 1.33, k >= 128 0.76 / 1.21,
    // k >= 160 0.755 / 1.15, k >= 200 0.79 / 1.13; hence 0.45 C (32 at
    // least).  GBA_SCHUR_HEAVY_K in the environment sets the threshold
    // (0: off) and skips the share test.
    {
        int kh = std::max(32, (9 * C + 19) / 20);
        bool forced = false;
        if (const char* env = std::getenv("GBA_SCHUR_HEAVY_K")) {
            kh = std::max(0, std::atoi(env));
            forced = true;
        }
        if (!sep_diag_ || env_) kh = 0;  // the rhs terms need the diagonal kernel
        std::vector<int> hp;
        if (kh > 0) {
            double all = 0.0, heavy = 0.0;
            for (int p = 0; p < P; ++p) {
                const double k = pt_off[p + 1] - pt_off[p];
                all += k * (k - 1) / 2;
                if (k >= kh) {
                    heavy += k * (k - 1) / 2;
                    hp.push_back(p);
                }
            }
            const size_t kp = (3 * hp.size() + 15) / 16 * 16;
            const size_t bytes = kp * static_cast<size_t>(pb_.npad + 127) * (mixed_ ? 4 : 8);
            if (hp.empty() || (!forced && heavy < 0.2 * all) || bytes > (size_t{1} << 30)) {
                kh = 0;
                hp.clear();
            }
        }
        pb_.heavy_k = kh;
        nheavy_ = static_cast<int>(hp.size());
        heavy_kp_ = (3 * nheavy_ + 15) / 16 * 16;
        heavy_ld_ = (pb_.npad + 127) / 128 * 128;
        if (nheavy_ > 0) {
            upload(hpts_, hp, st_);
            const size_t ny = static_cast<size_t>(heavy_kp_) * heavy_ld_;
            if (mixed_) {
                Yt32_.alloc(ny);
                GBA_HIP(hipMemsetAsync(Yt32_.p, 0, sizeof(float) * ny, st_));
            } else {
                Yt64_.alloc(ny);
                GBA_HIP(hipMemsetAsync(Yt64_.p, 0, sizeof(double) * ny, st_));
            }
        }
    }
    // Schur task and pair lists, built on the device from the sorted layout
    std::vector<SchurTask> stasks;
    build_schur_tasks_device(st_, C, sshape_.cb, cam_off, cam_off_.p, cam_list_.p, spt_.p, pt_off_.p, scam_.p,
                             stasks, spairs_, !sep_diag_, pb_.heavy_k);
    num_stasks_ = static_cast<int>(stasks.size());
    for (SchurTask& t : stasks) t.ld = pb_.npad;
    upload(stasks_, stasks, st_);
    upload(chunks_, chunks, st_);
    upload(cam_chunk_off_, cam_chunk_off, st_);
    pb_.scam = scam_.p;
    pb_.spt = spt_.p;
    pb_.spx = spx_.p;
    pb_.perm = perm_.p;
    pb_.pt_off = pt_off_.p;
    pb_.units = units_.p;
    pb_.cam_list = cam_list_.p;
    pb_.cam_off = cam_off_.p;
    pb_.chunks = chunks_.p;
    pb_.cam_chunk_off = cam_chunk_off_.p;
    build_cam_order_copies(st_, O, cam_list_.p, spt_.p, spx_.p, cspt_, cspx_);
    pb_.cspt = cspt_.p;
    pb_.cspx = spx_.p ? cspx_.p : nullptr;
    for (int i = 0; i < 2; ++i) {
        cams_[i].alloc(9 * static_cast<size_t>(C));
        pts_[i].alloc(3 * static_cast<size_t>(P));
    }
    if (mixed_) W32_.alloc(kWStride * static_cast<size_t>(O));
    else W64_.alloc(kWStride * static_cast<size_t>(O));
    VB_.alloc(9 * static_cast<size_t>(P));
    Vinv_.alloc(6 * static_cast<size_t>(P));
    U_.alloc(81 * static_cast<size_t>(C));
    bc_.alloc(9 * static_cast<size_t>(C));
    const size_t nu = std::max(1, pb_.num_units);
    unit_sse_.alloc(nu);
    unit_den_.alloc(nu);
    unit_anom_.alloc(nu);
    gram_.alloc(55 * static_cast<size_t>(std::max(1, pb_.num_chunks)));
    blk_sse_.alloc(static_cast<size_t>(cost_partials(pb_)) + 1);
    blk_anom_.alloc(static_cast<size_t>(cost_partials(pb_)) + 1);
    const size_t np = static_cast<size_t>(pb_.npad);
    if (mixed_) {
        S32_.alloc(np * np);
        S32low_.alloc(np * np);
        if (!precond_trsv_) {
            X32_.alloc(np * np);
            T32_.alloc(np * np);
        }
        D32_.alloc(2 * static_cast<size_t>(nblk_) * kTile * kTile);
    } else {
        S64_.alloc(np * np);
        D64_.alloc(2 * static_cast<size_t>(nblk_) * kTile * kTile);
    }
    rhs_.alloc(np);
    if (mixed_) {
        ref_res_.alloc(2 * np);  // the residual, then the magnitudes of its terms
        ref_dx_.alloc(np);
        ref_z_.alloc(3 * static_cast<size_t>(P));
        ref_rot_.alloc(4 * static_cast<size_t>(C));
        neg_one_.alloc(np);
    }
    y_.alloc(np);
    x_.alloc(np);
    dc_.alloc(np);
    dscale_.alloc(np);
    pcgw_.alloc(pcg_work(pb_.npad));
    sync_.alloc(2 * static_cast<size_t>(nblk_) + 2);
    {
        // the one-launch factor is opt-in: at C=356 it measured 2.0 ms (FP32
        // with the inverse) / 1.8 ms (FP64) against 1.13 / 1.47 ms for the
        // launch per tile column
        const char* env = std::getenv("GBA_FACTOR");
        persist_ = env && std::string(env) == "persistent";
        if (persist_) {
            std::vector<long long> cs;
            persist_plan(nblk_, mixed_, !mixed_, cs, pgrid_);
            pntasks_ = cs.back();
            upload(pcol_, cs, st_);
            psync_.alloc(persist_sync_ints(nblk_));
        }
    }
    flags_.alloc(1);
    status_.alloc(1);
    if (mixed_) setup_cg();  // after every buffer its kernel arguments point at
    GBA_HIP(hipHostMalloc(&h_status_, sizeof(PassStatus), hipHostMallocDefault));
    std::memset(h_status_, 0, sizeof(PassStatus));
    if (mixed_) GBA_HIP(hipHostMalloc(&h_res_, sizeof(double) * 4 * np, hipHostMallocDefault));
    // points without observations keep V = 0, bp = 0; padding of S is identity;
    // rhs padding stays zero
    launch_fill_double(st_, VB_.p, VB_.n, 0.0);
    launch_fill_double(st_, Vinv_.p, Vinv_.n, 0.0);
    launch_fill_double(st_, rhs_.p, rhs_.n, 0.0);
    if (mixed_) {
        launch_fill_double(st_, ref_res_.p, ref_res_.n, 0.0);
        launch_fill_double(st_, neg_one_.p, neg_one_.n, -1.0);
    }
    launch_fill_double(st_, dc_.p, dc_.n, 0.0);
    launch_fill_double(st_, x_.p, x_.n, 0.0);
    if (mixed_) {
        launch_identity_pad(st_, nullptr, S32_.p, pb_.npad, pb_.n, pb_.npad);
        launch_identity_pad(st_, nullptr, S32low_.p, pb_.npad, pb_.n, pb_.npad);
    } else {
        // envelope: the tiles outside the profile are never written, so they
        // (and the transposed tiles the update of a diagonal tile reads when
        // its row starts after the column) stay zero
        if (env_) launch_fill_double(st_, S64_.p, S64_.n, 0.0);
        launch_identity_pad(st_, S64_.p, nullptr, pb_.npad, pb_.n, pb_.npad);
    }
    launch_clear_flags(st_, flags_.p);
    reset_params();
    GBA_HIP(hipGetLastError());
    GBA_HIP(hipStreamSynchronize(st_));
}

Solver::~Solver() {
    if (st_) (void)hipStreamSynchronize(st_);
    if (h_status_) (void)hipHostFree(h_status_);
    if (h_res_) (void)hipHostFree(h_res_);
    if (st_) (void)hipStreamDestroy(st_);
}

void Solver::cams_to_internal(const double* in, double* out) const {
    const size_t C = static_cast<size_t>(pb_.C);
    for (size_t c = 0; c < C; ++c) {
        const size_t to = env_ ? static_cast<size_t>(pos_of_cam_[c]) : c;
        std::memcpy(out + 9 * to, in + 9 * c, 9 * sizeof(double));
    }
}

void Solver::cams_to_caller(const double* in, double* out) const {
    const size_t C = static_cast<size_t>(pb_.C);
    for (size_t c = 0; c < C; ++c) {
        const size_t from = env_ ? static_cast<size_t>(pos_of_cam_[c]) : c;
        std::memcpy(out + 9 * c, in + 9 * from, 9 * sizeof(double));
    }
}

void Solver::reset_params(const double* cams, const double* pts) {
    cur_ = 0;
    std::vector<double> internal;
    if (cams && env_) {
        internal.resize(init_cams_.size());
        cams_to_internal(cams, internal.data());
        cams = internal.data();
    }
    GBA_HIP(hipMemcpyAsync(cams_[0].p, cams ? cams : init_cams_.data(), sizeof(double) * cams_[0].n,
                           hipMemcpyHostToDevice, st_));
    // both point buffers: the back-substitution writes trial coordinates only
    // for observed points, so a point without observations keeps its value in
    // either buffer across accepted steps (the reference's dp = 0 for it,
    // src/lm.cu:289-301,645-646)
    for (int i = 0; i < 2; ++i)
        GBA_HIP(hipMemcpyAsync(pts_[i].p, pts ? pts : init_pts_.data(), sizeof(double) * pts_[i].n,
                               hipMemcpyHostToDevice, st_));
    if (!internal.empty()) GBA_HIP(hipStreamSynchronize(st_));
}

void Solver::sync_status() {
    GBA_HIP(hipGetLastError());
    GBA_HIP(hipMemcpyAsync(h_status_, status_.p, sizeof(PassStatus), hipMemcpyDeviceToHost, st_));
    GBA_HIP(hipStreamSynchronize(st_));
}

void Solver::linearize(double* sse, unsigned long long* anomalies) {
    launch_lin_points(st_, pb_, cams_[cur_].p, pts_[cur_].p, W64_.p, W32_.p, VB_.p, unit_sse_.p,
                      unit_anom_.p);
    launch_lin_cameras(st_, pb_, cams_[cur_].p, pts_[cur_].p, gram_.p);
    launch_lin_reduce(st_, pb_, gram_.p, unit_sse_.p, unit_anom_.p, U_.p, bc_.p, status_.p);
    sync_status();
    *sse = h_status_->sse;
    *anomalies = h_status_->anomalies;
}

void Solver::cost(double* sse, unsigned long long* anomalies) {
    launch_cost(st_, pb_, cams_[cur_].p, pts_[cur_].p, blk_sse_.p, blk_anom_.p, status_.p);
    sync_status();
    *sse = h_status_->sse;
    *anomalies = h_status_->anomalies;
}

void Solver::enqueue_schur(double lambda, bool fac_copy) {
    // mixed: the factor copy and the lower copy of the scaled S, or the lower
    // copy alone (written through the factor-copy argument)
    float* sfac = mixed_ ? (fac_copy ? S32_.p : S32low_.p) : nullptr;
    float* slow = mixed_ && fac_copy ? S32low_.p : nullptr;
    if (sep_diag_)
        launch_schur_diag(st_, pb_, nullptr, W64_.p, W32_.p, Vinv_.p, VB_.p, U_.p, bc_.p, lambda,
                          mixed_ ? dscale_.p : nullptr, S64_.p, sfac, slow, pb_.npad, rhs_.p, flags_.p);
    if (num_stasks_ > 0)
        launch_schur(st_, sshape_, pb_, stasks_.p, num_stasks_, spairs_.p, W64_.p, W32_.p, Vinv_.p,
                 VB_.p, U_.p, bc_.p, lambda, mixed_ ? dscale_.p : nullptr, S64_.p, sfac,
                 slow, pb_.npad, rhs_.p, flags_.p);
    if (nheavy_ > 0) {
        // after both kernels above have written every block of S
        launch_heavy_y(st_, pb_, hpts_.p, nheavy_, W64_.p, W32_.p, Vinv_.p, mixed_ ? dscale_.p : nullptr,
                       Yt64_.p, Yt32_.p, heavy_ld_);
        launch_heavy_syrk(st_, nblk_, heavy_kp_, heavy_ld_, pb_.npad, Yt64_.p, Yt32_.p, S64_.p, sfac, slow,
                          flags_.p);
    }
}

void Solver::setup_cg() {
    int mode = cfg_.mixed_precond;
    if (const char* env = std::getenv("GBA_MIXED_PRECOND")) {
        if (std::string(env) == "factor") mode = GBA_PRECOND_FACTOR;
        else if (std::string(env) == "jacobi") mode = GBA_PRECOND_JACOBI;
    }
    if (mode != GBA_PRECOND_AUTO && mode != GBA_PRECOND_FACTOR && mode != GBA_PRECOND_JACOBI)
        throw Error("unknown mixed preconditioner");
    if (mode == GBA_PRECOND_AUTO)
        mode = nblk_ >= kJacobiMinBlk ? GBA_PRECOND_JACOBI : GBA_PRECOND_FACTOR;
    cg_mode_ = mode == GBA_PRECOND_JACOBI;
    if (!cg_mode_) return;
    const int nblk = nblk_, npad = pb_.npad;
    // tiles per matrix-product workgroup: pairs while the tiles fill the chip
    // (one round of workgroups at C=356), runs of eight beyond (fewer partial
    // sums to reduce)
    int tpw = nblk <= kJacobiPairBlk ? 2 : 8;
    if (const char* env = std::getenv("GBA_CG_TPW")) tpw = std::max(1, std::atoi(env));
    std::vector<int2> tab;
    std::vector<int> rowstart(static_cast<size_t>(nblk) + 1);
    for (int i = 0; i < nblk; ++i) {
        rowstart[i] = static_cast<int>(tab.size());
        for (int c = 0; c < (i + tpw) / tpw; ++c) tab.push_back(make_int2(i, c));
    }
    rowstart[nblk] = static_cast<int>(tab.size());
    upload(cg_wgtab_, tab, st_);
    upload(cg_rowstart_, rowstart, st_);
    const int nwg_m = static_cast<int>(tab.size()), nwg_v = cg_vector_groups(npad);
    const size_t np = static_cast<size_t>(npad);
    cg_minv_.alloc(81 * static_cast<size_t>(pb_.C));
    cg_vec_.alloc(4 * np);
    cg_rp_.alloc(static_cast<size_t>(nwg_m) * kTile);
    cg_cp_.alloc(static_cast<size_t>(nblk) * (nblk + 1) / 2 * kTile);
    cg_dpart_.alloc(static_cast<size_t>(nwg_m));
    cg_gpart_.alloc(2 * kCgPartials * static_cast<size_t>(nwg_v));
    cg_st_.alloc(16);
    launch_fill_double(st_, cg_st_.p, cg_st_.n, 0.0);
    launch_fill_double(st_, cg_gpart_.p, cg_gpart_.n, 0.0);
    CgArgs& g = cg_;
    g.A = S32low_.p;
    g.ld = npad;
    g.npad = npad;
    g.nblk = nblk;
    g.ncam = pb_.C;
    g.tpw = tpw;
    g.wgtab = cg_wgtab_.p;
    g.rowstart = cg_rowstart_.p;
    g.nwg_m = nwg_m;
    g.nwg_v = nwg_v;
    g.b = rhs_.p;
    g.x = x_.p;
    g.r = cg_vec_.p;
    g.u = cg_vec_.p + np;
    g.p = cg_vec_.p + 2 * np;
    g.s = cg_vec_.p + 3 * np;
    g.minv = cg_minv_.p;
    g.rp = cg_rp_.p;
    g.cp = cg_cp_.p;
    g.dpart = cg_dpart_.p;
    g.gpart = cg_gpart_.p;
    g.st = cg_st_.p;
    g.status = status_.p;
    g.flags = flags_.p;
    // deflation by the similarity gauge (on unless the configuration turns it
    // off, or GBA_CG_DEFLATE=0 in the environment)
    bool defl = cfg_.cg_deflation >= 0;
    if (const char* env = std::getenv("GBA_CG_DEFLATE")) defl = std::atoi(env) != 0;
    g.zd = g.azd = g.einv = nullptr;
    g.zbpart = nullptr;
    if (defl) {
        const size_t nv = static_cast<size_t>(nwg_v), nb = static_cast<size_t>(nblk);
        const size_t na = nb * static_cast<size_t>(cg_gauge_runs(nblk)) * kTile * 8;
        cg_defl_.alloc(3 * 8 * np + 2 * 8 * nv + 49 * nb + 64 + na);
        double* q = cg_defl_.p;
        cg_dw_.zraw = q;
        cg_dw_.zd = q + 8 * np;
        cg_dw_.azd = q + 16 * np;
        cg_dw_.npart = q + 24 * np;
        g.zbpart = q + 24 * np + 8 * nv;
        cg_dw_.epart = q + 24 * np + 16 * nv;
        cg_dw_.einv = cg_dw_.epart + 49 * nb;
        cg_dw_.apart = cg_dw_.einv + 64;
        launch_fill_double(st_, cg_defl_.p, cg_defl_.n, 0.0);
        g.zd = cg_dw_.zd;
        g.azd = cg_dw_.azd;
        g.einv = cg_dw_.einv;
    }
}

// Every device pointer of the block-Jacobi kernels' arguments must be set
// before a launch (a null one would fault the GPU): checked on the host.
static void check_cg_args(const CgArgs& g) {
    const void* ptrs[] = {g.A, g.wgtab, g.rowstart, g.b, g.x, g.r, g.u, g.p, g.s, g.minv,
                          g.rp, g.cp, g.dpart, g.gpart, g.st, g.status, g.flags};
    for (const void* q : ptrs)
        if (!q) throw Error("block-Jacobi refinement: unset device buffer");
    // the deflation's buffers are all set or all null
    const bool any = g.zd || g.azd || g.einv || g.zbpart, all = g.zd && g.azd && g.einv && g.zbpart;
    if (any && !all) throw Error("block-Jacobi refinement: partly set deflation buffers");
    if (g.nwg_m <= 0 || g.nwg_v <= 0 || g.tpw <= 0 || g.nblk * kTile != g.npad || 9 * g.ncam > g.npad)
        throw Error("block-Jacobi refinement: inconsistent shape");
}

int Solver::cg_solve(const double* b, double* x, double tol, int max_it, int* last, double* rel, bool* converged) {
    CgArgs g = cg_;
    g.b = b;
    g.x = x;
    g.tol2 = tol * tol;
    g.max_it = max_it;
    check_cg_args(g);
    launch_cg_start(st_, g);
    // steps k = 0 .. max_it (the launch of step max_it only ends the loop),
    // the first batch sized by the last such solve's count, then batches of 8
    int enq = 0, batch = std::max(4, *last + std::max(2, *last / 8));
    for (;;) {
        const int nb = std::min(batch, max_it + 1 - enq);
        for (int k = enq; k < enq + nb; ++k) launch_cg_step(st_, g, k);
        enq += nb;
        sync_status();
        if (h_status_->pcg_done || enq > max_it) break;
        batch = 8;
    }
    const PassStatus& h = *h_status_;
    const TrialFlags f = h.flags;
    *rel = 0.0;
    *converged = true;
    if (f.nonspd || f.chol_fail || f.solve_err) return 0;
    const double bb = h.pcg_bb;
    if (!(bb > 0.0)) return 0;
    *rel = std::sqrt(h.pcg_rr / bb);
    *converged = !h.pcg_stop && h.pcg_rr <= g.tol2 * bb;
    if (*converged) *last = h.pcg_it;
    return h.pcg_it;
}

void Solver::enqueue_trial_linear_algebra(double lambda, std::vector<double>* hook_S,
                                          std::vector<double>* hook_rhs, TrialOut& out) {
    const int ld = pb_.npad;
    launch_clear_flags(st_, flags_.p);
    launch_damp_invert_v(st_, pb_.P, VB_.p, lambda, Vinv_.p, flags_.p);
    if (mixed_) launch_diag_scale(st_, pb_.n, pb_.npad, U_.p, lambda, dscale_.p);
    enqueue_schur(lambda, !cg_active());
    if (hook_S || hook_rhs) {
        GBA_HIP(hipGetLastError());
        GBA_HIP(hipStreamSynchronize(st_));
        const size_t np = static_cast<size_t>(pb_.npad), n = static_cast<size_t>(pb_.n);
        std::vector<double> scale(np, 1.0);
        if (mixed_)
            GBA_HIP(hipMemcpy(scale.data(), dscale_.p, sizeof(double) * np, hipMemcpyDeviceToHost));
        if (hook_S) {
            std::vector<double> S(np * np);
            if (mixed_) {
                std::vector<float> Sf(np * np);
                GBA_HIP(hipMemcpy(Sf.data(), cg_active() ? S32low_.p : S32_.p, sizeof(float) * np * np,
                                  hipMemcpyDeviceToHost));
                for (size_t i = 0; i < np * np; ++i) S[i] = Sf[i];
            } else {
                GBA_HIP(hipMemcpy(S.data(), S64_.p, sizeof(double) * np * np, hipMemcpyDeviceToHost));
            }
            // caller-order row r <-> internal row ir(r); the internal lower
            // triangle holds S, so an entry above it is read mirrored
            auto ir = [&](size_t r) { return env_ ? 9 * static_cast<size_t>(pos_of_cam_[r / 9]) + r % 9 : r; };
            hook_S->assign(n * n, 0.0);
            for (size_t r = 0; r < n; ++r)
                for (size_t c = 0; c <= r && c < n; ++c) {  // lower triangle, column-major out
                    const size_t a = ir(r), b = ir(c);
                    const double v = a >= b ? S[a * np + b] : S[b * np + a];
                    (*hook_S)[r + c * n] = v / (scale[a] * scale[b]);
                }
        }
        if (hook_rhs) {
            std::vector<double> rr(np);
            GBA_HIP(hipMemcpy(rr.data(), rhs_.p, sizeof(double) * np, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i) rr[i] /= scale[i];
            hook_rhs->assign(n, 0.0);
            cams_to_caller(rr.data(), hook_rhs->data());
        }
    }
    if (!mixed_) {
        enqueue_factor();
        launch_chol_solve(st_, S64_.p, nullptr, ld, nblk_, D64_.p, nullptr, rhs_.p, y_.p, dc_.p,
                          sync_.p, flags_.p, env_ ? first_.p : nullptr);
        return;
    }
    mixed_solve(lambda, out);
}

// Rounding level of the FP64 residual evaluation relative to the magnitudes
// of its terms: 32 units of 2^-53 (per-lane sums of a few terms, then a
// tree of about a dozen levels).
constexpr double kResidualFloor = 32.0 * 0x1p-53;
// Inner tolerances of the FP64 refinement: the first solve when corrections
// follow, and each correction relative to its right-hand side.  At C=356 the
// FP32-stored system's solution has an FP64 residual of about 3e-8 at lambda
// 1e-4, so solving it further than 1e-8 buys nothing, and one correction to
// 1e-2 takes that to about 1e-11 (measured).
constexpr double kInnerTol = 1e-8;
constexpr double kCorrTol = 1e-2;

void Solver::mixed_solve(double lambda, TrialOut& out) {
    const int ld = pb_.npad;
    const size_t np = static_cast<size_t>(pb_.npad);
    const double tol = cfg_.refine_tol > 0 ? cfg_.refine_tol : 1e-10;
    const int jmax = cfg_.jacobi_max_iters > 0 ? cfg_.jacobi_max_iters : 200;
    const int max_it = cfg_.refine_max_iters > 0 ? cfg_.refine_max_iters : 12;
    const int corrections = cfg_.refine_corrections == 0 ? 1 : std::max(0, cfg_.refine_corrections);
    // GBA_PCG_BATCH=1 in the environment enqueues one PCG step per
    // synchronisation (the default: the previous trial's step count)
    static const bool one_step = [] {
        const char* e = std::getenv("GBA_PCG_BATCH");
        return e && std::atoi(e) == 1;
    }();
    bool factor = !cg_active();
    if (factor) {
        out.refined_by_factor = true;
        enqueue_factor();
    } else {
        launch_cg_block_inverse(st_, S32low_.p, pb_.npad, pb_.C, cg_minv_.p, flags_.p);
        if (cg_.zd) launch_cg_deflation_setup(st_, cg_, cams_[cur_].p, dscale_.p, cg_dw_);
    }
    // inner solve of A d = b on the FP32-stored system to tolerance t
    // relative to |b|; false when the block-Jacobi CG missed it within its cap
    auto inner = [&](const double* b, double* d, double t, bool first) {
        bool conv = true;
        double rel = 0.0;
        if (!factor) {
            out.refine_iters += cg_solve(b, d, t, jmax, first ? &cg_last_ : &cg_last_corr_, &rel, &conv);
        } else {
            int& last = first ? last_refine_ : pcg_last_corr_;
            const int it = pcg_solve(st_, S32low_.p, precond(), ld, pb_.npad, b, d, pcgw_.p, flags_.p, status_.p,
                                     h_status_, t, max_it, &rel, one_step ? 1 : last);
            out.refine_iters += it;
            if (it > 0) last = it;
        }
        return conv;
    };
    // First solve, of the system the Schur kernels formed (S and rhs from
    // FP32 W): to 1e-8 when corrections follow (the FP32 system is itself
    // only that close to the FP64 one), to tol when none do.
    const double tin = corrections > 0 ? std::max(tol, kInnerTol) : tol;
    if (!inner(rhs_.p, x_.p, tin, true) && !factor) {
        // missed the tolerance within the cap: this trial and the rest of the
        // solve are refined with the FP32 factor (its copy of S from the lower one)
        cg_off_ = true;
        factor = true;
        out.refined_by_factor = true;
        GBA_HIP(hipMemcpyAsync(S32_.p, S32low_.p, sizeof(float) * S32_.n, hipMemcpyDeviceToDevice, st_));
        enqueue_factor();
        inner(rhs_.p, x_.p, tin, true);
    }
    const TrialFlags f = h_status_->flags;
    launch_vec_scale(st_, pb_.n, pb_.npad, x_.p, dscale_.p, dc_.p);
    const double bb = h_status_->pcg_bb;  // |D rhs|^2
    if (f.nonspd || f.chol_fail || f.solve_err || !(bb > 0.0)) return;  // no step to refine
    // FP64 iterative refinement: the residual D (rhs - S dc) of the FP64
    // system (Jacobians evaluated again, W not read) and the magnitudes of its
    // terms into host slot k % 2 (norms read after the next synchronisation);
    // the correction A d = D r to kCorrTol relative to |D r| on the FP32
    // system, x += d.  The chunk partials reuse the linearisation's Gram
    // buffer (55 doubles per camera chunk).  The first correction is enqueued
    // without a host round trip (the first solve's residual is far above any
    // tolerance it could meet: about 3e-8 at best, measured).
    auto residual = [&](int slot) {
        launch_refine_residual(st_, pb_, cams_[cur_].p, pts_[cur_].p, Vinv_.p, VB_.p, U_.p, bc_.p, lambda, dc_.p,
                               ref_z_.p, ref_rot_.p, gram_.p, dscale_.p, ref_res_.p, ref_res_.p + np, flags_.p);
        GBA_HIP(hipMemcpyAsync(h_res_ + 2 * np * slot, ref_res_.p, sizeof(double) * 2 * np, hipMemcpyDeviceToHost,
                               st_));
        out.refine_outer++;
    };
    auto norms = [&](int slot, double* rr, double* mm) {
        const double* h = h_res_ + 2 * np * slot;
        *rr = *mm = 0.0;
        for (size_t i = 0; i < np; ++i) {
            *rr += h[i] * h[i];
            *mm += h[np + i] * h[np + i];
        }
    };
    auto correct = [&](double sign) {
        if (sign > 0) inner(ref_res_.p, ref_dx_.p, kCorrTol, false);
        launch_vec_axpy_scaled(st_, pb_.npad, x_.p, ref_dx_.p, sign > 0 ? nullptr : neg_one_.p);
        launch_vec_scale(st_, pb_.n, pb_.npad, x_.p, dscale_.p, dc_.p);
    };
    residual(0);
    double rr = 0.0, mm = 0.0, rr_prev = bb;
    int slot = 0;
    if (corrections > 0) {
        correct(+1.0);
        residual(1);
    }
    GBA_HIP(hipGetLastError());
    GBA_HIP(hipStreamSynchronize(st_));
    norms(0, &rr, &mm);
    if (corrections > 0) {
        rr_prev = rr;
        double rr1 = 0.0, mm1 = 0.0;
        norms(1, &rr1, &mm1);
        if (!(rr1 < rr)) {
            // the correction did not help (the FP32 system's error along the
            // least determined directions is of the order of the damping):
            // take it back and keep the first solve's residual
            correct(-1.0);
            out.refine_floor = std::max(out.refine_floor, kResidualFloor * std::sqrt(mm / bb));
            out.refine_rel = std::sqrt(rr / bb);
            return;
        }
        rr = rr1;
        mm = mm1;
        slot = 1;
    }
    // further corrections (refine_corrections > 1), one round trip each
    for (int k = corrections > 0 ? 1 : 0;; ++k) {
        // the residual cannot be evaluated below the rounding of its terms
        // (kResidualFloor |terms|); a refinement whose residual reaches that
        // floor, or stops shrinking, has gone as far as FP64 can tell
        const double floor2 = kResidualFloor * kResidualFloor * mm;
        out.refine_rel = std::sqrt(rr / bb);
        out.refine_floor = std::max(out.refine_floor, std::sqrt(floor2 / bb));
        const bool done = !(rr > tol * tol * bb) || !(rr > floor2) || (k > 0 && !(rr < 0.25 * rr_prev));
        if (done || k >= corrections) break;
        rr_prev = rr;
        correct(+1.0);
        slot = 1 - slot;
        residual(slot);
        GBA_HIP(hipGetLastError());
        GBA_HIP(hipStreamSynchronize(st_));
        norms(slot, &rr, &mm);
    }
}

void Solver::enqueue_factor() {
    const int ld = pb_.npad;
    if (env_) {
        launch_batch_cholesky(st_, 1, etab_, edesc_.p, S64_.p, D64_.p, flags_.p, eactive_.p);
        return;
    }
    if (persist_) {
        if (mixed_)
            launch_cholesky_persistent(st_, nullptr, S32_.p, ld, nblk_, nullptr, D32_.p, flags_.p, pcol_.p, pntasks_,
                                       pgrid_, psync_.p, X32_.p, T32_.p);
        else
            launch_cholesky_persistent(st_, S64_.p, nullptr, ld, nblk_, D64_.p, nullptr, flags_.p, pcol_.p, pntasks_,
                                       pgrid_, psync_.p);
        return;
    }
    if (mixed_)
        launch_cholesky(st_, nullptr, S32_.p, ld, nblk_, nullptr, D32_.p, flags_.p, X32_.p, T32_.p);  // X32_ null: no inverse
    else
        launch_cholesky(st_, S64_.p, nullptr, ld, nblk_, D64_.p, nullptr, flags_.p);
}

TrialOut Solver::trial(double lambda, std::vector<double>* hook_S, std::vector<double>* hook_rhs) {
    TrialOut out;
    enqueue_trial_linear_algebra(lambda, hook_S, hook_rhs, out);
    const int nxt = 1 - cur_;
    launch_trial_cams(st_, pb_.C, cams_[cur_].p, dc_.p, cams_[nxt].p);
    launch_backsub(st_, pb_, W64_.p, W32_.p, Vinv_.p, VB_.p, dc_.p, pts_[cur_].p, cams_[nxt].p,
                   pts_[nxt].p, lambda, unit_sse_.p, unit_den_.p, unit_anom_.p, flags_.p);
    launch_trial_finalize(st_, pb_, unit_sse_.p, unit_den_.p, unit_anom_.p, dc_.p, U_.p, bc_.p, lambda,
                          status_.p, flags_.p);
    sync_status();
    out.flags = h_status_->flags;
    out.trial_sse = h_status_->sse;
    out.denom = h_status_->denom;
    out.anomalies = h_status_->anomalies;
    if (out.flags.solve_err) throw Error("dense factor or triangular solve gave up waiting for a tile");
    return out;
}

void Solver::accept_trial() { cur_ = 1 - cur_; }

void Solver::download_params(double* cams, double* pts) {
    std::vector<double> internal(env_ ? cams_[cur_].n : 0);
    GBA_HIP(hipMemcpyAsync(env_ ? internal.data() : cams, cams_[cur_].p, sizeof(double) * cams_[cur_].n,
                           hipMemcpyDeviceToHost, st_));
    GBA_HIP(hipMemcpyAsync(pts, pts_[cur_].p, sizeof(double) * pts_[cur_].n, hipMemcpyDeviceToHost, st_));
    GBA_HIP(hipStreamSynchronize(st_));
    if (env_) cams_to_caller(internal.data(), cams);
}

void Solver::download_linearization(double* U, double* V, double* W, double* bc, double* bp) {
    const size_t C = pb_.C, P = pb_.P, O = pb_.O;
    std::vector<double> hU(81 * C), hVB(9 * P), hbc(9 * C);
    std::vector<int> hperm(O);
    GBA_HIP(hipStreamSynchronize(st_));
    GBA_HIP(hipMemcpy(hU.data(), U_.p, sizeof(double) * hU.size(), hipMemcpyDeviceToHost));
    GBA_HIP(hipMemcpy(hVB.data(), VB_.p, sizeof(double) * hVB.size(), hipMemcpyDeviceToHost));
    GBA_HIP(hipMemcpy(hbc.data(), bc_.p, sizeof(double) * hbc.size(), hipMemcpyDeviceToHost));
    GBA_HIP(hipMemcpy(hperm.data(), perm_.p, sizeof(int) * O, hipMemcpyDeviceToHost));
    // camera records in the caller's order (internal index pos_of_cam_[c])
    auto cpos = [&](size_t c) { return env_ ? static_cast<size_t>(pos_of_cam_[c]) : c; };
    if (U)
        for (size_t c = 0; c < C; ++c)
            for (int i = 0; i < 9; ++i)
                for (int j = 0; j < 9; ++j) U[81 * c + 9 * i + j] = j <= i ? hU[81 * cpos(c) + 9 * i + j] : 0.0;
    if (bc)
        for (size_t c = 0; c < C; ++c) std::memcpy(bc + 9 * c, &hbc[9 * cpos(c)], 9 * sizeof(double));
    for (size_t p = 0; p < P; ++p) {
        const double* vb = &hVB[9 * p];
        if (V) {
            double* o = V + 9 * p;
            o[0] = vb[0]; o[1] = 0.0;   o[2] = 0.0;
            o[3] = vb[1]; o[4] = vb[2]; o[5] = 0.0;
            o[6] = vb[3]; o[7] = vb[4]; o[8] = vb[5];
        }
        if (bp) {
            bp[3 * p] = vb[6];
            bp[3 * p + 1] = vb[7];
            bp[3 * p + 2] = vb[8];
        }
    }
    if (W) {
        // device records are kWStride values apart (27 used)
        std::vector<double> hW(kWStride * O);
        if (mixed_) {
            std::vector<float> hw(kWStride * O);
            GBA_HIP(hipMemcpy(hw.data(), W32_.p, sizeof(float) * hw.size(), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < hw.size(); ++i) hW[i] = hw[i];
        } else {
            GBA_HIP(hipMemcpy(hW.data(), W64_.p, sizeof(double) * hW.size(), hipMemcpyDeviceToHost));
        }
        for (size_t k = 0; k < O; ++k)
            std::memcpy(W + 27 * static_cast<size_t>(hperm[k]), &hW[kWStride * k], 27 * sizeof(double));
    }
}

void Solver::download_step(double* dc, double* dp) {
    if (dp) {
        DevBuf<double> ddp;
        ddp.alloc(3 * static_cast<size_t>(pb_.P));
        launch_fill_double(st_, ddp.p, ddp.n, 0.0);
        launch_point_step(st_, pb_, W64_.p, W32_.p, Vinv_.p, VB_.p, dc_.p, ddp.p);
        GBA_HIP(hipGetLastError());
        GBA_HIP(hipStreamSynchronize(st_));
        GBA_HIP(hipMemcpy(dp, ddp.p, sizeof(double) * ddp.n, hipMemcpyDeviceToHost));
    }
    if (dc) {
        GBA_HIP(hipStreamSynchronize(st_));
        std::vector<double> internal(static_cast<size_t>(pb_.n));
        GBA_HIP(hipMemcpy(internal.data(), dc_.p, sizeof(double) * pb_.n, hipMemcpyDeviceToHost));
        cams_to_caller(internal.data(), dc);
    }
}

double Solver::time_kernel(int kernel, double lambda, int reps) {
    CgArgs cgt = cg_;
    int cg_k = 0;
    hipEvent_t a, b;
    GBA_HIP(hipEventCreate(&a));
    GBA_HIP(hipEventCreate(&b));
    const int ld = pb_.npad;
    auto enqueue = [&]() {
        switch (kernel) {
            case GBA_KERNEL_EDGE_POINT:
                launch_lin_points(st_, pb_, cams_[cur_].p, pts_[cur_].p, W64_.p, W32_.p, VB_.p,
                                  unit_sse_.p, unit_anom_.p);
                break;
            case GBA_KERNEL_EDGE_CAMERA:
                launch_lin_cameras(st_, pb_, cams_[cur_].p, pts_[cur_].p, gram_.p);
                break;
            case GBA_KERNEL_SCHUR:
                enqueue_schur(lambda);
                break;
            case GBA_KERNEL_HEAVY_SYRK: {
                float* sfac = mixed_ ? (cg_active() ? S32low_.p : S32_.p) : nullptr;
                float* slow = mixed_ && !cg_active() ? S32low_.p : nullptr;
                launch_heavy_syrk(st_, nblk_, heavy_kp_, heavy_ld_, pb_.npad, Yt64_.p, Yt32_.p, S64_.p, sfac, slow,
                                  flags_.p);
                break;
            }
            case GBA_KERNEL_FACTOR:
                // mixed: factor + explicit inverse (what a trial runs before PCG)
                enqueue_factor();
                break;
            case GBA_KERNEL_SOLVE:
                // FP64: both triangular solves; mixed: one preconditioner application
                if (mixed_)
                {
                    if (cg_active()) throw Error("no factor preconditioner on a block-Jacobi handle");
                    const PcgPrecond M = precond();
                    if (M.X32)
                        launch_precond_apply(st_, M.X32, ld, pb_.npad, rhs_.p, x_.p,
                                             pcgw_.p + 5 * static_cast<size_t>(pb_.npad));
                    else
                        launch_chol_solve(st_, nullptr, M.L32, ld, nblk_, nullptr, M.D32, rhs_.p, y_.p, x_.p, sync_.p,
                                          flags_.p);
                }
                else
                    launch_chol_solve(st_, S64_.p, nullptr, ld, nblk_, D64_.p, nullptr, rhs_.p, y_.p, x_.p,
                                      sync_.p, flags_.p, env_ ? first_.p : nullptr);
                break;
            case GBA_KERNEL_BACKSUB:
                launch_backsub(st_, pb_, W64_.p, W32_.p, Vinv_.p, VB_.p, dc_.p, pts_[cur_].p,
                               cams_[1 - cur_].p, pts_[1 - cur_].p, lambda, unit_sse_.p, unit_den_.p,
                               unit_anom_.p, flags_.p);
                break;
            case GBA_KERNEL_COST:
                launch_cost(st_, pb_, cams_[cur_].p, pts_[cur_].p, blk_sse_.p, blk_anom_.p, status_.p);
                break;
            case GBA_KERNEL_CG_SETUP:
                launch_cg_block_inverse(st_, S32low_.p, pb_.npad, pb_.C, cg_minv_.p, flags_.p);
                if (cg_.zd) launch_cg_deflation_setup(st_, cg_, cams_[cur_].p, dscale_.p, cg_dw_);
                break;
            case GBA_KERNEL_CG_STEP:
                launch_cg_step(st_, cgt, cg_k++);
                break;
            case GBA_KERNEL_REFINE_RESIDUAL:
                if (!mixed_) throw Error("the FP64 refinement residual is a mixed-precision kernel");
                launch_refine_residual(st_, pb_, cams_[cur_].p, pts_[cur_].p, Vinv_.p, VB_.p, U_.p, bc_.p, lambda,
                                       dc_.p, ref_z_.p, ref_rot_.p, gram_.p, dscale_.p, ref_res_.p,
                                       ref_res_.p + pb_.npad, flags_.p);
                break;
            default:
                throw Error("unknown kernel id");
        }
    };
    // prepare a valid state: a linearisation and one trial at lambda
    double sse;
    unsigned long long an;
    linearize(&sse, &an);
    (void)trial(lambda);
    if ((kernel == GBA_KERNEL_CG_STEP || kernel == GBA_KERNEL_CG_SETUP) && !cg_active())
        throw Error("the block-Jacobi refinement is not in use on this handle");
    if (kernel == GBA_KERNEL_HEAVY_SYRK && nheavy_ == 0)
        throw Error("this handle's Schur assembly has no heavy points");
    if (kernel == GBA_KERNEL_CG_STEP) {
        check_cg_args(cgt);
        // steps that never stop: no tolerance, no cap
        cgt.tol2 = 0.0;
        cgt.max_it = INT_MAX;
        launch_cg_block_inverse(st_, S32low_.p, pb_.npad, pb_.C, cg_minv_.p, flags_.p);
        if (cgt.zd) launch_cg_deflation_setup(st_, cgt, cams_[cur_].p, dscale_.p, cg_dw_);
        launch_cg_start(st_, cgt);
    }
    if (kernel == GBA_KERNEL_FACTOR) {
        // refactor the same matrix each repetition: re-assemble first
        double total = 0.0;
        for (int r = 0; r < reps; ++r) {
            launch_clear_flags(st_, flags_.p);
            launch_damp_invert_v(st_, pb_.P, VB_.p, lambda, Vinv_.p, flags_.p);
            if (mixed_) launch_diag_scale(st_, pb_.n, pb_.npad, U_.p, lambda, dscale_.p);
            enqueue_schur(lambda, true);
            GBA_HIP(hipEventRecord(a, st_));
            enqueue();
            GBA_HIP(hipEventRecord(b, st_));
            GBA_HIP(hipEventSynchronize(b));
            float ms = 0.f;
            GBA_HIP(hipEventElapsedTime(&ms, a, b));
            total += ms;
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return total / reps;
    }
    enqueue();  // warm-up
    GBA_HIP(hipEventRecord(a, st_));
    for (int r = 0; r < reps; ++r) enqueue();
    GBA_HIP(hipEventRecord(b, st_));
    GBA_HIP(hipEventSynchronize(b));
    GBA_HIP(hipGetLastError());
    float ms = 0.f;
    GBA_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return static_cast<double>(ms) / reps;
}

void Solver::solve(const gba_lm_options& opts, gba_lm_result* res, double t_entry) {
    const double t_solve0 = now_s();
    reset_params();
    double sse = 0.0;
    unsigned long long an = 0, counter0 = 0;
    linearize(&sse, &an);
    counter0 += an;
    res->initial_cost = sse / 2.0;
    res->iterations = 0;
    res->trials = 0;
    res->refine_iterations = 0;
    res->refine_outer_steps = 0;
    res->refine_floor_trials = 0;
    res->max_refine_floor = 0.0;
    res->max_refine_residual = 0.0;
    res->refine_factor_trials = 0;
    cg_off_ = false;
    double lambda = opts.initial_lambda;
    bool converged = false, failed = false;
    double rel_decrease = 0.0;
    int ntrace = 0;
    while (res->iterations < opts.max_iterations && !converged && !failed) {
        double nu = 2.0;
        bool accepted = false;
        for (int trial_i = 0; trial_i < opts.max_inner_trials; ++trial_i) {
            const unsigned long long before0 = counter0;
            const double lam_used = lambda;
            TrialOut t = trial(lambda);
            res->trials++;
            res->refine_iterations += t.refine_iters;
            res->refine_outer_steps += t.refine_outer;
            if (t.refine_rel > (cfg_.refine_tol > 0 ? cfg_.refine_tol : 1e-10)) res->refine_floor_trials++;
            res->max_refine_floor = std::max(res->max_refine_floor, t.refine_floor);
            if (t.refined_by_factor) res->refine_factor_trials++;
            if (!t.flags.nonspd && !t.flags.chol_fail && !t.flags.solve_err)
                res->max_refine_residual = std::max(res->max_refine_residual, t.refine_rel);
            int code;
            double rho = NAN, tcost = NAN;
            if (t.flags.nonspd) {
                code = GBA_TRIAL_NONSPD_V;
            } else if (t.flags.chol_fail) {
                code = GBA_TRIAL_CHOL_FAIL;
            } else if (!(t.denom > 0.0)) {
                code = GBA_TRIAL_DENOM;
            } else {
                counter0 += t.anomalies;
                tcost = t.trial_sse / 2.0;
                if (counter0 != before0 || !std::isfinite(t.trial_sse)) {
                    code = GBA_TRIAL_NONFINITE;
                } else {
                    rho = (sse - t.trial_sse) / (0.5 * t.denom);
                    code = (rho > 0.0 && t.trial_sse < sse) ? GBA_TRIAL_ACCEPT : GBA_TRIAL_REJECT;
                }
            }
            if (opts.verbose)
                std::fprintf(stderr, "  iter %d trial: lambda %.3g cost %.9g rho %.3g %s\n",
                             res->iterations, lam_used, tcost, rho,
                             code == GBA_TRIAL_ACCEPT ? "accept" : "reject");
            if (code == GBA_TRIAL_ACCEPT) {
                accept_trial();
                rel_decrease = (sse - t.trial_sse) / std::max(sse, 1.0);
                sse = t.trial_sse;
                lambda *= std::max(1.0 / 3.0, 1.0 - std::pow(2.0 * rho - 1.0, 3.0));
                accepted = true;
            }
            if (cfg_.trace && ntrace < cfg_.trace_capacity) {
                double* tr = cfg_.trace + GBA_TRACE_STRIDE * static_cast<size_t>(ntrace);
                tr[0] = res->iterations;
                tr[1] = lam_used;
                tr[2] = tcost;
                tr[3] = rho;
                tr[4] = code;
                tr[5] = now_s() - t_entry;
                ++ntrace;
            }
            if (accepted) break;
            lambda *= nu;
            nu *= 2.0;
        }
        if (!accepted) {
            failed = true;
            break;
        }
        ++res->iterations;
        if (rel_decrease < opts.rel_decrease_tol) {
            converged = true;
            break;
        }
        double lin_sse;  // the current cost stays the accepted trial cost
        linearize(&lin_sse, &an);
        counter0 += an;
    }
    res->final_cost = sse / 2.0;
    res->anomalies = counter0;
    res->termination = converged ? GBA_CONVERGENCE : (failed ? GBA_FAILURE : GBA_MAX_ITERATIONS);
    res->solve_time_s = now_s() - t_solve0;
}

}  // namespace gba_amd
