// capi.cpp -- the extern "C" boundary of libgba_amd.so (declared in
// include/gba_c.h).  Every entry point catches library exceptions and turns
// them into a non-zero return code plus gba_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <vector>

#include "gba_c.h"
#include "gba_internal.hpp"
#include "launchers.hpp"
#include "snavely_model.hpp"
#include "solver.hpp"
#include "batch.hpp"

namespace gba_amd {
const char* last_error();
int bal_header(const char* path, int* C, int* P, int* O);
int load_bal_into(const char* path, gba_problem_view* out);
int write_bal(const char* path, const gba_problem_view* p);
}  // namespace gba_amd

using namespace gba_amd;

struct gba_solver {
    std::unique_ptr<Solver> s;
};

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        return set_error(e.what());
    } catch (...) {
        return set_error("unknown error");
    }
}

// Argument checks of the entry points: a null pointer where one is required
// is an error return, not a crash.
void need(bool ok, const char* what) {
    if (!ok) throw Error(what);
}

gba_solver_config config_or_default(const gba_solver_config* cfg) {
    gba_solver_config c;
    gba_default_config(&c);
    if (cfg) c = *cfg;
    return c;
}

}  // namespace

extern "C" {

void gba_default_options(gba_lm_options* o) {
    o->max_iterations = 50;
    o->initial_lambda = 1e-3;
    o->rel_decrease_tol = 1e-6;
    o->max_inner_trials = 16;
    o->verbose = 0;
}

void gba_default_config(gba_solver_config* c) {
    std::memset(c, 0, sizeof *c);
    c->precision = GBA_PRECISION_FP64;
    c->refine_tol = 1e-10;
    c->refine_max_iters = 12;
}

int gba_solve_bal(gba_problem_view* prob, const gba_lm_options* opts, const gba_solver_config* cfg,
                  gba_lm_result* res) {
    return guarded([&] {
        need(prob && res, "gba_solve_bal: null problem or result");
        const double t0 = now_s();
        gba_lm_options o;
        gba_default_options(&o);
        if (opts) o = *opts;
        std::memset(res, 0, sizeof *res);
        Solver s(*prob, config_or_default(cfg));
        res->setup_time_s = now_s() - t0;
        s.solve(o, res, t0);
        s.download_params(prob->cameras, prob->points);
        res->total_time_s = now_s() - t0;
        return 0;
    });
}

int gba_solver_create(const gba_problem_view* prob, const gba_solver_config* cfg, gba_solver** out) {
    return guarded([&] {
        need(prob && out, "gba_solver_create: null problem or handle");
        *out = nullptr;
        auto h = std::make_unique<gba_solver>();
        h->s = std::make_unique<Solver>(*prob, config_or_default(cfg));
        *out = h.release();
        return 0;
    });
}

int gba_solver_solve(gba_solver* s, const gba_lm_options* opts, gba_lm_result* res) {
    return guarded([&] {
        need(s && res, "gba_solver_solve: null handle or result");
        const double t0 = now_s();
        gba_lm_options o;
        gba_default_options(&o);
        if (opts) o = *opts;
        std::memset(res, 0, sizeof *res);
        s->s->solve(o, res, t0);
        res->total_time_s = now_s() - t0;
        return 0;
    });
}

int gba_solver_params(gba_solver* s, double* cameras, double* points) {
    return guarded([&] {
        need(s && cameras && points, "gba_solver_params: null argument");
        s->s->download_params(cameras, points);
        return 0;
    });
}

int gba_solver_time_kernel(gba_solver* s, int kernel, double lambda, int reps, double* avg_ms) {
    return guarded([&] {
        need(s && avg_ms, "gba_solver_time_kernel: null argument");
        *avg_ms = s->s->time_kernel(kernel, lambda, reps < 1 ? 1 : reps);
        return 0;
    });
}

int gba_solver_stats(gba_solver* s, double* pair_count, int* n, int* n_padded) {
    return guarded([&] {
        need(s, "gba_solver_stats: null handle");
        if (pair_count) *pair_count = s->s->pair_count();
        if (n) *n = s->s->n();
        if (n_padded) *n_padded = s->s->npad();
        return 0;
    });
}

int gba_solver_heavy(gba_solver* s, int* heavy_k, int* heavy_points, int* width) {
    return guarded([&] {
        need(s, "gba_solver_heavy: null handle");
        int k = 0, np = 0, w = 0;
        s->s->heavy(&k, &np, &w);
        if (heavy_k) *heavy_k = k;
        if (heavy_points) *heavy_points = np;
        if (width) *width = w;
        return 0;
    });
}

int gba_solver_layout(gba_solver* s, int* layout, double* factor_work_ratio, double* graph_density) {
    return guarded([&] {
        need(s, "gba_solver_layout: null handle");
        int l = 0;
        double w = 1.0, d = 1.0;
        s->s->layout(&l, &w, &d);
        if (layout) *layout = l;
        if (factor_work_ratio) *factor_work_ratio = w;
        if (graph_density) *graph_density = d;
        return 0;
    });
}

int gba_solver_mixed_precond(gba_solver* s, int* mode) {
    return guarded([&] {
        need(s && mode, "gba_solver_mixed_precond: null argument");
        *mode = s->s->mixed_precond();
        return 0;
    });
}

void gba_solver_destroy(gba_solver* s) { delete s; }

struct gba_batch {
    std::unique_ptr<BatchSolver> s;
};

int gba_batch_create_precision(const gba_problem_view* probs, int count, int precision, gba_batch** out) {
    return guarded([&] {
        need(out, "gba_batch_create: null handle");
        *out = nullptr;
        if (!probs || count <= 0) return set_error("gba_batch_create: no problems");
        if (precision != GBA_PRECISION_FP64 && precision != GBA_PRECISION_MIXED)
            return set_error("gba_batch_create: unknown precision");
        auto h = std::make_unique<gba_batch>();
        h->s = std::make_unique<BatchSolver>(probs, count, /*autotune=*/true, precision == GBA_PRECISION_MIXED);
        *out = h.release();
        return 0;
    });
}

int gba_batch_create(const gba_problem_view* probs, int count, gba_batch** out) {
    return gba_batch_create_precision(probs, count, GBA_PRECISION_FP64, out);
}

int gba_batch_solve(gba_batch* b, const gba_lm_options* opts, gba_lm_result* results, int* rounds) {
    return guarded([&] {
        need(b && results, "gba_batch_solve: null handle or results");
        gba_lm_options o;
        gba_default_options(&o);
        if (opts) o = *opts;
        b->s->solve(o, results);
        if (rounds) *rounds = b->s->rounds();
        return 0;
    });
}

int gba_batch_params(gba_batch* b, int index, double* cameras, double* points) {
    return guarded([&] {
        need(b && cameras && points, "gba_batch_params: null argument");
        b->s->download_params(index, cameras, points);
        return 0;
    });
}

int gba_batch_info(gba_batch* b, int* tile_factor, double* tune_ms, int* graphs) {
    return guarded([&] {
        need(b, "gba_batch_info: null handle");
        if (tile_factor) *tile_factor = b->s->tile_factor() ? 1 : 0;
        if (tune_ms) {
            tune_ms[0] = b->s->tune_ms(0);
            tune_ms[1] = b->s->tune_ms(1);
        }
        if (graphs) *graphs = b->s->graphs() ? 1 : 0;
        return 0;
    });
}

void gba_batch_destroy(gba_batch* b) { delete b; }

int gba_solve_bal_batched(gba_problem_view* probs, int count, const gba_lm_options* opts,
                          gba_lm_result* results) {
    return guarded([&] {
        const double t0 = now_s();
        if (!probs || count <= 0) return set_error("gba_solve_bal_batched: no problems");
        need(results, "gba_solve_bal_batched: null results");
        BatchSolver bs(probs, count);
        gba_lm_options o;
        gba_default_options(&o);
        if (opts) o = *opts;
        bs.solve(o, results);
        for (int i = 0; i < count; ++i) bs.download_params(i, probs[i].cameras, probs[i].points);
        const double t1 = now_s();
        for (int i = 0; i < count; ++i) {
            results[i].setup_time_s = (t1 - t0) - results[i].solve_time_s;
            results[i].total_time_s = t1 - t0;
        }
        return 0;
    });
}

int gba_linearize(const gba_problem_view* prob, int precision, double* U, double* V, double* W,
                  double* bc, double* bp, double* sse, unsigned long long* anomalies) {
    return guarded([&] {
        need(prob && sse && anomalies, "gba_linearize: null argument");
        gba_solver_config c;
        gba_default_config(&c);
        c.precision = precision;
        Solver s(*prob, c);
        s.linearize(sse, anomalies);
        s.download_linearization(U, V, W, bc, bp);
        return 0;
    });
}

int gba_schur_step(const gba_problem_view* prob, int precision, double lambda, double* dc, double* dp,
                   double* denom, int* status, double* S_out, double* rhs_out) {
    return guarded([&] {
        need(prob && dc && dp && denom && status, "gba_schur_step: null argument");
        gba_solver_config c;
        gba_default_config(&c);
        c.precision = precision;
        Solver s(*prob, c);
        double sse;
        unsigned long long an;
        s.linearize(&sse, &an);
        std::vector<double> hS, hr;
        TrialOut t = s.trial(lambda, S_out ? &hS : nullptr, rhs_out ? &hr : nullptr);
        if (S_out) std::memcpy(S_out, hS.data(), sizeof(double) * hS.size());
        if (rhs_out) std::memcpy(rhs_out, hr.data(), sizeof(double) * hr.size());
        *denom = t.denom;
        if (t.flags.nonspd) *status = GBA_TRIAL_NONSPD_V;
        else if (t.flags.chol_fail) *status = GBA_TRIAL_CHOL_FAIL;
        else if (!(t.denom > 0.0)) *status = GBA_TRIAL_DENOM;
        else *status = GBA_TRIAL_ACCEPT;
        if (*status == GBA_TRIAL_ACCEPT || *status == GBA_TRIAL_DENOM) s.download_step(dc, dp);
        return 0;
    });
}

int gba_encode_obs_fp16(const gba_problem_view* prob, float* cam_scale, unsigned short* px16) {
    return guarded([&] {
        if (!prob || !cam_scale || !px16) return set_error("gba_encode_obs_fp16: null argument");
        if (prob->num_cameras < 0 || prob->num_observations < 0)
            return set_error("gba_encode_obs_fp16: negative size");
        for (int e = 0; e < prob->num_observations; ++e)
            if (prob->obs_camera[e] < 0 || prob->obs_camera[e] >= prob->num_cameras)
                return set_error("gba_encode_obs_fp16: camera index out of range");
        encode_obs_fp16(*prob, cam_scale, reinterpret_cast<uint16_t*>(px16));
        return 0;
    });
}

int gba_cost(const gba_problem_view* prob, double* sse, unsigned long long* anomalies) {
    return gba_cost_cfg(prob, nullptr, sse, anomalies);
}

int gba_linearize_cfg(const gba_problem_view* prob, const gba_solver_config* cfg, double* U, double* V,
                      double* W, double* bc, double* bp, double* sse, unsigned long long* anomalies) {
    return guarded([&] {
        need(prob && sse && anomalies, "gba_linearize_cfg: null argument");
        Solver s(*prob, config_or_default(cfg));
        s.linearize(sse, anomalies);
        s.download_linearization(U, V, W, bc, bp);
        return 0;
    });
}

int gba_cost_cfg(const gba_problem_view* prob, const gba_solver_config* cfg, double* sse,
                 unsigned long long* anomalies) {
    return guarded([&] {
        need(prob && sse && anomalies, "gba_cost: null argument");
        Solver s(*prob, config_or_default(cfg));
        s.cost(sse, anomalies);
        return 0;
    });
}

int gba_dense_spd_solve(int n, int precision, const double* A, double* b, int* ok, int* refine_iters) {
    return guarded([&] {
        if (n <= 0) return set_error("gba_dense_spd_solve: n must be positive");
        need(A && b && ok, "gba_dense_spd_solve: null argument");
        const int npad = (n + kTile - 1) / kTile * kTile, nblk = npad / kTile;
        const size_t np = static_cast<size_t>(npad);
        const bool mixed = precision == GBA_PRECISION_MIXED;
        DevBuf<double> S64, D64, rhs, y, x, work;
        DevBuf<float> S32, S32low, D32, X32, T32;
        DevBuf<int> sync;
        DevBuf<TrialFlags> flags;
        DevBuf<PassStatus> status;
        // declared after the buffers: on an error it drains and goes first
        struct StreamGuard {
            hipStream_t s = nullptr;
            ~StreamGuard() {
                if (s) {
                    (void)hipStreamSynchronize(s);
                    (void)hipStreamDestroy(s);
                }
            }
        } guard;
        GBA_HIP(hipStreamCreateWithFlags(&guard.s, hipStreamNonBlocking));
        hipStream_t st = guard.s;
        // host copy: row-major with identity padding; A is column-major
        std::vector<double> h(np * np, 0.0);
        for (size_t rr = 0; rr < np; ++rr)
            for (size_t cc = 0; cc < np; ++cc)
                h[rr * np + cc] = (rr < static_cast<size_t>(n) && cc < static_cast<size_t>(n))
                                      ? A[rr + cc * static_cast<size_t>(n)]
                                      : (rr == cc ? 1.0 : 0.0);
        std::vector<double> hb(np, 0.0);
        std::memcpy(hb.data(), b, sizeof(double) * n);
        rhs.alloc(np);
        y.alloc(np);
        x.alloc(np);
        sync.alloc(2 * static_cast<size_t>(nblk) + 2);
        flags.alloc(1);
        status.alloc(1);
        GBA_HIP(hipMemcpyAsync(rhs.p, hb.data(), sizeof(double) * np, hipMemcpyHostToDevice, st));
        launch_clear_flags(st, flags.p);
        int iters = 0;
        if (!mixed) {
            S64.alloc(np * np);
            D64.alloc(2 * static_cast<size_t>(nblk) * kTile * kTile);
            GBA_HIP(hipMemcpyAsync(S64.p, h.data(), sizeof(double) * np * np, hipMemcpyHostToDevice, st));
            launch_cholesky(st, S64.p, nullptr, npad, nblk, D64.p, nullptr, flags.p);
            launch_chol_solve(st, S64.p, nullptr, npad, nblk, D64.p, nullptr, rhs.p, y.p, x.p, sync.p,
                              flags.p);
        } else {
            std::vector<float> hf(np * np);
            for (size_t i = 0; i < np * np; ++i) hf[i] = static_cast<float>(h[i]);
            S32.alloc(np * np);
            S32low.alloc(np * np);
            D32.alloc(2 * static_cast<size_t>(nblk) * kTile * kTile);
            work.alloc(pcg_work(npad));
            PassStatus* hs = nullptr;
            GBA_HIP(hipHostMalloc(&hs, sizeof(PassStatus), hipHostMallocDefault));
            GBA_HIP(hipMemcpyAsync(S32.p, hf.data(), sizeof(float) * np * np, hipMemcpyHostToDevice, st));
            GBA_HIP(hipMemcpyAsync(S32low.p, hf.data(), sizeof(float) * np * np, hipMemcpyHostToDevice, st));
            X32.alloc(np * np);
            T32.alloc(np * np);
            launch_cholesky(st, nullptr, S32.p, npad, nblk, nullptr, D32.p, flags.p, X32.p, T32.p);
            double rel = 0.0;
            try {
                PcgPrecond M;
                M.X32 = X32.p;
                iters = pcg_solve(st, S32low.p, M, npad, npad, rhs.p, x.p, work.p, flags.p, status.p,
                                  hs, 1e-12, 50, &rel);
            } catch (...) {
                (void)hipHostFree(hs);
                throw;
            }
            (void)hipHostFree(hs);
        }
        TrialFlags hf{};
        GBA_HIP(hipGetLastError());
        GBA_HIP(hipMemcpyAsync(&hf, flags.p, sizeof hf, hipMemcpyDeviceToHost, st));
        std::vector<double> hx(np);
        GBA_HIP(hipMemcpyAsync(hx.data(), x.p, sizeof(double) * np, hipMemcpyDeviceToHost, st));
        GBA_HIP(hipStreamSynchronize(st));
        if (hf.solve_err) return set_error("triangular solve timed out");
        *ok = hf.chol_fail ? 0 : 1;
        if (*ok) std::memcpy(b, hx.data(), sizeof(double) * n);
        if (refine_iters) *refine_iters = iters;
        return 0;
    });
}

int gba_dense_spd_solve_noalloc(int n, const double* A, double* b, int* ok, void* stream) {
    return guarded([&] {
        need(A && b && ok, "gba_dense_spd_solve_noalloc: null argument");
        dense_spd_solve_noalloc(n, A, b, ok, static_cast<hipStream_t>(stream));
        return 0;
    });
}

int gba_batch_chol9_factor_solve(const double* A, const double* b, double* x, int* ok, int n,
                                 void* stream) {
    return guarded([&] {
        hipStream_t st = static_cast<hipStream_t>(stream);
        launch_batch_chol9(st, A, b, x, ok, n);
        GBA_HIP(hipGetLastError());
        return 0;
    });
}

int gba_batch_inverse3_spd(const double* A, double* Ainv, int* ok, int n, void* stream) {
    return guarded([&] {
        hipStream_t st = static_cast<hipStream_t>(stream);
        launch_batch_inverse3(st, A, Ainv, ok, n);
        GBA_HIP(hipGetLastError());
        return 0;
    });
}

void gba_snavely_residual(const double* camera9, const double* point3, const double* observed2,
                          double* out2) {
    snavely_residual_host(camera9, point3, observed2, out2);
}

double gba_total_squared_error(const gba_problem_view* p) {
    double sum = 0.0, r[2];
    for (int i = 0; i < p->num_observations; ++i) {
        snavely_residual_host(p->cameras + 9 * static_cast<size_t>(p->obs_camera[i]),
                              p->points + 3 * static_cast<size_t>(p->obs_point[i]),
                              p->obs_pixel + 2 * static_cast<size_t>(i), r);
        sum += r[0] * r[0] + r[1] * r[1];
    }
    return sum;
}

void gba_edge_residual_jacobian_host(const double* camera9, const double* point3,
                                     const double* observed2, double* out_residual2, double* out_Jc,
                                     double* out_Jp) {
    snavely_eval<true>(camera9, point3, observed2[0], observed2[1], out_residual2, out_Jc, out_Jp);
}

int gba_bal_header(const char* path, int* C, int* P, int* O) {
    return guarded([&] { return bal_header(path, C, P, O); });
}

int gba_load_bal(const char* path, gba_problem_view* out) {
    return guarded([&] { return load_bal_into(path, out); });
}

int gba_write_bal(const char* path, const gba_problem_view* prob) {
    return guarded([&] { return write_bal(path, prob); });
}

int gba_plan_envelope(const gba_problem_view* prob, double max_density, int* usable, int* cam_order, int* first,
                      double* work_ratio, double* density) {
    return guarded([&] {
        need(prob && usable, "gba_plan_envelope: null argument");
        EnvelopeLayout lay;
        *usable = plan_envelope(*prob, max_density, lay) ? 1 : 0;
        if (*usable) {
            if (cam_order) std::copy(lay.cam_of_pos.begin(), lay.cam_of_pos.end(), cam_order);
            if (first) std::copy(lay.first.begin(), lay.first.end(), first);
        }
        if (work_ratio) *work_ratio = lay.work_ratio;
        if (density) *density = lay.density;
        return 0;
    });
}

int gba_synth_preset(int config, unsigned long long seed, gba_synth_spec* s) {
    std::memset(s, 0, sizeof *s);
    s->kmin = 2;
    s->law = GBA_LAW_GEOMETRIC;
    s->alpha = 1.5;
    s->ring_radius = 30.0;
    s->pixel_noise = 1.0;
    s->perturb = 0.01;
    switch (config) {
        case 1:
            s->num_cameras = 16; s->num_points = 22106; s->num_observations = 83718; s->kmax = 16;
            return 0;
        case 2:
            s->num_cameras = 49; s->num_points = 7776; s->num_observations = 31843; s->kmax = 49;
            return 0;
        case 3:
            s->num_cameras = 356; s->num_points = 226730; s->num_observations = 1255268;
            s->kmax = 300; s->law = GBA_LAW_PARETO;
            return 0;
        case 4:
            s->num_cameras = 1936; s->num_points = 525000; s->num_observations = 3400000;
            s->kmax = 300; s->law = GBA_LAW_PARETO;
            return 0;
        case 5: {
            // sizes drawn from the seed (splitmix64 steps)
            unsigned long long z = seed;
            auto next = [&z]() {
                z += 0x9e3779b97f4a7c15ULL;
                unsigned long long x = z;
                x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
                x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
                return x ^ (x >> 31);
            };
            s->num_cameras = 8 + static_cast<int>(next() % 57);
            s->num_points = 2000 + static_cast<int>(next() % 18001);
            s->num_observations = static_cast<int>(std::llround(4.5 * s->num_points));
            s->kmax = s->num_cameras;
            return 0;
        }
        case 6:
            s->num_cameras = 400; s->num_points = 60000; s->num_observations = 300000;
            s->kmax = 16; s->view_window = 12;
            return 0;
        case 7:
            s->num_cameras = 2000; s->num_points = 400000; s->num_observations = 2000000;
            s->kmax = 16; s->view_window = 12;
            return 0;
        default:
            return set_error("gba_synth_preset: config must be 1..7");
    }
}

int gba_synth_generate(const gba_synth_spec* spec, unsigned long long seed, gba_problem_view* out) {
    return guarded([&] {
        if (out->num_cameras != spec->num_cameras || out->num_points != spec->num_points ||
            out->num_observations != spec->num_observations)
            return set_error("gba_synth_generate: view sizes do not match the spec");
        return synth_generate(*spec, seed, out);
    });
}

int gba_synth_launch_bounds(int C, int P, unsigned long long seed, gba_problem_view* out) {
    return guarded([&] {
        if (out->num_cameras != C || out->num_points != P || out->num_observations != 2 * P)
            return set_error("gba_synth_launch_bounds: view must be C x P x 2P");
        return synth_launch_bounds(C, P, seed, out);
    });
}

int gba_device_synchronize(void) {
    return guarded([&] {
        GBA_HIP(hipDeviceSynchronize());
        return 0;
    });
}

const char* gba_last_error(void) { return last_error(); }

const char* gba_version(void) { return "gba_amd 0.1 (gfx950)"; }

}  // extern "C"
