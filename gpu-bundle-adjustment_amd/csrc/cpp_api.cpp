// cpp_api.cpp -- the reference's C++ API (gba::solve_bal_gpu, gba::load_bal,
// ...) on top of the C ABI of this library.
#include <stdexcept>
#include <string>

#include "gba/bal_problem.hpp"
#include "gba/lm.hpp"
#include "gba/schur.hpp"
#include "gba_c.h"

namespace gba {

namespace {
gba_problem_view view_of(BALProblem& p) {
    return gba_problem_view{p.num_cameras,       p.num_points,       p.num_observations,
                            p.cameras.data(),    p.points.data(),    p.obs_camera.data(),
                            p.obs_point.data(),  p.obs_pixel.data()};
}
[[noreturn]] void raise() { throw std::runtime_error(gba_last_error()); }
}  // namespace

BALProblem load_bal(const std::string& path) {
    BALProblem p;
    if (gba_bal_header(path.c_str(), &p.num_cameras, &p.num_points, &p.num_observations)) raise();
    p.cameras.resize(9 * static_cast<size_t>(p.num_cameras));
    p.points.resize(3 * static_cast<size_t>(p.num_points));
    p.obs_camera.resize(p.num_observations);
    p.obs_point.resize(p.num_observations);
    p.obs_pixel.resize(2 * static_cast<size_t>(p.num_observations));
    gba_problem_view v = view_of(p);
    if (gba_load_bal(path.c_str(), &v)) raise();
    return p;
}

void snavely_residual(const double* camera9, const double* point3, const double* observed2,
                      double* out2) {
    gba_snavely_residual(camera9, point3, observed2, out2);
}

double total_squared_error(const BALProblem& p) {
    gba_problem_view v = view_of(const_cast<BALProblem&>(p));
    return gba_total_squared_error(&v);
}

LMResult solve_bal_gpu(BALProblem& prob, const LMOptions& opts) {
    gba_lm_options o{opts.max_iterations, opts.initial_lambda, opts.rel_decrease_tol,
                     opts.max_inner_trials, opts.verbose ? 1 : 0};
    gba_problem_view v = view_of(prob);
    gba_lm_result r{};
    if (gba_solve_bal(&v, &o, nullptr, &r)) raise();
    LMResult out;
    out.initial_cost = r.initial_cost;
    out.final_cost = r.final_cost;
    out.iterations = r.iterations;
    out.termination = r.termination == GBA_CONVERGENCE      ? "CONVERGENCE"
                      : r.termination == GBA_MAX_ITERATIONS ? "MAX_ITERATIONS"
                                                            : "FAILURE";
    out.anomalies = r.anomalies;
    out.total_time_s = r.total_time_s;
    return out;
}

void edge_residual_jacobian_host(const double* camera9, const double* point3,
                                 const double* observed2, double* out_residual2, double* out_Jc,
                                 double* out_Jp) {
    gba_edge_residual_jacobian_host(camera9, point3, observed2, out_residual2, out_Jc, out_Jp);
}

void batch_chol9_factor_solve(const double* A, const double* b, double* x, int* ok, int n,
                              hipStream_t stream) {
    if (gba_batch_chol9_factor_solve(A, b, x, ok, n, stream)) raise();
}

void batch_inverse3_spd(const double* A, double* Ainv, int* ok, int n, hipStream_t stream) {
    if (gba_batch_inverse3_spd(A, Ainv, ok, n, stream)) raise();
}

}  // namespace gba
