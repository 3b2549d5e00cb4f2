// lm.hpp -- LM solver API; same names and meaning as the reference's
// include/gba/lm.hpp:13-42, implemented by libgba_amd.so on one MI355X.
#pragma once

#include <string>

#include "gba/bal_problem.hpp"

namespace gba {

// Field names, types and defaults are the interface of the reference
// (include/gba/lm.hpp:13-29), so callers compile unchanged.
struct LMOptions {
    int max_iterations = 50;        // cap on accepted steps
    double initial_lambda = 1e-3;   // starting damping, scaled by each diagonal entry
    double rel_decrease_tol = 1e-6; // stop once an accepted step lowers the cost by less than this fraction
    int max_inner_trials = 16;      // damped solves tried per step before the solve gives up
    bool verbose = false;           // one line per trial on stderr
};

struct LMResult {
    double initial_cost = 0.0;      // half the sum of squared residuals at the start
    double final_cost = 0.0;        // the same at the last accepted parameters
    int iterations = 0;             // number of accepted steps
    std::string termination;        // "CONVERGENCE", "MAX_ITERATIONS" or "FAILURE"
    unsigned long long anomalies = 0;  // observations skipped as non-finite, summed over every pass
    double total_time_s = 0.0;      // seconds in solve_bal_gpu, device set-up and copies included
};

// The whole LM solve runs device-resident on the current HIP device (plain
// FP64).  Modifies prob's cameras and points to the last accepted state.
// Throws std::runtime_error on a GPU runtime failure.
LMResult solve_bal_gpu(BALProblem& prob, const LMOptions& opts = {});

// Host entry to the residual + analytic Jacobian the kernels use.  Jc is 2x9
// row-major (rodrigues[3], t[3], f, k1, k2), Jp is 2x3 row-major.
void edge_residual_jacobian_host(const double* camera9, const double* point3,
                                 const double* observed2, double* out_residual2,
                                 double* out_Jc, double* out_Jp);

}  // namespace gba
