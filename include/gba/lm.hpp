// lm.hpp -- LM solver API; same names and meaning as the reference's
// include/gba/lm.hpp:13-42, implemented by libgba_amd.so on one MI355X.
#pragma once

#include <string>

#include "gba/bal_problem.hpp"

namespace gba {

struct LMOptions {
    int max_iterations = 50;
    double initial_lambda = 1e-3;   // Marquardt diagonal damping
    double rel_decrease_tol = 1e-6; // convergence: relative cost decrease
    int max_inner_trials = 16;      // lambda escalations per outer iteration
    bool verbose = false;           // per-trial log to stderr
};

struct LMResult {
    double initial_cost = 0.0;
    double final_cost = 0.0;
    int iterations = 0;             // accepted outer iterations
    std::string termination;        // "CONVERGENCE" | "MAX_ITERATIONS" | "FAILURE"
    unsigned long long anomalies = 0;  // cumulative non-finite edge count
    double total_time_s = 0.0;      // wall time of the solve, allocation and upload included
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
