// bal_problem.hpp -- BAL problem type and host camera model; same API as the
// reference's include/gba/bal_problem.hpp:15-50 (implemented in libgba_amd.so).
#pragma once

#include <string>
#include <vector>

namespace gba {

// Structure of arrays; camera block = [rodrigues(3), t(3), f, k1, k2].
struct BALProblem {
    int num_cameras = 0;
    int num_points = 0;
    int num_observations = 0;

    std::vector<double> cameras;    // 9 * num_cameras
    std::vector<double> points;     // 3 * num_points
    std::vector<int> obs_camera;    // num_observations
    std::vector<int> obs_point;     // num_observations
    std::vector<double> obs_pixel;  // 2 * num_observations (x, y interleaved)

    const double* camera(int i) const { return cameras.data() + 9 * i; }
    double* camera(int i) { return cameras.data() + 9 * i; }
    const double* point(int i) const { return points.data() + 3 * i; }
    double* point(int i) { return points.data() + 3 * i; }
};

// Loads a .txt or .bz2 (through bzcat) BAL file; throws std::runtime_error on
// any format violation (reference bal_problem.hpp:36).
BALProblem load_bal(const std::string& path);

// Reprojection residual, BAL model: P = R X + t, p = -P / P.z,
// r = 1 + k1 |p|^2 + k2 |p|^4, residual = f r p - observed (reference :45-46).
void snavely_residual(const double* camera9, const double* point3, const double* observed2,
                      double* out2);

// Sum of squared residuals over all observations (reference :50).
double total_squared_error(const BALProblem& p);

}  // namespace gba
