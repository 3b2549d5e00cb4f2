// ref_shim.cpp -- TEST INFRASTRUCTURE. C-ABI binding (own code) over the
// reference's host-only reader/camera-model sources (src/bal_reader.cpp,
// include/gba/bal_problem.hpp), compiled in place from /root/reference by
// oracle/Makefile into oracle/_ref/.  Used only to write and check golden
// vectors for the oracle; never part of the product.
#include <cstring>
#include <exception>
#include <string>

#include "gba/bal_problem.hpp"

extern "C" {

void ref_snavely_residual(const double* cam9, const double* X3, const double* obs2, double* r2) {
    gba::snavely_residual(cam9, X3, obs2, r2);
}

double ref_total_squared_error(int C, int P, int O, const double* cams, const double* pts,
                               const int* oc, const int* op, const double* px) {
    gba::BALProblem p;
    p.num_cameras = C;
    p.num_points = P;
    p.num_observations = O;
    p.cameras.assign(cams, cams + 9 * static_cast<size_t>(C));
    p.points.assign(pts, pts + 3 * static_cast<size_t>(P));
    p.obs_camera.assign(oc, oc + O);
    p.obs_point.assign(op, op + O);
    p.obs_pixel.assign(px, px + 2 * static_cast<size_t>(O));
    return gba::total_squared_error(p);
}

// Returns 0 on success, 1 when the reference loader throws (message in err).
int ref_load_bal_header(const char* path, int* C, int* P, int* O, char* err, int errlen) {
    try {
        gba::BALProblem p = gba::load_bal(path);
        *C = p.num_cameras;
        *P = p.num_points;
        *O = p.num_observations;
        return 0;
    } catch (const std::exception& e) {
        if (err && errlen > 0) {
            std::strncpy(err, e.what(), static_cast<size_t>(errlen) - 1);
            err[errlen - 1] = 0;
        }
        return 1;
    }
}

int ref_load_bal(const char* path, double* cams, double* pts, int* oc, int* op, double* px) {
    try {
        gba::BALProblem p = gba::load_bal(path);
        std::memcpy(cams, p.cameras.data(), sizeof(double) * p.cameras.size());
        std::memcpy(pts, p.points.data(), sizeof(double) * p.points.size());
        std::memcpy(oc, p.obs_camera.data(), sizeof(int) * p.obs_camera.size());
        std::memcpy(op, p.obs_point.data(), sizeof(int) * p.obs_point.size());
        std::memcpy(px, p.obs_pixel.data(), sizeof(double) * p.obs_pixel.size());
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

}  // extern "C"
