// obs_fp16.cpp -- FP16 storage of the observed pixels with a per-camera FP32
// scale (host side: scale choice and encoding; the kernels decode through
// obs_px in device_common.hpp).
//
// The binary16 conversion is written from IEEE 754-2008 (binary16: 1 sign,
// 5 exponent bits with bias 15, 10 fraction bits; subnormals below 2^-14;
// round to nearest, ties to even), so that the encoding does not depend on
// the host compiler's or runtime's half-precision support.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "solver.hpp"

namespace gba_amd {

uint16_t double_to_half_rne(double x) {
    uint64_t bits;
    std::memcpy(&bits, &x, sizeof bits);
    const uint16_t sign = static_cast<uint16_t>((bits >> 48) & 0x8000u);
    const int exp = static_cast<int>((bits >> 52) & 0x7ff);
    const uint64_t frac = bits & ((uint64_t{1} << 52) - 1);
    if (exp == 0x7ff) {  // infinity, or NaN (kept quiet)
        if (frac == 0) return sign | 0x7c00u;
        return sign | 0x7e00u | static_cast<uint16_t>((frac >> 42) & 0x1ffu);
    }
    if (exp == 0) return sign;  // zero or FP64 subnormal (far below 2^-25)
    const int e = exp - 1023;
    if (e >= 16) return sign | 0x7c00u;  // |x| >= 2^16 overflows
    uint64_t q, rem, half;
    if (e >= -14) {
        // normal binary16: biased exponent e + 15, the top 10 fraction bits;
        // a carry out of the fraction moves into the exponent (up to infinity)
        q = (static_cast<uint64_t>(e + 15) << 10) | (frac >> 42);
        rem = frac & ((uint64_t{1} << 42) - 1);
        half = uint64_t{1} << 41;
    } else {
        // subnormal binary16: units of 2^-24; value = M 2^(e-52), M = 1.frac
        const int shift = 28 - e;  // >= 43
        if (shift > 54) return sign;  // below 2^-25: rounds to zero
        const uint64_t M = (uint64_t{1} << 52) | frac;
        q = M >> shift;
        rem = M & ((uint64_t{1} << shift) - 1);
        half = uint64_t{1} << (shift - 1);
    }
    if (rem > half || (rem == half && (q & 1))) ++q;
    return sign | static_cast<uint16_t>(q);
}

void encode_obs_fp16(const gba_problem_view& prob, float* cam_scale, uint16_t* px16) {
    const int C = prob.num_cameras, O = prob.num_observations;
    std::vector<double> m(static_cast<size_t>(C), 0.0);
    for (int e = 0; e < O; ++e) {
        const int c = prob.obs_camera[e];
        for (int i = 0; i < 2; ++i) {
            const double v = std::fabs(prob.obs_pixel[2 * static_cast<size_t>(e) + i]);
            if (std::isfinite(v) && v > m[c]) m[c] = v;
        }
    }
    for (int c = 0; c < C; ++c) {
        // the largest magnitude maps to about 2^15, inside binary16's range
        // (65504) with room for the rounding of the scale itself
        const float s = static_cast<float>(m[c] / 32768.0);
        cam_scale[c] = (s > 0.0f && std::isfinite(s)) ? s : 1.0f;
    }
    for (int e = 0; e < O; ++e) {
        const double s = static_cast<double>(cam_scale[prob.obs_camera[e]]);
        for (int i = 0; i < 2; ++i)
            px16[2 * static_cast<size_t>(e) + i] =
                double_to_half_rne(prob.obs_pixel[2 * static_cast<size_t>(e) + i] / s);
    }
}

}  // namespace gba_amd
