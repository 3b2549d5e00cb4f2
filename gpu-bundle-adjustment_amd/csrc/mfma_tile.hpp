// mfma_tile.hpp -- wave64 helpers shared by the dense-matrix kernels:
// v_readlane broadcasts, 16x16x4 MFMA wrappers and their accumulator layout,
// Newton-refined reciprocal square root / reciprocal.  Device code only;
// included by one translation unit each (internal linkage).
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace gba_amd {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double readlane(double v, int l) {
    int2 p = __builtin_bit_cast(int2, v);
    p.x = __builtin_amdgcn_readlane(p.x, l);
    p.y = __builtin_amdgcn_readlane(p.y, l);
    return __builtin_bit_cast(double, p);
}
__device__ __forceinline__ float readlane(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

__device__ __forceinline__ v4d mfma16(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma16(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// 1/sqrt(d) from the hardware estimate refined by Newton steps
// r <- r (3 - d r^2) / 2: two steps for FP64; none for FP32, whose estimate
// (v_rsq_f32) is within about one ulp -- the FP32 factor only preconditions
// an FP64 refinement, and each step removed is four dependent operations on
// the pivot chain.  Used for the pivots of the tile factor: the pivot is d r
// and the column scale is r, with no division in the chain.
__device__ __forceinline__ double rsqrt_refined(double d) {
    double r = __builtin_amdgcn_rsq(d);
    r = r * (1.5 - 0.5 * d * r * r);
    r = r * (1.5 - 0.5 * d * r * r);
    return r;
}
__device__ __forceinline__ float rsqrt_refined(float d) { return __builtin_amdgcn_rsqf(d); }

// 1/x from the hardware estimate refined by Newton steps r <- r (2 - x r)
// (FP64: two; FP32: none, v_rcp_f32 is within about one ulp).
__device__ __forceinline__ double recip_refined(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = r * (2.0 - x * r);
    r = r * (2.0 - x * r);
    return r;
}
__device__ __forceinline__ float recip_refined(float x) { return __builtin_amdgcn_rcpf(x); }

template <typename T> struct Acc4;
template <> struct Acc4<double> { typedef v4d type; };
template <> struct Acc4<float> { typedef v4f type; };
// Row of accumulator element q held by lane group g = lane / 16 (column =
// lane % 16): v_mfma_f32_16x16x4f32 uses row 4g + q, v_mfma_f64_16x16x4f64
// uses row g + 4q.
template <typename T> __device__ __forceinline__ int acc_row(int g, int q);
template <> __device__ __forceinline__ int acc_row<float>(int g, int q) { return 4 * g + q; }
template <> __device__ __forceinline__ int acc_row<double>(int g, int q) { return g + 4 * q; }

__device__ __forceinline__ bool any_fail(const TrialFlags* f) {
    return f->nonspd != 0 || f->chol_fail != 0;
}

}  // namespace
}  // namespace gba_amd
