// schur.hpp -- host-callable launchers of the batched small-block kernels
// (the reference exposes them as __global__ kernels in include/gba/schur.cuh:
// 104-116; a launch of another translation unit's kernel is not a stable ABI,
// so these take device pointers and a hipStream_t instead).
#pragma once

#include <hip/hip_runtime.h>

namespace gba {

// Factor each symmetric 9x9 A_i (row-major, lower triangle read) and solve
// A_i x_i = b_i; ok[i] = 0 where A_i is not SPD.  Throws on a launch error.
void batch_chol9_factor_solve(const double* A, const double* b, double* x, int* ok, int n,
                              hipStream_t stream = nullptr);

// Invert each SPD 3x3 A_i (full row-major output); ok[i] = 0 where not SPD.
void batch_inverse3_spd(const double* A, double* Ainv, int* ok, int n, hipStream_t stream = nullptr);

}  // namespace gba
