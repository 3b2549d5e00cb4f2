// k_blocks.hip -- batched small SPD blocks and small utility kernels.
//
// Replaces the reference's damp_invert_v_kernel (src/lm.cu:289-302) and the
// batched kernels of include/gba/schur.cuh:98-116.  One lane owns one block, so
// a wavefront packs 64 blocks; blocks are staged through LDS so the global
// loads are coalesced.  Non-SPD blocks are reported, never silently used.
#include "device_common.hpp"
#include "launchers.hpp"

namespace gba_amd {

namespace {

__global__ void k_clear_flags(TrialFlags* f) {
    f->nonspd = 0;
    f->chol_fail = 0;
    f->solve_err = 0;
    f->pad = 0;
}

__global__ void k_fill_int(int* p, int n, int v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_double(double* p, size_t n, double v) {
    const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// V' = V + lambda diag'(V) per point, inverted (lower-6 packed in and out).
__global__ __launch_bounds__(256) void k_damp_invert_v(int P, const double* __restrict__ VB,
                                                       double lambda, double* __restrict__ Vinv6,
                                                       TrialFlags* flags, const int* __restrict__ pt_prob,
                                                       const double* __restrict__ lam,
                                                       const int* __restrict__ active) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= P) return;
    int q = 0;
    if (pt_prob) {  // batched problems: this point's problem
        q = pt_prob[p];
        if (active && !active[q]) return;
        lambda = lam[q];
    }
    const double* v = VB + 9 * static_cast<size_t>(p);
    const double a[6] = {v[0] + lambda * damp_diag(v[0]), v[1], v[2] + lambda * damp_diag(v[2]),
                         v[3], v[4], v[5] + lambda * damp_diag(v[5])};
    double inv[6];
    if (spd3_inverse(a, inv)) {
        double* o = Vinv6 + 6 * static_cast<size_t>(p);
#pragma unroll
        for (int i = 0; i < 6; ++i) o[i] = inv[i];
    } else {
        atomicAdd(&flags[q].nonspd, 1);
    }
}

// 9x9: packed lower triangle in registers, row-oriented Cholesky, then the two
// triangular solves.  64 blocks per workgroup (one wavefront).
__global__ __launch_bounds__(64) void k_batch_chol9(const double* __restrict__ A,
                                                    const double* __restrict__ b,
                                                    double* __restrict__ x, int* __restrict__ ok,
                                                    int n) {
    __shared__ double sA[64 * 81];
    const int base = blockIdx.x * 64;
    const int cnt = min(64, n - base);
    for (int t = threadIdx.x; t < cnt * 81; t += 64) sA[t] = A[81 * static_cast<size_t>(base) + t];
    __syncthreads();
    const int i = threadIdx.x;
    if (i >= cnt) return;
    const double* a = sA + 81 * i;
    double L[45];
    bool good = true;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
#pragma unroll
        for (int c = 0; c <= r; ++c) {
            double s = a[9 * r + c];
#pragma unroll
            for (int k = 0; k < c; ++k) s -= L[r * (r + 1) / 2 + k] * L[c * (c + 1) / 2 + k];
            if (r == c) {
                good = good && (s > 0.0);
                L[r * (r + 1) / 2 + r] = sqrt(s);
            } else {
                L[r * (r + 1) / 2 + c] = s / L[c * (c + 1) / 2 + c];
            }
        }
    }
    const size_t g = static_cast<size_t>(base) + i;
    ok[g] = good ? 1 : 0;
    if (!good) return;
    double y[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        double s = b[9 * g + r];
#pragma unroll
        for (int k = 0; k < r; ++k) s -= L[r * (r + 1) / 2 + k] * y[k];
        y[r] = s / L[r * (r + 1) / 2 + r];
    }
#pragma unroll
    for (int r = 8; r >= 0; --r) {
        double s = y[r];
#pragma unroll
        for (int k = r + 1; k < 9; ++k) s -= L[k * (k + 1) / 2 + r] * y[k];
        y[r] = s / L[r * (r + 1) / 2 + r];
    }
#pragma unroll
    for (int r = 0; r < 9; ++r) x[9 * g + r] = y[r];
}

// 3x3 SPD inverse of full row-major blocks (lower triangle read), full output.
__global__ __launch_bounds__(256) void k_batch_inverse3(const double* __restrict__ A,
                                                        double* __restrict__ Ainv,
                                                        int* __restrict__ ok, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double* a = A + 9 * static_cast<size_t>(i);
    const double lo[6] = {a[0], a[3], a[4], a[6], a[7], a[8]};
    double inv[6];
    const bool good = spd3_inverse(lo, inv);
    ok[i] = good ? 1 : 0;
    if (!good) return;
    double* o = Ainv + 9 * static_cast<size_t>(i);
    o[0] = inv[0]; o[1] = inv[1]; o[2] = inv[3];
    o[3] = inv[1]; o[4] = inv[2]; o[5] = inv[4];
    o[6] = inv[3]; o[7] = inv[4]; o[8] = inv[5];
}

// Padding rows/columns [n, npad) of the dense matrix: identity.
template <typename T>
__global__ void k_identity_pad(T* S, int ld, int n, int npad) {
    const int r = n + blockIdx.x;
    if (r >= npad) return;
    for (int c = threadIdx.x; c < npad; c += blockDim.x) {
        S[static_cast<size_t>(r) * ld + c] = (c == r) ? T(1) : T(0);
        S[static_cast<size_t>(c) * ld + r] = (c == r) ? T(1) : T(0);
    }
}

}  // namespace

void launch_clear_flags(hipStream_t st, TrialFlags* flags) {
    hipLaunchKernelGGL(k_clear_flags, dim3(1), dim3(1), 0, st, flags);
}

void launch_fill_int(hipStream_t st, int* p, int n, int v) {
    if (n > 0) hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
}

void launch_fill_double(hipStream_t st, double* p, size_t n, double v) {
    if (n > 0)
        hipLaunchKernelGGL(k_fill_double, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                           st, p, n, v);
}

void launch_damp_invert_v(hipStream_t st, int P, const double* VB, double lambda, double* Vinv6,
                          TrialFlags* flags) {
    if (P > 0)
        hipLaunchKernelGGL(k_damp_invert_v, dim3((P + 255) / 256), dim3(256), 0, st, P, VB, lambda,
                           Vinv6, flags, static_cast<const int*>(nullptr),
                           static_cast<const double*>(nullptr), static_cast<const int*>(nullptr));
}

void launch_damp_invert_v_batched(hipStream_t st, int P, const double* VB, const int* pt_prob,
                                  const double* lam, const int* active, double* Vinv6,
                                  TrialFlags* flags) {
    if (P > 0)
        hipLaunchKernelGGL(k_damp_invert_v, dim3((P + 255) / 256), dim3(256), 0, st, P, VB, 0.0, Vinv6,
                           flags, pt_prob, lam, active);
}

void launch_batch_chol9(hipStream_t st, const double* A, const double* b, double* x, int* ok, int n) {
    if (n > 0) hipLaunchKernelGGL(k_batch_chol9, dim3((n + 63) / 64), dim3(64), 0, st, A, b, x, ok, n);
}

void launch_batch_inverse3(hipStream_t st, const double* A, double* Ainv, int* ok, int n) {
    if (n > 0)
        hipLaunchKernelGGL(k_batch_inverse3, dim3((n + 255) / 256), dim3(256), 0, st, A, Ainv, ok, n);
}

void launch_identity_pad(hipStream_t st, double* S64, float* S32, int ld, int n, int npad) {
    if (npad <= n) return;
    if (S64) hipLaunchKernelGGL(k_identity_pad<double>, dim3(npad - n), dim3(256), 0, st, S64, ld, n, npad);
    if (S32) hipLaunchKernelGGL(k_identity_pad<float>, dim3(npad - n), dim3(256), 0, st, S32, ld, n, npad);
}

}  // namespace gba_amd
