// k_precond.hip -- the mixed-precision preconditioner M^-1 r = X^T (X r) with
// the explicit inverse X = L^-1 of the FP32 tile Cholesky factor.
//
// The reference solves the reduced camera system once per trial with
// Eigen::LLT (src/lm.cu:597-604).  The mixed form here solves it with FP64
// conjugate gradients preconditioned by the FP32 factor, which applies
// (L L^T)^-1 several times per trial.  Block triangular substitution is a
// chain of 64-row steps bounded by inter-workgroup latency; applying an
// explicit inverse is two tile-parallel matrix-vector passes bounded by HBM
// bandwidth instead.  X itself is formed inside the factor's launches
// (launch_cholesky with X32 set).
//
// Layout (row-major, leading dimension ld, 64x64 tiles): X's lower tiles hold
// X_ij, the upper tile (j, i) holds X_ij^T, and a diagonal tile holds X_ii
// below and X_ii^T above its diagonal -- so both passes read their operand
// rows contiguously (one lane per column, no cross-lane reduction).
#include "device_common.hpp"
#include "launchers.hpp"

namespace gba_amd {

namespace {

// One lower tile (i, j) per workgroup, one lane per column, four wavefronts
// of 16 rows, FP64 accumulation.
//   mode 0 (y = X v): rows of the upper tile (j, i) (= X_ij^T) times v_j,
//                     partial of output block i; diagonal: rows c <= column.
//   mode 1 (y = X^T v): rows of the lower tile (i, j) times v_i, partial of
//                     output block j; diagonal: rows r >= column.
template <typename T>
__global__ __launch_bounds__(256) void k_tile_colmv(const T* __restrict__ X, int ld, int mode,
                                                    const double* __restrict__ v,
                                                    double* __restrict__ part, const double* done) {
    __shared__ double s_part[4][kTile];
    if (pcg_skip(done)) return;
    const int t = blockIdx.x;
    int i = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while (i * (i + 1) / 2 > t) --i;
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    const int j = t - i * (i + 1) / 2;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const T* tile = (mode == 0) ? X + static_cast<size_t>(j * kTile) * ld + i * kTile
                                : X + static_cast<size_t>(i * kTile) * ld + j * kTile;
    const double* vs = v + ((mode == 0) ? j : i) * kTile;
    const bool diag = (i == j);
    double acc = 0.0;
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) {
        const int r = 16 * w + rr;
        const double x = static_cast<double>(tile[static_cast<size_t>(r) * ld + lane]);
        const bool keep = !diag || (mode == 0 ? r <= lane : r >= lane);
        acc += keep ? x * vs[r] : 0.0;
    }
    s_part[w][lane] = acc;
    __syncthreads();
    if (threadIdx.x < kTile)
        part[static_cast<size_t>(t) * kTile + threadIdx.x] =
            (s_part[0][threadIdx.x] + s_part[1][threadIdx.x]) + (s_part[2][threadIdx.x] + s_part[3][threadIdx.x]);
}

// out_b = sum of the partials of output block b in a fixed order: wavefront
// g sums the partials with index = g mod 4, the four sums are combined in
// order.
__global__ __launch_bounds__(256) void k_tile_colmv_reduce(int nblk, int mode,
                                                           const double* __restrict__ part,
                                                           double* __restrict__ out, const double* done) {
    __shared__ double s_part[4][kTile];
    if (pcg_skip(done)) return;
    const int b = blockIdx.x, lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    double s = 0.0;
    if (mode == 0) {
#pragma unroll 4
        for (int j = g; j <= b; j += 4) s += part[static_cast<size_t>(b * (b + 1) / 2 + j) * kTile + lane];
    } else {
#pragma unroll 4
        for (int i = b + g; i < nblk; i += 4) s += part[static_cast<size_t>(i * (i + 1) / 2 + b) * kTile + lane];
    }
    s_part[g][lane] = s;
    __syncthreads();
    if (g == 0) out[b * kTile + lane] = (s_part[0][lane] + s_part[1][lane]) + (s_part[2][lane] + s_part[3][lane]);
}

}  // namespace

size_t precond_work(int npad) {
    const size_t nblk = static_cast<size_t>(npad / kTile);
    return (nblk * (nblk + 1) / 2) * kTile + static_cast<size_t>(npad);
}

void launch_precond_apply(hipStream_t st, const float* X32, int ld, int npad, const double* r,
                          double* z, double* work, const double* done) {
    const int nblk = npad / kTile;
    const int tiles = nblk * (nblk + 1) / 2;
    double* part = work;
    double* y = work + static_cast<size_t>(tiles) * kTile;
    hipLaunchKernelGGL(k_tile_colmv<float>, dim3(tiles), dim3(256), 0, st, X32, ld, 0, r, part, done);
    hipLaunchKernelGGL(k_tile_colmv_reduce, dim3(nblk), dim3(256), 0, st, nblk, 0, part, y, done);
    hipLaunchKernelGGL(k_tile_colmv<float>, dim3(tiles), dim3(256), 0, st, X32, ld, 1, y, part, done);
    hipLaunchKernelGGL(k_tile_colmv_reduce, dim3(nblk), dim3(256), 0, st, nblk, 1, part, z, done);
}

}  // namespace gba_amd
