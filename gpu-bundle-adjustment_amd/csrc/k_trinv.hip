// k_trinv.hip -- explicit inverse X = L^-1 of the tiled Cholesky factor and
// its application as the PCG preconditioner M^-1 r = X^T (X r).
//
// The reference solves the reduced camera system once per trial with
// Eigen::LLT (src/lm.cu:597-604).  The mixed form here solves it with FP64
// conjugate gradients preconditioned by the FP32 factor, which applies
// (L L^T)^-1 several times per trial.  Block triangular substitution is a
// chain of 64-row steps bounded by inter-workgroup latency; applying an
// explicit inverse is two tile-parallel matrix-vector passes bounded by HBM
// bandwidth instead.  X is built once per factorisation by recursive
// doubling over aligned tile blocks: with A = [s, s+h) and B = [s+h, s+2h),
// X_BA = -X_BB (L_BA X_AA), every merge of one level in the same two launches
// (T = L_BA X_AA, then X_BA = -X_BB T), all products as 64x64 MFMA tiles.
//
// Layout (row-major, leading dimension ld, 64x64 tiles): X's lower tiles hold
// X_ij, the upper tile (j, i) holds X_ij^T, and a diagonal tile holds X_ii
// below and X_ii^T above its diagonal -- so both passes of the preconditioner
// read their operand rows contiguously (one lane per column, no cross-lane
// reduction).
#include "device_common.hpp"
#include "launchers.hpp"
#include "mfma_tile.hpp"

namespace gba_amd {

namespace {

// 256 threads copy a 64x64 tile into LDS; with lower_only the strictly upper
// part is replaced by zeros (a triangular operand).
template <typename T>
__device__ __forceinline__ void stage_tile(T (*s)[kTile + 1], const T* __restrict__ src, int ld,
                                           bool lower_only) {
    T tmp[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int q = threadIdx.x + 256 * u;
        tmp[u] = src[static_cast<size_t>(q >> 6) * ld + (q & 63)];
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int q = threadIdx.x + 256 * u;
        const int r = q >> 6, c = q & 63;
        s[r][c] = (lower_only && c > r) ? T(0) : tmp[u];
    }
}

// acc += A B for one 64x64 tile pair: A from LDS, B read row by row from
// global memory (a 16-lane group reads 16 consecutive entries of row kk).
// With b_lower_only, B is a diagonal tile whose strictly upper part is
// treated as zero.  Four wavefronts, 2x2 16x16 MFMA blocks each.
template <typename T>
__device__ __forceinline__ void tile_mfma_acc(typename Acc4<T>::type (&acc)[2][2],
                                              const T (*sA)[kTile + 1], const T* __restrict__ B,
                                              int ld, bool b_lower_only) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int kk = 0; kk < kTile; kk += 4) {
        const T* brow = B + static_cast<size_t>(kk + lk) * ld;
        T b0 = brow[wc + li], b1 = brow[wc + 16 + li];
        if (b_lower_only) {
            if (kk + lk < wc + li) b0 = T(0);
            if (kk + lk < wc + 16 + li) b1 = T(0);
        }
        const T a0 = sA[wr + li][kk + lk], a1 = sA[wr + 16 + li][kk + lk];
        acc[0][0] = mfma16(a0, b0, acc[0][0]);
        acc[0][1] = mfma16(a0, b1, acc[0][1]);
        acc[1][0] = mfma16(a1, b0, acc[1][0]);
        acc[1][1] = mfma16(a1, b1, acc[1][1]);
    }
}

template <typename T>
__device__ __forceinline__ void acc_zero(typename Acc4<T>::type (&acc)[2][2]) {
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[r][c] = typename Acc4<T>::type{0, 0, 0, 0};
}

// Block of a level-h merge: merge m covers tiles [2hm, 2hm + 2h); the block
// owns output tile (i, j), i in the upper half B, j in the lower half A.
__device__ __forceinline__ void merge_tile(int h, int& s, int& i, int& j) {
    const int per = h * h;
    const int m = blockIdx.x / per, rem = blockIdx.x - m * per;
    s = 2 * h * m;
    i = s + h + rem / h;
    j = s + rem % h;
}

// Diagonal tiles of X from the stored diagonal inverses (DinvT[b] holds
// (L_bb^-1)^T): X_bb below the diagonal, X_bb^T above it.
template <typename T>
__global__ __launch_bounds__(256) void k_trinv_diag(const T* __restrict__ DinvT, T* __restrict__ X,
                                                    int ld, const TrialFlags* flags) {
    if (flags->nonspd || flags->chol_fail) return;
    const int b = blockIdx.x;
    const T* it = DinvT + static_cast<size_t>(b) * kTile * kTile;
    T* xt = X + static_cast<size_t>(b * kTile) * ld + b * kTile;
    for (int q = threadIdx.x; q < kTile * kTile; q += 256) {
        const int r = q >> 6, c = q & 63;
        // Linv[r][c] = DinvT[c][r]; entry (r, c) is Linv[max][min]
        xt[static_cast<size_t>(r) * ld + c] = (r >= c) ? it[c * kTile + r] : it[r * kTile + c];
    }
}

// T_ij = sum_{j <= k < s+h} L_ik X_kj for the merges of level h.
// kStage k steps share one staging round (their A tiles in LDS at once),
// which divides the barriers and multiplies the loads in flight of these
// latency-bound chains.
constexpr int kStage = 2;
template <typename T>
__global__ __launch_bounds__(256) void k_trinv_lx(const T* __restrict__ L, const T* __restrict__ X,
                                                  T* __restrict__ Tb, int ld, int nblk, int h,
                                                  const TrialFlags* flags) {
    __shared__ T sA[kStage][kTile][kTile + 1];
    if (flags->nonspd || flags->chol_fail) return;
    int s, i, j;
    merge_tile(h, s, i, j);
    if (i >= nblk) return;
    typename Acc4<T>::type acc[2][2];
    acc_zero<T>(acc);
    for (int k = j; k < s + h; k += kStage) {
        const int n = min(kStage, s + h - k);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kStage; ++u)
            if (u < n) stage_tile<T>(sA[u], L + static_cast<size_t>(i * kTile) * ld + (k + u) * kTile, ld, false);
        __syncthreads();
        for (int u = 0; u < n; ++u)
            tile_mfma_acc<T>(acc, sA[u], X + static_cast<size_t>((k + u) * kTile) * ld + j * kTile, ld,
                             k + u == j);
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    T* out = Tb + static_cast<size_t>(i * kTile) * ld + j * kTile;
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                out[static_cast<size_t>(wr + 16 * r + acc_row<T>(lk, q)) * ld + wc + 16 * c + li] =
                    acc[r][c][q];
}

// X_ij = -sum_{s+h <= k <= i} X_ik T_kj; written to the lower tile (i, j)
// and, transposed, to the upper tile (j, i).
template <typename T>
__global__ __launch_bounds__(256) void k_trinv_xt(T* __restrict__ X, const T* __restrict__ Tb,
                                                  int ld, int nblk, int h, const TrialFlags* flags) {
    __shared__ T sA[kStage][kTile][kTile + 1];
    if (flags->nonspd || flags->chol_fail) return;
    int s, i, j;
    merge_tile(h, s, i, j);
    if (i >= nblk) return;
    typename Acc4<T>::type acc[2][2];
    acc_zero<T>(acc);
    for (int k = s + h; k <= i; k += kStage) {
        const int n = min(kStage, i + 1 - k);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kStage; ++u)
            if (u < n)
                stage_tile<T>(sA[u], X + static_cast<size_t>(i * kTile) * ld + (k + u) * kTile, ld, k + u == i);
        __syncthreads();
        for (int u = 0; u < n; ++u)
            tile_mfma_acc<T>(acc, sA[u], Tb + static_cast<size_t>((k + u) * kTile) * ld + j * kTile, ld, false);
    }
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    __syncthreads();  // every wave is done reading sA
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) sA[0][wr + 16 * r + acc_row<T>(lk, q)][wc + 16 * c + li] = -acc[r][c][q];
    __syncthreads();
    T* low = X + static_cast<size_t>(i * kTile) * ld + j * kTile;
    T* up = X + static_cast<size_t>(j * kTile) * ld + i * kTile;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int q = threadIdx.x + 256 * u;
        const int r = q >> 6, c = q & 63;
        low[static_cast<size_t>(r) * ld + c] = sA[0][r][c];
        up[static_cast<size_t>(r) * ld + c] = sA[0][c][r];
    }
}

// One lower tile (i, j) per workgroup, one lane per column, four wavefronts
// of 16 rows, FP64 accumulation.
//   mode 0 (y = X v): rows of the upper tile (j, i) (= X_ij^T) times v_j,
//                     partial of output block i; diagonal: rows c <= column.
//   mode 1 (y = X^T v): rows of the lower tile (i, j) times v_i, partial of
//                     output block j; diagonal: rows r >= column.
template <typename T>
__global__ __launch_bounds__(256) void k_tile_colmv(const T* __restrict__ X, int ld, int mode,
                                                    const double* __restrict__ v,
                                                    double* __restrict__ part) {
    __shared__ double s_part[4][kTile];
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
                                                           double* __restrict__ out) {
    __shared__ double s_part[4][kTile];
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

void launch_trinv(hipStream_t st, const float* L32, const float* D32, int ld, int nblk, float* X32,
                  float* T32, const TrialFlags* flags) {
    const float* DinvT = D32 + static_cast<size_t>(nblk) * kTile * kTile;
    hipLaunchKernelGGL(k_trinv_diag<float>, dim3(nblk), dim3(256), 0, st, DinvT, X32, ld, flags);
    for (int h = 1; h < nblk; h *= 2) {
        const int merges = (nblk + 2 * h - 1) / (2 * h);
        const int grid = merges * h * h;
        hipLaunchKernelGGL(k_trinv_lx<float>, dim3(grid), dim3(256), 0, st, L32, X32, T32, ld, nblk, h,
                           flags);
        hipLaunchKernelGGL(k_trinv_xt<float>, dim3(grid), dim3(256), 0, st, X32, T32, ld, nblk, h, flags);
    }
}

size_t precond_work(int npad) {
    const size_t nblk = static_cast<size_t>(npad / kTile);
    return (nblk * (nblk + 1) / 2) * kTile + static_cast<size_t>(npad);
}

void launch_precond_apply(hipStream_t st, const float* X32, int ld, int npad, const double* r,
                          double* z, double* work) {
    const int nblk = npad / kTile;
    const int tiles = nblk * (nblk + 1) / 2;
    double* part = work;
    double* y = work + static_cast<size_t>(tiles) * kTile;
    hipLaunchKernelGGL(k_tile_colmv<float>, dim3(tiles), dim3(256), 0, st, X32, ld, 0, r, part);
    hipLaunchKernelGGL(k_tile_colmv_reduce, dim3(nblk), dim3(256), 0, st, nblk, 0, part, y);
    hipLaunchKernelGGL(k_tile_colmv<float>, dim3(tiles), dim3(256), 0, st, X32, ld, 1, y, part);
    hipLaunchKernelGGL(k_tile_colmv_reduce, dim3(nblk), dim3(256), 0, st, nblk, 1, part, z);
}

}  // namespace gba_amd
