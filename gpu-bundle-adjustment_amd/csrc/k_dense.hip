// k_dense.hip -- dense SPD factorisation and solves of the reduced camera
// system on the GPU (replaces Eigen::LLT<MatrixXd, Lower> + solve at
// src/lm.cu:597-604, which the reference runs single-threaded on the host).
//
// S is row-major, leading dimension ld = npad (9C rounded up to 64); only the
// lower triangle is read; padding rows/columns are the identity, so the factor
// of the padded matrix is block-diagonal [L, I].
//
// Right-looking blocked Cholesky on 64x64 tiles, per panel k:
//   panel  : one wavefront factors the diagonal tile in registers (lane = row,
//            pivots broadcast with v_readlane), every other wavefront of the
//            launch solves one tile of the panel column against it (TRSM,
//            lane = row, L_kk broadcast from LDS);
//   update : A_ij -= L_ik L_jk^T for all trailing lower tiles, one workgroup per
//            tile, four wavefronts of 2x2 MFMA 16x16 blocks
//            (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32), operands from LDS.
// A pivot that is not > 0 sets TrialFlags::chol_fail (the LM then escalates
// lambda, as the reference does on LLT failure).
//
// Triangular solves: one persistent launch per direction; workgroups take
// tile rows in dependency order through an atomic ticket, so a workgroup only
// ever waits on tile rows already owned by running workgroups; tile-row
// results are published with release/acquire flags at agent scope.  Waits are
// bounded (TrialFlags::solve_err on time-out), so every wave exits.
#include "device_common.hpp"
#include "launchers.hpp"

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

template <typename T>
__global__ __launch_bounds__(256) void k_chol_panel(T* __restrict__ S, int ld, int nblk, int k,
                                                    T* __restrict__ Ldiag, TrialFlags* flags) {
    __shared__ T Lkk[kTile][kTile + 1];
    __shared__ T invd[kTile];
    if (any_fail(flags)) return;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (w == 0) {
        const T* row = S + static_cast<size_t>(k * kTile + lane) * ld + k * kTile;
        T a[kTile];
#pragma unroll
        for (int c = 0; c < kTile; ++c) a[c] = row[c];
        bool bad = false;
#pragma unroll
        for (int j = 0; j < kTile; ++j) {
            const T d = readlane(a[j], j);
            bad = bad || !(d > T(0));
            const T piv = sqrt(d);
            const T inv = T(1) / piv;
            a[j] = (lane == j) ? piv : a[j] * inv;
#pragma unroll
            for (int m = j + 1; m < kTile; ++m) a[m] -= a[j] * readlane(a[j], m);
        }
#pragma unroll
        for (int c = 0; c < kTile; ++c) Lkk[lane][c] = (c <= lane) ? a[c] : T(0);
        invd[lane] = T(1) / a[lane];
        // every workgroup read A_kk above; only the diagonal-tile array
        // receives L_kk, so no workgroup can observe a half-written tile
        if (blockIdx.x == 0) {
            T* dst = Ldiag + static_cast<size_t>(k) * kTile * kTile + lane * kTile;
#pragma unroll
            for (int c = 0; c < kTile; ++c) dst[c] = (c <= lane) ? a[c] : T(0);
            if (bad && lane == 0) atomicOr(&flags->chol_fail, 1);
        }
    }
    __syncthreads();
    const int i = k + 1 + blockIdx.x * 4 + w;
    if (i >= nblk) return;
    T* row = S + static_cast<size_t>(i * kTile + lane) * ld + k * kTile;
    T x[kTile];
#pragma unroll
    for (int c = 0; c < kTile; ++c) x[c] = row[c];
    // x L_kk^T = a, column by column
#pragma unroll
    for (int j = 0; j < kTile; ++j) {
        x[j] *= invd[j];
#pragma unroll
        for (int m = j + 1; m < kTile; ++m) x[m] -= x[j] * Lkk[m][j];
    }
#pragma unroll
    for (int c = 0; c < kTile; ++c) row[c] = x[c];
}

template <typename T>
__global__ __launch_bounds__(256) void k_chol_update(T* __restrict__ S, int ld, int k,
                                                     const TrialFlags* flags) {
    __shared__ T Ai[kTile][kTile + 2];
    __shared__ T Aj[kTile][kTile + 2];
    if (any_fail(flags)) return;
    // tile index -> (ii, jj), jj <= ii, over the trailing lower triangle
    const int t = blockIdx.x;
    int ii = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while (ii * (ii + 1) / 2 > t) --ii;
    while ((ii + 1) * (ii + 2) / 2 <= t) ++ii;
    const int jj = t - ii * (ii + 1) / 2;
    const int i = k + 1 + ii, j = k + 1 + jj;
    const T* Lik = S + static_cast<size_t>(i * kTile) * ld + k * kTile;
    const T* Ljk = S + static_cast<size_t>(j * kTile) * ld + k * kTile;
    for (int q = threadIdx.x; q < kTile * kTile; q += 256) {
        const int r = q >> 6, c = q & 63;
        Ai[r][c] = Lik[static_cast<size_t>(r) * ld + c];
        Aj[r][c] = Ljk[static_cast<size_t>(r) * ld + c];
    }
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    typename Acc4<T>::type acc[2][2];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[r][c] = typename Acc4<T>::type{0, 0, 0, 0};
#pragma unroll 4
    for (int kk = 0; kk < kTile; kk += 4) {
        const T a0 = Ai[wr + li][kk + lk], a1 = Ai[wr + 16 + li][kk + lk];
        const T b0 = Aj[wc + li][kk + lk], b1 = Aj[wc + 16 + li][kk + lk];
        acc[0][0] = mfma16(a0, b0, acc[0][0]);
        acc[0][1] = mfma16(a0, b1, acc[0][1]);
        acc[1][0] = mfma16(a1, b0, acc[1][0]);
        acc[1][1] = mfma16(a1, b1, acc[1][1]);
    }
    T* Aij = S + static_cast<size_t>(i * kTile) * ld + j * kTile;
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = wr + 16 * r + acc_row<T>(lk, q), col = wc + 16 * c + li;
                Aij[static_cast<size_t>(row) * ld + col] -= acc[r][c][q];
            }
}

// Forward solve L y = b.  Workgroup with ticket t owns tile row i = t.
template <typename T>
__global__ __launch_bounds__(256) void k_trsv_fwd(const T* __restrict__ L, int ld, int nblk,
                                                  const T* __restrict__ Ldiag,
                                                  const double* __restrict__ b, double* y, int* done,
                                                  int* ticket, TrialFlags* flags) {
    __shared__ int s_blk, s_skip;
    __shared__ double s_part[4][16];
    __shared__ double s_rhs[kTile];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    if (tid == 0) {
        s_blk = atomicAdd(ticket, 1);
        s_skip = (any_fail(flags) || flags->solve_err) ? 1 : 0;
    }
    __syncthreads();
    if (s_skip) return;
    const int i = s_blk;
    double acc[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0;
    for (int j = 0; j < i; ++j) {
        if (tid == 0 && !wait_flag(done + j, 1)) atomicOr(&flags->solve_err, 1);
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const double yj = y[j * kTile + lane];
        const T* Lr = L + static_cast<size_t>(i * kTile + 16 * w) * ld + j * kTile + lane;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] += static_cast<double>(Lr[static_cast<size_t>(r) * ld]) * yj;
    }
    double mine = 0.0;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const double t = wave_sum(acc[r]);
        if (lane == r) mine = t;
    }
    if (lane < 16) s_part[w][lane] = mine;
    __syncthreads();
    if (tid < kTile) s_rhs[tid] = b[i * kTile + tid] - s_part[tid >> 4][tid & 15];
    __syncthreads();
    if (w == 0) {
        const T* Lrow = Ldiag + static_cast<size_t>(i) * kTile * kTile + lane * kTile;
        T Lr[kTile];
#pragma unroll
        for (int c = 0; c < kTile; ++c) Lr[c] = Lrow[c];
        double v = s_rhs[lane];
#pragma unroll
        for (int j = 0; j < kTile; ++j) {
            const double yj = readlane(v, j) / static_cast<double>(readlane(Lr[j], j));
            if (lane == j) v = yj;
            else if (lane > j) v -= static_cast<double>(Lr[j]) * yj;
        }
        y[i * kTile + lane] = v;
        if (lane == 0) __hip_atomic_store(done + i, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Backward solve L^T x = y.  Ticket t owns tile row i = nblk - 1 - t.
template <typename T>
__global__ __launch_bounds__(256) void k_trsv_bwd(const T* __restrict__ L, int ld, int nblk,
                                                  const T* __restrict__ Ldiag,
                                                  const double* __restrict__ y, double* x, int* done,
                                                  int* ticket, TrialFlags* flags) {
    __shared__ int s_blk, s_skip;
    __shared__ double s_part[4][kTile];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    if (tid == 0) {
        s_blk = nblk - 1 - atomicAdd(ticket, 1);
        s_skip = (any_fail(flags) || flags->solve_err) ? 1 : 0;
    }
    __syncthreads();
    if (s_skip) return;
    const int i = s_blk;
    double acc = 0.0;
    for (int j = nblk - 1; j > i; --j) {
        if (tid == 0 && !wait_flag(done + j, 1)) atomicOr(&flags->solve_err, 1);
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const T* Lc = L + static_cast<size_t>(j * kTile + 16 * w) * ld + i * kTile + lane;
        const double* xj = x + j * kTile + 16 * w;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc += static_cast<double>(Lc[static_cast<size_t>(r) * ld]) * xj[r];
    }
    s_part[w][lane] = acc;
    __syncthreads();
    if (w == 0) {
        double s = y[i * kTile + lane] - ((s_part[0][lane] + s_part[1][lane]) + (s_part[2][lane] + s_part[3][lane]));
        // lane r holds column r of the diagonal tile: Lc[j] = L[j][r], j >= r
        const T* Lcol = Ldiag + static_cast<size_t>(i) * kTile * kTile + lane;
        T Lc[kTile];
#pragma unroll
        for (int j = 0; j < kTile; ++j) Lc[j] = Lcol[j * kTile];
#pragma unroll
        for (int j = kTile - 1; j >= 0; --j) {
            const double xjv = readlane(s, j) / static_cast<double>(readlane(Lc[j], j));
            if (lane == j) s = xjv;
            else if (lane < j) s -= static_cast<double>(Lc[j]) * xjv;
        }
        x[i * kTile + lane] = s;
        if (lane == 0) __hip_atomic_store(done + i, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// r = b - A x (A full symmetric FP32, FP64 accumulation), one wave per row;
// per-workgroup partial of |r|^2 (and |b|^2 when asked).
__global__ __launch_bounds__(256) void k_residual(const float* __restrict__ A, int ld, int npad,
                                                  const double* __restrict__ b,
                                                  const double* __restrict__ x, double* __restrict__ r,
                                                  double* __restrict__ blk) {
    __shared__ double s_r[4], s_b[4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + w;
    double rr = 0.0, bb = 0.0;
    if (row < npad) {
        const float* Ar = A + static_cast<size_t>(row) * ld;
        double s = 0.0;
        for (int c = lane; c < npad; c += 64) s += static_cast<double>(Ar[c]) * x[c];
        s = wave_sum(s);
        const double bv = b[row];
        const double v = bv - s;
        if (lane == 0) r[row] = v;
        rr = v * v;
        bb = bv * bv;
    }
    if (lane == 0) {
        s_r[w] = rr;
        s_b[w] = bb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        blk[2 * blockIdx.x] = (s_r[0] + s_r[1]) + (s_r[2] + s_r[3]);
        blk[2 * blockIdx.x + 1] = (s_b[0] + s_b[1]) + (s_b[2] + s_b[3]);
    }
}

__global__ __launch_bounds__(256) void k_residual_reduce(int n, const double* __restrict__ blk,
                                                         PassStatus* status, int also_rhs,
                                                         const TrialFlags* flags) {
    __shared__ double s_r[256], s_b[256];
    const int tid = threadIdx.x;
    double a = 0.0, c = 0.0;
    for (int u = tid; u < n; u += 256) {
        a += blk[2 * u];
        c += blk[2 * u + 1];
    }
    s_r[tid] = a;
    s_b[tid] = c;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) {
            s_r[tid] += s_r[tid + h];
            s_b[tid] += s_b[tid + h];
        }
        __syncthreads();
    }
    if (tid == 0) {
        status->res_norm2 = s_r[0];
        if (also_rhs) status->rhs_norm2 = s_b[0];
        status->flags = *flags;
    }
}

__global__ void k_vec_axpy(int n, double* __restrict__ x, const double* __restrict__ d) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] += d[i];
}

__global__ void k_vec_scale(int n, int npad, const double* __restrict__ in,
                            const double* __restrict__ s, double* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < npad) out[i] = i < n ? in[i] * s[i] : 0.0;
}

}  // namespace

void launch_cholesky_panel(hipStream_t st, double* S64, float* S32, int ld, int nblk, int k,
                           double* D64, float* D32, TrialFlags* flags) {
    const int rest = nblk - k - 1;
    const int grid = rest > 0 ? (rest + 3) / 4 : 1;
    if (S64)
        hipLaunchKernelGGL(k_chol_panel<double>, dim3(grid), dim3(256), 0, st, S64, ld, nblk, k, D64,
                           flags);
    else
        hipLaunchKernelGGL(k_chol_panel<float>, dim3(grid), dim3(256), 0, st, S32, ld, nblk, k, D32,
                           flags);
}

void launch_cholesky_update(hipStream_t st, double* S64, float* S32, int ld, int nblk, int k,
                            TrialFlags* flags) {
    const int m = nblk - k - 1;
    if (m <= 0) return;
    const int tiles = m * (m + 1) / 2;
    if (S64)
        hipLaunchKernelGGL(k_chol_update<double>, dim3(tiles), dim3(256), 0, st, S64, ld, k, flags);
    else
        hipLaunchKernelGGL(k_chol_update<float>, dim3(tiles), dim3(256), 0, st, S32, ld, k, flags);
}

void launch_cholesky(hipStream_t st, double* S64, float* S32, int ld, int nblk, double* D64,
                     float* D32, TrialFlags* flags) {
    for (int k = 0; k < nblk; ++k) {
        launch_cholesky_panel(st, S64, S32, ld, nblk, k, D64, D32, flags);
        launch_cholesky_update(st, S64, S32, ld, nblk, k, flags);
    }
}

void launch_chol_solve(hipStream_t st, const double* L64, const float* L32, int ld, int nblk,
                       const double* D64, const float* D32, const double* b, double* y, double* x,
                       int* sync, TrialFlags* flags) {
    launch_fill_int(st, sync, 2 * nblk + 2, 0);
    int* done_f = sync;
    int* done_b = sync + nblk;
    int* tickets = sync + 2 * nblk;
    if (L64) {
        hipLaunchKernelGGL(k_trsv_fwd<double>, dim3(nblk), dim3(256), 0, st, L64, ld, nblk, D64, b,
                           y, done_f, tickets, flags);
        hipLaunchKernelGGL(k_trsv_bwd<double>, dim3(nblk), dim3(256), 0, st, L64, ld, nblk, D64, y,
                           x, done_b, tickets + 1, flags);
    } else {
        hipLaunchKernelGGL(k_trsv_fwd<float>, dim3(nblk), dim3(256), 0, st, L32, ld, nblk, D32, b,
                           y, done_f, tickets, flags);
        hipLaunchKernelGGL(k_trsv_bwd<float>, dim3(nblk), dim3(256), 0, st, L32, ld, nblk, D32, y,
                           x, done_b, tickets + 1, flags);
    }
}

int residual_partials(int npad) { return (npad + 3) / 4; }

void launch_residual(hipStream_t st, const float* A, int ld, int npad, const double* b,
                     const double* x, double* r, double* blk_norm2, PassStatus* status,
                     bool also_rhs_norm, const TrialFlags* flags) {
    const int grid = residual_partials(npad);
    hipLaunchKernelGGL(k_residual, dim3(grid), dim3(256), 0, st, A, ld, npad, b, x, r, blk_norm2);
    hipLaunchKernelGGL(k_residual_reduce, dim3(1), dim3(256), 0, st, grid, blk_norm2, status,
                       also_rhs_norm ? 1 : 0, flags);
}

void launch_vec_axpy(hipStream_t st, int n, double* x, const double* d) {
    hipLaunchKernelGGL(k_vec_axpy, dim3((n + 255) / 256), dim3(256), 0, st, n, x, d);
}

void launch_vec_scale(hipStream_t st, int n, int npad, const double* in, const double* s,
                      double* out) {
    hipLaunchKernelGGL(k_vec_scale, dim3((npad + 255) / 256), dim3(256), 0, st, n, npad, in, s, out);
}

}  // namespace gba_amd
