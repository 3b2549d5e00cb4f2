// k_dense.hip -- dense SPD factorisation and solves of the reduced camera
// system on the GPU (replaces Eigen::LLT<MatrixXd, Lower> + solve at
// src/lm.cu:597-604, which the reference runs single-threaded on the host).
//
// S is row-major, leading dimension ld = npad (9C rounded up to 64); only the
// lower triangle is read; padding rows/columns are the identity, so the factor
// of the padded matrix is block-diagonal [L, I].
//
// Right-looking blocked Cholesky on 64x64 tiles, per panel k:
//   potrf : one wavefront factors the diagonal tile in registers (lane = row,
//           pivots broadcast with v_readlane) and inverts it (lane = column,
//           forward substitution against the factor in LDS); L_kk and
//           L_kk^-T are kept in separate tile arrays;
//   trsm  : L_ik = A_ik L_kk^-T for every tile below, an MFMA GEMM per tile;
//   update: A_ij -= L_ik L_jk^T for all trailing lower tiles, one workgroup per
//           tile, four wavefronts of 2x2 MFMA 16x16 blocks
//           (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32), operands from LDS;
//           in the FP32 form, extra workgroups of update k form row k of the
//           explicit inverse X = L^-1 (the mixed form's preconditioner).
// A pivot that is not > 0 sets TrialFlags::chol_fail (the LM then escalates
// lambda, as the reference does on LLT failure).
//
// Triangular solves: one persistent launch per direction; workgroups take
// tile rows in dependency order through an atomic ticket, so a workgroup only
// ever waits on tile rows already owned by running workgroups; tile-row
// results are published with release/acquire flags at agent scope; the
// diagonal block is applied as a product with the stored inverse tile.  Waits
// are bounded (TrialFlags::solve_err on time-out), so every wave exits.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "device_common.hpp"
#include "launchers.hpp"
#include "mfma_tile.hpp"
#include "persist_tasks.hpp"
#include "solver.hpp"

namespace gba_amd {

namespace {

// Cholesky factor and inverse of the symmetric tile in a (lower triangle
// read), 256 threads or more (the trailing blocks are dealt to every
// wavefront, the inverse stages use the first four).  On return a holds L
// (lower; the upper part is stale) and b holds L^-1 (lower, zero above the diagonal inside the 16x16 diagonal
// blocks; the other upper blocks are scratch).  Blocked by 16 columns:
//   panel  : wavefront 0, one row per lane, 16 pivot steps; the pivot is
//            broadcast with v_readlane, the scaled column with v_readlane
//            (FP32) or one LDS store and broadcast reads (FP64, where each
//            entry would take two v_readlane);
//   update : the rank-16 update of the trailing triangle in 16x16 MFMA blocks;
//   inverse: each wavefront inverts one 16x16 diagonal block by forward
//            substitution (lane = column of the inverse, L entries read from
//            LDS as broadcasts), then by block distance d = 1..3 with MFMA:
//            Y = sum_{c <= m < r} L_rm Linv_mc, Linv_rc = -Linv_rr Y.
// Returns false on a pivot that is not > 0.
template <typename T>
__device__ __forceinline__ bool tile_potrf_inv(T (*a)[kTile + 2], T (*b)[kTile + 2]) {
    typedef typename Acc4<T>::type acc_t;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    __shared__ int s_bad;
    __shared__ T s_col[16];
    if (tid == 0) s_bad = 0;
    __syncthreads();
    for (int p = 0; p < 4; ++p) {
        const int c0 = 16 * p;
        if (w == 0) {
            T v[16];
#pragma unroll
            for (int m = 0; m < 16; ++m) v[m] = a[lane][c0 + m];
            bool bad = false;
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
                const T d = readlane(v[jj], c0 + jj);
                bad = bad || !(d > T(0));
                const T rinv = rsqrt_refined(d);
                const T l = (lane == c0 + jj) ? d * rinv : v[jj] * rinv;
                v[jj] = l;
                if constexpr (sizeof(T) == 8) {
                    // FP64: column jj of the diagonal block to every lane
                    // through LDS (one store, broadcast reads; LDS operations
                    // of a wavefront complete in order) -- fewer instructions
                    // than two v_readlane per entry
                    if (lane > c0 + jj && lane < c0 + 16) s_col[lane - c0] = l;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
                    for (int m = jj + 1; m < 16; ++m) v[m] -= l * s_col[m];
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                } else {
#pragma unroll
                    for (int m = jj + 1; m < 16; ++m) v[m] -= l * readlane(l, c0 + m);
                }
            }
            if (lane >= c0) {
#pragma unroll
                for (int m = 0; m < 16; ++m) a[lane][c0 + m] = v[m];
            }
            if (bad && lane == 0) s_bad = 1;
        }
        __syncthreads();
        // trailing blocks (rb, cb), cb <= rb, of the rows/columns after the panel
        const int nb = 3 - p;
        for (int bi = w; bi < nb * (nb + 1) / 2; bi += static_cast<int>(blockDim.x >> 6)) {
            int rb = 0;
            while ((rb + 1) * (rb + 2) / 2 <= bi) ++rb;
            const int cb = bi - rb * (rb + 1) / 2;
            const int R0 = c0 + 16 + 16 * rb, C0 = c0 + 16 + 16 * cb;
            acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
            for (int kk = 0; kk < 16; kk += 4)
                acc = mfma16(a[R0 + li][c0 + kk + lk], a[C0 + li][c0 + kk + lk], acc);
#pragma unroll
            for (int q = 0; q < 4; ++q) a[R0 + acc_row<T>(lk, q)][C0 + li] -= acc[q];
        }
        __syncthreads();
    }
    if (w < 4) {
        // inverse of diagonal block w by forward substitution, one lane per
        // column c of the inverse (L x = e_c; x_i = 0 for i < c), the entries
        // of L read from LDS as broadcasts; the lane groups of the wavefront
        // compute the same columns and the first writes them
        const int o = 16 * w, c = li;
        T rd[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) rd[i] = recip_refined(a[o + i][o + i]);
        T x[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            T s = (i == c) ? T(1) : T(0);
#pragma unroll
            for (int j = 0; j < i; ++j) s -= a[o + i][o + j] * x[j];
            x[i] = s * rd[i];
        }
        if (lk == 0) {
#pragma unroll
            for (int i = 0; i < 16; ++i) b[o + i][o + c] = (i >= c) ? x[i] : T(0);
        }
    }
    __syncthreads();
    for (int d = 1; d < 4; ++d) {
        if (w < 4 - d) {
            const int c = w, r = c + d;
            const int R0 = 16 * r, C0 = 16 * c;
            acc_t y = acc_t{0, 0, 0, 0};
            for (int mb = c; mb < r; ++mb) {
#pragma unroll
                for (int kk = 0; kk < 16; kk += 4)
                    y = mfma16(a[R0 + li][16 * mb + kk + lk], b[16 * mb + kk + lk][C0 + li], y);
            }
            // park Y transposed in the upper block (c, r) of b
#pragma unroll
            for (int q = 0; q < 4; ++q) b[C0 + li][R0 + acc_row<T>(lk, q)] = y[q];
            __builtin_amdgcn_wave_barrier();
            acc_t xr = acc_t{0, 0, 0, 0};
#pragma unroll
            for (int kk = 0; kk < 16; kk += 4)
                xr = mfma16(b[R0 + li][R0 + kk + lk], b[C0 + li][R0 + kk + lk], xr);
#pragma unroll
            for (int q = 0; q < 4; ++q) b[R0 + acc_row<T>(lk, q)][C0 + li] = -xr[q];
        }
        __syncthreads();
    }
    return s_bad == 0;
}

// Writes the factored diagonal tile k: D[k] = L_kk (row-major, upper zero),
// DinvT[k] = (L_kk^-1)^T; raises chol_fail on a bad pivot.
template <typename T>
__device__ __forceinline__ void store_diag_tile(const T (*a)[kTile + 2], const T (*b)[kTile + 2], bool ok, int k,
                                T* __restrict__ D, T* __restrict__ DinvT, TrialFlags* flags) {
    if (!ok && threadIdx.x == 0) atomicOr(&flags->chol_fail, 1);
    T* d = D + static_cast<size_t>(k) * kTile * kTile;
    T* it = DinvT + static_cast<size_t>(k) * kTile * kTile;
    for (int q = threadIdx.x; q < kTile * kTile; q += blockDim.x) {
        const int r = q >> 6, c = q & 63;
        d[q] = (c <= r) ? a[r][c] : T(0);
        it[q] = (r <= c) ? b[c][r] : T(0);
    }
}

// Batched use of the per-tile factor kernels: with tab set, workgroup w works
// on problem tab[w].x (its S at desc.soff with leading dimension desc.npad,
// its diagonal tiles and their inverses at desc.doff, its own flags) and
// takes tab[w].y as its tile index; problems switched off by active skip.
struct TileBatch {
    const int2* tab;
    const BatchDesc* desc;
    const int* active;
};

// Resolves the workgroup's problem; false when it has nothing to do.
template <typename T>
__device__ __forceinline__ bool tile_batch_bind(const TileBatch& tb, T* __restrict__& S, int& ld,
                                                T* __restrict__& D, T* __restrict__& DinvT, TrialFlags*& flags,
                                                int& unit) {
    unit = blockIdx.x;
    if (!tb.tab) return true;
    const int2 e = tb.tab[blockIdx.x];
    if (!tb.active[e.x]) return false;
    const BatchDesc d = tb.desc[e.x];
    S += d.soff;
    ld = d.npad;
    D += d.doff;
    DinvT = D + static_cast<size_t>(d.npad / kTile) * kTile * kTile;
    flags += e.x;
    unit = e.y;
    return true;
}

// Factor + invert diagonal tile k of S (the first panel; later diagonal tiles
// are factored by block 0 of the trailing update that finishes them).
template <typename T>
__global__ __launch_bounds__(256) void k_potrf_tile(T* __restrict__ S, int ld, int k,
                                                   T* __restrict__ D, T* __restrict__ DinvT,
                                                   TrialFlags* flags, TileBatch tb) {
    __shared__ T a[kTile][kTile + 2];
    __shared__ T b[kTile][kTile + 2];
    int unit;
    if (!tile_batch_bind(tb, S, ld, D, DinvT, flags, unit)) return;
    if (any_fail(flags)) return;
    const T* Akk = S + static_cast<size_t>(k * kTile) * ld + k * kTile;
    for (int q = threadIdx.x; q < kTile * kTile; q += 256) a[q >> 6][q & 63] = Akk[static_cast<size_t>(q >> 6) * ld + (q & 63)];
    __syncthreads();
    const bool ok = tile_potrf_inv<T>(a, b);
    store_diag_tile<T>(a, b, ok, k, D, DinvT, flags);
}

// 256 threads copy a 64x64 tile into LDS, all loads in flight at once.
template <typename T>
__device__ __forceinline__ void tile_to_lds(T (*s)[kTile + 2], const T* __restrict__ src, int ld) {
    T tmp[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int q = threadIdx.x + 256 * u;
        tmp[u] = src[static_cast<size_t>(q >> 6) * ld + (q & 63)];
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int q = threadIdx.x + 256 * u;
        s[q >> 6][q & 63] = tmp[u];
    }
}

// ---- explicit inverse X = L^-1 formed alongside the factor ----------------
// Row k of X needs row k of L (finished by TRSM k-1) and Linv_kk (factored by
// the lookahead of update launch k-1), so it is formed by extra workgroups of
// update launch k, which otherwise waits on its block 0:
//   X_kk = Linv_kk,   X_kj = -Linv_kk (R_kj + L_{k,k-1} X_{k-1,j})   (j < k),
//   R_kj = sum_{j <= m <= k-2} L_km X_mj,
// with R accumulated one term per launch: launch k also forms
// R_ij += L_{i,k-1} X_{k-1,j} for i > k, j < k (stored, not added, for the
// first term m = j = k-1).  Within a launch these roles read only row k-1 of
// X, row k of R and L tiles that no workgroup of the launch writes.  X layout:
// lower tile (i, j) X_ij, upper tile (j, i) X_ij^T, diagonal tile X_ii below
// and X_ii^T above its diagonal.  R lives in a separate n_pad^2 buffer.

// Roles of launch k: row k of X (k + 1 tiles), then the R tiles.
__host__ __device__ constexpr int inv_roles(int k, int nblk) { return (k + 1) + (nblk - 1 - k) * k; }

// acc = A B (or acc += A B with add set) for two 64x64 tiles staged in LDS.
template <typename T>
__device__ __forceinline__ void acc_lds_lds(typename Acc4<T>::type (&acc)[2][2], const T (*sA)[kTile + 2],
                                            const T (*sB)[kTile + 2], bool add = false) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    if (!add) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c) acc[r][c] = typename Acc4<T>::type{0, 0, 0, 0};
    }
#pragma unroll 4
    for (int kk = 0; kk < kTile; kk += 4) {
        const T a0 = sA[wr + li][kk + lk], a1 = sA[wr + 16 + li][kk + lk];
        const T b0 = sB[kk + lk][wc + li], b1 = sB[kk + lk][wc + 16 + li];
        acc[0][0] = mfma16(a0, b0, acc[0][0]);
        acc[0][1] = mfma16(a0, b1, acc[0][1]);
        acc[1][0] = mfma16(a1, b0, acc[1][0]);
        acc[1][1] = mfma16(a1, b1, acc[1][1]);
    }
}

// Role r (0 <= r < inv_roles(k, nblk)) of launch k, 256 threads; sa, sb are
// two LDS tiles of the calling kernel.
// kCoh: the one-launch factor's form -- R read and X / R written with
// agent-scope atomic accesses (other workgroups of the same launch read them).
// Grouped factor (kg < k: the launch closes the group of tile columns kg ..
// k): R holds the terms m <= kg - 2 only, and a role takes the terms
// m = max(j, kg - 1) .. k - 1 itself -- X_kj = -Linv_kk (R_kj + sum_m L_km
// X_mj), R_ij += sum_m L_im X_mj -- so inside a group the launches form the
// rows of X without touching R, and the launch that closes the group adds the
// group's terms to R at once.  kg = k is the column-by-column form (one term).
template <typename T, bool kCoh = false>
__device__ void inv_role(int r, int k, int nblk, const T* __restrict__ S, int ld, const T* __restrict__ DinvT,
                         T* __restrict__ X, T* __restrict__ R, T (*sa)[kTile + 2], T (*sb)[kTile + 2],
                         int kg) {
    auto put = [](T* p, T v) {
        if constexpr (kCoh) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else *p = v;
    };
    auto get = [](const T* p) -> T {
        if constexpr (kCoh) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else return *p;
    };
    typedef typename Acc4<T>::type acc_t;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    const size_t tile = static_cast<size_t>(kTile) * kTile;
    if (r == k) {
        // X_kk from DinvT[k] = (Linv_kk)^T
        const T* it = DinvT + static_cast<size_t>(k) * tile;
        T* xt = X + static_cast<size_t>(k * kTile) * ld + k * kTile;
#pragma unroll 4
        for (int q = threadIdx.x; q < kTile * kTile; q += 256) {
            const int rr = q >> 6, c = q & 63;
            put(xt + static_cast<size_t>(rr) * ld + c, it[(rr >= c) ? c * kTile + rr : rr * kTile + c]);
        }
        return;
    }
    int i, j;
    if (r < k) {
        i = k;
        j = r;
    } else {
        const int b = r - (k + 1);
        i = k + 1 + b / k;
        j = b % k;
    }
    acc_t acc[2][2];
    if (kCoh || kg == k) {
        // column-by-column form (and the one-launch factor): the one term
        // L_{i,k-1} X_{k-1,j}, staged through LDS -- A from the transposed
        // copy in the upper tile (k-1, i), B = X_{k-1,j} (lower part only for
        // a diagonal tile of X)
        const T* A = S + static_cast<size_t>((k - 1) * kTile) * ld + i * kTile;
        const T* B = X + static_cast<size_t>((k - 1) * kTile) * ld + j * kTile;
        const bool bl = (j == k - 1);
#pragma unroll
        for (int h = 0; h < 16; h += 8) {
            T ta[8], tb[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int q = threadIdx.x + 256 * (h + u);
                ta[u] = A[static_cast<size_t>(q >> 6) * ld + (q & 63)];
                tb[u] = B[static_cast<size_t>(q >> 6) * ld + (q & 63)];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int q = threadIdx.x + 256 * (h + u);
                const int rr = q >> 6, c = q & 63;
                sa[c][rr] = ta[u];
                sb[rr][c] = (bl && c > rr) ? T(0) : tb[u];
            }
        }
        __syncthreads();
        acc_lds_lds<T>(acc, sa, sb);
    } else {
        // acc = sum over m of L_{i,m} X_{m,j}, the MFMA operands read
        // straight from memory: A = L_{i,m} from its transposed copy in the
        // upper tile (m, i), which every factor form writes (16 lanes read 16
        // consecutive values), B = X_{m,j} by rows (lower part only for a
        // diagonal tile of X)
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int c = 0; c < 2; ++c) acc[a][c] = acc_t{0, 0, 0, 0};
        const int mlo = j > kg - 1 ? j : kg - 1;
        for (int m = mlo; m < k; ++m) {
            const T* LT = S + static_cast<size_t>(m * kTile) * ld + i * kTile;
            const T* XB = X + static_cast<size_t>(m * kTile) * ld + j * kTile;
            const bool bl = (j == m);
#pragma unroll 4
            for (int kk = 0; kk < kTile; kk += 4) {
                const int row = kk + lk;
                const size_t ro = static_cast<size_t>(row) * ld;
                const T a0 = LT[ro + wr + li], a1 = LT[ro + wr + 16 + li];
                T b0 = XB[ro + wc + li], b1 = XB[ro + wc + 16 + li];
                if (bl) {
                    b0 = (wc + li > row) ? T(0) : b0;
                    b1 = (wc + 16 + li > row) ? T(0) : b1;
                }
                acc[0][0] = mfma16(a0, b0, acc[0][0]);
                acc[0][1] = mfma16(a0, b1, acc[0][1]);
                acc[1][0] = mfma16(a1, b0, acc[1][0]);
                acc[1][1] = mfma16(a1, b1, acc[1][1]);
            }
        }
    }
    T* Rij = R + static_cast<size_t>(i * kTile) * ld + j * kTile;
    if (j <= kg - 2) {
        T part[2][2][4];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    part[a][c][q] = get(Rij + static_cast<size_t>(wr + 16 * a + acc_row<T>(lk, q)) * ld + wc + 16 * c + li);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[a][c][q] += part[a][c][q];
    }
    if (i > k) {
        // R role: R_ij = acc
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    put(Rij + static_cast<size_t>(wr + 16 * a + acc_row<T>(lk, q)) * ld + wc + 16 * c + li, acc[a][c][q]);
        return;
    }
    // X row role: X_kj = -Linv_kk Z with Z = acc; Linv_kk (rows of DinvT[k]
    // transposed) into sa, Z into sb
    __syncthreads();  // every wave is done reading sa and sb
    {
        const T* it = DinvT + static_cast<size_t>(k) * tile;  // it[c][rr] = Linv_kk[rr][c]
#pragma unroll
        for (int h = 0; h < 16; h += 8) {
            T tl[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) tl[u] = it[threadIdx.x + 256 * (h + u)];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int q = threadIdx.x + 256 * (h + u);
                sa[q & 63][q >> 6] = tl[u];
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) sb[wr + 16 * a + acc_row<T>(lk, q)][wc + 16 * c + li] = acc[a][c][q];
    __syncthreads();
    acc_lds_lds<T>(acc, sa, sb);
    __syncthreads();  // every wave is done reading sa
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) sa[wr + 16 * a + acc_row<T>(lk, q)][wc + 16 * c + li] = -acc[a][c][q];
    __syncthreads();
    T* low = X + static_cast<size_t>(k * kTile) * ld + j * kTile;
    T* up = X + static_cast<size_t>(j * kTile) * ld + k * kTile;
#pragma unroll 4
    for (int u = 0; u < 16; ++u) {
        const int q = threadIdx.x + 256 * u;
        const int rr = q >> 6, c = q & 63;
        put(low + static_cast<size_t>(rr) * ld + c, sa[rr][c]);
        put(up + static_cast<size_t>(rr) * ld + c, sa[c][rr]);
    }
}

// Row nblk-1 of X, after the last trailing update (its Linv comes from that
// update's lookahead).
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_inv_last_row(const T* __restrict__ S, int ld, int nblk,
                                                         const T* __restrict__ DinvT, T* __restrict__ X,
                                                         T* __restrict__ R, const TrialFlags* flags) {
    __shared__ T sa[kTile][kTile + 2];
    __shared__ T sb[kTile][kTile + 2];
    if (any_fail(flags)) return;
    inv_role<T>(blockIdx.x, nblk - 1, nblk, S, ld, DinvT, X, R, sa, sb, nblk - 1);
}

// L_ik = A_ik (L_kk^-1)^T for tile rows i > k: an MFMA GEMM per tile, A_ik
// staged in LDS, the B operand read straight from DinvT (row kk of DinvT is
// column kk of L_kk^-1, so a 16-lane group reads 128 contiguous bytes).  The
// result is written twice: L_ik over the lower tile (the triangular solves
// read it) and L_ik^T into the unused upper tile (k, i), from which the
// trailing update reads both MFMA operands with coalesced loads.
template <typename T>
__global__ __launch_bounds__(256) void k_trsm_tiles(T* __restrict__ S, int ld, int k, T* __restrict__ D,
                                                    T* __restrict__ DinvT, TrialFlags* flags, TileBatch tb) {
    __shared__ T Ai[kTile][kTile + 1];
    int unit;
    if (!tile_batch_bind(tb, S, ld, D, DinvT, flags, unit)) return;
    const int i = k + 1 + unit;
    T* Aik = S + static_cast<size_t>(i * kTile) * ld + k * kTile;
    T* LTki = S + static_cast<size_t>(k * kTile) * ld + i * kTile;
    const T* inv = DinvT + static_cast<size_t>(k) * kTile * kTile;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    // FP32: all operand loads first (in bounds whatever the flags say), then
    // the failure check, so the latencies overlap; FP64 keeps the inverse
    // fragments for after the barrier (the extra registers cost it more)
    constexpr bool kEarly = sizeof(T) == 4;
    T bv[16][2];
    auto load_inv = [&]() {
#pragma unroll
        for (int kk = 0; kk < kTile; kk += 4) {
            bv[kk >> 2][0] = inv[(kk + lk) * kTile + wc + li];
            bv[kk >> 2][1] = inv[(kk + lk) * kTile + wc + 16 + li];
        }
    };
    {
        if (!kEarly && any_fail(flags)) return;
        T tmp[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = threadIdx.x + 256 * u;
            tmp[u] = Aik[static_cast<size_t>(q >> 6) * ld + (q & 63)];
        }
        if constexpr (kEarly) {
            load_inv();
            if (any_fail(flags)) return;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = threadIdx.x + 256 * u;
            Ai[q >> 6][q & 63] = tmp[u];
        }
    }
    __syncthreads();
    if constexpr (!kEarly) load_inv();
    typename Acc4<T>::type acc[2][2];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[r][c] = typename Acc4<T>::type{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < kTile; kk += 4) {
        const T a0 = Ai[wr + li][kk + lk], a1 = Ai[wr + 16 + li][kk + lk];
        const T b0 = bv[kk >> 2][0], b1 = bv[kk >> 2][1];
        acc[0][0] = mfma16(a0, b0, acc[0][0]);
        acc[0][1] = mfma16(a0, b1, acc[0][1]);
        acc[1][0] = mfma16(a1, b0, acc[1][0]);
        acc[1][1] = mfma16(a1, b1, acc[1][1]);
    }
    __syncthreads();  // every wave has read Ai before it is overwritten
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) Ai[wr + 16 * r + acc_row<T>(lk, q)][wc + 16 * c + li] = acc[r][c][q];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int q = threadIdx.x + 256 * u;
        const int r = q >> 6, c = q & 63;
        Aik[static_cast<size_t>(r) * ld + c] = Ai[r][c];
        LTki[static_cast<size_t>(r) * ld + c] = Ai[c][r];
    }
}

// Trailing update A_ij -= L_ik L_jk^T for k < j <= i, one workgroup per tile,
// four wavefronts of 2x2 MFMA 16x16 blocks; both operands come straight from
// the transposed panel tiles L_ik^T, L_jk^T (upper tiles (k, i), (k, j)).
// Block 0 owns tile (k+1, k+1), which this update completes: it then factors
// and inverts it (lookahead), so the next panel needs no separate factor
// launch.  With X set (FP32), the workgroups from nup on form row k of
// X = L^-1 (inv_role).  With fused set (FP32, the late columns), there was
// no TRSM launch for column k: each workgroup forms the L_ik, L_jk it needs
// from the raw tiles below the diagonal and Linv_kk, and the workgroup of
// column k+1 writes L_ik^T to the upper tile (k, i), the only copy later
// launches read (the lower tiles keep A).
// Grouped form (large matrices, fused off): the update applies the panel
// columns kc0 .. k together, A_ij -= sum_c L_ic L_jc^T, and covers the tiles
// (i, j), k < j <= i, of the first jw + 1 trailing columns only (jw >= the
// trailing width: all of them), so a group of panel columns reads and writes
// the trailing matrix once instead of once per column.  kg: the first column
// of the launch's group, for the inverse roles (inv_role; kg = k otherwise).
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2))) void k_chol_update(
    T* __restrict__ S, int ld, int k, T* __restrict__ D, T* __restrict__ DinvT, TrialFlags* flags, TileBatch tb,
    int nup, int nblk, T* __restrict__ X, T* __restrict__ R, int fused, int kc0, int jw, int kg) {
    __shared__ T Ai[kTile][kTile + 2];
    __shared__ T Aj[kTile][kTile + 2];
    int t;
    if (!tile_batch_bind(tb, S, ld, D, DinvT, flags, t)) return;
    if constexpr (sizeof(T) == 4) {
        if (X && t >= nup) {
            // workgroups past the update tiles form row k of X = L^-1 (see inv_role)
            if (any_fail(flags)) return;
            inv_role<T>(t - nup, k, nblk, S, ld, DinvT, X, R, Ai, Aj, kg);
            return;
        }
    }
    // the failure check below comes after this workgroup's first loads are
    // issued (they are in bounds either way), so its latency overlaps them
    // tile index -> (ii, jj), jj <= ii, over the trailing lower triangle
    // (rows of the triangle of the first jw + 1 columns, then rows of jw + 1
    // tiles)
    int ii, jj;
    if (t < (jw + 1) * (jw + 2) / 2) {
        ii = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
        while (ii * (ii + 1) / 2 > t) --ii;
        while ((ii + 1) * (ii + 2) / 2 <= t) ++ii;
        jj = t - ii * (ii + 1) / 2;
    } else {
        const int t2 = t - (jw + 1) * (jw + 2) / 2;
        ii = jw + 1 + t2 / (jw + 1);
        jj = t2 % (jw + 1);
    }
    const int i = k + 1 + ii, j = k + 1 + jj;
    const T* LTi = S + static_cast<size_t>(k * kTile) * ld + i * kTile;
    const T* LTj = S + static_cast<size_t>(k * kTile) * ld + j * kTile;
    T* Aij = S + static_cast<size_t>(i * kTile) * ld + j * kTile;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    // block 0 reads its tile (k+1, k+1) up front: no load of this launch
    // writes it, and its latency then overlaps the operand loads
    // (FP32 only: in FP64 the sixteen extra registers cost more than the
    // overlap gains)
    constexpr bool kPrefetch = sizeof(T) == 4;
    T pd[16];
    if (kPrefetch && t == 0) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = threadIdx.x + 256 * u;
            pd[u] = Aij[static_cast<size_t>(q >> 6) * ld + (q & 63)];
        }
    }
    typename Acc4<T>::type acc[2][2];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[r][c] = typename Acc4<T>::type{0, 0, 0, 0};
    if (fused) {
        // L_ik = A_ik Linv_kk^T and L_jk likewise, into Ai and Aj; every
        // operand load is issued before the first barrier
        const T* inv = DinvT + static_cast<size_t>(k) * kTile * kTile;
        const bool two = (j != i);
        T bv[16][2];
#pragma unroll
        for (int kk = 0; kk < kTile; kk += 4) {
            bv[kk >> 2][0] = inv[(kk + lk) * kTile + wc + li];
            bv[kk >> 2][1] = inv[(kk + lk) * kTile + wc + 16 + li];
        }
        {
            const T* srci = S + static_cast<size_t>(i * kTile) * ld + k * kTile;
            const T* srcj = S + static_cast<size_t>(j * kTile) * ld + k * kTile;
            T ta[16], tb[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int q = threadIdx.x + 256 * u;
                ta[u] = srci[static_cast<size_t>(q >> 6) * ld + (q & 63)];
                if (two) tb[u] = srcj[static_cast<size_t>(q >> 6) * ld + (q & 63)];
            }
            if (any_fail(flags)) return;
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int q = threadIdx.x + 256 * u;
                Ai[q >> 6][q & 63] = ta[u];
                if (two) Aj[q >> 6][q & 63] = tb[u];
            }
        }
        __syncthreads();
        typename Acc4<T>::type acc2[2][2];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c) acc2[r][c] = typename Acc4<T>::type{0, 0, 0, 0};
#pragma unroll
        for (int kk = 0; kk < kTile; kk += 4) {
            const T b0 = bv[kk >> 2][0], b1 = bv[kk >> 2][1];
            const T a0 = Ai[wr + li][kk + lk], a1 = Ai[wr + 16 + li][kk + lk];
            acc[0][0] = mfma16(a0, b0, acc[0][0]);
            acc[0][1] = mfma16(a0, b1, acc[0][1]);
            acc[1][0] = mfma16(a1, b0, acc[1][0]);
            acc[1][1] = mfma16(a1, b1, acc[1][1]);
            if (two) {
                const T c0 = Aj[wr + li][kk + lk], c1 = Aj[wr + 16 + li][kk + lk];
                acc2[0][0] = mfma16(c0, b0, acc2[0][0]);
                acc2[0][1] = mfma16(c0, b1, acc2[0][1]);
                acc2[1][0] = mfma16(c1, b0, acc2[1][0]);
                acc2[1][1] = mfma16(c1, b1, acc2[1][1]);
            }
        }
        __syncthreads();  // every wave has read the raw tiles
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = wr + 16 * r + acc_row<T>(lk, q), col = wc + 16 * c + li;
                    Ai[row][col] = acc[r][c][q];
                    if (two) Aj[row][col] = acc2[r][c][q];
                    acc[r][c][q] = T(0);
                }
        __syncthreads();
        if (jj == 0) {
            T* up = S + static_cast<size_t>(k * kTile) * ld + i * kTile;  // L_ik^T
#pragma unroll 4
            for (int u = 0; u < 16; ++u) {
                const int q = threadIdx.x + 256 * u;
                const int r = q >> 6, c = q & 63;
                up[static_cast<size_t>(r) * ld + c] = Ai[c][r];
            }
        }
        // acc = L_ik L_jk^T from LDS
        const T (*Bj)[kTile + 2] = two ? Aj : Ai;
#pragma unroll 4
        for (int kk = 0; kk < kTile; kk += 4) {
            const T a0 = Ai[wr + li][kk + lk], a1 = Ai[wr + 16 + li][kk + lk];
            const T b0 = Bj[wc + li][kk + lk], b1 = Bj[wc + 16 + li][kk + lk];
            acc[0][0] = mfma16(a0, b0, acc[0][0]);
            acc[0][1] = mfma16(a0, b1, acc[0][1]);
            acc[1][0] = mfma16(a1, b0, acc[1][0]);
            acc[1][1] = mfma16(a1, b1, acc[1][1]);
        }
        __syncthreads();  // every wave is done with Ai / Aj before the lookahead reuses them
    } else {
        if (any_fail(flags)) return;
        for (int c = kc0; c <= k; ++c) {
            // panel column c: its transposed tiles L_ic^T, L_jc^T sit in the upper tiles (c, i), (c, j)
            const T* Ti = LTi - static_cast<size_t>((k - c) * kTile) * ld;
            const T* Tj = LTj - static_cast<size_t>((k - c) * kTile) * ld;
#pragma unroll
            for (int kk = 0; kk < kTile; kk += 4) {
                const size_t ro = static_cast<size_t>(kk + lk) * ld;
                const T a0 = Ti[ro + wr + li], a1 = Ti[ro + wr + 16 + li];
                const T b0 = Tj[ro + wc + li], b1 = Tj[ro + wc + 16 + li];
                acc[0][0] = mfma16(a0, b0, acc[0][0]);
                acc[0][1] = mfma16(a0, b1, acc[0][1]);
                acc[1][0] = mfma16(a1, b0, acc[1][0]);
                acc[1][1] = mfma16(a1, b1, acc[1][1]);
            }
        }
    }
    if (t == 0) {
        // lookahead: finish tile (k+1, k+1) in LDS (the tile, prefetched via
        // Aj in FP32, minus the product) and factor it
        if constexpr (kPrefetch) {
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int q = threadIdx.x + 256 * u;
                Aj[q >> 6][q & 63] = pd[u];
            }
            __syncthreads();
        }
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int row = wr + 16 * r + acc_row<T>(lk, q), col = wc + 16 * c + li;
                    Ai[row][col] = (kPrefetch ? Aj[row][col] : Aij[static_cast<size_t>(row) * ld + col]) - acc[r][c][q];
                }
        __syncthreads();
        const bool ok = tile_potrf_inv<T>(Ai, Aj);
        store_diag_tile<T>(Ai, Aj, ok, k + 1, D, DinvT, flags);
        return;
    }
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

// ---- batched small systems: one workgroup per problem ----------------------

// Dense factor and solve of every active problem's S (FP64), one workgroup of
// kBatchThreads threads (8 wavefronts) per problem, tile by tile: diagonal tile factored and inverted
// in LDS (tile_potrf_inv), TRSM of the tiles below as MFMA products with the
// stored inverse (L_ik written below and L_ik^T above the diagonal), trailing
// update from the transposed tiles, then forward and backward substitution
// with the diagonal inverses.  dc[9 c0 + r] receives the solution; a bad
// pivot sets flags[problem].chol_fail.  npad <= 1024.
constexpr int kBatchThreads = 512;
constexpr int kBatchWaves = kBatchThreads / 64;
// work unit of the TRSM / update: a strip of 16 kRB rows of one 64x64 tile
constexpr int kRB = kBatchWaves >= 16 ? 1 : 2;
constexpr int kStrips = 4 / kRB;

// kFactor = false: S already holds the factor (launch_batch_cholesky), only
// the substitutions run.
template <bool kFactor>
__global__ __launch_bounds__(kBatchThreads) void k_batch_chol_solve(const BatchDesc* __restrict__ desc,
                                                          double* __restrict__ S,
                                                          double* __restrict__ D,
                                                          const double* __restrict__ rhs,
                                                          double* __restrict__ dc, TrialFlags* flags,
                                                          const int* __restrict__ active) {
    typedef Acc4<double>::type acc_t;
    __shared__ double sa[kTile][kTile + 2];
    __shared__ double sb[kTile][kTile + 2];
    __shared__ double vec[1024];
    __shared__ double part[kBatchThreads];
    __shared__ double tmp[kTile];
    const int b = blockIdx.x;
    if (!active[b] || (kFactor ? flags[b].nonspd != 0 : any_fail(flags + b))) return;
    const BatchDesc d = desc[b];
    double* A = S + d.soff;
    const int ld = d.npad, nblk = d.npad / kTile;
    double* Lkk = D + d.doff;
    double* DinvT = Lkk + static_cast<size_t>(nblk) * kTile * kTile;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    for (int k = 0; kFactor && k < nblk; ++k) {
        const double* Akk = A + static_cast<size_t>(k * kTile) * ld + k * kTile;
        for (int q = tid; q < kTile * kTile; q += kBatchThreads)
            sa[q >> 6][q & 63] = Akk[static_cast<size_t>(q >> 6) * ld + (q & 63)];
        __syncthreads();
        const bool ok = tile_potrf_inv<double>(sa, sb);
        store_diag_tile<double>(sa, sb, ok, k, Lkk, DinvT, flags + b);
        __syncthreads();
        if (!ok) return;
        const double* inv = DinvT + static_cast<size_t>(k) * kTile * kTile;
        // TRSM: wavefront w takes every kBatchWaves-th row strip (16 kRB rows)
        // of the tiles below: L_ik = A_ik Linv^T (kRB x 4 MFMA blocks), written
        // below the diagonal and, transposed, into the upper tile (k, i)
        for (int h = w; h < kStrips * (nblk - k - 1); h += kBatchWaves) {
            const int i = k + 1 + h / kStrips, r0 = 16 * kRB * (h % kStrips);
            double* Aik = A + static_cast<size_t>(i * kTile) * ld + k * kTile;
            double* LTki = A + static_cast<size_t>(k * kTile) * ld + i * kTile;
            acc_t acc[kRB][4];
#pragma unroll
            for (int r = 0; r < kRB; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] = acc_t{0, 0, 0, 0};
#pragma unroll 2
            for (int kk = 0; kk < kTile; kk += 4) {
                double av[kRB], bv[4];
#pragma unroll
                for (int r = 0; r < kRB; ++r) av[r] = Aik[static_cast<size_t>(r0 + 16 * r + li) * ld + kk + lk];
#pragma unroll
                for (int c = 0; c < 4; ++c) bv[c] = inv[(kk + lk) * kTile + 16 * c + li];
#pragma unroll
                for (int r = 0; r < kRB; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c) acc[r][c] = mfma16(av[r], bv[c], acc[r][c]);
            }
#pragma unroll
            for (int r = 0; r < kRB; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int row = r0 + 16 * r + acc_row<double>(lk, q), col = 16 * c + li;
                        Aik[static_cast<size_t>(row) * ld + col] = acc[r][c][q];
                        LTki[static_cast<size_t>(col) * ld + row] = acc[r][c][q];
                    }
        }
        __syncthreads();
        // trailing update: wavefront w takes every kBatchWaves-th row strip of
        // the tiles (i, j), k < j <= i; operands from the transposed tiles
        {
            const int m = nblk - k - 1;
            for (int h = w; h < kStrips * (m * (m + 1) / 2); h += kBatchWaves) {
                const int t = h / kStrips, r0 = 16 * kRB * (h % kStrips);
                int ii = 0;
                while ((ii + 1) * (ii + 2) / 2 <= t) ++ii;
                const int jj = t - ii * (ii + 1) / 2;
                const int i = k + 1 + ii, j = k + 1 + jj;
                const double* LTi = A + static_cast<size_t>(k * kTile) * ld + i * kTile;
                const double* LTj = A + static_cast<size_t>(k * kTile) * ld + j * kTile;
                acc_t acc[kRB][4];
#pragma unroll
                for (int r = 0; r < kRB; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c) acc[r][c] = acc_t{0, 0, 0, 0};
#pragma unroll 2
                for (int kk = 0; kk < kTile; kk += 4) {
                    const size_t ro = static_cast<size_t>(kk + lk) * ld;
                    double av[kRB], bv[4];
#pragma unroll
                    for (int r = 0; r < kRB; ++r) av[r] = LTi[ro + r0 + 16 * r + li];
#pragma unroll
                    for (int c = 0; c < 4; ++c) bv[c] = LTj[ro + 16 * c + li];
#pragma unroll
                    for (int r = 0; r < kRB; ++r)
#pragma unroll
                        for (int c = 0; c < 4; ++c) acc[r][c] = mfma16(av[r], bv[c], acc[r][c]);
                }
                double* Aij = A + static_cast<size_t>(i * kTile) * ld + j * kTile;
#pragma unroll
                for (int r = 0; r < kRB; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c)
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const int row = r0 + 16 * r + acc_row<double>(lk, q), col = 16 * c + li;
                            Aij[static_cast<size_t>(row) * ld + col] -= acc[r][c][q];
                        }
            }
        }
        __syncthreads();
    }
    // Triangular solves, one sweep per tile row: thread (r, g) = (lane, wave)
    // sums row r over its eighth of every tile's columns, the eight partial
    // sums are combined in a fixed order.
    constexpr int kSlice = kTile / kBatchWaves;
    const int r = lane, c0 = kSlice * w;
    auto combine = [&](int row) {
        double t = 0.0;
#pragma unroll
        for (int g = 0; g < kBatchWaves; ++g) t += part[g * kTile + row];
        return t;
    };
    // forward: y_i = Linv_ii (b_i - sum_{j<i} L_ij y_j); vec holds y
    const double* bb = rhs + 9 * static_cast<size_t>(d.c0);
    for (int i = 0; i < nblk; ++i) {
        double s = 0.0;
        const double* row = A + static_cast<size_t>(i * kTile + r) * ld + c0;
#pragma unroll 4
        for (int j = 0; j < i; ++j) {
#pragma unroll
            for (int c = 0; c < kSlice; ++c) s += row[j * kTile + c] * vec[j * kTile + c0 + c];
        }
        part[w * kTile + r] = s;
        __syncthreads();
        if (tid < kTile) {
            const int gr = i * kTile + tid;
            tmp[tid] = (gr < d.n ? bb[gr] : 0.0) - combine(tid);
        }
        __syncthreads();
        // Linv s: DinvT[c][r] = Linv[r][c]
        const double* it = DinvT + static_cast<size_t>(i) * kTile * kTile;
        s = 0.0;
#pragma unroll
        for (int c = 0; c < kSlice; ++c) s += it[(c0 + c) * kTile + r] * tmp[c0 + c];
        part[w * kTile + r] = s;
        __syncthreads();
        if (tid < kTile) vec[i * kTile + tid] = combine(tid);
        __syncthreads();
    }
    // backward: x_i = Linv_ii^T (y_i - sum_{j>i} L_ji^T x_j), in place over vec;
    // L_ji^T is stored in the upper tile (i, j)
    for (int i = nblk - 1; i >= 0; --i) {
        double s = 0.0;
        const double* row = A + static_cast<size_t>(i * kTile + r) * ld + c0;
#pragma unroll 4
        for (int j = i + 1; j < nblk; ++j) {
#pragma unroll
            for (int c = 0; c < kSlice; ++c) s += row[j * kTile + c] * vec[j * kTile + c0 + c];
        }
        part[w * kTile + r] = s;
        __syncthreads();
        if (tid < kTile) tmp[tid] = vec[i * kTile + tid] - combine(tid);
        __syncthreads();
        // Linv^T s: DinvT[r][c] = Linv[c][r]
        const double* it = DinvT + static_cast<size_t>(i) * kTile * kTile;
        s = 0.0;
#pragma unroll
        for (int c = 0; c < kSlice; ++c) s += it[r * kTile + c0 + c] * tmp[c0 + c];
        part[w * kTile + r] = s;
        __syncthreads();
        if (tid < kTile) vec[i * kTile + tid] = combine(tid);
        __syncthreads();
    }
    for (int q = tid; q < d.n; q += kBatchThreads) dc[9 * static_cast<size_t>(d.c0) + q] = vec[q];
}

// Triangular solves: the results are their own completion flags.  Before a
// solve its output vector is filled with kUnset, a signalling NaN -- an
// arithmetic result is never one (IEEE operations return quiet NaNs) -- and
// each tile row stores its 64 results with agent-scope stores (written through
// to the coherence point).  A workgroup that needs a finished row polls the
// very values it uses with agent-scope loads until none is kUnset: no separate
// flag, no release / acquire fences, one memory round trip per row.  Waits are
// bounded: no progress for kWaitTicks of wall clock, or the error flag, and
// the row gives up (solve_err; never expected -- the tickets hand out rows in
// dependency order, so every awaited row has a running owner).
constexpr unsigned long long kUnset = 0x7FF0DEADBEEF0001ULL;

__device__ __forceinline__ double ld_result(const double* p) {
    return __longlong_as_double(static_cast<long long>(
        __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
}
__device__ __forceinline__ void st_result(double* p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(__double_as_longlong(v)),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool is_unset(double v) {
    return static_cast<unsigned long long>(__double_as_longlong(v)) == kUnset;
}
// Polls v = *p (one element per lane; lanes with use unset read nothing)
// until no lane of the wavefront holds kUnset; false when it gives up.
__device__ __forceinline__ bool wait_result(const double* p, bool use, double& v, const int* err) {
    const long long t0 = wall_clock64();
    for (;;) {
        v = use ? ld_result(p) : 0.0;
        if (__ballot(use && is_unset(v)) == 0) return true;
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
            wall_clock64() - t0 > kWaitTicks)
            return false;
        __builtin_amdgcn_s_sleep(1);
    }
}

// Forward solve L y = b (y filled with kUnset).  Workgroup with ticket t owns
// tile row i = t; lane = column of the L tiles, wavefront w = rows 16w..16w+15.
// DinvT[i] holds (L_ii^-1)^T.
// first (optional, envelope factor): L_ij = 0 for j < first[i], so row i
// neither reads nor waits for those tiles.
template <typename T>
__global__ __launch_bounds__(256) void k_trsv_fwd(const T* __restrict__ L, int ld, int nblk,
                                                  const T* __restrict__ DinvT,
                                                  const double* __restrict__ b, double* y,
                                                  int* ticket, TrialFlags* flags, const int* __restrict__ first,
                                                  const double* skip) {
    __shared__ int s_blk, s_skip;
    __shared__ double s_part[4][16];
    if (skip && *skip != 0.0) return;  // a refinement step after the stop
    __shared__ double s_rhs[kTile];
    __shared__ T s_inv[kTile][kTile + 1];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    if (tid == 0) {
        s_blk = atomicAdd(ticket, 1);
        s_skip = (any_fail(flags) || flags->solve_err) ? 1 : 0;
    }
    __syncthreads();
    if (s_skip) return;
    const int i = s_blk;
    const int lo = first ? first[i] : 0;
    {
        const T* inv = DinvT + static_cast<size_t>(i) * kTile * kTile;
        for (int q = tid; q < kTile * kTile; q += 256) s_inv[q >> 6][q & 63] = inv[q];
    }
    double acc[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0;
    bool gave_up = false;
    for (int j = lo; j < i; ++j) {
        // the tile's loads do not depend on y_j: issued before the wait
        const T* Lr = L + static_cast<size_t>(i * kTile + 16 * w) * ld + j * kTile + lane;
        T lv[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) lv[r] = Lr[static_cast<size_t>(r) * ld];
        double yj;
        if (!wait_result(y + j * kTile + lane, true, yj, &flags->solve_err)) {
            gave_up = true;
            break;
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] += static_cast<double>(lv[r]) * yj;
    }
    // after a failed wait the row is abandoned: its results stay unset (the
    // waiters on it see solve_err and give up)
    if (gave_up && lane == 0) {
        atomicOr(&flags->solve_err, 1);
        s_skip = 1;
    }
    __syncthreads();
    if (s_skip) return;
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
        // y_r = sum_c Linv[r][c] s_c = sum_c DinvT[c][r] s_c
        double v = 0.0;
#pragma unroll 16
        for (int c = 0; c < kTile; ++c) v += static_cast<double>(s_inv[c][lane]) * s_rhs[c];
        st_result(y + i * kTile + lane, v);
    }
}

// Backward solve L^T x = y (x filled with kUnset).  Ticket t owns tile row
// i = nblk - 1 - t.
template <typename T>
__global__ __launch_bounds__(256) void k_trsv_bwd(const T* __restrict__ L, int ld, int nblk,
                                                  const T* __restrict__ DinvT,
                                                  const double* __restrict__ y, double* x,
                                                  int* ticket, TrialFlags* flags, const int* __restrict__ first,
                                                  const double* skip) {
    __shared__ int s_blk, s_skip;
    __shared__ double s_part[4][kTile];
    if (skip && *skip != 0.0) return;
    __shared__ double s_rhs[kTile];
    __shared__ T s_inv[kTile][kTile + 1];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    if (tid == 0) {
        s_blk = nblk - 1 - atomicAdd(ticket, 1);
        s_skip = (any_fail(flags) || flags->solve_err) ? 1 : 0;
    }
    __syncthreads();
    if (s_skip) return;
    const int i = s_blk;
    {
        const T* inv = DinvT + static_cast<size_t>(i) * kTile * kTile;
        for (int q = tid; q < kTile * kTile; q += 256) s_inv[q >> 6][q & 63] = inv[q];
    }
    double acc = 0.0;
    bool gave_up = false;
    for (int j = nblk - 1; j > i; --j) {
        if (first && first[j] > i) continue;  // L_ji = 0 (uniform across the workgroup)
        const T* Lc = L + static_cast<size_t>(j * kTile + 16 * w) * ld + i * kTile + lane;
        T lv[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) lv[r] = Lc[static_cast<size_t>(r) * ld];
        // x_j rows 16w .. 16w+15: lane r < 16 polls row r, then every lane
        // takes the sixteen values
        double xr;
        if (!wait_result(x + j * kTile + 16 * w + (lane & 15), lane < 16, xr, &flags->solve_err)) {
            gave_up = true;
            break;
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) acc += static_cast<double>(lv[r]) * readlane(xr, r);
    }
    if (gave_up && lane == 0) {
        atomicOr(&flags->solve_err, 1);
        s_skip = 1;
    }
    __syncthreads();
    if (s_skip) return;
    s_part[w][lane] = acc;
    __syncthreads();
    if (tid < kTile)
        s_rhs[tid] = y[i * kTile + tid] -
                     ((s_part[0][tid] + s_part[1][tid]) + (s_part[2][tid] + s_part[3][tid]));
    __syncthreads();
    if (w == 0) {
        // x_r = sum_c Linv[c][r] s_c = sum_c DinvT[r][c] s_c
        double v = 0.0;
#pragma unroll 16
        for (int c = 0; c < kTile; ++c) v += static_cast<double>(s_inv[lane][c]) * s_rhs[c];
        st_result(x + i * kTile + lane, v);
    }
}

// Symmetric product y = A x from the lower triangle of A (FP32, row-major,
// FP64 accumulation), no atomics.  Pass 1: one workgroup per lower tile (i, j)
// writes the tile's row partial A_ij x_j (block i) and column partial
// A_ij^T x_i (block j; strictly lower part for diagonal tiles).  Pass 2: one
// workgroup per 64-block sums its partials in a fixed order.
__global__ __launch_bounds__(256) void k_symv_tiles(const float* __restrict__ A, int ld,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ rp, double* __restrict__ cp,
                                                    const double* done) {
    __shared__ float s_a[kTile][kTile + 1];
    __shared__ double s_xi[kTile], s_xj[kTile];
    __shared__ double s_col[4][kTile];
    if (pcg_skip(done)) return;
    const int t = blockIdx.x, tid = threadIdx.x;
    int i = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while (i * (i + 1) / 2 > t) --i;
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    const int j = t - i * (i + 1) / 2;
    const bool diag = (i == j);
    {
        // the tile into LDS with coalesced loads, all in flight at once
        const float* src = A + static_cast<size_t>(i * kTile) * ld + j * kTile;
        float tmp[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = tid + 256 * u;
            tmp[u] = src[static_cast<size_t>(q >> 6) * ld + (q & 63)];
        }
        if (tid < kTile) s_xj[tid] = x[j * kTile + tid];
        else if (tid < 2 * kTile) s_xi[tid - kTile] = x[i * kTile + tid - kTile];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = tid + 256 * u;
            s_a[q >> 6][q & 63] = tmp[u];
        }
    }
    __syncthreads();
    // row partials: four threads per row, sixteen columns each, combined by
    // two lane exchanges (the four are adjacent lanes)
    {
        const int r = tid >> 2, c0 = 16 * (tid & 3);
        double rs = 0.0;
#pragma unroll
        for (int c = c0; c < c0 + 16; ++c)
            rs += (!diag || c <= r) ? static_cast<double>(s_a[r][c]) * s_xj[c] : 0.0;
        rs += __shfl_xor(rs, 1, kWave);
        rs += __shfl_xor(rs, 2, kWave);
        if ((tid & 3) == 0) rp[static_cast<size_t>(t) * kTile + r] = rs;
    }
    // column partials (strictly lower part of a diagonal tile): one thread per
    // column and sixteen rows, the four row groups combined in order
    {
        const int c = tid & 63, r0 = 16 * (tid >> 6);
        double cs = 0.0;
#pragma unroll
        for (int r = r0; r < r0 + 16; ++r)
            cs += (!diag || c < r) ? static_cast<double>(s_a[r][c]) * s_xi[r] : 0.0;
        s_col[tid >> 6][c] = cs;
    }
    __syncthreads();
    if (tid < kTile)
        cp[static_cast<size_t>(t) * kTile + tid] = (s_col[0][tid] + s_col[1][tid]) + (s_col[2][tid] + s_col[3][tid]);
}

// y_b = sum_{j <= b} rp[(b, j)] + sum_{i > b} cp[(i, b)]: wavefront g sums
// the partials with index = g mod 4, the four sums are combined in order.
__global__ __launch_bounds__(256) void k_symv_reduce(int nblk, const double* __restrict__ rp,
                                                     const double* __restrict__ cp,
                                                     double* __restrict__ y, const double* done) {
    __shared__ double s_part[4][kTile];
    if (pcg_skip(done)) return;
    const int b = blockIdx.x, lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    double s = 0.0;
#pragma unroll 4
    for (int j = g; j <= b; j += 4) s += rp[static_cast<size_t>(b * (b + 1) / 2 + j) * kTile + lane];
#pragma unroll 4
    for (int i = b + g; i < nblk; i += 4) s += cp[static_cast<size_t>(i * (i + 1) / 2 + b) * kTile + lane];
    s_part[g][lane] = s;
    __syncthreads();
    if (g == 0) y[b * kTile + lane] = (s_part[0][lane] + s_part[1][lane]) + (s_part[2][lane] + s_part[3][lane]);
}

// Per-workgroup partials of a.b and c.d (FP64).
__global__ __launch_bounds__(256) void k_dot2(int n, const double* __restrict__ a,
                                              const double* __restrict__ b,
                                              const double* __restrict__ c,
                                              const double* __restrict__ d, double* __restrict__ blk,
                                              const double* done) {
    __shared__ double s0[4], s1[4];
    if (pcg_skip(done)) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    double u = 0.0, v = 0.0;
    if (i < n) {
        u = a[i] * b[i];
        v = c[i] * d[i];
    }
    u = wave_sum(u);
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) {
        s0[threadIdx.x >> 6] = u;
        s1[threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        blk[2 * blockIdx.x] = (s0[0] + s0[1]) + (s0[2] + s0[3]);
        blk[2 * blockIdx.x + 1] = (s1[0] + s1[1]) + (s1[2] + s1[3]);
    }
}

// Refinement scalars and the stopping test kept on the device, so the host
// enqueues several PCG steps per synchronisation:
// sc[0] = r.z, sc[1] = alpha, sc[2] = beta, sc[3] = stop (p.q or r.z not > 0),
// sc[4] = done, sc[5] = b.b, sc[6] = steps counted, sc[7] = r.r of the last
// counted step.
//   mode 0: s0 = r.z, s1 = b.b of the start; done at once when the trial has
//           failed, b.b is not > 0 or no step is allowed.
//   mode 1: s0 = p.q -> alpha = r.z / p.q (0 and stop when either is not > 0).
//   mode 2: s0 = r.z (new), s1 = r.r; unless stopped, beta = r.z (new) / r.z,
//           r.z is replaced and the step is counted; done on a stop
//           (uncounted), a non-finite r.r, r.r <= tol2 b.b or max_it steps.
// After done, modes 1 and 2 do nothing.  status gets the dots, the flags and
// the loop state (pcg_stop, pcg_done, pcg_it, pcg_rr).
__global__ __launch_bounds__(256) void k_pcg_reduce(int nb, const double* __restrict__ blk, int mode,
                                                    double* __restrict__ sc, PassStatus* status,
                                                    const TrialFlags* flags, double tol2, int max_it) {
    __shared__ double s0[256], s1[256];
    if (mode != 0 && sc[4] != 0.0) return;
    const int tid = threadIdx.x;
    double u = 0.0, v = 0.0;
    for (int q = tid; q < nb; q += 256) {
        u += blk[2 * q];
        v += blk[2 * q + 1];
    }
    s0[tid] = u;
    s1[tid] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) {
            s0[tid] += s0[tid + h];
            s1[tid] += s1[tid + h];
        }
        __syncthreads();
    }
    if (tid != 0) return;
    const double a = s0[0], c = s1[0];
    if (mode == 0) {
        const TrialFlags f = *flags;
        sc[0] = a;
        sc[3] = 0.0;
        sc[5] = c;
        sc[6] = 0.0;
        sc[7] = c;
        sc[4] = (f.nonspd || f.chol_fail || f.solve_err || !(c > 0.0) || max_it <= 0) ? 1.0 : 0.0;
    } else if (mode == 1) {
        const bool bad = !(a > 0.0) || !(sc[0] > 0.0);
        sc[1] = bad ? 0.0 : sc[0] / a;
        if (bad) sc[3] = 1.0;
        return;
    } else if (sc[3] != 0.0) {
        sc[4] = 1.0;  // stopped: the step is not counted
    } else {
        sc[2] = a / sc[0];
        sc[0] = a;
        const double it = sc[6] + 1.0;
        sc[6] = it;
        sc[7] = c;
        if (!isfinite(c) || c <= tol2 * sc[5] || it >= static_cast<double>(max_it)) sc[4] = 1.0;
    }
    status->dot[0] = a;
    status->dot[1] = c;
    status->flags = *flags;
    status->pcg_stop = sc[3] != 0.0 ? 1 : 0;
    status->pcg_done = sc[4] != 0.0 ? 1 : 0;
    status->pcg_it = static_cast<int>(sc[6]);
    status->pcg_rr = sc[7];
    status->pcg_bb = sc[5];
}

// x += alpha p ; r -= alpha q  (alpha = sc[1])
__global__ void k_pcg_update(int n, const double* __restrict__ sc, double* __restrict__ x,
                             const double* __restrict__ p, double* __restrict__ r,
                             const double* __restrict__ q) {
    if (sc[4] != 0.0) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const double alpha = sc[1];
    if (i < n) {
        x[i] += alpha * p[i];
        r[i] -= alpha * q[i];
    }
}

// p = z + beta p  (beta = sc[2])
__global__ void k_pcg_direction(int n, const double* __restrict__ sc, const double* __restrict__ z,
                                double* __restrict__ p) {
    if (sc[4] != 0.0) return;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const double beta = sc[2];
    if (i < n) p[i] = z[i] + beta * p[i];
}

__global__ void k_vec_copy(int n, const double* __restrict__ a, double* __restrict__ b) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = a[i];
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

// ---- one-launch tile Cholesky ----------------------------------------------
// Workgroup 0 runs the critical chain (per tile column k: TRSM of tile
// (k, k-1), the update of diagonal tile (k, k) by column k-1 from LDS, the
// factor and inverse of tile (k, k)); every other workgroup takes tasks --
// the remaining TRSMs, the trailing updates and, in the FP32 form, the rows
// of X = L^-1 -- through one atomic ticket, in the order and with the
// dependencies of persist_tasks.hpp.  Results are the launch-per-column
// factor's: L_ik below the diagonal, L_ik^T in the upper tile (k, i), D / DinvT,
// X and R.  A task's results are published with release stores after every
// thread's agent-scope release fence and a barrier; a waiter polls with
// acquire loads (thread 0), then every thread takes an agent-scope acquire
// fence.  Waits end on the flag, on a failure flag (a bad pivot or a non-SPD
// V, then the launch drains) or after kWaitTicks of wall clock (solve_err;
// never expected: the tickets are ordered so every wait has a running owner).

// Loads and stores of data that workgroups exchange inside the launch.  The
// L2 of an XCD is not coherent with the other XCDs' L2s: agent-scope atomic
// accesses (sc1) write through and read from the coherence point instead.
// Tiles that several workgroups read-modify-write (the trailing tiles A_ij,
// the partial sums R_ij) are read and written this way; tiles written once
// per launch (L_ik, L_ik^T, the diagonal tiles and inverses, X_kj) are written
// this way and read with ordinary loads -- no cache holds a copy of them
// before they are written, since nothing reads them earlier in the launch and
// the caches are invalidated at launch.  So no waiter needs an L2
// invalidation, and a publication needs only its stores to have completed.
template <typename T>
__device__ __forceinline__ T ld_coh(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_coh(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool p_failed(const TrialFlags* f) {
    return (__hip_atomic_load(&f->nonspd, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) |
            __hip_atomic_load(&f->chol_fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) |
            __hip_atomic_load(&f->solve_err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0;
}

// Thread 0 only: waits for each flag in d; false when the launch must stop.
__device__ bool p_wait_all(const PFlag* d, int nd, int* const* fl, TrialFlags* flags) {
    for (int q = 0; q < nd; ++q) {
        const int* p = fl[d[q].arr] + d[q].idx;
        const long long t0 = wall_clock64();
        for (;;) {
            if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= d[q].want) break;
            if (p_failed(flags)) return false;
            if (wall_clock64() - t0 > kWaitTicks) {
                atomicOr(&flags->solve_err, 1);
                return false;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return true;
}

// Every thread: the workgroup's wait on d (thread 0 polls, the barrier
// orders every later load of the workgroup after the flag it saw); false =
// stop.
__device__ __forceinline__ bool p_wg_wait(const PFlag* d, int nd, int* const* fl, TrialFlags* flags, int* s_ok) {
    if (threadIdx.x == 0) *s_ok = p_wait_all(d, nd, fl, flags) ? 1 : 0;
    __syncthreads();
    return *s_ok != 0;
}

// Every thread: publish f once every wave's stores have completed (s_waitcnt
// vmcnt(0), then the barrier; thread 0's release store adds the L2 write-back
// for any ordinary store).
__device__ __forceinline__ void p_wg_publish(const PFlag& f, int* const* fl) {
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0) on gfx9-family targets
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(fl[f.arr] + f.idx, f.want, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// 256 threads copy a 64x64 tile into LDS with coherent loads.
template <typename T>
__device__ __forceinline__ void tile_to_lds_coh(T (*s)[kTile + 2], const T* __restrict__ src, int ld) {
    T tmp[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int q = threadIdx.x + 256 * u;
        tmp[u] = ld_coh(src + static_cast<size_t>(q >> 6) * ld + (q & 63));
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int q = threadIdx.x + 256 * u;
        s[q >> 6][q & 63] = tmp[u];
    }
}

// L_ik = A_ik (Linv_kk)^T into Ai (LDS), L_ik over the lower tile and L_ik^T
// into the upper tile (k, i).
template <typename T>
__device__ void p_trsm(T* __restrict__ S, int ld, int i, int k, const T* __restrict__ DinvT, T (*Ai)[kTile + 2]) {
    T* Aik = S + static_cast<size_t>(i * kTile) * ld + k * kTile;
    T* LTki = S + static_cast<size_t>(k * kTile) * ld + i * kTile;
    const T* inv = DinvT + static_cast<size_t>(k) * kTile * kTile;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    T bv[16][2];
    {
        T tmp[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = threadIdx.x + 256 * u;
            tmp[u] = ld_coh(Aik + static_cast<size_t>(q >> 6) * ld + (q & 63));
        }
#pragma unroll
        for (int kk = 0; kk < kTile; kk += 4) {
            bv[kk >> 2][0] = inv[(kk + lk) * kTile + wc + li];
            bv[kk >> 2][1] = inv[(kk + lk) * kTile + wc + 16 + li];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int q = threadIdx.x + 256 * u;
            Ai[q >> 6][q & 63] = tmp[u];
        }
    }
    __syncthreads();
    typename Acc4<T>::type acc[2][2];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[r][c] = typename Acc4<T>::type{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < kTile; kk += 4) {
        const T a0 = Ai[wr + li][kk + lk], a1 = Ai[wr + 16 + li][kk + lk];
        const T b0 = bv[kk >> 2][0], b1 = bv[kk >> 2][1];
        acc[0][0] = mfma16(a0, b0, acc[0][0]);
        acc[0][1] = mfma16(a0, b1, acc[0][1]);
        acc[1][0] = mfma16(a1, b0, acc[1][0]);
        acc[1][1] = mfma16(a1, b1, acc[1][1]);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) Ai[wr + 16 * r + acc_row<T>(lk, q)][wc + 16 * c + li] = acc[r][c][q];
    __syncthreads();
#pragma unroll 4
    for (int u = 0; u < 16; ++u) {
        const int q = threadIdx.x + 256 * u;
        const int r = q >> 6, c = q & 63;
        st_coh(Aik + static_cast<size_t>(r) * ld + c, Ai[r][c]);
        st_coh(LTki + static_cast<size_t>(r) * ld + c, Ai[c][r]);
    }
}

// A_ij -= L_ik L_jk^T, both operands straight from the transposed tiles
// (k, i), (k, j).
template <typename T>
__device__ void p_update(T* __restrict__ S, int ld, int i, int j, int k) {
    const T* LTi = S + static_cast<size_t>(k * kTile) * ld + i * kTile;
    const T* LTj = S + static_cast<size_t>(k * kTile) * ld + j * kTile;
    T* Aij = S + static_cast<size_t>(i * kTile) * ld + j * kTile;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    // the tile's current values first: sixteen independent coherent loads in
    // flight while the operands load and the products run (a coherent load
    // after a coherent store would wait for it)
    T old[2][2][4];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                old[r][c][q] = ld_coh(Aij + static_cast<size_t>(wr + 16 * r + acc_row<T>(lk, q)) * ld + wc + 16 * c + li);
    typename Acc4<T>::type acc[2][2];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[r][c] = typename Acc4<T>::type{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < kTile; kk += 4) {
        const size_t ro = static_cast<size_t>(kk + lk) * ld;
        const T a0 = LTi[ro + wr + li], a1 = LTi[ro + wr + 16 + li];
        const T b0 = LTj[ro + wc + li], b1 = LTj[ro + wc + 16 + li];
        acc[0][0] = mfma16(a0, b0, acc[0][0]);
        acc[0][1] = mfma16(a0, b1, acc[0][1]);
        acc[1][0] = mfma16(a1, b0, acc[1][0]);
        acc[1][1] = mfma16(a1, b1, acc[1][1]);
    }
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = wr + 16 * r + acc_row<T>(lk, q), col = wc + 16 * c + li;
                st_coh(Aij + static_cast<size_t>(row) * ld + col, old[r][c][q] - acc[r][c][q]);
            }
}

// store_diag_tile with coherent stores
template <typename T>
__device__ __forceinline__ void store_diag_tile_coh(const T (*a)[kTile + 2], const T (*b)[kTile + 2], bool ok, int k,
                                                    T* __restrict__ D, T* __restrict__ DinvT, TrialFlags* flags) {
    if (!ok && threadIdx.x == 0) atomicOr(&flags->chol_fail, 1);
    T* d = D + static_cast<size_t>(k) * kTile * kTile;
    T* it = DinvT + static_cast<size_t>(k) * kTile * kTile;
    for (int q = threadIdx.x; q < kTile * kTile; q += blockDim.x) {
        const int r = q >> 6, c = q & 63;
        st_coh(d + q, (c <= r) ? a[r][c] : T(0));
        st_coh(it + q, (r <= c) ? b[c][r] : T(0));
    }
}

struct PersistArgs {
    int* ticket;
    int* fl[kPNumArrays];
    const long long* col_start;
    long long ntasks;
    int inverse;
};

template <typename T>
__global__ __launch_bounds__(256) void k_chol_persistent(T* __restrict__ S, int ld, int nblk, T* __restrict__ D,
                                                         T* __restrict__ DinvT, TrialFlags* flags, PersistArgs pa,
                                                         T* __restrict__ X, T* __restrict__ R) {
    __shared__ T buf0[kTile][kTile + 2];
    __shared__ T buf1[kTile][kTile + 2];
    __shared__ int s_ok;
    __shared__ long long s_t;
    if (blockIdx.x == 0) {
        // ---- critical chain ----
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
        const int li = lane & 15, lk = lane >> 4;
        for (int k = 0; k < nblk; ++k) {
            if (k == 0) {
                if (threadIdx.x == 0) s_ok = p_failed(flags) ? 0 : 1;
                __syncthreads();
                if (!s_ok) return;
                tile_to_lds_coh<T>(buf1, S, ld);
            } else {
                const PFlag wo = pcrit_wait_offdiag(k, nblk);
                if (!p_wg_wait(&wo, 1, pa.fl, flags, &s_ok)) return;
                p_trsm<T>(S, ld, k, k - 1, DinvT, buf0);
                p_wg_publish(pcrit_publish_trsm(k, nblk), pa.fl);
                const PFlag wd = pcrit_wait_diag(k, nblk);
                if (!p_wg_wait(&wd, 1, pa.fl, flags, &s_ok)) return;
                tile_to_lds_coh<T>(buf1, S + static_cast<size_t>(k * kTile) * ld + k * kTile, ld);
                __syncthreads();
                // A_kk -= L_{k,k-1} L_{k,k-1}^T from LDS
                typename Acc4<T>::type acc[2][2];
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int c = 0; c < 2; ++c) acc[r][c] = typename Acc4<T>::type{0, 0, 0, 0};
#pragma unroll 4
                for (int kk = 0; kk < kTile; kk += 4) {
                    const T a0 = buf0[wr + li][kk + lk], a1 = buf0[wr + 16 + li][kk + lk];
                    const T b0 = buf0[wc + li][kk + lk], b1 = buf0[wc + 16 + li][kk + lk];
                    acc[0][0] = mfma16(a0, b0, acc[0][0]);
                    acc[0][1] = mfma16(a0, b1, acc[0][1]);
                    acc[1][0] = mfma16(a1, b0, acc[1][0]);
                    acc[1][1] = mfma16(a1, b1, acc[1][1]);
                }
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int c = 0; c < 2; ++c)
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            buf1[wr + 16 * r + acc_row<T>(lk, q)][wc + 16 * c + li] -= acc[r][c][q];
            }
            __syncthreads();
            const bool ok = tile_potrf_inv<T>(buf1, buf0);
            store_diag_tile_coh<T>(buf1, buf0, ok, k, D, DinvT, flags);
            if (!ok) return;  // chol_fail is set: every waiter drains
            p_wg_publish(pcrit_publish_diag(k), pa.fl);
        }
        return;
    }
    // ---- workers: tasks in ticket order; the next ticket is taken while
    // the current task runs ----
    if (threadIdx.x == 0) s_t = atomicAdd(pa.ticket, 1);
    __syncthreads();
    long long t = s_t;
    while (t < pa.ntasks) {
        __syncthreads();  // every thread has read s_t
        if (threadIdx.x == 0) s_t = atomicAdd(pa.ticket, 1);
        const PTask p = ptask_decode(t, pa.col_start, nblk, pa.inverse != 0);
        PFlag d[5];
        const int nd = ptask_deps(p, nblk, d);
        if (!p_wg_wait(d, nd, pa.fl, flags, &s_ok)) return;
        switch (p.type) {
            case kPTrsm:
                p_trsm<T>(S, ld, p.i, p.k, DinvT, buf0);
                break;
            case kPUpd:
                p_update<T>(S, ld, p.i, p.j, p.k);
                break;
            default: {
                const int r = p.type == kPXDiag ? p.k : (p.type == kPXRow ? p.j : (p.k + 1) + (p.i - p.k - 1) * p.k + p.j);
                inv_role<T, true>(r, p.k, nblk, S, ld, DinvT, X, R, buf0, buf1, p.k);
                break;
            }
        }
        p_wg_publish(ptask_publish(p, nblk), pa.fl);
        __syncthreads();  // s_t holds the next ticket
        t = s_t;
    }
}

}  // namespace

// Number of tiles of a trailing update below tile column k (m = nblk - k - 1
// tiles below the diagonal) covering the first jw + 1 trailing columns.
static int update_tiles(int m, int jw) {
    if (jw >= m - 1) return m * (m + 1) / 2;
    return (jw + 1) * (jw + 2) / 2 + (m - 1 - jw) * (jw + 1);
}

// Grouping of the right-looking factor for large matrices: while more than
// tail tile columns would remain after it, a group of g panel columns takes
// its TRSMs and the updates inside the group column by column, then one
// update of the rest of the trailing matrix by the whole group (rank 64 g).
// Environment: GBA_FACTOR_GROUP (g, default 8; 1: never), GBA_FACTOR_GROUP_TAIL
// (tile columns left to the column-by-column form, default 64).  Measured at
// n = 17,424 (272 tile columns): FP64 factor 101.8 ms column by column, 48.1
// with g = 4, 42.3 with g = 8, 41.7 with g = 12; the FP32 factor with the
// explicit inverse 96 -> 65 ms.
static void factor_grouping(int& g, int& tail) {
    g = 8;
    tail = 64;
    if (const char* e = std::getenv("GBA_FACTOR_GROUP")) g = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("GBA_FACTOR_GROUP_TAIL")) tail = std::max(0, std::atoi(e));
}

template <typename T>
static void factor_columns(hipStream_t st, T* S, int ld, int nblk, T* D, T* Dinv, TrialFlags* flags, T* X,
                           T* R, int fuse_cols) {
    const TileBatch none{nullptr, nullptr, nullptr};
    int g = 8, tail = 64;
    factor_grouping(g, tail);
    auto trsm = [&](int k) {
        hipLaunchKernelGGL(k_trsm_tiles<T>, dim3(nblk - k - 1), dim3(256), 0, st, S, ld, k, D, Dinv, flags, none);
    };
    // update of the trailing tiles from panel columns kc0 .. k, first jw + 1
    // trailing columns; hosts the inverse roles of column k when X is set
    // (in a group kg .. : row k of X only, or, in the launch that closes the
    // group, row k and the group's terms of R)
    auto update = [&](int kc0, int k, int jw, int fused, int kg, bool rterms) {
        const int m = nblk - k - 1, nup = update_tiles(m, jw);
        const int grid = nup + (X ? (rterms ? inv_roles(k, nblk) : k + 1) : 0);
        hipLaunchKernelGGL(k_chol_update<T>, dim3(grid), dim3(256), 0, st, S, ld, k, D, Dinv, flags, none, nup,
                           nblk, X, R, fused, kc0, std::min(jw, m - 1), kg);
    };
    hipLaunchKernelGGL(k_potrf_tile<T>, dim3(1), dim3(256), 0, st, S, ld, 0, D, Dinv, flags, none);
    int k = 0;
    while (k + 1 < nblk) {
        if (g > 1 && nblk - (k + g) > tail) {
            const int k0 = k;
            for (int c = k0; c < k0 + g; ++c) {
                trsm(c);
                // columns c+1 .. k0+g-1 of the group (none after the last)
                if (c + 1 < k0 + g) update(c, c, k0 + g - 2 - c, 0, k0, false);
            }
            update(k0, k0 + g - 1, nblk, 0, k0, true);
            k = k0 + g;
            continue;
        }
        const int m = nblk - k - 1;
        const int fused = m <= fuse_cols ? 1 : 0;
        if (!fused) trsm(k);
        update(k, k, nblk, fused, k, true);
        ++k;
    }
    if (X)
        hipLaunchKernelGGL(k_inv_last_row<T>, dim3(inv_roles(nblk - 1, nblk)), dim3(256), 0, st,
                           static_cast<const T*>(S), ld, nblk, static_cast<const T*>(Dinv), X, R,
                           static_cast<const TrialFlags*>(flags));
}

void launch_cholesky(hipStream_t st, double* S64, float* S32, int ld, int nblk, double* D64,
                     float* D32, TrialFlags* flags, float* X32, float* R32) {
    const size_t tile = static_cast<size_t>(kTile) * kTile;
    if (S64) {
        factor_columns<double>(st, S64, ld, nblk, D64, D64 + static_cast<size_t>(nblk) * tile, flags, nullptr,
                               nullptr, 0);
    } else {
        // the last columns (at most kFuseCols tiles below the diagonal) take
        // their TRSM inside the update (measured at n = 3,204: 1.32 -> 1.26 ms;
        // fusing every column, 1.30 ms: the early updates have many tiles)
        // (only with X: fused columns keep A in their lower tiles, which the
        // triangular solves read)
        constexpr int kFuseCols = 32;
        factor_columns<float>(st, S32, ld, nblk, D32, D32 + static_cast<size_t>(nblk) * tile, flags, X32, R32,
                              X32 ? kFuseCols : 0);
    }
}

size_t persist_sync_ints(int nblk) {
    return 64 + static_cast<size_t>(nblk) + 4 * static_cast<size_t>(nblk) * nblk;
}

void persist_plan(int nblk, bool inverse, bool fp64, std::vector<long long>& col_start, int& grid) {
    col_start.assign(static_cast<size_t>(nblk) + 1, 0);
    ptask_col_start(nblk, inverse, col_start.data());
    int dev = 0, cus = 0, per_cu = 0;
    GBA_HIP(hipGetDevice(&dev));
    GBA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (fp64)
        GBA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_chol_persistent<double>, 256, 0));
    else
        GBA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_chol_persistent<float>, 256, 0));
    // one workgroup for the critical chain plus workers; all of them fit on
    // the device at once (the protocol does not need it, but a worker that is
    // not resident takes no ticket and does no work)
    const long long want = 1 + col_start[nblk];
    grid = static_cast<int>(std::max<long long>(1, std::min<long long>(want, static_cast<long long>(std::max(1, per_cu)) * cus)));
}

void launch_cholesky_persistent(hipStream_t st, double* S64, float* S32, int ld, int nblk, double* D64, float* D32,
                                TrialFlags* flags, const long long* col_start, long long ntasks, int grid, int* sync,
                                float* X32, float* R32) {
    const size_t tile = static_cast<size_t>(kTile) * kTile;
    launch_fill_int(st, sync, static_cast<int>(persist_sync_ints(nblk)), 0);
    PersistArgs pa;
    pa.ticket = sync;
    const size_t nn = static_cast<size_t>(nblk) * nblk;
    pa.fl[kPDiag] = sync + 64;
    pa.fl[kPTrsmDone] = sync + 64 + nblk;
    pa.fl[kPVer] = pa.fl[kPTrsmDone] + nn;
    pa.fl[kPXDone] = pa.fl[kPVer] + nn;
    pa.fl[kPRVer] = pa.fl[kPXDone] + nn;
    pa.col_start = col_start;
    pa.ntasks = ntasks;
    if (S64) {
        pa.inverse = 0;
        hipLaunchKernelGGL(k_chol_persistent<double>, dim3(grid), dim3(256), 0, st, S64, ld, nblk, D64,
                           D64 + static_cast<size_t>(nblk) * tile, flags, pa, static_cast<double*>(nullptr),
                           static_cast<double*>(nullptr));
    } else {
        pa.inverse = X32 ? 1 : 0;
        hipLaunchKernelGGL(k_chol_persistent<float>, dim3(grid), dim3(256), 0, st, S32, ld, nblk, D32,
                           D32 + static_cast<size_t>(nblk) * tile, flags, pa, X32, R32);
    }
}

void launch_batch_cholesky(hipStream_t st, int count, const BatchFactorTables& t, const BatchDesc* desc,
                           double* S, double* D, TrialFlags* flags, const int* active) {
    if (count <= 0) return;
    TileBatch tb{t.potrf0, desc, active};
    hipLaunchKernelGGL(k_potrf_tile<double>, dim3(count), dim3(256), 0, st, S, 0, 0, D, D, flags, tb);
    for (int k = 0; k + 1 < t.nsteps; ++k) {
        const int ntr = t.trsm_off[k + 1] - t.trsm_off[k], nup = t.upd_off[k + 1] - t.upd_off[k];
        if (ntr > 0) {
            tb.tab = t.trsm + t.trsm_off[k];
            hipLaunchKernelGGL(k_trsm_tiles<double>, dim3(ntr), dim3(256), 0, st, S, 0, k, D, D, flags, tb);
        }
        if (nup > 0) {
            tb.tab = t.upd + t.upd_off[k];
            hipLaunchKernelGGL(k_chol_update<double>, dim3(nup), dim3(256), 0, st, S, 0, k, D, D, flags, tb, nup, 0,
                               static_cast<double*>(nullptr), static_cast<double*>(nullptr), 0, k, 1 << 14, k);
        }
    }
}

void launch_batch_chol_solve(hipStream_t st, int count, const BatchDesc* desc, double* S, double* D,
                             const double* rhs, double* dc, TrialFlags* flags, const int* active,
                             bool factor) {
    if (count <= 0) return;
    if (factor)
        hipLaunchKernelGGL(k_batch_chol_solve<true>, dim3(count), dim3(kBatchThreads), 0, st, desc, S, D, rhs, dc,
                           flags, active);
    else
        hipLaunchKernelGGL(k_batch_chol_solve<false>, dim3(count), dim3(kBatchThreads), 0, st, desc, S, D, rhs,
                           dc, flags, active);
}

void launch_chol_solve(hipStream_t st, const double* L64, const float* L32, int ld, int nblk,
                       const double* D64, const float* D32, const double* b, double* y, double* x,
                       int* sync, TrialFlags* flags, const int* first, const double* skip) {
    launch_fill_int(st, sync, 2 * nblk + 2, 0);
    int* tickets = sync + 2 * nblk;
    const size_t inv_off = static_cast<size_t>(nblk) * kTile * kTile;
    const size_t np = static_cast<size_t>(nblk) * kTile;
    const double unset = __builtin_bit_cast(double, kUnset);
    // the outputs start unset (x after the forward pass: b and x may alias)
    launch_fill_double(st, y, np, unset);
    if (L64)
        hipLaunchKernelGGL(k_trsv_fwd<double>, dim3(nblk), dim3(256), 0, st, L64, ld, nblk,
                           D64 + inv_off, b, y, tickets, flags, first, skip);
    else
        hipLaunchKernelGGL(k_trsv_fwd<float>, dim3(nblk), dim3(256), 0, st, L32, ld, nblk,
                           D32 + inv_off, b, y, tickets, flags, first, skip);
    launch_fill_double(st, x, np, unset);
    if (L64)
        hipLaunchKernelGGL(k_trsv_bwd<double>, dim3(nblk), dim3(256), 0, st, L64, ld, nblk,
                           D64 + inv_off, y, x, tickets + 1, flags, first, skip);
    else
        hipLaunchKernelGGL(k_trsv_bwd<float>, dim3(nblk), dim3(256), 0, st, L32, ld, nblk,
                           D32 + inv_off, y, x, tickets + 1, flags, first, skip);
}

int residual_partials(int npad) { return (npad + 255) / 256; }

size_t symv_work(int npad) {
    const size_t nblk = static_cast<size_t>(npad / kTile);
    return 2 * (nblk * (nblk + 1) / 2) * kTile;
}

void launch_symv(hipStream_t st, const float* A, int ld, int npad, const double* x, double* y,
                 double* work, const double* done) {
    const int nblk = npad / kTile;
    const int tiles = nblk * (nblk + 1) / 2;
    double* rp = work;
    double* cp = work + static_cast<size_t>(tiles) * kTile;
    hipLaunchKernelGGL(k_symv_tiles, dim3(tiles), dim3(256), 0, st, A, ld, x, rp, cp, done);
    hipLaunchKernelGGL(k_symv_reduce, dim3(nblk), dim3(256), 0, st, nblk, rp, cp, y, done);
}

void launch_pcg_dot(hipStream_t st, int n, const double* a, const double* b, const double* c,
                    const double* d, double* blk, int mode, double* sc, PassStatus* status,
                    const TrialFlags* flags, double tol2, int max_it) {
    const int grid = residual_partials(n);
    hipLaunchKernelGGL(k_dot2, dim3(grid), dim3(256), 0, st, n, a, b, c, d, blk,
                       mode == 0 ? static_cast<const double*>(nullptr) : sc + 4);
    hipLaunchKernelGGL(k_pcg_reduce, dim3(1), dim3(256), 0, st, grid, blk, mode, sc, status, flags, tol2, max_it);
}

void launch_pcg_update(hipStream_t st, int n, const double* sc, double* x, const double* p, double* r,
                       const double* q) {
    hipLaunchKernelGGL(k_pcg_update, dim3((n + 255) / 256), dim3(256), 0, st, n, sc, x, p, r, q);
}

void launch_pcg_direction(hipStream_t st, int n, const double* sc, const double* z, double* p) {
    hipLaunchKernelGGL(k_pcg_direction, dim3((n + 255) / 256), dim3(256), 0, st, n, sc, z, p);
}

void launch_vec_copy(hipStream_t st, int n, const double* a, double* b) {
    hipLaunchKernelGGL(k_vec_copy, dim3((n + 255) / 256), dim3(256), 0, st, n, a, b);
}


void launch_vec_axpy(hipStream_t st, int n, double* x, const double* d) {
    hipLaunchKernelGGL(k_vec_axpy, dim3((n + 255) / 256), dim3(256), 0, st, n, x, d);
}

void launch_vec_scale(hipStream_t st, int n, int npad, const double* in, const double* s,
                      double* out) {
    hipLaunchKernelGGL(k_vec_scale, dim3((npad + 255) / 256), dim3(256), 0, st, n, npad, in, s, out);
}

}  // namespace gba_amd
