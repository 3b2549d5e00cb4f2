// k_schur.hip -- device assembly of the reduced camera system
//   S = U + lambda diag'(U) - sum_p W_p V'_p^-1 W_p^T,   rhs = bc - sum_e T_e bp_p
// (replaces the host Phase 1 / Phase 2 / merge loops of src/lm.cu:483-594).
//
// Work is a fixed list of tasks built once per problem (the sparsity never
// changes during a solve).  Task (a, lo, hi) owns the blocks S[a][b] for
// lo <= b < hi <= a + 1, accumulated in LDS; its observation pairs (e, f) --
// e an observation of camera a, f an observation of the same point by a camera
// b in [lo, hi) -- are stored contiguously, grouped by e and ordered by b.  One
// lane per pair: the lanes of a wavefront that share e hit distinct blocks, so
// the LDS adds of one instruction rarely meet on one address.  Per pair the
// lane forms Z = V'^-1 W_f^T (3x9) and adds -W_e Z row by row.  Each block of
// S is written by exactly one workgroup with plain stores; no global atomics.
// The rhs term T_e bp comes from the pairs with e == f, in FP64.
// FP64 path: everything double.  Mixed path: W is FP32 (widened; products in
// FP64), the result is written diagonally scaled (D S D,
// D = (U_ii + lambda U_ii)^-1/2) into the factor copy and a second lower copy
// used by the FP64 refinement residual, and the LDS sums are 64-bit fixed
// point (2^-46) of the scaled products, added with integer LDS atomics
// (on gfx950 ds_add_u64 runs at 1.3x the rate of ds_add_f64, and the sums do
// not depend on the order of the adds).  No partial sum can leave [-1, 1]: by
// Cauchy-Schwarz in the V'^-1 inner product, sum_p |w_ei V'^-1 w_fj| is at
// most sqrt(U_a,ii U_b,jj), since each point's W_p V'_p^-1 W_p^T is dominated
// by its share of U, and D_i <= U_ii^-1/2.  A pair with a non-finite input
// (which FP64 sums would carry into the factor as a failed pivot) raises
// chol_fail instead.
#include "device_common.hpp"
#include "launchers.hpp"

#include <algorithm>
#include <cstdlib>
#include <type_traits>

namespace gba_amd {

namespace {

// WT: stored W type; AT: LDS accumulator type (int: fixed point of the
// scaled products, which are then formed in FP64; otherwise also the product
// type); ST: type of S.
constexpr double kFix = 70368744177664.0;  // 2^46: fixed-point scale of the mixed sums
// Round-to-nearest of |x| < 2^51 to an integer by one FP64 add: x + 1.5 2^52
// keeps round(x) in the low mantissa bits, read back as a 64-bit integer.
constexpr double kRoundMagic = 6755399441055744.0;  // 1.5 * 2^52
constexpr long long kRoundBits = 0x4338000000000000LL;
template <int NT, typename WT, typename AT, typename ST>
__global__ __launch_bounds__(NT) void k_schur(
    DevProblem pb, const SchurTask* __restrict__ tasks, const int2* __restrict__ pairs,
    const WT* __restrict__ W, const double* __restrict__ Vinv6, const double* __restrict__ VB,
    const double* __restrict__ U, const double* __restrict__ bc, double lambda,
    const double* __restrict__ dscale, ST* __restrict__ Sfac, ST* __restrict__ Slow, int ld,
    double* __restrict__ rhs, TrialFlags* flags) {
    constexpr bool kFixed = std::is_same<AT, long long>::value;
    typedef typename std::conditional<kFixed, double, AT>::type PT;
    extern __shared__ __align__(16) unsigned char smem[];
    AT* acc = reinterpret_cast<AT*>(smem);
    __shared__ double s_rhs[NT / 64][9];
    const SchurTask task = tasks[blockIdx.x];
    const int a = task.a, lo = task.lo, hi = task.hi;
    const int q = prob_of_cam(pb, a);
    if (prob_off(pb, q) || flags[q].nonspd) return;
    lambda = lam_of(pb, q, lambda);
    Sfac += task.soff;
    if (Slow) Slow += task.soff;
    ld = task.ld;
    const bool own_rhs = (hi == a + 1);
    const int tid = threadIdx.x;
    const int nb = hi - lo;
    for (int t = tid; t < nb * 81; t += NT) acc[t] = AT(0);
    PT da[9];  // fixed point: row scales of camera a
#pragma unroll
    for (int i = 0; i < 9; ++i) da[i] = kFixed ? static_cast<PT>(dscale[9 * a + i]) : PT(1);
    __syncthreads();
    double racc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) racc[i] = 0.0;
    bool bad = false;  // fixed point: a product out of range
    for (long long q = task.begin + tid; q < task.end; q += NT) {
        const int2 ef = pairs[q];
        const int e = ef.x, f = ef.y;
        const int p = pb.spt[e];
        const int b = pb.scam[f];
        const double* vi = Vinv6 + 6 * static_cast<size_t>(p);
        const double v00 = vi[0], v10 = vi[1], v11 = vi[2], v20 = vi[3], v21 = vi[4], v22 = vi[5];
        // Z = V'^-1 W_f^T, column j = V'^-1 (row j of W_f)
        const PT a00 = static_cast<PT>(v00), a10 = static_cast<PT>(v10), a11 = static_cast<PT>(v11),
                 a20 = static_cast<PT>(v20), a21 = static_cast<PT>(v21), a22 = static_cast<PT>(v22);
        PT z0[9], z1[9], z2[9];
        {
            WT wf[28];
            load_w_raw(W, static_cast<size_t>(f), wf);
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const PT w0 = static_cast<PT>(wf[3 * j]), w1 = static_cast<PT>(wf[3 * j + 1]),
                         w2 = static_cast<PT>(wf[3 * j + 2]);
                z0[j] = a00 * w0 + a10 * w1 + a20 * w2;
                z1[j] = a10 * w0 + a11 * w1 + a21 * w2;
                z2[j] = a20 * w0 + a21 * w1 + a22 * w2;
                if constexpr (kFixed) {
                    // column scale of camera b, with the fixed-point scale folded in
                    const PT sb = static_cast<PT>(dscale[9 * b + j]) * kFix;
                    z0[j] *= sb;
                    z1[j] *= sb;
                    z2[j] *= sb;
                }
            }
        }
        // W_e rows are read where used: the lanes of an e group share them
        const WT* we = W + kWStride * static_cast<size_t>(e);
        if (own_rhs && e == f) {
            // T_e bp = W_e (V'^-1 bp), in FP64
            const double* bpp = VB + 9 * static_cast<size_t>(p) + 6;
            const double vb0 = v00 * bpp[0] + v10 * bpp[1] + v20 * bpp[2];
            const double vb1 = v10 * bpp[0] + v11 * bpp[1] + v21 * bpp[2];
            const double vb2 = v20 * bpp[0] + v21 * bpp[1] + v22 * bpp[2];
#pragma unroll
            for (int i = 0; i < 9; ++i)
                racc[i] += static_cast<double>(we[3 * i]) * vb0 + static_cast<double>(we[3 * i + 1]) * vb1 +
                           static_cast<double>(we[3 * i + 2]) * vb2;
        }
        AT* dst = acc + (b - lo) * 81;
        if constexpr (kFixed) {
            // the inputs of this pair are finite (a sum of magnitudes is);
            // then no scaled product leaves the fixed-point range
            PT mag = PT(0);
#pragma unroll
            for (int j = 0; j < 9; ++j) mag += fabs(z0[j]) + fabs(z1[j]) + fabs(z2[j]);
#pragma unroll
            for (int i = 0; i < 27; ++i) mag += fabs(static_cast<PT>(we[i]));
            if (!(mag < 1e300)) {
                bad = true;
                continue;
            }
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const PT x0 = static_cast<PT>(we[3 * i]) * da[i], x1 = static_cast<PT>(we[3 * i + 1]) * da[i],
                     x2 = static_cast<PT>(we[3 * i + 2]) * da[i];
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const PT prod = -(x0 * z0[j] + x1 * z1[j] + x2 * z2[j]);
                if constexpr (kFixed) {
                    const long long v = __double_as_longlong(prod + kRoundMagic) - kRoundBits;
                    atomicAdd(reinterpret_cast<unsigned long long*>(dst + 9 * i + j),
                              static_cast<unsigned long long>(v));
                } else {
                    atomicAdd(dst + 9 * i + j, prod);
                }
            }
        }
    }
    if (kFixed && bad) atomicOr(&flags[q].chol_fail, 1);
    if (own_rhs) {
        const int w = tid >> 6, lane = tid & 63;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const double t = wave_sum(racc[i]);
            if (lane == 0) s_rhs[w][i] = t;
        }
    }
    __syncthreads();
    if (own_rhs && tid < 9) {
        double sum = 0.0;
        for (int w = 0; w < NT / 64; ++w) sum += s_rhs[w][tid];
        const int row = 9 * a + tid;
        const double v = bc[row] - sum;
        rhs[row] = dscale ? v * dscale[row] : v;
    }
    for (int t = tid; t < nb * 81; t += NT) {
        const int bb = t / 81, ij = t - 81 * bb;
        const int i = ij / 9, j = ij - 9 * i;
        const int b = lo + bb;
        double v = static_cast<double>(acc[t]) * (kFixed ? 1.0 / kFix : 1.0);
        double u = 0.0;
        if (b == a) {
            u = U[81 * static_cast<size_t>(a) + 9 * i + j];
            if (i == j) u += lambda * damp_diag(U[81 * static_cast<size_t>(a) + 10 * i]);
        }
        if (kFixed)
            v += u * dscale[9 * a + i] * dscale[9 * b + j];  // the sums are scaled already
        else if (dscale)
            v = (v + u) * dscale[9 * a + i] * dscale[9 * b + j];
        else
            v += u;
        const int row = 9 * (a - task.a0) + i, col = 9 * (b - task.a0) + j;
        Sfac[static_cast<size_t>(row) * ld + col] = static_cast<ST>(v);
        if (Slow) Slow[static_cast<size_t>(row) * ld + col] = static_cast<ST>(v);
    }
}

// Diagonal block a and rhs segment a, one workgroup per camera (used when the
// pair list leaves out the pairs e == f, as the batched solver does: for
// problems with few observations per point those pairs are a large share of
// the list and all land on one LDS block):
//   S_aa = U_aa + lambda diag'(U_aa) - sum_{e in a} W_e V'^-1 W_e^T,
//   rhs_a = bc_a - sum_{e in a} W_e V'^-1 bp.
// Each thread sums its strided share of camera a's observations in FP64
// registers (45 lower entries + 9 rhs entries), then a fixed-order reduction
// (wavefront shuffles, then the four wavefronts in order): no atomics.
// desc (batched problems) gives the problem's S offset, order and first
// camera; without it S has leading dimension ld and starts at camera 0.
template <typename WT, typename ST>
__global__ __launch_bounds__(256) void k_schur_diag(
    DevProblem pb, const BatchDesc* __restrict__ desc, const WT* __restrict__ W,
    const double* __restrict__ Vinv6, const double* __restrict__ VB, const double* __restrict__ U,
    const double* __restrict__ bc, double lambda, const double* __restrict__ dscale,
    ST* __restrict__ S, ST* __restrict__ Slow, int ld, double* __restrict__ rhs,
    const TrialFlags* flags) {
    __shared__ double s_red[4][54];
    __shared__ double s_dup[45];  // repeated observations (see below)
    const int a = blockIdx.x, tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    if (tid < 45) s_dup[tid] = 0.0;
    __syncthreads();
    const int q = prob_of_cam(pb, a);
    if (prob_off(pb, q) || flags[q].nonspd) return;
    lambda = lam_of(pb, q, lambda);
    int a0 = 0;
    if (desc) {
        S += desc[q].soff;
        ld = desc[q].npad;
        a0 = desc[q].c0;
    }
    double acc[45], racc[9];
#pragma unroll
    for (int i = 0; i < 45; ++i) acc[i] = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) racc[i] = 0.0;
    for (int idx = pb.cam_off[a] + tid; idx < pb.cam_off[a + 1]; idx += 256) {
        const int e = pb.cam_list[idx];
        const int p = cam_obs_point(pb, idx);
        const double* vi = Vinv6 + 6 * static_cast<size_t>(p);
        const double v00 = vi[0], v10 = vi[1], v11 = vi[2], v20 = vi[3], v21 = vi[4], v22 = vi[5];
        const double* bpp = VB + 9 * static_cast<size_t>(p) + 6;
        const double vb0 = v00 * bpp[0] + v10 * bpp[1] + v20 * bpp[2];
        const double vb1 = v10 * bpp[0] + v11 * bpp[1] + v21 * bpp[2];
        const double vb2 = v20 * bpp[0] + v21 * bpp[1] + v22 * bpp[2];
        double wv[27];
        load_w(W, static_cast<size_t>(e), wv);
        // a heavy point's S terms come from the dense product (k_heavy.hip)
        const bool heavy = pb.heavy_k > 0 && pb.pt_off[p + 1] - pb.pt_off[p] >= pb.heavy_k;
        if (!heavy) {
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                const double t0 = wv[3 * i] * v00 + wv[3 * i + 1] * v10 + wv[3 * i + 2] * v20;
                const double t1 = wv[3 * i] * v10 + wv[3 * i + 1] * v11 + wv[3 * i + 2] * v21;
                const double t2 = wv[3 * i] * v20 + wv[3 * i + 1] * v21 + wv[3 * i + 2] * v22;
#pragma unroll
                for (int j = 0; j <= i; ++j)
                    acc[i * (i + 1) / 2 + j] += t0 * wv[3 * j] + t1 * wv[3 * j + 1] + t2 * wv[3 * j + 2];
            }
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) racc[i] += wv[3 * i] * vb0 + wv[3 * i + 1] * vb1 + wv[3 * i + 2] * vb2;
        // a point observed more than once by camera a: the pairs e != f of
        // that run (adjacent in the point's camera-ordered list) also land
        // on S_aa, W_e V'^-1 W_f^T (src/lm.cu:515-559 takes every f with
        // camera(f) <= camera(e)).  Rare, so it stays out of the registers
        // of the common path: rows of W_e and W_f read from memory, sums
        // into s_dup with LDS atomics.
        if (!pb.repeated_obs || heavy) continue;
        const int k0 = pb.pt_off[p], k1 = pb.pt_off[p + 1];
        if ((e > k0 && pb.scam[e - 1] == a) || (e + 1 < k1 && pb.scam[e + 1] == a)) {
            int f = e;
            while (f > k0 && pb.scam[f - 1] == a) --f;
            const WT* we = W + kWStride * static_cast<size_t>(e);
            for (; f < k1 && pb.scam[f] == a; ++f) {
                if (f == e) continue;
                const WT* wf = W + kWStride * static_cast<size_t>(f);
#pragma unroll 1
                for (int i = 0; i < 9; ++i) {
                    const double e0 = static_cast<double>(we[3 * i]), e1 = static_cast<double>(we[3 * i + 1]),
                                 e2 = static_cast<double>(we[3 * i + 2]);
                    const double t0 = e0 * v00 + e1 * v10 + e2 * v20;
                    const double t1 = e0 * v10 + e1 * v11 + e2 * v21;
                    const double t2 = e0 * v20 + e1 * v21 + e2 * v22;
#pragma unroll 1
                    for (int j = 0; j <= i; ++j)
                        atomicAdd(&s_dup[i * (i + 1) / 2 + j],
                                  t0 * static_cast<double>(wf[3 * j]) + t1 * static_cast<double>(wf[3 * j + 1]) +
                                      t2 * static_cast<double>(wf[3 * j + 2]));
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 54; ++i) {
        const double t = wave_sum(i < 45 ? acc[i] : racc[i - 45]);
        if (lane == 0) s_red[w][i] = t;
    }
    __syncthreads();
    if (tid < 54) {
        double sum = ((s_red[0][tid] + s_red[1][tid]) + s_red[2][tid]) + s_red[3][tid];
        if (tid < 45) sum += s_dup[tid];
        if (tid >= 45) {
            const int row = 9 * a + (tid - 45);
            const double v = bc[row] - sum;
            rhs[row] = dscale ? v * dscale[row] : v;
        } else {
            int i = 0;
            while ((i + 1) * (i + 2) / 2 <= tid) ++i;
            const int j = tid - i * (i + 1) / 2;
            double v = U[81 * static_cast<size_t>(a) + 9 * i + j] - sum;
            if (i == j) v += lambda * damp_diag(U[81 * static_cast<size_t>(a) + 10 * i]);
            if (dscale) v *= dscale[9 * a + i] * dscale[9 * a + j];
            const int row = 9 * (a - a0) + i, col = 9 * (a - a0) + j;
            S[static_cast<size_t>(row) * ld + col] = static_cast<ST>(v);
            S[static_cast<size_t>(col) * ld + row] = static_cast<ST>(v);
            if (Slow) {
                Slow[static_cast<size_t>(row) * ld + col] = static_cast<ST>(v);
                Slow[static_cast<size_t>(col) * ld + row] = static_cast<ST>(v);
            }
        }
    }
}

// D_i = (U_ii + lambda diag'(U_ii))^-1/2 for i < n, 1 for the padding;
// batched (cam_prob set): lambda of row i's problem, lam[cam_prob[i / 9]].
__global__ void k_diag_scale(int n, int npad, const double* __restrict__ U, double lambda,
                             const int* __restrict__ cam_prob, const double* __restrict__ lam,
                             double* __restrict__ d) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= npad) return;
    if (i >= n) {
        d[i] = 1.0;
        return;
    }
    if (cam_prob) lambda = lam[cam_prob[i / 9]];
    const double u = U[81 * static_cast<size_t>(i / 9) + 10 * (i % 9)];
    const double v = u + lambda * damp_diag(u);
    d[i] = v > 0.0 ? 1.0 / sqrt(v) : 1.0;
}

template <int NT, typename WT, typename AT, typename ST>
void launch_one(hipStream_t st, const SchurShape& sh, const DevProblem& pb, const SchurTask* tasks,
                int num_tasks, const int2* pairs, const WT* W, const double* Vinv6,
                const double* VB, const double* U, const double* bc, double lambda,
                const double* dscale, ST* Sfac, ST* Slow, int ld, double* rhs,
                TrialFlags* flags) {
    const size_t lds = static_cast<size_t>(sh.cb) * 81 * sizeof(AT);
    if (lds > 64 * 1024 - 4096)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_schur<NT, WT, AT, ST>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    hipLaunchKernelGGL((k_schur<NT, WT, AT, ST>), dim3(num_tasks), dim3(NT), lds, st, pb, tasks,
                       pairs, W, Vinv6, VB, U, bc, lambda, dscale, Sfac, Slow, ld, rhs, flags);
}

}  // namespace

SchurShape schur_shape(int C, bool /*w32*/) {
    SchurShape sh;
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        v = 64 * 1024;
    (void)hipGetLastError();
    // Accumulator LDS per workgroup: 128 KB (one workgroup of 1024 threads
    // per CU) unless GBA_SCHUR_LDS_KB asks otherwise; FP64 sums in the FP64
    // form, 64-bit fixed point in the mixed form (on gfx950, LDS atomics
    // measured 7.5 lanes/clk/CU for ds_add_f64, 9.75 for ds_add_u64, 13.9
    // for ds_add_u32 and 0.33 for ds_add_f32; 32-bit fixed point, 2^-30,
    // was faster still but too coarse for the small entries).
    int bytes = 128 * 1024;
    if (const char* env = std::getenv("GBA_SCHUR_LDS_KB")) {
        const int kb = std::atoi(env);
        if (kb >= 8) bytes = kb * 1024;
    }
    bytes = std::min(bytes, v - 4096);
    const int acc_bytes = 8;  // FP64 or 64-bit fixed-point LDS sums
    sh.cb = std::max(1, std::min(C, bytes / (81 * acc_bytes)));
    sh.threads = (static_cast<size_t>(sh.cb) * 81 * acc_bytes > 80 * 1024) ? 1024 : 512;
    return sh;
}

void launch_schur(hipStream_t st, const SchurShape& sh, const DevProblem& pb,
                  const SchurTask* tasks, int num_tasks, const int2* pairs, const double* W64,
                  const float* W32, const double* Vinv6, const double* VB, const double* U,
                  const double* bc, double lambda, const double* dscale, double* S64, float* S32fac,
                  float* S32low, int ld, double* rhs, TrialFlags* flags) {
    if (W32 && !S32fac) {
        // FP32 W with the fixed-point sums, S stored FP64 (the batched mixed form)
        if (sh.threads == 1024)
            launch_one<1024, float, long long, double>(st, sh, pb, tasks, num_tasks, pairs, W32, Vinv6, VB, U, bc,
                                                       lambda, dscale, S64, static_cast<double*>(nullptr), ld,
                                                       rhs, flags);
        else
            launch_one<512, float, long long, double>(st, sh, pb, tasks, num_tasks, pairs, W32, Vinv6, VB, U, bc,
                                                      lambda, dscale, S64, static_cast<double*>(nullptr), ld,
                                                      rhs, flags);
    } else if (W32) {
        if (sh.threads == 1024)
            launch_one<1024, float, long long, float>(st, sh, pb, tasks, num_tasks, pairs, W32, Vinv6, VB,
                                                      U, bc, lambda, dscale, S32fac, S32low, ld, rhs, flags);
        else
            launch_one<512, float, long long, float>(st, sh, pb, tasks, num_tasks, pairs, W32, Vinv6, VB, U,
                                                     bc, lambda, dscale, S32fac, S32low, ld, rhs, flags);
    } else {
        if (sh.threads == 1024)
            launch_one<1024, double, double, double>(st, sh, pb, tasks, num_tasks, pairs, W64, Vinv6, VB, U,
                                                     bc, lambda, dscale, S64, static_cast<double*>(nullptr),
                                                     ld, rhs, flags);
        else
            launch_one<512, double, double, double>(st, sh, pb, tasks, num_tasks, pairs, W64, Vinv6, VB, U,
                                                    bc, lambda, dscale, S64, static_cast<double*>(nullptr),
                                                    ld, rhs, flags);
    }
}

void launch_schur_diag(hipStream_t st, const DevProblem& pb, const BatchDesc* desc, const double* W64,
                       const float* W32, const double* Vinv6, const double* VB, const double* U,
                       const double* bc, double lambda, const double* dscale, double* S64,
                       float* S32fac, float* S32low, int ld, double* rhs, const TrialFlags* flags) {
    if (pb.C <= 0) return;
    if (W32 && !S32fac)
        hipLaunchKernelGGL((k_schur_diag<float, double>), dim3(pb.C), dim3(256), 0, st, pb, desc, W32, Vinv6,
                           VB, U, bc, lambda, dscale, S64, static_cast<double*>(nullptr), ld, rhs, flags);
    else if (W32)
        hipLaunchKernelGGL((k_schur_diag<float, float>), dim3(pb.C), dim3(256), 0, st, pb, desc, W32, Vinv6,
                           VB, U, bc, lambda, dscale, S32fac, S32low, ld, rhs, flags);
    else
        hipLaunchKernelGGL((k_schur_diag<double, double>), dim3(pb.C), dim3(256), 0, st, pb, desc, W64,
                           Vinv6, VB, U, bc, lambda, dscale, S64, static_cast<double*>(nullptr), ld, rhs,
                           flags);
}

void launch_diag_scale(hipStream_t st, int n, int npad, const double* U, double lambda,
                       double* dscale, const int* cam_prob, const double* lam) {
    hipLaunchKernelGGL(k_diag_scale, dim3((npad + 255) / 256), dim3(256), 0, st, n, npad, U, lambda,
                       cam_prob, lam, dscale);
}

}  // namespace gba_amd
