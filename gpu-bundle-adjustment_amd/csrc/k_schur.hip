// k_schur.hip -- device assembly of the reduced camera system
//   S = U + lambda diag'(U) - sum_p W_p V'_p^-1 W_p^T,   rhs = bc - sum_e T_e bp_p
// (replaces the host Phase 1 / Phase 2 / merge loops of src/lm.cu:483-594).
//
// Workgroup (a, chunk) owns block row a of S restricted to the camera columns
// [chunk*kSchurCB, ...] below the diagonal: it walks camera a's observations,
// forms T_e = W_e V'^-1 once per observation and, for every other observation f
// of the same point whose camera b falls in the chunk, adds -T_e W_f^T into an
// LDS accumulator of kSchurCB 9x9 blocks.  Each block of S is written by
// exactly one workgroup with plain stores; no global atomics.  FP64 path:
// everything double.  Mixed path: W, T and the accumulator are FP32, the
// result is written diagonally scaled (D S D, D = (U_ii + lambda U_ii)^-1/2)
// into the factor copy (lower blocks) and a full symmetric copy used by the
// FP64 refinement residual; rhs stays FP64.
#include "device_common.hpp"
#include "launchers.hpp"

namespace gba_amd {

namespace {

constexpr int kSchurCB = 16;

template <typename WT, typename ST>
__global__ __launch_bounds__(256) void k_schur(DevProblem pb, const WT* __restrict__ W,
                                               const double* __restrict__ Vinv6,
                                               const double* __restrict__ VB,
                                               const double* __restrict__ U,
                                               const double* __restrict__ bc, double lambda,
                                               const double* __restrict__ dscale,
                                               ST* __restrict__ Sfac, ST* __restrict__ Sfull, int ld,
                                               double* __restrict__ rhs, const TrialFlags* flags) {
    __shared__ ST acc[kSchurCB * 81];
    __shared__ double s_rhs[4][9];
    if (flags->nonspd) return;
    const int a = blockIdx.x;
    const int lo = blockIdx.y * kSchurCB;
    if (lo > a) return;
    const int hi = min(lo + kSchurCB, a + 1);
    const bool own_rhs = (a >= lo && a < hi);
    const int tid = threadIdx.x;
    for (int t = tid; t < kSchurCB * 81; t += 256) acc[t] = ST(0);
    __syncthreads();
    double racc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) racc[i] = 0.0;
    const int beg = pb.cam_off[a], end = pb.cam_off[a + 1];
    for (int idx = beg + tid; idx < end; idx += 256) {
        const int e = pb.cam_list[idx];
        const int p = pb.spt[e];
        const double* vi = Vinv6 + 6 * static_cast<size_t>(p);
        const double V[9] = {vi[0], vi[1], vi[3], vi[1], vi[2], vi[4], vi[3], vi[4], vi[5]};
        const WT* We = W + 27 * static_cast<size_t>(e);
        double T[27];
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const double w0 = We[3 * i], w1 = We[3 * i + 1], w2 = We[3 * i + 2];
#pragma unroll
            for (int j = 0; j < 3; ++j) T[3 * i + j] = w0 * V[j] + w1 * V[3 + j] + w2 * V[6 + j];
        }
        if (own_rhs) {
            const double* bpp = VB + 9 * static_cast<size_t>(p) + 6;
#pragma unroll
            for (int i = 0; i < 9; ++i)
                racc[i] += T[3 * i] * bpp[0] + T[3 * i + 1] * bpp[1] + T[3 * i + 2] * bpp[2];
        }
        ST Ts[27];
#pragma unroll
        for (int i = 0; i < 27; ++i) Ts[i] = static_cast<ST>(T[i]);
        const int f0 = pb.pt_off[p], f1 = pb.pt_off[p + 1];
        for (int f = f0; f < f1; ++f) {
            const int b = pb.scam[f];
            if (b < lo || b >= hi) continue;
            const WT* Wf = W + 27 * static_cast<size_t>(f);
            ST wf[27];
#pragma unroll
            for (int i = 0; i < 27; ++i) wf[i] = static_cast<ST>(Wf[i]);
            ST* dst = acc + (b - lo) * 81;
#pragma unroll
            for (int i = 0; i < 9; ++i)
#pragma unroll
                for (int j = 0; j < 9; ++j) {
                    const ST v = Ts[3 * i] * wf[3 * j] + Ts[3 * i + 1] * wf[3 * j + 1] +
                                 Ts[3 * i + 2] * wf[3 * j + 2];
                    atomicAdd(dst + 9 * i + j, -v);
                }
        }
    }
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
        const int row = 9 * a + tid;
        const double v = bc[row] - ((s_rhs[0][tid] + s_rhs[1][tid]) + (s_rhs[2][tid] + s_rhs[3][tid]));
        rhs[row] = dscale ? v * dscale[row] : v;
    }
    const int nb = hi - lo;
    for (int t = tid; t < nb * 81; t += 256) {
        const int bb = t / 81, ij = t - 81 * bb;
        const int i = ij / 9, j = ij - 9 * i;
        const int b = lo + bb;
        double v = static_cast<double>(acc[t]);
        if (b == a) {
            v += U[81 * static_cast<size_t>(a) + 9 * i + j];
            if (i == j) v += lambda * damp_diag(U[81 * static_cast<size_t>(a) + 10 * i]);
        }
        const int row = 9 * a + i, col = 9 * b + j;
        if (dscale) v *= dscale[row] * dscale[col];
        Sfac[static_cast<size_t>(row) * ld + col] = static_cast<ST>(v);
        if (Sfull) {
            Sfull[static_cast<size_t>(row) * ld + col] = static_cast<ST>(v);
            Sfull[static_cast<size_t>(col) * ld + row] = static_cast<ST>(v);
        }
    }
}

// D_i = (U_ii + lambda diag'(U_ii))^-1/2 for i < n, 1 for the padding.
__global__ void k_diag_scale(int n, int npad, const double* __restrict__ U, double lambda,
                             double* __restrict__ d) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= npad) return;
    if (i >= n) {
        d[i] = 1.0;
        return;
    }
    const double u = U[81 * static_cast<size_t>(i / 9) + 10 * (i % 9)];
    const double v = u + lambda * damp_diag(u);
    d[i] = v > 0.0 ? 1.0 / sqrt(v) : 1.0;
}

}  // namespace

void launch_schur(hipStream_t st, const DevProblem& pb, const double* W64, const float* W32,
                  const double* Vinv6, const double* VB, const double* U, const double* bc,
                  double lambda, const double* dscale, double* S64, float* S32fac, float* S32full,
                  int ld, double* rhs, const TrialFlags* flags) {
    dim3 grid(pb.C, (pb.C + kSchurCB - 1) / kSchurCB);
    if (W32)
        hipLaunchKernelGGL((k_schur<float, float>), grid, dim3(256), 0, st, pb, W32, Vinv6, VB, U, bc,
                           lambda, dscale, S32fac, S32full, ld, rhs, flags);
    else
        hipLaunchKernelGGL((k_schur<double, double>), grid, dim3(256), 0, st, pb, W64, Vinv6, VB, U,
                           bc, lambda, dscale, S64, static_cast<double*>(nullptr), ld, rhs, flags);
}

void launch_diag_scale(hipStream_t st, int n, int npad, const double* U, double lambda,
                       double* dscale) {
    hipLaunchKernelGGL(k_diag_scale, dim3((npad + 255) / 256), dim3(256), 0, st, n, npad, U, lambda,
                       dscale);
}

}  // namespace gba_amd
