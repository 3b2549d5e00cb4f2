// k_schur.hip -- device assembly of the reduced camera system
//   S = U + lambda diag'(U) - sum_p W_p V'_p^-1 W_p^T,   rhs = bc - sum_e T_e bp_p
// (replaces the host Phase 1 / Phase 2 / merge loops of src/lm.cu:483-594).
//
// Workgroup (a, chunk) owns block row a of S restricted to a chunk of camera
// columns b <= a wide enough that the whole chunk lives in LDS (up to 128 KB:
// 202 cameras in FP64, 404 in FP32).  Its 1024 threads walk camera a's
// observations (pre-sorted by the point's observation count, so a wavefront
// gets similar work), form T_e = W_e V'^-1 once per observation and, for every
// other observation f of the same point with camera b in the chunk (a point's
// observations are sorted by camera: skip to lo, stop at hi), add -T_e W_f^T
// into the LDS accumulator (ds_add_f64; FP64 accumulation in both forms).  Each block of S is
// written by exactly one workgroup with plain stores; no global atomics.
// FP64 path: everything double.  Mixed path: W is FP32, the result is written diagonally scaled (D S D, D = (U_ii + lambda
// U_ii)^-1/2) into the factor copy and a second lower copy used by the FP64
// refinement residual; rhs stays FP64.
#include "device_common.hpp"
#include "launchers.hpp"

#include <algorithm>
#include <cstdlib>

namespace gba_amd {

namespace {

constexpr int kSchurThreads = 1024;
constexpr int kSchurLdsCap = 128 * 1024;  // accumulator bytes per workgroup at most

// WT: stored W type; AT: LDS accumulator type; ST: output type of S.
template <typename WT, typename AT, typename ST>
__global__ __launch_bounds__(kSchurThreads) void k_schur(
    DevProblem pb, const int* __restrict__ slist, const WT* __restrict__ W,
    const double* __restrict__ Vinv6, const double* __restrict__ VB, const double* __restrict__ U,
    const double* __restrict__ bc, double lambda, const double* __restrict__ dscale,
    ST* __restrict__ Sfac, ST* __restrict__ Slow, int ld, double* __restrict__ rhs,
    const TrialFlags* flags, int CB) {
    extern __shared__ __align__(16) unsigned char smem[];
    AT* acc = reinterpret_cast<AT*>(smem);
    __shared__ double s_rhs[kSchurThreads / 64][9];
    if (flags->nonspd) return;
    const int a = blockIdx.x;
    const int lo = blockIdx.y * CB;
    if (lo > a) return;
    const int hi = min(lo + CB, a + 1);
    const bool own_rhs = (a >= lo && a < hi);
    const int tid = threadIdx.x;
    const int nb = hi - lo;
    for (int t = tid; t < nb * 81; t += kSchurThreads) acc[t] = AT(0);
    __syncthreads();
    double racc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) racc[i] = 0.0;
    const int beg = pb.cam_off[a], end = pb.cam_off[a + 1];
    for (int idx = beg + tid; idx < end; idx += kSchurThreads) {
        const int e = slist[idx];
        const int p = pb.spt[e];
        const double* vi = Vinv6 + 6 * static_cast<size_t>(p);
        const double V[9] = {vi[0], vi[1], vi[3], vi[1], vi[2], vi[4], vi[3], vi[4], vi[5]};
        const WT* We = W + 27 * static_cast<size_t>(e);
        if (own_rhs) {
            // T_e bp = W_e (V'^-1 bp), in FP64
            const double* bpp = VB + 9 * static_cast<size_t>(p) + 6;
            const double vb0 = V[0] * bpp[0] + V[1] * bpp[1] + V[2] * bpp[2];
            const double vb1 = V[3] * bpp[0] + V[4] * bpp[1] + V[5] * bpp[2];
            const double vb2 = V[6] * bpp[0] + V[7] * bpp[1] + V[8] * bpp[2];
#pragma unroll
            for (int i = 0; i < 9; ++i)
                racc[i] += static_cast<double>(We[3 * i]) * vb0 + static_cast<double>(We[3 * i + 1]) * vb1 +
                           static_cast<double>(We[3 * i + 2]) * vb2;
        }
        AT Ts[27];  // T_e = W_e V'^-1 in the accumulation precision
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const AT w0 = static_cast<AT>(We[3 * i]), w1 = static_cast<AT>(We[3 * i + 1]),
                     w2 = static_cast<AT>(We[3 * i + 2]);
#pragma unroll
            for (int j = 0; j < 3; ++j)
                Ts[3 * i + j] = w0 * static_cast<AT>(V[j]) + w1 * static_cast<AT>(V[3 + j]) +
                                w2 * static_cast<AT>(V[6 + j]);
        }
        const int f1 = pb.pt_off[p + 1];
        int f = pb.pt_off[p];
        while (f < f1 && pb.scam[f] < lo) ++f;
        for (; f < f1; ++f) {
            const int b = pb.scam[f];
            if (b >= hi) break;
            const WT* Wf = W + 27 * static_cast<size_t>(f);
            AT* dst = acc + (b - lo) * 81;
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const AT w0 = static_cast<AT>(Wf[3 * j]), w1 = static_cast<AT>(Wf[3 * j + 1]),
                         w2 = static_cast<AT>(Wf[3 * j + 2]);
#pragma unroll
                for (int i = 0; i < 9; ++i)
                    atomicAdd(dst + 9 * i + j, -(Ts[3 * i] * w0 + Ts[3 * i + 1] * w1 + Ts[3 * i + 2] * w2));
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
        double sum = 0.0;
        for (int w = 0; w < kSchurThreads / 64; ++w) sum += s_rhs[w][tid];
        const int row = 9 * a + tid;
        const double v = bc[row] - sum;
        rhs[row] = dscale ? v * dscale[row] : v;
    }
    for (int t = tid; t < nb * 81; t += kSchurThreads) {
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
        if (Slow) Slow[static_cast<size_t>(row) * ld + col] = static_cast<ST>(v);
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

namespace {
// Accumulator bytes a workgroup may use: the device's per-block LDS limit
// (160 KB on gfx950) minus the static part, capped at kSchurLdsCap.
int schur_lds_bytes() {
    static int bytes = -1;
    if (bytes < 0) {
        int dev = 0, v = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            v = 64 * 1024;
        bytes = std::min(kSchurLdsCap, v - 4096);
        // tuning knob: accumulator LDS per workgroup in KB (smaller chunks,
        // more workgroups per CU)
        if (const char* env = std::getenv("GBA_SCHUR_LDS_KB")) {
            const int kb = std::atoi(env);
            if (kb >= 16) bytes = std::min(bytes, kb * 1024);
        }
        if (bytes > 64 * 1024 - 4096) {
            if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_schur<float, double, float>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess ||
                hipFuncSetAttribute(reinterpret_cast<const void*>(&k_schur<double, double, double>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess)
                bytes = 64 * 1024 - 4096;
        }
        (void)hipGetLastError();
    }
    return bytes;
}
}  // namespace

void launch_schur(hipStream_t st, const DevProblem& pb, const int* slist, const double* W64,
                  const float* W32, const double* Vinv6, const double* VB, const double* U,
                  const double* bc, double lambda, const double* dscale, double* S64, float* S32fac,
                  float* S32low, int ld, double* rhs, const TrialFlags* flags) {
    const int bytes = schur_lds_bytes();
    if (W32) {
        const int CB = std::min(pb.C, bytes / static_cast<int>(81 * sizeof(double)));
        const int nch = (pb.C + CB - 1) / CB;
        const size_t lds = static_cast<size_t>(CB) * 81 * sizeof(double);
        hipLaunchKernelGGL((k_schur<float, double, float>), dim3(pb.C, nch), dim3(kSchurThreads), lds, st, pb,
                           slist, W32, Vinv6, VB, U, bc, lambda, dscale, S32fac, S32low, ld, rhs,
                           flags, CB);
    } else {
        const int CB = std::min(pb.C, bytes / static_cast<int>(81 * sizeof(double)));
        const int nch = (pb.C + CB - 1) / CB;
        const size_t lds = static_cast<size_t>(CB) * 81 * sizeof(double);
        hipLaunchKernelGGL((k_schur<double, double, double>), dim3(pb.C, nch), dim3(kSchurThreads), lds, st,
                           pb, slist, W64, Vinv6, VB, U, bc, lambda, dscale, S64,
                           static_cast<double*>(nullptr), ld, rhs, flags, CB);
    }
}

void launch_diag_scale(hipStream_t st, int n, int npad, const double* U, double lambda,
                       double* dscale) {
    hipLaunchKernelGGL(k_diag_scale, dim3((npad + 255) / 256), dim3(256), 0, st, n, npad, U, lambda,
                       dscale);
}

}  // namespace gba_amd
