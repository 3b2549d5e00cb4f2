// k_batch.hip -- per-problem reductions and accepted-step copies of the
// batched solver (many independent problems in one concatenated layout; see
// BatchSolver).  The per-edge, per-point and per-camera kernels of the single
// solver run unchanged on the concatenated layout, masked per problem; these
// kernels close a pass per problem.
#include "device_common.hpp"
#include "launchers.hpp"

#include <algorithm>

namespace gba_amd {

namespace {

// One workgroup per active problem: fixed-order sums of the work-unit
// partials of its point pass (SSE, anomalies; with trial = 1 also the
// point part of the gain denominator) and, for a trial, the camera part
// sum_i dc_i (lambda diag'(U_ii) dc_i + bc_i).
__global__ __launch_bounds__(256) void k_batch_status(const BatchDesc* __restrict__ desc, int trial,
                                                      const double* __restrict__ unit_sse,
                                                      const double* __restrict__ unit_den,
                                                      const unsigned* __restrict__ unit_anom,
                                                      const double* __restrict__ dc,
                                                      const double* __restrict__ U,
                                                      const double* __restrict__ bc,
                                                      const double* __restrict__ lam,
                                                      const TrialFlags* __restrict__ flags,
                                                      const int* __restrict__ active,
                                                      PassStatus* __restrict__ status) {
    __shared__ double s_a[256], s_b[256];
    __shared__ unsigned s_n[256];
    const int q = blockIdx.x, tid = threadIdx.x;
    if (!active[q]) return;
    const BatchDesc d = desc[q];
    const bool skipped = trial && (flags[q].nonspd || flags[q].chol_fail);
    double sse = 0.0, den = 0.0;
    unsigned an = 0;
    if (!skipped) {
        for (int u = d.u0 + tid; u < d.u1; u += 256) {
            sse += unit_sse[u];
            if (trial) den += unit_den[u];
            an += unit_anom[u];
        }
        if (trial) {
            const double lambda = lam[q];
            for (int i = 9 * d.c0 + tid; i < 9 * d.c1; i += 256) {
                const double ud = U[81 * static_cast<size_t>(i / 9) + 10 * (i % 9)];
                const double x = dc[i];
                den += x * (lambda * damp_diag(ud) * x + bc[i]);
            }
        }
    }
    s_a[tid] = sse;
    s_b[tid] = den;
    s_n[tid] = an;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) {
            s_a[tid] += s_a[tid + h];
            s_b[tid] += s_b[tid + h];
            s_n[tid] += s_n[tid + h];
        }
        __syncthreads();
    }
    if (tid == 0) {
        status[q].sse = s_a[0];
        status[q].denom = s_b[0];
        status[q].anomalies = s_n[0];
        status[q].flags = flags[q];
    }
}

// Current parameters of every problem with accept[problem] set become the
// trial parameters.
__global__ void k_batch_accept(int nc, int np, const int* __restrict__ cam_prob,
                               const int* __restrict__ pt_prob, const int* __restrict__ accept,
                               const double* __restrict__ tcams, const double* __restrict__ tpts,
                               double* __restrict__ cams, double* __restrict__ pts) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < 9 * nc && accept[cam_prob[i / 9]]) cams[i] = tcams[i];
    if (i < 3 * np && accept[pt_prob[i / 3]]) pts[i] = tpts[i];
}

}  // namespace

void launch_batch_status(hipStream_t st, int count, const BatchDesc* desc, int trial,
                         const double* unit_sse, const double* unit_den, const unsigned* unit_anom,
                         const double* dc, const double* U, const double* bc, const double* lam,
                         const TrialFlags* flags, const int* active, PassStatus* status) {
    if (count > 0)
        hipLaunchKernelGGL(k_batch_status, dim3(count), dim3(256), 0, st, desc, trial, unit_sse, unit_den,
                           unit_anom, dc, U, bc, lam, flags, active, status);
}

void launch_batch_accept(hipStream_t st, int nc, int np, const int* cam_prob, const int* pt_prob,
                         const int* accept, const double* tcams, const double* tpts, double* cams,
                         double* pts) {
    const int n = std::max(9 * nc, 3 * np);
    if (n > 0)
        hipLaunchKernelGGL(k_batch_accept, dim3((n + 255) / 256), dim3(256), 0, st, nc, np, cam_prob, pt_prob,
                           accept, tcams, tpts, cams, pts);
}

}  // namespace gba_amd
