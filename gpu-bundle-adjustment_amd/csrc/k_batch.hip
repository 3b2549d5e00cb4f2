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

// Sum of an int over the 1024 threads of a workgroup (fixed order).
__device__ __forceinline__ int block_count(int v, int* s_tmp) {
    v = static_cast<int>(wave_sum_u(static_cast<unsigned>(v)));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_tmp[threadIdx.x >> 6] = v;
    __syncthreads();
    int t = 0;
    for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) t += s_tmp[w];
    return t;
}

__global__ void k_batch_lm_init(int count, double lambda0, BatchLmState* __restrict__ st, double* __restrict__ lam,
                                int* __restrict__ lin_active, int* __restrict__ trial_active,
                                int* __restrict__ accept, int* __restrict__ counters) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q == 0) {
        counters[0] = 0;
        counters[1] = count;
    }
    if (q >= count) return;
    BatchLmState s{};
    s.t_start = wall_clock64();
    s.lambda = lambda0;
    s.nu = 2.0;
    s.need_lin = 1;
    st[q] = s;
    lam[q] = lambda0;
    lin_active[q] = 1;
    trial_active[q] = 0;
    accept[q] = 0;
}

// After a linearisation pass: the problems in lin_active take its SSE (the
// starting cost on the first one) and anomaly count; the reference's outer
// loop test (iterations < max_iterations) and its inner loop's trial limit
// (none allowed: FAILURE, src/lm.cu:462,704-707) decide who evaluates a trial.
__global__ __launch_bounds__(1024) void k_batch_lm_lin(int count, gba_lm_options opts,
                                                       const PassStatus* __restrict__ lin_status,
                                                       BatchLmState* __restrict__ st,
                                                       const int* __restrict__ lin_active,
                                                       int* __restrict__ trial_active, int* __restrict__ counters) {
    __shared__ int s_tmp[16];
    int running = 0;
    const int round = counters[0] + 1;
    for (int q = threadIdx.x; q < count; q += blockDim.x) {
        BatchLmState s = st[q];
        const int was_done = s.done;
        if (lin_active[q]) {
            s.counter += lin_status[q].anomalies;
            if (!s.started) {
                s.sse = lin_status[q].sse;
                s.cost0 = s.sse / 2.0;
                s.started = 1;
            }
            s.need_lin = 0;
            if (s.iterations >= opts.max_iterations) s.done = 1;
        }
        if (!s.done && opts.max_inner_trials <= 0) {
            s.failed = 1;
            s.done = 1;
        }
        if (s.done && !was_done) {
            s.t_done = wall_clock64();
            s.done_round = round;
        }
        trial_active[q] = s.done ? 0 : 1;
        running += s.done ? 0 : 1;
        st[q] = s;
    }
    running = block_count(running, s_tmp);
    if (threadIdx.x == 0) {
        if (running > 0) counters[0] = round;
        counters[1] = running;
    }
}

// After a trial pass, per problem in trial_active: the reference's gain
// ratio rho = (sse - sse_t) / (0.5 denom) (src/lm.cu:672), accept iff rho > 0
// and the SSE decreases (:678), then lambda *= max(1/3, 1 - (2 rho - 1)^3)
// (:683) and CONVERGENCE when (sse - sse_t) / max(sse, 1) < tol (:681,696);
// any failed trial (non-SPD V, failed factor, denominator not > 0,
// non-finite trial) multiplies lambda by nu and doubles nu (:478-479 ...
// :687-688), and max_inner_trials failures in a row end with FAILURE
// (:704-707); nu is 2 again at each accepted iteration (:461).
__global__ __launch_bounds__(1024) void k_batch_lm_trial(int count, gba_lm_options opts,
                                                         const PassStatus* __restrict__ trial_status,
                                                         BatchLmState* __restrict__ st,
                                                         const int* __restrict__ trial_active,
                                                         int* __restrict__ accept, int* __restrict__ lin_active,
                                                         double* __restrict__ lam, int* __restrict__ counters) {
    __shared__ int s_tmp[16];
    int running = 0;
    const int round = counters[0];
    for (int q = threadIdx.x; q < count; q += blockDim.x) {
        BatchLmState s = st[q];
        int acc = 0;
        if (trial_active[q]) {
            const PassStatus t = trial_status[q];
            s.trials++;
            if (!t.flags.nonspd && !t.flags.chol_fail && t.denom > 0.0) {
                const unsigned long long before = s.counter;
                s.counter += t.anomalies;
                if (s.counter == before && isfinite(t.sse)) {
                    const double rho = (s.sse - t.sse) / (0.5 * t.denom);
                    if (rho > 0.0 && t.sse < s.sse) {
                        acc = 1;
                        const double rel = (s.sse - t.sse) / fmax(s.sse, 1.0);
                        s.sse = t.sse;
                        s.lambda *= fmax(1.0 / 3.0, 1.0 - pow(2.0 * rho - 1.0, 3.0));
                        s.iterations++;
                        s.inner = 0;
                        s.nu = 2.0;
                        if (rel < opts.rel_decrease_tol) {
                            s.converged = 1;
                            s.done = 1;
                        } else {
                            s.need_lin = 1;
                        }
                    }
                }
            }
            if (!acc) {
                s.lambda *= s.nu;
                s.nu *= 2.0;
                if (++s.inner >= opts.max_inner_trials) {
                    s.failed = 1;
                    s.done = 1;
                }
            }
            if (s.done) {
                s.t_done = wall_clock64();
                s.done_round = round;
            }
            st[q] = s;
            lam[q] = s.lambda;
        }
        accept[q] = acc;
        lin_active[q] = (!s.done && s.need_lin) ? 1 : 0;
        running += s.done ? 0 : 1;
    }
    running = block_count(running, s_tmp);
    if (threadIdx.x == 0) counters[1] = running;
}

}  // namespace

void launch_batch_lm_init(hipStream_t st, int count, double lambda0, BatchLmState* state, double* lam,
                          int* lin_active, int* trial_active, int* accept, int* counters) {
    hipLaunchKernelGGL(k_batch_lm_init, dim3((count + 255) / 256), dim3(256), 0, st, count, lambda0, state, lam,
                       lin_active, trial_active, accept, counters);
}

void launch_batch_lm_lin(hipStream_t st, int count, const gba_lm_options& opts, const PassStatus* lin_status,
                         BatchLmState* state, const int* lin_active, int* trial_active, int* counters) {
    hipLaunchKernelGGL(k_batch_lm_lin, dim3(1), dim3(1024), 0, st, count, opts, lin_status, state, lin_active,
                       trial_active, counters);
}

void launch_batch_lm_trial(hipStream_t st, int count, const gba_lm_options& opts, const PassStatus* trial_status,
                           BatchLmState* state, const int* trial_active, int* accept, int* lin_active,
                           double* lam, int* counters) {
    hipLaunchKernelGGL(k_batch_lm_trial, dim3(1), dim3(1024), 0, st, count, opts, trial_status, state,
                       trial_active, accept, lin_active, lam, counters);
}

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
