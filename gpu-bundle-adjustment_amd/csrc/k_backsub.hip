// k_backsub.hip -- point back-substitution fused with the trial cost
// (replaces the host loop of src/lm.cu:606-643 plus the trial update and
// cost_kernel, :647-670).
//
// One wavefront per point-pass work unit: every lane forms W_e^T dc_c(e) for
// its observation, a segmented all-reduce gives each lane its point's sum,
// dp = V'^-1 (bp - sum) is formed redundantly per lane, the segment's first
// lane stores the trial point and the point's gain-denominator term, and every
// lane evaluates its residual at the trial parameters.  Per-unit partial sums
// are reduced in a fixed order by k_trial_finalize, together with the camera
// part of the denominator  sum_i dc_i (lambda diag'(U_ii) dc_i + bc_i).
#include "device_common.hpp"
#include "launchers.hpp"
#include "snavely_model.hpp"

namespace gba_amd {

namespace {

__global__ void k_trial_cams(int n, const double* __restrict__ cams, const double* __restrict__ dc,
                             double* __restrict__ tcams) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) tcams[i] = cams[i] + dc[i];
}

__device__ __forceinline__ bool trial_skipped(const TrialFlags* f) {
    return f->nonspd != 0 || f->chol_fail != 0 || f->solve_err != 0;
}

template <typename WT>
__device__ __forceinline__ void wt_dc(const WT* W, int e, const double* dcc, double (&m)[3]) {
    double w[27];
    load_w(W, static_cast<size_t>(e), w);
    m[0] = m[1] = m[2] = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const double d = dcc[i];
        m[0] += w[3 * i] * d;
        m[1] += w[3 * i + 1] * d;
        m[2] += w[3 * i + 2] * d;
    }
}

__device__ __forceinline__ void point_step(const double* VB, const double* Vinv6, int p,
                                           const double (&m)[3], double lambda, double (&dp)[3],
                                           double& den) {
    const double* vb = VB + 9 * static_cast<size_t>(p);
    const double* vi = Vinv6 + 6 * static_cast<size_t>(p);
    const double r0 = vb[6] - m[0], r1 = vb[7] - m[1], r2 = vb[8] - m[2];
    dp[0] = vi[0] * r0 + vi[1] * r1 + vi[3] * r2;
    dp[1] = vi[1] * r0 + vi[2] * r1 + vi[4] * r2;
    dp[2] = vi[3] * r0 + vi[4] * r1 + vi[5] * r2;
    den = dp[0] * (lambda * damp_diag(vb[0]) * dp[0] + vb[6]) +
          dp[1] * (lambda * damp_diag(vb[2]) * dp[1] + vb[7]) +
          dp[2] * (lambda * damp_diag(vb[5]) * dp[2] + vb[8]);
}

template <typename WT>
__global__ __launch_bounds__(256) void k_backsub(DevProblem pb, const WT* __restrict__ W,
                                                 const double* __restrict__ Vinv6,
                                                 const double* __restrict__ VB,
                                                 const double* __restrict__ dc,
                                                 const double* __restrict__ pts,
                                                 const double* __restrict__ tcams,
                                                 double* __restrict__ tpts, double lambda,
                                                 double* __restrict__ unit_sse,
                                                 double* __restrict__ unit_den,
                                                 unsigned* __restrict__ unit_anom,
                                                 const TrialFlags* flags) {
    const int lane = lane_id();
    const int u = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (u >= pb.num_units) return;
    const int2 rg = pb.units[u];
    const int q = prob_of_pt(pb, pb.spt[rg.x]);
    if (prob_off(pb, q) || trial_skipped(flags + q)) return;
    lambda = lam_of(pb, q, lambda);
    double sse = 0.0, den = 0.0;
    unsigned anom = 0;
    if (rg.y - rg.x <= kWave) {
        const int e = rg.x + lane;
        const bool active = e < rg.y;
        int key = -1, c = 0;
        double m[3] = {0.0, 0.0, 0.0};
        if (active) {
            key = pb.spt[e];
            c = pb.scam[e];
            wt_dc(W, e, dc + 9 * c, m);
        }
        seg_allreduce<3>(m, key);
        const int prev = __shfl_up(key, 1, kWave);
        if (active) {
            double dp[3], d;
            point_step(VB, Vinv6, key, m, lambda, dp, d);
            const double* X = pts + 3 * static_cast<size_t>(key);
            const double Xt[3] = {X[0] + dp[0], X[1] + dp[1], X[2] + dp[2]};
            if (lane == 0 || prev != key) {
                double* out = tpts + 3 * static_cast<size_t>(key);
                out[0] = Xt[0];
                out[1] = Xt[1];
                out[2] = Xt[2];
                den = d;
            }
            const double2 o = obs_px(pb, e, c);
            double r[2];
            if (snavely_eval<false>(tcams + 9 * c, Xt, o.x, o.y, r, nullptr, nullptr))
                sse = robust_rho(r[0] * r[0] + r[1] * r[1], pb.huber);
            else
                anom = 1;
        }
    } else {
        const int p = pb.spt[rg.x];
        double m[3] = {0.0, 0.0, 0.0};
        for (int base = rg.x; base < rg.y; base += kWave) {
            const int e = base + lane;
            if (e < rg.y) {
                double t[3];
                wt_dc(W, e, dc + 9 * pb.scam[e], t);
                m[0] += t[0];
                m[1] += t[1];
                m[2] += t[2];
            }
        }
        m[0] = wave_sum(m[0]);
        m[1] = wave_sum(m[1]);
        m[2] = wave_sum(m[2]);
        double dp[3], d;
        point_step(VB, Vinv6, p, m, lambda, dp, d);
        const double* X = pts + 3 * static_cast<size_t>(p);
        const double Xt[3] = {X[0] + dp[0], X[1] + dp[1], X[2] + dp[2]};
        if (lane == 0) {
            double* out = tpts + 3 * static_cast<size_t>(p);
            out[0] = Xt[0];
            out[1] = Xt[1];
            out[2] = Xt[2];
            den = d;
        }
        for (int base = rg.x; base < rg.y; base += kWave) {
            const int e = base + lane;
            if (e < rg.y) {
                const int c = pb.scam[e];
                const double2 o = obs_px(pb, e, c);
                double r[2];
                if (snavely_eval<false>(tcams + 9 * c, Xt, o.x, o.y, r, nullptr, nullptr))
                    sse += robust_rho(r[0] * r[0] + r[1] * r[1], pb.huber);
                else
                    anom += 1;
            }
        }
    }
    sse = wave_sum(sse);
    den = wave_sum(den);
    anom = wave_sum_u(anom);
    if (lane == 0) {
        unit_sse[u] = sse;
        unit_den[u] = den;
        unit_anom[u] = anom;
    }
}

// dp only (tests): same arithmetic as k_backsub.
template <typename WT>
__global__ __launch_bounds__(256) void k_point_step(DevProblem pb, const WT* __restrict__ W,
                                                    const double* __restrict__ Vinv6,
                                                    const double* __restrict__ VB,
                                                    const double* __restrict__ dc,
                                                    double* __restrict__ dp_out) {
    const int lane = lane_id();
    const int u = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (u >= pb.num_units) return;
    const int2 rg = pb.units[u];
    double m[3] = {0.0, 0.0, 0.0};
    int key = -1;
    if (rg.y - rg.x <= kWave) {
        const int e = rg.x + lane;
        if (e < rg.y) {
            key = pb.spt[e];
            wt_dc(W, e, dc + 9 * pb.scam[e], m);
        }
        seg_allreduce<3>(m, key);
    } else {
        key = pb.spt[rg.x];
        for (int base = rg.x; base < rg.y; base += kWave) {
            const int e = base + lane;
            if (e < rg.y) {
                double t[3];
                wt_dc(W, e, dc + 9 * pb.scam[e], t);
                m[0] += t[0];
                m[1] += t[1];
                m[2] += t[2];
            }
        }
        m[0] = wave_sum(m[0]);
        m[1] = wave_sum(m[1]);
        m[2] = wave_sum(m[2]);
        if (lane != 0) key = -1;
    }
    const int prev = __shfl_up(key, 1, kWave);
    if (key >= 0 && (lane == 0 || prev != key)) {
        double dp[3], d;
        point_step(VB, Vinv6, key, m, 0.0, dp, d);
        double* out = dp_out + 3 * static_cast<size_t>(key);
        out[0] = dp[0];
        out[1] = dp[1];
        out[2] = dp[2];
    }
}

constexpr int kFinThreads = 1024;

__global__ __launch_bounds__(kFinThreads) void k_trial_finalize(DevProblem pb, const double* __restrict__ unit_sse,
                                                        const double* __restrict__ unit_den,
                                                        const unsigned* __restrict__ unit_anom,
                                                        const double* __restrict__ dc,
                                                        const double* __restrict__ U,
                                                        const double* __restrict__ bc, double lambda,
                                                        PassStatus* status, const TrialFlags* flags) {
    __shared__ double s_a[kFinThreads], s_b[kFinThreads], s_c[kFinThreads];
    __shared__ unsigned s_n[kFinThreads];
    const int tid = threadIdx.x;
    if (tid == 0) status->flags = *flags;
    if (trial_skipped(flags)) return;
    double sse = 0.0, den = 0.0, cden = 0.0;
    unsigned an = 0;
#pragma unroll 4
    for (int u = tid; u < pb.num_units; u += kFinThreads) {
        sse += unit_sse[u];
        den += unit_den[u];
        an += unit_anom[u];
    }
    for (int i = tid; i < pb.n; i += kFinThreads) {
        const double ud = U[81 * static_cast<size_t>(i / 9) + 10 * (i % 9)];
        const double d = dc[i];
        cden += d * (lambda * damp_diag(ud) * d + bc[i]);
    }
    s_a[tid] = sse;
    s_b[tid] = den;
    s_c[tid] = cden;
    s_n[tid] = an;
    __syncthreads();
    for (int h = kFinThreads / 2; h > 0; h >>= 1) {
        if (tid < h) {
            s_a[tid] += s_a[tid + h];
            s_b[tid] += s_b[tid + h];
            s_c[tid] += s_c[tid + h];
            s_n[tid] += s_n[tid + h];
        }
        __syncthreads();
    }
    if (tid == 0) {
        status->sse = s_a[0];
        status->denom = s_c[0] + s_b[0];
        status->anomalies = s_n[0];
    }
}

}  // namespace

void launch_trial_cams(hipStream_t st, int C, const double* cams, const double* dc, double* tcams) {
    const int n = 9 * C;
    hipLaunchKernelGGL(k_trial_cams, dim3((n + 255) / 256), dim3(256), 0, st, n, cams, dc, tcams);
}

void launch_backsub(hipStream_t st, const DevProblem& pb, const double* W64, const float* W32,
                    const double* Vinv6, const double* VB, const double* dc, const double* pts,
                    const double* tcams, double* tpts, double lambda, double* unit_sse,
                    double* unit_den, unsigned* unit_anom, const TrialFlags* flags) {
    const int grid = (pb.num_units + 3) / 4;
    if (grid == 0) return;
    if (W32)
        hipLaunchKernelGGL(k_backsub<float>, dim3(grid), dim3(256), 0, st, pb, W32, Vinv6, VB, dc, pts,
                           tcams, tpts, lambda, unit_sse, unit_den, unit_anom, flags);
    else
        hipLaunchKernelGGL(k_backsub<double>, dim3(grid), dim3(256), 0, st, pb, W64, Vinv6, VB, dc,
                           pts, tcams, tpts, lambda, unit_sse, unit_den, unit_anom, flags);
}

void launch_trial_finalize(hipStream_t st, const DevProblem& pb, const double* unit_sse,
                           const double* unit_den, const unsigned* unit_anom, const double* dc,
                           const double* U, const double* bc, double lambda, PassStatus* status,
                           const TrialFlags* flags) {
    hipLaunchKernelGGL(k_trial_finalize, dim3(1), dim3(kFinThreads), 0, st, pb, unit_sse, unit_den, unit_anom,
                       dc, U, bc, lambda, status, flags);
}

void launch_point_step(hipStream_t st, const DevProblem& pb, const double* W64, const float* W32,
                       const double* Vinv6, const double* VB, const double* dc, double* dp) {
    const int grid = (pb.num_units + 3) / 4;
    if (grid == 0) return;
    if (W32)
        hipLaunchKernelGGL(k_point_step<float>, dim3(grid), dim3(256), 0, st, pb, W32, Vinv6, VB, dc, dp);
    else
        hipLaunchKernelGGL(k_point_step<double>, dim3(grid), dim3(256), 0, st, pb, W64, Vinv6, VB, dc, dp);
}

}  // namespace gba_amd
