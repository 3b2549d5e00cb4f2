// k_linearize.hip -- linearisation and cost passes (replaces the reference's
// edge_kernel and cost_kernel, src/lm.cu:205-281).
//
// Point pass (one wave per work unit, lane = observation, point-major order):
// residual + analytic Jacobian in FP64; W_e = Jc^T Jp stored (FP64, or FP32 in
// mixed precision); V (lower 6) and bp = -Jp^T r reduced per point with a
// segmented wave64 shuffle scan; SSE and the anomaly count per unit.
//
// Camera pass (one wave per camera chunk, camera-major order): the same
// Jacobian row pair [Jc | r] of every observation is staged in LDS and the
// 10x10 Gram matrix [Jc | r]^T [Jc | r] accumulated with v_mfma_f64_16x16x4f64:
// its 9x9 block is U, its last column is -bc.  Chunk partials are reduced per
// camera in a fixed order, so U / bc / SSE are deterministic.
#include "device_common.hpp"
#include "launchers.hpp"
#include "snavely_model.hpp"

namespace gba_amd {

namespace {

template <typename WT>
__device__ __forceinline__ void store_w(WT* W, size_t e, const double* Jc, const double* Jp, bool ok) {
    double w[27];
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) w[3 * i + j] = ok ? Jc[i] * Jp[j] + Jc[9 + i] * Jp[3 + j] : 0.0;
    store_w_rec(W, e, w);
}

__device__ __forceinline__ void point_terms(const double* r, const double* Jp, bool ok,
                                            double (&acc)[9]) {
    if (!ok) {
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] = 0.0;
        return;
    }
    acc[0] = Jp[0] * Jp[0] + Jp[3] * Jp[3];
    acc[1] = Jp[1] * Jp[0] + Jp[4] * Jp[3];
    acc[2] = Jp[1] * Jp[1] + Jp[4] * Jp[4];
    acc[3] = Jp[2] * Jp[0] + Jp[5] * Jp[3];
    acc[4] = Jp[2] * Jp[1] + Jp[5] * Jp[4];
    acc[5] = Jp[2] * Jp[2] + Jp[5] * Jp[5];
    acc[6] = -(Jp[0] * r[0] + Jp[3] * r[1]);
    acc[7] = -(Jp[1] * r[0] + Jp[4] * r[1]);
    acc[8] = -(Jp[2] * r[0] + Jp[5] * r[1]);
}

template <typename WT>
__global__ __launch_bounds__(256) void k_lin_points(DevProblem pb, const double* __restrict__ cams,
                                                    const double* __restrict__ pts,
                                                    WT* __restrict__ W, double* __restrict__ VB,
                                                    double* __restrict__ unit_sse,
                                                    unsigned* __restrict__ unit_anom) {
    const int lane = lane_id();
    const int u = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (u >= pb.num_units) return;
    const int2 rg = pb.units[u];
    if (prob_off(pb, prob_of_pt(pb, pb.spt[rg.x]))) return;
    double sse = 0.0;
    unsigned anom = 0;
    if (rg.y - rg.x <= kWave) {
        const int e = rg.x + lane;
        const bool active = e < rg.y;
        int key = -1;
        double acc[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] = 0.0;
        if (active) {
            key = pb.spt[e];
            const int c = pb.scam[e];
            const double2 o = obs_px(pb, e, c);
            double r[2], Jc[18], Jp[6];
            const bool ok = snavely_eval<true>(cams + 9 * c, pts + 3 * key, o.x, o.y, r, Jc, Jp);
            const double rho = ok ? robustify(pb.huber, r, Jc, Jp) : 0.0;
            store_w(W, static_cast<size_t>(e), Jc, Jp, ok);
            point_terms(r, Jp, ok, acc);
            if (ok) sse = rho;
            else anom = 1;
        }
        seg_suffix_sum<9>(acc, key);
        const int prev = __shfl_up(key, 1, kWave);
        if (active && (lane == 0 || prev != key)) {
            double* out = VB + 9 * static_cast<size_t>(key);
#pragma unroll
            for (int i = 0; i < 9; ++i) out[i] = acc[i];
        }
    } else {
        const int p = pb.spt[rg.x];
        double acc[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] = 0.0;
        for (int base = rg.x; base < rg.y; base += kWave) {
            const int e = base + lane;
            if (e < rg.y) {
                const int c = pb.scam[e];
                const double2 o = obs_px(pb, e, c);
                double r[2], Jc[18], Jp[6], t[9];
                const bool ok = snavely_eval<true>(cams + 9 * c, pts + 3 * p, o.x, o.y, r, Jc, Jp);
                const double rho = ok ? robustify(pb.huber, r, Jc, Jp) : 0.0;
                store_w(W, static_cast<size_t>(e), Jc, Jp, ok);
                point_terms(r, Jp, ok, t);
#pragma unroll
                for (int i = 0; i < 9; ++i) acc[i] += t[i];
                if (ok) sse += rho;
                else anom += 1;
            }
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] = wave_sum(acc[i]);
        if (lane == 0) {
            double* out = VB + 9 * static_cast<size_t>(p);
#pragma unroll
            for (int i = 0; i < 9; ++i) out[i] = acc[i];
        }
    }
    sse = wave_sum(sse);
    anom = wave_sum_u(anom);
    if (lane == 0) {
        unit_sse[u] = sse;
        unit_anom[u] = anom;
    }
}

typedef double v4d __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_lin_cameras(DevProblem pb, const double* __restrict__ cams,
                                                     const double* __restrict__ pts,
                                                     double* __restrict__ gram_part) {
    // per wave: 128 Jacobian rows of 10 values (Jc row, residual)
    __shared__ double rows[4][2 * kWave * 10];
    const int lane = lane_id();
    const int w = threadIdx.x >> 6;
    const int chunk = blockIdx.x * 4 + w;
    if (chunk >= pb.num_chunks) return;
    const int4 ch = pb.chunks[chunk];
    if (prob_off(pb, prob_of_cam(pb, ch.x))) return;
    const double* cam = cams + 9 * ch.x;
    double* R = rows[w];
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    const int col = lane & 15, rsub = lane >> 4;
    for (int base = ch.y; base < ch.z; base += kWave) {
        const int i = base + lane;
        double r[2] = {0.0, 0.0}, Jc[18], Jp[6];
#pragma unroll
        for (int k = 0; k < 18; ++k) Jc[k] = 0.0;
        if (i < ch.z) {
            const double2 o = cam_obs_px(pb, i, ch.x);
            const int p = cam_obs_point(pb, i);
            const bool ok = snavely_eval<true>(cam, pts + 3 * p, o.x, o.y, r, Jc, Jp);
            if (ok) robustify(pb.huber, r, Jc, nullptr);
            if (!ok) {
                r[0] = r[1] = 0.0;
#pragma unroll
                for (int k = 0; k < 18; ++k) Jc[k] = 0.0;
            }
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            R[(2 * lane) * 10 + k] = Jc[k];
            R[(2 * lane + 1) * 10 + k] = Jc[9 + k];
        }
        R[(2 * lane) * 10 + 9] = r[0];
        R[(2 * lane + 1) * 10 + 9] = r[1];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 4
        for (int s = 0; s < 32; ++s) {
            // A^T A with A = 4 rows x 16 cols (cols >= 10 zero): both operands
            // read element (row 4s + lane/16, col lane%16).
            const double val = col < 10 ? R[(4 * s + rsub) * 10 + col] : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(val, val, acc, 0, 0, 0);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // v_mfma_f64_16x16x4f64 leaves G[lane/16 + 4q][lane%16] in acc[q]; keep
    // the lower triangle of the leading 10x10 block (55 values, row by row).
    double* out = gram_part + 55 * static_cast<size_t>(chunk);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int gi = rsub + 4 * q, gj = col;
        if (gi < 10 && gj <= gi) out[gi * (gi + 1) / 2 + gj] = acc[q];
    }
}

// Per camera (one wavefront, lanes 0..54): fixed-order sum of its chunk
// partials -> U (full 9x9), bc.  The last workgroup reduces the unit SSE /
// anomaly partials.
__global__ __launch_bounds__(256) void k_lin_reduce(DevProblem pb, const double* __restrict__ gram_part,
                                                    const double* __restrict__ unit_sse,
                                                    const unsigned* __restrict__ unit_anom,
                                                    double* __restrict__ U, double* __restrict__ bc,
                                                    PassStatus* status) {
    const int tid = threadIdx.x;
    if (blockIdx.x == static_cast<unsigned>((pb.C + 3) / 4)) {
        __shared__ double s_sse[256];
        __shared__ unsigned s_an[256];
        double s = 0.0;
        unsigned a = 0;
#pragma unroll 8
        for (int u = tid; u < pb.num_units; u += 256) {
            s += unit_sse[u];
            a += unit_anom[u];
        }
        s_sse[tid] = s;
        s_an[tid] = a;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (tid < h) {
                s_sse[tid] += s_sse[tid + h];
                s_an[tid] += s_an[tid + h];
            }
            __syncthreads();
        }
        if (tid == 0) {
            status->sse = s_sse[0];
            status->anomalies = s_an[0];
        }
        return;
    }
    const int c = blockIdx.x * 4 + (tid >> 6);
    const int lane = tid & 63;
    if (c >= pb.C || lane >= 55 || prob_off(pb, prob_of_cam(pb, c))) return;
    double s = 0.0;
    for (int k = pb.cam_chunk_off[c]; k < pb.cam_chunk_off[c + 1]; ++k)
        s += gram_part[55 * static_cast<size_t>(k) + lane];
    // packed index tid -> (i, j), j <= i
    int i = 0;
    while ((i + 1) * (i + 2) / 2 <= lane) ++i;
    const int j = lane - i * (i + 1) / 2;
    if (i < 9) {
        U[81 * static_cast<size_t>(c) + 9 * i + j] = s;
        U[81 * static_cast<size_t>(c) + 9 * j + i] = s;
    } else if (j < 9) {
        bc[9 * static_cast<size_t>(c) + j] = -s;
    }
}

// Residual-only pass (trial cost without back-substitution): one thread per
// observation, partial sums per workgroup.
__global__ __launch_bounds__(256) void k_cost(DevProblem pb, const double* __restrict__ cams,
                                              const double* __restrict__ pts,
                                              double* __restrict__ blk_sse,
                                              unsigned* __restrict__ blk_anom) {
    __shared__ double s_sse[4];
    __shared__ unsigned s_an[4];
    const int e = blockIdx.x * 256 + threadIdx.x;
    double sse = 0.0;
    unsigned an = 0;
    if (e < pb.O) {
        const int c = pb.scam[e], p = pb.spt[e];
        const double2 o = obs_px(pb, e, c);
        double r[2];
        if (snavely_eval<false>(cams + 9 * c, pts + 3 * p, o.x, o.y, r, nullptr, nullptr))
            sse = robust_rho(r[0] * r[0] + r[1] * r[1], pb.huber);
        else
            an = 1;
    }
    sse = wave_sum(sse);
    an = wave_sum_u(an);
    if (lane_id() == 0) {
        s_sse[threadIdx.x >> 6] = sse;
        s_an[threadIdx.x >> 6] = an;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        blk_sse[blockIdx.x] = (s_sse[0] + s_sse[1]) + (s_sse[2] + s_sse[3]);
        blk_anom[blockIdx.x] = s_an[0] + s_an[1] + s_an[2] + s_an[3];
    }
}

// Fixed-order reduction of n partials into status (sse, anomalies).
__global__ __launch_bounds__(256) void k_reduce_partials(int n, const double* __restrict__ sse_part,
                                                         const unsigned* __restrict__ an_part,
                                                         PassStatus* status) {
    __shared__ double s_sse[256];
    __shared__ unsigned s_an[256];
    const int tid = threadIdx.x;
    double s = 0.0;
    unsigned a = 0;
    for (int u = tid; u < n; u += 256) {
        s += sse_part[u];
        a += an_part[u];
    }
    s_sse[tid] = s;
    s_an[tid] = a;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) {
            s_sse[tid] += s_sse[tid + h];
            s_an[tid] += s_an[tid + h];
        }
        __syncthreads();
    }
    if (tid == 0) {
        status->sse = s_sse[0];
        status->anomalies = s_an[0];
    }
}

}  // namespace

void launch_lin_points(hipStream_t st, const DevProblem& pb, const double* cams, const double* pts,
                       double* W64, float* W32, double* VB, double* unit_sse, unsigned* unit_anom) {
    const int grid = (pb.num_units + 3) / 4;
    if (grid == 0) return;
    if (W32)
        hipLaunchKernelGGL(k_lin_points<float>, dim3(grid), dim3(256), 0, st, pb, cams, pts, W32, VB,
                           unit_sse, unit_anom);
    else
        hipLaunchKernelGGL(k_lin_points<double>, dim3(grid), dim3(256), 0, st, pb, cams, pts, W64, VB,
                           unit_sse, unit_anom);
}

void launch_lin_cameras(hipStream_t st, const DevProblem& pb, const double* cams, const double* pts,
                        double* gram_part) {
    const int grid = (pb.num_chunks + 3) / 4;
    if (grid == 0) return;
    hipLaunchKernelGGL(k_lin_cameras, dim3(grid), dim3(256), 0, st, pb, cams, pts, gram_part);
}

void launch_lin_reduce(hipStream_t st, const DevProblem& pb, const double* gram_part,
                       const double* unit_sse, const unsigned* unit_anom, double* U, double* bc,
                       PassStatus* status) {
    // the extra last workgroup sums the whole-problem SSE into *status; callers
    // that sum per problem themselves pass status == nullptr
    const int grid = (pb.C + 3) / 4 + (status ? 1 : 0);
    if (grid > 0)
        hipLaunchKernelGGL(k_lin_reduce, dim3(grid), dim3(256), 0, st, pb, gram_part, unit_sse, unit_anom, U,
                           bc, status);
}

void launch_cost(hipStream_t st, const DevProblem& pb, const double* cams, const double* pts,
                 double* blk_sse, unsigned* blk_anom, PassStatus* status) {
    const int grid = (pb.O + 255) / 256;
    if (grid > 0)
        hipLaunchKernelGGL(k_cost, dim3(grid), dim3(256), 0, st, pb, cams, pts, blk_sse, blk_anom);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, st, grid, blk_sse, blk_anom, status);
}

int cost_partials(const DevProblem& pb) { return (pb.O + 255) / 256; }

}  // namespace gba_amd
