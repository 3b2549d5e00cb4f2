// device_common.hpp -- device layout of a problem and wave64 helpers.
//
// HBM layout of one problem (all arrays device resident for the whole solve):
//   observations sorted by point (stable in the caller's order):
//     scam[O] int, spt[O] int, spx[O] double2, perm[O] (sorted -> caller index)
//     pt_off[P+1]           edges of point p are [pt_off[p], pt_off[p+1])
//   point-pass work units (int2 edge ranges): either whole points packed into
//     at most 64 edges (one wave, one lane per edge) or a single point with more
//     than 64 edges (one wave looping over it), so every per-point reduction
//     stays inside one wavefront: no atomics, fixed summation order.
//   camera-major index: cam_list[O] (sorted-edge indices grouped by camera,
//     ascending point), cam_off[C+1]; camera chunks (int4: camera, begin,
//     end, 0) of at most kCamChunk entries of cam_list, cam_chunk_off[C+1].
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace gba_amd {

constexpr int kWave = 64;
constexpr int kCamChunk = 256;  // camera-pass edges per wave
constexpr int kTile = 64;       // dense factor tile edge
// W_e (9x3 row-major, 27 values) is stored with a stride of 28 values, so
// every record starts on a 16-byte boundary and moves as 16-byte vectors
// (14 for FP64, 7 for FP32); the 28th value is zero padding.
constexpr int kWStride = 28;

struct DevProblem {
    int C, P, O;
    int n, npad;  // 9C and 9C rounded up to kTile
    int num_units, num_chunks;
    const int* scam;
    const int* spt;
    const double2* spx;
    const int* perm;
    const int* pt_off;
    const int2* units;
    const int* cam_list;
    const int* cam_off;
    const int4* chunks;
    const int* cam_chunk_off;
    // Batched solves (independent problems side by side in one layout; all
    // null for a single problem): problem of each camera and point, lambda
    // per problem, and a per-problem switch selecting the problems a pass
    // runs for.  TrialFlags and PassStatus are then per problem as well.
    const int* cam_prob;
    const int* pt_prob;
    const double* lam;
    const int* active;
    // FP16 pixel storage (null for FP64): per sorted observation the two
    // binary16 pixels packed in one word (x in the low half), and the FP32
    // scale of each camera.
    const unsigned* spx16;
    const float* cam_scale;
    // Huber loss threshold on the residual norm (<= 0: plain least squares).
    double huber;
    // 1 when some point is observed more than once by one camera (set up
    // from the sorted layout; the diagonal Schur blocks then add the cross
    // terms of those observations)
    int repeated_obs;
    // Camera-ordered copies (entry i for observation cam_list[i]): its point
    // and its pixel, read contiguously by the camera passes instead of
    // through cam_list (null: read through cam_list; cspx is null with FP16
    // pixel storage)
    const int* cspt;
    const double2* cspx;
    // Points with at least heavy_k observations (0: none) are left out of the
    // pair and diagonal-block sums; their share of S is one dense product
    // (k_heavy.hip)
    int heavy_k;
};

// Point of the observation at camera-ordered position idx.
__device__ __forceinline__ int cam_obs_point(const DevProblem& pb, int idx) {
    return pb.cspt ? pb.cspt[idx] : pb.spt[pb.cam_list[idx]];
}

// Huber loss on the squared residual norm s = |r|^2, as the reference's
// Ceres harness applies HuberLoss(delta) (benchmarks/bench_ceres.cpp:120-131):
// rho(s) = s for s <= delta^2, else 2 delta sqrt(s) - delta^2; the cost is
// sum rho / 2.  The linearisation scales r and J by sqrt(rho'(s)) (IRLS);
// the second-order correction of a robustified Gauss-Newton step vanishes
// where rho'' <= 0, which holds for Huber everywhere.
__device__ __forceinline__ double robust_rho(double s, double delta) {
    return (delta > 0.0 && s > delta * delta) ? 2.0 * delta * sqrt(s) - delta * delta : s;
}
__device__ __forceinline__ double robust_sqrt_drho(double s, double delta) {
    return (delta > 0.0 && s > delta * delta) ? sqrt(delta / sqrt(s)) : 1.0;
}
// Scales the residual and the Jacobian rows (Jc 2x9, Jp 2x3; either may be
// null) of one observation for the loss; returns rho(|r|^2).
__device__ __forceinline__ double robustify(double delta, double* r, double* Jc, double* Jp) {
    const double s = r[0] * r[0] + r[1] * r[1];
    if (!(delta > 0.0) || s <= delta * delta) return s;
    const double w = robust_sqrt_drho(s, delta);
    r[0] *= w;
    r[1] *= w;
    if (Jc)
        for (int k = 0; k < 18; ++k) Jc[k] *= w;
    if (Jp)
        for (int k = 0; k < 6; ++k) Jp[k] *= w;
    return robust_rho(s, delta);
}

// Observed pixel of sorted observation e (camera c) in FP64.  With FP16
// storage the value is (double)half * (double)scale: an 11-bit times a 24-bit
// significand, exact in FP64.
__device__ __forceinline__ double2 obs_px(const DevProblem& pb, int e, int c) {
    if (pb.spx16) {
        const unsigned u = pb.spx16[e];
        const double s = static_cast<double>(pb.cam_scale[c]);
        const _Float16 hx = __builtin_bit_cast(_Float16, static_cast<unsigned short>(u & 0xffffu));
        const _Float16 hy = __builtin_bit_cast(_Float16, static_cast<unsigned short>(u >> 16));
        return make_double2(static_cast<double>(hx) * s, static_cast<double>(hy) * s);
    }
    return pb.spx[e];
}

// Pixel of the observation at camera-ordered position idx (camera c).
__device__ __forceinline__ double2 cam_obs_px(const DevProblem& pb, int idx, int c) {
    return pb.cspx ? pb.cspx[idx] : obs_px(pb, pb.cam_list[idx], c);
}

// The 27 values of W record e, widened to FP64, read as 16-byte vectors.
__device__ __forceinline__ void load_w(const double* __restrict__ W, size_t e, double (&w)[27]) {
    const double2* v = reinterpret_cast<const double2*>(W + kWStride * e);
#pragma unroll
    for (int k = 0; k < 13; ++k) {
        const double2 t = v[k];
        w[2 * k] = t.x;
        w[2 * k + 1] = t.y;
    }
    w[26] = v[13].x;
}
__device__ __forceinline__ void load_w(const float* __restrict__ W, size_t e, double (&w)[27]) {
    const float4* v = reinterpret_cast<const float4*>(W + kWStride * e);
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const float4 t = v[k];
        w[4 * k] = t.x;
        w[4 * k + 1] = t.y;
        w[4 * k + 2] = t.z;
        if (4 * k + 3 < 27) w[4 * k + 3] = t.w;
    }
}
// W record e in its stored type (28 values, the last one padding).
__device__ __forceinline__ void load_w_raw(const double* __restrict__ W, size_t e, double (&w)[28]) {
    const double2* v = reinterpret_cast<const double2*>(W + kWStride * e);
#pragma unroll
    for (int k = 0; k < 14; ++k) {
        const double2 t = v[k];
        w[2 * k] = t.x;
        w[2 * k + 1] = t.y;
    }
}
__device__ __forceinline__ void load_w_raw(const float* __restrict__ W, size_t e, float (&w)[28]) {
    const float4* v = reinterpret_cast<const float4*>(W + kWStride * e);
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const float4 t = v[k];
        w[4 * k] = t.x;
        w[4 * k + 1] = t.y;
        w[4 * k + 2] = t.z;
        w[4 * k + 3] = t.w;
    }
}
// Stores the 27 values (and the zero padding) of W record e as 16-byte vectors.
__device__ __forceinline__ void store_w_rec(double* __restrict__ W, size_t e, const double (&w)[27]) {
    double2* v = reinterpret_cast<double2*>(W + kWStride * e);
#pragma unroll
    for (int k = 0; k < 13; ++k) v[k] = make_double2(w[2 * k], w[2 * k + 1]);
    v[13] = make_double2(w[26], 0.0);
}
__device__ __forceinline__ void store_w_rec(float* __restrict__ W, size_t e, const double (&w)[27]) {
    float4* v = reinterpret_cast<float4*>(W + kWStride * e);
#pragma unroll
    for (int k = 0; k < 7; ++k)
        v[k] = make_float4(static_cast<float>(w[4 * k]), static_cast<float>(w[4 * k + 1]),
                           static_cast<float>(w[4 * k + 2]), 4 * k + 3 < 27 ? static_cast<float>(w[4 * k + 3]) : 0.0f);
}

__device__ __forceinline__ int prob_of_pt(const DevProblem& pb, int p) { return pb.pt_prob ? pb.pt_prob[p] : 0; }
__device__ __forceinline__ int prob_of_cam(const DevProblem& pb, int c) { return pb.cam_prob ? pb.cam_prob[c] : 0; }
__device__ __forceinline__ bool prob_off(const DevProblem& pb, int q) { return pb.active && !pb.active[q]; }
__device__ __forceinline__ double lam_of(const DevProblem& pb, int q, double lambda) {
    return pb.lam ? pb.lam[q] : lambda;
}

// Schur assembly task: blocks S[a][b], lo <= b < hi <= a + 1, from the
// observation pairs [begin, end) of the pair list (int2: e, f in sorted edge
// indices, same point, camera(e) = a, lo <= camera(f) < hi).  Block (a, b) is
// written at rows 9 (a - a0), columns 9 (b - a0) of the problem's S.
struct SchurTask {
    int a, lo, hi, a0;     // a0: first camera of the task's problem
    long long begin, end;
    long long soff;        // offset of the problem's S (leading dimension ld)
    int ld, pad;
};

// One problem of a batch: its camera and work-unit ranges in the
// concatenated layout, and its dense S (npad x npad, row-major) and diagonal
// tile store (2 npad/64 tiles) offsets.
struct BatchDesc {
    int c0, c1;
    int u0, u1;
    long long soff, doff;
    int n, npad;
};

// Per-trial device flags, cleared by a fill kernel at the start of a trial.
struct TrialFlags {
    int nonspd;     // damped V blocks that are not SPD
    int chol_fail;  // dense factor met a pivot that is not > 0
    int solve_err;  // triangular solve gave up waiting (never expected)
    int pad;
};

// Reduced results of a pass, copied to pinned host memory once per trial.
struct PassStatus {
    double sse;        // sum of squared residuals (finite edges)
    double denom;      // gain-ratio denominator (trial pass only)
    double dot[2];     // refinement dot products (mixed precision)
    unsigned long long anomalies;  // non-finite edges in this pass
    TrialFlags flags;
    int pcg_stop;      // refinement: the last step found p.q or r.z not > 0
    int pcg_done;      // refinement: no further step (converged, stopped, or the step limit)
    int pcg_it;        // refinement: steps counted
    double pcg_rr;     // refinement: r.r after the last counted step (b.b before the first)
    double pcg_bb;     // refinement: b.b
};

// A launch whose done flag (a device scalar, optional) is set does nothing:
// the refinement enqueues several steps between host synchronisations and
// the steps after its stopping test become empty launches.
__device__ __forceinline__ bool pcg_skip(const double* done) { return done && *done != 0.0; }

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}
// N independent wavefront sums, one butterfly level for all of them at a
// time so that their cross-lane exchanges overlap.
template <int N>
__device__ __forceinline__ void wave_sum_n(double* v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] += __shfl_xor(v[i], d, kWave);
}
__device__ __forceinline__ unsigned wave_sum_u(unsigned v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// Segmented suffix sum over lanes with non-decreasing keys inside the active
// prefix (inactive lanes carry a key no active lane has): afterwards the first
// lane of every segment holds the segment total.
template <int N>
__device__ __forceinline__ void seg_suffix_sum(double (&v)[N], int key) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int okey = __shfl_down(key, d, kWave);
        const bool take = (lane + d < kWave) && okey == key;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const double o = __shfl_down(v[i], d, kWave);
            if (take) v[i] += o;
        }
    }
}

// Segmented all-reduce: every lane ends with its segment's total.
template <int N>
__device__ __forceinline__ void seg_allreduce(double (&v)[N], int key) {
    const int lane = lane_id();
    double suf[N];
#pragma unroll
    for (int i = 0; i < N; ++i) suf[i] = v[i];
    seg_suffix_sum<N>(suf, key);
    // prefix (exclusive of suffix overlap): inclusive prefix minus own value
    double pre[N];
#pragma unroll
    for (int i = 0; i < N; ++i) pre[i] = v[i];
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int okey = __shfl_up(key, d, kWave);
        const bool take = (lane >= d) && okey == key;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const double o = __shfl_up(pre[i], d, kWave);
            if (take) pre[i] += o;
        }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = pre[i] + suf[i] - v[i];
}

__device__ __forceinline__ double damp_diag(double x) { return x > 0.0 ? x : 1e-12; }

// Inverse of a symmetric 3x3 given by its lower triangle
// (a00, a10, a11, a20, a21, a22) through A^-1 = L^-T L^-1; returns false when a
// pivot is not > 0 (the reference's SPD test, include/gba/schur.cuh:66-96).
__device__ __forceinline__ bool spd3_inverse(const double* a, double* inv6) {
    const double l00s = a[0];
    if (!(l00s > 0.0)) return false;
    const double l00 = sqrt(l00s);
    const double l10 = a[1] / l00;
    const double l20 = a[3] / l00;
    const double l11s = a[2] - l10 * l10;
    if (!(l11s > 0.0)) return false;
    const double l11 = sqrt(l11s);
    const double l21 = (a[4] - l20 * l10) / l11;
    const double l22s = a[5] - l20 * l20 - l21 * l21;
    if (!(l22s > 0.0)) return false;
    const double l22 = sqrt(l22s);
    // M = L^-1 (lower)
    const double m00 = 1.0 / l00, m11 = 1.0 / l11, m22 = 1.0 / l22;
    const double m10 = -l10 * m00 / l11;
    const double m21 = -l21 * m11 / l22;
    const double m20 = -(l20 * m00 + l21 * m10) / l22;
    // A^-1 = M^T M
    inv6[0] = m00 * m00 + m10 * m10 + m20 * m20;
    inv6[1] = m11 * m10 + m21 * m20;
    inv6[2] = m11 * m11 + m21 * m21;
    inv6[3] = m22 * m20;
    inv6[4] = m22 * m21;
    inv6[5] = m22 * m22;
    return true;
}

// Wall-clock limit of one wait: 10^7 ticks of the constant-rate counter
// (wall_clock64), 100 ms at the 100 MHz of MI300-class GPUs.
constexpr long long kWaitTicks = 10000000LL;

// Bounded wait for a device-scope flag written by another workgroup.  Gives
// up (returns false) when *err becomes non-zero -- another waiter already
// gave up, so the producer chain is broken -- or after kWaitTicks of wall
// clock; the caller then raises *err, so every later wait on the same chain
// returns at once instead of running out its own limit.
__device__ __forceinline__ bool wait_flag(const int* flag, int want, const int* err) {
    const long long t0 = wall_clock64();
    for (;;) {
        if (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) >= want) return true;
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
        if (wall_clock64() - t0 > kWaitTicks) return false;
        __builtin_amdgcn_s_sleep(2);
    }
}

}  // namespace gba_amd
