// k_refine.hip -- residual of the damped reduced camera system in FP64
// without S or W:
//   r = bc - (U + lambda diag'(U)) x - sum_{e in c} W_e z_{p(e)},
//   z_p = V'^-1 (bp_p - sum_{f in p} W_f^T x_{c(f)}),
// which is rhs - S x for S = U + lambda diag'(U) - W V'^-1 W^T and
// rhs = bc - W V'^-1 bp.  W_e = Jc_e^T Jp_e is not read: each pass evaluates
// the observation's Jacobians again at the linearisation point (the edge
// model of the linearisation, robust weight included), and applies them as
// Jp^T (Jc x) and Jc^T (Jp z) in product form (edge_jc_x and the others: one
// rotation-sized term per product, no Jacobian entries formed).  The
// residual is then that of the FP64 system whatever precision W and S were
// stored in, for about 50 bytes of reads per observation and pass (pixel,
// indices, parameters) instead of W.  An observation whose residual or
// products are not finite contributes nothing (its W is 0).  The rotation
// terms (sine, cosine, square root, divisions) depend on the camera only:
// the point pass reads them from a per-camera table, the camera pass
// evaluates them once per thread.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "launchers.hpp"
#include "snavely_model.hpp"

namespace gba_amd {
namespace {

__device__ __forceinline__ bool trial_failed(const TrialFlags* f) { return f->nonspd != 0 || f->chol_fail != 0; }

// Square of the robust weight of an observation with residual r (the
// linearisation scales Jc and Jp by it, W by its square)
__device__ __forceinline__ double robust_w2(double delta, const double* r) {
    const double s = r[0] * r[0] + r[1] * r[1];
    return (delta > 0.0 && s > delta * delta) ? delta / sqrt(s) : 1.0;
}

// m = Jp^T (Jc xc) for observation e (camera c, point p); 0 when not finite
__device__ __forceinline__ void jp_t_jc_x(const DevProblem& pb, const double* cams, const double* rot,
                                          const double* pts, int e, int c, int p, const double* xc, double (&m)[3]) {
    const double2 o = obs_px(pb, e, c);
    double r[2];
    EdgeLin el;
    m[0] = m[1] = m[2] = 0.0;
    const double4 t4 = reinterpret_cast<const double4*>(rot)[c];
    const RotTerms rt{t4.x, t4.y, t4.z, t4.w != 0.0};
    snavely_project<true>(cams + 9 * static_cast<size_t>(c), rt, pts + 3 * static_cast<size_t>(p), o.x, o.y, r, el);
    if (!(finite_d(r[0]) && finite_d(r[1]))) return;
    const double w2 = robust_w2(pb.huber, r);
    double jx[2];
    edge_jc_x(el, xc, jx);
    const Vec3 t = edge_jpT_m(el, w2 * jx[0], w2 * jx[1]);
    if (!(finite_d(t.x) && finite_d(t.y) && finite_d(t.z))) return;
    m[0] = t.x;
    m[1] = t.y;
    m[2] = t.z;
}

// Rotation terms of every camera: rot[4 c ..] = (A, B, G, first order).
__global__ void k_rot_terms(int C, const double* __restrict__ cams, double* __restrict__ rot) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= C) return;
    const RotTerms t = rot_terms(cams + 9 * static_cast<size_t>(c));
    reinterpret_cast<double4*>(rot)[c] = make_double4(t.A, t.B, t.G, t.first_order ? 1.0 : 0.0);
}

// Point pass over the work units (a wavefront per unit, as the
// back-substitution): z_p for every observed point of a running problem.
__global__ __launch_bounds__(256) void k_refine_points(DevProblem pb, const double* __restrict__ cams,
                                                       const double* __restrict__ rot,
                                                       const double* __restrict__ pts,
                                                       const double* __restrict__ Vinv6,
                                                       const double* __restrict__ VB, const double* __restrict__ x,
                                                       double* __restrict__ z, const TrialFlags* __restrict__ flags) {
    const int lane = lane_id();
    const int u = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (u >= pb.num_units) return;
    const int2 rg = pb.units[u];
    const int q = prob_of_pt(pb, pb.spt[rg.x]);
    if (prob_off(pb, q) || trial_failed(flags + q)) return;
    double m[3] = {0.0, 0.0, 0.0};
    int key = -1;
    if (rg.y - rg.x <= kWave) {
        const int e = rg.x + lane;
        if (e < rg.y) {
            key = pb.spt[e];
            const int c = pb.scam[e];
            jp_t_jc_x(pb, cams, rot, pts, e, c, key, x + 9 * static_cast<size_t>(c), m);
        }
        seg_allreduce<3>(m, key);
    } else {
        key = pb.spt[rg.x];
        for (int base = rg.x; base < rg.y; base += kWave) {
            const int e = base + lane;
            if (e < rg.y) {
                const int c = pb.scam[e];
                double t[3];
                jp_t_jc_x(pb, cams, rot, pts, e, c, key, x + 9 * static_cast<size_t>(c), t);
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
        const double* vb = VB + 9 * static_cast<size_t>(key);
        const double* vi = Vinv6 + 6 * static_cast<size_t>(key);
        const double r0 = vb[6] - m[0], r1 = vb[7] - m[1], r2 = vb[8] - m[2];
        double* out = z + 3 * static_cast<size_t>(key);
        out[0] = vi[0] * r0 + vi[1] * r1 + vi[3] * r2;
        out[1] = vi[1] * r0 + vi[2] * r1 + vi[4] * r2;
        out[2] = vi[3] * r0 + vi[4] * r1 + vi[5] * r2;
    }
}

// Camera pass, one workgroup per camera: strided share of its observations
// per thread, then a fixed-order reduction (wavefront sums, the four
// wavefronts in order).  out = r, or D r with dscale.
__global__ __launch_bounds__(256) void k_refine_cams(DevProblem pb, const double* __restrict__ cams,
                                                     const double* __restrict__ pts, const double* __restrict__ U,
                                                     const double* __restrict__ bc, double lambda,
                                                     const double* __restrict__ x, const double* __restrict__ z,
                                                     const double* __restrict__ dscale, double* __restrict__ out,
                                                     const TrialFlags* __restrict__ flags) {
    __shared__ double s_red[4][9];
    const int a = blockIdx.x, tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int q = prob_of_cam(pb, a);
    if (prob_off(pb, q) || trial_failed(flags + q)) return;
    lambda = lam_of(pb, q, lambda);
    double acc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) acc[i] = 0.0;
    const double* cam = cams + 9 * static_cast<size_t>(a);
    const RotTerms rt = rot_terms(cam);
    for (int idx = pb.cam_off[a] + tid; idx < pb.cam_off[a + 1]; idx += 256) {
        const int e = pb.cam_list[idx];
        const int p = pb.spt[e];
        const double2 o = obs_px(pb, e, a);
        double r[2];
        EdgeLin el;
        snavely_project<true>(cam, rt, pts + 3 * static_cast<size_t>(p), o.x, o.y, r, el);
        if (!(finite_d(r[0]) && finite_d(r[1]))) continue;
        const double w2 = robust_w2(pb.huber, r);
        const double* zp = z + 3 * static_cast<size_t>(p);
        double jz[2], t[9];
        edge_jp_z(el, Vec3{zp[0], zp[1], zp[2]}, jz);
        edge_jcT_t(el, w2 * jz[0], w2 * jz[1], t);
        bool ok = true;
#pragma unroll
        for (int i = 0; i < 9; ++i) ok = ok && finite_d(t[i]);
        if (!ok) continue;
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] += t[i];
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const double t = wave_sum(acc[i]);
        if (lane == 0) s_red[w][i] = t;
    }
    __syncthreads();
    if (tid < 9) {
        const size_t row = 9 * static_cast<size_t>(a) + tid;
        const double s = ((s_red[0][tid] + s_red[1][tid]) + s_red[2][tid]) + s_red[3][tid];
        const double* Ua = U + 81 * static_cast<size_t>(a) + 9 * tid;
        const double* xa = x + 9 * static_cast<size_t>(a);
        double ux = lambda * damp_diag(Ua[tid]) * xa[tid];
#pragma unroll
        for (int j = 0; j < 9; ++j) ux += Ua[j] * xa[j];
        const double r = bc[row] - ux - s;
        out[row] = dscale ? r * dscale[row] : r;
    }
}

// x_i += s_i d_i (s NULL: 1)
__global__ void k_vec_axpy_scaled(int n, double* __restrict__ x, const double* __restrict__ d,
                                  const double* __restrict__ s) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] += s ? s[i] * d[i] : d[i];
}

}  // namespace

void launch_refine_residual(hipStream_t st, const DevProblem& pb, const double* cams, const double* pts,
                            const double* Vinv6, const double* VB, const double* U, const double* bc, double lambda,
                            const double* x, double* z, double* rot, const double* dscale, double* out,
                            const TrialFlags* flags) {
    const int grid = (pb.num_units + 3) / 4;
    if (pb.C > 0) hipLaunchKernelGGL(k_rot_terms, dim3((pb.C + 255) / 256), dim3(256), 0, st, pb.C, cams, rot);
    if (grid > 0)
        hipLaunchKernelGGL(k_refine_points, dim3(grid), dim3(256), 0, st, pb, cams, rot, pts, Vinv6, VB, x, z,
                           flags);
    if (pb.C > 0)
        hipLaunchKernelGGL(k_refine_cams, dim3(pb.C), dim3(256), 0, st, pb, cams, pts, U, bc, lambda, x, z, dscale,
                           out, flags);
}

void launch_vec_axpy_scaled(hipStream_t st, int n, double* x, const double* d, const double* s) {
    if (n > 0) hipLaunchKernelGGL(k_vec_axpy_scaled, dim3((n + 255) / 256), dim3(256), 0, st, n, x, d, s);
}

}  // namespace gba_amd
