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
// both passes read them from a per-camera table.  The camera pass runs over
// the camera chunks (a workgroup each, one observation per lane) with
// per-chunk partial sums, summed per row in a fixed order: no atomics, one
// result for one input.
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
// back-substitution): z_p for every observed point of a running problem
// (x NULL: x = 0, z_p = V'^-1 bp).
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
            if (x) jp_t_jc_x(pb, cams, rot, pts, e, c, key, x + 9 * static_cast<size_t>(c), m);
        }
        seg_allreduce<3>(m, key);
    } else {
        key = pb.spt[rg.x];
        for (int base = rg.x; x && base < rg.y; base += kWave) {
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

// Camera pass over the camera chunks (runs of at most 256 observations of
// one camera), one workgroup of four wavefronts per chunk and one observation
// per lane: Jc^T (Jp z_p) and its magnitudes, wavefront sums, then the four
// wavefronts' sums combined in order in LDS: 18 values per chunk into part.
__global__ __launch_bounds__(256) void k_refine_cam_chunks(DevProblem pb, const double* __restrict__ cams,
                                                           const double* __restrict__ rot,
                                                           const double* __restrict__ pts,
                                                           const double* __restrict__ z, double* __restrict__ part,
                                                           const TrialFlags* __restrict__ flags) {
    static_assert(kCamChunk == 256, "one observation per thread of a 256-thread workgroup");
    __shared__ double s_part[4][18];
    const int chunk = blockIdx.x;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int4 ch = pb.chunks[chunk];
    const int a = ch.x;
    const int q = prob_of_cam(pb, a);
    if (prob_off(pb, q) || trial_failed(flags + q)) return;
    double t[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) t[i] = 0.0;
    const int idx = ch.y + tid;
    if (idx < ch.z) {
        const int p = cam_obs_point(pb, idx);
        const double2 o = cam_obs_px(pb, idx, a);
        const double4 t4 = reinterpret_cast<const double4*>(rot)[a];
        const RotTerms rt{t4.x, t4.y, t4.z, t4.w != 0.0};
        double r[2];
        EdgeLin el;
        snavely_project<true>(cams + 9 * static_cast<size_t>(a), rt, pts + 3 * static_cast<size_t>(p), o.x, o.y, r,
                              el);
        if (finite_d(r[0]) && finite_d(r[1])) {
            const double w2 = robust_w2(pb.huber, r);
            const double* zp = z + 3 * static_cast<size_t>(p);
            double jz[2];
            edge_jp_z(el, Vec3{zp[0], zp[1], zp[2]}, jz);
            edge_jcT_t(el, w2 * jz[0], w2 * jz[1], t);
            bool ok = true;
#pragma unroll
            for (int i = 0; i < 9; ++i) ok = ok && finite_d(t[i]);
            if (!ok) {
#pragma unroll
                for (int i = 0; i < 9; ++i) t[i] = 0.0;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const double s = wave_sum(t[i]);
        const double m = wave_sum(fabs(t[i]));
        if (lane == 0) {
            s_part[w][i] = s;
            s_part[w][9 + i] = m;
        }
    }
    __syncthreads();
    if (tid < 18)
        part[18 * static_cast<size_t>(chunk) + tid] =
            ((s_part[0][tid] + s_part[1][tid]) + s_part[2][tid]) + s_part[3][tid];
}

// One thread per camera row: the row's chunk sums in order, then
// out = D (bc - (U + lambda diag'(U)) x - sum) and mag (optional) = D (|bc|
// + |terms of U x| + sum of magnitudes).  The rounding of the residual's
// evaluation in FP64 is a small multiple of 2^-53 times mag; near a minimum
// the residual is a small difference of such terms (the reduced gradient),
// and that is its floor.  x NULL: x = 0.
__global__ void k_refine_cam_rows(DevProblem pb, const double* __restrict__ U, const double* __restrict__ bc,
                                  double lambda, const double* __restrict__ x, const double* __restrict__ part,
                                  const double* __restrict__ dscale, double* __restrict__ out,
                                  double* __restrict__ mag, const TrialFlags* __restrict__ flags) {
    const int row = blockIdx.x * 256 + threadIdx.x;
    if (row >= 9 * pb.C) return;
    const int a = row / 9, i = row - 9 * a;
    const int q = prob_of_cam(pb, a);
    if (prob_off(pb, q) || trial_failed(flags + q)) return;
    lambda = lam_of(pb, q, lambda);
    double s = 0.0, sm = 0.0;
    for (int c = pb.cam_chunk_off[a]; c < pb.cam_chunk_off[a + 1]; ++c) {
        s += part[18 * static_cast<size_t>(c) + i];
        sm += part[18 * static_cast<size_t>(c) + 9 + i];
    }
    double ux = 0.0, um = 0.0;
    if (x) {
        const double* Ua = U + 81 * static_cast<size_t>(a) + 9 * i;
        const double* xa = x + 9 * static_cast<size_t>(a);
        ux = lambda * damp_diag(Ua[i]) * xa[i];
        um = fabs(ux);
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            ux += Ua[j] * xa[j];
            um += fabs(Ua[j] * xa[j]);
        }
    }
    const double r = bc[row] - ux - s;
    const double sc = dscale ? dscale[row] : 1.0;
    out[row] = r * sc;
    if (mag) mag[row] = (fabs(bc[row]) + um + sm) * sc;
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
                            const double* x, double* z, double* rot, double* part, const double* dscale, double* out,
                            double* mag, const TrialFlags* flags) {
    const int grid = (pb.num_units + 3) / 4;
    if (pb.C > 0) hipLaunchKernelGGL(k_rot_terms, dim3((pb.C + 255) / 256), dim3(256), 0, st, pb.C, cams, rot);
    if (grid > 0)
        hipLaunchKernelGGL(k_refine_points, dim3(grid), dim3(256), 0, st, pb, cams, rot, pts, Vinv6, VB, x, z,
                           flags);
    if (pb.num_chunks > 0)
        hipLaunchKernelGGL(k_refine_cam_chunks, dim3(pb.num_chunks), dim3(256), 0, st, pb, cams, rot, pts, z, part,
                           flags);
    if (pb.C > 0)
        hipLaunchKernelGGL(k_refine_cam_rows, dim3((9 * pb.C + 255) / 256), dim3(256), 0, st, pb, U, bc, lambda, x,
                           part, dscale, out, mag, flags);
}

void launch_vec_axpy_scaled(hipStream_t st, int n, double* x, const double* d, const double* s) {
    if (n > 0) hipLaunchKernelGGL(k_vec_axpy_scaled, dim3((n + 255) / 256), dim3(256), 0, st, n, x, d, s);
}

}  // namespace gba_amd
