// k_cg.hip -- factor-free refinement of the mixed-precision reduced camera
// system (the reference solves it once per trial with Eigen::LLT,
// src/lm.cu:597-604).
//
// The mixed form stores the diagonally scaled Schur complement A = D S D in
// FP32 (row-major, leading dimension ld = npad; the lower 64x64 tiles are
// valid, and each 9x9 camera block on the diagonal is written in both
// triangles) and solves A x = b in FP64 by conjugate gradients.  Here the
// preconditioner is block Jacobi over the camera blocks, M^-1 =
// blockdiag(A_cc^-1) (padding rows: identity), inverted in FP64 once per
// trial.  It needs tens of steps where the FP32 Cholesky preconditioner needs
// two or three, but a step is one pass over the 64x64 tiles of A while the
// factor is a chain of dependent 64-column steps.
//
// The iteration is the Chronopoulos-Gear form of preconditioned CG, which
// needs one matrix product and one round of inner products per step:
//   p_i = u_i + beta_{i-1} p_{i-1},   s_i = w_i + beta_{i-1} s_{i-1}  (= A p_i)
//   x_{i+1} = x_i + alpha_i p_i,      r_{i+1} = r_i - alpha_i s_i
//   u_{i+1} = M^-1 r_{i+1},           w_{i+1} = A u_{i+1}
//   gamma_{i+1} = r_{i+1}.u_{i+1},    delta_{i+1} = w_{i+1}.u_{i+1}
//   beta_i = gamma_{i+1} / gamma_i
//   alpha_{i+1} = gamma_{i+1} / (delta_{i+1} - beta_i gamma_{i+1} / alpha_i)
// with x_0 = 0, r_0 = b, alpha_0 = gamma_0 / delta_0, beta_{-1} = 0 (in exact
// arithmetic the same iterates as the textbook PCG; Chronopoulos and Gear,
// J. Comput. Appl. Math. 25 (1989)).  Two launches per step:
//   k_cg_matvec: w = A u.  A workgroup takes a run of tiles (i, j0 .. j1-1)
//                of one tile row: the run's row partial sum_j A_ij u_j, each
//                tile's column partial A_ij^T u_i (strictly lower part of a
//                diagonal tile), and its share of delta = u.w, which every
//                tile contributes to as u_i.(A_ij u_j) + u_j.(A_ij^T u_i).
//   k_cg_vector: one workgroup per seven cameras (63 rows): alpha and beta
//                from the partial sums, w of its rows from the tile
//                partials, the vector updates, u = M^-1 r and its partials
//                of gamma and r.r.
// Each launch starts by reducing the partials it needs in a fixed order (every
// workgroup the same sums), so the results do not depend on scheduling and no
// workgroup waits on another.  The stopping test (r.r <= tol^2 b.b, the step
// cap, a failed trial) runs at the start of k_cg_matvec; gamma or the alpha
// denominator not > 0 stops the iteration in k_cg_vector.  Once stopped, every
// later launch of the batch returns at once, so the host enqueues steps in
// batches and synchronises once per batch.
#include "device_common.hpp"
#include "launchers.hpp"

namespace gba_amd {

namespace {

constexpr int kCgRows = 63;  // rows per k_cg_vector workgroup: seven cameras

// Device state of one solve (doubles):
//   [0] done, [1] stopped on a breakdown, [2] steps (x_k current), [3] b.b,
//   [4] r.r of the current iterate, [8 + 4 q] alpha, [9 + 4 q] gamma of the
//   step with parity q.
constexpr int kStDone = 0, kStStop = 1, kStSteps = 2, kStBB = 3, kStRR = 4, kStSlot = 8;

__device__ __forceinline__ bool trial_failed(const TrialFlags* f) {
    return f->nonspd || f->chol_fail || f->solve_err;
}

// Camera block c of A (9x9; the lower triangle is read) -> FP64 Cholesky ->
// inverse, stored in full (81 values, row-major).  A pivot that is not > 0
// raises chol_fail: the scaled S is then not numerically SPD and the trial
// escalates lambda as a failed factorisation does.
__global__ __launch_bounds__(64) void k_cg_block_inverse(const float* __restrict__ A, int ld, int ncam,
                                                         double* __restrict__ minv, TrialFlags* flags) {
    const int c = blockIdx.x * 64 + threadIdx.x;
    if (c >= ncam) return;
    const float* blk = A + static_cast<size_t>(9 * c) * ld + 9 * c;
    double l[45];  // packed lower: (i, j), j <= i at i (i + 1) / 2 + j
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) l[i * (i + 1) / 2 + j] = static_cast<double>(blk[static_cast<size_t>(i) * ld + j]);
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        double d = l[j * (j + 1) / 2 + j];
#pragma unroll
        for (int k = 0; k < j; ++k) d -= l[j * (j + 1) / 2 + k] * l[j * (j + 1) / 2 + k];
        ok = ok && (d > 0.0);
        const double ljj = sqrt(d > 0.0 ? d : 1.0);
        const double rinv = 1.0 / ljj;
        l[j * (j + 1) / 2 + j] = ljj;
#pragma unroll
        for (int i = j + 1; i < 9; ++i) {
            double v = l[i * (i + 1) / 2 + j];
#pragma unroll
            for (int k = 0; k < j; ++k) v -= l[i * (i + 1) / 2 + k] * l[j * (j + 1) / 2 + k];
            l[i * (i + 1) / 2 + j] = v * rinv;
        }
    }
    // m = L^-1 (lower), column by column
    double m[45];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
#pragma unroll
        for (int i = j; i < 9; ++i) {
            double v = (i == j) ? 1.0 : 0.0;
#pragma unroll
            for (int k = j; k < i; ++k) v -= l[i * (i + 1) / 2 + k] * m[k * (k + 1) / 2 + j];
            m[i * (i + 1) / 2 + j] = v / l[i * (i + 1) / 2 + i];
        }
    }
    // A_cc^-1 = m^T m
    double* out = minv + 81 * static_cast<size_t>(c);
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j) {
            double v = 0.0;
#pragma unroll
            for (int k = i; k < 9; ++k) v += m[k * (k + 1) / 2 + i] * m[k * (k + 1) / 2 + j];
            out[9 * i + j] = v;
            out[9 * j + i] = v;
        }
    if (!ok) atomicOr(&flags->chol_fail, 1);
}

// Sum over q < n of v[stride q] by thread t of the workgroup: q = t, t + nt,
// ... in order, eight loads in flight at a time.
__device__ __forceinline__ double strided_partial(const double* __restrict__ v, int n, int stride, int t, int nt) {
    double a = 0.0;
    for (int q0 = t; q0 < n; q0 += 8 * nt) {
        double x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int q = q0 + u * nt;
            x[u] = q < n ? v[static_cast<size_t>(stride) * q] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) a += x[u];
    }
    return a;
}

// Matrix-product workgroup: w = A u over a run of tiles (see the top).  The
// stopping test uses one wavefront (the r.r partials of k_cg_vector); its
// decision reaches the workgroup through LDS at the first barrier, which also
// publishes the first tile.
__global__ __launch_bounds__(256) void k_cg_matvec(CgArgs g, int k) {
    __shared__ float s_a[kTile][kTile + 1];
    __shared__ double s_ui[kTile], s_uj[kTile];
    __shared__ double s_col[4][kTile];
    __shared__ double s_tmp[4];
    __shared__ int s_done;
    const int tid = threadIdx.x;
    const int2 wi = g.wgtab[blockIdx.x];
    const int i = wi.x, j0 = wi.y * g.tpw, j1 = min(j0 + g.tpw, i + 1);
    float tmp[16];
    {
        const float* src = g.A + static_cast<size_t>(i * kTile) * g.ld + j0 * kTile;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = tid + 256 * q;
            tmp[q] = src[static_cast<size_t>(e >> 6) * g.ld + (e & 63)];
        }
    }
    if (tid < kTile) {
        s_ui[tid] = g.u[i * kTile + tid];
    } else if (tid < 2 * kTile) {
        s_uj[tid - kTile] = g.u[j0 * kTile + tid - kTile];
    } else if (tid < 3 * kTile) {
        // stopping test on r.r of the current iterate x_k (wavefront 2)
        const double* gp = g.gpart + 2 * static_cast<size_t>(g.nwg_v) * (k & 1);
        const double rr = wave_sum(strided_partial(gp + 1, g.nwg_v, 2, tid - 2 * kTile, kTile));
        if (tid == 2 * kTile) {
            // exit decisions come from one thread through LDS, never from a
            // per-wavefront read of state another workgroup may be writing
            const bool stopped = g.st[kStDone] != 0.0;
            const double bb = (k == 0) ? rr : g.st[kStBB];
            const bool done = stopped || trial_failed(g.flags) || !(bb > 0.0) || !isfinite(rr) ||
                              rr <= g.tol2 * bb || k >= g.max_it;
            s_done = done ? 1 : 0;
            if (blockIdx.x == 0 && !stopped) {
                if (k == 0) g.st[kStBB] = bb;
                g.st[kStSteps] = k;
                g.st[kStRR] = rr;
                if (done) g.st[kStDone] = 1.0;
                PassStatus* ps = g.status;
                ps->flags = *g.flags;
                ps->pcg_stop = 0;
                ps->pcg_done = done ? 1 : 0;
                ps->pcg_it = k;
                ps->pcg_rr = rr;
                ps->pcg_bb = bb;
            }
        }
    }
    const int r4 = tid >> 2, c4 = 16 * (tid & 3);   // row partials: four threads per row
    const int cc = tid & 63, rg = 16 * (tid >> 6);  // column partials: one thread per column, 16 rows
    double rs = 0.0, dsum = 0.0;
    for (int j = j0; j < j1; ++j) {
        const bool diag = (j == i);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = tid + 256 * q;
            s_a[e >> 6][e & 63] = tmp[q];
        }
        if (j > j0 && tid < kTile) s_uj[tid] = g.u[j * kTile + tid];
        __syncthreads();
        if (s_done) return;
        if (j + 1 < j1) {
            const float* src = g.A + static_cast<size_t>(i * kTile) * g.ld + (j + 1) * kTile;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int e = tid + 256 * q;
                tmp[q] = src[static_cast<size_t>(e >> 6) * g.ld + (e & 63)];
            }
        }
#pragma unroll 4
        for (int c = c4; c < c4 + 16; ++c)
            rs += (!diag || c <= r4) ? static_cast<double>(s_a[r4][c]) * s_uj[c] : 0.0;
        double cs = 0.0;
#pragma unroll 4
        for (int r = rg; r < rg + 16; ++r)
            cs += (!diag || cc < r) ? static_cast<double>(s_a[r][cc]) * s_ui[r] : 0.0;
        s_col[tid >> 6][cc] = cs;
        __syncthreads();
        if (tid < kTile) {
            const double v = (s_col[0][tid] + s_col[1][tid]) + (s_col[2][tid] + s_col[3][tid]);
            g.cp[static_cast<size_t>(i * (i + 1) / 2 + j) * kTile + tid] = v;
            dsum += s_uj[tid] * v;
        }
        __syncthreads();
    }
    rs += __shfl_xor(rs, 1, kWave);
    rs += __shfl_xor(rs, 2, kWave);
    if ((tid & 3) == 0) {
        g.rp[static_cast<size_t>(blockIdx.x) * kTile + r4] = rs;
        dsum += s_ui[r4] * rs;
    }
    // delta share: threads < 64 hold the column terms, every fourth thread a
    // row term; fixed-order reduction
    dsum = wave_sum(dsum);
    if ((tid & 63) == 0) s_tmp[tid >> 6] = dsum;
    __syncthreads();
    if (tid == 0) g.dpart[blockIdx.x] = (s_tmp[0] + s_tmp[1]) + (s_tmp[2] + s_tmp[3]);
}

// k = -1: the start (x = 0, r = b, p = s = 0); k >= 0: step k.  Every load a
// step needs (the tile partials of w, the partial sums of gamma and delta,
// the vectors and the camera blocks' inverses) is issued before the first
// barrier.
__global__ __launch_bounds__(256) void k_cg_vector(CgArgs g, int k) {
    __shared__ double s_g[4], s_d[4];
    __shared__ double s_w[4][64];
    __shared__ double s_r[64];
    __shared__ double s_ab[2];
    __shared__ int s_brk;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int row0 = kCgRows * blockIdx.x;
    const int row = row0 + lane;
    const bool mine = lane < kCgRows && row < g.npad;
    const bool cam_row = mine && row < 9 * g.ncam;
    const int cl = lane / 9, ci = lane - 9 * cl;  // row0 is a multiple of 9
    // wavefront 0's operands of the vector update
    double vu = 0.0, vp = 0.0, vs = 0.0, vr = 0.0, mrow[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) mrow[q] = 0.0;
    if (w == 0 && mine) {
        if (k < 0) {
            vr = g.b[row];
        } else {
            vu = g.u[row];
            vp = g.p[row];
            vs = g.s[row];
            vr = g.r[row];
        }
        if (cam_row) {
            const double* m = g.minv + 81 * static_cast<size_t>(row0 / 9 + cl) + 9 * ci;
#pragma unroll
            for (int q = 0; q < 9; ++q) mrow[q] = m[q];
        }
    }
    double alpha = 0.0, beta = 0.0;
    if (k < 0) {
        if (blockIdx.x == 0 && tid == 0) {
            g.st[kStDone] = 0.0;
            g.st[kStStop] = 0.0;
            g.st[kStSteps] = 0.0;
        }
    } else {
        // w of this workgroup's rows from the tile partials: wavefront w sums
        // the terms with index = w mod 4, the four sums are combined in order
        double acc = 0.0;
        if (mine) {
            const int I = row >> 6, rr = row & 63;
            const int nc = (I + g.tpw) / g.tpw;  // runs of tile row I
            const double* rpI = g.rp + static_cast<size_t>(g.rowstart[I]) * kTile + rr;
            // runs w, w + 4, ... of the nc: ceil((nc - w) / 4) terms
            acc = strided_partial(rpI + static_cast<size_t>(w) * kTile, nc > w ? (nc - w + 3) / 4 : 0,
                                  4 * kTile, 0, 1);
            // column partials of tiles (J, I), J = I + w, I + w + 4, ...
            double x[8];
            for (int J0 = I + w; J0 < g.nblk; J0 += 32) {
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int J = J0 + 4 * u;
                    x[u] = J < g.nblk ? g.cp[static_cast<size_t>(J * (J + 1) / 2 + I) * kTile + rr] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) acc += x[u];
            }
        }
        s_w[w][lane] = acc;
        // gamma_k and delta_k partial sums: each thread its strided share,
        // then wavefront sums and the four wavefronts in order
        const double* gp = g.gpart + 2 * static_cast<size_t>(g.nwg_v) * (k & 1);
        const double ga = wave_sum(strided_partial(gp, g.nwg_v, 2, tid, 256));
        const double da = wave_sum(strided_partial(g.dpart, g.nwg_m, 1, tid, 256));
        if (lane == 0) {
            s_g[w] = ga;
            s_d[w] = da;
        }
        __syncthreads();
        if (tid == 0) {
            const double gam = (s_g[0] + s_g[1]) + (s_g[2] + s_g[3]);
            const double del = (s_d[0] + s_d[1]) + (s_d[2] + s_d[3]);
            const bool stopped = g.st[kStDone] != 0.0;
            double al, be = 0.0, den = del;
            if (k > 0) {
                const double ap = g.st[kStSlot + 4 * ((k - 1) & 1)];
                const double gprev = g.st[kStSlot + 4 * ((k - 1) & 1) + 1];
                be = gam / gprev;
                den = del - be * gam / ap;
            }
            al = gam / den;
            const bool brk = !(gam > 0.0) || !(den > 0.0) || !isfinite(al) || !isfinite(be);
            s_brk = (stopped || brk) ? 1 : 0;
            s_ab[0] = al;
            s_ab[1] = be;
            if (blockIdx.x == 0 && !stopped) {
                if (brk) {
                    g.st[kStStop] = 1.0;
                    g.st[kStDone] = 1.0;
                    g.status->pcg_stop = 1;
                    g.status->pcg_done = 1;
                } else {
                    g.st[kStSlot + 4 * (k & 1)] = al;
                    g.st[kStSlot + 4 * (k & 1) + 1] = gam;
                }
            }
        }
        __syncthreads();
        if (s_brk) return;
        alpha = s_ab[0];
        beta = s_ab[1];
    }
    if (w != 0) return;
    double r = 0.0, u = 0.0;
    if (mine) {
        if (k < 0) {
            r = vr;
            g.x[row] = 0.0;
            g.p[row] = 0.0;
            g.s[row] = 0.0;
        } else {
            const double wv = (s_w[0][lane] + s_w[1][lane]) + (s_w[2][lane] + s_w[3][lane]);
            const double pv = vu + beta * vp;
            const double sv = wv + beta * vs;
            g.p[row] = pv;
            g.s[row] = sv;
            g.x[row] += alpha * pv;
            r = vr - alpha * sv;
        }
        g.r[row] = r;
    }
    s_r[lane] = r;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (mine) {
        if (cam_row) {
#pragma unroll
            for (int q = 0; q < 9; ++q) u += mrow[q] * s_r[9 * cl + q];
        } else {
            u = r;  // padding: identity
        }
        g.u[row] = u;
    }
    const double gsum = wave_sum(r * u), rsum = wave_sum(r * r);
    if (lane == 0) {
        double* gp = g.gpart + 2 * static_cast<size_t>(g.nwg_v) * ((k + 1) & 1);
        gp[2 * blockIdx.x] = gsum;
        gp[2 * blockIdx.x + 1] = rsum;
    }
}

}  // namespace

int cg_vector_groups(int npad) { return (npad + kCgRows - 1) / kCgRows; }

void launch_cg_block_inverse(hipStream_t st, const float* A, int ld, int ncam, double* minv, TrialFlags* flags) {
    if (ncam <= 0) return;
    hipLaunchKernelGGL(k_cg_block_inverse, dim3((ncam + 63) / 64), dim3(64), 0, st, A, ld, ncam, minv, flags);
}

void launch_cg_start(hipStream_t st, const CgArgs& g) {
    hipLaunchKernelGGL(k_cg_vector, dim3(g.nwg_v), dim3(256), 0, st, g, -1);
}

void launch_cg_step(hipStream_t st, const CgArgs& g, int k) {
    hipLaunchKernelGGL(k_cg_matvec, dim3(g.nwg_m), dim3(256), 0, st, g, k);
    hipLaunchKernelGGL(k_cg_vector, dim3(g.nwg_v), dim3(256), 0, st, g, k);
}

}  // namespace gba_amd
