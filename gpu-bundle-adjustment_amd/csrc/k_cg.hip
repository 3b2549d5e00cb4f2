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
// Deflation (when the basis is given): the preconditioned operator of these
// systems has seven eigenvalues of the order of lambda, one per direction of
// the similarity gauge (a rotation, translation or scaling of the whole scene
// moves no projection, so its camera part is a null vector of the undamped S),
// and the rest in a narrow band (about [0.25, 1.1] at C=356).  CG then spends
// most of its steps on those seven.  With Z the gauge basis (D^-1 times the
// camera part, columns normalised) and E = Z^T A Z, deflated CG (Saad, Yeung,
// Erhel and Guyomarc'h, SIAM J. Sci. Comput. 21 (2000)) starts from
// x_0 = Z E^-1 Z^T b, so that Z^T r_0 = 0, and takes the directions
// u~ = u - Z mu, mu = E^-1 (A Z)^T u, A-orthogonal to Z.  In the form above
// that is p_i = u~_i + beta p_{i-1}, s_i = w~_i + beta s_{i-1} with
// w~ = w - (A Z) mu, and delta~ = u~.w~ = u.w - g.E^-1 g with g = (A Z)^T u,
// whose partial sums the vector update forms with gamma's.  gamma = r.u~ =
// r.u because Z^T r = 0.  A Z and E are formed once per trial.
// Each launch starts by reducing the partials it needs in a fixed order (every
// workgroup the same sums), so the results do not depend on scheduling and no
// workgroup waits on another.  The stopping test (r.r <= tol^2 b.b, the step
// cap, a failed trial) runs at the start of k_cg_matvec; gamma or the alpha
// denominator not > 0 stops the iteration in k_cg_vector.  Once stopped, every
// later launch of the batch returns at once, so the host enqueues steps in
// batches and synchronises once per batch.
#include "device_common.hpp"
#include "launchers.hpp"
#include "snavely_model.hpp"

namespace gba_amd {

namespace {

constexpr int kCgRows = 63;  // rows per k_cg_vector workgroup: seven cameras

// Device state of one solve (doubles):
//   [0] done, [1] stopped on a breakdown, [2] steps (x_k current), [3] b.b,
//   [4] r.r of the current iterate, [8 + 4 q] alpha, [9 + 4 q] gamma of the
//   step with parity q.
constexpr int kStDone = 0, kStStop = 1, kStSteps = 2, kStBB = 3, kStRR = 4, kStSlot = 8;
// Partial sums per vector workgroup (kCgPartials, parity-double-buffered):
// [0] gamma, [1] r.r, [2 + j] (A Z)^T u, column j of the gauge basis.
constexpr int kPgGam = 0, kPgRR = 1, kPgZ = 2;
constexpr int kGauge = 7;  // rotation (3), translation (3), scale

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
        const double* gp = g.gpart + kCgPartials * static_cast<size_t>(g.nwg_v) * (k & 1);
        const double rr = wave_sum(strided_partial(gp + kPgRR, g.nwg_v, kCgPartials, tid - 2 * kTile, kTile));
        // b.b: r_0 = b unless the deflated start moved x_0 off zero
        const double bb0 = (k == 0 && g.zd)
                               ? wave_sum(strided_partial(g.zbpart + kGauge, g.nwg_v, 8, tid - 2 * kTile, kTile))
                               : rr;
        if (tid == 2 * kTile) {
            // exit decisions come from one thread through LDS, never from a
            // per-wavefront read of state another workgroup may be writing
            const bool stopped = g.st[kStDone] != 0.0;
            const double bb = (k == 0) ? bb0 : g.st[kStBB];
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

// k = -2: the deflated start's partial sums of Z^T b and b.b; k = -1: the
// start (x = 0, r = b, p = s = 0, or with the deflation x = Z mu_0,
// r = b - (A Z) mu_0, mu_0 = E^-1 Z^T b); k >= 0: step k.  Every load a step
// needs (the tile partials of w, the partial sums of gamma, delta and g, the
// vectors, the camera blocks' inverses and the basis rows) is issued before
// the first barrier.
__global__ __launch_bounds__(256) void k_cg_vector(CgArgs g, int k) {
    __shared__ double s_g[4], s_d[4];
    __shared__ double s_z[4][kGauge];
    __shared__ double s_mu[kGauge];
    __shared__ double s_einv[kGauge * kGauge];
    __shared__ double s_w[4][64];
    __shared__ double s_r[64];
    __shared__ double s_ab[2];
    __shared__ int s_brk;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int row0 = kCgRows * blockIdx.x;
    const int row = row0 + lane;
    const bool mine = lane < kCgRows && row < g.npad;
    const bool cam_row = mine && row < 9 * g.ncam;
    const bool defl = g.zd != nullptr;
    const int cl = lane / 9, ci = lane - 9 * cl;  // row0 is a multiple of 9
    if (k == -2) {
        if (w != 0) return;
        const double bv = mine ? g.b[row] : 0.0;
#pragma unroll
        for (int j = 0; j < kGauge; ++j) {
            const double t = wave_sum(mine ? g.zd[8 * static_cast<size_t>(row) + j] * bv : 0.0);
            if (lane == 0) g.zbpart[8 * static_cast<size_t>(blockIdx.x) + j] = t;
        }
        const double t = wave_sum(bv * bv);
        if (lane == 0) g.zbpart[8 * static_cast<size_t>(blockIdx.x) + kGauge] = t;
        return;
    }
    if (defl && tid < kGauge * kGauge) s_einv[tid] = g.einv[tid];
    // wavefront 0's operands of the vector update
    double vu = 0.0, vp = 0.0, vs = 0.0, vr = 0.0, mrow[9], zr[kGauge], azr[kGauge];
#pragma unroll
    for (int q = 0; q < 9; ++q) mrow[q] = 0.0;
#pragma unroll
    for (int j = 0; j < kGauge; ++j) zr[j] = azr[j] = 0.0;
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
        if (defl) {
#pragma unroll
            for (int j = 0; j < kGauge; ++j) {
                zr[j] = g.zd[8 * static_cast<size_t>(row) + j];
                azr[j] = g.azd[8 * static_cast<size_t>(row) + j];
            }
        }
    }
    double alpha = 0.0, beta = 0.0;
    if (k < 0) {
        if (blockIdx.x == 0 && tid == 0) {
            g.st[kStDone] = 0.0;
            g.st[kStStop] = 0.0;
            g.st[kStSteps] = 0.0;
        }
        if (defl) {
            // mu_0 = E^-1 Z^T b from every workgroup's partials, in a fixed order
#pragma unroll
            for (int j = 0; j < kGauge; ++j) {
                const double t = wave_sum(strided_partial(g.zbpart + j, g.nwg_v, 8, tid, 256));
                if (lane == 0) s_z[w][j] = t;
            }
            __syncthreads();
            if (tid < kGauge) {
                double m = 0.0;
                for (int h = 0; h < kGauge; ++h)
                    m += s_einv[kGauge * tid + h] * ((s_z[0][h] + s_z[1][h]) + (s_z[2][h] + s_z[3][h]));
                s_mu[tid] = m;
            }
            __syncthreads();
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
        // gamma_k, delta_k (and g_k) partial sums: each thread its strided
        // share, then wavefront sums and the four wavefronts in order
        // (one pass over the vector workgroups' records)
        const double* gp = g.gpart + kCgPartials * static_cast<size_t>(g.nwg_v) * (k & 1);
        double ga = 0.0, gz[kGauge];
#pragma unroll
        for (int j = 0; j < kGauge; ++j) gz[j] = 0.0;
        for (int q = tid; q < g.nwg_v; q += 256) {
            const double* rec = gp + kCgPartials * static_cast<size_t>(q);
            ga += rec[kPgGam];
            if (defl) {
#pragma unroll
                for (int j = 0; j < kGauge; ++j) gz[j] += rec[kPgZ + j];
            }
        }
        double red[2 + kGauge];
        red[0] = ga;
        red[1] = strided_partial(g.dpart, g.nwg_m, 1, tid, 256);
#pragma unroll
        for (int j = 0; j < kGauge; ++j) red[2 + j] = gz[j];
        if (defl)
            wave_sum_n<2 + kGauge>(red);
        else
            wave_sum_n<2>(red);
        if (lane == 0) {
            s_g[w] = red[0];
            s_d[w] = red[1];
            if (defl) {
#pragma unroll
                for (int j = 0; j < kGauge; ++j) s_z[w][j] = red[2 + j];
            }
        }
        __syncthreads();
        if (tid == 0) {
            const double gam = (s_g[0] + s_g[1]) + (s_g[2] + s_g[3]);
            double del = (s_d[0] + s_d[1]) + (s_d[2] + s_d[3]);
            if (defl) {
                // mu = E^-1 g, delta~ = delta - g.mu
                double gt[kGauge], gmu = 0.0;
#pragma unroll
                for (int j = 0; j < kGauge; ++j) gt[j] = (s_z[0][j] + s_z[1][j]) + (s_z[2][j] + s_z[3][j]);
                for (int j = 0; j < kGauge; ++j) {
                    double m = 0.0;
#pragma unroll
                    for (int h = 0; h < kGauge; ++h) m += s_einv[kGauge * j + h] * gt[h];
                    s_mu[j] = m;
                    gmu += gt[j] * m;
                }
                del -= gmu;
            }
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
            double xv = 0.0;
            r = vr;
            if (defl) {
#pragma unroll
                for (int j = 0; j < kGauge; ++j) {
                    xv += zr[j] * s_mu[j];
                    r -= azr[j] * s_mu[j];
                }
            }
            g.x[row] = xv;
            g.p[row] = 0.0;
            g.s[row] = 0.0;
        } else {
            double wv = (s_w[0][lane] + s_w[1][lane]) + (s_w[2][lane] + s_w[3][lane]);
            if (defl) {
                // u~ = u - Z mu, w~ = w - (A Z) mu
#pragma unroll
                for (int j = 0; j < kGauge; ++j) {
                    vu -= zr[j] * s_mu[j];
                    wv -= azr[j] * s_mu[j];
                }
            }
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
    // gamma, r.r and g = (A Z)^T u partials of this workgroup
    double red[2 + kGauge];
    red[kPgGam] = r * u;
    red[kPgRR] = r * r;
#pragma unroll
    for (int j = 0; j < kGauge; ++j) red[kPgZ + j] = azr[j] * u;
    if (defl)
        wave_sum_n<2 + kGauge>(red);
    else
        wave_sum_n<2>(red);
    double* gp = g.gpart + kCgPartials * (static_cast<size_t>(g.nwg_v) * ((k + 1) & 1) + blockIdx.x);
    if (lane == 0) {
        gp[kPgGam] = red[kPgGam];
        gp[kPgRR] = red[kPgRR];
        if (defl) {
#pragma unroll
            for (int j = 0; j < kGauge; ++j) gp[kPgZ + j] = red[kPgZ + j];
        }
    }
}

// Gauge basis rows of the scaled system and the partial sums of its column
// norms: one lane per row, 63 rows (seven cameras) per workgroup.  Camera c
// with angle-axis a and translation t moves under a rotation Q = exp([th]x)
// of the scene as R' = R Q^T, i.e. a' = a - J_r(a)^-1 th, under a
// translation tau as t' = t - R tau, and under a scaling s as t' = s t (the
// projection -P.xy / P.z does not see the scale), so the rows of camera c
// are: angle-axis rows -J_r^-1(a) (rotation columns), translation rows -R(a)
// (translation columns) and t (scale column), focal and distortion rows zero,
// each divided by the row's D (A = D S D).  J_r^-1 = I + K/2 + c K^2 with
// K = [a]x and c = 1/th^2 - (1 + cos th) / (2 th sin th) = (1 - A / (2B)) / th^2
// (A = sin th / th, B = (1 - cos th) / th^2; near 0 the series
// 1/12 + th^2/720 + th^4/30240); R = I + A K + B K^2.
__device__ __forceinline__ double skew_entry(const double* a, int i, int k) {
    // [a]x: (0, -a2, a1; a2, 0, -a0; -a1, a0, 0)
    if (i == k) return 0.0;
    const int m = 3 - i - k;  // the third index
    const double v = a[m];
    return ((i + 1) % 3 == k) ? -v : v;
}

__global__ __launch_bounds__(64) void k_cg_gauge(int npad, int ncam, const double* __restrict__ cams,
                                                 const double* __restrict__ dscale, double* __restrict__ zraw,
                                                 double* __restrict__ npart) {
    const int lane = threadIdx.x, row = kCgRows * blockIdx.x + lane;
    const bool mine = lane < kCgRows && row < npad;
    double z[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) z[j] = 0.0;
    if (mine && row < 9 * ncam) {
        const int c = row / 9, i = row - 9 * c;
        const double* cm = cams + 9 * static_cast<size_t>(c);
        const double a[3] = {cm[0], cm[1], cm[2]};
        const double th2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        const RotTerms rt = rot_terms(a);
        if (i < 3) {
            double cr;
            if (rt.first_order || th2 < 1e-4)
                cr = 1.0 / 12.0 + th2 / 720.0 + th2 * th2 / 30240.0;
            else
                cr = (1.0 - rt.A / (2.0 * rt.B)) / th2;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                z[k] = -((i == k ? 1.0 : 0.0) + 0.5 * skew_entry(a, i, k) + cr * (a[i] * a[k] - (i == k ? th2 : 0.0)));
        } else if (i < 6) {
            const int ii = i - 3;
#pragma unroll
            for (int k = 0; k < 3; ++k)
                z[3 + k] = -((ii == k ? 1.0 : 0.0) + rt.A * skew_entry(a, ii, k) +
                             rt.B * (a[ii] * a[k] - (ii == k ? th2 : 0.0)));
            z[6] = cm[3 + ii];
        }
        const double inv = 1.0 / dscale[row];
#pragma unroll
        for (int j = 0; j < kGauge; ++j) z[j] *= inv;
    }
    if (mine) {
#pragma unroll
        for (int j = 0; j < 8; ++j) zraw[8 * static_cast<size_t>(row) + j] = z[j];
    }
#pragma unroll
    for (int j = 0; j < kGauge; ++j) {
        const double t = wave_sum(z[j] * z[j]);
        if (lane == 0) npart[8 * static_cast<size_t>(blockIdx.x) + j] = t;
    }
}

// A Z (unnormalised basis), partial over a run of kGaugeRun tiles:
// workgroup (run c, tile row I) takes tiles J = kGaugeRun c ... of row I
// (A_IJ = A_JI^T above the diagonal, the diagonal tile's lower triangle
// mirrored), the basis rows of tile J staged in LDS; the next tile and basis
// rows are loaded while the current ones are used.
constexpr int kGaugeRun = 4;
__global__ __launch_bounds__(256) void k_cg_gauge_product(const float* __restrict__ A, int ld, int nblk,
                                                          const double* __restrict__ zraw,
                                                          double* __restrict__ apart) {
    __shared__ float s_a[kTile][kTile + 1];
    __shared__ double s_zj[kTile][8];
    const int tid = threadIdx.x, I = blockIdx.y, run = blockIdx.x;
    const int J0 = kGaugeRun * run, J1 = min(J0 + kGaugeRun, nblk);
    auto tile = [&](int J) {
        return J <= I ? A + static_cast<size_t>(I * kTile) * ld + J * kTile
                      : A + static_cast<size_t>(J * kTile) * ld + I * kTile;
    };
    float tmp[16];
    double zt[2];
    {
        const float* src = tile(J0);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = tid + 256 * q;
            tmp[q] = src[static_cast<size_t>(e >> 6) * ld + (e & 63)];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) zt[h] = zraw[8 * static_cast<size_t>(J0) * kTile + tid + 256 * h];
    }
    const int r = tid >> 2, c0 = 16 * (tid & 3);
    double acc[kGauge];
#pragma unroll
    for (int j = 0; j < kGauge; ++j) acc[j] = 0.0;
    for (int J = J0; J < J1; ++J) {
        __syncthreads();  // the previous tile is consumed
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = tid + 256 * q;
            s_a[e >> 6][e & 63] = tmp[q];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int t = tid + 256 * h;
            s_zj[t >> 3][t & 7] = zt[h];
        }
        __syncthreads();
        if (J + 1 < J1) {
            const float* src = tile(J + 1);
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int e = tid + 256 * q;
                tmp[q] = src[static_cast<size_t>(e >> 6) * ld + (e & 63)];
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) zt[h] = zraw[8 * static_cast<size_t>(J + 1) * kTile + tid + 256 * h];
        }
        for (int c = c0; c < c0 + 16; ++c) {
            float av;
            if (J < I)
                av = s_a[r][c];
            else if (J > I)
                av = s_a[c][r];
            else
                av = c <= r ? s_a[r][c] : s_a[c][r];
            const double a = static_cast<double>(av);
#pragma unroll
            for (int j = 0; j < kGauge; ++j) acc[j] += a * s_zj[c][j];
        }
    }
    // the four column quarters of a row
#pragma unroll
    for (int j = 0; j < kGauge; ++j) {
        acc[j] += __shfl_xor(acc[j], 1, kWave);
        acc[j] += __shfl_xor(acc[j], 2, kWave);
    }
    if ((tid & 3) == 0) {
        double* out = apart + (static_cast<size_t>(I) * gridDim.x + run) * (kTile * 8) + 8 * r;
#pragma unroll
        for (int j = 0; j < kGauge; ++j) out[j] = acc[j];
        out[kGauge] = 0.0;
    }
}

// Tile row I: the column normalisation (every workgroup sums the norm
// partials in the same order), A Z from the runs' partials (in run order),
// both scaled, and the row's share of Z^T A Z.
__global__ __launch_bounds__(256) void k_cg_gauge_rows(int nruns, int nwg_v, const double* __restrict__ zraw,
                                                       const double* __restrict__ npart,
                                                       const double* __restrict__ apart, double* __restrict__ zd,
                                                       double* __restrict__ azd, double* __restrict__ epart) {
    __shared__ double s_zi[kTile][8], s_az[kTile][8];
    __shared__ double s_np[32][8];
    __shared__ double s_sc[8];
    const int tid = threadIdx.x, I = blockIdx.x;
    {
        const int j = tid & 7, part = tid >> 3;
        double sum = 0.0;
        for (int q = part; q < nwg_v; q += 32) sum += npart[8 * static_cast<size_t>(q) + j];
        s_np[part][j] = sum;
    }
    __syncthreads();
    if (tid < 8) {
        double sum = 0.0;
        for (int part = 0; part < 32; ++part) sum += s_np[part][tid];
        s_sc[tid] = (tid < kGauge && sum > 0.0) ? 1.0 / sqrt(sum) : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int t = tid + 256 * h, rr = t >> 3, j = t & 7;
        const double* ap = apart + static_cast<size_t>(I) * nruns * (kTile * 8) + t;
        double x[4], az = 0.0;
        for (int c0 = 0; c0 < nruns; c0 += 4) {
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = c0 + u < nruns ? ap[static_cast<size_t>(c0 + u) * (kTile * 8)] : 0.0;
#pragma unroll
            for (int u = 0; u < 4; ++u) az += x[u];
        }
        const size_t row = static_cast<size_t>(I) * kTile + rr;
        const double zv = zraw[8 * row + j] * s_sc[j];
        az *= s_sc[j];
        zd[8 * row + j] = zv;
        azd[8 * row + j] = az;
        s_zi[rr][j] = zv;
        s_az[rr][j] = az;
    }
    __syncthreads();
    if (tid < kGauge * kGauge) {
        const int a = tid / kGauge, b = tid - kGauge * a;
        double e = 0.0;
        for (int rr = 0; rr < kTile; ++rr) e += s_zi[rr][a] * s_az[rr][b];
        epart[static_cast<size_t>(I) * (kGauge * kGauge) + tid] = e;
    }
}

// E = Z^T A Z from the tile rows' partials (fixed order), symmetrised, and its
// inverse by Cholesky; a pivot not above 1e-12 of the largest diagonal entry
// (a rank-deficient basis, e.g. cameras that do not fix a scale) or a
// non-finite entry leaves einv zero, which turns the deflation off.
__global__ __launch_bounds__(64) void k_cg_gauge_finish(int nblk, const double* __restrict__ epart,
                                                        double* __restrict__ einv) {
    __shared__ double s_e[kGauge * kGauge];
    const int tid = threadIdx.x;
    if (tid < kGauge * kGauge) s_e[tid] = strided_partial(epart + tid, nblk, kGauge * kGauge, 0, 1);
    __syncthreads();
    if (tid != 0) return;
    double l[kGauge][kGauge], m[kGauge][kGauge];
    double dmax = 0.0;
#pragma unroll
    for (int i = 0; i < kGauge; ++i) dmax = fmax(dmax, s_e[(kGauge + 1) * i]);
    bool ok = dmax > 0.0 && isfinite(dmax);
#pragma unroll
    for (int j = 0; j < kGauge; ++j) {
        double d = s_e[(kGauge + 1) * j];
#pragma unroll
        for (int q = 0; q < j; ++q) d -= l[j][q] * l[j][q];
        ok = ok && isfinite(d) && d > 1e-12 * dmax;
        l[j][j] = sqrt(d > 0.0 ? d : 1.0);
#pragma unroll
        for (int i = j + 1; i < kGauge; ++i) {
            double v = 0.5 * (s_e[kGauge * i + j] + s_e[kGauge * j + i]);
#pragma unroll
            for (int q = 0; q < j; ++q) v -= l[i][q] * l[j][q];
            l[i][j] = v / l[j][j];
        }
    }
    // m = L^-1 (lower), then E^-1 = m^T m
#pragma unroll
    for (int j = 0; j < kGauge; ++j)
#pragma unroll
        for (int i = j; i < kGauge; ++i) {
            double v = (i == j) ? 1.0 : 0.0;
#pragma unroll
            for (int q = j; q < i; ++q) v -= l[i][q] * m[q][j];
            m[i][j] = v / l[i][i];
        }
#pragma unroll
    for (int i = 0; i < kGauge; ++i)
#pragma unroll
        for (int j = 0; j < kGauge; ++j) {
            double v = 0.0;
#pragma unroll
            for (int q = (i > j ? i : j); q < kGauge; ++q) v += m[q][i] * m[q][j];
            ok = ok && isfinite(v);
            einv[kGauge * i + j] = v;
        }
    if (!ok)
        for (int t = 0; t < kGauge * kGauge; ++t) einv[t] = 0.0;
}

}  // namespace

int cg_vector_groups(int npad) { return (npad + kCgRows - 1) / kCgRows; }

int cg_gauge_runs(int nblk) { return (nblk + kGaugeRun - 1) / kGaugeRun; }

void launch_cg_block_inverse(hipStream_t st, const float* A, int ld, int ncam, double* minv, TrialFlags* flags) {
    if (ncam <= 0) return;
    hipLaunchKernelGGL(k_cg_block_inverse, dim3((ncam + 63) / 64), dim3(64), 0, st, A, ld, ncam, minv, flags);
}

void launch_cg_deflation_setup(hipStream_t st, const CgArgs& g, const double* cams, const double* dscale,
                               const CgDeflWork& w) {
    hipLaunchKernelGGL(k_cg_gauge, dim3(g.nwg_v), dim3(64), 0, st, g.npad, g.ncam, cams, dscale, w.zraw, w.npart);
    const int nruns = cg_gauge_runs(g.nblk);
    hipLaunchKernelGGL(k_cg_gauge_product, dim3(nruns, g.nblk), dim3(256), 0, st, g.A, g.ld, g.nblk, w.zraw,
                       w.apart);
    hipLaunchKernelGGL(k_cg_gauge_rows, dim3(g.nblk), dim3(256), 0, st, nruns, g.nwg_v, w.zraw, w.npart, w.apart,
                       w.zd, w.azd, w.epart);
    hipLaunchKernelGGL(k_cg_gauge_finish, dim3(1), dim3(64), 0, st, g.nblk, w.epart, w.einv);
}

void launch_cg_start(hipStream_t st, const CgArgs& g) {
    if (g.zd) hipLaunchKernelGGL(k_cg_vector, dim3(g.nwg_v), dim3(256), 0, st, g, -2);
    hipLaunchKernelGGL(k_cg_vector, dim3(g.nwg_v), dim3(256), 0, st, g, -1);
}

void launch_cg_step(hipStream_t st, const CgArgs& g, int k) {
    hipLaunchKernelGGL(k_cg_matvec, dim3(g.nwg_m), dim3(256), 0, st, g, k);
    hipLaunchKernelGGL(k_cg_vector, dim3(g.nwg_v), dim3(256), 0, st, g, k);
}

}  // namespace gba_amd
