// k_heavy.hip -- the heavy points' share of the reduced camera system as one
// dense product on MFMA (the reference sums every observation pair of every
// point, src/lm.cu:515-559).
//
// A point seen by k cameras contributes k(k-1)/2 observation pairs to the
// off-diagonal blocks of S.  On the benchmark's camera graphs the few points
// with the heaviest tails hold most of the pairs (at C=356, the 494 points
// with k >= 128 -- 0.2% of the points -- hold 57% of them) and are seen by a
// large fraction of all cameras, so their contribution
//     sum_{p heavy} W_p V'_p^-1 W_p^T = Y Y^T,   Y = [.. D W_p L_p ..],
// L_p L_p^T = V'_p^-1, is a dense rank-3|H| update of S: Y is n x 3|H| with
// the 9x3 block D_a W_{a,p} L_p where camera a sees point p and zeros
// elsewhere.  The pair kernels and the diagonal-block kernel leave the heavy
// points out; here k_heavy_y forms Y once per trial (the structure is fixed,
// so only the observed blocks are rewritten) and k_heavy_syrk subtracts Y Y^T
// from the lower tiles with v_mfma_f64_16x16x4f64 (plain form) or
// v_mfma_f32_16x16x4f32 (mixed form, whose S is stored FP32).
#include "device_common.hpp"
#include "launchers.hpp"
#include "mfma_tile.hpp"

#include <cmath>

namespace gba_amd {

namespace {

// One workgroup per heavy point h (hpts[h] = p).  A thread takes an
// observation e of p that starts its camera's run (the point's observations
// are ordered by camera; a run longer than one is a camera observing the
// point repeatedly, whose pairs all land on the same blocks) and writes
// Y_{a,p} = D_a (sum over the run of W_e) L_p into rows 9a .. 9a + 8 of the
// columns 3h .. 3h + 2 of Yt, stored k-major (Yt[k * ld + row]).
template <typename WT, typename YT>
__global__ __launch_bounds__(256) void k_heavy_y(DevProblem pb, const int* __restrict__ hpts,
                                                 const WT* __restrict__ W, const double* __restrict__ Vinv6,
                                                 const double* __restrict__ dscale, YT* __restrict__ Yt, int ld) {
    const int h = blockIdx.x;
    const int p = hpts[h];
    const int e0 = pb.pt_off[p], e1 = pb.pt_off[p + 1];
    const double* vi = Vinv6 + 6 * static_cast<size_t>(p);
    // lower Cholesky factor of V'^-1 = (v00; v10 v11; v20 v21 v22); a
    // non-positive pivot (V' not SPD: the trial is failed by its flags
    // already) is taken as zero
    const double l00 = sqrt(fmax(vi[0], 0.0));
    const double i00 = l00 > 0.0 ? 1.0 / l00 : 0.0;
    const double l10 = vi[1] * i00, l20 = vi[3] * i00;
    const double l11 = sqrt(fmax(vi[2] - l10 * l10, 0.0));
    const double l21 = l11 > 0.0 ? (vi[4] - l20 * l10) / l11 : 0.0;
    const double l22 = sqrt(fmax(vi[5] - l20 * l20 - l21 * l21, 0.0));
    for (int e = e0 + static_cast<int>(threadIdx.x); e < e1; e += 256) {
        const int a = pb.scam[e];
        if (e > e0 && pb.scam[e - 1] == a) continue;
        double w[27];
        load_w(W, static_cast<size_t>(e), w);
        for (int f = e + 1; f < e1 && pb.scam[f] == a; ++f) {
            double w2[27];
            load_w(W, static_cast<size_t>(f), w2);
#pragma unroll
            for (int i = 0; i < 27; ++i) w[i] += w2[i];
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const int row = 9 * a + i;
            const double d = dscale ? dscale[row] : 1.0;
            const double y0 = (w[3 * i] * l00 + w[3 * i + 1] * l10 + w[3 * i + 2] * l20) * d;
            const double y1 = (w[3 * i + 1] * l11 + w[3 * i + 2] * l21) * d;
            const double y2 = (w[3 * i + 2] * l22) * d;
            Yt[static_cast<size_t>(3 * h) * ld + row] = static_cast<YT>(y0);
            Yt[static_cast<size_t>(3 * h + 1) * ld + row] = static_cast<YT>(y1);
            Yt[static_cast<size_t>(3 * h + 2) * ld + row] = static_cast<YT>(y2);
        }
    }
}

// S -= Y Y^T over the lower tiles, 128x128 per workgroup: block (I2, J2),
// I2 >= J2 (blockIdx = I2 (I2 + 1) / 2 + J2), covering the 64x64 tiles
// (2 I2 + qr, 2 J2 + qc); wavefront w = 2 qr + qc forms that tile as a 4x4
// set of 16x16 MFMA accumulators (a tile above the diagonal or past the last
// tile row is skipped).  k runs in blocks of 16: the block's rows of Yt for
// the 128 rows and the 128 columns (k-major, so 16 k rows of 128 consecutive
// entries: 8 per thread, vector loads) are staged in LDS once for the four
// wavefronts, the next block's loads issued before the current block's
// MFMAs.  Fragments: lane l supplies A[l % 16][l / 16] = Y[row][k0 + l / 16]
// and B[l / 16][l % 16] = Y[col][k0 + l / 16].  Yt has kp rows (a multiple of
// 16) of ldy entries (a multiple of 128, zero past the matrix).  A diagonal
// tile updates its lower triangle and the upper halves of the 9x9 camera
// blocks inside it (the layout every consumer of S reads); S2, when given,
// is a second copy receiving the same values.
constexpr int kSyrkPad = 8;  // LDS row padding (bank spread across k rows)
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_heavy_syrk(
    const T* __restrict__ Yt, int kp, int ldy, int nblk, int ld,
                                                    T* __restrict__ S, T* __restrict__ S2,
                                                    const TrialFlags* __restrict__ flags) {
    __shared__ T s_a[16][128 + kSyrkPad], s_b[16][128 + kSyrkPad];
    if (flags->nonspd) return;
    typedef typename Acc4<T>::type acc_t;
    const int t = blockIdx.x;
    int I2 = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((I2 + 1) * (I2 + 2) / 2 <= t) ++I2;
    while (I2 * (I2 + 1) / 2 > t) --I2;
    const int J2 = t - I2 * (I2 + 1) / 2;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int qr = w >> 1, qc = w & 1;
    const int ti = 2 * I2 + qr, tj = 2 * J2 + qc;  // this wavefront's tile
    const bool live = ti < nblk && tj < nblk && ti >= tj;
    // staging: thread -> (k row lk, 8 entries from column c8)
    const int lk = tid >> 4, c8 = 8 * (tid & 15);
    const T* ga = Yt + static_cast<size_t>(lk) * ldy + 128 * I2 + c8;
    const T* gb = Yt + static_cast<size_t>(lk) * ldy + 128 * J2 + c8;
    T ra[8], rb[8];
    auto fetch = [&](int k0) {
        const size_t o = static_cast<size_t>(k0) * ldy;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            ra[u] = ga[o + u];
            rb[u] = gb[o + u];
        }
    };
    acc_t acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
    const int kk = lane >> 4, ii = lane & 15;
    fetch(0);
    for (int k0 = 0; k0 < kp; k0 += 16) {
        __syncthreads();  // the previous block is consumed
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            s_a[lk][c8 + u] = ra[u];
            s_b[lk][c8 + u] = rb[u];
        }
        __syncthreads();
        if (k0 + 16 < kp) fetch(k0 + 16);
        if (live) {
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                T av[4], bv[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    av[a] = s_a[4 * s + kk][64 * qr + 16 * a + ii];
                    bv[a] = s_b[4 * s + kk][64 * qc + 16 * a + ii];
                }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) acc[a][b] = mfma16(av[a], bv[b], acc[a][b]);
            }
        }
    }
    if (!live) return;
    const int g = lane >> 4, c16 = lane & 15;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = 16 * a + acc_row<T>(g, q), c = 16 * b + c16;
                const int R = kTile * ti + r, Cc = kTile * tj + c;
                if (ti == tj && r < c && R / 9 != Cc / 9) continue;
                const size_t o = static_cast<size_t>(R) * ld + Cc;
                S[o] = S[o] - acc[a][b][q];
                if (S2) S2[o] = S2[o] - acc[a][b][q];
            }
}

}  // namespace

void launch_heavy_y(hipStream_t st, const DevProblem& pb, const int* hpts, int nh, const double* W64,
                    const float* W32, const double* Vinv6, const double* dscale, double* Yt64, float* Yt32,
                    int ld) {
    if (nh <= 0) return;
    if (W32 && Yt32)
        hipLaunchKernelGGL((k_heavy_y<float, float>), dim3(nh), dim3(256), 0, st, pb, hpts, W32, Vinv6, dscale, Yt32,
                           ld);
    else if (W64 && Yt64)
        hipLaunchKernelGGL((k_heavy_y<double, double>), dim3(nh), dim3(256), 0, st, pb, hpts, W64, Vinv6, dscale,
                           Yt64, ld);
}

void launch_heavy_syrk(hipStream_t st, int nblk, int kp, int ldy, int ld, const double* Yt64, const float* Yt32,
                       double* S64, float* S32, float* S32b, const TrialFlags* flags) {
    const int n2 = (nblk + 1) / 2;
    const int nbk = n2 * (n2 + 1) / 2;
    if (nbk <= 0 || kp <= 0) return;
    if (Yt32)
        hipLaunchKernelGGL(k_heavy_syrk<float>, dim3(nbk), dim3(256), 0, st, Yt32, kp, ldy, nblk, ld, S32, S32b,
                           flags);
    else
        hipLaunchKernelGGL(k_heavy_syrk<double>, dim3(nbk), dim3(256), 0, st, Yt64, kp, ldy, nblk, ld, S64,
                           static_cast<double*>(nullptr), flags);
}

}  // namespace gba_amd
