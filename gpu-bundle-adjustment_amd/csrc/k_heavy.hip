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

// S -= Y Y^T over the lower tiles: tile (I, J), I >= J, one workgroup each
// (blockIdx = I (I + 1) / 2 + J).  The four wavefronts take the four 32x32
// quarters of the tile, each a 2x2 set of 16x16 MFMA accumulators over k in
// steps of four: lane l supplies A[l % 16][l / 16] = Y[row][k0 + l / 16] and
// B[l / 16][l % 16] = Y[col][k0 + l / 16], i.e. 16 consecutive entries of one
// k row of Yt -- coalesced loads straight from memory, four k steps of them
// issued ahead of their MFMAs.  kp is a multiple of 16 (zero columns pad it).
// A diagonal tile updates its lower triangle and the upper halves of the 9x9
// camera blocks inside it (the layout every consumer of S reads); S2, when
// given, is a second copy receiving the same values.
template <typename T>
__global__ __launch_bounds__(256) void k_heavy_syrk(const T* __restrict__ Yt, int kp, int ld, T* __restrict__ S,
                                                    T* __restrict__ S2, const TrialFlags* __restrict__ flags) {
    if (flags->nonspd) return;
    typedef typename Acc4<T>::type acc_t;
    const int t = blockIdx.x;
    int I = static_cast<int>((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= t) ++I;
    while (I * (I + 1) / 2 > t) --I;
    const int J = t - I * (I + 1) / 2;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = w >> 1, wc = w & 1;
    const int kk = lane >> 4, ii = lane & 15;
    const T* ya = Yt + static_cast<size_t>(kk) * ld + kTile * I + 32 * wr + ii;
    const T* yb = Yt + static_cast<size_t>(kk) * ld + kTile * J + 32 * wc + ii;
    acc_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
    const size_t step = static_cast<size_t>(4) * ld;
    for (int k0 = 0; k0 < kp; k0 += 16) {
        T av[4][2], bv[4][2];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const size_t o = static_cast<size_t>(k0) * ld + s * step;
            av[s][0] = ya[o];
            av[s][1] = ya[o + 16];
            bv[s][0] = yb[o];
            bv[s][1] = yb[o + 16];
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) acc[a][b] = mfma16(av[s][a], bv[s][b], acc[a][b]);
    }
    const int g = lane >> 4, c16 = lane & 15;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = 32 * wr + 16 * a + acc_row<T>(g, q), c = 32 * wc + 16 * b + c16;
                const int R = kTile * I + r, Cc = kTile * J + c;
                if (I == J && r < c && R / 9 != Cc / 9) continue;
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

void launch_heavy_syrk(hipStream_t st, int nblk, int kp, int ld, const double* Yt64, const float* Yt32, double* S64,
                       float* S32, float* S32b, const TrialFlags* flags) {
    const int ntiles = nblk * (nblk + 1) / 2;
    if (ntiles <= 0 || kp <= 0) return;
    if (Yt32)
        hipLaunchKernelGGL(k_heavy_syrk<float>, dim3(ntiles), dim3(256), 0, st, Yt32, kp, ld, S32, S32b, flags);
    else
        hipLaunchKernelGGL(k_heavy_syrk<double>, dim3(ntiles), dim3(256), 0, st, Yt64, kp, ld, S64,
                           static_cast<double*>(nullptr), flags);
}

}  // namespace gba_amd
