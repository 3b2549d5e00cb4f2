// synth.cpp -- seeded synthetic Snavely-model problems.
//
// The generator is defined here (it has no reference counterpart):
//   * RNG: xoshiro256** seeded through splitmix64(seed); normals by the
//     Box-Muller transform; uniforms from the top 53 bits.
//   * Cameras sit on a horizontal ring of radius 30 around the origin.  The
//     rotation is the look-at rotation (camera -z axis towards the origin,
//     BAL convention p = -P.xy / P.z) composed on the left with Exp(w),
//     w ~ N(0, 0.1^2 I) rad; the stored angle-axis is Log of the product.
//     The translation is -R c + n with n ~ N(0, I).  f ~ U(400, 600),
//     k1 ~ N(0, 0.05^2), k2 ~ N(0, 0.005^2).
//   * Points ~ U([-5, 5]^3).  With the ring radius 30 every point is in front
//     of every camera (P.z < -15).
//   * Observations per point k_p = kmin + K, K either geometric or a floored
//     Pareto(alpha) draw, clamped to kmax; the law's free parameter is solved
//     by bisection so that E[k] = O / P; the total is then made exactly O by
//     +-1 moves on uniformly drawn points that stay inside [kmin, kmax].
//     Each point picks k_p distinct cameras uniformly, or, with view_window
//     W > 0, uniformly among the 2W + 1 cameras within ring distance W of its
//     home camera -- the camera nearest to the point's azimuth
//     atan2(x, z) -- so cameras share points only within ring distance 2W
//     (kmax is then at most 2W + 1).
//   * Pixels: exact projection + N(0, pixel_noise^2) per coordinate.  The
//     solver starts from parameters perturbed by x <- x (1 + perturb N(0,1)).
//   * Observations are emitted camera-major (camera, then point ascending).
//
// The reference's own data-free end-to-end case (tests/test_launch_bounds.cpp:
// 28-80) is regenerated bit-for-bit by gba_synth_launch_bounds with the same
// libstdc++ engine and distributions.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "gba_c.h"
#include "gba_internal.hpp"

namespace gba_amd {
namespace {

struct Rng {
    uint64_t s[4];
    bool have_spare = false;
    double spare = 0.0;
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; ++i) {
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t x = z;
            x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
            x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
            s[i] = x ^ (x >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }  // [0,1)
    double uniform_open() { return (static_cast<double>(next() >> 11) + 0.5) * 0x1.0p-53; }  // (0,1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal() {
        if (have_spare) {
            have_spare = false;
            return spare;
        }
        double u1 = uniform_open(), u2 = uniform();
        double rad = std::sqrt(-2.0 * std::log(u1));
        double ang = 2.0 * M_PI * u2;
        spare = rad * std::sin(ang);
        have_spare = true;
        return rad * std::cos(ang);
    }
    uint32_t below(uint32_t n) { return static_cast<uint32_t>((next() >> 32) * n >> 32); }
};

// Exp of an angle-axis vector (Rodrigues), row-major 3x3.
void so3_exp(const double* w, double* R) {
    double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    double th = std::sqrt(th2);
    double a, b;
    if (th < 1e-8) {
        a = 1.0 - th2 / 6.0;
        b = 0.5 - th2 / 24.0;
    } else {
        a = std::sin(th) / th;
        b = (1.0 - std::cos(th)) / th2;
    }
    double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double kk = 0.0;
            for (int m = 0; m < 3; ++m) kk += K[3 * i + m] * K[3 * m + j];
            R[3 * i + j] = (i == j ? 1.0 : 0.0) + a * K[3 * i + j] + b * kk;
        }
}

// Log of a rotation matrix via its unit quaternion (w >= 0), angle in [0, pi].
void so3_log(const double* R, double* out) {
    double tr = R[0] + R[4] + R[8];
    double q[4];  // w, x, y, z
    if (tr > 0.0) {
        double s = std::sqrt(tr + 1.0) * 2.0;
        q[0] = 0.25 * s;
        q[1] = (R[7] - R[5]) / s;
        q[2] = (R[2] - R[6]) / s;
        q[3] = (R[3] - R[1]) / s;
    } else if (R[0] > R[4] && R[0] > R[8]) {
        double s = std::sqrt(1.0 + R[0] - R[4] - R[8]) * 2.0;
        q[0] = (R[7] - R[5]) / s;
        q[1] = 0.25 * s;
        q[2] = (R[1] + R[3]) / s;
        q[3] = (R[2] + R[6]) / s;
    } else if (R[4] > R[8]) {
        double s = std::sqrt(1.0 + R[4] - R[0] - R[8]) * 2.0;
        q[0] = (R[2] - R[6]) / s;
        q[1] = (R[1] + R[3]) / s;
        q[2] = 0.25 * s;
        q[3] = (R[5] + R[7]) / s;
    } else {
        double s = std::sqrt(1.0 + R[8] - R[0] - R[4]) * 2.0;
        q[0] = (R[3] - R[1]) / s;
        q[1] = (R[2] + R[6]) / s;
        q[2] = (R[5] + R[7]) / s;
        q[3] = 0.25 * s;
    }
    if (q[0] < 0.0)
        for (double& v : q) v = -v;
    double vn = std::sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (vn < 1e-15) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    double ang = 2.0 * std::atan2(vn, q[0]);
    for (int i = 0; i < 3; ++i) out[i] = ang * q[1 + i] / vn;
}

// E[min(K, M)] for the two laws, as a function of their free parameter.
double mean_geometric(double q, int M) {  // P(K >= m) = (1-q)^m
    double s = 0.0, t = 1.0;
    for (int m = 1; m <= M; ++m) {
        t *= (1.0 - q);
        s += t;
    }
    return s;
}
double mean_pareto(double beta, double alpha, int M) {  // P(K >= m) = (1 + m/beta)^-alpha
    double s = 0.0;
    for (int m = 1; m <= M; ++m) s += std::pow(1.0 + m / beta, -alpha);
    return s;
}

}  // namespace

int synth_generate(const gba_synth_spec& sp, uint64_t seed, gba_problem_view* out) {
    const int C = sp.num_cameras, P = sp.num_points, O = sp.num_observations;
    if (C <= 0 || P <= 0 || O <= 0) return set_error("synth: sizes must be positive");
    const int kmin = std::max(1, sp.kmin);
    const int win = sp.view_window > 0 ? std::min(sp.view_window, (C - 1) / 2) : 0;
    const int kmax = win > 0 ? std::min({sp.kmax, C, 2 * win + 1}) : std::min(sp.kmax, C);
    if (kmax < kmin) return set_error("synth: kmax < kmin");
    if (static_cast<long long>(P) * kmin > O || static_cast<long long>(P) * kmax < O)
        return set_error("synth: O outside [P*kmin, P*kmax]");
    Rng rng(seed);
    std::vector<double> cams(9 * static_cast<size_t>(C)), pts(3 * static_cast<size_t>(P));
    const double ring = sp.ring_radius > 0 ? sp.ring_radius : 30.0;
    for (int i = 0; i < C; ++i) {
        double phi = 2.0 * M_PI * i / C;
        double c[3] = {ring * std::sin(phi), 0.0, ring * std::cos(phi)};
        double R0[9] = {std::cos(phi), 0.0, -std::sin(phi), 0.0, 1.0, 0.0,
                        std::sin(phi), 0.0, std::cos(phi)};
        double w[3] = {0.1 * rng.normal(), 0.1 * rng.normal(), 0.1 * rng.normal()};
        double E[9], R[9];
        so3_exp(w, E);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double s = 0.0;
                for (int m = 0; m < 3; ++m) s += E[3 * a + m] * R0[3 * m + b];
                R[3 * a + b] = s;
            }
        double* cam = &cams[9 * static_cast<size_t>(i)];
        so3_log(R, cam);
        for (int a = 0; a < 3; ++a)
            cam[3 + a] = -(R[3 * a] * c[0] + R[3 * a + 1] * c[1] + R[3 * a + 2] * c[2]) + rng.normal();
        cam[6] = rng.uniform(400.0, 600.0);
        cam[7] = 0.05 * rng.normal();
        cam[8] = 0.005 * rng.normal();
    }
    for (auto& v : pts) v = rng.uniform(-5.0, 5.0);

    // observations per point
    const double target = static_cast<double>(O) / P - kmin;
    const int M = kmax - kmin;
    std::vector<int> k(static_cast<size_t>(P));
    if (sp.law == GBA_LAW_PARETO) {
        double alpha = sp.alpha > 0 ? sp.alpha : 1.5;
        double lo = 1e-6, hi = 1e6;
        for (int it = 0; it < 200; ++it) {
            double mid = std::sqrt(lo * hi);
            (mean_pareto(mid, alpha, M) < target ? lo : hi) = mid;
        }
        double beta = std::sqrt(lo * hi);
        for (int p = 0; p < P; ++p) {
            double y = std::pow(rng.uniform_open(), -1.0 / alpha);
            double kk = std::floor(beta * (y - 1.0));
            k[p] = kmin + static_cast<int>(std::min<double>(kk, M));
        }
    } else {
        double lo = 1e-9, hi = 1.0 - 1e-12;
        for (int it = 0; it < 200; ++it) {
            double mid = 0.5 * (lo + hi);
            (mean_geometric(mid, M) > target ? lo : hi) = mid;
        }
        double q = 0.5 * (lo + hi);
        double lq = std::log1p(-q);
        for (int p = 0; p < P; ++p) {
            double g = std::floor(std::log(rng.uniform_open()) / lq);
            k[p] = kmin + static_cast<int>(std::min<double>(g, M));
        }
    }
    long long total = 0;
    for (int v : k) total += v;
    while (total < O) {
        int p = static_cast<int>(rng.below(static_cast<uint32_t>(P)));
        if (k[p] < kmax) {
            ++k[p];
            ++total;
        }
    }
    while (total > O) {
        int p = static_cast<int>(rng.below(static_cast<uint32_t>(P)));
        if (k[p] > kmin) {
            --k[p];
            --total;
        }
    }

    // camera choice: partial Fisher-Yates with undo, O(k_p) per point
    std::vector<int> perm(static_cast<size_t>(C));
    for (int i = 0; i < C; ++i) perm[i] = i;
    std::vector<int> cam_count(static_cast<size_t>(C), 0);
    std::vector<int> chosen(static_cast<size_t>(O));
    std::vector<int> swaps;
    size_t pos = 0;
    std::vector<int> window(2 * static_cast<size_t>(win) + 1);
    for (int p = 0; p < P; ++p) {
        if (win > 0) {
            // k_p of the 2W + 1 cameras around the home camera
            const double az = std::atan2(pts[3 * static_cast<size_t>(p)], pts[3 * static_cast<size_t>(p) + 2]);
            const int home = static_cast<int>(std::lround(az / (2.0 * M_PI) * C));
            for (int d = -win; d <= win; ++d) window[d + win] = ((home + d) % C + C) % C;
            for (int j = 0; j < k[p]; ++j) {
                int r = j + static_cast<int>(rng.below(static_cast<uint32_t>(2 * win + 1 - j)));
                std::swap(window[j], window[r]);
                chosen[pos + j] = window[j];
                cam_count[window[j]]++;
            }
            pos += k[p];
            continue;
        }
        swaps.clear();
        for (int j = 0; j < k[p]; ++j) {
            int r = j + static_cast<int>(rng.below(static_cast<uint32_t>(C - j)));
            std::swap(perm[j], perm[r]);
            swaps.push_back(r);
            chosen[pos + j] = perm[j];
            cam_count[perm[j]]++;
        }
        for (int j = k[p] - 1; j >= 0; --j) std::swap(perm[j], perm[swaps[j]]);
        pos += k[p];
    }
    // camera-major emission: counting sort by camera, points ascending
    std::vector<long long> cstart(static_cast<size_t>(C) + 1, 0);
    for (int i = 0; i < C; ++i) cstart[i + 1] = cstart[i] + cam_count[i];
    std::vector<long long> fill(cstart.begin(), cstart.end() - 1);
    pos = 0;
    for (int p = 0; p < P; ++p) {
        for (int j = 0; j < k[p]; ++j) {
            int c = chosen[pos + j];
            long long e = fill[c]++;
            out->obs_camera[e] = c;
            out->obs_point[e] = p;
        }
        pos += k[p];
    }
    const double noise = sp.pixel_noise;
    for (long long e = 0; e < O; ++e) {
        double pred[2];
        double zero[2] = {0.0, 0.0};
        snavely_residual_host(&cams[9 * static_cast<size_t>(out->obs_camera[e])],
                             &pts[3 * static_cast<size_t>(out->obs_point[e])], zero, pred);
        out->obs_pixel[2 * e] = pred[0] + noise * rng.normal();
        out->obs_pixel[2 * e + 1] = pred[1] + noise * rng.normal();
    }
    const double pert = sp.perturb;
    for (size_t i = 0; i < cams.size(); ++i) out->cameras[i] = cams[i] * (1.0 + pert * rng.normal());
    for (size_t i = 0; i < pts.size(); ++i) out->points[i] = pts[i] * (1.0 + pert * rng.normal());
    return 0;
}

// tests/test_launch_bounds.cpp:28-80, regenerated with the same libstdc++
// engine and distributions so the inputs are identical.
int synth_launch_bounds(int C, int P, uint64_t seed, gba_problem_view* out) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> xy(-3.0, 3.0), zz(4.0, 12.0), pert(-0.05, 0.05);
    std::vector<double> truth_cams(9 * static_cast<size_t>(C)), truth_pts(3 * static_cast<size_t>(P));
    for (int i = 0; i < C; ++i) {
        double* c = &truth_cams[9 * static_cast<size_t>(i)];
        c[0] = 0.02 * (i + 1);
        c[1] = -0.01 * i;
        c[2] = 0.03;
        c[3] = 0.3 * i;
        c[4] = -0.1 * i;
        c[5] = 0.05 * i;
        c[6] = 500.0 + 10.0 * i;
        c[7] = -0.05;
        c[8] = 0.001;
    }
    for (int p = 0; p < P; ++p) {
        truth_pts[3 * static_cast<size_t>(p)] = xy(rng);
        truth_pts[3 * static_cast<size_t>(p) + 1] = xy(rng);
        truth_pts[3 * static_cast<size_t>(p) + 2] = zz(rng);
    }
    std::memcpy(out->cameras, truth_cams.data(), sizeof(double) * truth_cams.size());
    std::memcpy(out->points, truth_pts.data(), sizeof(double) * truth_pts.size());
    for (int p = 0; p < P; ++p)
        for (int j = 0; j < 3; ++j) out->points[3 * static_cast<size_t>(p) + j] += pert(rng);
    for (int p = 0; p < P; ++p) {
        const int cams2[2] = {p % C, (p + 1) % C};
        for (int m = 0; m < 2; ++m) {
            size_t e = 2 * static_cast<size_t>(p) + m;
            out->obs_camera[e] = cams2[m];
            out->obs_point[e] = p;
            // the reference passes the (zero-initialised) pixel slot as the
            // observation, so the residual equals the prediction
            double zero[2] = {0.0, 0.0}, pred[2];
            snavely_residual_host(&truth_cams[9 * static_cast<size_t>(cams2[m])],
                                  &truth_pts[3 * static_cast<size_t>(p)], zero, pred);
            out->obs_pixel[2 * e] = pred[0];
            out->obs_pixel[2 * e + 1] = pred[1];
        }
    }
    return 0;
}

}  // namespace gba_amd
