// Host check of the Jacobian product forms of snavely_model.hpp (edge_jc_x,
// edge_jpT_m, edge_jp_z, edge_jcT_t) against products with the explicit
// Jacobians of snavely_eval<true>, over random observations in the three
// rotation regimes (series for theta < 1e-2, closed form, first-order model
// at theta = 0).  Prints the largest relative difference per product.
#include <cmath>
#include <cstdio>
#include <random>

#include "snavely_model.hpp"

using namespace gba_amd;

int main() {
    std::mt19937_64 rng(7);
    std::normal_distribution<double> n01(0.0, 1.0);
    double worst[4] = {0, 0, 0, 0};
    int cases = 0;
    for (int it = 0; it < 30000; ++it) {
        double cam[9], X[3], x[9], z[3], m[2];
        const int regime = it % 3;  // 0: closed form, 1: series, 2: first order
        const double scale = regime == 0 ? 0.3 : (regime == 1 ? 3e-3 : 0.0);
        for (int k = 0; k < 3; ++k) cam[k] = scale * n01(rng);
        for (int k = 3; k < 6; ++k) cam[k] = n01(rng);
        cam[5] -= 8.0;  // the point lies in front of the camera (P_z < 0)
        cam[6] = 500.0 + 50.0 * n01(rng);
        cam[7] = 0.05 * n01(rng);
        cam[8] = 0.005 * n01(rng);
        for (int k = 0; k < 3; ++k) X[k] = n01(rng);
        for (int k = 0; k < 9; ++k) x[k] = n01(rng);
        for (int k = 0; k < 3; ++k) z[k] = n01(rng);
        m[0] = n01(rng);
        m[1] = n01(rng);
        double r[2], Jc[18], Jp[6], r2[2];
        if (!snavely_eval<true>(cam, X, 1.0, -2.0, r, Jc, Jp)) continue;
        EdgeLin e;
        snavely_core<true>(cam, X, 1.0, -2.0, r2, e);
        ++cases;
        // Jc x
        double a[2], b[2];
        edge_jc_x(e, x, a);
        for (int i = 0; i < 2; ++i) {
            b[i] = 0.0;
            double mag = 0.0;
            for (int k = 0; k < 9; ++k) {
                b[i] += Jc[9 * i + k] * x[k];
                mag += std::fabs(Jc[9 * i + k] * x[k]);
            }
            worst[0] = std::fmax(worst[0], std::fabs(a[i] - b[i]) / mag);
        }
        // Jp^T m
        const Vec3 t = edge_jpT_m(e, m[0], m[1]);
        const double tv[3] = {t.x, t.y, t.z};
        for (int k = 0; k < 3; ++k) {
            const double ref = Jp[k] * m[0] + Jp[3 + k] * m[1];
            const double mag = std::fabs(Jp[k] * m[0]) + std::fabs(Jp[3 + k] * m[1]);
            worst[1] = std::fmax(worst[1], std::fabs(tv[k] - ref) / mag);
        }
        // Jp z
        edge_jp_z(e, Vec3{z[0], z[1], z[2]}, a);
        for (int i = 0; i < 2; ++i) {
            double ref = 0.0, mag = 0.0;
            for (int k = 0; k < 3; ++k) {
                ref += Jp[3 * i + k] * z[k];
                mag += std::fabs(Jp[3 * i + k] * z[k]);
            }
            worst[2] = std::fmax(worst[2], std::fabs(a[i] - ref) / mag);
        }
        // Jc^T t
        double o9[9];
        edge_jcT_t(e, m[0], m[1], o9);
        for (int k = 0; k < 9; ++k) {
            const double ref = Jc[k] * m[0] + Jc[9 + k] * m[1];
            const double mag = std::fabs(Jc[k] * m[0]) + std::fabs(Jc[9 + k] * m[1]);
            worst[3] = std::fmax(worst[3], std::fabs(o9[k] - ref) / mag);
        }
        if (r[0] != r2[0] || r[1] != r2[1]) {
            std::printf("residual differs\n");
            return 1;
        }
    }
    std::printf("cases %d jc_x %.3e jpT_m %.3e jp_z %.3e jcT_t %.3e\n", cases, worst[0], worst[1], worst[2], worst[3]);
    return 0;
}
