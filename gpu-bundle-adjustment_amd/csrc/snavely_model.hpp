// snavely_model.hpp -- Snavely (BAL) camera model: residual and analytic
// Jacobian, shared by the HIP kernels and the host entry point.
//
// Camera c = [a (angle-axis, 3), t (3), f, k1, k2]; point X (3); pixel o (2).
//   P = R(a) X + t,  (u, v) = -(P_x, P_y) / P_z,  s = u^2 + v^2,
//   d = 1 + k1 s + k2 s^2,  prediction f d (u, v),  residual = prediction - o
// (reference model: src/lm.cu:58-82, src/bal_reader.cpp:100-128).
//
// Derivation used here (independent of the reference's, same values):
//   R(a) = I + A [a]x + B [a]x^2,  A = sin th / th,  B = (1 - cos th) / th^2
//   d(R X)/da = -[R X]x J(a),  J = I + B [a]x + G [a]x^2,  G = (th - sin th) / th^3
// (J is the left Jacobian of SO(3)).  For a row vector g: g [y]x = g x y, so
//   g d(RX)/da = h + B (h x a) + G ((h x a) x a)   with h = (R X) x g,
//   g R        = g + A (g x a) + B ((g x a) x a).
// Series expansions of A, B, G for th < 1e-2 keep them accurate near zero.
// For th^2 <= 1e-24 the first-order model R X = X + a x X, R = I,
// d(RX)/da = -[X]x is used, as the reference does (src/lm.cu:63,72-76,142-148).
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GBA_HD __host__ __device__ __forceinline__
#else
#define GBA_HD inline
#endif

#include <cmath>

namespace gba_amd {

struct Vec3 {
    double x, y, z;
};

GBA_HD bool finite_d(double x) { return __builtin_isfinite(x); }

GBA_HD void sin_cos(double x, double* s, double* c) {
#if defined(__HIP_DEVICE_COMPILE__)
    sincos(x, s, c);
#else
    *s = std::sin(x);
    *c = std::cos(x);
#endif
}

GBA_HD Vec3 cross3(Vec3 p, Vec3 q) {
    return {p.y * q.z - p.z * q.y, p.z * q.x - p.x * q.z, p.x * q.y - p.y * q.x};
}

// Evaluates the residual r[2]; with kJac also Jc (2x9 row-major, columns
// a, t, f, k1, k2) and Jp (2x3 row-major).  Returns true when every output
// is finite (the reference's anomaly test, src/lm.cu:219-225).
template <bool kJac>
GBA_HD bool snavely_eval(const double* cam, const double* Xp, double ox, double oy, double* r,
                         double* Jc, double* Jp) {
    const Vec3 a{cam[0], cam[1], cam[2]};
    const Vec3 X{Xp[0], Xp[1], Xp[2]};
    const double th2 = a.x * a.x + a.y * a.y + a.z * a.z;
    const bool first_order = !(th2 > 1e-24);
    const Vec3 aX = cross3(a, X);
    double A = 1.0, B = 0.0, G = 0.0;
    Vec3 RX;
    if (!first_order) {
        const double th = std::sqrt(th2);
        if (th < 1e-2) {
            const double t4 = th2 * th2, t6 = t4 * th2;
            A = 1.0 - th2 / 6.0 + t4 / 120.0 - t6 / 5040.0;
            B = 0.5 - th2 / 24.0 + t4 / 720.0 - t6 / 40320.0;
            G = 1.0 / 6.0 - th2 / 120.0 + t4 / 5040.0 - t6 / 362880.0;
        } else {
            double sn, cs;
            sin_cos(th, &sn, &cs);
            A = sn / th;
            B = (1.0 - cs) / th2;
            G = (th - sn) / (th2 * th);
        }
        const Vec3 aaX = cross3(a, aX);
        RX = {X.x + A * aX.x + B * aaX.x, X.y + A * aX.y + B * aaX.y, X.z + A * aX.z + B * aaX.z};
    } else {
        RX = {X.x + aX.x, X.y + aX.y, X.z + aX.z};
    }
    const double Px = RX.x + cam[3], Py = RX.y + cam[4], Pz = RX.z + cam[5];
    const double iz = 1.0 / Pz;
    const double u = -Px / Pz, v = -Py / Pz;
    const double s = u * u + v * v;
    const double f = cam[6], k1 = cam[7], k2 = cam[8];
    const double d = 1.0 + k1 * s + k2 * s * s;
    r[0] = f * d * u - ox;
    r[1] = f * d * v - oy;
    bool ok = finite_d(r[0]) && finite_d(r[1]);
    if constexpr (kJac) {
        // D = d(prediction)/d(u, v) (symmetric), M = d(u, v)/dP = -iz [[1,0,u],[0,1,v]]
        const double L2 = 2.0 * (k1 + 2.0 * k2 * s);
        const double D00 = f * (d + L2 * u * u), D01 = f * L2 * u * v, D11 = f * (d + L2 * v * v);
        const double m = -iz;
        const Vec3 g0{m * D00, m * D01, m * (D00 * u + D01 * v)};  // row 0 of D M
        const Vec3 g1{m * D01, m * D11, m * (D01 * u + D11 * v)};  // row 1 of D M
        Vec3 rot0, rot1, pt0, pt1;
        if (!first_order) {
            Vec3 h0 = cross3(RX, g0), h1 = cross3(RX, g1);
            Vec3 ha0 = cross3(h0, a), ha1 = cross3(h1, a);
            Vec3 haa0 = cross3(ha0, a), haa1 = cross3(ha1, a);
            rot0 = {h0.x + B * ha0.x + G * haa0.x, h0.y + B * ha0.y + G * haa0.y,
                    h0.z + B * ha0.z + G * haa0.z};
            rot1 = {h1.x + B * ha1.x + G * haa1.x, h1.y + B * ha1.y + G * haa1.y,
                    h1.z + B * ha1.z + G * haa1.z};
            Vec3 ga0 = cross3(g0, a), ga1 = cross3(g1, a);
            Vec3 gaa0 = cross3(ga0, a), gaa1 = cross3(ga1, a);
            pt0 = {g0.x + A * ga0.x + B * gaa0.x, g0.y + A * ga0.y + B * gaa0.y,
                   g0.z + A * ga0.z + B * gaa0.z};
            pt1 = {g1.x + A * ga1.x + B * gaa1.x, g1.y + A * ga1.y + B * gaa1.y,
                   g1.z + A * ga1.z + B * gaa1.z};
        } else {
            rot0 = cross3(X, g0);
            rot1 = cross3(X, g1);
            pt0 = g0;
            pt1 = g1;
        }
        Jc[0] = rot0.x; Jc[1] = rot0.y; Jc[2] = rot0.z;
        Jc[3] = g0.x;   Jc[4] = g0.y;   Jc[5] = g0.z;
        Jc[6] = d * u;  Jc[7] = f * s * u;  Jc[8] = f * s * s * u;
        Jc[9] = rot1.x; Jc[10] = rot1.y; Jc[11] = rot1.z;
        Jc[12] = g1.x;  Jc[13] = g1.y;   Jc[14] = g1.z;
        Jc[15] = d * v; Jc[16] = f * s * v; Jc[17] = f * s * s * v;
        Jp[0] = pt0.x; Jp[1] = pt0.y; Jp[2] = pt0.z;
        Jp[3] = pt1.x; Jp[4] = pt1.y; Jp[5] = pt1.z;
#pragma unroll
        for (int i = 0; i < 18; ++i) ok = ok && finite_d(Jc[i]);
#pragma unroll
        for (int i = 0; i < 6; ++i) ok = ok && finite_d(Jp[i]);
    }
    return ok;
}

}  // namespace gba_amd
