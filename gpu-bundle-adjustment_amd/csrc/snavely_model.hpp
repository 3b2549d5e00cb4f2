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

// What the residual and the Jacobian of one observation are built from: the
// angle-axis a, the rotation terms, the rotated point, the rows g0, g1 of
// d(prediction)/dP and the intrinsic columns.
struct EdgeLin {
    Vec3 a, X, RX;
    double A, B, G;
    bool first_order;
    Vec3 g0, g1;
    double d, u, v, s, f;
};

// Rotation terms of a camera's angle-axis a[3] (they depend on the camera
// only): A, B, G, and whether the first-order model applies.
struct RotTerms {
    double A, B, G;
    bool first_order;
};

GBA_HD RotTerms rot_terms(const double* a3) {
    const double th2 = a3[0] * a3[0] + a3[1] * a3[1] + a3[2] * a3[2];
    RotTerms t{1.0, 0.0, 0.0, !(th2 > 1e-24)};
    if (!t.first_order) {
        const double th = std::sqrt(th2);
        if (th < 1e-2) {
            const double t4 = th2 * th2, t6 = t4 * th2;
            t.A = 1.0 - th2 / 6.0 + t4 / 120.0 - t6 / 5040.0;
            t.B = 0.5 - th2 / 24.0 + t4 / 720.0 - t6 / 40320.0;
            t.G = 1.0 / 6.0 - th2 / 120.0 + t4 / 5040.0 - t6 / 362880.0;
        } else {
            double sn, cs;
            sin_cos(th, &sn, &cs);
            t.A = sn / th;
            t.B = (1.0 - cs) / th2;
            t.G = (th - sn) / (th2 * th);
        }
    }
    return t;
}

// The residual r[2] with the camera's rotation terms given; with kJac also
// e.g0, e.g1 (the rest of e always).
template <bool kJac>
GBA_HD void snavely_project(const double* cam, const RotTerms& rt, const double* Xp, double ox, double oy, double* r,
                            EdgeLin& e) {
    const Vec3 a{cam[0], cam[1], cam[2]};
    const Vec3 X{Xp[0], Xp[1], Xp[2]};
    const bool first_order = rt.first_order;
    const Vec3 aX = cross3(a, X);
    const double A = rt.A, B = rt.B, G = rt.G;
    Vec3 RX;
    if (!first_order) {
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
    e.a = a;
    e.X = X;
    e.RX = RX;
    e.A = A;
    e.B = B;
    e.G = G;
    e.first_order = first_order;
    e.d = d;
    e.u = u;
    e.v = v;
    e.s = s;
    e.f = f;
    if constexpr (kJac) {
        // D = d(prediction)/d(u, v) (symmetric), M = d(u, v)/dP = -iz [[1,0,u],[0,1,v]]
        const double L2 = 2.0 * (k1 + 2.0 * k2 * s);
        const double D00 = f * (d + L2 * u * u), D01 = f * L2 * u * v, D11 = f * (d + L2 * v * v);
        const double m = -iz;
        e.g0 = {m * D00, m * D01, m * (D00 * u + D01 * v)};  // row 0 of D M
        e.g1 = {m * D01, m * D11, m * (D01 * u + D11 * v)};  // row 1 of D M
    }
}

template <bool kJac>
GBA_HD void snavely_core(const double* cam, const double* Xp, double ox, double oy, double* r, EdgeLin& e) {
    snavely_project<kJac>(cam, rot_terms(cam), Xp, ox, oy, r, e);
}

// Evaluates the residual r[2]; with kJac also Jc (2x9 row-major, columns
// a, t, f, k1, k2) and Jp (2x3 row-major).  Returns true when every output
// is finite (the reference's anomaly test, src/lm.cu:219-225).
template <bool kJac>
GBA_HD bool snavely_eval(const double* cam, const double* Xp, double ox, double oy, double* r,
                         double* Jc, double* Jp) {
    EdgeLin e;
    snavely_core<kJac>(cam, Xp, ox, oy, r, e);
    bool ok = finite_d(r[0]) && finite_d(r[1]);
    if constexpr (kJac) {
        const Vec3 a = e.a, RX = e.RX, g0 = e.g0, g1 = e.g1;
        const double A = e.A, B = e.B, G = e.G, d = e.d, u = e.u, v = e.v, s = e.s, f = e.f;
        Vec3 rot0, rot1, pt0, pt1;
        if (!e.first_order) {
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
            rot0 = cross3(e.X, g0);
            rot1 = cross3(e.X, g1);
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

// Products with an observation's Jacobians without forming them (e from
// snavely_core<true>).  Row r of Jc is (g_r d(RX)/da, g_r, q_r) with the
// intrinsic columns q_r = (d, f s, f s^2) (u, v)_r, and row r of Jp is g_r R.
// With g d(RX)/da = h + B (h x a) + G ((h x a) x a), h = RX x g (first-order
// model: X x g), a dot product with y is (g d(RX)/da) . y = g . ((J y) x RX)
// for J y = y + B (a x y) + G (a x (a x y)); and g R . z = g . (R z) with
// R z = z + A (a x z) + B (a x (a x z)) (first order: R = I in the
// Jacobian).  So each product costs one rotation-sized term instead of the
// 24 Jacobian entries.
GBA_HD Vec3 rotate_terms(const Vec3& a, double A, double B, const Vec3& z) {  // z + A a x z + B a x (a x z)
    const Vec3 az = cross3(a, z), aaz = cross3(a, az);
    return {z.x + A * az.x + B * aaz.x, z.y + A * az.y + B * aaz.y, z.z + A * az.z + B * aaz.z};
}

GBA_HD double dot3(const Vec3& p, const Vec3& q) { return p.x * q.x + p.y * q.y + p.z * q.z; }

// Jc x (2 values) for the camera step x[9]
GBA_HD void edge_jc_x(const EdgeLin& e, const double* x, double* out) {
    const Vec3 y{x[0], x[1], x[2]};
    Vec3 q;
    if (!e.first_order) {
        const Vec3 w = rotate_terms(e.a, e.B, e.G, y);
        q = cross3(w, e.RX);
    } else {
        q = cross3(y, e.X);
    }
    q = {q.x + x[3], q.y + x[4], q.z + x[5]};
    const double k = e.d * x[6] + e.f * e.s * x[7] + e.f * e.s * e.s * x[8];
    out[0] = dot3(e.g0, q) + e.u * k;
    out[1] = dot3(e.g1, q) + e.v * k;
}

// Jp^T m (3 values) for m[2]
GBA_HD Vec3 edge_jpT_m(const EdgeLin& e, double m0, double m1) {
    const Vec3 g{m0 * e.g0.x + m1 * e.g1.x, m0 * e.g0.y + m1 * e.g1.y, m0 * e.g0.z + m1 * e.g1.z};
    if (e.first_order) return g;
    // (g R)^T = R^T g = g + A (g x a) + B ((g x a) x a) = rotation by -a
    const Vec3 na{-e.a.x, -e.a.y, -e.a.z};
    return rotate_terms(na, e.A, e.B, g);
}

// Jp z (2 values) for the point step z
GBA_HD void edge_jp_z(const EdgeLin& e, const Vec3& z, double* out) {
    const Vec3 Rz = e.first_order ? z : rotate_terms(e.a, e.A, e.B, z);
    out[0] = dot3(e.g0, Rz);
    out[1] = dot3(e.g1, Rz);
}

// Jc^T t (9 values) for t[2]
GBA_HD void edge_jcT_t(const EdgeLin& e, double t0, double t1, double* out) {
    const Vec3 g{t0 * e.g0.x + t1 * e.g1.x, t0 * e.g0.y + t1 * e.g1.y, t0 * e.g0.z + t1 * e.g1.z};
    Vec3 rot;
    if (!e.first_order) {
        const Vec3 h = cross3(e.RX, g);
        const Vec3 ha = cross3(h, e.a), haa = cross3(ha, e.a);
        rot = {h.x + e.B * ha.x + e.G * haa.x, h.y + e.B * ha.y + e.G * haa.y, h.z + e.B * ha.z + e.G * haa.z};
    } else {
        rot = cross3(e.X, g);
    }
    const double k = t0 * e.u + t1 * e.v;
    out[0] = rot.x;
    out[1] = rot.y;
    out[2] = rot.z;
    out[3] = g.x;
    out[4] = g.y;
    out[5] = g.z;
    out[6] = e.d * k;
    out[7] = e.f * e.s * k;
    out[8] = e.f * e.s * e.s * k;
}

}  // namespace gba_amd
