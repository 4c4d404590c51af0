// Differentiable D3Q27 brick step for the adjoint: the cumulant collision with
// Boussinesq buoyancy and Guo forcing, A-B pull streaming with halfway
// bounce-back, a per-cell velocity inlet on the x- face, convective outflow on
// the x+ face, periodic y and z, and the benchmark-mode temperature update. The
// same code runs in FP32 / FP64 and on forward-mode dual numbers, on the device
// and on the host, so the adjoint kernels differentiate exactly the forward they
// replay (ref: native/src/aero_lbm_jni.cpp:2209-2413 collide, 1475-1602 face
// rules, 2145-2188 temperature; the Smagorinsky term is off in this model).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace aero {
namespace adj {

#define AD_HD __host__ __device__ __forceinline__

// forward-mode dual number: value and one tangent
template <class T>
struct Dual {
    T v, d;
    AD_HD Dual() : v(0), d(0) {}
    AD_HD Dual(T a) : v(a), d(0) {}
    AD_HD Dual(T a, T b) : v(a), d(b) {}
};
template <class T> AD_HD Dual<T> operator+(Dual<T> a, Dual<T> b) { return {a.v + b.v, a.d + b.d}; }
template <class T> AD_HD Dual<T> operator-(Dual<T> a, Dual<T> b) { return {a.v - b.v, a.d - b.d}; }
template <class T> AD_HD Dual<T> operator-(Dual<T> a) { return {-a.v, -a.d}; }
template <class T> AD_HD Dual<T> operator*(Dual<T> a, Dual<T> b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
template <class T> AD_HD Dual<T> operator/(Dual<T> a, Dual<T> b) {
    const T q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
template <class T> AD_HD Dual<T> operator+(Dual<T> a, T b) { return {a.v + b, a.d}; }
template <class T> AD_HD Dual<T> operator+(T a, Dual<T> b) { return {a + b.v, b.d}; }
template <class T> AD_HD Dual<T> operator-(Dual<T> a, T b) { return {a.v - b, a.d}; }
template <class T> AD_HD Dual<T> operator-(T a, Dual<T> b) { return {a - b.v, -b.d}; }
template <class T> AD_HD Dual<T> operator*(Dual<T> a, T b) { return {a.v * b, a.d * b}; }
template <class T> AD_HD Dual<T> operator*(T a, Dual<T> b) { return {a * b.v, a * b.d}; }
template <class T> AD_HD Dual<T> operator/(Dual<T> a, T b) { return {a.v / b, a.d / b}; }
template <class T> AD_HD Dual<T>& operator+=(Dual<T>& a, Dual<T> b) { return a = a + b; }
template <class T> AD_HD Dual<T>& operator-=(Dual<T>& a, Dual<T> b) { return a = a - b; }
template <class T> AD_HD Dual<T>& operator*=(Dual<T>& a, Dual<T> b) { return a = a * b; }

template <class T> AD_HD T val(T x) { return x; }
template <class T> AD_HD T val(Dual<T> x) { return x.v; }
template <class T> AD_HD T tan_(T) { return T(0); }
template <class T> AD_HD T tan_(Dual<T> x) { return x.d; }
AD_HD float sqrt_(float x) { return sqrtf(x); }
AD_HD double sqrt_(double x) { return sqrt(x); }
template <class T> AD_HD Dual<T> sqrt_(Dual<T> x) {
    const T s = sqrt_(x.v);
    return {s, x.d / (T(2) * s)};
}
// min / max / clamp choose by value (the derivative of the chosen branch)
template <class R> AD_HD R rmin(R a, R b) { return val(b) < val(a) ? b : a; }
template <class R> AD_HD R rmax(R a, R b) { return val(a) < val(b) ? b : a; }
template <class R, class S> AD_HD R rclamp(R v, S lo, S hi) { return rmin(R(hi), rmax(R(lo), v)); }

struct Params {
    int nx, ny, nz;
    double tau_shear, tau_normal;  // relaxation times of the off-diagonal and diagonal second moments
    float beta;                   // Boussinesq coefficient
    int g_axis, g_sign;           // gravity axis and sign (force = -g_sign * beta T along g_axis)
    float alpha;                  // thermal diffusivity (lattice)
    float kmax;                   // velocity cap
    float rho_in;                 // inlet density
    float u0[3];                  // initial velocity
};

AD_HD int cxq(int q) { return q / 9 - 1; }
AD_HD int cyq(int q) { return (q / 3) % 3 - 1; }
AD_HD int czq(int q) { return q % 3 - 1; }
AD_HD int oppq(int q) { return 26 - q; }
template <class R> AD_HD R wq(int q) {
    const int n = (cxq(q) != 0) + (cyq(q) != 0) + (czq(q) != 0);
    return n == 0 ? R(8.0 / 27.0) : (n == 1 ? R(2.0 / 27.0) : (n == 2 ? R(1.0 / 54.0) : R(1.0 / 216.0)));
}
template <class R> AD_HD R feq(int q, R rho, R ux, R uy, R uz) {
    const R cu = R(3) * (R(float(cxq(q))) * ux + R(float(cyq(q))) * uy + R(float(czq(q))) * uz);
    const R uu = ux * ux + uy * uy + uz * uz;
    return wq<R>(q) * rho * (R(1) + cu + R(0.5) * cu * cu - R(1.5) * uu);
}

#define AD_IDX(a, b, c) (((a) * 3 + (b)) * 3 + (c))

// populations (c = -1, 0, +1) <-> raw moments (order 0, 1, 2) along one line
template <class R> AD_HD void fwd_line(R& a, R& b, R& c) {
    const R s = a + c, t = c - a;
    a = s + b;
    b = t;
    c = s;
}
template <class R> AD_HD void inv_line(R& m0, R& m1, R& m2) {
    const R h2 = R(0.5) * m2, h1 = R(0.5) * m1;
    const R c0 = m0 - m2;
    m0 = h2 - h1;  // c = -1
    m1 = c0;       // c = 0
    m2 = h2 + h1;  // c = +1
}
template <class R> AD_HD void shift_line(R& m0, R& m1, R& m2, R u) {
    const R k1 = m1 - u * m0;
    m2 = m2 - R(2) * u * m1 + u * u * m0;
    m1 = k1;
}
template <class R> AD_HD void unshift_line(R& k0, R& k1, R& k2, R u) {
    k2 = k2 + R(2) * u * k1 + u * u * k0;
    k1 = k1 + u * k0;
}
// apply a line transform along axis a (0 = x, 1 = y, 2 = z) to all 9 lines
template <class R, class Fn> AD_HD void lines(R m[27], int axis, Fn fn) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const int a0 = axis == 0 ? AD_IDX(0, i, j) : axis == 1 ? AD_IDX(i, 0, j) : AD_IDX(i, j, 0);
            const int st = axis == 0 ? 9 : axis == 1 ? 3 : 1;
            fn(m[a0], m[a0 + st], m[a0 + 2 * st]);
        }
}

// One fluid cell: f (27 incoming populations) -> post-collision populations; u_out
// receives the collision velocity. T: the cell's new temperature.
// The collision depends on the populations only through their raw moments of
// order <= 2 (the closure sets every higher central moment). kLow of them:
// m000, m100, m010, m001, m200, m020, m002, m110, m101, m011.
constexpr int kLow = 10;
AD_HD int low_a(int k) { return k == 1 || k == 4 || k == 7 || k == 8 ? (k == 4 ? 2 : 1) : 0; }
AD_HD int low_b(int k) { return k == 2 || k == 5 || k == 7 || k == 9 ? (k == 5 ? 2 : 1) : 0; }
AD_HD int low_c(int k) { return k == 3 || k == 6 || k == 8 || k == 9 ? (k == 6 ? 2 : 1) : 0; }

template <class R> AD_HD void low_moments(const R f[27], R m[kLow]) {
    R t[27];
    for (int q = 0; q < 27; ++q) t[q] = f[q];
    lines(t, 2, [](R& a, R& b, R& c) { fwd_line(a, b, c); });
    lines(t, 1, [](R& a, R& b, R& c) { fwd_line(a, b, c); });
    lines(t, 0, [](R& a, R& b, R& c) { fwd_line(a, b, c); });
    for (int k = 0; k < kLow; ++k) m[k] = t[AD_IDX(low_a(k), low_b(k), low_c(k))];
}

// post-collision populations f and collision velocity from the low raw moments m
// and the new temperature T
template <class R> AD_HD void collide_low(const Params& P, const R m[kLow], R T, R f[27], R u_out[3]) {
    const R rho = m[0];
    const R rho_s = rmax(R(1e-6f), rho);
    R u[3] = {m[1] / rho_s, m[2] / rho_s, m[3] / rho_s};
    // Boussinesq: F / rho = clamp(beta T, +-0.02) against gravity
    R b = rclamp(R(P.beta) * T, -0.02f, 0.02f);
    if (P.g_sign > 0) b = -b;
    u[P.g_axis] = u[P.g_axis] + R(0.5) * b;
    {
        const R s2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        if (val(s2) > P.kmax * P.kmax) {
            const R sc = R(P.kmax) / sqrt_(s2);
            u[0] = u[0] * sc;
            u[1] = u[1] * sc;
            u[2] = u[2] * sc;
        }
    }
    for (int k = 0; k < 3; ++k) u_out[k] = u[k];
    const R ux = u[0], uy = u[1], uz = u[2];
    // second central moments about u, in the operation order of the x, y, z line
    // shifts applied to the full moment set
    const R k100 = m[1] - ux * m[0], k010 = m[2] - uy * m[0];
    const R c200 = m[4] - R(2) * ux * m[1] + ux * ux * m[0];
    const R c020 = m[5] - R(2) * uy * m[2] + uy * uy * m[0];
    const R c002 = m[6] - R(2) * uz * m[3] + uz * uz * m[0];
    const R c110 = (m[7] - ux * m[2]) - uy * k100;
    const R c101 = (m[8] - ux * m[3]) - uz * k100;
    const R c011 = (m[9] - uy * m[3]) - uz * k010;
    const R cs2 = R(1.0 / 3.0);
    // relaxation rates clamped to [0, 1.95] in the closure (ref: jni.cpp:1384-1430);
    // the Guo prefactor takes the unclamped shear rate
    const double om_off = 1.0 / P.tau_shear;
    const R sd = R(fmin(fmax(1.0 / P.tau_normal, 0.0), 1.95)), so = R(fmin(fmax(om_off, 0.0), 1.95));
    const R eq2 = rho_s * cs2;
    const R d200 = c200 + sd * (eq2 - c200);
    const R d020 = c020 + sd * (eq2 - c020);
    const R d002 = c002 + sd * (eq2 - c002);
    const R om = R(1) - so;
    const R d110 = om * c110, d101 = om * c101, d011 = om * c011;
    const R ir = R(1) / rho_s;
    // Gaussian closure of orders 4 and 6, zero odd central moments (ref: jni.cpp:1384-1430)
    const R d220 = (d200 * d020 + R(2) * d110 * d110) * ir;
    const R d202 = (d200 * d002 + R(2) * d101 * d101) * ir;
    const R d022 = (d020 * d002 + R(2) * d011 * d011) * ir;
    const R d211 = (d200 * d011 + R(2) * d110 * d101) * ir;
    const R d121 = (d020 * d101 + R(2) * d110 * d011) * ir;
    const R d112 = (d002 * d110 + R(2) * d101 * d011) * ir;
    const R d222 = (d200 * d020 * d002 + R(2) * d110 * d110 * d002 + R(2) * d101 * d101 * d020 +
                    R(2) * d011 * d011 * d200 + R(8) * d110 * d101 * d011) * ir * ir;
    for (int k = 0; k < 27; ++k) f[k] = R(0);
    f[AD_IDX(0, 0, 0)] = rho_s;
    f[AD_IDX(2, 0, 0)] = d200;
    f[AD_IDX(0, 2, 0)] = d020;
    f[AD_IDX(0, 0, 2)] = d002;
    f[AD_IDX(1, 1, 0)] = d110;
    f[AD_IDX(1, 0, 1)] = d101;
    f[AD_IDX(0, 1, 1)] = d011;
    f[AD_IDX(2, 2, 0)] = d220;
    f[AD_IDX(2, 0, 2)] = d202;
    f[AD_IDX(0, 2, 2)] = d022;
    f[AD_IDX(2, 1, 1)] = d211;
    f[AD_IDX(1, 2, 1)] = d121;
    f[AD_IDX(1, 1, 2)] = d112;
    f[AD_IDX(2, 2, 2)] = d222;
    lines(f, 2, [uz](R& a, R& b, R& c) { unshift_line(a, b, c, uz); });
    lines(f, 1, [uy](R& a, R& b, R& c) { unshift_line(a, b, c, uy); });
    lines(f, 0, [ux](R& a, R& b, R& c) { unshift_line(a, b, c, ux); });
    lines(f, 0, [](R& a, R& b, R& c) { inv_line(a, b, c); });
    lines(f, 1, [](R& a, R& b, R& c) { inv_line(a, b, c); });
    lines(f, 2, [](R& a, R& b, R& c) { inv_line(a, b, c); });
    // Guo source S_q = P w_q [3 (c - u).F + 9 (c.u)(c.F)], F = rho b along the axis.
    // The branch depends on beta, not on b: at T = 0 the term is zero but its
    // derivative with respect to T is not.
    if (P.beta != 0.0f) {
        const R pref = rmax(R(0), R(1) - R(0.5) * R(fmin(fmax(om_off, 0.0), 1.999)));
        const R fa = pref * rho_s * b;
        for (int q = 0; q < 27; ++q) {
            const R c[3] = {R(float(cxq(q))), R(float(cyq(q))), R(float(czq(q)))};
            const R cu = c[0] * ux + c[1] * uy + c[2] * uz;
            f[q] = f[q] + wq<R>(q) * fa * (R(3) * (c[P.g_axis] - u[P.g_axis]) + R(9) * cu * c[P.g_axis]);
        }
    }
}

// Reverse mode of collide_low by hand, for R = float or double: given the adjoints
// fbar of the 27 post-collision populations and ubar of the collision velocity,
// the adjoints of the 10 low raw moments (mbar) and of T (*tbar). The forward
// values are recomputed here; every branch (density floor, buoyancy clamp, speed
// cap) takes the side collide_low takes, so this is the transposed Jacobian of
// exactly that map.
template <class R>
AD_HD void collide_low_adjoint(const Params& P, const R m[kLow], R T, const R fbar[27], const R ubar_in[3],
                               R mbar[kLow], R* tbar) {
    // ---- forward values
    const R rho = m[0];
    const bool rho_live = R(1e-6f) < rho;  // rmax(1e-6, rho) returns rho
    const R rho_s = rho_live ? rho : R(1e-6f);
    const R u0[3] = {m[1] / rho_s, m[2] / rho_s, m[3] / rho_s};
    const R bt = R(P.beta) * T;
    const R lo_ = R(-0.02f), hi_ = R(0.02f);
    const R b_in = lo_ < bt ? bt : lo_;
    const bool b_live = (lo_ < bt) && (b_in < hi_);
    R b = b_in < hi_ ? b_in : hi_;
    const R sb = P.g_sign > 0 ? R(-1) : R(1);
    b = sb * b;
    R u1[3] = {u0[0], u0[1], u0[2]};
    u1[P.g_axis] = u1[P.g_axis] + R(0.5) * b;
    const R s2 = u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2];
    const bool capped = val(s2) > P.kmax * P.kmax;  // the comparison collide_low makes
    const R sc = capped ? R(P.kmax) / sqrt_(s2) : R(1);
    const R u[3] = {capped ? u1[0] * sc : u1[0], capped ? u1[1] * sc : u1[1], capped ? u1[2] * sc : u1[2]};
    const R ux = u[0], uy = u[1], uz = u[2];
    const R k100 = m[1] - ux * m[0], k010 = m[2] - uy * m[0];
    const R c200 = m[4] - R(2) * ux * m[1] + ux * ux * m[0];
    const R c020 = m[5] - R(2) * uy * m[2] + uy * uy * m[0];
    const R c002 = m[6] - R(2) * uz * m[3] + uz * uz * m[0];
    const R c110 = (m[7] - ux * m[2]) - uy * k100;
    const R c101 = (m[8] - ux * m[3]) - uz * k100;
    const R c011 = (m[9] - uy * m[3]) - uz * k010;
    const R cs2 = R(1.0 / 3.0);
    const double om_off = 1.0 / P.tau_shear;
    const R sd = R(fmin(fmax(1.0 / P.tau_normal, 0.0), 1.95)), so = R(fmin(fmax(om_off, 0.0), 1.95));
    const R eq2 = rho_s * cs2;
    const R d200 = c200 + sd * (eq2 - c200);
    const R d020 = c020 + sd * (eq2 - c020);
    const R d002 = c002 + sd * (eq2 - c002);
    const R om = R(1) - so;
    const R d110 = om * c110, d101 = om * c101, d011 = om * c011;
    const R ir = R(1) / rho_s;
    // central moments after the collision, then the three unshift passes (kept for
    // the velocity derivative of each pass)
    R D0[27];
    for (int k = 0; k < 27; ++k) D0[k] = R(0);
    D0[AD_IDX(0, 0, 0)] = rho_s;
    D0[AD_IDX(2, 0, 0)] = d200;
    D0[AD_IDX(0, 2, 0)] = d020;
    D0[AD_IDX(0, 0, 2)] = d002;
    D0[AD_IDX(1, 1, 0)] = d110;
    D0[AD_IDX(1, 0, 1)] = d101;
    D0[AD_IDX(0, 1, 1)] = d011;
    D0[AD_IDX(2, 2, 0)] = (d200 * d020 + R(2) * d110 * d110) * ir;
    D0[AD_IDX(2, 0, 2)] = (d200 * d002 + R(2) * d101 * d101) * ir;
    D0[AD_IDX(0, 2, 2)] = (d020 * d002 + R(2) * d011 * d011) * ir;
    D0[AD_IDX(2, 1, 1)] = (d200 * d011 + R(2) * d110 * d101) * ir;
    D0[AD_IDX(1, 2, 1)] = (d020 * d101 + R(2) * d110 * d011) * ir;
    D0[AD_IDX(1, 1, 2)] = (d002 * d110 + R(2) * d101 * d011) * ir;
    const R p222 = d200 * d020 * d002 + R(2) * d110 * d110 * d002 + R(2) * d101 * d101 * d020 +
                   R(2) * d011 * d011 * d200 + R(8) * d110 * d101 * d011;
    D0[AD_IDX(2, 2, 2)] = p222 * ir * ir;
    R D1[27], D2[27];
    for (int k = 0; k < 27; ++k) D1[k] = D0[k];
    lines(D1, 2, [uz](R& a, R& bb, R& c) { unshift_line(a, bb, c, uz); });
    for (int k = 0; k < 27; ++k) D2[k] = D1[k];
    lines(D2, 1, [uy](R& a, R& bb, R& c) { unshift_line(a, bb, c, uy); });
    // ---- reverse
    R fb[27];
    for (int q = 0; q < 27; ++q) fb[q] = fbar[q];
    R ub[3] = {ubar_in[0], ubar_in[1], ubar_in[2]};
    R rsb = R(0), bb = R(0);
    if (P.beta != 0.0f) {  // Guo source
        const R pref = rmax(R(0), R(1) - R(0.5) * R(fmin(fmax(om_off, 0.0), 1.999)));
        const R fa = pref * rho_s * b;
        R fab = R(0);
        for (int q = 0; q < 27; ++q) {
            const R c[3] = {R(float(cxq(q))), R(float(cyq(q))), R(float(czq(q)))};
            const R cu = c[0] * ux + c[1] * uy + c[2] * uz;
            const R wf = wq<R>(q) * fb[q];
            fab = fab + wf * (R(3) * (c[P.g_axis] - u[P.g_axis]) + R(9) * cu * c[P.g_axis]);
            ub[P.g_axis] = ub[P.g_axis] - R(3) * wf * fa;
            const R cub = R(9) * wf * fa * c[P.g_axis];
            for (int k = 0; k < 3; ++k) ub[k] = ub[k] + cub * c[k];
        }
        rsb = rsb + fab * pref * b;
        bb = bb + fab * pref * rho_s;
    }
    // inverse line transforms (forward: x, y, z), transposed in reverse order
    auto inv_t = [](R& o0, R& o1, R& o2) {
        const R a = o0, b1 = o1, c = o2;
        o0 = b1;
        o1 = R(0.5) * (c - a);
        o2 = R(0.5) * (a + c) - b1;
    };
    lines(fb, 2, inv_t);
    lines(fb, 1, inv_t);
    lines(fb, 0, inv_t);
    // unshift passes (forward: z on D0, y on D1, x on D2), transposed in reverse order
    R uxb = R(0), uyb = R(0), uzb = R(0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const int a0 = AD_IDX(0, i, j);
            const R b0 = fb[a0], b1 = fb[a0 + 9], b2 = fb[a0 + 18];
            uxb = uxb + b1 * D2[a0] + b2 * (R(2) * D2[a0 + 9] + R(2) * ux * D2[a0]);
            fb[a0] = b0 + ux * b1 + ux * ux * b2;
            fb[a0 + 9] = b1 + R(2) * ux * b2;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const int a0 = AD_IDX(i, 0, j);
            const R b0 = fb[a0], b1 = fb[a0 + 3], b2 = fb[a0 + 6];
            uyb = uyb + b1 * D1[a0] + b2 * (R(2) * D1[a0 + 3] + R(2) * uy * D1[a0]);
            fb[a0] = b0 + uy * b1 + uy * uy * b2;
            fb[a0 + 3] = b1 + R(2) * uy * b2;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const int a0 = AD_IDX(i, j, 0);
            const R b0 = fb[a0], b1 = fb[a0 + 1], b2 = fb[a0 + 2];
            uzb = uzb + b1 * D0[a0] + b2 * (R(2) * D0[a0 + 1] + R(2) * uz * D0[a0]);
            fb[a0] = b0 + uz * b1 + uz * uz * b2;
            fb[a0 + 1] = b1 + R(2) * uz * b2;
        }
    // closure (orders 4 and 6)
    R d200b = fb[AD_IDX(2, 0, 0)], d020b = fb[AD_IDX(0, 2, 0)], d002b = fb[AD_IDX(0, 0, 2)];
    R d110b = fb[AD_IDX(1, 1, 0)], d101b = fb[AD_IDX(1, 0, 1)], d011b = fb[AD_IDX(0, 1, 1)];
    R irb = R(0);
    rsb = rsb + fb[AD_IDX(0, 0, 0)];
    {
        const R x = fb[AD_IDX(2, 2, 0)];
        d200b = d200b + x * ir * d020; d020b = d020b + x * ir * d200; d110b = d110b + x * ir * R(4) * d110;
        irb = irb + x * (d200 * d020 + R(2) * d110 * d110);
    }
    {
        const R x = fb[AD_IDX(2, 0, 2)];
        d200b = d200b + x * ir * d002; d002b = d002b + x * ir * d200; d101b = d101b + x * ir * R(4) * d101;
        irb = irb + x * (d200 * d002 + R(2) * d101 * d101);
    }
    {
        const R x = fb[AD_IDX(0, 2, 2)];
        d020b = d020b + x * ir * d002; d002b = d002b + x * ir * d020; d011b = d011b + x * ir * R(4) * d011;
        irb = irb + x * (d020 * d002 + R(2) * d011 * d011);
    }
    {
        const R x = fb[AD_IDX(2, 1, 1)];
        d200b = d200b + x * ir * d011; d011b = d011b + x * ir * d200;
        d110b = d110b + x * ir * R(2) * d101; d101b = d101b + x * ir * R(2) * d110;
        irb = irb + x * (d200 * d011 + R(2) * d110 * d101);
    }
    {
        const R x = fb[AD_IDX(1, 2, 1)];
        d020b = d020b + x * ir * d101; d101b = d101b + x * ir * d020;
        d110b = d110b + x * ir * R(2) * d011; d011b = d011b + x * ir * R(2) * d110;
        irb = irb + x * (d020 * d101 + R(2) * d110 * d011);
    }
    {
        const R x = fb[AD_IDX(1, 1, 2)];
        d002b = d002b + x * ir * d110; d110b = d110b + x * ir * d002;
        d101b = d101b + x * ir * R(2) * d011; d011b = d011b + x * ir * R(2) * d101;
        irb = irb + x * (d002 * d110 + R(2) * d101 * d011);
    }
    {
        const R x = fb[AD_IDX(2, 2, 2)];
        const R pb = x * ir * ir;
        irb = irb + x * R(2) * p222 * ir;
        d200b = d200b + pb * (d020 * d002 + R(2) * d011 * d011);
        d020b = d020b + pb * (d200 * d002 + R(2) * d101 * d101);
        d002b = d002b + pb * (d200 * d020 + R(2) * d110 * d110);
        d110b = d110b + pb * (R(4) * d110 * d002 + R(8) * d101 * d011);
        d101b = d101b + pb * (R(4) * d101 * d020 + R(8) * d110 * d011);
        d011b = d011b + pb * (R(4) * d011 * d200 + R(8) * d110 * d101);
    }
    rsb = rsb - irb * ir * ir;
    // relaxation
    const R c200b = d200b * (R(1) - sd), c020b = d020b * (R(1) - sd), c002b = d002b * (R(1) - sd);
    rsb = rsb + (d200b + d020b + d002b) * sd * cs2;
    const R c110b = om * d110b, c101b = om * d101b, c011b = om * d011b;
    // central moments about u
    R mb[kLow];
    for (int k = 0; k < kLow; ++k) mb[k] = R(0);
    mb[4] = c200b;
    mb[1] = mb[1] - R(2) * ux * c200b;
    mb[0] = mb[0] + ux * ux * c200b;
    uxb = uxb + c200b * (R(2) * ux * m[0] - R(2) * m[1]);
    mb[5] = c020b;
    mb[2] = mb[2] - R(2) * uy * c020b;
    mb[0] = mb[0] + uy * uy * c020b;
    uyb = uyb + c020b * (R(2) * uy * m[0] - R(2) * m[2]);
    mb[6] = c002b;
    mb[3] = mb[3] - R(2) * uz * c002b;
    mb[0] = mb[0] + uz * uz * c002b;
    uzb = uzb + c002b * (R(2) * uz * m[0] - R(2) * m[3]);
    R k100b = R(0), k010b = R(0);
    mb[7] = c110b;
    mb[2] = mb[2] - ux * c110b;
    uxb = uxb - m[2] * c110b;
    uyb = uyb - k100 * c110b;
    k100b = k100b - uy * c110b;
    mb[8] = c101b;
    mb[3] = mb[3] - ux * c101b;
    uxb = uxb - m[3] * c101b;
    uzb = uzb - k100 * c101b;
    k100b = k100b - uz * c101b;
    mb[9] = c011b;
    mb[3] = mb[3] - uy * c011b;
    uyb = uyb - m[3] * c011b;
    uzb = uzb - k010 * c011b;
    k010b = k010b - uz * c011b;
    mb[1] = mb[1] + k100b;
    uxb = uxb - m[0] * k100b;
    mb[0] = mb[0] - ux * k100b;
    mb[2] = mb[2] + k010b;
    uyb = uyb - m[0] * k010b;
    mb[0] = mb[0] - uy * k010b;
    ub[0] = ub[0] + uxb;
    ub[1] = ub[1] + uyb;
    ub[2] = ub[2] + uzb;
    // speed cap
    R u1b[3] = {ub[0], ub[1], ub[2]};
    if (capped) {
        const R scb = ub[0] * u1[0] + ub[1] * u1[1] + ub[2] * u1[2];
        const R s2b = scb * (R(-0.5) * sc / s2);
        for (int k = 0; k < 3; ++k) u1b[k] = ub[k] * sc + s2b * R(2) * u1[k];
    }
    // buoyancy and the division by the density
    bb = bb + R(0.5) * u1b[P.g_axis];
    for (int k = 0; k < 3; ++k) {
        mb[k + 1] = mb[k + 1] + u1b[k] / rho_s;
        rsb = rsb - u1b[k] * u0[k] / rho_s;
    }
    *tbar = b_live ? bb * sb * R(P.beta) : R(0);
    mb[0] = mb[0] + (rho_live ? rsb : R(0));
    for (int k = 0; k < kLow; ++k) mbar[k] = mb[k];
}

template <class R> AD_HD void collide(const Params& P, R f[27], R T, R u_out[3]) {
    R m[kLow];
    low_moments(f, m);
    collide_low(P, m, T, f, u_out);
}

AD_HD int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }
// Solid-flag queries: the flag of in-domain cell c (after periodic wrap), which lies
// at offset (dx, dy, dz) from the cell being updated. SolidArray reads the flag
// array; SolidNbrs answers offsets of at most one cell from the cell's 27-bit
// neighbour mask (bit (dx + 1) * 9 + (dy + 1) * 3 + dz + 1, built with the same
// wrap) and reads the array for farther cells.
struct SolidArray {
    const uint8_t* s;
    AD_HD bool operator()(int c, int, int, int) const { return s[c] != 0; }
};
struct SolidNbrs {
    uint32_t nb;
    const uint8_t* s;
    AD_HD bool operator()(int c, int dx, int dy, int dz) const {
        if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return s[c] != 0;
        return (nb >> ((dx + 1) * 9 + (dy + 1) * 3 + dz + 1)) & 1u;
    }
};
AD_HD SolidArray solid_query(const uint8_t* s) { return SolidArray{s}; }
AD_HD SolidNbrs solid_query(SolidNbrs q) { return q; }


AD_HD int cidx(const Params& P, int x, int y, int z) { return (x * P.ny + y) * P.nz + z; }

// incoming population q of fluid cell (x, y, z) from the previous post-collision
// field F (27 planes of `cells`): pull, halfway bounce-back at solid sources,
// velocity inlet (equilibrium at the cell's inflow parameter) on x-, convective
// outflow f1 + 0.15 (f1 - f2) on x+ (ref: jni.cpp:1494-1561)
template <class R, class Solid>
AD_HD R incoming(const Params& P, const R* F, Solid solid_in, const R* uin, int x, int y, int z, int q) {
    const auto solid = solid_query(solid_in);
    const int N = P.nx * P.ny * P.nz;
    const int cx = cxq(q);
    const int sx = x - cx, sy = wrap(y - cyq(q), P.ny), sz = wrap(z - czq(q), P.nz);
    const int c = cidx(P, x, y, z);
    if (sx < 0) {
        const R* u = uin + 3 * (y * P.nz + z);
        return feq<R>(q, R(P.rho_in), u[0], u[1], u[2]);
    }
    if (sx >= P.nx) {
        const int x1 = x + cx, x2 = x + 2 * cx;
        if (solid(cidx(P, x1, y, z), cx, 0, 0)) return F[(size_t)oppq(q) * N + c];
        const R f1 = F[(size_t)q * N + cidx(P, x1, y, z)];
        if (solid(cidx(P, x2, y, z), 2 * cx, 0, 0)) return f1;
        const R f2 = F[(size_t)q * N + cidx(P, x2, y, z)];
        return f1 + R(0.15) * (f1 - f2);
    }
    const int s = cidx(P, sx, sy, sz);
    if (solid(s, -cx, -cyq(q), -czq(q))) return F[(size_t)oppq(q) * N + c];
    return F[(size_t)q * N + s];
}

// temperature at (x, y, z) for the stencil of a cell whose own value is self:
// periodic in y and z; outside in x or solid -> self
template <class R, class Solid>
AD_HD R t_at(const Params& P, const R* T, const Solid& solid, int x, int y, int z, int dx, int dy, int dz, R self,
             int* idx) {
    y = wrap(y, P.ny);
    z = wrap(z, P.nz);
    if (x < 0 || x >= P.nx) {
        *idx = -1;
        return self;
    }
    const int c = cidx(P, x, y, z);
    if (solid(c, dx, dy, dz)) {
        *idx = -1;
        return self;
    }
    *idx = c;
    return T[c];
}

// first-order semi-Lagrangian advection + FTCS diffusion + source, clamped to
// +-1 (ref: jni.cpp:2145-2188). The 15 stencil indices (8 corners, 6 neighbours,
// the cell) and weights go to idx/w (-1: the read fell back to the cell itself).
template <class R, class SolidIn>
AD_HD R temperature(const Params& P, const R* T, SolidIn solid_in, R src, int x, int y, int z, R ux, R uy, R uz,
                    int idx[14], R w_corner[8], R* pre_clamp) {
    const auto solid = solid_query(solid_in);
    const int c = cidx(P, x, y, z);
    const R tc = T[c];
    const R px = R(float(x)) - ux, py = R(float(y)) - uy, pz = R(float(z)) - uz;
    const int x0 = (int)floor(val(px)), y0 = (int)floor(val(py)), z0 = (int)floor(val(pz));
    const R fx = px - R(float(x0)), fy = py - R(float(y0)), fz = pz - R(float(z0));
    R cv[8];
    for (int k = 0; k < 8; ++k)
        cv[k] = t_at(P, T, solid, x0 + (k >> 2), y0 + ((k >> 1) & 1), z0 + (k & 1), x0 + (k >> 2) - x,
                     y0 + ((k >> 1) & 1) - y, z0 + (k & 1) - z, tc, &idx[k]);
    // weights of corner k = (dx, dy, dz)
    for (int k = 0; k < 8; ++k) {
        const R wx = (k >> 2) ? fx : R(1) - fx;
        const R wy = ((k >> 1) & 1) ? fy : R(1) - fy;
        const R wz = (k & 1) ? fz : R(1) - fz;
        w_corner[k] = wx * wy * wz;
    }
    R adv = R(0);
    for (int k = 0; k < 8; ++k) adv = adv + w_corner[k] * cv[k];
    const int nb[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    R sum = R(0);
    for (int k = 0; k < 6; ++k)
        sum = sum + t_at(P, T, solid, x + nb[k][0], y + nb[k][1], z + nb[k][2], nb[k][0], nb[k][1], nb[k][2], tc,
                         &idx[8 + k]);
    const R pre = adv + (R(P.alpha) * (sum - R(6) * tc) + src);
    *pre_clamp = pre;
    return rclamp(pre, -1.0f, 1.0f);
}

// heat source of cell (x, y, z): trilinear sample of the m^3 parameter field at
// the cell centre, clamped to +-0.006; lo[3] and t[3] receive the coarse base
// node and fractions
AD_HD void source_weights(int n, int m, int x, float* pos) {
    const float p = ((float)x + 0.5f) * (float)m / (float)n - 0.5f;
    *pos = fminf(fmaxf(p, 0.0f), (float)(m - 1));
}

}  // namespace adj
}  // namespace aero
