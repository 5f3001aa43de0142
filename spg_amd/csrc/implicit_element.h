// Full element gradients of the five spg energies (all stencil slots), templated on the scalar, and the
// element Hessian as the forward-mode derivative of that gradient.
//
// The reference gets both from TinyAD's second-order dual numbers (energyGradient / energyHessian,
// spg/include/spg/sim/energy/energy.h:268-270,348-362; the spring from its own closed form,
// springEnergy.cpp:101-134). Here the gradient is written out per energy; instantiating it with a
// first-order dual number (value, one tangent) and seeding the tangent with the j-th unit vector gives
// column j of the Hessian. 3N such passes give the full 3N x 3N matrix with few registers, in FP64.
#pragma once

#include "element.h"

namespace spgvbd {

// Real square root and atan2 by scalar type (float for the FP32 vertex solve, double elsewhere).
SPG_HD float realSqrt(float a) { return sqrtf(a); }
SPG_HD double realSqrt(double a) { return ::sqrt(a); }
SPG_HD float realAtan2(float y, float x) { return atan2f(y, x); }
SPG_HD double realAtan2(double y, double x) { return ::atan2(y, x); }

// First-order dual number: value and one directional derivative. It lives in its own namespace so
// that its sqrt/atan2 are found only by argument-dependent lookup and never hide the float/double
// functions the other closed forms of this namespace call.
namespace ad {
template <class R>
struct DualT {
    R v, d;
    SPG_HD DualT() : v(0), d(0) {}
    SPG_HD DualT(R a) : v(a), d(0) {}  // NOLINT: constants inside expressions
    SPG_HD DualT(R a, R b) : v(a), d(b) {}
};
template <class R>
SPG_HD DualT<R> operator+(DualT<R> a, DualT<R> b)
{
    return {a.v + b.v, a.d + b.d};
}
template <class R>
SPG_HD DualT<R> operator-(DualT<R> a, DualT<R> b)
{
    return {a.v - b.v, a.d - b.d};
}
template <class R>
SPG_HD DualT<R> operator-(DualT<R> a)
{
    return {-a.v, -a.d};
}
template <class R>
SPG_HD DualT<R> operator*(DualT<R> a, DualT<R> b)
{
    return {a.v * b.v, a.v * b.d + a.d * b.v};
}
template <class R>
SPG_HD DualT<R> operator/(DualT<R> a, DualT<R> b)
{
    const R q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
template <class R>
SPG_HD DualT<R> sqrt(DualT<R> a)
{
    const R s = realSqrt(a.v);
    return {s, a.d / (R(2) * s)};
}
template <class R>
SPG_HD DualT<R> atan2(DualT<R> y, DualT<R> x)
{
    const R r2 = x.v * x.v + y.v * y.v;
    return {realAtan2(y.v, x.v), (x.v * y.d - y.v * x.d) / r2};
}
template <class R>
SPG_HD R valueOf(DualT<R> a)
{
    return a.v;
}
// Forward mode with three tangents at once: one pass over an element's gradient gives the three columns of a
// vertex's 3x3 block (the tangents seeded with the unit vectors of that vertex's coordinates).
template <class R>
struct Dual3T {
    R v, d[3];
    SPG_HD Dual3T() : v(0), d{0, 0, 0} {}
    SPG_HD Dual3T(R a) : v(a), d{0, 0, 0} {}  // NOLINT: constants inside expressions
    SPG_HD Dual3T(R a, R d0, R d1, R d2) : v(a), d{d0, d1, d2} {}
};
template <class R>
SPG_HD Dual3T<R> operator+(Dual3T<R> a, Dual3T<R> b)
{
    return {a.v + b.v, a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]};
}
template <class R>
SPG_HD Dual3T<R> operator-(Dual3T<R> a, Dual3T<R> b)
{
    return {a.v - b.v, a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]};
}
template <class R>
SPG_HD Dual3T<R> operator-(Dual3T<R> a)
{
    return {-a.v, -a.d[0], -a.d[1], -a.d[2]};
}
template <class R>
SPG_HD Dual3T<R> operator*(Dual3T<R> a, Dual3T<R> b)
{
    return {a.v * b.v, a.v * b.d[0] + a.d[0] * b.v, a.v * b.d[1] + a.d[1] * b.v, a.v * b.d[2] + a.d[2] * b.v};
}
template <class R>
SPG_HD Dual3T<R> operator/(Dual3T<R> a, Dual3T<R> b)
{
    const R q = a.v / b.v, ib = R(1) / b.v;
    return {q, (a.d[0] - q * b.d[0]) * ib, (a.d[1] - q * b.d[1]) * ib, (a.d[2] - q * b.d[2]) * ib};
}
template <class R>
SPG_HD Dual3T<R> sqrt(Dual3T<R> a)
{
    const R s = realSqrt(a.v);
    const R h = R(1) / (R(2) * s);
    return {s, a.d[0] * h, a.d[1] * h, a.d[2] * h};
}
template <class R>
SPG_HD Dual3T<R> atan2(Dual3T<R> y, Dual3T<R> x)
{
    const R ir2 = R(1) / (x.v * x.v + y.v * y.v);
    return {realAtan2(y.v, x.v), (x.v * y.d[0] - y.v * x.d[0]) * ir2, (x.v * y.d[1] - y.v * x.d[1]) * ir2,
            (x.v * y.d[2] - y.v * x.d[2]) * ir2};
}
template <class R>
SPG_HD R valueOf(Dual3T<R> a)
{
    return a.v;
}
}  // namespace ad
using Dual1 = ad::DualT<double>;
using Dual1f = ad::DualT<float>;
SPG_HD float valueOf(float a) { return a; }
SPG_HD double valueOf(double a) { return a; }
using ad::valueOf;

template <class T>
SPG_HD V3<T> v3c(const double *p)
{
    return {T(p[0]), T(p[1]), T(p[2])};
}

// springEnergy.cpp:101-115: s = x0 - x1, grad_0 = k (|s| - L0) s/|s|, grad_1 = -grad_0. p: (L0, k).
template <class T>
SPG_HD void fullGradSpring(const V3<T> *x, const double *p, V3<T> *g)
{
    const V3<T> s = x[0] - x[1];
    const T n = sqrt(dot(s, s));
    g[0] = ((T(p[1]) * (n - T(p[0]))) / n) * s;
    g[1] = -g[0];
}

// springAnchorEnergy.cpp:21-26: E = 1/2 k |x - a|^2, grad = k (x - a). p: (a, k).
template <class T>
SPG_HD void fullGradAnchor(const V3<T> *x, const double *p, V3<T> *g)
{
    g[0] = T(p[3]) * (x[0] - v3c<T>(p));
}

// membraneStvkEnergy.cpp:9-28 with energy.h:336-341: E = 1/2 C^T K C, C = (E00, E11, 2 E01) of the Green
// strain of F = [x1-x0, x2-x0] Dinv. With sigma = K C and S = [[s0, s2], [s2, s1]]: grad_a = F S gamma_a,
// gamma_1, gamma_2 the rows of Dinv and gamma_0 = -(gamma_1 + gamma_2). p: Dinv (2x2), K (3x3), row-major.
template <class T>
SPG_HD void fullGradStvk(const V3<T> *x, const double *p, V3<T> *g)
{
    const V3<T> u = x[1] - x[0], v = x[2] - x[0];
    const V3<T> f0 = T(p[0]) * u + T(p[2]) * v;
    const V3<T> f1 = T(p[1]) * u + T(p[3]) * v;
    const T C[3] = {T(0.5) * (dot(f0, f0) - T(1)), T(0.5) * (dot(f1, f1) - T(1)), dot(f0, f1)};
    T s[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) s[i] = T(p[4 + 3 * i]) * C[0] + T(p[5 + 3 * i]) * C[1] + T(p[6 + 3 * i]) * C[2];
    // F S = [f0 s0 + f1 s2, f0 s2 + f1 s1]
    const V3<T> c0 = s[0] * f0 + s[2] * f1, c1 = s[2] * f0 + s[1] * f1;
    const double ga[3][2] = {{-(p[0] + p[2]), -(p[1] + p[3])}, {p[0], p[1]}, {p[2], p[3]}};
#pragma unroll
    for (int a = 0; a < 3; ++a) g[a] = T(ga[a][0]) * c0 + T(ga[a][1]) * c1;
}

// discreteBendingEnergy.cpp:9-31 with energy.h:336-341: E = 1/2 k (theta - theta0)^2, the dihedral-angle
// gradients of the hinge x0-x1 with wings x2, x3 (e = x1 - x0, a = e x (x2 - x1), b = e x (x3 - x1)).
// p: (theta0, k).
template <class T>
SPG_HD void fullGradBending(const V3<T> *x, const double *p, V3<T> *g)
{
    const V3<T> e = x[1] - x[0], pp = x[2] - x[1], q = x[3] - x[1];
    const V3<T> a = cross(e, pp), b = cross(e, q);
    const T L = sqrt(dot(e, e));
    const T iL = T(1) / L;
    const T pe = dot(pp, e) * iL, qe = dot(q, e) * iL;
    const T ia2 = T(1) / dot(a, a), ib2 = T(1) / dot(b, b);
    const T theta = atan2(-dot(cross(a, b), e) * iL, -dot(a, b));
    const T kc = T(p[1]) * (theta - T(p[0]));
    g[0] = kc * ((-pe * ia2) * a + (qe * ib2) * b);
    g[1] = kc * (((pe + L) * ia2) * a + (-(qe + L) * ib2) * b);
    g[2] = kc * ((-L * ia2) * a);
    g[3] = kc * ((L * ib2) * b);
}

// stableNeoHookeanEnergy.cpp:33-61: E = 1/2 lambda (J - alpha)^2 + 1/2 mu (|F|^2 - 3), F = [x1-x0, x2-x0,
// x3-x0] Dinv, P = lambda (J - alpha) cof(F) + mu F, grad_{m+1} = sum_c P_c Dinv(m, c), grad_0 = -sum.
// p: Dinv (3x3 row-major), lambda, mu, alpha.
template <class T>
SPG_HD void fullGradSnh(const V3<T> *x, const double *p, V3<T> *g)
{
    const V3<T> u = x[1] - x[0], v = x[2] - x[0], w = x[3] - x[0];
    const V3<T> f0 = T(p[0]) * u + T(p[3]) * v + T(p[6]) * w;
    const V3<T> f1 = T(p[1]) * u + T(p[4]) * v + T(p[7]) * w;
    const V3<T> f2 = T(p[2]) * u + T(p[5]) * v + T(p[8]) * w;
    const V3<T> k0 = cross(f1, f2), k1 = cross(f2, f0), k2 = cross(f0, f1);
    const T J = dot(f0, k0);
    const T lj = T(p[9]) * (J - T(p[11]));
    const T mu = T(p[10]);
    const V3<T> P0 = lj * k0 + mu * f0, P1 = lj * k1 + mu * f1, P2 = lj * k2 + mu * f2;
    g[1] = T(p[0]) * P0 + T(p[1]) * P1 + T(p[2]) * P2;
    g[2] = T(p[3]) * P0 + T(p[4]) * P1 + T(p[5]) * P2;
    g[3] = T(p[6]) * P0 + T(p[7]) * P1 + T(p[8]) * P2;
    g[0] = -(g[1] + g[2] + g[3]);
}

// ---- The further particle energies (E = 1/2 C^T Keff C with a diagonal Keff unless noted) ----

// Slot weights of a triangle's F = [x1 - x0, x2 - x0] Dinv = [f0 f1]: f0 = sum a_s x_s, f1 = sum b_s x_s.
template <class T>
SPG_HD void triangleF(const V3<T> *x, const double *p, V3<T> &f0, V3<T> &f1)
{
    const V3<T> u = x[1] - x[0], v = x[2] - x[0];
    f0 = T(p[0]) * u + T(p[2]) * v;
    f1 = T(p[1]) * u + T(p[3]) * v;
}
template <class T>
SPG_HD void triangleScatter(const double *p, V3<T> dEdf0, V3<T> dEdf1, V3<T> *g)
{
    const double ga[3][2] = {{-(p[0] + p[2]), -(p[1] + p[3])}, {p[0], p[1]}, {p[2], p[3]}};
#pragma unroll
    for (int a = 0; a < 3; ++a) g[a] = T(ga[a][0]) * dEdf0 + T(ga[a][1]) * dEdf1;
}
// d/dv of 1/2 k (|v| - 1)^2 = k (|v| - 1) v / |v|
template <class T>
SPG_HD V3<T> unitStretch(V3<T> v, T k)
{
    const T n = sqrt(dot(v, v));
    return ((k * (n - T(1))) / n) * v;
}

// springContinuumEnergy.cpp:9-18: C = |d| / L0 - 1, d = x0 - x1. p: (L0, k L0).
template <class T>
SPG_HD void fullGradSpringContinuum(const V3<T> *x, const double *p, V3<T> *g)
{
    const V3<T> d = x[0] - x[1];
    const T n = sqrt(dot(d, d));
    const T L0 = T(p[0]);
    const T C = n / L0 - T(1);
    g[0] = ((T(p[1]) * C) / (L0 * n)) * d;
    g[1] = -g[0];
}
// springSquaredContinuumEnergy.cpp:9-19: C = |d|^2 / L0^2 - 1. p: (L0, k L0).
template <class T>
SPG_HD void fullGradSpringSquaredContinuum(const V3<T> *x, const double *p, V3<T> *g)
{
    const V3<T> d = x[0] - x[1];
    const T L2 = T(p[0]) * T(p[0]);
    const T C = dot(d, d) / L2 - T(1);
    g[0] = ((T(2) * T(p[1]) * C) / L2) * d;
    g[1] = -g[0];
}
// membraneBaraffWitkinEnergy.cpp:9-26: C = (|f0| - 1, |f1| - 1, f0 . f1). p: Dinv (4), Keff diagonal (3).
template <class T>
SPG_HD void fullGradMembraneBw(const V3<T> *x, const double *p, V3<T> *g)
{
    V3<T> f0, f1;
    triangleF(x, p, f0, f1);
    const T kb = T(p[6]) * dot(f0, f1);
    triangleScatter(p, unitStretch(f0, T(p[4])) + kb * f1, unitStretch(f1, T(p[5])) + kb * f0, g);
}
// membraneChoiEnergy.cpp:9-28: C = (|f0| - 1, |f1| - 1, |gp| - 1, |gm| - 1), gp = (f0 + f1) / sqrt 2,
// gm = (f0 - f1) / sqrt 2. p: Dinv (4), Keff diagonal (4).
template <class T>
SPG_HD void fullGradMembraneChoi(const V3<T> *x, const double *p, V3<T> *g)
{
    V3<T> f0, f1;
    triangleF(x, p, f0, f1);
    const T r = T(0.70710678118654752440);
    const V3<T> pp = unitStretch(r * (f0 + f1), T(p[6])), pm = unitStretch(r * (f0 - f1), T(p[7]));
    triangleScatter(p, unitStretch(f0, T(p[4])) + r * (pp + pm), unitStretch(f1, T(p[5])) + r * (pp - pm), g);
}
// quadraticBendingEnergy.cpp:9-30: C = sum c_s x_s (a 3-vector), E = 1/2 keff |C|^2. p: c0..c3, keff.
template <class T>
SPG_HD void fullGradBendingQuadratic(const V3<T> *x, const double *p, V3<T> *g)
{
    const V3<T> C = T(p[0]) * x[0] + T(p[1]) * x[1] + T(p[2]) * x[2] + T(p[3]) * x[3];
#pragma unroll
    for (int s = 0; s < 4; ++s) g[s] = (T(p[4]) * T(p[s])) * C;
}
// baraffWitkinBendingEnergy.cpp:9-36: C = theta - theta0, theta = atan2(s0 Y, X) with X = n1 . n2,
// Y = (n1 x n2) . e0, n1 = e0 x e3, n2 = -(e0 x e4) and s0 = 1/|e0| held constant (the reference
// differentiates through none of its normalisations; the common factor 1/(|n1||n2|) cancels in atan2).
// grad theta = (X grad Y - Y grad X) / (X^2 + Y^2), with grad(v . n1) and grad(v . n2) per vertex below.
template <class T>
SPG_HD void gradDotN1(V3<T> v, V3<T> e0, V3<T> e3, V3<T> *out)  // n1 = (x1 - x0) x (x2 - x1)
{
    const V3<T> a = cross(e3, v), b = cross(v, e0);
    out[0] = -a;
    out[1] = a - b;
    out[2] = b;
    out[3] = {T(0), T(0), T(0)};
}
template <class T>
SPG_HD void gradDotN2(V3<T> v, V3<T> e0, V3<T> e4, V3<T> *out)  // n2 = -(x1 - x0) x (x3 - x1)
{
    const V3<T> a = cross(e4, v), b = cross(v, e0);
    out[0] = a;
    out[1] = b - a;
    out[2] = {T(0), T(0), T(0)};
    out[3] = -b;
}
template <class T>
SPG_HD void fullGradBendingBw(const V3<T> *x, const double *p, V3<T> *g)
{
    const V3<T> e0 = x[1] - x[0], e3 = x[2] - x[1], e4 = x[3] - x[1];
    const V3<T> n1 = cross(e0, e3), n2 = -cross(e0, e4);
    const T s0 = T(1 / realSqrt(valueOf(dot(e0, e0))));
    const V3<T> m = cross(n1, n2);
    const T X = dot(n1, n2), Y = s0 * dot(m, e0);
    V3<T> dx1[4], dx2[4], dy1[4], dy2[4];
    gradDotN1(n2, e0, e3, dx1);
    gradDotN2(n1, e0, e4, dx2);
    gradDotN1(cross(n2, e0), e0, e3, dy1);
    gradDotN2(cross(e0, n1), e0, e4, dy2);
    const T theta = atan2(Y, X);
    const T w = (T(p[1]) * (theta - T(p[0]))) / (X * X + Y * Y);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        V3<T> gy = dy1[s] + dy2[s];
        if (s == 0) gy = gy - m;
        if (s == 1) gy = gy + m;
        g[s] = w * (X * (s0 * gy) - Y * (dx1[s] + dx2[s]));
    }
}
// F = [x1 - x0, x2 - x0, x3 - x0] Dinv as columns f0, f1, f2 (Dinv row-major in p[0..8]); grad of an energy
// with first Piola-Kirchhoff columns P_c: grad_{m+1} = sum_c P_c Dinv(m, c), grad_0 = -sum.
template <class T>
SPG_HD void tetF(const V3<T> *x, const double *p, V3<T> *f)
{
    const V3<T> u = x[1] - x[0], v = x[2] - x[0], w = x[3] - x[0];
#pragma unroll
    for (int c = 0; c < 3; ++c) f[c] = T(p[c]) * u + T(p[3 + c]) * v + T(p[6 + c]) * w;
}
template <class T>
SPG_HD void tetScatter(const double *p, const V3<T> *P, V3<T> *g)
{
    g[1] = T(p[0]) * P[0] + T(p[1]) * P[1] + T(p[2]) * P[2];
    g[2] = T(p[3]) * P[0] + T(p[4]) * P[1] + T(p[5]) * P[2];
    g[3] = T(p[6]) * P[0] + T(p[7]) * P[1] + T(p[8]) * P[2];
    g[0] = -(g[1] + g[2] + g[3]);
}
// stvkEnergy.cpp:41-63 (the energy form): E = 1/2 lambda (tr E)^2 + mu tr(E^T E), E = 1/2 (F^T F - I);
// P = F S with S = lambda tr(E) I + 2 mu E. p: Dinv (9), lambda, mu.
template <class T>
SPG_HD void fullGradStvkTet(const V3<T> *x, const double *p, V3<T> *g)
{
    V3<T> f[3];
    tetF(x, p, f);
    T E[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) E[i][j] = T(0.5) * (dot(f[i], f[j]) - T(i == j ? 1 : 0));
    const T lt = T(p[9]) * (E[0][0] + E[1][1] + E[2][2]);
    const T m2 = T(2) * T(p[10]);
    V3<T> P[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const T s0 = m2 * E[0][c] + (c == 0 ? lt : T(0));
        const T s1 = m2 * E[1][c] + (c == 1 ? lt : T(0));
        const T s2 = m2 * E[2][c] + (c == 2 ? lt : T(0));
        P[c] = s0 * f[0] + s1 * f[1] + s2 * f[2];
    }
    tetScatter(p, P, g);
}
// stableSquaredNeoHookeanEnergy.cpp:9-31: C = (det F - 1, tr F^T F - 3);
// P = lambda (J - 1) cof(F) + 2 mu (|F|^2 - 3) F. p: Dinv (9), lambda, mu.
template <class T>
SPG_HD void fullGradSsnh(const V3<T> *x, const double *p, V3<T> *g)
{
    V3<T> f[3];
    tetF(x, p, f);
    const V3<T> k0 = cross(f[1], f[2]), k1 = cross(f[2], f[0]), k2 = cross(f[0], f[1]);
    const T lj = T(p[9]) * (dot(f[0], k0) - T(1));
    const T mi = T(2) * T(p[10]) * (dot(f[0], f[0]) + dot(f[1], f[1]) + dot(f[2], f[2]) - T(3));
    const V3<T> P[3] = {lj * k0 + mi * f[0], lj * k1 + mi * f[1], lj * k2 + mi * f[2]};
    tetScatter(p, P, g);
}

// Type ids (spg_vbd.h): 1 spring, 2 membrane, 3 bending, 4 SNH, 5 anchor, 6 continuum spring, 7 squared
// continuum spring, 8 Baraff-Witkin membrane, 9 Choi membrane, 10 quadratic bending, 11 Baraff-Witkin
// bending, 12 StVK tetrahedron, 13 stable squared Neo-Hookean.
SPG_HD constexpr int fullStencilSize(int type)
{
    return type == 5 ? 1 : (type == 1 || type == 6 || type == 7) ? 2 : (type == 2 || type == 8 || type == 9) ? 3 : 4;
}
// prepared parameters per element
SPG_HD constexpr int fullParams(int type)
{
    switch (type) {
    case 1: return 2;
    case 2: return 13;
    case 3: return 2;
    case 4: return 12;
    case 5: return 4;
    case 6: return 2;
    case 7: return 2;
    case 8: return 7;
    case 9: return 8;
    case 10: return 5;
    case 11: return 2;
    case 12: return 11;
    case 13: return 11;
    }
    return 0;
}

// One further energy's gradient with the type fixed at compile time (types 6 ... 13).
template <int TYPE, class T>
SPG_HD void fullGradOf(const V3<T> *x, const double *p, V3<T> *g)
{
    if constexpr (TYPE == 6)
        fullGradSpringContinuum(x, p, g);
    else if constexpr (TYPE == 7)
        fullGradSpringSquaredContinuum(x, p, g);
    else if constexpr (TYPE == 8)
        fullGradMembraneBw(x, p, g);
    else if constexpr (TYPE == 9)
        fullGradMembraneChoi(x, p, g);
    else if constexpr (TYPE == 10)
        fullGradBendingQuadratic(x, p, g);
    else if constexpr (TYPE == 11)
        fullGradBendingBw(x, p, g);
    else if constexpr (TYPE == 12)
        fullGradStvkTet(x, p, g);
    else
        fullGradSsnh(x, p, g);
}

// The energies of stencil size NS (selected at compile time, so no path indexes past x[NS - 1]).
template <int NS, class T>
SPG_HD void fullGrad(int type, const V3<T> *x, const double *p, V3<T> *g)
{
    if constexpr (NS == 1) {
        fullGradAnchor(x, p, g);
    } else if constexpr (NS == 2) {
        if (type == 6)
            fullGradSpringContinuum(x, p, g);
        else if (type == 7)
            fullGradSpringSquaredContinuum(x, p, g);
        else
            fullGradSpring(x, p, g);
    } else if constexpr (NS == 3) {
        if (type == 8)
            fullGradMembraneBw(x, p, g);
        else if (type == 9)
            fullGradMembraneChoi(x, p, g);
        else
            fullGradStvk(x, p, g);
    } else {
        switch (type) {
        case 3: fullGradBending(x, p, g); break;
        case 10: fullGradBendingQuadratic(x, p, g); break;
        case 11: fullGradBendingBw(x, p, g); break;
        case 12: fullGradStvkTet(x, p, g); break;
        case 13: fullGradSsnh(x, p, g); break;
        default: fullGradSnh(x, p, g);
        }
    }
}

// Gradient (3N) and symmetrised Hessian (3N x 3N, row-major) of one element at xs (3N doubles).
template <int NS>
SPG_HD void elementGradHessFull(int type, const double *xs, const double *p, double *grad, double *hess)
{
    constexpr int N = 3 * NS;
    for (int j = 0; j < N; ++j) {
        V3<Dual1> x[NS], g[NS];
        for (int s = 0; s < NS; ++s)
            x[s] = {Dual1(xs[3 * s], 3 * s == j ? 1.0 : 0.0), Dual1(xs[3 * s + 1], 3 * s + 1 == j ? 1.0 : 0.0),
                    Dual1(xs[3 * s + 2], 3 * s + 2 == j ? 1.0 : 0.0)};
        fullGrad<NS, Dual1>(type, x, p, g);
        for (int s = 0; s < NS; ++s) {
            hess[(3 * s + 0) * N + j] = g[s].x.d;
            hess[(3 * s + 1) * N + j] = g[s].y.d;
            hess[(3 * s + 2) * N + j] = g[s].z.d;
            if (j == 0 && grad) {
                grad[3 * s + 0] = g[s].x.v;
                grad[3 * s + 1] = g[s].y.v;
                grad[3 * s + 2] = g[s].z.v;
            }
        }
    }
    for (int r = 0; r < N; ++r)
        for (int c = r + 1; c < N; ++c) {
            const double m = 0.5 * (hess[r * N + c] + hess[c * N + r]);
            hess[r * N + c] = m;
            hess[c * N + r] = m;
        }
}

}  // namespace spgvbd
