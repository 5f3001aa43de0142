// FP64 CPU oracle for the spg VBD step — TEST INFRASTRUCTURE ONLY (see vbd_oracle.h).
// Every function cites the reference file:line it restates. Nothing here is part of the product.
#include "vbd_oracle.h"

#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

// ---------------------------------------------------------------------------------------------
// Forward-mode dual numbers with dense gradient and (optionally) dense Hessian over N inputs.
// Textbook second-order forward AD (Griewank & Walther, "Evaluating Derivatives", ch. 13):
// every elementary op propagates value, gradient and Hessian by the chain rule, like
// TinyAD::Scalar<N, double, hess> does for the reference (energy.h:84-86,348-362).
// ---------------------------------------------------------------------------------------------
template <int N, bool H>
struct Dual {
    double v;
    double g[N];
    double h[H ? N * N : 1];
    Dual() = default;
    Dual(double c) : v(c)  // NOLINT: implicit on purpose (constants inside expressions)
    {
        std::memset(g, 0, sizeof(g));
        if (H) std::memset(h, 0, sizeof(h));
    }
    static Dual var(double c, int i)
    {
        Dual r(c);
        r.g[i] = 1.0;
        return r;
    }
};

template <int N, bool H>
inline Dual<N, H> chain1(const Dual<N, H> &u, double f, double fp, double fpp)
{
    Dual<N, H> r;
    r.v = f;
    for (int i = 0; i < N; ++i) r.g[i] = fp * u.g[i];
    if (H) {
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) r.h[i * N + j] = fp * u.h[i * N + j] + fpp * u.g[i] * u.g[j];
    }
    return r;
}

template <int N, bool H>
inline Dual<N, H> operator+(const Dual<N, H> &a, const Dual<N, H> &b)
{
    Dual<N, H> r;
    r.v = a.v + b.v;
    for (int i = 0; i < N; ++i) r.g[i] = a.g[i] + b.g[i];
    if (H)
        for (int i = 0; i < N * N; ++i) r.h[i] = a.h[i] + b.h[i];
    return r;
}
template <int N, bool H>
inline Dual<N, H> operator-(const Dual<N, H> &a, const Dual<N, H> &b)
{
    Dual<N, H> r;
    r.v = a.v - b.v;
    for (int i = 0; i < N; ++i) r.g[i] = a.g[i] - b.g[i];
    if (H)
        for (int i = 0; i < N * N; ++i) r.h[i] = a.h[i] - b.h[i];
    return r;
}
template <int N, bool H>
inline Dual<N, H> operator-(const Dual<N, H> &a)
{
    return chain1(a, -a.v, -1.0, 0.0);
}
template <int N, bool H>
inline Dual<N, H> operator*(const Dual<N, H> &a, const Dual<N, H> &b)
{
    Dual<N, H> r;
    r.v = a.v * b.v;
    for (int i = 0; i < N; ++i) r.g[i] = a.v * b.g[i] + b.v * a.g[i];
    if (H) {
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j)
                r.h[i * N + j] = a.v * b.h[i * N + j] + b.v * a.h[i * N + j] + a.g[i] * b.g[j] + b.g[i] * a.g[j];
    }
    return r;
}
template <int N, bool H>
inline Dual<N, H> operator*(double c, const Dual<N, H> &a)
{
    return chain1(a, c * a.v, c, 0.0);
}
template <int N, bool H>
inline Dual<N, H> operator*(const Dual<N, H> &a, double c)
{
    return chain1(a, c * a.v, c, 0.0);
}
template <int N, bool H>
inline Dual<N, H> operator+(const Dual<N, H> &a, double c)
{
    Dual<N, H> r = a;
    r.v += c;
    return r;
}
template <int N, bool H>
inline Dual<N, H> operator-(const Dual<N, H> &a, double c)
{
    Dual<N, H> r = a;
    r.v -= c;
    return r;
}
template <int N, bool H>
inline Dual<N, H> operator/(const Dual<N, H> &a, double c)
{
    Dual<N, H> r;
    r.v = a.v / c;
    for (int i = 0; i < N; ++i) r.g[i] = a.g[i] / c;
    if (H)
        for (int i = 0; i < N * N; ++i) r.h[i] = a.h[i] / c;
    return r;
}
template <int N, bool H>
inline Dual<N, H> operator/(const Dual<N, H> &a, const Dual<N, H> &b)
{
    const double iv = 1.0 / b.v;
    return a * chain1(b, iv, -iv * iv, 2.0 * iv * iv * iv);
}
template <int N, bool H>
inline Dual<N, H> sqrt(const Dual<N, H> &u)
{
    const double s = std::sqrt(u.v);
    return chain1(u, s, 0.5 / s, -0.25 / (s * u.v));
}
// atan2(y, x): partials f_y = x/r^2, f_x = -y/r^2, f_yy = -2xy/r^4, f_xx = 2xy/r^4, f_xy = (y^2-x^2)/r^4
template <int N, bool H>
inline Dual<N, H> atan2(const Dual<N, H> &y, const Dual<N, H> &x)
{
    Dual<N, H> r;
    const double r2 = x.v * x.v + y.v * y.v;
    const double r4 = r2 * r2;
    const double fy = x.v / r2, fx = -y.v / r2;
    r.v = std::atan2(y.v, x.v);
    for (int i = 0; i < N; ++i) r.g[i] = fy * y.g[i] + fx * x.g[i];
    if (H) {
        const double fyy = -2.0 * x.v * y.v / r4, fxx = 2.0 * x.v * y.v / r4, fxy = (y.v * y.v - x.v * x.v) / r4;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j)
                r.h[i * N + j] = fy * y.h[i * N + j] + fx * x.h[i * N + j] + fyy * y.g[i] * y.g[j] +
                                 fxx * x.g[i] * x.g[j] + fxy * (x.g[i] * y.g[j] + y.g[i] * x.g[j]);
    }
    return r;
}
inline double sqrt(double x) { return std::sqrt(x); }
inline double atan2(double y, double x) { return std::atan2(y, x); }
inline double valueOf(double s) { return s; }
template <int N, bool H>
inline double valueOf(const Dual<N, H> &s)
{
    return s.v;
}

template <class S>
struct V3 {
    S x, y, z;
};
template <class S>
V3<S> sub(const V3<S> &a, const V3<S> &b)
{
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class S>
S dot(const V3<S> &a, const V3<S> &b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class S>
V3<S> cross(const V3<S> &a, const V3<S> &b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class S>
V3<S> add(const V3<S> &a, const V3<S> &b)
{
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class S, class T>
V3<S> scale(const V3<S> &a, const T &c)
{
    return {a.x * c, a.y * c, a.z * c};
}
template <class S>
V3<S> normalized(const V3<S> &a)
{
    const S n = sqrt(dot(a, a));
    return {a.x / n, a.y / n, a.z / n};
}
template <class S>
V3<S> neg(const V3<S> &a)
{
    return {-a.x, -a.y, -a.z};
}

// ---------------------------------------------------------------------------------------------
// Element energies, written over a generic scalar S (double or Dual) exactly as the reference
// states them. p[] holds the prepared parameters in the layout of vbd_oracle.h.
// ---------------------------------------------------------------------------------------------

// springEnergy.cpp:20-30 — E = 1/2 k (|x0 - x1| - L0)^2
template <class S>
S energySpring(const V3<S> *x, const double *p)
{
    const V3<S> d = sub(x[0], x[1]);
    const S diff = sqrt(dot(d, d)) - p[0];
    return 0.5 * p[1] * diff * diff;
}

// springAnchorEnergy.cpp:21-26 — E = 1/2 k_model |x - a|^2
template <class S>
S energyAnchor(const V3<S> *x, const double *p)
{
    const V3<S> d = {x[0].x - p[0], x[0].y - p[1], x[0].z - p[2]};
    return 0.5 * p[3] * dot(d, d);
}

// membraneStvkEnergy.cpp:9-28 + energy.h:336-341 — C = (E00, E11, 2 E01), E = 1/2 C^T Keff C
template <class S>
S energyStvk(const V3<S> *x, const double *p)
{
    const double *Di = p;      // 2x2 row-major
    const double *K = p + 4;   // 3x3 row-major
    const V3<S> u = sub(x[1], x[0]);
    const V3<S> v = sub(x[2], x[0]);
    // F = [u v] * Dinv (3x2); column c: u * Di[0][c] + v * Di[1][c]
    const V3<S> f0 = {u.x * Di[0] + v.x * Di[2], u.y * Di[0] + v.y * Di[2], u.z * Di[0] + v.z * Di[2]};
    const V3<S> f1 = {u.x * Di[1] + v.x * Di[3], u.y * Di[1] + v.y * Di[3], u.z * Di[1] + v.z * Di[3]};
    const S e00 = 0.5 * (dot(f0, f0) - 1.0);
    const S e11 = 0.5 * (dot(f1, f1) - 1.0);
    const S e01 = 0.5 * dot(f0, f1);
    const S c[3] = {e00, e11, 2.0 * e01};
    S kc[3] = {S(0.0), S(0.0), S(0.0)};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) kc[i] = kc[i] + K[i * 3 + j] * c[j];
    return 0.5 * (c[0] * kc[0] + c[1] * kc[1] + c[2] * kc[2]);
}

// discreteBendingEnergy.cpp:9-30 + energy.h:336-341 — C = theta - theta0, E = 1/2 keff C^2
template <class S>
S bendingTheta(const V3<S> *x)
{
    const V3<S> e0 = sub(x[1], x[0]);
    const V3<S> e3 = sub(x[2], x[1]);
    const V3<S> e4 = sub(x[3], x[1]);
    const V3<S> n1 = normalized(cross(e0, e3));
    const V3<S> n2 = neg(normalized(cross(e0, e4)));
    return atan2(dot(cross(n1, n2), normalized(e0)), dot(n1, n2));
}
template <class S>
S energyBending(const V3<S> *x, const double *p)
{
    const S c = bendingTheta(x) - p[0];
    return 0.5 * c * (p[1] * c);
}

// stableNeoHookeanEnergy.cpp:33-61 — E = 1/2 lambda (det F - alpha)^2 + 1/2 mu (tr F^T F - 3)
template <class S>
S energySnh(const V3<S> *x, const double *p)
{
    const double *Di = p;  // 3x3 row-major
    const V3<S> u = sub(x[1], x[0]);
    const V3<S> v = sub(x[2], x[0]);
    const V3<S> w = sub(x[3], x[0]);
    V3<S> f[3];
    for (int c = 0; c < 3; ++c)
        f[c] = {u.x * Di[0 * 3 + c] + v.x * Di[1 * 3 + c] + w.x * Di[2 * 3 + c],
                u.y * Di[0 * 3 + c] + v.y * Di[1 * 3 + c] + w.y * Di[2 * 3 + c],
                u.z * Di[0 * 3 + c] + v.z * Di[1 * 3 + c] + w.z * Di[2 * 3 + c]};
    const S J = dot(f[0], cross(f[1], f[2]));
    const S hyd = J - p[11];
    const S dev = dot(f[0], f[0]) + dot(f[1], f[1]) + dot(f[2], f[2]) - 3.0;
    return 0.5 * p[9] * hyd * hyd + 0.5 * p[10] * dev;
}

// The eight further particle energies. Each is 1/2 C^T Keff C of its dConstraints (energy.h:336-341) unless
// noted; Keff is diagonal for all of them, so only its diagonal is kept.

// F = [x1 - x0, x2 - x0] Dinv (3x2) of a triangle; Dinv 2x2 row-major.
template <class S>
void triangleF(const V3<S> *x, const double *Di, V3<S> &f0, V3<S> &f1)
{
    const V3<S> u = sub(x[1], x[0]);
    const V3<S> v = sub(x[2], x[0]);
    f0 = {u.x * Di[0] + v.x * Di[2], u.y * Di[0] + v.y * Di[2], u.z * Di[0] + v.z * Di[2]};
    f1 = {u.x * Di[1] + v.x * Di[3], u.y * Di[1] + v.y * Di[3], u.z * Di[1] + v.z * Di[3]};
}
// F = [x1 - x0, x2 - x0, x3 - x0] Dinv (3x3) of a tetrahedron, as columns; Dinv 3x3 row-major.
template <class S>
void tetF(const V3<S> *x, const double *Di, V3<S> *f)
{
    const V3<S> u = sub(x[1], x[0]);
    const V3<S> v = sub(x[2], x[0]);
    const V3<S> w = sub(x[3], x[0]);
    for (int c = 0; c < 3; ++c)
        f[c] = {u.x * Di[0 * 3 + c] + v.x * Di[1 * 3 + c] + w.x * Di[2 * 3 + c],
                u.y * Di[0 * 3 + c] + v.y * Di[1 * 3 + c] + w.y * Di[2 * 3 + c],
                u.z * Di[0 * 3 + c] + v.z * Di[1 * 3 + c] + w.z * Di[2 * 3 + c]};
}
template <class S, int NC>
S diagQuadratic(const S *C, const double *k)
{
    S e = S(0.0);
    for (int i = 0; i < NC; ++i) e = e + C[i] * (k[i] * C[i]);
    return 0.5 * e;
}

// springContinuumEnergy.cpp:9-18 — C = |x0 - x1| / L0 - 1. p: (L0, k L0).
template <class S>
void constraintSpringContinuum(const V3<S> *x, const double *p, S *C)
{
    const V3<S> d = sub(x[0], x[1]);
    C[0] = sqrt(dot(d, d)) / p[0] - 1.0;
}
// springSquaredContinuumEnergy.cpp:9-19 — C = |x0 - x1|^2 / L0^2 - 1. p: (L0, k L0).
template <class S>
void constraintSpringSquaredContinuum(const V3<S> *x, const double *p, S *C)
{
    const V3<S> d = sub(x[0], x[1]);
    C[0] = dot(d, d) / (p[0] * p[0]) - 1.0;
}
// membraneBaraffWitkinEnergy.cpp:9-26 — C = (|F0| - 1, |F1| - 1, F0 . F1). p: Dinv (4), Keff diagonal (3).
template <class S>
void constraintMembraneBw(const V3<S> *x, const double *p, S *C)
{
    V3<S> f0, f1;
    triangleF(x, p, f0, f1);
    C[0] = sqrt(dot(f0, f0)) - 1.0;
    C[1] = sqrt(dot(f1, f1)) - 1.0;
    C[2] = dot(f0, f1);
}
// membraneChoiEnergy.cpp:9-28 — C = (|F0| - 1, |F1| - 1, |(F0 + F1)/sqrt 2| - 1, |(F0 - F1)/sqrt 2| - 1).
// p: Dinv (4), Keff diagonal (4).
template <class S>
void constraintMembraneChoi(const V3<S> *x, const double *p, S *C)
{
    V3<S> f0, f1;
    triangleF(x, p, f0, f1);
    const double r = 1.0 / std::sqrt(2.0);
    const V3<S> a = scale(add(f0, f1), r), b = scale(sub(f0, f1), r);
    C[0] = sqrt(dot(f0, f0)) - 1.0;
    C[1] = sqrt(dot(f1, f1)) - 1.0;
    C[2] = sqrt(dot(a, a)) - 1.0;
    C[3] = sqrt(dot(b, b)) - 1.0;
}
// quadraticBendingEnergy.cpp:9-30 — C = L x, L = [c0 I, c1 I, c2 I, c3 I] (3 x 12). p: c0..c3, keff.
template <class S>
void constraintQuadraticBending(const V3<S> *x, const double *p, S *C)
{
    C[0] = x[0].x * p[0] + x[1].x * p[1] + x[2].x * p[2] + x[3].x * p[3];
    C[1] = x[0].y * p[0] + x[1].y * p[1] + x[2].y * p[2] + x[3].y * p[3];
    C[2] = x[0].z * p[0] + x[1].z * p[1] + x[2].z * p[2] + x[3].z * p[3];
}
// baraffWitkinBendingEnergy.cpp:9-36 — C = theta - theta0 with the normalisations of e0, n1 and n2 taken
// as constants (their current values; no derivative flows through them). p: (theta0, k).
template <class S>
void constraintBendingBw(const V3<S> *x, const double *p, S *C)
{
    const V3<S> e0 = sub(x[1], x[0]);
    const V3<S> e3 = sub(x[2], x[1]);
    const V3<S> e4 = sub(x[3], x[1]);
    const double e0Inv = 1.0 / std::sqrt(valueOf(dot(e0, e0)));
    const V3<S> n1 = cross(e0, e3);
    const double n1Inv = 1.0 / std::sqrt(valueOf(dot(n1, n1)));
    const V3<S> n2 = neg(cross(e0, e4));
    const double n2Inv = 1.0 / std::sqrt(valueOf(dot(n2, n2)));
    const V3<S> a = scale(n1, n1Inv), b = scale(n2, n2Inv), e = scale(e0, e0Inv);
    C[0] = atan2(dot(cross(a, b), e), dot(a, b)) - p[0];
}
// Green strain E = 1/2 (F^T F - I) of a tetrahedron (stvkEnergy.cpp:27-29), symmetric, row-major.
template <class S>
void tetGreenStrain(const V3<S> *x, const double *Di, S *E)
{
    V3<S> f[3];
    tetF(x, Di, f);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) E[i * 3 + j] = 0.5 * (dot(f[i], f[j]) - (i == j ? 1.0 : 0.0));
}
// stvkEnergy.cpp:9-39 — C = (tr E, sqrt(2 tr(E E))), the second replaced by the constant 0 when it is 0.
template <class S>
void constraintStvkTet(const V3<S> *x, const double *p, S *C)
{
    S E[9];
    tetGreenStrain(x, p, E);
    S ee = S(0.0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) ee = ee + E[i * 3 + j] * E[j * 3 + i];
    C[0] = E[0] + E[4] + E[8];
    C[1] = sqrt(2.0 * ee);
    if (valueOf(C[1]) == 0) C[1] = S(0.0);
}
// stableSquaredNeoHookeanEnergy.cpp:9-31 — C = (det F - 1, tr F^T F - 3). p: Dinv (9), lambda, mu.
template <class S>
void constraintSsnh(const V3<S> *x, const double *p, S *C)
{
    V3<S> f[3];
    tetF(x, p, f);
    C[0] = dot(f[0], cross(f[1], f[2])) - 1.0;
    C[1] = dot(f[0], f[0]) + dot(f[1], f[1]) + dot(f[2], f[2]) - 3.0;
}

template <class S>
S energySpringContinuum(const V3<S> *x, const double *p)
{
    S C[1];
    constraintSpringContinuum(x, p, C);
    return diagQuadratic<S, 1>(C, p + 1);
}
template <class S>
S energySpringSquaredContinuum(const V3<S> *x, const double *p)
{
    S C[1];
    constraintSpringSquaredContinuum(x, p, C);
    return diagQuadratic<S, 1>(C, p + 1);
}
template <class S>
S energyMembraneBw(const V3<S> *x, const double *p)
{
    S C[3];
    constraintMembraneBw(x, p, C);
    return diagQuadratic<S, 3>(C, p + 4);
}
template <class S>
S energyMembraneChoi(const V3<S> *x, const double *p)
{
    S C[4];
    constraintMembraneChoi(x, p, C);
    return diagQuadratic<S, 4>(C, p + 4);
}
template <class S>
S energyQuadraticBending(const V3<S> *x, const double *p)
{
    S C[3];
    constraintQuadraticBending(x, p, C);
    const double k[3] = {p[4], p[4], p[4]};
    return diagQuadratic<S, 3>(C, k);
}
template <class S>
S energyBendingBw(const V3<S> *x, const double *p)
{
    S C[1];
    constraintBendingBw(x, p, C);
    return diagQuadratic<S, 1>(C, p + 1);
}
// stvkEnergy.cpp:41-63 — the energy form the reference differentiates (dEnergy override):
// E = 1/2 lambda (tr E)^2 + mu tr(E^T E).
template <class S>
S energyStvkTet(const V3<S> *x, const double *p)
{
    S E[9];
    tetGreenStrain(x, p, E);
    const S tr = E[0] + E[4] + E[8];
    S ee = S(0.0);
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) ee = ee + E[k * 3 + i] * E[k * 3 + i];
    return 0.5 * p[9] * tr * tr + p[10] * ee;
}
template <class S>
S energySsnh(const V3<S> *x, const double *p)
{
    S C[2];
    constraintSsnh(x, p, C);
    return diagQuadratic<S, 2>(C, p + 9);
}

int stencilSize(int type)
{
    switch (type) {
    case ORC_SPRING: return 2;
    case ORC_MEMBRANE_STVK: return 3;
    case ORC_DISCRETE_BENDING: return 4;
    case ORC_STABLE_NEO_HOOKEAN: return 4;
    case ORC_SPRING_ANCHOR: return 1;
    case ORC_SPRING_CONTINUUM: return 2;
    case ORC_SPRING_SQUARED_CONTINUUM: return 2;
    case ORC_MEMBRANE_BARAFF_WITKIN: return 3;
    case ORC_MEMBRANE_CHOI: return 3;
    case ORC_QUADRATIC_BENDING: return 4;
    case ORC_BARAFF_WITKIN_BENDING: return 4;
    case ORC_STVK: return 4;
    case ORC_STABLE_SQUARED_NEO_HOOKEAN: return 4;
    }
    return 0;
}
int rawParams(int type)
{
    switch (type) {
    case ORC_SPRING: return 2;
    case ORC_MEMBRANE_STVK: return 2;
    case ORC_DISCRETE_BENDING: return 2;
    case ORC_STABLE_NEO_HOOKEAN: return 2;
    case ORC_SPRING_ANCHOR: return 4;
    case ORC_SPRING_CONTINUUM: return 2;
    case ORC_SPRING_SQUARED_CONTINUUM: return 2;
    case ORC_MEMBRANE_BARAFF_WITKIN: return 3;
    case ORC_MEMBRANE_CHOI: return 3;
    case ORC_QUADRATIC_BENDING: return 2;
    case ORC_BARAFF_WITKIN_BENDING: return 2;
    case ORC_STVK: return 2;
    case ORC_STABLE_SQUARED_NEO_HOOKEAN: return 2;
    }
    return 0;
}
int preparedParams(int type)
{
    switch (type) {
    case ORC_SPRING: return 2;
    case ORC_MEMBRANE_STVK: return 13;
    case ORC_DISCRETE_BENDING: return 2;
    case ORC_STABLE_NEO_HOOKEAN: return 12;
    case ORC_SPRING_ANCHOR: return 4;
    case ORC_SPRING_CONTINUUM: return 2;
    case ORC_SPRING_SQUARED_CONTINUUM: return 2;
    case ORC_MEMBRANE_BARAFF_WITKIN: return 7;
    case ORC_MEMBRANE_CHOI: return 8;
    case ORC_QUADRATIC_BENDING: return 5;
    case ORC_BARAFF_WITKIN_BENDING: return 2;
    case ORC_STVK: return 11;
    case ORC_STABLE_SQUARED_NEO_HOOKEAN: return 11;
    }
    return 0;
}

template <class S>
S evalEnergy(int type, const V3<S> *x, const double *p)
{
    switch (type) {
    case ORC_SPRING: return energySpring(x, p);
    case ORC_MEMBRANE_STVK: return energyStvk(x, p);
    case ORC_DISCRETE_BENDING: return energyBending(x, p);
    case ORC_STABLE_NEO_HOOKEAN: return energySnh(x, p);
    case ORC_SPRING_ANCHOR: return energyAnchor(x, p);
    case ORC_SPRING_CONTINUUM: return energySpringContinuum(x, p);
    case ORC_SPRING_SQUARED_CONTINUUM: return energySpringSquaredContinuum(x, p);
    case ORC_MEMBRANE_BARAFF_WITKIN: return energyMembraneBw(x, p);
    case ORC_MEMBRANE_CHOI: return energyMembraneChoi(x, p);
    case ORC_QUADRATIC_BENDING: return energyQuadraticBending(x, p);
    case ORC_BARAFF_WITKIN_BENDING: return energyBendingBw(x, p);
    case ORC_STVK: return energyStvkTet(x, p);
    case ORC_STABLE_SQUARED_NEO_HOOKEAN: return energySsnh(x, p);
    }
    return S(0.0);
}

// Same as evalEnergy, restricted at compile time to the energies of stencil size NS.
template <int NS, class S>
S evalEnergyN(int type, const V3<S> *x, const double *p)
{
    if constexpr (NS == 1) {
        return energyAnchor(x, p);
    } else if constexpr (NS == 2) {
        if (type == ORC_SPRING_CONTINUUM) return energySpringContinuum(x, p);
        if (type == ORC_SPRING_SQUARED_CONTINUUM) return energySpringSquaredContinuum(x, p);
        return energySpring(x, p);
    } else if constexpr (NS == 3) {
        if (type == ORC_MEMBRANE_BARAFF_WITKIN) return energyMembraneBw(x, p);
        if (type == ORC_MEMBRANE_CHOI) return energyMembraneChoi(x, p);
        return energyStvk(x, p);
    } else {
        switch (type) {
        case ORC_DISCRETE_BENDING: return energyBending(x, p);
        case ORC_QUADRATIC_BENDING: return energyQuadraticBending(x, p);
        case ORC_BARAFF_WITKIN_BENDING: return energyBendingBw(x, p);
        case ORC_STVK: return energyStvkTet(x, p);
        case ORC_STABLE_SQUARED_NEO_HOOKEAN: return energySsnh(x, p);
        default: return energySnh(x, p);
        }
    }
}

// AD evaluation of the full element, as StencilBlockEnergy::dEnergyGradient / dEnergyHessian
// (energy.h:348-362). NS = stencil size; the gradient pass and the Hessian pass are separate, as
// accumulateVertexForce and accumulateVertexHessian are (energy.h:240-258).
template <int NS, bool H>
void adElement(int type, const double *xs, const double *p, double *E, double *grad, double *hess)
{
    constexpr int N = 3 * NS;
    using D = Dual<N, H>;
    V3<D> x[NS];
    for (int s = 0; s < NS; ++s) {
        x[s].x = D::var(xs[3 * s + 0], 3 * s + 0);
        x[s].y = D::var(xs[3 * s + 1], 3 * s + 1);
        x[s].z = D::var(xs[3 * s + 2], 3 * s + 2);
    }
    const D e = evalEnergyN<NS>(type, x, p);
    if (E) *E = e.v;
    if (grad)
        for (int i = 0; i < N; ++i) grad[i] = e.g[i];
    if (H && hess)
        for (int i = 0; i < N * N; ++i) hess[i] = e.h[i];
}

// ---------------------------------------------------------------------------------------------
// Closed-form per-vertex derivatives (gradient w.r.t. x_slot and its 3x3 diagonal Hessian block).
// ---------------------------------------------------------------------------------------------
using M3 = std::array<double, 9>;
using A3 = std::array<double, 3>;

inline A3 a3sub(const double *a, const double *b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline double a3dot(const A3 &a, const A3 &b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline A3 a3cross(const A3 &a, const A3 &b)
{
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
inline M3 skew(const A3 &a) { return {0, -a[2], a[1], a[2], 0, -a[0], -a[1], a[0], 0}; }
inline M3 mmul(const M3 &a, const M3 &b)
{
    M3 r{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) r[i * 3 + j] += a[i * 3 + k] * b[k * 3 + j];
    return r;
}
inline M3 mT(const M3 &a) { return {a[0], a[3], a[6], a[1], a[4], a[7], a[2], a[5], a[8]}; }
inline A3 mvec(const M3 &a, const A3 &v)
{
    return {a[0] * v[0] + a[1] * v[1] + a[2] * v[2], a[3] * v[0] + a[4] * v[1] + a[5] * v[2],
            a[6] * v[0] + a[7] * v[1] + a[8] * v[2]};
}
inline M3 outer(const A3 &a, const A3 &b)
{
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[i * 3 + j] = a[i] * b[j];
    return r;
}

// Spring: the reference's own closed form (springEnergy.cpp:101-134).
void cfSpring(const double *xs, const double *p, int slot, double *g, double *Hb)
{
    const double k = p[1], L0 = p[0];
    const A3 s = a3sub(xs, xs + 3);
    const double n = std::sqrt(a3dot(s, s));
    const double sg = slot == 0 ? 1.0 : -1.0;
    for (int i = 0; i < 3; ++i) g[i] = sg * k * (n - L0) * s[i] / n;
    const A3 sh = {s[0] / n, s[1] / n, s[2] / n};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double sst = sh[i] * sh[j];
            Hb[i * 3 + j] = k * ((1.0 - L0 / n) * ((i == j ? 1.0 : 0.0) - sst) + sst);
        }
}

void cfAnchor(const double *xs, const double *p, int, double *g, double *Hb)
{
    for (int i = 0; i < 3; ++i) g[i] = p[3] * (xs[i] - p[i]);
    for (int i = 0; i < 9; ++i) Hb[i] = (i % 4 == 0) ? p[3] : 0.0;
}

// StVK membrane: P = F S with S = [[s0, s2], [s2, s1]], s = Keff C; grad_s = P g_s;
// H_ss = (g_s^T S g_s) I + F (B^T K B) F^T with B = [[g0, 0], [0, g1], [g1, g0]].
void cfStvk(const double *xs, const double *p, int slot, double *g, double *Hb)
{
    const double *Di = p, *K = p + 4;
    const A3 u = a3sub(xs + 3, xs), v = a3sub(xs + 6, xs);
    A3 f0, f1;
    for (int i = 0; i < 3; ++i) {
        f0[i] = u[i] * Di[0] + v[i] * Di[2];
        f1[i] = u[i] * Di[1] + v[i] * Di[3];
    }
    const double C[3] = {0.5 * (a3dot(f0, f0) - 1.0), 0.5 * (a3dot(f1, f1) - 1.0), a3dot(f0, f1)};
    double sg[3];
    for (int i = 0; i < 3; ++i) sg[i] = K[i * 3] * C[0] + K[i * 3 + 1] * C[1] + K[i * 3 + 2] * C[2];
    double gs[2];
    if (slot == 0) {
        gs[0] = -(Di[0] + Di[2]);
        gs[1] = -(Di[1] + Di[3]);
    } else {
        gs[0] = Di[(slot - 1) * 2 + 0];
        gs[1] = Di[(slot - 1) * 2 + 1];
    }
    // P g_s = f0 (s0 g0 + s2 g1) + f1 (s2 g0 + s1 g1)
    const double a0 = sg[0] * gs[0] + sg[2] * gs[1], a1 = sg[2] * gs[0] + sg[1] * gs[1];
    for (int i = 0; i < 3; ++i) g[i] = f0[i] * a0 + f1[i] * a1;
    const double gSg = gs[0] * a0 + gs[1] * a1;
    // B (3x2)
    const double B[6] = {gs[0], 0.0, 0.0, gs[1], gs[1], gs[0]};
    double KB[6] = {0};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 3; ++k) KB[i * 2 + j] += K[i * 3 + k] * B[k * 2 + j];
    double M[4] = {0};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 3; ++k) M[i * 2 + j] += B[k * 2 + i] * KB[k * 2 + j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double FMFt = f0[i] * (M[0] * f0[j] + M[1] * f1[j]) + f1[i] * (M[2] * f0[j] + M[3] * f1[j]);
            Hb[i * 3 + j] = (i == j ? gSg : 0.0) + FMFt;
        }
}

// Stable Neo-Hookean: P = lambda (J - alpha) cof(F) + mu F; grad_s = P g_s;
// H_ss = mu |g_s|^2 I + lambda c c^T with c = cof(F) g_s (det is affine along a rank-one update,
// so its second derivative along x_s vanishes).
void cfSnh(const double *xs, const double *p, int slot, double *g, double *Hb)
{
    const double *Di = p;
    const double lambda = p[9], mu = p[10], alpha = p[11];
    const A3 u = a3sub(xs + 3, xs), v = a3sub(xs + 6, xs), w = a3sub(xs + 9, xs);
    A3 f[3];
    for (int c = 0; c < 3; ++c)
        for (int i = 0; i < 3; ++i) f[c][i] = u[i] * Di[0 * 3 + c] + v[i] * Di[1 * 3 + c] + w[i] * Di[2 * 3 + c];
    const A3 cof0 = a3cross(f[1], f[2]), cof1 = a3cross(f[2], f[0]), cof2 = a3cross(f[0], f[1]);
    const double J = a3dot(f[0], cof0);
    double gs[3];
    if (slot == 0) {
        for (int c = 0; c < 3; ++c) gs[c] = -(Di[0 * 3 + c] + Di[1 * 3 + c] + Di[2 * 3 + c]);
    } else {
        for (int c = 0; c < 3; ++c) gs[c] = Di[(slot - 1) * 3 + c];
    }
    A3 cg, fg;
    for (int i = 0; i < 3; ++i) {
        cg[i] = cof0[i] * gs[0] + cof1[i] * gs[1] + cof2[i] * gs[2];
        fg[i] = f[0][i] * gs[0] + f[1][i] * gs[1] + f[2][i] * gs[2];
    }
    const double hyd = lambda * (J - alpha);
    for (int i = 0; i < 3; ++i) g[i] = hyd * cg[i] + mu * fg[i];
    const double gg = gs[0] * gs[0] + gs[1] * gs[1] + gs[2] * gs[2];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Hb[i * 3 + j] = (i == j ? mu * gg : 0.0) + lambda * cg[i] * cg[j];
}

// Discrete bending. With a = e0 x e3, b = e0 x e4, the reference's angle is
// theta = atan2(s, c), s = -(a x b).e0/|e0|, c = -(a.b) (the common factor 1/(|a||b|) is positive).
// For the slot's position, a, b, e0 are affine: da = A dx, db = B dx, de0 = sigma dx, so s and c
// have closed-form gradients and Hessians; theta's follow from the atan2 chain rule.
void cfBending(const double *xs, const double *p, int slot, double *g, double *Hb)
{
    const double theta0 = p[0], k = p[1];
    const double *x0 = xs, *x1 = xs + 3, *x2 = xs + 6, *x3 = xs + 9;
    const A3 e = a3sub(x1, x0);
    const A3 a = a3cross(e, a3sub(x2, x1));
    const A3 b = a3cross(e, a3sub(x3, x1));
    M3 A{}, B{};
    double sigma = 0.0;
    switch (slot) {
    case 0:
        A = skew(a3sub(x2, x1));
        B = skew(a3sub(x3, x1));
        sigma = -1.0;
        break;
    case 1: {
        const A3 P = a3sub(x2, x0), Q = a3sub(x3, x0);
        A = skew({-P[0], -P[1], -P[2]});
        B = skew({-Q[0], -Q[1], -Q[2]});
        sigma = 1.0;
        break;
    }
    case 2: A = skew(e); break;
    case 3: B = skew(e); break;
    }
    const M3 At = mT(A), Bt = mT(B);
    // c = -a.b
    const double c = -a3dot(a, b);
    A3 gc;
    {
        const A3 t1 = mvec(At, b), t2 = mvec(Bt, a);
        for (int i = 0; i < 3; ++i) gc[i] = -(t1[i] + t2[i]);
    }
    M3 Hc;
    {
        const M3 m1 = mmul(At, B), m2 = mmul(Bt, A);
        for (int i = 0; i < 9; ++i) Hc[i] = -(m1[i] + m2[i]);
    }
    // T = (a x b).e
    const A3 axb = a3cross(a, b);
    const double T = a3dot(axb, e);
    A3 gT;
    {
        const A3 t1 = mvec(At, a3cross(b, e)), t2 = mvec(Bt, a3cross(e, a));
        for (int i = 0; i < 3; ++i) gT[i] = t1[i] + t2[i] + sigma * axb[i];
    }
    M3 HT;
    {
        const M3 m1 = mmul(mmul(At, skew(e)), B);
        const M3 m2 = mmul(At, skew(b));
        const M3 m3 = mmul(Bt, skew(a));
        M3 M;
        for (int i = 0; i < 9; ++i) M[i] = -m1[i] + sigma * m2[i] - sigma * m3[i];
        const M3 Mt = mT(M);
        for (int i = 0; i < 9; ++i) HT[i] = M[i] + Mt[i];
    }
    // s = -T / L
    const double L = std::sqrt(a3dot(e, e));
    const A3 eh = {e[0] / L, e[1] / L, e[2] / L};
    const double s = -T / L;
    A3 gs;
    for (int i = 0; i < 3; ++i) gs[i] = -gT[i] / L + T * sigma * eh[i] / (L * L);
    M3 Hs;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double id = i == j ? 1.0 : 0.0;
            Hs[i * 3 + j] = -HT[i * 3 + j] / L + sigma * (gT[i] * eh[j] + eh[i] * gT[j]) / (L * L) -
                            T * sigma * sigma * (3.0 * eh[i] * eh[j] - id) / (L * L * L);
        }
    // theta = atan2(s, c)
    const double r2 = s * s + c * c, r4 = r2 * r2;
    const double ts = c / r2, tc = -s / r2, tss = -2.0 * s * c / r4, tcc = 2.0 * s * c / r4,
                 tsc = (s * s - c * c) / r4;
    A3 gth;
    for (int i = 0; i < 3; ++i) gth[i] = ts * gs[i] + tc * gc[i];
    const double theta = bendingTheta<double>(reinterpret_cast<const V3<double> *>(xs));
    const double C = theta - theta0;
    for (int i = 0; i < 3; ++i) g[i] = k * C * gth[i];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double hth = ts * Hs[i * 3 + j] + tc * Hc[i * 3 + j] + tss * gs[i] * gs[j] +
                               tcc * gc[i] * gc[j] + tsc * (gs[i] * gc[j] + gc[i] * gs[j]);
            Hb[i * 3 + j] = k * (gth[i] * gth[j] + C * hth);
        }
}

void adVertex(int type, const double *xs, const double *p, int slot, double *g, double *Hb);

void cfVertex(int type, const double *xs, const double *p, int slot, double *g, double *Hb)
{
    switch (type) {
    case ORC_SPRING: cfSpring(xs, p, slot, g, Hb); return;
    case ORC_MEMBRANE_STVK: cfStvk(xs, p, slot, g, Hb); return;
    case ORC_DISCRETE_BENDING: cfBending(xs, p, slot, g, Hb); return;
    case ORC_STABLE_NEO_HOOKEAN: cfSnh(xs, p, slot, g, Hb); return;
    case ORC_SPRING_ANCHOR: cfAnchor(xs, p, slot, g, Hb); return;
    }
    adVertex(type, xs, p, slot, g, Hb);  // the further energies: dual numbers, as the reference (TinyAD)
}

// AD per-vertex: full element gradient pass + full element Hessian pass, then slice
// (energy.h:240-258). The spring keeps its closed-form override (springEnergy.cpp:101-134).
template <int NS>
void adVertexN(int type, const double *xs, const double *p, int slot, double *g, double *Hb)
{
    constexpr int N = 3 * NS;
    double grad[N];
    static thread_local std::vector<double> hess(144);
    adElement<NS, false>(type, xs, p, nullptr, grad, nullptr);
    adElement<NS, true>(type, xs, p, nullptr, nullptr, hess.data());
    for (int i = 0; i < 3; ++i) g[i] = grad[3 * slot + i];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Hb[i * 3 + j] = hess[(3 * slot + i) * N + 3 * slot + j];
}
void adVertex(int type, const double *xs, const double *p, int slot, double *g, double *Hb)
{
    switch (type) {
    case ORC_SPRING: cfSpring(xs, p, slot, g, Hb); return;
    case ORC_MEMBRANE_STVK: adVertexN<3>(type, xs, p, slot, g, Hb); return;
    case ORC_DISCRETE_BENDING: adVertexN<4>(type, xs, p, slot, g, Hb); return;
    case ORC_STABLE_NEO_HOOKEAN: adVertexN<4>(type, xs, p, slot, g, Hb); return;
    case ORC_SPRING_ANCHOR: adVertexN<1>(type, xs, p, slot, g, Hb); return;
    }
    switch (stencilSize(type)) {
    case 2: adVertexN<2>(type, xs, p, slot, g, Hb); return;
    case 3: adVertexN<3>(type, xs, p, slot, g, Hb); return;
    case 4: adVertexN<4>(type, xs, p, slot, g, Hb); return;
    }
    throw std::runtime_error("unknown energy type");
}

// Constraints per stencil: the TnConstraints of each energy's StencilBlockEnergy base
// (membraneStvkEnergy.h <3, 3>, stableNeoHookeanEnergy.h <4, 2>, the others 1).
int numConstraints(int type)
{
    switch (type) {
    case ORC_MEMBRANE_STVK: return 3;
    case ORC_STABLE_NEO_HOOKEAN: return 2;
    case ORC_SPRING:
    case ORC_DISCRETE_BENDING:
    case ORC_SPRING_ANCHOR: return 1;
    case ORC_SPRING_CONTINUUM:
    case ORC_SPRING_SQUARED_CONTINUUM:
    case ORC_BARAFF_WITKIN_BENDING: return 1;
    case ORC_MEMBRANE_BARAFF_WITKIN: return 3;
    case ORC_MEMBRANE_CHOI: return 4;
    case ORC_QUADRATIC_BENDING: return 3;
    case ORC_STVK:
    case ORC_STABLE_SQUARED_NEO_HOOKEAN: return 2;
    }
    return 0;
}

// Inverse of a 3x3 (row-major) by cofactors: inverse = adj / det.
M3 inverse3(const M3 &m)
{
    const double c00 = m[4] * m[8] - m[5] * m[7];
    const double c01 = m[5] * m[6] - m[3] * m[8];
    const double c02 = m[3] * m[7] - m[4] * m[6];
    const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    const M3 adj = {c00,
                    m[2] * m[7] - m[1] * m[8],
                    m[1] * m[5] - m[2] * m[4],
                    c01,
                    m[0] * m[8] - m[2] * m[6],
                    m[2] * m[3] - m[0] * m[5],
                    c02,
                    m[1] * m[6] - m[0] * m[7],
                    m[0] * m[4] - m[1] * m[3]};
    M3 r;
    for (int k = 0; k < 9; ++k) r[k] = adj[k] / det;
    return r;
}

// ---------------------------------------------------------------------------------------------
// Object: ParticleGroup + energies + VBD solver state
// ---------------------------------------------------------------------------------------------
struct Energy {
    int type = 0;
    int ssize = 0;
    std::vector<int32_t> st;   // flat stencils (Energy::flatStencils, energy.h:102-110)
    std::vector<double> raw;   // addStencil arguments
    std::vector<double> prm;   // prepared parameters
    std::vector<double> cmp;   // effectiveCompliance() per stencil, nC x nC row-major (XPBD)
    bool prepared = false;
    int64_t n() const { return ssize ? static_cast<int64_t>(st.size()) / ssize : 0; }
};

struct Object {
    int64_t nv = 0;
    std::vector<double> x, xInitial, x0, v, m;
    std::vector<Energy> energies;
    // VBD precomputation (vbd.cpp:203-268)
    bool precompDone = false;
    std::vector<int32_t> colors;
    std::vector<std::vector<int32_t>> groups;
    // incidence: per vertex, entries (energy, stencil, slot) in energy-major, ascending stencil order
    std::vector<int64_t> incPtr;
    std::vector<std::array<int32_t, 3>> inc;
    // XPBD stencil groups per energy (computeParallelStencilGroups, xpbd.cpp:18-21: first step only)
    bool xpbdGroupsDone = false;
    std::vector<std::vector<std::vector<int32_t>>> xpbdGroups;
};

struct OracleError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Restates the validation in each energy's addStencil (springEnergy.cpp:38-40,
// discreteBendingEnergy.cpp:38-40) and stores the addStencil arguments; the stiffness-matrix
// construction of addStencil (membraneStvkEnergy.cpp:35-42, stableNeoHookeanEnergy.cpp:71-79)
// happens together with preparePrecomputations below.
void addEnergy(Object &o, int type, int64_t n, const int32_t *st, const double *raw)
{
    Energy e;
    e.type = type;
    e.ssize = stencilSize(type);
    if (!e.ssize) throw OracleError("unknown energy type");
    const int nr = rawParams(type);
    e.st.assign(st, st + n * e.ssize);
    e.raw.assign(raw, raw + n * nr);
    for (int64_t i = 0; i < n; ++i) {
        for (int s = 0; s < e.ssize; ++s)
            if (e.st[i * e.ssize + s] < 0 || e.st[i * e.ssize + s] >= o.nv) throw OracleError("stencil index out of range");
        const bool springLike =
            type == ORC_SPRING || type == ORC_SPRING_CONTINUUM || type == ORC_SPRING_SQUARED_CONTINUUM;
        if (springLike && e.raw[i * nr] <= 0) throw OracleError("Zero/negative rest length invalid");
        const bool angled = type == ORC_DISCRETE_BENDING || type == ORC_BARAFF_WITKIN_BENDING;
        if (angled && (e.raw[i * nr] < -3.12 || e.raw[i * nr] > 3.12)) throw OracleError("Invalid restTheta");
        // quadraticBendingEnergy.cpp:37-39
        if (type == ORC_QUADRATIC_BENDING && e.raw[i * nr] != 0)
            throw OracleError("Rest theta != 0 not supported for quadratic bending");
    }
    o.energies.push_back(std::move(e));
}

// preparePrecomputations of each energy (membraneStvkEnergy.cpp:45-71, discreteBendingEnergy.cpp:48-72,
// stableNeoHookeanEnergy.cpp:82-108; spring/anchor only take the addStencil values, springEnergy.cpp:33-46,
// springAnchorEnergy.cpp:29-39). Uses the rest positions x0.
void prepareEnergy(const Object &o, Energy &e)
{
    if (e.prepared) throw OracleError("Energy was already prepared");  // energy.h:54-61
    const int64_t n = e.n();
    const int np = preparedParams(e.type);
    const int nc = numConstraints(e.type);
    e.prm.assign(n * np, 0.0);
    e.cmp.assign(n * nc * nc, 0.0);
    for (int64_t i = 0; i < n; ++i) {
        const int32_t *s = &e.st[i * e.ssize];
        const double *r = &e.raw[i * rawParams(e.type)];
        double *p = &e.prm[i * np];
        double *cm = &e.cmp[i * nc * nc];
        auto X0 = [&](int k) { return &o.x0[3 * static_cast<int64_t>(s[k])]; };
        switch (e.type) {
        case ORC_SPRING:  // springEnergy.cpp:42-45: compliance = StiffnessMat{k}.inverse()
            p[0] = r[0];
            p[1] = r[1];
            cm[0] = 1.0 / r[1];
            break;
        case ORC_SPRING_ANCHOR:  // springAnchorEnergy.cpp:34-37
            p[0] = r[0];
            p[1] = r[1];
            p[2] = r[2];
            p[3] = r[3];
            cm[0] = 1.0 / r[3];
            break;
        case ORC_MEMBRANE_STVK: {
            const double young = r[0], poisson = r[1];
            const double yc = young / (1 - poisson * poisson);
            const double K[9] = {yc, yc * poisson, 0, yc * poisson, yc, 0, 0, 0, yc * (1 - poisson) / 2};
            const double *a = X0(0), *b = X0(1), *c = X0(2);
            if (a[2] != 0.0 || b[2] != 0.0 || c[2] != 0.0)
                throw OracleError("Rest positions of triangles for membrane StVK energy must have coordinate z=0");
            const double u0 = b[0] - a[0], u1 = b[1] - a[1], v0 = c[0] - a[0], v1 = c[1] - a[1];
            // Dm = [u v] (columns), inverse of a 2x2
            const double det = u0 * v1 - v0 * u1;
            p[0] = v1 / det;
            p[1] = -v0 / det;
            p[2] = -u1 / det;
            p[3] = u0 / det;
            const double area = (u0 * v1 - v0 * u1) * 0.5;
            if (area <= 0) throw OracleError("Rest area of membrane StVK triangle is not positive");
            for (int k = 0; k < 9; ++k) p[4 + k] = K[k] * area;
            // membraneStvkEnergy.cpp:41-42,67: effectiveCompliance = C.inverse() / area
            const M3 Ci = inverse3({K[0], K[1], K[2], K[3], K[4], K[5], K[6], K[7], K[8]});
            for (int k = 0; k < 9; ++k) cm[k] = Ci[k] / area;
            break;
        }
        case ORC_DISCRETE_BENDING: {
            const double *a = X0(0), *b = X0(1), *c = X0(2), *d = X0(3);
            const A3 e0 = a3sub(b, a), e3 = a3sub(c, b), e4 = a3sub(d, b);
            const A3 c03 = a3cross(e0, e3), c04 = a3cross(e0, e4);
            const double factor =
                2. * 3. * a3dot(e0, e0) / (0.5 * (std::sqrt(a3dot(c03, c03)) + std::sqrt(a3dot(c04, c04))));
            const double compliance = (1.0 / r[1]) / factor;  // discreteBendingEnergy.cpp:44-45,63
            p[1] = 1.0 / compliance;
            cm[0] = compliance;
            const double xsr[12] = {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2], d[0], d[1], d[2]};
            p[0] = bendingTheta<double>(reinterpret_cast<const V3<double> *>(xsr));
            break;
        }
        case ORC_STABLE_NEO_HOOKEAN: {
            const double young = r[0], poisson = r[1];
            const double mu = young / (2 * (1 + poisson));
            const double lambda = mu / (1 - 2 * poisson);
            const double *a = X0(0), *b = X0(1), *c = X0(2), *d = X0(3);
            const A3 u = a3sub(b, a), v = a3sub(c, a), w = a3sub(d, a);
            // Dm columns u, v, w; Dm(r, c)
            const double Dm[9] = {u[0], v[0], w[0], u[1], v[1], w[1], u[2], v[2], w[2]};
            const double c00 = Dm[4] * Dm[8] - Dm[5] * Dm[7];
            const double c01 = Dm[5] * Dm[6] - Dm[3] * Dm[8];
            const double c02 = Dm[3] * Dm[7] - Dm[4] * Dm[6];
            const double det = Dm[0] * c00 + Dm[1] * c01 + Dm[2] * c02;
            const double volume = 1. / 6. * det;
            if (volume <= 0) throw OracleError("SNH tetrahedron rest volume is not positive");
            // inverse = adj / det, adj(i,j) = cofactor(j,i)
            const double cof[9] = {c00,
                                   c01,
                                   c02,
                                   Dm[2] * Dm[7] - Dm[1] * Dm[8],
                                   Dm[0] * Dm[8] - Dm[2] * Dm[6],
                                   Dm[1] * Dm[6] - Dm[0] * Dm[7],
                                   Dm[1] * Dm[5] - Dm[2] * Dm[4],
                                   Dm[2] * Dm[3] - Dm[0] * Dm[5],
                                   Dm[0] * Dm[4] - Dm[1] * Dm[3]};
            for (int rr = 0; rr < 3; ++rr)
                for (int cc = 0; cc < 3; ++cc) p[rr * 3 + cc] = cof[cc * 3 + rr] / det;
            p[9] = lambda * volume;
            p[10] = mu * volume;
            p[11] = 1 + mu / lambda;
            // stableNeoHookeanEnergy.cpp:74-78,104: effectiveCompliance = diag(lambda, mu).inverse() / volume
            const double det2 = lambda * mu;
            cm[0] = mu / det2 / volume;
            cm[1] = 0.0;
            cm[2] = 0.0;
            cm[3] = lambda / det2 / volume;
            break;
        }
        case ORC_SPRING_CONTINUUM:
        case ORC_SPRING_SQUARED_CONTINUUM:  // springContinuumEnergy.cpp:20-35: the stiffness integrated along L0
            p[0] = r[0];
            p[1] = r[1] * r[0];
            cm[0] = 1.0 / p[1];
            break;
        case ORC_MEMBRANE_BARAFF_WITKIN:
        case ORC_MEMBRANE_CHOI: {  // membraneBaraffWitkinEnergy.cpp:28-68, membraneChoiEnergy.cpp:30-70
            const bool choi = e.type == ORC_MEMBRANE_CHOI;
            const double *a = X0(0), *b = X0(1), *c = X0(2);
            if (a[2] != 0.0 || b[2] != 0.0 || c[2] != 0.0)
                throw OracleError(choi ? "Rest positions of triangles for membrane Choi energy must have coordinate z=0"
                                       : "Rest positions of triangles for membrane BaraffWitkin energy must have "
                                         "coordinate z=0");
            const double u0 = b[0] - a[0], u1 = b[1] - a[1], v0 = c[0] - a[0], v1 = c[1] - a[1];
            const double det = u0 * v1 - v0 * u1;
            p[0] = v1 / det;
            p[1] = -v0 / det;
            p[2] = -u1 / det;
            p[3] = u0 / det;
            const double area = (u0 * v1 - v0 * u1) * 0.5;
            if (area <= 0)
                throw OracleError(choi ? "Rest area of membrane Choi triangle is not positive"
                                       : "Rest area of membrane BaraffWitkin triangle is not positive");
            const double k[4] = {r[0], r[1], r[2], r[2]};
            const int nk = choi ? 4 : 3;
            for (int q = 0; q < nk; ++q) p[4 + q] = k[q] * area;
            // compliance: the inverse of the diagonal model stiffness (by cofactors, as the small fixed-size
            // inverses do), divided by the area
            double prod = 1.0;
            for (int q = 0; q < nk; ++q) prod *= k[q];
            for (int q = 0; q < nk; ++q) cm[q * nk + q] = (prod / k[q]) / prod / area;
            break;
        }
        case ORC_QUADRATIC_BENDING: {  // quadraticBendingEnergy.cpp:49-89
            const double *a = X0(0), *b = X0(1), *c = X0(2), *d = X0(3);
            const A3 e0 = a3sub(b, a), e1 = a3sub(c, a), e2 = a3sub(d, a), e3 = a3sub(c, b), e4 = a3sub(d, b);
            auto cot = [](const A3 &u, const A3 &v) {
                const A3 w = a3cross(u, v);
                return a3dot(u, v) / std::sqrt(a3dot(w, w));
            };
            const A3 me0 = {-e0[0], -e0[1], -e0[2]};
            const double c01 = cot(e0, e1), c02 = cot(e0, e2), c03 = cot(me0, e3), c04 = cot(me0, e4);
            p[0] = c03 + c04;
            p[1] = c01 + c02;
            p[2] = -c01 - c03;
            p[3] = -c02 - c04;
            const A3 n1 = a3cross(e0, e1), n2 = a3cross(e0, e2);
            const double keff = r[1] * (3 / (0.5 * (std::sqrt(a3dot(n1, n1)) + std::sqrt(a3dot(n2, n2)))));
            p[4] = keff;
            const M3 ci = inverse3({keff, 0, 0, 0, keff, 0, 0, 0, keff});
            for (int q = 0; q < 9; ++q) cm[q] = ci[q];
            break;
        }
        case ORC_BARAFF_WITKIN_BENDING: {  // baraffWitkinBendingEnergy.cpp:38-75: theta0 from the rest positions
            const double *a = X0(0), *b = X0(1), *c = X0(2), *d = X0(3);
            const double xsr[12] = {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2], d[0], d[1], d[2]};
            p[0] = bendingTheta<double>(reinterpret_cast<const V3<double> *>(xsr));
            p[1] = r[1];
            cm[0] = 1.0 / r[1];
            break;
        }
        case ORC_STVK:
        case ORC_STABLE_SQUARED_NEO_HOOKEAN: {  // stvkEnergy.cpp:66-104, stableSquaredNeoHookeanEnergy.cpp:34-70
            const double young = r[0], poisson = r[1];
            const double mu = young / (2 * (1 + poisson));
            const double lambda = young * poisson / ((1 + poisson) * (1 - 2 * poisson)) + 1e-20;
            const double *a = X0(0), *b = X0(1), *c = X0(2), *d = X0(3);
            const A3 u = a3sub(b, a), v = a3sub(c, a), w = a3sub(d, a);
            const M3 Dm = {u[0], v[0], w[0], u[1], v[1], w[1], u[2], v[2], w[2]};
            const double det = Dm[0] * (Dm[4] * Dm[8] - Dm[5] * Dm[7]) + Dm[1] * (Dm[5] * Dm[6] - Dm[3] * Dm[8]) +
                               Dm[2] * (Dm[3] * Dm[7] - Dm[4] * Dm[6]);
            const double volume = 1. / 6. * det;
            if (volume <= 0)
                throw OracleError(e.type == ORC_STVK ? "SNH tetrahedron rest volume is not positive"
                                                     : "SSNH tetrahedron rest volume is not positive");
            const M3 Di = inverse3(Dm);
            for (int q = 0; q < 9; ++q) p[q] = Di[q];
            p[9] = lambda * volume;
            p[10] = mu * volume;
            const double det2 = lambda * mu;
            cm[0] = mu / det2 / volume;
            cm[1] = 0.0;
            cm[2] = 0.0;
            cm[3] = lambda / det2 / volume;
            break;
        }
        }
    }
    e.prepared = true;
}

// coloring::colorVertices + Graph::greedyColoring (graphColoring.cpp:21-49,72-84): a clique over
// every stencil of every energy; node 0 gets colour 0; each later node takes the smallest colour
// not used by an already coloured (lower-index) neighbour. The per-node availability reset is
// done over the touched colours only (the reference fills a V-sized array), same result.
std::vector<int32_t> colorVertices(const Object &o)
{
    const int64_t nv = o.nv;
    std::vector<std::vector<int32_t>> adj(nv);
    for (const auto &e : o.energies) {
        const int ss = e.ssize;
        for (int64_t i = 0; i < static_cast<int64_t>(e.st.size()); ++i)
            for (int64_t j = i + 1; j % ss != 0; ++j) {
                adj[e.st[i]].push_back(e.st[j]);
                adj[e.st[j]].push_back(e.st[i]);
            }
    }
    std::vector<int32_t> col(nv, -1);
    if (nv == 0) return col;
    col[0] = 0;
    std::vector<char> used(nv + 1, 0);
    for (int64_t n = 1; n < nv; ++n) {
        for (int32_t q : adj[n])
            if (col[q] != -1) used[col[q]] = 1;
        int32_t c = 0;
        while (used[c]) ++c;
        col[n] = c;
        for (int32_t q : adj[n])
            if (col[q] != -1) used[col[q]] = 0;
    }
    return col;
}

void precompute(Object &o)
{
    // computeParallelVertexGroups (vbd.cpp:203-229): groups in ascending vertex index
    o.colors = colorVertices(o);
    o.groups.clear();
    for (int64_t i = 0; i < o.nv; ++i) {
        const int32_t c = o.colors[i];
        if (static_cast<int32_t>(o.groups.size()) < c + 1) o.groups.resize(c + 1);
        o.groups[c].push_back(static_cast<int32_t>(i));
    }
    // computeStencilInfoPerVertex (vbd.cpp:231-268): per vertex, per energy, (s/size, s%size) ascending s
    std::vector<std::vector<std::array<int32_t, 3>>> per(o.nv);
    for (int32_t ei = 0; ei < static_cast<int32_t>(o.energies.size()); ++ei) {
        const auto &e = o.energies[ei];
        for (int64_t s = 0; s < static_cast<int64_t>(e.st.size()); ++s)
            per[e.st[s]].push_back({ei, static_cast<int32_t>(s / e.ssize), static_cast<int32_t>(s % e.ssize)});
    }
    o.incPtr.assign(o.nv + 1, 0);
    o.inc.clear();
    for (int64_t v = 0; v < o.nv; ++v) {
        o.incPtr[v] = static_cast<int64_t>(o.inc.size());
        o.inc.insert(o.inc.end(), per[v].begin(), per[v].end());
    }
    o.incPtr[o.nv] = static_cast<int64_t>(o.inc.size());
    o.precompDone = true;
}

void gatherStencil(const Object &o, const double *x, const Energy &e, int64_t si, double *xs)
{
    for (int k = 0; k < e.ssize; ++k) {
        const int64_t vi = e.st[si * e.ssize + k];
        xs[3 * k + 0] = x[3 * vi + 0];
        xs[3 * k + 1] = x[3 * vi + 1];
        xs[3 * k + 2] = x[3 * vi + 2];
    }
}

// l_vertexDescent (vbd.cpp:99-147) for one particle.
void vertexDescent(Object &o, int64_t vi, const std::vector<double> &xt, double invdt2, int mode)
{
    const double mdt2 = o.m[vi] * invdt2;
    double f[3], H[9] = {0};
    for (int i = 0; i < 3; ++i) f[i] = -mdt2 * (o.x[3 * vi + i] - xt[3 * vi + i]);
    H[0] = H[4] = H[8] = mdt2;
    double xs[12], g[3], Hb[9];
    for (int64_t k = o.incPtr[vi]; k < o.incPtr[vi + 1]; ++k) {
        const auto &en = o.inc[k];
        const Energy &e = o.energies[en[0]];
        gatherStencil(o, o.x.data(), e, en[1], xs);
        const double *p = &e.prm[static_cast<int64_t>(en[1]) * preparedParams(e.type)];
        if (mode == ORC_MODE_AD)
            adVertex(e.type, xs, p, en[2], g, Hb);
        else
            cfVertex(e.type, xs, p, en[2], g, Hb);
        for (int i = 0; i < 3; ++i) f[i] -= g[i];
        for (int i = 0; i < 9; ++i) H[i] += Hb[i];
    }
    // H.determinant() > 1e-10 (vbd.cpp:22,139); cofactor expansion and adjugate inverse
    const double c00 = H[4] * H[8] - H[5] * H[7];
    const double c10 = H[5] * H[6] - H[3] * H[8];
    const double c20 = H[3] * H[7] - H[4] * H[6];
    const double det = H[0] * c00 + H[1] * c10 + H[2] * c20;
    if (det > 1e-10) {
        const double inv[9] = {c00,
                               H[2] * H[7] - H[1] * H[8],
                               H[1] * H[5] - H[2] * H[4],
                               c10,
                               H[0] * H[8] - H[2] * H[6],
                               H[2] * H[3] - H[0] * H[5],
                               c20,
                               H[1] * H[6] - H[0] * H[7],
                               H[0] * H[4] - H[1] * H[3]};
        for (int i = 0; i < 3; ++i)
            o.x[3 * vi + i] += (inv[i * 3] * f[0] + inv[i * 3 + 1] * f[1] + inv[i * 3 + 2] * f[2]) / det;
    }
}

void sweep(Object &o, const std::vector<double> &xt, double invdt2, int mode)
{
    // colour groups in order, OpenMP over the vertices of a group (vbd.cpp:148-158)
    for (const auto &grp : o.groups) {
        const int64_t n = static_cast<int64_t>(grp.size());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) vertexDescent(o, grp[i], xt, invdt2, mode);
    }
}

void residual(Object &o, const double *x, const double *xt, double dt, int mode, double *norm, double *rout)
{
    // r = -grad E(x) - (M/h^2)(x - xt) (implicitEulerBase.cpp:102-143 with x-form of
    // implicitEulerNewtonRobust.cpp:65-66), gathered per vertex over the incidence lists.
    if (!o.precompDone) precompute(o);
    const double invdt2 = 1.0 / (dt * dt);
    double sum = 0.0;
#pragma omp parallel for reduction(+ : sum) schedule(static)
    for (int64_t vi = 0; vi < o.nv; ++vi) {
        double r[3];
        for (int i = 0; i < 3; ++i) r[i] = -o.m[vi] * invdt2 * (x[3 * vi + i] - xt[3 * vi + i]);
        double xs[12], g[3], Hb[9];
        for (int64_t k = o.incPtr[vi]; k < o.incPtr[vi + 1]; ++k) {
            const auto &en = o.inc[k];
            const Energy &e = o.energies[en[0]];
            gatherStencil(o, x, e, en[1], xs);
            const double *p = &e.prm[static_cast<int64_t>(en[1]) * preparedParams(e.type)];
            if (mode == ORC_MODE_AD)
                adVertex(e.type, xs, p, en[2], g, Hb);
            else
                cfVertex(e.type, xs, p, en[2], g, Hb);
            for (int i = 0; i < 3; ++i) r[i] -= g[i];
        }
        if (rout)
            for (int i = 0; i < 3; ++i) rout[3 * vi + i] = r[i];
        sum += r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    }
    *norm = std::sqrt(sum);
}

void checkPrepared(const Object &o)
{
    for (const auto &e : o.energies)
        if (!e.prepared) throw OracleError("energy not prepared");
}

// One substep of VBD::step (vbd.cpp:46-185).
void substep(Object &o, double h, int iterations, const double *g, int mode, const double *startGuess,
             double *rhoCurve)
{
    const int64_t nv = o.nv;
    std::vector<double> xOld(o.x), xt(3 * nv);
    for (int64_t i = 0; i < nv; ++i)
        for (int d = 0; d < 3; ++d) {
            o.v[3 * i + d] += h * g[d];             // vbd.cpp:68 (every particle, also m = 0)
            o.x[3 * i + d] += o.v[3 * i + d] * h;   // particleGroup.cpp:57-63
            xt[3 * i + d] = o.x[3 * i + d];
        }
    double rho0 = 0.0;
    if (rhoCurve) {
        residual(o, xt.data(), xt.data(), h, mode, &rho0, nullptr);
        if (rho0 == 0.0) rho0 = 1.0;
    }
    if (startGuess) std::copy(startGuess, startGuess + 3 * nv, o.x.begin());
    const double invdt2 = 1.0 / (h * h);
    for (int it = 0; it < iterations; ++it) {
        sweep(o, xt, invdt2, mode);
        if (rhoCurve) {
            double r;
            residual(o, o.x.data(), xt.data(), h, mode, &r, nullptr);
            rhoCurve[it] = r / rho0;
        }
    }
    const double invdt = 1.0 / h;
    for (int64_t i = 0; i < 3 * nv; ++i) o.v[i] = (o.x[i] - xOld[i]) * invdt;  // particleGroup.cpp:65-72
}

// ---------------------------------------------------------------------------------------------
// XPBD (spg/source/sim/solver/xpbd.cpp:16-160): one constraint projection per stencil per substep,
// energies in addEnergy order, stencils in parallel colour groups (computeParallelStencilGroups,
// xpbd.cpp:134-160) or in index order (the serial Gauss-Seidel path, xpbd.cpp:99-104).
// ---------------------------------------------------------------------------------------------
// The dConstraints lambdas of the five energies over a generic scalar (double or Dual).
// springEnergy.cpp:10-18 — C = |x0 - x1| - L0
template <class S>
void constraintSpring(const V3<S> *x, const double *p, S *C)
{
    const V3<S> d = sub(x[0], x[1]);
    C[0] = sqrt(dot(d, d)) - p[0];
}
// springAnchorEnergy.cpp:8-20 — C = |x - a|, replaced by the constant 0 when it is exactly 0
template <class S>
void constraintAnchor(const V3<S> *x, const double *p, S *C)
{
    const V3<S> d = {x[0].x - p[0], x[0].y - p[1], x[0].z - p[2]};
    C[0] = sqrt(dot(d, d));
    if (valueOf(C[0]) == 0) C[0] = S(0.0);
}
// membraneStvkEnergy.cpp:9-28 — C = (E00, E11, 2 E01) of the Green strain E = 1/2 (F^T F - I)
template <class S>
void constraintStvk(const V3<S> *x, const double *p, S *C)
{
    const double *Di = p;
    const V3<S> u = sub(x[1], x[0]);
    const V3<S> v = sub(x[2], x[0]);
    const V3<S> f0 = {u.x * Di[0] + v.x * Di[2], u.y * Di[0] + v.y * Di[2], u.z * Di[0] + v.z * Di[2]};
    const V3<S> f1 = {u.x * Di[1] + v.x * Di[3], u.y * Di[1] + v.y * Di[3], u.z * Di[1] + v.z * Di[3]};
    C[0] = 0.5 * (dot(f0, f0) - 1.0);
    C[1] = 0.5 * (dot(f1, f1) - 1.0);
    C[2] = 2.0 * (0.5 * dot(f0, f1));
}
// discreteBendingEnergy.cpp:9-31 — C = theta - theta0
template <class S>
void constraintBending(const V3<S> *x, const double *p, S *C)
{
    C[0] = bendingTheta(x) - p[0];
}
// stableNeoHookeanEnergy.cpp:9-31 — C = (det F - alpha, sqrt(tr F^T F))
template <class S>
void constraintSnh(const V3<S> *x, const double *p, S *C)
{
    const double *Di = p;
    const V3<S> u = sub(x[1], x[0]);
    const V3<S> v = sub(x[2], x[0]);
    const V3<S> w = sub(x[3], x[0]);
    V3<S> f[3];
    for (int c = 0; c < 3; ++c)
        f[c] = {u.x * Di[0 * 3 + c] + v.x * Di[1 * 3 + c] + w.x * Di[2 * 3 + c],
                u.y * Di[0 * 3 + c] + v.y * Di[1 * 3 + c] + w.y * Di[2 * 3 + c],
                u.z * Di[0 * 3 + c] + v.z * Di[1 * 3 + c] + w.z * Di[2 * 3 + c]};
    C[0] = dot(f[0], cross(f[1], f[2])) - p[11];
    C[1] = sqrt(dot(f[0], f[0]) + dot(f[1], f[1]) + dot(f[2], f[2]));
}

template <int NS, int NC, class S>
void evalConstraintsN(int type, const V3<S> *x, const double *p, S *C)
{
    if constexpr (NS == 1) {
        constraintAnchor(x, p, C);
    } else if constexpr (NS == 2) {
        if (type == ORC_SPRING_CONTINUUM)
            constraintSpringContinuum(x, p, C);
        else if (type == ORC_SPRING_SQUARED_CONTINUUM)
            constraintSpringSquaredContinuum(x, p, C);
        else
            constraintSpring(x, p, C);
    } else if constexpr (NS == 3 && NC == 3) {
        if (type == ORC_MEMBRANE_BARAFF_WITKIN)
            constraintMembraneBw(x, p, C);
        else
            constraintStvk(x, p, C);
    } else if constexpr (NS == 3 && NC == 4) {
        constraintMembraneChoi(x, p, C);
    } else if constexpr (NS == 4 && NC == 1) {
        if (type == ORC_BARAFF_WITKIN_BENDING)
            constraintBendingBw(x, p, C);
        else
            constraintBending(x, p, C);
    } else if constexpr (NS == 4 && NC == 2) {
        if (type == ORC_STVK)
            constraintStvkTet(x, p, C);
        else if (type == ORC_STABLE_SQUARED_NEO_HOOKEAN)
            constraintSsnh(x, p, C);
        else
            constraintSnh(x, p, C);
    } else if constexpr (NS == 4 && NC == 3) {
        constraintQuadraticBending(x, p, C);
    }
}

// Constraint values and their gradient by forward AD, as StencilBlockEnergy::dConstraintsAndGradient
// does through TinyAD (energy.h:289-292). grad is (3 NS) x NC row-major.
template <int NS, int NC>
void adConstraints(int type, const double *xs, const double *p, double *c, double *grad)
{
    constexpr int N = 3 * NS;
    using D = Dual<N, false>;
    V3<D> x[NS];
    for (int s = 0; s < NS; ++s) {
        x[s].x = D::var(xs[3 * s + 0], 3 * s + 0);
        x[s].y = D::var(xs[3 * s + 1], 3 * s + 1);
        x[s].z = D::var(xs[3 * s + 2], 3 * s + 2);
    }
    D C[NC];
    evalConstraintsN<NS, NC>(type, x, p, C);
    for (int j = 0; j < NC; ++j) {
        c[j] = C[j].v;
        for (int i = 0; i < N; ++i) grad[i * NC + j] = C[j].g[i];
    }
}

// Fixed-size inverse of the NC x NC constraint matrix (1x1, 2x2 adjugate, 3x3 cofactors).
template <int NC>
void smallInverse(const double *A, double *Ai)
{
    if constexpr (NC == 1) {
        Ai[0] = 1.0 / A[0];
    } else if constexpr (NC == 2) {
        const double invdet = 1.0 / (A[0] * A[3] - A[1] * A[2]);
        Ai[0] = A[3] * invdet;
        Ai[1] = -A[1] * invdet;
        Ai[2] = -A[2] * invdet;
        Ai[3] = A[0] * invdet;
    } else if constexpr (NC == 3) {
        const M3 r = inverse3({A[0], A[1], A[2], A[3], A[4], A[5], A[6], A[7], A[8]});
        for (int k = 0; k < 9; ++k) Ai[k] = r[k];
    } else {
        // Gauss-Jordan elimination with partial pivoting on [A | I]
        double m[NC][2 * NC];
        for (int r = 0; r < NC; ++r)
            for (int c = 0; c < 2 * NC; ++c) m[r][c] = c < NC ? A[r * NC + c] : (c - NC == r ? 1.0 : 0.0);
        for (int c = 0; c < NC; ++c) {
            int piv = c;
            for (int r = c + 1; r < NC; ++r)
                if (std::fabs(m[r][c]) > std::fabs(m[piv][c])) piv = r;
            if (piv != c)
                for (int k = 0; k < 2 * NC; ++k) std::swap(m[c][k], m[piv][k]);
            const double inv = 1.0 / m[c][c];
            for (int k = 0; k < 2 * NC; ++k) m[c][k] *= inv;
            for (int r = 0; r < NC; ++r) {
                if (r == c) continue;
                const double f = m[r][c];
                for (int k = 0; k < 2 * NC; ++k) m[r][k] -= f * m[c][k];
            }
        }
        for (int r = 0; r < NC; ++r)
            for (int c = 0; c < NC; ++c) Ai[r * NC + c] = m[r][NC + c];
    }
}

// StencilBlockEnergy::projectPosition (energy.h:117-153): W = diag(inverse masses),
// A = grad^T W grad + compliance / dt^2 + 1e-10 I, dlambda = -A^-1 C, dx = W grad dlambda.
template <int NS, int NC>
void projectGeneric(int type, const double *xs, const double *w, const double *p, const double *cmp, double dt,
                    double *dx)
{
    constexpr int N = 3 * NS;
    double c[NC], grad[N * NC], Wg[N * NC], A[NC * NC], Ai[NC * NC], dl[NC];
    adConstraints<NS, NC>(type, xs, p, c, grad);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < NC; ++j) Wg[i * NC + j] = w[i / 3] * grad[i * NC + j];
    for (int a = 0; a < NC; ++a)
        for (int b = 0; b < NC; ++b) {
            double s = 0.0;
            for (int i = 0; i < N; ++i) s += grad[i * NC + a] * Wg[i * NC + b];
            A[a * NC + b] = s + cmp[a * NC + b] / (dt * dt) + (a == b ? 1e-10 : 0.0);
        }
    smallInverse<NC>(A, Ai);
    for (int a = 0; a < NC; ++a) {
        double s = 0.0;
        for (int b = 0; b < NC; ++b) s += Ai[a * NC + b] * c[b];
        dl[a] = -s;
    }
    for (int i = 0; i < N; ++i) {
        double s = 0.0;
        for (int a = 0; a < NC; ++a) s += Wg[i * NC + a] * dl[a];
        dx[i] = s;
    }
}

// SpringEnergy::projectPosition (springEnergy.cpp:48-78), the spring's own closed form.
void projectSpring(const double *xs, const double *w, const double *p, const double *cmp, double dt, double *dx)
{
    constexpr double eps = 1.0e-6;  // springEnergy.cpp:8
    for (int i = 0; i < 6; ++i) dx[i] = 0.0;
    const double wSum = w[0] + w[1];
    if (wSum < eps) return;
    const double alphaTilde = cmp[0] / (dt * dt);
    double n[3] = {xs[3] - xs[0], xs[4] - xs[1], xs[5] - xs[2]};
    const double d = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (d < eps) return;
    for (double &k : n) k = k / d;
    for (int k = 0; k < 3; ++k) {
        const double corr = n[k] * (d - p[0]) / (wSum + alphaTilde);
        dx[k] = w[0] * corr;
        dx[3 + k] = -(w[1] * corr);
    }
}

void projectStencilDx(int type, const double *xs, const double *w, const double *p, const double *cmp, double dt,
                      double *dx)
{
    switch (type) {
    case ORC_SPRING: projectSpring(xs, w, p, cmp, dt, dx); return;
    case ORC_SPRING_ANCHOR: projectGeneric<1, 1>(type, xs, w, p, cmp, dt, dx); return;
    case ORC_MEMBRANE_STVK: projectGeneric<3, 3>(type, xs, w, p, cmp, dt, dx); return;
    case ORC_DISCRETE_BENDING: projectGeneric<4, 1>(type, xs, w, p, cmp, dt, dx); return;
    case ORC_STABLE_NEO_HOOKEAN: projectGeneric<4, 2>(type, xs, w, p, cmp, dt, dx); return;
    case ORC_SPRING_CONTINUUM:
    case ORC_SPRING_SQUARED_CONTINUUM: projectGeneric<2, 1>(type, xs, w, p, cmp, dt, dx); return;
    case ORC_MEMBRANE_BARAFF_WITKIN: projectGeneric<3, 3>(type, xs, w, p, cmp, dt, dx); return;
    case ORC_MEMBRANE_CHOI: projectGeneric<3, 4>(type, xs, w, p, cmp, dt, dx); return;
    case ORC_QUADRATIC_BENDING: projectGeneric<4, 3>(type, xs, w, p, cmp, dt, dx); return;
    case ORC_BARAFF_WITKIN_BENDING: projectGeneric<4, 1>(type, xs, w, p, cmp, dt, dx); return;
    case ORC_STVK:
    case ORC_STABLE_SQUARED_NEO_HOOKEAN: projectGeneric<4, 2>(type, xs, w, p, cmp, dt, dx); return;
    }
    throw OracleError("unknown energy type");
}

// ParticleGroup::invMasses (particleGroup.cpp:81): 0 for zero mass, 1/m otherwise.
inline double invMass(double m) { return m == 0 ? 0 : 1 / m; }

// projectPosition(i) on the object's positions, then updateElementPositionFromDx per stencil slot
// (particleGroup.cpp:43-46), slots in order.
void projectStencil(Object &o, const Energy &e, int64_t i, double dt)
{
    double xs[12], w[4], dx[12];
    gatherStencil(o, o.x.data(), e, i, xs);
    for (int k = 0; k < e.ssize; ++k) w[k] = invMass(o.m[e.st[i * e.ssize + k]]);
    const int nc = numConstraints(e.type);
    projectStencilDx(e.type, xs, w, &e.prm[i * preparedParams(e.type)], &e.cmp[i * nc * nc], dt, dx);
    for (int k = 0; k < e.ssize; ++k) {
        const int64_t vi = e.st[i * e.ssize + k];
        for (int d = 0; d < 3; ++d) o.x[3 * vi + d] += dx[3 * k + d];
    }
}

// coloring::colorStencils (graphColoring.cpp:51-70): a graph node per stencil, an edge between every
// two stencils that share a vertex, then Graph::greedyColoring (graphColoring.cpp:21-49): node 0 takes
// colour 0, every later node the smallest colour not used by an already coloured neighbour.
std::vector<int32_t> colorStencils(const int32_t *flat, int64_t nStencils, int ssize)
{
    std::vector<int32_t> col(nStencils, -1);
    if (nStencils == 0) return col;
    int32_t maxV = -1;
    for (int64_t k = 0; k < nStencils * ssize; ++k) maxV = std::max(maxV, flat[k]);
    std::vector<std::vector<int32_t>> perVertex(maxV + 1);
    for (int64_t k = 0; k < nStencils * ssize; ++k) perVertex[flat[k]].push_back(static_cast<int32_t>(k / ssize));
    std::vector<std::vector<int32_t>> adj(nStencils);
    for (const auto &list : perVertex)
        for (size_t a = 0; a < list.size(); ++a)
            for (size_t b = a + 1; b < list.size(); ++b) {
                adj[list[a]].push_back(list[b]);
                adj[list[b]].push_back(list[a]);
            }
    // the availability flags are reset over the touched colours only (the reference refills an
    // n-sized array per node): same colours, O(edges) instead of O(n^2)
    col[0] = 0;
    std::vector<char> used(nStencils + 1, 0);
    for (int64_t n = 1; n < nStencils; ++n) {
        for (int32_t q : adj[n])
            if (col[q] != -1) used[col[q]] = 1;
        int32_t c = 0;
        while (used[c]) ++c;
        col[n] = c;
        for (int32_t q : adj[n])
            if (col[q] != -1) used[col[q]] = 0;
    }
    return col;
}

// One XPBD::step (xpbd.cpp:16-132) of nsub substeps on one object.
void xpbdStep(Object &o, double dtStep, int nsub, const double *g, bool parallel)
{
    const double dt = dtStep / nsub;  // xpbd.cpp:22
    const double invdt = 1 / dt;
    // computeParallelStencilGroups (xpbd.cpp:134-160): per energy, stencil ids per colour, ascending;
    // computed on the first step and kept (xpbd.cpp:18-21)
    if (parallel && !o.xpbdGroupsDone) {
        o.xpbdGroups.assign(o.energies.size(), {});
        for (size_t ei = 0; ei < o.energies.size(); ++ei) {
            const auto &e = o.energies[ei];
            const auto col = colorStencils(e.st.data(), e.n(), e.ssize);
            for (int64_t i = 0; i < e.n(); ++i) {
                if (static_cast<int32_t>(o.xpbdGroups[ei].size()) < col[i] + 1) o.xpbdGroups[ei].resize(col[i] + 1);
                o.xpbdGroups[ei][col[i]].push_back(static_cast<int32_t>(i));
            }
        }
        o.xpbdGroupsDone = true;
    }
    const auto &groups = o.xpbdGroups;
    const int64_t nv = o.nv;
    for (int s = 0; s < nsub; ++s) {
        const std::vector<double> xOld(o.x);  // xpbd.cpp:47-54
        const double dtg[3] = {dt * g[0], dt * g[1], dt * g[2]};
        for (int64_t i = 0; i < nv; ++i)
            for (int d = 0; d < 3; ++d) {
                o.v[3 * i + d] += dtg[d];              // xpbd.cpp:58-66 (every particle, also m = 0)
                o.x[3 * i + d] += o.v[3 * i + d] * dt;  // particleGroup.cpp:57-63
            }
        for (size_t ei = 0; ei < o.energies.size(); ++ei) {
            const Energy &e = o.energies[ei];
            if (parallel) {
                for (const auto &group : groups[ei]) {
                    const int64_t gs = static_cast<int64_t>(group.size());
#pragma omp parallel for schedule(static)
                    for (int64_t k = 0; k < gs; ++k) projectStencil(o, e, group[k], dt);
                }
            } else {
                for (int64_t i = 0; i < e.n(); ++i) projectStencil(o, e, i, dt);
            }
        }
        for (int64_t i = 0; i < 3 * nv; ++i) o.v[i] = (o.x[i] - xOld[i]) * invdt;  // particleGroup.cpp:65-72
    }
}

// ---------------------------------------------------------------------------------------------
// Implicit Euler family (spg/source/sim/solver/implicitEulerBase.cpp:102-240, implicitEulerBaraffWitkin.cpp,
// implicitEulerNewtonDv.cpp, implicitEulerNewtonDx.cpp, implicitEulerNewtonRobust.cpp, bdf2.cpp):
// global force vector, mass and stiffness matrices, and Eigen's Jacobi-preconditioned conjugate gradient.
// DOF layout: 3 per particle, particle-major (ParticleGroup::getPositions, particleGroup.cpp:7-14).
// ---------------------------------------------------------------------------------------------

// Full element gradient and Hessian (3N and 3N x 3N) as energyGradient / energyHessian do: the spring by
// its own closed form (springEnergy.cpp:101-134), the others by dual numbers over the energy (energy.h:268-270,
// 348-362; no positive-definite projection).
void elementGradHess(int type, const double *xs, const double *p, double *grad, double *hess)
{
    if (type == ORC_SPRING) {
        const A3 s = a3sub(xs, xs + 3);
        const double n = std::sqrt(a3dot(s, s));
        const double k = p[1], L0 = p[0];
        const A3 sh = {s[0] / n, s[1] / n, s[2] / n};
        for (int i = 0; i < 3; ++i) {
            const double g = k * (n - L0) * s[i] / n;
            if (grad) {
                grad[i] = g;
                grad[3 + i] = -g;
            }
        }
        if (hess)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    const double sst = sh[i] * sh[j];
                    const double b = k * ((1 - L0 / n) * ((i == j ? 1.0 : 0.0) - sst) + sst);
                    hess[i * 6 + j] = b;
                    hess[(3 + i) * 6 + 3 + j] = b;
                    hess[i * 6 + 3 + j] = -b;
                    hess[(3 + i) * 6 + j] = -b;
                }
        return;
    }
    switch (stencilSize(type)) {
    case 1:
        if (grad) adElement<1, false>(type, xs, p, nullptr, grad, nullptr);
        if (hess) adElement<1, true>(type, xs, p, nullptr, nullptr, hess);
        return;
    case 2:
        if (grad) adElement<2, false>(type, xs, p, nullptr, grad, nullptr);
        if (hess) adElement<2, true>(type, xs, p, nullptr, nullptr, hess);
        return;
    case 3:
        if (grad) adElement<3, false>(type, xs, p, nullptr, grad, nullptr);
        if (hess) adElement<3, true>(type, xs, p, nullptr, nullptr, hess);
        return;
    case 4:
        if (grad) adElement<4, false>(type, xs, p, nullptr, grad, nullptr);
        if (hess) adElement<4, true>(type, xs, p, nullptr, nullptr, hess);
        return;
    }
    throw OracleError("unknown energy type");
}

// Compressed sparse rows with sorted, unique columns (Eigen's setFromTriplets sums duplicates).
struct Csr {
    int64_t n = 0;
    std::vector<int64_t> ptr;
    std::vector<int64_t> col;
    std::vector<double> val;
};

struct Trip {
    int64_t r, c;
    double v;
};

Csr fromTriplets(int64_t n, std::vector<Trip> &t)
{
    std::stable_sort(t.begin(), t.end(), [](const Trip &a, const Trip &b) { return a.r != b.r ? a.r < b.r : a.c < b.c; });
    Csr m;
    m.n = n;
    m.ptr.assign(n + 1, 0);
    for (size_t k = 0; k < t.size();) {
        size_t j = k;
        double s = 0.0;
        while (j < t.size() && t[j].r == t[k].r && t[j].c == t[k].c) s += t[j++].v;
        m.col.push_back(t[k].c);
        m.val.push_back(s);
        ++m.ptr[t[k].r + 1];
        k = j;
    }
    for (int64_t r = 0; r < n; ++r) m.ptr[r + 1] += m.ptr[r];
    return m;
}

void spmv(const Csr &A, const std::vector<double> &x, std::vector<double> &y)
{
    y.assign(A.n, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < A.n; ++r) {
        double s = 0.0;
        for (int64_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k) s += A.val[k] * x[A.col[k]];
        y[r] = s;
    }
}

// ImplicitEulerBase::getSystemForce (implicitEulerBase.cpp:102-143): gravity * mass + sum of -grad E.
std::vector<double> systemForce(const Object &o, const double *x, const double *g)
{
    std::vector<double> f(3 * o.nv, 0.0);
    for (int64_t i = 0; i < o.nv; ++i)
        for (int d = 0; d < 3; ++d) f[3 * i + d] += g[d] * o.m[i];
    for (const auto &e : o.energies) {
        const int np = preparedParams(e.type);
        const int64_t n = e.n();
        std::vector<double> eg(n * 3 * e.ssize);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            double xs[12];
            gatherStencil(o, x, e, i, xs);
            elementGradHess(e.type, xs, &e.prm[i * np], &eg[i * 3 * e.ssize], nullptr);
        }
        for (int64_t i = 0; i < n; ++i)  // accumulateForces (energy.h:197-215), stencil order
            for (int s = 0; s < e.ssize; ++s)
                for (int d = 0; d < 3; ++d) f[3 * e.st[i * e.ssize + s] + d] -= eg[(i * e.ssize + s) * 3 + d];
    }
    return f;
}

// getSystemStiffnessMatrix (implicitEulerBase.cpp:180-222): triplets of -Hessian of every stencil.
Csr stiffnessMatrix(const Object &o, const double *x)
{
    std::vector<Trip> t;
    for (const auto &e : o.energies) {
        const int np = preparedParams(e.type);
        const int N = 3 * e.ssize;
        const int64_t n = e.n();
        std::vector<double> eh(n * N * N);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            double xs[12];
            gatherStencil(o, x, e, i, xs);
            elementGradHess(e.type, xs, &e.prm[i * np], nullptr, &eh[i * N * N]);
        }
        for (int64_t i = 0; i < n; ++i)  // negativeHessianTriplets (energy.h:218-236)
            for (int a = 0; a < e.ssize; ++a)
                for (int b = 0; b < e.ssize; ++b)
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c)
                            t.push_back({3 * int64_t(e.st[i * e.ssize + a]) + r, 3 * int64_t(e.st[i * e.ssize + b]) + c,
                                         -eh[i * N * N + (3 * a + r) * N + 3 * b + c]});
    }
    return fromTriplets(3 * o.nv, t);
}

// alpha * M + beta * K with M = diag(masses) (getSystemMassMatrix, implicitEulerBase.cpp:145-178).
Csr combine(const Object &o, double alpha, const Csr &K, double beta)
{
    std::vector<Trip> t;
    for (int64_t r = 0; r < K.n; ++r)
        for (int64_t k = K.ptr[r]; k < K.ptr[r + 1]; ++k) t.push_back({r, K.col[k], beta * K.val[k]});
    for (int64_t i = 0; i < o.nv; ++i)
        for (int d = 0; d < 3; ++d) t.push_back({3 * i + d, 3 * i + d, alpha * o.m[i]});
    return fromTriplets(K.n, t);
}

double dot(const std::vector<double> &a, const std::vector<double> &b)
{
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

// Eigen's ConjugateGradient<SparseMatrix, Lower|Upper> with its default DiagonalPreconditioner, as
// solveLinearSystem calls it (implicitEulerBase.cpp:224-239; tolerance 1e-5, 300 iterations,
// implicitEulerBase.h:26-27). Eigen is an empty submodule of the reference; this restates its published
// algorithm (Eigen 3 IterativeLinearSolvers, conjugate_gradient): zero initial guess, relative residual
// threshold max(tol^2 |b|^2, smallest normal double), inverse-diagonal preconditioner (1 where the
// diagonal is 0 or absent), the convergence test right after the residual update.
std::vector<double> conjugateGradient(const Csr &A, const std::vector<double> &b, double tol, int maxIters, int *iters,
                                      double *error)
{
    const int64_t n = A.n;
    std::vector<double> x(n, 0.0), r = b, p(n), z(n), tmp(n), invdiag(n, 1.0);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
            if (A.col[k] == i && A.val[k] != 0.0) invdiag[i] = 1.0 / A.val[k];
    const double rhsNorm2 = dot(b, b);
    if (rhsNorm2 == 0) {
        *iters = 0;
        *error = 0;
        return x;
    }
    const double threshold = std::max(tol * tol * rhsNorm2, std::numeric_limits<double>::min());
    double residualNorm2 = dot(r, r);
    if (residualNorm2 < threshold) {
        *iters = 0;
        *error = std::sqrt(residualNorm2 / rhsNorm2);
        return x;
    }
    for (int64_t i = 0; i < n; ++i) p[i] = invdiag[i] * r[i];
    double absNew = dot(r, p);
    int it = 0;
    while (it < maxIters) {
        spmv(A, p, tmp);
        const double alpha = absNew / dot(p, tmp);
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * tmp[i];
        }
        residualNorm2 = dot(r, r);
        if (residualNorm2 < threshold) break;
        for (int64_t i = 0; i < n; ++i) z[i] = invdiag[i] * r[i];
        const double absOld = absNew;
        absNew = dot(r, z);
        const double beta = absNew / absOld;
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        ++it;
    }
    *iters = it;
    *error = std::sqrt(residualNorm2 / rhsNorm2);
    return x;
}

enum { IMPL_BARAFF_WITKIN = 0, IMPL_NEWTON_DV = 1, IMPL_NEWTON_DX = 2, IMPL_NEWTON_ROBUST = 3, IMPL_BDF2 = 4 };

struct ImplicitState {
    // NewtonRobust (implicitEulerNewtonRobust.h:19-26)
    double currentDt = -1.0;
    int consecutiveSuccessful = 0;
    // BDF2 (bdf2.h:18-19)
    std::vector<double> xMinusOne, vMinusOne;
    // counters of the last step
    int cgIterations = 0, linearSolves = 0, newtonIterations = 0, failedSteps = 0;
};

const double kCgTol = 1e-5;
const int kCgMaxIters = 300;

std::vector<double> solveCounted(const Csr &A, const std::vector<double> &b, ImplicitState &st)
{
    int it = 0;
    double err = 0;
    auto x = conjugateGradient(A, b, kCgTol, kCgMaxIters, &it, &err);
    st.cgIterations += it;
    st.linearSolves += 1;
    return x;
}

// One substep of each solver on one object.
void implicitSubstep(Object &o, ImplicitState &st, int solver, double dtStep, int nsub, const double *g)
{
    const int64_t N = 3 * o.nv;
    const double dt = dtStep / nsub;
    auto integrate = [&](double h) {  // particleGroup.cpp:57-63
        for (int64_t i = 0; i < N; ++i) o.x[i] += o.v[i] * h;
    };
    if (solver == IMPL_BARAFF_WITKIN) {  // implicitEulerBaraffWitkin.cpp:31-53
        const std::vector<double> v0 = o.v;
        const auto f = systemForce(o, o.x.data(), g);
        const Csr K = stiffnessMatrix(o, o.x.data());
        const Csr LHS = combine(o, 1.0, K, -(dt * dt));
        std::vector<double> Kv;
        spmv(K, v0, Kv);
        std::vector<double> rhs(N);
        for (int64_t i = 0; i < N; ++i) rhs[i] = dt * (f[i] + dt * Kv[i]);
        const auto dv = solveCounted(LHS, rhs, st);
        for (int64_t i = 0; i < N; ++i) o.v[i] = v0[i] + dv[i];
        integrate(dt);
    } else if (solver == IMPL_NEWTON_DV) {  // implicitEulerNewtonDv.cpp:30-61
        const std::vector<double> x0 = o.x, v0 = o.v;
        integrate(dt);
        const auto f = systemForce(o, o.x.data(), g);
        const Csr K = stiffnessMatrix(o, o.x.data());
        const Csr LHS = combine(o, 1.0, K, -(dt * dt));
        std::vector<double> rhs(N);
        for (int64_t i = 0; i < N; ++i) rhs[i] = dt * f[i];
        const auto dv = solveCounted(LHS, rhs, st);
        o.x = x0;
        for (int64_t i = 0; i < N; ++i) o.v[i] = v0[i] + dv[i];
        integrate(dt);
    } else if (solver == IMPL_NEWTON_DX) {  // implicitEulerNewtonDx.cpp:30-58
        const std::vector<double> x0 = o.x;
        integrate(dt);
        const auto f = systemForce(o, o.x.data(), g);
        const Csr K = stiffnessMatrix(o, o.x.data());
        const double invdt = 1. / dt;
        const Csr LHS = combine(o, invdt * invdt, K, -1.0);
        const auto dx = solveCounted(LHS, f, st);
        for (int64_t i = 0; i < N; ++i) {  // integrateObjectsVelocitiesFromDx (implicitEulerBase.cpp:80-92)
            o.x[i] += dx[i];
            o.v[i] = (o.x[i] - x0[i]) * invdt;
        }
    } else if (solver == IMPL_BDF2) {  // bdf2.cpp:31-73
        const std::vector<double> x0 = o.x, v0 = o.v;
        if (st.xMinusOne.size() != x0.size() || st.vMinusOne.size() != v0.size()) {
            st.xMinusOne = x0;
            st.vMinusOne = v0;
        }
        const auto f = systemForce(o, o.x.data(), g);
        const Csr K = stiffnessMatrix(o, o.x.data());
        const Csr LHS = combine(o, 1.0, K, -4.0 / 9.0 * (dt * dt));
        std::vector<double> w(N), Kw;
        for (int64_t i = 0; i < N; ++i) w[i] = x0[i] - st.xMinusOne[i] + 2 * dt * v0[i];
        spmv(K, w, Kw);
        std::vector<double> rhs(N);
        for (int64_t i = 0; i < N; ++i)
            rhs[i] = 1.0 / 3.0 * o.m[i / 3] * (v0[i] - st.vMinusOne[i]) + 2.0 / 3.0 * dt * (f[i] + 1.0 / 3.0 * Kw[i]);
        const auto dv = solveCounted(LHS, rhs, st);
        for (int64_t i = 0; i < N; ++i) {
            o.v[i] = v0[i] + dv[i];
            o.x[i] = 4.0 / 3.0 * x0[i] - 1.0 / 3.0 * st.xMinusOne[i] + 2.0 / 3.0 * o.v[i] * dt;
        }
        st.xMinusOne = x0;
        st.vMinusOne = v0;
    } else {  // IMPL_NEWTON_ROBUST, implicitEulerNewtonRobust.cpp:31-122
        const double newtonThreshold = 1e-5, splitting = 0.75;
        const int maxNewton = 5, maxLineSearch = 5, increaseAfter = 5;
        bool ok = true;
        const double maxDt = dtStep / nsub;
        if (st.currentDt == -1. || st.currentDt > maxDt) st.currentDt = maxDt;
        const double h = st.currentDt;
        const std::vector<double> x0 = o.x, v0 = o.v;
        int newton = 0;
        double stepResidual = 0.;
        integrate(h);
        std::vector<double> vi = v0;
        do {
            auto f = systemForce(o, o.x.data(), g);
            const Csr K = stiffnessMatrix(o, o.x.data());
            const Csr LHS = combine(o, 1.0, K, -(h * h));
            std::vector<double> rhs(N);
            for (int64_t i = 0; i < N; ++i) rhs[i] = h * f[i] - o.m[i / 3] * (vi[i] - v0[i]);
            const double initialResidual = std::sqrt(dot(rhs, rhs));
            const auto dv = solveCounted(LHS, rhs, st);
            stepResidual = initialResidual;
            int ls = 0;
            double alpha = 1.;
            while (stepResidual >= initialResidual && ls < maxLineSearch) {
                std::vector<double> v(N);
                for (int64_t i = 0; i < N; ++i) v[i] = vi[i] + dv[i] * alpha;
                o.v = v;
                o.x = x0;
                integrate(h);
                f = systemForce(o, o.x.data(), g);
                std::vector<double> res(N);
                for (int64_t i = 0; i < N; ++i) res[i] = h * f[i] - o.m[i / 3] * (v[i] - v0[i]);
                stepResidual = std::sqrt(dot(res, res));
                if (stepResidual < initialResidual) vi = v;
                alpha *= 0.5;
                ++ls;
            }
            if (stepResidual >= initialResidual && stepResidual > newtonThreshold) ok = false;
            ++newton;
        } while (stepResidual > newtonThreshold && ok && newton < maxNewton);
        st.newtonIterations += newton;
        if (stepResidual > newtonThreshold) ok = false;
        if (!ok) {
            o.x = x0;
            o.v = v0;
            st.currentDt *= splitting;
            st.consecutiveSuccessful = 0;
            ++st.failedSteps;
        } else {
            ++st.consecutiveSuccessful;
            if (st.consecutiveSuccessful > increaseAfter && st.currentDt < maxDt) st.currentDt /= splitting;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Rigid-body VBD (vbd.cpp:38-185 with the RigidBodyGroup branches, rigidBodyGroup.cpp:19-243, and the
// energies rigidBodySpringEnergy.cpp, rigidBodySpringSquaredEnergy.cpp, rigidBodySpringAnchorEnergy.cpp,
// rigidBodyOrientationAlignmentEnergy.cpp). Each body has 6 DOFs: its centre of mass x and a rotation
// increment theta applied on the left of its rotation matrix R, (I + [theta]x) R, differentiated at 0.
// ---------------------------------------------------------------------------------------------
enum { RB_SPRING = 21, RB_SPRING_SQUARED = 22, RB_SPRING_ANCHOR = 23, RB_ORIENTATION_ALIGNMENT = 24 };

int rbStencilSize(int type) { return type == RB_SPRING_ANCHOR ? 1 : (type >= RB_SPRING && type <= 24) ? 2 : 0; }
int rbRawParams(int type)
{
    switch (type) {
    case RB_SPRING:
    case RB_SPRING_SQUARED: return 8;  // local points (2 x 3), rest length, stiffness
    case RB_SPRING_ANCHOR: return 7;   // local point, anchor, stiffness
    case RB_ORIENTATION_ALIGNMENT: return 1;
    }
    return 0;
}

// Eigen::AngleAxis(angle, axis).toRotationMatrix(): Rodrigues' formula; a zero vector is the identity
// (rigidBodyGroup.cpp:6-10).
M3 angleAxisToR(double angle, const A3 &a)
{
    const double c = std::cos(angle), sn = std::sin(angle);
    const A3 s = {sn * a[0], sn * a[1], sn * a[2]};
    const A3 c1 = {(1 - c) * a[0], (1 - c) * a[1], (1 - c) * a[2]};
    M3 r;
    for (int i = 0; i < 3; ++i) r[4 * i] = c1[i] * a[i] + c;
    double t = c1[0] * a[1];
    r[1] = t - s[2];
    r[3] = t + s[2];
    t = c1[0] * a[2];
    r[2] = t + s[1];
    r[6] = t - s[1];
    t = c1[1] * a[2];
    r[5] = t - s[0];
    r[7] = t + s[0];
    return r;
}

M3 aaToR(const A3 &t)
{
    const double angle = std::sqrt(a3dot(t, t));
    if (angle == 0) return {1, 0, 0, 0, 1, 0, 0, 0, 1};
    return angleAxisToR(angle, {t[0] / angle, t[1] / angle, t[2] / angle});
}

// Eigen::AngleAxis(R) as angle * axis (rigidBodyGroup.cpp:12-16): the unit quaternion of R by Shoemake's
// method (the largest of trace and diagonal picks the stable formula), then angle = 2 atan2(|v|, |w|) and
// axis = v / |v|, negated when w < 0; the identity maps to 0.
A3 rToAa(const M3 &R)
{
    double q[4];  // w, x, y, z
    const double tr = R[0] + R[4] + R[8];
    if (tr > 0) {
        double t = std::sqrt(tr + 1.0);
        q[0] = 0.5 * t;
        t = 0.5 / t;
        q[1] = (R[7] - R[5]) * t;
        q[2] = (R[2] - R[6]) * t;
        q[3] = (R[3] - R[1]) * t;
    } else {
        int i = 0;
        if (R[4] > R[0]) i = 1;
        if (R[8] > R[4 * i]) i = 2;
        const int j = (i + 1) % 3, k = (j + 1) % 3;
        double t = std::sqrt(R[4 * i] - R[4 * j] - R[4 * k] + 1.0);
        q[1 + i] = 0.5 * t;
        t = 0.5 / t;
        q[0] = (R[3 * k + j] - R[3 * j + k]) * t;
        q[1 + j] = (R[3 * j + i] + R[3 * i + j]) * t;
        q[1 + k] = (R[3 * k + i] + R[3 * i + k]) * t;
    }
    double n = std::sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (n == 0) return {0, 0, 0};
    const double angle = 2.0 * std::atan2(n, std::fabs(q[0]));
    if (q[0] < 0) n = -n;
    return {q[1] / n * angle, q[2] / n * angle, q[3] / n * angle};
}

M3 m3T(const M3 &a) { return {a[0], a[3], a[6], a[1], a[4], a[7], a[2], a[5], a[8]}; }
M3 m3mul(const M3 &a, const M3 &b)
{
    M3 r{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
    return r;
}
A3 m3vec(const M3 &a, const A3 &v)
{
    return {a[0] * v[0] + a[1] * v[1] + a[2] * v[2], a[3] * v[0] + a[4] * v[1] + a[5] * v[2],
            a[6] * v[0] + a[7] * v[1] + a[8] * v[2]};
}

struct RbObject {
    int64_t nb = 0;
    std::vector<A3> x, theta, v, omega;
    std::vector<M3> R, Iloc, IlocInv, I, Iinv;
    std::vector<double> m;
    Object shape;  // stencils of the energies, for the colouring (colorVertices)
    std::vector<std::vector<double>> params;  // per energy, raw addStencil arguments
    bool precompDone = false;
    std::vector<std::vector<int32_t>> groups;
    std::vector<std::vector<std::array<int32_t, 3>>> inc;  // per body: (energy, stencil, slot)
};

void rbUpdateInertia(RbObject &o, int64_t b)  // rigidBodyGroup.cpp:195-199
{
    o.I[b] = m3mul(m3mul(o.R[b], o.Iloc[b]), m3T(o.R[b]));
    o.Iinv[b] = m3mul(m3mul(o.R[b], o.IlocInv[b]), m3T(o.R[b]));
}

// (I + [theta]x) R l with theta a dual vector at 0: R l + theta x (R l)
template <class S>
V3<S> rbPoint(const V3<S> &com, const V3<S> &th, const M3 &R, const double *l)
{
    const A3 r = m3vec(R, {l[0], l[1], l[2]});
    const V3<S> rs = {S(r[0]), S(r[1]), S(r[2])};
    const V3<S> t = cross(th, rs);
    return {com.x + rs.x + t.x, com.y + rs.y + t.y, com.z + rs.z + t.z};
}

// The element energy over the stencil's 6 NS DOFs, as the reference differentiates it through TinyAD.
template <int NS, class S>
S rbEnergy(int type, const V3<S> *com, const V3<S> *th, const M3 *R, const double *p)
{
    if constexpr (NS == 1) {  // rigidBodySpringAnchorEnergy.cpp:35-45: 1/2 k |p - a|^2
        const V3<S> q = rbPoint(com[0], th[0], R[0], p);
        const V3<S> d = {q.x - p[3], q.y - p[4], q.z - p[5]};
        return 0.5 * p[6] * dot(d, d);
    } else {
        if (type == RB_ORIENTATION_ALIGNMENT) {  // C = sum_k |col_k(R0') x col_k(R1')|^2
            S c = S(0.0);
            for (int k = 0; k < 3; ++k) {
                const V3<S> u = {S(R[0][k]), S(R[0][3 + k]), S(R[0][6 + k])};
                const V3<S> w = {S(R[1][k]), S(R[1][3 + k]), S(R[1][6 + k])};
                const V3<S> a = {u.x + cross(th[0], u).x, u.y + cross(th[0], u).y, u.z + cross(th[0], u).z};
                const V3<S> b = {w.x + cross(th[1], w).x, w.y + cross(th[1], w).y, w.z + cross(th[1], w).z};
                const V3<S> ab = cross(a, b);
                c = c + dot(ab, ab);
            }
            return 0.5 * c * (p[0] * c);
        }
        const V3<S> q0 = rbPoint(com[0], th[0], R[0], p), q1 = rbPoint(com[1], th[1], R[1], p + 3);
        const V3<S> d = sub(q0, q1);
        const S c = type == RB_SPRING ? sqrt(dot(d, d)) - p[6] : dot(d, d) - p[6];
        return 0.5 * c * (p[7] * c);
    }
}

// Gradient (6 NS) and Hessian of one element at theta = 0 (6 NS x 6 NS, row-major).
template <int NS>
void rbElementAD(int type, const double *com, const M3 *R, const double *p, double *grad, double *hess)
{
    constexpr int N = 6 * NS;
    using D = Dual<N, true>;
    V3<D> c[NS], t[NS];
    for (int s = 0; s < NS; ++s) {
        c[s] = {D::var(com[3 * s], 6 * s), D::var(com[3 * s + 1], 6 * s + 1), D::var(com[3 * s + 2], 6 * s + 2)};
        t[s] = {D::var(0.0, 6 * s + 3), D::var(0.0, 6 * s + 4), D::var(0.0, 6 * s + 5)};
    }
    const D e = rbEnergy<NS>(type, c, t, R, p);
    for (int i = 0; i < N; ++i) grad[i] = e.g[i];
    if (hess)
        for (int i = 0; i < N * N; ++i) hess[i] = e.h[i];
}

// prepared parameters: the squared spring keeps L0^2 (rigidBodySpringSquaredEnergy.cpp:35-48)
void rbPrepareParams(int type, const double *raw, double *p)
{
    for (int k = 0; k < rbRawParams(type); ++k) p[k] = raw[k];
    if (type == RB_SPRING_SQUARED) p[6] = raw[6] * raw[6];
}

// Solve H dx = f by LU with partial pivoting (H.inverse() * f); returns det(H).
double solve6(double H[6][6], const double *f, double *dx)
{
    double a[6][7];
    for (int i = 0; i < 6; ++i) {
        for (int j = 0; j < 6; ++j) a[i][j] = H[i][j];
        a[i][6] = f[i];
    }
    double det = 1.0;
    for (int c = 0; c < 6; ++c) {
        int piv = c;
        for (int r = c + 1; r < 6; ++r)
            if (std::fabs(a[r][c]) > std::fabs(a[piv][c])) piv = r;
        if (piv != c) {
            for (int k = 0; k < 7; ++k) std::swap(a[c][k], a[piv][k]);
            det = -det;
        }
        det *= a[c][c];
        if (a[c][c] == 0) return 0.0;
        for (int r = c + 1; r < 6; ++r) {
            const double fct = a[r][c] / a[c][c];
            for (int k = c; k < 7; ++k) a[r][k] -= fct * a[c][k];
        }
    }
    for (int i = 5; i >= 0; --i) {
        double t = a[i][6];
        for (int k = i + 1; k < 6; ++k) t -= a[i][k] * dx[k];
        dx[i] = t / a[i][i];
    }
    return det;
}

void rbPrecompute(RbObject &o)
{
    const auto col = colorVertices(o.shape);
    int32_t nc = 0;
    for (auto c : col) nc = std::max(nc, c + 1);
    o.groups.assign(nc, {});
    for (int64_t b = 0; b < o.nb; ++b) o.groups[col[b]].push_back(static_cast<int32_t>(b));
    o.inc.assign(o.nb, {});
    for (size_t e = 0; e < o.shape.energies.size(); ++e) {
        const auto &en = o.shape.energies[e];
        for (int64_t i = 0; i < en.n(); ++i)
            for (int s = 0; s < en.ssize; ++s)
                o.inc[en.st[i * en.ssize + s]].push_back({static_cast<int32_t>(e), static_cast<int32_t>(i), s});
    }
    o.precompDone = true;
}

// omega from the rotation R R(thetaOld)^T over h (RigidBodyGroup::computeIntegratedOmega)
A3 rbIntegratedOmega(const RbObject &o, int64_t b, const A3 &thetaOld, double invdt)
{
    const A3 w = rToAa(m3mul(o.R[b], m3T(aaToR(thetaOld))));
    return {w[0] * invdt, w[1] * invdt, w[2] * invdt};
}

void rbVertexDescent(RbObject &o, int64_t b, const std::vector<A3> &xT, const std::vector<A3> &thOld, double h)
{
    const double invdt = 1 / h, invdt2 = invdt * invdt;
    const double mdt2 = o.m[b] * invdt2;
    double f[6], H[6][6] = {};
    for (int d = 0; d < 3; ++d) f[d] = -mdt2 * (o.x[b][d] - xT[b][d]);
    const A3 wi = rbIntegratedOmega(o, b, thOld[b], invdt);
    const A3 dw = {wi[0] - o.omega[b][0], wi[1] - o.omega[b][1], wi[2] - o.omega[b][2]};
    const A3 Idw = m3vec(o.I[b], dw);
    for (int d = 0; d < 3; ++d) f[3 + d] = -invdt * Idw[d];
    for (int d = 0; d < 3; ++d) {
        H[d][d] = mdt2;
        for (int q = 0; q < 3; ++q) H[3 + d][3 + q] = invdt2 * o.I[b][3 * d + q];
    }
    for (const auto &in : o.inc[b]) {
        const auto &en = o.shape.energies[in[0]];
        const int ns = en.ssize, N = 6 * ns;
        double com[6];
        M3 R[2];
        for (int s = 0; s < ns; ++s) {
            const int64_t bb = en.st[in[1] * ns + s];
            for (int d = 0; d < 3; ++d) com[3 * s + d] = o.x[bb][d];
            R[s] = o.R[bb];
        }
        double p[8];
        rbPrepareParams(en.type, &o.params[in[0]][in[1] * rbRawParams(en.type)], p);
        double g[12], hs[144];
        if (ns == 1)
            rbElementAD<1>(en.type, com, R, p, g, hs);
        else
            rbElementAD<2>(en.type, com, R, p, g, hs);
        const int sl = in[2];
        for (int i = 0; i < 6; ++i) {
            f[i] -= g[6 * sl + i];
            for (int j = 0; j < 6; ++j) H[i][j] += hs[(6 * sl + i) * N + 6 * sl + j];
        }
    }
    double dx[6];
    double Hc[6][6];
    std::memcpy(Hc, H, sizeof(H));
    if (solve6(Hc, f, dx) > 1e-10) {  // vbd.cpp:139-145, updateElementPositionFromDx (rigidBodyGroup.cpp:64-77)
        for (int d = 0; d < 3; ++d) o.x[b][d] += dx[d];
        const A3 dt = {dx[3], dx[4], dx[5]};
        if (std::sqrt(a3dot(dt, dt)) != 0) o.R[b] = m3mul(aaToR(dt), o.R[b]);
        o.theta[b] = rToAa(o.R[b]);
        rbUpdateInertia(o, b);
    }
}

void rbStep(RbObject &o, double dtStep, int nsub, int iterations, const double *g)
{
    if (!o.precompDone) rbPrecompute(o);
    const double h = dtStep / nsub, invdt = 1 / h;
    for (int sub = 0; sub < nsub; ++sub) {
        const std::vector<A3> xOld = o.x, thOld = o.theta;
        for (int64_t b = 0; b < o.nb; ++b) {  // vbd.cpp:62-72
            for (int d = 0; d < 3; ++d) o.v[b][d] += h * g[d];
            const A3 Iw = m3vec(o.I[b], o.omega[b]);
            const A3 c = a3cross(o.omega[b], Iw);
            const A3 acc = m3vec(o.Iinv[b], {-c[0], -c[1], -c[2]});
            for (int d = 0; d < 3; ++d) o.omega[b][d] += h * acc[d];
        }
        for (int64_t b = 0; b < o.nb; ++b) {  // integrateVelocities (rigidBodyGroup.cpp:97-113)
            for (int d = 0; d < 3; ++d) o.x[b][d] += o.v[b][d] * h;
            const double wn = std::sqrt(a3dot(o.omega[b], o.omega[b]));
            if (wn != 0)
                o.R[b] = m3mul(angleAxisToR(wn * h, {o.omega[b][0] / wn, o.omega[b][1] / wn, o.omega[b][2] / wn}),
                               o.R[b]);
        }
        for (int64_t b = 0; b < o.nb; ++b) {
            o.theta[b] = rToAa(o.R[b]);
            rbUpdateInertia(o, b);
        }
        const std::vector<A3> xT = o.x;
        for (int it = 0; it < iterations; ++it)
            for (const auto &grp : o.groups) {
                const int64_t gs = static_cast<int64_t>(grp.size());
#pragma omp parallel for schedule(static)
                for (int64_t k = 0; k < gs; ++k) rbVertexDescent(o, grp[k], xT, thOld, h);
            }
        for (int64_t b = 0; b < o.nb; ++b) {  // computeIntegratedVelocities (rigidBodyGroup.cpp:115-128)
            for (int d = 0; d < 3; ++d) o.v[b][d] = (o.x[b][d] - xOld[b][d]) * invdt;
            o.omega[b] = rbIntegratedOmega(o, b, thOld[b], invdt);
        }
    }
}

template <class F>
int guard(F &&f)
{
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return 1;
    }
}

}  // namespace

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

int orc_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

void *orc_new(int64_t nv, const double *x, const double *x0, const double *m)
{
    auto *o = new Object;
    o->nv = nv;
    o->x.assign(x, x + 3 * nv);
    o->xInitial = o->x;
    o->x0.assign(x0, x0 + 3 * nv);
    o->v.assign(3 * nv, 0.0);
    o->m.assign(m, m + nv);
    return o;
}

void orc_free(void *h) { delete static_cast<Object *>(h); }

int orc_add_energy(void *h, int type, int64_t n, const int32_t *stencils, const double *raw)
{
    return guard([&] {
        addEnergy(*static_cast<Object *>(h), type, n, stencils, raw);
        static_cast<Object *>(h)->precompDone = false;
        static_cast<Object *>(h)->xpbdGroupsDone = false;
    });
}

int orc_prepare(void *h)
{
    return guard([&] {
        auto &o = *static_cast<Object *>(h);
        for (auto &e : o.energies)
            if (!e.prepared) prepareEnergy(o, e);
    });
}

int orc_num_prepared_params(int type) { return preparedParams(type); }

int orc_get_prepared(void *h, int energy, double *out)
{
    return guard([&] {
        auto &o = *static_cast<Object *>(h);
        if (energy < 0 || energy >= static_cast<int>(o.energies.size())) throw OracleError("bad energy index");
        const auto &e = o.energies[energy];
        if (!e.prepared) throw OracleError("energy not prepared");
        std::copy(e.prm.begin(), e.prm.end(), out);
    });
}

int orc_set_prepared(void *h, int energy, const double *in)
{
    return guard([&] {
        auto &o = *static_cast<Object *>(h);
        if (energy < 0 || energy >= static_cast<int>(o.energies.size())) throw OracleError("bad energy index");
        auto &e = o.energies[energy];
        e.prm.assign(in, in + e.n() * preparedParams(e.type));
        e.prepared = true;
    });
}

int orc_colors(void *h, int32_t *out, int32_t *ncolors)
{
    return guard([&] {
        auto &o = *static_cast<Object *>(h);
        if (!o.precompDone) precompute(o);
        if (out) std::copy(o.colors.begin(), o.colors.end(), out);
        if (ncolors) *ncolors = static_cast<int32_t>(o.groups.size());
    });
}

int orc_incidence_count(void *h, int64_t *count)
{
    return guard([&] {
        auto &o = *static_cast<Object *>(h);
        if (!o.precompDone) precompute(o);
        *count = static_cast<int64_t>(o.inc.size());
    });
}

int orc_incidence(void *h, int32_t *vertex, int32_t *energy, int32_t *stencil, int32_t *slot)
{
    return guard([&] {
        auto &o = *static_cast<Object *>(h);
        if (!o.precompDone) precompute(o);
        for (int64_t v = 0; v < o.nv; ++v)
            for (int64_t k = o.incPtr[v]; k < o.incPtr[v + 1]; ++k) {
                vertex[k] = static_cast<int32_t>(v);
                energy[k] = o.inc[k][0];
                stencil[k] = o.inc[k][1];
                slot[k] = o.inc[k][2];
            }
    });
}

int orc_step(void *h, double dt, int nsubsteps, int iterations, const double *gravity, int mode)
{
    return guard([&] {
        auto &o = *static_cast<Object *>(h);
        checkPrepared(o);
        if (!o.precompDone) precompute(o);
        const double hh = dt / nsubsteps;  // vbd.cpp:19
        for (int s = 0; s < nsubsteps; ++s) substep(o, hh, iterations, gravity, mode, nullptr, nullptr);
    });
}

int orc_substep_curve(void *h, double dt, int iterations, const double *gravity, int mode, const double *start_guess,
                      double *rho_per_iter)
{
    return guard([&] {
        auto &o = *static_cast<Object *>(h);
        checkPrepared(o);
        if (!o.precompDone) precompute(o);
        substep(o, dt, iterations, gravity, mode, start_guess, rho_per_iter);
    });
}

int orc_get_state(void *h, double *x, double *v)
{
    auto &o = *static_cast<Object *>(h);
    if (x) std::copy(o.x.begin(), o.x.end(), x);
    if (v) std::copy(o.v.begin(), o.v.end(), v);
    return 0;
}

int orc_set_state(void *h, const double *x, const double *v)
{
    auto &o = *static_cast<Object *>(h);
    if (x) std::copy(x, x + 3 * o.nv, o.x.begin());
    if (v) std::copy(v, v + 3 * o.nv, o.v.begin());
    return 0;
}

int orc_reset(void *h)
{
    // ParticleGroup::reset (particleGroup.cpp:97-103) + VBD::reset (vbd.cpp:197-201)
    auto &o = *static_cast<Object *>(h);
    o.x = o.xInitial;
    std::fill(o.v.begin(), o.v.end(), 0.0);
    o.precompDone = false;
    o.xpbdGroupsDone = false;
    return 0;
}

int orc_residual(void *h, const double *x, const double *xt, double dt, int mode, double *norm, double *r_out)
{
    return guard([&] {
        auto &o = *static_cast<Object *>(h);
        checkPrepared(o);
        residual(o, x, xt, dt, mode, norm, r_out);
    });
}

int orc_energy(void *h, const double *x, double *eout)
{
    return guard([&] {
        auto &o = *static_cast<Object *>(h);
        checkPrepared(o);
        double sum = 0.0;
        for (const auto &e : o.energies) {
            const int np = preparedParams(e.type);
            for (int64_t i = 0; i < e.n(); ++i) {
                double xs[12];
                gatherStencil(o, x, e, i, xs);
                sum += evalEnergy<double>(e.type, reinterpret_cast<const V3<double> *>(xs), &e.prm[i * np]);
            }
        }
        *eout = sum;
    });
}

int orc_element(int type, const double *xs, const double *prepared, int mode, double *energy, double *grad,
                double *diag_blocks, double *full_hess)
{
    return guard([&] {
        const int ns = stencilSize(type);
        if (!ns) throw OracleError("unknown energy type");
        if (energy) *energy = evalEnergy<double>(type, reinterpret_cast<const V3<double> *>(xs), prepared);
        for (int s = 0; s < ns; ++s) {
            double g[3], Hb[9];
            if (mode == ORC_MODE_AD)
                adVertex(type, xs, prepared, s, g, Hb);
            else
                cfVertex(type, xs, prepared, s, g, Hb);
            if (grad)
                for (int i = 0; i < 3; ++i) grad[3 * s + i] = g[i];
            if (diag_blocks)
                for (int i = 0; i < 9; ++i) diag_blocks[9 * s + i] = Hb[i];
        }
        if (full_hess && mode == ORC_MODE_AD) {
            switch (ns) {
            case 1: adElement<1, true>(type, xs, prepared, nullptr, nullptr, full_hess); break;
            case 2: adElement<2, true>(type, xs, prepared, nullptr, nullptr, full_hess); break;
            case 3: adElement<3, true>(type, xs, prepared, nullptr, nullptr, full_hess); break;
            case 4: adElement<4, true>(type, xs, prepared, nullptr, nullptr, full_hess); break;
            }
        }
    });
}

int orc_bending_stencils(int64_t nf, const int32_t *faces, int64_t nv, int32_t *out, int64_t *count)
{
    // TriangleMesh constructor (triangleMesh.h:16-50) + createCloth's bending wiring
    // (apps/solverCompare.cpp:441-452: edges with exactly two opposing vertices).
    return guard([&] {
        std::map<std::array<int32_t, 2>, std::vector<int32_t>> edges;
        for (int64_t f = 0; f < nf; ++f) {
            const int32_t *fc = faces + 3 * f;
            for (int i = 0; i < 3; ++i)
                if (fc[i] < 0 || fc[i] >= nv) throw OracleError("Triangle mesh face contains index to invalid vertex");
            if (fc[0] == fc[1] || fc[1] == fc[2] || fc[2] == fc[0])
                throw OracleError("Triangle mesh face contains repeated vertex index");
            for (int i = 0; i < 3; ++i) {
                const int32_t a = fc[i], b = fc[(i + 1) % 3], c = fc[(i + 2) % 3];
                edges[{std::min(a, b), std::max(a, b)}].push_back(c);
            }
        }
        int64_t k = 0;
        for (const auto &kv : edges) {
            if (kv.second.size() != 2) continue;
            if (out) {
                out[4 * k + 0] = kv.first[0];
                out[4 * k + 1] = kv.first[1];
                out[4 * k + 2] = kv.second[0];
                out[4 * k + 3] = kv.second[1];
            }
            ++k;
        }
        *count = k;
    });
}

int orc_coloring_valid(void *h, const int32_t *colors)
{
    auto &o = *static_cast<Object *>(h);
    for (const auto &e : o.energies)
        for (int64_t i = 0; i < e.n(); ++i)
            for (int a = 0; a < e.ssize; ++a)
                for (int b = a + 1; b < e.ssize; ++b) {
                    const int32_t va = e.st[i * e.ssize + a], vb = e.st[i * e.ssize + b];
                    if (va != vb && colors[va] == colors[vb]) return 0;
                }
    return 1;
}

/* ---- XPBD ---- */
int orc_num_constraints(int type) { return numConstraints(type); }

int orc_color_stencils(int64_t n_stencils, int stencil_size, const int32_t *flat, int32_t *colors, int32_t *ncolors)
{
    return guard([&] {
        if (stencil_size <= 0) throw OracleError("invalid stencil size");
        for (int64_t k = 0; k < n_stencils * stencil_size; ++k)
            if (flat[k] < 0) throw OracleError("negative vertex index");
        const auto col = colorStencils(flat, n_stencils, stencil_size);
        int32_t nc = 0;
        for (int64_t i = 0; i < n_stencils; ++i) {
            if (colors) colors[i] = col[i];
            nc = std::max(nc, col[i] + 1);
        }
        if (ncolors) *ncolors = nc;
    });
}

int orc_get_compliance(void *h, int energy, double *out)
{
    return guard([&] {
        auto &o = *static_cast<Object *>(h);
        if (energy < 0 || energy >= static_cast<int>(o.energies.size())) throw OracleError("bad energy index");
        const auto &e = o.energies[energy];
        if (!e.prepared) throw OracleError("energy not prepared");
        std::copy(e.cmp.begin(), e.cmp.end(), out);
    });
}

int orc_xpbd_step(void *h, double dt, int nsubsteps, const double *gravity, int parallel)
{
    return guard([&] {
        auto &o = *static_cast<Object *>(h);
        checkPrepared(o);
        if (!(dt > 0) || nsubsteps < 1) throw OracleError("invalid dt/substeps");
        xpbdStep(o, dt, nsubsteps, gravity, parallel != 0);
    });
}

int orc_constraints(int type, const double *xs, const double *prepared, double *c, double *grad)
{
    return guard([&] {
        switch (type) {
        case ORC_SPRING: adConstraints<2, 1>(type, xs, prepared, c, grad); return;
        case ORC_SPRING_ANCHOR: adConstraints<1, 1>(type, xs, prepared, c, grad); return;
        case ORC_MEMBRANE_STVK: adConstraints<3, 3>(type, xs, prepared, c, grad); return;
        case ORC_DISCRETE_BENDING: adConstraints<4, 1>(type, xs, prepared, c, grad); return;
        case ORC_STABLE_NEO_HOOKEAN: adConstraints<4, 2>(type, xs, prepared, c, grad); return;
        case ORC_SPRING_CONTINUUM:
        case ORC_SPRING_SQUARED_CONTINUUM: adConstraints<2, 1>(type, xs, prepared, c, grad); return;
        case ORC_MEMBRANE_BARAFF_WITKIN: adConstraints<3, 3>(type, xs, prepared, c, grad); return;
        case ORC_MEMBRANE_CHOI: adConstraints<3, 4>(type, xs, prepared, c, grad); return;
        case ORC_QUADRATIC_BENDING: adConstraints<4, 3>(type, xs, prepared, c, grad); return;
        case ORC_BARAFF_WITKIN_BENDING: adConstraints<4, 1>(type, xs, prepared, c, grad); return;
        case ORC_STVK:
        case ORC_STABLE_SQUARED_NEO_HOOKEAN: adConstraints<4, 2>(type, xs, prepared, c, grad); return;
        }
        throw OracleError("unknown energy type");
    });
}

int orc_xpbd_project(int type, const double *xs, const double *inv_masses, const double *prepared,
                     const double *compliance, double dt, int generic_spring, double *dx)
{
    return guard([&] {
        if (type == ORC_SPRING && generic_spring)
            projectGeneric<2, 1>(type, xs, inv_masses, prepared, compliance, dt, dx);
        else
            projectStencilDx(type, xs, inv_masses, prepared, compliance, dt, dx);
    });
}

/* ---- implicit Euler family ---- */
int orc_system_force(void *h, const double *x, const double *gravity, double *f)
{
    return guard([&] {
        auto &o = *static_cast<Object *>(h);
        checkPrepared(o);
        const auto v = systemForce(o, x, gravity);
        std::copy(v.begin(), v.end(), f);
    });
}

/* -Hessian as CSR (rows 3 nv, sorted unique columns). Call with col/val NULL to get nnz. */
int orc_stiffness_csr(void *h, const double *x, int64_t *nnz, int64_t *ptr, int64_t *col, double *val)
{
    return guard([&] {
        auto &o = *static_cast<Object *>(h);
        checkPrepared(o);
        const Csr K = stiffnessMatrix(o, x);
        *nnz = static_cast<int64_t>(K.val.size());
        if (ptr) std::copy(K.ptr.begin(), K.ptr.end(), ptr);
        if (col) std::copy(K.col.begin(), K.col.end(), col);
        if (val) std::copy(K.val.begin(), K.val.end(), val);
    });
}

/* Eigen-style Jacobi-preconditioned CG on A = alpha M + beta K(x) (K = -Hessian at x). */
int orc_cg_solve(void *h, const double *x, double alpha, double beta, const double *b, double tol, int max_iters,
                 double *out, int *iters, double *error)
{
    return guard([&] {
        auto &o = *static_cast<Object *>(h);
        checkPrepared(o);
        const Csr A = combine(o, alpha, stiffnessMatrix(o, x), beta);
        const std::vector<double> bb(b, b + A.n);
        const auto r = conjugateGradient(A, bb, tol, max_iters, iters, error);
        std::copy(r.begin(), r.end(), out);
    });
}

/* Generic CG on a CSR matrix given by the caller (checks of the CG restatement itself). */
int orc_cg_csr(int64_t n, const int64_t *ptr, const int64_t *col, const double *val, const double *b, double tol,
               int max_iters, double *out, int *iters, double *error)
{
    return guard([&] {
        Csr A;
        A.n = n;
        A.ptr.assign(ptr, ptr + n + 1);
        A.col.assign(col, col + ptr[n]);
        A.val.assign(val, val + ptr[n]);
        const std::vector<double> bb(b, b + n);
        const auto r = conjugateGradient(A, bb, tol, max_iters, iters, error);
        std::copy(r.begin(), r.end(), out);
    });
}

void *orc_implicit_new(void) { return new ImplicitState; }
void orc_implicit_free(void *s) { delete static_cast<ImplicitState *>(s); }

/* One step (nsubsteps substeps) of solver 0 BaraffWitkin, 1 NewtonDv, 2 NewtonDx, 3 NewtonRobust, 4 BDF2
   on the object's state. counters (may be NULL): cg iterations, linear solves, newton iterations, failed
   substeps of this step; current_dt (may be NULL): NewtonRobust's dt after the step. */
int orc_implicit_step(void *h, void *state, int solver, double dt, int nsubsteps, const double *gravity,
                      int64_t *counters, double *current_dt)
{
    return guard([&] {
        auto &o = *static_cast<Object *>(h);
        auto &st = *static_cast<ImplicitState *>(state);
        checkPrepared(o);
        if (solver < 0 || solver > 4) throw OracleError("unknown solver");
        if (!(dt > 0) || nsubsteps < 1) throw OracleError("invalid dt/substeps");
        st.cgIterations = st.linearSolves = st.newtonIterations = st.failedSteps = 0;
        for (int s = 0; s < nsubsteps; ++s) implicitSubstep(o, st, solver, dt, nsubsteps, gravity);
        if (counters) {
            counters[0] = st.cgIterations;
            counters[1] = st.linearSolves;
            counters[2] = st.newtonIterations;
            counters[3] = st.failedSteps;
        }
        if (current_dt) *current_dt = st.currentDt;
    });
}

/* ---- rigid-body VBD ---- */
void *orc_rb_new(int64_t nb, const double *x, const double *theta, const double *m, const double *inertia)
{
    auto *o = new RbObject;
    o->nb = nb;
    o->shape.nv = nb;
    for (int64_t b = 0; b < nb; ++b) {
        o->x.push_back({x[3 * b], x[3 * b + 1], x[3 * b + 2]});
        o->theta.push_back({theta[3 * b], theta[3 * b + 1], theta[3 * b + 2]});
        o->v.push_back({0, 0, 0});
        o->omega.push_back({0, 0, 0});
        o->m.push_back(m[b]);
        M3 Il;
        for (int k = 0; k < 9; ++k) Il[k] = inertia[9 * b + k];
        o->Iloc.push_back(Il);
        o->IlocInv.push_back(inverse3(Il));
        o->R.push_back(aaToR(o->theta.back()));
        o->I.push_back(Il);
        o->Iinv.push_back(Il);
        rbUpdateInertia(*o, b);
    }
    return o;
}
void orc_rb_free(void *h) { delete static_cast<RbObject *>(h); }

int orc_rb_add_energy(void *h, int type, int64_t n, const int32_t *stencils, const double *raw)
{
    return guard([&] {
        auto &o = *static_cast<RbObject *>(h);
        const int ss = rbStencilSize(type), np = rbRawParams(type);
        if (!ss) throw OracleError("unknown rigid-body energy type");
        Energy e;
        e.type = type;
        e.ssize = ss;
        e.st.assign(stencils, stencils + n * ss);
        for (auto b : e.st)
            if (b < 0 || b >= o.nb) throw OracleError("stencil index out of range");
        for (int64_t i = 0; i < n; ++i) {
            if (type == RB_SPRING && raw[i * np + 6] <= 0) throw OracleError("Zero/negative rest length invalid");
            if (type == RB_SPRING_SQUARED && raw[i * np + 6] < 0) throw OracleError("Negative rest length invalid");
        }
        o.shape.energies.push_back(e);
        o.params.emplace_back(raw, raw + n * np);
        o.precompDone = false;
    });
}

int orc_rb_step(void *h, double dt, int nsubsteps, int iterations, const double *gravity)
{
    return guard([&] {
        if (!(dt > 0) || nsubsteps < 1 || iterations < 0) throw OracleError("invalid step arguments");
        rbStep(*static_cast<RbObject *>(h), dt, nsubsteps, iterations, gravity);
    });
}

int orc_rb_state(void *h, double *x, double *theta, double *v, double *omega)
{
    return guard([&] {
        const auto &o = *static_cast<RbObject *>(h);
        for (int64_t b = 0; b < o.nb; ++b)
            for (int d = 0; d < 3; ++d) {
                if (x) x[3 * b + d] = o.x[b][d];
                if (theta) theta[3 * b + d] = o.theta[b][d];
                if (v) v[3 * b + d] = o.v[b][d];
                if (omega) omega[3 * b + d] = o.omega[b][d];
            }
    });
}

int orc_rb_set_state(void *h, const double *x, const double *theta, const double *v, const double *omega)
{
    return guard([&] {
        auto &o = *static_cast<RbObject *>(h);
        for (int64_t b = 0; b < o.nb; ++b) {
            for (int d = 0; d < 3; ++d) {
                if (x) o.x[b][d] = x[3 * b + d];
                if (theta) o.theta[b][d] = theta[3 * b + d];
                if (v) o.v[b][d] = v[3 * b + d];
                if (omega) o.omega[b][d] = omega[3 * b + d];
            }
            o.R[b] = aaToR(o.theta[b]);  // setPositions: rotation and inertia from theta (rigidBodyGroup.cpp:41-51)
            rbUpdateInertia(o, b);
        }
    });
}

int orc_rb_colors(void *h, int32_t *colors, int32_t *ncolors)
{
    return guard([&] {
        auto &o = *static_cast<RbObject *>(h);
        if (!o.precompDone) rbPrecompute(o);
        for (size_t c = 0; c < o.groups.size(); ++c)
            for (auto b : o.groups[c]) colors[b] = static_cast<int32_t>(c);
        if (ncolors) *ncolors = static_cast<int32_t>(o.groups.size());
    });
}

/* gradient (6 size) and Hessian (6 size)^2 of one rigid-body element at theta = 0; com: 3 per body,
   rot: 9 per body (row-major), raw: the addStencil arguments */
int orc_rb_element(int type, const double *com, const double *rot, const double *raw, double *grad, double *hess)
{
    return guard([&] {
        const int ss = rbStencilSize(type);
        if (!ss) throw OracleError("unknown rigid-body energy type");
        M3 R[2];
        for (int s = 0; s < ss; ++s)
            for (int k = 0; k < 9; ++k) R[s][k] = rot[9 * s + k];
        double p[8];
        rbPrepareParams(type, raw, p);
        if (ss == 1)
            rbElementAD<1>(type, com, R, p, grad, hess);
        else
            rbElementAD<2>(type, com, R, p, grad, hess);
    });
}

/* Eigen's AngleAxis conversions as restated here (for tests) */
int orc_rb_aa_to_r(const double *theta, double *R)
{
    const M3 r = aaToR({theta[0], theta[1], theta[2]});
    std::copy(r.begin(), r.end(), R);
    return 0;
}
int orc_rb_r_to_aa(const double *R, double *theta)
{
    M3 r;
    std::copy(R, R + 9, r.begin());
    const A3 t = rToAa(r);
    std::copy(t.begin(), t.end(), theta);
    return 0;
}

}  // extern "C"
