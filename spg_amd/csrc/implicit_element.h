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

// First-order dual number: value and one directional derivative. It lives in its own namespace so
// that its sqrt/atan2 are found only by argument-dependent lookup and never hide the float/double
// functions the other closed forms of this namespace call.
namespace ad {
struct Dual1 {
    double v, d;
    SPG_HD Dual1() : v(0), d(0) {}
    SPG_HD Dual1(double a) : v(a), d(0) {}  // NOLINT: constants inside expressions
    SPG_HD Dual1(double a, double b) : v(a), d(b) {}
};
SPG_HD Dual1 operator+(Dual1 a, Dual1 b) { return {a.v + b.v, a.d + b.d}; }
SPG_HD Dual1 operator-(Dual1 a, Dual1 b) { return {a.v - b.v, a.d - b.d}; }
SPG_HD Dual1 operator-(Dual1 a) { return {-a.v, -a.d}; }
SPG_HD Dual1 operator*(Dual1 a, Dual1 b) { return {a.v * b.v, a.v * b.d + a.d * b.v}; }
SPG_HD Dual1 operator/(Dual1 a, Dual1 b)
{
    const double q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
SPG_HD Dual1 sqrt(Dual1 a)
{
    const double s = ::sqrt(a.v);
    return {s, a.d / (2.0 * s)};
}
SPG_HD Dual1 atan2(Dual1 y, Dual1 x)
{
    const double r2 = x.v * x.v + y.v * y.v;
    return {::atan2(y.v, x.v), (x.v * y.d - y.v * x.d) / r2};
}
}  // namespace ad
using ad::Dual1;

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
    for (int i = 0; i < 3; ++i) s[i] = T(p[4 + 3 * i]) * C[0] + T(p[5 + 3 * i]) * C[1] + T(p[6 + 3 * i]) * C[2];
    // F S = [f0 s0 + f1 s2, f0 s2 + f1 s1]
    const V3<T> c0 = s[0] * f0 + s[2] * f1, c1 = s[2] * f0 + s[1] * f1;
    const double ga[3][2] = {{-(p[0] + p[2]), -(p[1] + p[3])}, {p[0], p[1]}, {p[2], p[3]}};
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

// Type ids: 1 spring, 2 membrane, 3 bending, 4 SNH, 5 anchor.
SPG_HD constexpr int fullStencilSize(int type) { return type == 5 ? 1 : type == 1 ? 2 : type == 2 ? 3 : 4; }

// The energies of stencil size NS (selected at compile time, so no path indexes past x[NS - 1]).
template <int NS, class T>
SPG_HD void fullGrad(int type, const V3<T> *x, const double *p, V3<T> *g)
{
    if constexpr (NS == 1) {
        fullGradAnchor(x, p, g);
    } else if constexpr (NS == 2) {
        fullGradSpring(x, p, g);
    } else if constexpr (NS == 3) {
        fullGradStvk(x, p, g);
    } else {
        if (type == 3)
            fullGradBending(x, p, g);
        else
            fullGradSnh(x, p, g);
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
