// Closed-form XPBD projections of the five spg energies: one StencilBlockEnergy::projectPosition
// (spg/include/spg/sim/energy/energy.h:117-153) per call, returned as the displacement of each stencil
// vertex. With W the inverse masses, C the stencil's constraints and J = dC/dx:
//   A = J W J^T + alpha~ + 1e-10 I (alpha~ = effectiveCompliance / dt^2), dlambda = -A^-1 C, dx = W J^T dlambda.
// The reference differentiates the constraints with TinyAD; here every J is written out, so only
// dot products of a few vectors are formed and A is assembled from them.
//
// Templated on the scalar: float in the projection kernels, double for their ill-conditioned fallback
// and for the host check hook. Host and device.
#pragma once

#include <type_traits>

#include "element.h"

namespace spgvbd {

// SpringEnergy::projectPosition (springEnergy.cpp:48-78), the spring's own closed form: no 1e-10 term,
// no correction when w0 + w1 < 1e-6 or |x1 - x0| < 1e-6.
template <class R>
SPG_HD void xpbdSpring(V3<R> x0, V3<R> x1, R w0, R w1, R L0, R alphaTilde, V3<R> &d0, V3<R> &d1)
{
    const R eps = R(1.0e-6);
    d0 = v3<R>(R(0), R(0), R(0));
    d1 = d0;
    const R wSum = w0 + w1;
    if (wSum < eps) return;
    const V3<R> n = x1 - x0;
    const R d = sqrt(dot(n, n));
    if (d < eps) return;
    const R s = (d - L0) / (d * (wSum + alphaTilde));  // corr = (n / d) (d - L0) / (wSum + alpha~)
    d0 = (w0 * s) * n;
    d1 = (-w1 * s) * n;
}

// SpringAnchorEnergy (springAnchorEnergy.cpp:8-20): C = |x - a|, grad C = (x - a)/C; the reference
// zeroes value and gradient when C is exactly 0, which leaves x unchanged.
template <class R>
SPG_HD V3<R> xpbdAnchor(V3<R> x, R w, V3<R> a, R alphaTilde)
{
    const V3<R> r = x - a;
    const R C = sqrt(dot(r, r));
    if (C == R(0)) return v3<R>(R(0), R(0), R(0));
    const V3<R> g = (R(1) / C) * r;
    const R A = w * dot(g, g) + alphaTilde + R(1e-10);
    return (-w * C / A) * g;
}

// MembraneStvkEnergy (membraneStvkEnergy.cpp:9-28): F = [x1-x0, x2-x0] Dinv with columns f0, f1,
// C = (E00, E11, 2 E01) = ((f0.f0 - 1)/2, (f1.f1 - 1)/2, f0.f1). With a = (-(D00+D10), D00, D10) and
// b = (-(D01+D11), D01, D11) the slot weights of f0 and f1:
//   grad_k C0 = a_k f0, grad_k C1 = b_k f1, grad_k C2 = a_k f1 + b_k f0,
// so J W J^T needs only Saa = sum w a^2, Sbb, Sab and the dot products of f0, f1.
// Alpha~ (symmetric, already divided by dt^2): c00 c01 c02 c11 c12 c22.
// Returns det(A) / (A00 A11 A22), 1 for a diagonal A and small when A is close to singular.
template <class R>
SPG_HD R xpbdStvk(V3<R> x0, V3<R> x1, V3<R> x2, R w0, R w1, R w2, R D00, R D01, R D10, R D11, R c00, R c01,
                  R c02, R c11, R c12, R c22, V3<R> &o0, V3<R> &o1, V3<R> &o2)
{
    const V3<R> u = x1 - x0, v = x2 - x0;
    const V3<R> f0 = D00 * u + D10 * v;
    const V3<R> f1 = D01 * u + D11 * v;
    const R a0 = -(D00 + D10), a1 = D00, a2 = D10;
    const R b0 = -(D01 + D11), b1 = D01, b2 = D11;
    const R Saa = w0 * a0 * a0 + w1 * a1 * a1 + w2 * a2 * a2;
    const R Sbb = w0 * b0 * b0 + w1 * b1 * b1 + w2 * b2 * b2;
    const R Sab = w0 * a0 * b0 + w1 * a1 * b1 + w2 * a2 * b2;
    const R p00 = dot(f0, f0), p11 = dot(f1, f1), p01 = dot(f0, f1);
    const R C0 = R(0.5) * (p00 - R(1)), C1 = R(0.5) * (p11 - R(1)), C2 = p01;
    const R eps = R(1e-10);
    const R A00 = Saa * p00 + c00 + eps;
    const R A11 = Sbb * p11 + c11 + eps;
    const R A22 = Saa * p11 + R(2) * Sab * p01 + Sbb * p00 + c22 + eps;
    const R A01 = Sab * p01 + c01;
    const R A02 = Saa * p01 + Sab * p00 + c02;
    const R A12 = Sab * p11 + Sbb * p01 + c12;
    const R k00 = A11 * A22 - A12 * A12, k01 = A02 * A12 - A01 * A22, k02 = A01 * A12 - A02 * A11;
    const R k11 = A00 * A22 - A02 * A02, k12 = A01 * A02 - A00 * A12, k22 = A00 * A11 - A01 * A01;
    const R det = A00 * k00 + A01 * k01 + A02 * k02;
    const R id = R(-1) / det;
    const R l0 = (k00 * C0 + k01 * C1 + k02 * C2) * id;
    const R l1 = (k01 * C0 + k11 * C1 + k12 * C2) * id;
    const R l2 = (k02 * C0 + k12 * C1 + k22 * C2) * id;
    // dx_k = w_k [ (a_k l0 + b_k l2) f0 + (b_k l1 + a_k l2) f1 ]
    o0 = w0 * ((a0 * l0 + b0 * l2) * f0 + (b0 * l1 + a0 * l2) * f1);
    o1 = w1 * ((a1 * l0 + b1 * l2) * f0 + (b1 * l1 + a1 * l2) * f1);
    o2 = w2 * ((a2 * l0 + b2 * l2) * f0 + (b2 * l1 + a2 * l2) * f1);
    return det / (A00 * A11 * A22);
}

// DiscreteBendingEnergy (discreteBendingEnergy.cpp:9-31): C = theta - theta0 with hinge x0-x1 and
// wings x2, x3. With e = x1 - x0, p = x2 - x1, q = x3 - x1, a = e x p, b = e x q, L = |e|:
//   theta = atan2(-(a x b).e/L, -a.b)
//   grad_0 = -(p.e^)/|a|^2 a + (q.e^)/|b|^2 b,  grad_1 = (p.e^ + L)/|a|^2 a - (q.e^ + L)/|b|^2 b,
//   grad_2 = -L/|a|^2 a,                         grad_3 = L/|b|^2 b.
template <class R>
SPG_HD void xpbdBending(V3<R> x0, V3<R> x1, V3<R> x2, V3<R> x3, R w0, R w1, R w2, R w3, R theta0, R alphaTilde,
                        V3<R> &o0, V3<R> &o1, V3<R> &o2, V3<R> &o3)
{
    const V3<R> e = x1 - x0, p = x2 - x1, q = x3 - x1;
    const V3<R> a = cross(e, p), b = cross(e, q);
    const R L = sqrt(dot(e, e));
    const R iL = R(1) / L;
    const R pe = dot(p, e) * iL, qe = dot(q, e) * iL;
    const R ia2 = R(1) / dot(a, a), ib2 = R(1) / dot(b, b);
    const V3<R> g0 = (-pe * ia2) * a + (qe * ib2) * b;
    const V3<R> g1 = ((pe + L) * ia2) * a + (-(qe + L) * ib2) * b;
    const V3<R> g2 = (-L * ia2) * a;
    const V3<R> g3 = (L * ib2) * b;
    const R theta = atan2(-dot(cross(a, b), e) * iL, -dot(a, b));
    const R C = theta - theta0;
    const R A = w0 * dot(g0, g0) + w1 * dot(g1, g1) + w2 * dot(g2, g2) + w3 * dot(g3, g3) + alphaTilde + R(1e-10);
    const R dl = -C / A;
    o0 = (w0 * dl) * g0;
    o1 = (w1 * dl) * g1;
    o2 = (w2 * dl) * g2;
    o3 = (w3 * dl) * g3;
}

// StableNeoHookeanEnergy (stableNeoHookeanEnergy.cpp:9-31): F = [x1-x0, x2-x0, x3-x0] Dinv with columns
// f0, f1, f2, C = (det F - alpha, |F|_F). dC0/dF = [f1 x f2, f2 x f0, f0 x f1], dC1/dF = F / |F|_F, and
// grad_{m+1} C = sum_c (dC/dF)_c Dinv(m, c), grad_0 = -(grad_1 + grad_2 + grad_3).
// D: Dinv row-major (9); alpha~ (symmetric 2x2, divided by dt^2): c00 c01 c11.
// Returns det(A) / (A00 A11).
template <class R>
SPG_HD R xpbdSnh(V3<R> x0, V3<R> x1, V3<R> x2, V3<R> x3, R w0, R w1, R w2, R w3, const R *D, R c00, R c01, R c11,
                 R hydro, V3<R> &o0, V3<R> &o1, V3<R> &o2, V3<R> &o3)
{
    const V3<R> u = x1 - x0, v = x2 - x0, w = x3 - x0;
    const V3<R> f0 = D[0] * u + D[3] * v + D[6] * w;
    const V3<R> f1 = D[1] * u + D[4] * v + D[7] * w;
    const V3<R> f2 = D[2] * u + D[5] * v + D[8] * w;
    const V3<R> P0 = cross(f1, f2), P1 = cross(f2, f0), P2 = cross(f0, f1);
    const R J = dot(f0, P0);
    const R nF = sqrt(dot(f0, f0) + dot(f1, f1) + dot(f2, f2));
    const R inF = R(1) / nF;
    const R C0 = J - hydro, C1 = nF;
    const V3<R> h1 = D[0] * P0 + D[1] * P1 + D[2] * P2;
    const V3<R> h2 = D[3] * P0 + D[4] * P1 + D[5] * P2;
    const V3<R> h3 = D[6] * P0 + D[7] * P1 + D[8] * P2;
    const V3<R> h0 = -(h1 + h2 + h3);
    const V3<R> k1 = inF * (D[0] * f0 + D[1] * f1 + D[2] * f2);
    const V3<R> k2 = inF * (D[3] * f0 + D[4] * f1 + D[5] * f2);
    const V3<R> k3 = inF * (D[6] * f0 + D[7] * f1 + D[8] * f2);
    const V3<R> k0 = -(k1 + k2 + k3);
    const R eps = R(1e-10);
    const R A00 = w0 * dot(h0, h0) + w1 * dot(h1, h1) + w2 * dot(h2, h2) + w3 * dot(h3, h3) + c00 + eps;
    const R A01 = w0 * dot(h0, k0) + w1 * dot(h1, k1) + w2 * dot(h2, k2) + w3 * dot(h3, k3) + c01;
    const R A11 = w0 * dot(k0, k0) + w1 * dot(k1, k1) + w2 * dot(k2, k2) + w3 * dot(k3, k3) + c11 + eps;
    const R det = A00 * A11 - A01 * A01;
    const R id = R(-1) / det;
    const R l0 = (A11 * C0 - A01 * C1) * id;
    const R l1 = (A00 * C1 - A01 * C0) * id;
    o0 = w0 * (l0 * h0 + l1 * k0);
    o1 = w1 * (l0 * h1 + l1 * k1);
    o2 = w2 * (l0 * h2 + l1 * k2);
    o3 = w3 * (l0 * h3 + l1 * k3);
    return det / (A00 * A11);
}

// ---------------------------------------------------------------------------------------------
// FP32 projections with the FP64 fallback of the kernels: a membrane or SNH projection whose
// det(A) / prod(diag A) falls below kXpbdCondMin is recomputed in FP64 from the same FP32 inputs.
// The FP32 rounding of J then reaches dlambda amplified by about prod(diag A) / det(A): below 2e-2 that
// exceeds 5e-6 of the displacement. (Near F = I the two SNH constraint gradients are almost parallel
// and only alpha~ keeps A regular, so stiff tetrahedra at rest take this path.) They return true when
// the fallback ran.
// ---------------------------------------------------------------------------------------------
constexpr float kXpbdCondMin = 2e-2f;

SPG_HD V3<double> toDouble(V3<float> a) { return {double(a.x), double(a.y), double(a.z)}; }
SPG_HD V3<float> toFloat(V3<double> a) { return {float(a.x), float(a.y), float(a.z)}; }

SPG_HD bool xpbdStvkF(V3<float> x0, V3<float> x1, V3<float> x2, float w0, float w1, float w2, float D00, float D01,
                      float D10, float D11, float c00, float c01, float c02, float c11, float c12, float c22,
                      V3<float> &o0, V3<float> &o1, V3<float> &o2)
{
    const float cond =
        xpbdStvk<float>(x0, x1, x2, w0, w1, w2, D00, D01, D10, D11, c00, c01, c02, c11, c12, c22, o0, o1, o2);
    if (cond > kXpbdCondMin) return false;
    V3<double> e0, e1, e2;
    xpbdStvk<double>(toDouble(x0), toDouble(x1), toDouble(x2), w0, w1, w2, D00, D01, D10, D11, c00, c01, c02, c11, c12,
                     c22, e0, e1, e2);
    o0 = toFloat(e0);
    o1 = toFloat(e1);
    o2 = toFloat(e2);
    return true;
}

SPG_HD bool xpbdSnhF(V3<float> x0, V3<float> x1, V3<float> x2, V3<float> x3, float w0, float w1, float w2, float w3,
                     const float *D, float c00, float c01, float c11, float hydro, V3<float> &o0, V3<float> &o1,
                     V3<float> &o2, V3<float> &o3)
{
    const float cond = xpbdSnh<float>(x0, x1, x2, x3, w0, w1, w2, w3, D, c00, c01, c11, hydro, o0, o1, o2, o3);
    if (cond > kXpbdCondMin) return false;
    double Dd[9];
    for (int k = 0; k < 9; ++k) Dd[k] = D[k];
    V3<double> e0, e1, e2, e3;
    xpbdSnh<double>(toDouble(x0), toDouble(x1), toDouble(x2), toDouble(x3), w0, w1, w2, w3, Dd, c00, c01, c11, hydro, e0,
                    e1, e2, e3);
    o0 = toFloat(e0);
    o1 = toFloat(e1);
    o2 = toFloat(e2);
    o3 = toFloat(e3);
    return true;
}

// One projection from FP64 inputs in the C-ABI parameter layout of spg_xpbd_energy (x: size*3, w: size,
// p: the stencil's params), evaluated like the projection kernels (FP32 = true: inputs rounded to float,
// FP32 arithmetic with the FP64 fallback) or in FP64. Writes dx (size*3); returns whether the FP64
// fallback ran. Type ids: 1 spring, 2 membrane, 3 bending, 4 SNH, 5 anchor.
template <class R>
SPG_HD V3<R> xpbdLoad3(const double *p)
{
    return {R(p[0]), R(p[1]), R(p[2])};
}

template <class R>
SPG_HD void xpbdStore3(double *o, V3<R> v)
{
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
}

// ---- The further energies: a generic coupled projection from closed-form constraint gradients ----
// Parameters (spg_xpbd_energy layout): the geometry, then the compliance diagonal:
//   6, 7 continuum springs (L0, c); 8 Baraff-Witkin membrane (Dinv 2x2, c0..c2); 9 Choi membrane (Dinv 2x2,
//   c0..c3); 10 quadratic bending (c0..c3 of the Laplacian, compliance); 11 Baraff-Witkin bending (theta0, c);
//   12 StVK tetrahedron, 13 stable squared Neo-Hookean (Dinv 3x3, c_lambda, c_mu).
SPG_HD constexpr int xpbdFurtherConstraints(int type)
{
    return type == 8 ? 3 : type == 9 ? 4 : type == 10 ? 3 : (type == 12 || type == 13) ? 2 : 1;
}
SPG_HD constexpr int xpbdFurtherStencil(int type)
{
    return (type == 6 || type == 7) ? 2 : (type == 8 || type == 9) ? 3 : 4;
}

// Constraint values C[c] and gradients g[c][s] (w.r.t. stencil slot s) of one further-energy stencil.
template <class R>
SPG_HD void xpbdFurtherGradients(int type, const V3<R> *x, const R *p, R *C, V3<R> (*g)[4])
{
    const V3<R> z = v3<R>(R(0), R(0), R(0));
    for (int c = 0; c < 4; ++c)
        for (int s = 0; s < 4; ++s) g[c][s] = z;
    if (type == 6 || type == 7) {  // springContinuumEnergy.cpp:9-18, springSquaredContinuumEnergy.cpp:9-19
        const V3<R> d = x[0] - x[1];
        const R L0 = p[0];
        if (type == 6) {
            const R n = sqrt(dot(d, d));
            C[0] = n / L0 - R(1);
            g[0][0] = (R(1) / (L0 * n)) * d;
        } else {
            C[0] = dot(d, d) / (L0 * L0) - R(1);
            g[0][0] = (R(2) / (L0 * L0)) * d;
        }
        g[0][1] = -g[0][0];
    } else if (type == 8 || type == 9) {  // membraneBaraffWitkinEnergy.cpp:9-26, membraneChoiEnergy.cpp:9-28
        const V3<R> u = x[1] - x[0], v = x[2] - x[0];
        const V3<R> f0 = p[0] * u + p[2] * v, f1 = p[1] * u + p[3] * v;
        const R a[3] = {-(p[0] + p[2]), p[0], p[2]}, b[3] = {-(p[1] + p[3]), p[1], p[3]};
        const R n0 = sqrt(dot(f0, f0)), n1 = sqrt(dot(f1, f1));
        C[0] = n0 - R(1);
        C[1] = n1 - R(1);
        const V3<R> h0 = (R(1) / n0) * f0, h1 = (R(1) / n1) * f1;
        for (int s = 0; s < 3; ++s) {
            g[0][s] = a[s] * h0;
            g[1][s] = b[s] * h1;
        }
        if (type == 8) {
            C[2] = dot(f0, f1);
            for (int s = 0; s < 3; ++s) g[2][s] = a[s] * f1 + b[s] * f0;
        } else {
            const R r = R(0.70710678118654752440);
            const V3<R> gp = r * (f0 + f1), gm = r * (f0 - f1);
            const R np = sqrt(dot(gp, gp)), nm = sqrt(dot(gm, gm));
            C[2] = np - R(1);
            C[3] = nm - R(1);
            const V3<R> hp = (R(1) / np) * gp, hm = (R(1) / nm) * gm;
            for (int s = 0; s < 3; ++s) {
                g[2][s] = (r * (a[s] + b[s])) * hp;
                g[3][s] = (r * (a[s] - b[s])) * hm;
            }
        }
    } else if (type == 10) {  // quadraticBendingEnergy.cpp:9-30: C = L x, dC_d/dx_s = c_s e_d
        const R cx = p[0] * x[0].x + p[1] * x[1].x + p[2] * x[2].x + p[3] * x[3].x;
        const R cy = p[0] * x[0].y + p[1] * x[1].y + p[2] * x[2].y + p[3] * x[3].y;
        const R cz = p[0] * x[0].z + p[1] * x[1].z + p[2] * x[2].z + p[3] * x[3].z;
        C[0] = cx;
        C[1] = cy;
        C[2] = cz;
        for (int s = 0; s < 4; ++s) {
            g[0][s] = v3<R>(p[s], R(0), R(0));
            g[1][s] = v3<R>(R(0), p[s], R(0));
            g[2][s] = v3<R>(R(0), R(0), p[s]);
        }
    } else if (type == 11) {  // baraffWitkinBendingEnergy.cpp:9-36 with its normalisations held constant
        const V3<R> e0 = x[1] - x[0], e3 = x[2] - x[1], e4 = x[3] - x[1];
        const V3<R> n1 = cross(e0, e3), n2 = -cross(e0, e4);
        const R s0 = R(1) / sqrt(dot(e0, e0));
        const V3<R> m = cross(n1, n2);
        const R X = dot(n1, n2), Y = s0 * dot(m, e0);
        C[0] = atan2(Y, X) - p[0];
        // grad(v . n1) and grad(v . n2) per slot, as for the energy's gradient
        auto dN1 = [&](V3<R> vv, V3<R> *o) {
            const V3<R> q = cross(e3, vv), t = cross(vv, e0);
            o[0] = -q;
            o[1] = q - t;
            o[2] = t;
            o[3] = z;
        };
        auto dN2 = [&](V3<R> vv, V3<R> *o) {
            const V3<R> q = cross(e4, vv), t = cross(vv, e0);
            o[0] = q;
            o[1] = t - q;
            o[2] = z;
            o[3] = -t;
        };
        V3<R> dx1[4], dx2[4], dy1[4], dy2[4];
        dN1(n2, dx1);
        dN2(n1, dx2);
        dN1(cross(n2, e0), dy1);
        dN2(cross(e0, n1), dy2);
        const R w = R(1) / (X * X + Y * Y);
        for (int s = 0; s < 4; ++s) {
            V3<R> gy = dy1[s] + dy2[s];
            if (s == 0) gy = gy - m;
            if (s == 1) gy = gy + m;
            g[0][s] = w * (X * (s0 * gy) - Y * (dx1[s] + dx2[s]));
        }
    } else {  // 12 stvkEnergy.cpp:9-39, 13 stableSquaredNeoHookeanEnergy.cpp:9-31: F = [x1-x0 x2-x0 x3-x0] Dinv
        const V3<R> u = x[1] - x[0], v = x[2] - x[0], ww = x[3] - x[0];
        V3<R> f[3];
        for (int c = 0; c < 3; ++c) f[c] = p[c] * u + p[3 + c] * v + p[6 + c] * ww;
        V3<R> P[2][3];
        if (type == 12) {
            R E[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) E[i][j] = R(0.5) * (dot(f[i], f[j]) - R(i == j ? 1 : 0));
            R ee = R(0);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) ee += E[i][j] * E[j][i];
            C[0] = E[0][0] + E[1][1] + E[2][2];
            C[1] = sqrt(R(2) * ee);
            const R k = C[1] != R(0) ? R(2) / C[1] : R(0);  // the reference zeroes a vanishing deviatoric term
            for (int c = 0; c < 3; ++c) {
                P[0][c] = f[c];
                P[1][c] = k * (E[0][c] * f[0] + E[1][c] * f[1] + E[2][c] * f[2]);
            }
        } else {
            const V3<R> k0 = cross(f[1], f[2]), k1 = cross(f[2], f[0]), k2 = cross(f[0], f[1]);
            C[0] = dot(f[0], k0) - R(1);
            C[1] = dot(f[0], f[0]) + dot(f[1], f[1]) + dot(f[2], f[2]) - R(3);
            P[0][0] = k0;
            P[0][1] = k1;
            P[0][2] = k2;
            for (int c = 0; c < 3; ++c) P[1][c] = R(2) * f[c];
        }
        for (int q = 0; q < 2; ++q) {
            g[q][1] = p[0] * P[q][0] + p[1] * P[q][1] + p[2] * P[q][2];
            g[q][2] = p[3] * P[q][0] + p[4] * P[q][1] + p[5] * P[q][2];
            g[q][3] = p[6] * P[q][0] + p[7] * P[q][1] + p[8] * P[q][2];
            g[q][0] = -(g[q][1] + g[q][2] + g[q][3]);
        }
    }
}

// A = sum_s w_s g_a . g_b + alpha~ + 1e-10 I by Cholesky, dlambda = -A^-1 C, dx_s = w_s sum_a g_a dlambda_a.
// Returns det(A) / prod(diag A) (<= 0 when rounding made A indefinite).
template <class R>
SPG_HD R xpbdFurtherSolve(int nc, int ns, const R *C, const V3<R> (*g)[4], const R *w, const R *alphaTilde,
                          V3<R> *dx)
{
    R A[4][4], L[4][4];
    for (int a = 0; a < nc; ++a)
        for (int b = 0; b < nc; ++b) {
            R acc = R(0);
            for (int s = 0; s < ns; ++s) acc += w[s] * dot(g[a][s], g[b][s]);
            A[a][b] = acc + (a == b ? alphaTilde[a] + R(1e-10) : R(0));
        }
    R ratio = R(1);
    for (int j = 0; j < nc; ++j) {
        R d = A[j][j];
        for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
        ratio *= d / A[j][j];
        if (!(d > R(0))) {
            for (int s = 0; s < ns; ++s) dx[s] = v3<R>(R(0), R(0), R(0));
            return R(0);
        }
        L[j][j] = sqrt(d);
        for (int i = j + 1; i < nc; ++i) {
            R t = A[i][j];
            for (int k = 0; k < j; ++k) t -= L[i][k] * L[j][k];
            L[i][j] = t / L[j][j];
        }
    }
    R y[4], dl[4];
    for (int i = 0; i < nc; ++i) {
        R t = -C[i];
        for (int k = 0; k < i; ++k) t -= L[i][k] * y[k];
        y[i] = t / L[i][i];
    }
    for (int i = nc - 1; i >= 0; --i) {
        R t = y[i];
        for (int k = i + 1; k < nc; ++k) t -= L[k][i] * dl[k];
        dl[i] = t / L[i][i];
    }
    for (int s = 0; s < ns; ++s) {
        V3<R> acc = v3<R>(R(0), R(0), R(0));
        for (int a = 0; a < nc; ++a) acc = acc + dl[a] * g[a][s];
        dx[s] = w[s] * acc;
    }
    return ratio;
}

template <class R>
SPG_HD R xpbdFurther(int type, const double *x, const double *w, const double *p, double invdt2, double *dx)
{
    const int nc = xpbdFurtherConstraints(type), ns = xpbdFurtherStencil(type);
    const int ng = type == 12 || type == 13 ? 9 : (type == 8 || type == 9 || type == 10) ? 4 : 1;
    R pr[9], at[4], wr[4], C[4] = {R(0), R(0), R(0), R(0)};
    V3<R> xr[4], g[4][4], d[4];
    for (int k = 0; k < ng; ++k) pr[k] = R(p[k]);
    for (int a = 0; a < nc; ++a) at[a] = R(p[ng + (type == 10 ? 0 : a)]) * R(invdt2);  // quadratic bending: c I
    for (int s = 0; s < ns; ++s) {
        xr[s] = v3<R>(R(x[3 * s]), R(x[3 * s + 1]), R(x[3 * s + 2]));
        wr[s] = R(w[s]);
    }
    xpbdFurtherGradients<R>(type, xr, pr, C, g);
    const R ratio = xpbdFurtherSolve<R>(nc, ns, C, g, wr, at, d);
    for (int s = 0; s < ns; ++s) {
        dx[3 * s] = double(d[s].x);
        dx[3 * s + 1] = double(d[s].y);
        dx[3 * s + 2] = double(d[s].z);
    }
    return ratio;
}

// FP32 with the FP64 recomputation when det(A) / prod(diag A) < kXpbdCondMin; returns whether it fell back.
template <bool FP32>
SPG_HD bool xpbdFurtherRecord(int type, const double *x, const double *w, const double *p, double invdt2, double *dx)
{
    if constexpr (FP32) {
        if (xpbdFurther<float>(type, x, w, p, invdt2, dx) > kXpbdCondMin) return false;
    }
    (void)xpbdFurther<double>(type, x, w, p, invdt2, dx);
    return FP32;
}

template <bool FP32>
SPG_HD bool xpbdProjectRecord(int type, const double *x, const double *w, const double *p, double invdt2, double *dx)
{
    if (type >= 6) return xpbdFurtherRecord<FP32>(type, x, w, p, invdt2, dx);
    using R = typename std::conditional<FP32, float, double>::type;
    const R s = R(invdt2);
    bool fell = false;
    if (type == 1) {
        V3<R> d0, d1;
        xpbdSpring<R>(xpbdLoad3<R>(x), xpbdLoad3<R>(x + 3), R(w[0]), R(w[1]), R(p[0]), R(p[1]) * s, d0, d1);
        xpbdStore3(dx, d0);
        xpbdStore3(dx + 3, d1);
    } else if (type == 5) {
        xpbdStore3(dx, xpbdAnchor<R>(xpbdLoad3<R>(x), R(w[0]), xpbdLoad3<R>(p), R(p[3]) * s));
    } else if (type == 2) {
        V3<R> o0, o1, o2;
        const double *c = p + 4;
        const R cc[6] = {R(c[0]) * s, R(c[1]) * s, R(c[2]) * s, R(c[4]) * s, R(c[5]) * s, R(c[8]) * s};
        if constexpr (FP32)
            fell = xpbdStvkF(xpbdLoad3<R>(x), xpbdLoad3<R>(x + 3), xpbdLoad3<R>(x + 6), R(w[0]), R(w[1]), R(w[2]),
                             R(p[0]), R(p[1]), R(p[2]), R(p[3]), cc[0], cc[1], cc[2], cc[3], cc[4], cc[5], o0, o1, o2);
        else
            xpbdStvk<R>(xpbdLoad3<R>(x), xpbdLoad3<R>(x + 3), xpbdLoad3<R>(x + 6), R(w[0]), R(w[1]), R(w[2]), R(p[0]),
                        R(p[1]), R(p[2]), R(p[3]), cc[0], cc[1], cc[2], cc[3], cc[4], cc[5], o0, o1, o2);
        xpbdStore3(dx, o0);
        xpbdStore3(dx + 3, o1);
        xpbdStore3(dx + 6, o2);
    } else if (type == 3) {
        V3<R> o0, o1, o2, o3;
        xpbdBending<R>(xpbdLoad3<R>(x), xpbdLoad3<R>(x + 3), xpbdLoad3<R>(x + 6), xpbdLoad3<R>(x + 9), R(w[0]), R(w[1]),
                       R(w[2]), R(w[3]), R(p[0]), R(p[1]) * s, o0, o1, o2, o3);
        xpbdStore3(dx, o0);
        xpbdStore3(dx + 3, o1);
        xpbdStore3(dx + 6, o2);
        xpbdStore3(dx + 9, o3);
    } else {
        R D[9];
        for (int k = 0; k < 9; ++k) D[k] = R(p[k]);
        V3<R> o0, o1, o2, o3;
        if constexpr (FP32)
            fell = xpbdSnhF(xpbdLoad3<R>(x), xpbdLoad3<R>(x + 3), xpbdLoad3<R>(x + 6), xpbdLoad3<R>(x + 9), R(w[0]),
                            R(w[1]), R(w[2]), R(w[3]), D, R(p[9]) * s, R(p[10]) * s, R(p[12]) * s, R(p[13]), o0, o1,
                            o2, o3);
        else
            xpbdSnh<R>(xpbdLoad3<R>(x), xpbdLoad3<R>(x + 3), xpbdLoad3<R>(x + 6), xpbdLoad3<R>(x + 9), R(w[0]), R(w[1]),
                       R(w[2]), R(w[3]), D, R(p[9]) * s, R(p[10]) * s, R(p[12]) * s, R(p[13]), o0, o1, o2, o3);
        xpbdStore3(dx, o0);
        xpbdStore3(dx + 3, o1);
        xpbdStore3(dx + 6, o2);
        xpbdStore3(dx + 9, o3);
    }
    return fell;
}

}  // namespace spgvbd
