// Closed-form per-vertex derivatives of the five spg VBD energies.
//
// Each function returns, for the vertex being updated ("own" vertex), the gradient of the element
// energy w.r.t. that vertex and the 3x3 diagonal block of the element Hessian for it. That is all
// StencilBlockEnergy::accumulateVertexForce/Hessian keep of the full TinyAD derivative
// (spg/include/spg/sim/energy/energy.h:240-258). No PSD projection (energy.h:270,354).
//
// Templated on the scalar: float for the vertex solve, double for the residual kernel and for
// derivative checks. Host and device.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace spgvbd {

#define SPG_HD __host__ __device__ __forceinline__

template <class R>
struct V3 {
    R x, y, z;
};

template <class R>
struct Sym3 {  // symmetric 3x3: xx, xy, xz, yy, yz, zz
    R xx, xy, xz, yy, yz, zz;
};

template <class R>
SPG_HD V3<R> v3(R x, R y, R z)
{
    return {x, y, z};
}
template <class R>
SPG_HD V3<R> operator+(V3<R> a, V3<R> b)
{
    return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class R>
SPG_HD V3<R> operator-(V3<R> a, V3<R> b)
{
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class R>
SPG_HD V3<R> operator-(V3<R> a)
{
    return {-a.x, -a.y, -a.z};
}
template <class R>
SPG_HD V3<R> operator*(R s, V3<R> a)
{
    return {s * a.x, s * a.y, s * a.z};
}
template <class R>
SPG_HD R dot(V3<R> a, V3<R> b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class R>
SPG_HD V3<R> cross(V3<R> a, V3<R> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// s * (u v^T + v u^T)
template <class R>
SPG_HD void addSymOuter(Sym3<R> &H, R s, V3<R> u, V3<R> v)
{
    H.xx += s * (2 * u.x * v.x);
    H.yy += s * (2 * u.y * v.y);
    H.zz += s * (2 * u.z * v.z);
    H.xy += s * (u.x * v.y + v.x * u.y);
    H.xz += s * (u.x * v.z + v.x * u.z);
    H.yz += s * (u.y * v.z + v.y * u.z);
}
// s * u u^T
template <class R>
SPG_HD void addOuter(Sym3<R> &H, R s, V3<R> u)
{
    H.xx += s * u.x * u.x;
    H.yy += s * u.y * u.y;
    H.zz += s * u.z * u.z;
    H.xy += s * u.x * u.y;
    H.xz += s * u.x * u.z;
    H.yz += s * u.y * u.z;
}
template <class R>
SPG_HD void addDiag(Sym3<R> &H, R s)
{
    H.xx += s;
    H.yy += s;
    H.zz += s;
}
template <class R>
SPG_HD Sym3<R> zeroSym()
{
    return {R(0), R(0), R(0), R(0), R(0), R(0)};
}

// ---------------------------------------------------------------------------------------------
// Spring (springEnergy.cpp:101-134): E = 1/2 k (|s| - L0)^2 with s = x_own - x_other.
// grad = k (|s| - L0) s/|s|; H = k [(1 - L0/|s|)(I - s^ s^T) + s^ s^T]. Same for either slot.
// ---------------------------------------------------------------------------------------------
template <class R>
SPG_HD void springVertex(V3<R> xo, V3<R> xa, R L0, R k, V3<R> &g, Sym3<R> &H)
{
    const V3<R> s = xo - xa;
    const R n2 = dot(s, s);
    const R n = sqrt(n2);
    const R inv = R(1) / n;
    g = (k * (n - L0) * inv) * s;
    const R a = k * (R(1) - L0 * inv);  // coefficient of I
    const R b = (k - a) * inv * inv;    // coefficient of s s^T (k L0/|s|^3)
    H = zeroSym<R>();
    addDiag(H, a);
    addOuter(H, b, s);
}

// ---------------------------------------------------------------------------------------------
// Spring anchor (springAnchorEnergy.cpp:21-26): E = 1/2 k |x - a|^2; grad = k (x - a); H = k I.
// ---------------------------------------------------------------------------------------------
template <class R>
SPG_HD void anchorVertex(V3<R> xo, V3<R> a, R k, V3<R> &g, Sym3<R> &H)
{
    g = k * (xo - a);
    H = zeroSym<R>();
    addDiag(H, k);
}

// ---------------------------------------------------------------------------------------------
// StVK membrane (membraneStvkEnergy.cpp:9-28, energy.h:336-341).
// F = [x1-x0, x2-x0] Dinv = sum_s x_s g_s^T with g_1, g_2 the rows of Dinv and g_0 = -(g_1+g_2),
// so with the own vertex o and the two others a, b: F = (x_a - x_o) g_a^T + (x_b - x_o) g_b^T and
// g_o = -(g_a + g_b). C = (E00, E11, 2 E01) of the Green strain, sigma = K C with the StVK
// stiffness K = [[kd, ko, 0], [ko, kd, 0], [0, 0, ks]] (area-scaled).
// grad_o = F S g_o with S = [[s0, s2], [s2, s1]];
// H_oo = (g_o^T S g_o) I + F (B^T K B) F^T, B = [[g0, 0], [0, g1], [g1, g0]].
// ---------------------------------------------------------------------------------------------
template <class R>
SPG_HD void stvkVertex(V3<R> xo, V3<R> xa, V3<R> xb, R ga0, R ga1, R gb0, R gb1, R kd, R ko, R ks, V3<R> &g,
                       Sym3<R> &H)
{
    const V3<R> da = xa - xo, db = xb - xo;
    const V3<R> f0 = ga0 * da + gb0 * db;
    const V3<R> f1 = ga1 * da + gb1 * db;
    const R c0 = R(0.5) * (dot(f0, f0) - R(1));
    const R c1 = R(0.5) * (dot(f1, f1) - R(1));
    const R c2 = dot(f0, f1);
    const R s0 = kd * c0 + ko * c1;
    const R s1 = ko * c0 + kd * c1;
    const R s2 = ks * c2;
    const R g0 = -(ga0 + gb0), g1 = -(ga1 + gb1);
    const R a0 = s0 * g0 + s2 * g1;
    const R a1 = s2 * g0 + s1 * g1;
    g = a0 * f0 + a1 * f1;
    const R gSg = g0 * a0 + g1 * a1;
    const R m00 = kd * g0 * g0 + ks * g1 * g1;
    const R m01 = (ko + ks) * g0 * g1;
    const R m11 = kd * g1 * g1 + ks * g0 * g0;
    H = zeroSym<R>();
    addDiag(H, gSg);
    addOuter(H, m00, f0);
    addOuter(H, m11, f1);
    addSymOuter(H, m01, f0, f1);
}

// ---------------------------------------------------------------------------------------------
// Stable Neo-Hookean (stableNeoHookeanEnergy.cpp:33-61): E = 1/2 lambda (J - alpha)^2 + 1/2 mu (|F|^2 - 3).
// F = sum_{others} (x_a - x_o) g_a^T, g_o = -sum g_a. P = lambda (J - alpha) cof(F) + mu F.
// grad_o = P g_o. Along x_o, F moves by a rank-one update, on which det is affine, so
// H_oo = mu |g_o|^2 I + lambda c c^T with c = cof(F) g_o.
// ---------------------------------------------------------------------------------------------
template <class R>
SPG_HD void snhVertex(V3<R> xo, V3<R> xa, V3<R> xb, V3<R> xc, V3<R> ga, V3<R> gb, V3<R> gc, R lambda, R mu, R alpha,
                      V3<R> &g, Sym3<R> &H)
{
    const V3<R> da = xa - xo, db = xb - xo, dc = xc - xo;
    // columns of F
    const V3<R> f0 = ga.x * da + gb.x * db + gc.x * dc;
    const V3<R> f1 = ga.y * da + gb.y * db + gc.y * dc;
    const V3<R> f2 = ga.z * da + gb.z * db + gc.z * dc;
    const V3<R> k0 = cross(f1, f2), k1 = cross(f2, f0), k2 = cross(f0, f1);
    const R J = dot(f0, k0);
    const V3<R> go = -(ga + gb + gc);
    const V3<R> c = go.x * k0 + go.y * k1 + go.z * k2;
    const V3<R> fg = go.x * f0 + go.y * f1 + go.z * f2;
    g = (lambda * (J - alpha)) * c + mu * fg;
    H = zeroSym<R>();
    addDiag(H, mu * dot(go, go));
    addOuter(H, lambda, c);
}

// ---------------------------------------------------------------------------------------------
// Discrete bending, compact form used by the solve. Same angle theta = atan2(s, c) as below. With
// the signed angles alpha, beta of a and b about the hinge axis, theta = beta - alpha + pi and the
// frame rotation cancels in the difference, so grad_v theta = B^T (e^ x b)/|b|^2 - A^T (e^ x a)/|a|^2.
// For every slot this collapses to
//   grad theta = sg (ra.e^)/|a|^2 a - sg (rb.e^)/|b|^2 b
//   Hess theta = 1/2 sym(a, ua) + 1/2 sym(b, ub),
//   ua =  ra_perp/(L |a|^2) + 2 (ra.e^) (a x ra)/|a|^4,
//   ub = -rb_perp/(L |b|^2) - 2 (rb.e^) (b x rb)/|b|^4,
// (r_perp = r - (r.e^) e^; the skew parts of the Jacobian cancel, so only symmetric parts remain) with
//   slot 0: ra = x2 - x1, rb = x3 - x1, sg = -1     slot 1: ra = x2 - x0, rb = x3 - x0, sg = +1
//   slot 2: ra = e, rb = 0, sg = -1                 slot 3: ra = 0, rb = e, sg = -1
// ---------------------------------------------------------------------------------------------
// The hinge angle's atan2 and the hinge edge's length and inverse length, by scalar type. FP64 takes the
// library functions. FP32 (the vertex solve) takes:
//   * atan2 by octant reduction to t = min(|x|, |y|) / max(|x|, |y|) in [0, 1], then atan t = t P(t^2) with P of
//     degree 8 (coefficients fitted for this code by weighted least squares on Chebyshev nodes of [0, 1]; largest
//     error 1.1e-7 in FP32 Horner evaluation), then the octant's reflections; the arguments are finite (a
//     degenerate (0, 0) gives 0), so none of the library's special cases are needed;
//   * the length as |e|^2 times the reciprocal square root (one transcendental instead of a square root and a
//     reciprocal).
SPG_HD float hingeAtan2(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float t = mx > 0.f ? mn * (1.f / mx) : 0.f;
    const float s = t * t;
    float p = 0.0028429856f;
    p = p * s - 0.016038457f;
    p = p * s + 0.042637214f;
    p = p * s - 0.07499113f;
    p = p * s + 0.10638127f;
    p = p * s - 0.14202788f;
    p = p * s + 0.19992484f;
    p = p * s - 0.33333063f;
    p = p * s + 1.0f;
    float r = t * p;
    r = ay > ax ? 1.57079632679489662f - r : r;
    r = x < 0.f ? 3.14159265358979324f - r : r;
    return copysignf(r, y);
}
SPG_HD double hingeAtan2(double y, double x) { return atan2(y, x); }
SPG_HD void lengthAndInverse(float ee, float &L, float &iL)
{
#if defined(__HIP_DEVICE_COMPILE__)
    iL = rsqrtf(ee);
#else
    iL = 1.f / std::sqrt(ee);
#endif
    L = ee * iL;
}
SPG_HD void lengthAndInverse(double ee, double &L, double &iL)
{
    L = sqrt(ee);
    iL = 1.0 / L;
}

// Hinge slots 0 and 1 (the hinge edge's vertices): ra, rb of the compact form. Both slots share
// ra_perp = (x2 - x1)_perp and rb_perp = (x3 - x1)_perp since e_perp = 0.
template <class R>
SPG_HD void bendingEdge(V3<R> x0, V3<R> x1, V3<R> x2, V3<R> x3, bool slot1, R theta0, R k, V3<R> &g, Sym3<R> &H)
{
    const V3<R> e = x1 - x0;
    const V3<R> p = x2 - x1, q = x3 - x1;
    const V3<R> a = cross(e, p), b = cross(e, q);
    R L, iL;
    lengthAndInverse(dot(e, e), L, iL);
    const V3<R> eh = iL * e;
    const R pe = dot(p, eh), qe = dot(q, eh);
    const V3<R> pp = p - pe * eh, qp = q - qe * eh;  // perpendicular parts
    // slot 1 shifts the lever arms by e: s1 e with s1 = 0 or 1 adds exactly e or nothing, without selects
    const R s1 = slot1 ? R(1) : R(0);
    const R rae = s1 * L + pe, rbe = s1 * L + qe;
    const V3<R> ra = s1 * e + p, rb = s1 * e + q;
    const R sg = R(2) * s1 - R(1);
    const R ia2 = R(1) / dot(a, a), ib2 = R(1) / dot(b, b);
    const V3<R> gth = (sg * rae * ia2) * a + (-sg * rbe * ib2) * b;
    const V3<R> ua = (iL * ia2) * pp + (R(2) * rae * ia2 * ia2) * cross(a, ra);
    const V3<R> ub = (-iL * ib2) * qp + (R(-2) * rbe * ib2 * ib2) * cross(b, rb);
    // (a x b).e^ = (e.(p x q)) (e.e^) = (a.q) L, since (e x p) x (e x q) = (e.(p x q)) e
    const R theta = hingeAtan2(-dot(a, q) * L, -dot(a, b));
    const R kC = k * (theta - theta0);
    g = kC * gth;
    H = zeroSym<R>();
    addOuter(H, k, gth);
    addSymOuter(H, R(0.5) * kC, a, ua);
    addSymOuter(H, R(0.5) * kC, b, ub);
}

// Hinge slots 2 and 3 (the wing vertices): with w the normal of the own wing (a for slot 2, b for
// slot 3) and s_w = -1 (slot 2) / +1 (slot 3): grad theta = s_w L w/|w|^2,
// Hess theta = -s_w L sym(w, w x e)/|w|^4.
template <class R>
SPG_HD void bendingWing(V3<R> x0, V3<R> x1, V3<R> x2, V3<R> x3, bool slot3, R theta0, R k, V3<R> &g, Sym3<R> &H)
{
    const V3<R> e = x1 - x0;
    const V3<R> q = x3 - x1;
    const V3<R> a = cross(e, x2 - x1), b = cross(e, q);
    R L, iL;
    lengthAndInverse(dot(e, e), L, iL);
    // (a x b).e^ = (a.q) L, as in bendingEdge
    const R theta = hingeAtan2(-dot(a, q) * L, -dot(a, b));
    const V3<R> w = slot3 ? b : a;
    const R sw = slot3 ? R(1) : R(-1);
    const R iw2 = R(1) / dot(w, w);
    const V3<R> gth = (sw * L * iw2) * w;
    const R kC = k * (theta - theta0);
    g = kC * gth;
    H = zeroSym<R>();
    addOuter(H, k, gth);
    addSymOuter(H, -kC * sw * L * iw2 * iw2, w, cross(w, e));
}

// The hinge angle alone, as bendingEdge computes it: theta = atan2(-(a x b).e^, -a.b).
template <class R>
SPG_HD R bendingTheta(V3<R> x0, V3<R> x1, V3<R> x2, V3<R> x3)
{
    const V3<R> e = x1 - x0;
    const V3<R> a = cross(e, x2 - x1), b = cross(e, x3 - x1);
    const R iL = R(1) / sqrt(dot(e, e));
    return atan2(-dot(cross(a, b), iL * e), -dot(a, b));
}

// The gradient alone of the hinge energy at an edge slot (bendingEdge's g) for a given kC = k (theta - theta0),
// so that the angle, shared by the hinge's four slots, is evaluated once per hinge (the FP64 residual).
template <class R>
SPG_HD V3<R> bendingEdgeGrad(V3<R> x0, V3<R> x1, V3<R> x2, V3<R> x3, bool slot1, R kC)
{
    const V3<R> e = x1 - x0;
    const V3<R> p = x2 - x1, q = x3 - x1;
    const V3<R> a = cross(e, p), b = cross(e, q);
    const R L = sqrt(dot(e, e));
    const R iL = R(1) / L;
    const V3<R> eh = iL * e;
    const R pe = dot(p, eh), qe = dot(q, eh);
    const R rae = slot1 ? pe + L : pe, rbe = slot1 ? qe + L : qe;
    const R sg = slot1 ? R(1) : R(-1);
    const R ia2 = R(1) / dot(a, a), ib2 = R(1) / dot(b, b);
    return kC * ((sg * rae * ia2) * a + (-sg * rbe * ib2) * b);
}

// The same at a wing slot (bendingWing's g): kC s_w L w / |w|^2.
template <class R>
SPG_HD V3<R> bendingWingGrad(V3<R> x0, V3<R> x1, V3<R> x2, V3<R> x3, bool slot3, R kC)
{
    const V3<R> e = x1 - x0;
    const V3<R> w = cross(e, (slot3 ? x3 : x2) - x1);
    const R L = sqrt(dot(e, e));
    const R sw = slot3 ? R(1) : R(-1);
    return (kC * (sw * L / dot(w, w))) * w;
}

template <class R>
SPG_HD void bendingVertexCompact(V3<R> x0, V3<R> x1, V3<R> x2, V3<R> x3, int slot, R theta0, R k, V3<R> &g,
                                 Sym3<R> &H)
{
    if (slot < 2)
        bendingEdge<R>(x0, x1, x2, x3, slot == 1, theta0, k, g, H);
    else
        bendingWing<R>(x0, x1, x2, x3, slot == 3, theta0, k, g, H);
}

// Single-expression version of the compact form for every slot (kept for cross-checks).
template <class R>
SPG_HD void bendingVertexUnified(V3<R> x0, V3<R> x1, V3<R> x2, V3<R> x3, int slot, R theta0, R k, V3<R> &g,
                                 Sym3<R> &H)
{
    const V3<R> e = x1 - x0;
    const V3<R> p = x2 - x1, q = x3 - x1;
    const V3<R> a = cross(e, p), b = cross(e, q);
    const R L = sqrt(dot(e, e));
    const R iL = R(1) / L;
    const V3<R> eh = iL * e;
    const V3<R> zero = {R(0), R(0), R(0)};
    const V3<R> ra = slot == 0 ? p : (slot == 1 ? x2 - x0 : (slot == 2 ? e : zero));
    const V3<R> rb = slot == 0 ? q : (slot == 1 ? x3 - x0 : (slot == 3 ? e : zero));
    const R sg = slot == 1 ? R(1) : R(-1);
    const R ia2 = R(1) / dot(a, a), ib2 = R(1) / dot(b, b);
    const R rae = dot(ra, eh), rbe = dot(rb, eh);
    const V3<R> gth = (sg * rae * ia2) * a + (-sg * rbe * ib2) * b;
    const V3<R> ua = (iL * ia2) * (ra - rae * eh) + (R(2) * rae * ia2 * ia2) * cross(a, ra);
    const V3<R> ub = (-iL * ib2) * (rb - rbe * eh) + (R(-2) * rbe * ib2 * ib2) * cross(b, rb);
    // theta as atan2(s, c): s = -(a x b).e^, c = -a.b
    const R theta = atan2(-dot(cross(a, b), eh), -dot(a, b));
    const R C = theta - theta0;
    const R kC = k * C;
    g = kC * gth;
    H = zeroSym<R>();
    addOuter(H, k, gth);
    addSymOuter(H, R(0.5) * kC, a, ua);
    addSymOuter(H, R(0.5) * kC, b, ub);
}

// ---------------------------------------------------------------------------------------------
// Discrete bending (discreteBendingEnergy.cpp:9-30): hinge x0-x1, wings x2, x3;
// a = e x (x2 - x1), b = e x (x3 - x1), e = x1 - x0. The reference's angle
// theta = atan2((n1 x n2).e^, n1.n2) with n1 = a^, n2 = -b^ equals atan2(s, c) with
// s = -(a x b).e / |e| and c = -a.b. For the own slot, a, b and e are affine in x_own:
// da = [abar]x dx, db = [bbar]x dx, de = sigma dx with
//   slot 0: abar = x2 - x1, bbar = x3 - x1, sigma = -1
//   slot 1: abar = x0 - x2, bbar = x0 - x3, sigma = +1
//   slot 2: abar = e, bbar = 0; slot 3: abar = 0, bbar = e (sigma = 0)
// which gives with [u]x[v]x = v u^T - (u.v) I, T = (a x b).e:
//   grad c = abar x b + bbar x a,        Hess c = sym(abar, bbar) - 2 (abar.bbar) I
//   grad T = -abar x (b x e) - bbar x (e x a) + sigma a x b
//   Hess T = sym(abar x bbar, e) - sigma sym(b, abar) + sigma sym(a, bbar) + 2 sigma (abar.b - a.bbar) I
// (sym(u, v) = u v^T + v u^T), then s = -T/L (L = |e|) by the quotient rule and theta by the atan2
// chain rule. E = 1/2 k (theta - theta0)^2.
// ---------------------------------------------------------------------------------------------
template <class R>
SPG_HD void bendingVertex(V3<R> x0, V3<R> x1, V3<R> x2, V3<R> x3, int slot, R theta0, R k, V3<R> &g, Sym3<R> &H)
{
    const V3<R> e = x1 - x0;
    const V3<R> p = x2 - x1, q = x3 - x1;
    const V3<R> a = cross(e, p), b = cross(e, q);
    const V3<R> zero = {R(0), R(0), R(0)};
    V3<R> ab, bb;
    R sigma;
    if (slot == 0) {
        ab = p;
        bb = q;
        sigma = R(-1);
    } else if (slot == 1) {
        ab = x0 - x2;
        bb = x0 - x3;
        sigma = R(1);
    } else if (slot == 2) {
        ab = e;
        bb = zero;
        sigma = R(0);
    } else {
        ab = zero;
        bb = e;
        sigma = R(0);
    }
    const R c = -dot(a, b);
    const V3<R> gc = cross(ab, b) + cross(bb, a);
    const V3<R> axb = cross(a, b);
    const R T = dot(axb, e);
    const V3<R> gT = -cross(ab, cross(b, e)) - cross(bb, cross(e, a)) + sigma * axb;
    const R L2 = dot(e, e);
    const R L = sqrt(L2);
    const R iL = R(1) / L;
    const V3<R> eh = iL * e;
    const R s = -T * iL;
    const V3<R> gs = (-iL) * gT + (T * sigma * iL * iL) * eh;
    // Hs = -HT/L + sigma sym(gT, eh)/L^2 - T sigma^2 (3 eh eh^T - I)/L^3
    // Hc = sym(ab, bb) - 2 (ab.bb) I
    // Htheta = (c Hs - s Hc)/r2 - sym(gth, w)/r2 with w = s gs + c gc, gth = (c gs - s gc)/r2
    const R r2 = s * s + c * c;
    const R ir2 = R(1) / r2;
    const V3<R> gth = ir2 * (c * gs - s * gc);
    const V3<R> w = s * gs + c * gc;
    const R cs = c * ir2, ss = s * ir2;
    Sym3<R> Ht = zeroSym<R>();
    // c/r2 * Hs
    const R hsT = -cs * iL;                   // coefficient of HT
    addSymOuter(Ht, hsT, cross(ab, bb), e);   // HT term 1
    addSymOuter(Ht, -hsT * sigma, b, ab);     // HT term 2
    addSymOuter(Ht, hsT * sigma, a, bb);      // HT term 3
    addDiag(Ht, hsT * R(2) * sigma * (dot(ab, b) - dot(a, bb)));
    addSymOuter(Ht, cs * sigma * iL * iL, gT, eh);
    const R t3 = -cs * T * sigma * sigma * iL * iL * iL;
    addOuter(Ht, R(3) * t3, eh);
    addDiag(Ht, -t3);
    // -s/r2 * Hc
    addSymOuter(Ht, -ss, ab, bb);
    addDiag(Ht, ss * R(2) * dot(ab, bb));
    // -sym(gth, w)/r2
    addSymOuter(Ht, -ir2, gth, w);
    const R theta = atan2(s, c);
    const R C = theta - theta0;
    g = (k * C) * gth;
    H = zeroSym<R>();
    addOuter(H, k, gth);
    H.xx += k * C * Ht.xx;
    H.xy += k * C * Ht.xy;
    H.xz += k * C * Ht.xz;
    H.yy += k * C * Ht.yy;
    H.yz += k * C * Ht.yz;
    H.zz += k * C * Ht.zz;
}

}  // namespace spgvbd
