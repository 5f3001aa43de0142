// Evaluation of one element's per-vertex derivatives from the prepared parameters of spg_vbd.h,
// through the same device functions the solve uses. Shared by the GPU derivative kernel and the
// host instantiation used for derivative checks.
#pragma once

#include "element.h"

namespace spgvbd {

enum { kElemSpring = 1, kElemStvk = 2, kElemBending = 3, kElemSnh = 4, kElemAnchor = 5 };

SPG_HD int elemStencilSize(int type)
{
    return type == kElemSpring ? 2 : type == kElemStvk ? 3 : type == kElemAnchor ? 1 : (type == kElemBending || type == kElemSnh) ? 4 : 0;
}
SPG_HD int elemParams(int type)
{
    return type == kElemSpring ? 2 : type == kElemStvk ? 13 : type == kElemBending ? 2 : type == kElemSnh ? 12 : type == kElemAnchor ? 4 : 0;
}

// variant 0: the forms the solve uses (hinge edge/wing split); 1: the general (a, b, sigma) hinge form;
// 2: the single-expression compact hinge form.
template <class R>
SPG_HD void elementVertex(int type, const V3<R> *x, const double *p, int s, int variant, V3<R> &ge, Sym3<R> &He)
{
    ge = {R(0), R(0), R(0)};
    He = zeroSym<R>();
    if (type == kElemSpring) {
        springVertex<R>(x[s], x[1 - s], R(p[0]), R(p[1]), ge, He);
    } else if (type == kElemAnchor) {
        anchorVertex<R>(x[0], V3<R>{R(p[0]), R(p[1]), R(p[2])}, R(p[3]), ge, He);
    } else if (type == kElemStvk) {
        // rows of Dinv: g_1 = (p0, p1), g_2 = (p2, p3), g_0 = -(g_1 + g_2)
        const double g[3][2] = {{-(p[0] + p[2]), -(p[1] + p[3])}, {p[0], p[1]}, {p[2], p[3]}};
        const int ia = s == 0 ? 1 : 0, ib = s == 2 ? 1 : 2;
        stvkVertex<R>(x[s], x[ia], x[ib], R(g[ia][0]), R(g[ia][1]), R(g[ib][0]), R(g[ib][1]), R(p[4]), R(p[5]), R(p[12]),
                      ge, He);
    } else if (type == kElemBending) {
        if (variant == 1)
            bendingVertex<R>(x[0], x[1], x[2], x[3], s, R(p[0]), R(p[1]), ge, He);
        else if (variant == 2)
            bendingVertexUnified<R>(x[0], x[1], x[2], x[3], s, R(p[0]), R(p[1]), ge, He);
        else
            bendingVertexCompact<R>(x[0], x[1], x[2], x[3], s, R(p[0]), R(p[1]), ge, He);
    } else if (type == kElemSnh) {
        double gr[4][3];
        for (int c = 0; c < 3; ++c) {
            gr[1][c] = p[c];
            gr[2][c] = p[3 + c];
            gr[3][c] = p[6 + c];
            gr[0][c] = -(p[c] + p[3 + c] + p[6 + c]);
        }
        int o[3], m = 0;
        for (int q = 0; q < 4; ++q)
            if (q != s) o[m++] = q;
        const V3<R> g0 = {R(gr[o[0]][0]), R(gr[o[0]][1]), R(gr[o[0]][2])};
        const V3<R> g1 = {R(gr[o[1]][0]), R(gr[o[1]][1]), R(gr[o[1]][2])};
        const V3<R> g2 = {R(gr[o[2]][0]), R(gr[o[2]][1]), R(gr[o[2]][2])};
        snhVertex<R>(x[s], x[o[0]], x[o[1]], x[o[2]], g0, g1, g2, R(p[9]), R(p[10]), R(p[11]), ge, He);
    }
}

// All vertices of element i: xs (ns*3), params (np) -> grad (ns*3), hess (ns*9, row-major blocks).
template <class R>
SPG_HD void elementAll(int type, const double *xe, const double *p, int variant, double *grad, double *hess)
{
    const int ns = elemStencilSize(type);
    V3<R> x[4];
    for (int s = 0; s < ns; ++s) x[s] = {R(xe[3 * s]), R(xe[3 * s + 1]), R(xe[3 * s + 2])};
    for (int s = 0; s < ns; ++s) {
        V3<R> ge;
        Sym3<R> He;
        elementVertex<R>(type, x, p, s, variant, ge, He);
        double *gd = grad + s * 3;
        gd[0] = ge.x;
        gd[1] = ge.y;
        gd[2] = ge.z;
        double *hd = hess + s * 9;
        hd[0] = He.xx;
        hd[1] = He.xy;
        hd[2] = He.xz;
        hd[3] = He.xy;
        hd[4] = He.yy;
        hd[5] = He.yz;
        hd[6] = He.xz;
        hd[7] = He.yz;
        hd[8] = He.zz;
    }
}

}  // namespace spgvbd
