// Rigid-body pieces of the VBD solve, shared by the GPU kernels and the host entry point that evaluates them
// for tests: axis-angle maps, the closed-form derivatives of the four rigid-body energies per body slot, and the
// 6x6 solve with its determinant. All FP64.
//
// Parameterisation (rigidBodySpringEnergy.cpp:15-31 and the three sibling energies): a body's DOFs are its centre
// of mass c and a rotation increment t applied as (I + [t]x) R, differentiated at t = 0. A body point with local
// coordinates l sits at p = c + (I + [t]x) R l, so with r = R l its Jacobian is J = [I, -[r]x] and p is linear in
// (c, t): every element Hessian is J^T H_p J with H_p the Hessian of the energy in the points, and the gradient is
// J^T g_p = (g_p, r x g_p).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace spgrigid {

#define RB_HD __host__ __device__ __forceinline__

enum RigidType { kRbSpring = 21, kRbSpringSquared = 22, kRbAnchor = 23, kRbAlignment = 24 };
constexpr int kRbParamStride = 8;

RB_HD int rbStencil(int type) { return type == kRbAnchor ? 1 : 2; }
RB_HD int rbParams(int type)
{
    return type == kRbAnchor ? 7 : type == kRbAlignment ? 1 : 8;
}

RB_HD void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
RB_HD double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
RB_HD void matVec3(const double *M, const double *v, double *o)
{
    for (int i = 0; i < 3; ++i) o[i] = M[3 * i] * v[0] + M[3 * i + 1] * v[1] + M[3 * i + 2] * v[2];
}
// C = A B (3x3 row-major)
RB_HD void matMul3(const double *A, const double *B, double *C)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
// C = A B^T
RB_HD void matMulT3(const double *A, const double *B, double *C)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            C[3 * i + j] = A[3 * i] * B[3 * j] + A[3 * i + 1] * B[3 * j + 1] + A[3 * i + 2] * B[3 * j + 2];
}
RB_HD void skew3(const double *v, double *S)
{
    S[0] = 0;
    S[1] = -v[2];
    S[2] = v[1];
    S[3] = v[2];
    S[4] = 0;
    S[5] = -v[0];
    S[6] = -v[1];
    S[7] = v[0];
    S[8] = 0;
}

// Rotation by `angle` about the unit axis a: Rodrigues' formula, rounded as Eigen's AngleAxis::toRotationMatrix
// forms it (shared products (1 - c) a_i a_j for each symmetric pair).
RB_HD void angleAxisToR(double angle, const double *a, double *R)
{
    const double c = cos(angle), sn = sin(angle);
    const double s0 = sn * a[0], s1 = sn * a[1], s2 = sn * a[2];
    const double c0 = (1 - c) * a[0], c1 = (1 - c) * a[1], c2 = (1 - c) * a[2];
    R[0] = c0 * a[0] + c;
    R[4] = c1 * a[1] + c;
    R[8] = c2 * a[2] + c;
    double t = c0 * a[1];
    R[1] = t - s2;
    R[3] = t + s2;
    t = c0 * a[2];
    R[2] = t + s1;
    R[6] = t - s1;
    t = c1 * a[2];
    R[5] = t - s0;
    R[7] = t + s0;
}

// axisAngleToRotMatrix (rigidBodyGroup.cpp:6-10): the zero vector is the identity.
RB_HD void aaToR(const double *t, double *R)
{
    const double angle = sqrt(dot3(t, t));
    if (angle == 0) {
        for (int k = 0; k < 9; ++k) R[k] = (k % 4 == 0) ? 1.0 : 0.0;
        return;
    }
    const double a[3] = {t[0] / angle, t[1] / angle, t[2] / angle};
    angleAxisToR(angle, a, R);
}

// rotMatrixToAxisAngle (rigidBodyGroup.cpp:12-16): the unit quaternion of R (the trace, or the largest diagonal
// entry, selects the numerically stable formula), then angle 2 atan2(|v|, |w|) about v / |v| with the sign of w.
// Quaternion (w, x, y, z) of R when its largest diagonal entry R_II is used as the pivot; J, K follow I cyclically.
template <int I>
RB_HD void quatFromDiagonal(const double *R, double *q)
{
    constexpr int J = (I + 1) % 3, K = (J + 1) % 3;
    double s = sqrt(R[4 * I] - R[4 * J] - R[4 * K] + 1.0);
    q[1 + I] = 0.5 * s;
    s = 0.5 / s;
    q[0] = (R[3 * K + J] - R[3 * J + K]) * s;
    q[1 + J] = (R[3 * J + I] + R[3 * I + J]) * s;
    q[1 + K] = (R[3 * K + I] + R[3 * I + K]) * s;
}

RB_HD void rToAa(const double *R, double *t)
{
    double q[4];  // w, x, y, z
    const double tr = R[0] + R[4] + R[8];
    if (tr > 0) {
        double s = sqrt(tr + 1.0);
        q[0] = 0.5 * s;
        s = 0.5 / s;
        q[1] = (R[7] - R[5]) * s;
        q[2] = (R[2] - R[6]) * s;
        q[3] = (R[3] - R[1]) * s;
    } else {
        const int i0 = R[4] > R[0] ? 1 : 0;
        const int i = R[8] > (i0 ? R[4] : R[0]) ? 2 : i0;
        if (i == 0)
            quatFromDiagonal<0>(R, q);
        else if (i == 1)
            quatFromDiagonal<1>(R, q);
        else
            quatFromDiagonal<2>(R, q);
    }
    double n = sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (n == 0) {
        t[0] = t[1] = t[2] = 0;
        return;
    }
    const double angle = 2.0 * atan2(n, fabs(q[0]));
    if (q[0] < 0) n = -n;
    t[0] = q[1] / n * angle;
    t[1] = q[2] / n * angle;
    t[2] = q[3] / n * angle;
}

// Block storage of the accumulated 6x6 Hessian: the full row-major matrix (36 entries), or its upper triangle
// (21 entries, row-major) when the Hessian is known to be symmetric; entries below the diagonal are dropped then.
template <bool UPPER>
RB_HD void addB(double *B, int i, int j, double v)
{
    if constexpr (UPPER) {
        if (i <= j) B[i * 6 - i * (i - 1) / 2 + (j - i)] += v;
    } else {
        B[6 * i + j] += v;
    }
}

// Adds J^T H_p J (J = [I, -[r]x]) into the 6x6 block B and J^T g into f with the given sign (the solve
// accumulates forces, -gradient).
template <bool UPPER>
RB_HD void addPointTerms(const double *r, const double *gp, const double *Hp, double gsign, double *f, double *B)
{
    double rg[3];
    cross3(r, gp, rg);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        f[d] -= gsign * gp[d];
        f[3 + d] -= gsign * rg[d];
    }
    double S[9], HS[9], SH[9];
    skew3(r, S);
    // H (-S), S H, S H (-S)
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            HS[3 * i + j] = -(Hp[3 * i] * S[j] + Hp[3 * i + 1] * S[3 + j] + Hp[3 * i + 2] * S[6 + j]);
            SH[3 * i + j] = S[3 * i] * Hp[j] + S[3 * i + 1] * Hp[3 + j] + S[3 * i + 2] * Hp[6 + j];
        }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            addB<UPPER>(B, i, j, Hp[3 * i + j]);
            addB<UPPER>(B, i, 3 + j, HS[3 * i + j]);
            addB<UPPER>(B, 3 + i, j, SH[3 * i + j]);
            addB<UPPER>(B, 3 + i, 3 + j, -(SH[3 * i] * S[j] + SH[3 * i + 1] * S[3 + j] + SH[3 * i + 2] * S[6 + j]));
        }
}

// One element's force (-gradient) and diagonal Hessian block for the body in `slot`, accumulated into f and B.
// x0, R0 (x1, R1): centre and rotation of the stencil's first (second) body; p: the element's parameters
// (springs: l0, l1, L0 (squared spring: L0^2), k; anchor: l, a, k; alignment: k).
template <bool UPPER = false>
RB_HD void rbAccumulate(int type, int slot, const double *x0, const double *R0, const double *x1, const double *R1,
                        const double *p, double *f, double *B)
{
    if (type == kRbAnchor) {  // energy 1/2 k |p - a|^2 (rigidBodySpringAnchorEnergy.cpp:35-45, 85-100)
        double r[3];
        matVec3(R0, p, r);
        const double k = p[6];
        const double gp[3] = {k * (x0[0] + r[0] - p[3]), k * (x0[1] + r[1] - p[4]), k * (x0[2] + r[2] - p[5])};
        const double Hp[9] = {k, 0, 0, 0, k, 0, 0, 0, k};
        addPointTerms<UPPER>(r, gp, Hp, 1.0, f, B);
        return;
    }
    if (type == kRbAlignment) {
        // C = sum_k |u_k x w_k|^2 with u_k, w_k the k-th columns of (I + [t0]x) R0 and (I + [t1]x) R1; E = 1/2 k C^2
        // (rigidBodyOrientationAlignmentEnergy.cpp:15-31). With c_k = u_k x w_k: dc_k/dt0 = [w_k]x [u_k]x,
        // dc_k/dt1 = -[u_k]x [w_k]x, and c_k is bilinear in (t0, t1), so d2C/dts2 = 2 sum_k J_k^T J_k.
        double C = 0, g[3] = {0, 0, 0}, M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double u[3] = {R0[k], R0[3 + k], R0[6 + k]}, w[3] = {R1[k], R1[3 + k], R1[6 + k]};
            double c[3], Su[9], Sw[9], J[9];
            cross3(u, w, c);
            C += dot3(c, c);
            skew3(u, Su);
            skew3(w, Sw);
            if (slot == 0)
                matMul3(Sw, Su, J);
            else {
                matMul3(Su, Sw, J);
#pragma unroll
                for (int q = 0; q < 9; ++q) J[q] = -J[q];
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                g[i] += 2.0 * (J[i] * c[0] + J[3 + i] * c[1] + J[6 + i] * c[2]);
#pragma unroll
                for (int j = 0; j < 3; ++j) M[3 * i + j] += J[i] * J[j] + J[3 + i] * J[3 + j] + J[6 + i] * J[6 + j];
            }
        }
        const double k = p[0];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            f[3 + i] -= k * C * g[i];
#pragma unroll
            for (int j = 0; j < 3; ++j) addB<UPPER>(B, 3 + i, 3 + j, k * (g[i] * g[j] + 2.0 * C * M[3 * i + j]));
        }
        return;
    }
    // springs (rigidBodySpringEnergy.cpp:15-31, rigidBodySpringSquaredEnergy.cpp:15-33): d = p0 - p1
    double r0[3], r1[3];
    matVec3(R0, p, r0);
    matVec3(R1, p + 3, r1);
    const double d[3] = {x0[0] + r0[0] - x1[0] - r1[0], x0[1] + r0[1] - x1[1] - r1[1],
                         x0[2] + r0[2] - x1[2] - r1[2]};
    const double k = p[7];
    double gp[3], Hp[9];
    if (type == kRbSpring) {  // C = |d| - L0
        const double L = sqrt(dot3(d, d));
        const double n[3] = {d[0] / L, d[1] / L, d[2] / L};
        const double C = L - p[6], a = C / L;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            gp[i] = k * C * n[i];
#pragma unroll
            for (int j = 0; j < 3; ++j) Hp[3 * i + j] = k * ((1.0 - a) * n[i] * n[j] + (i == j ? a : 0.0));
        }
    } else {  // C = |d|^2 - L0^2
        const double C = dot3(d, d) - p[6];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            gp[i] = 2.0 * k * C * d[i];
#pragma unroll
            for (int j = 0; j < 3; ++j) Hp[3 * i + j] = k * (4.0 * d[i] * d[j] + (i == j ? 2.0 * C : 0.0));
        }
    }
    const double r[3] = {slot == 0 ? r0[0] : r1[0], slot == 0 ? r0[1] : r1[1], slot == 0 ? r0[2] : r1[2]};
    addPointTerms<UPPER>(r, gp, Hp, slot == 0 ? 1.0 : -1.0, f, B);
}

// H dx = f by LU with partial pivoting; returns det(H) (0 when a pivot vanishes, dx then unset). H is destroyed.
// The pivot search compares each lower row with the current pivot row and swaps on a strictly larger magnitude,
// so every index is a compile-time constant once the loops unroll and the matrix stays in registers.
RB_HD double solve6(double *H, const double *f, double *dx)
{
    double b[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) b[i] = f[i];
    double det = 1.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
#pragma unroll
        for (int r = c + 1; r < 6; ++r) {
            const bool sw = fabs(H[6 * r + c]) > fabs(H[6 * c + c]);
#pragma unroll
            for (int q = c; q < 6; ++q) {
                const double hc = H[6 * c + q], hr = H[6 * r + q];
                H[6 * c + q] = sw ? hr : hc;
                H[6 * r + q] = sw ? hc : hr;
            }
            const double bc = b[c], br = b[r];
            b[c] = sw ? br : bc;
            b[r] = sw ? bc : br;
            det = sw ? -det : det;
        }
        const double d = H[6 * c + c];
        det *= d;
        if (d == 0) return 0.0;
        const double inv = 1.0 / d;
#pragma unroll
        for (int r = c + 1; r < 6; ++r) {
            const double fct = H[6 * r + c] * inv;
#pragma unroll
            for (int q = c + 1; q < 6; ++q) H[6 * r + q] -= fct * H[6 * c + q];
            b[r] -= fct * b[c];
        }
    }
#pragma unroll
    for (int i = 5; i >= 0; --i) {
        double t = b[i];
#pragma unroll
        for (int q = i + 1; q < 6; ++q) t -= H[6 * i + q] * dx[q];
        dx[i] = t / H[6 * i + i];
    }
    return det;
}

}  // namespace spgrigid
