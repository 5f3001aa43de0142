// addStencil / preparePrecomputations of the five VBD energies and the mesh edge enumeration.
// The rest-state preparation is FP64, with the formulas and the validation (and messages) of:
//   springEnergy.cpp:33-46, membraneStvkEnergy.cpp:31-71, discreteBendingEnergy.cpp:33-72,
//   stableNeoHookeanEnergy.cpp:64-108, springAnchorEnergy.cpp:29-39, triangleMesh.h:16-50,
//   tetrahedralMesh.h:12-51.
#include <cmath>

#include <spg/geom/triangleMesh.h>
#include <spg/sim/energy/energies.h>

namespace spg {

namespace {
template <int N>
Matrix<N, N> scalarMat(Real s)
{
    Matrix<N, N> m{};
    m(0, 0) = s;
    return m;
}
Matrix1 inverse1(const Matrix1 &a) { return scalarMat<1>(1.0 / a(0, 0)); }

Matrix2 inverse2(const Matrix2 &a)
{
    const Real det = a(0, 0) * a(1, 1) - a(0, 1) * a(1, 0);
    Matrix2 r;
    r(0, 0) = a(1, 1) / det;
    r(0, 1) = -a(0, 1) / det;
    r(1, 0) = -a(1, 0) / det;
    r(1, 1) = a(0, 0) / det;
    return r;
}

Real det3(const Matrix3 &a)
{
    return a(0, 0) * (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1)) + a(0, 1) * (a(1, 2) * a(2, 0) - a(1, 0) * a(2, 2)) +
           a(0, 2) * (a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0));
}

Matrix3 inverse3(const Matrix3 &a)
{
    const Real det = det3(a);
    Matrix3 r;
    r(0, 0) = (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1)) / det;
    r(0, 1) = (a(0, 2) * a(2, 1) - a(0, 1) * a(2, 2)) / det;
    r(0, 2) = (a(0, 1) * a(1, 2) - a(0, 2) * a(1, 1)) / det;
    r(1, 0) = (a(1, 2) * a(2, 0) - a(1, 0) * a(2, 2)) / det;
    r(1, 1) = (a(0, 0) * a(2, 2) - a(0, 2) * a(2, 0)) / det;
    r(1, 2) = (a(0, 2) * a(1, 0) - a(0, 0) * a(1, 2)) / det;
    r(2, 0) = (a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0)) / det;
    r(2, 1) = (a(0, 1) * a(2, 0) - a(0, 0) * a(2, 1)) / det;
    r(2, 2) = (a(0, 0) * a(1, 1) - a(0, 1) * a(1, 0)) / det;
    return r;
}

template <int N>
Matrix<N, N> scaled(const Matrix<N, N> &a, Real s)
{
    Matrix<N, N> r = a;
    for (auto &v : r.m) v *= s;
    return r;
}
template <int N>
Matrix<N, N> divided(const Matrix<N, N> &a, Real s)
{
    Matrix<N, N> r = a;
    for (auto &v : r.m) v /= s;
    return r;
}

// signed hinge angle of discreteBendingEnergy.cpp:20-27
Real hingeAngle(const Vector3 &x0, const Vector3 &x1, const Vector3 &x2, const Vector3 &x3)
{
    const Vector3 e0 = x1 - x0, e3 = x2 - x1, e4 = x3 - x1;
    const Vector3 n1 = e0.cross(e3).normalized();
    const Vector3 n2 = e0.cross(e4).normalized() * -1.0;
    return std::atan2(n1.cross(n2).dot(e0.normalized()), n1.dot(n2));
}
}  // namespace

void SpringEnergy::addStencil(const std::array<int, s_stencilSize> &stencil, const Real restLength, const Real stiffness)
{
    m_stencils.push_back(stencil);
    if (restLength <= 0) {
        throw std::runtime_error("Zero/negative rest length invalid");
    }
    m_restLength.push_back(restLength);
    m_modelStiffness.push_back(scalarMat<1>(stiffness));
    m_effectiveStiffness.push_back(scalarMat<1>(stiffness));
    m_modelCompliance.push_back(inverse1(m_modelStiffness.back()));
    m_effectiveCompliance.push_back(inverse1(m_effectiveStiffness.back()));
}

void MembraneStvkEnergy::addStencil(const std::array<int, s_stencilSize> &stencil, const Real young, const Real poisson)
{
    m_stencils.push_back(stencil);
    const Real youngCommon = young / (1 - poisson * poisson);
    Matrix3 C{};
    C(0, 0) = youngCommon;
    C(0, 1) = youngCommon * poisson;
    C(1, 0) = youngCommon * poisson;
    C(1, 1) = youngCommon;
    C(2, 2) = youngCommon * (1 - poisson) / 2;
    m_modelStiffness.push_back(C);
    m_effectiveStiffness.push_back(C);
    m_modelCompliance.push_back(inverse3(C));
    m_effectiveCompliance.push_back(inverse3(C));
}

void MembraneStvkEnergy::preparePrecomputations(const ParticleGroup &pGroup)
{
    const int nstencils = static_cast<int>(m_stencils.size());
    m_inverseReferenceMat.resize(nstencils);
    for (int i = 0; i < nstencils; ++i) {
        const auto &x0 = pGroup.positions0()[m_stencils[i][0]];
        const auto &x1 = pGroup.positions0()[m_stencils[i][1]];
        const auto &x2 = pGroup.positions0()[m_stencils[i][2]];
        if (x0.z() != 0.0 || x1.z() != 0.0 || x2.z() != 0.0) {
            throw std::runtime_error("Rest positions of triangles for membrane StVK energy must have coordinate z=0");
        }
        const Real u0 = x1[0] - x0[0], u1 = x1[1] - x0[1];
        const Real v0 = x2[0] - x0[0], v1 = x2[1] - x0[1];
        Matrix2 Dm;
        Dm(0, 0) = u0;
        Dm(1, 0) = u1;
        Dm(0, 1) = v0;
        Dm(1, 1) = v1;
        m_inverseReferenceMat[i] = inverse2(Dm);
        const Real area = (u0 * v1 - v0 * u1) * 0.5;
        if (area <= 0) {
            throw std::runtime_error("Rest area of membrane StVK triangle is not positive");
        }
        m_effectiveCompliance[i] = divided(m_modelCompliance[i], area);
        m_effectiveStiffness[i] = scaled(m_modelStiffness[i], area);
    }
    StencilBlockEnergy<3, 3>::preparePrecomputations(pGroup);
}

void DiscreteBendingEnergy::addStencil(const std::array<int, s_stencilSize> &stencil, const Real restTheta,
                                       const Real stiffness)
{
    m_stencils.push_back(stencil);
    if (restTheta < -3.12 || restTheta > 3.12) {
        throw std::runtime_error("Invalid restTheta");
    }
    m_restTheta.push_back(restTheta);
    m_modelStiffness.push_back(scalarMat<1>(stiffness));
    m_effectiveStiffness.push_back(scalarMat<1>(stiffness));
    m_modelCompliance.push_back(inverse1(m_modelStiffness.back()));
    m_effectiveCompliance.push_back(inverse1(m_effectiveStiffness.back()));
}

void DiscreteBendingEnergy::preparePrecomputations(const ParticleGroup &pGroup)
{
    const int nstencils = static_cast<int>(m_stencils.size());
    for (int i = 0; i < nstencils; ++i) {
        const auto &x0 = pGroup.positions0()[m_stencils[i][0]];
        const auto &x1 = pGroup.positions0()[m_stencils[i][1]];
        const auto &x2 = pGroup.positions0()[m_stencils[i][2]];
        const auto &x3 = pGroup.positions0()[m_stencils[i][3]];
        const Vector3 e0 = x1 - x0, e3 = x2 - x1, e4 = x3 - x1;
        // 3 |e0|^2 / A, times 2 for the 1/2 of the energy (discreteBendingEnergy.cpp:60-64)
        const Real constantFactor = 2. * 3. * e0.squaredNorm() / (0.5 * (e0.cross(e3).norm() + e0.cross(e4).norm()));
        m_effectiveCompliance[i] = divided(m_modelCompliance[i], constantFactor);
        m_effectiveStiffness[i] = inverse1(m_effectiveCompliance[i]);
        // the rest angle always comes from the rest positions (discreteBendingEnergy.cpp:66-69)
        m_restTheta[i] = hingeAngle(x0, x1, x2, x3);
    }
    StencilBlockEnergy<4>::preparePrecomputations(pGroup);
}

void StableNeoHookeanEnergy::addStencil(const std::array<int, s_stencilSize> &stencil, const Real young,
                                        const Real poisson)
{
    m_stencils.push_back(stencil);
    // Smith et al. 2018 sec. 3.4 remap: mu = mu_lame, lambda = mu_lame + lambda_lame
    const Real mu = young / (2 * (1 + poisson));
    const Real lambda = mu / (1 - 2 * poisson);
    Matrix2 C{};
    C(0, 0) = lambda;
    C(1, 1) = mu;
    m_modelStiffness.push_back(C);
    m_effectiveStiffness.push_back(C);
    m_modelCompliance.push_back(inverse2(C));
    m_effectiveCompliance.push_back(inverse2(C));
    m_hydrostaticCorrectionTerm.push_back(1 + mu / lambda);
}

void StableNeoHookeanEnergy::preparePrecomputations(const ParticleGroup &pGroup)
{
    const int nstencils = static_cast<int>(m_stencils.size());
    m_inverseReferenceMat.resize(nstencils);
    for (int i = 0; i < nstencils; ++i) {
        const auto &x0 = pGroup.positions0()[m_stencils[i][0]];
        const Vector3 u = pGroup.positions0()[m_stencils[i][1]] - x0;
        const Vector3 v = pGroup.positions0()[m_stencils[i][2]] - x0;
        const Vector3 w = pGroup.positions0()[m_stencils[i][3]] - x0;
        Matrix3 Dm;
        for (int r = 0; r < 3; ++r) {
            Dm(r, 0) = u[r];
            Dm(r, 1) = v[r];
            Dm(r, 2) = w[r];
        }
        const Real volume = 1. / 6. * det3(Dm);
        if (volume <= 0) {
            throw std::runtime_error("SNH tetrahedron rest volume is not positive");
        }
        m_inverseReferenceMat[i] = inverse3(Dm);
        m_effectiveCompliance[i] = divided(m_modelCompliance[i], volume);
        m_effectiveStiffness[i] = scaled(m_modelStiffness[i], volume);
    }
    StencilBlockEnergy<4, 2>::preparePrecomputations(pGroup);
}

void SpringAnchorEnergy::addStencil(std::array<int, s_stencilSize> stencil, const Vector3 &anchor, const Real stiffness)
{
    m_stencils.push_back(stencil);
    m_modelStiffness.push_back(scalarMat<1>(stiffness));
    m_effectiveStiffness.push_back(scalarMat<1>(stiffness));
    m_modelCompliance.push_back(inverse1(m_modelStiffness.back()));
    m_effectiveCompliance.push_back(inverse1(m_effectiveStiffness.back()));
    m_anchor.push_back(anchor);
}

namespace {
// inverse of a diagonal N x N model stiffness, entry by entry
template <int N>
Matrix<N, N> inverseDiagonal(const Matrix<N, N> &a)
{
    Matrix<N, N> r{};
    for (int i = 0; i < N; ++i) r(i, i) = 1.0 / a(i, i);
    return r;
}
template <int N>
Matrix<N, N> diagonal(const std::array<Real, N> &d)
{
    Matrix<N, N> r{};
    for (int i = 0; i < N; ++i) r(i, i) = d[i];
    return r;
}
// F's rest matrix of a triangle in the z = 0 plane: inverse, and the signed area
Matrix2 restTriangleInverse(const Vector3 &x0, const Vector3 &x1, const Vector3 &x2, Real &area)
{
    const Real u0 = x1[0] - x0[0], u1 = x1[1] - x0[1];
    const Real v0 = x2[0] - x0[0], v1 = x2[1] - x0[1];
    Matrix2 Dm;
    Dm(0, 0) = u0;
    Dm(1, 0) = u1;
    Dm(0, 1) = v0;
    Dm(1, 1) = v1;
    area = (u0 * v1 - v0 * u1) * 0.5;
    return inverse2(Dm);
}
// F's rest matrix of a tetrahedron: inverse, and the signed volume
Matrix3 restTetInverse(const ParticleGroup &pGroup, const std::array<int, 4> &st, Real &volume)
{
    const auto &x0 = pGroup.positions0()[st[0]];
    const Vector3 u = pGroup.positions0()[st[1]] - x0;
    const Vector3 v = pGroup.positions0()[st[2]] - x0;
    const Vector3 w = pGroup.positions0()[st[3]] - x0;
    Matrix3 Dm;
    for (int r = 0; r < 3; ++r) {
        Dm(r, 0) = u[r];
        Dm(r, 1) = v[r];
        Dm(r, 2) = w[r];
    }
    volume = 1. / 6. * det3(Dm);
    return inverse3(Dm);
}
// Lame parameters of stvkEnergy.cpp:69-73 (lambda + 1e-20 keeps the compliance finite at nu = 0)
Matrix2 lameStiffness(Real young, Real poisson)
{
    const Real mu = young / (2 * (1 + poisson));
    constexpr Real epsilon = 1e-20;
    const Real lambda = young * poisson / ((1 + poisson) * (1 - 2 * poisson)) + epsilon;
    return diagonal<2>({lambda, mu});
}
}  // namespace

void SpringContinuumEnergy::addStencil(const std::array<int, s_stencilSize> &stencil, const Real restLength,
                                       const Real stiffness)
{
    m_stencils.push_back(stencil);
    if (restLength <= 0) {
        throw std::runtime_error("Zero/negative rest length invalid");
    }
    m_restLength.push_back(restLength);
    m_modelStiffness.push_back(scalarMat<1>(stiffness));
    // the stiffness integrated along the element
    m_effectiveStiffness.push_back(scalarMat<1>(stiffness * restLength));
    m_modelCompliance.push_back(inverse1(m_modelStiffness.back()));
    m_effectiveCompliance.push_back(inverse1(m_effectiveStiffness.back()));
}

void SpringSquaredContinuumEnergy::addStencil(const std::array<int, s_stencilSize> &stencil, const Real restLength,
                                              const Real stiffness)
{
    m_stencils.push_back(stencil);
    if (restLength <= 0) {
        throw std::runtime_error("Zero/negative rest length invalid");
    }
    m_restLength.push_back(restLength);
    m_modelStiffness.push_back(scalarMat<1>(stiffness));
    m_effectiveStiffness.push_back(scalarMat<1>(stiffness * restLength));
    m_modelCompliance.push_back(inverse1(m_modelStiffness.back()));
    m_effectiveCompliance.push_back(inverse1(m_effectiveStiffness.back()));
}

void MembraneBaraffWitkinEnergy::addStencil(const std::array<int, s_stencilSize> &stencil, const Real kWeft,
                                            const Real kWarp, const Real kBias)
{
    m_stencils.push_back(stencil);
    const Matrix3 C = diagonal<3>({kWeft, kWarp, kBias});
    m_modelStiffness.push_back(C);
    m_effectiveStiffness.push_back(C);
    m_modelCompliance.push_back(inverseDiagonal(C));
    m_effectiveCompliance.push_back(inverseDiagonal(C));
}

void MembraneBaraffWitkinEnergy::preparePrecomputations(const ParticleGroup &pGroup)
{
    const int nstencils = static_cast<int>(m_stencils.size());
    m_inverseReferenceMat.resize(nstencils);
    for (int i = 0; i < nstencils; ++i) {
        const auto &x0 = pGroup.positions0()[m_stencils[i][0]];
        const auto &x1 = pGroup.positions0()[m_stencils[i][1]];
        const auto &x2 = pGroup.positions0()[m_stencils[i][2]];
        if (x0.z() != 0.0 || x1.z() != 0.0 || x2.z() != 0.0) {
            throw std::runtime_error(
                "Rest positions of triangles for membrane BaraffWitkin energy must have coordinate z=0");
        }
        Real area = 0;
        m_inverseReferenceMat[i] = restTriangleInverse(x0, x1, x2, area);
        if (area <= 0) {
            throw std::runtime_error("Rest area of membrane BaraffWitkin triangle is not positive");
        }
        m_effectiveCompliance[i] = divided(m_modelCompliance[i], area);
        m_effectiveStiffness[i] = scaled(m_modelStiffness[i], area);
    }
    StencilBlockEnergy<3, 3>::preparePrecomputations(pGroup);
}

void MembraneChoiEnergy::addStencil(const std::array<int, s_stencilSize> &stencil, const Real kWeft,
                                    const Real kWarp, const Real kBias)
{
    m_stencils.push_back(stencil);
    const Matrix<4, 4> C = diagonal<4>({kWeft, kWarp, kBias, kBias});
    m_modelStiffness.push_back(C);
    m_effectiveStiffness.push_back(C);
    m_modelCompliance.push_back(inverseDiagonal(C));
    m_effectiveCompliance.push_back(inverseDiagonal(C));
}

void MembraneChoiEnergy::preparePrecomputations(const ParticleGroup &pGroup)
{
    const int nstencils = static_cast<int>(m_stencils.size());
    m_inverseReferenceMat.resize(nstencils);
    for (int i = 0; i < nstencils; ++i) {
        const auto &x0 = pGroup.positions0()[m_stencils[i][0]];
        const auto &x1 = pGroup.positions0()[m_stencils[i][1]];
        const auto &x2 = pGroup.positions0()[m_stencils[i][2]];
        if (x0.z() != 0.0 || x1.z() != 0.0 || x2.z() != 0.0) {
            throw std::runtime_error("Rest positions of triangles for membrane Choi energy must have coordinate z=0");
        }
        Real area = 0;
        m_inverseReferenceMat[i] = restTriangleInverse(x0, x1, x2, area);
        if (area <= 0) {
            throw std::runtime_error("Rest area of membrane Choi triangle is not positive");
        }
        m_effectiveCompliance[i] = divided(m_modelCompliance[i], area);
        m_effectiveStiffness[i] = scaled(m_modelStiffness[i], area);
    }
    StencilBlockEnergy<3, 4>::preparePrecomputations(pGroup);
}

void QuadraticBendingEnergy::addStencil(const std::array<int, s_stencilSize> &stencil, const Real restTheta,
                                        const Real stiffness)
{
    if (restTheta != 0) {
        throw std::runtime_error("Rest theta != 0 not supported for quadratic bending");
    }
    m_stencils.push_back(stencil);
    m_laplacians.push_back(Matrix<3, 12>{});
    m_stiffnessFactor.push_back(stiffness);
    m_modelCompliance.push_back(StiffnessMat{});
    m_effectiveCompliance.push_back(StiffnessMat{});
    m_modelStiffness.push_back(StiffnessMat{});
    m_effectiveStiffness.push_back(StiffnessMat{});
}

void QuadraticBendingEnergy::preparePrecomputations(const ParticleGroup &pGroup)
{
    // Bergou et al. 2006: cotangent weights of the hinge's two wings, keff = k * 3 / (wing areas)
    auto cotangent = [](const Vector3 &a, const Vector3 &b) { return a.dot(b) / a.cross(b).norm(); };
    const int nstencils = static_cast<int>(m_stencils.size());
    for (int i = 0; i < nstencils; ++i) {
        const auto &x0 = pGroup.positions0()[m_stencils[i][0]];
        const auto &x1 = pGroup.positions0()[m_stencils[i][1]];
        const auto &x2 = pGroup.positions0()[m_stencils[i][2]];
        const auto &x3 = pGroup.positions0()[m_stencils[i][3]];
        const Vector3 e0 = x1 - x0, e1 = x2 - x0, e2 = x3 - x0, e3 = x2 - x1, e4 = x3 - x1;
        const Vector3 me0 = e0 * -1.0;
        const Real c01 = cotangent(e0, e1), c02 = cotangent(e0, e2), c03 = cotangent(me0, e3),
                   c04 = cotangent(me0, e4);
        const Real c[4] = {c03 + c04, c01 + c02, -c01 - c03, -c02 - c04};
        Matrix<3, 12> L{};
        for (int s = 0; s < 4; ++s)
            for (int d = 0; d < 3; ++d) L(d, 3 * s + d) = c[s];
        m_laplacians[i] = L;
        const Real keff = m_stiffnessFactor[i] * (3 / (0.5 * (e0.cross(e1).norm() + e0.cross(e2).norm())));
        m_effectiveStiffness[i] = diagonal<3>({keff, keff, keff});
        m_modelStiffness[i] = m_effectiveStiffness[i];
        m_effectiveCompliance[i] = inverseDiagonal(m_effectiveStiffness[i]);
        m_modelCompliance[i] = m_effectiveCompliance[i];
    }
    StencilBlockEnergy<4, 3>::preparePrecomputations(pGroup);
}

void BaraffWitkinBendingEnergy::addStencil(const std::array<int, s_stencilSize> &stencil, const Real restTheta,
                                           const Real stiffness)
{
    m_stencils.push_back(stencil);
    if (restTheta < -3.12 || restTheta > 3.12) {
        throw std::runtime_error("Invalid restTheta");
    }
    m_restTheta.push_back(restTheta);
    m_modelStiffness.push_back(scalarMat<1>(stiffness));
    m_effectiveStiffness.push_back(scalarMat<1>(stiffness));
    m_modelCompliance.push_back(inverse1(m_modelStiffness.back()));
    m_effectiveCompliance.push_back(inverse1(m_effectiveStiffness.back()));
}

void BaraffWitkinBendingEnergy::preparePrecomputations(const ParticleGroup &pGroup)
{
    // no area weighting (the paper bends by curvature); the rest angle comes from the rest positions
    const int nstencils = static_cast<int>(m_stencils.size());
    for (int i = 0; i < nstencils; ++i)
        m_restTheta[i] = hingeAngle(pGroup.positions0()[m_stencils[i][0]], pGroup.positions0()[m_stencils[i][1]],
                                    pGroup.positions0()[m_stencils[i][2]], pGroup.positions0()[m_stencils[i][3]]);
    StencilBlockEnergy<4>::preparePrecomputations(pGroup);
}

void StvkEnergy::addStencil(const std::array<int, s_stencilSize> &stencil, const Real young, const Real poisson)
{
    m_stencils.push_back(stencil);
    const Matrix2 C = lameStiffness(young, poisson);
    m_modelStiffness.push_back(C);
    m_effectiveStiffness.push_back(C);
    m_modelCompliance.push_back(inverse2(C));
    m_effectiveCompliance.push_back(inverse2(C));
}

void StvkEnergy::preparePrecomputations(const ParticleGroup &pGroup)
{
    const int nstencils = static_cast<int>(m_stencils.size());
    m_inverseReferenceMat.resize(nstencils);
    for (int i = 0; i < nstencils; ++i) {
        Real volume = 0;
        m_inverseReferenceMat[i] = restTetInverse(pGroup, m_stencils[i], volume);
        if (volume <= 0) {
            throw std::runtime_error("SNH tetrahedron rest volume is not positive");  // the reference's text
        }
        m_effectiveCompliance[i] = divided(m_modelCompliance[i], volume);
        m_effectiveStiffness[i] = scaled(m_modelStiffness[i], volume);
    }
    StencilBlockEnergy<4, 2>::preparePrecomputations(pGroup);
}

void StableSquaredNeoHookeanEnergy::addStencil(const std::array<int, s_stencilSize> &stencil, const Real young,
                                               const Real poisson)
{
    m_stencils.push_back(stencil);
    const Matrix2 C = lameStiffness(young, poisson);
    m_modelStiffness.push_back(C);
    m_effectiveStiffness.push_back(C);
    m_modelCompliance.push_back(inverse2(C));
    m_effectiveCompliance.push_back(inverse2(C));
}

void StableSquaredNeoHookeanEnergy::preparePrecomputations(const ParticleGroup &pGroup)
{
    const int nstencils = static_cast<int>(m_stencils.size());
    m_inverseReferenceMat.resize(nstencils);
    for (int i = 0; i < nstencils; ++i) {
        Real volume = 0;
        m_inverseReferenceMat[i] = restTetInverse(pGroup, m_stencils[i], volume);
        if (volume <= 0) {
            throw std::runtime_error("SSNH tetrahedron rest volume is not positive");
        }
        m_effectiveCompliance[i] = divided(m_modelCompliance[i], volume);
        m_effectiveStiffness[i] = scaled(m_modelStiffness[i], volume);
    }
    StencilBlockEnergy<4, 2>::preparePrecomputations(pGroup);
}

TriangleMesh::TriangleMesh(const std::vector<Vector3> &vertices, const std::vector<Vector3> &vertices0,
                           const std::vector<Int3> &faces)
    : m_vertices(vertices), m_vertices0(vertices0), m_faces(faces)
{
    const int nVertices = static_cast<int>(m_vertices.size());
    for (const auto &face : m_faces)
        for (int idx : face)
            if (idx < 0 || idx >= nVertices) throw std::runtime_error("Triangle mesh face contains index to invalid vertex");
    for (const auto &face : m_faces)
        if (face[0] == face[1] || face[1] == face[2] || face[2] == face[0])
            throw std::runtime_error("Triangle mesh face contains repeated vertex index");
    std::map<Int2, std::vector<OpposingVertex>> edges;
    for (int f = 0; f < static_cast<int>(m_faces.size()); ++f) {
        const auto &face = m_faces[f];
        for (int i = 0; i < 3; ++i) {
            const int id0 = face[i], id1 = face[(i + 1) % 3], id2 = face[(i + 2) % 3];
            edges[{std::min(id0, id1), std::max(id0, id1)}].push_back({f, id2});
        }
    }
    for (const auto &kv : edges) {
        m_edges.push_back(kv.first);
        m_opposingVerticesPerEdge.push_back(kv.second);
    }
}

std::vector<Int4> TriangleMesh::bendingStencils() const
{
    std::vector<Int4> out;
    for (size_t i = 0; i < m_edges.size(); ++i)
        if (m_opposingVerticesPerEdge[i].size() == 2)
            out.push_back({m_edges[i][0], m_edges[i][1], m_opposingVerticesPerEdge[i][0].vertexId,
                           m_opposingVerticesPerEdge[i][1].vertexId});
    return out;
}

TetrahedralMesh::TetrahedralMesh(const std::vector<Vector3> &vertices, const std::vector<Vector3> &vertices0,
                                 const std::vector<Int4> &tets)
    : m_vertices(vertices), m_vertices0(vertices0), m_tets(tets)
{
    const int nVertices = static_cast<int>(m_vertices.size());
    for (const auto &tet : m_tets)
        for (int idx : tet)
            if (idx < 0 || idx >= nVertices)
                throw std::runtime_error("Tetrahedral mesh cell contains index to invalid vertex");
    for (const auto &t : m_tets)
        if (t[0] == t[1] || t[0] == t[2] || t[0] == t[3] || t[1] == t[2] || t[1] == t[3] || t[2] == t[3])
            throw std::runtime_error("Tetrahedral mesh cell contains repeated vertex index");
    std::set<Int2> edges;
    for (const auto &tet : m_tets)
        for (int i = 0; i < 3; ++i)
            for (int j = i + 1; j < 4; ++j) edges.insert({std::min(tet[i], tet[j]), std::max(tet[i], tet[j])});
    m_edges.assign(edges.begin(), edges.end());
}

}  // namespace spg
