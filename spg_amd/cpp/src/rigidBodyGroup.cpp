// spg::RigidBodyGroup of the C++ mirror (source/sim/simObject/rigidBodyGroup.cpp:21-243) and the addStencil of the
// four rigid-body energies (rigidBodySpringEnergy.cpp:38-53, rigidBodySpringSquaredEnergy.cpp:35-50,
// rigidBodySpringAnchorEnergy.cpp:63-75, rigidBodyOrientationAlignmentEnergy.cpp:33-40).
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include <spg/sim/energy/rigidBodyEnergies.h>
#include <spg/sim/simObject/rigidBodyGroup.h>

namespace spg {

namespace {

Matrix3 mul(const Matrix3 &a, const Matrix3 &b)
{
    Matrix3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r(i, j) = a(i, 0) * b(0, j) + a(i, 1) * b(1, j) + a(i, 2) * b(2, j);
    return r;
}
Matrix3 mulT(const Matrix3 &a, const Matrix3 &b)  // a b^T
{
    Matrix3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r(i, j) = a(i, 0) * b(j, 0) + a(i, 1) * b(j, 1) + a(i, 2) * b(j, 2);
    return r;
}
Matrix3 inverse(const Matrix3 &m)
{
    const Real c00 = m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1), c01 = m(1, 2) * m(2, 0) - m(1, 0) * m(2, 2),
               c02 = m(1, 0) * m(2, 1) - m(1, 1) * m(2, 0);
    const Real id = 1.0 / (m(0, 0) * c00 + m(0, 1) * c01 + m(0, 2) * c02);
    Matrix3 r;
    r(0, 0) = c00 * id;
    r(0, 1) = (m(0, 2) * m(2, 1) - m(0, 1) * m(2, 2)) * id;
    r(0, 2) = (m(0, 1) * m(1, 2) - m(0, 2) * m(1, 1)) * id;
    r(1, 0) = c01 * id;
    r(1, 1) = (m(0, 0) * m(2, 2) - m(0, 2) * m(2, 0)) * id;
    r(1, 2) = (m(0, 2) * m(1, 0) - m(0, 0) * m(1, 2)) * id;
    r(2, 0) = c02 * id;
    r(2, 1) = (m(0, 1) * m(2, 0) - m(0, 0) * m(2, 1)) * id;
    r(2, 2) = (m(0, 0) * m(1, 1) - m(0, 1) * m(1, 0)) * id;
    return r;
}

}  // namespace

// Rodrigues' formula with the shared (1 - c) a_i a_j products of Eigen's AngleAxis::toRotationMatrix
Matrix3 axisAngleToRotMatrix(const Vector3 &theta)
{
    const Real angle = theta.norm();
    if (angle == 0) return Matrix3::Identity();
    const Vector3 a = {theta[0] / angle, theta[1] / angle, theta[2] / angle};
    const Real c = std::cos(angle), sn = std::sin(angle);
    const Vector3 s = a * sn, c1 = a * (1 - c);
    Matrix3 R;
    for (int i = 0; i < 3; ++i) R(i, i) = c1[i] * a[i] + c;
    Real t = c1[0] * a[1];
    R(0, 1) = t - s[2];
    R(1, 0) = t + s[2];
    t = c1[0] * a[2];
    R(0, 2) = t + s[1];
    R(2, 0) = t - s[1];
    t = c1[1] * a[2];
    R(1, 2) = t - s[0];
    R(2, 1) = t + s[0];
    return R;
}

// Eigen::AngleAxis(R): the unit quaternion of R (trace, or the largest diagonal entry, as pivot), then
// 2 atan2(|v|, |w|) about v / |v| with the sign of w
Vector3 rotMatrixToAxisAngle(const Matrix3 &R)
{
    Real q[4];  // w, x, y, z
    const Real tr = R(0, 0) + R(1, 1) + R(2, 2);
    if (tr > 0) {
        Real s = std::sqrt(tr + 1.0);
        q[0] = 0.5 * s;
        s = 0.5 / s;
        q[1] = (R(2, 1) - R(1, 2)) * s;
        q[2] = (R(0, 2) - R(2, 0)) * s;
        q[3] = (R(1, 0) - R(0, 1)) * s;
    } else {
        int i = 0;
        if (R(1, 1) > R(0, 0)) i = 1;
        if (R(2, 2) > R(i, i)) i = 2;
        const int j = (i + 1) % 3, k = (j + 1) % 3;
        Real s = std::sqrt(R(i, i) - R(j, j) - R(k, k) + 1.0);
        q[1 + i] = 0.5 * s;
        s = 0.5 / s;
        q[0] = (R(k, j) - R(j, k)) * s;
        q[1 + j] = (R(j, i) + R(i, j)) * s;
        q[1 + k] = (R(k, i) + R(i, k)) * s;
    }
    Real n = std::sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (n == 0) return Vector3::Zero();
    const Real angle = 2.0 * std::atan2(n, std::fabs(q[0]));
    if (q[0] < 0) n = -n;
    return {q[1] / n * angle, q[2] / n * angle, q[3] / n * angle};
}

void RigidBodyGroup::getPositions(VectorX &pos, int offsetIndex) const
{
    for (int b = 0; b < size(); ++b)
        for (int d = 0; d < 3; ++d) {
            pos[b * 6 + offsetIndex + d] = m_x[b][d];
            pos[b * 6 + offsetIndex + 3 + d] = m_theta[b][d];
        }
}

void RigidBodyGroup::getVelocities(VectorX &vel, int offsetIndex) const
{
    for (int b = 0; b < size(); ++b)
        for (int d = 0; d < 3; ++d) {
            vel[b * 6 + offsetIndex + d] = m_v[b][d];
            vel[b * 6 + offsetIndex + 3 + d] = m_omega[b][d];
        }
}

void RigidBodyGroup::setPositions(const VectorX &pos, int offsetIndex)
{
    for (int b = 0; b < size(); ++b)
        for (int d = 0; d < 3; ++d) {
            m_x[b][d] = pos[b * 6 + offsetIndex + d];
            m_theta[b][d] = pos[b * 6 + offsetIndex + 3 + d];
        }
    updateRotationMatrices();
    updateInertias();
}

void RigidBodyGroup::setVelocities(const VectorX &vel, int offsetIndex)
{
    for (int b = 0; b < size(); ++b)
        for (int d = 0; d < 3; ++d) {
            m_v[b][d] = vel[b * 6 + offsetIndex + d];
            m_omega[b][d] = vel[b * 6 + offsetIndex + 3 + d];
        }
}

void RigidBodyGroup::addBody(const Vector3 &p, const Vector3 &p0, Real mass, const Vector3 &theta,
                             const Vector3 &theta0, const Matrix3 &localInertia, const TriangleMesh &mesh)
{
    m_x.push_back(p);
    m_xInitial.push_back(p);
    m_x0.push_back(p0);
    m_v.push_back(Vector3::Zero());
    m_m.push_back(mass);
    m_w.push_back(mass == 0 ? 0 : 1 / mass);
    m_theta.push_back(theta);
    m_thetaInitial.push_back(theta);
    m_omega.push_back(Vector3::Zero());
    m_theta0.push_back(theta0);
    m_localInertia.push_back(localInertia);
    m_localInertiaInv.push_back(inverse(localInertia));
    m_rotationMatrix.push_back(axisAngleToRotMatrix(theta));
    m_inertia.emplace_back();
    m_inertiaInv.emplace_back();
    updateInertia(size() - 1);
    m_visualMesh.push_back(mesh);
}

void RigidBodyGroup::addEnergy(std::shared_ptr<EnergyT> energy) { m_energies.push_back(energy); }

void RigidBodyGroup::removeEnergy(std::shared_ptr<EnergyT> energy)
{
    m_energies.erase(std::remove(m_energies.begin(), m_energies.end(), energy), m_energies.end());
}

void RigidBodyGroup::reset()
{
    m_x = m_xInitial;
    for (auto &v : m_v) v.setZero();
    m_theta = m_thetaInitial;
    updateRotationMatrices();
    updateInertias();
    for (auto &w : m_omega) w.setZero();
}

void RigidBodyGroup::scaleMasses(Real scaleFactor)
{
    for (int i = 0; i < size(); ++i) {
        if (m_m[i] != 0) {
            m_m[i] *= scaleFactor;
            m_w[i] /= scaleFactor;
            for (auto &e : m_localInertia[i].m) e *= scaleFactor;
            for (auto &e : m_localInertiaInv[i].m) e /= scaleFactor;
        }
    }
}

Vector3 RigidBodyGroup::globalToLocalRigidBodyFrame(const Vector3 &globalPos, int bodyId) const
{
    const Vector3 d = globalPos - m_x[bodyId];
    const Matrix3 &R = m_rotationMatrix[bodyId];
    return {R(0, 0) * d[0] + R(1, 0) * d[1] + R(2, 0) * d[2], R(0, 1) * d[0] + R(1, 1) * d[1] + R(2, 1) * d[2],
            R(0, 2) * d[0] + R(1, 2) * d[1] + R(2, 2) * d[2]};
}

void RigidBodyGroup::updateInertia(int i)
{
    m_inertia[i] = mulT(mul(m_rotationMatrix[i], m_localInertia[i]), m_rotationMatrix[i]);
    m_inertiaInv[i] = mulT(mul(m_rotationMatrix[i], m_localInertiaInv[i]), m_rotationMatrix[i]);
}

void RigidBodyGroup::updateInertias()
{
    for (int i = 0; i < size(); ++i) updateInertia(i);
}

void RigidBodyGroup::updateRotationMatrix(int i) { m_rotationMatrix[i] = axisAngleToRotMatrix(m_theta[i]); }

void RigidBodyGroup::updateRotationMatrices()
{
    for (int i = 0; i < size(); ++i) updateRotationMatrix(i);
}

void RigidBodyGroup::updateTheta(int i) { m_theta[i] = rotMatrixToAxisAngle(m_rotationMatrix[i]); }

void RigidBodyGroup::updateThetas()
{
    for (int i = 0; i < size(); ++i) updateTheta(i);
}

// ---- energies ----

void RBSpringEnergy::addStencil(std::array<int, 2> stencil, const std::array<Vector3, 2> &localRigidBodyPoints,
                                const Real restLength, const Real stiffness)
{
    m_stencils.push_back(stencil);
    if (restLength <= 0) {
        throw std::runtime_error("Zero/negative rest length invalid");
    }
    m_restLength.push_back(restLength);
    m_localRBPoints.push_back(localRigidBodyPoints);
    pushStiffness(stiffness);
}

void RBSpringSquaredEnergy::addStencil(std::array<int, 2> stencil,
                                       const std::array<Vector3, 2> &localRigidBodyPoints, const Real restLength,
                                       const Real stiffness)
{
    m_stencils.push_back(stencil);
    if (restLength < 0) {
        throw std::runtime_error("Negative rest length invalid");
    }
    m_squaredRestLength.push_back(restLength * restLength);
    m_restLengthArg.push_back(restLength);
    m_localRBPoints.push_back(localRigidBodyPoints);
    pushStiffness(stiffness);
}

void RBSpringAnchorEnergy::addStencil(std::array<int, 1> stencil, const Vector3 &localRigidBodyPoint,
                                      const Vector3 &anchor, const Real stiffness)
{
    m_stencils.push_back(stencil);
    m_localRBPoint.push_back(localRigidBodyPoint);
    m_anchor.push_back(anchor);
    pushStiffness(stiffness);
}

void RBOrientationAlignmentEnergy::addStencil(std::array<int, 2> stencil, const Real stiffness)
{
    m_stencils.push_back(stencil);
    pushStiffness(stiffness);
}

}  // namespace spg
