// Mirror of spg::RigidBodyGroup (spg/include/spg/sim/simObject/rigidBodyGroup.h:11-86,
// source/sim/simObject/rigidBodyGroup.cpp:21-243): bodies with centre of mass, axis-angle orientation, rotation
// matrix, linear and angular velocity, mass and inertia, and their energies. Same accessors and semantics; the
// VBD step on these bodies runs on the GPU (spg::solver::VBDGpu through include/spg_rigid.h).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include <spg/geom/triangleMesh.h>
#include <spg/types.h>

namespace spg {

template <class TSimObject>
class Energy;

class RigidBodyGroup
{
public:
    static constexpr int s_nDOFs = 6;
    using EnergyT = Energy<RigidBodyGroup>;

    const std::vector<Vector3> &positions() const { return m_x; }
    std::vector<Vector3> &positions() { return m_x; }
    const std::vector<Vector3> &positions0() const { return m_x0; }
    const std::vector<Vector3> &thetas() const { return m_theta; }
    std::vector<Vector3> &thetas() { return m_theta; }
    const std::vector<Vector3> &thetas0() const { return m_theta0; }
    std::vector<Vector3> &velocities() { return m_v; }
    const std::vector<Vector3> &velocities() const { return m_v; }
    // positions are (x, theta) and velocities (v, omega), 6 per body from offsetIndex (rigidBodyGroup.cpp:21-62);
    // setPositions rebuilds the rotation matrices and inertias from theta
    void getPositions(VectorX &pos, int offsetIndex) const;
    void getVelocities(VectorX &vel, int offsetIndex) const;
    void setPositions(const VectorX &pos, int offsetIndex);
    void setVelocities(const VectorX &vel, int offsetIndex);
    std::vector<Vector3> &omegas() { return m_omega; }
    const std::vector<Vector3> &omegas() const { return m_omega; }
    const std::vector<Real> &invMasses() const { return m_w; }
    const std::vector<Real> &masses() const { return m_m; }
    const std::vector<Matrix3> &invInertias() const { return m_inertiaInv; }
    const std::vector<Matrix3> &inertias() const { return m_inertia; }
    // local (body-frame) inertia tensors as passed to addBody; the reference keeps them in m_localInertia
    const std::vector<Matrix3> &localInertias() const { return m_localInertia; }
    std::vector<Matrix3> &rotationMatrices() { return m_rotationMatrix; }
    const std::vector<Matrix3> &rotationMatrices() const { return m_rotationMatrix; }
    const std::vector<TriangleMesh> &visualMeshes() const { return m_visualMesh; }
    std::vector<std::shared_ptr<EnergyT>> &energies() { return m_energies; }
    const std::vector<std::shared_ptr<EnergyT>> &energies() const { return m_energies; }
    int nDOF() const { return static_cast<int>(m_x.size() * 3 + m_theta.size() * 3); }
    int size() const { return static_cast<int>(m_x.size()); }

    void addBody(const Vector3 &p, const Vector3 &p0, Real mass, const Vector3 &theta, const Vector3 &theta0,
                 const Matrix3 &localInertia, const TriangleMesh &mesh);
    void addEnergy(std::shared_ptr<EnergyT> energy);
    void removeEnergy(std::shared_ptr<EnergyT> energy);
    void reset();
    void scaleMasses(Real scaleFactor);
    void setName(std::string name) { m_name = std::move(name); }
    void updateInertia(int i);
    void updateInertias();
    void updateRotationMatrix(int i);
    void updateRotationMatrices();
    void updateTheta(int i);
    void updateThetas();

    Vector3 globalToLocalRigidBodyFrame(const Vector3 &globalPos, int bodyId) const;

protected:

    std::string m_name;
    std::vector<Vector3> m_x, m_xInitial, m_x0, m_v;
    std::vector<Vector3> m_theta, m_thetaInitial, m_theta0, m_omega;
    std::vector<Real> m_m, m_w;
    std::vector<Matrix3> m_localInertia, m_localInertiaInv, m_inertia, m_inertiaInv, m_rotationMatrix;
    std::vector<TriangleMesh> m_visualMesh;
    std::vector<std::shared_ptr<EnergyT>> m_energies;
};

// axisAngleToRotMatrix / rotMatrixToAxisAngle of rigidBodyGroup.cpp:6-16 (Eigen::AngleAxis arithmetic)
Matrix3 axisAngleToRotMatrix(const Vector3 &theta);
Vector3 rotMatrixToAxisAngle(const Matrix3 &R);

}  // namespace spg
