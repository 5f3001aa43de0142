// Mirror of spg's four rigid-body energies (spg/include/spg/sim/energy/rigidBodySpringEnergy.h,
// rigidBodySpringSquaredEnergy.h, rigidBodySpringAnchorEnergy.h, rigidBodyOrientationAlignmentEnergy.h): stencil
// storage, addStencil validation and the accessors the GPU solver reads. Their derivatives are evaluated on the GPU
// in closed form (spg_amd/csrc/rigid_element.h), not by TinyAD.
#pragma once

#include <array>
#include <limits>
#include <vector>

#include <spg/sim/energy/energy.h>
#include <spg/sim/simObject/rigidBodyGroup.h>

namespace spg {

template <int TstencilSize>
class RBEnergyBase : public StencilBlockEnergy<TstencilSize, 1, RigidBodyGroup>
{
protected:
    // 1x1 model/effective stiffness and their inverses, as every rigid-body addStencil stores them
    void pushStiffness(Real k)
    {
        Matrix1 s;
        s(0, 0) = k;
        Matrix1 c;
        c(0, 0) = k == 0 ? std::numeric_limits<Real>::infinity() : 1.0 / k;
        this->m_modelStiffness.push_back(s);
        this->m_effectiveStiffness.push_back(s);
        this->m_modelCompliance.push_back(c);
        this->m_effectiveCompliance.push_back(c);
    }
};

class RBSpringEnergy : public RBEnergyBase<2>
{
public:
    // throws "Zero/negative rest length invalid" (rigidBodySpringEnergy.cpp:38-53)
    void addStencil(std::array<int, 2> stencil, const std::array<Vector3, 2> &localRigidBodyPoints, Real restLength,
                    Real stiffness = 0);
    const std::vector<std::array<Vector3, 2>> &localRBPoints() const { return m_localRBPoints; }
    const std::vector<Real> &restLengths() const { return m_restLength; }

protected:
    std::vector<std::array<Vector3, 2>> m_localRBPoints;
    std::vector<Real> m_restLength;
};

class RBSpringSquaredEnergy : public RBEnergyBase<2>
{
public:
    // throws "Negative rest length invalid" (rigidBodySpringSquaredEnergy.cpp:35-50)
    void addStencil(std::array<int, 2> stencil, const std::array<Vector3, 2> &localRigidBodyPoints, Real restLength,
                    Real stiffness = 0);
    const std::vector<std::array<Vector3, 2>> &localRBPoints() const { return m_localRBPoints; }
    const std::vector<Real> &squaredRestLengths() const { return m_squaredRestLength; }
    // the addStencil rest lengths, kept so the solver receives the reference's arguments unrounded
    const std::vector<Real> &restLengthArguments() const { return m_restLengthArg; }

protected:
    std::vector<std::array<Vector3, 2>> m_localRBPoints;
    std::vector<Real> m_squaredRestLength;
    std::vector<Real> m_restLengthArg;
};

class RBSpringAnchorEnergy : public RBEnergyBase<1>
{
public:
    void addStencil(std::array<int, 1> stencil, const Vector3 &localRigidBodyPoint, const Vector3 &anchor,
                    Real stiffness = 0);
    const std::vector<Vector3> &anchors() const { return m_anchor; }
    const std::vector<Vector3> &localRBPoints() const { return m_localRBPoint; }

protected:
    std::vector<Vector3> m_localRBPoint;
    std::vector<Vector3> m_anchor;
};

class RBOrientationAlignmentEnergy : public RBEnergyBase<2>
{
public:
    void addStencil(std::array<int, 2> stencil, Real stiffness = 0);
};

}  // namespace spg
