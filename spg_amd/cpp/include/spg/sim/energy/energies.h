// The five energies of the VBD path, with the constructors, addStencil signatures, accessors and
// exceptions of their spg counterparts:
//   SpringEnergy            spg/include/spg/sim/energy/springEnergy.h:9-30, source/sim/energy/springEnergy.cpp:33-46
//   MembraneStvkEnergy      membraneStvkEnergy.h:9-21, membraneStvkEnergy.cpp:31-71
//   DiscreteBendingEnergy   discreteBendingEnergy.h:9-21, discreteBendingEnergy.cpp:33-72
//   StableNeoHookeanEnergy  stableNeoHookeanEnergy.h:10-26, stableNeoHookeanEnergy.cpp:64-108
//   SpringAnchorEnergy      springAnchorEnergy.h:7-21, springAnchorEnergy.cpp:29-39
#pragma once

#include <spg/sim/energy/energy.h>

namespace spg {

class SpringEnergy : public StencilBlockEnergy<2>
{
public:
    SpringEnergy() { m_name = "SpringEnergy"; }
    void addStencil(const std::array<int, s_stencilSize> &stencil, Real restLength, Real stiffness);
    const std::vector<Real> &restLengths() const { return m_restLength; }

protected:
    std::vector<Real> m_restLength;
};

class MembraneStvkEnergy : public StencilBlockEnergy<3, 3>
{
public:
    MembraneStvkEnergy() { m_name = "MembraneStvkEnergy"; }
    void addStencil(const std::array<int, s_stencilSize> &stencil, Real young, Real poisson);
    const std::vector<Matrix2> &inverseReferenceMats() const { return m_inverseReferenceMat; }
    void preparePrecomputations(const ParticleGroup &pGroup) override;

protected:
    std::vector<Matrix2> m_inverseReferenceMat;
};

class DiscreteBendingEnergy : public StencilBlockEnergy<4>
{
public:
    DiscreteBendingEnergy() { m_name = "DiscreteBendingEnergy"; }
    void addStencil(const std::array<int, s_stencilSize> &stencil, Real restTheta, Real stiffness);
    void preparePrecomputations(const ParticleGroup &pGroup) override;
    const std::vector<Real> &restThetas() const { return m_restTheta; }

protected:
    std::vector<Real> m_restTheta;
};

class StableNeoHookeanEnergy : public StencilBlockEnergy<4, 2>
{
public:
    StableNeoHookeanEnergy() { m_name = "StableNeoHookeanEnergy"; }
    void addStencil(const std::array<int, s_stencilSize> &stencil, Real young, Real poisson);
    void preparePrecomputations(const ParticleGroup &pGroup) override;
    const std::vector<Matrix3> &inverseReferenceMats() const { return m_inverseReferenceMat; }
    const std::vector<Real> &hydrostaticCorrectionTerms() const { return m_hydrostaticCorrectionTerm; }

protected:
    std::vector<Matrix3> m_inverseReferenceMat;
    std::vector<Real> m_hydrostaticCorrectionTerm;
};

class SpringAnchorEnergy : public StencilBlockEnergy<1>
{
public:
    SpringAnchorEnergy() { m_name = "SpringAnchorEnergy"; }
    void addStencil(std::array<int, s_stencilSize> stencil, const Vector3 &anchor, Real stiffness = 0);
    const std::vector<Vector3> &anchors() const { return m_anchor; }
    void updateAnchor(int i, const Vector3 &anchor) { m_anchor[i] = anchor; }

protected:
    std::vector<Vector3> m_anchor;
};

}  // namespace spg
