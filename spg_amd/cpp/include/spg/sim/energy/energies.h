// The five energies of the VBD path, with the constructors, addStencil signatures, accessors and
// exceptions of their spg counterparts:
//   SpringEnergy            spg/include/spg/sim/energy/springEnergy.h:9-30, source/sim/energy/springEnergy.cpp:33-46
//   MembraneStvkEnergy      membraneStvkEnergy.h:9-21, membraneStvkEnergy.cpp:31-71
//   DiscreteBendingEnergy   discreteBendingEnergy.h:9-21, discreteBendingEnergy.cpp:33-72
//   StableNeoHookeanEnergy  stableNeoHookeanEnergy.h:10-26, stableNeoHookeanEnergy.cpp:64-108
//   SpringAnchorEnergy      springAnchorEnergy.h:7-21, springAnchorEnergy.cpp:29-39
// and the further particle energies:
//   SpringContinuumEnergy          springContinuumEnergy.h:9-21, springContinuumEnergy.cpp:20-35
//   SpringSquaredContinuumEnergy   springSquaredContinuumEnergy.h, springSquaredContinuumEnergy.cpp:21-36
//   MembraneBaraffWitkinEnergy     membraneBaraffWitkinEnergy.h:9-23, membraneBaraffWitkinEnergy.cpp:28-68
//   MembraneChoiEnergy             membraneChoiEnergy.h:12-26, membraneChoiEnergy.cpp:30-70
//   QuadraticBendingEnergy         quadraticBendingEnergy.h:9-24, quadraticBendingEnergy.cpp:32-89
//   BaraffWitkinBendingEnergy      baraffWitkinBendingEnergy.h:9-23, baraffWitkinBendingEnergy.cpp:38-75
//   StvkEnergy                     stvkEnergy.h:9-24, stvkEnergy.cpp:66-104
//   StableSquaredNeoHookeanEnergy  stableSquaredNeoHookeanEnergy.h:11-25, stableSquaredNeoHookeanEnergy.cpp:34-70
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

class SpringContinuumEnergy : public StencilBlockEnergy<2>
{
public:
    SpringContinuumEnergy() { m_name = "SpringContinuumEnergy"; }
    void addStencil(const std::array<int, s_stencilSize> &stencil, Real restLength, Real stiffness);
    const std::vector<Real> &restLengths() const { return m_restLength; }

protected:
    std::vector<Real> m_restLength;
};

class SpringSquaredContinuumEnergy : public StencilBlockEnergy<2>
{
public:
    SpringSquaredContinuumEnergy() { m_name = "SpringSquaredContinuumEnergy"; }
    void addStencil(const std::array<int, s_stencilSize> &stencil, Real restLength, Real stiffness);
    const std::vector<Real> &restLengths() const { return m_restLength; }

protected:
    std::vector<Real> m_restLength;
};

class MembraneBaraffWitkinEnergy : public StencilBlockEnergy<3, 3>
{
public:
    MembraneBaraffWitkinEnergy() { m_name = "MembraneBaraffWitkinEnergy"; }
    void addStencil(const std::array<int, s_stencilSize> &stencil, Real kWeft, Real kWarp, Real kBias);
    void preparePrecomputations(const ParticleGroup &pGroup) override;
    const std::vector<Matrix2> &inverseReferenceMats() const { return m_inverseReferenceMat; }

protected:
    std::vector<Matrix2> m_inverseReferenceMat;
};

class MembraneChoiEnergy : public StencilBlockEnergy<3, 4>
{
public:
    MembraneChoiEnergy() { m_name = "MembraneChoiEnergy"; }
    void addStencil(const std::array<int, s_stencilSize> &stencil, Real kWeft, Real kWarp, Real kBias);
    void preparePrecomputations(const ParticleGroup &pGroup) override;
    const std::vector<Matrix2> &inverseReferenceMats() const { return m_inverseReferenceMat; }

protected:
    std::vector<Matrix2> m_inverseReferenceMat;
};

class QuadraticBendingEnergy : public StencilBlockEnergy<4, 3>
{
public:
    QuadraticBendingEnergy() { m_name = "QuadraticBendingEnergy"; }
    void addStencil(const std::array<int, s_stencilSize> &stencil, Real restTheta, Real stiffness);
    void preparePrecomputations(const ParticleGroup &pGroup) override;
    const std::vector<Matrix<3, 12>> &laplacians() const { return m_laplacians; }

protected:
    std::vector<Matrix<3, 12>> m_laplacians;
    std::vector<Real> m_stiffnessFactor;
};

class BaraffWitkinBendingEnergy : public StencilBlockEnergy<4>
{
public:
    BaraffWitkinBendingEnergy() { m_name = "BaraffWitkinBendingEnergy"; }
    void addStencil(const std::array<int, s_stencilSize> &stencil, Real restTheta, Real stiffness);
    void preparePrecomputations(const ParticleGroup &pGroup) override;
    const std::vector<Real> &restThetas() const { return m_restTheta; }

protected:
    std::vector<Real> m_restTheta;
};

class StvkEnergy : public StencilBlockEnergy<4, 2>
{
public:
    StvkEnergy() { m_name = "StvkEnergy"; }
    void addStencil(const std::array<int, s_stencilSize> &stencil, Real young, Real poisson);
    void preparePrecomputations(const ParticleGroup &pGroup) override;
    const std::vector<Matrix3> &inverseReferenceMats() const { return m_inverseReferenceMat; }

protected:
    std::vector<Matrix3> m_inverseReferenceMat;
};

class StableSquaredNeoHookeanEnergy : public StencilBlockEnergy<4, 2>
{
public:
    StableSquaredNeoHookeanEnergy() { m_name = "StableSquaredNeoHookeanEnergy"; }
    void addStencil(const std::array<int, s_stencilSize> &stencil, Real young, Real poisson);
    void preparePrecomputations(const ParticleGroup &pGroup) override;
    const std::vector<Matrix3> &inverseReferenceMats() const { return m_inverseReferenceMat; }

protected:
    std::vector<Matrix3> m_inverseReferenceMat;
};

}  // namespace spg
