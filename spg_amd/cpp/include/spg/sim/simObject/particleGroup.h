// Mirror of spg::ParticleGroup (spg/include/spg/sim/simObject/particleGroup.h:14-60): the particle
// state that the VBD solver reads and writes, with the same accessors and semantics.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include <spg/types.h>

namespace spg {

template <class TSimObject>
class Energy;

class ParticleGroup
{
public:
    static constexpr int s_nDOFs = 3;
    using EnergyT = Energy<ParticleGroup>;

    const std::vector<Vector3> &positions() const { return m_x; }
    std::vector<Vector3> &positions() { return m_x; }
    const std::vector<Vector3> &positions0() const { return m_x0; }
    std::vector<Vector3> &velocities() { return m_v; }
    const std::vector<Vector3> &velocities() const { return m_v; }
    void getPositions(VectorX &pos, int offsetIndex) const;
    void getVelocities(VectorX &vel, int offsetIndex) const;
    void setPositions(const VectorX &pos, int offsetIndex);
    void setVelocities(const VectorX &vel, int offsetIndex);
    void updateElementPositionFromDx(const Vector3 &dx, int elementId);
    void integrateVelocities(Real dt);
    void computeIntegratedVelocities(const VectorX &oldPos, int offsetIndex, Real invdt);
    const std::vector<Real> &invMasses() const { return m_w; }
    const std::vector<Real> &masses() const { return m_m; }
    std::vector<std::shared_ptr<EnergyT>> &energies() { return m_energies; }
    const std::vector<std::shared_ptr<EnergyT>> &energies() const { return m_energies; }
    int nDOF() const { return static_cast<int>(m_x.size()) * 3; }
    int size() const { return static_cast<int>(m_x.size()); }

    void addParticle(const Vector3 &position, const Vector3 &position0, Real mass);
    void addEnergy(std::shared_ptr<EnergyT> energy);
    void removeEnergy(std::shared_ptr<EnergyT> energy);
    void scaleMasses(Real scaleFactor);
    void setName(std::string name) { m_name = std::move(name); }
    void reset();

protected:
    std::string m_name;
    std::vector<Vector3> m_x;
    std::vector<Vector3> m_xInitial;
    std::vector<Vector3> m_x0;
    std::vector<Vector3> m_v;
    std::vector<Real> m_m;
    std::vector<Real> m_w;
    std::vector<std::shared_ptr<EnergyT>> m_energies;
};

}  // namespace spg
