// spg::ParticleGroup mirror (reference: spg/source/sim/simObject/particleGroup.cpp:7-113).
#include <algorithm>

#include <spg/sim/energy/energy.h>
#include <spg/sim/simObject/particleGroup.h>

namespace spg {

void ParticleGroup::getPositions(VectorX &pos, int offsetIndex) const
{
    for (int i = 0; i < size(); ++i)
        for (int d = 0; d < 3; ++d) pos[offsetIndex + 3 * i + d] = m_x[i][d];
}

void ParticleGroup::getVelocities(VectorX &vel, int offsetIndex) const
{
    for (int i = 0; i < size(); ++i)
        for (int d = 0; d < 3; ++d) vel[offsetIndex + 3 * i + d] = m_v[i][d];
}

void ParticleGroup::setPositions(const VectorX &pos, int offsetIndex)
{
    for (int i = 0; i < size(); ++i)
        for (int d = 0; d < 3; ++d) m_x[i][d] = pos[offsetIndex + 3 * i + d];
}

void ParticleGroup::setVelocities(const VectorX &vel, int offsetIndex)
{
    for (int i = 0; i < size(); ++i)
        for (int d = 0; d < 3; ++d) m_v[i][d] = vel[offsetIndex + 3 * i + d];
}

void ParticleGroup::updateElementPositionFromDx(const Vector3 &dx, int elementId) { m_x[elementId] += dx; }

void ParticleGroup::integrateVelocities(Real dt)
{
    for (int i = 0; i < size(); ++i) m_x[i] += m_v[i] * dt;
}

void ParticleGroup::computeIntegratedVelocities(const VectorX &oldPos, int offsetIndex, Real invdt)
{
    for (int i = 0; i < size(); ++i)
        for (int d = 0; d < 3; ++d) m_v[i][d] = (m_x[i][d] - oldPos[offsetIndex + 3 * i + d]) * invdt;
}

void ParticleGroup::addParticle(const Vector3 &p, const Vector3 &p0, const Real mass)
{
    m_x.push_back(p);
    m_xInitial.push_back(p);
    m_x0.push_back(p0);
    m_v.push_back(Vector3::Zero());
    m_m.push_back(mass);
    m_w.push_back(mass == 0 ? 0 : 1 / mass);
}

void ParticleGroup::addEnergy(std::shared_ptr<EnergyT> energy) { m_energies.push_back(std::move(energy)); }

void ParticleGroup::removeEnergy(std::shared_ptr<EnergyT> energy)
{
    m_energies.erase(std::remove_if(m_energies.begin(), m_energies.end(),
                                    [&energy](auto &currentEnergy) { return energy == currentEnergy; }),
                     m_energies.end());
}

void ParticleGroup::reset()
{
    m_x = m_xInitial;
    for (auto &v : m_v) v.setZero();
}

void ParticleGroup::scaleMasses(const Real scaleFactor)
{
    for (size_t i = 0; i < m_m.size(); ++i) {
        if (m_m[i] != 0) {
            m_m[i] *= scaleFactor;
            m_w[i] /= scaleFactor;
        }
    }
}

}  // namespace spg
