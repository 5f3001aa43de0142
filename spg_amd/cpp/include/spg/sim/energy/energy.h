// Mirror of the part of spg's energy interface that the VBD path consumes
// (spg/include/spg/sim/energy/energy.h:21-69 Energy<TSimObject>, 75-115 StencilBlockEnergy<N, C>):
// stencil storage and queries, stiffness/compliance matrices and preparePrecomputations semantics.
// The derivative machinery (TinyAD) is not mirrored: the GPU solver evaluates closed forms from the
// public data of the five energies, which VBDGpu reads through these accessors.
#pragma once

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include <spg/sim/simObject/particleGroup.h>
#include <spg/types.h>

namespace spg {

template <class TSimObject>
class Energy
{
public:
    static constexpr int s_nDOFs = TSimObject::s_nDOFs;
    virtual ~Energy() = default;
    virtual int nStencils() const = 0;
    virtual int stencilSize() const = 0;
    virtual std::vector<int> flatStencils() const = 0;
    virtual int nDOFs() const = 0;

    // energy.h:54-61: a single preparation is accepted
    virtual void preparePrecomputations(const TSimObject & /*obj*/)
    {
        if (m_prepared) {
            throw std::runtime_error("Energy was already prepared");
        }
        m_prepared = true;
    }
    bool prepared() const { return m_prepared; }

    void setName(const std::string &name) { m_name = name; }
    std::string name() { return m_name; }

protected:
    bool m_prepared{false};
    std::string m_name;
};

template <int TstencilSize, int TnConstraints = 1, class TSimObject = ParticleGroup>
class StencilBlockEnergy : public Energy<TSimObject>
{
protected:
    static constexpr int s_nDOFs = Energy<TSimObject>::s_nDOFs;
    static constexpr int s_stencilSize{TstencilSize};
    using StiffnessMat = Matrix<TnConstraints, TnConstraints>;

public:
    int nStencils() const override { return static_cast<int>(m_stencils.size()); }
    int stencilSize() const override { return TstencilSize; }
    int nDOFs() const override { return s_nDOFs; }
    std::vector<int> flatStencils() const override
    {
        std::vector<int> flattened;
        flattened.reserve(m_stencils.size() * TstencilSize);
        for (const auto &stencil : m_stencils) {
            flattened.insert(flattened.end(), stencil.begin(), stencil.end());
        }
        return flattened;
    }
    const std::vector<std::array<int, TstencilSize>> &stencils() const { return m_stencils; }
    const std::vector<StiffnessMat> &effectiveStiffness() const { return m_effectiveStiffness; }
    const std::vector<StiffnessMat> &modelStiffness() const { return m_modelStiffness; }
    const std::vector<StiffnessMat> &effectiveCompliance() const { return m_effectiveCompliance; }
    const std::vector<StiffnessMat> &modelCompliance() const { return m_modelCompliance; }

protected:
    StencilBlockEnergy() = default;

    std::vector<std::array<int, TstencilSize>> m_stencils;
    std::vector<StiffnessMat> m_effectiveStiffness;
    std::vector<StiffnessMat> m_modelStiffness;
    std::vector<StiffnessMat> m_effectiveCompliance;
    std::vector<StiffnessMat> m_modelCompliance;
};

}  // namespace spg
