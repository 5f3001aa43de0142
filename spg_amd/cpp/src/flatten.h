// Flattening of one mirrored spg energy into the per-stencil parameter layouts of the C-ABI:
// spg_vbd_energy (include/spg_vbd.h, stiffness form) and spg_xpbd_energy (include/spg_xpbd.h,
// compliance form). The energy is identified by its concrete type, since the GPU solvers cannot call
// host virtuals; other energy types throw std::runtime_error.
#pragma once

#include <cstdint>
#include <vector>

#include <spg/sim/simObject/particleGroup.h>

namespace spg::solver {

struct FlatEnergy {
    int32_t type = 0;
    std::vector<int32_t> stencils;
    std::vector<double> params;
};

FlatEnergy flattenVbd(const ParticleGroup::EnergyT *energy);
FlatEnergy flattenXpbd(const ParticleGroup::EnergyT *energy);
int64_t flatStencilSize(int32_t type);

}  // namespace spg::solver
