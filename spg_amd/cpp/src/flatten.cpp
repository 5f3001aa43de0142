#include "flatten.h"

#include <stdexcept>

#include <spg/sim/energy/energies.h>

#include "spg_vbd.h"

namespace spg::solver {

namespace {

template <class E>
void requirePrepared(const E &e, const char *name)
{
    if (!e.prepared()) throw std::runtime_error(std::string(name) + " used before preparePrecomputations");
}

}  // namespace

int64_t flatStencilSize(int32_t type)
{
    switch (type) {
    case SPG_VBD_SPRING_ANCHOR: return 1;
    case SPG_VBD_SPRING: return 2;
    case SPG_VBD_MEMBRANE_STVK: return 3;
    default: return 4;
    }
}

FlatEnergy flattenVbd(const ParticleGroup::EnergyT *energy)
{
    FlatEnergy f;
    if (auto *e = dynamic_cast<const SpringEnergy *>(energy)) {
        f.type = SPG_VBD_SPRING;
        f.stencils = e->flatStencils();
        for (int i = 0; i < e->nStencils(); ++i) {
            f.params.push_back(e->restLengths()[i]);
            f.params.push_back(e->effectiveStiffness()[i](0, 0));
        }
    } else if (auto *e = dynamic_cast<const MembraneStvkEnergy *>(energy)) {
        requirePrepared(*e, "MembraneStvkEnergy");
        f.type = SPG_VBD_MEMBRANE_STVK;
        f.stencils = e->flatStencils();
        for (int i = 0; i < e->nStencils(); ++i) {
            const auto &D = e->inverseReferenceMats()[i];
            const auto &K = e->effectiveStiffness()[i];
            f.params.insert(f.params.end(), D.m.begin(), D.m.end());
            f.params.insert(f.params.end(), K.m.begin(), K.m.end());
        }
    } else if (auto *e = dynamic_cast<const DiscreteBendingEnergy *>(energy)) {
        requirePrepared(*e, "DiscreteBendingEnergy");
        f.type = SPG_VBD_DISCRETE_BENDING;
        f.stencils = e->flatStencils();
        for (int i = 0; i < e->nStencils(); ++i) {
            f.params.push_back(e->restThetas()[i]);
            f.params.push_back(e->effectiveStiffness()[i](0, 0));
        }
    } else if (auto *e = dynamic_cast<const StableNeoHookeanEnergy *>(energy)) {
        requirePrepared(*e, "StableNeoHookeanEnergy");
        f.type = SPG_VBD_STABLE_NEO_HOOKEAN;
        f.stencils = e->flatStencils();
        for (int i = 0; i < e->nStencils(); ++i) {
            const auto &D = e->inverseReferenceMats()[i];
            f.params.insert(f.params.end(), D.m.begin(), D.m.end());
            f.params.push_back(e->effectiveStiffness()[i](0, 0));
            f.params.push_back(e->effectiveStiffness()[i](1, 1));
            f.params.push_back(e->hydrostaticCorrectionTerms()[i]);
        }
    } else if (auto *e = dynamic_cast<const SpringAnchorEnergy *>(energy)) {
        f.type = SPG_VBD_SPRING_ANCHOR;
        f.stencils = e->flatStencils();
        for (int i = 0; i < e->nStencils(); ++i) {
            const auto &a = e->anchors()[i];
            f.params.insert(f.params.end(), {a[0], a[1], a[2], e->modelStiffness()[i](0, 0)});
        }
    } else {
        throw std::runtime_error("VBDGpu: unsupported energy type for the GPU solver");
    }
    return f;
}

FlatEnergy flattenXpbd(const ParticleGroup::EnergyT *energy)
{
    // the data StencilBlockEnergy::projectPosition reads (energy.h:117-153): effectiveCompliance()
    FlatEnergy f;
    if (auto *e = dynamic_cast<const SpringEnergy *>(energy)) {
        f.type = SPG_VBD_SPRING;
        f.stencils = e->flatStencils();
        for (int i = 0; i < e->nStencils(); ++i) {
            f.params.push_back(e->restLengths()[i]);
            f.params.push_back(e->effectiveCompliance()[i](0, 0));
        }
    } else if (auto *e = dynamic_cast<const MembraneStvkEnergy *>(energy)) {
        requirePrepared(*e, "MembraneStvkEnergy");
        f.type = SPG_VBD_MEMBRANE_STVK;
        f.stencils = e->flatStencils();
        for (int i = 0; i < e->nStencils(); ++i) {
            const auto &D = e->inverseReferenceMats()[i];
            const auto &C = e->effectiveCompliance()[i];
            f.params.insert(f.params.end(), D.m.begin(), D.m.end());
            f.params.insert(f.params.end(), C.m.begin(), C.m.end());
        }
    } else if (auto *e = dynamic_cast<const DiscreteBendingEnergy *>(energy)) {
        requirePrepared(*e, "DiscreteBendingEnergy");
        f.type = SPG_VBD_DISCRETE_BENDING;
        f.stencils = e->flatStencils();
        for (int i = 0; i < e->nStencils(); ++i) {
            f.params.push_back(e->restThetas()[i]);
            f.params.push_back(e->effectiveCompliance()[i](0, 0));
        }
    } else if (auto *e = dynamic_cast<const StableNeoHookeanEnergy *>(energy)) {
        requirePrepared(*e, "StableNeoHookeanEnergy");
        f.type = SPG_VBD_STABLE_NEO_HOOKEAN;
        f.stencils = e->flatStencils();
        for (int i = 0; i < e->nStencils(); ++i) {
            const auto &D = e->inverseReferenceMats()[i];
            const auto &C = e->effectiveCompliance()[i];
            f.params.insert(f.params.end(), D.m.begin(), D.m.end());
            f.params.insert(f.params.end(), C.m.begin(), C.m.end());
            f.params.push_back(e->hydrostaticCorrectionTerms()[i]);
        }
    } else if (auto *e = dynamic_cast<const SpringAnchorEnergy *>(energy)) {
        f.type = SPG_VBD_SPRING_ANCHOR;
        f.stencils = e->flatStencils();
        for (int i = 0; i < e->nStencils(); ++i) {
            const auto &a = e->anchors()[i];
            f.params.insert(f.params.end(), {a[0], a[1], a[2], e->effectiveCompliance()[i](0, 0)});
        }
    } else {
        throw std::runtime_error("XPBDGpu: unsupported energy type for the GPU solver");
    }
    return f;
}

}  // namespace spg::solver
