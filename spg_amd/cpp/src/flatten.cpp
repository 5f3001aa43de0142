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
    case SPG_VBD_SPRING:
    case SPG_VBD_SPRING_CONTINUUM:
    case SPG_VBD_SPRING_SQUARED_CONTINUUM: return 2;
    case SPG_VBD_MEMBRANE_STVK:
    case SPG_VBD_MEMBRANE_BARAFF_WITKIN:
    case SPG_VBD_MEMBRANE_CHOI: return 3;
    default: return 4;
    }
}

namespace {

// The further energies in either layout: `stiffness` selects effectiveStiffness() (spg_vbd_energy) or
// effectiveCompliance() (spg_xpbd_energy); both are diagonal for these energies, so only the diagonal goes.
template <int N>
void pushDiagonal(std::vector<double> &out, const Matrix<N, N> &m)
{
    for (int i = 0; i < N; ++i) out.push_back(m(i, i));
}

bool flattenFurther(const ParticleGroup::EnergyT *energy, bool stiffness, FlatEnergy &f)
{
    auto mats = [&](const auto *e, int i) -> const auto & {
        return stiffness ? e->effectiveStiffness()[i] : e->effectiveCompliance()[i];
    };
    if (auto *e = dynamic_cast<const SpringContinuumEnergy *>(energy)) {
        f.type = SPG_VBD_SPRING_CONTINUUM;
        for (int i = 0; i < e->nStencils(); ++i) f.params.insert(f.params.end(), {e->restLengths()[i], mats(e, i)(0, 0)});
    } else if (auto *e = dynamic_cast<const SpringSquaredContinuumEnergy *>(energy)) {
        f.type = SPG_VBD_SPRING_SQUARED_CONTINUUM;
        for (int i = 0; i < e->nStencils(); ++i) f.params.insert(f.params.end(), {e->restLengths()[i], mats(e, i)(0, 0)});
    } else if (auto *e = dynamic_cast<const MembraneBaraffWitkinEnergy *>(energy)) {
        requirePrepared(*e, "MembraneBaraffWitkinEnergy");
        f.type = SPG_VBD_MEMBRANE_BARAFF_WITKIN;
        for (int i = 0; i < e->nStencils(); ++i) {
            const auto &D = e->inverseReferenceMats()[i];
            f.params.insert(f.params.end(), D.m.begin(), D.m.end());
            pushDiagonal(f.params, mats(e, i));
        }
    } else if (auto *e = dynamic_cast<const MembraneChoiEnergy *>(energy)) {
        requirePrepared(*e, "MembraneChoiEnergy");
        f.type = SPG_VBD_MEMBRANE_CHOI;
        for (int i = 0; i < e->nStencils(); ++i) {
            const auto &D = e->inverseReferenceMats()[i];
            f.params.insert(f.params.end(), D.m.begin(), D.m.end());
            pushDiagonal(f.params, mats(e, i));
        }
    } else if (auto *e = dynamic_cast<const QuadraticBendingEnergy *>(energy)) {
        requirePrepared(*e, "QuadraticBendingEnergy");
        f.type = SPG_VBD_QUADRATIC_BENDING;
        for (int i = 0; i < e->nStencils(); ++i) {
            const auto &L = e->laplacians()[i];  // [c0 I, c1 I, c2 I, c3 I]
            for (int s = 0; s < 4; ++s) f.params.push_back(L(0, 3 * s));
            f.params.push_back(mats(e, i)(0, 0));
        }
    } else if (auto *e = dynamic_cast<const BaraffWitkinBendingEnergy *>(energy)) {
        requirePrepared(*e, "BaraffWitkinBendingEnergy");
        f.type = SPG_VBD_BARAFF_WITKIN_BENDING;
        for (int i = 0; i < e->nStencils(); ++i) f.params.insert(f.params.end(), {e->restThetas()[i], mats(e, i)(0, 0)});
    } else if (auto *e = dynamic_cast<const StvkEnergy *>(energy)) {
        requirePrepared(*e, "StvkEnergy");
        f.type = SPG_VBD_STVK;
        for (int i = 0; i < e->nStencils(); ++i) {
            const auto &D = e->inverseReferenceMats()[i];
            f.params.insert(f.params.end(), D.m.begin(), D.m.end());
            pushDiagonal(f.params, mats(e, i));
        }
    } else if (auto *e = dynamic_cast<const StableSquaredNeoHookeanEnergy *>(energy)) {
        requirePrepared(*e, "StableSquaredNeoHookeanEnergy");
        f.type = SPG_VBD_STABLE_SQUARED_NEO_HOOKEAN;
        for (int i = 0; i < e->nStencils(); ++i) {
            const auto &D = e->inverseReferenceMats()[i];
            f.params.insert(f.params.end(), D.m.begin(), D.m.end());
            pushDiagonal(f.params, mats(e, i));
        }
    } else {
        return false;
    }
    f.stencils = energy->flatStencils();
    return true;
}

}  // namespace

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
    } else if (!flattenFurther(energy, true, f)) {
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
    } else if (!flattenFurther(energy, false, f)) {
        throw std::runtime_error("XPBDGpu: unsupported energy type for the GPU solver");
    }
    return f;
}

}  // namespace spg::solver
