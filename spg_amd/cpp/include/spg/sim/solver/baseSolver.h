// Mirror of spg::solver::BaseSolver (spg/include/spg/sim/solver/baseSolver.h:14-41,
// source/sim/solver/baseSolver.cpp:10-35) for particle groups, and the GPU drop-ins for
// spg::solver::VBD (spg/include/spg/sim/solver/vbd.h:14-41), spg::solver::XPBD
// (spg/include/spg/sim/solver/xpbd.h:15-30) and the implicit Euler family
// (spg/include/spg/sim/solver/implicitEuler*.h, bdf2.h).
#pragma once

#include <memory>
#include <vector>

#include <spg/sim/simObject/particleGroup.h>
#include <spg/sim/simObject/rigidBodyGroup.h>
#include <spg/types.h>

struct spg_vbd_ctx;
struct spg_rigid_ctx;
struct spg_xpbd_ctx;
struct spg_implicit_ctx;

namespace spg {

enum class Verbosity { None, Performance, Debug };  // spg/include/spg/constants.h:5

namespace solver {

class BaseSolver
{
public:
    virtual ~BaseSolver() = default;
    virtual void step() = 0;
    // copies the group; its energies stay shared (baseSolver.cpp:10-14)
    void addObject(const ParticleGroup &object) { m_objects.push_back(object); }
    void addObject(const RigidBodyGroup &object) { m_rbObjects.push_back(object); }
    std::vector<ParticleGroup> &particleGroups() { return m_objects; }
    const std::vector<ParticleGroup> &particleGroups() const { return m_objects; }
    std::vector<RigidBodyGroup> &rbGroups() { return m_rbObjects; }
    const std::vector<RigidBodyGroup> &rbGroups() const { return m_rbObjects; }
    int numSimObjects() const { return static_cast<int>(m_objects.size() + m_rbObjects.size()); }
    void setDt(Real dt) { m_dtStep = dt; }
    Real dt() const { return m_dtStep; }
    void setNumSubsteps(int nsubsteps) { m_nsubsteps = nsubsteps; }
    int numSubsteps() const { return m_nsubsteps; }
    virtual void reset();
    void setVerbosity(Verbosity verbosity) { m_verbosity = verbosity; }
    void setGravity(const Vector3 &gravity) { m_gravity = gravity; }

protected:
    // rejects rigid-body groups in the solvers that do not handle them
    void requireNoRigidBodies(const char *solverName) const;

    std::vector<ParticleGroup> m_objects;
    std::vector<RigidBodyGroup> m_rbObjects;
    Vector3 m_gravity{0, -9.8, 0};
    Real m_time{0};
    Real m_dtStep{0.01};
    int m_nsubsteps{1};
    Verbosity m_verbosity{Verbosity::None};
};

// VBD on one MI355X. Same call pattern as spg::solver::VBD: add objects, set dt/substeps/gravity,
// call step() in a loop. The first step (or the first after reset()/requirePrecomputationUpdate())
// colours the graph, builds the incidence streams and uploads them (vbd.cpp:14-18).
//
// Every particle group's energies must be among the thirteen particle energies of spg_vbd.h, and every
// RigidBodyGroup's among the four rigid-body energies of spg_rigid.h; any other energy type throws
// std::runtime_error at the first step (there is no CPU fallback). Rigid-body groups run in their own FP64
// context (include/spg_rigid.h), stepped with the same dt, substeps, iterations and gravity; no energy couples
// the two kinds of object, so the order between them does not change the result.
//
// State ownership: by default the host ParticleGroups stay the source of truth, as with the CPU
// solver: step() uploads their positions and velocities, advances on the GPU and writes the result
// back. setResidentState(true) keeps the state on the GPU between steps (no per-step transfers);
// syncToHost() then copies it into particleGroups() on demand, and markHostStateDirty() uploads host
// edits before the next step.
class VBDGpu : public BaseSolver
{
public:
    enum class InitialGuessType { InertialWithAcceleration };
    explicit VBDGpu(int device = 0);
    ~VBDGpu() override;
    VBDGpu(const VBDGpu &) = delete;
    VBDGpu &operator=(const VBDGpu &) = delete;

    void step() override;
    void reset() override;
    void setInitialGuessType(const InitialGuessType type) { m_initialGuessType = type; }
    void requirePrecomputationUpdate() { m_precomputationUpdateRequired = true; }
    // m_iterations of the reference (vbd.h:33) has no setter there; here it has one (default 1).
    void setIterations(int iterations) { m_iterations = iterations; }
    int iterations() const { return m_iterations; }
    void setResidentState(bool resident) { m_resident = resident; }
    void syncToHost();
    void markHostStateDirty() { m_hostDirty = true; }
    // The C-ABI context (include/spg_vbd.h) behind this solver, or nullptr before the first step.
    spg_vbd_ctx *context() { return m_ctx; }
    // The rigid-body context (include/spg_rigid.h), or nullptr before the first step / without rigid bodies.
    spg_rigid_ctx *rigidContext() { return m_rbCtx; }

protected:
    void buildContext();
    void uploadState();
    void buildRigidContext();
    void uploadRigidState();
    void syncRigidToHost();

    InitialGuessType m_initialGuessType{InitialGuessType::InertialWithAcceleration};
    int m_iterations{1};
    int m_device{0};
    bool m_precomputationUpdateRequired{true};
    bool m_resident{false};
    bool m_hostDirty{true};
    spg_vbd_ctx *m_ctx{nullptr};
    spg_rigid_ctx *m_rbCtx{nullptr};
};

// XPBD on one MI355X, the GPU drop-in for spg::solver::XPBD: XPBDGpu(true) is the parallel Gauss-Seidel
// solver over stencil colours, XPBDGpu(false) the serial Gauss-Seidel order (run as dependency levels,
// which reproduce it). The first step (or the first after reset()/requirePrecomputationUpdate()) colours
// the stencils and uploads the particle groups (xpbd.cpp:18-21). Supported energies and state ownership
// are those of VBDGpu; the projection uses each energy's effectiveCompliance() and the particles'
// invMasses().
class XPBDGpu : public BaseSolver
{
public:
    explicit XPBDGpu(bool useParallelGaussSeidel = true, int device = 0);
    ~XPBDGpu() override;
    XPBDGpu(const XPBDGpu &) = delete;
    XPBDGpu &operator=(const XPBDGpu &) = delete;

    void step() override;
    void reset() override;
    void requirePrecomputationUpdate() { m_parallelGroupsDirty = true; }
    void setResidentState(bool resident) { m_resident = resident; }
    void syncToHost();
    void markHostStateDirty() { m_hostDirty = true; }
    // The C-ABI context (include/spg_xpbd.h) behind this solver, or nullptr before the first step.
    spg_xpbd_ctx *context() { return m_ctx; }

protected:
    void buildContext();
    void uploadState();

    bool m_useParallelGaussSeidel{true};
    bool m_parallelGroupsDirty{true};
    int m_device{0};
    bool m_resident{false};
    bool m_hostDirty{true};
    spg_xpbd_ctx *m_ctx{nullptr};
};

// The implicit Euler family on one MI355X, in FP64: ImplicitEulerBase's force, mass and stiffness
// assembly and its conjugate gradient (implicitEulerBase.cpp:102-240) on the GPU, driven by the step of
// the chosen solver. The concrete classes below carry the reference's class names with a Gpu suffix.
// The first step (or the first after reset() / a new object) builds the block-sparse pattern; the state
// handling is that of VBDGpu.
class ImplicitEulerGpu : public BaseSolver
{
public:
    enum class Kind { BaraffWitkin = 0, NewtonDv = 1, NewtonDx = 2, NewtonRobust = 3, BDF2 = 4 };
    ImplicitEulerGpu(Kind kind, int device);
    ~ImplicitEulerGpu() override;
    ImplicitEulerGpu(const ImplicitEulerGpu &) = delete;
    ImplicitEulerGpu &operator=(const ImplicitEulerGpu &) = delete;

    void step() override;
    void reset() override;
    void requirePrecomputationUpdate() { m_rebuild = true; }
    void setResidentState(bool resident) { m_resident = resident; }
    void syncToHost();
    void markHostStateDirty() { m_hostDirty = true; }
    // The C-ABI context (include/spg_implicit.h) behind this solver, or nullptr before the first step.
    spg_implicit_ctx *context() { return m_ctx; }

protected:
    void buildContext();
    void uploadState();

    Kind m_kind;
    int m_device{0};
    bool m_rebuild{true};
    bool m_resident{false};
    bool m_hostDirty{true};
    spg_implicit_ctx *m_ctx{nullptr};
};

class ImplicitEulerBaraffWitkinGpu : public ImplicitEulerGpu
{
public:
    explicit ImplicitEulerBaraffWitkinGpu(int device = 0) : ImplicitEulerGpu(Kind::BaraffWitkin, device) {}
};
class ImplicitEulerNewtonDvGpu : public ImplicitEulerGpu
{
public:
    explicit ImplicitEulerNewtonDvGpu(int device = 0) : ImplicitEulerGpu(Kind::NewtonDv, device) {}
};
class ImplicitEulerNewtonDxGpu : public ImplicitEulerGpu
{
public:
    explicit ImplicitEulerNewtonDxGpu(int device = 0) : ImplicitEulerGpu(Kind::NewtonDx, device) {}
};
class ImplicitEulerNewtonRobustGpu : public ImplicitEulerGpu
{
public:
    explicit ImplicitEulerNewtonRobustGpu(int device = 0) : ImplicitEulerGpu(Kind::NewtonRobust, device) {}
};
class BDF2Gpu : public ImplicitEulerGpu
{
public:
    explicit BDF2Gpu(int device = 0) : ImplicitEulerGpu(Kind::BDF2, device) {}
};

}  // namespace solver
}  // namespace spg
