"""GPU parity of VBD on rigid bodies (include/spg_rigid.h) against the FP64 oracle, on the reference app's three
rigid-body scenes (anchored box, articulated row, spring chain) and a lattice of boxes.

Bars (both sides FP64, closed forms on the GPU against the oracle's dual numbers, so only rounding differs):
colours bit-exact; the element derivatives within 1e-12 relative; after the steps, centres within 1e-9 relative
(max-norm error / max-norm of the centres' displacement), orientations within 1e-9 rad, velocities within 1e-8
of their max-norm.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation as Rot

import oracle_lib as ol
from spg_amd import native, scenes

pytestmark = pytest.mark.gpu

SCENES = {
    "anchored": lambda: scenes.rb_anchored(),
    "articulated": lambda: scenes.rb_articulated(8),
    "chain": lambda: scenes.rb_chain(12),
    "lattice": lambda: scenes.rb_lattice(16),
}


def compare(name, ctx, o, x_start, group=0, tol=1e-9):
    x, th, v, w = ctx.state(group)
    xo, tho, vo, wo = o.state()
    ex = np.abs(x - xo).max() / max(np.abs(xo - x_start).max(), 1e-300)
    # orientation error as the angle of R_gpu R_oracle^T
    eth = max(np.linalg.norm((Rot.from_rotvec(a) * Rot.from_rotvec(b).inv()).as_rotvec()) for a, b in zip(th, tho))
    ev = max(np.abs(v - vo).max() / max(np.abs(vo).max(), 1e-300), np.abs(w - wo).max() / max(np.abs(wo).max(), 1e-300))
    print(f"{name}: centre rel {ex:.2e}, orientation {eth:.2e} rad, velocity rel {ev:.2e}", flush=True)
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(th))
    assert ex < tol and eth < tol and ev < 10 * tol


@pytest.mark.parametrize("kind", [21, 22, 23, 24])
def test_element_derivatives_device_matches_oracle(kind):
    import test_rigid_oracle as T
    rng = np.random.default_rng(200 + kind)
    ns = ol.RB_STENCIL_SIZE[kind]
    els = [T.random_element(kind, rng) for _ in range(64)]
    com = np.array([c.ravel() for c, _, _ in els])
    rot = np.array([np.array(R).ravel() for _, R, _ in els])
    raw = np.array([r for _, _, r in els])
    g, h = native.rigid_element_derivatives(kind, com, rot, raw, device=0)
    for i, (c, R, r) in enumerate(els):
        go, ho = ol.rb_element(kind, c, np.array(R), r)
        hb = np.stack([ho[6 * s:6 * s + 6, 6 * s:6 * s + 6] for s in range(ns)])
        assert np.abs(g[i] - go).max() <= 1e-12 * max(1.0, np.abs(go).max())
        assert np.abs(h[i] - hb).max() <= 1e-12 * max(1.0, np.abs(hb).max())


@pytest.mark.parametrize("name", list(SCENES))
def test_rigid_vbd_steps(name, oracle_threads):
    s = SCENES[name]()
    ctx = native.RigidContext([s])
    o = ol.rb_from_scene(s)
    c, nc = ctx.colors()
    co, nco = o.colors()
    assert nc == nco and np.array_equal(c, co)
    x_start = s.x.copy()
    for _ in range(10):
        ctx.step(s.dt, s.nsubsteps, s.iterations, s.gravity)
        o.step(s.dt, s.nsubsteps, s.iterations, s.gravity)
    compare(name, ctx, o, x_start)
    assert ctx.stats()["skipped_blocks"] == 0


def test_rigid_vbd_substeps_and_graph_replay(oracle_threads):
    """Two substeps per step; the captured graph is replayed and re-captured when the arguments change; eager
    launches give the graph's result bitwise."""
    s = scenes.rb_lattice(12)
    ctx = native.RigidContext([s])
    eager = native.RigidContext([s], graph=False)
    o = ol.rb_from_scene(s)
    for k in range(6):
        it = 5 if k < 3 else 7
        ctx.step(s.dt, 2, it, s.gravity)
        eager.step(s.dt, 2, it, s.gravity)
        o.step(s.dt, 2, it, s.gravity)
    for a, b in zip(ctx.state(), eager.state()):
        assert np.array_equal(a, b)
    compare("lattice 2 substeps", ctx, o, s.x)


def test_rigid_groups_and_state_roundtrip(oracle_threads):
    """Three groups in one context (BaseSolver::addObject order) each equal their own oracle run; set_state with a
    new orientation rebuilds rotation and inertia (setPositions) and the next step matches the oracle's."""
    ss = [scenes.rb_chain(6), scenes.rb_anchored(), scenes.rb_lattice(6)]
    ctx = native.RigidContext(ss)
    os_ = [ol.rb_from_scene(s) for s in ss]
    rng = np.random.default_rng(9)
    th = rng.normal(size=(ss[2].nb, 3)) * 0.1
    w = rng.normal(size=(ss[2].nb, 3))
    ctx.set_state(2, theta=th, omega=w)
    os_[2].set_state(theta=th, omega=w)
    x, th2, v, w2 = ctx.state(2)
    assert np.array_equal(th2, th) and np.array_equal(w2, w) and np.array_equal(x, ss[2].x)
    for _ in range(4):
        ctx.step(ss[0].dt, 1, 10, ss[0].gravity)
        for s, o in zip(ss, os_):
            o.step(s.dt, 1, 10, s.gravity)
    for g, (s, o) in enumerate(zip(ss, os_)):
        compare(f"group {g}", ctx, o, s.x, group=g)


def test_rigid_lattice_large(oracle_threads):
    """A 96 x 96 lattice (9216 bodies), one step, against the oracle."""
    s = scenes.rb_lattice(96)
    ctx = native.RigidContext([s])
    o = ol.rb_from_scene(s)
    ctx.step(s.dt, 1, s.iterations, s.gravity)
    o.step(s.dt, 1, s.iterations, s.gravity)
    compare("lattice 96", ctx, o, s.x)


def test_rigid_profile_substep_advances_like_step():
    s = scenes.rb_lattice(10)
    a = native.RigidContext([s], graph=False)
    b = native.RigidContext([s], graph=False)
    ms, n = a.profile_substep(s.dt, s.iterations, s.gravity)
    b.step(s.dt, 1, s.iterations, s.gravity)
    assert n == s.iterations * a.stats()["n_colors"] and ms > 0
    for u, v in zip(a.state(), b.state()):
        assert np.array_equal(u, v)


def test_rigid_validation_messages():
    s = scenes.rb_chain(3)

    def with_energy(e):
        return scenes.RigidScene(x=s.x, theta=s.theta, m=s.m, inertia=s.inertia, energies=[e])

    with pytest.raises(native.NativeError, match="Zero/negative rest length invalid"):
        native.RigidContext([with_energy(("rb_spring", [[0, 1]], [[0, 0, 0, 0, 0, 0, 0.0, 1.0]]))])
    with pytest.raises(native.NativeError, match="Negative rest length invalid"):
        native.RigidContext([with_energy(("rb_spring_squared", [[0, 1]], [[0, 0, 0, 0, 0, 0, -1.0, 1.0]]))])
    with pytest.raises(native.NativeError, match="stencil index out of range"):
        native.RigidContext([with_energy(("rb_alignment", [[0, 3]], [[1.0]]))])
    with pytest.raises(native.NativeError, match="stencil index out of range"):
        native.RigidContext([with_energy(("rb_anchor", [[5]], [[0, 0, 0, 0, 0, 0, 1.0]]))])
    native.RigidContext([with_energy(("rb_spring_squared", [[0, 1]], [[0, 0, 0, 0, 0, 0, 0.0, 1.0]]))])


def test_vbd_gpu_mirror_with_rigid_groups(oracle_threads):
    """spg::solver::VBDGpu of the C++ mirror holding a cloth and two RigidBodyGroups: the rigid groups equal the
    spg_rigid C-ABI bitwise and the oracle to the bars above, the cloth equals a VBDGpu without rigid bodies
    bitwise, and the host groups receive the state including the composed rotation matrices."""
    cloth = scenes.cloth(16, young=1e4, pin="top_row", iterations=10)
    rbs = [scenes.rb_chain(5, iterations=10), scenes.rb_lattice(6, iterations=10)]
    solver = native.VBDGpu()
    solver.add_object(cloth.particle_group())
    for s in rbs:
        solver.add_rb_object(native.RigidBodyGroup(s))
    solver.set_dt(cloth.dt)
    solver.set_gravity(cloth.gravity)
    solver.set_iterations(10)
    plain = native.VBDGpu()
    plain.add_object(cloth.particle_group())
    plain.set_dt(cloth.dt)
    plain.set_gravity(cloth.gravity)
    plain.set_iterations(10)
    ctx = native.RigidContext(rbs)
    os_ = [ol.rb_from_scene(s) for s in rbs]
    for _ in range(3):
        solver.step()
        plain.step()
        ctx.step(cloth.dt, 1, 10, cloth.gravity)
        for s, o in zip(rbs, os_):
            o.step(s.dt, 1, 10, cloth.gravity)
    assert np.array_equal(solver.state(0)[0], plain.state(0)[0])
    for g, (s, o) in enumerate(zip(rbs, os_)):
        x, th, v, w, R = solver.rb_state(g)
        for a, b in zip((x, th, v, w), ctx.state(g)):
            assert np.array_equal(a, b)
        assert np.array_equal(R.reshape(-1, 9), ctx.rotations(g))
        compare(f"mirror group {g}", ctx, o, s.x, group=g)


def test_rigid_edge_cases(oracle_threads):
    """An empty group beside a free body with no energies (ballistic centre, gyroscopic spin of an anisotropic body),
    a body tied to itself by a squared spring (both stencil slots on one body, as addStencil allows), and a step of
    zero iterations (prediction and velocity update only): each equals the oracle."""
    empty = scenes.RigidScene(x=np.zeros((0, 3)), theta=np.zeros((0, 3)), m=np.zeros(0), inertia=np.zeros((0, 3, 3)))
    free = scenes.RigidScene(x=np.array([[0.3, 1.0, -0.2]]), theta=np.array([[0.2, -0.4, 0.1]]), m=np.array([2.0]),
                             inertia=np.diag([0.2, 0.5, 0.9])[None])
    selftied = scenes.rb_chain(3)
    selftied.energies.append(("rb_spring_squared", np.array([[1, 1]], np.int32),
                              np.array([[0.5, 0.0, 0.0, -0.5, 0.0, 0.0, 0.2, 50.0]])))
    groups = [empty, free, selftied]
    ctx = native.RigidContext(groups)
    w0 = np.array([[1.5, -0.7, 2.0]])
    ctx.set_state(1, v=np.array([[1.0, 0.0, 0.5]]), omega=w0)
    assert ctx.state(0)[0].shape == (0, 3)
    os_ = [None, ol.rb_from_scene(free), ol.rb_from_scene(selftied)]
    os_[1].set_state(v=np.array([[1.0, 0.0, 0.5]]), omega=w0)
    for it in (10, 0, 10):
        ctx.step(1 / 60, 2, it, (0.0, -9.81, 0.0))
        for o in os_[1:]:
            o.step(1 / 60, 2, it, (0.0, -9.81, 0.0))
    for g in (1, 2):
        compare(f"edge group {g}", ctx, os_[g], groups[g].x, group=g)
    assert ctx.stats()["skipped_blocks"] == 0


@pytest.mark.parametrize("resident", [False, True])
def test_vbd_gpu_mirror_rebuild_continues_bitwise(resident):
    """requirePrecomputationUpdate() between steps rebuilds the contexts from the groups' current state, including
    the composed rotation matrices (and, with resident state, after bringing the device state back): the steps
    that follow equal those of a solver that was never rebuilt, bitwise."""
    cloth = scenes.cloth(12, young=1e4, pin="top_row", iterations=10)
    rbs = [scenes.rb_chain(5, iterations=10), scenes.rb_lattice(4, iterations=10)]
    solvers = []
    for _ in range(2):
        sv = native.VBDGpu()
        sv.add_object(cloth.particle_group())
        for s in rbs:
            sv.add_rb_object(native.RigidBodyGroup(s))
        sv.set_dt(cloth.dt)
        sv.set_gravity(cloth.gravity)
        sv.set_iterations(10)
        sv.set_resident_state(resident)
        solvers.append(sv)
    rebuilt, plain = solvers
    for k in range(4):
        if k == 2:
            rebuilt.require_precomputation_update()
        rebuilt.step()
        plain.step()
    if resident:
        rebuilt.sync_to_host()
        plain.sync_to_host()
    assert np.array_equal(rebuilt.state(0)[0], plain.state(0)[0])
    assert np.array_equal(rebuilt.state(0)[1], plain.state(0)[1])
    for g in range(len(rbs)):
        for a, b in zip(rebuilt.rb_state(g), plain.rb_state(g)):
            assert np.array_equal(a, b)
