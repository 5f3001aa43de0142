"""GPU parity of the XPBD path (through the C-ABI of include/spg_xpbd.h) against the FP64 oracle's XPBD
restatement on the same inputs.

Bars:
  * stencil groups: bit-exact integer equality with the oracle's greedy stencil colouring
    (coloring::colorStencils); serial levels satisfy the order property that reproduces XPBD(false);
  * projections: the kernels' device functions on the GPU within 1e-12 (FP64) / 1e-5 (FP32, with the
    FP64 fallback) of the oracle's dual-number projection, relative to the displacement;
  * steps: positions within 1e-5 relative (max-norm error / max-norm of the positions) of the FP64
    oracle started from the same FP32-rounded state, velocities within 1e-5 relative to the velocity
    scale max|x| / h that one FP32 rounding of the positions maps to;
  * determinism: every launch shape, graph replay and checkpoint/restore give bitwise equal states.
"""
import numpy as np
import pytest

import oracle_lib as ol
from gpu_common import rel_err
from spg_amd import native, scenes

pytestmark = pytest.mark.gpu


def xpbd_context(scene_list, parallel=True, **launch):
    if not isinstance(scene_list, (list, tuple)):
        scene_list = [scene_list]
    ctx = native.XpbdContext([s.xpbd_object() for s in scene_list], parallel_gauss_seidel=parallel)
    if launch:
        ctx.set_launch(**launch)
    return ctx


def oracle_from_gpu(scene, ctx, obj=0):
    o = ol.from_scene(scene)
    x, v = ctx.state(obj)
    o.set_state(x, v)
    return o


def check_step(ctx, o, dt, nsub, obj=0, tol=1e-5, label=""):
    xg, vg = ctx.state(obj)
    xo, vo = o.state()
    ex = rel_err(xg, xo)
    ev = float(np.max(np.abs(vg - vo)) / (np.max(np.abs(xo)) * nsub / dt))
    print(f"{label}: position rel {ex:.2e}, velocity rel {ev:.2e}", flush=True)
    assert ex < tol and ev < tol, (ex, ev)


def test_stencil_colors_bit_exact_config1_and_512():
    for s in (scenes.cloth_config1(), scenes.cloth_config2(seed=1)):
        ctx = xpbd_context(s)
        for e, (kind, st, _) in enumerate(s.energies):
            ref, nc = ol.color_stencils(st)
            assert ctx.num_groups(0, e) == nc
            assert np.array_equal(ctx.groups(0, e, len(st)), ref), kind
    # 14 bending colours, 7 membrane colours and 1 anchor group; the two 255-hinge tail colours run as one
    # launch, and so do the 254-triangle tail colour and the 512 anchors
    ctx.set_launch(group_runs=0)
    assert ctx.num_launches() == 14 + 7 + 1
    ctx.set_launch(group_runs=1)
    assert ctx.num_launches() == 12 + 1 + 6 + 1


@pytest.mark.parametrize("parallel", [True, False])
def test_group_runs_bitwise_equal_single_launches(parallel):
    """Consecutive small stencil groups run in one workgroup give bitwise the steps of one launch per group
    (512x512 cloth: two tail runs per substep; serial levels of a 48x48 cloth: many small levels)."""
    for s in (scenes.cloth_config2(seed=3), scenes.cloth(48, young=1e4, pin="top_row")):
        if not parallel and s.n > 10000:
            continue
        out = []
        for runs in (0, 1):
            ctx = xpbd_context(s, parallel=parallel)
            ctx.set_launch(group_runs=runs)
            ctx.step(s.dt, 4, s.gravity)
            out.append(ctx.state())
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_serial_levels_order_property():
    """XPBD(false): every two stencils that share a particle run in index order (level(s) < level(t) for
    s < t), and every level is the smallest with that property."""
    s = scenes.cloth(12, young=1e4)
    ctx = xpbd_context(s, parallel=False)
    for e, (kind, st, _) in enumerate(s.energies):
        lev = ctx.groups(0, e, len(st))
        last = {}
        for i, row in enumerate(st):
            need = max((last.get(int(v), -1) for v in row), default=-1) + 1
            assert lev[i] == need, (kind, i)
            for v in row:
                last[int(v)] = lev[i]


@pytest.mark.parametrize("kind", ["spring", "anchor", "stvk", "bending", "snh"])
def test_projection_device_functions_on_gpu(kind):
    import test_native_cpu as T
    cases = [c for c in T.PROJ if c[0] == kind]
    for k, xs, w, p, cm in cases:
        ref = ol.xpbd_project(k, xs.ravel(), w, p, cm, 1 / 60)
        xp = T._xpbd_layout(k, p, cm)
        scale = max(np.abs(ref).max(), 1e-300)
        for prec, tol in ((1, 1e-12), (0, 1e-5)):
            got = native.xpbd_project_stencils(k, xs.ravel()[None], w[None], xp[None], 1 / 60, precision=prec,
                                               device=0)[0]
            assert np.abs(got - ref).max() <= tol * scale, (k, prec)


@pytest.mark.parametrize("nsub", [1, 4])
def test_config1_single_step(nsub, oracle_threads):
    s = scenes.cloth_config1()
    ctx = xpbd_context(s)
    o = oracle_from_gpu(s, ctx)
    ctx.step(s.dt, nsub, s.gravity)
    o.xpbd_step(s.dt, nsub, s.gravity)
    check_step(ctx, o, s.dt, nsub, label=f"32^2 cloth, {nsub} substeps")


def test_config1_trajectory(oracle_threads):
    s = scenes.cloth_config1()
    ctx = xpbd_context(s)
    o = oracle_from_gpu(s, ctx)
    for _ in range(10):
        ctx.step(s.dt, 5, s.gravity)
        o.xpbd_step(s.dt, 5, s.gravity)
    check_step(ctx, o, s.dt, 5, tol=1e-4, label="32^2 cloth, 10 steps x 5 substeps")


@pytest.mark.slow
def test_config2_single_step_512(oracle_threads):
    s = scenes.cloth_config2(seed=1)
    ctx = xpbd_context(s)
    o = oracle_from_gpu(s, ctx)
    ctx.step(s.dt, 2, s.gravity)
    o.xpbd_step(s.dt, 2, s.gravity)
    check_step(ctx, o, s.dt, 2, label="512^2 cloth, 2 substeps")


def test_cube_snh_single_step(oracle_threads):
    s = scenes.cube(8)
    ctx = xpbd_context(s)
    o = oracle_from_gpu(s, ctx)
    ctx.step(s.dt, 4, s.gravity)
    o.xpbd_step(s.dt, 4, s.gravity)
    check_step(ctx, o, s.dt, 4, label="8^3 SNH cube, 4 substeps")
    st = ctx.stats()
    print(f"FP64 solves: {st['fp64_solves']}")


def test_springs_single_step(oracle_threads):
    x0, tets = scenes.kuhn_cube_mesh(6)
    edges = np.unique(np.sort(np.concatenate([tets[:, [a, b]] for a in range(4) for b in range(a + 1, 4)]), axis=1),
                      axis=0)
    L0 = np.linalg.norm(x0[edges[:, 0]] - x0[edges[:, 1]], axis=1)
    pins = np.flatnonzero(x0[:, 0] == 0.0)
    x = x0 * np.array([1.0, 1.03, 1.0])
    s = scenes.Scene(x=x, x0=x0, m=np.full(len(x0), 1e-2),
                     energies=[("anchor", pins[:, None], np.column_stack([x0[pins], np.full(len(pins), 1e5)])),
                               ("spring", edges, np.column_stack([L0, np.full(len(edges), 1e2)]))])
    ctx = xpbd_context(s)
    o = oracle_from_gpu(s, ctx)
    for _ in range(3):
        ctx.step(s.dt, 2, s.gravity)
        o.xpbd_step(s.dt, 2, s.gravity)
    check_step(ctx, o, s.dt, 2, label="spring net, 3 steps x 2 substeps")


def test_zero_mass_particles_match_oracle(oracle_threads):
    """Zero-mass particles are never moved by projections but still fall under gravity (xpbd.cpp:58-66)."""
    s = scenes.cloth(16, young=1e4, pin="top_row")
    s.m[s.x[:, 1] > 0.99] = 0.0
    ctx = xpbd_context(s)
    o = oracle_from_gpu(s, ctx)
    for _ in range(2):
        ctx.step(s.dt, 3, s.gravity)
        o.xpbd_step(s.dt, 3, s.gravity)
    check_step(ctx, o, s.dt, 3, label="cloth with zero-mass top row")


def test_serial_path_matches_oracle(oracle_threads):
    s = scenes.cloth(20, young=1e4, pin="top_row")
    ctx = xpbd_context(s, parallel=False)
    o = oracle_from_gpu(s, ctx)
    ctx.step(s.dt, 2, s.gravity)
    o.xpbd_step(s.dt, 2, s.gravity, parallel=False)
    check_step(ctx, o, s.dt, 2, label="serial Gauss-Seidel, 20^2 cloth")
    # and it is a different Gauss-Seidel order than the colour groups
    o2 = oracle_from_gpu(s, xpbd_context(s))
    o2.xpbd_step(s.dt, 2, s.gravity, parallel=True)
    assert not np.array_equal(o.state()[0], o2.state()[0])


def test_batched_objects_match_independent_oracles(oracle_threads):
    batch = scenes.cloth_batch(count=5, seed=3, lo=8, hi=20)
    cube = scenes.cube(4)
    objs = batch + [cube]
    ctx = xpbd_context(objs)
    oracles = [oracle_from_gpu(sc, ctx, k) for k, sc in enumerate(objs)]
    for _ in range(2):
        ctx.step(scenes.DT, 3, scenes.GRAVITY)
        for o in oracles:
            o.xpbd_step(scenes.DT, 3, scenes.GRAVITY)
    for k, o in enumerate(oracles):
        check_step(ctx, o, scenes.DT, 3, obj=k, label=f"batch object {k}")


@pytest.mark.parametrize("block", [64, 128, 256])
def test_launch_shapes_bitwise_equal(block):
    """Threads of a launch touch disjoint particles: the block size cannot change a single bit."""
    s = scenes.cloth(40, young=1e4, pin="top_row")
    ref = xpbd_context(s, block_size=256, use_graph=0)
    ctx = xpbd_context(s, block_size=block, use_graph=0)
    for _ in range(3):
        ref.step(s.dt, 2, s.gravity)
        ctx.step(s.dt, 2, s.gravity)
    assert all(np.array_equal(a, b) for a, b in zip(ref.state(), ctx.state()))


def test_graph_replay_bitwise():
    s = scenes.cloth_config1()
    a, b = xpbd_context(s, use_graph=0), xpbd_context(s, use_graph=1)
    for _ in range(5):
        a.step(s.dt, 3, s.gravity)
        b.step(s.dt, 3, s.gravity)
        assert np.array_equal(a.state()[0], b.state()[0])  # eager work between replays
    assert np.array_equal(a.state()[1], b.state()[1])


def test_checkpoint_restore_bitwise():
    s = scenes.cloth_config1()
    ctx = xpbd_context(s)
    ctx.step(s.dt, 2, s.gravity)
    ctx.save_state()
    saved = ctx.state()
    ctx.step(s.dt, 2, s.gravity)
    first = ctx.state()
    for _ in range(3):
        ctx.step(s.dt, 2, s.gravity)
    ctx.restore_state()
    assert all(np.array_equal(a, b) for a, b in zip(ctx.state(), saved))
    ctx.step(s.dt, 2, s.gravity)
    assert all(np.array_equal(a, b) for a, b in zip(ctx.state(), first))


def test_xpbdgpu_mirror_matches_capi():
    """spg::solver::XPBDGpu (C++ mirror of the BaseSolver API) gives the C-ABI's result."""
    s = scenes.cloth_config1()
    for parallel in (True, False):
        pg = s.particle_group()
        solver = native.XPBDGpu(parallel_gauss_seidel=parallel)
        solver.add_object(pg)
        solver.set_dt(s.dt)
        solver.set_num_substeps(4)
        solver.set_gravity(s.gravity)
        ctx = xpbd_context(s, parallel=parallel)
        for _ in range(3):
            solver.step()
            ctx.step(s.dt, 4, s.gravity)
        xm, vm = solver.state()
        xc, vc = ctx.state()
        assert np.array_equal(xm, xc) and np.array_equal(vm, vc)


def test_stats_count_launches_and_bytes():
    s = scenes.cloth_config1()
    ctx = xpbd_context(s)
    ctx.set_timing(True)
    ctx.step(s.dt, 2, s.gravity)
    st = ctx.stats()
    n_st = sum(len(stc) for _, stc, _ in s.energies)
    assert st["projection_launches"] == 2 * ctx.num_launches()
    assert st["stencil_projections"] == 2 * n_st
    tri, hin, anc = (len(stc) for _, stc, _ in [e for e in s.energies if e[0] == "stvk"] +
                     [e for e in s.energies if e[0] == "bending"] + [e for e in s.energies if e[0] == "anchor"])
    assert st["algorithmic_bytes"] == 2 * (tri * (56 + 96) + hin * (24 + 128) + anc * (20 + 32))
    assert st["projection_ms"] > 0


def test_invalid_inputs_rejected():
    s = scenes.cloth_config1()
    ctx = xpbd_context(s)
    x, v = ctx.state()
    x[3, 1] = np.nan
    with pytest.raises(native.NativeError) as e:
        ctx.set_state(0, x, v)
    assert e.value.code == native.ERR_NONFINITE
    obj = s.xpbd_object()
    kind, st, prm = obj["energies"][1]
    st = st.copy()
    st[0, 1] = st[0, 0]
    obj["energies"][1] = (kind, st, prm)
    with pytest.raises(native.NativeError, match="names one particle twice"):
        native.XpbdContext([obj])
    with pytest.raises(native.NativeError):
        ctx.step(-1.0, 1, s.gravity)
