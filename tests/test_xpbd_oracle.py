"""Pins the FP64 XPBD restatement of the oracle (oracle/vbd_oracle.cpp, XPBD section) before it is used
as the checker of the GPU XPBD path. The reference has no tests or golden vectors for XPBD, so every
check here comes from mathematics the projection must satisfy, or from an independent restatement:

  * constraint gradients (dual numbers) against central finite differences;
  * the spring's own projection (springEnergy.cpp:48-78) against the generic projection
    (energy.h:117-153) applied to the spring constraint;
  * zero compliance: the projected positions satisfy the linearised constraint (J dx = -C), and a
    spring is put exactly at its rest length;
  * with compliance: dx minimises 1/2 |dx|_M^2 + 1/2 (C + J dx)^T alpha~^-1 (C + J dx), the implicit
    step of the quadratic constraint energy, i.e. M dx + J^T alpha~^-1 (C + J dx) = 0;
  * momentum: translation- and rotation-invariant constraints change neither the linear nor the
    angular momentum of their stencil;
  * closed forms of whole steps: a free particle, a particle on an anchor spring, and the hanging
    equilibrium of the anchor spring, which is a fixed point of the step;
  * colouring: a brute-force pure-Python greedy colouring (bit-exact), validity, and that the
    colour-group path equals the serial path run over the stencils sorted colour-major.
"""
import numpy as np
import pytest

import oracle_lib as ol
from spg_amd import scenes

RNG = np.random.default_rng(20240611)


def _rand_triangle(scale=0.05):
    x0 = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.2, 0.9, 0.0]]) * scale
    u, v = x0[1, :2] - x0[0, :2], x0[2, :2] - x0[0, :2]
    D = np.linalg.inv(np.column_stack([u, v]))
    xs = x0 + RNG.normal(0, 0.1 * scale, size=(3, 3))
    return xs, D


def _rand_tet(scale=0.05):
    x0 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * scale
    D = np.linalg.inv(np.column_stack([x0[1] - x0[0], x0[2] - x0[0], x0[3] - x0[0]]))
    xs = x0 + RNG.normal(0, 0.1 * scale, size=(4, 3))
    return xs, D


def _rand_hinge():
    xs = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.4, 0.8, 0.0], [0.6, -0.7, 0.0]])
    return xs + RNG.normal(0, 0.15, size=(4, 3))


def _cases():
    """(kind, xs (N,3), prepared params) with non-degenerate random geometry."""
    out = []
    for _ in range(3):
        xs = RNG.normal(size=(2, 3))
        out.append(("spring", xs, np.array([0.7, 50.0])))
        a = RNG.normal(size=3)
        out.append(("anchor", RNG.normal(size=(1, 3)), np.array([*a, 1e3])))
        xs, D = _rand_triangle()
        out.append(("stvk", xs, np.concatenate([D.ravel(), np.eye(3).ravel()])))
        out.append(("bending", _rand_hinge(), np.array([0.3, 1e-3])))
        xs, D = _rand_tet()
        out.append(("snh", xs, np.concatenate([D.ravel(), [1e4, 3e3, 1.3]])))
    return out


CASES = _cases()


@pytest.mark.parametrize("case", range(len(CASES)))
def test_constraint_gradient_matches_finite_differences(case):
    kind, xs, p = CASES[case]
    c, g = ol.constraints(kind, xs.ravel(), p)
    n = xs.size
    fd = np.zeros_like(g)
    for i in range(n):
        h = 1e-6 * max(1.0, abs(xs.ravel()[i]))
        xp, xm = xs.ravel().copy(), xs.ravel().copy()
        xp[i] += h
        xm[i] -= h
        fd[i] = (ol.constraints(kind, xp, p)[0] - ol.constraints(kind, xm, p)[0]) / (2 * h)
    assert np.allclose(g, fd, rtol=1e-6, atol=1e-6 * np.abs(g).max()), (kind, g, fd)


def test_constraint_values_closed_forms():
    # spring: |x0 - x1| - L
    xs = np.array([0.0, 0, 0, 3.0, 4.0, 0])
    assert ol.constraints("spring", xs, [2.0, 1.0])[0][0] == pytest.approx(3.0)
    # anchor at its anchor point: value and gradient are exactly zero
    c, g = ol.constraints("anchor", np.array([1.0, 2.0, 3.0]), [1.0, 2.0, 3.0, 5.0])
    assert c[0] == 0.0 and not np.any(g)
    # membrane: rest shape -> C = 0; uniform stretch by s along x -> E00 = (s^2 - 1)/2
    x0 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    D = np.linalg.inv(np.column_stack([x0[1, :2] - x0[0, :2], x0[2, :2] - x0[0, :2]]))
    p = np.concatenate([D.ravel(), np.zeros(9)])
    assert np.allclose(ol.constraints("stvk", x0.ravel(), p)[0], 0, atol=1e-15)
    xs = x0 * np.array([1.2, 1.0, 1.0])
    assert np.allclose(ol.constraints("stvk", xs.ravel(), p)[0], [(1.44 - 1) / 2, 0, 0], atol=1e-15)
    # bending: flat hinge, one wing rotated by phi about the hinge axis -> theta = +-phi
    phi = 0.4
    xs = np.array([[0, 0, 0], [1, 0, 0], [0.5, 1, 0], [0.5, -np.cos(phi), np.sin(phi)]])
    th = ol.constraints("bending", xs.ravel(), [0.0, 1.0])[0][0]
    assert abs(abs(th) - phi) < 1e-14
    # SNH at F = I: C = (1 - alpha, sqrt(3))
    x0 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    p = np.concatenate([np.eye(3).ravel(), [1.0, 1.0, 1.25]])
    assert np.allclose(ol.constraints("snh", x0.ravel(), p)[0], [-0.25, np.sqrt(3)], atol=1e-15)


def _compliance(kind, scale=1.0):
    nc = ol.num_constraints(kind)
    a = np.eye(nc) * scale
    if nc == 3:
        a = a + 0.1 * scale * (np.ones((3, 3)) - np.eye(3))
    return a.ravel()


def test_spring_closed_form_equals_generic_projection():
    for _ in range(5):
        xs = RNG.normal(size=(2, 3))
        w = RNG.uniform(0.5, 2.0, size=2)
        p = [0.5, 1.0]
        cm = [RNG.uniform(1e-4, 1e-2)]
        own = ol.xpbd_project("spring", xs.ravel(), w, p, cm, 1 / 60)
        gen = ol.xpbd_project("spring", xs.ravel(), w, p, cm, 1 / 60, generic_spring=True)
        assert np.allclose(own, gen, rtol=1e-9, atol=1e-14)


def test_spring_guards():
    xs = np.array([0.0, 0, 0, 1.0, 0, 0])
    # both ends immovable: no correction
    assert not np.any(ol.xpbd_project("spring", xs, [0.0, 0.0], [0.5, 1.0], [0.0], 0.01))
    # coincident ends: no correction
    assert not np.any(ol.xpbd_project("spring", np.zeros(6), [1.0, 1.0], [0.5, 1.0], [0.0], 0.01))


def test_zero_compliance_spring_reaches_rest_length():
    for _ in range(5):
        xs = RNG.normal(size=(2, 3))
        w = RNG.uniform(0.1, 3.0, size=2)
        dx = ol.xpbd_project("spring", xs.ravel(), w, [0.8, 1.0], [0.0], 0.01)
        y = xs + dx
        assert np.linalg.norm(y[1] - y[0]) == pytest.approx(0.8, rel=1e-13)
        # the correction splits in the ratio of the inverse masses
        assert np.allclose(dx[0] * w[1], -dx[1] * w[0], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_zero_compliance_satisfies_linearised_constraint(case):
    kind, xs, p = CASES[case]
    ns = xs.shape[0]
    w = RNG.uniform(0.5, 2.0, size=ns)
    nc = ol.num_constraints(kind)
    c, J = ol.constraints(kind, xs.ravel(), p)
    dx = ol.xpbd_project(kind, xs.ravel(), w, p, np.zeros(nc * nc), 1 / 60).ravel()
    lin = c + J.T @ dx
    assert np.max(np.abs(lin)) <= 1e-8 * max(np.max(np.abs(c)), 1e-12), (kind, lin, c)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_compliant_projection_is_the_implicit_minimiser(case):
    kind, xs, p = CASES[case]
    ns = xs.shape[0]
    w = RNG.uniform(0.5, 2.0, size=ns)
    dt = 1 / 60
    c, J = ol.constraints(kind, xs.ravel(), p)
    scale = float(np.mean(np.sum(J * J, axis=0) * np.mean(w))) * dt * dt  # alpha~ comparable to J W J^T
    cm = _compliance(kind, scale)
    nc = ol.num_constraints(kind)
    dx = ol.xpbd_project(kind, xs.ravel(), w, p, cm, dt).ravel()
    at = cm.reshape(nc, nc) / dt ** 2 + 1e-10 * np.eye(nc)
    Minv = np.repeat(w, 3)
    grad = dx / Minv + J @ np.linalg.solve(at, c + J.T @ dx)
    ref = np.abs(J @ np.linalg.solve(at, c)).max()
    assert np.max(np.abs(grad)) <= 1e-9 * ref, (kind, grad)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_projection_conserves_momentum(case):
    kind, xs, p = CASES[case]
    if kind == "anchor":
        pytest.skip("an anchor is not translation invariant")
    ns = xs.shape[0]
    m = RNG.uniform(0.5, 2.0, size=ns)
    nc = ol.num_constraints(kind)
    dx = ol.xpbd_project(kind, xs.ravel(), 1.0 / m, p, _compliance(kind, 1e-3), 1 / 60)
    lin = (m[:, None] * dx).sum(axis=0)
    ang = (m[:, None] * np.cross(xs, dx)).sum(axis=0)
    ref = np.abs(m[:, None] * dx).max()
    assert np.abs(lin).max() <= 1e-12 * ref
    assert np.abs(ang).max() <= 1e-11 * ref * np.abs(xs).max()


def _single(x, m, energies=(), v=None):
    o = ol.Oracle(np.atleast_2d(x), np.atleast_2d(x), np.atleast_1d(m))
    for kind, st, raw in energies:
        o.add_energy(kind, st, raw)
    o.prepare()
    if v is not None:
        o.set_state(np.atleast_2d(x), np.atleast_2d(v))
    return o


def test_free_particle_closed_form():
    x0, v0, g, dt, n = np.array([0.1, 0.2, 0.3]), np.array([1.0, -2.0, 0.5]), np.array([0, -9.81, 0]), 0.01, 4
    o = _single(x0, 1.0, v=v0)
    o.xpbd_step(dt * n, nsubsteps=n, gravity=g)
    x, v = o.state()
    assert np.allclose(x[0], x0 + n * dt * v0 + dt * dt * g * n * (n + 1) / 2, rtol=1e-13, atol=1e-15)
    assert np.allclose(v[0], v0 + n * dt * g, rtol=1e-13)


def test_zero_mass_particle_falls_freely():
    """Gravity reaches every particle, also those of zero mass (xpbd.cpp:58-66); projections never move them."""
    a = np.array([0.0, 1.0, 0.0])
    o = _single([[0.0, 1.0, 0.0], [0.0, 0.5, 0.0]], [0.0, 1.0],
                energies=[("spring", np.array([[0, 1]], np.int32), np.array([[0.5, 1e3]]))])
    o.xpbd_step(0.01, 1, (0, -9.81, 0))
    x, _ = o.state()
    assert x[0][1] == pytest.approx(1.0 - 0.01 * 0.01 * 9.81, rel=1e-15)


def test_anchor_step_closed_form():
    m, k, dt = 0.02, 40.0, 1 / 60
    a = np.array([0.0, 0.0, 0.0])
    x0, v0 = np.array([0.05, -0.1, 0.02]), np.array([0.3, 0.0, -0.1])
    g = np.array([0.0, -9.81, 0.0])
    o = _single(x0, m, energies=[("anchor", np.array([[0]], np.int32), np.array([[*a, k]]))], v=v0)
    o.xpbd_step(dt, 1, g)
    x, v = o.state()
    xt = x0 + dt * (v0 + dt * g)
    A = 1 / m + (1 / k) / dt ** 2 + 1e-10
    ref = xt - (1 / m) * (xt - a) / A
    assert np.allclose(x[0], ref, rtol=1e-13, atol=1e-16)
    assert np.allclose(v[0], (ref - x0) / dt, rtol=1e-11)


def test_hanging_anchor_equilibrium_is_a_fixed_point():
    m, k, dt = 0.05, 20.0, 1 / 120
    g = np.array([0.0, -9.81, 0.0])
    a = np.array([0.0, 1.0, 0.0])
    xeq = a + m * g / k
    o = _single(xeq, m, energies=[("anchor", np.array([[0]], np.int32), np.array([[*a, k]]))])
    for _ in range(5):
        o.xpbd_step(dt, 1, g)
    x, v = o.state()
    assert np.allclose(x[0], xeq, rtol=0, atol=1e-14)
    assert np.abs(v).max() < 1e-11


def _greedy_colors_reference(st):
    """Independent restatement: node n takes the smallest colour absent from the stencils j < n that
    share a vertex with it (node 0 takes 0)."""
    n = len(st)
    sets = [set(map(int, s)) for s in st]
    col = []
    for i in range(n):
        used = {col[j] for j in range(i) if sets[i] & sets[j]}
        c = 0
        while c in used:
            c += 1
        col.append(c)
    return np.array(col, dtype=np.int32)


@pytest.mark.parametrize("ssize", [1, 2, 3, 4])
def test_color_stencils_bit_exact_brute_force(ssize):
    for trial in range(6):
        nv = 12 + 7 * trial
        ns = 5 + 9 * trial
        st = np.array([RNG.choice(nv, size=ssize, replace=False) for _ in range(ns)], dtype=np.int32)
        col, nc = ol.color_stencils(st)
        assert np.array_equal(col, _greedy_colors_reference(st))
        assert nc == col.max() + 1
        for c in range(nc):  # no vertex twice in one colour
            verts = st[col == c].ravel()
            assert len(verts) == len(set(verts.tolist()))


def test_color_stencils_cloth_counts():
    s = scenes.cloth_config1()
    counts = {}
    for kind, st, _ in s.energies:
        col, nc = ol.color_stencils(st)
        counts[kind] = nc
        assert np.array_equal(col, _greedy_colors_reference(st)) if len(st) < 2500 else True
    assert counts == {"bending": 14, "stvk": 7, "anchor": 1}


def test_parallel_path_equals_serial_over_colour_sorted_stencils():
    """The colour-group path (XPBD(true)) is the serial Gauss-Seidel path over the stencils reordered
    colour-major (ascending index within a colour); within a colour the order cannot matter."""
    s = scenes.cloth_config1()
    o_par = ol.from_scene(s)
    o_ser = ol.Oracle(s.x, s.x0, s.m)
    for kind, st, raw in s.energies:
        col, _ = ol.color_stencils(st)
        order = np.lexsort((np.arange(len(st)), col))
        o_ser.add_energy(kind, st[order], raw[order])
    o_ser.prepare()
    for _ in range(2):
        o_par.xpbd_step(s.dt, 2, s.gravity, parallel=True)
        o_ser.xpbd_step(s.dt, 2, s.gravity, parallel=False)
    xa, va = o_par.state()
    xb, vb = o_ser.state()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)


def test_serial_and_parallel_paths_differ_but_agree_closely():
    """Different Gauss-Seidel orders give different (not bitwise equal) but close one-step results."""
    s = scenes.cloth_config1()
    a, b = ol.from_scene(s), ol.from_scene(s)
    a.xpbd_step(s.dt, 1, s.gravity, parallel=True)
    b.xpbd_step(s.dt, 1, s.gravity, parallel=False)
    xa, xb = a.state()[0], b.state()[0]
    assert not np.array_equal(xa, xb)
    assert np.abs(xa - xb).max() < 1e-3


def test_compliance_layout():
    s = scenes.cloth_config1()
    o = ol.from_scene(s)
    kinds = [k for k, _, _ in s.energies]
    cb = o.compliance(kinds.index("bending"))
    pb = o.prepared(kinds.index("bending"))
    assert np.allclose(cb[:, 0] * pb[:, 1], 1.0, rtol=1e-14)  # keff = 1 / compliance
    cs = o.compliance(kinds.index("stvk")).reshape(-1, 3, 3)
    ps = o.prepared(kinds.index("stvk"))[:, 4:].reshape(-1, 3, 3)
    assert np.allclose(np.einsum("nij,njk->nik", cs, ps), np.eye(3), atol=1e-12)  # C^-1/area . C area = I
    ca = o.compliance(kinds.index("anchor"))
    assert np.allclose(ca[:, 0], 1e-5)
