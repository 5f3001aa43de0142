"""Pins of the oracle's rigid-body VBD (RigidBodyGroup + the four rigid-body energies + the VBD branches for
6-DOF blocks) against what the mathematics fixes, independently of the oracle's own code:

  - the axis-angle maps (Eigen::AngleAxis, rigidBodyGroup.cpp:6-16) against scipy's rotation vectors;
  - element gradients and Hessians (TinyAD over centre of mass and a rotation increment applied as
    (I + [theta]x) R, differentiated at theta = 0) against central differences of a numpy energy of the same
    parameterisation, and the anchor against its closed form J^T k J (rigidBodySpringAnchorEnergy.cpp:63-100);
  - a free body: ballistic centre, constant spin of an isotropic body, rotation composed by scipy;
  - the descent's fixed point: after many iterations on one substep the 6-DOF block residual, rebuilt in numpy
    from the inertial prediction (vbd.cpp:62-72,105-119) and the numpy energy's gradient, vanishes;
  - colours, validation messages.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation as Rot

import oracle_lib as ol
from spg_amd import scenes


def skew(v):
    return np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])


def rb_energy_np(kind, q, R, raw):
    """The element energy over q = (com_s, theta_s) per body, with rotation (I + [theta]x) R."""
    ns = ol.RB_STENCIL_SIZE[kind]
    q = np.asarray(q).reshape(ns, 6)
    Rl = [(np.eye(3) + skew(q[s, 3:])) @ R[s] for s in range(ns)]
    if kind == ol.RB_SPRING_ANCHOR:
        p = q[0, :3] + Rl[0] @ raw[:3]
        return 0.5 * raw[6] * np.dot(p - raw[3:6], p - raw[3:6])
    if kind == ol.RB_ORIENTATION_ALIGNMENT:
        c = sum(np.dot(np.cross(Rl[0][:, k], Rl[1][:, k]), np.cross(Rl[0][:, k], Rl[1][:, k])) for k in range(3))
        return 0.5 * raw[0] * c * c
    d = q[0, :3] + Rl[0] @ raw[:3] - q[1, :3] - Rl[1] @ raw[3:6]
    c = np.linalg.norm(d) - raw[6] if kind == ol.RB_SPRING else np.dot(d, d) - raw[6] ** 2
    return 0.5 * raw[7] * c * c


def fd_grad_hess(f, q, eps=1e-5):
    n = len(q)
    g = np.zeros(n)
    H = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = eps
        g[i] = (f(q + e) - f(q - e)) / (2 * eps)
        for j in range(n):
            e2 = np.zeros(n)
            e2[j] = eps
            H[i, j] = (f(q + e + e2) - f(q + e - e2) - f(q - e + e2) + f(q - e - e2)) / (4 * eps * eps)
    return g, H


def random_element(kind, rng):
    ns = ol.RB_STENCIL_SIZE[kind]
    com = rng.normal(size=(ns, 3))
    R = [Rot.from_rotvec(rng.normal(size=3)).as_matrix() for _ in range(ns)]
    if kind == ol.RB_SPRING_ANCHOR:
        raw = np.concatenate([rng.normal(size=3) * 0.5, rng.normal(size=3), [7.0]])
    elif kind == ol.RB_ORIENTATION_ALIGNMENT:
        raw = np.array([3.0])
    else:
        raw = np.concatenate([rng.normal(size=6) * 0.5, [0.8, 5.0]])
    return com, R, raw


def test_axis_angle_maps_match_scipy():
    rng = np.random.default_rng(0)
    cases = [rng.normal(size=3) for _ in range(50)]
    cases += [np.array([np.pi - 1e-6, 0, 0]), np.array([0, 0, 3.1]), np.array([1e-9, -2e-9, 0]),
              np.array([2.0, -2.0, 1.0]), np.array([0, 3.14, 0.01])]
    for t in cases:
        t = t if np.linalg.norm(t) < np.pi else t / np.linalg.norm(t) * (np.pi - 0.1)
        R = ol.rb_aa_to_r(t)
        assert np.abs(R - Rot.from_rotvec(t).as_matrix()).max() < 1e-15 * 8
        back = ol.rb_r_to_aa(R)
        assert np.abs(back - t).max() < 1e-12 * max(1.0, np.abs(t).max()) * (1e3 if np.linalg.norm(t) > 3.1 else 1)
    assert np.array_equal(ol.rb_aa_to_r(np.zeros(3)), np.eye(3))
    assert np.array_equal(ol.rb_r_to_aa(np.eye(3)), np.zeros(3))
    # trace < 0 branches of the quaternion extraction: rotations by ~pi about each axis
    for ax in range(3):
        t = np.zeros(3)
        t[ax] = 3.0
        assert np.allclose(ol.rb_r_to_aa(Rot.from_rotvec(t).as_matrix()), t, atol=1e-12)


@pytest.mark.parametrize("kind", [ol.RB_SPRING, ol.RB_SPRING_SQUARED, ol.RB_SPRING_ANCHOR,
                                  ol.RB_ORIENTATION_ALIGNMENT])
def test_element_derivatives_match_finite_differences(kind):
    rng = np.random.default_rng(kind)
    for _ in range(4):
        com, R, raw = random_element(kind, rng)
        ns = len(R)
        g, H = ol.rb_element(kind, com, np.array(R), raw)
        q0 = np.concatenate([np.concatenate([com[s], np.zeros(3)]) for s in range(ns)])
        gf, Hf = fd_grad_hess(lambda q: rb_energy_np(kind, q, R, raw), q0)
        scale_g = max(1.0, np.abs(g).max())
        scale_h = max(1.0, np.abs(H).max())
        assert np.abs(g - gf).max() < 1e-7 * scale_g, (kind, g, gf)
        assert np.abs(H - Hf).max() < 1e-4 * scale_h, (kind, np.abs(H - Hf).max())
        assert np.abs(H - H.T).max() <= 1e-12 * scale_h


def test_anchor_matches_closed_form():
    """energyGradient / energyHessian of the anchor: J^T k (p - a) and J^T k J with J = [I, -[R l]x]."""
    rng = np.random.default_rng(5)
    com, R, raw = random_element(ol.RB_SPRING_ANCHOR, rng)
    g, H = ol.rb_element(ol.RB_SPRING_ANCHOR, com, np.array(R), raw)
    r = R[0] @ raw[:3]
    J = np.hstack([np.eye(3), -skew(r)])
    p = com[0] + r
    assert np.allclose(g, J.T @ (raw[6] * (p - raw[3:6])), rtol=1e-13, atol=1e-13)
    assert np.allclose(H, raw[6] * J.T @ J, rtol=1e-13, atol=1e-13)


def test_free_body_motion():
    """No energies: the centre falls ballistically (symplectic Euler), an isotropic body keeps its spin and
    its orientation is the initial one composed with exp(omega t)."""
    x0 = np.array([[0.5, 2.0, -1.0]])
    th0 = np.array([[0.3, -0.2, 0.1]])
    o = ol.RigidOracle(x0, th0, [2.0], np.eye(3)[None] * 0.7)
    w0 = np.array([[0.4, 1.1, -0.6]])
    v0 = np.array([[1.0, 0.0, 0.5]])
    o.set_state(v=v0, omega=w0)
    dt, nsub, steps, g = 1 / 60, 2, 10, np.array([0.0, -9.81, 0.0])
    for _ in range(steps):
        o.step(dt, nsub, 5, g)
    x, th, v, w = o.state()
    h = dt / nsub
    n = steps * nsub
    xe = x0[0] + sum((v0[0] + (k + 1) * h * g) * h for k in range(n))
    assert np.allclose(x[0], xe, rtol=1e-12, atol=1e-12)
    assert np.allclose(v[0], v0[0] + n * h * g, rtol=1e-12, atol=1e-12)
    assert np.allclose(w[0], w0[0], rtol=1e-10, atol=1e-12)
    Re = Rot.from_rotvec(w0[0] * h * n) * Rot.from_rotvec(th0[0])
    assert np.allclose(Rot.from_rotvec(th[0]).as_matrix(), Re.as_matrix(), atol=1e-12)


def test_free_body_gyroscopic_predictor():
    """Anisotropic free body, one substep: omega becomes omega + h I^-1 (-omega x I omega) with I the world
    inertia at the start (vbd.cpp:66-69), then the rotation integrates it (rigidBodyGroup.cpp:97-113) and the
    integrated omega (log of the rotation increment over h) equals it."""
    Il = np.diag([0.2, 0.5, 0.9])
    th0 = np.array([0.4, -0.3, 0.8])
    o = ol.RigidOracle([[0, 0, 0]], [th0], [1.0], Il[None])
    w0 = np.array([1.5, -0.7, 2.0])
    o.set_state(omega=[w0])
    h = 0.01
    o.step(h, 1, 3, [0, 0, 0])
    _, th, _, w = o.state()
    R0 = Rot.from_rotvec(th0).as_matrix()
    I = R0 @ Il @ R0.T
    we = w0 + h * np.linalg.solve(I, -np.cross(w0, I @ w0))
    assert np.allclose(w[0], we, rtol=1e-11, atol=1e-13)
    Re = Rot.from_rotvec(we * h).as_matrix() @ R0
    assert np.allclose(Rot.from_rotvec(th[0]).as_matrix(), Re, atol=1e-13)


def block_residual(scene, x_start, th_start, v_start, w_start, x, th, h, g):
    """-dE/dq of the incremental potential at (x, th) for every body, rebuilt in numpy (scipy rotations,
    finite-difference element gradients); forces of the VBD block system at its fixed point."""
    nb = scene.nb
    res = np.zeros((nb, 6))
    Rs = [Rot.from_rotvec(t).as_matrix() for t in th]
    for b in range(nb):
        R0 = Rot.from_rotvec(th_start[b]).as_matrix()
        I0 = R0 @ scene.inertia[b] @ R0.T
        vp = v_start[b] + h * g
        wp = w_start[b] + h * np.linalg.solve(I0, -np.cross(w_start[b], I0 @ w_start[b]))
        xt = x_start[b] + vp * h
        I = Rs[b] @ scene.inertia[b] @ Rs[b].T
        wi = Rot.from_matrix(Rs[b] @ R0.T).as_rotvec() / h
        res[b, :3] = -scene.m[b] / h ** 2 * (x[b] - xt)
        res[b, 3:] = -(1 / h) * I @ (wi - wp)
    for kind, st, raw in scene.energies:
        kind = ol.RB_KIND[kind]
        for e in range(len(st)):
            R = [Rs[b] for b in st[e]]
            q0 = np.concatenate([np.concatenate([x[b], np.zeros(3)]) for b in st[e]])
            gf = np.zeros(len(q0))
            for i in range(len(q0)):
                d = np.zeros(len(q0))
                d[i] = 1e-6
                gf[i] = (rb_energy_np(kind, q0 + d, R, raw[e]) - rb_energy_np(kind, q0 - d, R, raw[e])) / 2e-6
            for s, b in enumerate(st[e]):
                res[b] -= gf[6 * s:6 * s + 6]
    return res


@pytest.mark.parametrize("name", ["anchored", "chain"])
def test_descent_fixed_point(name):
    s = scenes.rb_anchored() if name == "anchored" else scenes.rb_chain(4)
    o = ol.rb_from_scene(s)
    g = np.array(s.gravity)
    for _ in range(3):
        o.step(s.dt, 1, 10, g)
    x0, th0, v0, w0 = o.state()
    o.step(s.dt, 1, 400, g)
    x, th, _, _ = o.state()
    res = block_residual(s, x0, th0, v0, w0, x, th, s.dt, g)
    scale = np.abs(np.array(s.m)).max() / s.dt ** 2 * np.abs(x - x0).max()
    print(name, np.abs(res).max(), scale)
    assert np.abs(res).max() < 1e-6 * scale


def test_colors_and_validation():
    s = scenes.rb_lattice(6)
    o = ol.rb_from_scene(s)
    c, nc = o.colors()
    for kind, st, _ in s.energies:
        if st.shape[1] == 2:
            assert np.all(c[st[:, 0]] != c[st[:, 1]])
    assert nc == 2
    bad = ol.RigidOracle([[0, 0, 0], [1, 0, 0]], np.zeros((2, 3)), [1, 1], np.repeat(np.eye(3)[None], 2, 0))
    with pytest.raises(ol.OracleError, match="Zero/negative rest length invalid"):
        bad.add_energy("rb_spring", [[0, 1]], [[0, 0, 0, 0, 0, 0, 0.0, 1]])
    with pytest.raises(ol.OracleError, match="Negative rest length invalid"):
        bad.add_energy("rb_spring_squared", [[0, 1]], [[0, 0, 0, 0, 0, 0, -1.0, 1]])
    bad.add_energy("rb_spring_squared", [[0, 1]], [[0, 0, 0, 0, 0, 0, 0.0, 1]])


# ---- the solver's closed forms (host build of the GPU code) against the oracle's dual numbers ----

@pytest.mark.parametrize("kind", [ol.RB_SPRING, ol.RB_SPRING_SQUARED, ol.RB_SPRING_ANCHOR,
                                  ol.RB_ORIENTATION_ALIGNMENT])
def test_closed_forms_match_oracle(kind):
    """spg_rigid_element_derivatives on the host (the kernels' device functions compiled for the CPU): gradient
    and per-body diagonal Hessian blocks within 1e-12 relative of the oracle's TinyAD-style derivatives."""
    from spg_amd import native
    rng = np.random.default_rng(100 + kind)
    ns = ol.RB_STENCIL_SIZE[kind]
    coms, rots, raws, gs, hs = [], [], [], [], []
    for _ in range(20):
        com, R, raw = random_element(kind, rng)
        g, H = ol.rb_element(kind, com, np.array(R), raw)
        coms.append(com.ravel())
        rots.append(np.array(R).ravel())
        raws.append(raw)
        gs.append(g)
        hs.append(np.stack([H[6 * s:6 * s + 6, 6 * s:6 * s + 6] for s in range(ns)]))
    g, h = native.rigid_element_derivatives(kind, np.array(coms), np.array(rots), np.array(raws), device=-1)
    for i in range(20):
        sg, sh = max(1.0, np.abs(gs[i]).max()), max(1.0, np.abs(hs[i]).max())
        assert np.abs(g[i] - gs[i]).max() <= 1e-12 * sg
        assert np.abs(h[i] - hs[i]).max() <= 1e-12 * sh



# ---- the C++ mirror of RigidBodyGroup and the rigid-body energies (host only) ----

def test_mirror_rigid_body_group():
    """addBody / setPositions of the mirror: rotation matrices from theta equal the oracle's Eigen restatement
    bitwise; addStencil validation texts are the reference's; XPBDGpu rejects rigid-body groups at step()."""
    from spg_amd import native
    s = scenes.rb_lattice(5)
    rb = native.RigidBodyGroup(s)
    x, th, v, w, R = rb.state()
    assert np.array_equal(x, s.x) and np.array_equal(th, s.theta)
    assert not v.any() and not w.any()
    for i in range(s.nb):
        assert np.array_equal(R[i], ol.rb_aa_to_r(s.theta[i]))
    th2 = np.random.default_rng(1).normal(size=(s.nb, 3))
    rb.set_state(theta=th2)
    for i in range(s.nb):
        assert np.array_equal(rb.state()[4][i], ol.rb_aa_to_r(th2[i]))
    g = native.RigidBodyGroup()
    g.add_bodies(np.zeros((2, 3)), np.ones(2), np.zeros((2, 3)), np.tile(np.eye(3).ravel(), (2, 1)))
    with pytest.raises(native.NativeError, match="Zero/negative rest length invalid"):
        g.add_energy("rb_spring", [[0, 1]], [[0, 0, 0, 0, 0, 0, 0.0, 1.0]])
    with pytest.raises(native.NativeError, match="Negative rest length invalid"):
        g.add_energy("rb_spring_squared", [[0, 1]], [[0, 0, 0, 0, 0, 0, -0.5, 1.0]])
    solver = native.XPBDGpu()
    solver.add_rb_object(g)
    with pytest.raises(native.NativeError, match="VBDGpu only"):
        solver.step()
