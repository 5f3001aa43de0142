"""Pins the oracle's restatement of spg's further particle energies (oracle/vbd_oracle.cpp) before they check
the GPU paths: continuum springs (springContinuumEnergy.cpp, springSquaredContinuumEnergy.cpp), Baraff-Witkin
and Choi membranes (membraneBaraffWitkinEnergy.cpp, membraneChoiEnergy.cpp), quadratic and Baraff-Witkin
bending (quadraticBendingEnergy.cpp, baraffWitkinBendingEnergy.cpp), the StVK tetrahedron (stvkEnergy.cpp)
and the stable squared Neo-Hookean tetrahedron (stableSquaredNeoHookeanEnergy.cpp). The reference has no
tests or vectors for them, so the checks come from mathematics:

  * energies at configurations with a closed-form value (uniform stretch, pure shear, a folded hinge), and
    zero energy with zero gradient at rest;
  * invariance under rigid motions;
  * the dual-number gradient and Hessian against central differences of the energy;
  * the Baraff-Witkin bending gradient against central differences of the angle with the normalisations
    frozen at the base point (the reference differentiates only through the unnormalised vectors);
  * the StVK tetrahedron's energy form (what VBD and the implicit solvers differentiate) against
    1/2 C^T K C of its constraint form (what XPBD projects);
  * the addStencil / preparePrecomputations validation messages.
"""
import numpy as np
import pytest

import oracle_lib as ol

RNG = np.random.default_rng(11)
TRI0 = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.1, 0.0]])
TET0 = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.1, 0.0], [0.0, 0.0, 0.1]])
HINGE0 = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.05, 0.08, 0.0], [0.04, -0.07, 0.0]])
SEG0 = np.array([[0.0, 0.0, 0.0], [0.2, 0.0, 0.0]])

RAW = {"spring_continuum": [0.2, 50.0], "spring_squared_continuum": [0.2, 50.0],
       "membrane_bw": [30.0, 20.0, 5.0], "membrane_choi": [30.0, 20.0, 5.0], "bending_quadratic": [0.0, 1e-2],
       "bending_bw": [0.0, 1e-2], "stvk_tet": [1e4, 0.3], "ssnh": [1e4, 0.3]}
REST = {"spring_continuum": SEG0, "spring_squared_continuum": SEG0, "membrane_bw": TRI0, "membrane_choi": TRI0,
        "bending_quadratic": HINGE0, "bending_bw": HINGE0, "stvk_tet": TET0, "ssnh": TET0}
KINDS = list(RAW)


def prepared(kind, x0=None, raw=None):
    x0 = REST[kind] if x0 is None else x0
    o = ol.Oracle(x0, x0, np.ones(len(x0)))
    o.add_energy(kind, np.arange(len(x0))[None, :], np.asarray(raw if raw is not None else RAW[kind])[None, :])
    o.prepare()
    return o.prepared(0)[0]


def energy(kind, x, p):
    return ol.element(kind, np.asarray(x).ravel(), p, mode=ol.AD)[0]


def rotation(seed):
    q = np.random.default_rng(seed).normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def deformed(kind, scale=0.03, seed=0):
    return REST[kind] + np.random.default_rng(seed).normal(0, scale * 0.1, size=REST[kind].shape)


@pytest.mark.parametrize("kind", KINDS)
def test_rest_state_has_zero_energy_and_gradient(kind):
    p = prepared(kind)
    E, g, _ = ol.element(kind, REST[kind].ravel(), p, mode=ol.AD)
    assert abs(E) < 1e-20 and np.abs(g).max() < 1e-9


@pytest.mark.parametrize("kind", KINDS)
def test_rigid_motion_invariance(kind):
    p = prepared(kind)
    x = deformed(kind)
    R = rotation(3)
    y = x @ R.T + np.array([0.3, -0.2, 0.7])
    E0, E1 = energy(kind, x, p), energy(kind, y, p)
    assert E0 > 0 and abs(E1 - E0) <= 1e-9 * E0


@pytest.mark.parametrize("kind", [k for k in KINDS if k != "bending_bw"])
def test_gradient_and_hessian_match_central_differences(kind):
    """(The Baraff-Witkin bending's derivatives are not those of its energy: see the frozen-norm test.)"""
    p = prepared(kind)
    x = deformed(kind).ravel()
    _, g, _, H = ol.element(kind, x, p, mode=ol.AD, full_hess=True)
    eps = 1e-6
    n = x.size
    gfd = np.zeros(n)
    Hfd = np.zeros((n, n))
    for i in range(n):
        xp, xm = x.copy(), x.copy()
        xp[i] += eps
        xm[i] -= eps
        gfd[i] = (energy(kind, xp, p) - energy(kind, xm, p)) / (2 * eps)
        Hfd[:, i] = (ol.element(kind, xp, p, mode=ol.AD)[1] - ol.element(kind, xm, p, mode=ol.AD)[1]) / (2 * eps)
    assert np.abs(g - gfd).max() <= 1e-6 * np.abs(g).max()
    assert np.abs(H - Hfd).max() <= 1e-5 * np.abs(H).max()
    assert np.abs(H - H.T).max() <= 1e-12 * np.abs(H).max()


def test_continuum_springs_closed_forms():
    L0, k = RAW["spring_continuum"]
    p = prepared("spring_continuum")
    ps = prepared("spring_squared_continuum")
    assert p[0] == L0 and p[1] == k * L0
    for length in (0.13, 0.2, 0.31):
        x = np.array([[0.0, 0.0, 0.0], [length, 0.0, 0.0]]) @ rotation(5).T
        # continuum spring = a spring of stiffness k / L0 (E = k L0 / 2 (l / L0 - 1)^2 = k / (2 L0) (l - L0)^2)
        e_spring = energy("spring", x, [L0, k / L0])
        assert energy("spring_continuum", x, p) == pytest.approx(e_spring, rel=1e-12, abs=1e-300)
        # squared: C = (l / L0)^2 - 1
        assert energy("spring_squared_continuum", x, ps) == pytest.approx(
            0.5 * k * L0 * ((length / L0) ** 2 - 1) ** 2, rel=1e-12, abs=1e-300)


@pytest.mark.parametrize("kind", ["membrane_bw", "membrane_choi"])
def test_membrane_stretch_and_shear(kind):
    kw, kp, kb = RAW[kind]
    area = 0.5 * 0.1 * 0.1
    p = prepared(kind)
    R = rotation(7)
    s, g = 1.3, 0.25
    # stretch s along the material x axis: F = [s e_x, e_y]
    x = TRI0 * np.array([s, 1.0, 1.0])
    E = energy(kind, x @ R.T, p)
    if kind == "membrane_bw":
        expect = 0.5 * area * kw * (s - 1) ** 2
    else:
        d = np.hypot(s, 1.0) / np.sqrt(2.0) - 1
        expect = 0.5 * area * (kw * (s - 1) ** 2 + 2 * kb * d * d)
    assert E == pytest.approx(expect, rel=1e-12)
    # shear: x -> x + g y, F = [e_x, (g, 1, 0)]
    x = TRI0.copy()
    x[:, 0] += g * TRI0[:, 1]
    E = energy(kind, x @ R.T, p)
    f1 = np.hypot(g, 1.0)
    if kind == "membrane_bw":
        expect = 0.5 * area * (kp * (f1 - 1) ** 2 + kb * g * g)
    else:
        dp = np.linalg.norm([1 + g, 1.0]) / np.sqrt(2.0) - 1
        dm = np.linalg.norm([1 - g, -1.0]) / np.sqrt(2.0) - 1
        expect = 0.5 * area * (kp * (f1 - 1) ** 2 + kb * (dp * dp + dm * dm))
    assert E == pytest.approx(expect, rel=1e-12)


def test_quadratic_bending_is_a_constant_quadratic_form():
    p = prepared("bending_quadratic")
    c, keff = p[:4], p[4]
    assert abs(c.sum()) < 1e-12  # translations are in the kernel of L
    # the flat rest hinge is in the kernel of L (the cotangent Laplacian annihilates linear fields)
    assert np.abs(c @ HINGE0).max() < 1e-12
    # Hessian = keff (c c^T) (x) I3 at any configuration
    for seed in (0, 1):
        x = deformed("bending_quadratic", scale=0.3, seed=seed).ravel()
        H = ol.element("bending_quadratic", x, p, mode=ol.AD, full_hess=True)[3]
        assert np.allclose(H, keff * np.kron(np.outer(c, c), np.eye(3)), rtol=1e-12, atol=1e-12 * keff)
    # Bergou's stiffness: keff = k * 3 / (area of the two wings)
    e0, e1, e2 = HINGE0[1] - HINGE0[0], HINGE0[2] - HINGE0[0], HINGE0[3] - HINGE0[0]
    wings = 0.5 * (np.linalg.norm(np.cross(e0, e1)) + np.linalg.norm(np.cross(e0, e2)))
    assert keff == pytest.approx(RAW["bending_quadratic"][1] * 3 / wings, rel=1e-14)


def _frozen_theta(x, base):
    """atan2 of the Baraff-Witkin bending with the three normalisations taken from `base`."""
    def parts(y):
        e0, e3, e4 = y[1] - y[0], y[2] - y[1], y[3] - y[1]
        return e0, np.cross(e0, e3), -np.cross(e0, e4)
    b0, bn1, bn2 = parts(base)
    s0, s1, s2 = 1 / np.linalg.norm(b0), 1 / np.linalg.norm(bn1), 1 / np.linalg.norm(bn2)
    e0, n1, n2 = parts(x)
    a, b = n1 * s1, n2 * s2
    return np.arctan2(np.cross(a, b) @ (e0 * s0), a @ b)


def test_baraff_witkin_bending_differentiates_with_frozen_norms():
    theta0, k = 0.0, RAW["bending_bw"][1]
    p = prepared("bending_bw")
    assert p[0] == pytest.approx(theta0, abs=1e-15) and p[1] == k
    x = HINGE0.copy()
    x[2, 2] += 0.03
    x[3, 2] += 0.01
    E, g, _ = ol.element("bending_bw", x.ravel(), p, mode=ol.AD)
    theta = _frozen_theta(x, x)
    # the value is the dihedral angle itself
    assert E == pytest.approx(0.5 * k * (theta - theta0) ** 2, rel=1e-12)
    eps = 1e-7
    gfd = np.zeros(12)
    for i in range(12):
        xp, xm = x.ravel().copy(), x.ravel().copy()
        xp[i] += eps
        xm[i] -= eps
        gfd[i] = (_frozen_theta(xp.reshape(4, 3), x) - _frozen_theta(xm.reshape(4, 3), x)) / (2 * eps)
    assert np.abs(g - k * (theta - theta0) * gfd).max() <= 1e-6 * np.abs(g).max()
    # the Hessian is the second derivative of 1/2 k (theta_frozen - theta0)^2 at the base point
    H = ol.element("bending_bw", x.ravel(), p, mode=ol.AD, full_hess=True)[3]
    ef = lambda y: 0.5 * k * (_frozen_theta(y.reshape(4, 3), x) - theta0) ** 2  # noqa: E731
    h = 1e-5
    Hfd = np.zeros((12, 12))
    for i in range(12):
        for j in range(12):
            d = np.zeros((4, 12))
            d[:, i] += h * np.array([1, 1, -1, -1])
            d[:, j] += h * np.array([1, -1, 1, -1])
            v = [ef(x.ravel() + d[q]) for q in range(4)]
            Hfd[i, j] = (v[0] - v[1] - v[2] + v[3]) / (4 * h * h)
    assert np.abs(H - Hfd).max() <= 1e-4 * np.abs(H).max()
    # and this is not the derivative of the true angle (which moves the normalisations too)
    _, gd, _ = ol.element("bending", x.ravel(), [theta0, k], mode=ol.AD)
    assert np.abs(gd - g).max() > 1e-3 * np.abs(g).max()


def test_rest_angle_is_recomputed_from_rest_positions():
    x0 = HINGE0.copy()
    x0[3, 2] = 0.02
    p = prepared("bending_bw", x0=x0, raw=[1.0, 1e-2])  # the addStencil angle is replaced
    assert p[0] == pytest.approx(_frozen_theta(x0, x0), rel=1e-14)


@pytest.mark.parametrize("kind", ["stvk_tet", "ssnh"])
def test_tetrahedra_uniform_stretch(kind):
    young, nu = RAW[kind]
    mu = young / (2 * (1 + nu))
    lam = young * nu / ((1 + nu) * (1 - 2 * nu)) + 1e-20
    vol = 0.1 ** 3 / 6
    p = prepared(kind)
    assert p[9] == pytest.approx(lam * vol, rel=1e-15) and p[10] == pytest.approx(mu * vol, rel=1e-15)
    R = rotation(9)
    for s in (0.9, 1.2):
        x = (TET0 * s) @ R.T
        E = energy(kind, x, p)
        if kind == "stvk_tet":
            e = 0.5 * (s * s - 1)  # E = e I
            expect = vol * (0.5 * lam * (3 * e) ** 2 + mu * 3 * e * e)
        else:
            expect = 0.5 * vol * (lam * (s ** 3 - 1) ** 2 + mu * (3 * s * s - 3) ** 2)
        assert E == pytest.approx(expect, rel=1e-12)


def test_stvk_tet_energy_form_equals_constraint_form():
    p = prepared("stvk_tet")
    for seed in range(3):
        x = deformed("stvk_tet", scale=0.5, seed=seed).ravel()
        c, _ = ol.constraints("stvk_tet", x, p)
        assert energy("stvk_tet", x, p) == pytest.approx(0.5 * (p[9] * c[0] ** 2 + p[10] * c[1] ** 2), rel=1e-12)


@pytest.mark.parametrize("kind,raw,x0,msg", [
    ("spring_continuum", [0.0, 1.0], SEG0, "Zero/negative rest length invalid"),
    ("spring_squared_continuum", [-1.0, 1.0], SEG0, "Zero/negative rest length invalid"),
    ("bending_quadratic", [0.1, 1.0], HINGE0, "Rest theta != 0 not supported for quadratic bending"),
    ("bending_bw", [3.2, 1.0], HINGE0, "Invalid restTheta"),
    ("membrane_bw", [1, 1, 1], TRI0 + [0, 0, 0.1], "must have coordinate z=0"),
    ("membrane_choi", [1, 1, 1], TRI0[[0, 2, 1]], "Rest area of membrane Choi triangle is not positive"),
    ("stvk_tet", [1e4, 0.3], TET0[[0, 2, 1, 3]], "SNH tetrahedron rest volume is not positive"),
    ("ssnh", [1e4, 0.3], TET0[[0, 2, 1, 3]], "SSNH tetrahedron rest volume is not positive"),
])
def test_validation_messages(kind, raw, x0, msg):
    with pytest.raises(RuntimeError, match=msg):
        prepared(kind, x0=np.asarray(x0, dtype=float), raw=raw)


@pytest.mark.parametrize("kind", [k for k in KINDS if k != "membrane_choi"])
def test_xpbd_projection_zero_compliance_solves_linearised_constraints(kind):
    """With zero compliance the coupled projection makes J dx = -C up to the 1e-10 regularisation. (Choi's
    four constraint gradients are linearly dependent, |F0|^2 + |F1|^2 = |G+|^2 + |G-|^2 being the
    parallelogram law, so its linearised constraints cannot all be met: see the next test.)"""
    p = prepared(kind)
    x = deformed(kind, scale=0.2).ravel()
    n = len(REST[kind])
    c, J = ol.constraints(kind, x, p)
    nc = len(c)
    w = np.linspace(0.5, 2.0, n)
    dx = ol.xpbd_project(kind, x, w, p, np.zeros((nc, nc)), 1 / 60).ravel()
    assert np.abs(J.T @ dx + c).max() <= 1e-6 * max(np.abs(c).max(), 1e-12)


def test_choi_projection_with_compliance():
    """The four coupled Choi constraints: rank-3 gradient, and with the prepared compliance the projection is
    dx = W J dl with (J^T W J + compliance / dt^2 + 1e-10 I) dl = -C (a 4x4 solve, checked with numpy)."""
    kind = "membrane_choi"
    p = prepared(kind)
    o = ol.Oracle(TRI0, TRI0, np.ones(3))
    o.add_energy(kind, [[0, 1, 2]], [RAW[kind]])
    o.prepare()
    cmp = o.compliance(0).reshape(4, 4)
    x = deformed(kind, scale=0.2).ravel()
    c, J = ol.constraints(kind, x, p)
    assert np.linalg.svd(J, compute_uv=False)[3] < 1e-12 * np.linalg.norm(J)
    w = np.linspace(0.5, 2.0, 3)
    dt = 1 / 60
    W = np.repeat(w, 3)
    A = J.T @ (W[:, None] * J) + cmp / dt ** 2 + 1e-10 * np.eye(4)
    expect = W * (J @ -np.linalg.solve(A, c))
    dx = ol.xpbd_project(kind, x, w, p, cmp, dt).ravel()
    assert np.abs(dx - expect).max() <= 1e-10 * np.abs(expect).max()

