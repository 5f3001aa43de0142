"""Pins the FP64 implicit-Euler restatement of the oracle (oracle/vbd_oracle.cpp, implicit section) before it
checks the GPU path. The reference has no tests or vectors for these solvers, and its Eigen submodule
(conjugate gradient, sparse assembly) is empty, so every check comes from mathematics:

  * the stiffness matrix is the derivative of the system force (central differences), is symmetric, and
    the force is gravity * mass minus the energy gradient (against the element gradients);
  * the conjugate gradient restatement: solutions of random SPD systems against a dense solve, Eigen's
    iteration-count convention (a Jacobi-exact diagonal system stops at iteration 0), zero right-hand
    side, and the relative-residual stopping rule;
  * closed forms of whole steps: a free particle under every solver (semi-implicit Euler for
    Baraff-Witkin / Newton, the BDF2 start-up step), and a particle on an anchor spring, whose quadratic
    energy makes one Newton step exact, so Baraff-Witkin, NewtonDv, NewtonDx and NewtonRobust all give
    the implicit Euler solution.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_lib as ol
from spg_amd import scenes

RNG = np.random.default_rng(99)
G = np.array([0.0, -9.81, 0.0])


def test_stiffness_is_force_derivative():
    s = scenes.cloth(5, young=1e3, pin="corners", noise=2e-3)
    o = ol.from_scene(s)
    x = s.x + RNG.normal(0, 1e-3, size=s.x.shape)
    K = o.stiffness(x).toarray()
    assert np.allclose(K, K.T, rtol=0, atol=1e-9 * np.abs(K).max())
    fd = np.zeros_like(K)
    h = 1e-7
    for j in range(K.shape[0]):
        xp, xm = x.ravel().copy(), x.ravel().copy()
        xp[j] += h
        xm[j] -= h
        fd[:, j] = (o.system_force(xp, G) - o.system_force(xm, G)) / (2 * h)
    assert np.abs(K - fd).max() <= 2e-5 * np.abs(K).max()


def test_system_force_is_gravity_minus_gradient():
    s = scenes.cloth(4, young=1e3, pin="corners")
    o = ol.from_scene(s)
    f = o.system_force(s.x, G)
    grad = np.zeros(3 * s.n)
    for e, (kind, st, _) in enumerate(s.energies):
        p = o.prepared(e)
        for i, row in enumerate(st):
            _, g, _ = ol.element(kind, s.x[row].ravel(), p[i], mode=ol.AD)
            for k, v in enumerate(row):
                grad[3 * v:3 * v + 3] += g[3 * k:3 * k + 3]
    ref = np.repeat(s.m, 3) * np.tile(G, s.n) - grad
    assert np.allclose(f, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def _spd(n, cond=1e3):
    q, _ = np.linalg.qr(RNG.normal(size=(n, n)))
    d = np.logspace(0, np.log10(cond), n)
    return (q * d) @ q.T


def _csr(A):
    m = sp.csr_matrix(A)
    m.sort_indices()
    return m.indptr, m.indices, m.data


@pytest.mark.parametrize("n", [5, 30, 120])
def test_cg_solves_spd_systems(n):
    A = _spd(n)
    b = RNG.normal(size=n)
    x, it, err = ol.cg_csr(*_csr(A), b, tol=1e-12, max_iters=10 * n)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-10)
    assert err <= 1e-12
    r = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert abs(r - err) <= 1e-9  # the reported error is the true relative residual


def test_cg_stopping_rule_and_iteration_convention():
    A = _spd(60, cond=1e4)
    b = RNG.normal(size=60)
    x, it, err = ol.cg_csr(*_csr(A), b, tol=1e-5, max_iters=300)
    assert err < 1e-5 and 0 < it < 300
    # one iteration fewer does not reach the tolerance
    x2, it2, err2 = ol.cg_csr(*_csr(A), b, tol=1e-5, max_iters=it)
    assert it2 == it and err2 >= 1e-5
    # Jacobi-exact (diagonal) system: the first update solves it and the loop stops at iteration 0
    D = np.diag(RNG.uniform(1, 10, size=20))
    bd = RNG.normal(size=20)
    xd, itd, errd = ol.cg_csr(*_csr(D), bd)
    assert itd == 0 and np.allclose(xd, bd / np.diag(D), rtol=1e-15)
    # zero right-hand side: zero solution, no iteration
    xz, itz, errz = ol.cg_csr(*_csr(A), np.zeros(60))
    assert itz == 0 and errz == 0 and not np.any(xz)


def _particle(m, energies=(), x=(0.1, 0.5, -0.2), v=(0.3, 0.0, 0.1)):
    x = np.array([x], float)
    o = ol.Oracle(x, x, np.array([m]))
    for kind, st, raw in energies:
        o.add_energy(kind, st, raw)
    o.prepare()
    o.set_state(x, np.array([v], float))
    return o


@pytest.mark.parametrize("solver", [ol.BARAFF_WITKIN, ol.NEWTON_DV, ol.NEWTON_DX, ol.NEWTON_ROBUST])
def test_free_particle_semi_implicit_euler(solver):
    dt, n = 0.01, 3
    o = _particle(0.5)
    x0, v0 = o.state()
    o.implicit_step(solver, dt * n, nsubsteps=n, gravity=G)
    x, v = o.state()
    assert np.allclose(v[0], v0[0] + n * dt * G, rtol=1e-13)
    assert np.allclose(x[0], x0[0] + n * dt * v0[0] + dt * dt * G * n * (n + 1) / 2, rtol=1e-13, atol=1e-15)


def test_free_particle_bdf2_first_step():
    dt = 0.02
    o = _particle(0.5)
    x0, v0 = o.state()
    o.implicit_step(ol.BDF2, dt, 1, G)
    x, v = o.state()
    v1 = v0[0] + 2.0 / 3.0 * dt * G
    assert np.allclose(v[0], v1, rtol=1e-14)
    assert np.allclose(x[0], x0[0] + 2.0 / 3.0 * dt * v1, rtol=1e-14)


@pytest.mark.parametrize("solver", [ol.BARAFF_WITKIN, ol.NEWTON_DV, ol.NEWTON_DX, ol.NEWTON_ROBUST])
def test_anchored_particle_exact_implicit_euler(solver):
    """Quadratic energy: one Newton step is exact. Implicit Euler: (m/h^2 + k)(x1 - xt) = f(xt), xt = x0 + h v0."""
    m, k, dt = 0.05, 30.0, 1 / 60
    a = np.array([0.0, 1.0, 0.0])
    o = _particle(m, energies=[("anchor", np.array([[0]], np.int32), np.array([[*a, k]]))])
    x0, v0 = o.state()
    info = o.implicit_step(solver, dt, 1, G)
    x, v = o.state()
    xt = x0[0] + dt * v0[0]
    f = m * G - k * (xt - a)
    x1 = xt + f / (m / dt ** 2 + k)
    if solver == ol.BARAFF_WITKIN:
        # Baraff-Witkin linearises at x0 (not xt): (m + h^2 k) dv = h (f(x0) - h k v0)
        f0 = m * G - k * (x0[0] - a)
        dv = dt * (f0 - dt * k * v0[0]) / (m + dt * dt * k)
        x1 = x0[0] + dt * (v0[0] + dv)
    assert np.allclose(x[0], x1, rtol=1e-12, atol=1e-15), (x[0], x1)
    assert np.allclose(v[0], (x1 - x0[0]) / dt, rtol=1e-10)
    assert info["linear_solves"] >= 1
    if solver == ol.NEWTON_ROBUST:
        assert info["failed_substeps"] == 0 and info["newton_iterations"] <= 2


def test_robust_newton_cloth_step_and_counters():
    s = scenes.cloth(8, young=1e4, pin="top_row")
    o = ol.from_scene(s)
    info = o.implicit_step(ol.NEWTON_ROBUST, s.dt, 1, s.gravity)
    x, v = o.state()
    assert np.all(np.isfinite(x))
    assert info["newton_iterations"] >= 1 and info["cg_iterations"] >= 1
    assert info["current_dt"] == pytest.approx(s.dt)


def test_baraff_witkin_cloth_solves_its_linear_system():
    """The step is v1 = v0 + dv with (M - h^2 K) dv = h (f + h K v0) solved by CG to a relative residual of
    1e-5, then x1 = x0 + h v1: check the residual of dv = v1 - v0 in that system (built here from the
    oracle's K and f with numpy) and that the CG stopped on its tolerance."""
    s = scenes.cloth(6, young=1e4, pin="top_row")
    o = ol.from_scene(s)
    x0, v0 = o.state()
    v0 = RNG.normal(0, 0.05, size=v0.shape)
    o.set_state(x0, v0)
    dt = s.dt
    K = o.stiffness(x0).toarray()
    f = o.system_force(x0, s.gravity)
    M = np.diag(np.repeat(s.m, 3))
    A = M - dt * dt * K
    rhs = dt * (f + dt * K @ v0.ravel())
    info = o.implicit_step(ol.BARAFF_WITKIN, dt, 1, s.gravity)
    x, v = o.state()
    dv = v.ravel() - v0.ravel()
    assert info["cg_iterations"] < 300
    assert np.linalg.norm(A @ dv - rhs) <= 1.0001e-5 * np.linalg.norm(rhs)
    assert np.allclose(x.ravel(), x0.ravel() + dt * v.ravel(), rtol=1e-14, atol=1e-16)
    # and a dense solve of the same system gives the same step to within its conditioning
    ref = np.linalg.solve(A, rhs)
    assert np.linalg.norm(dv - ref) <= 1e-5 * np.linalg.cond(A) * np.linalg.norm(ref)
