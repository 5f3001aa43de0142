"""GPU tests of the learned warm start, through the C-ABI of include/spg_vbd_learned.h:
  * network forward pass and loss gradient (BF16 MFMA kernels, FP32 accumulation) against the FP64
    oracle of oracle/mlp_oracle.py, both with the oracle rounding to bfloat16 where the kernels do
    (tight) and in exact FP64 (loose: the bfloat16 rounding itself);
  * parameters round-trip exactly; training reduces the loss of a smooth regression;
  * the solver's feature rows against the oracle's restatement from the downloaded states;
  * the gate: a prediction with a larger residual than the inertial guess is discarded exactly;
  * what the trained network gains: fewer VBD iterations to a 1e-4 relative residual than plain VBD.
"""
import os
import sys

import numpy as np
import pytest

from spg_amd import learned, native, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu


def _net_with_random_inputs(seed, n):
    rng = np.random.default_rng(seed)
    net = native.Mlp(seed=seed)
    mean = rng.normal(size=24).astype(np.float32) * 0.2
    std = rng.uniform(0.5, 2.0, size=24).astype(np.float32)
    net.set_normalization(mean, std)
    f = (rng.normal(size=(n, 24)) * std + mean).astype(np.float32)
    t = rng.normal(size=(n, 3)).astype(np.float32)
    return net, f, t, mean, std


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_params_roundtrip_exact():
    net = native.Mlp(seed=5)
    p = np.random.default_rng(1).normal(size=native.MLP_NUM_PARAMS).astype(np.float32) * 0.1
    net.set_params(p)
    assert np.array_equal(net.params(), p)
    m, s = np.arange(24, dtype=np.float32), np.arange(1, 25, dtype=np.float32)
    net.set_normalization(m, s)
    m2, s2 = net.normalization()
    assert np.array_equal(m2, m) and np.array_equal(s2, s)


@pytest.mark.parametrize("n", [1, 64, 333, 5000])
def test_forward_matches_oracle(n):
    net, f, _, mean, std = _net_with_random_inputs(10 + n, n)
    y = net.forward(f)
    p = net.params()
    ye = mo.forward(p, f, mean, std, emulate=True)
    yx = mo.forward(p, f, mean, std)
    print(f"n={n} vs bf16-emulating oracle {_rel(y, ye):.2e}, vs FP64 {_rel(y, yx):.2e}", flush=True)
    assert _rel(y, ye) < 5e-3
    assert _rel(y, yx) < 3e-2


@pytest.mark.parametrize("n", [64, 500, 4096])
def test_loss_gradient_matches_oracle(n):
    net, f, t, mean, std = _net_with_random_inputs(20 + n, n)
    loss, g = net.loss_grad(f, t)
    p = net.params()
    le, ge = mo.loss_grad(p, f, t, mean, std, emulate=True)
    lx, gx = mo.loss_grad(p, f, t, mean, std)
    print(f"n={n} loss {loss:.6f} emul {le:.6f} fp64 {lx:.6f}; grad vs emul {_rel(g, ge):.2e} vs fp64 {_rel(g, gx):.2e}",
          flush=True)
    assert abs(loss - le) <= 2e-3 * le
    assert _rel(g, ge) < 1e-2
    assert _rel(g, gx) < 5e-2
    # every layer's block on its own (a transposed or misplaced block fails here)
    k = 0
    for o, i in mo.SHAPES:
        for size in (o * i, o):
            assert _rel(g[k:k + size], ge[k:k + size]) < 3e-2, (k, size)
            k += size


def test_training_reduces_loss():
    rng = np.random.default_rng(3)
    n = 40000
    f = rng.normal(size=(n, 24)).astype(np.float32)
    a = rng.normal(size=(24, 3)) / np.sqrt(24)
    t = np.tanh(f @ a).astype(np.float32)
    net = native.Mlp(seed=4)
    l0, _ = net.loss_grad(f[:4096], t[:4096])
    losses = net.train(f, t, epochs=6, batch=1024, lr=2e-3, seed=5)
    l1, _ = net.loss_grad(f[:4096], t[:4096])
    print(f"loss {l0:.4f} -> {l1:.5f}, epochs {np.round(losses, 5).tolist()}", flush=True)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < losses[0]
    assert l1 < 0.05 * l0


def _step_states(scene, steps):
    ctx = native.VbdContext([scene.vbd_object()])
    states = [ctx.state()]
    for _ in range(steps):
        ctx.step(scene.dt, 1, scene.iterations, scene.gravity)
        states.append(ctx.state())
    return ctx, states


def test_features_match_oracle():
    s = scenes.cloth(20, young=3e3, bending=3e-4, pin="top_row", seed=4, iterations=6, mass=2e-3)
    ctx, st = _step_states(s, 3)
    h, g = s.dt, np.array(s.gravity)
    (xp, vp), (x, v) = st[-2], st[-1]
    xt_prev = xp + h * (vp + h * g)
    corr = x - xt_prev
    ctx.begin_substep(h, s.gravity)
    F = ctx.features(0)
    ref = mo.features(x + h * (v + h * g), v + h * g, corr, mo.static_tables(s), h, g)
    err = np.abs(F - ref)
    print(f"max |F - ref| per column {np.round(err.max(axis=0), 6).tolist()}", flush=True)
    assert np.allclose(F[:, 18:21], ref[:, 18:21], rtol=1e-5, atol=1e-5)
    assert np.allclose(F, ref, rtol=2e-3, atol=2e-3)
    # the inertial-offset columns are not trivially zero in this state
    assert np.abs(ref[:, 0:3]).max() > 0.1
    ctx.end_substep()


def test_features_batched_and_checkpointed():
    a = scenes.cloth(12, young=2e3, pin="top_row", seed=1, iterations=5)
    b = scenes.cloth(17, young=5e4, bending=1e-3, pin="corners", seed=2, iterations=5)
    ctx = native.VbdContext([a.vbd_object(), b.vbd_object()])
    for _ in range(2):
        ctx.step(a.dt, 1, 5, a.gravity)
    ctx.save_state()
    ctx.begin_substep(a.dt, a.gravity)
    F = ctx.features(-1)
    Fa, Fb = ctx.features(0), ctx.features(1)
    assert np.array_equal(F, np.concatenate([Fa, Fb]))
    ctx.iterate(5)
    ctx.end_substep()
    ctx.restore_state()
    ctx.begin_substep(a.dt, a.gravity)
    assert np.array_equal(ctx.features(-1), F)  # the checkpoint carries the substep history
    ctx.end_substep()
    # set_state clears the history of that object only
    x, v = ctx.state(0)
    ctx.set_state(0, x, v)
    ctx.begin_substep(a.dt, a.gravity)
    F2 = ctx.features(-1)
    assert np.all(F2[:a.n, 0:3] == 0) and np.all(F2[:a.n, 6:12] == 0)
    ctx.end_substep()


def _constant_net(bias):
    p = np.zeros(native.MLP_NUM_PARAMS, dtype=np.float32)
    p[-3:] = bias
    net = native.Mlp(seed=1)
    net.set_params(p)
    return net


def test_warm_start_zero_prediction_is_the_inertial_guess():
    s = scenes.cloth(16, young=1e4, pin="top_row", seed=3, iterations=5)
    ctx = native.VbdContext([s.vbd_object()])
    ctx.begin_substep(s.dt, s.gravity)
    acc, rho = ctx.warm_start(_constant_net([0.0, 0.0, 0.0]))
    assert acc[0] and rho[0] == 1.0
    assert np.array_equal(ctx.residual(), ctx.residual_at_inertial())
    ctx.end_substep()


def test_warm_start_gate_discards_a_worse_prediction():
    s = scenes.cloth(16, young=1e4, pin="top_row", seed=3, iterations=5)
    ctx = native.VbdContext([s.vbd_object()])
    ctx.save_state()
    ctx.begin_substep(s.dt, s.gravity)
    acc, rho = ctx.warm_start(_constant_net([0.0, 50.0, 0.0]), gate=False)
    assert acc[0] and rho[0] > 1.0  # an upward jump of 50 h^2 g: worse than x~, kept without the gate
    ctx.end_substep()
    ctx.restore_state()
    ctx.begin_substep(s.dt, s.gravity)
    acc, rho = ctx.warm_start(_constant_net([0.0, 50.0, 0.0]), gate=True)
    assert not acc[0] and rho[0] > 1.0
    assert np.array_equal(ctx.residual(), ctx.residual_at_inertial())
    x, _ = ctx.state()
    ctx.end_substep()
    # with the prediction discarded the solve continues from x~ exactly like plain VBD
    ctx.restore_state()
    ctx.begin_substep(s.dt, s.gravity)
    x2, _ = ctx.state()
    ctx.end_substep()
    assert np.array_equal(x, x2)


def test_trained_network_needs_fewer_iterations():
    """Train on converged trajectories of small random cloths (config-4 style) and compare iterations
    to a 1e-4 relative force residual on an unseen cloth of the 512x512 construction at 96x96."""
    train_set = scenes.cloth_batch(count=24, seed=11, lo=16, hi=48)
    f, t = learned.collect(train_set, steps=8, samples=120_000, rel_tol=1e-5, max_iterations=800, check_every=5)
    net, losses = learned.train(f, t, epochs=10, batch=2048, lr=2e-3, seed=2)
    print(f"{f.shape[0]} samples, losses {np.round(losses, 5).tolist()}", flush=True)
    assert losses[-1] < 0.5 * losses[0]
    scene = scenes.cloth_config2(seed=3, n=96)
    rows = learned.compare(scene, net, steps=6, rel_tol=1e-4, max_iterations=3000, check_every=1,
                           log=lambda m: print(m, flush=True))
    sm = learned.summarize(rows)
    print(sm, flush=True)
    assert sm["all_converged"]
    assert all(r["finite"] for r in rows)
    assert sm["learned_iterations"] < sm["plain_iterations"]


def test_iterate_to_tolerance_reports_the_residual_ratio():
    s = scenes.cloth(24, young=1e4, pin="top_row", seed=5, iterations=5)
    ctx = native.VbdContext([s.vbd_object()])
    ctx.begin_substep(s.dt, s.gravity)
    it, rho = ctx.iterate_to_tolerance(1e-4, max_iterations=2000, check_every=3)
    assert rho <= 1e-4 and it % 3 == 0 and it > 0
    assert rho == pytest.approx(ctx.residual()[0] / ctx.residual_at_inertial()[0], rel=1e-12)
    ctx.end_substep()


def test_iterate_to_tolerance_drops_a_prediction_that_is_worse_than_x_tilde():
    """An accepted start whose residual is above x~'s goes back to x~ at the first check and the solve
    continues exactly as plain VBD (bitwise the same state and iteration count)."""
    s = scenes.cloth(24, young=1e4, pin="top_row", seed=5, iterations=5)
    ctx = native.VbdContext([s.vbd_object()])
    ctx.save_state()
    ctx.begin_substep(s.dt, s.gravity)
    it_p, rho_p = ctx.iterate_to_tolerance(1e-4, max_iterations=2000, check_every=1)
    ctx.end_substep()
    plain = ctx.state()
    ctx.restore_state()
    ctx.reset_stats()
    ctx.begin_substep(s.dt, s.gravity)
    acc, rho0 = ctx.warm_start(_constant_net([0.0, 50.0, 0.0]), gate=False)
    assert acc[0] and rho0[0] > 1.0
    it_l, rho_l = ctx.iterate_to_tolerance(1e-4, max_iterations=2000, check_every=1)
    ctx.end_substep()
    assert ctx.stats()["prediction_fallbacks"] == 1
    assert (it_l, rho_l) == (it_p, rho_p)
    x, v = ctx.state()
    assert np.array_equal(x, plain[0]) and np.array_equal(v, plain[1])


def test_warm_start_smoothing_matches_restatement():
    """x = x~ + s * (one-ring averaged)^k network output, against numpy on the host forward pass."""
    sc = scenes.cloth(14, young=5e3, pin="top_row", seed=8, iterations=5)
    ctx = native.VbdContext([sc.vbd_object()])
    for _ in range(2):
        ctx.step(sc.dt, 1, 5, sc.gravity)
    ctx.begin_substep(sc.dt, sc.gravity)
    xt, _ = ctx.state()
    net = native.Mlp(seed=7)
    y = net.forward(ctx.features(0)).astype(np.float64)
    nb, _, _ = mo.static_tables(sc)
    for _ in range(2):
        y = np.stack([(y[i] + y[j].sum(axis=0)) / (1 + len(j)) for i, j in enumerate(nb)])
    ctx.warm_start(net, gate=False, smoothing=2)
    x, _ = ctx.state()
    s = sc.dt ** 2 * np.linalg.norm(sc.gravity)
    assert np.allclose(x, xt + s * y, rtol=0, atol=2e-6)
    ctx.end_substep()


def test_pipelined_tolerance_loop_equals_plain_sweeps():
    """iterate_to_tolerance checks on a second stream while sweeping ahead and rolls back to the
    passing snapshot: the result is bitwise that of exactly n sweeps, n is the first sweep count whose
    residual passes, and the reported rho is the residual ratio of the final state."""
    sc = scenes.cloth(40, young=2e4, pin="top_row", seed=6, iterations=5)
    ctx = native.VbdContext([sc.vbd_object()])
    ctx.step(sc.dt, 1, 5, sc.gravity)
    ctx.save_state()
    ctx.begin_substep(sc.dt, sc.gravity)
    n, rho = ctx.iterate_to_tolerance(1e-4, max_iterations=3000, check_every=1)
    ctx.end_substep()
    xa, va = ctx.state()
    assert rho <= 1e-4 and n > 2
    ctx.restore_state()
    ctx.begin_substep(sc.dt, sc.gravity)
    ctx.iterate(n)
    assert ctx.residual()[0] / ctx.residual_at_inertial()[0] == pytest.approx(rho, rel=1e-12)
    ctx.end_substep()
    xb, vb = ctx.state()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    ctx.restore_state()
    ctx.begin_substep(sc.dt, sc.gravity)
    ctx.iterate(n - 1)
    assert ctx.residual()[0] / ctx.residual_at_inertial()[0] > 1e-4
    ctx.end_substep()
    # a cap below n stops at the cap with the residual of that state
    ctx.restore_state()
    ctx.begin_substep(sc.dt, sc.gravity)
    m, rho_m = ctx.iterate_to_tolerance(1e-4, max_iterations=n - 2, check_every=1)
    assert m == n - 2 and rho_m > 1e-4
    ctx.end_substep()


def test_deferred_gate_rejects_in_the_first_check():
    """gate='deferred': the first check of iterate_to_tolerance is the gate; a worse prediction goes back
    to x~ and the solve ends bitwise like plain VBD."""
    s = scenes.cloth(24, young=1e4, pin="top_row", seed=5, iterations=5)
    ctx = native.VbdContext([s.vbd_object()])
    ctx.save_state()
    ctx.begin_substep(s.dt, s.gravity)
    it_p, rho_p = ctx.iterate_to_tolerance(1e-4, max_iterations=2000, check_every=1)
    ctx.end_substep()
    plain = ctx.state()
    ctx.restore_state()
    ctx.reset_stats()
    ctx.begin_substep(s.dt, s.gravity)
    acc, rho0 = ctx.warm_start(_constant_net([0.0, 50.0, 0.0]), gate="deferred")
    assert acc[0] and np.isnan(rho0[0])
    it_l, rho_l = ctx.iterate_to_tolerance(1e-4, max_iterations=2000, check_every=1)
    ctx.end_substep()
    assert ctx.stats()["prediction_fallbacks"] == 1
    assert (it_l, rho_l) == (it_p, rho_p)
    x, v = ctx.state()
    assert np.array_equal(x, plain[0]) and np.array_equal(v, plain[1])


def _per_object_net(scale):
    """A network whose output y = (0, scale * silu(silu(silu(f18 - 3))), 0) depends on the log10 Young's
    feature alone (identity normalisation): exactly 0 for Y = 1e3 (silu(0) = 0 through every layer),
    an upward lift for stiffer cloths."""
    p = np.zeros(native.MLP_NUM_PARAMS, dtype=np.float32)
    (w1, b1), (w2, b2), (w3, b3), (w4, b4) = native.mlp_unflatten(p)  # views into p
    w1[0, 18] = 1.0
    b1[0] = -3.0
    w2[0, 0] = 1.0
    w3[0, 0] = 1.0
    w4[1, 0] = scale
    net = native.Mlp(seed=1)
    net.set_params(p)
    return net


def test_batched_gate_and_tolerance_are_per_object():
    """Two cloths in one context. The prediction is exactly x~ for the soft cloth (Y t = 1e3: accepted,
    rho = 1) and a large lift for the stiff one (Y t = 5e4: rejected), so both start from x~ and the
    batched learned solve equals the batched plain solve bitwise, with rho the largest per-object ratio."""
    soft = scenes.cloth(14, young=1e3 / 1e-3, thickness=1e-3, pin="top_row", seed=2, iterations=5)
    stiff = scenes.cloth(18, young=5e4 / 1e-3, thickness=1e-3, pin="top_row", seed=3, iterations=5)
    objs = [stiff.vbd_object(), soft.vbd_object()]
    ctx = native.VbdContext(objs)
    ctx.begin_substep(soft.dt, soft.gravity)
    f = ctx.features(-1)
    assert abs(f[0, 18] - np.log10(5e4)) < 1e-5 and f[-1, 18] == np.float32(3.0)
    acc, rho = ctx.warm_start(_per_object_net(scale=20.0), gate=True)
    print(f"accepted {acc.tolist()} rho {rho.tolist()}", flush=True)
    assert acc.tolist() == [False, True] and rho[0] > 1.0 and rho[1] == 1.0
    n, r = ctx.iterate_to_tolerance(1e-4, max_iterations=3000, check_every=1)
    res, res0 = ctx.residual(), ctx.residual_at_inertial()
    assert r <= 1e-4 and r == pytest.approx(np.max(res / res0), rel=1e-12)
    ctx.end_substep()
    ref = native.VbdContext(objs)
    ref.begin_substep(soft.dt, soft.gravity)
    n_ref, r_ref = ref.iterate_to_tolerance(1e-4, max_iterations=3000, check_every=1)
    ref.end_substep()
    assert (n, r) == (n_ref, r_ref)
    for o in range(2):
        assert all(np.array_equal(a, b) for a, b in zip(ctx.state(o), ref.state(o)))


def test_history_start_is_the_previous_offset_and_gated():
    """spg_vbd_history_start: without history the start is x~ exactly (rho 1); after a substep it is
    x~ + (x* - x~)_previous in FP32; the deferred gate's measurement equals the immediate gate's."""
    s = scenes.cloth(20, young=3e3, pin="top_row", seed=4, iterations=6)
    ctx = native.VbdContext([s.vbd_object()])
    ctx.begin_substep(s.dt, s.gravity)
    xt0, _ = ctx.state()
    acc, rho = ctx.history_start(gate=True)
    assert acc[0] and rho[0] == 1.0
    assert np.array_equal(ctx.state()[0], xt0)
    ctx.iterate(6)
    x1, _ = ctx.state()
    ctx.end_substep()
    ctx.save_state()
    ctx.begin_substep(s.dt, s.gravity)
    xt1, _ = ctx.state()
    ctx.history_start(gate=False)
    xh, _ = ctx.state()
    f32 = np.float32
    expect = xt1.astype(f32) + (x1.astype(f32) - xt0.astype(f32))
    assert np.array_equal(xh, expect.astype(np.float64))
    assert np.abs(xh - xt1).max() > 0
    ctx.end_substep()
    ctx.restore_state()
    ctx.begin_substep(s.dt, s.gravity)
    _, rho_now = ctx.history_start(gate=True)
    ctx.end_substep()
    ctx.restore_state()
    ctx.begin_substep(s.dt, s.gravity)
    acc_d, rho_d = ctx.history_start(gate="deferred")
    assert acc_d[0] and np.isnan(rho_d[0])
    ctx.iterate_to_tolerance(1e-4, max_iterations=2000, check_every=1)
    ctx.end_substep()
    print(f"history start rho: immediate gate {rho_now[0]:.6e}, deferred gate {ctx.start_rho()[0]:.6e}", flush=True)
    assert ctx.start_rho()[0] == pytest.approx(rho_now[0], rel=1e-12)
