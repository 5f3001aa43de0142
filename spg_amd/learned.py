"""Learned warm start for the GPU VBD solver: training data from the solver's own trajectories,
training with the library's GPU kernels, and the plain-versus-learned comparison.

All compute runs in libspg_vbd.so (include/spg_vbd_learned.h): the feature rows come from the
solver's feature kernel, training is the library's BF16 MFMA forward / backward and Adam, and the
prediction enters the solve through spg_vbd_warm_start, which keeps it only when its FP64 force
residual is no larger than the inertial guess's.

Training targets: for a substep with features F (computed after the inertial predict) the target is
(x* - x~) / (h^2 |g|), where x* is the substep's solution iterated to a relative residual rel_tol. This
is exactly the next substep's inertial-offset feature (columns 0-2), so a trajectory of K substeps
gives K (features, target) sets.
"""
import time

import numpy as np

from . import native, scenes


def collect(scene_list, steps=10, samples=2_000_000, rel_tol=1e-5, max_iterations=600, check_every=10, seed=1,
            device=0, launch=None, log=None):
    """(features (N, 24), targets (N, 3)) from `steps` converged substeps of the scenes advanced together
    in one context, about samples / steps random vertices per substep."""
    ctx = native.VbdContext([s.vbd_object() for s in scene_list], device=device)
    if launch:
        ctx.set_launch(**launch)
    h, g = scene_list[0].dt, scene_list[0].gravity
    rng = np.random.default_rng(seed)
    per_step = max(1, samples // steps)
    feats, targets = [], []
    prev = None
    for t in range(steps + 1):
        ctx.begin_substep(h, g)
        F = ctx.features(-1)
        if prev is not None:
            idx = rng.choice(F.shape[0], size=min(per_step, F.shape[0]), replace=False)
            idx.sort()
            feats.append(prev[idx])
            targets.append(F[idx, 0:3].copy())
        if t == steps:
            ctx.end_substep()
            break
        it, rho = ctx.iterate_to_tolerance(rel_tol, max_iterations, check_every)
        ctx.end_substep()
        if log:
            log(f"collect substep {t}: {it} iterations, rho {rho:.2e}")
        prev = F
    return np.concatenate(feats), np.concatenate(targets)


def train(features, targets, epochs=12, batch=4096, lr=2e-3, seed=1, device=0):
    """A new network trained on (features, targets); returns (net, per-epoch loss)."""
    net = native.Mlp(device=device, seed=seed)
    losses = net.train(features, targets, epochs=epochs, batch=batch, lr=lr, seed=seed, standardize=True)
    return net, losses


def training_scenes(count=256, seed=7, iterations=15):
    """Config-4-style random cloths (side U{16..128}, Y log-U[1e3, 1e5], bending log-U[1e-5, 1e-3])."""
    return scenes.cloth_batch(count=count, seed=seed, iterations=iterations)


def compare(scene, net, steps=10, rel_tol=1e-4, max_iterations=3000, check_every=1, gate="deferred", smoothing=2,
            launch=None, device=0, reference_tol=1e-5, reference_iterations=2000, log=None):
    """Per substep of `scene`'s trajectory: iterations and wall time to reach rho <= rel_tol from the
    inertial guess (plain VBD) and from the network's gated prediction (learned), both from the same
    state (device checkpoint). The learned path keeps plain VBD as the fallback: the gate (by default
    deferred to the first convergence check) sends a prediction whose residual is above x~'s back to
    x~, and an accepted one whose residual later rises above x~'s is dropped for x~ as well (both
    counted in fell_back, the wasted iterations included).

    Both paths are measured from the states of one reference trajectory, advanced by plain VBD towards
    the tighter reference_tol (at most reference_iterations sweeps), so that neither path's
    tolerance-sized error steers the states the other is measured on."""
    ctx = native.VbdContext([scene.vbd_object()], device=device)
    if launch:
        ctx.set_launch(**launch)
    ctx.set_launch(use_graph=0)
    h, g = scene.dt, scene.gravity
    rows = []
    fallbacks_before = ctx.stats()["prediction_fallbacks"]
    for t in range(steps):
        ctx.save_state()
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.begin_substep(h, g)
        it_p, rho_p = ctx.iterate_to_tolerance(rel_tol, max_iterations, check_every)
        ctx.end_substep()
        ctx.synchronize()
        t_plain = time.perf_counter() - t0
        ctx.restore_state()
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.begin_substep(h, g)
        t1 = time.perf_counter()
        acc, rho0 = ctx.warm_start(net, gate=gate, smoothing=smoothing)
        t_ws = time.perf_counter() - t1
        it_l, rho_l = ctx.iterate_to_tolerance(rel_tol, max_iterations, check_every)
        ctx.end_substep()
        ctx.synchronize()
        t_learned = time.perf_counter() - t0
        x, _ = ctx.state()
        fallbacks = ctx.stats()["prediction_fallbacks"]
        row = {"step": t, "plain_iterations": it_p, "plain_rho": rho_p, "plain_ms": t_plain * 1e3,
               "learned_iterations": it_l, "learned_rho": rho_l, "learned_ms": t_learned * 1e3,
               "warm_start_ms": t_ws * 1e3,
               "prediction_rho": float(rho0[0]) if np.isfinite(rho0[0]) else None, "accepted": bool(acc[0]),
               "fell_back": fallbacks > fallbacks_before, "finite": bool(np.isfinite(x).all())}
        fallbacks_before = fallbacks
        # advance the reference trajectory
        ctx.restore_state()
        ctx.begin_substep(h, g)
        it_r, rho_r = ctx.iterate_to_tolerance(reference_tol, reference_iterations, 10)
        ctx.end_substep()
        row["reference_iterations"], row["reference_rho"] = it_r, rho_r
        rows.append(row)
        if log:
            log(f"step {t}: plain {it_p} it {t_plain * 1e3:.2f} ms | learned {it_l} it {t_learned * 1e3:.2f} ms "
                f"(warm start {t_ws * 1e3:.2f} ms), prediction rho {rho0[0]:.3e} accepted {bool(acc[0])} "
                f"fell back {row['fell_back']} | reference {it_r} it rho {rho_r:.1e}")
    return rows


def summarize(rows, rel_tol=1e-4):
    pi = sum(r["plain_iterations"] for r in rows)
    li = sum(r["learned_iterations"] for r in rows)
    pm = sum(r["plain_ms"] for r in rows)
    lm = sum(r["learned_ms"] for r in rows)
    return {"steps": len(rows), "plain_iterations": pi, "learned_iterations": li,
            "plain_ms": pm, "learned_ms": lm, "iteration_ratio": pi / max(li, 1), "time_ratio": pm / max(lm, 1e-9),
            "accepted": sum(r["accepted"] for r in rows), "fell_back": sum(r["fell_back"] for r in rows),
            "all_converged": all(r["plain_rho"] <= rel_tol and r["learned_rho"] <= rel_tol and r["finite"]
                                 for r in rows)}
