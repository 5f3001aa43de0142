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


def train(features, targets, epochs=12, batch=4096, lr=2e-3, seed=1, device=0, history_base=False):
    """A new network trained on (features, targets); returns (net, per-epoch loss). history_base: the network
    learns the change of the previous substep's offset, targets - features[:, 0:3] (the offset columns), and
    its warm start adds the prediction to the history start x~ + (x* - x~)_previous."""
    net = native.Mlp(device=device, seed=seed)
    if history_base:
        targets = targets - features[:, 0:3]
        net.set_history_base(True)
    losses = net.train(features, targets, epochs=epochs, batch=batch, lr=lr, seed=seed, standardize=True)
    return net, losses


def training_scenes(count=256, seed=7, iterations=15):
    """Config-4-style random cloths (side U{16..128}, Y log-U[1e3, 1e5], bending log-U[1e-5, 1e-3])."""
    return scenes.cloth_batch(count=count, seed=seed, iterations=iterations)


def _timed_solve(ctx, h, g, rel_tol, max_iterations, check_every, start=None):
    """One substep from the context's checkpoint: begin, the start (None: x~; "history"; or a network with its
    deferred gate), iterate to rel_tol, end; host-clock time of the whole substep. Returns (iterations, rho, ms,
    start rho as the gate measured it (NaN for x~), warm-start ms)."""
    ctx.restore_state()
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.begin_substep(h, g)
    t1 = time.perf_counter()
    if start == "history":
        ctx.history_start(gate="deferred")
    elif start is not None:
        ctx.warm_start(start[0], gate="deferred", smoothing=start[1])
    t_ws = time.perf_counter() - t1
    it, rho = ctx.iterate_to_tolerance(rel_tol, max_iterations, check_every)
    ctx.end_substep()
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    return it, rho, ms, float(ctx.start_rho()[0]), t_ws * 1e3


def _nets(net):
    """{row-key prefix: network}: a single network is "learned"; a dict names its networks."""
    return net if isinstance(net, dict) else {"learned": net}


def compare(scene, net, steps=10, rel_tol=1e-4, max_iterations=3000, check_every=1, gate="deferred", smoothing=2,
            launch=None, device=0, reference_tol=1e-5, reference_iterations=2000, cadences=(1,), log=None):
    """Per substep of `scene`'s trajectory: iterations and wall time to reach rho <= rel_tol from the
    inertial guess (plain VBD, once per convergence-check cadence in `cadences`; the row's plain_* values are
    those of `check_every`), from the history extrapolation x~ + (x* - x~)_previous (the non-learned start
    the network is measured against) and from each network's prediction (`net`: one network, row keys
    learned_*, or a dict of named networks), the starts behind the deferred gate and checked every
    `check_every` sweeps, all from the same state (device checkpoint). The learned paths keep plain VBD as the
    fallback: the gate sends a start whose residual is above x~'s back to x~, and an accepted one whose
    residual later rises above x~'s is dropped for x~ as well (both counted in *_fell_back, the wasted
    iterations included).

    All paths are measured from the states of one reference trajectory, advanced by plain VBD towards the
    tighter reference_tol (at most reference_iterations sweeps), so that no path's tolerance-sized error
    steers the states the others are measured on."""
    if gate != "deferred":
        raise ValueError("compare measures the deferred gate")
    nets = _nets(net)
    ctx = native.VbdContext([scene.vbd_object()], device=device)
    if launch:
        ctx.set_launch(**launch)
    ctx.set_launch(use_graph=0)
    h, g = scene.dt, scene.gravity
    cadences = tuple(sorted(set(cadences) | {check_every}))
    rows = []
    for t in range(steps):
        ctx.save_state()
        by_cadence = {}
        for ce in cadences:
            it_c, rho_c, ms_c, _, _ = _timed_solve(ctx, h, g, rel_tol, max_iterations, ce)
            by_cadence[ce] = {"iterations": it_c, "rho": rho_c, "ms": ms_c}
        plain = by_cadence[check_every]
        row = {"step": t, "plain_iterations": plain["iterations"], "plain_rho": plain["rho"], "plain_ms": plain["ms"],
               "plain_by_cadence": by_cadence}
        finite = True
        starts = [("history", "history")] + [(k, (n, smoothing)) for k, n in nets.items()]
        for key, start in starts:
            f0 = ctx.stats()["prediction_fallbacks"]
            it, rho, ms, srho, ms_ws = _timed_solve(ctx, h, g, rel_tol, max_iterations, check_every, start)
            finite = finite and bool(np.isfinite(ctx.state()[0]).all())
            row.update({key + "_iterations": it, key + "_rho": rho, key + "_ms": ms, key + "_start_rho": srho,
                        key + "_fell_back": ctx.stats()["prediction_fallbacks"] > f0})
            if key != "history":
                row[key + "_warm_start_ms"] = ms_ws
        # the primary network's fields under their historical names
        first = next(iter(nets))
        row.update({"learned_iterations": row[first + "_iterations"], "learned_rho": row[first + "_rho"],
                    "learned_ms": row[first + "_ms"], "warm_start_ms": row[first + "_warm_start_ms"],
                    "prediction_rho": row[first + "_start_rho"],
                    "accepted": bool(np.isfinite(row[first + "_start_rho"]) and row[first + "_start_rho"] <= 1.0),
                    "fell_back": row[first + "_fell_back"], "finite": finite})
        # advance the reference trajectory
        ctx.restore_state()
        ctx.begin_substep(h, g)
        it_r, rho_r = ctx.iterate_to_tolerance(reference_tol, reference_iterations, 10)
        ctx.end_substep()
        row["reference_iterations"], row["reference_rho"] = it_r, rho_r
        rows.append(row)
        if log:
            log(f"step {t}: plain {plain['iterations']} it {plain['ms']:.2f} ms | " + " | ".join(
                f"{k} {row[k + '_iterations']} it {row[k + '_ms']:.2f} ms (start rho {row[k + '_start_rho']:.3e}"
                f"{', fell back' if row[k + '_fell_back'] else ''})" for k, _ in starts) +
                f" | reference {it_r} it rho {rho_r:.3e}")
    return rows


def deploy(scene, net, steps=20, rel_tol=1e-4, max_iterations=3000, check_every=1, smoothing=2, launch=None,
           device=0, log=None):
    """Deployment: contexts advance the scene from its initial state, each continuing from its own rel_tol
    solves: plain VBD from x~, the history extrapolation, and each network's prediction (the starts behind the
    deferred gate). Returns per step the iterations and host-clock ms of each."""
    nets = _nets(net)
    paths = [("plain", None), ("history", "history")] + [(k, n) for k, n in nets.items()]
    ctxs = {}
    for name, _ in paths:
        c = native.VbdContext([scene.vbd_object()], device=device)
        if launch:
            c.set_launch(**launch)
        c.set_launch(use_graph=0)
        ctxs[name] = c
    h, g = scene.dt, scene.gravity
    rows = []
    for t in range(steps):
        row = {"step": t}
        for name, start in paths:
            c = ctxs[name]
            c.synchronize()
            t0 = time.perf_counter()
            c.begin_substep(h, g)
            if start == "history":
                c.history_start(gate="deferred")
            elif start is not None:
                c.warm_start(start, gate="deferred", smoothing=smoothing)
            it, rho = c.iterate_to_tolerance(rel_tol, max_iterations, check_every)
            c.end_substep()
            c.synchronize()
            row[name + "_iterations"] = it
            row[name + "_rho"] = rho
            row[name + "_ms"] = (time.perf_counter() - t0) * 1e3
            row[name + "_finite"] = bool(np.isfinite(c.state()[0]).all())
        rows.append(row)
        if log:
            log(f"deploy step {t}: " + " | ".join(f"{n} {row[n + '_iterations']} it {row[n + '_ms']:.2f} ms"
                                                 for n, _ in paths))
    return rows


def summarize(rows, rel_tol=1e-4):
    pi = sum(r["plain_iterations"] for r in rows)
    li = sum(r["learned_iterations"] for r in rows)
    pm = sum(r["plain_ms"] for r in rows)
    lm = sum(r["learned_ms"] for r in rows)
    out = {"steps": len(rows), "plain_iterations": pi, "learned_iterations": li,
           "plain_ms": pm, "learned_ms": lm, "iteration_ratio": pi / max(li, 1), "time_ratio": pm / max(lm, 1e-9),
           "accepted": sum(r["accepted"] for r in rows), "fell_back": sum(r["fell_back"] for r in rows),
           "all_converged": all(r["plain_rho"] <= rel_tol and r["learned_rho"] <= rel_tol and r["finite"]
                                for r in rows)}
    for k in rows[0] if rows else []:
        if k.endswith("_warm_start_ms") and k != "warm_start_ms":
            name = k[:-len("_warm_start_ms")]
            m = sum(r[name + "_ms"] for r in rows)
            out[name] = {"iterations": sum(r[name + "_iterations"] for r in rows), "ms": m,
                         "all_converged": all(r[name + "_rho"] <= rel_tol for r in rows),
                         "fell_back": sum(r[name + "_fell_back"] for r in rows)}
    if rows and "history_ms" in rows[0]:
        hm = sum(r["history_ms"] for r in rows)
        hi = sum(r["history_iterations"] for r in rows)
        cad = sorted(rows[0]["plain_by_cadence"])
        per_cad = {ce: sum(r["plain_by_cadence"][ce]["ms"] for r in rows) for ce in cad}
        best = min(per_cad, key=per_cad.get)
        out.update({"history_iterations": hi, "history_ms": hm,
                    "history_all_converged": all(r["history_rho"] <= rel_tol for r in rows),
                    "plain_ms_by_cadence": per_cad,
                    "plain_iterations_by_cadence": {ce: sum(r["plain_by_cadence"][ce]["iterations"] for r in rows)
                                                    for ce in cad},
                    "best_plain_cadence": best, "best_plain_ms": per_cad[best],
                    "time_ratio_vs_best_plain": per_cad[best] / max(lm, 1e-9),
                    "time_ratio_vs_history": hm / max(lm, 1e-9), "iteration_ratio_vs_history": hi / max(li, 1)})
    return out


def summarize_deploy(rows, rel_tol=1e-4):
    out = {"steps": len(rows)}
    names = [k[:-len("_ms")] for k in rows[0] if k.endswith("_ms")] if rows else []
    for name in names:
        out[name + "_iterations"] = sum(r[name + "_iterations"] for r in rows)
        out[name + "_ms"] = sum(r[name + "_ms"] for r in rows)
        out[name + "_all_converged"] = all(r[name + "_rho"] <= rel_tol and r[name + "_finite"] for r in rows)
    for name in names:
        if name not in ("plain", "history"):
            out[name + "_time_ratio_vs_plain"] = out["plain_ms"] / max(out[name + "_ms"], 1e-9)
            out[name + "_time_ratio_vs_history"] = out["history_ms"] / max(out[name + "_ms"], 1e-9)
    return out
