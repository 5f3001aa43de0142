"""The batched and learned configurations at their own workload sizes, through the C-ABI.

  * Batched VBD: 256 cloths of side U{16..128} (seed 7; 1,558,702 vertices), Y log-U[1e3, 1e5], bending
    log-U[1e-5, 1e-3], top row anchored, 15 iterations. The autotuner picks the launch shape (the context's
    state comes back bitwise), the step is replayed from a hipGraph, and every instance equals its own
    single-object FP64 oracle run (the reference's multi-object VBD loop cannot be the oracle: vbd.cpp:91-96
    never advances objId) to the positions bar of the single cloth, with the colours bit-exact.
  * Learned warm start on the 512x512 cloth: the fused BF16 MFMA inference on all 262,144 feature rows against
    the bfloat16-emulating FP64 network oracle, and learned-vs-plain iterations to a 1e-4 relative residual over
    three steps with a reduced training set.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib as ol
from gpu_common import gpu_context, rel_err
from spg_amd import learned, native, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402

pytestmark = pytest.mark.gpu


def test_config4_batch_autotuned_graph_matches_independent_oracles(oracle_threads):
    batch = scenes.cloth_batch(count=256, seed=7)
    assert sum(s.n for s in batch) == 1558702
    ctx = gpu_context(batch)
    start = [ctx.state(k) for k in range(len(batch))]
    tuned = ctx.autotune(trials=2)
    print(f"autotuned launch {tuned}", flush=True)
    for k in range(len(batch)):  # autotuning leaves the state as it found it, bit for bit
        x, v = ctx.state(k)
        assert np.array_equal(x, start[k][0]) and np.array_equal(v, start[k][1])
    ctx.set_launch(use_graph=1)
    ctx.step(scenes.DT, 1, 15, scenes.GRAVITY)
    worst, worst_k = 0.0, -1
    for k, sc in enumerate(batch):
        o = ol.from_scene(sc)
        o.set_state(*start[k])
        assert np.array_equal(ctx.colors(k), o.colors()[0]), k
        o.step(scenes.DT, 1, 15, scenes.GRAVITY)
        e = rel_err(ctx.state(k)[0], o.state()[0])
        if e > worst:
            worst, worst_k = e, k
    print(f"256 instances: worst max-norm rel error {worst:.2e} (instance {worst_k}, side {batch[worst_k].meta['n']})",
          flush=True)
    assert worst < 1e-5


def test_config5_fused_inference_on_all_512_rows():
    s = scenes.cloth_config2(seed=3)
    ctx = gpu_context(s)
    for _ in range(2):  # a substep history, so every feature column carries data
        ctx.step(s.dt, 1, s.iterations, s.gravity)
    ctx.begin_substep(s.dt, s.gravity)
    F = ctx.features(0)
    ctx.end_substep()
    assert F.shape == (262144, 24)
    net = native.Mlp(seed=9)
    mean = F.mean(axis=0).astype(np.float32)
    std = (F.std(axis=0) + 1e-3).astype(np.float32)
    net.set_normalization(mean, std)
    y = net.forward(F)
    ye = mo.forward(net.params(), F, mean, std, emulate=True)
    rel = float(np.linalg.norm(y.astype(np.float64) - ye) / np.linalg.norm(ye))
    print(f"262144 rows: rel error vs the bfloat16-emulating oracle {rel:.2e}", flush=True)
    assert np.all(np.isfinite(y))
    assert rel < 5e-3


def test_config5_learned_vs_plain_on_512_cloth():
    """Six steps of the 512x512 cloth's reference trajectory. The first two substeps have no substep history
    (the inertial-offset features are zero), so the network's gain shows from the third on; the prediction
    stays gated by the residual, and every step of both paths reaches the tolerance."""
    train_set = scenes.cloth_batch(count=48, seed=11, lo=16, hi=64)
    f, t = learned.collect(train_set, steps=8, samples=240_000, rel_tol=1e-5, max_iterations=800, check_every=5)
    net, losses = learned.train(f, t, epochs=10, batch=2048, lr=2e-3, seed=2)
    assert losses[-1] < 0.5 * losses[0]
    scene = scenes.cloth_config2(seed=3)
    rows = learned.compare(scene, net, steps=6, rel_tol=1e-4, max_iterations=3000, check_every=1,
                           log=lambda m: print(m, flush=True))
    sm = learned.summarize(rows)
    print(sm, flush=True)
    assert sm["all_converged"]
    later = rows[2:]
    assert sum(r["learned_iterations"] for r in later) < sum(r["plain_iterations"] for r in later)
    assert sm["learned_iterations"] < sm["plain_iterations"]
