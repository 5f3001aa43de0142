"""spg_vbd_residual_noalloc (the allocation-free variant of spg_vbd_residual) against spg_vbd_residual, through
the C-ABI: bitwise-equal per-object norms on two calls in a row, the second on a smaller context (so stale work
memory from the first call would show), on a multi-object context, and the size limits rejected with a clear
error."""
import numpy as np
import pytest

from gpu_common import gpu_context
from spg_amd import native, scenes

pytestmark = pytest.mark.gpu


def _open_and_sweep(ctx, s, sweeps):
    ctx.begin_substep(s.dt, s.gravity)
    ctx.iterate(sweeps)


def test_residual_noalloc_matches_residual_large_then_small():
    big = scenes.cloth_config2(seed=3, n=256)
    small = scenes.cloth_config1()
    for s in (big, small):
        ctx = gpu_context(s)
        _open_and_sweep(ctx, s, 3)
        r_alloc = ctx.residual()
        r_free = ctx.residual_noalloc()
        print(f"{s.n} particles: residual {r_alloc[0]:.17e} allocation-free {r_free[0]:.17e}", flush=True)
        assert np.all(np.isfinite(r_alloc)) and r_alloc[0] > 0
        assert np.array_equal(r_alloc, r_free)
        # a second call, after more sweeps, still agrees (nothing cached)
        ctx.iterate(2)
        assert np.array_equal(ctx.residual(), ctx.residual_noalloc())
        ctx.end_substep()


def test_residual_noalloc_multi_object():
    ss = scenes.cloth_batch(count=12, seed=5, lo=8, hi=40, iterations=5)
    ctx = gpu_context(ss)
    _open_and_sweep(ctx, ss[0], 4)
    r_alloc = ctx.residual()
    r_free = ctx.residual_noalloc()
    assert r_alloc.shape == (12,) and np.all(r_alloc > 0)
    assert np.array_equal(r_alloc, r_free)
    ctx.end_substep()


def test_residual_noalloc_limits_and_state():
    s = scenes.cloth_config1()
    ctx = gpu_context(s)
    with pytest.raises(native.NativeError, match="open substep"):
        ctx.residual_noalloc()
    # 4097 single free particles: more objects than the static work memory holds
    objs = [{"x": np.array([[0.0, float(i), 0.0]]), "v": None, "m": np.array([1.0]), "x0": None, "energies": []}
            for i in range(4097)]
    many = native.VbdContext(objs)
    many.begin_substep(s.dt, s.gravity)
    with pytest.raises(native.NativeError, match="more than 4096 objects"):
        many.residual_noalloc()
    assert many.residual().shape == (4097,)
    many.end_substep()
