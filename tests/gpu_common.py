"""Helpers shared by the GPU parity tests: run a scene on the GPU and on the FP64 oracle from the
same initial state and compare."""
import numpy as np

import oracle_lib as ol
from spg_amd import native


def gpu_context(scenes, **launch):
    if not isinstance(scenes, (list, tuple)):
        scenes = [scenes]
    ctx = native.VbdContext([s.vbd_object() for s in scenes])
    if launch:
        ctx.set_launch(**launch)
    return ctx


def oracle_from_gpu_start(scene, ctx, obj=0):
    """Oracle object for `scene` whose initial state is the GPU's (FP32-rounded) initial state, so the
    comparison isolates the arithmetic of the solve."""
    o = ol.from_scene(scene)
    x, v = ctx.state(obj)
    o.set_state(x, v)
    return o


def rel_err(a, b):
    """max-norm error relative to the max-norm of the reference positions."""
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def disp_rel_err(a, b, start):
    """max-norm error relative to the max-norm of the reference displacement."""
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b - start)), 1e-300))
