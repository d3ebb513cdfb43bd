"""brt_step_noalloc (the step with its scratch in a static device array, no
allocation) against brt_step through the C ABI: two calls in a row on one
context, the second with a smaller input and another step shape (then again
without the sorted smoothed-acceleration copy), each equal bit for bit to
brt_step on the same input; a call above the size limit is rejected with an
error."""
import ctypes as C

import numpy as np
import pytest

from behavert_amd import BrtError
from behavert_amd import workloads as W
from behavert_amd.engine import Flock, current_stream

from ._parity import first_mismatch

pytestmark = pytest.mark.gpu

NOALLOC_MAX_AGENTS = 1 << 20


def noalloc_step(fl, dt, step, feature_rate=0.0):
    b = fl.buffers()
    fl.ctx.check(fl.ctx._lib.brt_step_noalloc(fl.ctx.handle, C.byref(fl.grid), C.byref(fl.prox), C.byref(fl.steer),
                                              float(feature_rate), float(dt), int(step), C.byref(b),
                                              current_stream()))
    if fl.features is not None and feature_rate != 0.0:
        fl.features, fl.features_out = fl.features_out, fl.features


def same_step(ctx, w, steps, sorted_smooth_accel=True):
    a = Flock(ctx, w.grid, w.prox, w.steer, w.state, features=w.features, wander=w.wander)
    b = Flock(ctx, w.grid, w.prox, w.steer, w.state, features=w.features, wander=w.wander,
              sorted_smooth_accel=sorted_smooth_accel)
    for t in range(steps):
        if w.moving_targets:
            a.steer.set_targets(w.targets_at(t))
            b.steer.set_targets(w.targets_at(t))
        a.step(w.dt, t, w.feature_rate)
        noalloc_step(b, w.dt, t, w.feature_rate)
    sa, sb = a.state(), b.state()
    for k in sa:
        if sa[k] is None:
            continue
        assert first_mismatch(sb[k].view(np.uint32), sa[k].view(np.uint32)) is None, (w.name, k)
    ga, gb = a.grid_outputs(), b.grid_outputs()
    for k in ("hash", "cell_start", "cell_end", "lists"):
        assert first_mismatch(gb[k], ga[k]) is None, (w.name, k)


def test_noalloc_step_equals_step(ctx):
    # first call: the largest flock allowed (the fused step with the packed positions)
    w = W.blobs_1m()
    assert w.n == NOALLOC_MAX_AGENTS
    print(f"no-alloc step: {w.name} ({w.n} agents)", flush=True)
    same_step(ctx, w, 2)
    # second call, smaller: 65,536 agents, K 64 (the two-pass step), features
    # gathered into the scratch, targets and wander
    w = W.full_stack_4m(n=1 << 16)
    w.side = 640.0 * (1 / 64) ** (1 / 3)
    w.state = W.uniform_flock(w.n, w.side, 0)
    w.grid = W.cube_grid(w.side)
    print(f"no-alloc step: {w.name} at {w.n} agents", flush=True)
    same_step(ctx, w, 2)
    # third call, without the sorted smoothed-acceleration copy
    print(f"no-alloc step: {w.name} at {w.n} agents, no sorted smoothed acceleration", flush=True)
    same_step(ctx, w, 2, sorted_smooth_accel=False)


def test_noalloc_step_rejects_more_than_its_maximum(ctx):
    w = W.plain_4096(n=NOALLOC_MAX_AGENTS + 1)
    fl = Flock(ctx, w.grid, w.prox, w.steer, w.state)
    with pytest.raises(BrtError, match="no-allocation"):
        noalloc_step(fl, w.dt, 0)
