"""Batched multi-flock step (brt_batch_*): every flock of a batch must produce,
bit for bit, what brt_step produces for that flock alone (hash keys and IDs
shifted by the flock's cell base / offset, lists by its offset), whatever the
launch configuration, with and without the hipGraph replay; the autotuner must
leave the caller's state untouched and its cache must round-trip."""
import math
import os

import numpy as np
import pytest
import torch

from behavert_amd import BrtError
from behavert_amd import workloads as W
from behavert_amd.engine import Batch, Flock
from behavert_amd.types import LIST_PAD, ProxParams, SteerParams

from ._parity import first_mismatch

pytestmark = pytest.mark.gpu


def small_batch(seed=0, sizes=(1, 63, 64, 65, 700, 2048, 0, 3000, 129, 5000), density=0.5, K=16):
    g = W.rng(seed)
    sizes = np.asarray(sizes, np.int64)
    sides = [max((max(int(n), 1) / density) ** (1 / 3), 9.0) for n in sizes]
    grids = [W.cube_grid(float(sd)) for sd in sides]
    weights = g.uniform(0.5, 1.5, (len(sizes), 3)).astype(np.float32)
    states = [W.uniform_flock(int(n), float(sd), 100 + f) for f, (n, sd) in enumerate(zip(sizes, sides))]
    return sizes, grids, weights, states, ProxParams.make(K, 4.0)


def steer_with(weights, base: SteerParams | None = None):
    sp = SteerParams()
    import ctypes as C
    C.memmove(C.byref(sp), C.byref(base if base is not None else W.plain_steer()), C.sizeof(SteerParams))
    sp.separation[0], sp.alignment[0], sp.cohesion[0] = (float(x) for x in weights)
    return sp


def concat(states):
    return {k: np.concatenate([s[k] for s in states]) for k in ("pos", "forward", "smooth_accel")}


def compare_flocks(batch, flocks, grids):
    bh = batch.hash.cpu().numpy().view(np.uint32)
    bs = batch.cell_start.cpu().numpy().view(np.uint32)
    be = batch.cell_end.cpu().numpy().view(np.uint32)
    bl = batch.lists.cpu().numpy().view(np.uint32)
    bst = {k: getattr(batch, k).cpu().numpy() for k in ("pos", "forward", "smooth_accel")}
    for f, fl in enumerate(flocks):
        if fl is None:
            continue
        sl = batch.flock_slice(f)
        off = np.uint32(batch.offsets[f])
        cb = np.uint32(batch.cell_bases[f])
        go = fl.grid_outputs()
        h = bh[sl].copy()
        h[:, 0] -= cb
        h[:, 1] -= off
        assert first_mismatch(h, go["hash"]) is None, (f, first_mismatch(h, go["hash"]))
        cs = bs[batch.cell_slice(f, grids[f].num_cells)].copy()
        ce = be[batch.cell_slice(f, grids[f].num_cells)].copy()
        ne = cs != 0xFFFFFFFF
        cs[ne] -= off
        ce[ne] -= off
        assert first_mismatch(cs, go["cell_start"]) is None, f
        assert first_mismatch(ce, go["cell_end"]) is None, f
        L = bl[sl].copy()
        cnt = L[:, 0]
        for r in range(L.shape[0]):
            L[r, 1:1 + cnt[r]] -= off
        assert first_mismatch(L, go["lists"]) is None, (f, first_mismatch(L, go["lists"]))
        st = fl.state()
        for k in ("pos", "forward", "smooth_accel"):
            assert first_mismatch(bst[k][sl].view(np.uint32), st[k].view(np.uint32)) is None, (f, k)


def run_pair(ctx, sizes, grids, weights, states, prox, steps, base_steer=None, wander=None, cfg=None,
             feature_rate=0.0, features=None, dt=0.05):
    steer = base_steer if base_steer is not None else W.plain_steer()
    b = Batch(ctx, grids, sizes, weights, prox, steer, concat(states),
              features=None if features is None else np.concatenate(features),
              wander=None if wander is None else np.concatenate(wander))
    if cfg:
        b.set_config(**cfg)
    flocks = []
    for f, n in enumerate(sizes):
        if n == 0:
            flocks.append(None)
            continue
        flocks.append(Flock(ctx, grids[f], prox, steer_with(weights[f], steer), states[f],
                            features=None if features is None else features[f],
                            wander=None if wander is None else wander[f]))
    for t in range(steps):
        b.step(dt, t, feature_rate)
        for fl in flocks:
            if fl is not None:
                fl.step(dt, t, feature_rate)
    torch.cuda.synchronize()
    return b, flocks


def test_batch_equals_each_flock_alone(ctx):
    sizes, grids, weights, states, prox = small_batch()
    print("flocks", list(sizes), flush=True)
    b, flocks = run_pair(ctx, sizes, grids, weights, states, prox, steps=3)
    compare_flocks(b, flocks, grids)
    b.close()


@pytest.mark.parametrize("cfg", [
    dict(prox_block=64, uniform_max_cells=0, xcd_remap=0, sort_items=8, lds_lists=1, use_graph=0),
    dict(prox_block=128, uniform_max_cells=2, xcd_remap=1, sort_items=16, lds_lists=0, use_graph=1),
    dict(prox_block=256, uniform_max_cells=8, xcd_remap=1, sort_items=8, lds_lists=1, use_graph=1),
    dict(prox_block=64, uniform_max_cells=1, xcd_remap=1, sort_items=16, lds_lists=0, use_graph=0),
])
def test_batch_configurations_agree(ctx, cfg):
    sizes, grids, weights, states, prox = small_batch(seed=1, sizes=(4000, 1, 2500, 777, 64), density=2.0, K=32)
    print("config", cfg, flush=True)
    b, flocks = run_pair(ctx, sizes, grids, weights, states, prox, steps=2, cfg=cfg)
    compare_flocks(b, flocks, grids)
    b.close()


def test_batch_wander_targets_features(ctx):
    sizes, grids, weights, states, prox = small_batch(seed=2, sizes=(1500, 300, 2200))
    steer = W.plain_steer()
    steer.target_weight = 0.5
    steer.set_targets(W.moving_targets(0.0, 10.0, 4))
    steer.wander_weight = 0.3
    steer.wander_seed = 77
    g = W.rng(5)
    wander = [np.zeros((int(n), 2), np.float32) for n in sizes]
    feats = [g.random((int(n), 4)).astype(np.float32) for n in sizes]
    b, flocks = run_pair(ctx, sizes, grids, weights, states, prox, steps=4, base_steer=steer, wander=wander,
                         features=feats, feature_rate=0.1)
    compare_flocks(b, flocks, grids)
    bw = b.wander.cpu().numpy()
    bf = b.features.cpu().numpy()
    for f, fl in enumerate(flocks):
        st = fl.state()
        sl = b.flock_slice(f)
        assert first_mismatch(bw[sl].view(np.uint32), st["wander"].view(np.uint32)) is None, f
        assert first_mismatch(bf[sl].view(np.uint32), st["features"].view(np.uint32)) is None, f
    b.close()


def test_batch_first_flock_against_oracle(ctx, orc):
    sizes, grids, weights, states, prox = small_batch(seed=3, sizes=(3000, 1000))
    b = Batch(ctx, grids, sizes, weights, prox, W.plain_steer(), concat(states))
    b.step(0.05, 0)
    torch.cuda.synchronize()
    for f in range(2):
        hf = orc.HostFlock(grids[f], prox, states[f])
        hf.step(steer_with(weights[f]), 0.05, 0)
        sl = b.flock_slice(f)
        for k, a in (("pos", hf.pos), ("forward", hf.forward), ("smooth_accel", hf.smooth_accel)):
            got = getattr(b, k).cpu().numpy()[sl]
            assert first_mismatch(got.view(np.uint32), a.view(np.uint32)) is None, (f, k)
    b.close()


def test_graph_replay_matches_direct_launches(ctx):
    sizes, grids, weights, states, prox = small_batch(seed=4, sizes=(2000, 5000, 100))
    outs = []
    for use_graph in (1, 0):
        b = Batch(ctx, grids, sizes, weights, prox, W.plain_steer(), concat(states))
        b.set_config(use_graph=use_graph)
        for t in range(5):
            b.step(0.05, t)
            torch.cuda.synchronize()     # eager work between replays
        outs.append((b.pos.cpu().numpy().copy(), b.lists.cpu().numpy().copy(), b.hash.cpu().numpy().copy()))
        b.close()
    for a, c in zip(outs[0], outs[1]):
        assert first_mismatch(a.view(np.uint32), c.view(np.uint32)) is None


@pytest.mark.parametrize("use_graph", [0, 1])
def test_batch_without_sorted_smooth_accel_is_the_same(ctx, use_graph):
    """A batch stepped with sorted_smooth_accel NULL (own smoothed acceleration
    read by ID) equals the one with the sorted copy, bit for bit, direct and
    graph-replayed, with wander and features."""
    sizes, grids, weights, states, prox = small_batch(seed=6, sizes=(1500, 300, 2200), K=32)
    steer = W.plain_steer()
    steer.wander_weight = 0.3
    steer.wander_seed = 5
    g = W.rng(6)
    wander = np.zeros((int(sizes.sum()), 2), np.float32)
    feats = g.random((int(sizes.sum()), 4)).astype(np.float32)
    outs = []
    for keep in (True, False):
        b = Batch(ctx, grids, sizes, weights, prox, steer, concat(states), features=feats, wander=wander,
                  sorted_smooth_accel=keep)
        b.set_config(use_graph=use_graph)
        for t in range(4):
            b.step(0.05, t, 0.1)
            torch.cuda.synchronize()
        outs.append([getattr(b, k).cpu().numpy().copy() for k in
                     ("pos", "forward", "smooth_accel", "features", "wander", "hash", "lists")])
        b.close()
    for k, (a, c) in enumerate(zip(*outs)):
        assert first_mismatch(a.view(np.uint32), c.view(np.uint32)) is None, k


def test_autotune_restores_state_and_caches(ctx, tmp_path):
    sizes, grids, weights, states, prox = small_batch(seed=5, sizes=(3000, 1200, 4000, 64))
    b = Batch(ctx, grids, sizes, weights, prox, W.plain_steer(), concat(states))
    before = b.pos.cpu().numpy().copy(), b.forward.cpu().numpy().copy()
    cache = str(tmp_path / "tune.txt")
    c = b.autotune(0.05, steps_per_trial=1, cache_path=cache)
    print("tuned", c.as_dict(), flush=True)
    assert c.prox_block in (64, 128, 256) and c.sort_items in (8, 16) and c.uniform_max_cells <= 2
    after = b.pos.cpu().numpy(), b.forward.cpu().numpy()
    for x, y in zip(before, after):
        assert first_mismatch(x.view(np.uint32), y.view(np.uint32)) is None
    assert os.path.exists(cache)
    lines = open(cache).read().split("\n")
    assert lines[0].startswith("gfx950-4-")
    # a second batch of the same shape takes the cached choice without timing
    b2 = Batch(ctx, grids, sizes, weights, prox, W.plain_steer(), concat(states))
    c2 = b2.autotune(0.05, steps_per_trial=1, cache_path=cache)
    assert c2.as_dict() == c.as_dict()
    assert len(open(cache).read().strip().split("\n")) == 1
    b.close()
    b2.close()


def test_batch_error_paths(ctx):
    sizes, grids, weights, states, prox = small_batch(seed=6, sizes=(100, 200))
    with pytest.raises(BrtError):
        Batch(ctx, grids, sizes, weights, ProxParams.make(16, 4.0, use_knn=True), W.plain_steer(), concat(states))
    b = Batch(ctx, grids, sizes, weights, prox, W.plain_steer(), concat(states))
    with pytest.raises(BrtError):
        b.set_config(prox_block=96)
    with pytest.raises(BrtError):
        b.set_config(sort_items=4)
    buf = b.buffers()
    buf.n = 299
    import ctypes as C
    rc = ctx._lib.brt_batch_step(b.handle, C.byref(b.steer), 0.0, 0.05, 0, C.byref(buf), None)
    assert rc != 0
    b.close()
