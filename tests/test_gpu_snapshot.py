"""Binary snapshots on the device path: what brt_snapshot_save writes is the
flock's state as the independent numpy reader sees it, and a run resumed from a
snapshot continues bit for bit."""
import numpy as np
import pytest
import torch

from behavert_amd import workloads as W
from behavert_amd.engine import BrtError, Flock
from behavert_amd.types import SNAP_FEATURES, SNAP_LISTS, SNAP_WANDER, ProxParams, SteerParams
from oracle import snapshot as S

pytestmark = pytest.mark.gpu


def full_flock(ctx, seed=0):
    w = W.plain_4096(seed=seed)
    rng = np.random.default_rng(seed + 100)
    steer = SteerParams.make(weights=(1.0, 0.8, 0.6), radius=4.0, targets=W.moving_targets(0.0, 50.0),
                             target_weight=0.3, wander_weight=0.2, wander_seed=3)
    feats = rng.uniform(0, 1, (w.n, 4)).astype(np.float32)
    wander = rng.uniform(-1, 1, (w.n, 2)).astype(np.float32)
    return w, Flock(ctx, w.grid, w.prox, steer, w.state, features=feats, wander=wander)


def test_snapshot_contents_match_the_flock(ctx, tmp_path):
    w, fl = full_flock(ctx)
    for t in range(3):
        fl.step(w.dt, t, 0.1)
    path = str(tmp_path / "a.brts")
    fl.save(path, 3)
    snap = S.read(path)
    st, go = fl.state(), fl.grid_outputs()
    assert (snap["n"], snap["step"], snap["flags"]) == (w.n, 3, SNAP_LISTS | SNAP_FEATURES | SNAP_WANDER)
    for k in ("pos", "forward", "smooth_accel", "features", "wander"):
        np.testing.assert_array_equal(snap[k].view(np.uint32), st[k].view(np.uint32), err_msg=k)
    np.testing.assert_array_equal(snap["hash"], go["hash"])
    np.testing.assert_array_equal(snap["lists"], go["lists"])
    with open(path, "rb") as f:
        data = f.read()
    assert S.fnv1a64(data[S.HEADER:]) == snap["checksum"]


def test_resume_from_snapshot_is_bit_exact(ctx, tmp_path):
    w, fl = full_flock(ctx)
    for t in range(3):
        fl.step(w.dt, t, 0.1)
    path = str(tmp_path / "b.brts")
    fl.save(path, 3)
    for t in range(3, 7):
        fl.step(w.dt, t, 0.1)
    a = fl.state()
    _, other = full_flock(ctx, seed=5)          # different state, same size
    info = other.load(path, SNAP_LISTS | SNAP_FEATURES | SNAP_WANDER)
    assert info.step == 3
    for t in range(info.step, 7):
        other.step(w.dt, t, 0.1)
    b = other.state()
    for k in ("pos", "forward", "smooth_accel", "features", "wander"):
        np.testing.assert_array_equal(a[k].view(np.uint32), b[k].view(np.uint32), err_msg=k)


def test_snapshot_load_errors(ctx, tmp_path):
    w, fl = full_flock(ctx)
    fl.step(w.dt, 0, 0.1)
    path = tmp_path / "c.brts"
    fl.save(str(path), 1, flags=0)
    with pytest.raises(BrtError):                      # sections the file lacks
        fl.load(str(path), SNAP_LISTS)
    small = W.plain_4096(n=1024)
    fl2 = Flock(ctx, small.grid, small.prox, small.steer, small.state)
    with pytest.raises(BrtError):                      # another flock size
        fl2.load(str(path))
    raw = bytearray(path.read_bytes())
    raw[S.HEADER + 5] ^= 0xFF
    bad = tmp_path / "corrupt.brts"
    bad.write_bytes(bytes(raw))
    with pytest.raises(BrtError):                      # checksum
        fl.load(str(bad))
    fl.load(str(path))                                 # the intact file still loads


def test_snapshot_lists_need_the_same_k(ctx, tmp_path):
    """Lists saved with K = 64 (W = 68) cannot be copied into lists sized for
    K = 32 (W = 36), nor the other way round: the load is refused before any
    copy, and the state sections still load without BRT_SNAP_LISTS."""
    w = W.plain_4096(n=2048)
    big = Flock(ctx, w.grid, ProxParams.make(64, 4.0), w.steer, w.state)
    big.step(w.dt, 0)
    path = tmp_path / "k64.brts"
    big.save(str(path), 1)
    small = Flock(ctx, w.grid, ProxParams.make(32, 4.0), w.steer, w.state)
    with pytest.raises(BrtError) as e:
        small.load(str(path), SNAP_LISTS)
    assert e.value.code == -1 and "K = 64" in str(e.value)
    small.load(str(path))                               # state only
    torch.cuda.synchronize()
    np.testing.assert_array_equal(small.state()["pos"].view(np.uint32), big.state()["pos"].view(np.uint32))
    small.step(w.dt, 1)
    p32 = tmp_path / "k32.brts"
    small.save(str(p32), 2)
    with pytest.raises(BrtError):
        big.load(str(p32), SNAP_LISTS)
