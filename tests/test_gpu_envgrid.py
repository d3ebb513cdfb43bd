"""GPU parity of EnvGrid3D (hash, onesweep sort, reorder, cell ranges) through the
C ABI against the CPU oracle: bit-exact."""
import numpy as np
import pytest
import torch

from behavert_amd import workloads as W
from behavert_amd.types import GridParams

from ._parity import dev, empty_f32, empty_u32, f32, first_mismatch, u32

pytestmark = pytest.mark.gpu


def pos4(xyz):
    xyz = np.asarray(xyz, np.float32).reshape(-1, 3)
    p = np.ones((len(xyz), 4), np.float32)
    p[:, :3] = xyz
    return p


def test_hash_bit_exact_including_edges(ctx, orc):
    g = W.cube_grid(100.0)
    rng = np.random.default_rng(3)
    xyz = rng.uniform(-10, 110, (10000, 3)).astype(np.float32)
    xyz[:25] = np.float32(4.0) * np.arange(25, dtype=np.float32)[:, None]
    xyz[25] = [np.nan, 5, 5]
    xyz[26] = [3e38, -3e38, np.inf]
    p = pos4(xyz)
    h = empty_u32(len(p), 2)
    ctx.calc_hash(g, dev(p), h)
    assert first_mismatch(u32(h), orc.calc_hash(g, p)) is None


def test_hash_non_cubic_grid(ctx, orc):
    g = GridParams.direct(origin=(-10.0, 5.0, 2.5), cell=(2.0, 3.0, 0.5), size=(7, 11, 13))
    rng = np.random.default_rng(4)
    p = pos4(rng.uniform(-20, 40, (5000, 3)))
    h = empty_u32(len(p), 2)
    ctx.calc_hash(g, dev(p), h)
    assert first_mismatch(u32(h), orc.calc_hash(g, p)) is None


SORT_CASES = [(1, 10), (2, 2), (1000, 7), (4095, 15625), (4096, 15625), (4097, 15625), (65543, 1_000_000),
              (300_000, 256), (300_000, 257), (1 << 20, 1_000_000), (1 << 20, 4_096_000), (200_003, 2 ** 31 - 1),
              (100_000, 1)]


@pytest.mark.parametrize("n,cells", SORT_CASES)
def test_sort_bit_exact(ctx, n, cells):
    rng = np.random.default_rng(n + cells)
    k = rng.integers(0, cells, n, dtype=np.uint64).astype(np.uint32)
    h = np.stack([k, np.arange(n, dtype=np.uint32)], 1)
    d = dev(h.view(np.int32))
    ctx.sort(d, cells)
    expect = h[np.lexsort((h[:, 1], h[:, 0]))]
    assert first_mismatch(u32(d), expect) is None


def test_sort_skewed_single_hot_key(ctx):
    n = 500_000
    rng = np.random.default_rng(9)
    k = np.where(rng.random(n) < 0.9, 123456, rng.integers(0, 1_000_000, n)).astype(np.uint32)
    h = np.stack([k, np.arange(n, dtype=np.uint32)], 1)
    d = dev(h.view(np.int32))
    ctx.sort(d, 1_000_000)
    assert first_mismatch(u32(d), h[np.lexsort((h[:, 1], h[:, 0]))]) is None


@pytest.mark.parametrize("make", [W.plain_4096, W.blobs_1m])
def test_hash_sort_reorder_ranges_bit_exact(ctx, orc, make):
    w = make()
    st = w.state
    n = w.n
    hd = empty_u32(n, 2)
    pos = dev(st["pos"])
    fwd = dev(st["forward"])
    sa = dev(st["smooth_accel"])
    ctx.hash_sort(w.grid, pos, hd)
    h = orc.sort_radix(orc.calc_hash(w.grid, st["pos"]), w.grid.num_cells)
    assert first_mismatch(u32(hd), h) is None
    outs = [empty_f32(n, 4) for _ in range(3)]
    cs = empty_u32(w.grid.num_cells)
    ce = empty_u32(w.grid.num_cells)
    ctx.reorder_ranges(hd, [pos, fwd, sa], outs, w.grid.num_cells, cs, ce)
    s, e = orc.cell_ranges(h, w.grid.num_cells)
    assert first_mismatch(u32(cs), s) is None
    assert first_mismatch(u32(ce), e) is None
    for o, key in zip(outs, ("pos", "forward", "smooth_accel")):
        assert first_mismatch(f32(o).view(np.uint32), orc.reorder(h, st[key]).view(np.uint32)) is None
    # the separate entry points agree with the fused one
    outs2 = [empty_f32(n, 4) for _ in range(2)]
    ctx.reorder(hd, [pos, fwd], outs2)
    assert np.array_equal(f32(outs2[0]), f32(outs[0])) and np.array_equal(f32(outs2[1]), f32(outs[1]))
    cs2 = empty_u32(w.grid.num_cells)
    ce2 = empty_u32(w.grid.num_cells)
    ctx.cell_ranges(hd, w.grid.num_cells, cs2, ce2)
    assert np.array_equal(u32(cs2), s) and np.array_equal(u32(ce2), e)


def test_sort_repeated_calls_reuse_scratch(ctx):
    """Back-to-back sorts of different sizes on one context (scratch reuse and
    re-clearing of look-back state)."""
    rng = np.random.default_rng(21)
    for n in (100_000, 5000, 300_000, 1):
        k = rng.integers(0, 40_000, n).astype(np.uint32)
        h = np.stack([k, np.arange(n, dtype=np.uint32)], 1)
        d = dev(h.view(np.int32))
        ctx.sort(d, 40_000)
        assert first_mismatch(u32(d), h[np.lexsort((h[:, 1], h[:, 0]))]) is None
    torch.cuda.synchronize()


def test_sort_lookback_flag_is_reported(ctx):
    """The radix sort's bounded look-back raises a sticky device flag instead of
    hanging; the host reports it at the next synchronising call.  The flag is
    raised here from the host (no kernel is made to stall): brt_context_check,
    brt_phase_times and a flock's state read then fail with BrtError, and the
    check clears it."""
    import ctypes as C

    from behavert_amd import BrtError, lib
    from behavert_amd.engine import Flock
    from behavert_amd.types import DEVFLAG_SORT_LOOKBACK, OPT_DEVICE_FLAGS

    L = lib()
    v = C.c_int(-1)
    ctx.check_device()
    assert L.brt_context_get_option(ctx.handle, OPT_DEVICE_FLAGS, C.byref(v)) == 0 and v.value == 0
    w = W.plain_4096(n=2048)
    fl = Flock(ctx, w.grid, w.prox, w.steer, w.state)
    fl.step(w.dt, 0)
    fl.state()                                      # no flag: reads fine
    ctx.check(L.brt_context_set_option(ctx.handle, OPT_DEVICE_FLAGS, DEVFLAG_SORT_LOOKBACK))
    assert L.brt_context_get_option(ctx.handle, OPT_DEVICE_FLAGS, C.byref(v)) == 0
    assert v.value == DEVFLAG_SORT_LOOKBACK
    with pytest.raises(BrtError) as e:
        fl.state()
    assert e.value.code == -2 and "look-back" in str(e.value)
    ctx.check_device()                              # cleared by the failed check
    ctx.check(L.brt_context_set_option(ctx.handle, OPT_DEVICE_FLAGS, DEVFLAG_SORT_LOOKBACK))
    ms = (C.c_double * 5)()
    steps = C.c_uint32()
    assert L.brt_phase_times(ctx.handle, ms, C.byref(steps)) == -2
    assert L.brt_context_check(ctx.handle) == 0
    fl.step(w.dt, 1)
    fl.state()


def test_steps_keep_sort_metadata_zero_across_standalone_sorts():
    """A step leaves the sort metadata zero (its passes and reorder return what
    they used), so the next step skips the clearing launch; a standalone sort
    leaves it dirty, and the next step clears the whole buffer.  Steps of
    two sizes interleaved with a larger standalone sort and a standalone hash
    on one context give the same bits as the same steps each on a fresh context."""
    from behavert_amd.engine import Context, Flock

    big, small = W.blobs_1m(), W.plain_4096()
    shared = Context(0)
    fb = Flock(shared, big.grid, big.prox, big.steer, big.state)
    fs = Flock(shared, small.grid, small.prox, small.steer, small.state)
    ref_b = Flock(Context(0), big.grid, big.prox, big.steer, big.state)
    ref_s = Flock(Context(0), small.grid, small.prox, small.steer, small.state)
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 60_000, 2_000_000).astype(np.uint32)
    pairs = dev(np.stack([keys, np.arange(len(keys), dtype=np.uint32)], 1).view(np.int32))

    def same(a, b):
        for k in ("hash", "lists", "cell_start", "cell_end"):
            x, y = a.grid_outputs()[k], b.grid_outputs()[k]
            assert np.array_equal(x, y), k
        sa, sb = a.state(), b.state()
        for k in sa:
            assert np.array_equal(np.asarray(sa[k]).view(np.uint32), np.asarray(sb[k]).view(np.uint32)), k

    fb.step(big.dt, 0)
    ref_b.step(big.dt, 0)
    same(fb, ref_b)
    shared.sort(pairs, 60_000)              # 2M keys: more tiles than the 1M step uses
    fs.step(small.dt, 0)                     # smaller step after a dirty standalone sort
    ref_s.step(small.dt, 0)
    same(fs, ref_s)
    fb.step(big.dt, 1)                       # the larger step after the smaller one
    ref_b.step(big.dt, 1)
    same(fb, ref_b)
    hd = empty_u32(small.n, 2)
    shared.calc_hash(small.grid, dev(small.state["pos"]), hd)   # histogram side product
    fb.step(big.dt, 2)
    ref_b.step(big.dt, 2)
    same(fb, ref_b)
    fs.step(small.dt, 1)
    ref_s.step(small.dt, 1)
    same(fs, ref_s)
    shared.check_device()
