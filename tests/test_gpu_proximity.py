"""GPU parity of Proximity3D neighbour lists through the C ABI against the CPU
oracle: bit-exact lists (count, canonical-order prefix, padding)."""
import numpy as np
import pytest

from behavert_amd import workloads as W
from behavert_amd.types import GridParams, ProxParams

from ._parity import dev, empty_u32, first_mismatch, oracle_grid_outputs, u32

pytestmark = pytest.mark.gpu


def gpu_lists(ctx, grid, prox, o):
    n = len(o["sorted_pos"])
    L = empty_u32(n, 4 * (prox.max_neighbors // 4 + 1))
    L.fill_(0x5A5A5A5A)   # every word must be written
    ctx.neighbors(grid, prox, dev(o["sorted_pos"]), dev(o["hash"].view(np.int32)),
                  dev(o["cell_start"].view(np.int32)), dev(o["cell_end"].view(np.int32)), L)
    return u32(L)


@pytest.fixture(params=[1, 0], ids=["lds_lists", "register_lists"])
def list_mode(ctx, request):
    """Runs the test with the lists staged in LDS and with the register-chunk
    output; results must be identical."""
    from behavert_amd.types import OPT_LDS_LISTS
    old = ctx.get_option(OPT_LDS_LISTS)
    ctx.set_option(OPT_LDS_LISTS, request.param)
    yield request.param
    ctx.set_option(OPT_LDS_LISTS, old)


@pytest.mark.parametrize("make", [W.plain_4096, W.blobs_1m])
def test_lists_bit_exact_on_workloads(ctx, orc, make, list_mode):
    w = make()
    o = oracle_grid_outputs(orc, w.grid, w.prox, w.state)
    got = gpu_lists(ctx, w.grid, w.prox, o)
    assert first_mismatch(got, o["lists"]) is None
    if w.name == "blobs_1m":
        # the cap binds in the blob cores
        assert (o["lists"][:, 0] == w.prox.max_neighbors).mean() > 0.3


@pytest.mark.parametrize("K", [1, 3, 4, 16, 63, 64, 252])
def test_lists_bit_exact_across_caps(ctx, orc, K, list_mode):
    rng = np.random.default_rng(K)
    grid = GridParams.direct((0.0,) * 3, (4.0,) * 3, (12,) * 3)
    st = W.make_state(rng.uniform(0, 48, (20000, 3)).astype(np.float32),
                      *W.directions_and_speeds(rng, 20000))
    prox = ProxParams.make(K, 4.0)
    o = oracle_grid_outputs(orc, grid, prox, st)
    assert first_mismatch(gpu_lists(ctx, grid, prox, o), o["lists"]) is None


def test_edges_out_of_grid_nan_coincident_maxbodies(ctx, orc, list_mode):
    rng = np.random.default_rng(5)
    for maxb in (0xFFFFFFFF, 3):
        grid = GridParams.direct((0.0,) * 3, (4.0,) * 3, (8,) * 3, max_bodies=maxb)
        xyz = rng.uniform(-6, 38, (6000, 3)).astype(np.float32)
        xyz[0] = [np.nan, 2, 2]
        xyz[1] = xyz[2] = [5.0, 5.0, 5.0]
        st = W.make_state(xyz, *W.directions_and_speeds(rng, 6000))
        prox = ProxParams.make(12, 4.0)
        o = oracle_grid_outputs(orc, grid, prox, st)
        assert first_mismatch(gpu_lists(ctx, grid, prox, o), o["lists"]) is None


def test_knn_mode_bit_exact(ctx, orc):
    rng = np.random.default_rng(6)
    grid = GridParams.direct((0.0,) * 3, (4.0,) * 3, (6,) * 3)
    st = W.make_state(rng.uniform(0, 24, (8000, 3)).astype(np.float32), *W.directions_and_speeds(rng, 8000))
    prox = ProxParams.make(8, 4.0, use_knn=True)
    o = oracle_grid_outputs(orc, grid, prox, st)
    assert first_mismatch(gpu_lists(ctx, grid, prox, o), o["lists"]) is None


def test_radius_boundary_exact(ctx, orc):
    """d == r accepted, d == nextafter(r) rejected, for many axis-aligned pairs."""
    grid = GridParams.direct((0.0,) * 3, (4.0,) * 3, (10,) * 3)
    base = np.float32(0.5) + np.float32(4.0) * np.arange(8, dtype=np.float32)
    xyz = []
    for b in base:
        xyz.append([b, 2.0, 2.0])
        xyz.append([b + np.float32(4.0), 6.0, 2.0 + 4.0])
    a = np.array(xyz, np.float32)
    far = a.copy()
    far[:, 0] = np.nextafter(far[:, 0] + np.float32(4.0), np.float32(100))
    near = a.copy()
    near[:, 0] = near[:, 0] + np.float32(4.0)
    pts = np.concatenate([a, near, far])
    rng = np.random.default_rng(7)
    st = W.make_state(pts, *W.directions_and_speeds(rng, len(pts)))
    prox = ProxParams.make(64, 4.0)
    o = oracle_grid_outputs(orc, grid, prox, st)
    assert first_mismatch(gpu_lists(ctx, grid, prox, o), o["lists"]) is None


@pytest.mark.parametrize("uniform_max", [0, 1, 4, 64])
def test_dense_with_clamped_boundaries_all_scan_modes(ctx, orc, uniform_max, list_mode):
    """Dense flock overflowing a small grid on every side: boundary cells hold
    clamped out-of-grid agents (the cell-box skip must not apply to them); the
    scan mode (uniform / per lane) must not change a single bit."""
    from behavert_amd.types import OPT_UNIFORM_MAX_CELLS
    rng = np.random.default_rng(99)
    grid = GridParams.direct((0.0,) * 3, (4.0,) * 3, (8,) * 3)
    xyz = rng.uniform(-3, 35, (40000, 3)).astype(np.float32)
    xyz[:20] = [[np.nan, 3.0, 3.0]] * 20
    st = W.make_state(xyz, *W.directions_and_speeds(rng, len(xyz)))
    prox = ProxParams.make(32, 4.0)
    o = oracle_grid_outputs(orc, grid, prox, st)
    old = ctx.get_option(OPT_UNIFORM_MAX_CELLS)
    try:
        ctx.set_option(OPT_UNIFORM_MAX_CELLS, uniform_max)
        assert first_mismatch(gpu_lists(ctx, grid, prox, o), o["lists"]) is None
    finally:
        ctx.set_option(OPT_UNIFORM_MAX_CELLS, old)


@pytest.mark.parametrize("K,maxb", [(32, 0xFFFFFFFF), (252, 0xFFFFFFFF), (252, 100), (64, 40)])
def test_crowded_cells_long_streams(ctx, orc, K, maxb, list_mode):
    """About 280 agents per cell: a cell group's stream of 27 capped ranges runs
    to thousands of candidates, so with a large K the dense-wave scan goes past
    the chunks it keeps in registers; maxBodiesPerCell cuts ranges (and some
    agents' own entries) short."""
    rng = np.random.default_rng(K + maxb % 1000)
    grid = GridParams.direct((0.0,) * 3, (4.0,) * 3, (6,) * 3, max_bodies=maxb)
    xyz = rng.uniform(0, 24, (60000, 3)).astype(np.float32)
    st = W.make_state(xyz, *W.directions_and_speeds(rng, len(xyz)))
    prox = ProxParams.make(K, 4.0)
    o = oracle_grid_outputs(orc, grid, prox, st)
    assert first_mismatch(gpu_lists(ctx, grid, prox, o), o["lists"]) is None


def test_non_power_of_two_cells_offset_origin(ctx, orc, list_mode):
    """Cell size 2.5 and an offset origin: cell boxes are not exact in float, so
    the skip margins are exercised."""
    rng = np.random.default_rng(123)
    grid = GridParams.direct((-13.37, 7.25, 100.5), (2.5, 2.5, 2.5), (20, 20, 20))
    lo = np.array([-13.37, 7.25, 100.5])
    xyz = (lo + rng.uniform(0, 50, (60000, 3))).astype(np.float32)
    st = W.make_state(xyz, *W.directions_and_speeds(rng, len(xyz)))
    for r in (2.5, 2.4999, 1.7):
        prox = ProxParams.make(24, r)
        o = oracle_grid_outputs(orc, grid, prox, st)
        assert first_mismatch(gpu_lists(ctx, grid, prox, o), o["lists"]) is None, r
