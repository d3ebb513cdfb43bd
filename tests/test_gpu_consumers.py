"""GPU parity of the small list consumers through the C ABI, bit for bit
against the FP32 oracle: steerAndSlowToAvoidNeighbors, the Drawable3D
neighbourhood colouring and the Shapes3D target attraction."""
import numpy as np
import pytest
import torch

from behavert_amd.engine import BrtError
from behavert_amd.types import ProxParams, SteerParams

from ._parity import dev, empty_f32, f32, first_mismatch, oracle_grid_outputs
from .test_gpu_clustering import mixed_workload

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("K", [16, 32])
def test_slow_avoid_bit_exact(ctx, orc, K):
    """K = 32 lists often hold more than 16 agents (the amplified branch)."""
    w = mixed_workload(n=30000, K=K, seed=5)
    o = oracle_grid_outputs(orc, w.grid, w.prox, w.state)
    assert (o["lists"][:, 0] > 16).any() == (K > 16)
    sp = SteerParams.make(weights=(1.3, 0.8, 0.6), radius=3.0, cos_max=-0.5, min_radius=0.5)
    fin = np.random.default_rng(2).standard_normal((w.n, 4)).astype(np.float32)
    fwd_out, f_out = empty_f32(w.n, 4), empty_f32(w.n, 4)
    ctx.steer_and_slow_to_avoid_neighbors(w.prox, sp, dev(o["sorted_pos"]), dev(o["sorted_forward"]),
                                          dev(o["lists"].view(np.int32)), fwd_out, dev(fin), f_out)
    rfwd, rf = orc.slow_avoid(w.prox, sp, o["sorted_pos"], o["sorted_forward"], o["lists"], fin)
    assert first_mismatch(f32(fwd_out).view(np.uint32), rfwd.view(np.uint32)) is None
    assert first_mismatch(f32(f_out).view(np.uint32), rf.view(np.uint32)) is None
    changed = rfwd[:, 3] != o["sorted_forward"][:, 3]
    print(f"K={K}: speed changed for {changed.mean():.3f} of the agents")
    assert changed.any()
    # in place (forward and force updated where they lie)
    fw, ff = dev(o["sorted_forward"]), dev(fin)
    ctx.steer_and_slow_to_avoid_neighbors(w.prox, sp, dev(o["sorted_pos"]), fw, dev(o["lists"].view(np.int32)),
                                          fw, ff, ff)
    assert first_mismatch(f32(fw).view(np.uint32), rfwd.view(np.uint32)) is None
    assert first_mismatch(f32(ff).view(np.uint32), rf.view(np.uint32)) is None


@pytest.mark.parametrize("vertexes", [1, 3])
def test_neighborhood_color_bit_exact(ctx, orc, vertexes):
    w = mixed_workload(n=30000, K=16, seed=6)
    o = oracle_grid_outputs(orc, w.grid, w.prox, w.state)
    base = (0.1, -0.2, 0.3)
    out = torch.full((w.n, vertexes, 4), float("nan"), dtype=torch.float32, device="cuda")
    ctx.neighborhood_color(w.prox, base, vertexes, dev(o["hash"].view(np.int32)), dev(o["lists"].view(np.int32)),
                           out)
    ref = orc.neighborhood_color(w.prox, base, vertexes, o["hash"], o["lists"])
    assert first_mismatch(f32(out).view(np.uint32), ref.view(np.uint32)) is None


@pytest.mark.parametrize("shift", [0, 1, 12345, 29999])
def test_move_towards_target_bit_exact(ctx, orc, shift):
    w = mixed_workload(n=30000, K=16, seed=7)
    o = oracle_grid_outputs(orc, w.grid, w.prox, w.state)
    rng = np.random.default_rng(shift)
    targets = np.c_[rng.uniform(0, 40, (w.n, 3)), rng.uniform(0, 2, w.n)].astype(np.float32)
    fin = rng.standard_normal((w.n, 4)).astype(np.float32)
    out = empty_f32(w.n, 4)
    ctx.move_towards_target(shift, (1.0, 2.0, 3.0), dev(o["hash"].view(np.int32)), dev(o["sorted_pos"]),
                            dev(targets), dev(fin), out)
    ref = orc.move_towards_target(shift, (1.0, 2.0, 3.0), o["hash"], o["sorted_pos"], targets, fin)
    assert first_mismatch(f32(out).view(np.uint32), ref.view(np.uint32)) is None


def test_consumer_argument_checks(ctx):
    prox = ProxParams.make(8, 4.0)
    h = torch.zeros((4, 2), dtype=torch.int32, device="cuda")
    L = torch.zeros((4, 12), dtype=torch.int32, device="cuda")
    p = torch.zeros((4, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(BrtError):
        ctx.neighborhood_color(prox, (0, 0, 0), 2, h, L, p)
    with pytest.raises(BrtError):
        ctx.move_towards_target(4, (0, 0, 0), h, p, p, p, p)
    with pytest.raises(BrtError):
        ctx.move_towards_target(-1, (0, 0, 0), h, p, p, p, p)
