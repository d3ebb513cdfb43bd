"""The workload configurations at their own sizes, through the C ABI against the
CPU oracle (the 4,096-boid flock runs in the steering, golden and smoke tests):

* blobs_1m (1,048,576 boids, 64 Gaussian blobs, K 32): 10 steps bit-exact
  against the FP32 oracle (state, sort, lists), and positions and headings
  within 1e-3 of the FP64 trajectory that uses the same sorts, lists and
  discrete decisions (smoothed acceleration: 99.9 % of agents within 1e-3,
  all within 5e-3);
* full_stack_4m (4,194,304 agents, K 64, targets + wander + features): 2 steps
  bit-exact against the oracle;
* batched_256 (256 flocks, 4,643,977 agents, autotuned launch): 2 batched steps,
  every flock bit-equal to brt_step run on that flock alone, and the smallest
  and the largest flock bit-equal to the oracle;
* differentiable_65k (65,536 boids, 200 steps, checkpoints every 20): the
  gradients bit-identical for checkpoint intervals 20 and 200, linear in the
  cotangent, and, at the horizons where central differences converge (10 and
  20 steps), within the stated tolerance of FP64 central differences (eps 1e-4
  for v0, 1e-5 for the weights, 16 coordinates).
"""
import numpy as np
import pytest
import torch

from behavert_amd import workloads as W
from behavert_amd.autograd import Rollout
from behavert_amd.engine import Batch, Flock

from ._parity import dev, first_mismatch, floored_rel
from .test_gpu_adjoint import check, rollout_case, rollout_fd_errors
from .test_gpu_batch import steer_with
from .test_gpu_steer import FLOORS, run_gpu_steps, run_oracle_steps

pytestmark = pytest.mark.gpu


def test_blobs_1m_ten_steps_bit_exact_and_within_1e3_of_fp64(ctx, orc, record_property):
    w = W.blobs_1m()
    print(f"{w.name}: {w.n} boids, 10 steps", flush=True)
    fl = run_gpu_steps(ctx, w, 10)
    got, go = fl.state(), fl.grid_outputs()
    hf = orc.HostFlock(w.grid, w.prox, w.state)
    hist = []
    for i in range(10):
        hf.step(w.steer, w.dt, i)
        hist.append((hf.hash.copy(), hf.lists.copy(), hf.sorted_pos.copy(), hf.sorted_forward.copy()))
    ref = hf.state()
    for k in ("pos", "forward", "smooth_accel"):
        assert first_mismatch(got[k].view(np.uint32), ref[k].view(np.uint32)) is None, k
    assert first_mismatch(go["hash"], hf.hash) is None
    assert first_mismatch(go["lists"], hf.lists) is None
    # the FP64 trajectory with the FP32 run's sorts, lists and discrete
    # decisions (radius/angle predicates, world box: pred)
    s64 = {k: w.state[k].astype(np.float64) for k in ("pos", "forward", "smooth_accel")}
    kappa = np.zeros(w.n)   # by agent ID: the largest condition number of its normalisations over the steps
    for i, (h, L, sp32, sf32) in enumerate(hist):
        out = orc.step_f64_fixed(w.grid, w.prox, w.steer, w.dt, i, s64, h, L, pred=(sp32, sf32))
        s64 = {k: out[k] for k in ("pos", "forward", "smooth_accel")}
        ids = h[:, 1].astype(np.int64)
        kappa[ids] = np.maximum(kappa[ids], out["cond"])
    for k in ("pos", "forward", "smooth_accel"):
        err = floored_rel(got[k][:, :3], s64[k][:, :3], FLOORS[k])
        record_property(f"blobs_1m_traj10_{k}_max_err", float(err.max()))
        q = np.quantile(err, [0.5, 0.99, 0.999])
        print(f"blobs_1m 10-step {k}: floored rel err vs FP64 median {q[0]:.2e} p99 {q[1]:.2e} p99.9 {q[2]:.2e} "
              f"max {err.max():.2e}; {(err > 1e-3).sum()} agents above 1e-3", flush=True)
        for c in (10.0, 100.0, 1000.0):
            sel = kappa <= c
            print(f"   agents with condition number <= {c:g} at every step: {sel.mean():.4f}, "
                  f"max err {err[sel].max():.2e}", flush=True)
        bad = np.argsort(err)[-5:]
        print("   worst agents: err", err[bad], "cond", kappa[bad], flush=True)
        if k == "smooth_accel":
            # the smoothed acceleration (a lerp of the truncated force) of a
            # few agents in 10^6 drifts past 1e-3 of its 0.4 scale
            assert np.quantile(err, 0.999) < 1e-3 and err.max() < 5e-3, k
        else:
            assert err.max() < 1e-3, k


def test_full_stack_4m_two_steps_bit_exact(ctx, orc):
    w = W.full_stack_4m()
    print(f"{w.name}: {w.n} agents, K {w.prox.max_neighbors}, 2 steps", flush=True)
    fl = run_gpu_steps(ctx, w, 2)
    got, go = fl.state(), fl.grid_outputs()
    hf, _ = run_oracle_steps(orc, w, 2)
    ref = hf.state()
    for k in ("pos", "forward", "smooth_accel", "features", "wander"):
        assert first_mismatch(got[k].view(np.uint32), ref[k].view(np.uint32)) is None, k
    assert first_mismatch(go["hash"], hf.hash) is None
    assert first_mismatch(go["lists"], hf.lists) is None


def test_batched_256_every_flock_equals_its_own_step(ctx, orc):
    w = W.batched_flocks()
    print(f"{w.name}: {len(w.sizes)} flocks, {w.n} agents", flush=True)
    b = Batch(ctx, w.grids, w.sizes, w.weights, w.prox, w.steer, w.concat_state())
    cfg = b.autotune(w.dt, 1)
    print("tuned", cfg.as_dict(), flush=True)
    for t in range(2):
        b.step(w.dt, t)
    torch.cuda.synchronize()
    bh = b.hash.cpu().numpy().view(np.uint32)
    bl = b.lists.cpu().numpy().view(np.uint32)
    bst = {k: getattr(b, k).cpu().numpy() for k in ("pos", "forward", "smooth_accel")}
    cols = np.arange(bl.shape[1])[None, :]
    for f in range(len(w.sizes)):
        fl = Flock(ctx, w.grids[f], w.prox, steer_with(w.weights[f], w.steer), w.states[f])
        for t in range(2):
            fl.step(w.dt, t)
        st, go = fl.state(), fl.grid_outputs()
        sl = b.flock_slice(f)
        off, cb = np.uint32(b.offsets[f]), np.uint32(b.cell_bases[f])
        h = bh[sl].copy()
        h[:, 0] -= cb
        h[:, 1] -= off
        assert first_mismatch(h, go["hash"]) is None, f
        L = bl[sl].copy()
        entries = (cols >= 1) & (cols <= L[:, :1])
        L[entries] -= off
        assert first_mismatch(L, go["lists"]) is None, f
        for k in ("pos", "forward", "smooth_accel"):
            assert first_mismatch(bst[k][sl].view(np.uint32), st[k].view(np.uint32)) is None, (f, k)
        del fl
    for f in (int(np.argmin(w.sizes)), int(np.argmax(w.sizes))):
        hf = orc.HostFlock(w.grids[f], w.prox, w.states[f])
        for t in range(2):
            hf.step(steer_with(w.weights[f], w.steer), w.dt, t)
        sl = b.flock_slice(f)
        for k, a in (("pos", hf.pos), ("forward", hf.forward), ("smooth_accel", hf.smooth_accel)):
            assert first_mismatch(bst[k][sl].view(np.uint32), a.view(np.uint32)) is None, (f, k)
    b.close()


def test_differentiable_65k_checkpoint_interval_and_linearity(ctx, orc):
    """200 steps of the 65,536-boid rollout: the gradients with checkpoints every
    20 steps and with one 200-step segment are the same bits; and the adjoint
    is linear in the cotangent (to FP32 rounding of the sums)."""
    w = W.differentiable_65k()
    res = []
    for every in (w.extra["checkpoint_every"], w.steps):
        print(f"rollout: {w.n} boids, {w.steps} steps, checkpoints every {every}", flush=True)
        _, _, gw, gv, _, _ = rollout_case(ctx, orc, 0, 0.0, w.steps, every, w=W.differentiable_65k())
        res.append((gw, gv))
    assert np.array_equal(res[0][0].view(np.uint32), res[1][0].view(np.uint32))
    assert np.array_equal(res[0][1].view(np.uint32), res[1][1].view(np.uint32))
    assert np.isfinite(res[0][1]).all() and np.abs(res[0][1]).max() > 0
    # linearity: g(c1 + 2 c2) = g(c1) + 2 g(c2)
    ro = Rollout(ctx, w.grid, w.prox, W.plain_steer(w.extra["start_weights"]), w.n, w.dt, w.steps,
                 w.extra["checkpoint_every"])
    pos0, f0, sa0 = dev(w.state["pos"]), dev(w.state["forward"]), dev(w.state["smooth_accel"])
    ro.forward(list(w.extra["start_weights"]), pos0, f0, sa0)
    rng = np.random.default_rng(3)
    c1, c2 = (np.zeros((w.n, 4), np.float32) for _ in range(2))
    c1[:, :3] = rng.standard_normal((w.n, 3)) / w.n
    c2[:, :3] = rng.standard_normal((w.n, 3)) / w.n
    z = torch.zeros((w.n, 4), dtype=torch.float32, device="cuda")
    g1 = ro.backward(dev(c1), z, z)
    g2 = ro.backward(dev(c2), z, z)
    g12 = ro.backward(dev(c1 + 2 * c2), z, z)
    for a, b, c in zip(g1, g2, g12):
        a, b, c = (x.cpu().numpy().astype(np.float64) for x in (a, b, c))
        lin = a + 2 * b
        err = np.abs(c - lin).max() / np.abs(lin).max()
        print(f"linearity: max |g(c1+2c2) - g(c1) - 2g(c2)| / max = {err:.2e}", flush=True)
        assert err < 1e-4


@pytest.mark.parametrize("steps", [10, 20])
def test_differentiable_65k_against_fp64_fd(ctx, orc, steps, record_property):
    """The c5 objective (mean squared distance of the final positions to the
    (1.2, 0.6, 0.9) rollout's) on the 65,536-boid workload at the horizons
    where its central differences converge (beyond ~20 steps the force and
    speed truncation kinks of the fixed-decision objective stop them
    converging, tests/test_oracle_adjoint_fd.py).  dL/dweights against the
    differences of the objective itself; dL/dweights and dL/dv0 against those
    of the objective linearised at the FP32 forward (the cotangent
    2 (final - target) / n fixed): the adjoint's gradient is J^T of that
    cotangent, and the residual is small enough that the FP32 rounding of the
    final positions otherwise dominates the per-agent dL/dv0."""
    w = W.differentiable_65k()
    every = min(steps, w.extra["checkpoint_every"])
    print(f"c5 objective: {w.n} boids, {steps} steps", flush=True)
    (gw, wfd), _ = rollout_fd_errors(ctx, orc, w, steps, every, coords=0)
    record_property(f"c5_{steps}_weights", [float(x) for x in gw])
    check(gw, wfd, f"c5 {steps} steps dL/dweights (objective)", rel=1e-3)
    (gw, wfd), (gad, gfd) = rollout_fd_errors(ctx, orc, W.differentiable_65k(), steps, every, coords=16,
                                              linearised=True)
    check(gw, wfd, f"c5 {steps} steps dL/dweights (linearised)", rel=1e-3)
    check(gad, gfd, f"c5 {steps} steps dL/dv0 (linearised)", rel=1e-3 if steps <= 10 else 5e-3)
