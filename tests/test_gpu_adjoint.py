"""The hand-written adjoint of the steering + integration step (brt_osw_adjoint_step)
and the checkpointed autograd rollout (behavert_amd.autograd), against FP64
central finite differences of the FP64 oracle with the sort, the neighbour
lists and the discrete decisions fixed as the FP32 forward made them.
Tolerance: |g_adjoint - g_fd| <= 1e-3 * max(|g_fd|, 1e-2 * max_k |g_fd_k|).

Steps: 1e-4 for state and velocity coordinates; 1e-5 for the behaviour
weights.  A weight perturbation moves every agent's force, and at +-1e-4 a few
agents cross the force or speed truncation kink, which biases the central
difference by ~1e-3 (the difference converges once the step is <= 3e-5)."""
import numpy as np
import pytest
import torch

from behavert_amd import workloads as W
from behavert_amd.autograd import Rollout, forward_from_velocity, rollout_final_pos
from behavert_amd.types import GridParams, ProxParams, SteerParams

from ._adjoint_common import record_forward, small_flock, velocity_state
from ._parity import dev, empty_f32, f32

pytestmark = pytest.mark.gpu

EPS = 1e-4
EPS_W = 1e-5


def check(gad, gfd, label, rel=1e-3):
    gad = np.asarray(gad, np.float64)
    gfd = np.asarray(gfd, np.float64)
    floor = 1e-2 * np.abs(gfd).max()
    err = np.abs(gad - gfd) / np.maximum(np.abs(gfd), floor)
    print(f"{label}: max rel err {err.max():.2e} over {len(gfd)} coordinates (|g| max {np.abs(gfd).max():.3e})")
    assert err.max() <= rel, (label, list(zip(gad, gfd)))


def adjoint_step(ctx, grid, prox, steer, dt, spos, sfwd, ssa, h, lists, gpos_out, gfwd_out, gsa_out, wg):
    """brt_osw_adjoint_step on host arrays; returns the three input adjoints
    (device tensors) and adds to wg (device)."""
    import ctypes as C
    from behavert_amd import lib
    import oracle
    held = [dev(np.ascontiguousarray(a, np.float32)) for a in (spos, sfwd, ssa, gpos_out, gfwd_out, gsa_out)]
    cs, ce = oracle.cell_ranges(np.ascontiguousarray(h, np.uint32), grid.num_cells)
    held += [dev(np.ascontiguousarray(a, np.uint32).view(np.int32)) for a in (h, lists, cs, ce)]
    n = len(h)
    gi = [empty_f32(n, 4) for _ in range(3)]
    P = [t.data_ptr() for t in held]
    ctx.check(lib().brt_osw_adjoint_step(ctx.handle, C.byref(grid), C.byref(prox), C.byref(steer), dt, n,
                                         P[0], P[1], P[2], P[6], P[7], P[8], P[9], P[3], P[4], P[5],
                                         gi[0].data_ptr(), gi[1].data_ptr(), gi[2].data_ptr(), wg.data_ptr(),
                                         torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    del held
    return gi


def test_one_step_vjp_against_fd(ctx, orc):
    w = small_flock(3, 3000, 24.0)
    st = w.state
    rng = np.random.default_rng(4)
    st["forward"][:300, 3] = rng.uniform(0.05, 0.19, 300).astype(np.float32)   # cone-limit branch
    st["smooth_accel"][:, :3] = rng.uniform(-0.2, 0.2, (w.n, 3)).astype(np.float32)
    st["pos"][:100, :3] = rng.uniform(-3, 0, (100, 3)).astype(np.float32)        # outside the world box
    steer = SteerParams.make(weights=(1.0, 0.8, 0.6), seek_world_center=True,
                             targets=W.moving_targets(0.3, 12.0), target_weight=0.4)
    _, recs = record_forward(orc, w, steer, 1)
    h, L, spos32, sfwd32 = recs[0]
    ids = h[:, 1].astype(np.int64)
    ssa32 = st["smooth_accel"][ids]
    cpos = rng.standard_normal((w.n, 4))
    cfwd = rng.standard_normal((w.n, 4))
    csa = rng.standard_normal((w.n, 4))
    cpos[:, 3] = 0.0
    csa[:, 3] = 0.0
    # GPU adjoint (every input tensor held until the kernel has run)
    wg = torch.zeros(3, dtype=torch.float32, device="cuda")
    gi = adjoint_step(ctx, w.grid, w.prox, steer, w.dt, spos32, sfwd32, ssa32, h, L,
                      cpos.astype(np.float32), cfwd.astype(np.float32), csa.astype(np.float32), wg)
    gpos, gfwd, gsa = (f32(t).astype(np.float64) for t in gi)
    gw = f32(wg).astype(np.float64)
    # FP64 objective with decisions fixed
    s64 = {k: st[k].astype(np.float64) for k in ("pos", "forward", "smooth_accel")}

    def loss(state, sp=steer):
        out = orc.step_f64_fixed(w.grid, w.prox, sp, w.dt, 0, state, h, L, pred=(spos32, sfwd32))
        cp = cpos.astype(np.float32).astype(np.float64)
        cf = cfwd.astype(np.float32).astype(np.float64)
        ca = csa.astype(np.float32).astype(np.float64)
        return (out["pos"] * cp).sum() + (out["forward"] * cf).sum() + (out["smooth_accel"] * ca).sum()

    coords = [(k, int(a), int(c)) for k in ("pos", "forward", "smooth_accel")
              for a, c in zip(rng.integers(0, w.n, 12), rng.integers(0, 4 if k == "forward" else 3, 12))]
    coords += [("forward", a, 3) for a in range(5)]   # slow agents: speed derivative through the cone
    gad, gfd = [], []
    for k, a, c in coords:
        sp_, sm_ = {q: v.copy() for q, v in s64.items()}, {q: v.copy() for q, v in s64.items()}
        sp_[k][a, c] += EPS
        sm_[k][a, c] -= EPS
        gfd.append((loss(sp_) - loss(sm_)) / (2 * EPS))
        gad.append({"pos": gpos, "forward": gfwd, "smooth_accel": gsa}[k][a, c])
    check(gad, gfd, "one-step state VJP")
    wfd = []
    for b in range(3):
        vals, ws = [], []
        for sgn in (1, -1):
            sp = SteerParams.make(weights=(1.0, 0.8, 0.6), seek_world_center=True,
                                  targets=W.moving_targets(0.3, 12.0), target_weight=0.4)
            arr = (sp.separation, sp.alignment, sp.cohesion)[b]
            arr[0] = float(np.float64(arr[0]) + sgn * EPS_W)
            ws.append(float(arr[0]))            # the float32 value actually used
            vals.append(loss(s64, sp))
        wfd.append((vals[0] - vals[1]) / (ws[0] - ws[1]))
    check(gw, wfd, "one-step weight gradient")


@pytest.mark.parametrize("T,tol", [(1, 1e-3), (5, 1e-3), (20, 5e-3)])
def test_adjoint_chain_on_fp64_trajectory(ctx, orc, T, tol):
    """T adjoint steps chained along an FP64 trajectory (each step linearised at
    that trajectory's state rounded to FP32) against FP64 central differences of
    the same fixed-decision rollout: checks the adjoint and its composition
    without the FP32-vs-FP64 trajectory drift of a long forward.  Loss: a random
    linear functional of the final positions.  At 20 steps the central
    difference itself only converges to ~1e-3 as the step shrinks (the kinks
    of the truncations accumulate), hence the looser bound there."""
    n, side = 4096, 48.0
    w = small_flock(7, n, side, K=32)
    st = w.state
    v64 = (st["forward"][:, :3] * st["forward"][:, 3:4]).astype(np.float32).astype(np.float64)
    w0 = np.array([1.0, 0.8, 0.6], np.float32).astype(np.float64)
    steer0 = W.plain_steer(tuple(w0))
    s0 = velocity_state(st, v64)
    s = {k: v.copy() for k, v in s0.items()}
    recs, lin = [], []
    for t in range(T):
        # the sort and the lists of the state being linearised at (the adjoint
        # takes lists that are those of the sorted positions it is given)
        pos32 = s["pos"].astype(np.float32)
        h = orc.sort_radix(orc.calc_hash(w.grid, pos32), w.grid.num_cells)
        cs, ce = orc.cell_ranges(h, w.grid.num_cells)
        ids = h[:, 1].astype(np.int64)
        spos, sfwd, ssa = (s[k][ids].astype(np.float32) for k in ("pos", "forward", "smooth_accel"))
        L = orc.neighbors(w.grid, w.prox, spos, h, cs, ce)
        recs.append((h, L, spos, sfwd))
        lin.append((spos, sfwd, ssa))
        out = orc.step_f64_fixed(w.grid, w.prox, steer0, w.dt, t, s, h, L, pred=(spos, sfwd))
        s = {k: out[k] for k in ("pos", "forward", "smooth_accel")}
    cot = np.random.default_rng(8).standard_normal((n, 3)).astype(np.float32).astype(np.float64)

    def loss(v, weights):
        out = orc.rollout_f64_fixed(w.grid, w.prox, W.plain_steer(tuple(weights)), w.dt, recs, velocity_state(st, v))
        return (out["pos"][:, :3] * cot).sum() / n

    # adjoint chain
    g_out = [np.zeros((n, 4), np.float32) for _ in range(3)]
    g_out[0][:, :3] = (cot / n).astype(np.float32)
    wg = torch.zeros(3, dtype=torch.float32, device="cuda")
    for t in range(T - 1, -1, -1):
        h, L = recs[t][0], recs[t][1]
        gi = adjoint_step(ctx, w.grid, w.prox, steer0, w.dt, *lin[t], h, L, *g_out, wg)
        g_out = [f32(x) for x in gi]
    gf = g_out[1].astype(np.float64)
    u = s0["forward"][:, :3]
    spd = s0["forward"][:, 3:4]
    gv = (gf[:, :3] - u * (u * gf[:, :3]).sum(1, keepdims=True)) / spd + u * gf[:, 3:4]
    gw = f32(wg).astype(np.float64)
    wfd = []
    for b in range(3):
        e = np.zeros(3)
        e[b] = EPS_W
        wp = (w0 + e).astype(np.float32).astype(np.float64)
        wm = (w0 - e).astype(np.float32).astype(np.float64)
        wfd.append((loss(v64, wp) - loss(v64, wm)) / (wp[b] - wm[b]))
    check(gw, wfd, f"chain({T}) dL/dweights", rel=tol)
    rng = np.random.default_rng(9)
    gad, gfd = [], []
    for a, c in zip(rng.integers(0, n, 16), rng.integers(0, 3, 16)):
        vp, vm = v64.copy(), v64.copy()
        vp[a, c] += EPS
        vm[a, c] -= EPS
        gfd.append((loss(vp, w0) - loss(vm, w0)) / (2 * EPS))
        gad.append(gv[a, c])
    check(gad, gfd, f"chain({T}) dL/dv0", rel=tol)


def rollout_case(ctx, orc, n, side, steps, every, seed=0, w=None, offset=0.0):
    """The c5-style objective on the GPU: the loss is the mean squared distance
    of the final positions to those of a (1.2, 0.6, 0.9) rollout (shifted by
    `offset` times fixed standard normals), the start uses (1.0, 0.8, 0.6);
    gradients by the autograd Function over the adjoint."""
    if w is None:
        w = small_flock(seed, n, side, K=32)
    n = w.n
    w.steps = steps
    target_steer = SteerParams.make(weights=(1.2, 0.6, 0.9))
    start = (1.0, 0.8, 0.6)
    ro = Rollout(ctx, w.grid, w.prox, W.plain_steer(start), n, w.dt, steps, every)
    pos0 = dev(w.state["pos"])
    sa0 = dev(w.state["smooth_accel"])
    f0 = w.state["forward"]
    v0 = torch.tensor(f0[:, :3] * f0[:, 3:4], dtype=torch.float32, device="cuda", requires_grad=True)
    weights = torch.tensor(start, dtype=torch.float32, device="cuda", requires_grad=True)
    tro = Rollout(ctx, w.grid, w.prox, target_steer, n, w.dt, steps, every)
    with torch.no_grad():
        target = rollout_final_pos(tro, torch.tensor((1.2, 0.6, 0.9)), pos0, dev(f0), sa0).clone()
        if offset:
            shift = np.random.default_rng(11).standard_normal((n, 3)).astype(np.float32) * np.float32(offset)
            target += dev(shift)
    fwd0 = forward_from_velocity(v0)
    final = rollout_final_pos(ro, weights, pos0, fwd0, sa0)
    loss = ((final - target) ** 2).sum(dim=1).mean()
    loss.backward()
    w.extra["fwd0_used"] = fwd0.detach().cpu().numpy()
    w.extra["final"] = final.detach().cpu().numpy()
    return w, ro, weights.grad.detach().cpu().numpy(), v0.grad.detach().cpu().numpy(), target.cpu().numpy(), \
        v0.detach().cpu().numpy()


def rollout_fd_errors(ctx, orc, w, steps, every, coords=16, offset=0.0, linearised=False):
    """GPU autograd gradients and FP64 central differences of the fixed-decision
    rollout along the FP32 forward's sorts and lists, for dL/dweights and for
    dL/dv0 on `coords` coordinates.  linearised: the differences are those of
    the loss linearised at the FP32 forward, sum(c * final) with the cotangent
    c = 2 (final32 - target) / n fixed -- the gradient the adjoint computes is
    J^T c, so this leaves out the FP32 rounding of the residual (final -
    target), which for a small residual dominates a per-agent gradient."""
    w, ro, gw, gv, target, v0 = rollout_case(ctx, orc, 0, 0.0, steps, every, w=w, offset=offset)
    n = w.n
    st = dict(w.state)
    st["forward"] = w.extra["fwd0_used"]      # the FP32 forward state the GPU rollout started from
    _, recs = record_forward(orc, w, W.plain_steer((1.0, 0.8, 0.6)), steps, state=st)
    v64 = v0.astype(np.float64)
    tgt = target.astype(np.float64)
    cot = 2.0 * (w.extra["final"].astype(np.float64) - tgt) / n

    def loss(v, weights):
        out = orc.rollout_f64_fixed(w.grid, w.prox, W.plain_steer(tuple(weights)), w.dt, recs,
                                    velocity_state(st, v))
        if linearised:
            return (out["pos"][:, :3] * cot).sum()
        return ((out["pos"][:, :3] - tgt) ** 2).sum(axis=1).mean()

    w0 = np.array([1.0, 0.8, 0.6], np.float32).astype(np.float64)
    wfd = []
    for b in range(3):
        e = np.zeros(3)
        e[b] = EPS_W
        wp = (w0 + e).astype(np.float32).astype(np.float64)   # weights travel as float32
        wm = (w0 - e).astype(np.float32).astype(np.float64)
        wfd.append((loss(v64, wp) - loss(v64, wm)) / (wp[b] - wm[b]))
    rng = np.random.default_rng(5)
    gad, gfd = [], []
    for a, c in zip(rng.integers(0, n, coords), rng.integers(0, 3, coords)):
        vp, vm = v64.copy(), v64.copy()
        vp[a, c] += EPS
        vm[a, c] -= EPS
        gfd.append((loss(vp, w0) - loss(vm, w0)) / (2 * EPS))
        gad.append(gv[a, c])
    return (gw, np.array(wfd)), (np.array(gad), np.array(gfd))


def test_rollout_gradient_against_fd(ctx, orc):
    """10 steps (checkpoints every 5), 4,096 boids: the whole autograd path
    (GPU forward, checkpoint replay, adjoint) within 1e-3 of FP64 FD.  The
    target is offset by O(1) so the residual is far above the FP32 rounding of
    the final positions, which the FP32 loss and the FP64 loss otherwise differ
    by (the c5-style objective itself: next test)."""
    (gw, wfd), (gad, gfd) = rollout_fd_errors(ctx, orc, small_flock(0, 4096, 64.0, K=32), 10, 5, offset=0.5)
    check(gw, wfd, "rollout dL/dweights")
    check(gad, gfd, "rollout dL/dv0")


def test_rollout_gradient_agreement_by_horizon(ctx, orc, record_property):
    """The c5-style objective (target = the (1.2, 0.6, 0.9) rollout itself) over
    10, 20 and 40 steps.  Its residual is small, so the FP32 rounding of the
    final positions weighs in the per-agent dL/dv0, and with the horizon the
    fixed-decision objective stops being smooth at the FD step (its central
    differences stop converging, tests/test_oracle_adjoint_fd.py), so there is
    no FD reference at long horizons.  Recorded; weights held to 1e-3 at 10
    steps and 1e-2 at 20."""
    out = {}
    for steps in (10, 20, 40):
        (gw, wfd), (gad, gfd) = rollout_fd_errors(ctx, orc, small_flock(0, 4096, 64.0, K=32), steps, 10, coords=8)
        ew = np.abs(gw - wfd) / np.maximum(np.abs(wfd), 1e-2 * np.abs(wfd).max())
        ev = np.abs(gad - gfd) / np.maximum(np.abs(gfd), 1e-2 * np.abs(gfd).max())
        out[steps] = (float(ew.max()), float(ev.max()))
        print(f"horizon {steps}: dL/dweights max rel err {ew.max():.2e}, dL/dv0 max rel err {ev.max():.2e}")
        record_property(f"horizon_{steps}", out[steps])
    assert out[10][0] <= 1e-3 and out[20][0] <= 1e-2


def test_checkpoint_interval_does_not_change_gradients(ctx, orc):
    """The replayed segments equal the forward and the adjoint has no atomics,
    so the gradients are bit-identical for every checkpoint interval."""
    res = [rollout_case(ctx, orc, 2048, 50.0, 12, every) for every in (1, 5, 12)]
    for _, _, gw, gv, _, _ in res[1:]:
        assert np.array_equal(gw.view(np.uint32), res[0][2].view(np.uint32))
        assert np.array_equal(gv.view(np.uint32), res[0][3].view(np.uint32))


def test_adjoint_step_is_deterministic_and_matches_a_scatter_of_the_pair_terms(ctx, orc):
    """Two calls on the same inputs give the same bits; and with one agent's
    cotangent set, the neighbour terms land on exactly the agents of its
    list (the gather finds every list that holds an agent, including full
    lists cut by K and lists cut by maxBodiesPerCell)."""
    rng = np.random.default_rng(12)
    for maxb, K, side in ((0xFFFFFFFF, 32, 30.0), (6, 8, 14.0)):
        grid = W.cube_grid(side, max_bodies=maxb)
        n = 3000
        # agents below and above the grid on every axis: the hash keys those
        # below at the far face (unsigned clamp), and their lists still hold
        # agents of the near face
        st = W.make_state(rng.uniform(-3.0, side + 3.0, (n, 3)).astype(np.float32),
                          *W.directions_and_speeds(rng, n))
        prox = ProxParams.make(K, 4.0)
        w = W.Workload("det", n, side, grid, prox, W.plain_steer(), 0.05, 1, st)
        _, recs = record_forward(orc, w, w.steer, 1)
        h, L, spos, sfwd = recs[0]
        ids = h[:, 1].astype(np.int64)
        ssa = st["smooth_accel"][ids]
        if K == 8:
            assert (L[:, 0] == K).mean() > 0.3, "the case must have full lists"
        cot = [rng.standard_normal((n, 4)).astype(np.float32) for _ in range(3)]
        outs = []
        for _ in range(2):
            wg = torch.zeros(3, dtype=torch.float32, device="cuda")
            gi = adjoint_step(ctx, grid, prox, w.steer, w.dt, spos, sfwd, ssa, h, L, *cot, wg)
            outs.append([f32(x) for x in gi] + [f32(wg)])
        for a, b in zip(*outs):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        # one agent's cotangent: the neighbour terms reach exactly its list
        # (a full list, a list of an agent below the grid, the longest list)
        below = np.nonzero((spos[:, :3] < 0).any(axis=1) & (L[:, 0] > 0))[0]
        picks = {int(np.argmax(L[:, 0])), int(below[0])}
        if K == 8:
            picks.add(int(np.nonzero(L[:, 0] == K)[0][-1]))
        for i in sorted(picks):
            one = [np.zeros((n, 4), np.float32) for _ in range(3)]
            one[0][h[i, 1], :3] = 1.0
            wg = torch.zeros(3, dtype=torch.float32, device="cuda")
            gi = adjoint_step(ctx, grid, prox, w.steer, w.dt, spos, sfwd, ssa, h, L, *one, wg)
            gp = f32(gi[0])
            touched = set(np.nonzero(np.abs(gp[:, :3]).sum(1) > 0)[0].tolist())
            expect = {int(h[j, 1]) for j in L[i, 1:1 + L[i, 0]]} | {int(h[i, 1])}
            assert touched == expect, (i, sorted(touched - expect), sorted(expect - touched))


def test_adjoint_rejects_unsupported_options(ctx):
    from behavert_amd import BrtError
    w = small_flock(1, 256, 16.0)
    sp = W.plain_steer()
    sp.wander_weight = 0.5
    with pytest.raises(ValueError):
        Rollout(ctx, w.grid, w.prox, sp, w.n, w.dt, 2, 1)
    import ctypes as C
    from behavert_amd import lib
    z = empty_f32(w.n, 4)
    rc = lib().brt_osw_adjoint_step(ctx.handle, C.byref(w.grid), C.byref(w.prox), C.byref(sp), w.dt, w.n,
                                    z.data_ptr(), z.data_ptr(), z.data_ptr(), z.data_ptr(), z.data_ptr(),
                                    z.data_ptr(), z.data_ptr(), z.data_ptr(), z.data_ptr(), z.data_ptr(),
                                    z.data_ptr(), z.data_ptr(), z.data_ptr(), z.data_ptr(),
                                    torch.cuda.current_stream().cuda_stream)
    assert rc != 0
    with pytest.raises(BrtError):
        ctx.check(rc)


def test_two_forwards_then_backward_of_the_first(ctx, orc):
    """A Rollout reused by a second forward (other weights) before the first
    graph's backward: the first backward replays its own checkpoints and
    weights, so its gradients equal those of a lone forward + backward."""
    w = small_flock(2, 2048, 50.0, K=32)
    start, other = (1.0, 0.8, 0.6), (0.3, 1.4, 0.2)
    pos0, f0, sa0 = dev(w.state["pos"]), dev(w.state["forward"]), dev(w.state["smooth_accel"])

    def grads(ro, seq):
        outs = []
        for ws in seq:
            wt = torch.tensor(ws, dtype=torch.float32, device="cuda", requires_grad=True)
            p0 = pos0.clone().requires_grad_(True)
            final = rollout_final_pos(ro, wt, p0, f0, sa0)
            outs.append((wt, p0, (final ** 2).sum(dim=1).mean()))
        wt, p0, loss = outs[0]
        loss.backward()
        return wt.grad.cpu().numpy(), p0.grad.cpu().numpy()

    ro = Rollout(ctx, w.grid, w.prox, W.plain_steer(start), w.n, w.dt, 8, 4)
    g_alone = grads(ro, [start])
    g_shared = grads(ro, [start, other])
    np.testing.assert_allclose(g_shared[0], g_alone[0], rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(g_shared[1], g_alone[1], rtol=1e-4, atol=1e-7 * np.abs(g_alone[1]).max())


def test_overlapped_replay_gives_the_same_gradients(ctx):
    """The backward with each segment's replay on a second stream and context,
    overlapping the adjoint of the later segment (the default), returns the
    same bits as the serial backward, with checkpoints every 7 of 40 steps (a
    short last segment) and twice in a row."""
    w = small_flock(5, 3000, 40.0, K=32)
    pos0, fw0, sa0 = dev(w.state["pos"]), dev(w.state["forward"]), dev(w.state["smooth_accel"])
    rng = np.random.default_rng(2)
    gp = np.zeros((w.n, 4), np.float32)
    gp[:, :3] = rng.standard_normal((w.n, 3)) / w.n
    z = torch.zeros((w.n, 4), dtype=torch.float32, device="cuda")
    outs = []
    for overlap in (False, True, True):
        ro = Rollout(ctx, w.grid, w.prox, W.plain_steer((1.0, 0.8, 0.6)), w.n, w.dt, 40, 7, overlap=overlap)
        ro.forward([1.0, 0.8, 0.6], pos0, fw0, sa0)
        g = ro.backward(dev(gp), z, z)
        torch.cuda.synchronize()
        outs.append([x.cpu().numpy() for x in g])
    for run in outs[1:]:
        for a, b in zip(outs[0], run):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.abs(outs[0][1]).max() > 0
