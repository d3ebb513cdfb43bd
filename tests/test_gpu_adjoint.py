"""The hand-written adjoint of the steering + integration step (brt_osw_adjoint_step)
and the checkpointed autograd rollout (behavert_amd.autograd), against FP64
central finite differences (eps 1e-4) of the FP64 oracle with the sort, the
neighbour lists and the discrete decisions fixed as the FP32 forward made them.
Tolerance: |g_adjoint - g_fd| <= 1e-3 * max(|g_fd|, 1e-2 * max_k |g_fd_k|)."""
import numpy as np
import pytest
import torch

from behavert_amd import workloads as W
from behavert_amd.autograd import Rollout, forward_from_velocity, rollout_final_pos
from behavert_amd.types import GridParams, ProxParams, SteerParams

from ._parity import dev, empty_f32, f32

pytestmark = pytest.mark.gpu

EPS = 1e-4


def check(gad, gfd, label, rel=1e-3):
    gad = np.asarray(gad, np.float64)
    gfd = np.asarray(gfd, np.float64)
    floor = 1e-2 * np.abs(gfd).max()
    err = np.abs(gad - gfd) / np.maximum(np.abs(gfd), floor)
    print(f"{label}: max rel err {err.max():.2e} over {len(gfd)} coordinates (|g| max {np.abs(gfd).max():.3e})")
    assert err.max() <= rel, (label, list(zip(gad, gfd)))


def record_forward(orc, w, steer, steps, state=None):
    hf = orc.HostFlock(w.grid, w.prox, state if state is not None else w.state)
    recs = []
    for t in range(steps):
        hf.step(steer, w.dt, t)
        recs.append((hf.hash.copy(), hf.lists.copy(), hf.sorted_pos.copy(), hf.sorted_forward.copy()))
    return hf, recs


def small_flock(seed, n, side, K=16):
    grid = W.cube_grid(side)
    return W.Workload("adj", n, side, grid, ProxParams.make(K, 4.0), W.plain_steer(), 0.05, 1,
                      W.uniform_flock(n, side, seed))


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
    # GPU adjoint
    gi = [empty_f32(w.n, 4) for _ in range(3)]
    wg = torch.zeros(3, dtype=torch.float32, device="cuda")
    import ctypes as C
    from behavert_amd import lib
    ctx.check(lib().brt_osw_adjoint_step(ctx.handle, C.byref(w.grid), C.byref(w.prox), C.byref(steer), w.dt, w.n,
                                         dev(spos32).data_ptr(), dev(sfwd32).data_ptr(), dev(ssa32).data_ptr(),
                                         dev(h.view(np.int32)).data_ptr(), dev(L.view(np.int32)).data_ptr(),
                                         dev(cpos.astype(np.float32)).data_ptr(),
                                         dev(cfwd.astype(np.float32)).data_ptr(),
                                         dev(csa.astype(np.float32)).data_ptr(),
                                         gi[0].data_ptr(), gi[1].data_ptr(), gi[2].data_ptr(), wg.data_ptr(),
                                         torch.cuda.current_stream().cuda_stream))
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
            arr[0] = float(np.float64(arr[0]) + sgn * EPS)
            ws.append(float(arr[0]))            # the float32 value actually used
            vals.append(loss(s64, sp))
        wfd.append((vals[0] - vals[1]) / (ws[0] - ws[1]))
    check(gw, wfd, "one-step weight gradient")


def rollout_case(ctx, orc, n, side, steps, every, seed=0):
    w = small_flock(seed, n, side, K=32)
    w.steps = steps
    target_steer = SteerParams.make(weights=(1.2, 0.6, 0.9))
    start = (1.0, 0.8, 0.6)
    ro = Rollout(ctx, w.grid, w.prox, W.plain_steer(start), n, w.dt, steps, every)
    pos0 = dev(w.state["pos"])
    sa0 = dev(w.state["smooth_accel"])
    f0 = w.state["forward"]
    v0 = torch.tensor(f0[:, :3] * f0[:, 3:4], dtype=torch.float32, device="cuda", requires_grad=True)
    weights = torch.tensor(start, dtype=torch.float32, device="cuda", requires_grad=True)
    # target final positions from the (1.2, 0.6, 0.9) rollout
    tro = Rollout(ctx, w.grid, w.prox, target_steer, n, w.dt, steps, every)
    with torch.no_grad():
        target = rollout_final_pos(tro, torch.tensor((1.2, 0.6, 0.9)), pos0, dev(f0), sa0).clone()
    fwd0 = forward_from_velocity(v0)
    final = rollout_final_pos(ro, weights, pos0, fwd0, sa0)
    loss = ((final - target) ** 2).sum(dim=1).mean()
    loss.backward()
    w.extra["fwd0_used"] = fwd0.detach().cpu().numpy()
    return w, ro, weights.grad.detach().cpu().numpy(), v0.grad.detach().cpu().numpy(), target.cpu().numpy(), \
        v0.detach().cpu().numpy()


def test_rollout_gradient_against_fd(ctx, orc):
    n, side, steps, every = 4096, 64.0, 40, 10
    w, ro, gw, gv, target, v0 = rollout_case(ctx, orc, n, side, steps, every)
    steer = W.plain_steer((1.0, 0.8, 0.6))
    st = dict(w.state)
    st["forward"] = w.extra["fwd0_used"]      # the FP32 forward state the GPU rollout started from
    _, recs = record_forward(orc, w, steer, steps, state=st)
    base = {"pos": st["pos"].astype(np.float64), "smooth_accel": st["smooth_accel"].astype(np.float64)}
    v64 = v0.astype(np.float64)
    tgt = target.astype(np.float64)

    def loss(v, weights):
        s = dict(base)
        sp = np.linalg.norm(v, axis=1, keepdims=True)
        fw = np.zeros((n, 4))
        fw[:, :3] = v / sp
        fw[:, 3] = sp[:, 0]
        s["forward"] = fw
        stp = W.plain_steer(tuple(weights))
        out = orc.rollout_f64_fixed(w.grid, w.prox, stp, w.dt, recs, s)
        return ((out["pos"][:, :3] - tgt) ** 2).sum(axis=1).mean()

    w0 = np.array([1.0, 0.8, 0.6], np.float32).astype(np.float64)
    wfd = []
    for b in range(3):
        e = np.zeros(3)
        e[b] = EPS
        wp = (w0 + e).astype(np.float32).astype(np.float64)   # weights travel as float32
        wm = (w0 - e).astype(np.float32).astype(np.float64)
        wfd.append((loss(v64, wp) - loss(v64, wm)) / (wp[b] - wm[b]))
    check(gw, wfd, "rollout dL/dweights")
    rng = np.random.default_rng(5)
    gad, gfd = [], []
    for a, c in zip(rng.integers(0, n, 16), rng.integers(0, 3, 16)):
        vp, vm = v64.copy(), v64.copy()
        vp[a, c] += EPS
        vm[a, c] -= EPS
        gfd.append((loss(vp, w0) - loss(vm, w0)) / (2 * EPS))
        gad.append(gv[a, c])
    check(gad, gfd, "rollout dL/dv0")


def test_checkpoint_interval_does_not_change_gradients(ctx, orc):
    res = [rollout_case(ctx, orc, 2048, 50.0, 12, every) for every in (1, 5, 12)]
    for _, _, gw, gv, _, _ in res[1:]:
        np.testing.assert_allclose(gw, res[0][2], rtol=1e-5, atol=1e-9)
        np.testing.assert_allclose(gv, res[0][3], rtol=1e-4, atol=1e-9 * np.abs(res[0][3]).max())


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
                                    z.data_ptr(), z.data_ptr(), torch.cuda.current_stream().cuda_stream)
    assert rc != 0
    with pytest.raises(BrtError):
        ctx.check(rc)
