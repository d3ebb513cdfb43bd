"""Where finite differences can check a gradient of the rollout at all (CPU,
FP64 oracle, decisions and lists fixed along an FP32 forward).

One step: the central difference of the loss in a behaviour weight converges
as the step shrinks (the objective is smooth at that scale), so it is a valid
reference for the adjoint.  Forty steps: the fixed-decision objective has
accumulated so many force/speed truncation kinks, amplified by the flock's
dynamics, that the central difference keeps changing by tens of percent as
the step shrinks; no finite-difference reference exists there, which is why
the GPU gradient tests compare at horizons of 1-20 steps."""
import numpy as np

from behavert_amd import workloads as W

from ._adjoint_common import record_forward, small_flock, velocity_state


def weight_fd(orc, steps, eps_list, b=2):
    w = small_flock(0, 4096, 64.0, K=32)
    st = w.state
    v64 = (st["forward"][:, :3] * st["forward"][:, 3:4]).astype(np.float32).astype(np.float64)
    s0 = velocity_state(st, v64)
    st32 = {k: s0[k].astype(np.float32) for k in s0}
    w0 = np.array([1.0, 0.8, 0.6], np.float32).astype(np.float64)
    _, recs = record_forward(orc, w, W.plain_steer(tuple(w0)), steps, state=st32)
    tgt = orc.rollout_f64_fixed(w.grid, w.prox, W.plain_steer((1.2, 0.6, 0.9)), w.dt, recs, s0)["pos"][:, :3]

    def loss(weights):
        out = orc.rollout_f64_fixed(w.grid, w.prox, W.plain_steer(tuple(weights)), w.dt, recs, s0)
        return ((out["pos"][:, :3] - tgt) ** 2).sum(axis=1).mean()

    out = []
    for eps in eps_list:
        e = np.zeros(3)
        e[b] = eps
        wp = (w0 + e).astype(np.float32).astype(np.float64)
        wm = (w0 - e).astype(np.float32).astype(np.float64)
        out.append((loss(wp) - loss(wm)) / (wp[b] - wm[b]))
    return np.array(out)


def test_fd_converges_for_one_step(orc):
    g = weight_fd(orc, 1, (1e-5, 3e-6))
    assert abs(g[0] - g[1]) <= 1e-4 * abs(g[1]), g


def test_fd_does_not_converge_at_forty_steps(orc):
    g = weight_fd(orc, 40, (1e-5, 3e-6))
    assert abs(g[0] - g[1]) >= 0.1 * abs(g[1]), g
