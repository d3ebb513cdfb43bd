"""Helpers shared by the adjoint tests (CPU and GPU)."""
import numpy as np

from behavert_amd import workloads as W
from behavert_amd.types import ProxParams


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


def velocity_state(st, v64):
    s = {"pos": st["pos"].astype(np.float64), "smooth_accel": st["smooth_accel"].astype(np.float64)}
    sp = np.linalg.norm(v64, axis=1, keepdims=True)
    fw = np.zeros((len(v64), 4))
    fw[:, :3] = v64 / sp
    fw[:, 3] = sp[:, 0]
    s["forward"] = fw
    return s
