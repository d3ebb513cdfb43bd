"""Parameters and seeded inputs of the committed golden cases (small flocks whose
expected outputs tests/golden/make_golden.py wrote with the FP32 oracle into
golden_small.npz).  The inputs are stored in the fixture as well, so a change to
the generators shows up as a fixture mismatch rather than silently new data."""
from __future__ import annotations

import numpy as np

from behavert_amd import workloads as W
from behavert_amd.types import GridParams, ProxParams

DT = 0.05


def special_state(n: int, side: float, seed: int) -> dict:
    """A uniform flock with hand-placed agents: on cell faces (p = 4k), outside the
    grid on both sides, an exact duplicate, and a pair exactly r = 4 apart."""
    st = W.uniform_flock(n, side, seed)
    p = st["pos"]
    p[0, :3] = (8.0, 12.0, 16.0)          # on faces of three axes
    p[1, :3] = (8.0, 12.0, 16.0)          # duplicate of agent 0 (d = 0)
    p[2, :3] = (12.0, 12.0, 16.0)         # exactly r from agent 0 along x
    p[3, :3] = (-3.0, 5.0, 5.0)           # below the grid on x (clamps to the last cell)
    p[4, :3] = (side + 5.0, 5.0, 5.0)     # above the grid on x
    p[5, :3] = (0.0, 0.0, 0.0)            # grid corner
    p[6, :3] = (np.nextafter(np.float32(4.0), np.float32(0.0)), 4.0, 4.0)
    return st


def cases() -> dict:
    side = 40.0
    grid = W.cube_grid(side)
    capped = W.cube_grid(side, max_bodies=3)
    st = special_state(600, side, 11)
    g = W.rng(12)
    feats = g.random((600, 4)).astype(np.float32)
    full = W.plain_steer((1.0, 0.8, 0.6))
    full.seek_world_center = 1
    full.target_weight = 0.5
    full.set_targets(W.moving_targets(0.0, 20.0, 4) * 0.25)
    full.wander_weight = 0.2
    full.wander_seed = 99
    return {
        "plain": dict(grid=grid, prox=ProxParams.make(16, 4.0), steer=W.plain_steer(), state=st, steps=3),
        "knn": dict(grid=grid, prox=ProxParams.make(4, 4.0, use_knn=True), steer=W.plain_steer(), state=st,
                    steps=1),
        "max_bodies": dict(grid=capped, prox=ProxParams.make(16, 4.0), steer=W.plain_steer(), state=st, steps=1),
        "full": dict(grid=grid, prox=ProxParams.make(16, 4.0), steer=full, state=st, steps=2, features=feats,
                     feature_rate=0.1, wander=np.zeros((600, 2), np.float32)),
    }


FIELDS = ("hash", "cell_start", "cell_end", "lists", "pos", "forward", "smooth_accel")
