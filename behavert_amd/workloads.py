"""Seeded synthetic flocks and the workload configurations of the step path.

Generators use numpy's PCG64 bit stream with uniform doubles only (Box-Muller for
normals), so every input is a deterministic function of (seed, sizes).

State layout (all float32, agent-ID order):
  pos          (N, 4)  x, y, z, 1
  forward      (N, 4)  unit heading xyz, speed w
  smooth_accel (N, 4)  zeros
Directions are uniform on the sphere (normalised 3-D normals), speed U[0.5, 1.0].
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .types import GridParams, ProxParams, SteerParams, UNLIMITED

SEP_ALIGN_COH = (1.0, 0.8, 0.6)
MAX_FORCE = 0.5
MAX_SPEED = 1.0
DT = 0.05
CELL = 4.0
RADIUS = 4.0


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def normals(g: np.random.Generator, shape) -> np.ndarray:
    u1 = 1.0 - g.random(shape)
    u2 = g.random(shape)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def directions_and_speeds(g: np.random.Generator, n: int, smin=0.5, smax=1.0):
    v = normals(g, (n, 3))
    nrm = np.linalg.norm(v, axis=1, keepdims=True)
    nrm[nrm == 0] = 1.0
    d = (v / nrm).astype(np.float32)
    s = (smin + (smax - smin) * g.random(n)).astype(np.float32)
    return d, s


def make_state(xyz: np.ndarray, d: np.ndarray, s: np.ndarray) -> dict:
    n = len(xyz)
    pos = np.ones((n, 4), np.float32)
    pos[:, :3] = xyz
    fwd = np.zeros((n, 4), np.float32)
    fwd[:, :3] = d
    fwd[:, 3] = s
    return {"pos": pos, "forward": fwd, "smooth_accel": np.zeros((n, 4), np.float32)}


def uniform_flock(n: int, side: float, seed: int) -> dict:
    """Positions uniform in [0, side)^3."""
    g = rng(seed)
    xyz = (g.random((n, 3)) * side).astype(np.float32)
    xyz = np.minimum(xyz, np.nextafter(np.float32(side), np.float32(0)))
    d, s = directions_and_speeds(g, n)
    return make_state(xyz, d, s)


def blob_flock(n: int, side: float, blobs: int, sigma: float, seed: int) -> dict:
    """Mixture of isotropic Gaussian blobs (centres uniform in the cube), each
    point clamped to [0, nextafter(side, 0)]; blob k gets agents i with
    perm[i] % blobs == k, so the ID order is spatially random."""
    g = rng(seed)
    centres = g.random((blobs, 3)) * side
    comp = g.permutation(n) % blobs
    xyz = centres[comp] + sigma * normals(g, (n, 3))
    hi = float(np.nextafter(np.float32(side), np.float32(0)))
    xyz = np.clip(xyz, 0.0, hi).astype(np.float32)
    d, s = directions_and_speeds(g, n)
    return make_state(xyz, d, s)


def cube_grid(side: float, cell: float = CELL, max_bodies: int = UNLIMITED) -> GridParams:
    """World = [0, side)^3 exactly as EnvGrid3DPlugIn::setupGrid maps it:
    worldCenter = side/2, worldRadius = side/2 - 2*cell -> origin 0,
    gridSize = side/cell (PlugIns/EnvGrid3D/include/EnvGrid3DPlugIn.h:478-508)."""
    ncell = int(math.ceil(side / cell))
    half = 0.5 * ncell * cell
    return GridParams.direct((0.0, 0.0, 0.0), (cell,) * 3, (ncell,) * 3, center=(half,) * 3,
                             radius=(half - 2 * cell,) * 3, max_bodies=max_bodies)


def moving_targets(t: float, centre: float, count: int = 8) -> np.ndarray:
    """Analytic target motion for the multi-target seek: target k on a tilted
    circular orbit of radius 60+15k around the world centre, angular speed
    0.1+0.02k rad per time unit, phase 2*pi*k/count, tilt pi*k/count."""
    out = np.zeros((count, 3), np.float64)
    for k in range(count):
        r = 60.0 + 15.0 * k
        a = (0.1 + 0.02 * k) * t + 2.0 * math.pi * k / count
        tilt = math.pi * k / count
        out[k] = (centre + r * math.cos(a), centre + r * math.sin(a) * math.cos(tilt),
                  centre + r * math.sin(a) * math.sin(tilt))
    return out.astype(np.float32)


@dataclass
class Workload:
    name: str
    n: int
    side: float
    grid: GridParams
    prox: ProxParams
    steer: SteerParams
    dt: float
    steps: int
    state: dict
    feature_rate: float = 0.0
    features: np.ndarray | None = None
    wander: np.ndarray | None = None
    moving_targets: bool = False
    extra: dict = field(default_factory=dict)

    def targets_at(self, step: int) -> np.ndarray | None:
        if not self.moving_targets:
            return None
        return moving_targets(step * self.dt, 0.5 * self.side)

    def describe(self) -> dict:
        return {"workload": self.name, "n": self.n, "world": [0.0, self.side],
                "grid": list(self.grid.grid_size), "cell": self.grid.cell_size[0],
                "radius": self.prox.search_radius, "max_neighbors": self.prox.max_neighbors,
                "steps": self.steps, "dt": self.dt}


def plain_steer(weights=SEP_ALIGN_COH) -> SteerParams:
    return SteerParams.make(weights=weights, radius=RADIUS, cos_max=-1.0, mass=1.0,
                            max_speed=MAX_SPEED, max_force=MAX_FORCE)


def plain_4096(seed: int = 0, n: int = 4096) -> Workload:
    """4,096 boids uniform in [0,100)^3, grid 25^3 of 4, r 4, K 16, 100 steps."""
    side = 100.0
    return Workload("plain_4096", n, side, cube_grid(side), ProxParams.make(16, RADIUS),
                    plain_steer(), DT, 100, uniform_flock(n, side, seed))


def blobs_1m(seed: int = 0, n: int = 1 << 20) -> Workload:
    """1,048,576 boids in [0,400)^3 from 64 Gaussian blobs (sigma 10), grid 100^3
    of 4, r 4, K 32, 200 steps."""
    side = 400.0
    return Workload("blobs_1m", n, side, cube_grid(side), ProxParams.make(32, RADIUS),
                    plain_steer(), DT, 200, blob_flock(n, side, 64, 10.0, seed))


def full_stack_4m(seed: int = 0, n: int = 1 << 22) -> Workload:
    """4,194,304 agents uniform in [0,640)^3, K 64, separation/alignment/cohesion +
    seek toward 8 moving targets + wander + 4-D feature averaging (rate 0.1),
    100 steps."""
    side = 640.0
    steer = plain_steer()
    steer.target_weight = 0.5
    steer.wander_weight = 0.2
    steer.wander_seed = 12345
    w = Workload("full_stack_4m", n, side, cube_grid(side), ProxParams.make(64, RADIUS),
                 steer, DT, 100, uniform_flock(n, side, seed), feature_rate=0.1,
                 moving_targets=True)
    g = rng(seed + 1)
    w.features = g.random((n, 4)).astype(np.float32)
    w.wander = np.zeros((n, 2), np.float32)
    w.steer.set_targets(w.targets_at(0))
    return w


def differentiable_65k(seed: int = 0, n: int = 65536) -> Workload:
    """65,536 boids uniform in [0,160)^3, r 4, K 32, dt 0.05, 200 steps,
    checkpoints every 20; start weights (1.0, 0.8, 0.6), target (1.2, 0.6, 0.9)."""
    side = 160.0
    w = Workload("differentiable_65k", n, side, cube_grid(side), ProxParams.make(32, RADIUS),
                 plain_steer(), DT, 200, uniform_flock(n, side, seed))
    w.extra = {"checkpoint_every": 20, "target_weights": (1.2, 0.6, 0.9),
               "start_weights": SEP_ALIGN_COH}
    return w


def batched_flock_sizes(seed: int = 0, flocks: int = 256, lo: int = 1024, hi: int = 65536) -> np.ndarray:
    """Flock sizes round(exp(U(ln lo, ln hi)))."""
    g = rng(seed)
    return np.round(np.exp(g.uniform(math.log(lo), math.log(hi), flocks))).astype(np.int64)


@dataclass
class BatchWorkload:
    """Independent flocks advanced together; flock f is sizes[f] agents uniform in
    a cube of side (n/density)^(1/3) with its own grid and weights."""
    name: str
    sizes: np.ndarray
    sides: np.ndarray
    grids: list
    weights: np.ndarray      # (F, 3) float32
    prox: ProxParams
    steer: SteerParams
    dt: float
    steps: int
    states: list             # per-flock state dicts

    @property
    def n(self) -> int:
        return int(self.sizes.sum())

    def concat_state(self) -> dict:
        return {k: np.concatenate([st[k] for st in self.states]) for k in ("pos", "forward", "smooth_accel")}

    def describe(self) -> dict:
        return {"workload": self.name, "flocks": len(self.sizes), "n": self.n,
                "flock_sizes": [int(self.sizes.min()), int(self.sizes.max())],
                "density": 0.5, "cell": CELL, "radius": self.prox.search_radius,
                "max_neighbors": self.prox.max_neighbors, "steps": self.steps, "dt": self.dt}


def batched_flocks(seed: int = 0, flocks: int = 256, lo: int = 1024, hi: int = 65536,
                   density: float = 0.5, steps: int = 200) -> BatchWorkload:
    """256 flocks with sizes log-uniform in [1024, 65536], each in its own cube at
    0.5 agents per unit^3, cell 4, r 4, K 32, weights U[0.5, 1.5] per flock and
    behaviour, 200 steps.  Flock f's state uses seed 1000*(seed+1)+f."""
    sizes = batched_flock_sizes(seed, flocks, lo, hi)
    sides = (sizes / density) ** (1.0 / 3.0)
    grids = [cube_grid(float(sd)) for sd in sides]
    g = rng(seed + 7)
    weights = g.uniform(0.5, 1.5, (flocks, 3)).astype(np.float32)
    states = [uniform_flock(int(n), float(sd), 1000 * (seed + 1) + f) for f, (n, sd) in enumerate(zip(sizes, sides))]
    return BatchWorkload("batched_256", sizes, sides, grids, weights, ProxParams.make(32, RADIUS), plain_steer(),
                         DT, steps, states)


WORKLOADS = {
    "plain_4096": plain_4096,
    "blobs_1m": blobs_1m,
    "full_stack_4m": full_stack_4m,
    "differentiable_65k": differentiable_65k,
}
