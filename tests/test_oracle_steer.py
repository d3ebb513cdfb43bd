"""Pins the oracle's steering + integration (OpenSteerWrapper_resources.cu:67-379,
OpenSteerWrapper_kernel.cu:62-333, Core/common_resources.cu:140-271) with
hand-derived cases and the FP64 evaluation, and the Philox stream with the
Random123 known-answer vectors."""
import math

import numpy as np
import pytest

from behavert_amd import workloads as W
from behavert_amd.types import GridParams, ProxParams, SteerParams


def agents(xyz, fwd, speed, sa=None):
    n = len(xyz)
    pos = np.ones((n, 4), np.float32)
    pos[:, :3] = xyz
    f = np.zeros((n, 4), np.float32)
    f[:, :3] = fwd
    f[:, 3] = speed
    s = np.zeros((n, 4), np.float32)
    if sa is not None:
        s[:, :3] = sa
    return pos, f, s


def one_step(orc, grid, prox, steer, pos, fwd, sa, dt=0.05, f64=False):
    h = orc.sort_radix(orc.calc_hash(grid, pos), grid.num_cells)
    s, e = orc.cell_ranges(h, grid.num_cells)
    spos, sf, ssa = orc.reorder(h, pos), orc.reorder(h, fwd), orc.reorder(h, sa)
    L = orc.neighbors(grid, prox, spos, h, s, e)
    fn = orc.steer_f64 if f64 else orc.steer_f32
    out = fn(grid, prox, steer, dt, 0, spos, sf, ssa, h, L)
    out["hash"], out["lists"] = h, L
    return out


GRID = GridParams.direct((0.0,) * 3, (4.0,) * 3, (5,) * 3)
PROX = ProxParams.make(16, 4.0)


def test_two_agents_on_x_axis_closed_forms(orc):
    """A at x=8 heading +y, B at x=10 heading +z.  For A:
    separation = normalize((2,0,0) / -4) = (-1,0,0),
    cohesion   = normalize(B - A) = (+1,0,0),
    alignment  = normalize(fB - fA) = (0,-1,1)/sqrt(2);
    force = w_s*sep + w_a*align + w_c*coh."""
    pos, f, sa = agents([[8.0, 8.0, 8.0], [10.0, 8.0, 8.0]], [[0, 1, 0], [0, 0, 1]], [0.5, 0.5])
    for w in [(1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (1.0, 0.8, 0.6)]:
        steer = SteerParams.make(weights=w, max_force=100.0)
        out = one_step(orc, GRID, PROX, steer, pos, f, sa)
        iA = int(np.nonzero(out["hash"][:, 1] == 0)[0][0])
        F = out["force"][iA, :3].astype(np.float64)
        expect = (w[0] * np.array([-1.0, 0, 0]) + w[1] * np.array([0, -1.0, 1.0]) / math.sqrt(2)
                  + w[2] * np.array([1.0, 0, 0]))
        np.testing.assert_allclose(F, expect, rtol=0, atol=2e-7)


def test_zero_force_single_agent_is_euler_with_speed_clamp(orc):
    pos, f, sa = agents([[8.0, 8.0, 8.0]], [[0.6, 0.8, 0.0]], [0.9], sa=[[0.1, -0.2, 0.3]])
    steer = SteerParams.make()
    out = one_step(orc, GRID, PROX, steer, pos, f, sa, dt=0.05, f64=True)
    f64 = lambda x: np.float32(x).astype(np.float64)  # noqa: E731  (inputs are float32)
    dt, rate = f64(0.05), f64(0.4)
    sa1 = f64([0.1, -0.2, 0.3]) * (1 - rate)
    v = f64([0.6, 0.8, 0.0]) * f64(0.9) + sa1 * dt
    sp = np.linalg.norm(v)
    np.testing.assert_allclose(out["smooth_accel"][0, :3], sa1, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(out["pos"][0, :3], f64([8.0, 8.0, 8.0]) + v * dt, rtol=1e-14)
    np.testing.assert_allclose(out["forward"][0, :3], v / sp, rtol=1e-14)
    assert out["forward"][0, 3] == pytest.approx(sp, rel=1e-14)
    # speed above maxSpeed is truncated
    pos, f, sa = agents([[8.0, 8.0, 8.0]], [[1.0, 0.0, 0.0]], [1.5])
    out = one_step(orc, GRID, PROX, steer, pos, f, sa, f64=True)
    assert out["forward"][0, 3] == pytest.approx(1.0, rel=1e-12)
    assert out["pos"][0, 0] == pytest.approx(8.0 + 1.0 * f64(0.05), rel=1e-14)


def test_force_truncation_and_mass(orc):
    # two agents 1 apart: |separation| = 1 (unit), weight 3 -> |F| = 3 > maxForce 0.5
    pos, f, sa = agents([[8.0, 8.0, 8.0], [9.0, 8.0, 8.0]], [[0, 1, 0], [0, 1, 0]], [0.5, 0.5])
    steer = SteerParams.make(weights=(3.0, 0.0, 0.0), max_force=0.5, mass=2.0)
    out = one_step(orc, GRID, PROX, steer, pos, f, sa, f64=True)
    iA = int(np.nonzero(out["hash"][:, 1] == 0)[0][0])
    np.testing.assert_allclose(out["force"][iA, :3], [-3.0, 0, 0], atol=1e-12)
    # clipped (-0.5,0,0)/mass 2 -> accel (-0.25,0,0); sa = 0 + 0.4*accel
    np.testing.assert_allclose(out["smooth_accel"][0, :3], [-0.1, 0, 0], atol=1e-12)


def test_slow_agent_cone_limit_branch(orc):
    """speed <= 0.2*maxSpeed: the force is limited to a cone around forward with
    cos = 1 - 2*(speed/(0.2*maxSpeed))^20 (adjustRawSteeringForce)."""
    pos, f, sa = agents([[8.0, 8.0, 8.0], [9.0, 8.0, 8.0]], [[0, 1, 0], [0, 1, 0]], [0.19, 0.19])
    steer = SteerParams.make(weights=(0.1, 0.0, 0.0), max_force=10.0)
    out = one_step(orc, GRID, PROX, steer, pos, f, sa, f64=True)
    iA = int(np.nonzero(out["hash"][:, 1] == 0)[0][0])
    F = out["force"][iA, :3]
    r = np.float64(np.float32(0.19)) / np.float64(np.float32(0.2))
    cos = 1.0 - 2.0 * r ** 20
    # F = (-0.1,0,0) is at 90 deg to forward (0,1,0): limited onto the cone
    w = np.float64(np.float32(0.1))
    expect = (np.array([0, 1.0, 0]) * cos + np.array([-1.0, 0, 0]) * math.sqrt(1 - cos * cos)) * w
    accel = expect  # mass 1, |expect| = 0.1 < maxForce
    rate = np.float64(np.float32(0.4))  # clamp(9*dt, 0.15f, 0.4f) evaluated on float constants
    np.testing.assert_allclose(out["smooth_accel"][0, :3], rate * accel, rtol=1e-12, atol=1e-15)
    assert np.allclose(F, [-0.1, 0, 0])


def test_seek_world_center_and_targets(orc):
    grid = W.cube_grid(100.0)  # world centre 50, radius 42: box (8, 92)
    pos, f, sa = agents([[5.0, 50.0, 50.0], [50.0, 50.0, 50.0]], [[1, 0, 0], [1, 0, 0]], [0.5, 0.5])
    steer = SteerParams.make(weights=(0.0, 0.0, 0.0), seek_world_center=True, max_force=100.0)
    out = one_step(orc, grid, PROX, steer, pos, f, sa, f64=True)
    ids = out["hash"][:, 1]
    i0, i1 = int(np.nonzero(ids == 0)[0][0]), int(np.nonzero(ids == 1)[0][0])
    # outside: truncate(centre - pos, maxSpeed) - fwd*speed = (1,0,0) - (0.5,0,0)
    np.testing.assert_allclose(out["force"][i0, :3], [0.5, 0, 0], atol=1e-12)
    np.testing.assert_allclose(out["force"][i1, :3], [0, 0, 0], atol=0)
    steer = SteerParams.make(weights=(0.0, 0.0, 0.0), targets=[[5.0, 60.0, 50.0], [50.0, 50.0, 40.0]],
                             target_weight=2.0, max_force=100.0)
    out = one_step(orc, grid, PROX, steer, pos, f, sa, f64=True)
    # id 0 seeks target 0: (0,1,0) - (0.5,0,0); id 1 seeks target 1: (0,0,-1) - (0.5,0,0)
    np.testing.assert_allclose(out["force"][i0, :3], 2.0 * np.array([-0.5, 1.0, 0]), atol=1e-12)
    np.testing.assert_allclose(out["force"][i1, :3], 2.0 * np.array([-0.5, 0, -1.0]), atol=1e-12)


def floored_rel(a32, a64, floor):
    a32 = np.asarray(a32, np.float64)
    d = np.linalg.norm(a32 - a64, axis=1)
    return d / np.maximum(np.linalg.norm(a64, axis=1), floor)


def test_fp32_mirror_within_tolerance_of_fp64(orc):
    """One step of the 4,096-boid workload: force and state within 1e-5 of the
    FP64 evaluation (norm-wise per agent, floored at each quantity's scale)."""
    w = W.plain_4096()
    st = w.state
    o32 = one_step(orc, w.grid, w.prox, w.steer, st["pos"], st["forward"], st["smooth_accel"])
    h, L = o32["hash"], o32["lists"]
    ids = h[:, 1]
    o64 = orc.steer_f64(w.grid, w.prox, w.steer, w.dt, 0, st["pos"][ids].astype(np.float64),
                        st["forward"][ids].astype(np.float64), st["smooth_accel"][ids].astype(np.float64), h, L)
    assert floored_rel(o32["force"][:, :3], o64["force"][:, :3], 1.0).max() < 1e-5
    assert floored_rel(o32["pos"][:, :3], o64["pos"][:, :3], 1.0).max() < 1e-6
    assert floored_rel(o32["forward"][:, :3], o64["forward"][:, :3], 1.0).max() < 1e-6
    # smooth_accel = rate * force / mass: its scale is rate * (force scale 1.0) / mass = 0.4
    assert floored_rel(o32["smooth_accel"][:, :3], o64["smooth_accel"][:, :3], 0.4).max() < 1e-5


# Random123 kat_vectors, philox4x32 with 10 rounds: (ctr, key) -> out
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(orc, ctr, key, expect):
    assert list(orc.philox(ctr, key)) == expect


def test_feature_update_fixed_point_and_two_agent_form(orc):
    prox = ProxParams.make(8, 4.0)
    grid = GRID
    pos, f, sa = agents([[8.0, 8.0, 8.0], [9.0, 8.0, 8.0], [19.0, 19.0, 19.0]], [[0, 1, 0]] * 3, [0.5] * 3)
    h = orc.sort_radix(orc.calc_hash(grid, pos), grid.num_cells)
    s, e = orc.cell_ranges(h, grid.num_cells)
    L = orc.neighbors(grid, prox, orc.reorder(h, pos), h, s, e)
    same = np.tile(np.float32([0.25, 0.5, 0.75, 1.0]), (3, 1))
    assert np.array_equal(orc.feature_update_f32(prox, 0.1, h, L, same), same)
    fin = np.float32([[0.0, 0.0, 0.0, 0.0], [1.0, 2.0, 3.0, 4.0], [5.0, 5.0, 5.0, 5.0]])
    out = orc.feature_update_f64(prox, 0.1, h, L, fin.astype(np.float64))
    rate = np.float64(np.float32(0.1))
    f = fin.astype(np.float64)
    np.testing.assert_allclose(out[0], rate * f[1], rtol=1e-15)
    np.testing.assert_allclose(out[1], f[1] + rate * (f[0] - f[1]), rtol=1e-15)
    np.testing.assert_allclose(out[2], fin[2])   # no neighbours: unchanged


def _alone(orc, grid, prox, steer, xyz, fwd, speed, sa=None, step=0, wander=None, f64=True):
    """One agent with no neighbour: sorted order = ID order."""
    pos, f, s = agents([xyz], [fwd], [speed], sa=None if sa is None else [sa])
    h = orc.sort_radix(orc.calc_hash(grid, pos), grid.num_cells)
    st, en = orc.cell_ranges(h, grid.num_cells)
    L = orc.neighbors(grid, prox, orc.reorder(h, pos), h, st, en)
    assert L[0, 0] == 0
    fn = orc.steer_f64 if f64 else orc.steer_f32
    dt = np.float32(0.05)
    wd = None if wander is None else np.array([wander], np.float64 if f64 else np.float32)
    return fn(grid, prox, steer, dt, step, pos, f, s, h, L, wander=wd)


def _walk(word, w0, dt=np.float32(0.05)):
    """OpenSteer scalarRandomWalk(w0, 12*dt, -1, +1) (Utilities.h:158-167) with
    frandom01() = the top 24 bits of a Philox word / 2^24."""
    u = np.float64((int(word) >> 8) / 16777216.0)
    nxt = np.float64(w0) + (u * 2.0 - 1.0) * (12.0 * np.float64(dt))
    return min(max(nxt, -1.0), 1.0)


# heading -> (side, up) of OpenSteer's right-handed local space with the world
# up as the previous up (LocalSpace.h:277-305: side = fwd x up, up = side x fwd),
# derived by hand:
#   (1,0,0):        side (0,0,1),            up (0,1,0)
#   (0.6,0.8,0):    side (0,0,1),            up (-0.8,0.6,0)
#   (2,3,6)/7:      side (-3,0,1)/sqrt(10),  up (-3,20,-9)/(7 sqrt(10))
WANDER_BASES = [
    ((1.0, 0.0, 0.0), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0)),
    ((0.6, 0.8, 0.0), (0.0, 0.0, 1.0), (-0.8, 0.6, 0.0)),
    ((2 / 7, 3 / 7, 6 / 7), tuple(np.array([-3.0, 0.0, 1.0]) / math.sqrt(10)),
     tuple(np.array([-3.0, 20.0, -9.0]) / (7 * math.sqrt(10)))),
]


@pytest.mark.parametrize("fwd,side,up", WANDER_BASES)
def test_wander_known_answer(orc, fwd, side, up):
    """steerForWander (SteerLibrary.h:319-331): force = side*ws + up*wu, times
    the wander weight, after one random-walk step of (ws, wu) driven by the
    Philox words of (ctr = (step,0,0,0), key = (seed, ID)); the walk state is
    written back.  A mirrored side axis or a swapped cross product fails."""
    seed, step, weight = 7, 3, 2.0
    steer = SteerParams.make(weights=(0.0, 0.0, 0.0), wander_weight=weight, wander_seed=seed, max_force=100.0)
    r = orc.philox([step, 0, 0, 0], [seed, 0])
    for w0 in [(0.25, -0.5), (1.0, -1.0), (-0.95, 0.97)]:
        out = _alone(orc, GRID, PROX, steer, (8.0, 8.0, 8.0), fwd, 0.5, step=step, wander=w0)
        ws, wu = _walk(r[0], w0[0]), _walk(r[1], w0[1])
        np.testing.assert_allclose(out["wander"][0], [ws, wu], rtol=0, atol=1e-12)
        expect = weight * (ws * np.array(side) + wu * np.array(up))
        # the heading is stored in FP32: the basis moves by ~1e-7
        np.testing.assert_allclose(out["force"][0, :3], expect, rtol=0, atol=1e-6)
        o32 = _alone(orc, GRID, PROX, steer, (8.0, 8.0, 8.0), fwd, 0.5, step=step, wander=w0, f64=False)
        np.testing.assert_allclose(o32["force"][0, :3], expect, rtol=0, atol=2e-6)


def test_wander_is_lateral_and_keyed_by_id_and_step(orc):
    """The wander force is perpendicular to the heading for any state, and the
    stream changes with the step and the agent ID."""
    steer = SteerParams.make(weights=(0.0, 0.0, 0.0), wander_weight=1.0, wander_seed=11, max_force=100.0)
    rng = np.random.default_rng(5)
    for _ in range(20):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        out = _alone(orc, GRID, PROX, steer, (8.0, 8.0, 8.0), d, 0.5, step=int(rng.integers(0, 1000)),
                     wander=rng.uniform(-1, 1, 2))
        F = out["force"][0, :3]
        assert abs(np.dot(F, np.float32(d).astype(np.float64))) < 1e-6 * max(1.0, np.linalg.norm(F))
    assert list(orc.philox([1, 0, 0, 0], [11, 0])) != list(orc.philox([2, 0, 0, 0], [11, 0]))
    assert list(orc.philox([1, 0, 0, 0], [11, 0])) != list(orc.philox([1, 0, 0, 0], [11, 1]))


def test_2d_projection_known_answer(orc):
    """use2DProjection (OpenSteerWrapper_kernel.cu:296-313,
    OpenSteerWrapper_resources.cu:365-368): the force's y is zeroed before the
    integrator, the speed is taken from the full new velocity, the position
    moves by the full velocity and then y is set to 0.5, and the heading is the
    new velocity with y zeroed.  Agent at (10,10,10), heading (0.6,0.8,0) at
    speed 0.5, seeking a target at (10,20,10):
      force  = truncate((0,10,0), 1) - (0.3,0.4,0) = (-0.3, 0.6, 0)
      2-D    : (-0.3, 0, 0);  sa = 0.4 * (-0.3,0,0) = (-0.12, 0, 0)
      v      = (0.3,0.4,0) + 0.05 * sa = (0.294, 0.4, 0);  speed = |v| = sqrt(0.246436)
      pos    = (10 + 0.05*0.294, 0.5, 10);  forward = (1, 0, 0)."""
    steer = SteerParams.make(weights=(0.0, 0.0, 0.0), targets=[[10.0, 20.0, 10.0]], target_weight=1.0,
                             max_force=10.0, use_2d=True)
    out = _alone(orc, GRID, PROX, steer, (10.0, 10.0, 10.0), (0.6, 0.8, 0.0), 0.5)
    tol = dict(rtol=0, atol=1e-6)
    np.testing.assert_allclose(out["force"][0, :3], [-0.3, 0.6, 0.0], **tol)
    np.testing.assert_allclose(out["smooth_accel"][0, :3], [-0.12, 0.0, 0.0], **tol)
    np.testing.assert_allclose(out["pos"][0, :3], [10.0 + 0.05 * 0.294, 0.5, 10.0], **tol)
    np.testing.assert_allclose(out["forward"][0, :3], [1.0, 0.0, 0.0], **tol)
    assert out["forward"][0, 3] == pytest.approx(math.sqrt(0.246436), abs=1e-6)
    # the same step in 3-D keeps the y components
    steer3 = SteerParams.make(weights=(0.0, 0.0, 0.0), targets=[[10.0, 20.0, 10.0]], target_weight=1.0,
                              max_force=10.0)
    o3 = _alone(orc, GRID, PROX, steer3, (10.0, 10.0, 10.0), (0.6, 0.8, 0.0), 0.5)
    v3 = np.array([0.3, 0.4, 0.0]) + 0.05 * 0.4 * np.array([-0.3, 0.6, 0.0])
    np.testing.assert_allclose(o3["pos"][0, :3], np.array([10.0, 10.0, 10.0]) + 0.05 * v3, **tol)
    np.testing.assert_allclose(o3["forward"][0, :3], v3 / np.linalg.norm(v3), **tol)


def test_sensing_min_radius_counts_a_neighbour_behind(orc):
    """inBoidNeighborhood (OpenSteerWrapper_resources.cu:67-92): a neighbour
    closer than sensing_min_radius counts whatever its direction; past it the
    angle test applies.  A at (8,8,8) heading +x, B straight behind at (7,8,8)
    heading +y, cosMaxAngle 0.5 (a +-60 degree front cone):
      min radius 0 or exactly 1 (dist < minR is strict): B is outside the cone,
        no behaviour sees it, force 0;
      min radius 1.5: separation = normalize((-1,0,0) / -1) = (1,0,0),
        alignment = normalize((0,1,0) - (1,0,0)) = (-1,1,0)/sqrt(2),
        cohesion = (-1,0,0)."""
    pos, f, sa = agents([[8.0, 8.0, 8.0], [7.0, 8.0, 8.0]], [[1, 0, 0], [0, 1, 0]], [0.5, 0.5])
    w = (1.0, 0.8, 0.6)
    for min_r, counted in [(0.0, False), (1.0, False), (1.5, True)]:
        steer = SteerParams.make(weights=w, cos_max=0.5, min_radius=min_r, max_force=100.0)
        for f64 in (True, False):
            out = one_step(orc, GRID, PROX, steer, pos, f, sa, f64=f64)
            iA = int(np.nonzero(out["hash"][:, 1] == 0)[0][0])
            F = out["force"][iA, :3].astype(np.float64)
            if counted:
                expect = (w[0] * np.array([1.0, 0, 0]) + w[1] * np.array([-1.0, 1.0, 0]) / math.sqrt(2)
                          + w[2] * np.array([-1.0, 0, 0]))
            else:
                expect = np.zeros(3)
            np.testing.assert_allclose(F, expect, rtol=0, atol=1e-6)
