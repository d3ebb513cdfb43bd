"""Pins of the oracle's small list consumers (steerToAvoidNeighbors,
extractColorFromNeighborhoodD, moveTowardsTargetD) by hand-derived known
answers on hand-built lists."""
import math

import numpy as np
import pytest

from behavert_amd.types import ProxParams, SteerParams, list_stride


def lists_for(rows, K):
    W = list_stride(K)
    L = np.full((len(rows), W), 0xFFFFFFFF, np.uint32)
    for i, r in enumerate(rows):
        L[i, 0] = len(r)
        L[i, 1:1 + len(r)] = r
    return L


@pytest.mark.parametrize("other, factor", [([1.0, 0, 0], 0.9), ([-1.0, 0.1, 0], 1.02), ([0, 1.0, 0], 1.0)])
def test_slow_avoid_speed_rule(orc, other, factor):
    """One neighbour ahead (+x), behind (-x, slightly off axis: straight behind
    fails the strict cos > -1 cone test) or beside (+y) of an agent heading +x
    at speed 0.5: separation points away from it, so the speed is multiplied by
    0.9 (opposed), 1.02 (agreeing) or kept (orthogonal); the force is the unit
    separation times its weight 2 plus the incoming force."""
    K = 8
    spos = np.array([[0, 0, 0, 1], other + [1]], np.float32)
    sfwd = np.array([[1, 0, 0, 0.5], [1, 0, 0, 0.5]], np.float32)
    L = lists_for([[1], []], K)
    sp = SteerParams.make(weights=(2.0, 0.0, 0.0), radius=4.0)
    fin = np.array([[0.1, 0.2, 0.3, 7.0], [0, 0, 0, 0]], np.float32)
    fwd, force = orc.slow_avoid(ProxParams.make(K, 4.0), sp, spos, sfwd, L, fin)
    assert fwd[0, 3] == np.float32(0.5 * factor)
    away = -np.array(other, np.float64) / np.linalg.norm(other)
    np.testing.assert_allclose(force[0, :3], 2.0 * away + fin[0, :3], rtol=0, atol=5e-7)   # a few FP32 ulps at 2
    assert force[0, 3] == 0.0
    np.testing.assert_array_equal(fwd[1], sfwd[1])          # no neighbours: speed kept
    np.testing.assert_array_equal(force[1, :3], [0, 0, 0])


def test_slow_avoid_crowd_amplification(orc):
    """20 listed neighbours on the +x side (separation -x, weight 1): above 16
    the steering is scaled by (20 - 16) * 0.5 = 2, and the speed drops by 10 %."""
    K = 24
    n = 21
    rng = np.random.default_rng(0)
    pos = np.zeros((n, 4), np.float32)
    pos[:, 3] = 1
    # neighbours on the +x axis at distinct distances: every offset is along +x
    pos[1:, 0] = np.linspace(0.5, 3.5, 20).astype(np.float32)
    rng.shuffle(pos[1:, 0])
    fwd = np.tile(np.array([1, 0, 0, 0.8], np.float32), (n, 1))
    L = lists_for([list(range(1, n))] + [[]] * (n - 1), K)
    _, force = orc.slow_avoid(ProxParams.make(K, 4.0), SteerParams.make(weights=(1.0, 0, 0), radius=4.0),
                              pos, fwd, L, np.zeros((n, 4), np.float32))
    np.testing.assert_allclose(force[0, :3], [-2.0, 0, 0], rtol=0, atol=1e-6)
    fwd_out, _ = orc.slow_avoid(ProxParams.make(K, 4.0), SteerParams.make(weights=(1.0, 0, 0), radius=4.0),
                                pos, fwd, L, np.zeros((n, 4), np.float32))
    assert fwd_out[0, 3] == np.float32(0.8 * 0.9)


@pytest.mark.parametrize("vertexes", [1, 3])
def test_neighborhood_color_known_answer(orc, vertexes):
    """count 8 of K 16 -> f = 1/2: (b.x + 1/2, b.y + 0.3 + 1 - 1/2, b.z + 1/4, 1)
    by ID; count 0 -> f = 0; the three-vertex form darkens the first by 0.5."""
    K = 16
    L = lists_for([list(range(1, 9)), [], [0]], K)
    h = np.array([[0, 2], [0, 0], [0, 1]], np.uint32)   # sorted 0 -> ID 2, 1 -> ID 0, 2 -> ID 1
    base = np.array([0.25, 0.5, -0.5], np.float32)
    c = orc.neighborhood_color(ProxParams.make(K, 4.0), base, vertexes, h, L)
    half = np.array([0.75, np.float32(0.5 + 0.3 + 1 - 0.5), -0.25, 1.0], np.float32)
    zero = np.array([0.25, np.float32(0.5 + 0.3 + 1 - 0.0), -0.5, 1.0], np.float32)
    np.testing.assert_array_equal(c[2, -1], half)
    np.testing.assert_array_equal(c[0, -1], zero)
    if vertexes == 3:
        np.testing.assert_array_equal(c[2, 0], half - np.array([0.5, 0.5, 0.5, 0], np.float32))
        np.testing.assert_array_equal(c[2, 1], half)


def test_move_towards_target_known_answers(orc):
    """n = 4, shift 2: ID 0 -> target 2, ID 1 -> 3, ID 2 -> index 4 = n (reads as
    zeros: the incoming force passes through), ID 3 -> 5 - 4 = 1.  Force =
    12 * t.w * unit((t + base) - pos) + force_in."""
    n = 4
    h = np.array([[0, 3], [0, 2], [0, 1], [0, 0]], np.uint32)    # sorted i -> ID 3 - i
    spos = np.zeros((n, 4), np.float32)
    targets = np.array([[9, 9, 9, 1], [3, 0, 0, 0.5], [0, 4, 0, 2], [0, 0, -1, 1]], np.float32)
    base = np.array([0, 0, 1], np.float32)
    fin = np.tile(np.array([1, 1, 1, 5], np.float32), (n, 1))
    f = orc.move_towards_target(2, base, h, spos, targets, fin)
    # sorted 3 = ID 0 -> target 2 = (0, 4, 0) + base -> unit (0, 4, 1)/sqrt(17), weight 24
    np.testing.assert_allclose(f[3, :3], 24 * np.array([0, 4, 1]) / math.sqrt(17) + 1, rtol=0, atol=1e-5)
    # sorted 2 = ID 1 -> target 3 = (0, 0, -1) + base = 0: normalize(0) is NaN, times w = 1
    assert np.isnan(f[2, :3]).all()
    # sorted 1 = ID 2 -> index 4 = n: zero target, w = 0 -> NaN * 0 only if the
    # offset is zero; here base - pos = (0, 0, 1) -> unit * 0 = 0
    np.testing.assert_array_equal(f[1, :3], [1, 1, 1])
    # sorted 0 = ID 3 -> 5 > 4 -> index 1 = (3, 0, 0) + base, weight 0.5 * 12 = 6
    np.testing.assert_allclose(f[0, :3], 6 * np.array([3, 0, 1]) / math.sqrt(10) + 1, rtol=0, atol=1e-6)
    assert (f[:, 3] == 0).all()
