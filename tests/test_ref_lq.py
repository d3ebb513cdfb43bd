"""The reference tree's OpenSteer locality-query database (Samples/OpensteerDemo/
src/lq.c), built into oracle/_ref by oracle/build_ref.sh, as the bench's
reference-code CPU neighbour search: its neighbour counts equal a brute-force
count with lq's own test (d^2 < r^2, strict, self excluded) for agents in the
lattice's box, on bin faces and coincident, on 1 and 4 threads.  (Agents
outside the box are left out: lq truncates bin coordinates toward zero, so a
sphere reaching less than a bin below the origin does not visit lq's list of
outside objects -- lq.c:469-483 -- a quirk of the reference code, not of the
driver.)"""
import numpy as np
import pytest

import oracle
from behavert_amd import workloads as W


def brute_count(pos, r):
    p = pos[:, :3].astype(np.float32)
    total = 0
    for i in range(p.shape[0]):
        d = p - p[i]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        total += int(np.count_nonzero(d2 < np.float32(r * r))) - 1
    return total


@pytest.mark.skipif(not __import__("os").path.exists(oracle.REF_LQ_PATH), reason="oracle/_ref not built (no reference checkout)")
@pytest.mark.parametrize("threads", [1, 4])
def test_lq_counts_match_brute_force(threads):
    side = 40.0
    g = W.cube_grid(side)
    st = W.uniform_flock(2000, side, 3)
    pos = st["pos"].copy()
    pos[:20, 0] = 4.0 * np.arange(20) % side   # on bin faces
    pos[20:40, 1] = 0.0                         # on the box's low face
    pos[40:60] = pos[60:80]                     # coincident pairs (distance 0)
    res = oracle.lq_flock_baseline(pos, g, 4.0, threads)
    assert res is not None
    assert res["neighbours"] == brute_count(pos, 4.0)
    assert res["build_s"] >= 0 and res["query_s"] >= 0
