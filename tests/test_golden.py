"""The committed golden fixtures (tests/golden/golden_small.npz, written by
tests/golden/make_golden.py with the FP32 oracle): the seeded inputs still match
the generators, and the oracle still reproduces every stored output bit for bit.
The GPU side of the same fixtures is tests/test_gpu_golden.py."""
import os

import numpy as np
import pytest

from tests.golden.cases import DT, FIELDS, cases

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_small.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(FIXTURE, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", list(cases()))
def test_inputs_match_generators(gold, name):
    st = cases()[name]["state"]
    for k in ("pos", "forward", "smooth_accel"):
        np.testing.assert_array_equal(gold[f"{name}/in/{k}"].view(np.uint32), st[k].view(np.uint32))


@pytest.mark.parametrize("name", list(cases()))
def test_oracle_reproduces_fixture(orc, gold, name):
    c = cases()[name]
    st = {k: gold[f"{name}/in/{k}"] for k in ("pos", "forward", "smooth_accel")}
    hf = orc.HostFlock(c["grid"], c["prox"], st, features=c.get("features"), wander=c.get("wander"))
    for t in range(c["steps"]):
        hf.step(c["steer"], DT, t, c.get("feature_rate", 0.0))
        for f in FIELDS + (("features",) if c.get("features") is not None else ()) + \
                (("wander",) if c.get("wander") is not None else ()):
            np.testing.assert_array_equal(np.asarray(getattr(hf, f)).view(np.uint32),
                                          gold[f"{name}/step{t}/{f}"].view(np.uint32), err_msg=f"{name} {t} {f}")


def test_fixture_covers_the_special_agents(gold):
    """The hand-placed agents: duplicate and exact-r neighbours are listed, the
    out-of-grid agents carry the clamped last-cell key."""
    h = gold["plain/step0/hash"]
    L = gold["plain/step0/lists"]
    pos = {int(i): k for k, i in enumerate(h[:, 1])}
    row0 = L[pos[0]]
    members = set(int(x) for x in row0[1:1 + row0[0]])
    assert pos[1] in members and pos[2] in members
    G = 10
    key3 = int(h[pos[3], 0])
    assert key3 % G == G - 1          # x < 0 clamps to the last column (unsigned min)
    key4 = int(h[pos[4], 0])
    assert key4 % G == G - 1
