"""The device step path against the committed golden fixtures (bit for bit): the
same seeded inputs, the same steps, every stored output."""
import os

import numpy as np
import pytest

from behavert_amd.engine import Flock
from tests.golden.cases import DT, FIELDS, cases

from ._parity import first_mismatch

pytestmark = pytest.mark.gpu
FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_small.npz")


@pytest.mark.parametrize("name", list(cases()))
def test_device_step_reproduces_fixture(ctx, name):
    with np.load(FIXTURE, allow_pickle=False) as z:
        gold = {k: z[k] for k in z.files}
    c = cases()[name]
    st = {k: gold[f"{name}/in/{k}"] for k in ("pos", "forward", "smooth_accel")}
    fl = Flock(ctx, c["grid"], c["prox"], c["steer"], st, features=c.get("features"), wander=c.get("wander"))
    for t in range(c["steps"]):
        print(name, "step", t, flush=True)
        fl.step(DT, t, c.get("feature_rate", 0.0))
        out = {**fl.grid_outputs(), **fl.state()}
        extra = (("features",) if c.get("features") is not None else ()) + \
                (("wander",) if c.get("wander") is not None else ())
        for f in FIELDS + extra:
            m = first_mismatch(np.asarray(out[f]).view(np.uint32), gold[f"{name}/step{t}/{f}"].view(np.uint32))
            assert m is None, (name, t, f, m)
