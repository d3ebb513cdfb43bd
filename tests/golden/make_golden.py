"""Writes tests/golden/golden_small.npz: for every case of cases.py, the inputs and,
after each step, the FP32 oracle's hash (sorted key/ID pairs), cell ranges,
neighbour lists and new state (plus features / wander where the case has them).

Run:  python tests/golden/make_golden.py   (needs oracle/liboracle.so)."""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from tests.golden.cases import DT, FIELDS, cases  # noqa: E402

OUT = os.path.join(HERE, "golden_small.npz")


def main():
    arrays = {}
    for name, c in cases().items():
        st = c["state"]
        for k in ("pos", "forward", "smooth_accel"):
            arrays[f"{name}/in/{k}"] = st[k]
        hf = oracle.HostFlock(c["grid"], c["prox"], st, features=c.get("features"), wander=c.get("wander"))
        for t in range(c["steps"]):
            hf.step(c["steer"], DT, t, c.get("feature_rate", 0.0))
            for f in FIELDS:
                arrays[f"{name}/step{t}/{f}"] = np.array(getattr(hf, f))
            if c.get("features") is not None:
                arrays[f"{name}/step{t}/features"] = np.array(hf.features)
            if c.get("wander") is not None:
                arrays[f"{name}/step{t}/wander"] = np.array(hf.wander)
    np.savez_compressed(OUT, **arrays)
    print(OUT, os.path.getsize(OUT), "bytes,", len(arrays), "arrays")


if __name__ == "__main__":
    main()
