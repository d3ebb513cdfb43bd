"""The C++ plug-in layer (behavert_amd/plugins: the reference's mixin chain over the
C ABI).  CPU: the configuration parser keeps BehaveRT's semantics
(Core/src/UtilityConfigFile.cpp:69-160, UtilityStringConverter.cpp:86-130).
GPU: a BehaveRT_Flocking-style chain driven through the plug-in API gives the
oracle's state bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from behavert_amd import build as B
from behavert_amd import workloads as W

CFG = os.path.join(os.path.dirname(B.EXAMPLE_SRC), "BehaveRT_Flocking.cfg")


def demo():
    if not os.path.exists(B.EXAMPLE_BIN):
        pytest.fail("plug-in example not built (run __graft_entry__.build())")
    return B.EXAMPLE_BIN


def dump(cfg_path, *pairs):
    out = subprocess.run([demo(), "--dump-config", cfg_path, *pairs], capture_output=True, text=True, check=True)
    return [line.split("|") for line in out.stdout.strip().splitlines()]


def test_config_parser_semantics(tmp_path):
    p = tmp_path / "t.cfg"
    p.write_bytes(b"# comment\r\n@ also comment\r\n[A]\r\nx = 1.5\r\nflag:yes\r\nv\t1 2 3\r\nbad = 1 2\r\n"
                  b"[B]\nk=7\nempty =\n")
    rows = dump(str(p), "A", "x", "A", "flag", "A", "v", "A", "bad", "B", "k", "B", "empty", "A", "missing",
                "C", "x")
    assert rows[0][2:6] == ["1.5", "1.5", "1", "1"]
    assert rows[1][2] == "yes" and rows[1][5] == "1"
    assert rows[2][2] == "1 2 3" and rows[2][6] == "1 2 3"
    assert rows[3][6] == "0 0 0"           # parseFloat3 needs three numbers
    assert rows[4][4] == "7"
    assert rows[5][2] == "" and rows[5][3] == "0"
    assert rows[6][2] == "" and rows[6][5] == "0"   # missing key -> "" -> 0 / false
    assert rows[7][2] == ""                # missing section


def test_shipped_config_matches_plain_workload():
    rows = {(r[0], r[1]): r for r in dump(CFG, "Body3DPlugIn", "numBodies", "EnvGrid3DPlugIn", "worldCenter",
                                          "EnvGrid3DPlugIn", "worldRadius", "Proximity3DPlugIn", "maxNeighbors",
                                          "Proximity3DPlugIn", "commonSearchRadius")}
    w = W.plain_4096()
    assert int(rows[("Body3DPlugIn", "numBodies")][4]) == w.n
    assert rows[("EnvGrid3DPlugIn", "worldCenter")][6] == "50 50 50"
    assert rows[("EnvGrid3DPlugIn", "worldRadius")][6] == "42 42 42"
    assert int(rows[("Proximity3DPlugIn", "maxNeighbors")][4]) == w.prox.max_neighbors
    assert float(rows[("Proximity3DPlugIn", "commonSearchRadius")][3]) == w.prox.search_radius


@pytest.mark.gpu
def test_plugin_chain_matches_oracle(tmp_path, orc):
    w = W.plain_4096()
    st = w.state
    init = tmp_path / "init.bin"
    out = tmp_path / "out.bin"
    with open(init, "wb") as f:
        for k in ("pos", "forward", "smooth_accel"):
            f.write(np.ascontiguousarray(st[k], np.float32).tobytes())
    steps = 10
    r = subprocess.run([demo(), CFG, str(steps), str(init), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, np.float32).reshape(3, w.n, 4)
    hf = orc.HostFlock(w.grid, w.prox, st)
    for i in range(steps):
        hf.step(w.steer, w.dt, i)
    for a, k in zip(got, ("pos", "forward", "smooth_accel")):
        assert np.array_equal(a.view(np.uint32), getattr(hf, k).view(np.uint32)), k


@pytest.mark.gpu
def test_clustering_plugin_chain_matches_oracle(tmp_path, orc):
    """BehaveRT_Clustering's update sequence through the plug-in API
    (BehaviorClusteringPlugIn::steerForClustering between the flocking
    behaviours and the world-centre seek, computeClusterMembership after it)
    against the oracle's passes run in the same order, bit for bit after 5
    steps."""
    from behavert_amd.types import BcParams, SteerParams

    from ._parity import oracle_grid_outputs

    if not os.path.exists(B.CLUSTERING_BIN):
        pytest.fail("clustering example not built (run __graft_entry__.build())")
    cfg = os.path.join(os.path.dirname(B.CLUSTERING_SRC), "BehaveRT_Clustering.cfg")
    w = W.plain_4096()
    F = 6
    rng = np.random.default_rng(5)
    feats = rng.uniform(0, 1, (w.n, F)).astype(np.float32)
    stats = np.tile(np.array([1.0, 0.5, 0.0, 0.0], np.float32), (w.n, 1))
    init = tmp_path / "init.bin"
    out = tmp_path / "out.bin"
    with open(init, "wb") as f:
        for a in (w.state["pos"], w.state["forward"], w.state["smooth_accel"], feats, stats):
            f.write(np.ascontiguousarray(a, np.float32).tobytes())
    steps = 5
    r = subprocess.run([B.CLUSTERING_BIN, cfg, str(steps), str(init), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, np.float32)
    got = raw[:5 * w.n * 4].reshape(5, w.n, 4)
    memb = raw[5 * w.n * 4:].reshape(w.n, 3)

    sp = SteerParams.make(weights=(1.0, 0.8, 0.6), seek_world_center=True)
    bc = BcParams.make(F, adaptive=True, stats_params=(0.1, 0.1, 0.001), cluster_weight=0.9, alignment_weight=0.7)
    st = {k: np.array(w.state[k], np.float32) for k in ("pos", "forward", "smooth_accel")}
    for _ in range(steps):
        o = oracle_grid_outputs(orc, w.grid, w.prox, st)
        f0, _ = orc.steer_forces(w.prox, sp, w.dt, 0, o["sorted_pos"], o["sorted_forward"], o["hash"], o["lists"])
        sims, stats = orc.bc_similarities(w.prox, bc, o["hash"], o["lists"], feats, stats)
        f1 = orc.bc_cluster_steer(w.prox, sp, bc, o["sorted_pos"], o["sorted_forward"], o["lists"], sims, f0)
        m, _ = orc.cluster_membership(w.prox, 10, o["hash"], o["lists"], o["sorted_pos"])
        res = orc.apply_steering_force(w.grid, sp, w.dt, o["sorted_pos"], o["sorted_forward"],
                                       o["sorted_smooth_accel"], o["hash"], f1)
        st = {k: res[k] for k in ("pos", "forward", "smooth_accel")}
    ref = (st["pos"], st["forward"], st["smooth_accel"], stats, res["force"])
    for a, b, k in zip(got, ref, ("pos", "forward", "smooth_accel", "stats", "force")):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k
    assert np.array_equal(memb.view(np.uint32), m.view(np.uint32)), "cluster membership"
