"""The C ABI library loads without a GPU and exports every entry point that
include/behavert_amd.h declares; host-only helpers are checked against closed
forms.  No device work is launched here."""
import ctypes as C
import math
import os
import re
import subprocess

import numpy as np
import pytest

import behavert_amd as B
from behavert_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")]


def declared_symbols():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(brt_[a-z0-9_]+)\s*\(", text))
    return names


@pytest.fixture(scope="module")
def L():
    if not B.is_built():
        pytest.fail("libbehavert_amd.so is not built (run __graft_entry__.build())")
    return B.lib()


def test_every_declared_symbol_is_exported_and_bound(L):
    names = declared_symbols()
    assert len(names) >= 19
    out = subprocess.run(["nm", "-D", "--defined-only", B.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (brt_[a-z0-9_]+)", out.stdout))
    missing = names - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"
    unbound = names - set(B.SIGNATURES)
    assert not unbound, f"declared but not bound in behavert_amd._lib: {sorted(unbound)}"
    for n in names:
        assert getattr(L, n) is not None


def test_version_and_list_stride(L):
    assert b"gfx950" in L.brt_version()
    for k in range(1, 253):
        assert L.brt_list_stride(k) == 4 * (k // 4 + 1) == B.list_stride(k)
        assert L.brt_list_stride(k) >= k + 1


@pytest.mark.parametrize("r", [4.0, 1.0, 0.5, 3.3, 1e-20, 1e20, 7.77])
def test_radius_threshold_is_exact_boundary(L, r):
    t = np.float32(L.brt_radius_threshold(r))
    rf = np.float32(r)
    assert np.sqrt(t) <= rf
    nxt = np.nextafter(t, np.float32(np.inf))
    assert np.sqrt(nxt) > rf


def test_radius_threshold_special_values(L):
    assert L.brt_radius_threshold(0.0) == 0.0
    assert L.brt_radius_threshold(-1.0) == -math.inf
    assert L.brt_radius_threshold(math.inf) == math.inf
    assert L.brt_radius_threshold(math.nan) == math.inf


def test_grid_setup_matches_oracle(L, orc):
    for c, r, cell in ((50, 42, 4), (200, 192, 4), (320, 312, 4), (10.5, 3.25, 0.75)):
        g = B.GridParams()
        fc = (C.c_float * 3)(c, c, c)
        fr = (C.c_float * 3)(r, r, r)
        fs = (C.c_float * 3)(cell, cell, cell)
        assert L.brt_grid_setup(fc, fr, fs, 0, C.byref(g)) == 0
        o = orc.grid_setup([c] * 3, [r] * 3, [cell] * 3)
        assert list(g.grid_size) == list(o.grid_size)
        assert list(g.world_origin) == list(o.world_origin)
        assert g.num_cells == o.num_cells
        assert g.max_bodies_per_cell == B.UNLIMITED


def test_context_create_without_gpu_fails_cleanly(L):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = L.brt_context_create(0, C.byref(h))
    assert rc != 0 and not h.value


def test_workload_generators_are_deterministic():
    a = W.plain_4096(seed=0)
    b = W.plain_4096(seed=0)
    for k in ("pos", "forward", "smooth_accel"):
        assert np.array_equal(a.state[k], b.state[k])
    f = a.state["forward"]
    assert np.allclose(np.linalg.norm(f[:, :3], axis=1), 1.0, atol=1e-6)
    assert f[:, 3].min() >= 0.5 and f[:, 3].max() <= 1.0
    p = a.state["pos"]
    assert p[:, :3].min() >= 0 and p[:, :3].max() < 100 and np.all(p[:, 3] == 1)
    blobs = W.blob_flock(1 << 14, 400.0, 64, 10.0, 0)
    assert blobs["pos"][:, :3].max() < 400 and blobs["pos"][:, :3].min() >= 0


def test_build_flags_name_existing_sources():
    """Every per-file compiler flag set in the build names a source that is
    built, and no flag set changes arithmetic (only scheduler options)."""
    from behavert_amd import build as hb
    for src, flags in hb.SOURCE_FLAGS.items():
        assert src in hb.SOURCES and os.path.exists(os.path.join(hb.CSRC, src)), src
        opts = [f for f in flags if f != "-mllvm"]
        assert opts and all(o.startswith("-amdgpu-sched-strategy=") or o == "-amdgpu-use-amdgpu-trackers"
                            for o in opts), (src, flags)
    for f in ("-ffp-contract=off", "-fno-fast-math", "-packed-fp32-ops"):
        assert f in hb.HIPCC_FLAGS
