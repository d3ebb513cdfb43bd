"""Compiles the gfx950 HIP library (libbehavert_amd.so) in-tree with hipcc.

Every kernel TU is built with -ffp-contract=off (no FMA contraction: the
reference arithmetic is restated with one rounding per operation) and with the
packed-FP32 target feature disabled.  The TUs are compiled one by one (in
parallel) and linked, so a TU can carry flags of its own (SOURCE_FLAGS).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbehavert_amd.so")
SOURCES = ["brt_envgrid.hip", "brt_proximity.hip", "brt_opensteer.hip", "brt_clustering.hip", "brt_consumers.hip",
           "brt_batch.hip", "brt_adjoint.hip", "brt_snapshot.hip", "brt_capi.hip"]
ARCH = os.environ.get("BRT_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]

# Per-TU flags, each measured against the default scheduler on its own
# workload (A/B on one box): the neighbour kernels (brt_proximity.hip), the
# sort (brt_envgrid.hip), the adjoint (brt_adjoint.hip) and the list
# consumers (brt_consumers.hip) are scheduled for
# instruction-level parallelism (LLVM's max-ilp machine scheduler: fused
# neighbour kernel 421 -> 402 us on the 1M-boid blob flock at the same 108
# VGPRs, onesweep pass 19.7 -> 19.4 us, adjoint step 56 -> 54 us,
# steer-and-slow-to-avoid 179 -> 167 us); the steering
# and clustering kernels (brt_opensteer.hip, brt_clustering.hip) schedule with
# the AMDGPU register-pressure trackers (4M full-stack steering + feature pass
# 832 -> 789 us, similarities 343 -> 324 us), where max-ilp measured 1-4 %
# slower.  Scheduling reorders instructions only: every result bit is the same.
_ILP = ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]
_TRACKERS = ["-mllvm", "-amdgpu-use-amdgpu-trackers"]
SOURCE_FLAGS = {
    "brt_proximity.hip": _ILP,
    "brt_envgrid.hip": _ILP,
    "brt_adjoint.hip": _ILP,
    "brt_consumers.hip": _ILP,
    "brt_opensteer.hip": _TRACKERS,
    "brt_clustering.hip": _TRACKERS,
}


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found (ROCm is required to build behavert_amd)")
    return path


EXAMPLE_SRC = os.path.join(HERE, "examples", "flocking_demo.cpp")
EXAMPLE_BIN = os.path.join(HERE, "examples", "flocking_demo")
CLUSTERING_SRC = os.path.join(HERE, "examples", "clustering_demo.cpp")
CLUSTERING_BIN = os.path.join(HERE, "examples", "clustering_demo")


def build_example(verbose: bool = False) -> str:
    """The plug-in layer's BehaveRT_Flocking- and BehaveRT_Clustering-style
    drivers, linked to the library."""
    for src, out in ((EXAMPLE_SRC, EXAMPLE_BIN), (CLUSTERING_SRC, CLUSTERING_BIN)):
        cmd = [hipcc(), "-std=c++17", "-O2", "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
               "-packed-fp32-ops", "-Wall", "-Wno-unused-command-line-argument", "-o", out, src,
               "-L", HERE, "-lbehavert_amd", "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        proc = subprocess.run(cmd, capture_output=True, text=True)
        if proc.returncode != 0:
            sys.stderr.write(proc.stdout + proc.stderr)
            raise RuntimeError(f"building {os.path.basename(src)} failed with exit code {proc.returncode}")
    return EXAMPLE_BIN


def _run(cmd, what: str, verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"{what} failed with exit code {proc.returncode}")
    if verbose and proc.stderr.strip():
        sys.stderr.write(proc.stderr)


def build(verbose: bool = False) -> str:
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    with tempfile.TemporaryDirectory(prefix="brt_build_") as objdir:
        objs = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in srcs]
        cmds = [[hipcc(), *HIPCC_FLAGS, *SOURCE_FLAGS.get(s, []), "-c", "-o", o, os.path.join(CSRC, s)]
                for s, o in zip(srcs, objs)]
        jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
        with ThreadPoolExecutor(max_workers=jobs) as pool:
            for f in [pool.submit(_run, c, f"compiling {s}", verbose) for c, s in zip(cmds, srcs)]:
                f.result()
        _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], "linking", verbose)
    os.replace(tmp, LIB)
    build_example(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
