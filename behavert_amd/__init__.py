"""behavert_amd: MI355X-native (gfx950) implementation of BehaveRT's per-step
agent simulation path behind a C ABI (include/behavert_amd.h).

The compute path is the HIP library libbehavert_amd.so built in-tree by
behavert_amd/build.py; importing the package does not require a GPU, but every
call that launches device work needs the built library and a ROCm device.
"""
from ._lib import LIB_PATH, SIGNATURES, BrtError, is_built, lib
from .types import (EMPTY_CELL, LIST_PAD, MAX_NEIGHBORS, MAX_TARGETS, NUM_PHASES, PHASE_NAMES, UNLIMITED,
                    GridParams, ProxParams, SteerParams, StepBuffers, list_stride)

__all__ = [
    "LIB_PATH", "SIGNATURES", "BrtError", "is_built", "lib", "EMPTY_CELL", "LIST_PAD", "MAX_NEIGHBORS",
    "MAX_TARGETS", "NUM_PHASES", "PHASE_NAMES", "UNLIMITED", "GridParams", "ProxParams", "SteerParams",
    "StepBuffers", "list_stride",
]
