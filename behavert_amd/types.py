"""ctypes mirrors of the POD parameter structs declared in include/behavert_amd.h.

These carry the reference's plug-in parameters:
  GridParams   <- EnvGrid3DParams   (PlugIns/EnvGrid3D/include/envgrid3d_kernel.cuh:38-65)
  ProxParams   <- Proximity3DParams (PlugIns/Proximity3D/include/Proximity3D_kernel.cuh:40-73)
  SteerParams  <- OpenSteerWrapperParams + Body3DParams
                  (PlugIns/OpenSteerWrapper/include/OpenSteerWrapperPlugIn.h:134-142,
                   PlugIns/Body/include/body3d_kernel.cuh:39-58)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

BRT_OK = 0
BRT_ERR_INVALID_ARGUMENT = -1
BRT_ERR_HIP = -2
BRT_ERR_OUT_OF_MEMORY = -3
BRT_ERR_UNSUPPORTED = -4
BRT_ERR_DEPENDENCY = -5

EMPTY_CELL = 0xFFFFFFFF
LIST_PAD = 0xFFFFFFFF
UNLIMITED = 0xFFFFFFFF
MAX_TARGETS = 16
MAX_NEIGHBORS = 252
NUM_PHASES = 5
OPT_FUSED_STEP = 1
OPT_UNIFORM_MAX_CELLS = 2
OPT_XCD_REMAP = 3
OPT_LDS_LISTS = 4
OPT_DEVICE_FLAGS = 5
OPT_STEP_GRAPH = 6
DEVFLAG_SORT_LOOKBACK = 1
PHASE_NAMES = ("hash_sort", "reorder_ranges", "neighbors", "steer_integrate", "features")


class GridParams(C.Structure):
    _fields_ = [
        ("world_origin", C.c_float * 3),
        ("cell_size", C.c_float * 3),
        ("grid_size", C.c_uint32 * 3),
        ("num_cells", C.c_uint32),
        ("max_bodies_per_cell", C.c_uint32),
        ("world_center", C.c_float * 3),
        ("world_radius", C.c_float * 3),
    ]

    @classmethod
    def direct(cls, origin, cell, size, center=None, radius=None, max_bodies=UNLIMITED):
        g = cls()
        for a in range(3):
            g.world_origin[a] = float(origin[a])
            g.cell_size[a] = float(cell[a])
            g.grid_size[a] = int(size[a])
        g.num_cells = int(size[0]) * int(size[1]) * int(size[2])
        g.max_bodies_per_cell = int(max_bodies)
        if center is None:
            center = [origin[a] + 0.5 * cell[a] * size[a] for a in range(3)]
        if radius is None:
            radius = [0.5 * cell[a] * size[a] - 2.0 * cell[a] for a in range(3)]
        for a in range(3):
            g.world_center[a] = float(center[a])
            g.world_radius[a] = float(radius[a])
        return g

    def as_dict(self):
        return {k: (list(getattr(self, k)) if hasattr(getattr(self, k), "__len__") else getattr(self, k))
                for k, _ in self._fields_}


class ProxParams(C.Structure):
    _fields_ = [
        ("max_neighbors", C.c_uint32),
        ("search_radius", C.c_float),
        ("use_knn", C.c_uint32),
    ]

    @classmethod
    def make(cls, max_neighbors, radius, use_knn=False):
        p = cls()
        p.max_neighbors = int(max_neighbors)
        p.search_radius = float(radius)
        p.use_knn = 1 if use_knn else 0
        return p


class SteerParams(C.Structure):
    _fields_ = [
        ("separation", C.c_float * 3),
        ("alignment", C.c_float * 3),
        ("cohesion", C.c_float * 3),
        ("mass", C.c_float),
        ("max_speed", C.c_float),
        ("max_force", C.c_float),
        ("sensing_min_radius", C.c_float),
        ("seek_world_center", C.c_uint32),
        ("use_2d_projection", C.c_uint32),
        ("num_targets", C.c_uint32),
        ("target_weight", C.c_float),
        ("targets", (C.c_float * 4) * MAX_TARGETS),
        ("wander_weight", C.c_float),
        ("wander_seed", C.c_uint32),
    ]

    @classmethod
    def make(cls, weights=(1.0, 0.8, 0.6), radius=4.0, cos_max=-1.0, mass=1.0, max_speed=1.0,
             max_force=0.5, min_radius=0.0, seek_world_center=False, use_2d=False,
             targets=None, target_weight=0.0, wander_weight=0.0, wander_seed=0):
        s = cls()
        for arr, w in ((s.separation, weights[0]), (s.alignment, weights[1]), (s.cohesion, weights[2])):
            arr[0] = float(w)
            arr[1] = float(radius)
            arr[2] = float(cos_max)
        s.mass = float(mass)
        s.max_speed = float(max_speed)
        s.max_force = float(max_force)
        s.sensing_min_radius = float(min_radius)
        s.seek_world_center = 1 if seek_world_center else 0
        s.use_2d_projection = 1 if use_2d else 0
        s.set_targets(targets, target_weight)
        s.wander_weight = float(wander_weight)
        s.wander_seed = int(wander_seed) & 0xFFFFFFFF
        return s

    def set_targets(self, targets, weight=None):
        if weight is not None:
            self.target_weight = float(weight)
        if targets is None:
            self.num_targets = 0
            return
        t = np.asarray(targets, dtype=np.float32).reshape(-1, 3)
        if len(t) > MAX_TARGETS:
            raise ValueError(f"at most {MAX_TARGETS} targets")
        self.num_targets = len(t)
        for k in range(len(t)):
            for a in range(3):
                self.targets[k][a] = float(t[k, a])
            self.targets[k][3] = 0.0

    def weights(self):
        return (self.separation[0], self.alignment[0], self.cohesion[0])


class StepBuffers(C.Structure):
    _fields_ = [
        ("n", C.c_uint32),
        ("pos", C.c_void_p),
        ("forward", C.c_void_p),
        ("smooth_accel", C.c_void_p),
        ("sorted_pos", C.c_void_p),
        ("sorted_forward", C.c_void_p),
        ("sorted_smooth_accel", C.c_void_p),
        ("hash", C.c_void_p),
        ("cell_start", C.c_void_p),
        ("cell_end", C.c_void_p),
        ("lists", C.c_void_p),
        ("features", C.c_void_p),
        ("features_out", C.c_void_p),
        ("wander", C.c_void_p),
    ]


MAX_FEATURES = 32

SNAP_LISTS = 1
SNAP_FEATURES = 2
SNAP_WANDER = 4


class SnapshotInfo(C.Structure):
    """brt_snapshot_info: the header of a state snapshot."""
    _fields_ = [
        ("version", C.c_uint32),
        ("n", C.c_uint32),
        ("max_neighbors", C.c_uint32),
        ("list_stride", C.c_uint32),
        ("flags", C.c_uint32),
        ("step", C.c_uint32),
        ("grid", GridParams),
        ("prox", ProxParams),
        ("payload_bytes", C.c_uint64),
        ("checksum", C.c_uint64),
    ]


class BcParams(C.Structure):
    """BehaviorClusteringParams subset read by the similarity and cluster-steering
    kernels (PlugIns/BehaviorClustering/include/BehaviorClustering_kernel.cuh:40-59)."""
    _fields_ = [
        ("feature_count", C.c_uint32),
        ("adaptive", C.c_uint32),
        ("avg_similarity_controller", C.c_float),
        ("stats_params", C.c_float * 3),
        ("cluster_weight", C.c_float),
        ("alignment_weight", C.c_float),
    ]

    @classmethod
    def make(cls, feature_count, adaptive=True, stats_params=(0.1, 0.1, 0.001), cluster_weight=1.0,
             alignment_weight=1.0, avg_similarity_controller=1.0):
        b = cls()
        b.feature_count = int(feature_count)
        b.adaptive = 1 if adaptive else 0
        b.avg_similarity_controller = float(avg_similarity_controller)
        for k in range(3):
            b.stats_params[k] = float(stats_params[k])
        b.cluster_weight = float(cluster_weight)
        b.alignment_weight = float(alignment_weight)
        return b


class FlockDesc(C.Structure):
    """One flock of a batch: its own grid, agent count and behaviour weights."""
    _fields_ = [("grid", GridParams), ("n", C.c_uint32), ("weights", C.c_float * 3)]


class BatchConfig(C.Structure):
    """Tunables of the batched step (brt_batch_config)."""
    _fields_ = [
        ("prox_block", C.c_uint32),
        ("uniform_max_cells", C.c_uint32),
        ("xcd_remap", C.c_uint32),
        ("sort_items", C.c_uint32),
        ("lds_lists", C.c_uint32),
        ("use_graph", C.c_uint32),
    ]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


def list_stride(k: int) -> int:
    """W = 4*(K/4+1) uint slots per neighbour list (Proximity3DPlugIn.h:148-149)."""
    return 4 * (int(k) // 4 + 1)
