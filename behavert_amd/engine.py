"""Python host layer over the C ABI: a context and a flock whose device buffers are
torch tensors (torch is used only for device memory and the current HIP stream).

Mirrors the reference's per-step plug-in calls (Samples/OpensteerDemo/plugins/
BehaveRT_Flocking.cpp:177-196):
  Flock.reorder_sim_data()       EnvGrid3DPlugIn::reorderSimData
  Flock.compute_neighborhoods()  Proximity3DPlugIn::computeNeighborhoods
  Flock.apply_steering(dt)       OpenSteerWrapperPlugIn::resetSteerForce ...
                                 applySteeringForces (fused)
  Flock.update_features(rate)    BehaviorClustering feature update
  Flock.step(dt, i)              the whole step in one C call (brt_step)
  Flock.save(path) / load(path)  binary state snapshots (brt_snapshot_*)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch

from ._lib import BrtError, lib
from .types import (NUM_PHASES, PHASE_NAMES, SNAP_FEATURES, SNAP_LISTS, SNAP_WANDER, BatchConfig, FlockDesc,
                    GridParams, ProxParams, SnapshotInfo, SteerParams, StepBuffers, list_stride)


def _ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensors passed to the C ABI must be contiguous")
    return t.data_ptr()


def current_stream() -> int:
    return torch.cuda.current_stream().cuda_stream


class Context:
    """Owns a brt_context (sort scratch, phase timers) on one device."""

    def __init__(self, device: int = 0):
        self._lib = lib()
        h = C.c_void_p()
        rc = self._lib.brt_context_create(device, C.byref(h))
        if rc != 0:
            raise BrtError(rc, f"brt_context_create(device={device}) failed")
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            self._lib.brt_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int):
        if rc != 0:
            msg = self._lib.brt_context_last_error(self.handle)
            raise BrtError(rc, msg.decode() if msg else "")

    def check_device(self):
        """Synchronises the device and raises BrtError if a kernel recorded a
        failure since the last check (brt_context_check)."""
        self.check(self._lib.brt_context_check(self.handle))

    # ------------------------------------------------------------ EnvGrid3D
    def calc_hash(self, grid: GridParams, pos, hash_out):
        self.check(self._lib.brt_envgrid_calc_hash(self.handle, C.byref(grid), _ptr(pos), pos.shape[0],
                                                   _ptr(hash_out), current_stream()))

    def sort(self, hash_, num_cells: int):
        self.check(self._lib.brt_envgrid_sort(self.handle, _ptr(hash_), hash_.shape[0], num_cells,
                                              current_stream()))

    def hash_sort(self, grid: GridParams, pos, hash_out):
        self.check(self._lib.brt_envgrid_hash_sort(self.handle, C.byref(grid), _ptr(pos), pos.shape[0],
                                                   _ptr(hash_out), current_stream()))

    def reorder(self, hash_, ins, outs):
        n = len(ins)
        a = (C.c_void_p * max(n, 1))(*[_ptr(t) for t in ins])
        b = (C.c_void_p * max(n, 1))(*[_ptr(t) for t in outs])
        self.check(self._lib.brt_envgrid_reorder(self.handle, _ptr(hash_), hash_.shape[0], n, a, b,
                                                 current_stream()))

    def cell_ranges(self, hash_, num_cells, cell_start, cell_end):
        self.check(self._lib.brt_prox_cell_ranges(self.handle, _ptr(hash_), hash_.shape[0], num_cells,
                                                  _ptr(cell_start), _ptr(cell_end), current_stream()))

    def reorder_ranges(self, hash_, ins, outs, num_cells, cell_start, cell_end):
        n = len(ins)
        a = (C.c_void_p * max(n, 1))(*[_ptr(t) for t in ins])
        b = (C.c_void_p * max(n, 1))(*[_ptr(t) for t in outs])
        self.check(self._lib.brt_envgrid_reorder_ranges(self.handle, _ptr(hash_), hash_.shape[0], n, a, b,
                                                        num_cells, _ptr(cell_start), _ptr(cell_end),
                                                        current_stream()))

    # ------------------------------------------------------------ Proximity3D
    def neighbors(self, grid, prox, sorted_pos, hash_, cell_start, cell_end, lists):
        self.check(self._lib.brt_prox_neighbors(self.handle, C.byref(grid), C.byref(prox), _ptr(sorted_pos),
                                                _ptr(hash_), _ptr(cell_start), _ptr(cell_end),
                                                sorted_pos.shape[0], _ptr(lists), current_stream()))

    # ------------------------------------------------------------ OpenSteerWrapper
    def steer_integrate(self, grid, prox, steer, dt, step, sorted_pos, sorted_fwd, sorted_sa, hash_, lists,
                        pos_out, fwd_out, sa_out, wander=None, force_out=None):
        self.check(self._lib.brt_osw_steer_integrate(
            self.handle, C.byref(grid), C.byref(prox), C.byref(steer), float(dt), int(step),
            sorted_pos.shape[0], _ptr(sorted_pos), _ptr(sorted_fwd), _ptr(sorted_sa), _ptr(hash_), _ptr(lists),
            _ptr(wander), _ptr(pos_out), _ptr(fwd_out), _ptr(sa_out), _ptr(force_out), current_stream()))

    def steer_forces(self, prox, steer, dt, step, sorted_pos, sorted_fwd, hash_, lists, force_out, wander=None):
        self.check(self._lib.brt_osw_steer_forces(
            self.handle, C.byref(prox), C.byref(steer), float(dt), int(step), sorted_pos.shape[0], _ptr(sorted_pos),
            _ptr(sorted_fwd), _ptr(hash_), _ptr(lists), _ptr(wander), _ptr(force_out), current_stream()))

    def apply_steering_force(self, grid, steer, dt, sorted_pos, sorted_fwd, sorted_sa, hash_, force_in, pos_out,
                             fwd_out, sa_out, force_out=None):
        self.check(self._lib.brt_osw_apply_steering_force(
            self.handle, C.byref(grid), C.byref(steer), float(dt), sorted_pos.shape[0], _ptr(sorted_pos),
            _ptr(sorted_fwd), _ptr(sorted_sa), _ptr(hash_), _ptr(force_in), _ptr(pos_out), _ptr(fwd_out),
            _ptr(sa_out), _ptr(force_out), current_stream()))

    def bc_similarities(self, prox, bc, hash_, lists, features, stats_in, stats_out, sims):
        self.check(self._lib.brt_bc_neigh_similarities(
            self.handle, C.byref(prox), C.byref(bc), hash_.shape[0], _ptr(hash_), _ptr(lists), _ptr(features),
            _ptr(stats_in), _ptr(stats_out), _ptr(sims), current_stream()))

    def bc_steer_for_clustering(self, prox, steer, bc, sorted_pos, sorted_fwd, lists, sims, force_in, force_out):
        self.check(self._lib.brt_bc_steer_for_clustering(
            self.handle, C.byref(prox), C.byref(steer), C.byref(bc), sorted_pos.shape[0], _ptr(sorted_pos),
            _ptr(sorted_fwd), _ptr(lists), _ptr(sims), _ptr(force_in), _ptr(force_out), current_stream()))

    def bc_similarities_steer(self, prox, steer, bc, hash_, lists, features, stats_in, stats_out, sims,
                              sorted_pos, sorted_fwd, force_in, force_out):
        self.check(self._lib.brt_bc_similarities_steer(
            self.handle, C.byref(prox), C.byref(steer), C.byref(bc), hash_.shape[0], _ptr(hash_), _ptr(lists),
            _ptr(features), _ptr(stats_in), _ptr(stats_out), _ptr(sims), _ptr(sorted_pos), _ptr(sorted_fwd),
            _ptr(force_in), _ptr(force_out), current_stream()))

    def cluster_membership(self, prox, iterations, hash_, lists, sorted_pos, membership_in, membership_out,
                           connector=None):
        self.check(self._lib.brt_bc_cluster_membership(
            self.handle, C.byref(prox), hash_.shape[0], int(iterations), _ptr(hash_), _ptr(lists), _ptr(sorted_pos),
            _ptr(membership_in), _ptr(membership_out), _ptr(connector), current_stream()))

    def cluster_hit_rate(self, prox, radius, hash_, lists, sorted_pos, membership_in, membership_out):
        self.check(self._lib.brt_bc_cluster_hit_rate(
            self.handle, C.byref(prox), float(radius), hash_.shape[0], _ptr(hash_), _ptr(lists), _ptr(sorted_pos),
            _ptr(membership_in), _ptr(membership_out), current_stream()))

    def steer_and_slow_to_avoid_neighbors(self, prox, steer, sorted_pos, sorted_fwd, lists, forward_out, force_in,
                                          force_out):
        self.check(self._lib.brt_osw_steer_and_slow_to_avoid_neighbors(
            self.handle, C.byref(prox), C.byref(steer), sorted_pos.shape[0], _ptr(sorted_pos), _ptr(sorted_fwd),
            _ptr(lists), _ptr(forward_out), _ptr(force_in), _ptr(force_out), current_stream()))

    def neighborhood_color(self, prox, color_base, vertexes, hash_, lists, color):
        base = (C.c_float * 3)(*[float(x) for x in color_base])
        self.check(self._lib.brt_draw_neighborhood_color(self.handle, C.byref(prox), base, int(vertexes),
                                                         hash_.shape[0], _ptr(hash_), _ptr(lists), _ptr(color),
                                                         current_stream()))

    def move_towards_target(self, index_shift, target_base, hash_, sorted_pos, targets, force_in, force_out):
        base = (C.c_float * 3)(*[float(x) for x in target_base])
        self.check(self._lib.brt_shapes_move_towards_target(self.handle, int(index_shift), base, hash_.shape[0],
                                                            _ptr(hash_), _ptr(sorted_pos), _ptr(targets),
                                                            _ptr(force_in), _ptr(force_out), current_stream()))

    def feature_update(self, prox, rate, hash_, lists, fin, fout):
        self.check(self._lib.brt_bc_feature_update(self.handle, C.byref(prox), float(rate), hash_.shape[0],
                                                   _ptr(hash_), _ptr(lists), _ptr(fin), _ptr(fout),
                                                   current_stream()))

    # ------------------------------------------------------------ options
    def set_option(self, option: int, value: int):
        self.check(self._lib.brt_context_set_option(self.handle, int(option), int(value)))

    def step_shape(self, prox) -> tuple[bool, bool]:
        """(fused, lds_lists) that brt_step uses for these proximity parameters."""
        f, l = C.c_int(), C.c_int()
        self.check(self._lib.brt_context_step_shape(self.handle, C.byref(prox), C.byref(f), C.byref(l)))
        return bool(f.value), bool(l.value)

    def get_option(self, option: int) -> int:
        v = C.c_int()
        self.check(self._lib.brt_context_get_option(self.handle, int(option), C.byref(v)))
        return v.value

    # ------------------------------------------------------------ timing
    def phase_timing(self, enable: bool):
        self.check(self._lib.brt_phase_timing_enable(self.handle, 1 if enable else 0))

    def phase_times(self) -> tuple[dict, int]:
        ms = (C.c_double * NUM_PHASES)()
        steps = C.c_uint32()
        self.check(self._lib.brt_phase_times(self.handle, ms, C.byref(steps)))
        return {PHASE_NAMES[p]: ms[p] for p in range(NUM_PHASES)}, steps.value


def _alloc_buffers(obj, state: dict, num_cells: int, max_neighbors: int, features, wander, device):
    """The per-agent state (ID order), its sorted copies, the hash, cell ranges and
    neighbour lists of one step, as torch device tensors on obj."""
    n = state["pos"].shape[0]
    obj.n = n
    dev = torch.device(device)
    f32 = dict(dtype=torch.float32, device=dev)
    obj.pos = torch.from_numpy(np.ascontiguousarray(state["pos"], np.float32)).to(dev)
    obj.forward = torch.from_numpy(np.ascontiguousarray(state["forward"], np.float32)).to(dev)
    obj.smooth_accel = torch.from_numpy(np.ascontiguousarray(state["smooth_accel"], np.float32)).to(dev)
    obj.sorted_pos = torch.empty((n, 4), **f32)
    obj.sorted_forward = torch.empty((n, 4), **f32)
    obj.sorted_smooth_accel = torch.empty((n, 4), **f32)
    u32 = dict(dtype=torch.int32, device=dev)  # uint32 bit patterns
    obj.hash = torch.empty((n, 2), **u32)
    obj.cell_start = torch.empty((num_cells,), **u32)
    obj.cell_end = torch.empty((num_cells,), **u32)
    obj.W = list_stride(max_neighbors)
    obj.lists = torch.empty((n, obj.W), **u32)
    obj.features = None
    obj.features_out = None
    if features is not None:
        obj.features = torch.from_numpy(np.ascontiguousarray(features, np.float32)).to(dev)
        obj.features_out = torch.empty_like(obj.features)
    obj.wander = None
    if wander is not None:
        obj.wander = torch.from_numpy(np.ascontiguousarray(wander, np.float32)).to(dev)


class Flock:
    """Device buffers of one flock (the features the reference's plug-ins allocate:
    EnvGrid3DPlugIn.h:157-173, Proximity3DPlugIn.h:167-174, OpenSteerWrapperPlugIn.h:153-178)."""

    def __init__(self, ctx: Context, grid: GridParams, prox: ProxParams, steer: SteerParams, state: dict,
                 features: np.ndarray | None = None, wander: np.ndarray | None = None, device="cuda",
                 sorted_smooth_accel: bool = True):
        """sorted_smooth_accel=False: brt_step leaves sorted_smooth_accel out
        (NULL) and reads each agent's own smoothed acceleration by ID; the
        state, hash and lists are the same bits."""
        self.ctx = ctx
        self.grid = grid
        self.prox = prox
        self.steer = steer
        self.keep_sorted_smooth_accel = bool(sorted_smooth_accel)
        _alloc_buffers(self, state, grid.num_cells, prox.max_neighbors, features, wander, device)

    def buffers(self) -> StepBuffers:
        b = StepBuffers()
        b.n = self.n
        b.pos = _ptr(self.pos)
        b.forward = _ptr(self.forward)
        b.smooth_accel = _ptr(self.smooth_accel)
        b.sorted_pos = _ptr(self.sorted_pos)
        b.sorted_forward = _ptr(self.sorted_forward)
        b.sorted_smooth_accel = _ptr(self.sorted_smooth_accel) if self.keep_sorted_smooth_accel else None
        b.hash = _ptr(self.hash)
        b.cell_start = _ptr(self.cell_start)
        b.cell_end = _ptr(self.cell_end)
        b.lists = _ptr(self.lists)
        b.features = _ptr(self.features)
        b.features_out = _ptr(self.features_out)
        b.wander = _ptr(self.wander)
        return b

    # the reference's per-step functionalities, one C call each
    def reorder_sim_data(self):
        self.ctx.hash_sort(self.grid, self.pos, self.hash)
        self.ctx.reorder_ranges(self.hash, [self.pos, self.forward, self.smooth_accel],
                                [self.sorted_pos, self.sorted_forward, self.sorted_smooth_accel],
                                self.grid.num_cells, self.cell_start, self.cell_end)

    def compute_neighborhoods(self):
        self.ctx.neighbors(self.grid, self.prox, self.sorted_pos, self.hash, self.cell_start, self.cell_end,
                           self.lists)

    def apply_steering(self, dt: float, step: int = 0, force_out=None):
        self.ctx.steer_integrate(self.grid, self.prox, self.steer, dt, step, self.sorted_pos,
                                 self.sorted_forward, self.sorted_smooth_accel, self.hash, self.lists,
                                 self.pos, self.forward, self.smooth_accel, wander=self.wander,
                                 force_out=force_out)

    def update_features(self, rate: float):
        self.ctx.feature_update(self.prox, rate, self.hash, self.lists, self.features, self.features_out)
        self.features, self.features_out = self.features_out, self.features

    def step(self, dt: float, step: int, feature_rate: float = 0.0):
        b = self.buffers()
        self.ctx.check(self.ctx._lib.brt_step(self.ctx.handle, C.byref(self.grid), C.byref(self.prox),
                                              C.byref(self.steer), float(feature_rate), float(dt), int(step),
                                              C.byref(b), current_stream()))
        if self.features is not None and feature_rate != 0.0:
            self.features, self.features_out = self.features_out, self.features

    def save(self, path: str, step: int, flags: int | None = None):
        """Binary snapshot of the flock (brt_snapshot_save): the state by ID, the
        last step's hash and lists, and the features and wander state when the
        flock has them (or the BRT_SNAP_* flags given)."""
        if flags is None:
            flags = SNAP_LISTS | (SNAP_FEATURES if self.features is not None else 0) | \
                (SNAP_WANDER if self.wander is not None else 0)
        b = self.buffers()
        self.ctx.check(self.ctx._lib.brt_snapshot_save(self.ctx.handle, os.fsencode(path), C.byref(self.grid),
                                                       C.byref(self.prox), int(step), int(flags), C.byref(b),
                                                       current_stream()))

    def load(self, path: str, flags: int = 0) -> SnapshotInfo:
        """Restores the state (and the sections in `flags`) from a snapshot of a
        flock of the same size; returns its header (the step it was saved at)."""
        info = SnapshotInfo()
        b = self.buffers()
        self.ctx.check(self.ctx._lib.brt_snapshot_load(self.ctx.handle, os.fsencode(path), int(flags),
                                                       C.byref(self.prox), C.byref(b),
                                                       C.byref(info), current_stream()))
        return info

    def state(self) -> dict:
        self.ctx.check_device()
        out = {"pos": self.pos.cpu().numpy(), "forward": self.forward.cpu().numpy(),
               "smooth_accel": self.smooth_accel.cpu().numpy()}
        if self.features is not None:
            out["features"] = self.features.cpu().numpy()
        if self.wander is not None:
            out["wander"] = self.wander.cpu().numpy()
        return out

    def grid_outputs(self) -> dict:
        self.ctx.check_device()
        return {"hash": self.hash.cpu().numpy().view(np.uint32),
                "cell_start": self.cell_start.cpu().numpy().view(np.uint32),
                "cell_end": self.cell_end.cpu().numpy().view(np.uint32),
                "lists": self.lists.cpu().numpy().view(np.uint32),
                "sorted_pos": self.sorted_pos.cpu().numpy(),
                "sorted_forward": self.sorted_forward.cpu().numpy(),
                "sorted_smooth_accel": self.sorted_smooth_accel.cpu().numpy()}


class Batch(Flock):
    """Independent flocks advanced by one batched step (brt_batch_*): the state
    arrays are the flocks' arrays concatenated in order; flock f owns agent IDs
    and sorted indices [offsets[f], offsets[f] + sizes[f]) and cell keys
    [cell_bases[f], cell_bases[f] + grids[f].num_cells)."""

    def __init__(self, ctx: Context, grids, sizes, weights, prox: ProxParams, steer: SteerParams, state: dict,
                 features: np.ndarray | None = None, wander: np.ndarray | None = None, device="cuda",
                 sorted_smooth_accel: bool = True):
        self.ctx = ctx
        self.prox = prox
        self.steer = steer
        self.grid = None
        self.keep_sorted_smooth_accel = bool(sorted_smooth_accel)
        F = len(grids)
        descs = (FlockDesc * F)()
        for f in range(F):
            descs[f].grid = grids[f]
            descs[f].n = int(sizes[f])
            for k in range(3):
                descs[f].weights[k] = float(weights[f][k])
        h = C.c_void_p()
        ctx.check(ctx._lib.brt_batch_create(ctx.handle, descs, F, C.byref(prox), C.byref(h)))
        self.handle = h
        tn, tc = C.c_uint32(), C.c_uint32()
        offs, cbs = (C.c_uint32 * F)(), (C.c_uint32 * F)()
        ctx.check(ctx._lib.brt_batch_layout(h, C.byref(tn), C.byref(tc), offs, cbs))
        self.total_cells = int(tc.value)
        self.offsets = np.array(offs[:], np.int64)
        self.cell_bases = np.array(cbs[:], np.int64)
        self.sizes = np.asarray(sizes, np.int64)
        if state["pos"].shape[0] != tn.value:
            raise ValueError(f"state holds {state['pos'].shape[0]} agents, the flocks {tn.value}")
        _alloc_buffers(self, state, self.total_cells, prox.max_neighbors, features, wander, device)

    def close(self):
        if getattr(self, "handle", None):
            self.ctx._lib.brt_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def config(self) -> BatchConfig:
        c = BatchConfig()
        self.ctx.check(self.ctx._lib.brt_batch_get_config(self.handle, C.byref(c)))
        return c

    def set_config(self, **kw):
        c = self.config()
        for k, v in kw.items():
            setattr(c, k, int(v))
        self.ctx.check(self.ctx._lib.brt_batch_set_config(self.handle, C.byref(c)))

    def autotune(self, dt: float, steps_per_trial: int = 2, cache_path: str | None = None) -> BatchConfig:
        c = BatchConfig()
        b = self.buffers()
        self.ctx.check(self.ctx._lib.brt_batch_autotune(
            self.handle, C.byref(self.steer), float(dt), C.byref(b), int(steps_per_trial),
            None if cache_path is None else cache_path.encode(), C.byref(c), current_stream()))
        return c

    def step(self, dt: float, step: int, feature_rate: float = 0.0):
        b = self.buffers()
        self.ctx.check(self.ctx._lib.brt_batch_step(self.handle, C.byref(self.steer), float(feature_rate), float(dt),
                                                    int(step), C.byref(b), current_stream()))
        if self.features is not None and feature_rate != 0.0:
            self.features, self.features_out = self.features_out, self.features

    def flock_slice(self, f: int) -> slice:
        return slice(int(self.offsets[f]), int(self.offsets[f] + self.sizes[f]))

    def cell_slice(self, f: int, num_cells: int) -> slice:
        return slice(int(self.cell_bases[f]), int(self.cell_bases[f] + num_cells))
