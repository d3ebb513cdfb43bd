"""CPU oracle for the BehaveRT per-step path (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
import this package, and only as the checker or as the timed CPU baseline; the
product (behavert_amd) never does.

numpy wrappers over oracle/liboracle.so (built from oracle/brt_oracle.cpp by
oracle/build.sh).  See brt_oracle.cpp for the reference file:line each function
restates.  Parity pins (tests/test_oracle_*.py):
  calc_hash ........ closed forms (face points p = 4k, clamping, NaN, mul24)
  sort_radix ....... std::stable_sort, bijection/order invariants
  cell_ranges ...... closed form from sorted keys
  neighbors ........ O(N^2) brute force in canonical order, KATs at d = r
  steer_f32 ........ FP64 evaluation, hand-derived 2/3-agent KATs
  philox ........... Random123 known-answer vectors
  feature_update ... closed form (uniform features are a fixed point)
  bc_similarities .. closed forms (parallel / opposite / orthogonal features),
                     hand-traced statistics updates, interpolation end points
  bc_cluster_steer . two-agent KATs (similarity 1 -> seek, 0 -> flee)
  steer_forces, apply_steering_force
                     bit-equal, together, to the fused steer_f32/f64
  cluster_membership labels = least sorted index within t list hops (breadth-
                     first sets); connected components (scipy) at convergence
  cluster_hit_rate . KAT at distances 1.0 / 1.5 (on the radius) / 1.6
  slow_avoid, neighborhood_color, move_towards_target
                     hand-derived KATs (tests/test_oracle_consumers.py)
  snapshot.py ...... numpy restatement of the snapshot format; FNV-1a 64 test
                     vectors (tests/test_snapshot.py)

oracle/_ref/liblq_ref.so (oracle/build_ref.sh, when the reference checkout is
present) is the reference tree's own OpenSteer locality-query database
(Samples/OpensteerDemo/src/lq.c) with a timing driver (oracle/lq_driver.c):
lq_flock_baseline() times it as the reference-code CPU neighbour search; its
counts are pinned against a brute-force count (tests/test_ref_lq.py).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from behavert_amd.types import BcParams, GridParams, ProxParams, SteerParams, StepBuffers, list_stride

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


REF_LQ_PATH = os.path.join(HERE, "_ref", "liblq_ref.so")


def build():
    subprocess.run(["bash", os.path.join(HERE, "build.sh")], check=True)
    # the reference's own lq.c, only where the reference checkout is present
    subprocess.run(["bash", os.path.join(HERE, "build_ref.sh")], check=True)


def lq_flock_baseline(pos4: np.ndarray, grid, radius: float, threads: int, stride: int = 1) -> dict | None:
    """OpenSteer's locality-query database (the reference tree's lq.c) over a
    flock: database build (one lqUpdateForNewLocation per agent, serial) and
    one sphere query (lqMapOverAllObjectsInLocality, `threads` OpenMP threads)
    for every stride-th agent.  The lattice is the grid's box and cells.
    None when oracle/_ref was not built."""
    if not os.path.exists(REF_LQ_PATH):
        return None
    L = C.CDLL(REF_LQ_PATH)
    f = L.lq_flock_baseline
    f.restype = C.c_int
    F3, I3 = C.c_float * 3, C.c_int * 3
    f.argtypes = [C.c_void_p, C.c_uint32, F3, F3, I3, C.c_float, C.c_uint32, C.c_int, C.POINTER(C.c_double),
                  C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
    pos4 = np.ascontiguousarray(pos4, np.float32)
    gs = [int(x) for x in grid.grid_size]
    origin = F3(*[float(x) for x in grid.world_origin])
    size = F3(*[float(grid.cell_size[a]) * gs[a] for a in range(3)])
    b, q, nb = C.c_double(), C.c_double(), C.c_uint64()
    rc = f(pos4.ctypes.data, pos4.shape[0], origin, size, I3(*gs), float(radius), int(stride), int(threads),
           C.byref(b), C.byref(q), C.byref(nb))
    if rc != 0:
        raise MemoryError("lq_flock_baseline: out of memory")
    queries = (pos4.shape[0] + stride - 1) // stride
    return {"build_s": b.value, "query_s": q.value, "neighbours": int(nb.value), "queries": queries}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        U = C.c_uint32
        F = C.c_float
        GP, PP, SP = C.POINTER(GridParams), C.POINTER(ProxParams), C.POINTER(SteerParams)
        sig = {
            "orc_set_threads": (None, [C.c_int]),
            "orc_num_threads": (C.c_int, []),
            "orc_philox4x32_10": (None, [P, P, P]),
            "orc_grid_setup": (None, [P, P, P, U, GP]),
            "orc_calc_grid_pos": (None, [GP, P, U, P]),
            "orc_calc_hash": (None, [GP, P, U, P]),
            "orc_sort_stable": (None, [P, U]),
            "orc_sort_radix": (None, [P, U, U]),
            "orc_reorder": (None, [P, U, P, P]),
            "orc_cell_ranges": (None, [P, U, U, P, P]),
            "orc_neighbors": (None, [GP, PP, P, P, P, P, U, P]),
            "orc_neighbors_brute": (None, [GP, PP, P, P, U, P]),
            "orc_steer_f32": (None, [GP, PP, SP, F, U, U, P, P, P, P, P, P, P, P, P, P]),
            "orc_steer_f64": (None, [GP, PP, SP, F, U, U, P, P, P, P, P, P, P, P, P, P, P, P, P]),
            "orc_feature_update_f32": (None, [PP, F, U, P, P, P, P]),
            "orc_feature_update_f64": (None, [PP, F, U, P, P, P, P]),
            "orc_step_f32": (None, [GP, PP, SP, F, F, U, C.POINTER(StepBuffers), P]),
            "orc_bc_similarities_f32": (None, [PP, C.POINTER(BcParams), U, P, P, P, P, P, P]),
            "orc_bc_similarities_f64": (None, [PP, C.POINTER(BcParams), U, P, P, P, P, P, P]),
            "orc_bc_cluster_steer_f32": (None, [PP, SP, C.POINTER(BcParams), U, P, P, P, P, P, P]),
            "orc_bc_cluster_steer_f64": (None, [PP, SP, C.POINTER(BcParams), U, P, P, P, P, P, P]),
            "orc_cluster_membership": (None, [PP, U, U, P, P, P, P, P, P]),
            "orc_slow_avoid": (None, [PP, SP, U, P, P, P, P, P, P]),
            "orc_neighborhood_color": (None, [PP, P, U, U, P, P, P]),
            "orc_move_towards_target": (None, [C.c_int, P, U, P, P, P, P, P]),
            "orc_cluster_hit_rate": (None, [PP, F, U, P, P, P, P, P]),
            "orc_steer_forces_f32": (None, [PP, SP, F, U, U, P, P, P, P, P, P]),
            "orc_steer_forces_f64": (None, [PP, SP, F, U, U, P, P, P, P, P, P]),
            "orc_apply_steering_force_f32": (None, [GP, SP, F, U, P, P, P, P, P, P, P, P, P]),
            "orc_apply_steering_force_f64": (None, [GP, SP, F, U, P, P, P, P, P, P, P, P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def num_threads() -> int:
    return lib().orc_num_threads()


def philox(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def grid_setup(center, radius, cell, max_bodies=0xFFFFFFFF) -> GridParams:
    g = GridParams()
    c = np.ascontiguousarray(center, np.float32)
    r = np.ascontiguousarray(radius, np.float32)
    s = np.ascontiguousarray(cell, np.float32)
    lib().orc_grid_setup(_p(c), _p(r), _p(s), int(max_bodies), C.byref(g))
    return g


def calc_grid_pos(grid, pos) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float32)
    out = np.zeros((len(pos), 3), np.int32)
    lib().orc_calc_grid_pos(C.byref(grid), _p(pos), len(pos), _p(out))
    return out


def calc_hash(grid, pos) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.float32)
    h = np.zeros((len(pos), 2), np.uint32)
    lib().orc_calc_hash(C.byref(grid), _p(pos), len(pos), _p(h))
    return h


def sort_stable(h) -> np.ndarray:
    h = np.array(h, np.uint32, copy=True, order="C")
    lib().orc_sort_stable(_p(h), len(h))
    return h


def sort_radix(h, num_cells) -> np.ndarray:
    h = np.array(h, np.uint32, copy=True, order="C")
    lib().orc_sort_radix(_p(h), len(h), int(num_cells))
    return h


def reorder(h, arr) -> np.ndarray:
    arr = np.ascontiguousarray(arr, np.float32)
    out = np.zeros_like(arr)
    lib().orc_reorder(_p(np.ascontiguousarray(h, np.uint32)), len(h), _p(arr), _p(out))
    return out


def cell_ranges(h, num_cells):
    s = np.zeros(num_cells, np.uint32)
    e = np.zeros(num_cells, np.uint32)
    lib().orc_cell_ranges(_p(np.ascontiguousarray(h, np.uint32)), len(h), int(num_cells), _p(s), _p(e))
    return s, e


def neighbors(grid, prox, spos, h, start, end) -> np.ndarray:
    n = len(spos)
    lists = np.zeros((n, list_stride(prox.max_neighbors)), np.uint32)
    lib().orc_neighbors(C.byref(grid), C.byref(prox), _p(np.ascontiguousarray(spos, np.float32)),
                        _p(np.ascontiguousarray(h, np.uint32)), _p(np.ascontiguousarray(start, np.uint32)),
                        _p(np.ascontiguousarray(end, np.uint32)), n, _p(lists))
    return lists


def neighbors_brute(grid, prox, spos, h) -> np.ndarray:
    n = len(spos)
    lists = np.zeros((n, list_stride(prox.max_neighbors)), np.uint32)
    lib().orc_neighbors_brute(C.byref(grid), C.byref(prox), _p(np.ascontiguousarray(spos, np.float32)),
                              _p(np.ascontiguousarray(h, np.uint32)), n, _p(lists))
    return lists


def _steer(fn, dtype, grid, prox, steer, dt, step, spos, sfwd, ssa, h, lists, wander=None, cond=False,
           pred=None):
    n = len(spos)
    spos = np.ascontiguousarray(spos, dtype)
    sfwd = np.ascontiguousarray(sfwd, dtype)
    ssa = np.ascontiguousarray(ssa, dtype)
    pos = np.zeros((n, 4), dtype)
    fwd = np.zeros((n, 4), dtype)
    sa = np.zeros((n, 4), dtype)
    force = np.zeros((n, 4), dtype)
    w = None if wander is None else np.array(wander, dtype, copy=True, order="C")
    args = [C.byref(grid), C.byref(prox), C.byref(steer), float(dt), int(step), n, _p(spos), _p(sfwd), _p(ssa),
            _p(np.ascontiguousarray(h, np.uint32)), _p(np.ascontiguousarray(lists, np.uint32)), _p(w), _p(pos),
            _p(fwd), _p(sa), _p(force)]
    out = {"pos": pos, "forward": fwd, "smooth_accel": sa, "force": force, "wander": w}
    if cond:
        k = np.zeros(n, dtype)
        args.append(_p(k))
        out["cond"] = k
        pp = None if pred is None else np.ascontiguousarray(pred[0], np.float32)
        pf = None if pred is None else np.ascontiguousarray(pred[1], np.float32)
        args += [_p(pp), _p(pf)]
    fn(*args)
    return out


def steer_f32(grid, prox, steer, dt, step, spos, sfwd, ssa, h, lists, wander=None):
    return _steer(lib().orc_steer_f32, np.float32, grid, prox, steer, dt, step, spos, sfwd, ssa, h, lists, wander)


def steer_f64(grid, prox, steer, dt, step, spos, sfwd, ssa, h, lists, wander=None, pred=None):
    """FP64 evaluation; out["cond"][i] = max over the agent's behaviours of
    sum|terms| / |sum of terms| (the condition number of its normalisations).
    pred = (sorted FP32 pos, sorted FP32 forward): take the discrete decisions
    (radius/angle predicates, world box) from that state instead."""
    return _steer(lib().orc_steer_f64, np.float64, grid, prox, steer, dt, step, spos, sfwd, ssa, h, lists, wander,
                  cond=True, pred=pred)


def steer_forces(prox, steer, dt, step, spos, sfwd, h, lists, wander=None, dtype=np.float32):
    """The behaviours' summed steer force (sorted (n, 4)); wander (by ID) is
    advanced in a copy, returned beside the force."""
    n = len(spos)
    spos = np.ascontiguousarray(spos, dtype)
    sfwd = np.ascontiguousarray(sfwd, dtype)
    w = None if wander is None else np.array(wander, dtype, copy=True, order="C")
    force = np.zeros((n, 4), dtype)
    fn = lib().orc_steer_forces_f32 if dtype == np.float32 else lib().orc_steer_forces_f64
    fn(C.byref(prox), C.byref(steer), float(dt), int(step), n, _p(spos), _p(sfwd),
       _p(np.ascontiguousarray(h, np.uint32)), _p(np.ascontiguousarray(lists, np.uint32)), _p(w), _p(force))
    return force, w


def apply_steering_force(grid, steer, dt, spos, sfwd, ssa, h, force_in, dtype=np.float32):
    """World-centre seek + integration of a given steer force: returns
    {pos, forward, smooth_accel (by ID), force (sorted, after the seek)}."""
    n = len(spos)
    arrs = [np.ascontiguousarray(a, dtype) for a in (spos, sfwd, ssa, force_in)]
    out = {k: np.zeros((n, 4), dtype) for k in ("pos", "forward", "smooth_accel", "force")}
    fn = lib().orc_apply_steering_force_f32 if dtype == np.float32 else lib().orc_apply_steering_force_f64
    fn(C.byref(grid), C.byref(steer), float(dt), n, _p(arrs[0]), _p(arrs[1]), _p(arrs[2]),
       _p(np.ascontiguousarray(h, np.uint32)), _p(arrs[3]), _p(out["pos"]), _p(out["forward"]),
       _p(out["smooth_accel"]), _p(out["force"]))
    return out


def feature_update_f32(prox, rate, h, lists, fin):
    fin = np.ascontiguousarray(fin, np.float32)
    out = np.zeros_like(fin)
    lib().orc_feature_update_f32(C.byref(prox), float(rate), len(h), _p(np.ascontiguousarray(h, np.uint32)),
                                 _p(np.ascontiguousarray(lists, np.uint32)), _p(fin), _p(out))
    return out


def feature_update_f64(prox, rate, h, lists, fin):
    fin = np.ascontiguousarray(fin, np.float64)
    out = np.zeros_like(fin)
    lib().orc_feature_update_f64(C.byref(prox), float(rate), len(h), _p(np.ascontiguousarray(h, np.uint32)),
                                 _p(np.ascontiguousarray(lists, np.uint32)), _p(fin), _p(out))
    return out


def bc_similarities(prox, bc, h, lists, features, stats=None, dtype=np.float32):
    """computeNeighNaiveSimilarities (+ computeAdaptiveSimilarities when
    bc.adaptive): returns (similarities (n, K+1) by sorted index, new stats
    (n, 4) by ID or None)."""
    n = len(h)
    K = prox.max_neighbors
    feats = np.ascontiguousarray(features, dtype)
    st_in = np.ascontiguousarray(stats if stats is not None else np.zeros((n, 4)), dtype)
    st_out = st_in.copy()
    sims = np.zeros((n, K + 1), dtype)
    fn = lib().orc_bc_similarities_f32 if dtype == np.float32 else lib().orc_bc_similarities_f64
    fn(C.byref(prox), C.byref(bc), n, _p(np.ascontiguousarray(h, np.uint32)),
       _p(np.ascontiguousarray(lists, np.uint32)), _p(feats), _p(st_in), _p(st_out), _p(sims))
    return sims, (st_out if bc.adaptive else None)


def bc_cluster_steer(prox, steer, bc, spos, sfwd, lists, sims, force_in, dtype=np.float32):
    """steerForClusterCohesion then steerForClusterAlignment blended onto force_in
    (sorted float4): returns the new force (n, 4)."""
    n = len(spos)
    arrs = [np.ascontiguousarray(a, dtype) for a in (spos, sfwd, sims, force_in)]
    out = np.zeros((n, 4), dtype)
    fn = lib().orc_bc_cluster_steer_f32 if dtype == np.float32 else lib().orc_bc_cluster_steer_f64
    fn(C.byref(prox), C.byref(steer), C.byref(bc), n, _p(arrs[0]), _p(arrs[1]),
       _p(np.ascontiguousarray(lists, np.uint32)), _p(arrs[2]), _p(arrs[3]), _p(out))
    return out


def cluster_membership(prox, iterations, h, lists, spos, membership=None, connector=False):
    """computeClusterMembership: (labels float3 (n, 3) by ID, connector
    (n, 2, 4) by ID or None); .y/.z carried from `membership` (zeros)."""
    n = len(h)
    m_in = np.ascontiguousarray(membership if membership is not None else np.zeros((n, 3)), np.float32)
    m_out = np.zeros((n, 3), np.float32)
    conn = np.zeros((n, 2, 4), np.float32) if connector else None
    lib().orc_cluster_membership(C.byref(prox), n, int(iterations), _p(np.ascontiguousarray(h, np.uint32)),
                                 _p(np.ascontiguousarray(lists, np.uint32)),
                                 _p(np.ascontiguousarray(spos, np.float32)), _p(m_in), _p(m_out), _p(conn))
    return m_out, conn


def cluster_hit_rate(prox, radius, h, lists, spos, membership):
    """clusterHitRate: float3 (label, hit rate, mean close label) by ID."""
    n = len(h)
    m_out = np.zeros((n, 3), np.float32)
    lib().orc_cluster_hit_rate(C.byref(prox), float(radius), n, _p(np.ascontiguousarray(h, np.uint32)),
                               _p(np.ascontiguousarray(lists, np.uint32)), _p(np.ascontiguousarray(spos, np.float32)),
                               _p(np.ascontiguousarray(membership, np.float32)), _p(m_out))
    return m_out


def slow_avoid(prox, steer, spos, sfwd, lists, force_in):
    """steerToAvoidNeighbors: (forward_out sorted, force_out sorted)."""
    n = len(spos)
    fwd = np.zeros((n, 4), np.float32)
    force = np.zeros((n, 4), np.float32)
    lib().orc_slow_avoid(C.byref(prox), C.byref(steer), n, _p(np.ascontiguousarray(spos, np.float32)),
                         _p(np.ascontiguousarray(sfwd, np.float32)), _p(np.ascontiguousarray(lists, np.uint32)),
                         _p(fwd), _p(np.ascontiguousarray(force_in, np.float32)), _p(force))
    return fwd, force


def neighborhood_color(prox, base, vertexes, h, lists):
    """extractColorFromNeighborhoodD: (n, vertexes, 4) colours by ID."""
    n = len(h)
    out = np.zeros((n, vertexes, 4), np.float32)
    lib().orc_neighborhood_color(C.byref(prox), _p(np.asarray(base, np.float32)), int(vertexes), n,
                                 _p(np.ascontiguousarray(h, np.uint32)), _p(np.ascontiguousarray(lists, np.uint32)),
                                 _p(out))
    return out


def move_towards_target(shift, base, h, spos, targets, force_in):
    """moveTowardsTargetD: the new force (n, 4) sorted."""
    n = len(h)
    out = np.zeros((n, 4), np.float32)
    lib().orc_move_towards_target(int(shift), _p(np.asarray(base, np.float32)), n,
                                  _p(np.ascontiguousarray(h, np.uint32)), _p(np.ascontiguousarray(spos, np.float32)),
                                  _p(np.ascontiguousarray(targets, np.float32)),
                                  _p(np.ascontiguousarray(force_in, np.float32)), _p(out))
    return out


class HostFlock:
    """Host buffers for orc_step_f32 (same layout as the device flock)."""

    def __init__(self, grid, prox, state, features=None, wander=None):
        n = len(state["pos"])
        self.n = n
        self.grid, self.prox = grid, prox
        self.pos = np.array(state["pos"], np.float32, copy=True, order="C")
        self.forward = np.array(state["forward"], np.float32, copy=True, order="C")
        self.smooth_accel = np.array(state["smooth_accel"], np.float32, copy=True, order="C")
        self.sorted_pos = np.zeros((n, 4), np.float32)
        self.sorted_forward = np.zeros((n, 4), np.float32)
        self.sorted_smooth_accel = np.zeros((n, 4), np.float32)
        self.hash = np.zeros((n, 2), np.uint32)
        self.cell_start = np.zeros(grid.num_cells, np.uint32)
        self.cell_end = np.zeros(grid.num_cells, np.uint32)
        self.lists = np.zeros((n, list_stride(prox.max_neighbors)), np.uint32)
        self.features = None if features is None else np.array(features, np.float32, copy=True, order="C")
        self.features_out = None if features is None else np.zeros_like(self.features)
        self.wander = None if wander is None else np.array(wander, np.float32, copy=True, order="C")
        self.phase_s = np.zeros(5, np.float64)

    def step(self, steer, dt, step, feature_rate=0.0):
        b = StepBuffers()
        b.n = self.n
        for name in ("pos", "forward", "smooth_accel", "sorted_pos", "sorted_forward", "sorted_smooth_accel",
                     "hash", "cell_start", "cell_end", "lists", "features", "features_out", "wander"):
            setattr(b, name, _p(getattr(self, name)))
        lib().orc_step_f32(C.byref(self.grid), C.byref(self.prox), C.byref(steer), float(feature_rate),
                           float(dt), int(step), C.byref(b), _p(self.phase_s))
        if self.features is not None and feature_rate != 0.0:
            self.features, self.features_out = self.features_out, self.features

    def state(self):
        out = {"pos": self.pos.copy(), "forward": self.forward.copy(), "smooth_accel": self.smooth_accel.copy()}
        if self.features is not None:
            out["features"] = self.features.copy()
        if self.wander is not None:
            out["wander"] = self.wander.copy()
        return out


def step_f64_fixed(grid, prox, steer, dt, step, state64, h, lists, wander64=None, pred=None):
    """One FP64 step from an FP64 state (agent-ID order) with the sort and the
    neighbour lists fixed to (h, lists) (e.g. those the FP32 step produced);
    pred = (sorted FP32 pos, sorted FP32 forward) also fixes the decisions."""
    ids = h[:, 1].astype(np.int64)
    spos = state64["pos"][ids]
    sfwd = state64["forward"][ids]
    ssa = state64["smooth_accel"][ids]
    out = steer_f64(grid, prox, steer, dt, step, spos, sfwd, ssa, h, lists, wander64, pred=pred)
    return out


def rollout_f64_fixed(grid, prox, steer, dt, records, state64, first_step=0):
    """FP64 rollout along recorded FP32 steps: records[t] = (hash, lists,
    sorted FP32 pos, sorted FP32 forward) of step t; sort, lists and discrete
    decisions are taken from the record, the arithmetic is FP64."""
    s = {k: np.array(state64[k], np.float64, copy=True) for k in ("pos", "forward", "smooth_accel")}
    for t, (h, lists, spos32, sfwd32) in enumerate(records):
        out = step_f64_fixed(grid, prox, steer, dt, first_step + t, s, h, lists, pred=(spos32, sfwd32))
        s = {k: out[k] for k in ("pos", "forward", "smooth_accel")}
    return s
