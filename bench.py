#!/usr/bin/env python3
"""Benchmark of the per-step simulation path on one MI355X.

Workload (default): the 1,048,576-boid Gaussian-blob flock (64 blobs, sigma 10, in
[0,400)^3, grid 100^3 of cell 4, radius 4, K 32, weights 1.0/0.8/0.6, dt 0.05).
A step = hash + sort + reorder + cell ranges + neighbour lists + fused steering
and integration (brt_step), inputs resident in HBM.

Prints ONE JSON line: agent-updates/s over the timed steps, per-phase times and
HBM roofline fractions (hipEvents recorded on the step's stream during a second
pass of the same number of steps), and the CPU oracle timed on the host cores on
a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes_per_agent(w, packed=False, sorted_sa=False) -> dict:
    """Compulsory single-touch HBM bytes per agent for each phase of brt_step.
    D digit passes, W list words, C/N cells per agent, F reordered float4s
    (position and heading, + the smoothed acceleration when its sorted copy
    is kept; without it the steering reads the agent's own by ID, the same 16
    bytes); packed: the reorder also writes the 12-byte x,y,z copy the
    LDS-list scan reads candidates from."""
    from behavert_amd.types import list_stride

    n = w.n
    C = w.grid.num_cells
    bits = max(1, int(C - 1).bit_length())
    D = (bits + 7) // 8
    W = list_stride(w.prox.max_neighbors)
    cpa = C / n
    b = {
        # prepare (empty-cell fill 8 B/cell) + hash (R pos 16, W pair 8) + sort (R 8 hist, D x (R 8 + W 8))
        "hash_sort": 24 + (8 + 16 * D) + 8 * cpa,
        # R pair 8 (ID + key), 3 x (gather 16 + write 16) [+ W packed x,y,z 12]
        "reorder_ranges": 8 + 32 * (3 if sorted_sa else 2) + (12 if packed else 0),
        # R own pos 16, W list 4W, R cell ranges 8 B/cell
        "neighbors": 16 + 4 * W + 8 * cpa,
        # R pos/fwd/sa 48 + list 4W + ID 4, W pos/fwd/sa 48 (neighbour gathers cache-resident)
        "steer_integrate": 100 + 4 * W,
        "features": 0.0,
    }
    if w.features is not None and w.feature_rate:
        b["features"] = 4 + 4 * W + 32  # R ID + list, R own feature 16, W 16
        # the two-pass step gathers the features into sorted order with the
        # state (R 16 by ID, W 16) and updates them in the steering pass (the
        # list is read once for both; R own feature 16, W 16 by ID)
        b["reorder_ranges_features"] = b["reorder_ranges"] + 32
        b["steer_features"] = b["steer_integrate"] + 32
    # fused neighbour lists + steering: the list is written but never re-read
    b["neighbors_steer"] = 16 + 4 * W + 8 * cpa + 36 + 48
    return b


PHASE_KERNELS = {"neighbors": "k_prox<false,false,*>", "steer_integrate": "k_steer",
                 "steer_features": "k_steer<true,true> (steering + feature update)",
                 "hash_sort": "k_prepare+k_hash_hist+k_onesweep", "reorder_ranges": "k_reorder_ranges",
                 "features": "k_features", "neighbors_steer": "k_prox<true,false,*>"}


def committed_metric(name: str, phase: str, workload: str):
    """A figure of the phase's kernels on the workload from the committed
    rocprofv3 summaries (profiles/*_<workload>.md, a line `name[phase]: N`),
    or None when there is none."""
    import glob

    key = f"{name}[{phase}]:"
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", f"*_{workload}.md"))):
        with open(path) as f:
            for line in f:
                if line.startswith(key):
                    return float(line.split(":", 1)[1].split()[0])
    return None


def committed_traffic(phase: str, workload: str):
    """HBM bytes per launch of the phase's kernels (FETCH_SIZE and WRITE_SIZE
    scaled by the copy-kernel calibration), or None."""
    return committed_metric("traffic_bytes_per_launch", phase, workload)


def committed_valu(phase: str, workload: str):
    """The VALU roofline of the phase's kernel from the committed SQ counters:
    VALU instructions per wave, issue fraction of the SIMDs' cycles, active
    lanes per instruction, lane-ops against 256 CUs x 64 lanes x 2.4 GHz."""
    keys = ("valu_insts_per_wave", "valu_issue_frac", "valu_active_lane_frac", "valu_lane_ops_frac_of_peak")
    v = {k[5:]: committed_metric(k, phase, workload) for k in keys}
    return v if all(x is not None for x in v.values()) else None


def with_traffic(entry: dict, phase: str, workload: str, n: int) -> dict:
    """Adds the committed counter traffic of the phase and its ratio to the
    algorithmic bytes."""
    t = committed_traffic(phase, workload)
    entry["traffic"] = t
    if t is not None and entry.get("bytes_per_agent"):
        entry["traffic_ratio"] = round(t / (entry["bytes_per_agent"] * n), 3)
    return entry


def phase_table(ctx, w, abytes, fused):
    """Average per-phase times of the steps timed since phase timing was
    enabled (HIP events recorded on the step's stream between the phases)."""
    ms, nsteps = ctx.phase_times()
    out = {}
    merged = not fused and "steer_features" in abytes
    for k, v in ms.items():
        name = "neighbors_steer" if (fused and k == "neighbors") else k
        if fused and k == "steer_integrate":
            continue
        if k == "features" and abytes["features"] == 0:
            continue  # no feature update ran: the phase is only the gap between two events
        if merged and k == "features":
            continue  # computed in the steering pass
        if merged and k == "steer_integrate":
            name = "steer_features"
        bkey = "reorder_ranges_features" if (merged and k == "reorder_ranges") else name
        avg = v / max(nsteps, 1)
        gbs = (abytes[bkey] * w.n / 1e9) / (avg / 1e3) if avg > 0 and abytes[bkey] > 0 else 0.0
        out[name] = with_traffic({"ms": round(avg, 5), "bytes_per_agent": round(abytes[bkey], 2),
                                  "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                                  "kernel": PHASE_KERNELS[name]}, name, w.name, w.n)
    return out


def cpu_model() -> str:
    """The host CPU's model name (lscpu's "Model name", from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, sample_steps: int) -> dict:
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    oracle.set_threads(threads)
    hf = oracle.HostFlock(w.grid, w.prox, w.state, features=w.features, wander=w.wander)
    hf.step(w.steer, w.dt, 0, w.feature_rate)  # first touch of the buffers, untimed
    t0 = time.perf_counter()
    for i in range(sample_steps):
        hf.step(w.steer, w.dt, 1 + i, w.feature_rate)
    dt = time.perf_counter() - t0
    out = {"value": w.n * sample_steps / dt, "unit": "agent-updates/s", "cores": oracle.num_threads(),
           "cpu_model": cpu_model(), "kind": "port",
           "sample": f"{sample_steps} steps of the {w.name} workload ({w.n} agents) after 1 untimed step, "
                     f"FP32 C++ oracle (hash, LSD radix sort, gather, ranges, 27-cell lists, steering), OpenMP",
           "ms_per_step": 1e3 * dt / sample_steps}
    # the reference tree's own CPU proximity code (OpenSteer's lq.c bin
    # lattice: uncapped sphere queries, no canonical order) on the same
    # positions: a reference-code figure for the neighbour search alone
    stride = 16
    lq = oracle.lq_flock_baseline(w.state["pos"], w.grid, w.prox.search_radius, threads, stride)
    if lq is not None:
        out["reference_neighbour_search"] = {
            "value": lq["queries"] / lq["query_s"], "unit": "neighbourhood queries/s", "cores": threads,
            "kind": "reference",
            "code": "Samples/OpensteerDemo/src/lq.c (OpenSteer locality-query bin lattice, built into oracle/_ref)",
            "build_ms": round(1e3 * lq["build_s"], 3), "query_ms": round(1e3 * lq["query_s"], 3),
            "neighbours_per_query": round(lq["neighbours"] / lq["queries"], 2),
            "sample": f"all {w.n} agents of the {w.name} initial state in the lattice (the grid's box and cells), "
                      f"one radius-{w.prox.search_radius} query for every {stride}th ({lq['queries']} queries)"}
    return out


def adjoint_ratio(steps=None, every=None) -> dict:
    """Differentiable rollout (65,536 boids in [0,160)^3, K 32): time of the plain
    forward step vs a forward step + its adjoint step, and of the whole
    checkpointed rollout (forward with checkpoints, then a backward that re-runs
    each segment and walks the adjoint) vs the plain forward rollout."""
    import torch

    from behavert_amd import workloads as W
    from behavert_amd.autograd import Rollout
    from behavert_amd.engine import Context

    w = W.differentiable_65k()
    T = steps or w.steps
    K = every or w.extra["checkpoint_every"]
    ctx = Context(0)
    ro = Rollout(ctx, w.grid, w.prox, w.steer, w.n, w.dt, T, K)
    dev = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    p0, f0, a0 = dev(w.state["pos"]), dev(w.state["forward"]), dev(w.state["smooth_accel"])
    wts = list(w.extra["start_weights"])

    def ev():
        return torch.cuda.Event(enable_timing=True)

    # plain forward rollout (the same steps, no checkpoints kept beyond the few copies)
    ro.forward(wts, p0, f0, a0)
    torch.cuda.synchronize()
    e0, e1, e2 = ev(), ev(), ev()
    e0.record()
    ro.forward(wts, p0, f0, a0)
    e1.record()
    g = torch.zeros_like(p0)
    g[:, :3] = 1.0 / w.n
    z = torch.zeros_like(p0)
    ro.backward(g, z, z)
    e2.record()
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1)
    bwd_ms = e1.elapsed_time(e2)
    # one step: forward step vs adjoint kernel alone, over the last recorded segment
    import ctypes as C

    from behavert_amd import lib

    L = lib()
    e3, e4 = ev(), ev()
    reps = 20
    e3.record()
    for _ in range(reps):
        go, gi = ro.g[0], ro.g[1]
        ctx.check(L.brt_osw_adjoint_step(ctx.handle, C.byref(ro.grid), C.byref(ro.prox), C.byref(ro.steer), ro.dt,
                                         ro.n, ro.r_spos[0, 0].data_ptr(), ro.r_sfwd[0, 0].data_ptr(),
                                         ro.r_ssa[0, 0].data_ptr(), ro.r_hash[0, 0].data_ptr(),
                                         ro.r_lists[0, 0].data_ptr(), ro.r_cs[0, 0].data_ptr(),
                                         ro.r_ce[0, 0].data_ptr(), go[0].data_ptr(),
                                         go[1].data_ptr(), go[2].data_ptr(), gi[0].data_ptr(), gi[1].data_ptr(),
                                         gi[2].data_ptr(), ro.wgrad.data_ptr(),
                                         torch.cuda.current_stream().cuda_stream))
    e4.record()
    torch.cuda.synchronize()
    adj_ms = e3.elapsed_time(e4) / reps
    step_ms = fwd_ms / T
    ctx.close()
    # the adjoint step's compulsory bytes per agent (K 32, W 36): k_adjoint_own
    # reads the hash 8, the sorted state 48, the list 4W and the cotangents by
    # ID 48, and writes its own terms 48 and neighbour-term vectors 48;
    # k_adjoint_gather reads the hash 8, the own terms 48 and the list 4W, and
    # writes the gradients by ID 48 (neighbours' data is cache-resident)
    from behavert_amd.types import list_stride

    Wl = list_stride(w.prox.max_neighbors)
    adj_bpa = (200 + 4 * Wl) + (104 + 4 * Wl)
    adj_gbs = adj_bpa * w.n / 1e9 / (adj_ms / 1e3)
    return {"workload": "differentiable_65k", "n": w.n, "steps": T, "checkpoint_every": K,
            "replay_overlaps_adjoint": ro.overlap,
            "forward_step_ms": round(step_ms, 5), "adjoint_step_ms": round(adj_ms, 5),
            "step_ratio": round((step_ms + adj_ms) / step_ms, 3),
            "forward_rollout_ms": round(fwd_ms, 3), "backward_rollout_ms": round(bwd_ms, 3),
            "rollout_ratio": round((fwd_ms + bwd_ms) / fwd_ms, 3),
            "roofline": {"bound": "hbm", "scope": "adjoint step (k_adjoint_own + k_adjoint_gather)",
                         "achieved": round(adj_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(adj_gbs / HBM_PEAK_GBS, 4), "bytes_per_agent": adj_bpa,
                         "traffic": committed_traffic("adjoint", "differentiable_65k")}}


def full_stack(steps=20, warmup=3) -> dict:
    """4,194,304 agents uniform in [0,640)^3, K 64: separation/alignment/cohesion +
    seek toward 8 moving targets + wander + 4-D feature averaging (rate 0.1)."""
    import torch

    from behavert_amd import workloads as W
    from behavert_amd.engine import Context, Flock

    w = W.full_stack_4m()
    ctx = Context(0)
    fl = Flock(ctx, w.grid, w.prox, w.steer, w.state, features=w.features, wander=w.wander,
               sorted_smooth_accel=False)

    def run(first, count):
        for i in range(first, first + count):
            fl.steer.set_targets(w.targets_at(i))
            fl.step(w.dt, i, w.feature_rate)

    run(0, warmup)
    torch.cuda.synchronize()
    ctx.phase_timing(True)
    t0 = time.perf_counter()
    run(warmup, steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    fused, lds = ctx.step_shape(w.prox)
    abytes = algorithmic_bytes_per_agent(w, packed=bool(lds) and not w.prox.use_knn, sorted_sa=False)
    ph = phase_table(ctx, w, abytes, fused=fused)
    ctx.phase_timing(False)
    ctx.close()
    dom = max((k for k in ph if isinstance(ph[k], dict) and ph[k]["bytes_per_agent"] > 0), key=lambda k: ph[k]["ms"])
    roofline = {"bound": "hbm", "kernel": ph[dom]["kernel"], "phase": dom, "achieved": ph[dom]["GB/s"],
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ph[dom]["frac"],
                "traffic": committed_traffic(dom, w.name), "bytes_per_agent": ph[dom]["bytes_per_agent"],
                "agents_per_launch": w.n, "avg_launch_ms": ph[dom]["ms"]}
    return {**w.describe(), "sorted_smooth_accel": False, "ms_per_step": round(1e3 * el / steps, 4),
            "agent_updates_per_s": w.n * steps / el,
            "phases": ph, "roofline": roofline, "timed_steps": steps, "step_shape": {"fused": fused, "lds_lists": lds}}


def clustering(reps=20, features=8, lp_iterations=10) -> dict:
    """BehaviorClustering on the 1,048,576-boid blob flock (K 32, 8 features by
    agent ID): the passes of steerForClustering in BehaveRT_Clustering's order
    (behaviour forces, similarities with adaptive statistics, cluster cohesion +
    alignment, world-centre seek + integration), the local propagation
    (10 iterations + connector) and the hit rate, each repeated `reps` times
    between HIP events on the launch stream, on the lists and sorted state of
    one brt_step."""
    import numpy as np
    import torch

    from behavert_amd import workloads as W
    from behavert_amd.engine import Context, Flock
    from behavert_amd.types import BcParams, SteerParams, list_stride

    w = W.blobs_1m()
    n, K = w.n, w.prox.max_neighbors
    Wd = list_stride(K)
    ctx = Context(0)
    fl = Flock(ctx, w.grid, w.prox, w.steer, w.state)
    fl.step(w.dt, 0)
    torch.cuda.synchronize()
    rng = np.random.default_rng(0)
    feats = torch.from_numpy(rng.standard_normal((n, features)).astype(np.float32)).cuda()
    stats = torch.tensor([1.0, 0.5, 0.0, 0.0], device="cuda").repeat(n, 1).contiguous()
    stats2 = torch.empty_like(stats)
    sims = torch.empty((n, K + 1), dtype=torch.float32, device="cuda")
    f0, f1, fo = (torch.empty((n, 4), dtype=torch.float32, device="cuda") for _ in range(3))
    outs = [torch.empty((n, 4), dtype=torch.float32, device="cuda") for _ in range(3)]
    memb = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    memb2 = torch.empty_like(memb)
    conn = torch.empty((n, 2, 4), dtype=torch.float32, device="cuda")
    sp = SteerParams.make(weights=(1.0, 0.8, 0.6), seek_world_center=True)
    bc = BcParams.make(features, adaptive=True, cluster_weight=0.9, alignment_weight=0.7)
    sargs = (fl.sorted_pos, fl.sorted_forward)
    passes = {
        "steer_forces": (lambda: ctx.steer_forces(w.prox, sp, w.dt, 0, *sargs, fl.hash, fl.lists, f0),
                         8 + 32 + 4 * Wd + 16, "k_steer<false>"),
        "similarities": (lambda: ctx.bc_similarities(w.prox, bc, fl.hash, fl.lists, feats, stats, stats2, sims),
                         8 + 4 * Wd + 4 * features + 4 * (K + 1) + 32, "k_bc_sort_features+k_bc_similarities<true>"),
        "cluster_steer": (lambda: ctx.bc_steer_for_clustering(w.prox, sp, bc, *sargs, fl.lists, sims, f0, f1),
                          32 + 4 * Wd + 4 * (K + 1) + 32, "k_bc_cluster_steer"),
        "similarities_steer": (lambda: ctx.bc_similarities_steer(w.prox, sp, bc, fl.hash, fl.lists, feats, stats,
                                                                 stats2, sims, *sargs, f0, f1),
                               8 + 4 * Wd + 4 * features + 4 * (K + 1) + 32 + 32 + 32,
                               "k_bc_sort_features+k_bc_similarities<true,true>"),
        "apply": (lambda: ctx.apply_steering_force(w.grid, sp, w.dt, *sargs, fl.sorted_smooth_accel, fl.hash, f1,
                                                   *outs, force_out=fo), 8 + 48 + 16 + 48 + 16, "k_apply"),
        # T iterations: list + own label per iteration, then by-ID float3 in/out,
        # own position and the connector pair
        "membership": (lambda: ctx.cluster_membership(w.prox, lp_iterations, fl.hash, fl.lists, fl.sorted_pos, memb,
                                                      memb2, conn),
                       lp_iterations * (4 * Wd + 8) + 8 + 24 + 16 + 32, f"k_lp_iterate x{lp_iterations}"),
        "hit_rate": (lambda: ctx.cluster_hit_rate(w.prox, 1.5, fl.hash, fl.lists, fl.sorted_pos, memb2, memb),
                     8 + 4 * Wd + 16 + 12 + 12, "k_hit_rate"),
    }
    res = {}
    for name, (fn, bpa, kern) in passes.items():
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        gbs = bpa * n / 1e9 / (ms / 1e3)
        res[name] = {"ms": round(ms, 5), "bytes_per_agent": bpa, "GB/s": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "kernel": kern}
    res["steer_for_clustering_ms"] = round(sum(res[k]["ms"] for k in ("steer_forces", "similarities_steer",
                                                                        "apply")), 5)
    ctx.close()
    return {"workload": w.name, "n": n, "K": K, "features": features, "lp_iterations": lp_iterations,
            "reps": reps, "passes": res}


def consumers(reps=20) -> dict:
    """The small list consumers on the 1,048,576-boid blob flock (K 32), on the
    lists and sorted state of one brt_step: steerAndSlowToAvoidNeighbors, the
    neighbourhood colouring and the shape-target attraction, each repeated
    `reps` times between HIP events on the launch stream."""
    import numpy as np
    import torch

    from behavert_amd import workloads as W
    from behavert_amd.engine import Context, Flock
    from behavert_amd.types import SteerParams, list_stride

    w = W.blobs_1m()
    n, K = w.n, w.prox.max_neighbors
    Wd = list_stride(K)
    ctx = Context(0)
    fl = Flock(ctx, w.grid, w.prox, w.steer, w.state)
    fl.step(w.dt, 0)
    torch.cuda.synchronize()
    rng = np.random.default_rng(0)
    targets = torch.from_numpy(np.c_[rng.uniform(0, 400, (n, 3)), np.ones(n)].astype(np.float32)).cuda()
    f0 = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    f1, fwd = torch.empty_like(f0), torch.empty_like(f0)
    color = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    sp = SteerParams.make()
    passes = {
        "slow_avoid": (lambda: ctx.steer_and_slow_to_avoid_neighbors(w.prox, sp, fl.sorted_pos, fl.sorted_forward,
                                                                     fl.lists, fwd, f0, f1),
                       32 + 4 * Wd + 16 + 32, "k_slow_avoid"),
        "neighborhood_color": (lambda: ctx.neighborhood_color(w.prox, (0.1, 0.2, 0.3), 1, fl.hash, fl.lists, color),
                               8 + 4 + 16, "k_neighborhood_color"),
        "move_towards_target": (lambda: ctx.move_towards_target(1, (0.0, 0.0, 0.0), fl.hash, fl.sorted_pos, targets,
                                                                f0, f1),
                                8 + 16 + 16 + 16 + 16, "k_move_towards_target"),
    }
    res = {}
    for name, (fn, bpa, kern) in passes.items():
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        gbs = bpa * n / 1e9 / (ms / 1e3)
        res[name] = {"ms": round(ms, 5), "bytes_per_agent": bpa, "GB/s": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "kernel": kern}
    ctx.close()
    return {"workload": w.name, "n": n, "K": K, "reps": reps, "passes": res}


def batched(steps=20, loop_steps=3, config: str | None = None) -> dict:
    """256 flocks (sizes log-uniform in [1024, 65536], 0.5 agents per unit^3, K 32,
    per-flock weights): the autotuned, graph-replayed batched step vs brt_step
    looped over the same flocks."""
    import ctypes as C

    import torch

    from behavert_amd import workloads as W
    from behavert_amd.engine import Batch, Context, Flock
    from behavert_amd.types import SteerParams

    w = W.batched_flocks()
    ctx = Context(0)
    b = Batch(ctx, w.grids, w.sizes, w.weights, w.prox, w.steer, w.concat_state(), sorted_smooth_accel=False)
    t0 = time.perf_counter()
    if config:
        # a fixed launch configuration "key=value,..." instead of the tuner
        # (counter passes see the same kernels as the traced run)
        b.set_config(**{k: int(v) for k, v in (kv.split("=") for kv in config.split(","))})
        cfg = b.config()
    else:
        cfg = b.autotune(w.dt, 2)
    tune_s = time.perf_counter() - t0

    def timed(fn, k):
        torch.cuda.synchronize()
        t = time.perf_counter()
        for i in range(k):
            fn(i)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) * 1e3 / k

    timed(lambda i: b.step(w.dt, i), 3)
    batch_ms = timed(lambda i: b.step(w.dt, 3 + i), steps)
    flocks = []
    for f in range(len(w.sizes)):
        sp = SteerParams()
        C.memmove(C.byref(sp), C.byref(w.steer), C.sizeof(SteerParams))
        sp.separation[0], sp.alignment[0], sp.cohesion[0] = (float(x) for x in w.weights[f])
        flocks.append(Flock(ctx, w.grids[f], w.prox, sp, w.states[f], sorted_smooth_accel=False))

    def loop(i):
        for fl in flocks:
            fl.step(w.dt, i)

    timed(loop, 1)
    loop_ms = timed(loop, loop_steps)
    b.close()
    ctx.close()
    # the batch's compulsory bytes per agent (the plain step's, with the
    # batch's own cells per agent and sort passes)
    from behavert_amd.types import list_stride

    cells = sum(int(g.num_cells) for g in w.grids)
    D = (max(1, int(cells - 1).bit_length()) + 7) // 8
    Wl, cpa = list_stride(w.prox.max_neighbors), cells / w.n
    packed = 12 if cfg.lds_lists else 0
    bpa = 24 + (8 + 16 * D) + 8 * cpa + (8 + 32 * 2 + packed) + (16 + 4 * Wl + 8 * cpa + 36 + 48)
    gbs = bpa * w.n / 1e9 / (batch_ms / 1e3)
    return {**w.describe(), "sorted_smooth_accel": False, "tuned": cfg.as_dict(), "tune_s": round(tune_s, 3),
            "roofline": {"bound": "hbm", "scope": "batched step", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_agent": round(bpa, 2),
                         "traffic": committed_traffic("batched", "batched_256")},
            "batched_ms_per_step": round(batch_ms, 4), "batched_agent_updates_per_s": w.n / (batch_ms / 1e3),
            "looped_ms_per_step": round(loop_ms, 4), "looped_agent_updates_per_s": w.n / (loop_ms / 1e3),
            "speedup_per_agent": round(loop_ms / batch_ms, 3), "timed_steps": steps, "looped_timed_steps": loop_steps}


def copy_calibration():
    """Three 256 MiB device-to-device copies: the kernel of known traffic the
    FETCH_SIZE / WRITE_SIZE counters are scaled by."""
    import torch

    x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    del x, y


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="blobs_1m")
    ap.add_argument("--cpu-steps", type=int, default=10, help="CPU baseline sample steps (0 disables)")
    ap.add_argument("--no-phases", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the adjoint and batched measurements")
    ap.add_argument("--only", choices=("adjoint", "batched", "full_stack", "clustering", "consumers"),
                    help="run one extra section alone and print its JSON (for profiling)")
    ap.add_argument("--batch-config", default=None,
                    help="fixed batched-step configuration, e.g. prox_block=256,uniform_max_cells=1 (skips the tuner)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with three 256 MiB device copies (the reference kernel for counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the step path runs on exactly one GPU: it has no multi-GPU form",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.only:
        print(json.dumps({args.only: {"adjoint": adjoint_ratio,
                                      "batched": lambda: batched(config=args.batch_config),
                                      "full_stack": full_stack, "clustering": clustering,
                                      "consumers": consumers}[args.only]()}),
              flush=True)
        if args.copy_calibration:
            copy_calibration()
        return

    import torch

    from behavert_amd import workloads as W
    from behavert_amd.engine import Context, Flock

    w = W.WORKLOADS[args.workload]()
    ctx = Context(0)
    fl = Flock(ctx, w.grid, w.prox, w.steer, w.state, features=w.features, wander=w.wander,
               sorted_smooth_accel=False)

    def run(first, count):
        for i in range(first, first + count):
            if w.moving_targets:
                fl.steer.set_targets(w.targets_at(i))
            fl.step(w.dt, i, w.feature_rate)

    abytes = algorithmic_bytes_per_agent(w, packed=bool(ctx.step_shape(w.prox)[1]) and not w.prox.use_knn,
                                         sorted_sa=False)
    run(0, args.warmup)
    torch.cuda.synchronize()
    # the timed steps: nothing between them but the steps themselves
    t0 = time.perf_counter()
    run(args.warmup, args.steps)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = w.n * args.steps / elapsed

    phases = {}
    roofline = None
    if not args.no_phases:
        # the phases: HIP events recorded between them on the step's stream,
        # over the following steps (the events add a few microseconds a step)
        from behavert_amd.types import OPT_FUSED_STEP

        first = args.warmup + args.steps
        count = min(args.steps, 100)
        ctx.phase_timing(True)
        run(first, count)
        phases["fused_step"] = phase_table(ctx, w, abytes, fused=True)
        phases["fused_step"]["steps"] = [first, first + count]
        ctx.phase_timing(False)
        # the two-kernel path (lists, then steering) on the following steps, for
        # the grid + neighbour phases on their own
        fused_opt = ctx.get_option(OPT_FUSED_STEP)
        ctx.set_option(OPT_FUSED_STEP, 0)
        first += count
        count = min(args.steps, 50)
        ctx.phase_timing(True)
        run(first, count)
        phases["unfused_step"] = phase_table(ctx, w, abytes, fused=False)
        phases["unfused_step"]["steps"] = [first, first + count]
        ctx.phase_timing(False)
        ctx.set_option(OPT_FUSED_STEP, fused_opt)
        fz = phases["fused_step"]
        fz["sum_ms"] = round(sum(v["ms"] for v in fz.values() if isinstance(v, dict)), 5)
        # the phases are kernel spans (events bound to each phase's first and
        # last kernel); their sum against the event-free timed steps
        fz["sum_over_ms_per_step"] = round(fz["sum_ms"] / ms_per_step, 4)
        dom = max((k for k in fz if isinstance(fz[k], dict) and abytes.get(k, 0) > 0), key=lambda k: fz[k]["ms"])
        roofline = {"bound": "hbm", "kernel": fz[dom]["kernel"], "phase": dom, "achieved": fz[dom]["GB/s"],
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": fz[dom]["frac"],
                    "traffic": committed_traffic(dom, w.name), "valu": committed_valu(dom, w.name),
                    "bytes_per_agent": round(abytes[dom], 2), "agents_per_launch": w.n,
                    "avg_launch_ms": fz[dom]["ms"], "timed_steps": fz["steps"]}
        uf = phases["unfused_step"]
        g_keys = ("hash_sort", "reorder_ranges", "neighbors")
        g_ms = sum(uf[k]["ms"] for k in g_keys)
        g_b = sum(abytes[k] for k in g_keys) * w.n
        phases["grid_neighbor_phases"] = {"ms": round(g_ms, 5), "GB/s": round(g_b / 1e9 / (g_ms / 1e3), 1),
                                          "frac": round(g_b / 1e9 / (g_ms / 1e3) / HBM_PEAK_GBS, 4),
                                          "phases": list(g_keys), "from": "unfused_step"}
    total_b = sum(abytes[k] for k in ("hash_sort", "reorder_ranges", "neighbors_steer", "features"))
    step_gbs = total_b * w.n / 1e9 / (ms_per_step / 1e3)

    ctx.close()
    del fl
    if args.copy_calibration:
        copy_calibration()
    extras = {}
    if not args.no_extras:
        extras["adjoint"] = adjoint_ratio()
        extras["batched"] = batched()
        extras["full_stack"] = full_stack()
        extras["clustering"] = clustering()
        extras["consumers"] = consumers()
    cpu = cpu_baseline(w, args.cpu_steps) if args.cpu_steps > 0 else None
    line = {
        "metric": "agent-updates/s at 1M boids (% HBM peak); fwd+adjoint vs plain step time",
        "value": value,
        "unit": "agent-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {**w.describe(), "sorted_smooth_accel": False},
        "hbm_frac_step": round(step_gbs / HBM_PEAK_GBS, 4),
        "roofline": roofline,
        "phases": phases,
        "cpu_baseline": cpu,
        **extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
