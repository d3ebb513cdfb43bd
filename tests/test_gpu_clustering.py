"""GPU parity of the decomposed OpenSteerWrapper steering and of
BehaviorClustering's steerForClustering (naive + adaptive similarities, cluster
cohesion and alignment) through the C ABI: bit-exact against the FP32 oracle,
the decomposed path bit-exact against the fused one, and the clustering step
within 1e-5 of the FP64 evaluation (per-agent floored relative error, as in the
steering tests) on agents where the FP64 reference is well conditioned: cluster
sums with condition number <= 100 and, for adaptive similarities, no branch of
the statistics update taken within 1e-5 of its threshold (the update is
discontinuous there) and interpolation slopes <= 100."""
import numpy as np
import pytest
import torch

from behavert_amd import workloads as W
from behavert_amd.engine import BrtError
from behavert_amd.types import MAX_FEATURES, BcParams, GridParams, ProxParams, SteerParams

from ._parity import dev, empty_f32, f32, first_mismatch, floored_rel, oracle_grid_outputs

pytestmark = pytest.mark.gpu

FLOORS = {"smooth_accel": 0.4, "forward": 1.0, "pos": 1.0}


def mixed_workload(n=30000, K=16, seed=41, slow=True):
    """Uniform agents in a 40^3 box; with slow, a tenth of them move below the
    speed where the integrator's cone limit applies."""
    rng = np.random.default_rng(seed)
    grid = GridParams.direct((0.0,) * 3, (4.0,) * 3, (10,) * 3)
    st = W.make_state(rng.uniform(0, 40, (n, 3)).astype(np.float32), *W.directions_and_speeds(rng, n))
    if slow:
        st["forward"][: n // 10, 3] = rng.uniform(0.0, 0.2, n // 10).astype(np.float32)
    st["smooth_accel"][:, :3] = rng.uniform(-0.1, 0.1, (n, 3)).astype(np.float32)
    prox = ProxParams.make(K, 4.0)
    return W.Workload("mixed", n, 40.0, grid, prox, W.plain_steer(), 0.05, 1, st)


def features_and_stats(n, F, seed):
    rng = np.random.default_rng(seed)
    feats = rng.standard_normal((n, F)).astype(np.float32)
    # a shared bias makes the similarities spread over (0, 1), not only around 1/2
    feats += rng.standard_normal(F).astype(np.float32)
    stats = np.zeros((n, 4), np.float32)
    stats[:, 0], stats[:, 1], stats[:, 2] = 1.0, 0.5, 0.0     # the plug-in's initial statistics
    stats[n // 2:, :3] = np.sort(rng.uniform(0, 1, (n - n // 2, 3)), axis=1).astype(np.float32)
    return feats, stats


def gpu_similarities(ctx, prox, bc, o, feats, stats):
    n, K = len(feats), prox.max_neighbors
    sims = torch.full((n, K + 1), float("nan"), dtype=torch.float32, device="cuda")
    st_in = dev(stats)
    st_out = torch.full((n, 4), float("nan"), dtype=torch.float32, device="cuda")
    ctx.bc_similarities(prox, bc, dev(o["hash"].view(np.int32)), dev(o["lists"].view(np.int32)), dev(feats),
                        st_in, st_out, sims)
    return f32(sims), f32(st_out)


def test_bc_params_validation(ctx):
    prox = ProxParams.make(8, 4.0)
    h = torch.zeros((4, 2), dtype=torch.int32, device="cuda")
    L = torch.zeros((4, 12), dtype=torch.int32, device="cuda")
    f = torch.zeros((4, MAX_FEATURES + 1), dtype=torch.float32, device="cuda")
    st = torch.zeros((4, 4), dtype=torch.float32, device="cuda")
    s = torch.zeros((4, 9), dtype=torch.float32, device="cuda")
    for bad in (BcParams.make(0), BcParams.make(MAX_FEATURES + 1)):
        with pytest.raises(BrtError):
            ctx.bc_similarities(prox, bad, h, L, f, st, st.clone(), s)
    with pytest.raises(BrtError):   # adaptive statistics cannot be updated in place
        ctx.bc_similarities(prox, BcParams.make(4), h, L, f, st, st, s)


@pytest.mark.parametrize("F", [1, 3, 8, MAX_FEATURES])
@pytest.mark.parametrize("adaptive", [False, True])
def test_similarities_bit_exact(ctx, orc, F, adaptive):
    w = mixed_workload()
    o = oracle_grid_outputs(orc, w.grid, w.prox, w.state)
    feats, stats = features_and_stats(w.n, F, seed=F)
    bc = BcParams.make(F, adaptive=adaptive, stats_params=(0.1, 0.1, 0.001), avg_similarity_controller=0.75)
    sims, st = gpu_similarities(ctx, w.prox, bc, o, feats, stats)
    rs, rst = orc.bc_similarities(w.prox, bc, o["hash"], o["lists"], feats, stats if adaptive else None)
    assert first_mismatch(sims.view(np.uint32), rs.view(np.uint32)) is None
    if adaptive:
        assert first_mismatch(st.view(np.uint32), rst.view(np.uint32)) is None
    if not adaptive:
        r64, _ = orc.bc_similarities(w.prox, bc, o["hash"], o["lists"], feats.astype(np.float64), dtype=np.float64)
        err = np.abs(sims - r64).max()
        print(f"F={F}: naive similarity max abs err vs FP64 = {err:.3e}")
        assert err < 1e-5
    assert o["lists"][:, 0].max() > 0


@pytest.mark.parametrize("K", [7, 31, 32, 64, 100])
def test_similarities_list_widths(ctx, orc, K):
    """List widths on both sides of the LDS row-tile switch (256 agents up to
    K = 39, one wave above), with an even row length (K + 1 = 32) and with the
    36-word rows the list walk loads whole (K = 32); the one-pass similarities
    + cluster steering as well."""
    w = mixed_workload(n=20000, K=K, seed=K)
    o = oracle_grid_outputs(orc, w.grid, w.prox, w.state)
    F = 5
    feats, stats = features_and_stats(w.n, F, seed=K)
    bc = BcParams.make(F, adaptive=True)
    sims, st = gpu_similarities(ctx, w.prox, bc, o, feats, stats)
    rs, rst = orc.bc_similarities(w.prox, bc, o["hash"], o["lists"], feats, stats)
    assert first_mismatch(sims.view(np.uint32), rs.view(np.uint32)) is None
    assert first_mismatch(st.view(np.uint32), rst.view(np.uint32)) is None
    sp = SteerParams.make()
    f0 = np.random.default_rng(1).standard_normal((w.n, 4)).astype(np.float32)
    s2 = torch.full((w.n, K + 1), float("nan"), dtype=torch.float32, device="cuda")
    st2 = empty_f32(w.n, 4)
    f1 = empty_f32(w.n, 4)
    ctx.bc_similarities_steer(w.prox, sp, bc, dev(o["hash"].view(np.int32)), dev(o["lists"].view(np.int32)),
                              dev(feats), dev(stats), st2, s2, dev(o["sorted_pos"]), dev(o["sorted_forward"]), dev(f0),
                              f1)
    rf = orc.bc_cluster_steer(w.prox, sp, bc, o["sorted_pos"], o["sorted_forward"], o["lists"], rs, f0)
    assert first_mismatch(f32(s2).view(np.uint32), rs.view(np.uint32)) is None
    assert first_mismatch(f32(f1).view(np.uint32), rf.view(np.uint32)) is None


@pytest.mark.parametrize("K", [16, 32])
@pytest.mark.parametrize("variant", range(4))
def test_decomposed_steering_equals_fused(ctx, orc, variant, K):
    """steer_forces -> apply_steering_force reproduces the fused steering pass bit
    for bit (the reference's kernel-per-behaviour sequence summing onto the
    steer force), and the behaviour force equals the FP32 oracle's; K 32 rows
    (36 words) are the ones the list walk loads whole."""
    w = mixed_workload(K=K)
    o = oracle_grid_outputs(orc, w.grid, w.prox, w.state)
    sp = [
        SteerParams.make(weights=(1.0, 0.8, 0.6)),
        SteerParams.make(weights=(1.0, 0.8, 0.6), seek_world_center=True),
        SteerParams.make(weights=(1.0, 0.8, 0.6), use_2d=True, mass=2.0, max_force=0.2),
        SteerParams.make(weights=(1.0, 0.8, 0.6), targets=W.moving_targets(1.5, 20.0), target_weight=0.5,
                         wander_weight=0.3, wander_seed=5),
    ][variant]
    n = w.n
    wander0 = np.random.default_rng(3).uniform(-1, 1, (n, 2)).astype(np.float32)
    args = (dev(o["sorted_pos"]), dev(o["sorted_forward"]))
    h, L = dev(o["hash"].view(np.int32)), dev(o["lists"].view(np.int32))
    ssa = dev(o["sorted_smooth_accel"])

    fused = [empty_f32(n, 4) for _ in range(4)]
    wd_f = dev(wander0)
    ctx.steer_integrate(w.grid, w.prox, sp, w.dt, 7, *args, ssa, h, L, *fused[:3], wander=wd_f, force_out=fused[3])

    force = empty_f32(n, 4)
    wd_d = dev(wander0)
    ctx.steer_forces(w.prox, sp, w.dt, 7, *args, h, L, force, wander=wd_d)
    dec = [empty_f32(n, 4) for _ in range(4)]
    ctx.apply_steering_force(w.grid, sp, w.dt, *args, ssa, h, force, *dec[:3], force_out=dec[3])
    for a, b, k in zip(fused, dec, ("pos", "forward", "smooth_accel", "force")):
        assert first_mismatch(f32(a).view(np.uint32), f32(b).view(np.uint32)) is None, k
    assert first_mismatch(f32(wd_f).view(np.uint32), f32(wd_d).view(np.uint32)) is None

    rf, rw = orc.steer_forces(w.prox, sp, w.dt, 7, o["sorted_pos"], o["sorted_forward"], o["hash"], o["lists"],
                              wander=wander0)
    assert first_mismatch(f32(force).view(np.uint32), rf.view(np.uint32)) is None
    assert first_mismatch(f32(wd_d).view(np.uint32), rw.view(np.uint32)) is None


def cluster_condition(prox, steer, spos, sfwd, lists, sims):
    """Per agent, the larger of the two cluster sums' condition numbers
    sum|terms| / |sum of terms| (FP64), which bounds how much the final
    normalisations amplify the rounding of the terms."""
    n = len(spos)
    K = prox.max_neighbors
    cnt = np.minimum(lists[:, 0], K).astype(np.int64)
    idx = lists[:, 1:K + 1].astype(np.int64)
    valid = np.arange(K)[None, :] < cnt[:, None]
    idx = np.where(valid, idx, 0)
    p = spos[:, :3].astype(np.float64)
    f = sfwd[:, :3].astype(np.float64)
    sp_ = sfwd[:, 3].astype(np.float64)
    s = sims[:, :K].astype(np.float64) * valid
    off = p[idx] - p[:, None, :]
    ln = np.linalg.norm(off, axis=-1, keepdims=True)
    tr = np.where(ln > steer.max_speed, off * (steer.max_speed / np.maximum(ln, 1e-300)), off)
    v = (f * sp_[:, None])[:, None, :]
    terms = ((tr - v) * s[..., None] + (-tr - v) * (1 - s[..., None])) * valid[..., None]
    c1 = np.abs(terms).sum(1).sum(-1) / np.maximum(np.linalg.norm(terms.sum(1), axis=-1), 1e-300)
    at = f[idx] * s[..., None] * valid[..., None]
    num = np.abs(at).sum(1).sum(-1) + cnt * np.abs(f).sum(-1)
    den = np.linalg.norm(at.sum(1) - cnt[:, None] * f, axis=-1)
    c2 = num / np.maximum(den, 1e-300)
    return np.where(cnt > 0, np.maximum(c1, c2), 1.0)


def adaptive_sensitivity(bc, naive, stats_sorted, counts, tie_tol=1e-5):
    """Replays the adaptive statistics along an FP64 trajectory of naive
    similarities to locate where the FP64 reference is not a continuous
    function of its inputs or strongly amplifies them, so that those agents
    are left out of the FP64 comparison (never to check values).  Returns, per
    sorted agent, the largest interpolation slope 1/2 / (avg - min) or
    1/2 / (max - avg) used, and whether any discontinuous branch of the update
    (sim < min, sim > max, min > avg, max < avg; the +-rate3 memory steps) was
    taken within tie_tol of its threshold."""
    r1, r2, r3 = (float(x) for x in bc.stats_params)
    ctrl = float(bc.avg_similarity_controller)
    st = np.asarray(stats_sorted, np.float64)[:, :3].copy()
    amp = np.zeros(len(st))
    tie = np.zeros(len(st), bool)
    for t in range(naive.shape[1] - 1):
        act = t < counts
        v = naive[:, t]
        mn, avg0, mx = st[:, 0], st[:, 1], st[:, 2]
        avg = np.clip(avg0 + (v * ctrl + ((mx + mn) / 2) * (1 - ctrl) - avg0) * r1, 0, 1)
        near = lambda a, b: np.abs(a - b) < tie_tol
        lo, hi = v < mn, v > mx
        t1 = near(v, mn) | near(mn, avg) | near(v, mx) | near(mx, avg)
        mn2 = np.where(lo, mn + (v - mn) * r2, np.where(mn < avg, mn + r3, mn))
        mx2 = np.where(hi, mx + (v - mx) * r2, np.where(mx > avg, mx - r3, mx))
        t1 |= (lo & near(mn2, avg)) | (hi & near(mx2, avg))
        mn2 = np.where(lo & (mn2 > avg), np.clip(mn2, 0, v), mn2)
        mx2 = np.where(hi & (mx2 < avg), np.clip(mx2, v, 1), mx2)
        st = np.where(act[:, None], np.stack([mn2, avg, mx2], 1), st)
        den = np.where(v > st[:, 1], st[:, 2] - st[:, 1], st[:, 1] - st[:, 0])
        amp = np.where(act, np.maximum(amp, 0.5 / np.maximum(np.abs(den), 1e-300)), amp)
        tie |= act & t1
    return amp, tie


@pytest.mark.parametrize("adaptive", [False, True])
def test_clustering_step_bit_exact_and_fp64(ctx, orc, adaptive, record_property):
    """The reference's BehaveRT_Clustering update: behaviours, steerForClustering,
    world-centre seek, integration (BehaveRT_Clustering.cpp:212-235)."""
    w = mixed_workload(n=60000, K=24, seed=7, slow=False)
    o = oracle_grid_outputs(orc, w.grid, w.prox, w.state)
    n, F = w.n, 6
    feats, stats = features_and_stats(n, F, seed=11)
    sp = SteerParams.make(weights=(1.0, 0.8, 0.6), seek_world_center=True)
    bc = BcParams.make(F, adaptive=adaptive, cluster_weight=0.9, alignment_weight=0.7)
    args = (dev(o["sorted_pos"]), dev(o["sorted_forward"]))
    h, L = dev(o["hash"].view(np.int32)), dev(o["lists"].view(np.int32))

    f0 = empty_f32(n, 4)
    ctx.steer_forces(w.prox, sp, w.dt, 0, *args, h, L, f0)
    sims = torch.empty((n, w.prox.max_neighbors + 1), dtype=torch.float32, device="cuda")
    st_out = empty_f32(n, 4)
    ctx.bc_similarities(w.prox, bc, h, L, dev(feats), dev(stats), st_out, sims)
    f1 = empty_f32(n, 4)
    ctx.bc_steer_for_clustering(w.prox, sp, bc, *args, L, sims, f0, f1)
    # the one-pass form (similarities + cluster steering) gives the same bits
    sims_b = torch.full_like(sims, float("nan"))
    st_b = torch.full_like(st_out, float("nan"))
    f1_b = torch.full_like(f1, float("nan"))
    ctx.bc_similarities_steer(w.prox, sp, bc, h, L, dev(feats), dev(stats), st_b, sims_b, *args, f0, f1_b)
    for a_, b_, k in ((sims, sims_b, "similarities"), (f1, f1_b, "force")) + (((st_out, st_b, "stats"),)
                                                                               if adaptive else ()):
        assert first_mismatch(f32(a_).view(np.uint32), f32(b_).view(np.uint32)) is None, k
    out = [empty_f32(n, 4) for _ in range(4)]
    ctx.apply_steering_force(w.grid, sp, w.dt, *args, dev(o["sorted_smooth_accel"]), h, f1, *out[:3],
                             force_out=out[3])
    got = dict(zip(("pos", "forward", "smooth_accel", "force"), (f32(t) for t in out)))
    got_f1 = f32(f1)

    # FP32 oracle, same sequence
    r0, _ = orc.steer_forces(w.prox, sp, w.dt, 0, o["sorted_pos"], o["sorted_forward"], o["hash"], o["lists"])
    rs, _ = orc.bc_similarities(w.prox, bc, o["hash"], o["lists"], feats, stats if adaptive else None)
    r1 = orc.bc_cluster_steer(w.prox, sp, bc, o["sorted_pos"], o["sorted_forward"], o["lists"], rs, r0)
    assert first_mismatch(got_f1.view(np.uint32), r1.view(np.uint32)) is None, "cluster force"
    ref = orc.apply_steering_force(w.grid, sp, w.dt, o["sorted_pos"], o["sorted_forward"],
                                   o["sorted_smooth_accel"], o["hash"], r1)
    for k in ("pos", "forward", "smooth_accel", "force"):
        assert first_mismatch(got[k].view(np.uint32), ref[k].view(np.uint32)) is None, k

    # FP64 evaluation of the same step (same FP32 inputs)
    d64 = lambda a: a.astype(np.float64)
    q0, _ = orc.steer_forces(w.prox, sp, w.dt, 0, d64(o["sorted_pos"]), d64(o["sorted_forward"]), o["hash"],
                             o["lists"], dtype=np.float64)
    qs, _ = orc.bc_similarities(w.prox, bc, o["hash"], o["lists"], d64(feats), d64(stats) if adaptive else None,
                                dtype=np.float64)
    q1 = orc.bc_cluster_steer(w.prox, sp, bc, d64(o["sorted_pos"]), d64(o["sorted_forward"]), o["lists"], qs, q0,
                              dtype=np.float64)
    r64 = orc.apply_steering_force(w.grid, sp, w.dt, d64(o["sorted_pos"]), d64(o["sorted_forward"]),
                                   d64(o["sorted_smooth_accel"]), o["hash"], q1, dtype=np.float64)
    kappa = cluster_condition(w.prox, sp, o["sorted_pos"], o["sorted_forward"], o["lists"], qs)
    ok_sorted = kappa <= 100.0
    ids = o["hash"][:, 1]
    if adaptive:
        naive, _ = orc.bc_similarities(w.prox, BcParams.make(F, adaptive=False), o["hash"], o["lists"], d64(feats),
                                       dtype=np.float64)
        amp, tie = adaptive_sensitivity(bc, naive, d64(stats)[ids], np.minimum(o["lists"][:, 0], w.prox.max_neighbors))
        print(f"adaptive statistics: {int(tie.sum())} agents with a branch tie, "
              f"{int((amp > 100).sum())} with an interpolation slope above 100")
        ok_sorted &= (amp <= 100.0) & ~tie
    ok = np.zeros(n, bool)
    ok[ids] = ok_sorted
    print(f"adaptive={adaptive}: cluster-sum condition median {np.median(kappa):.2f}, "
          f"p99 {np.quantile(kappa, 0.99):.1f}; {int((~ok_sorted).sum())} of {n} agents left out")
    assert ok_sorted.mean() > 0.9
    ferr = floored_rel(got_f1[:, :3], q1[:, :3], 1.0)
    print(f"cluster force: max floored rel err vs FP64 (well-conditioned) = {ferr[ok_sorted].max():.3e}, "
          f"all agents {ferr.max():.3e}")
    assert ferr[ok_sorted].max() < 1e-5
    for k, floor in FLOORS.items():
        err = floored_rel(got[k][:, :3], r64[k][:, :3], floor)
        record_property(f"clustering_{k}_max_err_vs_fp64", float(err[ok].max()))
        print(f"{k}: max floored rel err vs FP64 = {err[ok].max():.3e} (all agents {err.max():.3e})")
        assert err[ok].max() < 1e-5, k


@pytest.mark.parametrize("K", [16, 32])
@pytest.mark.parametrize("iterations", [0, 1, 2, 10, 25])
def test_cluster_membership_bit_exact(ctx, orc, iterations, K):
    """computeClusterMembership (reset + Jacobi min-label iterations +
    connector) against the oracle's by-ID restatement, labels and connector bit
    for bit; .y/.z carried; in-place update gives the same labels.  K 16 rows
    are read in prefetched pairs, K 32 rows whole."""
    w = mixed_workload(n=30000, K=K, seed=13)
    o = oracle_grid_outputs(orc, w.grid, w.prox, w.state)
    n = w.n
    m_in = np.random.default_rng(2).uniform(-1, 1, (n, 3)).astype(np.float32)
    h, L, sp = dev(o["hash"].view(np.int32)), dev(o["lists"].view(np.int32)), dev(o["sorted_pos"])
    m_out = torch.full((n, 3), float("nan"), dtype=torch.float32, device="cuda")
    conn = torch.full((n, 2, 4), float("nan"), dtype=torch.float32, device="cuda")
    ctx.cluster_membership(w.prox, iterations, h, L, sp, dev(m_in), m_out, conn)
    rm, rc = orc.cluster_membership(w.prox, iterations, o["hash"], o["lists"], o["sorted_pos"], m_in, connector=True)
    assert first_mismatch(f32(m_out).view(np.uint32), rm.view(np.uint32)) is None
    assert first_mismatch(f32(conn).view(np.uint32), rc.view(np.uint32)) is None
    labels = rm[:, 0]
    print(f"T={iterations}: {len(np.unique(labels))} distinct labels over {n} agents")
    m_io = dev(m_in)
    ctx.cluster_membership(w.prox, iterations, h, L, sp, m_io, m_io)
    assert first_mismatch(f32(m_io).view(np.uint32), rm.view(np.uint32)) is None


@pytest.mark.parametrize("K", [16, 32])
@pytest.mark.parametrize("radius", [1.5, 4.0])
def test_cluster_hit_rate_bit_exact(ctx, orc, radius, K):
    w = mixed_workload(n=30000, K=K, seed=13)
    o = oracle_grid_outputs(orc, w.grid, w.prox, w.state)
    m0 = np.random.default_rng(4).uniform(-1, 1, (w.n, 3)).astype(np.float32)
    m, _ = orc.cluster_membership(w.prox, 10, o["hash"], o["lists"], o["sorted_pos"], m0)
    h, L, sp = dev(o["hash"].view(np.int32)), dev(o["lists"].view(np.int32)), dev(o["sorted_pos"])
    out = torch.full((w.n, 3), float("nan"), dtype=torch.float32, device="cuda")
    ctx.cluster_hit_rate(w.prox, radius, h, L, sp, dev(m), out)
    ref = orc.cluster_hit_rate(w.prox, radius, o["hash"], o["lists"], o["sorted_pos"], m)
    assert first_mismatch(f32(out).view(np.uint32), ref.view(np.uint32)) is None
    hr = ref[:, 1][ref[:, 2] >= 0]
    print(f"r={radius}: mean hit rate {hr.mean():.3f} over {len(hr)} agents with close neighbours")
    assert 0 < hr.mean() <= 1


def test_cluster_membership_limits(ctx):
    prox = ProxParams.make(8, 4.0)
    h = torch.zeros((4, 2), dtype=torch.int32, device="cuda")
    L = torch.zeros((4, 12), dtype=torch.int32, device="cuda")
    p = torch.zeros((4, 4), dtype=torch.float32, device="cuda")
    m = torch.zeros((4, 3), dtype=torch.float32, device="cuda")
    import ctypes as C
    from behavert_amd.engine import _ptr
    rc = ctx._lib.brt_bc_cluster_membership(ctx.handle, C.byref(prox), (1 << 24) + 1, 10, _ptr(h), _ptr(L), _ptr(p),
                                            _ptr(m), _ptr(m), None, None)
    assert rc != 0 and b"2^24" in ctx._lib.brt_context_last_error(ctx.handle)
    with pytest.raises(BrtError):   # the hit rate reads neighbours' labels: no in-place update
        ctx.cluster_hit_rate(prox, 1.5, h, L, p, m, m)
