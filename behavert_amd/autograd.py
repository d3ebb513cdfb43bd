"""Differentiable rollouts: a torch.autograd.Function over the gfx950 step and its
hand-written adjoint (brt_osw_adjoint_step).

Forward: T plain steps (brt_step) from (pos0, forward0, smooth_accel0) with
behaviour weights w; the state is checkpointed every K steps.
Backward: for each segment, last first, the checkpoint is restored and the K
steps are re-run with their sorted state, sort, cell ranges and neighbour
lists recorded (the step is deterministic, so the records equal the
forward's), then the adjoint kernels walk the records backwards.  The sort and
the lists are thus held fixed as the forward built them, and the adjoint has
no atomics, so the gradients do not depend on the checkpoint interval.
The replay of a segment does not depend on the adjoint of the later one, so
(overlap=True) segment s-1 is replayed on a second stream, with a second
context and a second set of records, while the adjoint walks segment s: the
kernels of both are small at these sizes and share the GPU.

Gradients are returned for w (3 behaviour weights), pos0, forward0 (xyz = unit
heading, w = speed) and smooth_accel0.  Differentiating with respect to an
initial velocity v0 = f0 * s0 is done in torch around the Function
(forward0 = [v0/|v0|, |v0|]).
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import lib
from .engine import Context, current_stream
from .types import GridParams, ProxParams, SteerParams, StepBuffers, list_stride


def _p(t):
    return None if t is None else t.data_ptr()


class Rollout:
    """Buffers and settings of one differentiable rollout configuration."""

    def __init__(self, ctx: Context, grid: GridParams, prox: ProxParams, steer: SteerParams, n: int, dt: float,
                 steps: int, checkpoint_every: int, device="cuda", overlap: bool = True):
        if steer.use_2d_projection or steer.wander_weight != 0.0:
            raise ValueError("the adjoint supports neither 2-D projection nor wander")
        self.ctx, self.grid, self.prox, self.n, self.dt = ctx, grid, prox, n, float(dt)
        self.steer = SteerParams()
        C.memmove(C.byref(self.steer), C.byref(steer), C.sizeof(SteerParams))
        self.steps, self.K = int(steps), int(checkpoint_every)
        dev = torch.device(device)
        f4 = lambda *s: torch.empty(*s, 4, dtype=torch.float32, device=dev)  # noqa: E731
        i32 = dict(dtype=torch.int32, device=dev)
        self.W = list_stride(prox.max_neighbors)
        # working state (agent-ID order)
        self.pos, self.fwd, self.sa = f4(n), f4(n), f4(n)
        # scratch outputs of plain steps
        self.spos, self.sfwd, self.ssa = f4(n), f4(n), f4(n)
        self.hash = torch.empty((n, 2), **i32)
        self.lists = torch.empty((n, self.W), **i32)
        self.cell_start = torch.empty((grid.num_cells,), **i32)
        self.cell_end = torch.empty((grid.num_cells,), **i32)
        # per-step records of one segment (two sets when replays overlap the
        # adjoint: segment s-1 is recorded while segment s is read)
        self.overlap = bool(overlap)
        S = 2 if self.overlap else 1
        self.r_spos, self.r_sfwd, self.r_ssa = f4(S, self.K, n), f4(S, self.K, n), f4(S, self.K, n)
        self.r_hash = torch.empty((S, self.K, n, 2), **i32)
        self.r_lists = torch.empty((S, self.K, n, self.W), **i32)
        # cell ranges, rows padded to 16 bytes (the step's alignment)
        cpad = (grid.num_cells + 3) // 4 * 4
        self.r_cs = torch.empty((S, self.K, cpad), **i32)
        self.r_ce = torch.empty((S, self.K, cpad), **i32)
        # the replays' own context (scratch) and stream
        self.ctx_replay = Context(ctx.device) if self.overlap else ctx
        self.stream_replay = torch.cuda.Stream(device=dev) if self.overlap else None
        # adjoints (two sets, ping-pong) and the weight gradient
        self.g = [(f4(n), f4(n), f4(n)), (f4(n), f4(n), f4(n))]
        self.wgrad = torch.zeros(3, dtype=torch.float32, device=dev)
        self.checkpoints = []
        # argument blocks built once: the per-step calls only pass them on
        self._bufs = self._buffers(self.spos, self.sfwd, self.ssa, self.hash, self.lists, self.cell_start,
                                   self.cell_end)
        self._rec_bufs = [[self._buffers(self.r_spos[q, k], self.r_sfwd[q, k], self.r_ssa[q, k], self.r_hash[q, k],
                                         self.r_lists[q, k], self.r_cs[q, k], self.r_ce[q, k]) for k in range(self.K)]
                          for q in range(S)]
        self._adj_args = {}

    def set_weights(self, w):
        w = [float(x) for x in w]
        self.steer.separation[0], self.steer.alignment[0], self.steer.cohesion[0] = w

    def _buffers(self, spos, sfwd, ssa, hash_, lists, cell_start, cell_end) -> StepBuffers:
        b = StepBuffers()
        b.n = self.n
        b.pos, b.forward, b.smooth_accel = _p(self.pos), _p(self.fwd), _p(self.sa)
        b.sorted_pos, b.sorted_forward, b.sorted_smooth_accel = _p(spos), _p(sfwd), _p(ssa)
        b.hash, b.lists = _p(hash_), _p(lists)
        b.cell_start, b.cell_end = _p(cell_start), _p(cell_end)
        return b

    def _step(self, t, bufs: StepBuffers, ctx: Context | None = None):
        ctx = self.ctx if ctx is None else ctx
        ctx.check(lib().brt_step(ctx.handle, C.byref(self.grid), C.byref(self.prox), C.byref(self.steer),
                                 0.0, self.dt, int(t), C.byref(bufs), current_stream()))

    def _adjoint(self, q, k, cur):
        key = (q, k, cur)
        args = self._adj_args.get(key)
        if args is None:
            go, gi = self.g[cur], self.g[1 - cur]
            args = (self.ctx.handle, C.byref(self.grid), C.byref(self.prox), C.byref(self.steer), self.dt, self.n,
                    _p(self.r_spos[q, k]), _p(self.r_sfwd[q, k]), _p(self.r_ssa[q, k]), _p(self.r_hash[q, k]),
                    _p(self.r_lists[q, k]), _p(self.r_cs[q, k]), _p(self.r_ce[q, k]), _p(go[0]), _p(go[1]),
                    _p(go[2]), _p(gi[0]), _p(gi[1]), _p(gi[2]), _p(self.wgrad))
            self._adj_args[key] = args
        self.ctx.check(lib().brt_osw_adjoint_step(*args, current_stream()))

    def _replay(self, seg, q, ctx):
        """Re-runs the segment (t0, pos, fwd, sa) from its checkpoint into record set q."""
        t0, p, f, a = seg
        t1 = min(t0 + self.K, self.steps)
        self.pos.copy_(p)
        self.fwd.copy_(f)
        self.sa.copy_(a)
        for t in range(t0, t1):
            self._step(t, self._rec_bufs[q][t - t0], ctx)
        return t0, t1

    def forward(self, weights, pos0, fwd0, sa0):
        """Runs the rollout; keeps checkpoints; returns the final (pos, forward, smooth_accel)."""
        self.set_weights(weights)
        self.pos.copy_(pos0)
        self.fwd.copy_(fwd0)
        self.sa.copy_(sa0)
        self.checkpoints = []
        for t in range(self.steps):
            if t % self.K == 0:
                self.checkpoints.append((t, self.pos.clone(), self.fwd.clone(), self.sa.clone()))
            self._step(t, self._bufs)
        return self.pos.clone(), self.fwd.clone(), self.sa.clone()

    def backward(self, gpos, gfwd, gsa, checkpoints=None, weights=None):
        """dL/d(final state) -> dL/d(weights, pos0, forward0, smooth_accel0).
        checkpoints/weights: those of the forward run being differentiated
        (default: the last forward's); the segments are replayed with them."""
        if checkpoints is None:
            checkpoints = self.checkpoints
        if weights is not None:
            self.set_weights(weights)
        cur = 0
        self.g[cur][0].copy_(gpos)
        self.g[cur][1].copy_(gfwd)
        self.g[cur][2].copy_(gsa)
        self.wgrad.zero_()
        segs = list(reversed(checkpoints))
        if not self.overlap:
            for seg in segs:
                t0, t1 = self._replay(seg, 0, self.ctx)
                for t in range(t1 - 1, t0 - 1, -1):
                    self._adjoint(0, t - t0, cur)
                    cur = 1 - cur
        else:
            main, side = torch.cuda.current_stream(), self.stream_replay
            side.wait_stream(main)  # the checkpoints and the working state
            replayed, adjointed, spans = [], [], []

            def replay(i):
                with torch.cuda.stream(side):
                    if i >= 2:
                        side.wait_event(adjointed[i - 2])  # record set i % 2 was read by adjoint i - 2
                    spans.append(self._replay(segs[i], i % 2, self.ctx_replay))
                    ev = torch.cuda.Event()
                    ev.record(side)
                    replayed.append(ev)

            replay(0)
            for i in range(len(segs)):
                if i + 1 < len(segs):
                    replay(i + 1)  # enqueued before this segment's adjoint, so the two overlap
                main.wait_event(replayed[i])
                t0, t1 = spans[i]
                for t in range(t1 - 1, t0 - 1, -1):
                    self._adjoint(i % 2, t - t0, cur)
                    cur = 1 - cur
                ev = torch.cuda.Event()
                ev.record(main)
                adjointed.append(ev)
            main.wait_stream(side)  # leave no replay work pending on the side stream
        g = self.g[cur]
        return self.wgrad.clone(), g[0].clone(), g[1].clone(), g[2].clone()


class RolloutFunction(torch.autograd.Function):
    """final_pos (N,3) = rollout(weights (3,), pos0 (N,4), forward0 (N,4), smooth_accel0 (N,4))."""

    @staticmethod
    def forward(fctx, weights, pos0, fwd0, sa0, rollout: Rollout):
        fctx.rollout = rollout
        w = weights.detach().cpu().tolist()
        pos, _, _ = rollout.forward(w, pos0.detach(), fwd0.detach(), sa0.detach())
        # this run's checkpoints and weights: the Rollout's buffers are reused
        # by any later forward, and the backward must replay this one
        fctx.checkpoints = rollout.checkpoints
        fctx.weights = w
        return pos[:, :3].contiguous()

    @staticmethod
    def backward(fctx, gfinal):
        r: Rollout = fctx.rollout
        gpos = torch.zeros((r.n, 4), dtype=torch.float32, device=gfinal.device)
        gpos[:, :3] = gfinal
        z = torch.zeros_like(gpos)
        gw, gp0, gf0, ga0 = r.backward(gpos, z, z, checkpoints=fctx.checkpoints, weights=fctx.weights)
        return gw.to(gfinal.device), gp0, gf0, ga0, None


def rollout_final_pos(rollout: Rollout, weights, pos0, fwd0, sa0):
    return RolloutFunction.apply(weights, pos0, fwd0, sa0, rollout)


def forward_from_velocity(v0: torch.Tensor) -> torch.Tensor:
    """[v/|v|, |v|] as the (N,4) forward feature, differentiable in torch."""
    s = torch.linalg.norm(v0, dim=1, keepdim=True)
    return torch.cat([v0 / s, s], dim=1)
