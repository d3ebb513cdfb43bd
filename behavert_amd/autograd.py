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
                 steps: int, checkpoint_every: int, device="cuda"):
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
        # per-step records of one segment
        self.r_spos, self.r_sfwd, self.r_ssa = f4(self.K, n), f4(self.K, n), f4(self.K, n)
        self.r_hash = torch.empty((self.K, n, 2), **i32)
        self.r_lists = torch.empty((self.K, n, self.W), **i32)
        # cell ranges, rows padded to 16 bytes (the step's alignment)
        cpad = (grid.num_cells + 3) // 4 * 4
        self.r_cs = torch.empty((self.K, cpad), **i32)
        self.r_ce = torch.empty((self.K, cpad), **i32)
        # adjoints (two sets, ping-pong) and the weight gradient
        self.g = [(f4(n), f4(n), f4(n)), (f4(n), f4(n), f4(n))]
        self.wgrad = torch.zeros(3, dtype=torch.float32, device=dev)
        self.checkpoints = []
        # argument blocks built once: the per-step calls only pass them on
        self._bufs = self._buffers(self.spos, self.sfwd, self.ssa, self.hash, self.lists, self.cell_start,
                                   self.cell_end)
        self._rec_bufs = [self._buffers(self.r_spos[k], self.r_sfwd[k], self.r_ssa[k], self.r_hash[k],
                                        self.r_lists[k], self.r_cs[k], self.r_ce[k]) for k in range(self.K)]
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

    def _step(self, t, bufs: StepBuffers):
        self.ctx.check(lib().brt_step(self.ctx.handle, C.byref(self.grid), C.byref(self.prox), C.byref(self.steer),
                                      0.0, self.dt, int(t), C.byref(bufs), current_stream()))

    def _adjoint(self, k, cur):
        key = (k, cur)
        args = self._adj_args.get(key)
        if args is None:
            go, gi = self.g[cur], self.g[1 - cur]
            args = (self.ctx.handle, C.byref(self.grid), C.byref(self.prox), C.byref(self.steer), self.dt, self.n,
                    _p(self.r_spos[k]), _p(self.r_sfwd[k]), _p(self.r_ssa[k]), _p(self.r_hash[k]),
                    _p(self.r_lists[k]), _p(self.r_cs[k]), _p(self.r_ce[k]), _p(go[0]), _p(go[1]), _p(go[2]),
                    _p(gi[0]), _p(gi[1]), _p(gi[2]), _p(self.wgrad))
            self._adj_args[key] = args
        self.ctx.check(lib().brt_osw_adjoint_step(*args, current_stream()))

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
        for (t0, p, f, a) in reversed(checkpoints):
            t1 = min(t0 + self.K, self.steps)
            self.pos.copy_(p)
            self.fwd.copy_(f)
            self.sa.copy_(a)
            for t in range(t0, t1):
                self._step(t, self._rec_bufs[t - t0])
            for t in range(t1 - 1, t0 - 1, -1):
                self._adjoint(t - t0, cur)
                cur = 1 - cur
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
