"""Shared helpers of the GPU parity tests (device path through the C ABI vs the
CPU oracle)."""
from __future__ import annotations

import numpy as np
import torch


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def u32(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint32)


def f32(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def empty_u32(*shape):
    return torch.empty(shape, dtype=torch.int32, device="cuda")


def empty_f32(*shape):
    return torch.empty(shape, dtype=torch.float32, device="cuda")


def oracle_grid_outputs(orc, grid, prox, state):
    """hash, ranges, sorted state and lists of one step from the oracle."""
    h = orc.sort_radix(orc.calc_hash(grid, state["pos"]), grid.num_cells)
    s, e = orc.cell_ranges(h, grid.num_cells)
    spos = orc.reorder(h, state["pos"])
    sf = orc.reorder(h, state["forward"])
    ssa = orc.reorder(h, state["smooth_accel"])
    L = orc.neighbors(grid, prox, spos, h, s, e)
    return {"hash": h, "cell_start": s, "cell_end": e, "sorted_pos": spos, "sorted_forward": sf,
            "sorted_smooth_accel": ssa, "lists": L}


def floored_rel(a32, a64, floor):
    a32 = np.asarray(a32, np.float64)
    d = np.linalg.norm(a32 - a64, axis=-1)
    return d / np.maximum(np.linalg.norm(a64, axis=-1), floor)


def first_mismatch(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return f"shape {a.shape} vs {b.shape}"
    idx = np.argwhere(a != b)
    if len(idx) == 0:
        return None
    i = tuple(idx[0])
    return f"{len(idx)} mismatches, first at {i}: {a[i]} vs {b[i]}"
