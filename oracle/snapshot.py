"""Test infrastructure: a numpy restatement of the binary snapshot format
(256-byte header, then the sections), written independently of the library's
C implementation so the tests can check one against the other."""
from __future__ import annotations

import struct

import numpy as np

MAGIC = b"BRTSNAP\0"
HEADER = 256
LISTS, FEATURES, WANDER = 1, 2, 4


def fnv1a64(data: bytes) -> int:
    h = 0xCBF29CE484222325
    arr = np.frombuffer(data, np.uint8)
    # FNV-1a is sequential; process in Python integers (test-sized files only)
    for b in arr.tolist():
        h ^= b
        h = (h * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def _sections(n, W, flags):
    s = [("pos", (n, 4), np.float32), ("forward", (n, 4), np.float32), ("smooth_accel", (n, 4), np.float32)]
    if flags & LISTS:
        s += [("hash", (n, 2), np.uint32), ("lists", (n, W), np.uint32)]
    if flags & FEATURES:
        s += [("features", (n, 4), np.float32)]
    if flags & WANDER:
        s += [("wander", (n, 2), np.float32)]
    return s


def write(path, n, K, W, flags, step, grid_bytes: bytes, prox_bytes: bytes, arrays: dict, checksum=True):
    payload = b"".join(np.ascontiguousarray(arrays[name], dt).tobytes() for name, _, dt in _sections(n, W, flags))
    h = bytearray(HEADER)
    h[0:8] = MAGIC
    struct.pack_into("<IIIIIIII", h, 8, 1, HEADER, n, K, W, flags, step, len(grid_bytes))
    h[40:40 + len(grid_bytes)] = grid_bytes
    h[108:108 + len(prox_bytes)] = prox_bytes
    struct.pack_into("<QQ", h, 120, len(payload), fnv1a64(payload) if checksum else 0)
    with open(path, "wb") as f:
        f.write(bytes(h) + payload)


def read(path) -> dict:
    with open(path, "rb") as f:
        data = f.read()
    assert data[:8] == MAGIC
    version, hdr, n, K, W, flags, step, gb = struct.unpack_from("<IIIIIIII", data, 8)
    payload_bytes, checksum = struct.unpack_from("<QQ", data, 120)
    out = {"version": version, "n": n, "K": K, "W": W, "flags": flags, "step": step,
           "payload_bytes": payload_bytes, "checksum": checksum, "grid_bytes": data[40:40 + gb],
           "prox_bytes": data[108:120]}
    off = hdr
    for name, shape, dt in _sections(n, W, flags):
        cnt = int(np.prod(shape))
        out[name] = np.frombuffer(data, dt, cnt, off).reshape(shape).copy()
        off += cnt * np.dtype(dt).itemsize
    assert off == len(data) == hdr + payload_bytes
    return out
