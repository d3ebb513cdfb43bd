"""The binary snapshot format: the library's header parser (no device needed)
against the independent numpy writer/reader in oracle/snapshot.py, and the
FNV-1a 64 checksum against its published test vectors."""
import ctypes as C

import numpy as np
import pytest

from behavert_amd import workloads as W
from behavert_amd.types import SnapshotInfo, list_stride
from oracle import snapshot as S


def test_fnv1a64_vectors():
    # Fowler/Noll/Vo test vectors for FNV-1a 64
    assert S.fnv1a64(b"") == 0xCBF29CE484222325
    assert S.fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert S.fnv1a64(b"foobar") == 0x85944171F73967E8


def small_arrays(n, W, seed=0):
    rng = np.random.default_rng(seed)
    return {"pos": rng.standard_normal((n, 4)), "forward": rng.standard_normal((n, 4)),
            "smooth_accel": rng.standard_normal((n, 4)), "hash": rng.integers(0, 1 << 31, (n, 2)),
            "lists": rng.integers(0, n, (n, W)), "features": rng.standard_normal((n, 4)),
            "wander": rng.uniform(-1, 1, (n, 2))}


@pytest.mark.parametrize("flags", [0, S.LISTS, S.LISTS | S.FEATURES | S.WANDER, S.WANDER])
def test_library_reads_numpy_written_header(tmp_path, flags):
    from behavert_amd._lib import lib

    w = W.plain_4096()
    n, K = 257, w.prox.max_neighbors
    Wd = list_stride(K)
    path = str(tmp_path / "s.brts")
    arrays = small_arrays(n, Wd)
    S.write(path, n, K, Wd, flags, 77, bytes(w.grid), bytes(w.prox), arrays)
    info = SnapshotInfo()
    assert lib().brt_snapshot_read_info(path.encode(), C.byref(info)) == 0
    assert (info.version, info.n, info.max_neighbors, info.list_stride, info.flags, info.step) == \
        (1, n, K, Wd, flags, 77)
    assert bytes(info.grid) == bytes(w.grid) and bytes(info.prox) == bytes(w.prox)
    back = S.read(path)
    assert info.payload_bytes == back["payload_bytes"] and info.checksum == back["checksum"]
    for k in ("pos", "forward", "smooth_accel"):
        np.testing.assert_array_equal(back[k], arrays[k].astype(np.float32))


def test_library_rejects_bad_headers(tmp_path):
    from behavert_amd._lib import lib

    w = W.plain_4096()
    path = tmp_path / "s.brts"
    S.write(str(path), 8, 16, list_stride(16), 0, 0, bytes(w.grid), bytes(w.prox), small_arrays(8, 20))
    raw = bytearray(path.read_bytes())
    info = SnapshotInfo()
    bad = tmp_path / "bad_magic.brts"
    bad.write_bytes(b"X" + bytes(raw[1:]))
    assert lib().brt_snapshot_read_info(str(bad).encode(), C.byref(info)) != 0
    ver = bytearray(raw)
    ver[8] = 2
    (tmp_path / "v2.brts").write_bytes(bytes(ver))
    assert lib().brt_snapshot_read_info(str(tmp_path / "v2.brts").encode(), C.byref(info)) != 0
    (tmp_path / "short.brts").write_bytes(bytes(raw[:100]))
    assert lib().brt_snapshot_read_info(str(tmp_path / "short.brts").encode(), C.byref(info)) != 0
    assert lib().brt_snapshot_read_info(str(tmp_path / "missing.brts").encode(), C.byref(info)) != 0


def test_library_rejects_inconsistent_sizes(tmp_path):
    """A header whose payload size does not follow from its n, K and flags (a
    corrupt or hostile file) is rejected before anything is allocated, and so
    are an out-of-range K and a list stride that does not match it."""
    import struct

    from behavert_amd._lib import lib

    w = W.plain_4096()
    path = tmp_path / "s.brts"
    S.write(str(path), 8, 16, list_stride(16), S.LISTS, 0, bytes(w.grid), bytes(w.prox), small_arrays(8, 20))
    raw = bytearray(path.read_bytes())
    info = SnapshotInfo()
    assert lib().brt_snapshot_read_info(str(path).encode(), C.byref(info)) == 0
    cases = {
        "huge_payload": (120, "<Q", 1 << 62),
        "payload_off_by_4": (120, "<Q", info.payload_bytes + 4),
        "k_zero": (20, "<I", 0),
        "k_too_big": (20, "<I", 253),
        "stride_mismatch": (24, "<I", list_stride(16) + 4),
        "unknown_flag": (28, "<I", S.LISTS | 8),
        "n_grows": (16, "<I", 1 << 31),
    }
    for name, (off, fmt, val) in cases.items():
        bad = bytearray(raw)
        struct.pack_into(fmt, bad, off, val)
        p = tmp_path / f"{name}.brts"
        p.write_bytes(bytes(bad))
        assert lib().brt_snapshot_read_info(str(p).encode(), C.byref(info)) == -1, name
