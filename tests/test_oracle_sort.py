"""Pins the oracle's stable key/ID sort (BehaveRT::RadixSort,
PlugIns/EnvGrid3D/radixsort.cu:50-80): a stable LSD sort whose values are the
agent IDs 0..N-1 in order, i.e. the (key, ID) lexicographic order."""
import numpy as np
import pytest


def pairs(keys):
    keys = np.asarray(keys, np.uint32)
    return np.stack([keys, np.arange(len(keys), dtype=np.uint32)], 1)


@pytest.mark.parametrize("n,cells", [(1, 1), (2, 2), (1000, 7), (4096, 15625), (50000, 1_000_000),
                                     (70000, 4_096_000), (3000, 2 ** 31 - 1)])
def test_radix_equals_lexicographic_and_stable_sort(orc, n, cells):
    rng = np.random.default_rng(n)
    k = rng.integers(0, cells, n, dtype=np.uint64).astype(np.uint32)
    h = pairs(k)
    r = orc.sort_radix(h, cells)
    s = orc.sort_stable(h)
    lex = h[np.lexsort((h[:, 1], h[:, 0]))]
    assert np.array_equal(r, s)
    assert np.array_equal(r, lex)


def test_sort_invariants_with_heavy_duplicates(orc):
    rng = np.random.default_rng(5)
    k = rng.integers(0, 3, 20000).astype(np.uint32)
    r = orc.sort_radix(pairs(k), 3)
    assert np.all(np.diff(r[:, 0].astype(np.int64)) >= 0)
    same = r[1:, 0] == r[:-1, 0]
    assert np.all(r[1:, 1][same] > r[:-1, 1][same])
    assert np.array_equal(np.sort(r[:, 1]), np.arange(20000, dtype=np.uint32))


def test_cell_ranges_closed_form(orc):
    rng = np.random.default_rng(7)
    cells = 500
    k = rng.integers(0, cells, 3000).astype(np.uint32)
    r = orc.sort_radix(pairs(k), cells)
    s, e = orc.cell_ranges(r, cells)
    for c in range(cells):
        idx = np.nonzero(r[:, 0] == c)[0]
        if len(idx) == 0:
            assert s[c] == 0xFFFFFFFF and e[c] == 0xFFFFFFFF
        else:
            assert s[c] == idx[0] and e[c] == idx[-1] + 1
    occ = s != 0xFFFFFFFF
    assert (e[occ] - s[occ]).sum() == 3000   # ranges partition the sorted array


def test_reorder_is_exact_gather(orc):
    rng = np.random.default_rng(8)
    a = rng.standard_normal((1000, 4)).astype(np.float32)
    h = pairs(rng.integers(0, 50, 1000))
    r = orc.sort_radix(h, 50)
    out = orc.reorder(r, a)
    assert np.array_equal(out.view(np.uint32), a[r[:, 1]].view(np.uint32))
