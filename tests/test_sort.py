"""The device radix sort (m3dx_sort_pairs), the sort under every stage of
the step: grid tiers (src/world.cpp:186-229), pair runs, island lists and
the giant schedule.

The sort must be stable and exact, so it is checked against numpy's stable
argsort on the same keys: values are the input positions, so any reordering
of equal keys shows. Sizes straddle the 4096-pair tile (one tile, a partial
last tile, many tiles with the look-back chain), digit counts from 1 to 8
passes (odd pass counts end in the other buffer), and key sets with heavy
duplication (a few distinct keys), all-equal keys and keys that fill every
bit.
"""
import ctypes as C

import numpy as np
import pytest


def _sort(dev, keys, bits):
    n = keys.size
    vals = np.arange(n, dtype=np.uint32)
    ko = np.empty(n, dtype=np.uint64)
    vo = np.empty(n, dtype=np.uint32)
    rc = dev.m3dx_sort_pairs(keys.ctypes.data_as(C.c_void_p), vals.ctypes.data_as(C.c_void_p), n, bits,
                             ko.ctypes.data_as(C.c_void_p), vo.ctypes.data_as(C.c_void_p))
    assert rc == 0, dev.last_error()
    return ko, vo


def _expect(keys, bits):
    order = np.argsort(keys, kind="stable")
    return keys[order], order.astype(np.uint32)


CASES = [
    (1, 8, "random"),
    (2, 3, "random"),
    (4095, 12, "random"),
    (4096, 16, "random"),
    (4097, 16, "few"),
    (12289, 1, "random"),
    (100_003, 24, "random"),
    (100_003, 40, "few"),
    (131_072, 16, "random"),
    (135_169, 24, "few"),
    (262_144, 32, "equal"),
    (1_000_000, 64, "random"),
    (1_000_000, 20, "few"),
]


@pytest.mark.gpu
@pytest.mark.parametrize("n,bits,kind", CASES)
def test_sort_is_stable_and_exact(dev, n, bits, kind):
    rng = np.random.default_rng(n * 131 + bits)
    top = (1 << bits) - 1 if bits < 64 else (1 << 64) - 1
    if kind == "random":
        keys = rng.integers(0, top, size=n, dtype=np.uint64, endpoint=True)
    elif kind == "few":
        pool = rng.integers(0, top, size=7, dtype=np.uint64, endpoint=True)
        keys = pool[rng.integers(0, 7, size=n)]
    else:
        keys = np.full(n, top, dtype=np.uint64)
    ko, vo = _sort(dev, keys, bits)
    ek, ev = _expect(keys, bits)
    assert np.array_equal(ko, ek)
    assert np.array_equal(vo, ev)


@pytest.mark.gpu
def test_sort_rejects_bad_arguments(dev):
    k = np.zeros(4, dtype=np.uint64)
    v = np.zeros(4, dtype=np.uint32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    assert dev.m3dx_sort_pairs(p(k), p(v), 4, 65, p(k), p(v)) == -1
    assert dev.m3dx_sort_pairs(None, p(v), 4, 8, p(k), p(v)) == -1
    assert dev.m3dx_sort_pairs(None, None, 0, 8, None, None) == 0
