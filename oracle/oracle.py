"""numpy/ctypes front end of the oracle's C restatement (m3d_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker for the device
backend's discrete stage outputs; nothing in melon3d_amd imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "c", "liboracle.so")


class orc_entry(C.Structure):
    _fields_ = [("key", C.c_uint64), ("body", C.c_int32), ("pad", C.c_int32)]


class orc_body_flags(C.Structure):
    _fields_ = [("inUse", C.c_uint8), ("dynamic", C.c_uint8), ("awake", C.c_uint8), ("moved", C.c_uint8),
                ("activeTier", C.c_uint8), ("stableTier", C.c_uint8), ("pad0", C.c_uint8), ("pad1", C.c_uint8),
                ("category", C.c_uint32), ("mask", C.c_uint32), ("group", C.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise FileNotFoundError(f"{LIB} not built (make -C oracle oracle)")
        L = C.CDLL(LIB)
        L.orc_cell_key.restype = C.c_uint64
        L.orc_cell_key.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float]
        L.orc_build_tier.restype = C.c_int64
        L.orc_build_tier.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_float, C.POINTER(orc_entry), C.c_int64,
                                     C.c_void_p, C.POINTER(C.c_int32), C.c_void_p]
        L.orc_brute_pairs.restype = C.c_int64
        L.orc_brute_pairs.argtypes = [C.c_int32, C.c_void_p, C.POINTER(orc_body_flags), C.c_void_p, C.c_int64,
                                      C.c_void_p, C.c_int64, C.c_void_p, C.c_int64]
        L.orc_islands.restype = C.c_int32
        L.orc_islands.argtypes = [C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_greedy_color.restype = None
        L.orc_greedy_color.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                       C.c_void_p, C.c_void_p]
        L.orc_color_valid.restype = C.c_int32
        L.orc_color_valid.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_int32]
        L.orc_levels.restype = None
        L.orc_levels.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                 C.c_void_p]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def cell_key(x, y, z, cell):
    return int(lib().orc_cell_key(x, y, z, cell))


def build_tier(aabbs: np.ndarray, include: np.ndarray, cell: float):
    """aabbs float32[n,6]; include bool[n] -> (entries [(key, body)], oversized list, inGrid bool[n])."""
    n = len(include)
    a = np.ascontiguousarray(aabbs, dtype=np.float32)
    inc = np.ascontiguousarray(include, dtype=np.uint8)
    ov = np.zeros(max(n, 1), dtype=np.int32)
    ovn = C.c_int32(0)
    ing = np.zeros(max(n, 1), dtype=np.uint8)
    cnt = lib().orc_build_tier(n, _p(a), _p(inc), cell, None, 0, _p(ov), C.byref(ovn), _p(ing))
    buf = (orc_entry * max(cnt, 1))()
    lib().orc_build_tier(n, _p(a), _p(inc), cell, buf, cnt, _p(ov), C.byref(ovn), _p(ing))
    return [(buf[i].key, buf[i].body) for i in range(cnt)], list(ov[: ovn.value]), ing[:n].astype(bool)


def brute_pairs(aabbs, flags, existing, jointed):
    n = len(flags)
    fl = (orc_body_flags * max(n, 1))()
    for i, f in enumerate(flags):
        for k, v in f.items():
            setattr(fl[i], k, v)
    a = np.ascontiguousarray(aabbs, dtype=np.float32)
    ex = np.ascontiguousarray(sorted(existing), dtype=np.uint64)
    jo = np.ascontiguousarray(sorted(jointed), dtype=np.uint64)
    cnt = lib().orc_brute_pairs(n, _p(a), fl, _p(ex), len(ex), _p(jo), len(jo), None, 0)
    out = np.zeros(max(cnt, 1), dtype=np.uint64)
    lib().orc_brute_pairs(n, _p(a), fl, _p(ex), len(ex), _p(jo), len(jo), _p(out), cnt)
    return [int(k) for k in out[:cnt]]


def islands(dynamic: np.ndarray, edges):
    n = len(dynamic)
    d = np.ascontiguousarray(dynamic, dtype=np.uint8)
    ea = np.ascontiguousarray([e[0] for e in edges] or [0], dtype=np.int32)
    eb = np.ascontiguousarray([e[1] for e in edges] or [0], dtype=np.int32)
    out = np.zeros(max(n, 1), dtype=np.int32)
    cnt = lib().orc_islands(n, _p(d), len(edges), _p(ea), _p(eb), _p(out))
    return int(cnt), out[:n]


def greedy_color(bodyA, bodyB, dynA, dynB, body_count):
    k = len(bodyA)
    a = np.ascontiguousarray(bodyA, dtype=np.int32)
    b = np.ascontiguousarray(bodyB, dtype=np.int32)
    da = np.ascontiguousarray(dynA, dtype=np.uint8)
    db = np.ascontiguousarray(dynB, dtype=np.uint8)
    mask = np.zeros(max(body_count, 1), dtype=np.uint32)
    out = np.zeros(max(k, 1), dtype=np.int32)
    lib().orc_greedy_color(k, _p(a), _p(b), _p(da), _p(db), body_count, _p(mask), _p(out))
    return out[:k]


def color_valid(bodyA, bodyB, dynA, dynB, colors, body_count):
    a = np.ascontiguousarray(bodyA, dtype=np.int32)
    b = np.ascontiguousarray(bodyB, dtype=np.int32)
    da = np.ascontiguousarray(dynA, dtype=np.uint8)
    db = np.ascontiguousarray(dynB, dtype=np.uint8)
    c = np.ascontiguousarray(colors, dtype=np.int32)
    return bool(lib().orc_color_valid(len(a), _p(a), _p(b), _p(da), _p(db), _p(c), body_count))


def levels(bodyA, bodyB, dynA, dynB, body_count):
    k = len(bodyA)
    a = np.ascontiguousarray(bodyA, dtype=np.int32)
    b = np.ascontiguousarray(bodyB, dtype=np.int32)
    da = np.ascontiguousarray(dynA, dtype=np.uint8)
    db = np.ascontiguousarray(dynB, dtype=np.uint8)
    last = np.zeros(max(body_count, 1), dtype=np.int32)
    out = np.zeros(max(k, 1), dtype=np.int32)
    lib().orc_levels(k, _p(a), _p(b), _p(da), _p(db), body_count, _p(last), _p(out))
    return out[:k]
