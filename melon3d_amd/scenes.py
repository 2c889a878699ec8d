"""Scene builders for the benchmark workloads and parity tests.

All scenes are built through the C API, so the same builder drives the
device backend and the reference build. Random draws use the reference
benchmark's LCG (bench/bench.cpp:121-125): s = s*1664525 + 1013904223,
u = (s >> 8) / 0xFFFFFF, with a fixed per-body draw order documented at
each builder.
"""
from __future__ import annotations

import ctypes as C
import math

from .api import (M3D_BODY_DYNAMIC, M3D_BODY_STATIC, M3D_SHAPE_BOX, M3D_SHAPE_CAPSULE, M3D_SHAPE_SPHERE, Library,
                  m3d_body_def, m3d_body_id, m3d_quat, m3d_shape_def, m3d_vec3)


class Lcg:
    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFF

    def u(self) -> float:
        self.s = (self.s * 1664525 + 1013904223) & 0xFFFFFFFF
        return float(self.s >> 8) / float(0x00FFFFFF)

    def uniform(self, a: float, b: float) -> float:
        return a + (b - a) * self.u()

    def quat(self) -> m3d_quat:
        # Shoemake's uniform random rotation from three uniforms
        u1, u2, u3 = self.u(), self.u(), self.u()
        a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
        return m3d_quat(a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
                        b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3))

    def normal(self) -> float:
        # Box-Muller
        u1 = max(self.u(), 1e-12)
        u2 = self.u()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2 * math.pi * u2)


def add_ground(L: Library, w, half=(50.0, 1.0, 50.0), y=-1.0, friction=0.6):
    bd = L.m3d_body_def_default()
    bd.type = M3D_BODY_STATIC
    bd.position = m3d_vec3(0.0, y, 0.0)
    sd = L.m3d_shape_def_default()
    sd.type = M3D_SHAPE_BOX
    sd.halfExtents = m3d_vec3(*half)
    sd.friction = friction
    return L.m3d_body_create(w, C.byref(bd), C.byref(sd))


def add_box(L: Library, w, pos, half, friction=0.6, **body):
    bd = L.m3d_body_def_default()
    bd.type = M3D_BODY_DYNAMIC
    bd.position = m3d_vec3(*pos)
    for k, v in body.items():
        setattr(bd, k, v)
    sd = L.m3d_shape_def_default()
    sd.type = M3D_SHAPE_BOX
    sd.halfExtents = m3d_vec3(*half)
    sd.friction = friction
    return L.m3d_body_create(w, C.byref(bd), C.byref(sd))


def pyramid(L: Library, w, base=20, he=0.5, friction=0.6):
    """Box pyramid, base*(base+1)/2 boxes in one island (bench/bench.cpp:48-72).
    Ground is 100 x 1 x 100 m full extent with its top face at y = 0."""
    add_ground(L, w, half=(50.0, 0.5, 50.0), y=-0.5, friction=friction)
    ids = []
    for row in range(base):
        n = base - row
        y = he + 2.0 * he * row
        x0 = -he * (n - 1)
        for i in range(n):
            ids.append(add_box(L, w, (x0 + 2.0 * he * i, y, 0.0), (he, he, he), friction))
    return ids


def mixed_body(L: Library, w, rng: Lcg, kind: int, pos, vel=(0.0, 0.0, 0.0)):
    """One rain-pile body. Draw order: size draws for the kind, then a
    quaternion (3 draws), then friction (1 draw)."""
    bd, sd = mixed_defs(L, rng, kind, pos, vel)
    return L.m3d_body_create(w, C.byref(bd), C.byref(sd))


def mixed_defs(L: Library, rng: Lcg, kind: int, pos, vel=(0.0, 0.0, 0.0)):
    """The body and shape definitions of mixed_body (same draws)."""
    bd = L.m3d_body_def_default()
    bd.type = M3D_BODY_DYNAMIC
    bd.position = m3d_vec3(*pos)
    bd.linearVelocity = m3d_vec3(*vel)
    sd = L.m3d_shape_def_default()
    if kind == 0:
        sd.type = M3D_SHAPE_SPHERE
        sd.radius = rng.uniform(0.2, 0.5)
    elif kind == 1:
        sd.type = M3D_SHAPE_CAPSULE
        sd.radius = rng.uniform(0.15, 0.3)
        sd.halfHeight = rng.uniform(0.2, 0.6)
    else:
        sd.type = M3D_SHAPE_BOX
        sd.halfExtents = m3d_vec3(rng.uniform(0.2, 0.6), rng.uniform(0.2, 0.6), rng.uniform(0.2, 0.6))
    bd.orientation = rng.quat()
    sd.friction = rng.uniform(0.3, 0.8)
    return bd, sd


def rain_pile(L: Library, w, count=20000, seed=12345, column=(40.0, 40.0, 60.0)):
    """Rain pile: kind = i % 3 (sphere, capsule, box), positions uniform in a
    column x*z*y above the ground starting 1 m up; per body the position
    draws (x, y, z) come first, then mixed_body's draws."""
    gx, gz, gy = column
    add_ground(L, w, half=(max(50.0, gx), 1.0, max(50.0, gz)), y=-1.0)
    rng = Lcg(seed)
    ids = []
    for i in range(count):
        x = rng.uniform(-0.5 * gx, 0.5 * gx)
        y = 1.0 + rng.uniform(0.0, gy)
        z = rng.uniform(-0.5 * gz, 0.5 * gz)
        ids.append(mixed_body(L, w, rng, i % 3, (x, y, z)))
    return ids


class Stream:
    """Streaming scene: bodies in a 200 x 200 m arena (uniform in a
    200 x 50 x 200 m volume), each step destroys the `churn` oldest (FIFO ring,
    bench/bench.cpp:209-251) and spawns as many at y in [30, 60] with
    downward speed U[5, 20] m/s."""

    def __init__(self, L: Library, w, count=200000, churn=1000, seed=777, arena=200.0):
        self.L, self.w, self.churn, self.arena = L, w, churn, arena
        add_ground(L, w, half=(0.5 * arena + 10.0, 1.0, 0.5 * arena + 10.0), y=-1.0)
        self.rng = Lcg(seed)
        self.ring = []
        for i in range(count):
            x = self.rng.uniform(-0.5 * arena, 0.5 * arena)
            y = 1.0 + self.rng.uniform(0.0, 50.0)
            z = self.rng.uniform(-0.5 * arena, 0.5 * arena)
            self.ring.append(mixed_body(L, w, self.rng, i % 3, (x, y, z)))
        self.cursor = 0

    def churn_step(self, bulk: bool = False):
        """Destroy the churn oldest and spawn as many; bulk: as one
        m3dx_body_respawn call with the same draws."""
        L, w, rng = self.L, self.w, self.rng
        n = len(self.ring)
        if bulk:
            k = self.churn
            gone = (m3d_body_id * k)()
            bds = (m3d_body_def * k)()
            sds = (m3d_shape_def * k)()
            born = (m3d_body_id * k)()
            cur = self.cursor
            for j in range(k):
                gone[j] = self.ring[cur]
                x = rng.uniform(-0.5 * self.arena, 0.5 * self.arena)
                y = rng.uniform(30.0, 60.0)
                z = rng.uniform(-0.5 * self.arena, 0.5 * self.arena)
                vy = -rng.uniform(5.0, 20.0)
                bds[j], sds[j] = mixed_defs(L, rng, cur % 3, (x, y, z), (0.0, vy, 0.0))
                cur = (cur + 1) % n
            if L.m3dx_body_respawn(w, gone, bds, sds, k, born) != 0:
                raise RuntimeError("m3dx_body_respawn failed: " + L.last_error())
            for j in range(k):
                self.ring[self.cursor] = born[j]
                self.cursor = (self.cursor + 1) % n
            return
        for _ in range(self.churn):
            L.m3d_body_destroy(w, self.ring[self.cursor])
            x = rng.uniform(-0.5 * self.arena, 0.5 * self.arena)
            y = rng.uniform(30.0, 60.0)
            z = rng.uniform(-0.5 * self.arena, 0.5 * self.arena)
            vy = -rng.uniform(5.0, 20.0)
            self.ring[self.cursor] = mixed_body(L, w, rng, self.cursor % 3, (x, y, z), (0.0, vy, 0.0))
            self.cursor = (self.cursor + 1) % n


# ------------------------------------------------------------ native driver

class _SceneApi(C.Structure):
    _fields_ = [("body_def_default", C.c_void_p), ("shape_def_default", C.c_void_p),
                ("body_create", C.c_void_p), ("body_destroy", C.c_void_p), ("respawn", C.c_void_p)]


_driver = None


def _scene_driver():
    """libm3d_scenes.so (built by melon3d_amd/build.py)."""
    global _driver
    if _driver is None:
        import os
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libm3d_scenes.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing; run python melon3d_amd/build.py")
        d = C.CDLL(path)
        d.m3ds_rain.argtypes = [C.POINTER(_SceneApi), C.c_void_p, C.c_int, C.c_uint32, C.c_float, C.c_float, C.c_float]
        d.m3ds_rain.restype = C.c_int
        d.m3ds_stream_create.argtypes = [C.POINTER(_SceneApi), C.c_void_p, C.c_int, C.c_int, C.c_uint32, C.c_float]
        d.m3ds_stream_create.restype = C.c_void_p
        d.m3ds_stream_churn.argtypes = [C.c_void_p]
        d.m3ds_stream_churn.restype = None
        d.m3ds_stream_clone.argtypes = [C.c_void_p, C.POINTER(_SceneApi), C.c_void_p]
        d.m3ds_stream_clone.restype = C.c_void_p
        d.m3ds_batch_world.argtypes = [C.POINTER(_SceneApi), C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        d.m3ds_batch_world.restype = C.c_int
        d.m3ds_stream_free.argtypes = [C.c_void_p]
        d.m3ds_stream_free.restype = None
        _driver = d
    return _driver


def _api_table(L: Library, bulk: bool = False) -> _SceneApi:
    """Function table of a library for the native driver; bulk: churn through
    m3dx_body_respawn (one call per churn step) instead of one call per body."""
    t = _SceneApi()
    for name, fn in (("body_def_default", "m3d_body_def_default"), ("shape_def_default", "m3d_shape_def_default"),
                     ("body_create", "m3d_body_create"), ("body_destroy", "m3d_body_destroy")):
        setattr(t, name, C.cast(getattr(L.lib, fn), C.c_void_p).value)
    t.respawn = C.cast(L.lib.m3dx_body_respawn, C.c_void_p).value if bulk else None
    return t


def _world_ptr(w) -> int:
    return C.cast(w, C.c_void_p).value


def native_rain_pile(L: Library, w, count=20000, seed=12345, column=(40.0, 40.0, 60.0)) -> int:
    """rain_pile built in native code (same draws, same bodies)."""
    gx, gz, gy = column
    t = _api_table(L)
    return _scene_driver().m3ds_rain(C.byref(t), _world_ptr(w), count, seed, gx, gz, gy)


class NativeStream:
    """Stream with its per-step churn in native code (as the reference's own
    bench does it in C++)."""

    def __init__(self, L: Library, w, count=200000, churn=1000, seed=777, arena=200.0, _handle=None,
                 bulk: bool = False):
        """bulk: each churn step is one m3dx_body_respawn call (the same
        bodies, ids and results as the per-body calls)."""
        self._table = _api_table(L, bulk)
        self._h = _handle if _handle is not None else _scene_driver().m3ds_stream_create(
            C.byref(self._table), _world_ptr(w), count, churn, seed, arena)
        if not self._h:
            raise RuntimeError("stream driver allocation failed")

    def churn_step(self):
        _scene_driver().m3ds_stream_churn(self._h)

    def clone_onto(self, L: Library, w, bulk: bool = False) -> "NativeStream":
        """Continue this stream on a world imported from this one's snapshot."""
        t = _api_table(L, bulk)
        h = _scene_driver().m3ds_stream_clone(self._h, C.byref(t), _world_ptr(w))
        s = NativeStream.__new__(NativeStream)
        s._table, s._h = t, h
        return s

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _driver is not None:
            _driver.m3ds_stream_free(h)
            self._h = None


# ------------------------------------------------------------ batched worlds

def batch_world(L: Library, w, index: int):
    """World `index` of the batched workload built into world `w` (a plain
    world, or a batch with that sub-world selected). Returns (dynamic body
    count, is_pile)."""
    t = _api_table(L)
    kind = C.c_int(0)
    n = _scene_driver().m3ds_batch_world(C.byref(t), _world_ptr(w), index, C.byref(kind))
    return n, bool(kind.value)


def batch(L: Library, worlds: int, first: int = 0, wd=None):
    """A batch world holding workload worlds first .. first + worlds - 1 as
    sub-worlds 0 .. worlds - 1. Returns (world, [(count, is_pile)])."""
    if wd is None:
        wd = L.m3d_world_def_default()
    w = L.m3dx_batch_create(C.byref(wd), worlds)
    if not w:
        raise RuntimeError("m3dx_batch_create failed: " + L.last_error())
    t = _api_table(L)
    d = _scene_driver()
    info = []
    kind = C.c_int(0)
    for s in range(worlds):
        if L.m3dx_batch_select(w, s) != 0:
            raise RuntimeError(L.last_error())
        n = d.m3ds_batch_world(C.byref(t), _world_ptr(w), first + s, C.byref(kind))
        info.append((n, bool(kind.value)))
    return w, info
