"""m3dx_world_ray_cast_batch_noalloc: the allocation-free variant of
m3dx_world_ray_cast_batch (staging in the library's fixed per-GPU array)
returns exactly what the allocating entry point returns, on two calls in a
row (the second smaller, so stale slots of the first call would show), and
refuses what it cannot do without allocating."""
import ctypes as C

import numpy as np
import pytest

from melon3d_amd import scenes
from melon3d_amd.api import m3d_ray_result, m3d_vec3

DT = 1.0 / 60.0
MAX_RAYS = 65536


def _arrays(o, d, masks):
    n = len(o)
    O = (m3d_vec3 * n)(*[m3d_vec3(*map(float, p)) for p in o])
    D = (m3d_vec3 * n)(*[m3d_vec3(*map(float, p)) for p in d])
    M = (C.c_uint32 * n)(*masks) if masks is not None else None
    return O, D, M


def _rays(rng, n):
    o = np.column_stack([rng.uniform(-10, 10, n), rng.uniform(-2, 25, n), rng.uniform(-10, 10, n)])
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d *= rng.uniform(0.5, 40.0, n)[:, None]
    return o.astype(np.float32), d.astype(np.float32)


def _rows(out, n):
    return [(bool(out[k].hit), out[k].body.index, out[k].body.generation, np.float32(out[k].fraction),
             (out[k].point.x, out[k].point.y, out[k].point.z), (out[k].normal.x, out[k].normal.y, out[k].normal.z))
            for k in range(n)]


@pytest.mark.gpu
def test_noalloc_ray_cast_matches_allocating_entry_point(dev):
    w = dev.create_world()
    try:
        scenes.rain_pile(dev, w, count=1500, column=(12.0, 12.0, 20.0))
        for _ in range(40):
            dev.m3d_world_step(w, DT, 4)
        rng = np.random.default_rng(5)
        for n, masked in ((5000, True), (1200, False)):
            o, d = _rays(rng, n)
            masks = [int(x) for x in rng.choice([0xFFFFFFFF, 1, 0], size=n)] if masked else None
            O, D, M = _arrays(o, d, masks)
            a = (m3d_ray_result * n)()
            b = (m3d_ray_result * n)()
            # the allocating call refreshes a dirty broadphase first, as the
            # reference's ray cast does; the allocation-free one needs it current
            assert dev.m3dx_world_ray_cast_batch(w, O, D, M, n, a) == 0, dev.last_error()
            assert dev.m3dx_world_ray_cast_batch_noalloc(w, O, D, M, n, b) == 0, dev.last_error()
            ra, rb = _rows(a, n), _rows(b, n)
            assert ra == rb
            assert sum(r[0] for r in ra) > n // 10
            dev.m3d_world_step(w, DT, 4)
    finally:
        dev.m3d_world_destroy(w)


@pytest.mark.gpu
def test_noalloc_ray_cast_refusals(dev):
    w = dev.create_world()
    try:
        scenes.rain_pile(dev, w, count=200, column=(6.0, 6.0, 8.0))
        dev.m3d_world_step(w, DT, 4)
        one = (m3d_vec3 * 1)(m3d_vec3(0.0, 20.0, 0.0))
        down = (m3d_vec3 * 1)(m3d_vec3(0.0, -30.0, 0.0))
        out = (m3d_ray_result * 1)()
        # a pending API change would need a broadphase refresh
        scenes.add_box(dev, w, (0.0, 30.0, 0.0), (0.5, 0.5, 0.5))
        assert dev.m3dx_world_ray_cast_batch_noalloc(w, one, down, None, 1, out) == -1
        assert "not current" in dev.last_error()
        assert dev.m3dx_world_ray_cast_batch(w, one, down, None, 1, out) == 0
        assert dev.m3dx_world_ray_cast_batch_noalloc(w, one, down, None, 1, out) == 0
        assert out[0].hit
        # beyond the fixed maximum: refused before anything is touched
        n = MAX_RAYS + 1
        big = (m3d_vec3 * n)()
        bigout = (m3d_ray_result * n)()
        assert dev.m3dx_world_ray_cast_batch_noalloc(w, big, big, None, n, bigout) == -1
        assert "maximum" in dev.last_error()
        assert dev.m3dx_world_ray_cast_batch_noalloc(w, one, down, None, 0, out) == 0
    finally:
        dev.m3d_world_destroy(w)
