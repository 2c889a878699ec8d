"""Bulk despawn and spawn (m3dx_body_respawn) against the reference's
per-body calls (src/world.cpp:512-615).

The bulk call must leave exactly what the calls m3d_body_destroy(destroy[k])
then m3d_body_create(defs[k]) leave, in call order: the same ids (LIFO slot
reuse, generations), contacts in the same order, the same end events in the
same order, the same wakes and joint removals; every step after must then
match the reference as a world driven by single calls does.
"""
import ctypes as C

import pytest

from melon3d_amd import scenes
from melon3d_amd.api import m3d_ball_joint_def, m3d_body_def, m3d_body_id, m3d_shape_def, m3d_vec3
from parity import contact_pairs, events, state_diff

DT = 1.0 / 60.0
TOL = 1e-5


def _same(ref, rw, dev, dw, label):
    a, b = ref.export_state(rw), dev.export_state(dw)
    assert contact_pairs(a) == contact_pairs(b), f"{label}: contact order"
    assert events(ref, rw) == events(dev, dw), f"{label}: events"
    d = state_diff(a, b)
    assert d["pos"] <= TOL and d["rot"] <= TOL, f"{label}: {d}"
    assert d["awake_equal"], f"{label}: sleep state"
    return a, b


@pytest.mark.gpu
@pytest.mark.parametrize("idle_every", [0, 3])
def test_bulk_churn_lockstep(ref, dev, idle_every):
    """The streaming churn: per-body calls on the reference, one bulk call
    per step on the device (with steps that skip the churn when idle_every
    is set, so a bulk call also follows a step without despawns)."""
    rw, dw = ref.create_world(), dev.create_world()
    try:
        sr = scenes.Stream(ref, rw, count=1200, churn=12, arena=16.0)
        sd = scenes.Stream(dev, dw, count=1200, churn=12, arena=16.0)
        for step in range(45):
            if not idle_every or step % idle_every != 1:
                sr.churn_step()
                sd.churn_step(bulk=True)
                # ids equal the reference's, and end events are there before
                # the step clears them
                assert [(b.index, b.generation) for b in sr.ring] == [(b.index, b.generation) for b in sd.ring]
                assert events(ref, rw) == events(dev, dw), f"step {step}: despawn end events"
            ref.m3d_world_step(rw, DT, 4)
            dev.m3d_world_step(dw, DT, 4)
            _same(ref, rw, dev, dw, f"step {step}")
        assert dev.m3dx_step_failures(dw) == 0, dev.last_error()
    finally:
        ref.m3d_world_destroy(rw)
        dev.m3d_world_destroy(dw)


def _at(arr, k):
    """Address of element k of a ctypes array."""
    return C.c_void_p(C.addressof(arr) + k * C.sizeof(arr._type_))


def _defs(L, n, y0):
    bds = (m3d_body_def * n)()
    sds = (m3d_shape_def * n)()
    rng = scenes.Lcg(17)
    for k in range(n):
        bds[k], sds[k] = scenes.mixed_defs(L, rng, k % 3, (rng.uniform(-3, 3), y0 + 0.7 * k, rng.uniform(-3, 3)))
    return bds, sds


def _resting_scene(L, w):
    """A settled heap with a ball-jointed pair in it, stepped until most
    bodies sleep."""
    scenes.add_ground(L, w, half=(30.0, 0.5, 30.0), y=-0.5)
    ids = []
    for i in range(6):
        for k in range(6):
            for l in range(3):
                ids.append(scenes.add_box(L, w, (-3.0 + 1.05 * i, 0.5 + 1.0 * l, -3.0 + 1.05 * k), (0.5, 0.5, 0.5)))
    jd = m3d_ball_joint_def()
    jd.bodyA, jd.bodyB = ids[3], ids[4]
    jd.localAnchorA = m3d_vec3(0.0, 0.5, 0.0)
    jd.localAnchorB = m3d_vec3(0.0, -0.5, 0.0)
    L.m3d_joint_create_ball(w, C.byref(jd))
    return ids


@pytest.mark.gpu
def test_bulk_despawn_wakes_sleepers_and_mixes_with_single_calls(ref, dev):
    """Despawns inside a sleeping heap (partners wake, the jointed body's
    joint goes), stale and repeated ids, bodies created and destroyed within
    one call, and single calls before and after the bulk call in the same
    window between steps."""
    rw, dw = ref.create_world(), dev.create_world()
    try:
        ir = _resting_scene(ref, rw)
        idv = _resting_scene(dev, dw)
        for _ in range(150):
            ref.m3d_world_step(rw, DT, 4)
            dev.m3d_world_step(dw, DT, 4)
        a, b = _same(ref, rw, dev, dw, "settled")
        # (a snapshot's arrays are pointers: slice them with their counts)
        assert sum(not x.isAwake for x in a.bodies[:a.header.bodyCapacity] if x.inUse) > 50, "heap not asleep"
        # a single destroy before the bulk call
        ref.m3d_body_destroy(rw, ir[40])
        dev.m3d_body_destroy(dw, idv[40])
        n = 10
        bds_r, sds_r = _defs(ref, n, 8.0)
        bds_d, sds_d = _defs(dev, n, 8.0)
        # the jointed body, bodies deep in the heap, a stale id (the single
        # call's), a repeated id, and (k = 6) the body created at k = 5
        pick = [3, 20, 21, 40, 55, 0, 99, 70, 21, 90]
        gone_r = (m3d_body_id * n)(*[ir[p] for p in pick])
        gone_d = (m3d_body_id * n)(*[idv[p] for p in pick])
        born_r = (m3d_body_id * n)()
        born_d = (m3d_body_id * n)()
        for k in range(n):
            if k == 6:
                ref.m3d_body_destroy(rw, born_r[5])
            else:
                ref.m3d_body_destroy(rw, gone_r[k])
            born_r[k] = ref.m3d_body_create(rw, C.byref(bds_r[k]), C.byref(sds_r[k]))
        # the device does k = 0..5 and 7..9 in bulk and k = 6 (destroying
        # the body created at k = 5) as a second bulk call in between
        assert dev.m3dx_body_respawn(dw, gone_d, bds_d, sds_d, 6, born_d) == 0, dev.last_error()
        six = (m3d_body_id * 1)(born_d[5])
        one = (m3d_body_id * 1)()
        assert dev.m3dx_body_respawn(dw, six, _at(bds_d, 6), _at(sds_d, 6), 1, one) == 0
        born_d[6] = one[0]
        rest = (m3d_body_id * 3)(*[gone_d[k] for k in range(7, 10)])
        out = (m3d_body_id * 3)()
        assert dev.m3dx_body_respawn(dw, rest, _at(bds_d, 7), _at(sds_d, 7), 3, out) == 0
        for j in range(3):
            born_d[7 + j] = out[j]
        assert [(x.index, x.generation) for x in born_r] == [(x.index, x.generation) for x in born_d]
        # a single destroy after the bulk calls
        ref.m3d_body_destroy(rw, ir[100])
        dev.m3d_body_destroy(dw, idv[100])
        assert events(ref, rw) == events(dev, dw), "despawn end events"
        for step in range(60):
            ref.m3d_world_step(rw, DT, 4)
            dev.m3d_world_step(dw, DT, 4)
            _same(ref, rw, dev, dw, f"after despawn, step {step}")
        assert dev.m3dx_step_failures(dw) == 0, dev.last_error()
    finally:
        ref.m3d_world_destroy(rw)
        dev.m3d_world_destroy(dw)


@pytest.mark.gpu
def test_native_stream_bulk_equals_single_calls(dev):
    """The benchmark's native churn in bulk on the device equals its
    per-body churn on the device (same contacts, same states)."""
    a, b = dev.create_world(), dev.create_world()
    try:
        x = scenes.NativeStream(dev, a, count=20000, churn=100, arena=60.0)
        y = scenes.NativeStream(dev, b, count=20000, churn=100, arena=60.0, bulk=True)
        for _ in range(30):
            x.churn_step()
            y.churn_step()
            dev.m3d_world_step(a, DT, 4)
            dev.m3d_world_step(b, DT, 4)
        A, B = dev.export_state(a), dev.export_state(b)
        assert contact_pairs(A) == contact_pairs(B)
        d = state_diff(A, B)
        assert d["pos"] == 0.0 and d["rot"] <= 1e-12, d
    finally:
        dev.m3d_world_destroy(a)
        dev.m3d_world_destroy(b)


def test_bulk_respawn_is_the_single_call_sequence_on_the_reference(ref):
    """On the reference build the bulk entry point is the per-call loop."""
    a, b = ref.create_world(), ref.create_world()
    try:
        sa = scenes.Stream(ref, a, count=300, churn=20, arena=8.0)
        sb = scenes.Stream(ref, b, count=300, churn=20, arena=8.0)
        for _ in range(5):
            sa.churn_step()
            sb.churn_step(bulk=True)
            ref.m3d_world_step(a, DT, 4)
            ref.m3d_world_step(b, DT, 4)
        assert [(r.index, r.generation) for r in sa.ring] == [(r.index, r.generation) for r in sb.ring]
        assert contact_pairs(ref.export_state(a)) == contact_pairs(ref.export_state(b))
    finally:
        ref.m3d_world_destroy(a)
        ref.m3d_world_destroy(b)
