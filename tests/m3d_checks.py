"""The Melon3D behavioural test suite restated over the C API.

Each scenario follows the reference test of the same name
(test/test_main.cpp, cited per function) with the same bodies, steps and
pass thresholds; each returns a list of (check message, passed). The same
functions run against the reference build (CPU, pins the restatement) and
against the MI355X backend (GPU). The reference's thread-count determinism
checks become run-to-run determinism checks (the device backend has no
worker count).
"""
from __future__ import annotations

import ctypes as C
import math

from melon3d_amd.api import (M3D_BODY_DYNAMIC, M3D_BODY_STATIC, M3D_PI, M3D_SHAPE_BOX, M3D_SHAPE_CAPSULE,
                             M3D_SHAPE_SPHERE, m3d_aabb, m3d_ball_joint_def, m3d_distance_joint_def, m3d_vec3,
                             overlap_fn)

DT = 1.0 / 60.0


def v3(x, y, z):
    return m3d_vec3(x, y, z)


def vlen(v):
    return math.sqrt(v.x * v.x + v.y * v.y + v.z * v.z)


def vsub(a, b):
    return v3(a.x - b.x, a.y - b.y, a.z - b.z)


def rotate(q, v):
    # the header's m3d_rotate, in float64 (only used for measurements)
    ux, uy, uz = q.x, q.y, q.z
    tx, ty, tz = 2 * (uy * v.z - uz * v.y), 2 * (uz * v.x - ux * v.z), 2 * (ux * v.y - uy * v.x)
    return v3(v.x + q.w * tx + (uy * tz - uz * ty), v.y + q.w * ty + (uz * tx - ux * tz),
              v.z + q.w * tz + (ux * ty - uy * tx))


def world(L, **wdef):
    wd = L.m3d_world_def_default()
    for k, v in wdef.items():
        setattr(wd, k, v)
    return L.create_world(wd)


def ground(L, w):
    # test/test_main.cpp:23-32
    bd = L.m3d_body_def_default()
    bd.type = M3D_BODY_STATIC
    bd.position = v3(0.0, -1.0, 0.0)
    sd = L.m3d_shape_def_default()
    sd.type = M3D_SHAPE_BOX
    sd.halfExtents = v3(50.0, 1.0, 50.0)
    return L.m3d_body_create(w, C.byref(bd), C.byref(sd))


def box(L, w, pos, half, friction):
    # test/test_main.cpp:34-44
    bd = L.m3d_body_def_default()
    bd.type = M3D_BODY_DYNAMIC
    bd.position = pos
    sd = L.m3d_shape_def_default()
    sd.type = M3D_SHAPE_BOX
    sd.halfExtents = half
    sd.friction = friction
    return L.m3d_body_create(w, C.byref(bd), C.byref(sd))


def body(L, w, btype, pos, shape, **kw):
    bd = L.m3d_body_def_default()
    bd.type = btype
    bd.position = pos
    sd = L.m3d_shape_def_default()
    sd.type = shape
    for k, v in kw.items():
        if hasattr(bd, k) and k not in ("type",):
            setattr(bd, k, v)
        else:
            setattr(sd, k, v)
    return L.m3d_body_create(w, C.byref(bd), C.byref(sd))


def steps(L, w, n, dt=DT, sub=4):
    for _ in range(n):
        L.m3d_world_step(w, dt, sub)


# ----------------------------------------------------------------- tests


def sphere_rest(L):  # test/test_main.cpp:47-71
    w = world(L)
    ground(L, w)
    ball = body(L, w, M3D_BODY_DYNAMIC, v3(0, 3, 0), M3D_SHAPE_SPHERE, radius=0.5)
    steps(L, w, 240)
    p = L.m3d_body_position(w, ball)
    v = L.m3d_body_linear_velocity(w, ball)
    L.m3d_world_destroy(w)
    return [("sphere rests at y=radius", abs(p.y - 0.5) < 0.02),
            ("sphere does not drift", abs(p.x) < 0.01 and abs(p.z) < 0.01),
            ("sphere at rest has ~zero velocity", vlen(v) < 0.05)]


def capsule_rest(L):  # test/test_main.cpp:74-97
    w = world(L)
    ground(L, w)
    bd = L.m3d_body_def_default()
    bd.type = M3D_BODY_DYNAMIC
    bd.position = v3(0, 2, 0)
    bd.orientation = L.quat_axis_angle(v3(0, 0, 1), 0.5 * M3D_PI)
    sd = L.m3d_shape_def_default()
    sd.type = M3D_SHAPE_CAPSULE
    sd.radius = 0.3
    sd.halfHeight = 0.5
    cap = L.m3d_body_create(w, C.byref(bd), C.byref(sd))
    steps(L, w, 240)
    p = L.m3d_body_position(w, cap)
    L.m3d_world_destroy(w)
    return [("capsule rests at y=radius", abs(p.y - 0.3) < 0.02)]


def box_stack(L):  # test/test_main.cpp:100-131
    w = world(L)
    ground(L, w)
    boxes = [box(L, w, v3(0, 0.5 + i, 0), v3(0.5, 0.5, 0.5), 0.6) for i in range(10)]
    steps(L, w, 300)
    ok = True
    for i, b in enumerate(boxes):
        p = L.m3d_body_position(w, b)
        if abs(p.y - (0.5 + i)) > 0.1 or abs(p.x) > 0.2 or abs(p.z) > 0.2:
            ok = False
    L.m3d_world_destroy(w)
    return [("10-box stack remains standing after 5s", ok)]


def restitution(L):  # test/test_main.cpp:134-169
    w = world(L)
    ground(L, w)
    ball = body(L, w, M3D_BODY_DYNAMIC, v3(0, 2, 0), M3D_SHAPE_SPHERE, radius=0.5, restitution=0.8)
    max_h, bounced = 0.0, False
    for _ in range(240):
        L.m3d_world_step(w, DT, 4)
        p = L.m3d_body_position(w, ball)
        v = L.m3d_body_linear_velocity(w, ball)
        if v.y > 0.5:
            bounced = True
        if bounced and p.y > max_h:
            max_h = p.y
    L.m3d_world_destroy(w)
    return [("restitution 0.8 ball bounces", bounced), ("bounce reaches a reasonable height", max_h > 1.0)]


def sleep_and_wake(L):  # test/test_main.cpp:172-191
    w = world(L)
    ground(L, w)
    b = box(L, w, v3(0, 0.5, 0), v3(0.5, 0.5, 0.5), 0.6)
    steps(L, w, 180)
    r = [("resting box falls asleep", not L.m3d_body_is_awake(w, b))]
    L.m3d_body_apply_impulse(w, b, v3(0, 500, 0), L.m3d_body_position(w, b))
    r.append(("impulse wakes the box", L.m3d_body_is_awake(w, b)))
    L.m3d_world_step(w, DT, 4)
    r.append(("woken box moves upward", L.m3d_body_linear_velocity(w, b).y > 0.0))
    L.m3d_world_destroy(w)
    return r


def ray_cast(L):  # test/test_main.cpp:194-210
    w = world(L)
    ground(L, w)
    b = box(L, w, v3(0, 5, 0), v3(0.5, 0.5, 0.5), 0.6)
    r = L.m3d_world_ray_cast(w, v3(0, 10, 0), v3(0, -20, 0))
    miss = L.m3d_world_ray_cast(w, v3(100, 10, 100), v3(0, 1, 0))
    L.m3d_world_destroy(w)
    return [("downward ray hits", r.hit), ("ray hits the box first", r.body.index == b.index),
            ("ray hit point on top face", abs(r.point.y - 5.5) < 0.01), ("ray hit normal points up", r.normal.y > 0.99),
            ("ray away from all bodies misses", not miss.hit)]


def distance_joint(L):  # test/test_main.cpp:213-257
    w = world(L)
    anchor = body(L, w, M3D_BODY_STATIC, v3(0, 5, 0), M3D_SHAPE_SPHERE, radius=0.1)
    bob = body(L, w, M3D_BODY_DYNAMIC, v3(2, 5, 0), M3D_SHAPE_SPHERE, radius=0.25)
    jd = m3d_distance_joint_def()
    jd.bodyA, jd.bodyB, jd.length = anchor, bob, 2.0
    L.m3d_joint_create_distance(w, C.byref(jd))
    max_err = 0.0
    for _ in range(300):
        L.m3d_world_step(w, DT, 4)
        p = L.m3d_body_position(w, bob)
        max_err = max(max_err, abs(vlen(vsub(p, v3(0, 5, 0))) - 2.0))
    L.m3d_world_destroy(w)
    return [("distance joint keeps pendulum length within 5cm", max_err < 0.05)]


def ball_joint(L):  # test/test_main.cpp:260-304
    w = world(L)
    anchor = body(L, w, M3D_BODY_STATIC, v3(0, 5, 0), M3D_SHAPE_SPHERE, radius=0.1)
    rod = body(L, w, M3D_BODY_DYNAMIC, v3(0.5, 5, 0), M3D_SHAPE_BOX, halfExtents=v3(0.5, 0.1, 0.1))
    jd = m3d_ball_joint_def()
    jd.bodyA, jd.bodyB, jd.localAnchorB = anchor, rod, v3(-0.5, 0, 0)
    L.m3d_joint_create_ball(w, C.byref(jd))
    max_err = 0.0
    for _ in range(300):
        L.m3d_world_step(w, DT, 4)
        xf = L.m3d_body_transform(w, rod)
        rp = rotate(xf.q, v3(-0.5, 0, 0))
        pinned = v3(rp.x + xf.p.x, rp.y + xf.p.y, rp.z + xf.p.z)
        max_err = max(max_err, vlen(vsub(pinned, v3(0, 5, 0))))
    L.m3d_world_destroy(w)
    return [("ball joint keeps anchors pinned within 5cm", max_err < 0.05)]


def contact_events(L):  # test/test_main.cpp:307-330
    w = world(L)
    ground(L, w)
    b = box(L, w, v3(0, 2, 0), v3(0.5, 0.5, 0.5), 0.6)
    saw = False
    for _ in range(120):
        L.m3d_world_step(w, DT, 4)
        if L.m3d_world_contact_events(w).beginCount > 0:
            saw = True
    L.m3d_body_destroy(w, b)
    valid = L.m3d_body_is_valid(w, b)
    L.m3d_world_step(w, DT, 4)
    L.m3d_world_destroy(w)
    return [("begin-touch event fires on landing", saw), ("destroyed body id is invalid", not valid)]


def speculative_no_tunnel(L):  # test/test_main.cpp:334-364
    w = world(L)
    body(L, w, M3D_BODY_STATIC, v3(10, 0, 0), M3D_SHAPE_BOX, halfExtents=v3(0.05, 5, 5))
    bullet = body(L, w, M3D_BODY_DYNAMIC, v3(0, 0, 0), M3D_SHAPE_SPHERE, radius=0.2,
                  linearVelocity=v3(50, 0, 0), gravityScale=0.0)
    steps(L, w, 60)
    p = L.m3d_body_position(w, bullet)
    L.m3d_world_destroy(w)
    return [("fast sphere stopped by thin wall (no tunneling)", p.x < 10.0)]


def _towers_run(L):
    w = world(L)
    ground(L, w)
    top = None
    for i in range(5):
        for j in range(5):
            top = box(L, w, v3(2.0 * i - 4.0, 0.5 + j, 0), v3(0.45, 0.45, 0.45), 0.6)
    steps(L, w, 120)
    p = L.m3d_body_position(w, top)
    L.m3d_world_destroy(w)
    return p.tuple()


def determinism(L):  # test/test_main.cpp:367-398 (run-to-run on the device backend)
    return [("two runs are bit-identical", _towers_run(L) == _towers_run(L))]


def collision_filter(L):  # test/test_main.cpp:402-440
    w = world(L)
    ground(L, w)
    bd = L.m3d_body_def_default()
    bd.type = M3D_BODY_DYNAMIC
    bd.position = v3(0, 4, 0)
    sd = L.m3d_shape_def_default()
    sd.type = M3D_SHAPE_SPHERE
    sd.radius = 0.5
    sd.filter.groupIndex = -7
    upper = L.m3d_body_create(w, C.byref(bd), C.byref(sd))
    bd.position = v3(0, 1, 0)
    lower = L.m3d_body_create(w, C.byref(bd), C.byref(sd))
    bd.position = v3(5, 2, 0)
    gh = L.m3d_shape_def_default()
    gh.type = M3D_SHAPE_BOX
    gh.halfExtents = v3(0.4, 0.4, 0.4)
    gh.filter.maskBits = 0x2
    ghost = L.m3d_body_create(w, C.byref(bd), C.byref(gh))
    steps(L, w, 180)
    pu, pl = L.m3d_body_position(w, upper), L.m3d_body_position(w, lower)
    pg = L.m3d_body_position(w, ghost)
    L.m3d_world_destroy(w)
    return [("same negative group: bodies interpenetrate", abs(pu.y - 0.5) < 0.05 and abs(pl.y - 0.5) < 0.05),
            ("disjoint masks: body ignores the ground", pg.y < -5.0)]


def sensor(L):  # test/test_main.cpp:443-491
    w = world(L)
    ground(L, w)
    body(L, w, M3D_BODY_STATIC, v3(0, 2, 0), M3D_SHAPE_BOX, halfExtents=v3(1, 1, 1), isSensor=True)
    ball = body(L, w, M3D_BODY_DYNAMIC, v3(0, 6, 0), M3D_SHAPE_SPHERE, radius=0.3)
    nb = ne = 0
    for _ in range(240):
        L.m3d_world_step(w, DT, 4)
        ev = L.m3d_world_contact_events(w)
        nb += sum(1 for k in range(ev.beginCount) if ev.beginEvents[k].isSensor)
        ne += sum(1 for k in range(ev.endCount) if ev.endEvents[k].isSensor)
    p = L.m3d_body_position(w, ball)
    L.m3d_world_destroy(w)
    return [("sensor reported a begin event", nb >= 1), ("sensor reported an end event after the body passed", ne >= 1),
            ("sensor exerted no force: ball fell through the zone to the ground", abs(p.y - 0.3) < 0.05)]


def forces_and_torques(L):  # test/test_main.cpp:495-531
    w = world(L, gravity=v3(0, 0, 0))
    b = body(L, w, M3D_BODY_DYNAMIC, v3(0, 0, 0), M3D_SHAPE_BOX, halfExtents=v3(0.5, 0.5, 0.5), density=1000.0,
             angularDamping=0.0)
    for _ in range(60):
        L.m3d_body_apply_force(w, b, v3(2000, 0, 0))
        L.m3d_world_step(w, DT, 4)
    r = [("apply_force: v = a*t after 1 s", abs(L.m3d_body_linear_velocity(w, b).x - 2.0) < 0.05)]
    for _ in range(60):
        L.m3d_body_apply_torque(w, b, v3(0, 500, 0))
        L.m3d_world_step(w, DT, 4)
    r.append(("apply_torque spins the body", L.m3d_body_angular_velocity(w, b).y > 0.5))
    v0 = L.m3d_body_linear_velocity(w, b)
    L.m3d_world_step(w, DT, 4)
    v1 = L.m3d_body_linear_velocity(w, b)
    r.append(("force does not persist past the step", abs(v1.x - v0.x) < 1e-6))
    L.m3d_world_destroy(w)
    return r


def hinge_joint(L):  # test/test_main.cpp:535-618
    w = world(L)
    pivot = body(L, w, M3D_BODY_STATIC, v3(0, 5, 0), M3D_SHAPE_BOX, halfExtents=v3(0.1, 0.1, 0.1))
    bd = L.m3d_body_def_default()
    bd.type = M3D_BODY_DYNAMIC
    bd.position = v3(1, 5, 0)
    sd = L.m3d_shape_def_default()
    sd.type = M3D_SHAPE_BOX
    sd.halfExtents = v3(0.4, 0.1, 0.1)
    arm = L.m3d_body_create(w, C.byref(bd), C.byref(sd))
    hd = L.m3d_hinge_joint_def_default()
    hd.bodyA, hd.bodyB = pivot, arm
    hd.localAnchorA = v3(0, 0, 0)
    hd.localAnchorB = v3(-1, 0, 0)
    hd.localAxisA = v3(0, 0, 1)
    hinge = L.m3d_joint_create_hinge(w, C.byref(hd))
    max_z = 0.0
    for _ in range(240):
        L.m3d_world_step(w, DT, 4)
        max_z = max(max_z, abs(L.m3d_body_position(w, arm).z))
    r = [("hinge pendulum stays in the hinge plane", max_z < 0.02)]
    p = L.m3d_body_position(w, arm)
    r.append(("hinge point constraint holds the arm at the pivot", abs(vlen(vsub(p, v3(0, 5, 0))) - 1.0) < 0.05))
    L.m3d_joint_destroy(w, hinge)
    L.m3d_body_destroy(w, arm)
    bd.position = v3(1, 5, 0)
    arm = L.m3d_body_create(w, C.byref(bd), C.byref(sd))
    hd.bodyB = arm
    hd.enableLimit = True
    hd.lowerAngle = -0.25 * M3D_PI
    hd.upperAngle = 0.25 * M3D_PI
    hinge = L.m3d_joint_create_hinge(w, C.byref(hd))
    min_angle = 0.0
    for _ in range(240):
        L.m3d_world_step(w, DT, 4)
        min_angle = min(min_angle, L.m3d_joint_hinge_angle(w, hinge))
    r.append(("hinge limit holds the swing", min_angle > -0.25 * M3D_PI - 0.05))
    L.m3d_joint_destroy(w, hinge)
    L.m3d_body_destroy(w, arm)
    bd.position = v3(0, 5, 0)
    ws = L.m3d_shape_def_default()
    ws.type = M3D_SHAPE_SPHERE
    ws.radius = 0.5
    wheel = L.m3d_body_create(w, C.byref(bd), C.byref(ws))
    hd.bodyB = wheel
    hd.localAnchorB = v3(0, 0, 0)
    hd.enableLimit = False
    hd.enableMotor = True
    hd.motorSpeed = 5.0
    hd.maxMotorTorque = 1000.0
    L.m3d_joint_create_hinge(w, C.byref(hd))
    steps(L, w, 120)
    r.append(("hinge motor reaches target speed", abs(L.m3d_body_angular_velocity(w, wheel).z - 5.0) < 0.2))
    L.m3d_world_destroy(w)
    return r


def weld_joint(L):  # test/test_main.cpp:621-649
    w = world(L)
    ground(L, w)
    a = box(L, w, v3(0, 3, 0), v3(0.4, 0.4, 0.4), 0.6)
    b = box(L, w, v3(0.85, 3, 0), v3(0.4, 0.4, 0.4), 0.6)
    jd = L.m3d_weld_joint_def_default()
    jd.bodyA, jd.bodyB = a, b
    jd.localAnchorA = v3(0.425, 0, 0)
    jd.localAnchorB = v3(-0.425, 0, 0)
    L.m3d_joint_create_weld(w, C.byref(jd))
    steps(L, w, 300)
    pa, pb = L.m3d_body_position(w, a), L.m3d_body_position(w, b)
    qa, qb = L.m3d_body_rotation(w, a), L.m3d_body_rotation(w, b)
    rel_w = qa.w * qb.w + qa.x * qb.x + qa.y * qb.y + qa.z * qb.z  # (conj(qa) * qb).w
    L.m3d_world_destroy(w)
    return [("weld keeps the anchor distance", abs(vlen(vsub(pb, pa)) - 0.85) < 0.03),
            ("weld keeps the relative orientation", abs(rel_w) > 0.999)]


def sphere_cast_and_overlap(L):  # test/test_main.cpp:652-693
    w = world(L)
    ground(L, w)
    hit = L.m3d_world_sphere_cast(w, v3(0, 2, 0), 0.25, v3(0, -2, 0), 0xFFFFFFFF)
    miss = L.m3d_world_sphere_cast(w, v3(0, 5, 0), 0.25, v3(0, -2, 0), 0xFFFFFFFF)
    inside = box(L, w, v3(0, 0.5, 0), v3(0.4, 0.4, 0.4), 0.6)
    box(L, w, v3(20, 0.5, 0), v3(0.4, 0.4, 0.4), 0.6)
    seen = {"count": 0, "inside": False}

    @overlap_fn
    def cb(bid, ctx):
        seen["count"] += 1
        if bid.index == inside.index:
            seen["inside"] = True
        return True

    region = m3d_aabb(v3(-2, 0.1, -2), v3(2, 2, 2))
    L.m3d_world_overlap_aabb(w, region, cb, None)
    L.m3d_world_destroy(w)
    return [("sphere cast hits the ground", hit.hit),
            ("sphere cast fraction accounts for the radius", abs(hit.fraction - 0.875) < 0.01),
            ("sphere cast normal points up", hit.normal.y > 0.99), ("short sphere cast misses", not miss.hit),
            ("overlap query finds the box in the region", seen["inside"]),
            ("overlap query skips bodies outside the region", seen["count"] == 1)]


def thin_wall_high_speed(L):  # test/test_main.cpp:699-729
    w = world(L)
    ground(L, w)
    body(L, w, M3D_BODY_STATIC, v3(20, 2, 0), M3D_SHAPE_BOX, halfExtents=v3(0.025, 2, 2))
    pea = body(L, w, M3D_BODY_DYNAMIC, v3(12, 2, 0), M3D_SHAPE_SPHERE, radius=0.06, linearVelocity=v3(150, 0, 0),
               gravityScale=0.0)
    steps(L, w, 60)
    p = L.m3d_body_position(w, pea)
    L.m3d_world_destroy(w)
    return [("150 m/s pea ends on the near side of a 5 cm wall", p.x < 20.0)]


def _chain_run(L):
    w = world(L)
    ground(L, w)
    prev = body(L, w, M3D_BODY_STATIC, v3(0, 8, 0), M3D_SHAPE_BOX, halfExtents=v3(0.1, 0.1, 0.1))
    last = prev
    for i in range(6):
        link = box(L, w, v3(0.9 * (i + 1), 8, 0), v3(0.4, 0.15, 0.15), 0.6)
        hd = L.m3d_hinge_joint_def_default()
        hd.bodyA, hd.bodyB = prev, link
        hd.localAnchorA = v3(0, 0, 0) if i == 0 else v3(0.45, 0, 0)
        hd.localAnchorB = v3(-0.45, 0, 0)
        hd.localAxisA = v3(0, 0, 1)
        L.m3d_joint_create_hinge(w, C.byref(hd))
        prev = last = link
    for i in range(12):
        box(L, w, v3(0.6 * i - 2.0, 10.0 + 0.3 * i, 0.05 * i), v3(0.25, 0.25, 0.25), 0.6)
    steps(L, w, 180)
    p = L.m3d_body_position(w, last)
    L.m3d_world_destroy(w)
    return p.tuple()


def joint_determinism(L):  # test/test_main.cpp:732-782 (run-to-run)
    return [("hinge chain is bit-identical run to run", _chain_run(L) == _chain_run(L))]


def _incline(L, theta_deg, mu):
    th = theta_deg * M3D_PI / 180.0
    w = world(L)
    gd = L.m3d_body_def_default()
    gd.orientation = L.quat_axis_angle(v3(0, 0, 1), th)
    gs = L.m3d_shape_def_default()
    gs.type = M3D_SHAPE_BOX
    gs.halfExtents = v3(60, 1, 8)
    gs.friction = mu
    L.m3d_body_create(w, C.byref(gd), C.byref(gs))
    nrm = (-math.sin(th), math.cos(th))
    along = (math.cos(th), math.sin(th))
    bd = L.m3d_body_def_default()
    bd.type = M3D_BODY_DYNAMIC
    bd.position = v3(20 * along[0] + 1.402 * nrm[0], 20 * along[1] + 1.402 * nrm[1], 0)
    bd.orientation = gd.orientation
    sd = L.m3d_shape_def_default()
    sd.type = M3D_SHAPE_BOX
    sd.halfExtents = v3(0.4, 0.4, 0.4)
    sd.friction = mu
    return w, L.m3d_body_create(w, C.byref(bd), C.byref(sd))


def inclined_plane(L):  # test/test_main.cpp:786-851
    r = []
    w, b = _incline(L, 28.0, 0.6)
    p0 = L.m3d_body_position(w, b)
    steps(L, w, 480)
    r.append(("box holds below the friction cone", vlen(vsub(L.m3d_body_position(w, b), p0)) < 0.05))
    L.m3d_world_destroy(w)
    w, b = _incline(L, 38.0, 0.6)
    p0 = L.m3d_body_position(w, b)
    steps(L, w, 480)
    r.append(("box slides above the friction cone", vlen(vsub(L.m3d_body_position(w, b), p0)) > 0.5))
    L.m3d_world_destroy(w)
    w, b = _incline(L, 35.0, 0.3)
    steps(L, w, 60, dt=1.0 / 240.0, sub=1)
    v1 = L.m3d_body_linear_velocity(w, b)
    steps(L, w, 240, dt=1.0 / 240.0, sub=1)
    a_meas = vlen(vsub(L.m3d_body_linear_velocity(w, b), v1))
    th = 35.0 * M3D_PI / 180.0
    a_exp = 9.81 * (math.sin(th) - 0.3 * math.cos(th))
    r.append(("Coulomb sliding acceleration matches g(sin-mu cos)", abs(a_meas - a_exp) < 0.1 * a_exp))
    L.m3d_world_destroy(w)
    return r


def rolling_physics(L):  # test/test_main.cpp:856-927
    r = []
    th = 20.0 * M3D_PI / 180.0
    w = world(L)
    gd = L.m3d_body_def_default()
    gd.orientation = L.quat_axis_angle(v3(0, 0, 1), th)
    gs = L.m3d_shape_def_default()
    gs.type = M3D_SHAPE_BOX
    gs.halfExtents = v3(60, 1, 8)
    gs.friction = 0.8
    L.m3d_body_create(w, C.byref(gd), C.byref(gs))
    ball = body(L, w, M3D_BODY_DYNAMIC,
                v3(20 * math.cos(th) - 1.402 * math.sin(th), 20 * math.sin(th) + 1.402 * math.cos(th), 0),
                M3D_SHAPE_SPHERE, angularDamping=0.0, radius=0.4, friction=0.8, rollingResistance=0.0)
    steps(L, w, 120, dt=1.0 / 240.0, sub=1)
    v1 = L.m3d_body_linear_velocity(w, ball)
    steps(L, w, 240, dt=1.0 / 240.0, sub=1)
    a_meas = vlen(vsub(L.m3d_body_linear_velocity(w, ball), v1))
    a_exp = 5.0 / 7.0 * 9.81 * math.sin(th)
    r.append(("solid sphere rolls at 5/7 g sin(th)", abs(a_meas - a_exp) < 0.08 * a_exp))
    L.m3d_world_destroy(w)

    w = world(L)
    ground(L, w)
    ball = body(L, w, M3D_BODY_DYNAMIC, v3(0, 0.5, 0), M3D_SHAPE_SPHERE, angularVelocity=v3(0, 0, -10.0),
                angularDamping=0.0, radius=0.5, friction=0.5, rollingResistance=0.0)
    steps(L, w, 150, dt=1.0 / 240.0, sub=1)
    v_roll = L.m3d_body_linear_velocity(w, ball).x
    r.append(("backspin converts at v = 2/7 w r", abs(v_roll - 2.0 / 7.0 * 10.0 * 0.5) < 0.15))
    steps(L, w, 480, dt=1.0 / 240.0, sub=1)
    r.append(("free roller never self-accelerates", L.m3d_body_linear_velocity(w, ball).x <= v_roll + 0.05))
    L.m3d_world_destroy(w)
    return r


def dzhanibekov(L):  # test/test_main.cpp:931-953
    w = world(L, gravity=v3(0, 0, 0))
    brick = body(L, w, M3D_BODY_DYNAMIC, v3(0, 0, 0), M3D_SHAPE_BOX, angularDamping=0.0,
                 angularVelocity=v3(0.02, 8.0, 0.0), halfExtents=v3(0.5, 0.3, 0.1))
    min_dot = 1.0
    for _ in range(2400):
        L.m3d_world_step(w, 1.0 / 240.0, 1)
        min_dot = min(min_dot, rotate(L.m3d_body_rotation(w, brick), v3(0, 1, 0)).y)
    L.m3d_world_destroy(w)
    return [("Dzhanibekov flip about the intermediate axis", min_dot < -0.5)]


def _raft(L):  # test/test_main.cpp:960-985
    w = world(L)
    ground(L, w)
    ids = []
    for gx in range(15):
        for gz in range(15):
            ids.append(box(L, w, v3(1.0 * gx - 7.0, 0.5, 1.0 * gz - 7.0), v3(0.501, 0.5, 0.501), 0.6))
    for i in range(5):
        ids.append(box(L, w, v3(2.0 * i - 4.0, 3.0 + 0.7 * i, 0.5 * i - 1.0), v3(0.45, 0.45, 0.45), 0.6))
    return w, ids


def colored_path_determinism(L):  # test/test_main.cpp:987-1067 (run-to-run)
    results = []
    seen = 0
    for _ in range(2):
        w, ids = _raft(L)
        for _ in range(90):
            L.m3d_world_step(w, DT, 4)
            p = L.m3d_world_profile(w)
            if p.touchingContactCount >= 300 and p.islandCount <= 8:
                seen += 1
        results.append(L.m3d_body_position(w, ids[-1]).tuple())
        L.m3d_world_destroy(w)
    return [("raft scene actually reaches the colored-path contact threshold", seen > 0),
            ("colored path bit-identical run to run", results[0] == results[1])]


ALL = [sphere_rest, capsule_rest, box_stack, restitution, sleep_and_wake, ray_cast, distance_joint, ball_joint,
       contact_events, speculative_no_tunnel, determinism, collision_filter, sensor, forces_and_torques,
       hinge_joint, weld_joint, sphere_cast_and_overlap, thin_wall_high_speed, joint_determinism, inclined_plane,
       rolling_physics, dzhanibekov, colored_path_determinism]
