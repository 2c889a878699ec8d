"""Helpers comparing the device backend with the reference build."""
from __future__ import annotations

import math

import numpy as np


def bodies_array(snap):
    """(position[n,3], rotation[n,4], linVel[n,3], angVel[n,3], inUse[n]) as float64."""
    n = snap.header.bodyCapacity
    pos = np.zeros((n, 3))
    rot = np.zeros((n, 4))
    lv = np.zeros((n, 3))
    av = np.zeros((n, 3))
    use = np.zeros(n, dtype=bool)
    awake = np.zeros(n, dtype=bool)
    for i in range(n):
        b = snap.bodies[i]
        pos[i] = b.position[:]
        rot[i] = b.rotation[:]
        lv[i] = b.linearVelocity[:]
        av[i] = b.angularVelocity[:]
        use[i] = bool(b.inUse)
        awake[i] = bool(b.isAwake)
    return pos, rot, lv, av, use, awake


def contact_pairs(snap):
    return [(snap.contacts[c].bodyA, snap.contacts[c].bodyB) for c in range(snap.header.contactCount)]


def contact_touching(snap):
    return [bool(snap.contacts[c].touching) for c in range(snap.header.contactCount)]


def rot_angle(q1, q2):
    """Angle (rad) of the relative rotation q1 * conj(q2) between quaternion
    rows, from 2 atan2(|vec|, |w|); float32 products are exact in float64, so
    identical inputs give exactly 0 even when |q| != 1 by rounding."""
    x1, y1, z1, w1 = q1.T
    x2, y2, z2, w2 = (-q2[:, 0], -q2[:, 1], -q2[:, 2], q2[:, 3])
    rx = w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2
    ry = w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2
    rz = w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2
    rw = w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2
    return 2.0 * np.arctan2(np.sqrt(rx * rx + ry * ry + rz * rz), np.abs(rw))


def state_diff(a, b):
    pa, ra, va, wa, ua, aa = bodies_array(a)
    pb, rb, vb, wb, ub, ab = bodies_array(b)
    assert (ua == ub).all(), "in-use sets differ"
    m = ua
    dp = float(np.abs(pa[m] - pb[m]).max()) if m.any() else 0.0
    dr = float(rot_angle(ra[m], rb[m]).max()) if m.any() else 0.0
    dv = float(np.abs(va[m] - vb[m]).max()) if m.any() else 0.0
    dw = float(np.abs(wa[m] - wb[m]).max()) if m.any() else 0.0
    return {"pos": dp, "rot": dr, "vel": dv, "angvel": dw, "awake_equal": bool((aa[m] == ab[m]).all())}


def events(lib, world):
    ev = lib.m3d_world_contact_events(world)
    b = [(ev.beginEvents[k].bodyA.index, ev.beginEvents[k].bodyB.index, bool(ev.beginEvents[k].isSensor))
         for k in range(ev.beginCount)]
    e = [(ev.endEvents[k].bodyA.index, ev.endEvents[k].bodyB.index, bool(ev.endEvents[k].isSensor))
         for k in range(ev.endCount)]
    return b, e


def one_step_from(ref, dev, ref_world, dt=1.0 / 60.0, sub=4):
    """Import the reference world's full state into a fresh device world,
    step both once and return (ref snapshot, device snapshot, device world)."""
    snap = ref.export_state(ref_world)
    dw = dev.import_state(snap)
    rw2 = ref.import_state(snap)
    ref.m3d_world_step(rw2, dt, sub)
    dev.m3d_world_step(dw, dt, sub)
    a = ref.export_state(rw2)
    b = dev.export_state(dw)
    ref.m3d_world_destroy(rw2)
    return a, b, dw


# ---- vectorized views for full-size states (hundreds of thousands of rows)

def np_bodies(snap):
    """Structured numpy view of a snapshot's body slots."""
    return np.ctypeslib.as_array(snap._keep["bodies"])[: snap.header.bodyCapacity]


def np_contacts(snap):
    """Structured numpy view of a snapshot's contacts."""
    return np.ctypeslib.as_array(snap._keep["contacts"])[: snap.header.contactCount]


def state_diff_rows(A, B):
    """state_diff over two structured arrays of body slots."""
    assert (A["inUse"] == B["inUse"]).all(), "in-use sets differ"
    m = A["inUse"] != 0
    if not m.any():
        return {"pos": 0.0, "rot": 0.0, "vel": 0.0, "angvel": 0.0, "awake_equal": True, "identical": True}
    f = lambda k: float(np.abs(A[k][m].astype(np.float64) - B[k][m].astype(np.float64)).max())
    rot = float(rot_angle(A["rotation"][m].astype(np.float64), B["rotation"][m].astype(np.float64)).max())
    same = all(np.array_equal(A[k][m], B[k][m]) for k in ("position", "rotation", "linearVelocity", "angularVelocity"))
    return {"pos": f("position"), "rot": rot, "vel": f("linearVelocity"), "angvel": f("angularVelocity"),
            "awake_equal": bool((A["isAwake"][m] == B["isAwake"][m]).all()), "identical": same}


def state_diff_np(a, b):
    """state_diff over numpy views of two snapshots."""
    return state_diff_rows(np_bodies(a), np_bodies(b))


def pairs_np(snap, offset=0):
    """(contacts, 2) body pairs in contact order, body indices minus offset."""
    c = np_contacts(snap)
    return np.stack([c["bodyA"] - offset, c["bodyB"] - offset], axis=1)
