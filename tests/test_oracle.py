"""Pins the oracle's C restatement (oracle/m3d_oracle.c) before it is trusted
as a checker: against the reference build's own stage outputs, closed forms
and brute force. A plausible slip (axis order, floor vs trunc, sort
tie-break, first-fit vs best-fit, missing edge) fails one of these."""
import random

import numpy as np
import pytest

import oracle.oracle as orc
from melon3d_amd import scenes

DT = 1.0 / 60.0
OFF = 1 << 20


def _pack(x, y, z):
    return (((x + OFF) & 0x1FFFFF) << 42) | (((y + OFF) & 0x1FFFFF) << 21) | ((z + OFF) & 0x1FFFFF)


def test_cell_key_closed_form():
    assert orc.cell_key(0.0, 0.0, 0.0, 1.5) == _pack(0, 0, 0)
    # floor, not truncation, for negative coordinates; axis order x<<42|y<<21|z
    assert orc.cell_key(-0.1, 1.6, 3.0, 1.5) == _pack(-1, 1, 2)
    assert orc.cell_key(-1.5, -3.0, 0.75, 1.5) == _pack(-1, -2, 0)
    # the multiply by a rounded reciprocal is what the reference evaluates
    rng = np.random.default_rng(7)
    for v in rng.uniform(-200.0, 200.0, size=(200, 3)).astype(np.float32):
        inv = np.float32(1.0) / np.float32(1.5)
        c = [int(np.floor(np.float32(t) * inv)) for t in v]
        assert orc.cell_key(float(v[0]), float(v[1]), float(v[2]), 1.5) == _pack(*c)


def _fat(snap):
    n = snap.header.bodyCapacity
    return np.array([snap.bodies[i].fatAABB[:] for i in range(n)], dtype=np.float32)


@pytest.mark.parametrize("scene,steps", [("pyramid", 3), ("rain", 25)])
def test_build_tier_matches_reference(ref, scene, steps):
    w = ref.create_world()
    if scene == "pyramid":
        scenes.pyramid(ref, w, base=12)
    else:
        scenes.rain_pile(ref, w, count=800, column=(10.0, 10.0, 12.0))
    for _ in range(steps):
        ref.m3d_world_step(w, DT, 4)
    snap = ref.export_state(w)
    active = ref.grid_entries(w, 0)
    over = ref.oversized(w, 0)
    include = np.zeros(snap.header.bodyCapacity, dtype=bool)
    for _, b in active:
        include[b] = True
    include[over] = True
    entries, ov, _ = orc.build_tier(_fat(snap), include, snap.header.gridCellSize)
    assert entries == active
    assert ov == over
    ref.m3d_world_destroy(w)


def test_brute_pairs_match_reference_created_set(ref):
    """Step 1 of a rain pile: every body is active and moved, the ground sits
    in both coarse lists; the created contact set equals the brute-force
    admissible overlapping pairs."""
    w = ref.create_world()
    scenes.rain_pile(ref, w, count=400, column=(6.0, 6.0, 4.0))
    s0 = ref.export_state(w)
    ref.m3d_world_step(w, DT, 4)
    s1 = ref.export_state(w)
    n = s1.header.bodyCapacity
    flags = []
    for i in range(n):
        b0 = s0.bodies[i]
        flags.append(dict(inUse=b0.inUse, dynamic=int(b0.type == 2), awake=b0.isAwake, moved=1, activeTier=1,
                          stableTier=int(b0.inUse and not b0.isAwake), category=b0.categoryBits, mask=b0.maskBits,
                          group=b0.groupIndex))
    brute = orc.brute_pairs(_fat(s1), flags, existing=[], jointed=[])
    created = sorted((s1.contacts[c].bodyA << 32) | s1.contacts[c].bodyB for c in range(s1.header.contactCount))
    assert brute == created
    assert len(created) > 100
    ref.m3d_world_destroy(w)


def test_islands_match_reference(ref):
    w = ref.create_world()
    scenes.rain_pile(ref, w, count=900, column=(9.0, 9.0, 10.0))
    for _ in range(90):
        ref.m3d_world_step(w, DT, 4)
    snap = ref.export_state(w)
    n = snap.header.bodyCapacity
    dyn = np.array([snap.bodies[i].inUse and snap.bodies[i].type == 2 for i in range(n)])
    edges = [(snap.contacts[c].bodyA, snap.contacts[c].bodyB) for c in range(snap.header.contactCount)
             if snap.contacts[c].touching and not snap.contacts[c].isSensor]
    cnt, labels = orc.islands(dyn, edges)
    rcnt, rlab, _ = ref.islands(w, n)
    assert rcnt > 1
    assert cnt == rcnt
    assert list(labels) == rlab
    ref.m3d_world_destroy(w)


def test_islands_brute_force_property():
    """Labels equal connected components found by repeated relaxation."""
    rng = random.Random(3)
    for _ in range(20):
        n = rng.randint(1, 60)
        dyn = np.array([rng.random() < 0.8 for _ in range(n)])
        edges = [(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randint(0, 80))]
        cnt, lab = orc.islands(dyn, edges)
        comp = list(range(n))
        changed = True
        while changed:
            changed = False
            for a, b in edges:
                if dyn[a] and dyn[b] and comp[a] != comp[b]:
                    m = min(comp[a], comp[b])
                    comp[a] = comp[b] = m
                    changed = True
        roots = sorted({comp[i] for i in range(n) if dyn[i]})
        expect = [roots.index(comp[i]) if dyn[i] else -1 for i in range(n)]
        assert list(lab) == expect
        assert cnt == len(roots)


def test_greedy_color_matches_reference(ref):
    """Box pyramid: one island with > 300 contacts takes the colored path."""
    w = ref.create_world()
    scenes.pyramid(ref, w, base=20)
    for _ in range(5):
        ref.m3d_world_step(w, DT, 4)
    snap = ref.export_state(w)
    C_ = snap.header.contactCount
    ref_colors = ref.contact_colors(w, C_)
    idx = [c for c in range(C_) if snap.contacts[c].touching and not snap.contacts[c].isSensor]
    assert len(idx) >= 300 and all(ref_colors[c] >= 0 for c in idx)
    a = [snap.contacts[c].bodyA for c in idx]
    b = [snap.contacts[c].bodyB for c in idx]
    da = [snap.bodies[x].invMass > 0 for x in a]
    db = [snap.bodies[x].invMass > 0 for x in b]
    colors = orc.greedy_color(a, b, da, db, snap.header.bodyCapacity)
    assert list(colors) == [ref_colors[c] for c in idx]
    assert orc.color_valid(a, b, da, db, colors, snap.header.bodyCapacity)
    # a planted conflict must be caught by the validity checker
    bad = colors.copy()
    j = next(k for k in range(1, len(idx)) if (a[k] == a[0] and da[0]) or (b[k] == b[0] and db[0]) or
             (a[k] == b[0] and db[0]) or (b[k] == a[0] and da[0]))
    bad[j] = bad[0]
    assert not orc.color_valid(a, b, da, db, bad, snap.header.bodyCapacity)
    ref.m3d_world_destroy(w)


def test_levels_reproduce_serial_order():
    """Applying constraints level by level gives every body the same ordered
    sequence of updates as the serial order (non-commutative per-body logs)."""
    rng = random.Random(11)
    for _ in range(50):
        nb = rng.randint(2, 30)
        k = rng.randint(1, 120)
        a = [rng.randrange(nb) for _ in range(k)]
        b = [rng.randrange(nb) for _ in range(k)]
        da = [rng.random() < 0.85 for _ in range(k)]
        db = [rng.random() < 0.85 for _ in range(k)]
        lv = orc.levels(a, b, da, db, nb)
        serial = {i: [] for i in range(nb)}
        for c in range(k):
            if da[c]:
                serial[a[c]].append(c)
            if db[c]:
                serial[b[c]].append(c)
        leveled = {i: [] for i in range(nb)}
        for L in range(int(lv.max()) + 1 if k else 0):
            members = [c for c in range(k) if lv[c] == L]
            touched = []
            for c in members:
                for x, d in ((a[c], da[c]), (b[c], db[c])):
                    if d:
                        touched.append(x)
                        leveled[x].append(c)
            # one level never writes a body twice (unless one constraint
            # names the same body on both sides)
            dup = [x for x in set(touched) if touched.count(x) > 1]
            assert all(any(a[c] == b[c] == x for c in members) for x in dup)
        assert serial == leveled
