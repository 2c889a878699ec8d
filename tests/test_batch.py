"""Batched sub-worlds: every sub-world of a batch steps exactly as a world of
its own holding the same bodies.

The batch is new (the reference steps one world per call), so its oracle is
the single-world step: the plain device world, which matches the reference
engine step for step (test_parity.py), and the reference engine itself for a
few steps. A batch sub-world must equal its plain twin bit for bit: same
bodies, same contacts in the same order with the same manifolds, through
sleeping and waking (per-sub-world stable-tier timing) and contact removal.
"""
import ctypes as C

import numpy as np
import pytest

from melon3d_amd import scenes
from parity import bodies_array, events, rot_angle

DT = 1.0 / 60.0


def _counts(dev, w, S):
    out = (C.c_int32 * S)()
    assert dev.m3dx_batch_contact_counts(w, out, S) == S
    return list(out)


def _split(dev, w, S):
    """Per sub-world (bodies, contact rows) of a batch, body indices local."""
    snap = dev.export_state(w)
    pos, rot, lv, av, use, awake = bodies_array(snap)
    counts = _counts(dev, w, S)
    out = []
    c0 = 0
    for s in range(S):
        first, n = C.c_int32(0), C.c_int32(0)
        assert dev.m3dx_batch_body_range(w, s, C.byref(first), C.byref(n)) == 0
        b0, nb = first.value, n.value
        sl = slice(b0, b0 + nb)
        bodies = (pos[sl], rot[sl], lv[sl], av[sl], use[sl], awake[sl])
        rows = [_contact_row(snap.contacts[c], b0) for c in range(c0, c0 + counts[s])]
        out.append((bodies, rows))
        c0 += counts[s]
    assert c0 == snap.header.contactCount
    return out


def _contact_row(c, b0):
    pts = tuple((tuple(p.localAnchorA), tuple(p.localAnchorB), p.separation, p.id, p.lambdaN, p.lambdaT,
                 tuple(p.slipDir)) for p in list(c.points)[:c.pointCount])
    return (c.bodyA - b0, c.bodyB - b0, c.touching, c.pointCount, tuple(c.normal), pts)


def _plain(dev, w):
    snap = dev.export_state(w)
    rows = [_contact_row(snap.contacts[c], 0) for c in range(snap.header.contactCount)]
    return bodies_array(snap), rows


def _ranges(dev, w, S):
    out = []
    for s in range(S):
        first, n = C.c_int32(0), C.c_int32(0)
        assert dev.m3dx_batch_body_range(w, s, C.byref(first), C.byref(n)) == 0
        out.append((first.value, n.value))
    return out


def _events_by_sub(ev, ranges):
    """A batch's event list split by sub-world (of body A), body indices local."""
    starts = np.array([r[0] for r in ranges])
    out = [[] for _ in ranges]
    for a, b, sensor in ev:
        s = int(np.searchsorted(starts, a, side="right")) - 1
        out[s].append((a - starts[s], b - starts[s], sensor))
    return out


def _assert_same(batch_part, plain_part, s, step):
    (bb, brows), (pb, prows) = batch_part, plain_part
    for k, name in enumerate(("position", "rotation", "linear velocity", "angular velocity", "in use", "awake")):
        assert np.array_equal(bb[k], pb[k]), f"sub-world {s} step {step}: {name} differs"
    assert len(brows) == len(prows), f"sub-world {s} step {step}: {len(brows)} vs {len(prows)} contacts"
    for i, (a, b) in enumerate(zip(brows, prows)):
        assert a == b, f"sub-world {s} step {step}: contact {i} differs"


@pytest.mark.gpu
@pytest.mark.parametrize("first,S,steps,every", [(0, 12, 60, 10), (40, 6, 240, 40)])
def test_batch_equals_plain_worlds(dev, first, S, steps, every):
    """Bitwise equality with S plain worlds, including sleep (the second case
    runs 4 s, long enough for piles and pyramids to sleep, which rebuilds the
    stable tiers of sub-worlds at different steps). Every step, each
    sub-world's begin and end events (those of its bodies, in the batch's
    order) are its plain twin's, in order: the end events of removed
    contacts, then of touches that ended."""
    bw, info = scenes.batch(dev, S, first=first)
    plain = []
    for s in range(S):
        w = dev.create_world()
        n, pile = scenes.batch_world(dev, w, first + s)
        assert (n, pile) == info[s]
        plain.append(w)
    ranges = _ranges(dev, bw, S)
    asleep = 0
    ended = 0
    try:
        for step in range(1, steps + 1):
            dev.m3d_world_step(bw, DT, 4)
            for w in plain:
                dev.m3d_world_step(w, DT, 4)
            bb, be = events(dev, bw)
            sb, se = _events_by_sub(bb, ranges), _events_by_sub(be, ranges)
            for s in range(S):
                pb, pe = events(dev, plain[s])
                assert sb[s] == pb, f"sub-world {s} step {step}: begin events differ"
                assert se[s] == pe, f"sub-world {s} step {step}: end events differ"
            ended += len(be)
            if step % every == 0 or step == steps:
                parts = _split(dev, bw, S)
                for s in range(S):
                    _assert_same(parts[s], _plain(dev, plain[s]), s, step)
                asleep = sum(int((~p[0][5] & p[0][4]).sum()) for p in parts)
    finally:
        dev.m3d_world_destroy(bw)
        for w in plain:
            dev.m3d_world_destroy(w)
    if steps < 240:
        assert ended > 0, "no end events"
    if steps >= 240:
        assert asleep > 0, "no body fell asleep: the stable-tier path went untested"


@pytest.mark.gpu
def test_batch_removal_end_events(dev):
    """Touching contacts removed by the step (bodies teleported away from the
    heap they rest on) give each sub-world its plain twin's end events, in
    its order (descending contact position), sub-worlds in ascending order."""
    from melon3d_amd.api import m3d_body_id, m3d_vec3

    S, first = 4, 3
    bw, info = scenes.batch(dev, S, first=first)
    plain = []
    for s in range(S):
        w = dev.create_world()
        scenes.batch_world(dev, w, first + s)
        plain.append(w)
    ranges = _ranges(dev, bw, S)
    removed = 0
    try:
        for step in range(1, 91):
            if step in (40, 70):
                # every third body of each sub-world jumps 500 m up (its slot
                # k in the sub-world: the same body in the twin)
                for s in range(S):
                    b0, nb = ranges[s]
                    for k in range(1 + step % 3, nb, 3):
                        for w, idx in ((bw, b0 + k), (plain[s], k)):
                            bid = m3d_body_id(idx, 1)  # never destroyed: first generation
                            t = dev.m3d_body_transform(w, bid)
                            dev.m3d_body_set_transform(w, bid, m3d_vec3(t.p.x, t.p.y + 500.0, t.p.z), t.q)
            dev.m3d_world_step(bw, DT, 4)
            for w in plain:
                dev.m3d_world_step(w, DT, 4)
            bb, be = events(dev, bw)
            sb, se = _events_by_sub(bb, ranges), _events_by_sub(be, ranges)
            for s in range(S):
                pb, pe = events(dev, plain[s])
                assert sb[s] == pb, f"sub-world {s} step {step}: begin events differ"
                assert se[s] == pe, f"sub-world {s} step {step}: end events differ"
            if step in (40, 70):
                removed += len(be)
        parts = _split(dev, bw, S)
        for s in range(S):
            _assert_same(parts[s], _plain(dev, plain[s]), s, 90)
    finally:
        dev.m3d_world_destroy(bw)
        for w in plain:
            dev.m3d_world_destroy(w)
    assert removed > 0, "the teleports removed no touching contact"


@pytest.mark.gpu
def test_pass_workgroups_change_nothing(dev):
    """The solver passes' workgroup count (tuned by measurement, or forced)
    changes no result: a 48-world batch stepped with 32 and with 256 forced
    workgroups and with the tuner choosing is bit-identical after 30 steps."""
    S, steps = 48, 30
    worlds = []
    try:
        for blocks in (32, 256, 0):
            bw, _ = scenes.batch(dev, S, first=100)
            assert dev.m3dx_set_pass_blocks(bw, blocks) == 0
            worlds.append(bw)
        used = [set() for _ in worlds]
        for step in range(steps):
            for k, bw in enumerate(worlds):
                dev.m3d_world_step(bw, DT, 4)
                used[k].add(dev.m3dx_pass_blocks(bw))
        assert [dev.m3dx_step_failures(bw) for bw in worlds] == [0, 0, 0], dev.last_error()
        assert used[0] != used[1], used
        assert len(used[2]) > 1, used  # the tuner tried several counts
        parts = [_split(dev, bw, S) for bw in worlds]
        for k in (1, 2):
            for s in range(S):
                _assert_same(parts[k][s], parts[0][s], s, steps)
    finally:
        for bw in worlds:
            dev.m3d_world_destroy(bw)


@pytest.mark.gpu
def test_xcd_barrier_changes_nothing(dev):
    """The solver passes' grid barriers in XCD mode (one L2 write-back and
    invalidation per XCD) and with agent-scope fences in every workgroup
    give bit-identical batches after 30 steps, with 256 workgroups asked
    for (as many as the runtime lets run together, over 100: a dozen or more
    per XCD) and with 24 (a group of 3 on each XCD), and every step
    succeeds."""
    S, steps = 48, 30
    for blocks in (256, 24):
        worlds = []
        try:
            for xcd in (0, 1):
                bw, _ = scenes.batch(dev, S, first=300)
                assert dev.m3dx_set_pass_blocks(bw, blocks) == 0
                assert dev.m3dx_set_xcd_barrier(bw, xcd) == 0
                worlds.append(bw)
            for step in range(steps):
                for bw in worlds:
                    dev.m3d_world_step(bw, DT, 4)
            assert [dev.m3dx_step_failures(bw) for bw in worlds] == [0, 0], dev.last_error()
            used = dev.m3dx_pass_blocks(worlds[1])
            assert used == blocks if blocks < 100 else used > 100, used
            parts = [_split(dev, bw, S) for bw in worlds]
            for s in range(S):
                _assert_same(parts[1][s], parts[0][s], s, steps)
        finally:
            for bw in worlds:
                dev.m3d_world_destroy(bw)


@pytest.mark.gpu
def test_step_graph_changes_nothing(dev):
    """The batch's substep loop run as a captured hipGraph (the default once
    the workgroup count is settled) equals the directly launched loop bit for
    bit over 30 steps."""
    S, steps = 48, 30
    worlds = []
    try:
        for graphs in (0, 1):
            bw, _ = scenes.batch(dev, S, first=200)
            assert dev.m3dx_set_pass_blocks(bw, 64) == 0
            assert dev.m3dx_set_step_graphs(bw, graphs) == 0
            worlds.append(bw)
        for step in range(steps):
            for bw in worlds:
                dev.m3d_world_step(bw, DT, 4)
        assert [dev.m3dx_step_failures(bw) for bw in worlds] == [0, 0], dev.last_error()
        assert dev.m3dx_step_graph_launches(worlds[0]) == 0
        assert dev.m3dx_step_graph_launches(worlds[1]) >= steps - 2
        direct, graphed = (_split(dev, bw, S) for bw in worlds)
        for s in range(S):
            _assert_same(graphed[s], direct[s], s, steps)
    finally:
        for bw in worlds:
            dev.m3d_world_destroy(bw)


@pytest.mark.gpu
def test_batch_against_reference(dev, ref):
    """Four sub-worlds against the reference engine, 3 steps, 1e-5."""
    S, first = 4, 7
    bw, info = scenes.batch(dev, S, first=first)
    refs = []
    for s in range(S):
        w = ref.create_world()
        scenes.batch_world(ref, w, first + s)
        refs.append(w)
    try:
        for step in range(3):
            dev.m3d_world_step(bw, DT, 4)
            for w in refs:
                ref.m3d_world_step(w, DT, 4)
        parts = _split(dev, bw, S)
        for s in range(S):
            rs = ref.export_state(refs[s])
            pb, rb, vb, wb, ub, ab = bodies_array(rs)
            pa, ra, va, wa, ua, aa = parts[s][0]
            assert (ua == ub).all()
            assert float(np.abs(pa - pb).max()) <= 1e-5, s
            assert float(rot_angle(ra, rb).max()) <= 1e-5, s
            assert (aa == ab).all(), s
            assert len(parts[s][1]) == rs.header.contactCount
    finally:
        dev.m3d_world_destroy(bw)
        for w in refs:
            ref.m3d_world_destroy(w)


@pytest.mark.gpu
def test_batch_destroy_equals_plain_worlds(dev):
    """m3d_body_destroy in a batch is a destroy in the body's sub-world: after
    destroys at two steps (contacts removed in the sub-world's order, end
    events, partner wakes, the stable tier rebuilt) and creates into the last
    sub-world (its freed slots reused last-in first-out), every sub-world
    equals its plain twin: ids, bodies, contacts and events."""
    from melon3d_amd.api import m3d_body_id

    S, first = 4, 11
    bw, info = scenes.batch(dev, S, first=first)
    plain = []
    for s in range(S):
        w = dev.create_world()
        scenes.batch_world(dev, w, first + s)
        plain.append(w)
    ranges = _ranges(dev, bw, S)
    destroyed = []
    try:
        for step in range(1, 101):
            if step in (30, 56):
                for s in range(S):
                    b0, nb = ranges[s]
                    for k in range(2 + step % 5, nb, 5):
                        if (s, k) in destroyed:
                            continue
                        destroyed.append((s, k))
                        dev.m3d_body_destroy(bw, m3d_body_id(b0 + k, 1))
                        dev.m3d_body_destroy(plain[s], m3d_body_id(k, 1))
                        assert not dev.m3d_body_is_valid(bw, m3d_body_id(b0 + k, 1))
                    be = events(dev, bw)[1]
                    pe = events(dev, plain[s])[1]
                    assert _events_by_sub(be, ranges)[s] == pe, f"sub-world {s}: destroy end events differ"
            if step == 70:
                # the last sub-world is the selected one: creates reuse its slots
                s = S - 1
                b0 = ranges[s][0]
                gens = []
                for j in range(8):
                    pos = (0.3 * j, 8.0 + 1.1 * j, 0.2)
                    a = scenes.add_box(dev, bw, pos, (0.4, 0.4, 0.4))
                    p = scenes.add_box(dev, plain[s], pos, (0.4, 0.4, 0.4))
                    assert (a.index - b0, a.generation) == (p.index, p.generation), j
                    gens.append(a.generation)
                assert gens[0] == 2, "the first create did not reuse a destroyed slot"
            dev.m3d_world_step(bw, DT, 4)
            for w in plain:
                dev.m3d_world_step(w, DT, 4)
            bb, be = events(dev, bw)
            sb, se = _events_by_sub(bb, ranges), _events_by_sub(be, ranges)
            for s in range(S):
                pb, pe = events(dev, plain[s])
                assert sb[s] == pb, f"sub-world {s} step {step}: begin events differ"
                assert se[s] == pe, f"sub-world {s} step {step}: end events differ"
            if step % 10 == 0 or step in (30, 31, 56, 57, 70, 71):
                parts = _split(dev, bw, S)
                for s in range(S):
                    _assert_same(parts[s], _plain(dev, plain[s]), s, step)
    finally:
        dev.m3d_world_destroy(bw)
        for w in plain:
            dev.m3d_world_destroy(w)
    assert len(destroyed) > 20


@pytest.mark.gpu
def test_batch_rejects_cross_joints(dev):
    bw, info = scenes.batch(dev, 2, first=0)
    try:
        first, n = C.c_int32(0), C.c_int32(0)
        dev.m3dx_batch_body_range(bw, 1, C.byref(first), C.byref(n))
        from melon3d_amd.api import m3d_ball_joint_def, m3d_body_id
        a = m3d_body_id(1, 1)
        b = m3d_body_id(first.value + 1, 1)
        jd = m3d_ball_joint_def()
        jd.bodyA, jd.bodyB = a, b
        assert dev.m3d_joint_create_ball(bw, C.byref(jd)).index == -1
        jd.bodyB = m3d_body_id(2, 1)  # same sub-world
        assert dev.m3d_joint_create_ball(bw, C.byref(jd)).index >= 0
        assert dev.m3dx_batch_select(bw, 0) == -1  # below the current sub-world
    finally:
        dev.m3d_world_destroy(bw)


def test_batch_workload_generator(ref):
    """The batched workload: log-uniform counts in [16, 2048], piles and
    pyramids about half each, deterministic per world index."""
    counts, kinds = [], []
    for i in range(256):
        w = ref.create_world()
        n, p = scenes.batch_world(ref, w, i)
        ref.m3d_world_destroy(w)
        counts.append(n)
        kinds.append(p)
    assert min(counts) >= 16 and max(counts) <= 2048
    assert 90 <= sum(kinds) <= 166
    # log-uniform: about a third of the worlds below 64 bodies (ln 4 / ln 128)
    assert 55 <= sum(c < 64 for c in counts) <= 115
    w = ref.create_world()
    assert scenes.batch_world(ref, w, 5) == (counts[5], kinds[5])
    ref.m3d_world_destroy(w)
    # the reference engine has no batches
    assert not ref.m3dx_batch_create(None, 4)
