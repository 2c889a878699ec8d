"""Device backend vs the reference engine on the same inputs.

Bar: contact order, grid tiers, island labels and colors bit-exact; one step
from identical state within 1e-5 m and 1e-5 rad (both engines evaluate the
same strict-IEEE expressions, so any difference comes only from sinf, cosf
and atan2f, which may round differently on the device)."""
import ctypes as C

import numpy as np
import pytest

from melon3d_amd import scenes
from parity import contact_pairs, contact_touching, events, one_step_from, state_diff

TOL_POS = 1e-5
TOL_ROT = 1e-5
DT = 1.0 / 60.0


def _island_sets(lib, world, n):
    cnt, isl, awake = lib.islands(world, n)
    return cnt, isl, awake


def _check_step(ref, dev, rw, label):
    a, b, dw = one_step_from(ref, dev, rw)
    d = state_diff(a, b)
    assert contact_pairs(a) == contact_pairs(b), f"{label}: contact order differs"
    assert contact_touching(a) == contact_touching(b), f"{label}: touching flags differ"
    assert d["pos"] <= TOL_POS and d["rot"] <= TOL_ROT, f"{label}: {d}"
    assert d["awake_equal"], f"{label}: sleep state differs"
    dev.m3d_world_destroy(dw)
    return d


@pytest.mark.gpu
@pytest.mark.parametrize("at", [0, 1, 10, 60, 150, 300])
def test_pyramid_one_step(ref, dev, at):
    """Box pyramid, 20 levels (210 boxes, one colored island)."""
    rw = ref.create_world()
    scenes.pyramid(ref, rw, base=20)
    for _ in range(at):
        ref.m3d_world_step(rw, DT, 4)
    _check_step(ref, dev, rw, f"pyramid step {at}")
    ref.m3d_world_destroy(rw)


@pytest.mark.gpu
def test_pyramid_lockstep_stage_views(ref, dev):
    """Both engines run the pyramid from scratch; every step the created
    contact order, sorted grid tiers, coarse lists, island labels, colors and
    events must agree exactly while the trajectories stay within tolerance."""
    rw = ref.create_world()
    dw = dev.create_world()
    scenes.pyramid(ref, rw, base=20)
    scenes.pyramid(dev, dw, base=20)
    n = 211
    worst = 0.0
    for step in range(120):
        ref.m3d_world_step(rw, DT, 4)
        dev.m3d_world_step(dw, DT, 4)
        a, b = ref.export_state(rw), dev.export_state(dw)
        d = state_diff(a, b)
        worst = max(worst, d["pos"])
        assert contact_pairs(a) == contact_pairs(b), f"step {step}: contact order"
        assert ref.grid_entries(rw, 0) == dev.grid_entries(dw, 0), f"step {step}: active tier"
        assert ref.grid_entries(rw, 1) == dev.grid_entries(dw, 1), f"step {step}: stable tier"
        assert ref.oversized(rw, 0) == dev.oversized(dw, 0) and ref.oversized(rw, 1) == dev.oversized(dw, 1)
        assert _island_sets(ref, rw, n) == _island_sets(dev, dw, n), f"step {step}: islands"
        C_ = a.header.contactCount
        if ref.m3d_world_profile(rw).islandCount > 0:
            assert ref.contact_colors(rw, C_) == dev.contact_colors(dw, C_), f"step {step}: colors"
        assert events(ref, rw) == events(dev, dw), f"step {step}: events"
        pr, pd = ref.m3d_world_profile(rw), dev.m3d_world_profile(dw)
        for f in ("bodyCount", "awakeBodyCount", "contactCount", "touchingContactCount", "islandCount"):
            assert getattr(pr, f) == getattr(pd, f), f"step {step}: profile {f}"
        assert d["pos"] <= TOL_POS and d["rot"] <= TOL_ROT, f"step {step}: trajectories diverged {d}"
    ref.m3d_world_destroy(rw)
    dev.m3d_world_destroy(dw)


@pytest.mark.gpu
@pytest.mark.parametrize("at", [1, 30, 90, 180])
def test_rain_one_step(ref, dev, at):
    """Mixed sphere/capsule/box rain pile (1500 bodies, rain-pile size and
    shape distributions) at several stages of settling."""
    rw = ref.create_world()
    scenes.rain_pile(ref, rw, count=1500, column=(12.0, 12.0, 20.0))
    for _ in range(at):
        ref.m3d_world_step(rw, DT, 4)
    _check_step(ref, dev, rw, f"rain step {at}")
    ref.m3d_world_destroy(rw)


@pytest.mark.gpu
@pytest.mark.parametrize("at", [1, 40])
def test_giant_island_one_step(ref, dev, at):
    """Box pyramid, 90 levels (4095 boxes): one island with far more
    contacts than a workgroup schedules, solved by the whole-device path
    (greedy colors and overflow levels as device fixpoints)."""
    rw = ref.create_world()
    scenes.pyramid(ref, rw, base=90)
    for _ in range(at):
        ref.m3d_world_step(rw, DT, 4)
    a, b, dw = one_step_from(ref, dev, rw)
    st = dev.stats(dw)
    assert st.solvedContacts >= 4096 and st.sweeps > 0, "giant path not taken"
    assert contact_pairs(a) == contact_pairs(b)
    d = state_diff(a, b)
    assert d["pos"] <= TOL_POS and d["rot"] <= TOL_ROT, d
    assert d["awake_equal"]
    dev.m3d_world_destroy(dw)
    ref.m3d_world_destroy(rw)


@pytest.mark.gpu
def test_giant_island_lockstep_colors(ref, dev):
    """Both engines run the 90-level pyramid from scratch; every step the
    giant island's greedy colors must match the reference's buckets."""
    rw, dw = ref.create_world(), dev.create_world()
    scenes.pyramid(ref, rw, base=90)
    scenes.pyramid(dev, dw, base=90)
    for step in range(12):
        ref.m3d_world_step(rw, DT, 4)
        dev.m3d_world_step(dw, DT, 4)
        a, b = ref.export_state(rw), dev.export_state(dw)
        assert contact_pairs(a) == contact_pairs(b), f"step {step}: contact order"
        C_ = a.header.contactCount
        assert ref.contact_colors(rw, C_) == dev.contact_colors(dw, C_), f"step {step}: colors"
        d = state_diff(a, b)
        assert d["pos"] <= TOL_POS and d["rot"] <= TOL_ROT, f"step {step}: {d}"
    ref.m3d_world_destroy(rw)
    dev.m3d_world_destroy(dw)


@pytest.mark.gpu
def test_destroy_after_idle_step(ref, dev):
    """Churn with every third step run without destroys before it: the
    contact index the device downloads at the end of a step that followed
    destroys must not serve destroys after a later step (the contacts have
    changed since); contact order and events must match the reference."""
    rw, dw = ref.create_world(), dev.create_world()
    sr = scenes.Stream(ref, rw, count=1200, churn=12, arena=16.0)
    sd = scenes.Stream(dev, dw, count=1200, churn=12, arena=16.0)
    for step in range(45):
        if step % 3 != 1:
            sr.churn_step()
            sd.churn_step()
        ref.m3d_world_step(rw, DT, 4)
        dev.m3d_world_step(dw, DT, 4)
        a, b = ref.export_state(rw), dev.export_state(dw)
        assert contact_pairs(a) == contact_pairs(b), f"step {step}: contact order"
        (rb, re), (db, de) = events(ref, rw), events(dev, dw)
        assert rb == db, f"step {step}: begin events"
        assert re == de, f"step {step}: end events"
        d = state_diff(a, b)
        assert d["pos"] <= TOL_POS and d["rot"] <= TOL_ROT, f"step {step}: {d}"
    ref.m3d_world_destroy(rw)
    dev.m3d_world_destroy(dw)


@pytest.mark.gpu
def test_stream_churn_lockstep(ref, dev):
    """Streaming churn (destroy oldest + respawn every step) driven through
    the API on both engines; contact order and events must match."""
    rw, dw = ref.create_world(), dev.create_world()
    sr = scenes.Stream(ref, rw, count=1200, churn=12, arena=16.0)
    sd = scenes.Stream(dev, dw, count=1200, churn=12, arena=16.0)
    for step in range(40):
        sr.churn_step()
        sd.churn_step()
        ref.m3d_world_step(rw, DT, 4)
        dev.m3d_world_step(dw, DT, 4)
        a, b = ref.export_state(rw), dev.export_state(dw)
        assert contact_pairs(a) == contact_pairs(b), f"step {step}: contact order"
        (rb, re), (db, de) = events(ref, rw), events(dev, dw)
        assert rb == db, f"step {step}: begin events ref-only {sorted(set(rb) - set(db))} dev-only {sorted(set(db) - set(rb))}"
        assert re == de, f"step {step}: end events ref {re[:8]} dev {de[:8]} ref-only {sorted(set(re) - set(de))} dev-only {sorted(set(de) - set(re))}"
        d = state_diff(a, b)
        assert d["pos"] <= TOL_POS and d["rot"] <= TOL_ROT, f"step {step}: {d}"
    ref.m3d_world_destroy(rw)
    dev.m3d_world_destroy(dw)
