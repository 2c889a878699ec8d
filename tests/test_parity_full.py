"""The full-size workloads against the reference engine, and the device
against the reference built on its own system libm.

Bar (as in test_parity.py): contact order, touching flags, grid tiers,
island labels and colors exact; one step from identical state within
1e-5 m and 1e-5 rad.

- Rain pile: 20,000 mixed bodies in the 40 x 40 x 60 m column, stepped 300
  times on the device (one large island: the colored path), then one step
  on both engines from the device's state.
- Stream: 200,000 bodies in the 200 x 200 m arena after 200 steps of 1000
  spawns and 1000 despawns, then one more churn and step on both engines
  from the device's state (the churn continues on the reference by the
  same driver, so ids and free-list reuse are part of the comparison).
- Batch: all 4096 sub-worlds of the batched workload against 4096
  reference worlds, stepped side by side from their creation.
- libm: the device against the strict reference build that keeps the
  system's sinf, cosf, sincosf and atan2f (the parity build interposes
  correctly rounded ones, as the device evaluates them), one step from
  identical state on the pyramid, rain and churned stream scenes.
"""
import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from melon3d_amd import scenes
from parity import events, np_bodies, np_contacts, pairs_np, rot_angle, state_diff_np, state_diff_rows

TOL = 1e-5
DT = 1.0 / 60.0
CORES = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 8))


def _stage_views_equal(ref, rw, dev, dw, n, label):
    assert ref.grid_entries(rw, 0) == dev.grid_entries(dw, 0), f"{label}: active tier"
    assert ref.grid_entries(rw, 1) == dev.grid_entries(dw, 1), f"{label}: stable tier"
    assert ref.oversized(rw, 0) == dev.oversized(dw, 0), f"{label}: active coarse list"
    assert ref.oversized(rw, 1) == dev.oversized(dw, 1), f"{label}: stable coarse list"
    assert ref.islands(rw, n) == dev.islands(dw, n), f"{label}: islands"
    cc = ref.m3d_world_profile(rw).contactCount
    assert ref.contact_colors(rw, cc) == dev.contact_colors(dw, cc), f"{label}: colors"


def _compare_step(ref, rw, dev, dw, label, stage_views=True):
    a, b = ref.export_state(rw), dev.export_state(dw)
    assert a.header.contactCount == b.header.contactCount, f"{label}: contact count"
    assert np.array_equal(pairs_np(a), pairs_np(b)), f"{label}: contact order"
    ca, cb = np_contacts(a), np_contacts(b)
    assert np.array_equal(ca["touching"], cb["touching"]), f"{label}: touching flags"
    assert np.array_equal(ca["pointCount"], cb["pointCount"]), f"{label}: manifold point counts"
    d = state_diff_np(a, b)
    assert d["pos"] <= TOL and d["rot"] <= TOL, f"{label}: {d}"
    assert d["awake_equal"], f"{label}: sleep state"
    assert events(ref, rw) == events(dev, dw), f"{label}: events"
    if stage_views:
        _stage_views_equal(ref, rw, dev, dw, a.header.bodyCapacity, label)
    return d


def _per_body(a, b):
    """Per in-use body: max position difference and rotation angle."""
    A, B = np_bodies(a), np_bodies(b)
    m = A["inUse"] != 0
    dp = np.abs(A["position"][m].astype(np.float64) - B["position"][m].astype(np.float64)).max(axis=1)
    dr = rot_angle(A["rotation"][m].astype(np.float64), B["rotation"][m].astype(np.float64))
    return np.maximum(dp, dr)


def _libm_sensitivity(ref, rw, ref_libm, rw_libm, dev, dw, label):
    """The device against the reference on the system libm after the same
    step. Where the two reference builds (correctly rounded and system
    sinf/cosf/atan2f) agree within 1e-5 the device must too; bodies where a
    last-ulp difference of those functions tipped a discrete decision (a
    manifold feature, a re-collision) differ between the two reference builds
    themselves, and there the device's difference must be exactly theirs.
    Contact order and touching flags must agree everywhere."""
    a, c, b = ref.export_state(rw), ref_libm.export_state(rw_libm), dev.export_state(dw)
    assert np.array_equal(pairs_np(c), pairs_np(b)), f"{label}: contact order vs system libm"
    assert np.array_equal(np_contacts(c)["touching"], np_contacts(b)["touching"]), f"{label}: touching vs system libm"
    dev_libm = _per_body(b, c)
    ref_libm_d = _per_body(a, c)
    beyond = dev_libm > TOL
    assert np.array_equal(dev_libm[beyond], ref_libm_d[beyond]), f"{label}: device differs from both references"
    assert (ref_libm_d[beyond] > TOL).all()
    print(f"{label} vs reference on the system libm: {int(beyond.sum())} of {beyond.size} bodies beyond 1e-5 "
          f"(max {dev_libm.max():.3g}), the same bodies and amounts by which the two reference builds differ; "
          f"the rest within {float(dev_libm[~beyond].max(initial=0.0)):.3g}")


def _import_ref(ref, snap, workers=CORES):
    snap.header.workerCount = workers
    return ref.import_state(snap)


@pytest.mark.gpu
def test_rain_pile_20k_one_step_after_300(ref, ref_libm, dev):
    dw = dev.create_world()
    rws = []
    try:
        scenes.native_rain_pile(dev, dw, count=20000)
        for _ in range(300):
            dev.m3d_world_step(dw, DT, 4)
        assert dev.m3dx_step_failures(dw) == 0, dev.last_error()
        snap = dev.export_state(dw)
        for lib in (ref, ref_libm):
            rws.append((lib, _import_ref(lib, snap)))
        dev.m3d_world_step(dw, DT, 4)
        st = dev.stats(dw)
        assert st.coloredIslands >= 1 and st.sweeps > 0, "colored path not taken"
        for lib, rw in rws:
            lib.m3d_world_step(rw, DT, 4)
        d = _compare_step(ref, rws[0][1], dev, dw, "rain 20k, step 301")
        print(f"rain 20k vs reference (correctly rounded libm): {d}")
        _libm_sensitivity(ref, rws[0][1], ref_libm, rws[1][1], dev, dw, "rain 20k")
    finally:
        dev.m3d_world_destroy(dw)
        for lib, rw in rws:
            lib.m3d_world_destroy(rw)


@pytest.mark.gpu
def test_stream_200k_one_step_after_200_churn_steps(ref, ref_libm, dev):
    dw = dev.create_world()
    rws = []
    try:
        sd = scenes.NativeStream(dev, dw, count=200000, churn=1000, bulk=True)
        for _ in range(200):
            sd.churn_step()
            dev.m3d_world_step(dw, DT, 4)
        assert dev.m3dx_step_failures(dw) == 0, dev.last_error()
        snap = dev.export_state(dw)
        streams = []
        for lib in (ref, ref_libm):
            rw = _import_ref(lib, snap)
            rws.append((lib, rw))
            streams.append(sd.clone_onto(lib, rw))
        sd.churn_step()
        dev.m3d_world_step(dw, DT, 4)
        for (lib, rw), s in zip(rws, streams):
            s.churn_step()
            lib.m3d_world_step(rw, DT, 4)
        d = _compare_step(ref, rws[0][1], dev, dw, "stream 200k, step 201")
        print(f"stream 200k vs reference (correctly rounded libm): {d}")
        _libm_sensitivity(ref, rws[0][1], ref_libm, rws[1][1], dev, dw, "stream 200k")
    finally:
        dev.m3d_world_destroy(dw)
        for lib, rw in rws:
            lib.m3d_world_destroy(rw)


@pytest.mark.gpu
def test_batch_4096_worlds_against_reference(ref, dev):
    """Every sub-world of the 4096-world batch against its own reference
    world, side by side from creation: bodies within 1e-5 and contacts in
    the same order after every compared step."""
    S, steps = 4096, 24
    bw, info = scenes.batch(dev, S)
    refs = [None] * S

    def make(s):
        w = ref.create_world()
        scenes.batch_world(ref, w, s)
        refs[s] = w

    def step(s):
        ref.m3d_world_step(refs[s], DT, 4)

    try:
        with ThreadPoolExecutor(CORES) as ex:
            list(ex.map(make, range(S)))
            colored = 0
            for k in range(steps):
                dev.m3d_world_step(bw, DT, 4)
                colored = max(colored, dev.stats(bw).coloredIslands)
                list(ex.map(step, range(S)))
        assert dev.m3dx_step_failures(bw) == 0, dev.last_error()
        assert colored > 0, "no sub-world took the colored path"
        snap = dev.export_state(bw)
        counts = (C.c_int32 * S)()
        assert dev.m3dx_batch_contact_counts(bw, counts, S) == S
        allpairs = pairs_np(snap)
        bodies = np_bodies(snap)
        c0, worst, identical = 0, 0.0, 0
        for s in range(S):
            first, n = C.c_int32(0), C.c_int32(0)
            assert dev.m3dx_batch_body_range(bw, s, C.byref(first), C.byref(n)) == 0
            b0, nb = first.value, n.value
            rs = ref.export_state(refs[s])
            assert rs.header.bodyCapacity == nb, s
            assert rs.header.contactCount == counts[s], f"sub-world {s}: contact count"
            assert np.array_equal(allpairs[c0:c0 + counts[s]] - b0, pairs_np(rs)), f"sub-world {s}: contact order"
            c0 += counts[s]
            d = state_diff_rows(np_bodies(rs), bodies[b0:b0 + nb])
            assert d["pos"] <= TOL and d["rot"] <= TOL, f"sub-world {s}: {d}"
            assert d["awake_equal"], f"sub-world {s}: sleep state"
            worst = max(worst, d["pos"])
            identical += d["identical"]
        print(f"batch 4096 vs reference after {steps} steps: max |dx| {worst:.3g} m, bit-identical worlds {identical}/{S}")
    finally:
        dev.m3d_world_destroy(bw)
        for w in refs:
            if w:
                ref.m3d_world_destroy(w)


def _libm_check(ref_libm, dev, rw_lib, label):
    """One step of the reference world on both engines from its state."""
    snap = ref_libm.export_state(rw_lib)
    dw = dev.import_state(snap)
    r2 = ref_libm.import_state(snap)
    try:
        ref_libm.m3d_world_step(r2, DT, 4)
        dev.m3d_world_step(dw, DT, 4)
        return _compare_step(ref_libm, r2, dev, dw, label, stage_views=True)
    finally:
        dev.m3d_world_destroy(dw)
        ref_libm.m3d_world_destroy(r2)


@pytest.mark.gpu
@pytest.mark.parametrize("scene,at", [("pyramid", 0), ("pyramid", 60), ("pyramid", 300), ("rain", 30),
                                      ("rain", 180), ("giant", 40)])
def test_system_libm_reference_one_step(ref_libm, dev, scene, at):
    rw = ref_libm.create_world()
    try:
        if scene == "pyramid":
            scenes.pyramid(ref_libm, rw, base=20)
        elif scene == "giant":
            scenes.pyramid(ref_libm, rw, base=90)
        else:
            scenes.rain_pile(ref_libm, rw, count=1500, column=(12.0, 12.0, 20.0))
        for _ in range(at):
            ref_libm.m3d_world_step(rw, DT, 4)
        d = _libm_check(ref_libm, dev, rw, f"{scene} step {at} (system libm)")
        print(f"{scene} at {at} vs reference on the system libm: {d}")
    finally:
        ref_libm.m3d_world_destroy(rw)


@pytest.mark.gpu
def test_system_libm_reference_churned_stream(ref_libm, dev):
    """The small churn stream run on the system-libm reference for 30 steps,
    then one churn and step on both engines from its state."""
    rw = ref_libm.create_world()
    try:
        sr = scenes.Stream(ref_libm, rw, count=1200, churn=12, arena=16.0)
        for _ in range(30):
            sr.churn_step()
            ref_libm.m3d_world_step(rw, DT, 4)
        d = _libm_check(ref_libm, dev, rw, "churned stream (system libm)")
        print(f"churned stream vs reference on the system libm: {d}")
    finally:
        ref_libm.m3d_world_destroy(rw)
