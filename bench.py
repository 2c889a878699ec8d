#!/usr/bin/env python3
"""Melon3D world-step throughput on one MI355X.

A "step" is one m3d_world_step (dt = 1/60 s, 4 substeps) of the whole world.

Default workload: the streaming scene, 200 000 bodies in a 200 x 200 m arena
(1/3 spheres r~U[0.2,0.5], 1/3 capsules r~U[0.15,0.3] hh~U[0.2,0.6], 1/3
boxes he~U[0.2,0.6], uniform orientations, friction U[0.3,0.8], seed 777);
every step despawns the 1000 oldest bodies and spawns 1000 at y~U[30,60]
with downward speed U[5,20] m/s, in native code, inside the timed region
(as the reference's bench does, bench/bench.cpp:243-253). `--scene rain`
runs the 20 000-body rain pile instead.

W untimed warmup steps, then K timed steps. State is resident on the device
when timing starts; m3d_world_step returns after the device finished the
step, so the wall clock around the K steps is the step time.

roofline: the dominant kernel of the timed window: a device world imported
from the state the timed region started from is stepped again over the same
K steps, with the same churn, with HIP events around every timed kernel
(solver passes, prepare, body integration, ordering, narrowphase, pairs and
every radix sort); the kernel with the largest summed time is the dominant
one. achieved = its algorithmic bytes per launch (the library's byte models,
from the step counters) / its average launch duration. `traffic` is its
measured HBM bytes per launch (rocprofv3 PMC FETCH_SIZE and WRITE_SIZE,
scaled by a copy-kernel calibration) from the committed profile of this
workload and window, or null when none matches.

cpu_baseline: the reference engine built from its own sources with its
release flags (-O3 -mavx2 -mfma), workerCount = this host's cores, stepping
a bounded sample from the exact state the GPU had when its timed region
started (same churn sequence); cpu_baseline_1thread the same with
workerCount = 1 (bench/bench.cpp:266-273 reports both).

The stream line also carries the two other parts of the metric:
`batch` (4096 worlds of the batched workload: body-steps/s over steps
41-240, the same worlds looped as plain device worlds, and the reference
on a 256-world sample, one world per host thread) and `diff` (the
differentiable batch: forward and backward times, backward / forward, and
gradient errors against FP64 central differences and FP64 CPU tangents).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
DT = 1.0 / 60.0


def traffic_bytes(st, substeps: int) -> float:
    """Compulsory HBM bytes of one whole step (fp32 SoA, every live field
    read or written once per pass): per awake body per substep 184 B, per
    manifold point 52 B (+24 B friction anchors for box-box points), per
    solved contact 40 B; per body broadphase 88 B (+24 B on re-fat); 32 B per
    grid entry; 640 B per regenerated manifold, 120 B per cached contact;
    islands and coloring 16 B per solved contact + 12 B per body."""
    S = substeps
    N = st.bodyCapacity
    per_sub = 184 * st.awakeDynamicBodies + 52 * st.solvedPoints + 24 * st.boxBoxPoints + 40 * st.solvedContacts
    once = (88 * N + 24 * st.refattedBodies + 32 * st.activeEntries + 640 * st.regeneratedManifolds +
            120 * max(st.contacts - st.regeneratedManifolds, 0) + 16 * st.solvedContacts + 12 * N)
    return float(S * per_sub + once)


def measured_traffic(scene: str, size: int, kernel: str, window=None):
    """HBM bytes per launch of `kernel` measured with rocprofv3 PMC counters
    (FETCH_SIZE and WRITE_SIZE, scaled by a copy-kernel calibration) on this
    workload, from the committed profile summary: lines
    `traffic_bytes_per_launch <kernel> <bytes> <warmup> <steps>`, the profiled
    run's window, which must be this run's `window` when one is given. None
    when not measured."""
    files = {("stream", 200000): ("stream200k_window.md", "stream200k_steady.md"),
             ("batch", 4096): ("batch4096_kernels.md",), ("diff", 256): ("diff256_adjoint.md",)}
    for name in files.get((scene, size), ()):
        try:
            with open(os.path.join(ROOT, "profiles", name)) as f:
                for line in f:
                    parts = line.split()
                    if len(parts) >= 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                        if window is not None and (len(parts) < 5 or (int(parts[3]), int(parts[4])) != tuple(window)):
                            continue
                        return float(parts[2])
        except OSError:
            continue
    return None


def build_scene(lib, w, scene: str, bodies: int):
    from melon3d_amd import scenes

    if scene == "rain":
        scenes.native_rain_pile(lib, w, count=bodies, seed=12345, column=(40.0, 40.0, 60.0))
        return None, bodies
    if scene == "pyramid":
        ids = scenes.pyramid(lib, w, base=20)
        return None, len(ids)
    if scene == "stream":
        # the device takes each step's churn in one bulk call (m3dx_body_respawn:
        # the same ids and results as the per-body calls the reference makes)
        st = scenes.NativeStream(lib, w, count=bodies, churn=max(1, bodies // 200), seed=777, arena=200.0,
                                 bulk=True)
        return st, bodies
    raise SystemExit(f"unknown scene {scene}")


def batch_cpu_baseline(worlds: int, warmup: int, steps_max: int, budget_s: float, substeps: int, threads: int):
    """Reference engine on a sample of the batched workload: worlds 0 ..
    worlds-1, each its own world with workerCount = 1, spread over a pool of
    host threads (one world per task), advanced `warmup` untimed steps, then
    timed steps until the budget is spent."""
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libm3dref_fast.so")
    if not os.path.exists(ref_path):
        return None
    from concurrent.futures import ThreadPoolExecutor

    from melon3d_amd import api, scenes

    ref = api.load_library(ref_path)
    cores = threads
    wd = ref.m3d_world_def_default()
    wd.workerCount = 1
    ws, dyn = [], 0
    for i in range(worlds):
        w = ref.create_world(wd)
        n, _ = scenes.batch_world(ref, w, i)
        ws.append(w)
        dyn += n

    def run(w, k):
        for _ in range(k):
            ref.m3d_world_step(w, DT, substeps)

    with ThreadPoolExecutor(max_workers=cores) as ex:
        list(ex.map(lambda w: run(w, warmup), ws))
        t0 = time.perf_counter()
        k = 0
        while k < steps_max:
            list(ex.map(lambda w: run(w, 1), ws))
            k += 1
            if time.perf_counter() - t0 > budget_s:
                break
        dt = time.perf_counter() - t0
    for w in ws:
        ref.m3d_world_destroy(w)
    return {"value": dyn * k / dt, "unit": "body-steps/s", "cores": cores, "kind": "reference",
            "sample": f"batched-workload worlds 0..{worlds - 1} ({dyn} dynamic bodies), {k} steps after {warmup} "
                      f"untimed, one world per task on {cores} threads, workerCount=1, reference built "
                      f"-O3 -mavx2 -mfma"}


def batch_loop_plain(dev, worlds: int, warmup: int, steps: int, substeps: int):
    """The same sample stepped as plain device worlds, one m3d_world_step per
    world (the loop a batch replaces)."""
    from melon3d_amd import scenes

    ws, dyn = [], 0
    for i in range(worlds):
        w = dev.create_world()
        n, _ = scenes.batch_world(dev, w, i)
        ws.append(w)
        dyn += n
    for _ in range(warmup):
        for w in ws:
            dev.m3d_world_step(w, DT, substeps)
    t0 = time.perf_counter()
    for _ in range(steps):
        for w in ws:
            dev.m3d_world_step(w, DT, substeps)
    dt = time.perf_counter() - t0
    for w in ws:
        dev.m3d_world_destroy(w)
    return {"value": dyn * steps / dt, "unit": "body-steps/s", "sample": f"worlds 0..{worlds - 1} ({dyn} dynamic "
                                                                           f"bodies), {steps} steps after {warmup}"}


def run_diff(args, dev):
    """Differentiable worlds: forward (the engine's steps, recording each
    step's solve inputs) and backward (the parallel adjoint of all recorded
    steps, states recomputed from checkpoints every 10 steps) of the
    final-COM loss; the plain forward (no recording) for comparison; the
    gradient checked on a 16-world subset against FP64 central differences
    (3 steps, double replay) and against FP64 forward tangents on the host
    CPU along the engine's path (all steps)."""
    import ctypes as C

    import numpy as np

    from melon3d_amd import diff

    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    scene = diff.DiffScene(worlds=args.worlds)
    w = scene.build(dev)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dev.m3d_world_step(w, DT, args.substeps)
    fwd_plain = time.perf_counter() - t0
    check_steps(dev, w)
    dev.m3d_world_destroy(w)
    w = scene.build(dev)
    if dev.m3dx_diff_record(w, 1) != 0:
        raise SystemExit(dev.last_error())
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dev.m3d_world_step(w, DT, args.substeps)
    fwd = time.perf_counter() - t0
    check_steps(dev, w)
    dev.m3dx_diff_record(w, 0)
    engine = diff.final_positions(dev, w, scene)
    m = scene.mass[..., None]
    com = (m * engine).sum(1) / scene.mass.sum(1)[:, None]
    g_pos = (2.0 * (com - scene.target))[:, None, :] * m / scene.mass.sum(1)[:, None, None]
    print(f"# forward {fwd:.2f} s, backward ...", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    gv, gf, gm = diff.backward(dev, w, scene, g_pos)
    bwd = time.perf_counter() - t0
    # the same backward again with HIP events around every adjoint launch
    dev.m3dx_set_kernel_timing(w, 1)
    diff.backward(dev, w, scene, g_pos)
    kt = dev.kernel_timing(w)
    dev.m3dx_set_kernel_timing(w, 0)
    dev.m3d_world_destroy(w)
    launches, kms, kbytes = kt["adjoint"]
    avg_s = (kms / 1000.0) / launches if launches else 0.0
    achieved = (kbytes / launches) / avg_s / 1e9 if launches and avg_s > 0 else 0.0

    def loss_grads(L, cw, cs, steps):
        e = diff.final_positions(L, cw, cs)
        cm = cs.mass[..., None]
        msum = cs.mass.sum(1)[:, None, None]
        ccom = (cm * e).sum(1) / cs.mass.sum(1)[:, None]
        g_com = 2.0 * (ccom - cs.target)
        cgv, cgf, cgm = diff.backward(L, cw, cs, g_com[:, None, :] * cm / msum)
        return e, ccom, g_com, cgv, cgf, cgm

    # gradient check 1: per-world directional derivatives vs FP64 central
    # differences of the double replay, 3 steps
    cs = diff.DiffScene(worlds=16)
    cw = cs.build(dev)
    dev.m3dx_diff_record(cw, 1)
    for _ in range(3):
        dev.m3d_world_step(cw, DT, args.substeps)
    dev.m3dx_diff_record(cw, 0)
    e, ccom, g_com, cgv, cgf, cgm = loss_grads(dev, cw, cs, 3)
    msum = cs.mass.sum(1)[:, None, None]
    cgm = cgm + (g_com[:, None, :] * ((e - ccom[:, None, :]) / msum)).sum(2)
    rng = np.random.default_rng(11)
    dv, df, dm = rng.normal(size=cs.v0.shape), rng.normal(size=cs.friction.shape) * 0.05, rng.normal(
        size=cs.mass.shape) * 10.0
    adj = (cgv * dv).sum((1, 2)) + (cgf * df).sum(1) + (cgm * dm).sum(1)

    def losses(eps):
        st = diff.replay(dev, cw, cs, 3, precision=1, v0=cs.v0 + eps * dv, friction=cs.friction + eps * df,
                         mass=cs.mass + eps * dm)
        pos = st[cs.dynamic_index(), 0:3]
        mm = (cs.mass + eps * dm)[..., None]
        c = (mm * pos).sum(1) / mm.sum(1)
        return ((c - cs.target) ** 2).sum(1)

    fd = (losses(1e-4) - losses(-1e-4)) / 2e-4
    rel_fd = np.abs(adj - fd) / np.maximum(np.abs(fd), 1e-12)
    dev.m3d_world_destroy(cw)

    # gradient check 2: the same subset over all steps against FP64 forward
    # tangents on the host CPU along the engine's path, directions in v0,
    # friction and mass (the recorded values' derivatives)
    cw = cs.build(dev)
    dev.m3dx_diff_record(cw, 1)
    for _ in range(args.steps):
        dev.m3d_world_step(cw, DT, args.substeps)
    dev.m3dx_diff_record(cw, 0)
    e, ccom, g_com, cgv, cgf, cgm = loss_grads(dev, cw, cs, args.steps)
    n = cs.body_count()
    idx = cs.dynamic_index()
    st0 = np.zeros((n, 13), np.float32)
    dev.m3dx_diff_state(cw, 0, ptr(st0))
    u = np.zeros((n, 13))
    u[idx, 7:10] = dv
    dmass = np.zeros(n)
    dmass[idx] = dm
    dfric = np.zeros(n)
    dfric[idx] = df
    out = np.zeros((n, 13))
    dout = np.zeros((n, 13))
    t_cpu = time.perf_counter()
    if dev.m3dx_diff_jvp_cpu(cw, 0, args.steps, 0, ptr(st0.astype(np.float64)), ptr(u), None, ptr(dmass), None,
                             ptr(dfric), ptr(out), ptr(dout)) != 0:
        raise SystemExit(dev.last_error())
    t_cpu = time.perf_counter() - t_cpu
    dx = dout[idx, 0:3]
    mm = cs.mass[..., None]
    # loss along the direction: d|COM - t|^2 with COM = sum(m x) / sum(m), the
    # recorded masses' own perturbation entering the COM weights too
    dcom = (mm * dx).sum(1) / cs.mass.sum(1)[:, None] + (dm[..., None] * (e - ccom[:, None, :])).sum(1) / \
        cs.mass.sum(1)[:, None]
    jvp = (g_com * dcom).sum(1)
    cgm_total = cgm + (g_com[:, None, :] * ((e - ccom[:, None, :]) / msum)).sum(2)
    vjp = (cgv * dv).sum((1, 2)) + (cgf * df).sum(1) + (cgm_total * dm).sum(1)
    rel_cpu = np.abs(vjp - jvp) / np.maximum(np.abs(jvp), 1e-12)
    dev.m3d_world_destroy(cw)

    dyn = args.worlds * diff.BODIES
    out = {
        "metric": "body-steps/sec",
        "value": dyn * args.steps / fwd,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 0,
        "ms_per_step": 1000.0 * fwd / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"differentiable: {args.worlds} worlds x 64 bodies (spheres r 0.5, cubes he 0.5), "
                               f"v0 ~ N(0, 2), friction ~ U[0.2, 0.9], loss |COM - target|^2 after {args.steps} "
                               f"steps, dt 1/60 s, {args.substeps} substeps, checkpoints every 10 steps",
                   "worlds": args.worlds, "bodies": dyn, "parallelism": "single GPU"},
        "forward_s": fwd,
        "forward_plain_s": fwd_plain,
        "backward_s": bwd,
        "backward_over_forward": bwd / fwd,
        "backward_over_plain_forward": bwd / fwd_plain,
        "grad_check": {"subset_worlds": 16,
                       "fp64_central_differences": {"steps": 3, "h": 1e-4, "max_rel_err": float(rel_fd.max()),
                                                    "median_rel_err": float(np.median(rel_fd))},
                       "fp64_cpu_tangents": {"steps": args.steps, "max_rel_err": float(rel_cpu.max()),
                                             "median_rel_err": float(np.median(rel_cpu)), "cpu_s": t_cpu}},
        "grad_norms": {"v0": float(np.linalg.norm(gv)), "friction": float(np.linalg.norm(gf)),
                       "mass": float(np.linalg.norm(gm))},
        "roofline": {"bound": "hbm", "kernel": "adjoint", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("diff", args.worlds, "adjoint"),
                     "launches": launches,
                     "avg_launch_us": avg_s * 1e6, "bytes_per_launch": (kbytes / launches) if launches else 0.0},
        "kernel_ms_backward": {n: v[1] for n, v in kt.items() if v[0] > 0},
        "cpu_baseline": None,
    }
    return out


def run_batch(args, dev):
    from melon3d_amd import scenes

    t_build = time.perf_counter()
    w, info = scenes.batch(dev, args.worlds)
    dyn = sum(n for n, _ in info)
    print(f"# batch of {args.worlds} worlds ({dyn} dynamic bodies) built in {time.perf_counter() - t_build:.1f} s",
          file=sys.stderr, flush=True)
    for i in range(args.warmup):
        dev.m3d_world_step(w, DT, args.substeps)
        if (i + 1) % 20 == 0:
            print(f"# warmup {i + 1}/{args.warmup}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        dev.m3d_world_step(w, DT, args.substeps)
    elapsed = time.perf_counter() - t0
    check_steps(dev, w)
    prof = dev.m3d_world_profile(w)
    st = dev.stats(w)
    value = dyn * args.steps / elapsed
    pass_blocks = dev.m3dx_pass_blocks(w)
    graph_steps = dev.m3dx_step_graph_launches(w)
    dev.m3dx_set_kernel_timing(w, 1)
    for i in range(args.timing_steps):
        dev.m3d_world_step(w, DT, args.substeps)
    kt = dev.kernel_timing(w)
    dev.m3dx_set_kernel_timing(w, 0)
    dev.m3d_world_destroy(w)
    roof = roofline_of(kt, "batch", args.worlds, None)
    loop = None
    if not args.no_loop:
        print("# looping plain worlds ...", file=sys.stderr, flush=True)
        loop = batch_loop_plain(dev, args.loop_worlds, args.loop_warmup, args.loop_steps, args.substeps)
    cpu = None
    if not args.no_cpu:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        cpu = batch_cpu_baseline(args.cpu_worlds, args.warmup, args.cpu_steps, args.cpu_budget, args.substeps,
                                 threads=host_cores())
    out = {
        "metric": "body-steps/sec",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batch of {args.worlds} worlds (world w: seed 1000003(w+1), "
                               f"floor(exp U[ln 16, ln 2049)) bodies, pile or pyramid), dt 1/60 s, "
                               f"{args.substeps} substeps",
                   "worlds": args.worlds, "bodies": dyn, "mean_bodies": dyn / args.worlds,
                   "contacts": prof.contactCount, "islands": prof.islandCount, "awake": prof.awakeBodyCount,
                   "pass_workgroups": pass_blocks, "graph_steps": graph_steps,
                   "parallelism": "single GPU, one launch sequence for all worlds"},
        "roofline": roof,
        "kernel_ms_per_step": {n: v[1] / max(args.timing_steps, 1) for n, v in kt.items() if v[0] > 0},
        "kernel_timing": f"HIP events over {args.timing_steps} steps after the timed region",
        "loop_plain_gpu": loop,
        "cpu_baseline": cpu,
    }
    if loop:
        out["batched_over_loop"] = value / loop["value"]
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    return out


def check_steps(dev, w):
    """A step that stopped on an error would make the timing meaningless."""
    n = dev.m3dx_step_failures(w)
    if n:
        raise SystemExit(f"{n} steps failed: {dev.last_error()}")


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def make_rays(n: int, seed: int, arena: float):
    """n segments over the arena: origins x, z ~ U[-arena/2, arena/2], y ~
    U[1, 60]; unit directions with a downward bias; length 60 m."""
    import numpy as np

    rng = np.random.default_rng(seed)
    o = np.column_stack([rng.uniform(-arena / 2, arena / 2, n), rng.uniform(1.0, 60.0, n),
                         rng.uniform(-arena / 2, arena / 2, n)]).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d[:, 1] -= 1.5
    d = (60.0 * d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return np.ascontiguousarray(o), np.ascontiguousarray(d)


def ray_bench(dev, w, n: int, arena: float):
    """m3dx_world_ray_cast_batch of n rays: end-to-end rays/s (host arrays in
    and out) and the hit count."""
    import ctypes as C

    from melon3d_amd.api import m3d_ray_result

    o, d = make_rays(n, 99, arena)
    out = (m3d_ray_result * n)()
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    dev.m3dx_world_ray_cast_batch(w, p(o[:1]), p(d[:1]), None, 1, out)  # first-call setup
    t0 = time.perf_counter()
    if dev.m3dx_world_ray_cast_batch(w, p(o), p(d), None, n, out) != 0:
        raise SystemExit(dev.last_error())
    dt = time.perf_counter() - t0
    hits = sum(1 for k in range(0, n, 97) if out[k].hit)
    return n / dt, hits * 97 / n


def cpu_baseline(snap, stream, steps_max: int, budget_s: float, substeps: int, dyn: int, threads: int,
                 rays: int = 0, arena: float = 200.0):
    """The reference's release build (-O3 -mavx2 -mfma) stepping the GPU's
    timed-region start state with the same churn, workerCount = threads, on
    a bounded sample (steps until the budget is spent)."""
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libm3dref_fast.so")
    if not os.path.exists(ref_path):
        return None
    from melon3d_amd import api

    ref = api.load_library(ref_path)
    snap.header.workerCount = threads
    rw = ref.import_state(snap)
    rs = stream.clone_onto(ref, rw) if stream is not None else None
    t0 = time.perf_counter()
    k = 0
    while k < steps_max:
        if rs is not None:
            rs.churn_step()
        ref.m3d_world_step(rw, DT, substeps)
        k += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    ray_rate = None
    if rays:
        from melon3d_amd.api import m3d_vec3

        o, d = make_rays(rays, 99, arena)
        t1 = time.perf_counter()
        for j in range(rays):
            ref.m3d_world_ray_cast(rw, m3d_vec3(*map(float, o[j])), m3d_vec3(*map(float, d[j])))
        ray_rate = rays / (time.perf_counter() - t1)
    ref.m3d_world_destroy(rw)
    out = {"value": dyn * k / dt, "unit": "body-steps/s", "cores": threads, "kind": "reference",
           "sample": f"{k} steps ({'with the same churn, ' if rs else ''}from the GPU's timed-region start state), "
                     f"reference -O3 -mavx2 -mfma, workerCount={threads}"}
    if rays:
        out["ray_cast_rays_per_s"] = ray_rate
    return out


def kernel_window(dev, snap, stream, steps: int, substeps: int):
    """Per-kernel HIP-event times and algorithmic bytes of exactly the timed
    window: a device world imported from the state the timed region started
    from, stepped again with the same churn with kernel timing on (the events
    cost host time, so they stay out of the timed region; the replay is the
    same computation, the engine being deterministic)."""
    w = dev.import_state(snap)
    rs = stream.clone_onto(dev, w, bulk=True) if stream is not None else None
    dev.m3dx_set_kernel_timing(w, 1)
    for _ in range(steps):
        if rs is not None:
            rs.churn_step()
        dev.m3d_world_step(w, DT, substeps)
    check_steps(dev, w)
    kt = dev.kernel_timing(w)
    dev.m3dx_set_kernel_timing(w, 0)
    dev.m3d_world_destroy(w)
    return kt


def roofline_of(kt, scene: str, size: int, window):
    """The dominant kernel (largest summed event time over all timed kernels)
    with its achieved algorithmic bandwidth and, when a committed profile of
    this workload and window exists, its measured HBM traffic per launch."""
    name, (launches, kms, kbytes) = max(((n, v) for n, v in kt.items() if v[0] > 0), key=lambda x: x[1][1],
                                        default=("none", (0, 0.0, 0.0)))
    avg_s = (kms / 1000.0) / launches if launches else 0.0
    per_launch = kbytes / launches if launches else 0.0
    achieved = per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    return {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(scene, size, name, window),
            "launches": launches, "avg_launch_us": avg_s * 1e6, "bytes_per_launch": per_launch}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scene", default="stream", choices=["stream", "rain", "pyramid", "batch", "diff"])
    ap.add_argument("--worlds", type=int, default=4096, help="batch: number of worlds")
    ap.add_argument("--batch-warmup", type=int, default=40)
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--loop-worlds", type=int, default=256, help="batch: worlds stepped one by one for comparison")
    ap.add_argument("--loop-warmup", type=int, default=10)
    ap.add_argument("--loop-steps", type=int, default=5)
    ap.add_argument("--cpu-worlds", type=int, default=256, help="batch: worlds in the CPU baseline sample")
    ap.add_argument("--diff-worlds", type=int, default=256)
    ap.add_argument("--diff-steps", type=int, default=120)
    ap.add_argument("--no-loop", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="stream line without the batch sub-object")
    ap.add_argument("--no-diff", action="store_true", help="stream line without the diff sub-object")
    ap.add_argument("--bodies", type=int, default=None)
    ap.add_argument("--substeps", type=int, default=4)
    ap.add_argument("--timing-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 256 MiB device to device (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_probe:
        # 256 MiB device-to-device copies (the runtime's copy kernel): the
        # known bytes that calibrate FETCH_SIZE and WRITE_SIZE
        import ctypes as C

        hip = C.CDLL("libamdhip64.so")
        nbytes = 256 << 20
        a, b = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
        hip.hipMemset(a, 1, C.c_size_t(nbytes))
        for _ in range(3):
            hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
        hip.hipDeviceSynchronize()
        hip.hipFree(a)
        hip.hipFree(b)
        print(json.dumps({"copy_probe_bytes": nbytes}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                                   "(batched worlds would shard across GPUs by world range)"}))
        sys.exit(2)
    from melon3d_amd import api

    dev = api.load()
    if args.scene == "batch":
        args.warmup, args.steps = args.batch_warmup, args.batch_steps
        print(json.dumps(run_batch(args, dev)), flush=True)
        return
    if args.scene == "diff":
        args.worlds, args.steps = args.diff_worlds, args.diff_steps
        print(json.dumps(run_diff(args, dev)), flush=True)
        return
    out = run_single(args, dev)
    if args.scene == "stream" and not args.no_batch:
        print("# batch ...", file=sys.stderr, flush=True)
        b = argparse.Namespace(**vars(args))
        b.warmup, b.steps = args.batch_warmup, args.batch_steps
        out["batch"] = run_batch(b, dev)
    if args.scene == "stream" and not args.no_diff:
        print("# differentiable batch ...", file=sys.stderr, flush=True)
        d = argparse.Namespace(**vars(args))
        d.worlds, d.steps = args.diff_worlds, args.diff_steps
        out["diff"] = run_diff(d, dev)
    print(json.dumps(out), flush=True)


def run_single(args, dev):
    """One world (the stream by default): W untimed steps, K timed steps."""
    if args.bodies is None:
        args.bodies = {"stream": 200000, "rain": 20000, "pyramid": 210}[args.scene]
    w = dev.create_world()
    t_build = time.perf_counter()
    stream, dyn = build_scene(dev, w, args.scene, args.bodies)
    print(f"# scene built in {time.perf_counter() - t_build:.1f} s", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        if stream:
            stream.churn_step()
        dev.m3d_world_step(w, DT, args.substeps)
        if (i + 1) % 50 == 0:
            print(f"# warmup {i + 1}/{args.warmup}", file=sys.stderr, flush=True)
    # the state the timed region starts from, and the churn's state there
    snap = dev.export_state(w)
    stream0 = stream.clone_onto(dev, w) if stream else None

    model_bytes = 0.0
    churn_s = 0.0
    prof_ms = [0.0, 0.0, 0.0]
    t0 = time.perf_counter()
    for i in range(args.steps):
        if stream:
            tc = time.perf_counter()
            stream.churn_step()
            churn_s += time.perf_counter() - tc
        dev.m3d_world_step(w, DT, args.substeps)
        model_bytes += traffic_bytes(dev.stats(w), args.substeps)
        p = dev.m3d_world_profile(w)
        prof_ms[0] += p.broadphaseMs
        prof_ms[1] += p.narrowphaseMs
        prof_ms[2] += p.solverMs
    elapsed = time.perf_counter() - t0
    check_steps(dev, w)
    prof = dev.m3d_world_profile(w)
    st = dev.stats(w)
    ms = 1000.0 * elapsed / args.steps
    value = dyn * args.steps / elapsed

    print("# kernel timing of the timed window ...", file=sys.stderr, flush=True)
    kt = kernel_window(dev, snap, stream0, args.steps, args.substeps)
    roof = roofline_of(kt, args.scene, args.bodies, (args.warmup, args.steps))

    cpu = cpu1 = None
    if not args.no_cpu:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        arena = 200.0 if args.scene == "stream" else 40.0
        cpu = cpu_baseline(snap, stream0, args.cpu_steps, args.cpu_budget, args.substeps, dyn, host_cores(),
                           rays=2000, arena=arena)
        cpu1 = cpu_baseline(snap, stream0, args.cpu_steps, args.cpu_budget, args.substeps, dyn, 1)
    rays = None
    if args.scene in ("stream", "rain"):
        rate, frac = ray_bench(dev, w, 1 << 20, 200.0 if args.scene == "stream" else 40.0)
        rays = {"rays": 1 << 20, "device_batch_rays_per_s": rate, "hit_fraction": frac,
                "reference_cpu_rays_per_s": cpu.get("ray_cast_rays_per_s") if cpu else None,
                "query": "m3dx_world_ray_cast_batch (host arrays in and out) vs the reference's m3d_world_ray_cast "
                         "called once per ray through ctypes (2000 rays) on the same state; segments of 60 m over "
                         "the arena"}
    kernel_total = sum(v[1] for v in kt.values()) / args.steps
    out = {
        "metric": "body-steps/sec",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.scene} {dyn} bodies" + (f", {stream and max(1, dyn // 200)} spawned and "
                                                                   f"despawned per step" if stream else "")
                               + f", dt 1/60 s, {args.substeps} substeps",
                   "bodies": dyn, "contacts": prof.contactCount, "touching": prof.touchingContactCount,
                   "islands": prof.islandCount, "awake": prof.awakeBodyCount, "giant_sweeps": st.sweeps,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "step_traffic_model": {"bytes_per_step": model_bytes / args.steps,
                               "achieved_GBps": model_bytes / elapsed / 1e9},
        "kernel_ms_per_step": {n: v[1] / args.steps for n, v in kt.items() if v[0] > 0},
        "timed_kernels_ms_per_step": kernel_total,
        "host_split_ms_per_step": {"churn_api_calls": 1000.0 * churn_s / args.steps,
                                   "step_broadphase": prof_ms[0] / args.steps,
                                   "step_narrowphase": prof_ms[1] / args.steps,
                                   "step_solver": prof_ms[2] / args.steps},
        "cpu_baseline": cpu,
        "cpu_baseline_1thread": cpu1,
        "ray_cast": rays,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    if cpu1:
        out["gpu_over_cpu_1thread"] = value / cpu1["value"]
    dev.m3d_world_destroy(w)
    return out


if __name__ == "__main__":
    main()
