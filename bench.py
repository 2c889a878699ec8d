#!/usr/bin/env python3
"""Melon3D world-step throughput on one MI355X.

A "step" is one m3d_world_step (dt = 1/60 s, 4 substeps) of the whole world.
Default workload: the 20 000-body rain pile (1/3 spheres r~U[0.2,0.5],
1/3 capsules r~U[0.15,0.3] hh~U[0.2,0.6], 1/3 boxes he~U[0.2,0.6], uniform
orientations, friction U[0.3,0.8], dropped from a 40x40x60 m column);
W untimed warmup steps, then K timed steps. Inputs are resident on the
device when timing starts; m3d_world_step returns after the device
finished the step, so the wall clock around the K calls is the step time.

CPU baseline: the reference engine built from its own sources with its
release flags (-O3 -mavx2 -mfma), multithreaded on this host's cores,
stepping a bounded sample from the exact state the GPU had when its timed
region started.

Prints one JSON line (see keys below).
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


def traffic_bytes(st, substeps: int) -> float:
    """Compulsory HBM bytes of one step (fp32 SoA, every live field read or
    written once per pass): per awake body per substep 184 B, per manifold
    point 52 B (+24 B friction anchors for box-box points), per solved
    contact 40 B; per body broadphase 88 B (+24 B on re-fat); 32 B per grid
    entry; 640 B per regenerated manifold, 120 B per cached contact; islands
    and coloring 16 B per solved contact + 12 B per body."""
    S = substeps
    N = st.bodyCapacity
    per_sub = 184 * st.awakeDynamicBodies + 52 * st.solvedPoints + 24 * st.boxBoxPoints + 40 * st.solvedContacts
    once = (88 * N + 24 * st.refattedBodies + 32 * st.activeEntries + 640 * st.regeneratedManifolds +
            120 * max(st.contacts - st.regeneratedManifolds, 0) + 16 * st.solvedContacts + 12 * N)
    return float(S * per_sub + once)


def build_scene(lib, w, scene: str, bodies: int):
    from melon3d_amd import scenes

    if scene == "rain":
        scenes.rain_pile(lib, w, count=bodies, seed=12345, column=(40.0, 40.0, 60.0))
        return None, bodies
    if scene == "pyramid":
        ids = scenes.pyramid(lib, w, base=20)
        return None, len(ids)
    if scene == "stream":
        st = scenes.Stream(lib, w, count=bodies, churn=max(1, bodies // 200), seed=777)
        return st, bodies
    raise SystemExit(f"unknown scene {scene}")


def cpu_baseline(snap, steps_max: int, budget_s: float, substeps: int, dyn: int):
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libm3dref_fast.so")
    if not os.path.exists(ref_path):
        return None
    from melon3d_amd import api

    ref = api.load_library(ref_path)
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    snap.header.workerCount = cores
    rw = ref.import_state(snap)
    t0 = time.perf_counter()
    k = 0
    while k < steps_max:
        ref.m3d_world_step(rw, 1.0 / 60.0, substeps)
        k += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    ref.m3d_world_destroy(rw)
    return {"value": dyn * k / dt, "unit": "body-steps/s", "cores": cores, "kind": "reference",
            "sample": f"{k} steps of the same world from the GPU's timed-region start state, "
                      f"reference built -O3 -mavx2 -mfma, workerCount={cores}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--scene", default="rain", choices=["rain", "stream", "pyramid"])
    ap.add_argument("--bodies", type=int, default=20000)
    ap.add_argument("--substeps", type=int, default=4)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported "
                                   "(batched worlds shard across GPUs by world range; see DESIGN.md)"}))
        sys.exit(2)

    from melon3d_amd import api

    dev = api.load()
    w = dev.create_world()
    stream, dyn = build_scene(dev, w, args.scene, args.bodies)
    for i in range(args.warmup):
        if stream:
            stream.churn_step()
        dev.m3d_world_step(w, 1.0 / 60.0, args.substeps)
        if (i + 1) % 100 == 0:
            print(f"# warmup {i + 1}/{args.warmup}", file=sys.stderr, flush=True)
    snap = None if args.no_cpu else dev.export_state(w)

    traffic = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        if stream:
            stream.churn_step()
        dev.m3d_world_step(w, 1.0 / 60.0, args.substeps)
        traffic.append(traffic_bytes(dev.stats(w), args.substeps))
    elapsed = time.perf_counter() - t0
    prof = dev.m3d_world_profile(w)
    st = dev.stats(w)
    ms = 1000.0 * elapsed / args.steps
    value = dyn * args.steps / elapsed
    bytes_per_step = sum(traffic) / len(traffic)
    achieved = bytes_per_step / (elapsed / args.steps) / 1e9

    cpu = None
    if snap is not None:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(snap, args.cpu_steps, args.cpu_budget, args.substeps, dyn)

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
        "config": {"workload": f"{args.scene} {dyn} bodies, dt 1/60 s, {args.substeps} substeps",
                   "bodies": dyn, "contacts": prof.contactCount, "touching": prof.touchingContactCount,
                   "islands": prof.islandCount, "awake": prof.awakeBodyCount, "sweeps": st.sweeps,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "scope": "whole step (algorithmic bytes / step wall time)"},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    dev.m3d_world_destroy(w)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
