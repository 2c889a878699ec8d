"""Scene frame dumps in the reference's scene.bin format, for visual and
trajectory diffs of device and CPU runs (examples/dump_scene.cpp:212-237
writes the format; examples/render.py renders it).

Layout (little-endian, no padding):
  int32 frames, int32 bodies;
  per body: int32 shapeType, float radius, float halfHeight,
            float halfExtents x, y, z, int32 colorId, int32 isStatic;
  per frame, per body: 7 float32, position xyz then rotation xyzw.

`dump` steps any library exporting the C ABI (this backend by default) and
writes one frame per step, the poses gathered on the device in one call
(m3dx_world_transforms). `compare` diffs two dumps frame by frame.
`showcase_scene` is the example's scene (examples/dump_scene.cpp:25-210: a
wrecking ball on a chain of ball joints and a rigid distance rod, a box
pyramid, a rain of 46 mixed shapes), built through the C API with the same
float32 inputs, so a dump of it on this backend and the example's own
scene.bin describe the same run.

Command line:
  python -m melon3d_amd.scene_dump dump --out scene.bin [--frames 500]
      [--substeps 8] [--library PATH]
  python -m melon3d_amd.scene_dump compare a.bin b.bin [--tol 1e-5]
"""
from __future__ import annotations

import argparse
import ctypes as C
import math
import os
import struct
import sys

import numpy as np

from .api import (M3D_BODY_DYNAMIC, M3D_BODY_STATIC, M3D_SHAPE_BOX, M3D_SHAPE_CAPSULE, M3D_SHAPE_SPHERE, Library,
                  m3d_ball_joint_def, m3d_body_id, m3d_distance_joint_def, m3d_quat, m3d_transform, m3d_vec3)

BODY_DTYPE = np.dtype([("shapeType", "<i4"), ("radius", "<f4"), ("halfHeight", "<f4"), ("halfExtents", "<f4", (3,)),
                       ("colorId", "<i4"), ("isStatic", "<i4")])
POSE_FLOATS = 7


class SceneWriter:
    """Writes a scene.bin: the header and body table at once, then one frame
    per write_frame (exactly `frames` of them)."""

    def __init__(self, path: str, infos: np.ndarray, frames: int):
        infos = np.asarray(infos, dtype=BODY_DTYPE)
        self.n, self.frames, self.written = len(infos), int(frames), 0
        self.f = open(path, "wb")
        self.f.write(struct.pack("<ii", self.frames, self.n))
        self.f.write(infos.tobytes())

    def write_frame(self, poses: np.ndarray):
        poses = np.ascontiguousarray(poses, dtype="<f4")
        if poses.shape != (self.n, POSE_FLOATS):
            raise ValueError(f"frame of shape {poses.shape}, expected ({self.n}, {POSE_FLOATS})")
        if self.written >= self.frames:
            raise ValueError("more frames than the header declares")
        self.f.write(poses.tobytes())
        self.written += 1

    def close(self):
        self.f.close()
        if self.written != self.frames:
            raise ValueError(f"{self.written} frames written, the header declares {self.frames}")


def read_scene(path: str):
    """(body table as a BODY_DTYPE array, poses of shape (frames, bodies, 7))."""
    with open(path, "rb") as f:
        data = f.read()
    frames, n = struct.unpack_from("<ii", data, 0)
    if frames < 0 or n < 0:
        raise ValueError("bad scene.bin header")
    table_end = 8 + n * BODY_DTYPE.itemsize
    want = table_end + frames * n * POSE_FLOATS * 4
    if len(data) != want:
        raise ValueError(f"scene.bin of {len(data)} bytes, its header needs {want}")
    infos = np.frombuffer(data, dtype=BODY_DTYPE, count=n, offset=8).copy()
    poses = np.frombuffer(data, dtype="<f4", offset=table_end).reshape(frames, n, POSE_FLOATS).copy()
    return infos, poses


def compare(a: np.ndarray, b: np.ndarray, tol: float = 1e-5) -> dict:
    """Frame-by-frame difference of two pose arrays (frames, bodies, 7): the
    largest position difference and rotation angle per frame, the first
    frame beyond tol (-1: none) and whether the dumps are bit-identical."""
    if a.shape != b.shape:
        raise ValueError(f"pose arrays of shapes {a.shape} and {b.shape}")
    pa, pb = a[..., :3].astype(np.float64), b[..., :3].astype(np.float64)
    dpos = np.abs(pa - pb).max(axis=(1, 2)) if a.size else np.zeros(0)
    qa, qb = a[..., 3:].astype(np.float64), b[..., 3:].astype(np.float64)
    dot = np.abs((qa * qb).sum(-1)) / np.maximum(np.linalg.norm(qa, axis=-1) * np.linalg.norm(qb, axis=-1), 1e-30)
    drot = (2.0 * np.arccos(np.clip(dot, 0.0, 1.0))).max(axis=1) if a.size else np.zeros(0)
    beyond = np.nonzero((dpos > tol) | (drot > tol))[0]
    return {"frames": int(a.shape[0]), "max_pos": float(dpos.max(initial=0.0)), "max_rot": float(drot.max(initial=0.0)),
            "pos_per_frame": dpos, "rot_per_frame": drot, "first_beyond": int(beyond[0]) if len(beyond) else -1,
            "identical": bool(np.array_equal(a, b))}


def frame_poses(L: Library, w, ids) -> np.ndarray:
    """(bodies, 7) poses of the given body ids, gathered in one call."""
    n = len(ids)
    arr = (m3d_body_id * n)(*ids)
    out = (m3d_transform * n)()
    if L.m3dx_world_transforms(w, arr, n, out) != 0:
        raise RuntimeError("m3dx_world_transforms failed: " + L.last_error())
    raw = np.frombuffer(out, dtype="<f4").reshape(n, POSE_FLOATS)
    return raw.copy()


# ------------------------------------------------------------ the showcase

f32 = np.float32


def _sinf(x) -> np.float32:
    return f32(math.sin(float(x)))


def _cosf(x) -> np.float32:
    return f32(math.cos(float(x)))


def _v(x, y, z):
    return np.array([x, y, z], dtype=np.float32)


def _axis_angle(axis, angle) -> m3d_quat:
    """m3d_quat_axis_angle (include/melon3d.h:115-122) in float32."""
    axis = np.asarray(axis, dtype=np.float32)
    ln = np.sqrt(f32(f32(axis[0] * axis[0] + axis[1] * axis[1]) + axis[2] * axis[2]))
    u = np.zeros(3, np.float32) if ln < f32(1e-12) else (f32(1.0) / ln) * axis
    half = f32(0.5) * f32(angle)
    s = _sinf(half)
    return m3d_quat(float(s * u[0]), float(s * u[1]), float(s * u[2]), float(_cosf(half)))


class _Rand:
    """The example's rain draws: s = s * 1664525 + 1013904223, u = (s >> 8) /
    0xFFFFFF in float32 (examples/dump_scene.cpp:191-192)."""

    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFF

    def __call__(self) -> np.float32:
        self.s = (self.s * 1664525 + 1013904223) & 0xFFFFFFFF
        return f32(self.s >> 8) / f32(0x00FFFFFF)


def showcase_scene(L: Library, w):
    """The example's scene (examples/dump_scene.cpp:25-210). Returns the body
    ids in creation order and their BODY_DTYPE table."""
    ids, rows = [], []

    def add(bd, sd, shape, r, hh, e, color, static):
        ids.append(L.m3d_body_create(w, C.byref(bd), C.byref(sd)))
        rows.append((shape, r, hh, e, color, 1 if static else 0))
        return ids[-1]

    def add_box(p, he, color, static, q=None):
        bd = L.m3d_body_def_default()
        bd.type = M3D_BODY_STATIC if static else M3D_BODY_DYNAMIC
        bd.position = m3d_vec3(*map(float, p))
        bd.orientation = q if q is not None else m3d_quat(0.0, 0.0, 0.0, 1.0)
        sd = L.m3d_shape_def_default()
        sd.type = M3D_SHAPE_BOX
        sd.halfExtents = m3d_vec3(*map(float, he))
        sd.friction = 0.7
        return add(bd, sd, M3D_SHAPE_BOX, 0.0, 0.0, tuple(map(float, he)), color, static)

    def add_sphere(p, r, color, density, v):
        bd = L.m3d_body_def_default()
        bd.type = M3D_BODY_DYNAMIC
        bd.position = m3d_vec3(*map(float, p))
        bd.linearVelocity = m3d_vec3(*map(float, v))
        sd = L.m3d_shape_def_default()
        sd.type = M3D_SHAPE_SPHERE
        sd.radius = float(r)
        sd.friction = 0.5
        sd.restitution = 0.15
        sd.density = float(density)
        return add(bd, sd, M3D_SHAPE_SPHERE, float(r), 0.0, (0.0, 0.0, 0.0), color, False)

    def add_capsule(p, r, hh, color, q, density, v):
        bd = L.m3d_body_def_default()
        bd.type = M3D_BODY_DYNAMIC
        bd.position = m3d_vec3(*map(float, p))
        bd.orientation = q
        bd.linearVelocity = m3d_vec3(*map(float, v))
        sd = L.m3d_shape_def_default()
        sd.type = M3D_SHAPE_CAPSULE
        sd.radius = float(r)
        sd.halfHeight = float(hh)
        sd.friction = 0.6
        sd.density = float(density)
        return add(bd, sd, M3D_SHAPE_CAPSULE, float(r), float(hh), (0.0, 0.0, 0.0), color, False)

    # ground
    add_box(_v(0, -1, 0), _v(60, 1, 60), 0, True)
    # pyramid of boxes (base 6) at the origin
    he = f32(0.5)
    for row in range(6):
        n = 6 - row
        y = he + f32(2.0) * he * f32(row)
        x0 = -he * f32(n - 1)
        for i in range(n):
            bd = L.m3d_body_def_default()
            bd.type = M3D_BODY_DYNAMIC
            bd.position = m3d_vec3(float(x0 + f32(2.0) * he * f32(i)), float(y), 0.0)
            sd = L.m3d_shape_def_default()
            sd.type = M3D_SHAPE_BOX
            sd.halfExtents = m3d_vec3(0.5, 0.5, 0.5)
            sd.friction = 0.72
            add(bd, sd, M3D_SHAPE_BOX, 0.0, 0.0, (0.5, 0.5, 0.5), 1, False)
    # wrecking rig: gantry, chain of capsule links on ball joints, heavy ball
    anchor = _v(-2.0, 11.1, 0.0)
    link_count, link_r, link_hh, ball_r, theta = 6, f32(0.18), f32(0.55), f32(1.1), f32(1.25)
    link_len = f32(2.0) * (link_hh + link_r)
    add_box(_v(-2.0, 5.6, -2.6), _v(0.28, 5.6, 0.28), 7, True)
    add_box(_v(-2.0, 5.6, 2.6), _v(0.28, 5.6, 0.28), 7, True)
    beam = add_box(_v(-2.0, 11.4, 0.0), _v(0.3, 0.3, 2.9), 7, True)
    dirv = np.array([_sinf(theta), -_cosf(theta), f32(0.0)], dtype=np.float32)
    q = _axis_angle(_v(0, 0, 1), theta)
    omega = _v(0, 0, -0.55)

    def swing_vel(c):
        r = c - anchor
        return np.array([omega[1] * r[2] - omega[2] * r[1], omega[2] * r[0] - omega[0] * r[2],
                         omega[0] * r[1] - omega[1] * r[0]], dtype=np.float32)

    prev = beam
    for i in range(link_count):
        c = anchor + (f32(0.5) + f32(i)) * link_len * dirv
        link = add_capsule(c, link_r, link_hh, 6, q, 7800.0, swing_vel(c))
        jd = m3d_ball_joint_def()
        jd.bodyA, jd.bodyB = prev, link
        jd.localAnchorA = m3d_vec3(0.0, -0.3, 0.0) if i == 0 else m3d_vec3(0.0, float(-(link_hh + link_r)), 0.0)
        jd.localAnchorB = m3d_vec3(0.0, float(link_hh + link_r), 0.0)
        L.m3d_joint_create_ball(w, C.byref(jd))
        prev = link
    chain_len = f32(link_count) * link_len
    ball_c = anchor + (chain_len + ball_r * f32(0.85)) * dirv
    ball = add_sphere(ball_c, ball_r, 2, 5500.0, swing_vel(ball_c))
    jd = m3d_ball_joint_def()
    jd.bodyA, jd.bodyB = prev, ball
    jd.localAnchorA = m3d_vec3(0.0, float(-(link_hh + link_r)), 0.0)
    jd.localAnchorB = m3d_vec3(0.0, float(ball_r * f32(0.85)), 0.0)
    L.m3d_joint_create_ball(w, C.byref(jd))
    rod = m3d_distance_joint_def()
    rod.bodyA, rod.bodyB = beam, ball
    rod.localAnchorA = m3d_vec3(0.0, -0.3, 0.0)
    rod.localAnchorB = m3d_vec3(0.0, 0.0, 0.0)
    rod.length, rod.hertz, rod.dampingRatio = 0.0, 0.0, 0.0
    L.m3d_joint_create_distance(w, C.byref(rod))
    # rain of mixed shapes. The example draws the rotation's four numbers
    # inside one call's argument list, whose evaluation order C++ leaves to
    # the compiler; GCC on x86-64 evaluates them last to first (the angle,
    # then the axis z, y, x), which this follows.
    rnd = _Rand(20260708)
    for i in range(46):
        x = f32(5.0) * rnd() - f32(0.4)
        z = f32(4.4) * rnd() - f32(2.2)
        y = f32(19.0) + f32(2.0) * f32(i)
        kind = i % 3
        p = np.array([x, y, z], dtype=np.float32)
        if kind == 0:
            add_sphere(p, f32(0.32) + f32(0.15) * rnd(), 3, 1000.0, _v(0, 0, 0))
            continue
        angle = rnd() * f32(6.28)
        az, ay, ax = rnd(), rnd(), rnd()
        qr = _axis_angle(np.array([ax, ay, az], dtype=np.float32), angle)
        if kind == 1:
            add_capsule(p, f32(0.24), f32(0.34), 4, qr, 1000.0, _v(0, 0, 0))
        else:
            add_box(p, _v(0.33, 0.33, 0.33), 5, False, qr)
    return ids, np.array(rows, dtype=BODY_DTYPE)


def dump(L: Library, path: str, frames: int = 500, substeps: int = 8, dt: float = 1.0 / 60.0,
         scene=showcase_scene) -> np.ndarray:
    """Builds `scene` in a new world of L, steps it `frames` times and writes
    one frame per step to `path`. Returns the poses (frames, bodies, 7)."""
    w = L.create_world()
    try:
        ids, infos = scene(L, w)
        writer = SceneWriter(path, infos, frames)
        poses = np.zeros((frames, len(ids), POSE_FLOATS), dtype=np.float32)
        for k in range(frames):
            L.m3d_world_step(w, dt, substeps)
            poses[k] = frame_poses(L, w, ids)
            writer.write_frame(poses[k])
        writer.close()
        if hasattr(L, "m3dx_step_failures") and L.m3dx_step_failures(w) != 0:
            raise RuntimeError("a step failed: " + L.last_error())
        return poses
    finally:
        L.m3d_world_destroy(w)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m melon3d_amd.scene_dump")
    sub = ap.add_subparsers(dest="cmd", required=True)
    d = sub.add_parser("dump", help="run the showcase scene and write scene.bin")
    d.add_argument("--out", default="scene.bin")
    d.add_argument("--frames", type=int, default=500)
    d.add_argument("--substeps", type=int, default=8)
    d.add_argument("--library", default=None, help="a library exporting the C ABI (default: this backend)")
    c = sub.add_parser("compare", help="diff two scene.bin files frame by frame")
    c.add_argument("a")
    c.add_argument("b")
    c.add_argument("--tol", type=float, default=1e-5)
    args = ap.parse_args(argv)
    from . import api

    if args.cmd == "dump":
        L = api.load_library(args.library) if args.library else api.load()
        poses = dump(L, args.out, args.frames, args.substeps)
        print(f"wrote {args.out}: {poses.shape[0]} frames, {poses.shape[1]} bodies")
        return 0
    ia, pa = read_scene(args.a)
    ib, pb = read_scene(args.b)
    if not np.array_equal(ia, ib):
        print("body tables differ")
        return 1
    r = compare(pa, pb, args.tol)
    print(f"{r['frames']} frames: max |dx| {r['max_pos']:.3g} m, max rotation {r['max_rot']:.3g} rad, "
          f"first frame beyond {args.tol}: {r['first_beyond']}, bit-identical: {r['identical']}")
    return 0 if r["first_beyond"] < 0 else 1


if __name__ == "__main__":
    sys.exit(main())
