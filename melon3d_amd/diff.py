"""Differentiable worlds: the XPBD solve's adjoint exposed to PyTorch.

`rollout(v0, friction, mass)` steps a batch of worlds on the device and
returns the final body positions; its backward runs the adjoint of every
recorded step (`m3dx_diff_backward`): the solve with each step's contact set
and manifolds held fixed, reversed step by step from the last, each step
replayed from its recorded start state on a reverse-mode tape.

The scene (`DiffScene`): `worlds` worlds of 64 bodies (spheres of radius 0.5
and cubes of half-extent 0.5, alternating), on a 4 x 4 x 4 lattice 1.25 m
apart above a ground box, with initial velocities v0 ~ N(0, 2) per
component, friction ~ U[0.2, 0.9] and masses of density 1000 (the default).
Draws use the LCG of bench/bench.cpp:121-125 (Box-Muller for normals),
world w seeded 4242 + w. Sleep is off (a step's output is its solve's).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from .api import M3D_BODY_DYNAMIC, M3D_BODY_STATIC, M3D_SHAPE_BOX, M3D_SHAPE_SPHERE, Library, m3d_vec3
from .scenes import Lcg

BODIES = 64
DT = 1.0 / 60.0
SPHERE_VOLUME = 4.0 / 3.0 * math.pi * 0.5 ** 3
BOX_VOLUME = 1.0
OUT = 13  # per body state: x(3) q(4) v(3) w(3)


class DiffScene:
    def __init__(self, worlds: int = 256, seed: int = 4242):
        self.worlds = worlds
        W, B = worlds, BODIES
        self.pos = np.zeros((W, B, 3))
        self.v0 = np.zeros((W, B, 3))
        self.friction = np.zeros((W, B))
        self.sphere = np.array([k % 2 == 0 for k in range(B)])
        vol = np.where(self.sphere, SPHERE_VOLUME, BOX_VOLUME)
        self.mass = np.tile(1000.0 * vol, (W, 1))
        for w in range(W):
            rng = Lcg(seed + w)
            for k in range(B):
                i, j, l = k % 4, (k // 4) % 4, k // 16
                jit = [rng.uniform(-0.1, 0.1) for _ in range(3)]
                self.pos[w, k] = (1.25 * (i - 1.5) + jit[0], 0.6 + 1.25 * l + 0.5 * jit[1], 1.25 * (j - 1.5) + jit[2])
                self.v0[w, k] = [2.0 * rng.normal() for _ in range(3)]
                self.friction[w, k] = rng.uniform(0.2, 0.9)
        # loss target: 1.5 m along +x from the initial centre of mass
        com0 = (self.mass[..., None] * self.pos).sum(1) / self.mass.sum(1)[:, None]
        self.target = com0 + np.array([1.5, 0.0, 0.0])

    def dynamic_index(self) -> np.ndarray:
        """Body slot of (world, body): world w's ground first, then its 64."""
        w = np.arange(self.worlds)[:, None]
        return w * (BODIES + 1) + 1 + np.arange(BODIES)[None, :]

    def build(self, L: Library, v0=None, friction=None, mass=None):
        """A batch world of this scene with the given parameters (arrays of
        shape (worlds, 64, 3) / (worlds, 64); None keeps the scene's)."""
        v0 = self.v0 if v0 is None else np.asarray(v0, dtype=np.float64)
        friction = self.friction if friction is None else np.asarray(friction, dtype=np.float64)
        mass = self.mass if mass is None else np.asarray(mass, dtype=np.float64)
        wd = L.m3d_world_def_default()
        wd.enableSleep = False
        w = L.m3dx_batch_create(C.byref(wd), self.worlds)
        if not w:
            raise RuntimeError("m3dx_batch_create failed: " + L.last_error())
        for s in range(self.worlds):
            L.m3dx_batch_select(w, s)
            bd = L.m3d_body_def_default()
            bd.type = M3D_BODY_STATIC
            bd.position = m3d_vec3(0.0, -0.5, 0.0)
            sd = L.m3d_shape_def_default()
            sd.type = M3D_SHAPE_BOX
            sd.halfExtents = m3d_vec3(20.0, 0.5, 20.0)
            sd.friction = 0.6
            L.m3d_body_create(w, C.byref(bd), C.byref(sd))
            for k in range(BODIES):
                bd = L.m3d_body_def_default()
                bd.type = M3D_BODY_DYNAMIC
                bd.position = m3d_vec3(*self.pos[s, k])
                bd.linearVelocity = m3d_vec3(*v0[s, k])
                sd = L.m3d_shape_def_default()
                if self.sphere[k]:
                    sd.type = M3D_SHAPE_SPHERE
                    sd.radius = 0.5
                    sd.density = float(mass[s, k] / SPHERE_VOLUME)
                else:
                    sd.type = M3D_SHAPE_BOX
                    sd.halfExtents = m3d_vec3(0.5, 0.5, 0.5)
                    sd.density = float(mass[s, k] / BOX_VOLUME)
                sd.friction = float(friction[s, k])
                L.m3d_body_create(w, C.byref(bd), C.byref(sd))
        return w

    def body_count(self) -> int:
        return self.worlds * (BODIES + 1)

    def initial_state(self, v0=None) -> np.ndarray:
        """The (bodies, 13) float64 state at the first step: x, q, v, w."""
        v0 = self.v0 if v0 is None else v0
        n = self.body_count()
        st = np.zeros((n, OUT))
        idx = self.dynamic_index()
        st[:, 6] = 1.0
        st[::BODIES + 1, 1] = -0.5  # grounds
        st[idx, 0:3] = self.pos
        st[idx, 7:10] = v0
        return st

    def loss(self, final_pos, mass):
        """Sum over worlds of |COM - target|^2 (numpy or torch)."""
        m = mass[..., None]
        com = (m * final_pos).sum(1) / m.sum(1)
        d = com - (self.target if isinstance(final_pos, np.ndarray) else _torch().as_tensor(self.target,
                                                                                              dtype=final_pos.dtype))
        return (d * d).sum()


def _torch():
    import torch

    return torch


def final_positions(L: Library, w, scene: DiffScene) -> np.ndarray:
    snap = L.export_state(w)
    idx = scene.dynamic_index()
    out = np.zeros((scene.worlds, BODIES, 3))
    for s in range(scene.worlds):
        for k in range(BODIES):
            out[s, k] = snap.bodies[int(idx[s, k])].position[:]
    return out


def replay(L: Library, w, scene: DiffScene, steps: int, precision: int = 1, v0=None, friction=None, mass=None,
           recorded: bool = False):
    """Final (bodies, 13) state of the recorded steps replayed with the given
    parameters (fixed contact sets), in float (0) or double (1). recorded:
    start from the recorded state with the recorded masses and frictions
    (the engine's exact inputs) instead."""
    n = scene.body_count()
    dt = np.float64 if precision else np.float32
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    out = np.zeros((n, OUT), dtype=dt)
    if recorded:
        rc = L.m3dx_diff_replay(w, 0, steps, precision, None, None, None, ptr(out))
        if rc != 0:
            raise RuntimeError("replay failed: " + L.last_error())
        return out
    st = scene.initial_state(v0).astype(dt)
    m = np.zeros(n, dtype=dt)
    mu = np.full(n, 0.6, dtype=dt)
    idx = scene.dynamic_index()
    m[idx] = scene.mass if mass is None else mass
    mu[idx] = scene.friction if friction is None else friction
    rc = L.m3dx_diff_replay(w, 0, steps, precision, ptr(st), ptr(m), ptr(mu), ptr(out))
    if rc != 0:
        raise RuntimeError("replay failed: " + L.last_error())
    return out


def backward(L: Library, w, scene: DiffScene, grad_final_pos: np.ndarray, tape: bool = False):
    """(d/dv0, d/dfriction, d/dmass) of <grad_final_pos, final positions>, by
    the parallel adjoint (or, tape=True, by the reverse-mode tape)."""
    n = scene.body_count()
    idx = scene.dynamic_index()
    dt = np.float32 if tape else np.float64
    seed = np.zeros((n, OUT), dtype=dt)
    seed[idx, 0:3] = grad_final_pos
    adj = np.zeros((n, OUT), dtype=dt)
    mg = np.zeros(n, dtype=dt)
    fg = np.zeros(n, dtype=dt)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    fn = L.m3dx_diff_backward_tape if tape else L.m3dx_diff_backward_f64
    if fn(w, ptr(seed), ptr(adj), ptr(mg), ptr(fg)) != 0:
        raise RuntimeError("adjoint failed: " + L.last_error())
    return adj[idx, 7:10].astype(np.float64), fg[idx].astype(np.float64), mg[idx].astype(np.float64)


class _RecordedWorld:
    """Owns a recorded batch world for as long as the autograd graph that
    needs it: freed when the owner is collected (weakref.finalize), so a
    second backward (retain_graph=True) finds it alive and a forward whose
    output never gets a backward does not leak it."""

    def __init__(self, L: Library, w):
        import weakref

        self.L, self.w = L, w
        self._finalizer = weakref.finalize(self, L.m3d_world_destroy, w)

    def close(self):
        self._finalizer()


def record_rollout(L: Library, scene: DiffScene, v0, friction, mass, steps: int, substeps: int) -> _RecordedWorld:
    """Builds the scene's batch with these parameters and steps it `steps`
    times with recording on. Raises if a step failed or the recording was
    refused (joints: the adjoint would not describe the steps the engine
    ran)."""
    owner = _RecordedWorld(L, scene.build(L, v0, friction, mass))
    w = owner.w
    if L.m3dx_diff_record(w, 1) != 0:
        owner.close()
        raise RuntimeError("m3dx_diff_record failed: " + L.last_error())
    for _ in range(steps):
        L.m3d_world_step(w, DT, substeps)
    failures = L.m3dx_step_failures(w)
    rc = L.m3dx_diff_record(w, 0)
    if failures != 0 or rc != 0:
        msg = L.last_error()
        owner.close()
        raise RuntimeError(f"recorded rollout failed ({failures} failed steps): {msg}")
    return owner


def rollout_function():
    """torch.autograd.Function: (v0, friction, mass) -> final positions."""
    torch = _torch()

    class XpbdRollout(torch.autograd.Function):
        @staticmethod
        def forward(ctx, v0, friction, mass, L, scene, steps, substeps):
            owner = record_rollout(L, scene, v0.detach().cpu().double().numpy(),
                                   friction.detach().cpu().double().numpy(), mass.detach().cpu().double().numpy(),
                                   steps, substeps)
            out = final_positions(L, owner.w, scene)
            ctx.owner, ctx.scene = owner, scene
            return torch.as_tensor(out, dtype=v0.dtype, device=v0.device)

        @staticmethod
        def backward(ctx, g):
            owner = ctx.owner
            gv, gf, gm = backward(owner.L, owner.w, ctx.scene, g.detach().cpu().numpy())
            dev, dt = g.device, g.dtype
            return (torch.as_tensor(gv, dtype=dt, device=dev), torch.as_tensor(gf, dtype=dt, device=dev),
                    torch.as_tensor(gm, dtype=dt, device=dev), None, None, None, None)

    return XpbdRollout


def rollout(L: Library, scene: DiffScene, v0, friction, mass, steps: int = 120, substeps: int = 4):
    """Final positions (worlds, 64, 3) as a differentiable function of the
    torch tensors v0 (worlds, 64, 3), friction and mass (worlds, 64)."""
    return rollout_function().apply(v0, friction, mass, L, scene, steps, substeps)
