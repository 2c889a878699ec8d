"""ctypes binding of the Melon3D C API (include/melon3d.h) and the m3dx_*
extension (include/melon3d_ext.h).

`load()` opens libmelon3d_amd.so, the MI355X backend; it raises when the
library is missing, there is no silent fallback. `load_library(path)` opens
any library with the same exported C ABI.

Functions keep their C names, argument meaning and error behaviour, so
callers read like C callers: `m3d.m3d_body_position(world, body)`.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmelon3d_amd.so")

M3D_PI = 3.14159265358979323846

M3D_BODY_STATIC, M3D_BODY_KINEMATIC, M3D_BODY_DYNAMIC = 0, 1, 2
M3D_SHAPE_SPHERE, M3D_SHAPE_CAPSULE, M3D_SHAPE_BOX = 0, 1, 2


class m3d_vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]

    def tuple(self):
        return (self.x, self.y, self.z)

    def __repr__(self):
        return f"v3({self.x:.6g}, {self.y:.6g}, {self.z:.6g})"


class m3d_quat(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]

    def tuple(self):
        return (self.x, self.y, self.z, self.w)


class m3d_transform(C.Structure):
    _fields_ = [("p", m3d_vec3), ("q", m3d_quat)]


class m3d_aabb(C.Structure):
    _fields_ = [("lowerBound", m3d_vec3), ("upperBound", m3d_vec3)]


class m3d_body_id(C.Structure):
    _fields_ = [("index", C.c_int32), ("generation", C.c_uint32)]


class m3d_joint_id(C.Structure):
    _fields_ = [("index", C.c_int32), ("generation", C.c_uint32)]


class m3d_filter(C.Structure):
    _fields_ = [("categoryBits", C.c_uint32), ("maskBits", C.c_uint32), ("groupIndex", C.c_int32)]


class m3d_shape_def(C.Structure):
    _fields_ = [
        ("type", C.c_int),
        ("radius", C.c_float),
        ("halfHeight", C.c_float),
        ("halfExtents", m3d_vec3),
        ("density", C.c_float),
        ("friction", C.c_float),
        ("restitution", C.c_float),
        ("rollingResistance", C.c_float),
        ("filter", m3d_filter),
        ("isSensor", C.c_bool),
    ]


class m3d_body_def(C.Structure):
    _fields_ = [
        ("type", C.c_int),
        ("position", m3d_vec3),
        ("orientation", m3d_quat),
        ("linearVelocity", m3d_vec3),
        ("angularVelocity", m3d_vec3),
        ("linearDamping", C.c_float),
        ("angularDamping", C.c_float),
        ("gravityScale", C.c_float),
        ("enableSleep", C.c_bool),
        ("isAwake", C.c_bool),
        ("userData", C.c_uint64),
    ]


class m3d_world_def(C.Structure):
    _fields_ = [
        ("gravity", m3d_vec3),
        ("workerCount", C.c_int32),
        ("positionIterations", C.c_int32),
        ("gridCellSize", C.c_float),
        ("enableSleep", C.c_bool),
    ]


class m3d_distance_joint_def(C.Structure):
    _fields_ = [
        ("bodyA", m3d_body_id),
        ("bodyB", m3d_body_id),
        ("localAnchorA", m3d_vec3),
        ("localAnchorB", m3d_vec3),
        ("length", C.c_float),
        ("hertz", C.c_float),
        ("dampingRatio", C.c_float),
    ]


class m3d_ball_joint_def(C.Structure):
    _fields_ = [("bodyA", m3d_body_id), ("bodyB", m3d_body_id), ("localAnchorA", m3d_vec3), ("localAnchorB", m3d_vec3)]


class m3d_hinge_joint_def(C.Structure):
    _fields_ = [
        ("bodyA", m3d_body_id),
        ("bodyB", m3d_body_id),
        ("localAnchorA", m3d_vec3),
        ("localAnchorB", m3d_vec3),
        ("localAxisA", m3d_vec3),
        ("localAxisB", m3d_vec3),
        ("enableLimit", C.c_bool),
        ("lowerAngle", C.c_float),
        ("upperAngle", C.c_float),
        ("enableMotor", C.c_bool),
        ("motorSpeed", C.c_float),
        ("maxMotorTorque", C.c_float),
    ]


class m3d_weld_joint_def(C.Structure):
    _fields_ = [("bodyA", m3d_body_id), ("bodyB", m3d_body_id), ("localAnchorA", m3d_vec3), ("localAnchorB", m3d_vec3)]


class m3d_contact_event(C.Structure):
    _fields_ = [("bodyA", m3d_body_id), ("bodyB", m3d_body_id), ("isSensor", C.c_bool)]


class m3d_contact_events(C.Structure):
    _fields_ = [
        ("beginEvents", C.POINTER(m3d_contact_event)),
        ("beginCount", C.c_int32),
        ("endEvents", C.POINTER(m3d_contact_event)),
        ("endCount", C.c_int32),
    ]


class m3d_ray_result(C.Structure):
    _fields_ = [
        ("body", m3d_body_id),
        ("point", m3d_vec3),
        ("normal", m3d_vec3),
        ("fraction", C.c_float),
        ("hit", C.c_bool),
    ]


class m3d_profile(C.Structure):
    _fields_ = [
        ("stepMs", C.c_float),
        ("broadphaseMs", C.c_float),
        ("narrowphaseMs", C.c_float),
        ("solverMs", C.c_float),
        ("bodyCount", C.c_int32),
        ("awakeBodyCount", C.c_int32),
        ("contactCount", C.c_int32),
        ("touchingContactCount", C.c_int32),
        ("islandCount", C.c_int32),
    ]


# ------------------------------------------------------------- extension

F3 = C.c_float * 3
F4 = C.c_float * 4


class m3dx_body_state(C.Structure):
    _fields_ = [
        ("position", F3), ("rotation", F4), ("linearVelocity", F3), ("angularVelocity", F3),
        ("prevPosition", F3), ("prevRotation", F4), ("prevLinearVelocity", F3), ("prevAngularVelocity", F3),
        ("invMass", C.c_float), ("invInertiaLocal", F3), ("invInertiaWorld", C.c_float * 9),
        ("shapeType", C.c_int32), ("radius", C.c_float), ("halfHeight", C.c_float), ("halfExtents", F3),
        ("friction", C.c_float), ("restitution", C.c_float), ("rollingResistance", C.c_float), ("mass", C.c_float),
        ("force", F3), ("torque", F3),
        ("categoryBits", C.c_uint32), ("maskBits", C.c_uint32), ("groupIndex", C.c_int32), ("isSensor", C.c_int32),
        ("type", C.c_int32), ("linearDamping", C.c_float), ("angularDamping", C.c_float), ("gravityScale", C.c_float),
        ("sleepTime", C.c_float), ("smoothedEnergy", C.c_float), ("enableSleep", C.c_int32), ("isAwake", C.c_int32),
        ("inUse", C.c_int32), ("generation", C.c_uint32), ("userDataLo", C.c_uint32), ("userDataHi", C.c_uint32),
        ("fatAABB", C.c_float * 6), ("hasFatAABB", C.c_int32), ("inStableTier", C.c_int32),
        ("inStableCells", C.c_int32),
    ]


class m3dx_point_state(C.Structure):
    _fields_ = [
        ("localAnchorA", F3), ("localAnchorB", F3), ("frictionAnchorA", F3), ("frictionAnchorB", F3),
        ("separation", C.c_float), ("id", C.c_uint32), ("lambdaN", C.c_float), ("lambdaT", C.c_float),
        ("separationOffset", C.c_float), ("slipDir", F3), ("rA", F3), ("rB", F3), ("wN", C.c_float),
        ("angJA", F3), ("angJB", F3),
    ]


class m3dx_contact_state(C.Structure):
    _fields_ = [
        ("bodyA", C.c_int32), ("bodyB", C.c_int32), ("friction", C.c_float), ("restitution", C.c_float),
        ("rollingResistance", C.c_float), ("touching", C.c_int32), ("isSensor", C.c_int32),
        ("tunnelGuard", C.c_int32), ("curved", C.c_int32), ("hasCache", C.c_int32),
        ("cachePosA", F3), ("cachePosB", F3), ("cacheRotA", F4), ("cacheRotB", F4),
        ("normal", F3), ("pointCount", C.c_int32), ("points", m3dx_point_state * 4),
    ]


class m3dx_joint_state(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("bodyA", C.c_int32), ("bodyB", C.c_int32), ("localAnchorA", F3), ("localAnchorB", F3),
        ("length", C.c_float), ("hertz", C.c_float), ("dampingRatio", C.c_float), ("localAxisA", F3),
        ("localAxisB", F3), ("localRefA", F3), ("localRefB", F3), ("enableLimit", C.c_int32),
        ("lowerAngle", C.c_float), ("upperAngle", C.c_float), ("enableMotor", C.c_int32), ("motorSpeed", C.c_float),
        ("maxMotorTorque", C.c_float), ("relRot0", F4), ("compliance", C.c_float), ("lambda_", C.c_float),
        ("motorLambda", C.c_float), ("inUse", C.c_int32), ("generation", C.c_uint32),
    ]


class m3dx_grid_entry(C.Structure):
    _fields_ = [("key", C.c_uint64), ("body", C.c_int32), ("pad", C.c_int32)]


class m3dx_world_header(C.Structure):
    _fields_ = [
        ("bodyCapacity", C.c_int32), ("contactCount", C.c_int32), ("jointCapacity", C.c_int32),
        ("freeBodyCount", C.c_int32), ("freeJointCount", C.c_int32), ("stableEntryCount", C.c_int32),
        ("stableOversizedCount", C.c_int32), ("gridDirty", C.c_int32), ("stableDirty", C.c_int32),
        ("stableForce", C.c_int32), ("stepsSinceStableBuild", C.c_int32), ("anyForcesApplied", C.c_int32),
        ("jointedPairsDirty", C.c_int32), ("gravity", F3), ("workerCount", C.c_int32),
        ("positionIterations", C.c_int32), ("gridCellSize", C.c_float), ("enableSleep", C.c_int32),
    ]


class m3dx_snapshot(C.Structure):
    _fields_ = [
        ("header", m3dx_world_header),
        ("bodies", C.POINTER(m3dx_body_state)),
        ("contacts", C.POINTER(m3dx_contact_state)),
        ("joints", C.POINTER(m3dx_joint_state)),
        ("freeBodies", C.POINTER(C.c_int32)),
        ("freeJoints", C.POINTER(C.c_int32)),
        ("stableEntries", C.POINTER(m3dx_grid_entry)),
        ("stableOversized", C.POINTER(C.c_int32)),
    ]


class m3dx_step_stats(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "bodyCapacity", "awakeDynamicBodies", "contacts", "solvedContacts", "solvedPoints", "boxBoxPoints",
        "regeneratedManifolds", "activeEntries", "stableEntries", "refattedBodies", "islands", "coloredIslands",
        "sweeps", "colorRounds")]


class m3dx_kernel_time(C.Structure):
    _fields_ = [("launches", C.c_int32), ("ms", C.c_float), ("bytes", C.c_double)]


M3DX_KERNELS = ["grid_normals", "grid_friction", "grid_velocity", "grid_order", "joint_islands", "narrow", "pairs",
                "adjoint", "adjoint_forward", "grid_prepare", "grid_bodies", "sort"]

W = C.c_void_p  # m3d_world*
overlap_fn = C.CFUNCTYPE(C.c_bool, m3d_body_id, C.c_void_p)

_SIGS = {
    "m3d_world_def_default": (m3d_world_def, []),
    "m3d_body_def_default": (m3d_body_def, []),
    "m3d_shape_def_default": (m3d_shape_def, []),
    "m3d_hinge_joint_def_default": (m3d_hinge_joint_def, []),
    "m3d_weld_joint_def_default": (m3d_weld_joint_def, []),
    "m3d_world_create": (W, [C.POINTER(m3d_world_def)]),
    "m3d_world_destroy": (None, [W]),
    "m3d_world_step": (None, [W, C.c_float, C.c_int]),
    "m3d_body_create": (m3d_body_id, [W, C.POINTER(m3d_body_def), C.POINTER(m3d_shape_def)]),
    "m3d_body_destroy": (None, [W, m3d_body_id]),
    "m3d_body_is_valid": (C.c_bool, [W, m3d_body_id]),
    "m3d_body_position": (m3d_vec3, [W, m3d_body_id]),
    "m3d_body_rotation": (m3d_quat, [W, m3d_body_id]),
    "m3d_body_transform": (m3d_transform, [W, m3d_body_id]),
    "m3d_body_set_transform": (None, [W, m3d_body_id, m3d_vec3, m3d_quat]),
    "m3d_body_linear_velocity": (m3d_vec3, [W, m3d_body_id]),
    "m3d_body_angular_velocity": (m3d_vec3, [W, m3d_body_id]),
    "m3d_body_set_linear_velocity": (None, [W, m3d_body_id, m3d_vec3]),
    "m3d_body_set_angular_velocity": (None, [W, m3d_body_id, m3d_vec3]),
    "m3d_body_apply_impulse": (None, [W, m3d_body_id, m3d_vec3, m3d_vec3]),
    "m3d_body_apply_force": (None, [W, m3d_body_id, m3d_vec3]),
    "m3d_body_apply_force_at_point": (None, [W, m3d_body_id, m3d_vec3, m3d_vec3]),
    "m3d_body_apply_torque": (None, [W, m3d_body_id, m3d_vec3]),
    "m3d_body_apply_angular_impulse": (None, [W, m3d_body_id, m3d_vec3]),
    "m3d_body_mass": (C.c_float, [W, m3d_body_id]),
    "m3d_body_get_type": (C.c_int, [W, m3d_body_id]),
    "m3d_body_is_awake": (C.c_bool, [W, m3d_body_id]),
    "m3d_body_wake": (None, [W, m3d_body_id]),
    "m3d_body_user_data": (C.c_uint64, [W, m3d_body_id]),
    "m3d_body_set_user_data": (None, [W, m3d_body_id, C.c_uint64]),
    "m3d_body_friction": (C.c_float, [W, m3d_body_id]),
    "m3d_body_set_friction": (None, [W, m3d_body_id, C.c_float]),
    "m3d_body_restitution": (C.c_float, [W, m3d_body_id]),
    "m3d_body_set_restitution": (None, [W, m3d_body_id, C.c_float]),
    "m3d_world_gravity": (m3d_vec3, [W]),
    "m3d_world_set_gravity": (None, [W, m3d_vec3]),
    "m3d_joint_create_distance": (m3d_joint_id, [W, C.POINTER(m3d_distance_joint_def)]),
    "m3d_joint_create_ball": (m3d_joint_id, [W, C.POINTER(m3d_ball_joint_def)]),
    "m3d_joint_create_hinge": (m3d_joint_id, [W, C.POINTER(m3d_hinge_joint_def)]),
    "m3d_joint_create_weld": (m3d_joint_id, [W, C.POINTER(m3d_weld_joint_def)]),
    "m3d_joint_destroy": (None, [W, m3d_joint_id]),
    "m3d_joint_hinge_angle": (C.c_float, [W, m3d_joint_id]),
    "m3d_joint_hinge_set_motor": (None, [W, m3d_joint_id, C.c_bool, C.c_float, C.c_float]),
    "m3d_world_ray_cast": (m3d_ray_result, [W, m3d_vec3, m3d_vec3]),
    "m3d_world_ray_cast_filtered": (m3d_ray_result, [W, m3d_vec3, m3d_vec3, C.c_uint32]),
    "m3d_world_sphere_cast": (m3d_ray_result, [W, m3d_vec3, C.c_float, m3d_vec3, C.c_uint32]),
    "m3d_world_overlap_aabb": (None, [W, m3d_aabb, overlap_fn, C.c_void_p]),
    "m3d_world_contact_events": (m3d_contact_events, [W]),
    "m3d_world_profile": (m3d_profile, [W]),
    # extension
    "m3dx_world_export": (C.c_int, [W, C.POINTER(m3dx_snapshot)]),
    "m3dx_world_import": (W, [C.POINTER(m3dx_snapshot)]),
    "m3dx_debug_grid_entries": (C.c_int, [W, C.c_int, C.POINTER(m3dx_grid_entry), C.c_int]),
    "m3dx_debug_oversized": (C.c_int, [W, C.c_int, C.POINTER(C.c_int32), C.c_int]),
    "m3dx_debug_islands": (C.c_int, [W, C.POINTER(C.c_int32), C.c_int, C.POINTER(C.c_uint8), C.c_int]),
    "m3dx_debug_contact_colors": (C.c_int, [W, C.POINTER(C.c_int32), C.c_int]),
    "m3dx_world_stats": (None, [W, C.POINTER(m3dx_step_stats)]),
    "m3dx_set_kernel_timing": (C.c_int, [W, C.c_int]),
    "m3dx_set_pass_blocks": (C.c_int, [W, C.c_int]),
    "m3dx_pass_blocks": (C.c_int, [W]),
    "m3dx_step_failures": (C.c_int32, [W]),
    "m3dx_set_step_graphs": (C.c_int, [W, C.c_int]),
    "m3dx_set_xcd_barrier": (C.c_int, [W, C.c_int]),
    "m3dx_set_grid_launch": (C.c_int, [W, C.c_int, C.c_uint32]),
    "m3dx_grid_launch_mode": (C.c_int, [W]),
    "m3dx_step_graph_launches": (C.c_int64, [W]),
    "m3dx_kernel_timing": (C.c_int, [W, C.POINTER(m3dx_kernel_time)]),
    "m3dx_batch_create": (W, [C.POINTER(m3d_world_def), C.c_int32]),
    "m3dx_batch_select": (C.c_int, [W, C.c_int32]),
    "m3dx_batch_size": (C.c_int32, [W]),
    "m3dx_batch_body_range": (C.c_int, [W, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "m3dx_batch_contact_counts": (C.c_int, [W, C.POINTER(C.c_int32), C.c_int32]),
    "m3dx_diff_record": (C.c_int, [W, C.c_int]),
    "m3dx_diff_steps": (C.c_int32, [W]),
    "m3dx_diff_state": (C.c_int, [W, C.c_int, C.c_void_p]),
    "m3dx_diff_replay": (C.c_int, [W, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "m3dx_diff_backward": (C.c_int, [W, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "m3dx_diff_backward_tape": (C.c_int, [W, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "m3dx_diff_backward_f64": (C.c_int, [W, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "m3dx_diff_set_checkpoint": (C.c_int, [W, C.c_int]),
    "m3dx_diff_forward": (C.c_int, [W, C.c_int, C.c_int, C.c_void_p]),
    "m3dx_diff_jvp_cpu": (C.c_int, [W, C.c_int, C.c_int, C.c_int] + [C.c_void_p] * 8),
    "m3dx_world_ray_cast_batch": (C.c_int, [W, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "m3dx_world_ray_cast_batch_noalloc": (C.c_int, [W, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "m3dx_body_respawn": (C.c_int, [W, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "m3dx_world_transforms": (C.c_int, [W, C.c_void_p, C.c_int32, C.c_void_p]),
    "m3dx_sort_pairs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p]),
    "m3dx_world_overlap_aabb_batch": (C.c_int, [W, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32]),
    "m3dx_world_sphere_cast_batch": (C.c_int, [W, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                                 C.c_void_p]),
    "m3dx_last_error": (C.c_char_p, []),
    "m3dx_set_device": (C.c_int, []),
}
_SIGS["m3dx_set_device"] = (C.c_int, [C.c_int])

C_API_FUNCTIONS = [n for n in _SIGS if n.startswith("m3d_")]
EXT_FUNCTIONS = [n for n in _SIGS if n.startswith("m3dx_")]


class Library:
    """One loaded implementation of the C ABI (device backend or oracle)."""

    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run __graft_entry__.build())")
        self.path = path
        self.lib = C.CDLL(path, mode=C.RTLD_LOCAL)
        for name, (res, args) in _SIGS.items():
            fn = getattr(self.lib, name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    # ---- convenience mirrors of the header's inline helpers ----
    @staticmethod
    def v3(x, y, z):
        return m3d_vec3(x, y, z)

    @staticmethod
    def quat(x, y, z, w):
        return m3d_quat(x, y, z, w)

    def quat_axis_angle(self, axis, angle):
        import math
        ax = [axis.x, axis.y, axis.z]
        n = math.sqrt(sum(a * a for a in ax))
        ax = [a / n for a in ax] if n >= 1e-12 else [0.0, 0.0, 0.0]
        s = math.sin(0.5 * angle)
        return m3d_quat(s * ax[0], s * ax[1], s * ax[2], math.cos(0.5 * angle))

    def create_world(self, wd=None):
        if wd is None:
            wd = self.m3d_world_def_default()
        w = self.m3d_world_create(C.byref(wd))
        if not w:
            raise RuntimeError("m3d_world_create failed: " + self.last_error())
        return w

    def last_error(self) -> str:
        e = self.m3dx_last_error()
        return e.decode() if e else ""

    # ---- snapshot helpers ----
    def export_state(self, world):
        snap = m3dx_snapshot()
        if self.m3dx_world_export(world, C.byref(snap)) != 0:
            raise RuntimeError("export failed: " + self.last_error())
        h = snap.header
        arrays = {
            "bodies": (m3dx_body_state * max(h.bodyCapacity, 1))(),
            "contacts": (m3dx_contact_state * max(h.contactCount, 1))(),
            "joints": (m3dx_joint_state * max(h.jointCapacity, 1))(),
            "freeBodies": (C.c_int32 * max(h.freeBodyCount, 1))(),
            "freeJoints": (C.c_int32 * max(h.freeJointCount, 1))(),
            "stableEntries": (m3dx_grid_entry * max(h.stableEntryCount, 1))(),
            "stableOversized": (C.c_int32 * max(h.stableOversizedCount, 1))(),
        }
        for k, a in arrays.items():
            setattr(snap, k, C.cast(a, type(getattr(snap, k))))
        if self.m3dx_world_export(world, C.byref(snap)) != 0:
            raise RuntimeError("export failed: " + self.last_error())
        snap._keep = arrays
        return snap

    def import_state(self, snap):
        w = self.m3dx_world_import(C.byref(snap))
        if not w:
            raise RuntimeError("import failed: " + self.last_error())
        return w

    def grid_entries(self, world, tier):
        n = self.m3dx_debug_grid_entries(world, tier, None, 0)
        buf = (m3dx_grid_entry * max(n, 1))()
        self.m3dx_debug_grid_entries(world, tier, buf, n)
        return [(buf[i].key, buf[i].body) for i in range(n)]

    def oversized(self, world, tier):
        n = self.m3dx_debug_oversized(world, tier, None, 0)
        buf = (C.c_int32 * max(n, 1))()
        self.m3dx_debug_oversized(world, tier, buf, n)
        return list(buf[:n])

    def islands(self, world, body_cap):
        isl = (C.c_int32 * max(body_cap, 1))()
        n = self.m3dx_debug_islands(world, isl, body_cap, None, 0)
        awake = (C.c_uint8 * max(n, 1))()
        self.m3dx_debug_islands(world, isl, body_cap, awake, n)
        return n, list(isl[:body_cap]), list(awake[:n])

    def contact_colors(self, world, count):
        buf = (C.c_int32 * max(count, 1))()
        n = self.m3dx_debug_contact_colors(world, buf, count)
        return list(buf[: min(n, count)])

    def kernel_timing(self, world):
        """{kernel name: (launches, ms, algorithmic bytes)} since timing was enabled."""
        arr = (m3dx_kernel_time * len(M3DX_KERNELS))()
        self.m3dx_kernel_timing(world, arr)
        return {n: (arr[i].launches, arr[i].ms, arr[i].bytes) for i, n in enumerate(M3DX_KERNELS)}

    def stats(self, world):
        s = m3dx_step_stats()
        self.m3dx_world_stats(world, C.byref(s))
        return s


_loaded: dict[str, Library] = {}


def load_library(path: str) -> Library:
    path = os.path.abspath(path)
    if path not in _loaded:
        _loaded[path] = Library(path)
    return _loaded[path]


def load() -> Library:
    """The MI355X backend. Fails loudly when it is not built."""
    return load_library(LIB_PATH)
