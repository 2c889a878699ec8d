"""MI355X (gfx950) device backend of the Melon3D rigid-body engine's world
step, behind the unchanged m3d_* C API (include/melon3d.h).

`melon3d_amd.api.load()` opens the native library built by
`__graft_entry__.build()`.
"""
__all__ = ["api", "scenes"]
