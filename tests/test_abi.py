"""The C-ABI library loads and exports every function include/*.h declares
(no GPU needed: nothing is called)."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for h in ("melon3d.h", "melon3d_ext.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        names += re.findall(r"M3D_API\s+[^;(]*?\b(m3dx?_\w+)\s*\(", src)
    return names


def test_header_declares_full_reference_api():
    names = [n for n in _declared() if n.startswith("m3d_")]
    assert len(names) == 49 and len(set(names)) == 49


def test_library_exports_every_declared_symbol():
    from melon3d_amd import api

    lib = C.CDLL(api.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    from melon3d_amd import api

    assert sorted(_declared()) == sorted(api.C_API_FUNCTIONS + api.EXT_FUNCTIONS)
