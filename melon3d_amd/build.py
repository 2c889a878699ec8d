"""Builds libmelon3d_amd.so (gfx950) in-tree, and the oracle libraries.

The device library is compiled with strict IEEE single precision
(-ffp-contract=off, correctly rounded division and square root) so its
arithmetic matches the reference engine operation for operation, and with
packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "libmelon3d_amd.so")

SOURCES = ["k_prims.hip", "k_broad.hip", "k_narrow.hip", "k_solve.hip", "k_isolve.hip", "k_diff.hip", "k_adjoint.hip", "k_query.hip", "step.cpp", "world_api.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("M3D_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value", "-Wno-unknown-warning-option",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
    cmd = [HIPCC, *COMMON, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32" not in l)
        raise RuntimeError(f"compile failed: {src}\n{msg}")
    return obj


def build_library(jobs: int = 6) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-Wl,-Bsymbolic"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    return LIB


SCENES_LIB = os.path.join(HERE, "libm3d_scenes.so")


def build_scene_driver() -> str:
    """Native scene driver for the benchmark (plain C, host only)."""
    tmp = SCENES_LIB + ".tmp"
    cmd = [os.environ.get("CC", "gcc"), "-O2", "-std=c11", "-fPIC", "-shared", "-I" + os.path.join(ROOT, "include"),
           os.path.join(CSRC, "scene_driver.c"), "-o", tmp, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("scene driver build failed:\n" + r.stderr)
    os.replace(tmp, SCENES_LIB)
    return SCENES_LIB


def build_oracle(reference: str = "/root/reference") -> None:
    """C restatement always; the reference build only where its tree exists."""
    oracle = os.path.join(ROOT, "oracle")
    targets = ["oracle"]
    if os.path.isdir(os.path.join(reference, "src")):
        targets.append("ref")
    r = subprocess.run(["make", "-s", "-C", oracle, f"REF={reference}", *targets], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


if __name__ == "__main__":
    print(build_library())
    build_scene_driver()
    build_oracle()
    sys.exit(0)
