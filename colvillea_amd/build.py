"""Build libclpt_hip.so in-tree with hipcc for gfx950.

    python -m colvillea_amd.build

Flags: -ffp-contract=off (no FMA contraction: results follow the written
evaluation order, as the CPU oracle does), correctly rounded float32
divide/sqrt (the HIP default), -munsafe-fp-atomics (hardware float64 atomic
add for the accumulation buffer), and packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(REPO, "_build", "clpt")
LIB = os.path.join(OUT_DIR, "libclpt_hip.so")

ARCH = os.environ.get("CLPT_OFFLOAD_ARCH", "gfx950")
SOURCES = ["kernels.hip", "guide.hip", "clhip_api.cpp", "bvh_build.cpp", "image_io.cpp", "mesh_io.cpp"]
CFLAGS = [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-munsafe-fp-atomics",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in (r.stderr or "").splitlines() if noise not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{err}")
    return err


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    cc = hipcc()
    objs = [os.path.join(OBJ_DIR, s.rsplit(".", 1)[0] + ".o") for s in SOURCES]
    cmds = [[cc, *CFLAGS, "-c", os.path.join(CSRC, s), "-o", o] for s, o in zip(SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for msg in ex.map(_run, cmds):
            if verbose and msg:
                print(msg, file=sys.stderr)
    tmp = LIB + ".tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    build_host_example()
    return LIB


HOST_EXAMPLE = os.path.join(HERE, "host", "render_example")
SCENE_EXAMPLE = os.path.join(HERE, "host", "scene_example")


def build_host_example() -> str:
    """The C++ host examples over host/hip_integrator.h and host/hip_scene.h
    (plain g++, linked against the library with an rpath to
    colvillea_amd/lib)."""
    for name, out in (("render_example.cpp", HOST_EXAMPLE), ("scene_example.cpp", SCENE_EXAMPLE)):
        src = os.path.join(HERE, "host", name)
        _run(["g++", "-std=c++17", "-O2", "-Wall", f"-I{REPO}", src, f"-L{OUT_DIR}", "-lclpt_hip",
              "-Wl,-rpath,$ORIGIN/../lib", "-o", out])
    return HOST_EXAMPLE


if __name__ == "__main__":
    print(build(verbose=True))
