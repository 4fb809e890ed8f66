"""The C ABI library loads and exports what include/clhip.h declares (no GPU
compute here); host-side argument checks fail with a status, not a crash."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from colvillea_amd import _capi as capi
from colvillea_amd import scenes

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "clhip.h")


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s+(clhip_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 25
    lib = C.CDLL(capi.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(capi.SIGNATURES) == names


def test_exported_symbols_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s(clhip_\w+)$", out, re.M))
    assert set(header_functions()) <= exported


STRUCTS = ("clhip_mesh", "clhip_material", "clhip_emitter", "clhip_params", "clhip_stats", "clhip_view",
           "clhip_launch")


def test_struct_layouts_match_header(tmp_path):
    """sizeof and every field offset of the ctypes mirrors equal what a C
    compiler makes of include/clhip.h."""
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "clhip.h"', "int main(void) {"]
    for name in STRUCTS:
        lines.append(f'printf("{name} size %zu\\n", sizeof({name}));')
        for f, _ in getattr(capi, name)._fields_:
            lines.append(f'printf("{name} {f} %zu\\n", offsetof({name}, {f}));')
    lines += ["return 0; }"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for name in STRUCTS:
        ct = getattr(capi, name)
        assert got[(name, "size")] == C.sizeof(ct), name
        for f, _ in ct._fields_:
            assert got[(name, f)] == getattr(ct, f).offset, (name, f)


def test_invalid_arguments_return_status():
    lib = capi.lib()
    assert lib.clhip_create(0, 10, 2, C.byref(C.c_void_p())) == -1
    assert "bad film size" in capi.last_error()
    assert lib.clhip_create(8, 8, 7, C.byref(C.c_void_p())) == -1
    assert lib.clhip_render(None, 1) == -1
    assert lib.clhip_destroy(None) == 0


def test_create_without_device_fails_loudly():
    """On a machine without a HIP device, context creation reports it."""
    ctx = C.c_void_p()
    rc = capi.lib().clhip_create(8, 8, 2, C.byref(ctx))
    if rc == 0:  # a GPU is present (CPU suite run on a machine with a GPU)
        capi.lib().clhip_destroy(ctx)
        pytest.skip("a HIP device is present")
    assert rc == -2
    assert "no HIP device" in capi.last_error()


def test_scene_generators_are_deterministic():
    a = scenes.make_c0(width=16, height=16)
    b = scenes.make_c0(width=16, height=16)
    assert np.array_equal(a.instance_xform, b.instance_xform)
    assert np.array_equal(a.hdri, b.hdri)
    assert a.num_triangles() == 2 + 64 * 1280
    c = scenes.make_c1(width=16, height=16, grid=(64, 32), n_spheres=9, sphere_subdiv=2, hdri_w=64, hdri_h=32)
    assert c.num_triangles() == 2 * 63 * 31 + 9 * 320
    kinds = [type(m).__name__ for m in c.materials[1:]]
    assert kinds.count("DiffuseMtl") == kinds.count("MetalMtl") == kinds.count("GlassMtl") == 3
    assert scenes.icosphere(4).triangles.shape[0] == 5120
