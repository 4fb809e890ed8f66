"""BVH4 and BVH8 builders on the host (no GPU): child boxes (quantised, for
BVH8) contain their subtrees, leaves respect the leaf cap, and every
primitive lands in exactly one leaf (tests/bvh_check.cpp)."""
import os
import subprocess

from colvillea_amd import build as hipbuild

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bvh_boxes_are_conservative_and_leaves_capped(tmp_path):
    exe = str(tmp_path / "bvh_check")
    src = os.path.join(REPO, "tests", "bvh_check.cpp")
    lib = os.path.join(REPO, "colvillea_amd", "csrc", "bvh_build.cpp")
    subprocess.run([hipbuild.hipcc(), "-std=c++17", "-O2", "-ffp-contract=off", "-fno-gpu-rdc", src, lib, "-o", exe],
                   check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout
