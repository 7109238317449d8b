"""The BVH builder's node layout (dort_amd/csrc/bvh_build.cpp, host code), checked on
the CPU by tests/bvh_layout_check.cpp: reachability of every triangle and node,
child boxes containing their triangles (BVH4: as decoded from the 8-bit planes),
leaf sizes, depth, and the leaf-order triangle records.  The GPU traversal tests check
the device's decoding of the same layout against the oracle."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "dort_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_bvh_layout(tmp_path):
    exe = str(tmp_path / "bvh_layout_check")
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{CSRC}", os.path.join(HERE, "bvh_layout_check.cpp"),
                    os.path.join(CSRC, "bvh_build.cpp"), "-o", exe, "-lpthread"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
