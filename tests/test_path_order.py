"""The kernels' pixel order (dort_amd/csrc/path_order.hpp), checked on the CPU by
tests/path_order_check.cpp: tile_pixel is a bijection onto the pixels with tile_pos its
inverse, and a wave's 64 consecutive paths cover one 8x8 block of the image.  The GPU
render tests check that the order changes no image (random numbers are keyed by pixel
and sample, and every render is compared with the oracle pixel by pixel)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "dort_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_path_order(tmp_path):
    exe = str(tmp_path / "path_order_check")
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{CSRC}", os.path.join(HERE, "path_order_check.cpp"),
                    "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
