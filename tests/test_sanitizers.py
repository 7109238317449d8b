"""The host code and the oracle under the compilers' sanitizers (CPU only).

* dort_amd/csrc/io.cpp (PLY / RGBE / PNG readers, including the inflater) built with
  AddressSanitizer + UndefinedBehaviorSanitizer reads valid files and thousands of
  truncated and bit-flipped copies (tests/io_robustness_check.cpp);
* dort_amd/csrc/bvh_build.cpp under ASan + UBSan (every layout case of
  tests/bvh_layout_check.cpp) and under ThreadSanitizer (a 300k-triangle build whose top
  subtrees are built on threads of their own);
* the oracle (oracle/Makefile `san`: oracle/_san/liboracle.so) loaded into a Python
  process with the ASan runtime preloaded, running the oracle's own tests.

Any report aborts the program, which fails the test.  The GPU code is not built with
sanitizers: the device has none here.
"""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "dort_amd", "csrc")
ASAN_UBSAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def _build(exe, sources, flags):
    cmd = ["g++", "-std=c++17", *flags, f"-I{CSRC}", *sources, "-o", exe, "-lpthread"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)


def _run(cmd, env=None, timeout=600):
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-6000:])
    return r.stdout


def test_io_readers_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "io_check")
    _build(exe, [os.path.join(HERE, "io_robustness_check.cpp"), os.path.join(CSRC, "io.cpp"),
                 os.path.join(CSRC, "status.cpp")], ASAN_UBSAN)
    out = _run([exe, str(tmp_path)])
    assert out.startswith("OK")
    assert int(out.split()[1]) > 10000


def test_bvh_builder_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "bvh_asan")
    _build(exe, [os.path.join(HERE, "bvh_layout_check.cpp"), os.path.join(CSRC, "bvh_build.cpp")], ASAN_UBSAN)
    assert _run([exe]).startswith("OK")


def test_bvh_builder_threads_under_tsan(tmp_path):
    exe = str(tmp_path / "bvh_tsan")
    _build(exe, [os.path.join(HERE, "bvh_layout_check.cpp"), os.path.join(CSRC, "bvh_build.cpp")],
           ["-O1", "-g", "-fsanitize=thread"])
    # TSan reports a race as a warning and exits with status 66
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    # one threaded build per BVH width (2, 4, 8)
    assert _run([exe, "threaded"], env=env).startswith("OK 3")


def test_oracle_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    lib = os.path.join(ROOT, "oracle", "_san", "liboracle.so")
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0", DORT_ORACLE_LIB=lib)
    # the sanitized build is the one the process loads
    probe = ("from oracle import oracle; oracle.lib(); "
             "import sys; sys.exit(0 if '_san/liboracle.so' in open('/proc/self/maps').read() else 3)")
    _run([sys.executable, "-c", probe], env=env)
    out = _run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                os.path.join(HERE, "test_oracle.py"), os.path.join(HERE, "test_oracle_pins.py")],
               env=env, timeout=900)
    assert " passed" in out and "failed" not in out
