"""The backend's host C++ (generators.cpp, symbolic.cpp, hsd.cpp) built with
AddressSanitizer and UndefinedBehaviourSanitizer (g++, CPU only) and run by
tests/sanitize_host.cpp over every instance generator, both orderings, the
file formats and the HSD driver with the CPU oracle backend."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "help_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_host_code_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "sanitize_host")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fopenmp", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-I", os.path.join(ROOT, "oracle"),
           os.path.join(HERE, "sanitize_host.cpp"), os.path.join(CSRC, "symbolic.cpp"),
           os.path.join(CSRC, "generators.cpp"), os.path.join(CSRC, "hsd.cpp"),
           "-x", "c", os.path.join(ROOT, "oracle", "oracle.c"), "-x", "none", "-o", exe, "-lm"]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=900, env=env)
    print(r.stdout[-2000:])
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-6000:])
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
