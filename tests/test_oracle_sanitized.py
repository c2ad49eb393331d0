"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

oracle/mdas_oracle.cpp is compiled together with oracle/sanitize_driver.cpp
(every oracle entry point the tests use, over small grids of every dimension
and scene) with -fsanitize=address,undefined and no recovery, and run as its
own process: an out-of-bounds access, a use after free, a leak or undefined
arithmetic in the oracle ends it with an error report, which fails this
test.  ThreadSanitizer is not used: the system's OpenMP runtime is not built
for it, so it would report the runtime's own synchronisation."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_under_asan_ubsan(tmp_path):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no g++")
    exe = str(tmp_path / "sanitize_driver")
    cmd = [cxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fopenmp",
           "-ffp-contract=off", "-I", os.path.join(ROOT, "pmdas_amd", "csrc"),
           "-I", os.path.join(ROOT, "oracle"), os.path.join(ROOT, "oracle", "mdas_oracle.cpp"),
           os.path.join(ROOT, "oracle", "sanitize_driver.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    env = dict(os.environ, OMP_NUM_THREADS="4",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "sanitize_driver: ok" in r.stdout
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-4000:]
