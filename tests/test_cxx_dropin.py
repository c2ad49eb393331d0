"""The C++ header mirror of KDGrid compiles like the reference app's code
(CPU) and runs the reference's loop on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cxx", "drop_in.cpp")


def _build(out):
    cmd = ["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
           "-I", "/opt/rocm/include", SRC, "-o", out, "-L", os.path.join(ROOT, "pmdas_amd"),
           "-lmdas", "-L", "/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath," + os.path.join(ROOT, "pmdas_amd"), "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)


def test_drop_in_compiles_and_links(tmp_path):
    _build(str(tmp_path / "drop_in"))
    assert os.path.exists(tmp_path / "drop_in")


@pytest.mark.gpu
def test_drop_in_runs(tmp_path):
    exe = str(tmp_path / "drop_in")
    _build(exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "valid 1" in r.stdout
