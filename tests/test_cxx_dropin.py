"""The C++ header mirror of KDGrid takes the reference app's KDGrid lines
unchanged (tests/cxx/drop_in.cpp: initMdas, samplingPassMdas, integrateMdas
and the batch loop of optixMotionBlur.cpp, with its std::ofstream log and
float3* sample values).  On the CPU it compiles and links; on the GPU it runs
the app's loop and writes the reference's log lines to the app's stream."""
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
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_drop_in_compiles_and_links(tmp_path):
    _build(str(tmp_path / "drop_in"))
    assert os.path.exists(tmp_path / "drop_in")


def _log(path):
    lines = open(path).read().split("\n")
    keys = [(lines[i], lines[i + 1]) for i in range(0, len(lines) - 1, 2)]
    return keys


@pytest.mark.gpu
def test_drop_in_runs(tmp_path):
    exe = str(tmp_path / "drop_in")
    _build(exe)
    out = str(tmp_path / "frame.png")
    r = subprocess.run([exe, out], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "valid 1 values 1" in r.stdout
    summary = [ln for ln in r.stdout.splitlines() if ln.startswith("passes ")][0].split()
    passes, samples = int(summary[1]), int(summary[3])
    # the app's stream: its own WIDTH / HEIGHT, then the library's lines in the
    # order the reference writes them (kdgrid.cu:670-671, 713, 769-770, 816-817)
    kv = _log(out + ".log")
    keys = [k for k, _ in kv]
    want = (["WIDTH", "HEIGHT", "INITIAL SAMPLES", "INITIAL SAMPLING TIME"] +
            ["COMPUTE ERRORS TIME", "SAMPLES", "ADAPTIVE SAMPLING TIME"] * passes +
            ["TOTAL SAMPLES", "CONSTRUCT TIME", "INTEGRATE TIME"])
    assert keys == want
    d = dict(kv)
    assert d["WIDTH"] == "160" and d["HEIGHT"] == "90"
    assert int(d["TOTAL SAMPLES"]) == samples  # an integer, as the reference prints it
    assert int(kv[2][1]) == 4 << (5 + 6)  # INITIAL SAMPLES: 4 * 2^(b D + 2 e)
    new = [int(v) for k, v in kv if k == "SAMPLES"]
    assert sum(new) == samples - (4 << 11)
    for k, v in kv:
        if k.endswith("TIME"):
            assert float(v) >= 0.0
