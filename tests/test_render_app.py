"""The .env front-end executable (pmdas_amd/mdas_render): the reference
apps' main loop over libmdas."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "pmdas_amd", "mdas_render")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _log(path):
    lines = open(path).read().split("\n")
    out = {}
    for i in range(len(lines) - 1):
        if lines[i].isupper() and lines[i]:
            try:
                out.setdefault(lines[i], []).append(float(lines[i + 1]))
            except ValueError:
                pass
    return out


def test_usage_without_arguments():
    if not os.path.exists(EXE):
        pytest.skip("mdas_render not built")
    r = subprocess.run([EXE], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


def test_bad_options_are_refused():
    """An option without a value or an unknown option ends the run with
    status 2 before anything touches the GPU."""
    if not os.path.exists(EXE):
        pytest.skip("mdas_render not built")
    env = os.path.join(GOLDEN, "pool.env")
    r = subprocess.run([EXE, env, "--seed", "3", "--out"], capture_output=True, text=True)
    assert r.returncode == 2 and "--out has no value" in r.stderr
    r = subprocess.run([EXE, env, "--bogus", "1"], capture_output=True, text=True)
    assert r.returncode == 2 and "unknown option --bogus" in r.stderr


def _read_pfm(path):
    raw = open(path, "rb").read()
    head = raw.split(b"\n", 3)
    w, h = map(int, head[1].split())
    return np.frombuffer(head[3], "<f4").reshape(h, w, 3)


@pytest.mark.gpu
def test_pool_env_runs(tmp_path):
    out = str(tmp_path / "pool")
    r = subprocess.run([EXE, os.path.join(GOLDEN, "pool.env"), "--out", out, "--seed", "3"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    log = _log(out + ".log")
    assert log["WIDTH"] == [1024] and log["HEIGHT"] == [1024]
    assert log["TOTAL SAMPLES"][0] >= 4 * 1024 * 1024
    img = _read_pfm(out + ".pfm")
    assert img.shape == (1024, 1024, 3) and np.isfinite(img).all()
    from pmdas_amd import metrics
    exr = metrics.read_exr(out + ".exr")  # rows bottom to top, like the PFM
    assert np.array_equal(exr[..., :3], img) and (exr[..., 3] == 1).all()
    dens = _read_pfm(out + "-density.pfm")
    assert dens.shape == (1024, 1024, 3) and dens.max() > 0
    assert open(out + ".ppm", "rb").read(15).startswith(b"P6\n1024 1024\n")


@pytest.mark.gpu
def test_plain_sampler_env(tmp_path):
    env = tmp_path / "plain.env"
    env.write_text("Sampler {\n  mdas false\n  samples 8\n}\nFilm {\n  width 96\n  height 64\n}\n"
                   "Scene {\n  kind dof\n  seed 2\n}\n")
    out = str(tmp_path / "plain")
    r = subprocess.run([EXE, str(env), "--out", out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    img = _read_pfm(out + ".pfm")
    assert img.shape == (64, 96, 3) and np.isfinite(img).all() and img.max() > 0
    assert "TRACE TIME" in _log(out + ".log")
