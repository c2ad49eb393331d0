"""Ends of the device-side sampling loop through the C ABI.

- A target beyond what max_samples can hold: the loop stops before a pass
  that could overflow, N + L > max_samples (the app's interactive guard,
  optixMotionBlur.cpp:723-724).  The run ends cleanly, equal to the oracle's
  loop with the same rule, and the .env front-end still writes its image.
- No leaf selected.  Non-finite errors (a NaN integrand) or a zero maximum
  error (learned scores that underflow) make a block of passes add nothing.
  mdas_run and the batch then return MDAS_ERR_STATE instead of looping.
- A run of more passes than the device's per-pass ring holds keeps every
  pass's row: the pass statistics and the log's SAMPLES lines.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def mdas():
    import pmdas_amd
    from pmdas_amd import kdgrid, hip
    pmdas_amd.load()
    assert hip.device_count() >= 1, "no GPU visible"
    return kdgrid


def _same(gs, os_, what):
    for k in gs:
        a, r = gs[k], os_[k]
        eq = a.view(np.uint32) == r.view(np.uint32) if a.dtype == np.float32 else a == r
        assert eq.all(), (what, k, np.argwhere(~eq)[:3].tolist())


@pytest.mark.parametrize("cap", [300, 3000, 20000])
def test_run_beyond_capacity_ends_cleanly(mdas, cap):
    args = (3, cap, 4, 0, 3, 0.25, 32.0, 32.0)  # 64 initial leaves, 256 samples
    sc = mdas.Scene(1, 5, 3, 32.0, 32.0)
    g = mdas.KDGrid(*args)
    o = pyoracle.OracleGrid(*args)
    passes = g.Run(sc, 10**7)
    assert passes == o.run(sc.params_bytes(), 10**7)
    N, L = g.GetNumberOfSamples(), g.GetNumberOfNodes()
    assert N <= cap < N + L  # stopped by the guard, not by the target
    _same(g.download_state(), o.state(), f"cap {cap}")
    assert g.Validate()
    # the same in a batch
    g2 = mdas.KDGrid(*args)
    b = mdas.Batch([g2], [sc])
    b.run([10**7])
    _same(g2.download_state(), o.state(), f"cap {cap}, batch")
    b.close()
    for x in (g, g2, o, sc):
        x.close()


def test_render_app_beyond_capacity(tmp_path):
    exe = os.path.join(ROOT, "pmdas_amd", "mdas_render")
    env = tmp_path / "big.env"
    env.write_text("Sampler {\n  mdas true\n  samples 64\n}\nMdas {\n  bitsPerDim 0\n"
                   "  extraImgBits 4\n}\nFilm {\n  width 64\n  height 64\n}\n"
                   "Scene {\n  kind mbdof\n  seed 2\n}\n")
    out = str(tmp_path / "big")
    # 64 spp x 64 x 64 = 262144 samples asked for, 40000 allowed
    r = subprocess.run([exe, str(env), "--out", out, "--max-samples", "40000"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(out + ".pfm") and os.path.exists(out + ".exr")
    samples = int(r.stdout.split("samples ")[1].split(",")[0])
    assert 20000 < samples <= 40000


def _nan_scene(mdas):
    # the ramp integrand is (x / sX, y / sY, t): a NaN sX makes every x value NaN
    return mdas.Scene(5, 0, 3, float("nan"), 16.0)


def test_nan_integrand_stops_with_state_error(mdas):
    from pmdas_amd import MdasError
    args = (3, 1 << 16, 4, 0, 2, 0.25, 16.0, 16.0)
    sc = _nan_scene(mdas)
    g = mdas.KDGrid(*args)
    with pytest.raises(MdasError) as ei:
        g.Run(sc, 4096)
    assert ei.value.code == -4 and "no leaf selected" in str(ei.value)
    assert g.GetNumberOfSamples() == 64  # nothing was added to the initial grid
    g2 = mdas.KDGrid(*args)
    b = mdas.Batch([g2], [sc])
    with pytest.raises(MdasError) as ei:
        b.run([4096])
    assert ei.value.code == -4 and "no leaf selected" in str(ei.value)
    for x in (b, g, g2, sc):
        x.close()


def test_zero_scores_stop_with_state_error(mdas):
    """Learned scores e^16 of errors ~1e-5 underflow to 0: the maximum score
    is 0 and no leaf passes the selection test."""
    from pmdas_amd import MdasError
    args = (3, 1 << 16, 4, 0, 3, 0.25, 8.0, 8.0)  # 64 leaves of volume 1
    sc = mdas.Scene(0, 0, 3, 8.0, 8.0)  # constant: every error is 1e-5 x volume
    m = mdas.Mlp(1)
    g = mdas.KDGrid(*args)
    g.set_learned(2, m)
    g.set_learned_exponents(16.0, 0.0)
    with pytest.raises(MdasError) as ei:
        g.Run(sc, 4096)
    assert ei.value.code == -4 and "no leaf selected" in str(ei.value)
    g2 = mdas.KDGrid(*args)
    g2.set_learned(2, m)
    g2.set_learned_exponents(16.0, 0.0)
    b = mdas.Batch([g2], [sc])
    with pytest.raises(MdasError) as ei:
        b.run([4096])
    assert ei.value.code == -4
    for x in (b, g, g2, sc, m):
        x.close()


def test_pass_rows_past_the_ring(mdas, tmp_path):
    """7573 passes (more than the 4096-row device ring): every pass keeps its
    row, and the log has a nonzero SAMPLES line and a time per pass."""
    target = 60000
    args = (3, int(target * 1.6) + 68, 4, 0, 0, 1.0, 1.0, 1.0)
    sc = mdas.Scene(1, 3, 3, 1.0, 1.0)
    log = str(tmp_path / "run.log")
    g = mdas.KDGrid(*args, log=log)
    passes = g.Run(sc, target)
    assert passes > 4096
    rows = g.pass_stats(max_rows=passes + 10)
    assert rows.shape[0] == passes
    assert (rows[:, 2] > 0).all()  # selected samples of every pass
    # consecutive rows chain: N after pass p = N before pass p + 1
    assert np.array_equal(rows[1:, 1], rows[:-1, 1] + rows[:-1, 2])
    lines = open(log).read().split("\n")
    samples = [int(float(lines[i + 1])) for i in range(len(lines) - 1) if lines[i] == "SAMPLES"]
    assert len(samples) == passes and min(samples) > 0
    assert samples == rows[:, 2].tolist()
    # the timed launches are folded block by block: every pass keeps its
    # time, and the logged error-pass times add up to the kernel statistics'
    # total (which also counts the no-op passes after the limit in the last
    # block of passes)
    err_t = [float(lines[i + 1]) for i in range(len(lines) - 1) if lines[i] == "COMPUTE ERRORS TIME"]
    assert len(err_t) == passes and min(err_t) > 0.0
    launches, total_ms, _ = g.kernel_stats()["k_errors"]
    assert launches >= passes
    assert 0.9 * total_ms <= sum(err_t) <= total_ms * (1 + 1e-4) + 1e-3
    g.close()
    sc.close()
