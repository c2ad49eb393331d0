"""bench.py's command-line contract on a host without a GPU: more than one
GPU is refused with a message (the sampler runs on one GPU), and the help
text names the flags a benchmark harness passes (--gpus, --steps, --warmup)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2", "--steps", "1", "--warmup", "0")
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
    assert not r.stdout.strip().startswith("{")  # no result line


def test_help_names_the_harness_flags():
    r = _bench("--help")
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup"):
        assert flag in r.stdout


def test_phase_model_bytes_counts_each_term():
    """The whole-run byte model: a grid with one pass, checked term by term
    against counts written out for D = 5 (node 16 B, box 40 B, point 20 B,
    value 12 B, error and seed 4 B each)."""
    import numpy as np
    sys.path.insert(0, ROOT)
    import bench
    w = dict(dim=5, b=1, e=1, W=4, H=2)  # L0 = 2^(5 + 2) = 128 leaves
    # one pass over 128 leaves: 10 selected holding 40 samples, 10 splits,
    # 128 leaves re-estimated reading 512 values
    rows = np.array([[128, 512, 10, 10, 40, 128, 512, 0]], np.int32)
    b = bench.phase_model_bytes(w, rows, 138, 522, None)
    assert b["initial"] == 128 * (4 * 20 + 16 + 40 + 4 + 4)
    assert b["errors"] == 128 * 4 + 128 * (16 + 40 + 4) + 512 * 12
    assert b["adaptive"] == (128 * (4 + 40 + 4 + 4) + 10 * (16 + 20 + 16) + 40 * 20
                             + 10 * (16 + 2 * 40 + 2 * 4))
    assert b["recon"] == 138 * (16 + 40) + 522 * 12 + 8 * 16
    res = dict(kernels={"k_place": dict(launches=1, ms=2.0, bytes=4e9)}, mdas_ms=4.0,
               phase_model_bytes=b, phases_ms={k + "_ms": 1.0 for k in b})
    r = bench.roofline(res, os.path.join(ROOT, "no-such-file.json"))
    assert r["achieved"] == 2000.0 and r["frac"] == 0.25
    assert r["whole_run"]["bytes"] == round(sum(b.values()))
    assert abs(r["whole_run"]["achieved"] - sum(b.values()) / 4e-3 / 1e9) < 0.1


def test_roofline_traffic_rate(tmp_path):
    """With a traffic file for the dominant kernel, the line carries the
    counted traffic over the same launch time and over the copy probe; a
    file for another kernel is ignored."""
    import json
    sys.path.insert(0, ROOT)
    import bench
    res = dict(kernels={"k_place": dict(launches=4, ms=2.0, bytes=4e9),
                        "k_eval": dict(launches=4, ms=9.0, bytes=1e9)}, mdas_ms=4.0)
    f = tmp_path / "traffic.json"
    f.write_text(json.dumps({"kernel": "k_place", "hbm_bytes_per_launch": 1.5e9}))
    r = bench.roofline(res, str(f), copy_gbs=5000.0)
    assert r["kernel"] == "k_place"  # the integrand is not an MDAS kernel
    assert r["avg_launch_ms"] == 0.5 and r["traffic"] == 1.5e9
    assert r["traffic_GBps"] == 3000.0 and r["traffic_frac_of_copy"] == 0.6
    assert r["frac_of_copy"] == 0.4
    f.write_text(json.dumps({"kernel": "k_errors", "hbm_bytes_per_launch": 1.5e9}))
    r = bench.roofline(res, str(f), copy_gbs=5000.0)
    assert r["traffic"] is None and "traffic_GBps" not in r
