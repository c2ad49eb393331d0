"""Error metrics (parse_logs.py:43-50 definitions) by closed forms, and the
image writers by reading the files back (CPU)."""
import struct

import numpy as np

from pmdas_amd import metrics


def test_mse_closed_form():
    ref = np.zeros((2, 2, 4))
    img = np.zeros((2, 2, 4))
    img[0, 0, :3] = 1.0  # 3 of 12 channel values off by 1
    assert metrics.mse(img, ref) == 3 / 12
    img[..., 3] = 99.0  # alpha ignored
    assert metrics.mse(img, ref) == 3 / 12


def test_rel_mse_closed_form_and_nan():
    ref = np.full((1, 2, 3), 0.3)
    img = ref.copy()
    img[0, 0, 0] = 0.4
    want = (0.1 ** 2 / (0.09 + 1e-2)) / 6
    assert np.isclose(metrics.rel_mse(img, ref), want)
    img[0, 1, 1] = np.nan  # a NaN pixel counts as 0: error ref^2
    assert np.isclose(metrics.rel_mse(img, ref), want + (0.09 / (0.09 + 1e-2)) / 6)
    assert np.isclose(metrics.mse(img, ref), (0.01 + 0.09) / 6)
    assert not np.isnan(img[0, 0, 0]) and np.isnan(img[0, 1, 1])  # input untouched


def test_pfm_and_ppm(tmp_path):
    a = np.random.default_rng(0).random((3, 5, 4), dtype=np.float32)
    metrics.write_pfm(tmp_path / "a.pfm", a)
    raw = (tmp_path / "a.pfm").read_bytes()
    assert raw.startswith(b"PF\n5 3\n-1.0\n")
    back = np.frombuffer(raw[len(b"PF\n5 3\n-1.0\n"):], "<f4").reshape(3, 5, 3)
    assert np.array_equal(back, a[..., :3])
    b8 = (a * 255).astype(np.uint8)
    metrics.write_ppm(tmp_path / "a.ppm", b8)
    raw = (tmp_path / "a.ppm").read_bytes()
    body = np.frombuffer(raw[len(b"P6\n5 3\n255\n"):], np.uint8).reshape(3, 5, 3)
    assert np.array_equal(body, b8[::-1, :, :3])


def test_exr_round_trip_and_layout(tmp_path):
    a = np.random.default_rng(1).random((4, 6, 4), dtype=np.float32)
    metrics.write_exr(tmp_path / "a.exr", a)
    raw = (tmp_path / "a.exr").read_bytes()
    assert struct.unpack_from("<ii", raw, 0) == (20000630, 2)  # OpenEXR magic, version 2
    assert b"channels\0chlist\0" in raw and b"compression\0compression\0" in raw
    assert np.array_equal(metrics.read_exr(tmp_path / "a.exr"), a)


def _rows(curves, spps, seeds=(0, 1)):
    """Synthetic comparison rows: arm -> (samples(spp), error(samples))."""
    rows = []
    for s in seeds:
        for spp in spps:
            r = dict(seed=s, spp=spp)
            for arm, (n_of, e_of) in curves.items():
                n = n_of(spp)
                r[arm] = dict(mse=e_of(n), rel_mse=2 * e_of(n), samples=n, ms=n * 1e-6, integrand_ms=0.0)
            rows.append(r)
    return rows


def test_error_at_equal_samples_power_laws():
    """Arms with error c * n^-1 sampled at different counts: at equal samples
    the ratio is the ratio of the constants, whatever each arm's overshoot."""
    from pmdas_amd import learned
    spps = (4, 8, 16, 32, 64)
    rows = _rows({"plain": (lambda s: 1000.0 * s, lambda n: 3.0 / n),
                  "learned": (lambda s: 1300.0 * s, lambda n: 1.5 / n),
                  "same": (lambda s: 1100.0 * s if s > 4 else 1000.0 * s, lambda n: 2.0 / n)}, spps)
    r = learned.error_at_equal_samples(rows, "learned", "plain", min_spp=8)
    # plain's 8 spp (8000 samples) lies below learned's first point (10400): left out
    assert sorted(r["ratios"], key=int) == ["16", "32", "64"]
    assert all(abs(v - 0.5) < 1e-12 for v in r["ratios"].values())
    assert abs(r["geomean"] - 0.5) < 1e-12
    r = learned.error_at_equal_samples(rows, "learned", "same", key="rel_mse", min_spp=8)
    assert all(abs(v - 0.75) < 1e-12 for v in r["ratios"].values())
    # samples to reach plain's 32-spp error 3/32000 on error 2/n: n = 21333.3
    eq = learned.samples_to_equal(rows, 32, "mse", "same", min_spp=8)
    assert "bound" not in eq
    assert abs(eq["sample_ratio"] - (2 / 3)) < 1e-9
    assert abs(eq["time_ratio"] - (2 / 3)) < 1e-9
    # an arm already below the target at its first adaptive point is a bound
    # (learned at 8 spp: 1.5/10400 < plain's 16-spp 3/16000)
    eq = learned.samples_to_equal(rows, 16, "mse", "learned", min_spp=8)
    assert eq["bound"] == "upper" and abs(eq["sample_ratio"] - 10400 / 16000) < 1e-12
    # and the first point counts only from min_spp on
    eq = learned.samples_to_equal(rows, 8, "mse", "learned", min_spp=4)
    assert eq["bound"] == "upper" and abs(eq["sample_ratio"] - 5200 / 8000) < 1e-12
    # a crossing: learned reaches plain's 64-spp error 3/64000 at n = 32000
    eq = learned.samples_to_equal(rows, 64, "mse", "learned", min_spp=8)
    assert "bound" not in eq and abs(eq["sample_ratio"] - 0.5) < 1e-9
