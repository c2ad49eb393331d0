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
