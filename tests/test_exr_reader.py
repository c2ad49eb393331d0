"""The EXR files this build writes, read back by the reference's own EXR
reader: tinyexr (SDK/support/tinyexr/tinyexr.h, which the reference's sutil
loads and saves its images with, SDK/sutil/sutil.cpp:311, 698-728), compiled
from where it lies in the reference tree into oracle/_ref/ by
`make -C oracle ref`.  Both writers are checked, pixel for pixel and bit for
bit: pmdas_amd.metrics.write_exr and the .env front-end's C++ writer
(pmdas_amd/csrc/image_io.hpp), the latter also against its own PFM and PPM
output.  CPU only; skipped where the reference tree is absent."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/SDK/support/tinyexr/tinyexr.h"
LIB = os.path.join(ROOT, "oracle", "_ref", "libexr_ref.so")


@pytest.fixture(scope="module")
def tinyexr():
    if not os.path.exists(LIB):
        if not os.path.exists(REF):
            pytest.skip("the reference's tinyexr is not present")
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    lib = ctypes.CDLL(LIB)
    lib.exr_ref_load.restype = ctypes.c_int
    lib.exr_ref_load.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_longlong,
                                 ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]

    def load(path, cap=1 << 24):
        out = np.zeros(cap, np.float32)
        w, h = ctypes.c_int(), ctypes.c_int()
        rc = lib.exr_ref_load(path.encode(), out.ctypes.data, cap, ctypes.byref(w), ctypes.byref(h))
        assert rc == 0, f"tinyexr LoadEXR -> {rc}"
        return out[:w.value * h.value * 4].reshape(h.value, w.value, 4)
    return load


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_python_writer(tinyexr, tmp_path):
    from pmdas_amd import metrics
    rng = np.random.default_rng(3)
    img = rng.standard_normal((37, 53, 4)).astype(np.float32)  # row 0 at the bottom
    img[0, :5, 0] = [0.0, -0.0, np.finfo(np.float32).tiny / 8, np.finfo(np.float32).max, np.inf]
    p = str(tmp_path / "py.exr")
    metrics.write_exr(p, img)
    got = tinyexr(p)  # rows as stored: top first
    assert got.shape == img.shape
    assert np.array_equal(_bits(got), _bits(img[::-1]))
    metrics.write_exr(p, img[..., :3])  # RGB in: alpha 1
    rgb = tinyexr(p)
    assert np.array_equal(_bits(rgb[..., :3]), _bits(img[::-1, :, :3])) and (rgb[..., 3] == 1).all()


def test_cxx_writer(tinyexr, tmp_path):
    exe = str(tmp_path / "exr_write")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "pmdas_amd", "csrc"),
                        os.path.join(ROOT, "tests", "cxx", "exr_write.cpp"), "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    W, H = 61, 29
    out = str(tmp_path / "img")
    assert subprocess.run([exe, out, str(W), str(H)]).returncode == 0
    got = tinyexr(out + ".exr")[::-1]  # back to row 0 at the bottom
    y, x = np.mgrid[0:H, 0:W]
    r_ = (x.astype(np.float32) + y.astype(np.float32) / np.float32(1000.0)).astype(np.float32)
    r_.reshape(-1)[:5] = [0.0, -0.0, np.float32(2.0 ** -149), np.finfo(np.float32).max, np.inf]
    want = np.stack([r_, -y.astype(np.float32), np.ldexp(np.float32(1.0), (x % 40) - 20).astype(np.float32),
                     np.ones((H, W), np.float32)], axis=-1)
    assert np.array_equal(_bits(got), _bits(want))
    # the PFM holds the same RGB, bottom row first; the PPM the bytes, top row first
    raw = open(out + ".pfm", "rb").read()
    head = raw.split(b"\n", 3)
    assert head[0] == b"PF" and head[1] == b"%d %d" % (W, H) and float(head[2]) < 0
    pfm = np.frombuffer(head[3], "<f4").reshape(H, W, 3)
    assert np.array_equal(_bits(pfm), _bits(want[..., :3]))
    ppm = open(out + ".ppm", "rb").read()
    assert ppm.startswith(b"P6\n%d %d\n255\n" % (W, H))
    px = np.frombuffer(ppm[len(b"P6\n%d %d\n255\n" % (W, H)):], np.uint8).reshape(H, W, 3)[::-1]
    assert np.array_equal(px, ((x[..., None] * 7 + y[..., None] * 3 + np.arange(3)) & 255).astype(np.uint8))
