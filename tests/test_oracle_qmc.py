"""Pins of the oracle's plain QMC sampler (cuda/sampler.h:4-28,
random.h:74-89, whitted.cu:59-69), CPU only.

- halton: base-2 values are dyadic and exact; the first b^k points of the
  radical inverse in base b are a permutation of {j / b^k} (van der Corput).
- image: the ramp scene's red/green channels are the pixel jitter, so a
  1-sample image equals (px + offset0) / W with offset0 the first rnd of
  tea4(pixel, frame) -- recomputed here from the pinned TEA and LCG; the
  constant scene gives the constant exactly when spp is a power of two; the
  FP64 image agrees with the FP32 one to float rounding.
"""
import numpy as np

TEA_KEY = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)


def test_halton_base2_exact(orc):
    want = [0, 0.5, 0.25, 0.75, 0.125, 0.625, 0.375, 0.875, 0.0625]
    assert [orc.halton(0, i) for i in range(9)] == want


def test_halton_van_der_corput_permutation(orc):
    for dim, base, k in [(1, 3, 4), (2, 5, 3), (3, 7, 2), (4, 11, 2), (5, 13, 2)]:
        n = base ** k
        pts = np.array([orc.halton(dim, i) for i in range(n)], np.float64)
        assert np.allclose(np.sort(pts), np.arange(n) / n, atol=1e-6), dim
        # digit reversal: i = d0 + d1*b  ->  d0/b + d1/b^2
        for i in (1, base - 1, base, base + 2, n - 1):
            d = [(i // base ** j) % base for j in range(k)]
            assert abs(orc.halton(dim, i) - sum(dj / base ** (j + 1) for j, dj in enumerate(d))) < 1e-6


def _rnd_seq(orc, seed, n):
    out = []
    for _ in range(n):
        seed = (1664525 * seed + 1013904223) & 0xFFFFFFFF
        out.append((seed & 0xFFFFFF) / 2.0 ** 24)
    return out


def test_ramp_one_sample_is_offset(orc):
    W, H, sx, sy = 7, 5, 70.0, 25.0
    P = orc.make_scene(5, 1, 3, sx, sy)
    for frame in (0, 3):
        img = orc.qmc_image(P, W, H, 1, 0, frame)
        for py in range(H):
            for px in range(W):
                seed = orc.tea(py * W + px, frame, 4, TEA_KEY)[0]
                o = _rnd_seq(orc, seed, 3)
                r, g, b = img[py, px, :3]
                assert abs(r - (px + o[0]) / W) < 2e-6
                assert abs(g - (py + o[1]) / H) < 2e-6
                assert abs(b - o[2]) < 1e-7


def test_ramp_many_samples_near_pixel_centre(orc):
    W, H = 6, 4
    P = orc.make_scene(5, 1, 3, 30.0, 20.0)
    img = orc.qmc_image(P, W, H, 256, 0, 0)
    cx = (np.arange(W) + 0.5) / W
    cy = (np.arange(H) + 0.5) / H
    assert np.abs(img[..., 0] - cx[None, :]).max() < 0.01 / W
    assert np.abs(img[..., 1] - cy[:, None]).max() < 0.01 / H
    assert np.abs(img[..., 2] - 0.5).max() < 0.01


def test_constant_scene_exact(orc):
    P = orc.make_scene(0, 3, 4, 16.0, 8.0)
    img = orc.qmc_image(P, 8, 4, 8, 5, 1)
    c = orc.eval_scene(P, np.zeros((1, 4), np.float32), 4)[0]
    assert np.array_equal(img[..., :3], np.broadcast_to(c, (4, 8, 3)))


def test_fp64_close_to_fp32(orc):
    P = orc.make_scene(3, 2, 5, 40.0, 24.0)
    a = orc.qmc_image(P, 10, 6, 16, 0, 0)
    b = orc.qmc_image(P, 10, 6, 16, 0, 0, fp64=True)
    # geometric edges can flip a sample between float and double; the mean
    # difference stays at rounding level.
    assert np.mean(np.abs(a[..., :3] - b[..., :3])) < 2e-3
