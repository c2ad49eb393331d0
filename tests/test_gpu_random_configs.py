"""Parity over a seeded random matrix of small configurations: dimension 3..6,
bitsPerDim 0..2, extraImgBits, alpha (including 0 and 1), candidatesNum 1..8,
every analytic scene of the dimension, odd image sizes and sample targets.
Each runs the device-side loop (optixMotionBlur.cpp:768-791) and the
reconstruction, and must equal the oracle bit for bit: final state, pass
count, image and sRGB bytes.  MDAS_RANDOM_CONFIGS, MDAS_RANDOM_SEED and
MDAS_RANDOM_SCALE (image sides multiplied by it) set the matrix for a
one-off sweep; the default is 200 configurations from seed 2024 at scale 1
(about 5 s on one MI355X)."""
import os
import random

import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.gpu

SCENES = {3: (1, 5, 0), 4: (2, 5, 0), 5: (3, 5, 0), 6: (4, 0)}


def _configs(n=200, seed=2024, scale=1):
    rng = random.Random(seed)
    out = []
    for k in range(n):
        dim = rng.choice((3, 4, 5, 6))
        b = rng.choice((0, 0, 1) if dim >= 5 else (0, 1, 1, 2))
        e = rng.randint(1, 4)
        while b * dim + 2 * e > 12:
            e -= 1
        W, H = rng.randint(5, 120) * scale, rng.randint(3, 90) * scale
        spp = rng.choice((4, 6, 9, 16, 24))
        out.append(dict(dim=dim, b=b, e=e, W=W, H=H, spp=spp,
                        alpha=rng.choice((0.0, 0.125, 0.25, 0.5, 1.0)),
                        cand=rng.choice((1, 2, 4, 8)), kind=rng.choice(SCENES[dim]),
                        seed=rng.randint(0, 1 << 20), id=k))
    return out


CONFIGS = _configs(int(os.environ.get("MDAS_RANDOM_CONFIGS", "200")),
                   int(os.environ.get("MDAS_RANDOM_SEED", "2024")),
                   int(os.environ.get("MDAS_RANDOM_SCALE", "1")))


@pytest.fixture(scope="module")
def mdas():
    import pmdas_amd
    from pmdas_amd import kdgrid, hip
    pmdas_amd.load()
    assert hip.device_count() >= 1, "no GPU visible"
    return kdgrid


@pytest.mark.parametrize("c", CONFIGS, ids=[f"cfg{c['id']}-d{c['dim']}" for c in CONFIGS])
def test_random_config_bit_exact(mdas, c):
    print(c, flush=True)
    W, H, dim = c["W"], c["H"], c["dim"]
    target = max(1, int(c["spp"] * W * H))
    cap = int(target * 1.6) + (4 << (c["b"] * dim + 2 * c["e"])) + 64
    args = (dim, cap, c["cand"], c["b"], c["e"], c["alpha"], float(W), float(H))
    sc = mdas.Scene(c["kind"], c["seed"], dim, float(W), float(H))
    g = mdas.KDGrid(*args)
    passes = g.Run(sc, target)
    o = pyoracle.OracleGrid(*args)
    assert passes == o.run(sc.params_bytes(), target)
    assert (g.GetNumberOfSamples(), g.GetNumberOfNodes()) == (o.n_samples, o.n_leaves)
    gs, os_ = g.download_state(), o.state()
    for k in gs:
        a, r = gs[k], os_[k]
        eq = a.view(np.uint32) == r.view(np.uint32) if a.dtype == np.float32 else a == r
        assert eq.all(), (k, np.argwhere(~eq)[:3].tolist())
    px = g.integrate_image(W, H)
    want = o.integrate(W, H)
    assert np.array_equal(px.view(np.uint32), want.view(np.uint32))
    from pmdas_amd import hip
    bp, bb = hip.DeviceBuffer(px.nbytes), hip.DeviceBuffer(W * H * 4)
    bp.upload(px)
    g.ConvertToBytes(bp.ptr, bb.ptr, W, H)
    hip.synchronize()
    assert np.array_equal(bb.download(np.uint8, W * H * 4).reshape(-1, 4),
                          pyoracle.convert_to_bytes(want))
    bp.free()
    bb.free()
    assert g.Validate()
    g.close()
    o.close()
    sc.close()


# grids that start from a single leaf (bitsPerDim 0, extraImgBits 0) or
# reach their target with no pass at all, on 1- and 2-pixel sample spaces:
# most of the launch's chunks and workgroups have no leaf to work on
# dim, b, e, W, H, spp, scene, alpha, candidatesNum
TINY = [(3, 0, 0, 1, 1, 8, 5, 0.25, 4), (4, 0, 0, 1, 1, 100, 5, 0.25, 4),
        (5, 0, 0, 1, 1, 64, 0, 0.0, 1), (6, 0, 0, 2, 1, 50, 4, 1.0, 8),
        (5, 1, 0, 1, 1, 40, 3, 0.25, 4), (3, 0, 1, 2, 2, 30, 1, 0.5, 2)]


@pytest.mark.parametrize("c", TINY, ids=[f"tiny{i}-d{c[0]}" for i, c in enumerate(TINY)])
def test_tiny_grid_bit_exact(mdas, c):
    dim, b, e, W, H, spp, kind, alpha, cand = c
    print(c, flush=True)
    target = spp * W * H
    cap = int(target * 1.6) + (4 << (b * dim + 2 * e)) + 64
    args = (dim, cap, cand, b, e, alpha, float(W), float(H))
    sc = mdas.Scene(kind, 5, dim, float(W), float(H))
    g = mdas.KDGrid(*args)
    o = pyoracle.OracleGrid(*args)
    assert g.Run(sc, target) == o.run(sc.params_bytes(), target)
    gs, os_ = g.download_state(), o.state()
    for k in gs:
        a, r = gs[k], os_[k]
        eq = a.view(np.uint32) == r.view(np.uint32) if a.dtype == np.float32 else a == r
        assert eq.all(), (k, np.argwhere(~eq)[:3].tolist())
    for w, h in ((W, H), (3, 2)):
        px = g.integrate_image(w, h)
        assert np.array_equal(px.view(np.uint32), o.integrate(w, h).view(np.uint32))
    assert g.Validate()
    g.close()
    o.close()
    sc.close()
