"""The allocation-free reconstruction entry point (mdas_integrate_noalloc)
through the C ABI against mdas_integrate: two calls in a row, the second on
a smaller tree and image, each bit-exact; a reconstruction larger than the
scratch array is refused before anything runs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mdas():
    import pmdas_amd
    from pmdas_amd import kdgrid, hip
    pmdas_amd.load()
    assert hip.device_count() >= 1, "no GPU visible"
    return kdgrid


def _noalloc(g, buf, W, H):
    from pmdas_amd import _lib
    return _lib.load().mdas_integrate_noalloc(ctypes.c_void_p(g.handle), ctypes.c_void_p(buf.ptr),
                                              W, H, None)


def test_noalloc_equals_integrate_twice(mdas):
    from pmdas_amd import hip
    # (dim, W, H, spp, scene, seed, b, e): the second call's tree and image are smaller
    for dim, W, H, spp, kind, seed, b, e in [(5, 240, 136, 24, 3, 3, 1, 3),
                                             (4, 72, 40, 8, 2, 5, 1, 2)]:
        target = spp * W * H
        g = mdas.KDGrid(dim, int(target * 1.6) + (4 << (b * dim + 2 * e)), 4, b, e, 0.25,
                        float(W), float(H))
        sc = mdas.Scene(kind, seed, dim, float(W), float(H))
        g.Run(sc, target)
        want = g.integrate_image(W, H)
        buf = hip.DeviceBuffer(W * H * 16)
        buf.upload(np.full(W * H * 4, np.nan, np.float32))
        assert _noalloc(g, buf, W, H) == 0
        hip.synchronize()
        got = buf.download(np.float32, W * H * 4).reshape(H, W, 4)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (dim, W, H)
        buf.free()
        g.close()
        sc.close()


def test_noalloc_refuses_a_reconstruction_beyond_its_scratch(mdas):
    from pmdas_amd import _lib, hip
    # a 1920x1080 5D initial grid with extraImgBits 9: 8.4 M leaves, 268 MB of
    # leaf records alone
    g = mdas.KDGrid(5, (4 << 23) + 1024, 4, 1, 9, 0.25, 1920.0, 1080.0)
    g.InitialSampling()
    buf = hip.DeviceBuffer(1920 * 1080 * 16)
    assert _noalloc(g, buf, 1920, 1080) == _lib.MDAS_ERR_CAPACITY
    assert "scratch" in _lib.load().mdas_last_error().decode()
    buf.free()
    g.close()
