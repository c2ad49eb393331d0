"""Two grids driven from two host threads at once, each on its own stream:
the library keeps all state per handle (the reference keeps its counters and
error maximum in device globals, kdgrid.cu:19-23, and cannot run two grids
concurrently), so each result equals the same run done alone, bit for bit,
and each thread reads its own last error message."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# dim, W, H, spp, scene kind, seed
SPECS = [(5, 160, 96, 16, 3, 11), (4, 192, 108, 12, 2, 12)]


@pytest.fixture(scope="module")
def mdas():
    import pmdas_amd
    from pmdas_amd import kdgrid, hip
    pmdas_amd.load()
    assert hip.device_count() >= 1, "no GPU visible"
    return kdgrid


def _make(mdas, spec):
    dim, W, H, spp, kind, seed = spec
    target = spp * W * H
    g = mdas.KDGrid(dim, int(target * 1.6) + (4 << (dim + 4)), 4, 1, 2, 0.25, float(W), float(H))
    sc = mdas.Scene(kind, seed, dim, float(W), float(H))
    return g, sc, target


def _same(a, b, what):
    for k in a:
        x, y = a[k], b[k]
        eq = x.view(np.uint32) == y.view(np.uint32) if x.dtype == np.float32 else x == y
        assert eq.all(), (what, k)


def test_two_grids_from_two_threads(mdas):
    from pmdas_amd import hip
    made = [_make(mdas, s) for s in SPECS]
    alone = []
    for (g, sc, t), s in zip(made, SPECS):
        g.Run(sc, t)
        alone.append((g.download_state(), g.integrate_image(s[1], s[2])))
    streams = [hip.Stream() for _ in made]
    results, errors = [None, None], [None, None]

    def work(k):
        try:
            g, sc, t = made[k]
            g.stream = streams[k]
            for _ in range(3):
                g.Run(sc, t)
            img = g.integrate_image(SPECS[k][1], SPECS[k][2])
            results[k] = (g.download_state(), img)
        except Exception as e:  # noqa: BLE001 -- reported below
            errors[k] = e

    threads = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert errors == [None, None]
    for k in range(2):
        _same(results[k][0], alone[k][0], f"grid {k} state")
        assert np.array_equal(results[k][1].view(np.uint32), alone[k][1].view(np.uint32)), k
    for (g, sc, _), st in zip(made, streams):
        g.stream = None
        g.close()
        sc.close()
        st.destroy()


def test_last_error_is_per_thread(mdas):
    """Failing calls of two threads interleave; each sees its own message."""
    from pmdas_amd import MdasError
    made = [_make(mdas, s) for s in SPECS]
    for g, sc, t in made:
        g.InitialSampling()
    bad = []

    def work(k):
        g = made[k][0]
        for _ in range(300):
            try:
                if k == 0:
                    g.set_recon_tile(3, 5)  # not a power-of-two tile
                else:
                    g.set_launch_shape(1, 1024)  # too few workgroups
            except MdasError as e:
                want = "tile" if k == 0 else "launch shape"
                if want not in str(e):
                    bad.append((k, str(e)))
            else:
                bad.append((k, "no error"))

    threads = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not bad, bad[:3]
    for g, sc, _ in made:
        g.close()
        sc.close()
