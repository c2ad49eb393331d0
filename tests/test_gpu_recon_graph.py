"""Stream-ordered reconstruction (kdgrid.cu:781-820 has no host
synchronisation outside its logging).

- mdas_integrate, once its buffers are sized, can be captured into a
  caller's hipGraph: a host synchronisation or an allocation inside it would
  make the capture fail.  Every replay equals the eager reconstruction, with
  eager work and synchronisations between the replays.
- mdas_batch_integrate reconstructs every image of a batch as one captured
  graph.  The result equals each grid's own mdas_integrate, also on a replay
  after the batch has sampled again, and without the graph.
- A null-stream caller's batch graphs run on the batch's own stream, joined
  to the null stream on both sides: a plain hipMemcpy right after the call
  reads the finished image, and null-stream work issued before the call is
  done before the graph writes.
"""
import ctypes

import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mdas():
    import pmdas_amd
    from pmdas_amd import kdgrid, hip
    pmdas_amd.load()
    assert hip.device_count() >= 1, "no GPU visible"
    return kdgrid


def _grid(mdas, dim, W, H, spp, kind, seed, b=1, e=2):
    target = spp * W * H
    g = mdas.KDGrid(dim, int(target * 1.6) + (4 << (b * dim + 2 * e)), 4, b, e, 0.25, float(W), float(H))
    sc = mdas.Scene(kind, seed, dim, float(W), float(H))
    return g, sc, target


@pytest.mark.parametrize("tile", [(8, 8), (4, 4)])
def test_integrate_replays_from_a_captured_graph(mdas, tile):
    from pmdas_amd import hip
    W, H = 96, 64
    g, sc, target = _grid(mdas, 5, W, H, 16, 3, 4)
    g.set_recon_tile(*tile)
    g.Run(sc, target)
    eager = g.integrate_image(W, H)  # sizes the buffers
    o = pyoracle.OracleGrid(g.dim, g.GetMaxSamples(), 4, 1, 2, 0.25, float(W), float(H))
    o.run(sc.params_bytes(), target)
    assert np.array_equal(eager.view(np.uint32), o.integrate(W, H).view(np.uint32))
    st = hip.Stream()
    buf = hip.DeviceBuffer(W * H * 16)
    g.stream = st
    graph = hip.Graph(st)
    graph.begin()
    g.Integrate(buf.ptr, W, H)  # fails if it synchronises or allocates
    graph.end()
    g.stream = None
    junk = np.full(W * H * 4, np.nan, np.float32)
    for r in range(3):
        buf.upload(junk)  # eager work between replays
        graph.launch()
        st.synchronize()
        got = buf.download(np.float32, W * H * 4).reshape(H, W, 4)
        assert np.array_equal(got.view(np.uint32), eager.view(np.uint32)), f"replay {r}"
        again = g.integrate_image(W, H)  # an eager call between replays
        assert np.array_equal(again.view(np.uint32), eager.view(np.uint32))
    keys = g.recon_keys()
    assert np.array_equal(keys, o.recon_keys(W, H, *tile))
    graph.destroy()
    st.destroy()
    buf.free()
    for x in (g, sc, o):
        x.close()


def test_batch_integrate_graph_equals_each_image(mdas):
    from pmdas_amd import hip
    specs = [(3, 80, 48, 8, 1, 2), (4, 64, 64, 8, 2, 3), (5, 120, 72, 12, 3, 5), (5, 40, 40, 16, 3, 6)]
    grids, scenes, targets, sizes = [], [], [], []
    for dim, W, H, spp, kind, seed in specs:
        g, sc, t = _grid(mdas, dim, W, H, spp, kind, seed)
        grids.append(g)
        scenes.append(sc)
        targets.append(t)
        sizes.append((W, H))
    grids[1].set_recon_tile(4, 4)
    b = mdas.Batch(grids, scenes)
    bufs = [hip.DeviceBuffer(W * H * 16) for W, H in sizes]
    Ws, Hs = [s[0] for s in sizes], [s[1] for s in sizes]
    for rnd in range(2):
        b.run(targets)
        want = [g.integrate_image(W, H) for g, (W, H) in zip(grids, sizes)]
        for use_graph in (True, True, False):
            for buf, (W, H) in zip(bufs, sizes):
                buf.upload(np.full(W * H * 4, np.nan, np.float32))
            b.integrate(bufs, Ws, Hs, use_graph=use_graph)
            hip.synchronize()
            for i, (buf, (W, H)) in enumerate(zip(bufs, sizes)):
                got = buf.download(np.float32, W * H * 4).reshape(H, W, 4)
                assert np.array_equal(got.view(np.uint32), want[i].view(np.uint32)), (rnd, use_graph, i)
    b.close()
    for x in bufs + grids + scenes:
        x.close() if hasattr(x, "close") else x.free()


def test_batch_graph_keeps_null_stream_order(mdas):
    from pmdas_amd import hip
    specs = [(5, 192, 108, 16, 3, 7), (4, 160, 96, 8, 2, 8)]
    grids, scenes, targets, sizes = [], [], [], []
    for dim, W, H, spp, kind, seed in specs:
        g, sc, t = _grid(mdas, dim, W, H, spp, kind, seed)
        grids.append(g)
        scenes.append(sc)
        targets.append(t)
        sizes.append((W, H))
    b = mdas.Batch(grids, scenes)
    b.run(targets)  # null stream, graphs on the batch's stream
    want = [g.integrate_image(W, H) for g, (W, H) in zip(grids, sizes)]
    bufs = [hip.DeviceBuffer(W * H * 16) for W, H in sizes]
    Ws, Hs = [s[0] for s in sizes], [s[1] for s in sizes]
    for rep in range(3):
        # NaN bytes written on the null stream just before the call: the graph
        # must run after them
        for buf, (W, H) in zip(bufs, sizes):
            assert hip.lib().hipMemsetAsync(ctypes.c_void_p(buf.ptr), ctypes.c_int(0xFF),
                                            ctypes.c_size_t(W * H * 16), None) == 0
        b.integrate(bufs, Ws, Hs, use_graph=True)
        # no synchronisation: the download (hipMemcpy on the null stream)
        # must wait for the graph
        for i, (buf, (W, H)) in enumerate(zip(bufs, sizes)):
            got = buf.download(np.float32, W * H * 4).reshape(H, W, 4)
            assert np.array_equal(got.view(np.uint32), want[i].view(np.uint32)), (rep, i)
    b.close()
    for x in bufs + grids + scenes:
        x.close() if hasattr(x, "close") else x.free()
