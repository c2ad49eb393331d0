"""Call sequences and misuse through the C ABI: calls before InitialSampling,
mismatched scenes, out-of-range images and ranges, invalid tiles and launch
shapes are errors that leave the grid usable; a device-side run followed by
standalone SamplingPass calls, adaptive passes without a ComputeErrors in
between, and repeated runs, equal the oracle doing the same
(kdgrid.cu:633-779 called in any order)."""
import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.gpu

# kind, dim, b, e, W, H, spp, alpha, cand, seed
CASE = (3, 5, 1, 2, 48, 36, 10, 0.25, 4, 11)


@pytest.fixture(scope="module")
def mdas():
    import pmdas_amd
    from pmdas_amd import kdgrid, hip
    pmdas_amd.load()
    assert hip.device_count() >= 1, "no GPU visible"
    return kdgrid


def _args(c=CASE, cap_scale=3.0):
    kind, dim, b, e, W, H, spp, alpha, cand, seed = c
    target = spp * W * H
    cap = int(target * cap_scale) + (4 << (b * dim + 2 * e))
    return (dim, cap, cand, b, e, alpha, float(W), float(H)), target


def _same(gs, os_, what):
    for k in gs:
        a, r = gs[k], os_[k]
        eq = a.view(np.uint32) == r.view(np.uint32) if a.dtype == np.float32 else a == r
        assert eq.all(), (what, k, np.argwhere(~eq)[:3].tolist())


def test_calls_before_initial_sampling_are_errors(mdas):
    from pmdas_amd import MdasError, hip
    args, _ = _args()
    g = mdas.KDGrid(*args)
    for call in (g.ComputeErrors, g.AdaptiveSampling, g.SamplingPass):
        with pytest.raises(MdasError):
            call()
    buf = hip.DeviceBuffer(16 * 8 * 8)
    with pytest.raises(MdasError):
        g.Integrate(buf.ptr, 8, 8)
    buf.free()
    g.InitialSampling()  # the grid is still usable
    assert g.GetNumberOfSamples() == 4 << (CASE[2] * CASE[1] + 2 * CASE[3])
    g.close()


def test_invalid_arguments_leave_the_grid_usable(mdas):
    from pmdas_amd import MdasError, hip
    kind, dim, b, e, W, H, spp, alpha, cand, seed = CASE
    args, target = _args()
    g = mdas.KDGrid(*args)
    sc4 = mdas.Scene(2, seed, 4, float(W), float(H))  # a 4D scene for a 5D grid
    with pytest.raises(MdasError):
        g.Run(sc4, target)
    sc = mdas.Scene(kind, seed, dim, float(W), float(H))
    g.InitialSampling()
    with pytest.raises(MdasError):
        g.EvalIntegrand(sc4, 0, 16)
    with pytest.raises(MdasError):
        g.EvalIntegrand(sc, args[1] - 8, 16)  # past the sample capacity
    buf = hip.DeviceBuffer(16 * 8 * 8)
    for w, h in ((0, 8), (8, -1), (70000, 1)):
        with pytest.raises(MdasError):
            g.Integrate(buf.ptr, w, h)
    buf.free()
    for tw, th in ((3, 5), (0, 16), (32, 4)):
        with pytest.raises(MdasError):
            g.set_recon_tile(tw, th)
    for cb, eb in ((1, 1024), (4096, 0), (1 << 20, 1024)):
        with pytest.raises(MdasError):
            g.set_launch_shape(cb, eb)
    # and the run still equals the oracle
    passes = g.Run(sc, target)
    o = pyoracle.OracleGrid(*args)
    assert passes == o.run(sc.params_bytes(), target)
    _same(g.download_state(), o.state(), "after errors")
    g.close()
    o.close()
    sc.close()
    sc4.close()


def test_run_then_standalone_passes(mdas):
    """The device-side loop leaves the last pass's children marked stale;
    standalone SamplingPass calls continue from there like the reference."""
    kind, dim, b, e, W, H, spp, alpha, cand, seed = CASE
    args, target = _args()
    sc = mdas.Scene(kind, seed, dim, float(W), float(H))
    params = sc.params_bytes()
    g = mdas.KDGrid(*args)
    o = pyoracle.OracleGrid(*args)
    assert g.Run(sc, target) == o.run(params, target)
    _same(g.download_state(), o.state(), "after run")
    for p in range(3):
        g.SamplingPass()
        g.EvalNewSamples(sc)
        o.compute_errors()
        assert o.adaptive_sampling() == 0
        o.eval_scene(params, o.n_samples - o.new_samples, o.new_samples)
        assert g.GetNewSamples() == o.new_samples
        _same(g.download_state(), o.state(), f"pass {p} after run")
    assert g.Validate()
    g.close()
    o.close()
    sc.close()


def test_adaptive_sampling_without_compute_errors(mdas):
    """AdaptiveSampling keeps using g_error from the last ComputeErrors
    (kdgrid.cu:684-685 resets it only there) and sees the last pass's
    children still stale: two adaptive passes in a row, after a standalone
    ComputeErrors and after a device-side run, equal the oracle doing the
    same, and so does the maximum error the grid reports."""
    kind, dim, b, e, W, H, spp, alpha, cand, seed = CASE
    args, target = _args()
    sc = mdas.Scene(kind, seed, dim, float(W), float(H))
    params = sc.params_bytes()
    for after_run in (False, True):
        g = mdas.KDGrid(*args)
        o = pyoracle.OracleGrid(*args)
        if after_run:
            assert g.Run(sc, target // 2) == o.run(params, target // 2)
        else:
            g.InitialSampling()
            g.EvalNewSamples(sc)
            o.initial_sampling()
            o.eval_scene(params, 0, o.n_samples)
            g.ComputeErrors()
            o.compute_errors()
        assert g.GetMaxError() == o.max_error, after_run
        for p in range(2):
            g.AdaptiveSampling()
            g.EvalNewSamples(sc)
            assert o.adaptive_sampling() == 0
            o.eval_scene(params, o.n_samples - o.new_samples, o.new_samples)
            assert g.GetNewSamples() == o.new_samples, (after_run, p)
            assert g.GetMaxError() == o.max_error, (after_run, p)
            _same(g.download_state(), o.state(), f"adaptive pass {p}, after run {after_run}")
        g.close()
        o.close()
    sc.close()


def test_repeated_runs(mdas):
    """A run starts from the initial grid: a second run, and a smaller run
    after a larger one, equal fresh runs."""
    kind, dim, b, e, W, H, spp, alpha, cand, seed = CASE
    args, target = _args()
    sc = mdas.Scene(kind, seed, dim, float(W), float(H))
    g = mdas.KDGrid(*args)
    g.Run(sc, target)
    first = g.download_state()
    g.Run(sc, target)
    _same(g.download_state(), first, "second run")
    small = target // 3
    g.Run(sc, small)
    fresh = mdas.KDGrid(*args)
    fresh.Run(sc, small)
    _same(g.download_state(), fresh.download_state(), "smaller run")
    fresh.close()
    g.close()
    sc.close()


def test_run_of_many_passes(mdas):
    """A run that needs 21287 passes (alpha 1 on the disk scene from a single
    initial leaf selects a handful of leaves per pass) goes to its target
    like the oracle's loop, which has no pass limit; the same image in a
    batch too (every pass adds the sample of the max-error leaf, so the run
    always ends)."""
    target = 200000
    args = (3, int(target * 1.6) + 68, 4, 0, 0, 1.0, 1.0, 1.0)
    sc = mdas.Scene(1, 3, 3, 1.0, 1.0)
    g = mdas.KDGrid(*args)
    o = pyoracle.OracleGrid(*args)
    passes = g.Run(sc, target)
    assert passes == o.run(sc.params_bytes(), target) == 21287
    _same(g.download_state(), o.state(), "many passes")
    ref = g.download_state()
    g2 = mdas.KDGrid(*args)
    b = mdas.Batch([g2], [sc])
    b.run([target])
    _same(g2.download_state(), ref, "many passes, batch")
    b.close()
    g2.close()
    g.close()
    o.close()
    sc.close()
