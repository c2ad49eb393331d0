"""GPU parity: libmdas.so (HIP, gfx950) against the CPU oracle on the same
seeded inputs.  Bar: bit-exact for coordinates, slots, boxes, errors, seeds,
counters, reconstruction keys, pixels and bytes (the oracle fixes leaf-order
allocation and summation, and IEEE-754 FP32 without contraction); the
reference-style atomic scatter is compared within 1e-5 relative.
"""
import sys

import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mdas():
    import pmdas_amd
    from pmdas_amd import kdgrid, hip
    pmdas_amd.load()
    assert hip.device_count() >= 1, "no GPU visible"
    return kdgrid


def _pair(mdas, dim, max_samples, cand, b, e, alpha, sx, sy):
    g = mdas.KDGrid(dim, max_samples, cand, b, e, alpha, sx, sy)
    o = pyoracle.OracleGrid(dim, max_samples, cand, b, e, alpha, sx, sy)
    return g, o


def _same_state(gs, os_, what=""):
    for k in ("coords", "values", "idx", "box", "err", "seed"):
        a, b = gs[k], os_[k]
        assert a.shape == b.shape, (what, k, a.shape, b.shape)
        if a.dtype == np.float32:
            eq = a.view(np.uint32) == b.view(np.uint32)
        else:
            eq = a == b
        if not eq.all():
            bad = np.argwhere(~eq)[:5]
            raise AssertionError(f"{what}: {k} differs at {bad.tolist()}: "
                                 f"{a[tuple(bad[0])]} vs {b[tuple(bad[0])]}")


SMALL = [  # dim, b, e, sx, sy
    (3, 1, 3, 64.0, 48.0),
    (4, 1, 2, 40.0, 30.0),
    (5, 1, 2, 48.0, 27.0),
    (6, 1, 1, 32.0, 18.0),
    (3, 0, 5, 32.0, 32.0),
]


@pytest.mark.parametrize("dim,b,e,sx,sy", SMALL)
def test_initial_sampling_bit_exact(mdas, dim, b, e, sx, sy):
    _say(f"initial dim={dim} b={b} e={e}")
    g, o = _pair(mdas, dim, 1 << 18, 4, b, e, 0.25, sx, sy)
    g.InitialSampling()
    o.initial_sampling()
    assert g.GetNumberOfSamples() == o.n_samples and g.GetNumberOfNodes() == o.n_leaves
    gs, os_ = g.download_state(), o.state()
    os_["values"] = gs["values"]  # not produced yet
    _same_state(gs, os_, "initial")
    assert g.Validate()


SCENES = [(1, 3), (2, 4), (3, 5), (4, 6), (0, 4), (5, 3)]


@pytest.mark.parametrize("kind,dim", SCENES)
def test_integrand_bit_exact(mdas, kind, dim):
    _say(f"integrand kind={kind} dim={dim}")
    sx, sy = 96.0, 54.0
    g = mdas.KDGrid(dim, 1 << 19, 4, 1, 3, 0.25, sx, sy)
    sc = mdas.Scene(kind, 17, dim, sx, sy)
    g.InitialSampling()
    g.EvalIntegrand(sc, 0, g.GetNumberOfSamples())
    st = g.download_state()
    want = pyoracle.eval_scene(sc.params_bytes(), st["coords"], dim)
    assert np.array_equal(st["values"].view(np.uint32), want.view(np.uint32))
    assert np.isfinite(want).all() and (want >= 0).all()


RUNS = [  # kind, dim, b, e, W, H, spp, alpha, cand
    (1, 3, 0, 5, 64, 64, 8, 0.25, 4),
    (2, 4, 1, 3, 96, 64, 12, 0.25, 4),
    (3, 5, 1, 2, 80, 48, 16, 1 / 32, 4),
    (4, 6, 1, 2, 64, 40, 10, 1 / 64, 2),
    (3, 5, 0, 4, 64, 32, 9, 0.0, 1),
]


def _make(mdas, kind, dim, b, e, W, H, spp, alpha, cand, seed=5):
    target = int(spp * W * H)
    cap = int(target * 1.6) + (4 << (b * dim + 2 * e))
    g, o = _pair(mdas, dim, cap, cand, b, e, alpha, float(W), float(H))
    sc = mdas.Scene(kind, seed, dim, float(W), float(H))
    return g, o, sc, target


@pytest.mark.parametrize("kind,dim,b,e,W,H,spp,alpha,cand", RUNS)
def test_pass_by_pass_bit_exact(mdas, kind, dim, b, e, W, H, spp, alpha, cand):
    """The reference's call sequence (optixMotionBlur.cpp:768-791) through the
    C ABI, one entry point at a time, against the oracle after every call."""
    g, o, sc, target = _make(mdas, kind, dim, b, e, W, H, spp, alpha, cand)
    params = sc.params_bytes()
    g.InitialSampling()
    g.EvalIntegrand(sc, 0, g.GetNumberOfSamples())
    o.initial_sampling()
    o.eval_scene(params, 0, o.n_samples)
    p = 0
    while g.GetNumberOfSamples() < target:
        _say(f"pass {p} kind={kind} dim={dim} N={g.GetNumberOfSamples()}")
        g.ComputeErrors()
        o.compute_errors()
        assert np.float32(g.GetMaxError()) == np.float32(o.max_error)
        g.AdaptiveSampling()
        assert o.adaptive_sampling() == 0
        assert g.GetNewSamples() == o.new_samples and g.GetNewNodes() == o.new_nodes
        g.EvalNewSamples(sc)
        o.eval_scene(params, o.n_samples - o.new_samples, o.new_samples)
        _same_state(g.download_state(), o.state(), f"pass {p}")
        p += 1
        assert p < 200
    assert p > 0
    assert g.Validate()


@pytest.mark.parametrize("kind,dim,b,e,W,H,spp,alpha,cand", RUNS)
def test_run_loop_bit_exact(mdas, kind, dim, b, e, W, H, spp, alpha, cand):
    """mdas_run (device-side stopping rule) == the oracle's sampling loop."""
    _say(f"run kind={kind} dim={dim} {W}x{H} spp={spp}")
    g, o, sc, target = _make(mdas, kind, dim, b, e, W, H, spp, alpha, cand)
    passes = g.Run(sc, target)
    op = o.run(sc.params_bytes(), target)
    assert passes == op
    assert g.GetNumberOfSamples() == o.n_samples >= target
    _same_state(g.download_state(), o.state(), "run")
    stats = g.pass_stats()
    assert len(stats) == passes
    assert stats[-1, 1] + stats[-1, 2] == o.n_samples


@pytest.mark.parametrize("kind,dim,b,e,W,H,spp,alpha,cand", RUNS[:4])
@pytest.mark.parametrize("tile", [(8, 8), (16, 4), (4, 4), (8, 2)])
def test_integrate_bit_exact(mdas, kind, dim, b, e, W, H, spp, alpha, cand, tile):
    _say(f"integrate kind={kind} dim={dim} tile={tile}")
    g, o, sc, target = _make(mdas, kind, dim, b, e, W, H, spp, alpha, cand)
    g.Run(sc, target)
    o.run(sc.params_bytes(), target)
    g.set_recon_tile(*tile)
    px = g.integrate_image(W, H)
    want = o.integrate(W, H)
    assert np.array_equal(px.view(np.uint32), want.view(np.uint32))
    keys = g.recon_keys()
    assert np.array_equal(keys, o.recon_keys(W, H, *tile))
    sc_px = g.integrate_image(W, H, scatter=True)
    assert np.allclose(sc_px, want, rtol=1e-5, atol=1e-7)


def test_integrate_other_resolution(mdas):
    """Pixel grid finer and coarser than the sample space (scaleFactor != 1)."""
    g, o, sc, target = _make(mdas, 2, 4, 1, 2, 40, 30, 8, 0.25, 4)
    g.Run(sc, target)
    o.run(sc.params_bytes(), target)
    for W, H in [(80, 60), (20, 15), (37, 23), (1, 1), (7, 3), (1, 45), (300, 2)]:
        _say(f"integrate {W}x{H}")
        px = g.integrate_image(W, H)
        assert np.array_equal(px.view(np.uint32), o.integrate(W, H).view(np.uint32))
        for tile in ((8, 8), (4, 4), (16, 1)):
            g.set_recon_tile(*tile)
            assert np.array_equal(g.integrate_image(W, H).view(np.uint32),
                                  px.view(np.uint32)), (W, H, tile)
        g.set_recon_tile(8, 8)


def test_autotune_keeps_result(mdas):
    from pmdas_amd import hip
    g, o, sc, target = _make(mdas, 3, 5, 1, 2, 80, 48, 16, 0.25, 4)
    g.Run(sc, target)
    o.run(sc.params_bytes(), target)
    buf = hip.DeviceBuffer(80 * 48 * 16)
    tw, th = g.autotune_recon(buf.ptr, 80, 48)
    assert tw * th in (16, 64)
    buf.free()
    px = g.integrate_image(80, 48)
    assert np.array_equal(px.view(np.uint32), o.integrate(80, 48).view(np.uint32))


def test_convert_to_bytes_bit_exact(mdas):
    from pmdas_amd import hip
    rng = np.random.default_rng(1)
    n = 1 << 16
    t = np.float32(0.0031308)  # the sRGB toe's threshold and its neighbours
    special = np.array([np.nan, np.inf, -np.inf, -0.0, 0.0, 1e-45, 1.17e-38, 1.0, 1.0000001,
                        np.nextafter(t, np.float32(0)), t, np.nextafter(t, np.float32(1)),
                        0.9999999, 255.5 / 256], np.float32)
    special = np.resize(special, (len(special) + 3) // 4 * 4).reshape(-1, 4)
    px = np.concatenate([rng.random((n, 4), dtype=np.float32) * 1.2 - 0.1,
                         np.linspace(0, 1, 4096 * 4, dtype=np.float32).reshape(-1, 4),
                         special, special[:, ::-1].copy()])
    src = hip.DeviceBuffer(px.nbytes)
    dst = hip.DeviceBuffer(len(px) * 4)
    src.upload(px)
    g = mdas.KDGrid(3, 1 << 12, 4, 0, 2, 0.25, 4.0, 4.0)
    g.ConvertToBytes(src.ptr, dst.ptr, len(px), 1)
    hip.synchronize()
    out = dst.download(np.uint8, len(px) * 4).reshape(-1, 4)
    assert np.array_equal(out, pyoracle.convert_to_bytes(px))


def test_sampling_density_matches(mdas):
    g, o, sc, target = _make(mdas, 1, 3, 0, 4, 48, 32, 6, 0.25, 4)
    g.Run(sc, target)
    o.run(sc.params_bytes(), target)
    den = np.zeros(48 * 32 * 4, np.float32)
    g.SamplingDensity(den, 48, 32)
    assert np.array_equal(den.reshape(32, 48, 4), o.sampling_density(48, 32))


def test_capacity_error_is_reported(mdas):
    from pmdas_amd import MdasError
    g = mdas.KDGrid(3, 300, 4, 0, 3, 0.0, 8.0, 8.0)  # 256 initial samples, 64 leaves
    sc = mdas.Scene(0, 0, 3, 8.0, 8.0)
    g.InitialSampling()
    g.EvalIntegrand(sc, 0, 256)
    g.ComputeErrors()
    before = g.download_state()
    with pytest.raises(MdasError) as ei:
        g.AdaptiveSampling()
    assert ei.value.code == -2
    # the refused pass changed nothing, LCG states included, and the state is
    # the oracle's after its own refused pass
    after = g.download_state()
    assert set(after) == {"coords", "values", "idx", "box", "err", "seed"}
    for k in after:
        assert np.array_equal(before[k], after[k]), k
    o = pyoracle.OracleGrid(3, 300, 4, 0, 3, 0.0, 8.0, 8.0)
    o.initial_sampling()
    o.eval_scene(sc.params_bytes(), 0, 256)
    o.compute_errors()
    assert o.adaptive_sampling() != 0
    for k, v in o.state().items():
        assert np.array_equal(after[k], v), k
    # a second attempt draws the same numbers and is refused the same way
    with pytest.raises(MdasError):
        g.AdaptiveSampling()
    again = g.download_state()
    for k in after:
        assert np.array_equal(after[k], again[k]), k
    o.close()
    # the device-side loop never starts a pass that could overflow (N + L =
    # 320 > 300): it ends cleanly with the initial grid
    assert g.Run(sc, 10**6) == 0 and g.GetNumberOfSamples() == 256


def test_refused_pass_mid_run_then_retry(mdas):
    """A pass refused for capacity after a successful one: the state is the
    oracle's after its own refused pass; ComputeErrors again (the retry a
    SamplingPass makes) gives the oracle's errors and maximum, and the pass
    is refused again with nothing changed."""
    from pmdas_amd import MdasError
    args = (3, 400, 4, 0, 3, 0.0, 8.0, 8.0)  # 64 leaves; pass 1 -> 320 samples, pass 2 -> 448
    g = mdas.KDGrid(*args)
    o = pyoracle.OracleGrid(*args)
    sc = mdas.Scene(0, 5, 3, 8.0, 8.0)
    params = sc.params_bytes()
    g.InitialSampling()
    o.initial_sampling()
    g.EvalIntegrand(sc, 0, 256)
    o.eval_scene(params, 0, 256)
    g.SamplingPass()
    assert o.sampling_pass() == 0
    n = g.GetNumberOfSamples()
    assert n == o.n_samples == 320
    g.EvalIntegrand(sc, 256, n - 256)
    o.eval_scene(params, 256, n - 256)
    for attempt in range(2):
        g.ComputeErrors()
        o.compute_errors()
        assert g.GetMaxError() == o.max_error, attempt
        with pytest.raises(MdasError) as ei:
            g.AdaptiveSampling()
        assert ei.value.code == -2
        assert o.adaptive_sampling() != 0
        gs, os_ = g.download_state(), o.state()
        for k in os_:
            assert np.array_equal(gs[k], os_[k]), (attempt, k)
    g.close()
    o.close()
    sc.close()


@pytest.mark.parametrize("kind,dim", [(1, 3), (2, 4), (3, 5), (4, 6)])
def test_upload_state_then_pass(mdas, kind, dim):
    """An uploaded mid-run state reads back as uploaded (slots, wherever the
    line layout keeps them, with the reference's empty markers), and a pass
    on it matches the oracle on the same state."""
    g, o, sc, target = _make(mdas, kind, dim, 1, 1 if dim >= 5 else 2, 48, 32, 8, 0.25, 4)
    o.run(sc.params_bytes(), target // 2)
    st = o.state()
    g.upload_state(st["coords"], st["values"], st["idx"], st["box"], st["err"], st["seed"])
    _same_state(g.download_state(), st, "uploaded")
    g.ComputeErrors()
    o.compute_errors()
    g.AdaptiveSampling()
    o.adaptive_sampling()
    gs, os_ = g.download_state(), o.state()
    n_old = o.n_samples - o.new_samples
    gs["values"][n_old:] = 0
    os_["values"][n_old:] = 0
    _same_state(gs, os_, "uploaded pass")


def test_full_size_4d_dof_properties(mdas):
    """1920x1080 4D DOF, 16 spp: size-independent invariants on the GPU result."""
    W, H = 1920, 1080
    target = 16 * W * H
    g = mdas.KDGrid(4, int(target * 1.5) + (4 << 20), 4, 1, 8, 0.25, float(W), float(H))
    sc = mdas.Scene(2, 1, 4, float(W), float(H))
    _say("full-size run")
    passes = g.Run(sc, target)
    N, L = g.GetNumberOfSamples(), g.GetNumberOfNodes()
    assert N >= target and 2 * L <= N and passes > 0
    stats = g.pass_stats()
    assert stats[0, 0] == 1 << 20 and stats[0, 1] == 4 << 20
    for p in range(1, len(stats)):  # counters chain: N' = N + sel, L' = L + split
        assert stats[p, 1] == stats[p - 1, 1] + stats[p - 1, 2]
        assert stats[p, 0] == stats[p - 1, 0] + stats[p - 1, 3]
    _say("validate")
    assert g.Validate()
    px = g.integrate_image(W, H)
    assert np.isfinite(px).all() and (px[..., :3] >= 0).all()
    assert (px[..., 3] == 0).all()


@pytest.mark.parametrize("kind,dim,fp64", [(5, 3, False), (3, 5, False), (4, 6, False),
                                           (2, 4, True), (3, 5, True)])
def test_render_qmc_bit_exact(mdas, kind, dim, fp64):
    _say(f"render_qmc kind={kind} dim={dim} fp64={fp64}")
    W, H, sx, sy = 40, 24, 40.0, 24.0
    sc = mdas.Scene(kind, 11, dim, sx, sy)
    got = sc.render_qmc(W, H, 12, first_sample=3, frame=2, fp64=fp64)
    want = pyoracle.qmc_image(sc.params_bytes(), W, H, 12, 3, 2, fp64=fp64)
    u = np.uint64 if fp64 else np.uint32
    assert np.array_equal(got.view(u), want.view(u))


def _log_values(path, key):
    """All values logged under `key` (the KEY / value line pairs the
    reference's parse_logs.get_values reads)."""
    lines = open(path).read().split("\n")
    return [float(lines[i + 1]) for i in range(len(lines) - 1) if lines[i] == key]


def test_log_keys_per_pass(mdas, tmp_path):
    """mdas_run + mdas_integrate with a log: the reference's keys, one
    COMPUTE ERRORS / SAMPLES / ADAPTIVE SAMPLING / TRACE entry per pass."""
    log = str(tmp_path / "run.log")
    W, H, spp = 64, 48, 12
    target = spp * W * H
    g = mdas.KDGrid(5, int(target * 1.6) + (4 << 9), 4, 1, 2, 0.25, float(W), float(H), log=log)
    sc = mdas.Scene(3, 2, 5, float(W), float(H))
    passes = g.Run(sc, target)
    g.integrate_image(W, H)
    assert _log_values(log, "INITIAL SAMPLES") == [4 << 9]
    assert len(_log_values(log, "INITIAL SAMPLING TIME")) == 1
    for key in ("COMPUTE ERRORS TIME", "SAMPLES", "ADAPTIVE SAMPLING TIME"):
        assert len(_log_values(log, key)) == passes, key
    assert len(_log_values(log, "TRACE TIME")) == passes + 1
    assert sum(_log_values(log, "SAMPLES")) == g.GetNumberOfSamples() - (4 << 9)
    assert _log_values(log, "TOTAL SAMPLES") == [g.GetNumberOfSamples()]
    times = sum((_log_values(log, k) for k in ("COMPUTE ERRORS TIME", "ADAPTIVE SAMPLING TIME",
                                                "TRACE TIME", "INTEGRATE TIME")), [])
    assert all(t >= 0 for t in times) and sum(times) > 0


def _random_runs():
    from test_gpu_random_configs import _configs
    return _configs(16, 31)


@pytest.mark.parametrize("c", _random_runs(), ids=lambda c: f"rnd{c['id']}-d{c['dim']}")
def test_pass_by_pass_random_configs(mdas, c):
    """The one-call-at-a-time sequence (ComputeErrors, AdaptiveSampling, the
    integrand on the new slice) over seeded random configurations, against
    the oracle after every pass, for up to 60 passes."""
    g, o, sc, target = _make(mdas, c["kind"], c["dim"], c["b"], c["e"], c["W"], c["H"],
                             c["spp"], c["alpha"], c["cand"], seed=c["seed"])
    print(c, flush=True)
    params = sc.params_bytes()
    g.InitialSampling()
    g.EvalIntegrand(sc, 0, g.GetNumberOfSamples())
    o.initial_sampling()
    o.eval_scene(params, 0, o.n_samples)
    p = 0
    while g.GetNumberOfSamples() < target and p < 60:
        g.ComputeErrors()
        o.compute_errors()
        assert np.float32(g.GetMaxError()) == np.float32(o.max_error)
        g.AdaptiveSampling()
        assert o.adaptive_sampling() == 0
        g.EvalNewSamples(sc)
        o.eval_scene(params, o.n_samples - o.new_samples, o.new_samples)
        _same_state(g.download_state(), o.state(), f"pass {p}")
        p += 1
    assert g.Validate()
