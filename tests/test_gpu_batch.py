"""Batched runs (one launch per kernel and pass for all images of a
dimension, replayed from a captured hipGraph): every image must equal the
same image run alone, bit for bit, and the oracle."""
import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.gpu

# kind, dim, b, e, W, H, spp, alpha, cand, seed
IMAGES = [
    (3, 5, 1, 2, 64, 48, 12, 0.25, 4, 1),
    (3, 5, 1, 3, 96, 64, 10, 0.25, 4, 2),
    (3, 5, 0, 4, 40, 40, 16, 1 / 32, 4, 3),
    (3, 5, 1, 2, 128, 72, 8, 0.25, 2, 4),
    (2, 4, 1, 3, 80, 60, 12, 0.25, 4, 5),
    (4, 6, 1, 1, 48, 32, 9, 1 / 64, 4, 6),
]


def _make(kdgrid, spec):
    kind, dim, b, e, W, H, spp, alpha, cand, seed = spec
    target = int(spp * W * H)
    cap = int(target * 1.6) + (4 << (b * dim + 2 * e))
    g = kdgrid.KDGrid(dim, cap, cand, b, e, alpha, float(W), float(H))
    sc = kdgrid.Scene(kind, seed, dim, float(W), float(H))
    return g, sc, target, (dim, cap, cand, b, e, alpha, float(W), float(H))


@pytest.fixture(scope="module")
def kd():
    import pmdas_amd
    from pmdas_amd import kdgrid
    pmdas_amd.load()
    return kdgrid


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_single_runs(kd, use_graph):
    made = [_make(kd, s) for s in IMAGES]
    batch = kd.Batch([m[0] for m in made], [m[1] for m in made])
    targets = [m[2] for m in made]
    for replay in range(3):  # graph replays with eager work in between
        print(f"batch replay {replay} graph={use_graph}", flush=True)
        batch.run(targets, use_graph=use_graph)
        states = [m[0].download_state() for m in made]
        for (g, sc, target, args), st in zip(made, states):
            solo = kd.KDGrid(*args)
            solo.Run(sc, target)
            ref = solo.download_state()
            for k in st:
                assert np.array_equal(st[k], ref[k]), (replay, k, args)
            assert g.GetNumberOfSamples() >= target
            solo.close()
        # eager work on one of the grids between replays
        made[0][0].Run(made[0][1], targets[0] // 2)
    info = batch.info()
    assert info["groups"] == 3 and info["pass_blocks"] > 0
    batch.close()


def test_batch_matches_oracle(kd):
    made = [_make(kd, s) for s in IMAGES[:3]]
    batch = kd.Batch([m[0] for m in made], [m[1] for m in made])
    batch.run([m[2] for m in made], use_graph=True)
    for g, sc, target, args in made:
        o = pyoracle.OracleGrid(*args)
        o.run(sc.params_bytes(), target)
        st, ref = g.download_state(), o.state()
        for k in st:
            assert np.array_equal(st[k], ref[k]), k
    batch.close()


def test_batch_rejects_mismatched_scene(kd):
    from pmdas_amd import MdasError
    g, _, _, _ = _make(kd, IMAGES[0])
    sc4 = kd.Scene(2, 0, 4, 64.0, 48.0)
    with pytest.raises(MdasError):
        kd.Batch([g], [sc4])


def test_batch_learned_equals_single_runs(kd):
    """Learned members (MLP scoring fused into the batched error pass) next to
    a plain member of the same dimension: every image equals its solo run with
    the same mode, bit for bit, with graph replay."""
    from oracle import mlp_oracle as mo
    r = np.random.default_rng(9)
    p = (r.standard_normal(mo.N_FLOATS) * 0.15).astype(np.float32)
    p[mo.OFF_MU:mo.OFF_ISD] = 0.0
    p[mo.OFF_ISD:] = 0.2
    net = kd.Mlp(0)
    net.set_params(p)
    specs = IMAGES[:3] + [IMAGES[5]]
    learned = [True, False, True, True]
    made = [_make(kd, s) for s in specs]
    for (g, *_), on in zip(made, learned):
        if on:
            g.set_learned(2 | 8, net)
            g.set_learned_exponent(0.5)
    batch = kd.Batch([m[0] for m in made], [m[1] for m in made])
    targets = [m[2] for m in made]
    for replay in range(2):
        print(f"learned batch replay {replay}", flush=True)
        batch.run(targets, use_graph=True)
        for (g, sc, target, args), on in zip(made, learned):
            st = g.download_state()
            solo = kd.KDGrid(*args)
            if on:
                solo.set_learned(2 | 8, net)
                solo.set_learned_exponent(0.5)
            solo.Run(sc, target)
            ref = solo.download_state()
            for k in st:
                assert np.array_equal(st[k], ref[k]), (replay, k, args, on)
            if on:  # leaves split by the last pass have no features yet (err < 0)
                live = st["err"] >= 0
                fa, ea = g.leaf_features()
                fb, eb = solo.leaf_features()
                assert np.array_equal(fa[live], fb[live]) and np.array_equal(ea[live], eb[live])
            solo.close()
    batch.close()
    net.close()


def test_batch_autotune_keeps_results(kd):
    """The autotuner's candidate launch shapes all give the solo results; the
    chosen shape is one of the candidates."""
    made = [_make(kd, s) for s in IMAGES[:3]]
    batch = kd.Batch([m[0] for m in made], [m[1] for m in made])
    targets = [m[2] for m in made]
    shape = batch.autotune(targets)
    assert shape in (2048, 4096, 8192, 16384, 32768)
    batch.run(targets, use_graph=True)
    for g, sc, target, args in made:
        solo = kd.KDGrid(*args)
        solo.Run(sc, target)
        st, ref = g.download_state(), solo.download_state()
        for k in st:
            assert np.array_equal(st[k], ref[k]), k
        solo.close()
    batch.close()


def test_random_batch_matches_oracle(kd):
    """24 seeded random small images of every dimension and scene (the random
    parity matrix's generator, another seed) in one batch: each equals the
    oracle's run of it alone, bit for bit."""
    from test_gpu_random_configs import _configs
    made = []
    for c in _configs(24, 77):
        W, H, dim = c["W"], c["H"], c["dim"]
        target = max(1, int(c["spp"] * W * H))
        cap = int(target * 1.6) + (4 << (c["b"] * dim + 2 * c["e"])) + 64
        args = (dim, cap, c["cand"], c["b"], c["e"], c["alpha"], float(W), float(H))
        made.append((kd.KDGrid(*args), kd.Scene(c["kind"], c["seed"], dim, float(W), float(H)),
                     target, args))
    batch = kd.Batch([m[0] for m in made], [m[1] for m in made])
    batch.run([m[2] for m in made], use_graph=True)
    for g, sc, target, args in made:
        o = pyoracle.OracleGrid(*args)
        o.run(sc.params_bytes(), target)
        st, ref = g.download_state(), o.state()
        for k in st:
            assert np.array_equal(st[k], ref[k]), (k, args)
        o.close()
    batch.close()
    for g, sc, _, _ in made:
        g.close()
        sc.close()
