"""GPU tests of the learned proposal (MLP on MFMA, training kernels, the
fused error pass) against the CPU oracle (oracle/mlp_oracle.py).

Bars: the network and its gradient within float tolerances stated per test
(bf16 rounding of a layer input can flip when the FP32 accumulation of the
MFMA and the float64 oracle straddle a rounding boundary); collection mode
bit-exact to the plain sampler; collected features bit-exact except the
log2 features (1e-5 relative); labels within 1e-4.
"""
import sys

import numpy as np
import pytest

from oracle import mlp_oracle as mo
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


def _params(seed, scale=1.0):
    r = np.random.default_rng(seed)
    p = (r.standard_normal(mo.N_FLOATS) * 0.3 * scale).astype(np.float32)
    p[mo.OFF_MU:mo.OFF_ISD] = r.standard_normal(16).astype(np.float32)
    p[mo.OFF_ISD:] = (0.5 + r.random(16)).astype(np.float32)
    return p


def _flip_bound(p, X, margin=4e-3):
    """The oracle's outputs for rows X and, per row, a bound on how far a
    correct GPU evaluation can differ from them.  Both round the layer inputs
    to bf16; the GPU accumulates in float32 (MFMA) and takes the GELU's
    quotient on the hardware reciprocal, so a pre-rounding activation can
    land on the other side of a bf16 rounding tie when it lies within a
    small fraction of an ulp of one (`margin`, in ulps; the float32
    differences are ~1e-3 of that).  Each such flip moves that activation by
    one bf16 ulp: a second-layer flip moves the output by |w3_k| ulp; a
    first-layer flip moves every z2 by |W2_kj| ulp to first order (GELU'
    and w3 carry it on), and can flip any second-layer value whose tie lies
    within that move.  Rows with no value near a tie must agree to float32
    accumulation error."""
    y, c = mo.forward(p, X, keep=True)

    def parts(a):
        a32 = np.asarray(a, np.float32)
        u = a32.view(np.uint32)
        frac = (u & 0xFFFF).astype(np.float64) / 65536.0
        ulp = np.ldexp(1.0, ((u >> 23) & 0xFF).astype(np.int64) - 127 - 7)
        return np.abs(frac - 0.5), ulp
    t1, u1 = parts(mo.gelu(c["z1"]))
    t2, u2 = parts(mo.gelu(c["z2"]))
    d1 = np.where(t1 < margin, u1, 0.0)  # first-layer flips, (n, 64)
    d2 = np.where(t2 < margin, u2, 0.0)
    w3, W2 = np.abs(c["w3"]), np.abs(c["W2"])
    dz2 = d1 @ W2.T
    move2 = np.abs(mo.gelu_grad(c["z2"])) * dz2  # how far each a2 moves
    sec = np.where(t2 * u2 <= 1.5 * move2, u2, 0.0)  # flips that move can cause
    bound = (d2 + sec + 1.5 * move2) @ w3
    floor = 1e-5 * (1.0 + np.abs(c["a2"]) @ w3)
    return y, bound + floor, (d1 > 0).any(1) | (d2 > 0).any(1)


@pytest.mark.parametrize("n", [1, 16, 37, 1000])
def test_mlp_forward_matches_oracle(mdas, n):
    _say(f"mlp forward n={n}")
    p = _params(1)
    X = np.random.default_rng(n).standard_normal((n, 16)).astype(np.float32) * 2
    m = mdas.Mlp(0)
    m.set_params(p)
    assert np.array_equal(m.params(), p)
    y = m.forward(X)
    want, bound, near = _flip_bound(p, X)
    d = np.abs(y - want)
    _say(f"  rows near a bf16 tie {near.mean():.3f}, max diff {d.max():.2e}, "
         f"max diff / bound {np.max(d / bound):.3f}")
    assert (d <= bound).all(), (np.argmax(d / bound), float(np.max(d / bound)))
    assert np.median(d) < 1e-4
    m.close()


def test_mlp_gradient_matches_oracle(mdas):
    _say("mlp gradient")
    p = _params(2)
    r = np.random.default_rng(2)
    X = r.standard_normal((300, 16)).astype(np.float32)
    T = r.standard_normal(300).astype(np.float32)
    m = mdas.Mlp(0)
    m.set_params(p)
    g, loss = m.grad(X, T)
    gw, lw = mo.grad(p, X, T)
    assert abs(loss - lw) < 1e-3 * lw
    rel = np.linalg.norm(g - gw) / np.linalg.norm(gw)
    assert rel < 1e-2, rel
    for a, b in ((mo.OFF_W1, mo.OFF_B1), (mo.OFF_B1, mo.OFF_W2), (mo.OFF_W2, mo.OFF_B2),
                 (mo.OFF_B2, mo.OFF_W3), (mo.OFF_W3, mo.OFF_B3), (mo.OFF_B3, mo.N_PARAMS)):
        gb, wb = g[a:b], gw[a:b]
        assert np.linalg.norm(gb - wb) <= 2e-2 * np.linalg.norm(wb) + 1e-7, (a, b)
    m.close()


def test_adam_step_matches_oracle(mdas):
    _say("adam step")
    from pmdas_amd import hip
    p = _params(3)
    r = np.random.default_rng(3)
    X = r.standard_normal((256, 16)).astype(np.float32)
    T = r.standard_normal(256).astype(np.float32)
    m = mdas.Mlp(0)
    m.set_params(p)
    bx, bt = hip.DeviceBuffer(X.nbytes), hip.DeviceBuffer(T.nbytes)
    bx.upload(X)
    bt.upload(T)
    losses = m.train(bx, bt, 256, epochs=1, batch=256, lr=1e-3, seed=7)
    g, lw = mo.grad(p, X, T)
    assert abs(losses[0] - lw) < 1e-3 * lw
    p1, _, _ = mo.adam(p, np.zeros(mo.N_PARAMS), np.zeros(mo.N_PARAMS), g, 1, 1e-3)
    got = m.params()
    big = np.abs(g) > 1e-3 * np.abs(g).max()
    assert np.allclose(got[:mo.N_PARAMS][big], p1[:mo.N_PARAMS][big], rtol=0, atol=2e-6)
    assert np.array_equal(got[mo.N_PARAMS:], p[mo.N_PARAMS:])
    bx.free()
    bt.free()
    m.close()


def test_training_fits_and_is_deterministic(mdas):
    _say("training")
    from pmdas_amd import hip
    r = np.random.default_rng(4)
    n = 50000
    X = r.standard_normal((n, 16)).astype(np.float32)
    T = (np.tanh(X[:, 0] * X[:, 1]) + 0.5 * np.sin(2 * X[:, 2]) - 0.3 * X[:, 3]).astype(np.float32)
    bx, bt = hip.DeviceBuffer(X.nbytes), hip.DeviceBuffer(T.nbytes)
    bx.upload(X)
    bt.upload(T)
    out = []
    for _ in range(2):
        m = mdas.Mlp(11)
        m.fit_norm(bx, n)
        losses = m.train(bx, bt, n, epochs=30, batch=1024, lr=3e-3, seed=5)
        out.append((losses, m.params()))
        m.close()
    losses = out[0][0]
    _say("losses", losses[0], losses[-1], "target var", float(T.var()))
    assert losses[-1] < 0.1 * T.var()
    assert np.array_equal(out[0][1], out[1][1])  # same seed, same bits
    bx.free()
    bt.free()


CASE = dict(kind=3, dim=5, b=1, e=2, W=64, H=36, spp=12, alpha=0.25, cand=4, seed=5)


def _grid(mdas, c):
    target = c["spp"] * c["W"] * c["H"]
    cap = int(target * 1.6) + (4 << (c["b"] * c["dim"] + 2 * c["e"]))
    g = mdas.KDGrid(c["dim"], cap, c["cand"], c["b"], c["e"], c["alpha"], float(c["W"]),
                    float(c["H"]))
    sc = mdas.Scene(c["kind"], c["seed"], c["dim"], float(c["W"]), float(c["H"]))
    return g, sc, target, cap


def test_score_of_constant_network_equals_plain(mdas):
    """All weights zero, b3 = -3: every score is e / 8 exactly, so the
    selection -- pow(score / max, alpha) -- and the whole run equal the plain
    method bit for bit."""
    _say("constant network")
    g0, sc, target, _ = _grid(mdas, CASE)
    g0.Run(sc, target)
    s0 = g0.download_state()
    p = np.zeros(mo.N_FLOATS, np.float32)
    p[mo.OFF_ISD:] = 1.0
    p[mo.OFF_B3] = -3.0
    m = mdas.Mlp(0)
    m.set_params(p)
    g1, sc1, _, _ = _grid(mdas, CASE)
    g1.set_learned(2 | 8, m)  # score, keep the features
    g1.Run(sc1, target)
    s1 = g1.download_state()
    for k in ("coords", "values", "idx", "box", "seed"):
        assert np.array_equal(s0[k], s1[k]), k
    live = s0["err"] >= 0
    assert np.array_equal(s1["err"][live], s0["err"][live] * np.float32(0.125))
    f, pe = g1.leaf_features()
    assert np.array_equal(pe[live], s0["err"][live])
    m.close()


@pytest.mark.parametrize("label", ["reduction", "image"])
def test_collect_rows_match_oracle(mdas, label):
    """Collection mode: the run equals the plain oracle run bit for bit, and
    every row equals the oracle tracker's row (both label forms)."""
    _say(f"collect rows, {label} label")
    from pmdas_amd import hip
    c = CASE
    g, sc, target, cap = _grid(mdas, c)
    cap_rows = 200000
    bx, by = hip.DeviceBuffer(cap_rows * 64), hip.DeviceBuffer(cap_rows * 4)
    g.set_learned(1 if label == "reduction" else 1 | 4, None, bx, by, cap_rows)
    passes = g.Run(sc, target)
    nrows = g.learned_rows()
    X = bx.download(np.float32, nrows * 16).reshape(nrows, 16)
    Y = by.download(np.float32, nrows)
    # oracle replay, pass by pass
    o = pyoracle.OracleGrid(c["dim"], cap, c["cand"], c["b"], c["e"], c["alpha"], float(c["W"]),
                            float(c["H"]))
    params = sc.params_bytes()
    o.initial_sampling()
    o.eval_scene(params, 0, o.n_samples)
    tr = mo.CollectTracker(c["dim"], float(c["W"]), float(c["H"]), label)
    RX, RY = [], []
    p = 0
    while o.n_samples < target:
        before = o.state()
        stale = np.nonzero(before["err"] < 0)[0]
        o.compute_errors()
        st = o.state()
        rx, ry = tr.after_errors(st, stale)
        RX.append(rx)
        RY.append(ry)
        assert o.adaptive_sampling() == 0
        o.eval_scene(params, o.n_samples - o.new_samples, o.new_samples)
        tr.after_adaptive(st, o.state())
        p += 1
    assert p == passes
    RX, RY = np.concatenate(RX), np.concatenate(RY)
    assert nrows == len(RX) > 100
    gs, os_ = g.download_state(), o.state()
    for k in ("coords", "values", "idx", "box", "seed"):
        assert np.array_equal(gs[k], os_[k]), k
    exact = [6, 7, 8, 9, 10, 11, 15]
    assert np.array_equal(X[:, exact], RX[:, exact].astype(np.float32))
    logs = [0, 1, 2, 3, 4, 5, 12, 13, 14]
    assert np.allclose(X[:, logs], RX[:, logs], rtol=1e-5, atol=1e-5)
    assert np.allclose(Y, RY, atol=1e-4)
    if label == "image":  # the label varies (not a constant floor)
        assert np.std(Y) > 1.0
    bx.free()
    by.free()


def test_learned_scores_match_oracle(mdas):
    """Scoring mode: after the run, every live leaf's score equals
    perr * 2^clamp(network(features), -10, 0) with the oracle network."""
    _say("learned scores")
    p = _params(6, scale=0.5)
    m = mdas.Mlp(0)
    m.set_params(p)
    g, sc, target, _ = _grid(mdas, CASE)
    g.set_learned(2 | 8, m)  # score, keep the features
    g.Run(sc, target)
    assert g.GetNumberOfSamples() >= target
    assert g.Validate()
    st = g.download_state()
    f, pe = g.leaf_features()
    live = st["err"] >= 0
    y, bound, near = _flip_bound(p, f[live])
    y = np.clip(y, -10, 0)
    want = pe[live].astype(np.float64) * np.exp2(y)
    rel = np.abs(st["err"][live] - want) / want
    # score = perr * 2^y: a change dy of the output moves it by ln2 * dy
    # relative, plus the float32 rounding of the score itself
    lim = np.log(2.0) * bound * 1.001 + 2e-6
    _say(f"  leaves near a bf16 tie {near.mean():.3f}, max rel {rel.max():.2e}, "
         f"max rel / bound {np.max(rel / lim):.3f}")
    assert (rel <= lim).all(), (int(np.argmax(rel / lim)), float(np.max(rel / lim)))
    assert np.median(rel) < 1e-4
    m.close()


def test_keeping_features_does_not_change_the_run(mdas):
    """A scoring run that keeps only each leaf's depth equals one that keeps
    the full feature rows, bit for bit; reading features back needs them kept."""
    _say("depth-only scoring")
    p = _params(6, scale=0.5)
    m = mdas.Mlp(0)
    m.set_params(p)
    states = []
    for mode in (2, 2 | 8):
        g, sc, target, _ = _grid(mdas, CASE)
        g.set_learned(mode, m)
        g.Run(sc, target)
        states.append(g.download_state())
        if mode == 2:
            with pytest.raises(Exception):
                g.leaf_features()
        g.close()
        sc.close()
    for k in states[0]:
        assert np.array_equal(states[0][k], states[1][k]), k
    m.close()


def test_learned_gain_over_plain(mdas):
    """What the trained network gains, on small sizes: the network predicts
    the measured error reductions better than a constant on held-out scenes,
    and the learned form (score exponent chosen on validation scenes) reaches
    a lower image MSE against an FP64 ground truth than plain MDAS at the
    same sample budget.  Plain sampling with the learned form's effective
    exponent is measured beside it (printed): it isolates how much of the
    gain is the flatter selection rather than the network's ranking."""
    _say("learned gain")
    from pmdas_amd import learned as L
    from pmdas_amd.metrics import mse
    X, Y, n = L.collect(range(40), 96, 54, 32, cap_rows=1 << 23, log=_say)
    Xh, Yh, nh = L.collect(range(1000, 1004), 96, 54, 32, cap_rows=1 << 21, log=_say)
    m, losses = L.train(X, Y, n, epochs=10, batch=16384, lr=1e-3, seed=0, log=_say)
    net, const = L.heldout_loss(m, Xh, Yh, nh)
    _say(f"held-out loss {net:.3f} vs constant {const:.3f}")
    assert net < 0.9 * const
    kappa, _ = L.tune_kappa(m, (1100, 1101), 160, 90, 32, (1.0, 0.25, 0.125), log=_say)
    res = {"plain": [], "learned": [], "same": []}
    for s in (2100, 2101):
        sc = mdas.Scene(3, s, 5, 160.0, 90.0)
        gt = L.ground_truth(sc, 160, 90)
        res["plain"].append(mse(L.render(sc, 160, 90, 32, 5)[0], gt))
        res["learned"].append(mse(L.render(sc, 160, 90, 32, 5, m, kappa=kappa)[0], gt))
        res["same"].append(mse(L.render(sc, 160, 90, 32, 5, alpha=0.25 * kappa)[0], gt))
        sc.close()
    p, l, q = (float(np.mean(res[k])) for k in ("plain", "learned", "same"))
    _say(f"32 spp image mse: plain {p:.4e}, learned (kappa {kappa}) {l:.4e}, "
         f"plain with alpha*kappa {q:.4e}")
    assert l < p
    for b in (X, Y, Xh, Yh):
        b.free()
    m.close()
