"""Pins of the learned proposal's CPU oracle (oracle/mlp_oracle.py), CPU only.

- bf16 rounding and GELU against torch (round-to-nearest-even conversion,
  gelu(approximate='tanh')) and the published value gelu_tanh(1) = 0.841192;
- the forward and straight-through gradient against torch float64 autograd
  of the same network, and the unrounded gradient against central finite
  differences;
- Adam against torch.optim.Adam;
- features and labels on hand-built leaves against closed forms.
"""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo


def _params(seed=0, scale=1.0):
    r = np.random.default_rng(seed)
    p = (r.standard_normal(mo.N_FLOATS) * 0.3 * scale).astype(np.float32)
    p[mo.OFF_MU:mo.OFF_ISD] = r.standard_normal(16).astype(np.float32)
    p[mo.OFF_ISD:] = (0.5 + r.random(16)).astype(np.float32)
    return p


def test_bf16_matches_torch():
    r = np.random.default_rng(1)
    x = np.concatenate([r.standard_normal(10000).astype(np.float32) * 10.0 ** r.integers(-8, 8, 10000),
                        np.array([1.0, 1.00390625, 1.01171875, -2.5, 0.0], np.float32)]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(mo.bf16(x), want)


def test_gelu_matches_torch_and_published_value():
    assert abs(mo.gelu(1.0) - 0.8411919906) < 1e-9
    x = np.linspace(-8, 8, 1001)
    want = torch.nn.functional.gelu(torch.from_numpy(x), approximate="tanh").numpy()
    assert np.allclose(mo.gelu(x), want, atol=1e-10)
    xt = torch.from_numpy(x).requires_grad_()
    torch.nn.functional.gelu(xt, approximate="tanh").sum().backward()
    assert np.allclose(mo.gelu_grad(x), xt.grad.numpy(), atol=1e-9)


def _torch_grad(p, X, T):
    """Same network in torch float64 with straight-through bf16 roundings."""
    q = {k: torch.tensor(np.asarray(v, np.float64), requires_grad=True)
         for k, v in mo.unpack(p).items() if k not in ("mu", "isd")}

    def st(a):  # value bf16(a), gradient 1
        r = torch.from_numpy(mo.bf16(a.detach().numpy().astype(np.float32)).astype(np.float64))
        return a + (r - a).detach()

    def gelu(z):
        return torch.nn.functional.gelu(z, approximate="tanh")

    x = torch.from_numpy(mo.staged_inputs(p, X))
    a1 = st(gelu(x @ st(q["W1"]).T + q["b1"]))
    a2 = st(gelu(a1 @ st(q["W2"]).T + q["b2"]))
    y = a2 @ st(q["w3"]) + q["b3"]
    loss = ((y - torch.from_numpy(np.asarray(T, np.float64))) ** 2).mean()
    loss.backward()
    g = np.zeros(mo.N_PARAMS)
    for k, off in (("W1", mo.OFF_W1), ("b1", mo.OFF_B1), ("W2", mo.OFF_W2), ("b2", mo.OFF_B2),
                   ("w3", mo.OFF_W3), ("b3", mo.OFF_B3)):
        v = q[k].grad.numpy().ravel()
        g[off:off + v.size] = v
    return y.detach().numpy(), g, float(loss.detach())


def test_forward_and_gradient_match_torch_autograd():
    r = np.random.default_rng(2)
    p = _params(2)
    X = r.standard_normal((37, 16)).astype(np.float32) * 3
    T = r.standard_normal(37)
    y_t, g_t, l_t = _torch_grad(p, X, T)
    g, l = mo.grad(p, X, T)
    assert np.allclose(mo.forward(p, X), y_t, rtol=1e-12, atol=1e-12)
    assert np.allclose(g, g_t, rtol=1e-9, atol=1e-12)
    assert abs(l - l_t) < 1e-12


def test_gradient_finite_differences_unrounded():
    r = np.random.default_rng(3)
    p = _params(3).astype(np.float64)
    X = r.standard_normal((9, 16))
    T = r.standard_normal(9)
    g, _ = mo.grad(p, X, T, rounding=False)

    def loss(pp):
        return np.mean((mo.forward(pp, X, rounding=False) - T) ** 2)

    for k in list(range(0, 5313, 97)) + [mo.OFF_B3, mo.OFF_W3 + 5, mo.OFF_B1 + 7]:
        h = 1e-5
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (loss(pp) - loss(pm)) / (2 * h)
        assert abs(fd - g[k]) < 1e-6 * max(1.0, abs(fd)), k


def test_forward_closed_form_constant_network():
    p = np.zeros(mo.N_FLOATS, np.float32)
    p[mo.OFF_ISD:] = 1
    p[mo.OFF_B2:mo.OFF_W3] = 0.5
    p[mo.OFF_W3:mo.OFF_B3] = 0.25
    p[mo.OFF_B3] = -1.0
    X = np.random.default_rng(4).standard_normal((5, 16))
    a = float(mo.bf16(np.float32([mo.gelu(0.5)]))[0])
    assert np.allclose(mo.forward(p, X), 64 * 0.25 * a - 1.0, atol=1e-12)


def test_adam_matches_torch():
    r = np.random.default_rng(5)
    p = _params(5)
    X = r.standard_normal((20, 16)).astype(np.float32)
    T = r.standard_normal(20)
    w = torch.tensor(p[:mo.N_PARAMS].astype(np.float64), requires_grad=True)
    opt = torch.optim.Adam([w], lr=1e-3)
    pp = p.astype(np.float64)
    m = np.zeros(mo.N_PARAMS)
    v = np.zeros(mo.N_PARAMS)
    for t in range(1, 4):
        g, _ = mo.grad(pp, X, T, rounding=False)
        w.grad = torch.from_numpy(g.copy())
        opt.step()
        pp, m, v = mo.adam(pp, m, v, g, t, 1e-3)
        assert np.allclose(pp[:mo.N_PARAMS], w.detach().numpy(), rtol=1e-10, atol=1e-12)
    # first step moves each weight by lr * sign(g) (|g| >> eps)
    g, _ = mo.grad(p, X, T)
    p1, _, _ = mo.adam(p, np.zeros(mo.N_PARAMS), np.zeros(mo.N_PARAMS), g, 1, 1e-3)
    big = np.abs(g) > 1e-4
    assert np.allclose((p.astype(np.float64) - p1)[:mo.N_PARAMS][big], 1e-3 * np.sign(g[big]), rtol=1e-4)


def test_features_closed_form():
    vals = np.array([[1, 1, 1], [0, 0, 0], [0.5, 0.5, 0.5], [9, 9, 9]], np.float32)
    mn = np.array([0, 0, 0.25, 0.5, 0], np.float32)
    mx = np.array([4, 2, 0.5, 1.0, 0.125], np.float32)
    f = mo.leaf_features(vals, 3, mn, mx, np.float32(0.01), 7, 5, 64.0, 32.0)
    # grey values: luminance = value (weights sum to 1 up to rounding)
    assert np.allclose(f[:6], [np.log2(4 / 64), np.log2(2 / 32), -2, -1, -3, 0])
    assert f[6] == 3 and f[11] == 7 and f[15] == 5
    assert np.isclose(f[7], 0.5, atol=1e-6) and np.isclose(f[8], 1 / 6, atol=1e-6)
    assert np.isclose(f[9], 0.0, atol=1e-7) and np.isclose(f[10], 1.0, atol=1e-6)
    vol = 4 * 2 * 0.25 * 0.5 * 0.125
    assert np.isclose(f[14], np.log2(vol)) and np.isclose(f[12], np.log2(0.01))
    assert np.isclose(f[13], np.log2(0.01 / vol), atol=1e-6)


def test_split_label_closed_form():
    assert mo.split_label(1.0, 0.25, 0.25) == -1.0          # half the error removed
    assert mo.split_label(1.0, 0.75, 0.5) == -10.0          # no reduction: the floor
    assert mo.split_label(0.5, 0.0, 0.0) == 0.0             # all removed
    assert mo.split_label(0.0, 0.0, 0.0) == -10.0


def test_image_label_closed_form():
    # one child changes mean by 0.5 with volume 2 over area 4: 0.25 * 4 / 4
    assert mo.image_label(1.0, [1.5, 1.0], [2.0, 3.0], [4.0, 3.0]) == np.log2(np.float32(0.25 + 1e-30))
    # both children equal the parent: only the 1e-30 floor remains
    assert abs(mo.image_label(0.3, [0.3, 0.3], [1.0, 1.0], [1.0, 1.0]) - np.log2(1e-30)) < 1e-6
    # scaling the volumes by k scales the change by k^2
    a = mo.image_label(0.0, [1.0, 2.0], [1.0, 1.0], [1.0, 1.0])
    b = mo.image_label(0.0, [1.0, 2.0], [2.0, 2.0], [1.0, 1.0])
    assert abs((b - a) - 2.0) < 1e-6
