"""CPU oracle of the learned proposal -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.

The learned proposal has no counterpart in the reference; this module
restates the build's own definitions so the HIP kernels can be checked:

- network: 16 -> 64 -> 64 -> 1, GELU (tanh form), weights and layer inputs
  rounded to bf16 (round to nearest even), biases FP32; evaluated here in
  float64 on the rounded values.  Backward: straight-through for every
  rounding (d bf16(w)/dw = 1), the rounded W2 / w3 in the chain rule.
- Adam (beta 0.9 / 0.999, eps 1e-8, bias-corrected).
- leaf features and split labels, evaluated in float32 in the kernels'
  operation order (log2 in float64).
- a tracker that replays the sampling loop of the grid oracle
  (oracle.pyoracle.OracleGrid) pass by pass and yields the rows the
  collection mode writes.

Pins (tests/test_oracle_mlp.py): torch float64 autograd with the same
straight-through roundings, central finite differences on the unrounded
float64 network, torch.optim.Adam, torch's bf16 conversion and GELU, and
closed forms of the features on hand-built leaves.
"""
import numpy as np

OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 1024, 1088, 5184, 5248, 5312
N_PARAMS, OFF_MU, OFF_ISD, N_FLOATS = 5313, 5313, 5329, 5345
C_GELU = 0.7978845608  # sqrt(2/pi) as the kernels write it


def bf16(x):
    """Round float32 values to bf16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C_GELU * (x + 0.044715 * x ** 3)))


def gelu_grad(x):
    t = np.tanh(C_GELU * (x + 0.044715 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * C_GELU * (1.0 + 3.0 * 0.044715 * x * x)


def unpack(p, dtype=np.float32):
    p = np.asarray(p, dtype)
    return dict(W1=p[OFF_W1:OFF_B1].reshape(64, 16), b1=p[OFF_B1:OFF_W2],
                W2=p[OFF_W2:OFF_B2].reshape(64, 64), b2=p[OFF_B2:OFF_W3], w3=p[OFF_W3:OFF_B3],
                b3=p[OFF_B3], mu=p[OFF_MU:OFF_ISD], isd=p[OFF_ISD:N_FLOATS])


def staged_inputs(p, X):
    q = unpack(p)
    X = np.asarray(X, np.float32)
    return bf16((X - q["mu"]) * q["isd"]).astype(np.float64)


def forward(p, X, rounding=True, keep=False):
    """Network outputs for raw feature rows X (n, 16); float64.  Without
    rounding the parameters are taken in float64 as given."""
    q = unpack(p, np.float32 if rounding else np.float64)
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float32))
    x = staged_inputs(p, X) if rounding else (
        (np.asarray(X, np.float64) - q["mu"].astype(np.float64)) * q["isd"].astype(np.float64))
    W1 = r(q["W1"]).astype(np.float64)
    W2 = r(q["W2"]).astype(np.float64)
    w3 = r(q["w3"]).astype(np.float64)
    if not rounding:
        W1, W2, w3 = (q[k].astype(np.float64) for k in ("W1", "W2", "w3"))
    z1 = x @ W1.T + q["b1"].astype(np.float64)
    a1 = gelu(z1)
    if rounding:
        a1 = bf16(a1.astype(np.float32)).astype(np.float64)
    z2 = a1 @ W2.T + q["b2"].astype(np.float64)
    a2 = gelu(z2)
    if rounding:
        a2 = bf16(a2.astype(np.float32)).astype(np.float64)
    y = a2 @ w3 + float(q["b3"])
    if keep:
        return y, dict(x=x, z1=z1, a1=a1, z2=z2, a2=a2, W2=W2, w3=w3)
    return y


def grad(p, X, T, rounding=True):
    """Gradient (N_PARAMS,) of mean((y - T)^2) and the loss, float64."""
    y, c = forward(p, X, rounding, keep=True)
    n = len(y)
    dy = 2.0 * (y - np.asarray(T, np.float64)) / n
    g = np.zeros(N_PARAMS)
    g[OFF_W3:OFF_B3] = dy @ c["a2"]
    g[OFF_B3] = dy.sum()
    dz2 = dy[:, None] * c["w3"][None, :] * gelu_grad(c["z2"])
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    g[OFF_W2:OFF_B2] = (dz2.T @ c["a1"]).ravel()
    dz1 = (dz2 @ c["W2"]) * gelu_grad(c["z1"])
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    g[OFF_W1:OFF_B1] = (dz1.T @ c["x"]).ravel()
    return g, float(np.mean((y - T) ** 2))


def adam(p, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step on the first N_PARAMS entries; returns (p, m, v)."""
    p = np.array(p, np.float64)
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    p[:N_PARAMS] -= lr * mh / (np.sqrt(vh) + eps)
    return p, m, v


# ------------------------------------------------------------ features --
def _log2(x):
    return np.log2(np.maximum(np.float32(x), np.float32(1e-30)).astype(np.float64))


def leaf_features(vals, n, mn, mx, e, depth, D, sX, sY):
    """Features of one leaf: vals (4, 3) float32 by slot (first n used),
    box mn / mx (D,), plain error e, depth."""
    f32 = np.float32
    f = np.zeros(16, np.float64)
    for d in range(6):
        if d < D:
            dom = f32(sX) if d == 0 else (f32(sY) if d == 1 else f32(1.0))
            f[d] = _log2((f32(mx[d]) - f32(mn[d])) / dom)
    lum = [f32(f32(f32(0.2126) * f32(v[0])) + f32(f32(0.7152) * f32(v[1]))) + f32(f32(0.0722) * f32(v[2]))
           for v in vals]
    s = f32(0.0)
    for q in range(n):
        s = f32(s + lum[q])
    mean = f32(s / f32(n))
    var = f32(0.0)
    for q in range(n):
        dd = f32(lum[q] - mean)
        var = f32(var + f32(dd * dd))
    var = f32(var / f32(n))
    vol = f32(1.0)
    for d in range(D):
        vol = f32(vol * f32(f32(mx[d]) - f32(mn[d])))
    f[6] = n
    f[7] = mean
    f[8] = var
    f[9] = min(lum[:n])
    f[10] = max(lum[:n])
    f[11] = depth
    f[12] = _log2(e)
    f[13] = _log2(f32(f32(e) / vol))
    f[14] = _log2(vol)
    f[15] = D
    return f


def split_label(ep, el, er):
    f32 = np.float32
    ep, el, er = f32(ep), f32(el), f32(er)
    if not ep > 0:
        return -10.0
    d = f32(f32(ep - el) - er)
    r = f32(max(d, f32(ep * f32(0.0009765625))) / ep)
    return float(np.log2(np.float64(r)))


def image_label(mp, mc, vc, ac):
    """log2(sum_c (m_c - m_p)^2 v_c^2 / a_c + 1e-30) in the kernel's float32
    operation order (log2 in float64)."""
    f32 = np.float32
    s = f32(0.0)
    for h in range(2):
        d = f32(f32(mc[h]) - f32(mp))
        w = f32(f32(f32(vc[h]) * f32(vc[h])) / f32(ac[h]))
        s = f32(s + f32(f32(d * d) * w))
    return float(np.log2(np.float64(f32(s + f32(1e-30)))))


class CollectTracker:
    """Replays the collection mode on the grid oracle.  After each
    compute_errors call after_errors(state, stale leaves): it computes the
    features of the re-estimated leaves and returns the rows of the splits of
    the previous pass, in split order.  After each adaptive pass call
    after_adaptive(state before, state after) to record that pass's splits."""

    def __init__(self, D, sX, sY, label="reduction"):
        self.D, self.sX, self.sY = D, sX, sY
        self.label = label
        self.feat = None
        self.depth = None
        self.perr = None
        self.parents = []

    def _leaf(self, st, i):
        idx = st["idx"][i]
        n = int((idx >= 0).sum())
        vals = np.zeros((4, 3), np.float32)
        for q in range(4):
            if idx[q] >= 0:
                vals[q] = st["values"][idx[q]]
        D = self.D
        return vals, n, st["box"][i, :D], st["box"][i, D:]

    def after_adaptive(self, st_before, st_after):
        """Record the split parents of the adaptive pass (leaf order)."""
        L0 = len(st_before["err"])
        L1 = len(st_after["err"])
        parents = [i for i in range(L0) if st_after["err"][i] < 0]
        assert len(parents) == L1 - L0, (len(parents), L1 - L0)
        self.parents = parents
        self.L_old = L0
        d_new = np.zeros(L1, np.int64)
        d_new[:L0] = self.depth
        for j, i in enumerate(parents):
            d_new[i] = self.depth[i] + 1
            d_new[L0 + j] = self.depth[i] + 1
        self.depth_next = d_new

    def after_errors(self, st, stale):
        """st: state after compute_errors; stale: indices re-estimated."""
        L = len(st["err"])
        rows_X, rows_Y = [], []
        if self.feat is None:
            self.feat = np.zeros((L, 16))
            self.perr = np.zeros(L, np.float32)
            self.depth = np.zeros(L, np.int64)
        else:
            for j, i in enumerate(self.parents):
                m = self.L_old + j
                rows_X.append(self.feat[i].copy())
                if self.label == "reduction":
                    rows_Y.append(split_label(self.perr[i], st["err"][i], st["err"][m]))
                else:
                    mc, vc, ac = [], [], []
                    for c in (i, m):
                        vals, n, mn, mx = self._leaf(st, c)
                        fc = leaf_features(vals, n, mn, mx, st["err"][c], 0, self.D, self.sX,
                                           self.sY)
                        mc.append(np.float32(fc[7]))
                        vol = np.float32(1.0)
                        for d in range(self.D):
                            vol = np.float32(vol * np.float32(np.float32(mx[d]) - np.float32(mn[d])))
                        vc.append(vol)
                        ac.append(np.float32(np.float32(mx[0] - mn[0]) * np.float32(mx[1] - mn[1])))
                    rows_Y.append(image_label(np.float32(self.feat[i][7]), mc, vc, ac))
            feat = np.zeros((L, 16))
            feat[:len(self.feat)] = self.feat
            perr = np.zeros(L, np.float32)
            perr[:len(self.perr)] = self.perr
            self.feat, self.perr, self.depth = feat, perr, self.depth_next
        for i in stale:
            vals, n, mn, mx = self._leaf(st, i)
            self.feat[i] = leaf_features(vals, n, mn, mx, st["err"][i], self.depth[i], self.D,
                                         self.sX, self.sY)
            self.perr[i] = st["err"][i]
        return np.array(rows_X).reshape(-1, 16), np.array(rows_Y)
