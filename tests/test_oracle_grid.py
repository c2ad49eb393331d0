"""Pins of the oracle's grid algorithm (kdgrid.cu) by closed forms, brute
force and invariants -- chosen so that a dropped term, a wrong index or a
swapped operand in the oracle fails one of them.
"""
import itertools
import json
import os

import numpy as np
import pytest

from oracle import pyoracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _grid(dim=3, max_samples=1 << 16, cand=4, b=0, e=3, alpha=0.25, sx=8.0, sy=8.0):
    return pyoracle.OracleGrid(dim, max_samples, cand, b, e, alpha, sx, sy)


# --------------------------------------------------------------- Morton grid --
def _cells(orc, dim, b, e):
    lib = orc.load()
    L = 1 << (b * dim + 2 * e)
    xq = np.zeros((L, dim), np.uint32)
    off = np.zeros((L, dim), np.float32)
    ext = np.zeros((L, dim), np.float32)
    for i in range(L):
        lib.orc_morton_cell(i, dim, b, e, pyoracle.vp(xq[i]), pyoracle.vp(off[i]),
                            pyoracle.vp(ext[i]))
    return xq, off, ext


def test_morton_hand_table(orc):
    """Hand-derived cells of kdgrid.cu:141-164 for D=3, b=1, e=1."""
    with open(os.path.join(GOLDEN, "morton_d3_b1_e1.json")) as f:
        table = json.load(f)
    xq, _, _ = _cells(orc, 3, 1, 1)
    for row in table["cells"]:
        assert list(xq[row["leaf"]]) == row["xq"], row


@pytest.mark.parametrize("dim,b,e", [(3, 1, 1), (4, 0, 2), (3, 2, 1), (5, 1, 1), (6, 1, 0)])
def test_morton_bijection_and_tiling(orc, dim, b, e):
    xq, off, ext = _cells(orc, dim, b, e)
    # every cell of the 2^(b+e) x 2^(b+e) x 2^b ... grid appears exactly once
    sizes = [1 << (b + (e if d < 2 else 0)) for d in range(dim)]
    flat = np.ravel_multi_index(xq.T.astype(np.int64), sizes)
    assert len(np.unique(flat)) == len(flat) == int(np.prod(sizes))
    # extents are the powers of two; cells tile the unit cube exactly
    for d in range(dim):
        assert np.all(ext[:, d] == np.float32(1.0 / sizes[d]))
        assert np.all(off[:, d] == xq[:, d].astype(np.float32) * ext[:, d])
    assert np.isclose(np.prod(ext, axis=1).astype(np.float64).sum(), 1.0)


def test_morton_b0_is_textbook_z_order(orc):
    """b = 0: y takes the even index bits, x the odd ones (2D Z-order)."""
    xq, _, _ = _cells(orc, 3, 0, 3)
    for i in range(64):
        y = sum(((i >> (2 * k)) & 1) << k for k in range(3))
        x = sum(((i >> (2 * k + 1)) & 1) << k for k in range(3))
        assert (xq[i, 0], xq[i, 1]) == (x, y)


def test_initial_sampling_inside_cells(orc):
    g = _grid(dim=4, b=1, e=2, sx=6.0, sy=3.0)
    g.initial_sampling()
    st = g.state()
    L = g.n_leaves
    assert L == 1 << (4 + 4) and g.n_samples == 4 * L
    assert np.array_equal(st["idx"], np.arange(4 * L, dtype=np.int32).reshape(L, 4))
    mn, mx = st["box"][:, :4], st["box"][:, 4:]
    c = st["coords"].reshape(L, 4, 4)
    assert np.all((c >= mn[:, None, :]) & (c <= mx[:, None, :]))
    assert np.all(st["err"] == -1.0)
    vol = np.prod((mx - mn).astype(np.float64), axis=1).sum()
    assert np.isclose(vol, 6.0 * 3.0, rtol=1e-6)
    assert g.validate()


# --------------------------------------------------------------- leaf error --
def _err(orc, vals, vol=1.0):
    v = np.ascontiguousarray(vals, np.float32).reshape(-1)
    return orc.load().orc_leaf_error(pyoracle.vp(v), len(v) // 3, vol)


def test_error_constant_leaf(orc):
    for n in (2, 3, 4):
        assert _err(orc, [[0.3, 0.6, 0.9]] * n, 2.0) == np.float32(1e-5) * np.float32(2.0)


def test_error_zero_channel_skipped(orc):
    e = _err(orc, [[0, 1, 1], [0, 3, 1]], 1.0)
    # channel x: mean 0 -> skipped; y: mean 2, |dev| sum 2 -> 1; z: 0
    assert np.isclose(e, 1.0 / 2 + 1e-5, rtol=1e-6)


def test_error_mean_is_reciprocal_multiply(orc):
    """float3 /= n is x * (1/n) (sutil/vec_math.h:492-496), not x / n."""
    rng = np.random.default_rng(11)
    third = np.float32(1.0) / np.float32(3.0)
    found = 0
    for _ in range(20000):
        v = rng.random((3, 3), dtype=np.float32)
        s = (v[0] + v[1]) + v[2]
        if np.all(s * third == s / np.float32(3.0)):
            continue
        found += 1
        avg_r, avg_d = s * third, s / np.float32(3.0)

        def err_of(avg):
            e = np.float32(0.0)
            for ch in range(3):
                dif = np.float32(0.0)
                for k in range(3):
                    dif = dif + np.abs(v[k, ch] - avg[ch])
                if avg[ch] != 0:
                    e = e + dif / avg[ch]
            return (e / np.float32(3.0) + np.float32(1e-5)) * np.float32(1.0)

        got = _err(orc, v)
        if err_of(avg_r) != err_of(avg_d):
            assert got == err_of(avg_r)
            assert got != err_of(avg_d)
        if found > 200:
            break
    assert found > 0


def test_max_error_is_brute_force_max(orc):
    g = _grid(dim=3, b=0, e=4, sx=16, sy=16)
    g.initial_sampling()
    rng = np.random.default_rng(5)
    g.set_values(0, rng.random((g.n_samples, 3), dtype=np.float32))
    g.compute_errors()
    st = g.state()
    assert np.all(st["err"] > 0)
    assert g.max_error == st["err"].max()
    for i in range(0, g.n_leaves, 37):  # spot-check against the closed form
        vals = st["values"][st["idx"][i]]
        vol = np.prod(st["box"][i, 3:] - st["box"][i, :3], dtype=np.float32)
        assert st["err"][i] == _err(orc, vals, float(vol))


# ------------------------------------------------------ selection, placement --
def test_sort5_is_stable_sort(orc):
    lib = orc.load()
    order = np.zeros(5, np.int32)
    for keys in itertools.product([0.0, 1.0, 2.0], repeat=5):
        k = np.array(keys, np.float32)
        lib.orc_sort5(pyoracle.vp(k), pyoracle.vp(order))
        assert list(order) == sorted(range(5), key=lambda i: keys[i])


def test_largest_axis_ties_to_highest_index(orc):
    lib = orc.load()
    for diag, want in [([1, 1, 1], 2), ([2, 1, 1], 0), ([1, 2, 2, 1], 2), ([3, 3, 1, 1, 1], 1),
                       ([1, 1, 1, 1, 1, 1], 5), ([0.5, 0.25, 0.5, 0.5, 0.25], 3)]:
        d = np.array(diag, np.float32)
        assert lib.orc_largest_axis(pyoracle.vp(d), len(d)) == want


def test_alpha_zero_selects_every_leaf(orc):
    g = _grid(dim=3, b=0, e=3, alpha=0.0)
    g.initial_sampling()
    rng = np.random.default_rng(2)
    g.set_values(0, rng.random((g.n_samples, 3), dtype=np.float32))
    L, N = g.n_leaves, g.n_samples
    for p in range(3):
        full = int((g.state()["idx"][:, 3] >= 0).sum())
        assert g.sampling_pass() == 0
        assert g.new_samples == L and g.n_samples == N + L
        assert g.new_nodes == full and g.n_leaves == L + full
        g.set_values(N, rng.random((g.new_samples, 3), dtype=np.float32))
        L, N = g.n_leaves, g.n_samples
        assert g.validate()


def test_new_samples_inside_their_leaf_and_split_at_median(orc):
    g = _grid(dim=4, b=1, e=1, alpha=0.0, sx=4, sy=4)
    g.initial_sampling()
    rng = np.random.default_rng(4)
    g.set_values(0, rng.random((g.n_samples, 3), dtype=np.float32))
    before = g.state()
    g.sampling_pass()
    st = g.state()
    L0 = len(before["idx"])
    D = 4
    for i in range(L0):
        parent = before["box"][i]
        diag = parent[D:] - parent[:D]
        a = max(d for d in range(D) if diag[d] == diag.max())
        left, right = st["idx"][i], st["idx"][L0 + i]  # every leaf was full: one split each
        assert (left[2:] < 0).all() and (right[:3] >= 0).all() and right[3] < 0
        pos = st["box"][i, D + a]
        assert st["box"][L0 + i, a] == pos
        lc = st["coords"][left[:2], a]
        rc = st["coords"][right[:3], a]
        assert lc.max() <= pos <= rc.min()
        assert pos == np.float32(0.5) * (np.sort(np.concatenate([lc, rc]))[1] +
                                         np.sort(np.concatenate([lc, rc]))[2])
    assert g.validate()


def test_single_candidate_is_first_draw_in_box(orc):
    g = _grid(dim=3, cand=1, b=0, e=2, alpha=0.0, sx=4, sy=4)
    g.initial_sampling()
    g.set_values(0, np.ones((g.n_samples, 3), np.float32))
    before = g.state()
    g.sampling_pass()
    st = g.state()
    N0 = len(before["coords"])
    lib = orc.load()
    import ctypes
    for i in range(len(before["idx"])):
        s = ctypes.c_uint32(int(before["seed"][i]))
        lib.orc_rnd(ctypes.byref(s))  # selection draw
        r = np.array([lib.orc_rnd(ctypes.byref(s)) for _ in range(3)], np.float32)
        mn, mx = before["box"][i, :3], before["box"][i, 3:]
        assert np.array_equal(st["coords"][N0 + i], mn + r * (mx - mn))


def test_topology_invariants_over_passes(orc):
    params = pyoracle.make_scene(1, 0, 3, 64.0, 64.0)
    g = _grid(dim=3, max_samples=200000, b=0, e=4, sx=64, sy=64)
    passes = g.run(params, 8 * 64 * 64)
    st = g.state()
    N, L = g.n_samples, g.n_leaves
    assert passes > 0 and N >= 8 * 64 * 64
    cnt = (st["idx"] >= 0).sum(axis=1)
    assert cnt.min() >= 2 and cnt.max() <= 4 and cnt.sum() == N and 2 * L <= N
    ref = st["idx"][st["idx"] >= 0]
    assert np.array_equal(np.sort(ref), np.arange(N))
    assert g.validate()


def test_run_is_deterministic(orc):
    params = pyoracle.make_scene(1, 3, 3, 32.0, 32.0)
    out = []
    for threads in (1, 4):
        pyoracle.load().orc_set_threads(threads)
        g = _grid(dim=3, max_samples=50000, b=0, e=3, sx=32, sy=32)
        g.run(params, 8 * 32 * 32)
        out.append(g.state())
    pyoracle.load().orc_set_threads(0)
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k


# ------------------------------------------------------------ reconstruction --
def _recon64(st, dim, W, H, sx, sy):
    """Box filter in float64 over the same leaves (independent evaluation)."""
    mn, mx = st["box"][:, :dim].astype(np.float64), st["box"][:, dim:].astype(np.float64)
    valid = st["idx"] >= 0
    vals = st["values"][np.where(valid, st["idx"], 0)].astype(np.float64)
    avg = (vals * valid[..., None]).sum(1) / valid.sum(1)[:, None]
    out = np.zeros((H, W, 3))
    other = np.prod(mx[:, 2:] - mn[:, 2:], axis=1)
    px, py = sx / W, sy / H
    for i in range(len(mn)):
        x0, x1 = int(mn[i, 0] / px), min(int(mx[i, 0] / px) + 1, W)
        y0, y1 = int(mn[i, 1] / py), min(int(mx[i, 1] / py) + 1, H)
        for y in range(y0, y1):
            oy = min(mx[i, 1], (y + 1) * py) - max(mn[i, 1], y * py)
            if oy <= 0:
                continue
            for x in range(x0, x1):
                ox = min(mx[i, 0], (x + 1) * px) - max(mn[i, 0], x * px)
                if ox > 0:
                    out[y, x] += avg[i] * ox * oy * other[i] / (px * py)
    return out


def test_reconstruction_constant_integrand(orc):
    params = pyoracle.make_scene(0, 0, 4, 48.0, 40.0)
    g = _grid(dim=4, max_samples=100000, b=1, e=2, sx=48, sy=40)
    g.run(params, 4 * 48 * 40)
    px = g.integrate(48, 40)
    want = np.array([0.25, 0.5, 0.75], np.float32)
    assert np.allclose(px[..., :3], want, rtol=2e-6, atol=0)
    assert np.all(px[..., 3] == 0)


def test_reconstruction_cell_per_pixel_is_mean(orc):
    g = _grid(dim=3, b=0, e=3, sx=8, sy=8)
    g.initial_sampling()
    rng = np.random.default_rng(9)
    vals = rng.random((g.n_samples, 3), dtype=np.float32)
    g.set_values(0, vals)
    px = g.integrate(8, 8)
    st = g.state()
    for i in range(g.n_leaves):
        x, y = int(st["box"][i, 0]), int(st["box"][i, 1])
        s = vals[st["idx"][i]]
        mean = ((s[0] + s[1]) + s[2] + s[3]) * np.float32(0.25)
        assert np.array_equal(px[y, x, :3], mean)


def test_reconstruction_matches_float64(orc):
    params = pyoracle.make_scene(3, 1, 5, 40.0, 24.0)
    g = _grid(dim=5, max_samples=100000, b=1, e=1, sx=40, sy=24)
    g.run(params, 6 * 40 * 24)
    px = g.integrate(40, 24)
    ref = _recon64(g.state(), 5, 40, 24, 40.0, 24.0)
    assert np.allclose(px[..., :3], ref, rtol=1e-5, atol=1e-6)


def test_reconstruction_scaled_sample_space(orc):
    """scaleFactor != 1: sample space sf*W x sf*H maps onto the full image."""
    params = pyoracle.make_scene(1, 2, 3, 16.0, 12.0)
    g = _grid(dim=3, max_samples=60000, b=1, e=2, sx=16.0, sy=12.0)
    g.run(params, 10 * 32 * 24)
    px = g.integrate(32, 24)
    ref = _recon64(g.state(), 3, 32, 24, 16.0, 12.0)
    assert np.allclose(px[..., :3], ref, rtol=1e-5, atol=1e-6)


def test_recon_keys_sorted_and_cover_spans(orc):
    params = pyoracle.make_scene(2, 0, 4, 40.0, 30.0)
    g = _grid(dim=4, max_samples=60000, b=1, e=1, sx=40, sy=30)
    g.run(params, 5 * 40 * 30)
    keys = g.recon_keys(40, 30, 8, 8)
    assert np.all(np.diff(keys.astype(np.uint64)) > 0)
    leaves = (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    assert set(leaves.tolist()) == set(range(g.n_leaves))


# --------------------------------------------------------------------- bytes --
def _srgb8_64(x):
    x = np.clip(np.asarray(x, np.float64), 0.0, 1.0)
    s = np.where(x < 0.0031308, 12.92 * x, 1.055 * np.power(x, 1 / 2.4) - 0.055)
    return np.minimum((np.clip(s, 0, 1) * 256).astype(np.int64), 255), s * 256


def test_bytes_against_float64_srgb(orc):
    # inputs just around every quantisation boundary plus random values
    bounds = []
    for k in range(1, 256):
        t = k / 256.0
        lin = ((t + 0.055) / 1.055) ** 2.4 if t > 0.04045 else t / 12.92
        bounds.append(lin)
    b = np.array(bounds, np.float32)
    xs = np.concatenate([b, np.nextafter(b, 1), np.nextafter(b, 0),
                         np.random.default_rng(0).random(20000, dtype=np.float32),
                         np.array([-1, 0, 1, 2, np.float32(0.0031308)], np.float32)])
    px = np.zeros((len(xs), 4), np.float32)
    px[:, 0] = xs
    px[:, 1] = xs[::-1]
    px[:, 2] = 0.5
    out = pyoracle.convert_to_bytes(px)
    assert np.all(out[:, 3] == 255)
    for ch in range(2):
        q, scaled = _srgb8_64(px[:, ch])
        near = np.abs(scaled - np.round(scaled)) < 1e-3
        ok = (out[:, ch] == q) | near
        assert ok.all()


def test_sampling_density_histogram(orc):
    params = pyoracle.make_scene(1, 0, 3, 32.0, 32.0)
    g = _grid(dim=3, max_samples=40000, b=0, e=3, sx=32, sy=32)
    g.run(params, 6 * 32 * 32)
    den = g.sampling_density(32, 32)
    c = g.state()["coords"]
    x = np.minimum((c[:, 0] / np.float32(32.0) * np.float32(32.0)).astype(np.int64), 31)
    y = np.minimum((c[:, 1] / np.float32(32.0) * np.float32(32.0)).astype(np.int64), 31)
    hist = np.bincount(y * 32 + x, minlength=32 * 32).reshape(32, 32)
    assert np.allclose(den[..., 0], hist * 0.05, rtol=1e-5)
    assert np.all(den[..., 3] == 1.0)


def test_capacity_error_leaves_state(orc):
    g = _grid(dim=3, max_samples=300, b=0, e=3, alpha=0.0, sx=8, sy=8)  # 256 initial
    g.initial_sampling()
    g.set_values(0, np.ones((g.n_samples, 3), np.float32))
    g.compute_errors()
    st = g.state()
    assert g.adaptive_sampling() == -2
    st2 = g.state()
    for k in ("coords", "idx", "box", "err"):
        assert np.array_equal(st[k], st2[k])
