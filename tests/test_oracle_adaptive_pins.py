"""Independent recomputation of the oracle's adaptive pass
(kdgrid.cu:316-477): selection at alpha > 0 and best-candidate placement
with candidatesNum > 1, then insert / median split.

Nothing here calls the oracle's arithmetic.  The random streams come from a
pure-Python TEA / LCG restatement that this file first checks against the
published vectors: the 32-cycle TEA vector (all-zero key and block) and
Numerical Recipes' ranqd1 sequence.  pow is libm's double-precision pow,
and distances are float64.  The oracle's state before a pass (after its error
pass) is the input.  Its state after the pass is compared with the
recomputation: the selected set, the new sample's coordinates, slots, boxes,
errors and LCG states, for 3 passes over every D in 3..6,
candidatesNum in {2, 4, 8} and alpha in {1/64, 1/4, 1}.

Near-ties cannot be settled without the oracle's own pow and float32
distances.  These are a selection draw within 1e-6 of pow(err/max, alpha),
or two candidates whose nearest-sample distances agree to 1e-5.  The test
accepts either outcome there, and asserts that such cases are rare.  Every
other leaf is decided by the float64 computation alone, so a swapped
comparison, an inverted ratio or a dropped term changes the outcome of many
leaves and fails the test.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle

M32 = 0xFFFFFFFF
TEA_KEY = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)  # cuda/random.h:40-42


def tea(v0, v1, rounds, key):
    """TEA encipher of (v0, v1), `rounds` cycles (Wheeler & Needham)."""
    s = 0
    for _ in range(rounds):
        s = (s + 0x9E3779B9) & M32
        v0 = (v0 + ((((v1 << 4) + key[0]) & M32) ^ ((v1 + s) & M32) ^ (((v1 >> 5) + key[1]) & M32))) & M32
        v1 = (v1 + ((((v0 << 4) + key[2]) & M32) ^ ((v0 + s) & M32) ^ (((v0 >> 5) + key[3]) & M32))) & M32
    return v0, v1


class Lcg:
    """x <- 1664525 x + 1013904223 (mod 2^32); a draw is (x & 2^24-1) / 2^24."""

    def __init__(self, state):
        self.x = state & M32

    def draw(self):
        self.x = (1664525 * self.x + 1013904223) & M32
        return (self.x & 0xFFFFFF) / float(1 << 24)


def test_restated_rng_matches_published_vectors():
    assert tea(0, 0, 32, (0, 0, 0, 0)) == (0x41EA3A0A, 0x94BAA940)
    g = Lcg(0)
    seq = []
    for _ in range(4):
        g.draw()
        seq.append(g.x)
    assert seq == [0x3C6EF35F, 0x47502932, 0xD1CCF6E9, 0xAAF95334]


def _seed(seed, i):
    return int(seed) if int(seed) != 0 else tea(i, 0, 4, TEA_KEY)[0]


def _recompute(before, gmax, dim, cand, alpha):
    """The pass from `before` (oracle state after its error pass): returns the
    expected state after it plus the number of near-tie decisions taken from
    the oracle (`after` is consulted only for those)."""
    D = dim
    coords, idx, box, err, seed = (before[k] for k in ("coords", "idx", "box", "err", "seed"))
    L, N = len(idx), len(coords)
    f32 = np.float32
    sel, states, draws = [], [], []
    ambiguous_sel = []
    for i in range(L):
        r = Lcg(_seed(seed[i], i))
        u = r.draw()
        ratio = f32(err[i]) / f32(gmax)  # float division, as the kernel does
        p = math.pow(float(ratio), alpha) if alpha != 0 else 1.0
        sel.append(p >= u)
        ambiguous_sel.append(abs(p - u) < 1e-6)
        states.append(r)
    return sel, ambiguous_sel, states


def _candidate_rule(pts, mn, diag, r, cand, D):
    """Candidates mn + u * diag (float32 ops); the one with the largest float64
    distance to its nearest sample, earliest on ties.  Returns (all
    candidates, index of the best, indices tied with it to 1e-5)."""
    f32 = np.float32
    cands, dmin = [], []
    for _ in range(cand):
        u = [r.draw() for _ in range(D)]
        c = np.array([f32(mn[d]) + f32(u[d]) * f32(diag[d]) for d in range(D)], np.float32)
        cands.append(c)
        d2 = [math.sqrt(sum((float(c[d]) - float(p[d])) ** 2 for d in range(D))) for p in pts]
        dmin.append(min(d2))
    best = 0
    for j in range(1, cand):
        if dmin[j] > dmin[best]:
            best = j
    tied = [j for j in range(cand) if abs(dmin[j] - dmin[best]) <= 1e-5 * max(dmin[best], 1e-30)]
    return cands, best, tied


def _check_pass(before, after, gmax, dim, cand, alpha):
    D = dim
    f32 = np.float32
    L, N = len(before["idx"]), len(before["coords"])
    sel, amb, states = _recompute(before, gmax, dim, cand, alpha)
    # the oracle's selected set, read from its LCG states: a selected leaf
    # advanced 1 + cand * D draws, an unselected one 1
    osel = []
    for i in range(L):
        r1 = Lcg(_seed(before["seed"][i], i))
        r1.draw()
        osel.append(int(after["seed"][i]) != r1.x)
    n_amb = 0
    for i in range(L):
        if amb[i]:
            n_amb += 1
            sel[i] = osel[i]
        assert sel[i] == osel[i], f"leaf {i}: selection differs"
    n_sel = sum(sel)
    assert len(after["coords"]) == N + n_sel
    rank_s = rank_l = 0
    E = ~D
    n_tie = 0
    split_rank = []
    for i in range(L):
        r = states[i]
        if not sel[i]:
            assert int(after["seed"][i]) == r.x
            assert np.array_equal(after["idx"][i], before["idx"][i])
            assert np.array_equal(after["box"][i], before["box"][i])
            continue
        ids = [int(k) for k in before["idx"][i] if k >= 0]
        pts = [before["coords"][k] for k in ids]
        mn, mx = before["box"][i, :D], before["box"][i, D:]
        diag = np.array([f32(mx[d]) - f32(mn[d]) for d in range(D)], np.float32)
        cands, best, tied = _candidate_rule(pts, mn, diag, r, cand, D)
        k = N + rank_s
        rank_s += 1
        got = after["coords"][k]
        if len(tied) > 1:
            n_tie += 1
            assert any(np.array_equal(got, cands[j]) for j in tied), f"leaf {i}"
            newp = got
        else:
            assert np.array_equal(got, cands[best]), f"leaf {i}: best candidate differs"
            newp = cands[best]
        assert int(after["seed"][i]) == r.x, f"leaf {i}: LCG state differs"
        if len(ids) < 4:  # insert at the first free slot, error kept
            want = list(ids) + [k] + [E] * (3 - len(ids))
            assert list(after["idx"][i]) == want
            assert np.array_equal(after["box"][i], before["box"][i])
            assert after["err"][i] == before["err"][i]
            continue
        # split: the largest axis, ties to the highest index (aabb.h:34-45)
        axis = max(d for d in range(D) if diag[d] == diag.max())
        five = ids + [k]
        key = {s: (before["coords"][s][axis] if s < N else newp[axis]) for s in five}
        order = sorted(five, key=lambda s: key[s])  # stable
        pos = f32(0.5) * (f32(key[order[1]]) + f32(key[order[2]]))
        m = L + rank_l
        rank_l += 1
        split_rank.append(m)
        assert list(after["idx"][i]) == [order[0], order[1], E, E], f"leaf {i}: left child"
        assert list(after["idx"][m]) == [order[2], order[3], order[4], E], f"leaf {i}: right child"
        lbox = before["box"][i].copy()
        lbox[D + axis] = pos
        rbox = before["box"][i].copy()
        rbox[axis] = pos
        assert np.array_equal(after["box"][i], lbox)
        assert np.array_equal(after["box"][m], rbox)
        assert after["err"][i] == -1.0 and after["err"][m] == -1.0
        assert after["seed"][m] == 0
    assert len(after["idx"]) == L + rank_l
    return n_sel, rank_l, n_amb, n_tie


SCENE_FOR_DIM = {3: 1, 4: 2, 5: 3, 6: 3}


@pytest.mark.parametrize("dim", [3, 4, 5, 6])
@pytest.mark.parametrize("cand", [2, 4, 8])
@pytest.mark.parametrize("alpha", [1.0 / 64.0, 0.25, 1.0])
def test_adaptive_pass_recomputed(dim, cand, alpha):
    W, H = 24.0, 16.0
    b, e = (1, 1) if dim <= 4 else (0, 3)
    g = pyoracle.OracleGrid(dim, 1 << 15, cand, b, e, alpha, W, H)
    params = pyoracle.make_scene(SCENE_FOR_DIM[dim], 7 + dim, dim, W, H)
    g.initial_sampling()
    g.eval_scene(params, 0, g.n_samples)
    totals = [0, 0, 0, 0]
    for p in range(3):
        g.compute_errors()
        before = g.state()
        gmax = g.max_error
        assert gmax == before["err"].max()  # brute-force max
        assert g.adaptive_sampling() == 0
        after = g.state()
        res = _check_pass(before, after, gmax, dim, cand, alpha)
        totals = [a + r for a, r in zip(totals, res)]
        g.eval_scene(params, g.n_samples - g.new_samples, g.new_samples)
    n_sel, n_split, n_amb, n_tie = totals
    assert n_sel > 20 and n_split > 5, totals
    assert n_amb <= max(2, n_sel // 50) and n_tie <= max(2, n_sel // 50), totals
    if alpha == 1.0:
        assert n_amb == 0  # pow(x, 1) = x: every comparison is exact
