"""The oracle against the frozen tiny-grid trees (tests/golden/tiny_trees.json,
written by tests/golden/make_tiny_trees.py), and the fixture itself against
closed forms that hold whatever the oracle does: the initial grid has four
samples per leaf in Morton order, new samples extend the sample list in one
contiguous slice, every sample sits in exactly one leaf, and with alpha = 0
every leaf is selected (N' = N + L, L' = L + #full leaves, kdgrid.cu:326,
:392-470)."""
import json
import os
import sys

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
import tree_fixture as tf  # noqa: E402

with open(os.path.join(GOLDEN, "tiny_trees.json")) as _f:
    FIX = json.load(_f)
RUNS = {r["name"]: r["steps"] for r in FIX["runs"]}


def test_fixture_lists_every_config():
    assert [c[0] for c in tf.CONFIGS] == [r["name"] for r in FIX["runs"]]
    assert [list(c) for c in tf.CONFIGS] == FIX["configs"]


@pytest.mark.parametrize("cfg", tf.CONFIGS, ids=[c[0] for c in tf.CONFIGS])
def test_fixture_closed_forms(cfg):
    name, kind, seed, dim, b, e, alpha, cand, sx, sy, passes = cfg
    steps = RUNS[name]
    assert len(steps) == passes + 1
    L0 = 1 << (b * dim + 2 * e)
    assert (steps[0]["n_samples"], steps[0]["n_leaves"]) == (4 * L0, L0)
    for k, s in enumerate(steps):
        if "arrays" not in s:
            continue
        idx = tf.decode(s["arrays"]["idx"]).view(np.int32).reshape(-1, 4)
        assert len(idx) == s["n_leaves"]
        used = idx[idx >= 0]
        assert np.array_equal(np.sort(used), np.arange(s["n_samples"]))  # a partition
        if k == 0:
            assert np.array_equal(idx.reshape(-1), np.arange(4 * L0))
            err = tf.decode(s["arrays"]["err"]).view(np.float32)
            assert (err == -1).all()
    if alpha == 0.0:
        for a, c in zip(steps, steps[1:]):
            assert "arrays" in a
            cnt = (tf.decode(a["arrays"]["idx"]).view(np.int32).reshape(-1, 4) >= 0).sum(1)
            assert c["n_samples"] == a["n_samples"] + a["n_leaves"]
            assert c["n_leaves"] == a["n_leaves"] + int((cnt == 4).sum())


@pytest.mark.parametrize("cfg", tf.CONFIGS, ids=[c[0] for c in tf.CONFIGS])
def test_oracle_reproduces_fixture(orc, cfg):
    name, kind, seed, dim, b, e, alpha, cand, sx, sy, passes = cfg
    params = orc.make_scene(kind, seed, dim, sx, sy)
    o = orc.OracleGrid(dim, tf.MAX_SAMPLES, cand, b, e, alpha, sx, sy)
    o.initial_sampling()
    o.eval_scene(params, 0, o.n_samples)
    steps = RUNS[name]
    assert tf.mismatches(steps[0], o.state(), o.n_samples, o.n_leaves) == []
    for p in range(passes):
        o.compute_errors()
        m = o.max_error
        assert o.adaptive_sampling() == 0
        o.eval_scene(params, o.n_samples - o.new_samples, o.new_samples)
        assert tf.mismatches(steps[p + 1], o.state(), o.n_samples, o.n_leaves, m) == [], p
    o.close()
