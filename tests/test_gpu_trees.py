"""The HIP path, one C-ABI call at a time, against the frozen tiny-grid
trees (tests/golden/tiny_trees.json): the same inputs give the same state,
bit for bit, after InitialSampling + integrand and after every
ComputeErrors + AdaptiveSampling + integrand (optixMotionBlur.cpp:768-791)."""
import json
import os
import sys

import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)
import tree_fixture as tf  # noqa: E402

with open(os.path.join(GOLDEN, "tiny_trees.json")) as _f:
    RUNS = {r["name"]: r["steps"] for r in json.load(_f)["runs"]}

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mdas():
    import pmdas_amd
    from pmdas_amd import kdgrid, hip
    pmdas_amd.load()
    assert hip.device_count() >= 1, "no GPU visible"
    return kdgrid


@pytest.mark.parametrize("cfg", tf.CONFIGS, ids=[c[0] for c in tf.CONFIGS])
def test_hip_path_reproduces_fixture(mdas, cfg):
    name, kind, seed, dim, b, e, alpha, cand, sx, sy, passes = cfg
    print(f"{name}: dim={dim} b={b} e={e} passes={passes}", flush=True)
    g = mdas.KDGrid(dim, tf.MAX_SAMPLES, cand, b, e, alpha, sx, sy)
    sc = mdas.Scene(kind, seed, dim, sx, sy)
    steps = RUNS[name]
    g.InitialSampling()
    g.EvalIntegrand(sc, 0, g.GetNumberOfSamples())
    st = g.download_state()
    assert tf.mismatches(steps[0], st, g.GetNumberOfSamples(), g.GetNumberOfNodes()) == []
    for p in range(passes):
        g.ComputeErrors()
        m = g.GetMaxError()
        g.AdaptiveSampling()
        g.EvalNewSamples(sc)
        st = g.download_state()
        bad = tf.mismatches(steps[p + 1], st, g.GetNumberOfSamples(), g.GetNumberOfNodes(), m)
        assert bad == [], (p, bad)
    assert g.Validate()
    g.close()
    sc.close()
