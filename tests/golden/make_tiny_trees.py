"""Writes tiny_trees.json: the oracle's grid state, step by step, for the
tiny configurations of tree_fixture.CONFIGS -- after InitialSampling plus
the integrand, then after each ComputeErrors + AdaptiveSampling + integrand
on the new slice.  Run from the repository root after `make -C oracle`:

    python tests/golden/make_tiny_trees.py

The values freeze the oracle as it stands once its closed-form, brute-force
and invariant pins pass; tests check both the oracle and the HIP path
against them.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

from oracle import pyoracle  # noqa: E402
import tree_fixture as tf  # noqa: E402


def run(cfg):
    name, kind, seed, dim, b, e, alpha, cand, sx, sy, passes = cfg
    params = pyoracle.make_scene(kind, seed, dim, sx, sy)
    o = pyoracle.OracleGrid(dim, tf.MAX_SAMPLES, cand, b, e, alpha, sx, sy)
    o.initial_sampling()
    o.eval_scene(params, 0, o.n_samples)
    steps = [tf.encode(o.state(), o.n_samples, o.n_leaves)]
    for _ in range(passes):
        o.compute_errors()
        max_error = o.max_error
        assert o.adaptive_sampling() == 0
        o.eval_scene(params, o.n_samples - o.new_samples, o.new_samples)
        steps.append(tf.encode(o.state(), o.n_samples, o.n_leaves, max_error))
    o.close()
    return {"name": name, "steps": steps}


def main():
    out = {"configs": [list(c) for c in tf.CONFIGS], "runs": [run(c) for c in tf.CONFIGS]}
    with open(os.path.join(HERE, "tiny_trees.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
