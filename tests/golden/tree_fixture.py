"""Shared by make_tiny_trees.py and the tests that read tiny_trees.json: the
tiny-grid configurations whose state is frozen after the initial sampling
and after each of the first passes (kdgrid.cu:633-779, called in the order
of optixMotionBlur.cpp:768-791), and the state encoding.

State arrays are compared as raw 32-bit patterns: coords (N, D), values
(N, 3), idx (L, 4), box (L, 2D), err (L), seed (L).  Every step records the
counters, the SHA-256 of each array's little-endian bytes and, for states of
at most FULL_LIMIT words, the arrays themselves (base64 of those bytes).
"""
import base64
import hashlib

import numpy as np

# name, scene kind, scene seed, dim, bitsPerDim, extraImgBits, alpha,
# candidatesNum, scaleX, scaleY, passes
CONFIGS = [
    ("disks-d3-b1-e1", 1, 0, 3, 1, 1, 0.25, 4, 4.0, 4.0, 3),
    ("dof-d4-b0-e2", 2, 2, 4, 0, 2, 0.25, 4, 4.0, 4.0, 3),
    ("mbdof-d5-b1-e1", 3, 1, 5, 1, 1, 0.25, 4, 8.0, 8.0, 3),
    ("ao-d6-b0-e1", 4, 5, 6, 0, 1, 0.5, 2, 2.0, 2.0, 3),
    ("disks-d3-b1-e2-alpha0", 1, 7, 3, 1, 2, 0.0, 1, 16.0, 16.0, 1),
    ("mbdof-d5-b1-e3", 3, 4, 5, 1, 3, 0.25, 4, 64.0, 64.0, 1),
]
MAX_SAMPLES = 1 << 16
KEYS = ("coords", "values", "idx", "box", "err", "seed")
FULL_LIMIT = 6000


def words(a):
    return np.ascontiguousarray(a).reshape(-1).view(np.uint32)


def encode(state, n_samples, n_leaves, max_error=None):
    step = {"n_samples": int(n_samples), "n_leaves": int(n_leaves)}
    if max_error is not None:
        step["max_error"] = int(np.float32(max_error).view(np.uint32))
    raw = {k: words(state[k]).astype("<u4").tobytes() for k in KEYS}
    step["sha256"] = {k: hashlib.sha256(raw[k]).hexdigest() for k in KEYS}
    if sum(len(raw[k]) for k in KEYS) <= 4 * FULL_LIMIT:
        step["arrays"] = {k: base64.b64encode(raw[k]).decode() for k in KEYS}
    return step


def decode(text):
    return np.frombuffer(base64.b64decode(text), "<u4")


def mismatches(step, state, n_samples, n_leaves, max_error=None):
    """Names of what differs between a frozen step and a state (empty = equal)."""
    got = encode(state, n_samples, n_leaves, max_error)
    bad = [k for k in ("n_samples", "n_leaves", "max_error") if k in step and step[k] != got.get(k)]
    for k in KEYS:
        if step["sha256"][k] != got["sha256"][k]:
            want, have = (decode(step["arrays"][k]) if "arrays" in step else None), words(state[k])
            if want is not None and len(want) == len(have):
                d = np.flatnonzero(want != have)
                bad.append(f"{k} (first differing words {d[:4].tolist()})")
            else:
                bad.append(k)
    return bad
