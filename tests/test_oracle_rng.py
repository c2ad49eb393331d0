"""Pins of the oracle's random number generation (cuda/random.h:31-67).

- TEA: the round function is checked against the published 32-cycle TEA test
  vector (Wheeler & Needham's cipher, all-zero key and block); the reference's
  seed function is the same round function with 4 cycles and its own key.
- LCG: the first outputs of Numerical Recipes' ranqd1 generator from state 0
  (published in Numerical Recipes in C, 2nd ed., sec. 7.1), and k-step
  jump-ahead by composing the affine map x -> a x + c (mod 2^32), a method
  independent of stepping.
- rnd: 24 random bits scaled by 2^-24, exact in binary32.
"""
import ctypes
import random

import numpy as np

TEA_REF_KEY = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)  # cuda/random.h:40-42


def test_tea_published_vector(orc):
    assert orc.tea(0, 0, 32, (0, 0, 0, 0)) == (0x41EA3A0A, 0x94BAA940)


def test_tea4_is_four_rounds_with_reference_key(orc):
    lib = orc.load()
    for v0, v1 in [(0, 0), (1, 0), (12345, 0), (2**31 - 1, 0), (7, 9)]:
        assert lib.orc_tea4(v0, v1) == orc.tea(v0, v1, 4, TEA_REF_KEY)[0]


def test_tea4_distinct_on_leaf_indices(orc):
    lib = orc.load()
    seeds = {lib.orc_tea4(i, 0) for i in range(1 << 16)}
    assert len(seeds) == 1 << 16  # TEA is a permutation of the 64-bit block


def test_lcg_numerical_recipes_sequence(orc):
    lib = orc.load()
    s = ctypes.c_uint32(0)
    expect = [0x3C6EF35F, 0x47502932, 0xD1CCF6E9, 0xAAF95334, 0x6252E503, 0x9F2EC686,
              0x57FE6C2D, 0xA3D95FA8, 0x81FDBEE7, 0x94F0AF1A, 0xCBF633B1]
    got = []
    for _ in expect:
        r = lib.orc_lcg(ctypes.byref(s))
        got.append(s.value)
        assert r == s.value & 0xFFFFFF
    assert got == expect


def _jump(a, c, k, x):
    # compose (A, C) k times by squaring: x -> A x + C
    A, C = 1, 0
    pa, pc = a, c
    while k:
        if k & 1:
            A, C = (pa * A) % 2**32, (pa * C + pc) % 2**32
        pa, pc = (pa * pa) % 2**32, (pa * pc + pc) % 2**32
        k >>= 1
    return (A * x + C) % 2**32


def test_lcg_jump_ahead(orc):
    lib = orc.load()
    rng = random.Random(1)
    for _ in range(20):
        x0 = rng.getrandbits(32)
        k = rng.randint(1, 3000)
        s = ctypes.c_uint32(x0)
        for _ in range(k):
            lib.orc_lcg(ctypes.byref(s))
        assert s.value == _jump(1664525, 1013904223, k, x0)


def test_rnd_exact_24_bit(orc):
    lib = orc.load()
    s = ctypes.c_uint32(0xDEADBEEF)
    for _ in range(5000):
        before = s.value
        r = lib.orc_rnd(ctypes.byref(s))
        assert 0.0 <= r < 1.0
        assert np.float32(r) * np.float32(2**24) == np.float32(s.value & 0xFFFFFF)
        assert s.value == (1664525 * before + 1013904223) % 2**32
