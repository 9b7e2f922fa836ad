"""Pins random() (shaderCommon.h.glsl:39-45): an LCG step s' = 747796405 s + 1
(mod 2^32) followed by the RXS-M-XS output permutation, converted as
float(w) / 4294967295.0f, where the constant rounds to 2^32."""
import json
import os
from fractions import Fraction

import numpy as np

import oracle

A, C, M = 747796405, 1, 2 ** 32
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "pcg_vectors.json")


def test_golden_vectors():
    """Known-answer table written by tests/golden/make_pcg_vectors.py from the
    published RXS-M-XS definition (O'Neill, PCG, 2014)."""
    g = json.load(open(GOLDEN))
    for case in g["cases"]:
        st, val = oracle.pcg_stream(case["seed"], len(case["states"]))
        assert st.tolist() == case["states"]
        assert [int(v) for v in np.ascontiguousarray(val).view(np.uint32)] == case["value_bits"]


def test_lcg_jump_ahead():
    """s_n = A^n s_0 + C (A^n - 1)/(A - 1) mod 2^32 (closed form of the affine map)."""
    seed, n = 123456789, 5000
    st, _ = oracle.pcg_stream(seed, n)
    for k in (1, 2, 17, 999, 5000):
        geo = (A ** k - 1) // (A - 1)          # exact big-integer geometric sum
        expect = (pow(A, k, M) * seed + C * geo) % M
        assert int(st[k - 1]) == expect


def test_full_period_multiplier():
    # Hull-Dobell: c odd, a-1 divisible by 4 (every prime factor of 2^32) -> period 2^32
    assert C % 2 == 1 and (A - 1) % 4 == 0


def _output(s):
    w = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) % M
    return (w >> 22) ^ w


def _inv_output(w):
    """Inverse of the RXS-M-XS permutation (for the value-one known answer)."""
    # undo w ^= w >> 22
    x = w ^ (w >> 22)
    # undo multiply
    x = (x * pow(277803737, -1, M)) % M
    # undo x = (s >> ((s >> 28) + 4)) ^ s : the top 4 bits are unchanged
    top = x >> 28
    shift = top + 4
    s = x
    for _ in range(8):
        s = x ^ (s >> shift)
    return s


def test_float_conversion_and_value_one():
    """float(w) * 2^-32 in RNE; w >= 2^32 - 128 returns exactly 1.0."""
    s_star = _inv_output(M - 1)
    assert _output(s_star) == M - 1
    prev = ((s_star - C) * pow(A, -1, M)) % M
    st, val = oracle.pcg_stream(prev, 1)
    assert int(st[0]) == s_star
    assert float(val[0]) == 1.0
    rng = np.random.default_rng(3)
    seeds = rng.integers(0, M, 2000)
    for s0 in seeds[:200]:
        st, val = oracle.pcg_stream(int(s0), 1)
        w = _output(int(st[0]))
        exact = Fraction(int(np.float32(w)), M)        # RNE to float32, exact scaling by 2^-32
        assert Fraction(float(val[0])) == exact
