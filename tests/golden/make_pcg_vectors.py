"""Writes tests/golden/pcg_vectors.json: known-answer states and values of
Reina's random() for a few seeds, computed with Python integers from the
published PCG definitions (LCG step state * 747796405 + 1 mod 2^32, output
permutation RXS-M-XS-32: xorshift by (state >> 28) + 4, multiply by
277803737, xorshift by 22) and the float conversion float(word) / 2^32 in
round-to-nearest-even (numpy float32 cast).

Run: python tests/golden/make_pcg_vectors.py
"""
import json
import os

import numpy as np

M = 2 ** 32


def stream(seed, n):
    s = seed
    states, bits = [], []
    for _ in range(n):
        s = (s * 747796405 + 1) % M
        w = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) % M
        w = (w >> 22) ^ w
        v = np.float32(w) / np.float32(4294967296.0)
        states.append(s)
        bits.append(int(np.array([v], np.float32).view(np.uint32)[0]))
    return states, bits


def main():
    cases = []
    for seed in (0, 1, 2, 1920 * 1080 - 1, 4294967295, 123456789):
        st, b = stream(seed, 64)
        cases.append({"seed": seed, "states": st, "value_bits": b})
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pcg_vectors.json")
    json.dump({"generator": "tests/golden/make_pcg_vectors.py", "cases": cases}, open(out, "w"))
    print("wrote", out)


if __name__ == "__main__":
    main()
