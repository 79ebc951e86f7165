"""Writes tests/golden/rng_vectors.json: known-answer vectors of the per-pixel RNG.

The generator restates opencl_kernels.h:78-97 with Python integers (64-bit
product of the state word and 0xD256D193, split into high and low words),
independently of the C oracle.  Seeds follow acmmp_random_init
(opencl_kernels.h:1003-1005) for a few pixel indices and keys.
"""
import json
import os

import numpy as np

M = 0xD256D193
GOLD = 0x9E3779B9
MASK = 0xFFFFFFFF


def draw(state, key):
    x, y = state
    for i in range(4):
        prod = x * M
        hi, lo = prod >> 32, prod & MASK
        x, y = (y ^ ((key + i * GOLD) & MASK) ^ hi) & MASK, lo
    return (x, y), float(np.float32(x) / np.float32(4294967296.0))


def main():
    cases = []
    for idx in (0, 1, 2, 77, 12345, 1919999, 4294967):
        state = ((idx * 1099087573 + 2654435769) & MASK, (idx * 2246822519 + 13) & MASK)
        key = (idx + 7919) & MASK
        seq = []
        st = state
        for k in (key, key + 3, key + 5, key + 600, key + 600 + 7919, key + 0xFFFFFFF0):
            st, u = draw(st, k & MASK)
            seq.append({"key": k & MASK, "state": list(st), "u": u})
        cases.append({"idx": idx, "seed": list(state), "draws": seq})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_vectors.json")
    with open(path, "w") as f:
        json.dump(cases, f, indent=1)


if __name__ == "__main__":
    main()
