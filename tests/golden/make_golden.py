"""Writes tests/golden/golden.npz: small input/expected-output vectors for every operation.

Expected outputs come from numpy's reshape/transpose and slicing only (the textbook
definitions of the operations, test/cl_unittest.cc:138-166 and 549-551 of the reference),
never from the oracle or the library, so they pin both. Inputs are seeded bytes; elements of
4+ bytes carry their own index. Run: python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ES = (1, 2, 4, 8, 16)


def elements(count, es, seed):
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 256, size=(count, es), dtype=np.uint8)
    if es >= 4:
        w = min(es, 8)
        x[:, :w] = np.arange(count, dtype=np.uint64).view(np.uint8).reshape(count, 8)[:, :w]
    return x


def main():
    out = {}
    seed = 100
    for es in ES:
        for (A, a, B) in [(3, 4, 5), (2, 7, 3), (1, 64, 16), (5, 1, 9), (4, 16, 1), (2, 18, 72)]:
            seed += 1
            x = elements(A * a * B, es, seed)
            y = x.reshape(A, a, B, es).transpose(0, 2, 1, 3)
            key = f"010_es{es}_{A}_{a}_{B}"
            out[key + "_in"] = x.reshape(-1)
            out[key + "_out"] = np.ascontiguousarray(y).reshape(-1)
        for (A, B, b) in [(4, 5, 3), (7, 3, 1), (16, 8, 2), (3, 11, 5)]:
            seed += 1
            x = elements(A * B * b, es, seed)
            y = x.reshape(A, B, b, es).transpose(1, 0, 2, 3)
            key = f"100_es{es}_{A}_{B}_{b}"
            out[key + "_in"] = x.reshape(-1)
            out[key + "_out"] = np.ascontiguousarray(y).reshape(-1)
        for (A, a, B, b) in [(3, 4, 5, 2), (2, 6, 7, 1), (4, 3, 3, 3)]:
            seed += 1
            x = elements(A * a * B * b, es, seed)
            y = x.reshape(A, a, B, b, es).transpose(0, 2, 1, 3, 4)
            key = f"0100_es{es}_{A}_{a}_{B}_{b}"
            out[key + "_in"] = x.reshape(-1)
            out[key + "_out"] = np.ascontiguousarray(y).reshape(-1)
        for (M, N) in [(6, 10), (7, 13), (18, 72), (1, 9), (9, 1), (12, 12)]:
            seed += 1
            x = elements(M * N, es, seed)
            y = x.reshape(M, N, es).transpose(1, 0, 2)
            key = f"T_es{es}_{M}_{N}"
            out[key + "_in"] = x.reshape(-1)
            out[key + "_out"] = np.ascontiguousarray(y).reshape(-1)
        for (xs, ys, pad) in [(5, 7, 8), (3, 4, 3), (10, 3, 17), (1, 9, 2)]:
            seed += 1
            x = elements(ys * pad, es, seed)
            y = np.zeros((ys, pad, es), np.uint8)
            y[:, :xs] = x[: ys * xs].reshape(ys, xs, es)
            key = f"PAD_es{es}_{xs}_{ys}_{pad}"
            out[key + "_in"] = x.reshape(-1)
            out[key + "_out"] = y.reshape(-1)
            seed += 1
            x = elements(ys * pad, es, seed)
            y = np.zeros((ys * pad, es), np.uint8)
            y[: ys * xs] = x.reshape(ys, pad, es)[:, :xs].reshape(ys * xs, es)
            key = f"UNPAD_es{es}_{xs}_{ys}_{pad}"
            out[key + "_in"] = x.reshape(-1)
            out[key + "_out"] = y.reshape(-1)
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print(len(out) // 2, "vectors")


if __name__ == "__main__":
    main()
