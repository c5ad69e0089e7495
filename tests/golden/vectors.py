"""Iterates the committed golden vectors (tests/golden/golden.npz)."""
import os

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")


def cases():
    """Yields (op, es, dims, input_bytes, expected_bytes) for every vector."""
    z = np.load(PATH, allow_pickle=False)
    keys = sorted(k[:-3] for k in z.files if k.endswith("_in"))
    for k in keys:
        parts = k.split("_")
        op, es = parts[0], int(parts[1][2:])
        dims = tuple(int(p) for p in parts[2:])
        yield op, es, dims, z[k + "_in"], z[k + "_out"]
