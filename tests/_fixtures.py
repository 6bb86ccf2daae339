"""Loaders for the golden vectors extracted from the reference's test data."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _arr(t):
    e = np.asarray(t["els"], dtype=np.float64)
    return e[:, 0] + 1j * e[:, 1]


def contraction_cases(group):
    """Yield (name, t1, d1, t2, d2, t3, d3, wires) for a group of reference
    tests/gen/random-tensors.hpp (e.g. 'dense_vals')."""
    with open(os.path.join(GOLDEN, "random_tensors.json")) as f:
        data = json.load(f)
    for idx in data["groups"][group]:
        c = data["cases"]["v%d" % idx]
        yield ("v%d" % idx, _arr(c["t1"]), c["t1"]["dims"], _arr(c["t2"]), c["t2"]["dims"],
               _arr(c["t3"]), c["t3"]["dims"], [tuple(w) for w in c["wires"]])


def network_cases(group):
    """Yield (name, [(t, dims)], [(i1, i2, wires)], (result, dims)) for a group of
    reference tests/gen/random-tn.hpp (e.g. 'tn2_vals')."""
    with open(os.path.join(GOLDEN, "random_tn.json")) as f:
        data = json.load(f)
    for idx in data["groups"][group]:
        n = data["networks"]["v%d" % idx]
        yield ("v%d" % idx, [(_arr(t), t["dims"]) for t in n["tensors"]],
               [(b["t1"], b["t2"], [tuple(w) for w in b["wires"]]) for b in n["bonds"]],
               (_arr(n["result"]), n["result"]["dims"]))


def swap_tensor(n):
    """Dense form of the reference SwapTensor element rule (symm.cpp:246-284):
    T[a, b, c, d] = 1 if a == d and b == c."""
    t = np.zeros((n, n, n, n), dtype=np.complex128)
    for a in range(n):
        for b in range(n):
            t[a, b, b, a] = 1
    return t.reshape(-1)


def iden_tensor(dims):
    """Dense form of the reference IdenTensor rule (diag.cpp:432-440)."""
    m = int(np.prod(dims))
    return np.eye(m, dtype=np.complex128).reshape(-1)
