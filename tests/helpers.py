"""Shared helpers of the test suite (fixture parsing, comparisons)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def read_freqs(path):
    """freqs.dat -> (energies, assignment strings); line 0 holds the ZPE,
    later lines E_n - E_0 (Ham.cpp:448-502)."""
    e, a = [], []
    with open(path) as f:
        for line in f:
            head, _, rest = line.rstrip("\n").partition("   ")
            e.append(float(head))
            a.append(rest.strip())
    return np.array(e), a


def absolute(e):
    out = e.copy()
    out[1:] += e[0]
    return out


def csr_to_dict(rp, col, val, diag=None):
    d = {}
    n = len(rp) - 1
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            d[(i, int(col[k]))] = float(val[k])
        if diag is not None:
            d[(i, i)] = float(diag[i])
    return d
