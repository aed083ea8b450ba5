"""shared helpers for the parity tests"""
import os

import numpy as np

from conftest import ROOT


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b)) if b.size else 1.0
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0)) if a.size else 0.0


def sorted_rows(row, ent, key=None):
    """per-row sorted neighbour lists (optionally mapped through key)"""
    out = []
    for i in range(len(row) - 1):
        r = ent[row[i]:row[i + 1]]
        if key is not None:
            r = key[r]
        out.append(np.sort(r))
    return out
