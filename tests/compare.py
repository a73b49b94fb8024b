"""Comparison helpers shared by the parity tests."""
import numpy as np


def relerr(a, b):
    """Normwise relative difference max|a-b| / max|b| (absolute if b == 0)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.abs(b).max() if b.size else 0.0
    d = np.abs(a - b).max() if a.size else 0.0
    return d / scale if scale > 0 else d


def rows(a, P):
    """Drop the periodic halo rows j=0 and j=maxm (keep j=1..P)."""
    return a[:, 1:P + 1, ...]


def demean(p, maxl, P, maxn):
    """Pressure modulo its constant null-space component (interior mean)."""
    return p - p[1:maxl, 1:P + 1, 1:maxn].mean()
