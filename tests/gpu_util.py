"""Helpers for the GPU parity tests (HIP library vs oracle)."""
import numpy as np

from oracle import oracle as orc
from pmsim_amd.solver import Fem


def rel_err(a, b):
    a = np.asarray(a, float); b = np.asarray(b, float)
    den = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / den


def csr_dense(rowptr, cols, vals, n):
    K = np.zeros((n, n))
    for r in range(n):
        K[r, cols[rowptr[r]:rowptr[r + 1]]] = vals[rowptr[r]:rowptr[r + 1]]
    return K


def pair(d, cfg=None, keep_history=True):
    """Oracle and HIP handles for the same flat description."""
    return orc.Oracle(d), Fem(d, cfg, keep_history=keep_history)
