"""Shared numerics helpers for the tests."""
import numpy as np


def rel_fro(x, ref) -> float:
    """||x - ref||_F / ||ref||_F (absolute norm when ref is zero)."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.linalg.norm(ref)
    num = np.linalg.norm(x - ref)
    return float(num / den) if den > 0 else float(num)


def max_rel(x, ref) -> float:
    """max |x - ref| / max |ref|."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.max(np.abs(ref)) if ref.size else 0.0
    num = np.max(np.abs(x - ref)) if x.size else 0.0
    return float(num / den) if den > 0 else float(num)


def op_view(x, rows, cols, ld, trans):
    """op(X) as a (rows x cols) view of a flat row-major buffer:
    op(X)(i, p) = x[p*ld + i] if trans else x[i*ld + p]."""
    assert x.ndim == 1 and x.flags.c_contiguous
    if trans:
        return np.lib.stride_tricks.as_strided(x, shape=(rows, cols), strides=(8, ld * 8))
    return np.lib.stride_tricks.as_strided(x, shape=(rows, cols), strides=(ld * 8, 8))
