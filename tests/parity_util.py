"""Per-entry error measures of the device-vs-oracle comparisons.

Every entry is judged relative to its own magnitude.  An entry far below the other
entries of its 3x3 block is the result of cancellation, and its rounding error is set by
the block's magnitude, not by its own; such entries are judged against a floor of
``floor`` times the block's largest entry.  With floor 1e-6 and a bar of 1e-10, an entry at
the floor may carry a rounding error of 1e-16 of its block.

A gradient or matrix entry is a sum of terms from inertia, elements, contacts and friction,
which can cancel to far below every one of them, and each element's term is itself a sum
of products that can cancel.  The oracle reports, per entry, the magnitude of what was
summed into it (``grad_scale``: per term the magnitudes of the products inside an element
force, or the length of a contact or inertia vector; ``val_scale``: per block n times the
largest entry of the element or contact matrix of order n, a bound of its 2-norm).  A sum formed in another order, or with fused
multiply-adds, differs by up to a few hundred units in the last place of that magnitude.
An entry is judged relative to its own magnitude, floored at ``scale_floor`` of that scale:
1e-4 for the gradient (with a bar of 1e-10, an entry that cancels may differ by 1e-14 of
what was summed) and 1e-3 for the matrix, whose element and contact blocks come out of
eigendecompositions (the PSD projections, computed by different methods on the device and
in the oracle) whose error is a small multiple of the unit roundoff times the block's
2-norm (1e-13 of the scale).
"""
import numpy as np


def csr_columns(outer):
    outer = np.asarray(outer, np.int64)
    return np.repeat(np.arange(len(outer) - 1), np.diff(outer))


def csr_entry_err(outer, inner, a, b, floor=1e-6, scale=None, scale_floor=1e-3):
    """max over the CSR entries of |a - b| / max(|b|, floor * blockmax), where blockmax is
    the largest |b| in the entry's 3x3 block (vertex row, vertex column); with ``scale``
    (per entry, the summed magnitude of the blocks added into it) the floor is
    scale_floor * scale instead."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    if len(b) == 0:
        return 0.0
    if scale is not None:
        den = np.maximum(np.abs(b), scale_floor * np.asarray(scale, np.float64))
        den = np.where(den > 0, den, 1.0)
        return float((np.abs(a - b) / den).max())
    col = csr_columns(outer)
    row = np.asarray(inner, np.int64)
    nb = (int(col.max()) // 3 + 1) if len(col) else 1
    key = (row // 3) * nb + col // 3
    uk, inv = np.unique(key, return_inverse=True)
    bmax = np.zeros(len(uk))
    np.maximum.at(bmax, inv, np.abs(b))
    den = np.maximum(np.abs(b), floor * bmax[inv])
    den = np.where(den > 0, den, 1.0)
    return float((np.abs(a - b) / den).max())


def vec3_entry_err(a, b, floor=1e-6, scale=None, scale_floor=1e-4):
    """Per-entry relative error of (n,3) arrays with the floor taken per row, or, with
    ``scale`` (the per-entry magnitude of the terms summed into b), per entry from it."""
    a = np.asarray(a, np.float64).reshape(-1, 3)
    b = np.asarray(b, np.float64).reshape(-1, 3)
    if b.size == 0:
        return 0.0
    if scale is not None:
        low = scale_floor * np.asarray(scale, np.float64).reshape(-1, 3)
    else:
        low = floor * np.abs(b).max(axis=1, keepdims=True)
    den = np.maximum(np.abs(b), low)
    den = np.where(den > 0, den, 1.0)
    err = np.abs(a - b) / den
    return float(err.max())


def entry_err(a, b, floor=1e-6):
    """Per-entry relative error of two arrays, the floor taken from the largest |b|."""
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    if b.size == 0:
        return 0.0
    den = np.maximum(np.abs(b), floor * np.abs(b).max())
    den = np.where(den > 0, den, 1.0)
    return float((np.abs(a - b) / den).max())
