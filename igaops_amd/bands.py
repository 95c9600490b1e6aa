"""Conversion of the reference's CSR factors into the banded device tables.

``basis[k]`` is ``(n_q x n_c)`` CSR and ``weights[z]`` is ``(n_c x n_q)`` CSR
(``quadrature/utils/operations.py:97-105``, ``weighted_quad.py:353-363``).  On the
device every factor is a band: row ``i`` keeps ``width`` entries starting at column
``first[i]``, zero padded, with ``0 <= first[i]`` and ``first[i] + width <= ncols``.
"""
from __future__ import annotations

import weakref
from typing import Sequence, Tuple

import numpy as np
from scipy import sparse as sp


def _row_extent(m: sp.csr_array) -> Tuple[np.ndarray, np.ndarray]:
    """First and last stored column of every row (rows without entries: (+inf, -inf))."""
    m = sp.csr_array(m)
    m.sort_indices()
    nr = m.shape[0]
    lo = np.full(nr, np.iinfo(np.int64).max)
    hi = np.full(nr, -1)
    ptr, idx = m.indptr, m.indices
    nz = np.diff(ptr) > 0
    lo[nz] = idx[ptr[:-1][nz]]
    hi[nz] = idx[ptr[1:][nz] - 1]
    return lo, hi


def common_band(mats: Sequence) -> Tuple[np.ndarray, int, np.ndarray]:
    """Smallest band holding every stored entry of all ``mats`` (same shape).

    Returns ``first (nrows,) int32``, ``width`` and ``vals (len(mats), nrows, width)``."""
    mats = [sp.csr_array(m) for m in mats]
    nr, ncol = mats[0].shape
    lo = np.full(nr, np.iinfo(np.int64).max)
    hi = np.full(nr, -1)
    for m in mats:
        if m.shape != (nr, ncol):
            raise ValueError("band factors must share their shape")
        l, h = _row_extent(m)
        lo = np.minimum(lo, l)
        hi = np.maximum(hi, h)
    empty = hi < 0
    lo[empty] = 0
    hi[empty] = 0
    width = int(max(1, np.max(hi - lo + 1)))
    first = np.minimum(lo, ncol - width).astype(np.int64)
    first = np.maximum(first, 0)
    vals = np.zeros((len(mats), nr, width))
    for k, m in enumerate(mats):
        coo = m.tocoo()
        j = coo.col - first[coo.row]
        vals[k, coo.row, j] = coo.data
    return first.astype(np.int32), width, vals


def fused_direction_tables(quadrule, P: int = None):
    """Per-point band of the four factors of one direction, as the fused kernel
    wants them: ``first (n_q,)`` and ``tab (4, n_q, P)`` = B0, B1, W0^T, W1^T with
    ``W0 = weights[0]`` (test value) and ``W1 = weights[2]`` (test derivative)."""
    B0, B1 = quadrule.basis[0], quadrule.basis[1]
    W0t = sp.csr_array(quadrule.weights[0].T)
    W1t = sp.csr_array(quadrule.weights[2].T)
    first, width, vals = common_band([B0, B1, W0t, W1t])
    if P is not None and P > width:
        first = np.minimum(first, quadrule.nbctrlpts - P).astype(np.int32)
        first = np.maximum(first, 0).astype(np.int32)
        _, _, vals = _rebanded([B0, B1, W0t, W1t], first, P)
        width = P
    return first, width, vals


def _rebanded(mats, first, width):
    vals = np.zeros((len(mats), mats[0].shape[0], width))
    for k, m in enumerate(mats):
        coo = sp.csr_array(m).tocoo()
        j = coo.col - first[coo.row]
        if np.any(j < 0) or np.any(j >= width):
            raise ValueError("entries outside the requested band")
        vals[k, coo.row, j] = coo.data
    return first, width, vals


def same_matrix(a, b) -> bool:
    """Value equality of two sparse (or dense) matrices."""
    if a is b:
        return True
    if a.shape != b.shape:
        return False
    d = sp.csr_array(a) - sp.csr_array(b)
    return d.nnz == 0 or float(np.max(np.abs(d.data))) == 0.0


_ALPHA_ONLY = weakref.WeakKeyDictionary()


def test_weights_depend_on_alpha_only(quadrule) -> bool:
    """True for rules with ``weights = [W0, W0, W1, W1]`` (Gauss and WQ type "2"),
    the case handled by the fused kernel.  The answer is kept per rule object for as
    long as the rule holds the same weight matrices (rules are not modified in place
    after construction)."""
    w = quadrule.weights
    try:
        hit = _ALPHA_ONLY.get(quadrule)
    except TypeError:  # not weakly referenceable
        hit = None
    if hit is not None and len(hit[0]) == len(w) and all(a is b for a, b in zip(hit[0], w)):
        return hit[1]
    ans = len(w) >= 4 and same_matrix(w[0], w[1]) and same_matrix(w[2], w[3])
    try:
        _ALPHA_ONLY[quadrule] = (tuple(w), ans)
    except TypeError:
        pass
    return ans
