"""Constants, enums and Kronecker helpers shared by the operator, FD and solver layers.

Mirrors the small host-side utilities of the reference that sit on the hot path:
``Constants`` and ``validate_entry`` (``common/utils.py:5-62``), ``combine_arrays``
(``common/math.py:10-58``) and ``KronProduct.kron_nonzero_indices``
(``common/math.py:63-90``).  Ordering convention everywhere in this package: a
tensor-product index is ``a = a_1 + n_1 (a_2 + n_2 a_3)`` (direction 1 fastest),
the C-order flattening of an array of shape ``(n_3, n_2, n_1)``.
"""
from __future__ import annotations

from enum import Enum
from typing import Any, List, Optional, Sequence

import numpy as np


class Constants:
    """Numerical constants with the values of ``common/utils.py:5-21``."""

    INFTY = 1e14
    HUGE = 1e8
    BIG = 1e4
    SMALL = 1e-4
    TINY = 1e-8
    SAFEGUARD = 1e-14


class ParametricDirection(Enum):
    """Parametric direction labels (``common/utils.py:24-37``)."""

    XI = 0
    ETA = 1
    NU = 2
    TAU = 3
    ALL = -1


def validate_entry(value: str, reference: List[str]) -> Any:
    """Check that a string option is one of ``reference`` (``common/utils.py:53-62``)."""
    if not isinstance(value, str):
        raise TypeError(f"Value's type must be string, got {type(value)}")
    if value not in reference:
        raise ValueError(f"Invalid value. Got '{value}', but must be one of {reference}")
    return value


def kron_vectors(arrays: Sequence[np.ndarray]) -> np.ndarray:
    """``a_n (x) ... (x) a_1`` for 1-D arrays given as ``[a_1, ..., a_n]`` (direction 1 fastest)."""
    out = np.asarray(arrays[0])
    for cur in arrays[1:]:
        out = np.kron(np.asarray(cur), out)
    return out


def combine_arrays(
    arr: Sequence[np.ndarray],
    coefs: Sequence[float],
    brr: Optional[Sequence[np.ndarray]] = None,
) -> np.ndarray:
    """Kronecker sum ``sum_i c_i (b_n (x) .. a_i .. (x) b_1)`` (``common/math.py:10-58``).

    With ``brr`` omitted the ``b_i`` are vectors of ones, so the result is the
    eigenvalue tensor of a separable operator ``sum_i c_i M_n..K_i..M_1``.
    """
    ndim = len(arr)
    if ndim == 0:
        return np.array([])
    if brr is None:
        brr = [np.ones_like(np.asarray(a, dtype=float)) for a in arr]
    total = None
    for i in range(ndim):
        factors = [np.asarray(brr[d], dtype=float) for d in range(ndim)]
        factors[i] = np.asarray(arr[i], dtype=float) * coefs[i]
        term = kron_vectors(factors)
        total = term if total is None else total + term
    return total


def kron_nonzero_indices(indices_list: Sequence[Sequence[int]], nnz_list: Sequence[int]) -> List[int]:
    """Flat indices of the Kronecker product of index sets (``common/math.py:63-90``).

    ``indices_list`` and ``nnz_list`` are given slowest direction first (the reference
    passes reversed lists), and the result is ordered like ``itertools.product``.
    """
    nd = len(nnz_list)
    strides = [int(np.prod(nnz_list[i + 1:], dtype=np.int64)) for i in range(nd)]
    flat = np.zeros(1, dtype=np.int64)
    for idx, stride in zip(indices_list, strides):
        idx = np.asarray(list(idx), dtype=np.int64)
        flat = (flat[:, None] + idx[None, :] * stride).ravel()
    if any(len(i) == 0 for i in indices_list):
        return []
    return flat.tolist()


class KronProduct:
    """Namespace mirroring ``common/math.py:61`` for the hot-path helper."""

    kron_nonzero_indices = staticmethod(kron_nonzero_indices)
