"""Oracle: L-mode products ``(M_d (x) ... (x) M_1) x`` (``operators/utils/matrix_free.py:7-119``).

Pinned by: the exact integer known answer of ``tests/test_kron_product.py:145-165``
(``apply([A, B, C], v) == kron(C, kron(B, A)) @ v``), and brute-force ``np.kron`` on
random dense/sparse factors of every shape combination (tests/test_oracle_matrix_free.py).
"""
from __future__ import annotations

from typing import Sequence

import numpy as np


def mode_product(tensor: np.ndarray, matrix, mode: int, transpose: bool = False) -> np.ndarray:
    """n-mode product along ``mode`` (``matrix_free.py:39-75``): move the mode axis to
    the front, multiply as a 2-D matrix, move it back."""
    mat = matrix.T if transpose else matrix
    shp = tensor.shape
    moved = np.moveaxis(tensor, mode, 0).reshape(shp[mode], -1)
    out = np.asarray(mat @ moved)
    rest = [shp[i] for i in range(len(shp)) if i != mode]
    return np.moveaxis(out.reshape([mat.shape[0]] + rest), 0, mode)


def apply(matrix_list: Sequence, array_in: np.ndarray, is_transpose: bool = False) -> np.ndarray:
    """``MatrixFree.apply`` (``matrix_free.py:77-119``).  The list is given as
    ``[M_1, ..., M_d]``; reversing it makes the C-order reshape put direction 1 on the
    last (fastest) axis, so mode ``i`` of the reversed list is direction ``d - i``.
    Trailing batch dimensions of ``array_in`` are carried along."""
    mats = list(matrix_list)[::-1]
    shape = [m.shape[0 if is_transpose else 1] for m in mats]
    shape.extend(array_in.shape[1:] or (1,))
    tensor = np.reshape(array_in, shape)
    for i, m in enumerate(mats):
        tensor = mode_product(tensor, m, i, transpose=is_transpose)
    if array_in.ndim == 1:
        return tensor.ravel()
    return tensor.reshape((-1,) + array_in.shape[1:])
