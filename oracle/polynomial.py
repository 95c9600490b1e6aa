"""Oracle: GMRES polynomial preconditioner ``p(A) ~= A^{-1}``.

Restates ``solvers/polynomial/preconditioner.py``: an m-step Arnoldi process with
two classical Gram-Schmidt passes from ``b0`` (default ``default_rng(0)`` normal,
lines 96-108, 160-186), the harmonic Ritz values
``eig(H_m + |h_{m+1,m}|^2 (H_m^*)^{-1} e_m e_m^*)`` (188-225) as the roots of the
GMRES residual polynomial ``r(z) = prod (1 - z / theta)``, Leja ordering (230-300),
and the root-product application ``y += r / theta; r -= A r / theta`` (306-330).  For a
real operator the roots are stored in the operator's real dtype, which keeps only
their real parts (lines 218-223 cast to ``self.dtype``), so every step is a real one.

Pinned by (tests/test_polynomial_preconditioner.py): with degree = n on a diagonal
matrix of distinct eigenvalues the roots are the eigenvalues and ``p(A) = A^{-1}``;
the harmonic Ritz values of an SPD matrix lie in its spectral interval;
``b - A p(A) b = prod (I - A / theta) b`` equals the m-step GMRES residual computed
independently by least squares over the Krylov space; Leja order does not change
``p(A) x``; a dropped root is detected.  Test infrastructure only.
"""
from __future__ import annotations

from typing import Callable, List, Optional, Tuple

import numpy as np


def harmonic_ritz(matvec: Callable, n: int, degree: int, b0: Optional[np.ndarray] = None,
                  breakdown_tol: float = 1e-14) -> Tuple[np.ndarray, int]:
    """(roots, actual_m): ``preconditioner.py:96-108,150-225``, real operators."""
    m = min(int(degree), n)
    if b0 is None:
        b0 = np.random.default_rng(0).standard_normal(n)
    b0 = np.asarray(b0, dtype=float)
    V = np.zeros((n, m + 1))
    H = np.zeros((m + 1, m))
    V[:, 0] = b0 / np.linalg.norm(b0)
    actual_m = m
    for j in range(m):
        w = np.asarray(matvec(V[:, j]), dtype=float).copy()
        for _ in range(2):
            h = V[:, : j + 1].T @ w
            w -= V[:, : j + 1] @ h
            H[: j + 1, j] += h
        beta = np.linalg.norm(w)
        H[j + 1, j] = beta
        if beta <= breakdown_tol * max(1.0, np.linalg.norm(H[: j + 1, j])):
            actual_m = j + 1
            break
        V[:, j + 1] = w / beta
    Hm = H[:actual_m, :actual_m]
    hlast = H[1, 0] if actual_m == 1 else H[actual_m, actual_m - 1]
    e = np.zeros(actual_m)
    e[-1] = 1.0
    f = np.linalg.solve(Hm.T, e)
    roots = np.linalg.eigvals(Hm + abs(hlast) ** 2 * np.outer(f, e))
    return np.real(roots).astype(float), actual_m


def leja_order(roots: np.ndarray) -> List[float]:
    """``preconditioner.py:230-300`` for real roots: largest magnitude first, then the
    root maximising the sum of log distances to the roots already chosen."""
    cand = [float(t) for t in roots]
    chosen: List[float] = []
    while cand:
        if not chosen:
            j = int(np.argmax([abs(t) for t in cand]))
        else:
            scores = [sum(np.log(max(abs(t - q), np.finfo(float).tiny)) for q in chosen) for t in cand]
            j = int(np.argmax(scores))
        chosen.append(cand.pop(j))
    return chosen


def apply(matvec: Callable, steps: List[float], x: np.ndarray) -> np.ndarray:
    """``y = p(A) x`` by the root product (``preconditioner.py:306-330``)."""
    x = np.asarray(x, dtype=float)
    y = np.zeros_like(x)
    r = x.copy()
    for theta in steps:
        Ar = np.asarray(matvec(r), dtype=float)
        y += r / theta
        r -= Ar / theta
    return y
