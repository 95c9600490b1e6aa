"""Oracle: randomized low-rank preconditioners for ``M = mu I + S``
(``solvers/sketch_preconditioner.py:10-90``, ``solvers/sketching/template.py:20-224``,
``solvers/sketching/operations.py:13-250``, ``spd_preconditioner.py:8-55``,
``general_preconditioner.py:8-55``).  Test infrastructure only: a numpy restatement
used to check the library, never called by it.

Build: below ``small_threshold`` rows (or when the rank reaches the size) the dense
``mu I + S`` is LU-factored; otherwise a range basis ``Q`` of ``S`` comes from a fixed
Gaussian sketch with power iterations (``rank > 0``) or the adaptive range finder
(``rank <= 0``; Halko, Martinsson & Tropp 2011, Alg. 4.2 with power iterations), then
the SPD form takes ``eigh(Q^T S Q)`` and the general form the SVD of ``Q^T S``.
Apply: SPD ``x + U (((lambda_r + mu) / (Lambda + mu) - 1) o U^T x)`` (Frangella, Tropp &
Udell 2023); general, Woodbury ``x / mu - U R^{-1} V^T x / mu`` with
``R = mu diag(1 / Lambda) + V^T U``.

Pinned by (tests/test_sketch_preconditioner.py): the dense form is the inverse of
``mu I + S``; on an exactly rank-r SPD ``S`` the SPD form gives
``P^{-1} M = (lambda_r + mu) U U^T + mu (I - U U^T)``; on an exactly rank-r general ``S``
the Woodbury form is the exact inverse; the adaptive basis captures the range of a
low-rank ``S``; a planted sign error in the Woodbury correction fails.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import lu_factor, lu_solve, qr

SAFEGUARD = 1e-14


def power_iteration(A, Omega, power_iters=1):
    """``operations.py:13-48``: ``(A A^T)^p A Omega`` with QR between the products."""
    Y = A @ Omega
    for _ in range(power_iters):
        Qa = qr(Y, mode="economic")[0]
        Qb = qr(A.T @ Qa, mode="economic")[0]
        Y = A @ Qb
    return np.asarray(Y)


def gaussian_test_matrix(n, sample_size, rng):
    """``operations.py:161-191`` (not orthonormalised)."""
    return rng.standard_normal((n, sample_size)) / np.sqrt(n)


def adaptive_range_finder(A, n, sample_size, relative_tolerance, rng, power_iters=1, max_rank=None):
    """``operations.py:51-158``."""
    Om = rng.standard_normal((n, sample_size))
    Y = power_iteration(A, Om, power_iters)
    m = Y.shape[0]
    max_rank = int(min(max_rank if max_rank is not None else min(m, n), m, n))
    err = lambda X: np.max(np.linalg.norm(X, axis=0))
    thresh = relative_tolerance * err(Y) / (10 * np.sqrt(2 / np.pi))
    Q = np.empty((m, 0))
    for _ in range(max_rank):
        if err(Y) <= thresh:
            break
        y = Y[:, 0].copy()
        if Q.shape[1]:
            y -= Q @ (Q.T @ y)
        ny = np.linalg.norm(y)
        Om, Y = Om[:, 1:], Y[:, 1:]
        if ny >= SAFEGUARD:
            q = y / ny
            Q = np.column_stack((Q, q))
            if Y.shape[1]:
                Y -= np.outer(q, q @ Y)
        om = rng.standard_normal(n)
        yn = power_iteration(A, om[:, None], power_iters).ravel()
        if Q.shape[1]:
            yn -= Q @ (Q.T @ yn)
        Om = np.column_stack((Om, om))
        Y = np.column_stack((Y, yn))
    return np.asarray(qr(Q, mode="economic")[0])


class _Base:
    def __init__(self, S, mu, rank=24, small_threshold=576, random_state=0):
        self.S, self.mu, self.rank = S, float(mu), int(rank)
        self.small_threshold = int(small_threshold)
        self.n = S.shape[0]
        self.rng = np.random.default_rng(random_state)
        self.oversampling, self.tol, self.power_iters = 10, 1e-2, 1
        self.dense = False

    def range_basis(self):
        """``template.py:136-165``."""
        if self.rank <= 0:
            Q = adaptive_range_finder(self.S, self.n, self.oversampling, self.tol, self.rng, self.power_iters)
            return Q, Q.shape[1]
        r = min(self.rank, *self.S.shape)
        k = min(r + self.oversampling, *self.S.shape)
        Y = power_iteration(self.S, gaussian_test_matrix(self.n, k, self.rng), self.power_iters)
        return np.linalg.qr(Y, mode="reduced")[0], r

    def build(self):
        """``template.py:113-175``."""
        if self.n < self.small_threshold or self.rank >= self.n:
            self.LU = lu_factor(self.S @ np.eye(self.n) + self.mu * np.eye(self.n))
            self.dense = True
        else:
            self.low_rank()
        return self

    def apply(self, x):
        return lu_solve(self.LU, x) if self.dense else self.apply_low_rank(x)


class SPD(_Base):
    def low_rank(self):
        """``spd_preconditioner.py:23-30``, ``operations.py:197-235``."""
        Q, r = self.range_basis()
        B = Q.T @ (self.S @ Q)
        w, E = np.linalg.eigh(0.5 * (B + B.T))
        idx = np.argsort(w)[::-1]
        self.U = (Q @ E[:, idx])[:, :r]
        self.Lam = np.maximum(w[idx], 0.0)[:r]

    def apply_low_rank(self, x):
        """``spd_preconditioner.py:35-51``."""
        d = (self.Lam[-1] + self.mu) / (self.Lam + self.mu) - 1.0
        return x + self.U @ (d * (self.U.T @ x))


class General(_Base):
    def low_rank(self):
        """``general_preconditioner.py:22-30``, ``operations.py:238-250``."""
        Q, r = self.range_basis()
        Uh, sig, VT = np.linalg.svd((self.S.T @ Q).T, full_matrices=False)
        self.U, self.Lam, self.V = (Q @ Uh)[:, :r], sig[:r], VT.T[:, :r]
        self.R = lu_factor(self.mu * np.diag(1.0 / self.Lam) + self.V.T @ self.U)

    def apply_low_rank(self, x):
        """``general_preconditioner.py:35-51``."""
        xm = x / self.mu
        return xm - self.U @ lu_solve(self.R, self.V.T @ xm)
