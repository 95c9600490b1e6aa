"""GMRES polynomial preconditioner on the device (``solvers/polynomial/preconditioner.py``).

``p(A) ~= A^{-1}`` is the polynomial ``(1 - r_m(z)) / z`` of the GMRES residual
polynomial ``r_m(z) = prod_i (1 - z / theta_i)`` whose roots are the harmonic Ritz
values of an m-step Arnoldi process.  The Arnoldi vectors live in HBM (two classical
Gram-Schmidt passes per step, as the reference); the small projected matrix, its
eigenvalues and the Leja ordering of the roots are host work on ``m x m`` data.
Applying ``p(A)`` is ``m`` device matvecs with vector updates on the device.

``A`` is a device operator: an object with ``apply(x, out=...)`` (the operators of
this package), a callable ``A(x, out)``, or a matrix (``scipy.sparse`` or dense),
which is moved to the device once.  Float64, real operators; as in the reference,
the roots are kept in the operator's real dtype, i.e. their real parts
(``preconditioner.py:218-223``).
"""
from __future__ import annotations

from typing import List, Optional

import numpy as np

from . import device as dv


def _device_matvec(A):
    """``out = A x`` on device float64 vectors."""
    t = dv.torch()
    if hasattr(A, "apply") and not isinstance(A, np.ndarray):
        def mv(x, out):
            res = A.apply(x, out=out)
            if res is not None and res is not out:
                out.copy_(res.reshape(-1))
        return mv
    if callable(A) and not hasattr(A, "shape"):
        return lambda x, out: A(x, out)
    from scipy import sparse as sp

    if sp.issparse(A):
        M = sp.csr_array(A)
        Md = t.sparse_csr_tensor(t.as_tensor(M.indptr.astype(np.int64)), t.as_tensor(M.indices.astype(np.int64)),
                                 t.as_tensor(M.data.astype(np.float64)), size=M.shape).to(dv.device())
    else:
        Md = dv.to_device(np.asarray(A, dtype=np.float64), "f64")
    return lambda x, out: out.copy_((Md @ x.reshape(-1, 1)).reshape(-1))


def _shape(A, n: Optional[int]):
    if hasattr(A, "shape") and len(getattr(A, "shape")) == 2:
        return int(A.shape[0]), int(A.shape[1])
    if n is None:
        raise ValueError("the size of a callable operator must be given (n=...)")
    return int(n), int(n)


class GMRESPolynomialPreconditioner:
    """``GMRESPolynomialPreconditioner(A, degree=10, b0=None, breakdown_tol=1e-14,
    root_tol=1e-12)`` (``preconditioner.py:7-123``); ``apply_device(x, out)`` on device
    vectors, ``matvec(x)`` / ``@`` on numpy vectors."""

    def __init__(self, A, degree: int = 10, b0: Optional[np.ndarray] = None, breakdown_tol: float = 1e-14,
                 root_tol: float = 1e-12, n: Optional[int] = None):
        dv.require_gpu()
        self.shape = _shape(A, n)
        if self.shape[0] != self.shape[1]:
            raise ValueError("A must be square.")
        if not isinstance(degree, (int, np.integer)) or degree < 1:
            raise ValueError("degree must be a positive integer.")
        self.dtype = np.dtype(np.float64)
        self._mv = _device_matvec(A)
        self.A = A
        self.degree_requested = int(degree)
        self.breakdown_tol = breakdown_tol
        self.root_tol = root_tol
        nn = self.shape[0]
        self.degree = min(self.degree_requested, nn)
        if b0 is None:
            b0 = np.random.default_rng(0).standard_normal(nn)
        b0 = np.asarray(b0) if not isinstance(b0, dv.torch().Tensor) else b0
        if tuple(b0.shape) != (nn,):
            raise ValueError(f"b0 must be a vector of size {nn}.")
        self.roots = self._harmonic_ritz(b0)
        if self.roots.size == 0:
            raise np.linalg.LinAlgError("No harmonic Ritz values were computed.")
        zero_tol = self.root_tol * max(1.0, float(np.max(np.abs(self.roots))))
        if np.any(np.abs(self.roots) <= zero_tol):
            raise np.linalg.LinAlgError("Harmonic Ritz value(s) too close to zero; the polynomial is singular.")
        self._steps = self._leja_steps(self.roots)
        self.degree_actual = len(self.roots)

    def _harmonic_ritz(self, b0) -> np.ndarray:
        """``preconditioner.py:150-225``: Arnoldi on the device, then the harmonic
        Ritz values of the projected matrix on the host."""
        t = dv.torch()
        n, m = self.shape[0], self.degree
        b = dv.to_device(b0, "f64").reshape(-1)
        beta0 = float(t.linalg.vector_norm(b))
        if beta0 == 0:
            raise ValueError("b0 must be nonzero.")
        V = t.zeros((m + 1, n), dtype=t.float64, device=b.device)
        H = np.zeros((m + 1, m))
        V[0] = b / beta0
        w = t.empty(n, dtype=t.float64, device=b.device)
        actual_m = m
        for j in range(m):
            self._mv(V[j], w)
            for _ in range(2):
                h = V[: j + 1] @ w
                w -= V[: j + 1].T @ h
                H[: j + 1, j] += h.cpu().numpy()
            beta = float(t.linalg.vector_norm(w))
            H[j + 1, j] = beta
            if beta <= self.breakdown_tol * max(1.0, float(np.linalg.norm(H[: j + 1, j]))):
                actual_m = j + 1
                break
            V[j + 1] = w / beta
        Hm = H[:actual_m, :actual_m]
        hlast = H[1, 0] if actual_m == 1 else H[actual_m, actual_m - 1]
        e = np.zeros(actual_m)
        e[-1] = 1.0
        try:
            f = np.linalg.solve(Hm.T, e)
        except np.linalg.LinAlgError as exc:
            raise np.linalg.LinAlgError(
                "The projected Arnoldi matrix H_m is singular or numerically singular, so harmonic Ritz "
                "values could not be computed.") from exc
        roots = np.linalg.eigvals(Hm + abs(hlast) ** 2 * np.outer(f, e))
        return np.real(roots).astype(np.float64)

    @staticmethod
    def _leja_steps(roots: np.ndarray) -> List[float]:
        """``preconditioner.py:230-300`` (real roots)."""
        cand = [float(th) for th in roots]
        chosen: List[float] = []
        while cand:
            if not chosen:
                j = int(np.argmax([abs(th) for th in cand]))
            else:
                scores = [sum(np.log(max(abs(th - q), np.finfo(float).tiny)) for q in chosen) for th in cand]
                j = int(np.argmax(scores))
            chosen.append(cand.pop(j))
        return chosen

    def apply_device(self, x, out=None):
        """``out = p(A) x`` (``preconditioner.py:306-366``), float64 device vectors."""
        t = dv.torch()
        xv = x.reshape(-1)
        if xv.numel() != self.shape[1]:
            raise ValueError(f"Expected vector of size {self.shape[1]}, got {xv.numel()}.")
        y = out.reshape(-1) if out is not None else t.empty_like(xv)
        y.zero_()
        r = xv.clone()
        Ar = t.empty_like(r)
        for theta in self._steps:
            self._mv(r, Ar)
            y.add_(r, alpha=1.0 / theta)
            r.sub_(Ar, alpha=1.0 / theta)
        return y if out is None else out

    def matvec(self, x):
        t = dv.torch()
        if isinstance(x, t.Tensor):
            return self.apply_device(x.to(dtype=t.float64))
        return self.apply_device(dv.to_device(np.asarray(x, dtype=np.float64), "f64")).cpu().numpy()

    def __matmul__(self, x):
        return self.matvec(x)

    def __call__(self, x, out=None):
        return self.apply_device(x, out)
