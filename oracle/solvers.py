"""Oracle: the Krylov drivers of the path.

* ``gmres`` restates ``LinearSolver.solve`` + ``LinearSolver._GMRES``
  (``solvers/linear_solver.py:149-257,438-525``): Dirichlet masking of the operands,
  un-restarted right-preconditioned GMRES, modified Gram-Schmidt with a second full
  pass, a fresh least-squares solve of the whole Hessenberg each iteration, ``x``
  rebuilt from the preconditioned vectors, and the any-of stopping rule of
  ``ConvergenceManager`` (relative LS residual < tol, absolute LS residual < 1e-14,
  relative increment ``||dx|| / (||x|| + 1e-14)`` < tol;
  ``solvers/utils/convergence_manager.py:92-327``).
* ``GCRODR`` restates ``solvers/deflated/gcrodr.py:188-373`` with ``gmres1``,
  ``gmres2``, ``getHarmVecs1``, ``getHarmVecs2`` (``solvers/deflated/helpers.py:13-265``).

Pinned by (tests/test_oracle_solvers.py): agreement with a direct sparse solve
(``tests/test_fast_diagonalization.py:239-251``), the tridiagonal sequence of
``tests/test_gcrodr.py`` (convergence, match with GMRES to 1e-4, true residual
<= 10 tol), and the GCRODR invariants ``C^H C = I``, ``C = M A U``.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Dict, Optional

import numpy as np
import scipy.linalg as sla

SAFEGUARD = 1e-14


@dataclass
class Output:
    solution: np.ndarray
    residual: np.ndarray
    success: bool
    iterations: int = 0
    reason: str = ""


def _converged(tol, rel=None, absv=None, incr=None):
    """Any-of rule (``convergence_manager.py:281-327``); thresholds are ``max(ref, 1e-14)``."""
    thr = max(tol, SAFEGUARD)
    if absv is not None and absv < SAFEGUARD:
        return "absolute"
    if rel is not None and rel < thr:
        return "relative"
    if incr is not None and incr < thr:
        return "increment"
    return ""


def gmres(A: Callable, b: np.ndarray, P: Optional[Callable] = None, tol=1e-10, maxiters=100, cleandod=None, x0=None):
    """``LinearSolver(tol, maxiters, cleandod).solve(A, b, P, x0)`` with ``linear_type='gmres'``."""

    def mask(v):
        if cleandod is not None:
            v[cleandod] = 0.0

    def matvec(v):
        mask(v)
        y = np.asarray(A(v))
        mask(y)
        return y

    P = P if P is not None else (lambda v: v)

    def precond(v):
        mask(v)
        y = np.asarray(P(v))
        mask(y)
        return y

    b = np.array(b, dtype=float, copy=True)
    mask(b)
    if x0 is not None:
        x0 = np.array(x0, dtype=float, copy=True)
        mask(x0)
        b -= matvec(x0)
    out = _gmres_core(matvec, b, precond, tol, maxiters)
    if x0 is not None:
        out.solution = out.solution + x0
    return out


def _gmres_core(A, b, P, tol, maxiters):
    """``linear_solver.py:438-525``."""
    n = len(b)
    x = np.zeros_like(b)
    norm0 = float(np.linalg.norm(b))
    if norm0 < SAFEGUARD:
        return Output(x, np.array([1.0]), True, 0, "absolute")
    H = np.zeros((maxiters + 1, maxiters))
    V = np.zeros((maxiters + 1, n))
    Pv = np.zeros_like(V)
    V[0] = b / norm0
    res = [1.0]
    for it in range(1, maxiters + 1):
        p = np.asarray(P(V[it - 1].copy()))
        Pv[it - 1] = p
        w = np.asarray(A(p.copy()), dtype=float).copy()
        for _pass in range(2):
            for j in range(it):
                h = w.dot(V[j])
                H[j, it - 1] += h
                w -= h * V[j]
        nw = np.linalg.norm(w)
        H[it, it - 1] = nw
        if nw != 0:
            V[it] = w / nw
        mat = H[: it + 1, :it]
        rhs = np.zeros(it + 1)
        rhs[0] = norm0
        y = np.linalg.lstsq(mat, rhs, rcond=None)[0]
        xold = x
        x = Pv[: len(y)].T @ y
        incr = float(np.linalg.norm(x - xold) / (np.linalg.norm(x) + SAFEGUARD))
        n1 = float(np.linalg.norm(rhs - mat @ y))
        res.append(n1 / norm0)
        why = _converged(tol, rel=res[-1], absv=n1, incr=incr)
        if why:
            return Output(x, np.asarray(res), True, it, why)
    return Output(x, np.asarray(res), False, maxiters, "maxiters")


# --- GCRODR -------------------------------------------------------------------


def gmres1(A, x, r, m, M, tol):
    """``helpers.py:13-74``."""
    n = len(r)
    V = np.zeros((m + 1, n), dtype=complex)
    H = np.zeros((m + 1, m), dtype=complex)
    resvec = np.zeros(m)
    k = 1
    beta = np.linalg.norm(r)
    y = np.zeros(1, dtype=complex)
    res = np.array([beta], dtype=complex)
    if beta <= SAFEGUARD:
        return x, r, V, H, k, resvec[:k]
    V[0] = r / beta
    for k in range(1, m + 1):
        w = M(A(V[k - 1]))
        for _pass in range(2):
            for j in range(k):
                h = V[j].conj().dot(w)
                H[j, k - 1] += h
                w = w - h * V[j]
        nw = np.linalg.norm(w)
        H[k, k - 1] = nw
        if nw > 0:
            V[k] = w / nw
        mat = H[: k + 1, :k]
        rhs = np.zeros(k + 1, dtype=complex)
        rhs[0] = beta
        y = np.linalg.lstsq(mat, rhs, rcond=None)[0]
        res = rhs - mat @ y
        resvec[k - 1] = np.linalg.norm(res)
        if resvec[k - 1] < tol:
            break
    x = x + V[:k].T @ y
    r = V[: k + 1].T @ res
    return x, r, V, H, k, resvec[:k]


def gmres2(A, r, m, M, C, tol):
    """``helpers.py:82-146``: Arnoldi deflated against the columns of ``C``."""
    n = len(r)
    V = np.zeros((m + 1, n), dtype=complex)
    H = np.zeros((m + 1, m), dtype=complex)
    B = np.zeros((C.shape[1], m), dtype=complex)
    resvec = np.zeros(m)
    k = 1
    beta = np.linalg.norm(r)
    if beta <= SAFEGUARD:
        return V, H, B, k, resvec[:k]
    V[0] = r / beta
    for k in range(1, m + 1):
        w = M(A(V[k - 1]))
        B[:, k - 1] = C.conj().T @ w
        w = w - C @ B[:, k - 1]
        for _pass in range(2):
            for j in range(k):
                h = V[j].conj().dot(w)
                H[j, k - 1] += h
                w = w - h * V[j]
        nw = np.linalg.norm(w)
        H[k, k - 1] = nw
        if nw > 0:
            V[k] = w / nw
        mat = H[: k + 1, :k]
        rhs = np.zeros(k + 1, dtype=complex)
        rhs[0] = beta
        y = np.linalg.lstsq(mat, rhs, rcond=None)[0]
        resvec[k - 1] = np.linalg.norm(rhs - mat @ y)
        if resvec[k - 1] < tol:
            break
    return V, H, B, k, resvec[:k]


def harmonic_vectors_1(m, k, H):
    """``helpers.py:154-196``: the k harmonic Ritz vectors of smallest modulus."""
    Hm = H[:m, :m]
    em = np.zeros(m, dtype=complex)
    em[-1] = 1.0
    f = np.linalg.solve(Hm.conj().T, em)
    G = Hm + np.abs(H[m, m - 1]) ** 2 * np.outer(f, em.conj())
    vals, vecs = sla.eig(G)
    good = np.isfinite(vals)
    vals, vecs = vals[good], vecs[:, good]
    return vecs[:, np.argsort(np.abs(vals))[:k]]


def harmonic_vectors_2(m, k, G, V, U, C):
    """``helpers.py:204-265``: generalised harmonic Ritz problem of the augmented space."""
    p = m - k
    Vt = np.hstack([U, V[:p].T])
    Wt = np.hstack([C, V[: p + 1].T])
    Phi = Wt.conj().T @ Vt
    vals, vecs = sla.eig(G.conj().T @ G, G.conj().T @ Phi)
    good = np.isfinite(vals)
    vals, vecs = vals[good], vecs[:, good]
    return vecs[:, np.argsort(np.abs(vals))[:k]]


class GCRODR:
    """``solvers/deflated/gcrodr.py:50-373`` (left-preconditioned, complex arithmetic,
    recycled space persisted per ``reuse_name``)."""

    def __init__(self, tolerance, maxiters, krylov_size=40, recycled_size=20, max_cycles=10):
        if krylov_size < recycled_size:
            raise ValueError("Not consistent with the method")
        self.tol, self.maxiters = tolerance, maxiters
        self.m, self.k, self.max_cycles = krylov_size, recycled_size, max_cycles
        self.U_persist: Dict[str, np.ndarray] = {}
        self.last_C = None
        self.last_U = None

    def _check(self, tol, nr, norm0, nmv) -> bool:
        """Absolute, then relative criterion; only when neither holds does the
        iteration counter raise past ``maxiters`` (``convergence_manager.py:297-327``)."""
        if _converged(tol, rel=nr / norm0, absv=nr):
            return True
        if nmv > self.maxiters:
            raise RuntimeError("Maximum iterations exceeded. Algorithm did not converge.")
        return False

    def solve(self, Afun, b, Pfun=None, reuse_name="default") -> Output:
        Pfun = Pfun if Pfun is not None else (lambda v: v)

        def Aop(v):
            v = np.asarray(v)
            if np.iscomplexobj(v):
                return Afun(np.real(v).copy()) + 1j * Afun(np.imag(v).copy())
            return Afun(v)

        def Mop(v):
            v = np.asarray(v)
            if np.iscomplexobj(v):
                return Pfun(np.real(v).copy()) + 1j * Pfun(np.imag(v).copy())
            return Pfun(v)

        m, k, tol = self.m, self.k, self.tol
        b = np.asarray(b, dtype=complex)
        x = np.zeros_like(b)
        r = Mop(b)
        norm0 = float(np.linalg.norm(r))
        resvec = [norm0]

        def rel(v):
            v = np.abs(np.asarray(v))
            return v / max(v[0], SAFEGUARD)

        if norm0 < SAFEGUARD:
            return Output(np.real(x), rel(resvec), True, 0, "absolute")
        nmv = 0
        U = C = None
        if reuse_name in self.U_persist:
            Y = self.U_persist[reuse_name].copy()
            C = np.column_stack([Mop(Aop(Y[:, i])) for i in range(Y.shape[1])])
            C, R = np.linalg.qr(C, mode="reduced")
            U = sla.solve_triangular(R.T, Y.T, lower=True).T
            Cr = C.conj().T @ r
            x = x + U @ Cr
            r = r - C @ Cr
            resvec.append(float(np.linalg.norm(r)))
            p = np.inf
        else:
            x, r, V, H, p, rv = gmres1(Aop, x, r, m, Mop, tol * norm0)
            resvec.extend(rv.tolist())
            nmv += p
            if p >= k:
                P = harmonic_vectors_1(p, k, H)
                Y = V[:p].T @ P
                Q, R = np.linalg.qr(H[: p + 1, :p] @ P, mode="reduced")
                C = V[: p + 1].T @ Q
                U = sla.solve_triangular(R.T, Y.T, lower=True).T
        nr = float(np.linalg.norm(r))
        if self._check(tol, nr, norm0, nmv) or p < m:
            if U is not None:
                self.U_persist[reuse_name] = U
            self.last_C, self.last_U = C, U
            return Output(np.real(x), rel(resvec), True, nmv, "converged-before-cycling")
        success = False
        for cycle in range(self.max_cycles):
            V, Hin, B, p, rv = gmres2(Aop, r, m - k, Mop, C, tol * norm0)
            resvec.extend(rv.tolist())
            nmv += p
            D = np.diag(1.0 / np.sqrt(np.diag(U.conj().T @ U)))
            Ut = U @ D
            Vhat = np.hstack([Ut, V[:p].T])
            What = np.hstack([C, V[: p + 1].T])
            G = np.block([[D, B[:, :p]], [np.zeros((p + 1, k), dtype=complex), Hin[: p + 1, :p]]])
            y = np.linalg.lstsq(G, What.conj().T @ r, rcond=None)[0]
            x = x + Vhat @ y
            r = r - What @ (G @ y)
            nr = float(np.linalg.norm(r))
            if self._check(tol, nr, norm0, nmv) or p < m - k:
                success = True
                break
            P = harmonic_vectors_2(p + k, k, G, V[: p + 1], Ut, C)
            Y = Vhat @ P
            Q, R = np.linalg.qr(G @ P, mode="reduced")
            C = What @ Q
            U = sla.solve_triangular(R.T, Y.T, lower=True).T
        self.U_persist[reuse_name] = U
        self.last_C, self.last_U = C, U
        return Output(np.real(x), rel(resvec), success, nmv, "cycles" if success else "max_cycles")
