"""Oracle: fast-diagonalization (FD) preconditioner.

Spatial FD (``fastdiag/utils/space_precond.py:104-320``): per direction the Gauss
pencil ``(K_d, M_d) = (w1 B1, w0 B0)`` trimmed by the Dirichlet table, ``+1e-8 M``
when untrimmed, ``U_d, lambda_d = eigh(K_d, M_d)``; eigenvalues combined as
``c0 * m_corr + c1 * sum_d s_d lambda_d`` (Kronecker sum) and clipped at
``1e-14 * max``; apply ``y[free] = U (Lambda^{-1} o U^T x[free])``, zero elsewhere.

Space-time arrowhead FD (``fastdiag/utils/time_arrow.py:44-192``,
``fastdiag/utils/operations.py:5-93``, ``fastdiag/utils/time_template.py:46-71``):
the time pencil (advection ``w0 B1``, mass ``w0 B0``) with the first time DOF removed,
complex ``eig`` of its leading block plus an M-normalised last column gives ``U_t`` with
``U_t^H W_t U_t = D_t`` arrowhead; apply = complex Kronecker transform, arrowhead solve
per spatial eigenvalue through the Schur complement, back-transform, real part.

Legacy QZ space-time FD (``fastdiag/utils/time_legacy.py:16-171``): the complex QZ
decomposition ``W_t = Q S Z^H``, ``M_t = Q T Z^H`` of the same time pencil; apply =
``(U_s^T (x) Q^H)`` transform, per spatial eigenvalue the upper-triangular solve with
``adv S + lambda T``, ``(U_s (x) Z)`` back-transform, real part (transpose: ``Z^T``,
the transposed triangular solve, ``conj(Q)``).

Pinned by (tests/test_oracle_fastdiag.py): dense ``kron(U)`` reconstruction
(``tests/test_tensor_operations.py:130-156``), eigenvalues against an independent
``eigh`` + Kronecker sum (``tests/test_fast_diagonalization.py:20-119``), the exact
inverse of the separable stiffness on the Gauss cube (direct sparse solve), and for
the arrowhead: ``U_t^H M_t U_t = I``, ``U_t^H W_t U_t = D_t`` and a dense solve of the
assembled space-time system; for the QZ form the factorisation identities and the same
exact inverse of the separable space-time operator (apply and transpose).
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np
import scipy.linalg as sla

from . import matrix_free as mf

TINY = 1e-8
SAFEGUARD = 1e-14


def gauss_pencils(gauss_rules):
    """Dense 1-D mass and stiffness of each direction (``space_precond.py:139-145``)."""
    M = [(q.weights[0] @ q.basis[0]).toarray() for q in gauss_rules]
    K = [(q.weights[-1] @ q.basis[-1]).toarray() for q in gauss_rules]
    return M, K


class SpaceFD:
    """``SpaceFD`` for one DOF per node (``nbDoFsPerNode = 1``)."""

    def __init__(self, gauss_rules, table_dirichlet):
        table = np.asarray(table_dirichlet, dtype=bool)
        if table.ndim == 3:
            table = table[0]
        M, K = gauss_pencils(gauss_rules)
        self.nnz = [m.shape[0] for m in M]
        self.indices: List[np.ndarray] = []
        self.eigvecs: List[np.ndarray] = []
        self.eigvals: List[np.ndarray] = []
        for d in range(len(gauss_rules)):
            lo = 1 if table[d, 0] else 0
            hi = self.nnz[d] - (1 if table[d, 1] else 0)
            idx = np.arange(lo, hi)
            Md = M[d][np.ix_(idx, idx)]
            Kd = K[d][np.ix_(idx, idx)]
            if len(idx) == self.nnz[d]:
                Kd = Kd + TINY * Md
            lam, U = sla.eigh(Kd, Md)
            self.indices.append(idx)
            self.eigvecs.append(np.real(U))
            self.eigvals.append(np.real(lam))
        self.stiff_corr = np.ones(len(gauss_rules))
        self.mass_corr = 1.0
        self.scalar_coefs = (0.0, 1.0)

    def free_nodes(self) -> np.ndarray:
        """Flat indices of the free box, direction 1 fastest (``space_precond.py:104-122``)."""
        grids = np.meshgrid(*[idx for idx in self.indices[::-1]], indexing="ij")
        flat = np.zeros_like(grids[0])
        stride = 1
        for d in range(len(self.nnz)):
            flat = flat + grids[len(self.nnz) - 1 - d] * stride
            stride *= self.nnz[d]
        return flat.ravel()

    def eigenvalues(self) -> np.ndarray:
        """``update_space_eigenvalues`` (``space_precond.py:195-236``)."""
        c0, c1 = self.scalar_coefs
        nd = len(self.eigvals)
        total = None
        for i in range(nd):
            parts = [np.ones_like(l) for l in self.eigvals]
            parts[i] = self.eigvals[i] * self.stiff_corr[i]
            t = parts[0]
            for q in parts[1:]:
                t = np.kron(q, t)
            total = t if total is None else total + t
        lam = c0 * self.mass_corr * np.ones_like(total) + c1 * total
        return np.maximum(lam, SAFEGUARD * np.max(lam))

    def apply(self, x: np.ndarray) -> np.ndarray:
        """``space_precond.py:273-320`` (all-zero columns are left zero)."""
        free = self.free_nodes()
        lam = self.eigenvalues()
        out = np.zeros_like(x, dtype=float)
        xin = x[free]
        if xin.ndim == 1:
            if not np.any(xin != 0.0):
                return out
            t = mf.apply(self.eigvecs, xin, is_transpose=True) / lam
            out[free] = mf.apply(self.eigvecs, t)
            return out
        active = np.any(xin != 0.0, axis=0)
        if not np.any(active):
            return out
        t = mf.apply(self.eigvecs, xin[:, active], is_transpose=True) / lam[:, None]
        out[np.ix_(free, active)] = mf.apply(self.eigvecs, t)
        return out


# --- space-time arrowhead -----------------------------------------------------


def time_pencil(gauss_time):
    """Advection ``w0 B1`` and mass ``w0 B0`` of the time direction with the first time
    DOF removed (``time_template.py:46-59``)."""
    W = (gauss_time.weights[1] @ gauss_time.basis[1]).toarray()
    M = (gauss_time.weights[0] @ gauss_time.basis[0]).toarray()
    idx = np.arange(1, gauss_time.nbctrlpts)
    return W[np.ix_(idx, idx)], M[np.ix_(idx, idx)]


def arrowhead_decomposition(Wt: np.ndarray, Mt: np.ndarray):
    """``time_arrow.py:44-80``: returns ``U_t``, the diagonal of ``D_t`` and its border ``g``."""
    Wp, Mp = Wt[:-1, :-1], Mt[:-1, :-1]
    m = Mt[:-1, -1]
    w = Wt[:-1, -1]
    lam, V = sla.eig(Wp, Mp)
    for i in range(V.shape[1]):
        v = V[:, i].copy()
        V[:, i] /= np.sqrt(np.abs(v.conj() @ Mp @ v))
    v = np.linalg.solve(Mp, -m)
    v1 = np.concatenate([v, [1.0]])
    krho = v1 / np.sqrt(v1 @ (Mt @ v1))
    k = krho[:-1]
    n = Wt.shape[0]
    Ut = np.zeros((n, n), dtype=complex)
    Ut[:-1, :-1] = V
    Ut[:-1, -1] = k
    Ut[-1, -1] = krho[-1]
    g = V.conj().T @ (np.hstack([Wp, w[:, None]]) @ krho)
    sigma = krho @ (Wt @ krho)
    diag = np.concatenate([lam, [sigma]])
    return Ut, diag, g


def schur_complement(H: np.ndarray, b: np.ndarray) -> np.ndarray:
    """``operations.py:5-32``."""
    return H[-1] + np.sum((np.abs(b) ** 2)[:, None] / H[:-1], axis=0)


def solve_arrowhead(H: np.ndarray, b: np.ndarray, rhs: np.ndarray, S=None) -> np.ndarray:
    """``operations.py:35-93`` for one or several right-hand sides."""
    one = rhs.ndim == 1
    r = 1 if one else rhs.shape[1]
    R = rhs.reshape(H.shape + (r,))
    y = R[:-1] / H[:-1, :, None]
    if S is None:
        S = schur_complement(H, b)
    last = (R[-1] + np.tensordot(np.conj(b), y, axes=(0, 0))) / S[:, None]
    top = y - (b[:, None] / H[:-1])[..., None] * last[None]
    x = np.concatenate([top, last[None]], axis=0)
    return x.ravel() if one else x.reshape(-1, r)


class SpaceTimeFD:
    """Arrowhead space-time FD (``time_arrow.py:16-192``) for one DOF per node."""

    def __init__(self, space: SpaceFD, gauss_time):
        self.space = space
        Wt, Mt = time_pencil(gauss_time)
        self.Ut, self.Dt, self.g = arrowhead_decomposition(Wt, Mt)
        self.nt = gauss_time.nbctrlpts
        self.adv_corr = 1.0

    def free_nodes(self) -> np.ndarray:
        sp_free = self.space.free_nodes()
        nsp = int(np.prod(self.space.nnz))
        t_idx = np.arange(1, self.nt)
        return (t_idx[:, None] * nsp + sp_free[None, :]).ravel()

    def _blocks(self, transpose: bool):
        g = -self.g.conj() if transpose else self.g
        b = self.adv_corr * g
        H = np.add.outer(self.adv_corr * self.Dt, self.space.eigenvalues())
        return H, b, schur_complement(H, b)

    def apply(self, x: np.ndarray) -> np.ndarray:
        """``time_arrow.py:111-146``."""
        free = self.free_nodes()
        out = np.zeros_like(x, dtype=float)
        xin = x[free]
        if np.linalg.norm(xin) == 0.0:
            return out
        H, b, S = self._blocks(False)
        t = mf.apply(self.space.eigvecs + [self.Ut.conj()], xin.astype(complex), is_transpose=True)
        t = solve_arrowhead(H, b, t, S)
        out[free] = np.real(mf.apply(self.space.eigvecs + [self.Ut], t))
        return out

    def apply_transpose(self, x: np.ndarray) -> np.ndarray:
        """``time_arrow.py:148-192``."""
        free = self.free_nodes()
        out = np.zeros_like(x, dtype=float)
        xin = x[free]
        if not np.any(xin != 0.0):
            return out
        H, b, S = self._blocks(True)
        t = mf.apply(self.space.eigvecs + [self.Ut], xin.astype(complex), is_transpose=True)
        t = solve_arrowhead(H, b, t, S)
        out[free] = np.real(mf.apply(self.space.eigvecs + [self.Ut.conj()], t))
        return out


class SpaceTimeQZFD(SpaceTimeFD):
    """Legacy QZ space-time FD (``time_legacy.py:16-171``) for one DOF per node."""

    def __init__(self, space: SpaceFD, gauss_time):
        super().__init__(space, gauss_time)
        Wt, Mt = time_pencil(gauss_time)
        # time_legacy.py:42-48
        self.S, self.T, self.Q, self.Z = sla.qz(Wt, Mt, output="complex")

    def _solve(self, t: np.ndarray, transpose: bool) -> np.ndarray:
        """Per spatial eigenvalue ``(adv S + lambda T) y = t`` (transposed: lower solve
        with the transpose, ``time_legacy.py:97-101, 152-156``); ``t`` is
        ``(n_t - 1, n_space)`` with time slowest."""
        lam = self.space.eigenvalues()
        out = np.empty_like(t)
        for j, lj in enumerate(lam):
            mat = self.adv_corr * self.S + lj * self.T
            if transpose:
                out[:, j] = sla.solve_triangular(mat.T, t[:, j], lower=True)
            else:
                out[:, j] = sla.solve_triangular(mat, t[:, j], lower=False)
        return out

    def apply(self, x: np.ndarray) -> np.ndarray:
        """``time_legacy.py:62-108``."""
        free = self.free_nodes()
        out = np.zeros_like(x, dtype=float)
        xin = x[free]
        if not np.any(xin != 0.0):
            return out
        t = mf.apply(self.space.eigvecs + [self.Q.conj()], xin.astype(complex), is_transpose=True)
        t = self._solve(t.reshape(self.nt - 1, -1), False).ravel()
        out[free] = np.real(mf.apply(self.space.eigvecs + [self.Z], t))
        return out

    def apply_transpose(self, x: np.ndarray) -> np.ndarray:
        """``time_legacy.py:110-171``."""
        free = self.free_nodes()
        out = np.zeros_like(x, dtype=float)
        xin = x[free]
        if not np.any(xin != 0.0):
            return out
        t = mf.apply(self.space.eigvecs + [self.Z], xin.astype(complex), is_transpose=True)
        t = self._solve(t.reshape(self.nt - 1, -1), True).ravel()
        out[free] = np.real(mf.apply(self.space.eigvecs + [self.Q.conj()], t))
        return out
