"""Randomized low-rank preconditioners for ``M = mu I + S`` (the Schur-complement
preconditioners of the reference, ``solvers/sketch_preconditioner.py:10-90``).

``RandomPreconditioner(S, mu, random_type, rank, small_threshold, random_state)``
returns a ``RandSPDPreconditioner`` or ``RandGeneralPreconditioner`` with the
reference's interface (``build()``, ``apply(x)``, ``is_built``, ``is_dense``,
``Lambda``, ``low_rank_size``, ``low_rank_ratio``).  The build follows the reference's
steps (``solvers/sketching/template.py:113-175``, ``operations.py:13-250``) on the host,
where the caller's operator ``S`` lives: a dense LU of ``mu I + S`` below
``small_threshold`` rows, otherwise a range basis ``Q`` of ``S`` from a fixed Gaussian
sketch with power iterations (``rank > 0``) or the adaptive range finder (``rank <= 0``,
Halko-Martinsson-Tropp), and the eigen- or singular decomposition of ``S`` projected
on it.  The factors then live in HBM and ``apply`` runs on the device for device
vectors (numpy in, numpy out otherwise): the SPD form
``x + U (d o U^T x)``, ``d = (lambda_r + mu) / (Lambda + mu) - 1``
(``spd_preconditioner.py:35-51``), and the general form (Woodbury)
``x / mu - U R^{-1} V^T x / mu`` with ``R = mu diag(1 / Lambda) + V^T U``
(``general_preconditioner.py:35-51``): two passes over the ``r`` stored vectors
(``iga_dot_many`` then ``iga_axpy_many``), the ``r x r`` part on the device as well.
The dense form factors ``mu I + S`` on the device (rocSOLVER LU through torch).
"""
from __future__ import annotations

from typing import Literal, Optional

import numpy as np
from scipy.linalg import lu_factor, lu_solve, qr

from . import _lib
from . import device as dv
from .common import Constants, validate_entry


def _sketch_products(S, Omega, power_iters: int):
    """``(S S^T)^p S Omega`` with a QR between the products (``operations.py:13-48``)."""
    Y = np.asarray(S @ Omega)
    for _ in range(power_iters):
        Y = np.asarray(S @ qr(np.asarray(S.T @ qr(Y, mode="economic")[0]), mode="economic")[0])
    return Y


class _Template:
    """``solvers/sketching/template.py:20-224``."""

    def __init__(self, S, mu: float, rank: int = 24, small_threshold: int = 576, random_state: Optional[int] = 0):
        self.S = S
        self.mu = float(mu)
        self.rank = int(rank)
        self.small_threshold = int(small_threshold)
        self.opshape = S.shape
        self._built = False
        self._dense = False
        self._rng = np.random.default_rng(random_state)
        self._oversampling = 10
        self._low_rank_tolerance = 1e-2
        self._power_iters = 1
        self._LAMBDA = np.array([])
        self._lu = None
        self._dev = {}

    # -- properties of the reference --
    @property
    def is_built(self):
        return self._built

    @property
    def is_dense(self):
        return self._dense

    @property
    def auto_range_finder(self):
        return self.rank <= 0

    @property
    def should_build_dense(self):
        return self.opshape[0] < self.small_threshold or self.rank >= self.opshape[0]

    @property
    def choosen_method(self):
        return "Adaptive randomized" if self.auto_range_finder else "Full randomized"

    @property
    def Lambda(self):
        return self._LAMBDA

    @property
    def low_rank_size(self) -> int:
        return len(self._LAMBDA)

    @property
    def low_rank_ratio(self) -> float:
        if self.low_rank_size == 0:
            return 1.0
        return self._LAMBDA[-1] / (self._LAMBDA[0] + Constants.SAFEGUARD)

    @property
    def low_rank_tolerance(self) -> float:
        return self._low_rank_tolerance

    @property
    def power_iters(self) -> int:
        return self._power_iters

    # -- build --
    def _adaptive_basis(self):
        """The adaptive range finder (``operations.py:51-158``): columns are taken from
        a pool of sketched vectors, orthogonalised against the basis, and the pool is
        deflated and replenished until its largest column norm falls below the
        threshold set by the first pool."""
        n = self.opshape[0]
        pool_w = self._rng.standard_normal((n, self._oversampling))
        pool = _sketch_products(self.S, pool_w, self._power_iters)
        m = pool.shape[0]
        cap = int(min(m, n))
        colmax = lambda Y: np.max(np.linalg.norm(Y, axis=0))
        thresh = self._low_rank_tolerance * colmax(pool) / (10.0 * np.sqrt(2.0 / np.pi))
        basis = []
        for _ in range(cap):
            if colmax(pool) <= thresh:
                break
            Qm = np.column_stack(basis) if basis else None
            y = pool[:, 0].copy()
            if Qm is not None:
                y -= Qm @ (Qm.T @ y)
            ny = np.linalg.norm(y)
            pool_w, pool = pool_w[:, 1:], pool[:, 1:]
            if ny >= Constants.SAFEGUARD:
                q = y / ny
                basis.append(q)
                Qm = np.column_stack(basis)
                if pool.shape[1]:
                    pool = pool - np.outer(q, q @ pool)
            w = self._rng.standard_normal(n)
            yn = _sketch_products(self.S, w[:, None], self._power_iters).ravel()
            if Qm is not None:
                yn = yn - Qm @ (Qm.T @ yn)
            pool_w = np.column_stack((pool_w, w))
            pool = np.column_stack((pool, yn))
        Q = np.column_stack(basis) if basis else np.empty((m, 0))
        return np.asarray(qr(Q, mode="economic")[0])

    def _range_basis(self):
        """``template.py:136-165``."""
        if self.auto_range_finder:
            Q = self._adaptive_basis()
            return Q, Q.shape[1]
        r = min(self.rank, *self.opshape)
        k = min(r + self._oversampling, *self.opshape)
        Omega = self._rng.standard_normal((self.opshape[0], k)) / np.sqrt(self.opshape[0])
        return np.linalg.qr(_sketch_products(self.S, Omega, self._power_iters), mode="reduced")[0], r

    def build(self):
        """``template.py:167-175``."""
        if self.should_build_dense:
            n = self.opshape[0]
            M = np.asarray(self.S @ np.eye(n)) + self.mu * np.eye(n)
            self._lu = lu_factor(M)
            self._M = M
            self._dense = True
        else:
            self._build_low_rank()
            self._dense = False
        self._built = True
        self._dev = {}
        return self

    # -- device factors --
    def _rows(self, A: np.ndarray):
        """The columns of ``A`` (n x r) as rows of a device array with 16-byte row pitch."""
        t = dv.torch()
        n, r = A.shape
        ld = -(-n // 2) * 2
        R = t.zeros((max(r, 1), ld), dtype=t.float64, device=dv.device())
        if r:
            R[:r, :n].copy_(t.as_tensor(np.ascontiguousarray(A.T), device=dv.device()))
        return R, ld

    def _device(self):
        if not self._dev:
            self._dev = self._device_factors()
        return self._dev

    # -- apply --
    def apply(self, x):
        """``template.py:180-202``: ``P^{-1} x`` for a vector or an ``(n, k)`` array."""
        if not self._built:
            raise RuntimeError("Call build() before applying the preconditioner.")
        dv.require_gpu()
        t = dv.torch()
        was_np = not isinstance(x, t.Tensor)
        xd = dv.to_device(x, "f64")
        cols = xd.reshape(xd.shape[0], -1)
        if self._dense:
            d = self._device()
            y = t.linalg.lu_solve(d["LU"], d["piv"], cols.contiguous())
        else:
            y = t.empty_like(cols)
            for j in range(cols.shape[1]):
                y[:, j] = self._apply_device(cols[:, j].contiguous())
        y = y.reshape(xd.shape)
        return y.cpu().numpy() if was_np else y

    def _dense_factors(self):
        t = dv.torch()
        LU, piv = t.linalg.lu_factor(t.as_tensor(self._M, device=dv.device()))
        return {"LU": LU, "piv": piv}

    def _dot(self, rows, ld, r, x):
        """``rows[:r] . x`` on the device (r doubles)."""
        t = dv.torch()
        n = x.numel()
        out = t.empty(r, dtype=t.float64, device=x.device)
        work = dv.workspace(_lib.load().iga_reduce_workspace_bytes(n, r))
        _lib.call("iga_dot_many", _lib.IGA_F64, n, r, dv.ptr(rows), ld, dv.ptr(x), dv.ptr(out), dv.ptr(work),
                  dv.stream_ptr())
        return out

    def _axpy(self, rows, ld, r, h, alpha, y):
        """``y += alpha sum_i h_i rows[i]`` on the device."""
        _lib.call("iga_axpy_many", _lib.IGA_F64, y.numel(), r, dv.ptr(rows), ld, dv.ptr(h), float(alpha), dv.ptr(y),
                  dv.stream_ptr())

    def _get_info(self, name):
        if not self._built:
            return f"{name}(not built)"
        if self._dense:
            return f"{name}(dense, shape={self.opshape}, mu={self.mu:.2e})"
        tol = f"{self._low_rank_tolerance:.2e}" if self.auto_range_finder else "not concerned"
        return (f"{name}(sketch {self.choosen_method}, shape={self.opshape}, mu={self.mu:.2e}, "
                f"rank={self.low_rank_size}, spectral_ratio={self.low_rank_ratio:.2e}, tolerance={tol})")

    def __repr__(self):
        return self._get_info(type(self).__name__)


class RandSPDPreconditioner(_Template):
    """``solvers/sketching/spd_preconditioner.py:8-55``."""

    def _build_low_rank(self):
        Q, r = self._range_basis()
        B = Q.T @ np.asarray(self.S @ Q)
        w, E = np.linalg.eigh(0.5 * (B + B.T))  # operations.py:197-235
        order = np.argsort(w)[::-1]
        self._U = (Q @ E[:, order])[:, :r]
        self._LAMBDA = np.maximum(w[order], 0.0)[:r]

    def _device_factors(self):
        if self._dense:
            return self._dense_factors()
        t = dv.torch()
        rows, ld = self._rows(self._U)
        d = (self._LAMBDA[-1] + self.mu) / (self._LAMBDA + self.mu) - 1.0
        return {"U": rows, "ld": ld, "d": t.as_tensor(d, device=dv.device())}

    def _apply_device(self, x):
        dd = self._device()
        r = self.low_rank_size
        y = x.clone()
        if r:
            c = self._dot(dd["U"], dd["ld"], r, x)
            c.mul_(dd["d"])
            self._axpy(dd["U"], dd["ld"], r, c, 1.0, y)
        return y


class RandGeneralPreconditioner(_Template):
    """``solvers/sketching/general_preconditioner.py:8-55``."""

    def _build_low_rank(self):
        Q, r = self._range_basis()
        Uh, sig, VT = np.linalg.svd(np.asarray(self.S.T @ Q).T, full_matrices=False)  # operations.py:238-250
        self._U = (Q @ Uh)[:, :r]
        self._LAMBDA = sig[:r]
        self._V = VT.T[:, :r]

    def _device_factors(self):
        if self._dense:
            return self._dense_factors()
        t = dv.torch()
        r = self.low_rank_size
        U, ldu = self._rows(self._U)
        V, ldv = self._rows(self._V)
        R = self.mu * np.diag(1.0 / self._LAMBDA) + self._V.T @ self._U
        Rinv = lu_solve(lu_factor(R), np.eye(r))
        return {"U": U, "V": V, "ldu": ldu, "ldv": ldv, "Rinv": t.as_tensor(Rinv, device=dv.device())}

    def _apply_device(self, x):
        dd = self._device()
        r = self.low_rank_size
        y = x / self.mu
        if r:
            c = self._dot(dd["V"], dd["ldv"], r, x)
            coeff = (dd["Rinv"] @ c) / self.mu
            self._axpy(dd["U"], dd["ldu"], r, coeff, -1.0, y)
        return y


class RandomPreconditioner:
    """``solvers/sketch_preconditioner.py:10-90``: ``RandomPreconditioner(S, mu,
    random_type, rank=24, small_threshold=576, random_state=None)`` returns the SPD or
    the general form."""

    registry = {"spd": RandSPDPreconditioner, "general": RandGeneralPreconditioner}

    def __new__(cls, S, mu: float, random_type: Literal["spd", "general"], rank: int = 24,
                small_threshold: int = 576, random_state: Optional[int] = None):
        validate_entry(random_type, list(cls.registry))
        return cls.registry[random_type](S, mu, rank, small_threshold, random_state)
