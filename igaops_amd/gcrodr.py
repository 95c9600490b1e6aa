"""GCRO with deflated restarting on device vectors.

Follows ``solvers/deflated/gcrodr.py:50-373`` and ``solvers/deflated/helpers.py:13-265``
(Parks et al., SIAM J. Sci. Comput. 28(5), 2006): left preconditioning (the
residual is ``M (b - A x)``, relative to ``||M b||``), complex arithmetic, ``m``-step
cycles with ``k`` recycled vectors, harmonic Ritz extraction, the recycled space ``U``
kept per ``reuse_name`` across calls.

Device layout: a complex vector of length n is a complex128 tensor; viewed as real it
is an (n, 2) float64 array, i.e. two right-hand sides (real, imaginary part).  A real
operator is therefore applied to a complex vector in one launch with ``nrhs = 2``
(the reference applies it to the two parts separately, ``gcrodr.py:147-162``).  Tall
products (``V^H w``, ``W^H V``, ``V y``, ``Y R^{-1}``) go through the one-pass block
kernels ``iga_zgram`` / ``iga_zcombine`` (in an Arnoldi step the projection
coefficients stay on the device between them; one host round trip per step brings
them and the new norm for the least-squares residual); the small dense problems (least squares,
``eig``, QR of ``(p+1) x k`` matrices, Cholesky) run on the host.  The tall QR of the
warm start is CholQR2 (two Cholesky-based passes) instead of Householder: the same
factorisation up to the phases of the columns, which cancel in every later use.
"""
from __future__ import annotations

import ctypes as C
import logging
from typing import Callable, Dict, List, Optional

import numpy as np
import scipy.linalg as sla

from . import _lib
from . import device as dv
from .common import Constants
from .solvers import OutputArgs, converged

logger = logging.getLogger(__name__)


class ZBlas:
    """Complex tall-vector kernels on rows given as lists of complex device tensors."""

    def __init__(self, n: int, max_rows: int = 64):
        self.n = int(n)
        self.work = dv.workspace(_lib.load().iga_zgram_workspace_bytes(self.n, max_rows, max_rows))
        t = dv.torch()
        self._ptrbuf = t.empty(4 * max_rows + 8, dtype=t.int64, device=dv.device())
        self._coefbuf = t.empty(2 * 64 * 64 + 8, dtype=t.float64, device=dv.device())
        self._out = t.empty(2 * 64 * 64 + 8, dtype=t.float64, device=dv.device())
        self._host_out = t.empty(2 * 64 * 64 + 8, dtype=t.float64, pin_memory=True)

    def _ptrs(self, rows, offset):
        t = dv.torch()
        host = t.tensor([r.data_ptr() for r in rows], dtype=t.int64)
        dev = self._ptrbuf[offset:offset + len(rows)]
        dev.copy_(host)
        return dev

    def gram(self, X: List, Y: List) -> np.ndarray:
        """``X^H Y`` as a host complex array (len(X), len(Y))."""
        out = np.zeros((len(X), len(Y)), dtype=complex)
        # chunk rows so that one call stays within the kernel's limits
        for a0 in range(0, len(X), 32):
            for b0 in range(0, len(Y), 32):
                Xa, Yb = X[a0:a0 + 32], Y[b0:b0 + 32]
                px = self._ptrs(Xa, 0)
                py = self._ptrs(Yb, 64)
                na, nb = len(Xa), len(Yb)
                _lib.call("iga_zgram", self.n, na, dv.ptr(px), nb, dv.ptr(py), dv.ptr(self._out),
                          dv.ptr(self.work), dv.stream_ptr())
                h = self._host_out[: 2 * na * nb]
                h.copy_(self._out[: 2 * na * nb])
                v = h.numpy()
                out[a0:a0 + na, b0:b0 + nb] = (v[0::2] + 1j * v[1::2]).reshape(na, nb)
        return out

    def combine(self, V: List, coef: np.ndarray, out: List, beta: float = 0.0):
        """``out_c = beta out_c + sum_j coef[j, c] V_j`` (coef: host complex (len(V), len(out)))."""
        t = dv.torch()
        coef = np.asarray(coef, dtype=complex).reshape(len(V), len(out))
        for c0 in range(0, len(out), 24):
            oc = out[c0:c0 + 24]
            for j0 in range(0, len(V), 64):
                Vj = V[j0:j0 + 64]
                cf = np.ascontiguousarray(coef[j0:j0 + 64, c0:c0 + 24])
                flat = np.empty(cf.size * 2)
                flat[0::2] = cf.real.ravel()
                flat[1::2] = cf.imag.ravel()
                cdev = self._coefbuf[: flat.size]
                cdev.copy_(t.from_numpy(flat))
                pv = self._ptrs(Vj, 0)
                po = self._ptrs(oc, 128)
                b = beta if j0 == 0 else 1.0
                _lib.call("iga_zcombine", self.n, len(Vj), dv.ptr(pv), len(oc), dv.ptr(cdev), 1.0, float(b),
                          dv.ptr(po), dv.stream_ptr())

    # -- device-resident coefficients: rows given by a device array of their addresses
    #    (uploaded once), results and coefficients left on the device --
    def row_ptrs(self, rows):
        """Device int64 array of the rows' addresses (one upload)."""
        t = dv.torch()
        return t.tensor([r.data_ptr() for r in rows], dtype=t.int64).to(dv.device())

    def gram_into(self, px, na: int, py, out):
        """``out[a] = X_a^H y`` for a < na (one right-hand row y, ``py`` its address
        array) into the complex device tensor ``out``."""
        for a0 in range(0, na, 32):
            nc = min(32, na - a0)
            _lib.call("iga_zgram", self.n, nc, dv.ptr(px) + 8 * a0, 1, dv.ptr(py), dv.ptr(out) + 16 * a0,
                      dv.ptr(self.work), dv.stream_ptr())

    def combine_into(self, pv, nv: int, coef, pout, alpha: float = 1.0, beta: float = 0.0):
        """``out = beta out + alpha sum_j coef[j] V_j`` for one output row (``pout``),
        ``coef`` a complex device tensor of nv values."""
        for j0 in range(0, nv, 64):
            nc = min(64, nv - j0)
            _lib.call("iga_zcombine", self.n, nc, dv.ptr(pv) + 8 * j0, 1, dv.ptr(coef) + 16 * j0, float(alpha),
                      float(beta if j0 == 0 else 1.0), dv.ptr(pout), dv.stream_ptr())

    def norm(self, v) -> float:
        return float(np.sqrt(max(self.gram([v], [v])[0, 0].real, 0.0)))

    def scale(self, v, s: float, out=None):
        out = v if out is None else out
        _lib.call("iga_axpby", _lib.IGA_F64, 2 * self.n, float(s), None, 0, dv.ptr(dv.torch().view_as_real(v)), 0.0,
                  dv.ptr(dv.torch().view_as_real(out)), dv.stream_ptr())
        return out


def complex_apply(op: Callable):
    """Apply a real operator ``op(x, out)`` (or a device operator with ``apply_device``)
    to a complex device vector: both parts in one two-right-hand-side launch when the
    operator supports it."""
    t = dv.torch()

    def f(v, out):
        vr = t.view_as_real(v).reshape(-1)
        orr = t.view_as_real(out).reshape(-1)
        if hasattr(op, "apply_device"):
            op.apply_device(vr, out=orr, nrhs=2)
        else:
            re = t.view_as_real(v)[:, 0].contiguous()
            im = t.view_as_real(v)[:, 1].contiguous()
            ro, io = t.empty_like(re), t.empty_like(im)
            op(re, ro)
            op(im, io)
            t.view_as_real(out)[:, 0].copy_(ro)
            t.view_as_real(out)[:, 1].copy_(io)
        return out

    return f


class GCRODR:
    """``solvers/deflated/gcrodr.py:50-373`` with device vectors."""

    def __init__(self, tolerance: float, maxiters: int, krylov_size: int = 40, recycled_size: int = 20,
                 max_cycles: int = 10, verbose: bool = True):
        if tolerance < Constants.SAFEGUARD or tolerance >= 1:
            raise ValueError(f"Relative tolerance must be > 0 and < 1, got {tolerance}")
        if maxiters <= 0:
            raise ValueError("Max iterations must be > 0")
        if krylov_size <= 0:
            raise ValueError("Krylov size should be positive")
        if recycled_size < 0:
            raise ValueError("Recycle size should be greater than or equal to 0")
        if max_cycles <= 0:
            raise ValueError("Max no. cycles should greater than 0")
        if krylov_size < recycled_size:
            raise ValueError("Not consistent with the method")
        self.tol, self.maxiters = tolerance, maxiters
        self.m, self.k, self.max_cycles = krylov_size, recycled_size, max_cycles
        self.verbose = verbose
        self._U_persist: Dict[str, object] = {}
        self.last_C = None
        self.last_U = None

    def reset(self, reuse_name: str = "default"):
        self._U_persist.pop(reuse_name, None)

    def _check(self, nr, norm0, nmv) -> bool:
        if converged(self.tol, rel=nr / norm0, absv=nr):
            return True
        if nmv > self.maxiters:
            raise RuntimeError(f"Maximum iterations ({self.maxiters}) exceeded. Algorithm did not converge.")
        return False

    # -- Arnoldi cycles -------------------------------------------------------
    def _arnoldi(self, A, M, r, m, zb: ZBlas, tol, C_rows=None):
        """Up to m steps; returns V rows, H, B (C^H M A v), steps, residual estimates."""
        t = dv.torch()
        n = zb.n
        V = [t.empty(n, dtype=t.complex128, device=r.device) for _ in range(m + 1)]
        H = np.zeros((m + 1, m), dtype=complex)
        kc = 0 if C_rows is None else len(C_rows)
        B = np.zeros((kc, m), dtype=complex)
        resvec = np.zeros(m)
        beta = zb.norm(r)
        tmp = t.empty(n, dtype=t.complex128, device=r.device)
        w = t.empty(n, dtype=t.complex128, device=r.device)
        y = np.zeros(1, dtype=complex)
        res = np.array([beta], dtype=complex)
        k = 1
        if beta <= Constants.SAFEGUARD:
            return V, H, B, k, resvec[:k], y, res, beta
        zb.scale(r, 1.0 / beta, V[0])
        # the projections stay on the device: each gram writes its coefficients where
        # the combine that subtracts them reads them, and one copy per step brings the
        # step's coefficients and the new norm to the host ([C part | pass 1 | pass 2 | norm])
        pV, pw = zb.row_ptrs(V), zb.row_ptrs([w])
        pC = zb.row_ptrs(C_rows) if kc else None
        coef = t.empty(kc + 2 * m + 1, dtype=t.complex128, device=r.device)
        host = t.empty(kc + 2 * m + 1, dtype=t.complex128, pin_memory=True)
        for k in range(1, m + 1):
            A(V[k - 1], tmp)
            M(tmp, w)
            if kc:
                zb.gram_into(pC, kc, pw, coef[:kc])
                zb.combine_into(pC, kc, coef[:kc], pw, alpha=-1.0, beta=1.0)
            for p in range(2):
                seg = coef[kc + p * m: kc + p * m + k]
                zb.gram_into(pV, k, pw, seg)
                zb.combine_into(pV, k, seg, pw, alpha=-1.0, beta=1.0)
            zb.gram_into(pw, 1, pw, coef[kc + 2 * m:])
            host.copy_(coef, non_blocking=True)
            t.cuda.current_stream().synchronize()
            hv = host.numpy()
            if kc:
                B[:, k - 1] = hv[:kc]
            H[:k, k - 1] += hv[kc:kc + k]
            H[:k, k - 1] += hv[kc + m:kc + m + k]
            nw = float(np.sqrt(max(hv[kc + 2 * m].real, 0.0)))
            H[k, k - 1] = nw
            if nw > 0.0:
                zb.scale(w, 1.0 / nw, V[k])
            mat = H[: k + 1, :k]
            rhs = np.zeros(k + 1, dtype=complex)
            rhs[0] = beta
            y = np.linalg.lstsq(mat, rhs, rcond=None)[0]
            res = rhs - mat @ y
            resvec[k - 1] = np.linalg.norm(res)
            if resvec[k - 1] < tol:
                break
        return V, H, B, k, resvec[:k], y, res, beta

    # -- public solve -----------------------------------------------------------
    def solve_device(self, Afun: Callable, b, Pfun: Optional[Callable] = None, reuse_name: str = "default",
                     ) -> OutputArgs:
        """``Afun(x, out)``/``Pfun(x, out)`` real device operators (or objects with
        ``apply_device``); ``b`` a real float64 device vector."""
        t = dv.torch()
        A = complex_apply(Afun)
        M = complex_apply(Pfun) if Pfun is not None else (lambda v, out: out.copy_(v))
        n = b.numel()
        zb = ZBlas(n)
        m, k, tol = self.m, self.k, self.tol
        bc = b.to(t.complex128)
        x = t.zeros(n, dtype=t.complex128, device=b.device)
        r = t.empty_like(x)
        M(bc, r)
        norm0 = zb.norm(r)
        resvec = [norm0]

        def rel(v):
            v = np.abs(np.asarray(v))
            return v / max(v[0], Constants.SAFEGUARD)

        if converged(tol, absv=norm0):
            return OutputArgs(t.real(x).contiguous(), rel(resvec), True, 0, "absolute")
        nmv = 0
        U = Cm = None
        tmp = t.empty_like(x)
        if reuse_name in self._U_persist:
            Y = [u.clone() for u in self._U_persist[reuse_name]]
            Cm = [t.empty_like(x) for _ in Y]
            for yi, ci in zip(Y, Cm):
                A(yi, tmp)
                M(tmp, ci)
            Cm, R = self._cholqr2(Cm, zb)
            U = self._right_solve(Y, R, zb)
            cr = zb.gram(Cm, [r])[:, 0]
            zb.combine(U, cr[:, None], [x], beta=1.0)
            zb.combine(Cm, -cr[:, None], [r], beta=1.0)
            resvec.append(zb.norm(r))
            p = np.inf
        else:
            V, H, _, p, rv, y, res, beta = self._arnoldi(A, M, r, m, zb, tol * norm0)
            resvec.extend(rv.tolist())
            nmv += p
            if beta > Constants.SAFEGUARD:
                zb.combine(V[:p], y[:, None], [x], beta=1.0)
                zb.combine(V[: p + 1], res[:, None], [r], beta=0.0)
            if p >= k and k > 0:
                P = _harmonic_1(p, k, H)
                Yr = [t.empty_like(x) for _ in range(P.shape[1])]
                zb.combine(V[:p], P, Yr)
                Q, R = np.linalg.qr(H[: p + 1, :p] @ P, mode="reduced")
                Cm = [t.empty_like(x) for _ in range(Q.shape[1])]
                zb.combine(V[: p + 1], Q, Cm)
                U = self._right_solve(Yr, R, zb)
        nr = zb.norm(r)
        if self._check(nr, norm0, nmv) or p < m:
            if U is not None:
                self._U_persist[reuse_name] = U
            self.last_C, self.last_U = Cm, U
            return OutputArgs(t.real(x).contiguous(), rel(resvec), True, nmv, "converged-before-cycling")
        if U is None or Cm is None:
            raise RuntimeError("U and/or C are not initialized.")
        success = False
        for cycle in range(self.max_cycles):
            V, Hin, Bm, p, rv, _, _, _ = self._arnoldi(A, M, r, m - k, zb, tol * norm0, C_rows=Cm)
            resvec.extend(rv.tolist())
            nmv += p
            unorm = np.sqrt(np.real(np.array([zb.gram([u], [u])[0, 0] for u in U])))
            D = np.diag(1.0 / unorm)
            Ut = [t.empty_like(x) for _ in U]
            zb.combine(U, D, Ut)
            Vhat = Ut + V[:p]
            What = Cm + V[: p + 1]
            G = np.block([[D, Bm[:, :p]], [np.zeros((p + 1, k), dtype=complex), Hin[: p + 1, :p]]])
            rhs = zb.gram(What, [r])[:, 0]
            yv = np.linalg.lstsq(G, rhs, rcond=None)[0]
            zb.combine(Vhat, yv[:, None], [x], beta=1.0)
            zb.combine(What, -(G @ yv)[:, None], [r], beta=1.0)
            nr = zb.norm(r)
            if self._check(nr, norm0, nmv) or p < m - k:
                success = True
                break
            Phi = zb.gram(What, Vhat)
            P = _harmonic_2(G, Phi, k)
            Yr = [t.empty_like(x) for _ in range(P.shape[1])]
            zb.combine(Vhat, P, Yr)
            Q, R = np.linalg.qr(G @ P, mode="reduced")
            Cnew = [t.empty_like(x) for _ in range(Q.shape[1])]
            zb.combine(What, Q, Cnew)
            Cm = Cnew
            U = self._right_solve(Yr, R, zb)
        if not success:
            logger.warning("No convergence after %d cycles.", self.max_cycles)
        self._U_persist[reuse_name] = U
        self.last_C, self.last_U = Cm, U
        return OutputArgs(t.real(x).contiguous(), rel(resvec), success, nmv, "cycles" if success else "max_cycles")

    def solve(self, Afun, b, Pfun=None, reuse_name: str = "default") -> OutputArgs:
        """Reference signature (``gcrodr.py:188-216``); numpy in, numpy out."""
        from .solvers import device_callable

        t = dv.torch()
        dv.require_gpu()
        numpy_mode = not isinstance(b, t.Tensor)
        A = Afun if hasattr(Afun, "apply_device") else device_callable(Afun, numpy_mode)
        P = None if Pfun is None else (Pfun if hasattr(Pfun, "apply_device") else device_callable(Pfun, numpy_mode))
        bd = dv.to_device(b, "f64").reshape(-1)
        out = self.solve_device(A, bd, P, reuse_name)
        if numpy_mode:
            out.solution = out.solution.cpu().numpy().reshape(np.shape(b))
        return out

    # -- helpers ----------------------------------------------------------------
    def _cholqr2(self, Cm, zb: ZBlas):
        """Tall QR by two Cholesky passes: returns (Q rows, R)."""
        t = dv.torch()
        Rtot = np.eye(len(Cm), dtype=complex)
        cur = Cm
        for _ in range(2):
            G = zb.gram(cur, cur)
            G = 0.5 * (G + G.conj().T)
            L = np.linalg.cholesky(G)
            R = L.conj().T
            Rinv = sla.solve_triangular(R, np.eye(len(cur)), lower=False)
            nxt = [t.empty_like(c) for c in cur]
            zb.combine(cur, Rinv, nxt)
            cur = nxt
            Rtot = R @ Rtot
        return cur, Rtot

    def _right_solve(self, Y, R, zb: ZBlas):
        """``U = Y R^{-1}`` (``solve_triangular(R.T, Y.T, lower=True).T``)."""
        t = dv.torch()
        Rinv = sla.solve_triangular(R, np.eye(R.shape[0], dtype=complex), lower=False)
        U = [t.empty_like(y) for y in Y]
        zb.combine(Y, Rinv, U)
        return U


def _harmonic_1(m, k, H):
    """``getHarmVecs1`` (``helpers.py:154-196``)."""
    Hm = H[:m, :m]
    em = np.zeros(m, dtype=complex)
    em[-1] = 1.0
    f = np.linalg.solve(Hm.conj().T, em)
    G = Hm + np.abs(H[m, m - 1]) ** 2 * np.outer(f, em.conj())
    vals, vecs = sla.eig(G)
    good = np.isfinite(vals)
    vals, vecs = vals[good], vecs[:, good]
    return vecs[:, np.argsort(np.abs(vals))[:k]]


def _harmonic_2(G, Phi, k):
    """``getHarmVecs2`` (``helpers.py:204-265``) given the overlap ``Phi = W^H V``."""
    vals, vecs = sla.eig(G.conj().T @ G, G.conj().T @ Phi)
    good = np.isfinite(vals)
    vals, vecs = vals[good], vecs[:, good]
    return vecs[:, np.argsort(np.abs(vals))[:k]]
