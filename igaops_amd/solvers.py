"""Krylov drivers of the path with device-resident vectors.

``LinearSolver(tolerance, maxiters, cleandod=None, linear_type="gmres", verbose=True)
.solve(Afun, b, Pfun=None, x0=None, **mv_args) -> OutputArgs`` keeps the reference
signature and semantics (``solvers/linear_solver.py:52-257,438-525``):

* ``Afun``/``Pfun`` are applied to Dirichlet-masked vectors and their outputs are
  masked (``linear_solver.py:119-124,213-223``); ``b`` and ``x0`` are copied.
* GMRES is right preconditioned and, by default, not restarted.  Stop when any of:
  least-squares residual ``< tol * ||b||``, least-squares residual ``< 1e-14``, or
  ``||x_k - x_{k-1}|| / (||x_k|| + 1e-14) < tol`` (``convergence_manager.py:281-327``).

What changes on the device: the Krylov and preconditioned bases live in HBM;
orthogonalisation is classical Gram-Schmidt applied twice (CGS2, two many-dot and two
many-axpy kernels per pass instead of 2k dependent dot/axpy pairs; same span and
the same stability class as the reference's MGS with re-orthogonalisation); the
small least-squares problem is solved with Givens rotations on the host (the
residual estimate equals ``||beta e1 - H y||``); ``||x_k||`` and ``||x_k - x_{k-1}||``
come from the Gram matrix of the preconditioned vectors, so ``x`` is formed once at
the end.  One device-to-host copy of the new Hessenberg column per iteration.

``Afun``/``Pfun`` may be device operators (objects with ``apply_device(x, out, nrhs)``,
e.g. ``StiffnessMassOperator``, ``SingleFD``), callables on device tensors (when ``b``
is a device tensor) or numpy callables (when ``b`` is a numpy array; vectors then make
a host round trip per call, as the reference's callables would need).
"""
from __future__ import annotations

import logging
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from . import _lib
from . import device as dv
from .common import Constants, validate_entry

logger = logging.getLogger(__name__)


@dataclass
class OutputArgs:
    solution: object
    residual: np.ndarray
    success: bool = False
    iterations: int = 0
    reason: str = ""
    history: dict = field(default_factory=dict)


@dataclass
class SolverArgs:
    """``solvers/utils/argsclass.py:8-25``."""

    tolerance: float
    maxiters: int

    def __post_init__(self):
        if self.tolerance < Constants.SAFEGUARD or self.tolerance >= 1:
            raise ValueError(f"Relative tolerance must be > 0 and < 1, got {self.tolerance}")
        if self.maxiters <= 0:
            raise ValueError(f"Max iterations must be > 0, got {self.maxiters}")


def converged(tol: float, rel=None, absv=None, incr=None) -> str:
    """Any-of stopping rule; thresholds ``max(ref, 1e-14)``; returns the criterion met."""
    thr = max(tol, Constants.SAFEGUARD)
    if absv is not None and absv < Constants.SAFEGUARD:
        return "absolute"
    if rel is not None and rel < thr:
        return "relative"
    if incr is not None and incr < thr:
        return "increment"
    return ""


# ---------------------------------------------------------------------------
# Device vector kernels
# ---------------------------------------------------------------------------


class Krylov:
    """Reduction / update kernels on rows of a device matrix."""

    def __init__(self, n: int, kmax: int, stream=None):
        self.n = int(n)
        self.work = dv.workspace(_lib.load().iga_reduce_workspace_bytes(self.n, max(1, int(kmax))))
        self.stream = stream

    def dots(self, V, k: int, w, out):
        """``out[:k] = V[:k] @ w`` (device float64 out)."""
        _lib.call(
            "iga_dot_many", dv.iga_dtype(w), self.n, int(k), dv.ptr(V), V.stride(0) if V.dim() == 2 else self.n,
            dv.ptr(w), dv.ptr(out), dv.ptr(self.work), dv.stream_ptr(self.stream),
        )

    def axpy(self, V, k: int, h, alpha: float, w):
        """``w += alpha * V[:k]^T h``."""
        _lib.call(
            "iga_axpy_many", dv.iga_dtype(w), self.n, int(k), dv.ptr(V), V.stride(0) if V.dim() == 2 else self.n,
            dv.ptr(h), float(alpha), dv.ptr(w), dv.stream_ptr(self.stream),
        )

    def axpby(self, a: float, x, b: float, y, a_dev=None, invert_a=False):
        _lib.call(
            "iga_axpby", dv.iga_dtype(y), self.n, float(a), dv.ptr(a_dev), int(bool(invert_a)), dv.ptr(x), float(b),
            dv.ptr(y), dv.stream_ptr(self.stream),
        )

    def cast(self, x, y):
        _lib.call("iga_cast", dv.iga_dtype(x), dv.iga_dtype(y), self.n, dv.ptr(x), dv.ptr(y), dv.stream_ptr(self.stream))


# ---------------------------------------------------------------------------
# Operator adaptation
# ---------------------------------------------------------------------------


class Mask:
    """Dirichlet mask: an explicit index list or a free box of a tensor grid."""

    def __init__(self, cleandod=None, box=None):
        self.idx = None
        self.box = box
        if cleandod is not None and len(cleandod) > 0:
            self.idx = dv.torch().as_tensor(np.asarray(cleandod, dtype=np.int64), device=dv.device())
            self._keep = self.idx
        if box is not None:
            # ctypes views of the box, built once (the mask runs twice per iteration)
            self._box_arrays = [_lib.i64_array(a) for a in box]

    def __call__(self, v, stream=None):
        if self.idx is not None:
            _lib.call("iga_index_zero", dv.iga_dtype(v), int(self.idx.numel()), dv.ptr(self.idx), 1, dv.ptr(v),
                      dv.stream_ptr(stream))
        if self.box is not None:
            (na, npt), (la, lpt), (lna, lnpt) = self._box_arrays
            _lib.call("iga_box_mask", dv.iga_dtype(v), len(na), npt, lpt, lnpt, 1, dv.ptr(v), dv.stream_ptr(stream))


def device_callable(op, numpy_mode: bool, dtype: str = "f64"):
    """``f(x_dev, out_dev)`` writing ``op(x)`` into ``out`` (device vectors)."""
    t = dv.torch()
    if op is None:
        return lambda x, out: out.copy_(x)
    if hasattr(op, "apply_device"):
        return lambda x, out: op.apply_device(x, out=out)
    if numpy_mode:
        def f(x, out):
            y = np.asarray(op(x.double().cpu().numpy()))
            out.copy_(t.from_numpy(np.ascontiguousarray(y, dtype=np.float64)).to(out.device, out.dtype))
            return out
        return f

    def g(x, out):
        out.copy_(op(x))
        return out

    return g


class PrecisionAdapter:
    """Applies a lower-precision device operator to float64 Krylov vectors."""

    def __init__(self, op, n: int, dtype: str = "f32"):
        self.op = op
        self.dtype = dtype
        self.xin = dv.empty(n, dtype)
        self.yout = dv.empty(n, dtype)
        self.k = Krylov(n, 1)

    def apply_device(self, x, out=None, nrhs=1, stream=None):
        self.k.cast(x, self.xin)
        self.op.apply_device(self.xin, out=self.yout)
        if out is None:
            out = dv.torch().empty_like(x)
        self.k.cast(self.yout, out)
        return out


# ---------------------------------------------------------------------------
# GMRES
# ---------------------------------------------------------------------------


def _givens(a: float, b: float):
    if b == 0.0:
        return 1.0, 0.0
    r = np.hypot(a, b)
    return a / r, b / r


def gmres_device(
    A: Callable,
    b,
    P: Callable,
    tol: float,
    maxiters: int,
    mask: Optional[Mask] = None,
    restart: Optional[int] = None,
    x0=None,
    abs_floor: float = Constants.SAFEGUARD,
    use_increment: bool = True,
    stream=None,
) -> OutputArgs:
    """Right-preconditioned GMRES on device vectors (float64).

    ``A(x, out)`` and ``P(x, out)`` write into ``out``.  Without ``restart`` this is the
    reference's un-restarted GMRES; with it, cycles of ``restart`` iterations restart
    from the true residual ``b - A x``.  Without ``restart`` and without the
    increment criterion the Arnoldi steps run device-side (``runs.GmresRun``)."""
    if restart is None and not use_increment:
        from .runs import GmresRun

        return GmresRun(A, P, b, tol, maxiters, mask=mask, x0=x0, stream=stream).run()
    t = dv.torch()
    n = b.numel()
    m = min(maxiters, restart) if restart else maxiters
    V = t.empty((m + 1, n), dtype=t.float64, device=b.device)
    Z = t.empty((m, n), dtype=t.float64, device=b.device)
    w = t.empty(n, dtype=t.float64, device=b.device)
    dbuf = t.zeros(3 * m + 4, dtype=t.float64, device=b.device)
    hbuf = t.empty(3 * m + 4, dtype=t.float64, pin_memory=True)
    kr = Krylov(n, m + 1, stream)
    x = t.zeros(n, dtype=t.float64, device=b.device)
    r = t.empty(n, dtype=t.float64, device=b.device)
    r.copy_(b)
    if x0 is not None:
        x.copy_(x0)
        A(x, w)
        if mask:
            mask(w)
        r.sub_(w)

    def norm_dev(v):
        kr.dots(v.unsqueeze(0), 1, v, dbuf[:1])
        return float(dbuf[:1].cpu()[0]) ** 0.5

    # the relative criterion is measured against ||b - A x0||, as the reference solves
    # the shifted system b - A x0 from zero (linear_solver.py:228-236)
    beta = norm_dev(r)
    res_hist = [1.0]
    total_it = 0
    reason = ""
    success = False
    if beta < abs_floor:
        return OutputArgs(x, np.array([1.0]), True, 0, "absolute")
    ref = beta
    while True:
        V[0].copy_(r)
        kr.axpby(1.0 / beta, r, 0.0, V[0])
        H = np.zeros((m + 1, m))
        G = np.zeros((m, m))
        cs = np.zeros(m)
        sn = np.zeros(m)
        g = np.zeros(m + 1)
        g[0] = beta
        R = np.zeros((m, m))
        y_prev = np.zeros(0)
        k = 0
        for k in range(1, m + 1):
            total_it += 1
            P(V[k - 1], Z[k - 1])
            if mask:
                mask(Z[k - 1])
            A(Z[k - 1], w)
            if mask:
                mask(w)
            # Gram column of the preconditioned vectors (increment criterion)
            kr.dots(Z, k, Z[k - 1], dbuf[2 * m + 2:2 * m + 2 + k])
            # CGS2
            kr.dots(V, k, w, dbuf[0:k])
            kr.axpy(V, k, dbuf[0:k], -1.0, w)
            kr.dots(V, k, w, dbuf[m:m + k])
            kr.axpy(V, k, dbuf[m:m + k], -1.0, w)
            kr.dots(w.unsqueeze(0), 1, w, dbuf[2 * m + 1:2 * m + 2])
            hbuf.copy_(dbuf)  # one device-to-host copy per iteration (synchronising)
            hb = hbuf.numpy()
            h = hb[0:k] + hb[m:m + k]
            nw = float(np.sqrt(max(hb[2 * m + 1], 0.0)))
            gk = hb[2 * m + 2:2 * m + 2 + k]
            G[:k, k - 1] = gk
            G[k - 1, :k] = gk
            H[:k, k - 1] = h
            H[k, k - 1] = nw
            if nw != 0.0:
                kr.axpby(1.0 / nw, w, 0.0, V[k])
            # Givens update of the Hessenberg column
            col = H[: k + 1, k - 1].copy()
            for i in range(k - 1):
                a_, b_ = col[i], col[i + 1]
                col[i] = cs[i] * a_ + sn[i] * b_
                col[i + 1] = -sn[i] * a_ + cs[i] * b_
            c_, s_ = _givens(col[k - 1], col[k])
            cs[k - 1], sn[k - 1] = c_, s_
            col[k - 1] = c_ * col[k - 1] + s_ * col[k]
            col[k] = 0.0
            R[:k, k - 1] = col[:k]
            g[k] = -s_ * g[k - 1]
            g[k - 1] = c_ * g[k - 1]
            lsres = abs(g[k])
            # y_k and the increment criterion
            y = _back_substitute(R[:k, :k], g[:k])
            d = y.copy()
            d[: len(y_prev)] -= y_prev
            Gk = G[:k, :k]
            xn = float(np.sqrt(max(y @ Gk @ y, 0.0)))
            dn = float(np.sqrt(max(d @ Gk @ d, 0.0)))
            incr = dn / (xn + Constants.SAFEGUARD) if (use_increment and not restart) else None
            y_prev = y
            res_hist.append(lsres / ref)
            reason = converged(tol, rel=lsres / ref, absv=lsres, incr=incr)
            if reason or total_it >= maxiters or nw == 0.0:
                break
        # x += Z y
        yd = t.as_tensor(y_prev, dtype=t.float64, device=b.device)
        kr.axpy(Z, k, yd, 1.0, x)
        if reason:
            success = True
            break
        if total_it >= maxiters or not restart:
            break
        # restart from the true residual
        A(x, w)
        if mask:
            mask(w)
        r.copy_(b)
        r.sub_(w)
        if mask:
            mask(r)
        beta = norm_dev(r)
        if beta / ref < tol or beta < abs_floor:
            reason = "relative" if beta >= abs_floor else "absolute"
            success = True
            res_hist.append(beta / ref)
            break
    if not success:
        logger.warning("No convergence after %d iterations.", total_it)
    return OutputArgs(x, np.asarray(res_hist), success, total_it, reason or "maxiters")


def _back_substitute(R: np.ndarray, g: np.ndarray) -> np.ndarray:
    k = len(g)
    y = np.zeros(k)
    for i in range(k - 1, -1, -1):
        y[i] = (g[i] - R[i, i + 1:k] @ y[i + 1:k]) / R[i, i]
    return y


# ---------------------------------------------------------------------------
# Mixed precision: iterative refinement around a low-precision GMRES
# ---------------------------------------------------------------------------


def refinement_solve(
    A64: Callable,
    A_lo: Callable,
    P_lo: Callable,
    b,
    tol: float = 1e-10,
    inner_tol: float = 1e-4,
    inner_maxiters: int = 200,
    max_refinements: int = 30,
    mask: Optional[Mask] = None,
    restart: Optional[int] = None,
) -> OutputArgs:
    """``x += GMRES_lo(A_lo, r; P_lo)`` with ``r = b - A64 x`` in float64 until
    ``||r|| / ||b|| <= tol``.  ``A_lo``/``P_lo`` take float64 vectors (wrap float32
    operators in ``PrecisionAdapter``); the Arnoldi and least squares stay float64."""
    t = dv.torch()
    n = b.numel()
    kr = Krylov(n, 1)
    x = t.zeros(n, dtype=t.float64, device=b.device)
    r = b.clone()
    if mask:
        mask(r)
    buf = t.zeros(1, dtype=t.float64, device=b.device)

    def nrm(v):
        kr.dots(v.unsqueeze(0), 1, v, buf)
        return float(buf.cpu()[0]) ** 0.5

    nb = nrm(r)
    hist = [1.0]
    inner_its = []
    if nb < Constants.SAFEGUARD:
        return OutputArgs(x, np.array([1.0]), True, 0, "absolute")
    w = t.empty_like(b)
    success = False
    for _ in range(max_refinements):
        out = gmres_device(A_lo, r, P_lo, inner_tol, inner_maxiters, mask=mask, restart=restart, use_increment=False)
        inner_its.append(out.iterations)
        x.add_(out.solution)
        A64(x, w)
        if mask:
            mask(w)
        r.copy_(b)
        r.sub_(w)
        if mask:
            mask(r)
        rel = nrm(r) / nb
        hist.append(rel)
        if rel <= tol:
            success = True
            break
    return OutputArgs(x, np.asarray(hist), success, int(sum(inner_its)), "true-residual" if success else "max_refinements",
                      history={"inner_iterations": inner_its})


def mixed_solve(A64: Callable, A_lo: Callable, P_lo: Callable, b, tol: float = 1e-10,
                inner_tol: Optional[float] = None, inner_maxiters: int = 200, max_refinements: int = 30, mask: Optional[Mask] = None,
                lo_dtype: str = "f32") -> OutputArgs:
    """Iterative refinement with float64 residuals around inner GMRES solves whose
    vectors, operator ``A_lo`` and preconditioner ``P_lo`` are ``lo_dtype``
    (``inner_tol`` None: ``runs.REFINEMENT_INNER_TOL``)."""
    from .runs import REFINEMENT_INNER_TOL, RefinementRun

    itol = REFINEMENT_INNER_TOL if inner_tol is None else inner_tol
    return RefinementRun(A64, A_lo, P_lo, b, tol, itol, inner_maxiters, max_refinements, mask=mask,
                         lo_dtype=lo_dtype).run()


# ---------------------------------------------------------------------------
# Reference-signature front end
# ---------------------------------------------------------------------------


class LinearSolver:
    """``solvers/linear_solver.py:26-257`` for the GMRES and GCRODR types.

    CG, BiCGSTAB and block CG are not part of this path and raise
    ``NotImplementedError``."""

    def __init__(self, tolerance: float, maxiters: int, cleandod: Optional[List[int]] = None,
                 linear_type: str = "gmres", verbose: bool = True, restart: Optional[int] = None):
        self._config = SolverArgs(tolerance=tolerance, maxiters=maxiters)
        self.set_cleandod(cleandod)
        self._linear_type = validate_entry(linear_type, ["cg", "bicgstab", "gcrodr", "gmres", "block_cg"])
        self._verbose = verbose
        self._restart = restart
        self._last_simulation: Optional[OutputArgs] = None
        self._gcrodr_solver = None

    @property
    def config(self):
        return self._config

    @property
    def linear_type(self):
        return self._linear_type

    @property
    def last_simulation(self) -> OutputArgs:
        if self._last_simulation is None:
            return OutputArgs(solution=np.array([]), residual=np.array([]), success=False)
        return self._last_simulation

    # dense float64 LU on the device: N^2 doubles of HBM
    DIRECT_MAX_N = 40000

    @staticmethod
    def direct(A, b, verbose: bool = True) -> OutputArgs:
        """``linear_solver.py:595-630``: ``A x = b`` by LU.  The reference factors the
        sparse matrix with SuperLU on the host; here the matrix (scipy sparse, dense
        array or ``LinearOperator``) is expanded to a dense float64 matrix in HBM and
        factored on the device (rocSOLVER LU through ``torch.linalg.solve``), for
        N <= 40000."""
        import time as _time

        from scipy import sparse as sp
        from scipy.sparse.linalg import LinearOperator

        dv.require_gpu()
        t = dv.torch()
        start = _time.time()
        if isinstance(A, LinearOperator):
            A = np.asarray(A @ np.eye(A.shape[1]))
        n = A.shape[0]
        if A.shape[0] != A.shape[1]:
            raise ValueError("the matrix must be square")
        if n > LinearSolver.DIRECT_MAX_N:
            raise ValueError(f"direct solve on the device is limited to N <= {LinearSolver.DIRECT_MAX_N}")
        dense = t.zeros((n, n), dtype=t.float64, device=dv.device())
        if sp.issparse(A):
            coo = sp.coo_array(A)
            r = t.as_tensor(coo.row.astype(np.int64), device=dv.device())
            c = t.as_tensor(coo.col.astype(np.int64), device=dv.device())
            dense.index_put_((r, c), t.as_tensor(coo.data.astype(np.float64), device=dv.device()), accumulate=True)
        else:
            dense.copy_(dv.to_device(np.asarray(A, dtype=np.float64), "f64").reshape(n, n))
        was_np = not isinstance(b, t.Tensor)
        bd = dv.to_device(b, "f64")
        x = t.linalg.solve(dense, bd.reshape(n, -1)).reshape(bd.shape)
        if verbose:
            logger.info(f"Direct solver took {_time.time() - start:.2e} seconds.")
        return OutputArgs(solution=x.cpu().numpy() if was_np else x, residual=np.array([]), success=True)

    def set_cleandod(self, cleandod):
        if cleandod is not None and not isinstance(cleandod, list):
            raise TypeError("Cleandod should be list")
        self._cleandod = cleandod

    def update(self, tolerance=None, maxiters=None):
        kw = dict(self._config.__dict__)
        if tolerance is not None:
            kw["tolerance"] = tolerance
        if maxiters is not None:
            kw["maxiters"] = maxiters
        self._config = SolverArgs(**kw)

    def solve(self, Afun, b, Pfun=None, x0=None, **mv_args) -> OutputArgs:
        t = dv.torch()
        if Pfun is not None and not callable(Pfun) and not hasattr(Pfun, "apply_device"):
            Pfun = None
        if not (callable(Afun) or hasattr(Afun, "apply_device")):
            raise TypeError("Afun and Pfun should be callable")
        if self._linear_type not in ("gmres", "gcrodr"):
            raise NotImplementedError(f"{self._linear_type} is not part of the GPU path")
        dv.require_gpu()
        solver_args = {k: mv_args.pop(k) for k in ["krylov_size", "recycled_size", "max_cycles"] if k in mv_args}
        numpy_mode = not isinstance(b, t.Tensor)
        if mv_args and not hasattr(Afun, "apply_device"):
            base = Afun
            Afun = lambda v, _b=base: _b(v, **mv_args)
        A = device_callable(Afun, numpy_mode)
        P = device_callable(Pfun, numpy_mode)
        bd = dv.to_device(b, "f64").reshape(-1).clone()
        mask = Mask(self._cleandod) if self._cleandod is not None else None
        if mask:
            mask(bd)
        x0d = None
        if x0 is not None:
            # linear_solver.py:228-236: solve A d = b - A x0 from zero, then x = x0 + d
            x0d = dv.to_device(x0, "f64").reshape(-1).clone()
            if mask:
                mask(x0d)
            ax0 = t.empty_like(bd)
            A(x0d.clone(), ax0)
            if mask:
                mask(ax0)
            bd.sub_(ax0)
        start = time.time()
        if self._linear_type == "gmres":
            out = gmres_device(A, bd, P, self._config.tolerance, self._config.maxiters, mask=mask,
                               restart=self._restart)
        else:
            from .gcrodr import GCRODR

            if self._gcrodr_solver is None:
                self._gcrodr_solver = GCRODR(self._config.tolerance, self._config.maxiters,
                                             krylov_size=solver_args.get("krylov_size", 40),
                                             recycled_size=solver_args.get("recycled_size", 20),
                                             max_cycles=solver_args.get("max_cycles", 10), verbose=self._verbose)
            Am = (lambda x, out: (A(x, out), mask(out) if mask else None)[0]) if mask else A
            Pm = (lambda x, out: (P(x, out), mask(out) if mask else None)[0]) if mask else P
            out = self._gcrodr_solver.solve_device(Am, bd, Pm)
        if x0d is not None:
            out.solution = out.solution + x0d
        if numpy_mode:
            out.solution = out.solution.cpu().numpy().reshape(np.shape(b))
        self._last_simulation = out
        if self._verbose:
            logger.info("LINEAR SOLVER: %d equations, %d iterations, rel residual %.2e, %s, %.2e s",
                        bd.numel(), out.iterations, out.residual[-1], out.reason, time.time() - start)
        return out
