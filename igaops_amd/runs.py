"""Device-resident Krylov solves advanced step by step, without a host round trip
per iteration.

``GmresRun`` is the reference's un-restarted, right-preconditioned GMRES
(``solvers/linear_solver.py:438-525``) with its relative and absolute residual
criteria (``solvers/utils/convergence_manager.py:156-327``).  Each Arnoldi step is
queued on the run's stream: preconditioner, operator, the orthogonalisation
(classical Gram-Schmidt with the reorthogonalisation delayed by one step: two
passes over the basis per step), the Givens update of the least-squares problem
and the next basis vector all stay on the device (``iga_gmres_step``).  Step ``k``
yields the residual estimate ``|g|`` of Krylov dimension ``k - 1``; the host reads
it from a two-value asynchronous copy and an event, so it can queue the next
steps before the previous ones finish.  Steps queued past convergence do not
change the solution ``x0 + P (V[:k] y)``, formed from the first ``k`` basis
vectors only.

``RefinementRun`` is mixed-precision iterative refinement: inner ``GmresRun`` solves
of ``A_lo d = r`` in float32 vectors and operators, the residual ``r = b - A x`` and
the solution in float64, until ``||r|| / ||b|| <= tol``.

With a fused operator and an FD preconditioner (``NativeStep``), a run queues each
Arnoldi step with one library call (``iga_arnoldi_step``: FD, operator, mask,
orthogonalisation, and the residual estimate's copy to the host behind an event),
which keeps the host time per step well below the device time even for small
problems.

Both expose ``advance()``, which queues whatever work is possible and returns
whether anything happened, so that one host thread can interleave many runs on
separate streams (the serving loop does).
"""
from __future__ import annotations

import contextlib
import ctypes as C
from typing import Callable, Optional

import numpy as np

from . import _lib
from . import device as dv
from .common import Constants


def _back_substitute(R: np.ndarray, g: np.ndarray) -> np.ndarray:
    k = len(g)
    y = np.zeros(k)
    for i in range(k - 1, -1, -1):
        y[i] = (g[i] - R[i, i + 1:k] @ y[i + 1:k]) / R[i, i]
    return y


def _relative_or_absolute(tol: float, rel: float, absv: float) -> str:
    if rel < tol:
        return "relative"
    if absv < Constants.SAFEGUARD:
        return "absolute"
    return ""


class HostRecord:
    """Coherent pinned host memory (``iga_host_alloc``) that the device writes the
    residual estimates of a solve into: ``values[2 + 2d]``, ``values[3 + 2d]`` for
    hist entry d, and ``tag + d`` into the int64 at the front once they are written."""

    def __init__(self, ndoubles: int):
        ptr = C.c_void_p()
        _lib.call("iga_host_alloc", int(ndoubles) * 8, C.byref(ptr))
        self.ptr = int(ptr.value)
        self.size = int(ndoubles)
        self.values = np.ctypeslib.as_array((C.c_double * self.size).from_address(self.ptr))
        self.seq = self.values.view(np.int64)[:1]


_RECORDS: dict = {}  # free host records by size; a record outlives the solves that use it
_TAGS = iter(range(1, 1 << 30))


def _take_record(ndoubles: int) -> HostRecord:
    free = _RECORDS.setdefault(ndoubles, [])
    return free.pop() if free else HostRecord(ndoubles)


def _give_record(rec: HostRecord):
    _RECORDS.setdefault(rec.size, []).append(rec)


class NativeStep:
    """The pieces of a GMRES solve that the one-call Arnoldi step (``iga_arnoldi_*``)
    takes over from ``A``, ``P`` and ``mask``: a fused ``StiffnessMassOperator``, an
    FD plan (``None``: no preconditioner) with its workspace, and the mask's free box
    ``(n, lo, len)`` (``None``: no mask).  The FD writes zeros outside its own box, so
    that box must be the mask's."""

    def __init__(self, op, fd_plan=None, fd_work=None, box=None):
        if not getattr(op, "fused", False):
            raise ValueError("the one-call Arnoldi step needs a fused operator")
        if fd_plan is not None:
            if fd_work is None:
                raise ValueError("the FD plan needs its workspace")
            if box is None or not all(np.array_equal(a, b) for a, b in zip(fd_plan.box, box)):
                raise ValueError("the FD's free box must be the mask's box")
            if (op.dtype == "f64") != (fd_plan.dtype == "f64"):
                raise ValueError("operator and FD plan take different vector types")
        self.op, self.fd_plan, self.fd_work = op, fd_plan, fd_work
        self.box = None if box is None else tuple(np.ascontiguousarray(a, dtype=np.int64) for a in box)


class GmresRun:
    """One GMRES solve of ``A x = b`` right-preconditioned by ``P``.

    ``A(x, out)`` / ``P(x, out)`` write into ``out`` (vectors of ``dtype``); ``mask``
    zeroes the constrained entries (``linear_solver.py:119-124``).  ``maxiters`` is
    also the Krylov dimension (the reference does not restart).  With ``native`` (a
    :class:`NativeStep` describing the same ``A``, ``P`` and ``mask``) each Arnoldi step
    is queued by one library call instead of one call per kernel group; ``A``, ``P``
    and ``mask`` still serve the start and the solution update."""

    def __init__(self, A: Callable, P: Optional[Callable], b, tol: float, maxiters: int, mask=None, x0=None,
                 stream=None, lookahead: int = 2, dtype: Optional[str] = None, host_buf=None,
                 native: Optional[NativeStep] = None, increment: bool = False):
        t = dv.torch()
        self.A, self.P, self.mask = A, P, mask
        self.tol, self.m = float(tol), int(maxiters)
        if self.m < 1:
            raise ValueError("maxiters must be positive")
        self.stream = stream
        self.lookahead = max(1, int(lookahead))
        self.dtype = dtype or ("f32" if b.dtype == t.float32 else "f64")
        tdt = dv.torch_dtype(self.dtype)
        self.n = n = b.numel()
        lib = _lib.load()
        with self._on_stream():
            self.b = b
            # the basis grows by doubling as steps are queued (most solves stop far below
            # maxiters, and the allocator keeps what a solve has held); step k writes V[k].
            # Rows start on 16-byte boundaries (leading dimension padded to 16 bytes) so
            # that the reductions over the basis take their vector path for every n.
            per16 = 16 // t.empty(0, dtype=tdt).element_size()
            self.ldv = -(-n // per16) * per16
            self.V = t.empty((min(self.m, 64) + 2, self.ldv), dtype=tdt, device=b.device)
            self.z = t.empty(n, dtype=tdt, device=b.device)
            self.w = t.empty(n, dtype=tdt, device=b.device)
            self.x = t.zeros(n, dtype=tdt, device=b.device)
            self.x0 = x0
            self.nls = int(lib.iga_gmres_state_doubles(self.m))
            self.ls = t.zeros(self.nls, dtype=t.float64, device=b.device)
            offs = (C.c_int64 * 3)()
            _lib.check(lib.iga_gmres_layout(self.m, offs), "iga_gmres_layout")
            self._off_r, self._off_g, self._hist_off = int(offs[0]), int(offs[1]), int(offs[2])
            self.work = dv.workspace(lib.iga_reduce_workspace_bytes(n, self.m + 1))
        # one pinned host buffer per run (or lent by the caller, at least
        # host_doubles(maxiters) values): polled residuals, then the least-squares
        # factors and the solution coefficients
        need = self.host_doubles(self.m)
        if host_buf is None or host_buf.numel() < need:
            host_buf = t.zeros(need, dtype=t.float64, pin_memory=True)
        self.host = host_buf
        self.hist_host = host_buf[:2 * (self.m + 2)]
        self._sol_region = host_buf[2 * (self.m + 2):2 * (self.m + 2) + self.m * self.m + self.m]
        self._y_region = host_buf[2 * (self.m + 2) + self.m * self.m + self.m:need]
        self.launched = 0      # steps queued
        self.known = -1        # last step whose residual estimate has been read
        self.recorded = -1     # last step whose residual estimate is on its way to the host
        self.started = False
        self.events = {}
        self.kfinal: Optional[int] = None
        self.reason = ""
        self.res_hist = []
        self.ref = None
        self.beta = None
        self._sol_event = None
        self._sol_host = None
        self.result = None
        self._ctx = None
        self._rec = None
        self._inc = None
        if increment and native is None:
            raise ValueError("the increment criterion of the device steps needs the one-call step (native)")
        if native is not None:
            if native.op.dtype != self.dtype:
                raise ValueError("the native operator does not take the run's vector type")
            op, fd = native.op, native.fd_plan
            bx = native.box
            nd = 0 if bx is None else len(bx[0])
            bp = [None, None, None] if bx is None else [a.ctypes.data_as(_lib._i64p) for a in bx]
            self._native = native
            # the device publishes the residual estimates here under this solve's tag
            self._rec = _take_record(2 * (self.m + 2))
            self._tag = next(_TAGS)
            with self._on_stream():
                # the operator's partial blocks are written on this run's stream: take
                # them from its memory pool (a block from another stream's pool may still
                # be read by that stream's pending work)
                op_work = op.workspace(1)
            ctx = C.c_void_p()
            _lib.call("iga_arnoldi_create", C.byref(ctx), dv.DTYPES[self.dtype], n, self.m, op.plan._h,
                      dv.ptr(op.packed), dv.ptr(op.w), dv.ptr(op_work),
                      None if fd is None else fd._h, dv.ptr(native.fd_work), nd, *bp, dv.ptr(self.z),
                      dv.ptr(self.w), dv.ptr(self.ls), dv.ptr(self.work), self._rec.ptr, self._tag << 32)
            self._ctx = ctx
            if increment:
                self._setup_increment()

    def __del__(self):
        self._drop_native(give_back=False)

    # -- increment criterion (linear_solver.py:107-117,503-515) --
    def _setup_increment(self):
        """Keep the preconditioned vectors Z (rows like V) and read, per step, the
        increment record the device writes (``iga_arnoldi_set_increment``)."""
        t = dv.torch()
        m = self.m
        with self._on_stream():
            self.Z = t.empty(self.V.shape, dtype=self.V.dtype, device=self.V.device)
            self._gram = t.empty(m + 1, dtype=t.float64, device=self.V.device)
        nd = int(_lib.load().iga_arnoldi_increment_doubles(m))
        self._inc = _take_record(nd)
        self._inc_stride = 2 * m + 3
        self._incR = np.zeros((m + 1, m + 1))
        self._incg = np.zeros(m + 1)
        self._incG = np.zeros((m + 2, m + 2))
        self._inc_read = -1
        self._inc_y = np.zeros(0)
        self.increments = []
        self._bind_increment()

    def _bind_increment(self):
        _lib.call("iga_arnoldi_set_increment", self._ctx, dv.ptr(self.Z), self.ldv, dv.ptr(self._gram),
                  self._inc.ptr)

    def _increment(self, j: int) -> float:
        """||x_j - x_{j-1}|| / (||x_j|| + 1e-14) of Krylov dimension j >= 1, with
        x_j = Z_j y_j, y_j = R_j^-1 g_j, the norms from the Gram matrix of Z."""
        from scipy.linalg import solve_triangular

        vals = self._inc.values
        stride, m = self._inc_stride, self.m
        seq = int(self._inc.seq[0])
        if (seq >> 32) != self._tag or (seq & 0xFFFFFFFF) < j:
            raise RuntimeError(f"increment record {j} was not published by the device")
        while self._inc_read < j:
            e = self._inc_read + 1
            base = 1 + e * stride
            if e >= 1:
                self._incR[:e, e - 1] = vals[base:base + e]
                self._incg[e - 1] = vals[base + m]
            row = vals[base + m + 1:base + m + 2 + e]
            self._incG[e, :e + 1] = row
            self._incG[:e + 1, e] = row
            self._inc_read = e
        R = self._incR[:j, :j]
        y = solve_triangular(R, self._incg[:j]) if np.all(np.diag(R) != 0) else np.linalg.lstsq(R, self._incg[:j],
                                                                                                  rcond=None)[0]
        d = y.copy()
        d[:len(self._inc_y)] -= self._inc_y
        G = self._incG[:j, :j]
        num = max(float(d @ G @ d), 0.0) ** 0.5
        den = max(float(y @ G @ y), 0.0) ** 0.5
        self._inc_y = y
        return num / (den + Constants.SAFEGUARD)

    def _drop_native(self, give_back: bool = True):
        """Destroy the step context; with ``give_back`` the host record returns to the
        pool (only once every step of this solve has finished: the solution copy,
        queued after them, has been waited for)."""
        ctx = getattr(self, "_ctx", None)
        if ctx is not None and ctx.value and _lib._lib is not None:
            _lib.load().iga_arnoldi_destroy(ctx)
        self._ctx = None
        rec = getattr(self, "_rec", None)
        if rec is not None and give_back:
            _give_record(rec)
        self._rec = None
        inc = getattr(self, "_inc", None)
        if inc is not None and give_back:
            _give_record(inc)
        self._inc = None

    @staticmethod
    def host_doubles(m: int) -> int:
        return 2 * (m + 2) + m * m + 2 * m

    # -- plumbing --
    def _on_stream(self):
        return dv.torch().cuda.stream(self.stream) if self.stream is not None else contextlib.nullcontext()

    def _sp(self):
        return dv.stream_ptr(self.stream)

    def _record(self, step: int, src_off: int, dst_off: int, count: int):
        t = dv.torch()
        self.hist_host[dst_off:dst_off + count].copy_(self.ls[src_off:src_off + count], non_blocking=True)
        ev = t.cuda.Event()
        ev.record(self.stream if self.stream is not None else t.cuda.current_stream())
        self.events[step] = ev
        self.recorded = step

    def _published(self, j: int) -> bool:
        seq = int(self._rec.seq[0])
        return (seq >> 32) == self._tag and (seq & 0xFFFFFFFF) >= j

    def _arrived(self, j: int, block: bool) -> bool:
        """Whether residual estimate j is on the host (waiting for it with ``block``)."""
        if self._ctx is not None:
            if self._published(j):
                return True
            if block:
                (self.stream or dv.torch().cuda.current_stream()).synchronize()
                if not self._published(j):
                    raise RuntimeError(f"residual estimate {j} was not published by the device")
                return True
            return False
        ev = self.events[j]
        if ev.query():
            del self.events[j]
            return True
        if block:
            ev.synchronize()
            del self.events[j]
            return True
        return False

    # -- phases --
    def start(self):
        t = dv.torch()
        dt = dv.DTYPES[self.dtype]
        with self._on_stream():
            r = self.w
            r.copy_(self.b)
            if self.x0 is not None:
                self.x.copy_(self.x0)
                tmp = t.empty_like(self.w)
                self.A(self.x.clone(), tmp)
                if self.mask:
                    self.mask(tmp)
                r.sub_(tmp)
            if self._ctx is not None:
                _lib.call("iga_arnoldi_start", self._ctx, dv.ptr(r), dv.ptr(self.V), self._sp())
                self.recorded = 0
            else:
                _lib.call("iga_gmres_start", dt, self.n, self.m, dv.ptr(r), dv.ptr(self.row(0)), dv.ptr(self.ls),
                          dv.ptr(self.work), self._sp())
                self._record(0, self._hist_off, 0, 2)
        self.launched = 0
        self.started = True

    def can_launch(self) -> bool:
        # step k reports dimension k - 1: one more step in flight than the lookahead
        return self.kfinal is None and self.launched <= self.m and self.launched - self.known <= self.lookahead

    def row(self, i: int):
        """Basis vector i (a length-n view of row i)."""
        return self.V[i, : self.n]

    def _ensure_rows(self, k: int):
        """Rows V[0..k] for step k."""
        rows = self.V.shape[0]
        if k < rows:
            return
        t = dv.torch()
        while rows <= k:
            rows = min(self.m + 2, 2 * rows)
        V = t.empty((rows, self.ldv), dtype=self.V.dtype, device=self.V.device)
        V[:self.V.shape[0]].copy_(self.V)
        self.V = V
        if self._inc is not None:
            Z = t.empty((rows, self.ldv), dtype=self.Z.dtype, device=self.Z.device)
            Z[:self.Z.shape[0]].copy_(self.Z)
            self.Z = Z
            self._bind_increment()

    def prepare_step(self):
        """Capacity for step k = launched + 1; returns (k, V[k-1], z) so that a caller
        can fill z = P V[k-1] itself (batched with other runs) before
        ``launch(preconditioned=True)``."""
        k = self.launched + 1
        if k >= self.V.shape[0]:
            with self._on_stream():
                self._ensure_rows(k)
        return k, self.row(k - 1), self.z

    def launch(self, preconditioned: bool = False, applied: bool = False):
        """Queue Arnoldi step k = launched + 1 (with ``preconditioned``, ``z`` already
        holds P V[k-1]; with ``applied``, ``w`` already holds the masked A P V[k-1])."""
        k = self.launched + 1
        dt = dv.DTYPES[self.dtype]
        stage = 2 if applied else int(bool(preconditioned))
        if self._ctx is not None:
            if k >= self.V.shape[0]:  # the basis grows (on the run's stream) only now and then
                with self._on_stream():
                    self._ensure_rows(k)
            _lib.call("iga_arnoldi_step", self._ctx, k, dv.ptr(self.V), self.ldv, stage, self._sp())
            if k >= 2:
                self.recorded = k - 1
            self.launched = k
            return
        with self._on_stream():
            self._ensure_rows(k)
            z = self.z
            if stage == 0:
                if self.P is not None:
                    self.P(self.row(k - 1), z)
                else:
                    z.copy_(self.row(k - 1))
            if stage <= 1:
                if self.mask:
                    self.mask(z)
                self.A(z, self.w)
                if self.mask:
                    self.mask(self.w)
            _lib.call("iga_gmres_step", dt, self.n, k, self.m, dv.ptr(self.V), self.ldv, dv.ptr(self.w),
                      dv.ptr(self.ls), dv.ptr(self.work), self._sp())
            if k >= 2:
                self._record(k - 1, self._hist_off + 2 * (k - 1), 2 * (k - 1), 2)
        self.launched = k

    def _launched_native(self, k: int):
        """Bookkeeping of one-call step k queued by someone else (a group step)."""
        if k >= 2:
            self.recorded = k - 1
        self.launched = k

    def poll(self, block: bool = False) -> bool:
        """Read the residual estimates of finished steps; decide convergence."""
        progressed = False
        while self.kfinal is None and self.known + 1 <= self.recorded:
            j = self.known + 1
            if not self._arrived(j, block):
                break
            block = False
            self.known = j
            progressed = True
            if self._ctx is not None:
                g, nw = float(self._rec.values[2 + 2 * j]), float(self._rec.values[3 + 2 * j])
            else:
                g, nw = float(self.hist_host[2 * j]), float(self.hist_host[2 * j + 1])
            if j == 0:
                # relative to ||b - A x0|| (the reference shifts b by A x0 and starts
                # from zero, linear_solver.py:228-236)
                self.beta = abs(g)
                self.ref = self.beta
                self.res_hist.append(1.0)
                if self.beta < Constants.SAFEGUARD:
                    self.kfinal, self.reason = 0, "absolute"
                continue
            res = abs(g)
            self.res_hist.append(res / self.ref)
            reason = _relative_or_absolute(self.tol, res / self.ref, res)
            if self._inc is not None and nw != 0.0:
                incr = self._increment(j)
                self.increments.append(incr)
                if not reason and incr < self.tol:
                    reason = "increment"
            if reason or j >= self.m or nw == 0.0:
                self.kfinal = j
                self.reason = reason or ("breakdown" if nw == 0.0 else "maxiters")
        return progressed

    @property
    def converged_known(self) -> bool:
        return self.kfinal is not None

    def request_solution(self):
        """Queue the copy of the least-squares factors of the first kfinal steps."""
        t = dv.torch()
        k = self.kfinal
        m = self.m
        if k == 0:
            self._sol_host = None
            return
        with self._on_stream():
            off_r, off_g = self._off_r, self._off_g
            self._sol_host = self._sol_region[:k * m + k]
            self._sol_host[:k * m].copy_(self.ls[off_r:off_r + k * m], non_blocking=True)
            self._sol_host[k * m:].copy_(self.ls[off_g:off_g + k], non_blocking=True)
            self._sol_event = t.cuda.Event()
            self._sol_event.record(self.stream if self.stream is not None else t.cuda.current_stream())

    def solution_ready(self, block: bool = False) -> bool:
        if self._sol_event is None:
            return True
        if block:
            self._sol_event.synchronize()
            return True
        return self._sol_event.query()

    def complete(self):
        """x += P (V[:k]^T y) (queued); returns the solve's OutputArgs."""
        from .solvers import OutputArgs

        t = dv.torch()
        k = self.kfinal
        if k and k > 0:
            self.solution_ready(block=True)
            h = self._sol_host.numpy()
            R = h[:k * self.m].reshape(k, self.m).T[:k, :k]
            y = _back_substitute(R, h[k * self.m:k * self.m + k])
            with self._on_stream():
                # the pinned source and the device copy are kept with the run until it is dropped
                yh = self._y_region[:k]
                yh.copy_(t.from_numpy(y))
                self._y_dev = yh.to(self.x.device, non_blocking=True)
                v = self.w
                v.zero_()
                _lib.call("iga_axpy_many", dv.DTYPES[self.dtype], self.n, k, dv.ptr(self.V), self.ldv,
                          dv.ptr(self._y_dev), 1.0, dv.ptr(v), self._sp())
                z = self.z
                if self.P is not None:
                    self.P(v, z)
                else:
                    z.copy_(v)
                if self.mask:
                    self.mask(z)
                self.x.add_(z)
        success = self.reason in ("relative", "absolute", "increment")
        if self.reason == "breakdown":
            # a vanishing basis vector means the Krylov space is invariant, which in exact
            # arithmetic makes the least-squares solution exact; the true residual decides
            success = self._true_residual_meets_tol()
        self.result = OutputArgs(self.x, np.asarray(self.res_hist), success, int(k or 0), self.reason)
        if self._rec is not None and not k and self.launched > 0:
            # no solution copy was queued behind the steps already launched: wait for them
            # before the host record can serve another solve
            (self.stream or t.cuda.current_stream()).synchronize()
        self._drop_native()
        return self.result

    def _true_residual_meets_tol(self) -> bool:
        """``||b - A x|| <= tol * ||b - A x0||`` or below the absolute floor (synchronising)."""
        t = dv.torch()
        with self._on_stream():
            r = t.empty_like(self.w)
            self.A(self.x.clone(), r)
            r.sub_(self.b)
            if self.mask:
                self.mask(r)
            res = float(t.linalg.vector_norm(r.double()))
        return res <= self.tol * self.ref or res < Constants.SAFEGUARD

    def release(self):
        """Drop the Krylov basis (the solution tensor is kept)."""
        self.V = None

    # -- driver --
    def advance(self, launch: bool = True) -> bool:
        """Queue what can be queued without waiting; with ``launch=False`` no
        Arnoldi step is queued (the caller may batch them)."""
        progressed = False
        if not self.started:
            self.start()
            progressed = True
        if self.result is not None:
            return False
        progressed |= self.poll()
        while launch and self.can_launch():
            self.launch()
            progressed = True
        if self.kfinal is not None:
            if self._sol_event is None and self._sol_host is None:
                self.request_solution()
                progressed = True
            if self.solution_ready():
                self.complete()
                progressed = True
        return progressed

    def step_run(self) -> Optional["GmresRun"]:
        """This solve if it can take an Arnoldi step now."""
        if self.result is None and self.known >= 0 and self.can_launch():
            return self
        return None

    def stepping(self) -> bool:
        """Still taking Arnoldi steps (started, stopping step not yet known)."""
        return self.started and self.result is None and self.kfinal is None

    @property
    def done(self) -> bool:
        return self.result is not None

    def run(self):
        """Synchronous driver: the same steps with the host waiting on each event."""
        self.start()
        while self.kfinal is None:
            while self.can_launch():
                self.launch()
            self.poll(block=True)
        self.request_solution()
        return self.complete()


# Inner tolerance floor of the refinement's FP32 solves: each inner solve restarts GMRES,
# so a looser floor keeps the bases short and the orthogonalisation cheap; on a saturated
# 400-request serving stream 1e-4 to 1e-2 took 104.6-106.7 inner iterations and 84-86
# solves/s, 1e-6 109 and 78, 1e-7 115 and 73, 3e-2 112 and 78.
REFINEMENT_INNER_TOL = 1e-4


class RefinementRun:
    """``x += GMRES_lo(A_lo, r; P_lo)`` with ``r = b - A64 x`` in float64 until
    ``||r|| / ||b|| <= tol``; the inner solves use float32 vectors and operators."""

    def __init__(self, A64: Callable, A_lo: Callable, P_lo: Callable, b, tol: float = 1e-10,
                 inner_tol: float = REFINEMENT_INNER_TOL, inner_maxiters: int = 200, max_refinements: int = 30,
                 mask=None,
                 stream=None, lookahead: int = 2, lo_dtype: str = "f32", host_buf=None,
                 native_lo: Optional[NativeStep] = None):
        t = dv.torch()
        self.A64, self.A_lo, self.P_lo, self.mask = A64, A_lo, P_lo, mask
        self.native_lo = native_lo
        self.tol, self.inner_tol = float(tol), float(inner_tol)
        self.inner_maxiters, self.max_ref = int(inner_maxiters), int(max_refinements)
        self.stream, self.lookahead, self.lo_dtype = stream, lookahead, lo_dtype
        self.n = b.numel()
        with self._on_stream():
            self.b = b
            self.x = t.zeros(self.n, dtype=t.float64, device=b.device)
            self.r = b.clone()
            if mask:
                mask(self.r)
            self.r_lo = t.empty(self.n, dtype=dv.torch_dtype(lo_dtype), device=b.device)
            self.w = t.empty_like(self.x)
            self.nrm = t.zeros(2, dtype=t.float64, device=b.device)
            self.work = dv.workspace(_lib.load().iga_reduce_workspace_bytes(self.n, 1))
        need = GmresRun.host_doubles(self.inner_maxiters) + 2
        if host_buf is None or host_buf.numel() < need:
            host_buf = t.zeros(need, dtype=t.float64, pin_memory=True)
        self.nrm_host = host_buf[:2]
        self._inner_host = host_buf[2:need]
        self.phase = "norm"
        self.inner: Optional[GmresRun] = None
        self.inner_its = []
        self.hist = []
        self.nb = None
        self._ev = None
        self.result = None

    def _on_stream(self):
        return dv.torch().cuda.stream(self.stream) if self.stream is not None else contextlib.nullcontext()

    def _queue_norm(self, v, slot: int):
        t = dv.torch()
        with self._on_stream():
            _lib.call("iga_dot_many", _lib.IGA_F64, self.n, 1, dv.ptr(v), self.n, dv.ptr(v), dv.ptr(self.nrm[slot:]),
                      dv.ptr(self.work), dv.stream_ptr(self.stream))
            self.nrm_host[slot:slot + 1].copy_(self.nrm[slot:slot + 1], non_blocking=True)
            self._ev = t.cuda.Event()
            self._ev.record(self.stream if self.stream is not None else t.cuda.current_stream())

    def _start_inner(self):
        with self._on_stream():
            _lib.call("iga_cast", _lib.IGA_F64, dv.DTYPES[self.lo_dtype], self.n, dv.ptr(self.r), dv.ptr(self.r_lo),
                      dv.stream_ptr(self.stream))
        # each inner solve only needs to shrink the residual by what is left to reach
        # tol (with a factor 2 of margin), but not below inner_tol, the accuracy a
        # float32 solve can deliver
        rel = self.hist[-1] if self.hist else 1.0
        itol = min(1e-1, max(self.inner_tol, 0.5 * self.tol / max(rel, 1e-300)))
        self.inner = GmresRun(self.A_lo, self.P_lo, self.r_lo, itol, self.inner_maxiters, mask=self.mask,
                              stream=self.stream, lookahead=self.lookahead, dtype=self.lo_dtype,
                              host_buf=self._inner_host, native=self.native_lo)
        self.phase = "inner"

    def _finish(self, success: bool):
        from .solvers import OutputArgs

        self.result = OutputArgs(self.x, np.asarray(self.hist), success, int(sum(self.inner_its)),
                                 "true-residual" if success else "max_refinements",
                                 history={"inner_iterations": list(self.inner_its)})

    def step_run(self) -> Optional[GmresRun]:
        """The inner solve if it can take an Arnoldi step now."""
        if self.phase == "inner" and self.inner is not None and self.inner.result is None and self.inner.known >= 0 \
                and self.inner.can_launch():
            return self.inner
        return None

    def stepping(self) -> bool:
        """An inner solve is taking Arnoldi steps."""
        return self.phase == "inner" and self.inner is not None and self.inner.stepping()

    def advance(self, launch: bool = True) -> bool:
        if self.result is not None:
            return False
        if self.phase == "norm":  # ||b||
            self._queue_norm(self.r, 0)
            self.phase = "wait_nb"
            return True
        if self.phase == "wait_nb":
            if not self._ev.query():
                return False
            self.nb = float(self.nrm_host[0]) ** 0.5
            self.hist.append(1.0)
            if self.nb < Constants.SAFEGUARD:
                self._finish(True)
                return True
            self._start_inner()
            return True
        if self.phase == "inner":
            progressed = self.inner.advance(launch)
            if self.inner.done:
                out = self.inner.result
                self.inner_its.append(out.iterations)
                with self._on_stream():
                    _lib.call("iga_cast", dv.DTYPES[self.lo_dtype], _lib.IGA_F64, self.n, dv.ptr(out.solution),
                              dv.ptr(self.w), dv.stream_ptr(self.stream))
                    self.x.add_(self.w)
                    self.A64(self.x, self.w)
                    if self.mask:
                        self.mask(self.w)
                    self.r.copy_(self.b)
                    self.r.sub_(self.w)
                    if self.mask:
                        self.mask(self.r)
                self.inner.release()
                self._queue_norm(self.r, 1)
                self.phase = "wait_r"
                return True
            return progressed
        if self.phase == "wait_r":
            if not self._ev.query():
                return False
            rel = float(self.nrm_host[1]) ** 0.5 / self.nb
            self.hist.append(rel)
            if rel <= self.tol:
                self._finish(True)
            elif len(self.inner_its) >= self.max_ref:
                self._finish(False)
            else:
                self._start_inner()
            return True
        return False

    @property
    def done(self) -> bool:
        return self.result is not None

    def run(self):
        while not self.done:
            if not self.advance():
                # nothing to do until the stream catches up
                if self.phase == "inner" and self.inner is not None:
                    self.inner.poll(block=True)
                elif self._ev is not None:
                    self._ev.synchronize()
        return self.result


def launch_group(group, work: dict, batch_stream=None):
    """One Arnoldi step of every run of ``group`` (``GmresRun`` objects with one-call
    steps sharing the FD plan, the operator plan, the mask box and the vector type)
    queued by one library call (``iga_arnoldi_step_group``): their basis vectors as the
    columns of one multi-right-hand-side FD launch, one grouped operator launch into
    each run's ``w``, then the runs' orthogonalisations (one launch per kernel of the
    step for all of them).  The batch goes on ``batch_stream`` (the first run's stream if
    None), each run's stream ordered after it; ``work`` caches the batch scratch per
    group size (allocated on the batch stream)."""
    t = dv.torch()
    g = len(group)
    if g < 1 or g > 16:
        raise ValueError("a group step takes 1 to 16 runs")
    nv = group[0]._native
    bs = batch_stream if batch_stream is not None else group[0].stream
    for r in group:
        if r._ctx is None:
            raise ValueError("group steps need runs with one-call steps")
        r.prepare_step()
    key = (id(nv.fd_plan), id(nv.op.plan), g)
    if key not in work:
        lib = _lib.load()
        with (t.cuda.stream(bs) if bs is not None else contextlib.nullcontext()):
            work[key] = (dv.workspace(lib.iga_arnoldi_group_workspace_bytes(group[0]._ctx, g)),
                         nv.fd_plan.workspace(g), dv.workspace(nv.op.plan.workspace_bytes(g)))
    xy, fd_work, op_work = work[key]
    ks = [r.launched + 1 for r in group]
    ctxs = (C.c_void_p * g)(*[r._ctx.value for r in group])
    karr = (C.c_int * g)(*ks)
    Vs = (C.c_void_p * g)(*[dv.ptr(r.V) for r in group])
    ldvs = (C.c_int64 * g)(*[r.ldv for r in group])
    streams = (C.c_void_p * g)(*[r._sp() for r in group])
    _lib.call("iga_arnoldi_step_group", g, ctxs, karr, Vs, ldvs, streams, dv.ptr(xy), dv.ptr(fd_work),
              dv.ptr(op_work), dv.stream_ptr(bs))
    for r, k in zip(group, ks):
        r._launched_native(k)


def run_batched(runs, max_group: int = 16, batch_stream=None, lockstep: bool = True):
    """Drive several solves of one shape to completion with their Arnoldi steps
    batched: ``runs`` are ``GmresRun`` or ``RefinementRun`` objects whose one-call
    steps (``native`` / ``native_lo``) share the FD plan and the mask box and whose
    operators share one fused plan (the same operator for several right-hand sides, or
    one per solve).  Every pass, the runs that can take a step take it together (up to
    ``max_group`` per launch): the FD, operator and orthogonalisation launches of a
    group go on ``batch_stream`` (the runs' stream when they all share one), and each
    run's own later work (refinement residuals, solution updates) on its stream is
    ordered after them.  With ``lockstep`` a group step goes out only when every
    run that is still taking Arnoldi steps can take one, so that each FD and operator
    launch serves all of them (the runs' records arrive at slightly different times,
    and launching whatever is ready splits the steps into groups of one or two); no
    run waits for anything but the device, since a run that cannot step is waiting for
    its own queued steps' records.  Everything else (starts, refinement residuals,
    solution updates) stays per run.  Returns the results in order and the host clock
    (``time.perf_counter``) at which each completed."""
    import time

    ctxs = []
    for r in runs:
        inner_native = r.native_lo if isinstance(r, RefinementRun) else getattr(r, "_native", None)
        if inner_native is None or inner_native.fd_plan is None or inner_native.box is None:
            raise ValueError("batched steps need one-call steps with an FD plan and a mask box")
        ctxs.append(inner_native)
    nv0 = ctxs[0]
    for nv in ctxs:
        if nv.fd_plan is not nv0.fd_plan or nv.op.plan is not nv0.op.plan or \
                not all(np.array_equal(a, b) for a, b in zip(nv.box, nv0.box)):
            raise ValueError("batched runs share the FD plan, the operator plan and the mask box")
    t = dv.torch()
    if batch_stream is None:
        if len({id(r.stream) for r in runs}) != 1:
            raise ValueError("runs on different streams need a batch stream")
        batch_stream = runs[0].stream
    stream = batch_stream
    work: dict = {}
    done_at = [None] * len(runs)
    with (t.cuda.stream(stream) if stream is not None else contextlib.nullcontext()):
        while True:
            pending = [i for i, r in enumerate(runs) if not r.done]
            if not pending:
                break
            progressed = False
            for i in pending:
                progressed |= runs[i].advance(launch=False)
                if runs[i].done:
                    done_at[i] = time.perf_counter()
            ready = [s for s in (runs[i].step_run() for i in pending if not runs[i].done) if s is not None]
            if lockstep and len(ready) < sum(1 for i in pending if runs[i].stepping()):
                ready = []
            for j in range(0, len(ready), max_group):
                launch_group(ready[j:j + max_group], work, stream)
                progressed = True
            if not progressed:
                time.sleep(20e-6)  # nothing to queue: the device is busy with the queued steps
    return [r.result for r in runs], done_at
