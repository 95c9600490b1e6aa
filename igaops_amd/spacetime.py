"""Space-time arrowhead fast diagonalization: host set-up, device apply.

Set-up follows ``fastdiag/utils/time_arrow.py:44-97`` and
``fastdiag/utils/time_template.py:46-71``: the time pencil of the Gauss-rewritten
time rule (advection ``W_t = w0 B1``, mass ``M_t = w0 B0``) with the first time DOF
removed (the initial condition); the complex generalized eigenproblem of its leading
block, eigenvectors normalised in ``M``; the M-normalised completion ``[k, rho]`` of
the last column, giving ``U_t`` with ``U_t^H M_t U_t = I`` and
``U_t^H W_t U_t = D_t`` an arrowhead (diagonal plus last row/column ``g``, corner
``sigma``).  The free set is (time DOFs 1..n_t-1) x (spatial free box), time
slowest, so on the device it is one 4-D box.

The apply (``time_arrow.py:111-192``) runs in ``libigaops_hip.so``
(``iga_stfd_apply``): real space transforms and the time transforms as MFMA GEMMs,
and the arrowhead solve (``fastdiag/utils/operations.py:5-93``) in one kernel with
one thread per spatial eigen-index.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional

import numpy as np
import scipy.linalg as sla

from . import _lib
from . import device as dv
from .fastdiag import SpaceFD, rewrite_quadrature


def time_pencil(time_quadrule):
    """``(W_t, M_t)`` without the first time DOF (``time_template.py:46-59``)."""
    q = rewrite_quadrature([time_quadrule])[0]
    adv = (q.weights[1] @ q.basis[1]).toarray()
    mass = (q.weights[0] @ q.basis[0]).toarray()
    keep = np.arange(1, q.nbctrlpts)
    return adv[np.ix_(keep, keep)], mass[np.ix_(keep, keep)], q.nbctrlpts


def arrowhead_factors(Wt: np.ndarray, Mt: np.ndarray):
    """``time_arrow.py:44-80``: ``U_t``, ``diag(D_t)`` (eigenvalues then ``sigma``) and ``g``."""
    lead_w, lead_m = Wt[:-1, :-1], Mt[:-1, :-1]
    evals, evecs = sla.eig(lead_w, lead_m)
    mnorm = np.sqrt(np.abs(np.einsum("ki,kl,li->i", evecs.conj(), lead_m, evecs)))
    evecs = evecs / mnorm[None, :]
    # M-orthogonal completion of the last column: [v, 1] with M_p v = -M[:-1, -1]
    ext = np.append(np.linalg.solve(lead_m, -Mt[:-1, -1]), 1.0)
    ext = ext / np.sqrt(ext @ Mt @ ext)
    n = Wt.shape[0]
    Ut = np.zeros((n, n), dtype=complex)
    Ut[:-1, :-1] = evecs
    Ut[:, -1] = ext
    g = evecs.conj().T @ (Wt[:-1, :] @ ext)
    sigma = ext @ Wt @ ext
    return Ut, np.append(evals, sigma), g


class STFDPlan:
    """Device plan of the space-time FD (``iga_stfd_t``)."""

    def __init__(self, n, lo, ln, space_eigvecs, Ut, dtype: str):
        self.dtype = dtype
        self.n, self.lo, self.ln = (np.asarray(v, dtype=np.int64) for v in (n, lo, ln))
        na, npt = _lib.i64_array(self.n)
        la, lpt = _lib.i64_array(self.lo)
        lna, lnpt = _lib.i64_array(self.ln)
        keep = [np.ascontiguousarray(U, dtype=np.float64) for U in space_eigvecs]
        arr = (_lib._f64p * 3)(*[k.ctypes.data_as(_lib._f64p) for k in keep])
        ure, upr = _lib.f64_array(np.real(Ut))
        uim, upi = _lib.f64_array(np.imag(Ut))
        h = C.c_void_p()
        _lib.call("iga_stfd_create", C.byref(h), npt, lpt, lnpt, arr, upr, upi, dv.DTYPES[dtype])
        self._h = h
        self.version = None
        self._work = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib.load().iga_stfd_destroy(h)
            self._h = None

    def set_blocks(self, lam_space, D, g, adv: float, version):
        lam, lp = _lib.f64_array(lam_space)
        dr, drp = _lib.f64_array(np.real(D))
        di, dip = _lib.f64_array(np.imag(D))
        gr, grp = _lib.f64_array(np.real(g))
        gi, gip = _lib.f64_array(np.imag(g))
        _lib.call("iga_stfd_set_blocks", self._h, lp, drp, dip, grp, gip, float(adv))
        self.version = version

    def apply(self, x, out, nrhs: int = 1, transpose: bool = False, stream=None):
        w = self._work.get(nrhs)
        if w is None:
            w = dv.workspace(_lib.load().iga_stfd_workspace_bytes(self._h, int(nrhs)))
            self._work[nrhs] = w
        _lib.call("iga_stfd_apply", self._h, dv.ptr(x), dv.ptr(out), int(nrhs), int(bool(transpose)), dv.ptr(w),
                  dv.stream_ptr(stream))
        return out


class SpaceTimeFD:
    """``SpTimeFD`` (``fastdiag/utils/time_arrow.py:16-192``) for one DOF per node."""

    def __init__(self, space_fd: SpaceFD):
        self.space_fd = space_fd
        self._eigenvec_time = np.array([])
        self._eigenblocks_time: List[np.ndarray] = []
        self._advection_time_corrector = [1.0]
        self._nt = 0
        self._plans = {}

    @property
    def nbDoFsPerNode(self):
        return self.space_fd.nbDoFsPerNode

    @property
    def eigenvec_time(self):
        return self._eigenvec_time

    @property
    def eigenblocks_time(self):
        return self._eigenblocks_time

    @property
    def advection_time_corrector(self):
        return self._advection_time_corrector

    def compute_time_schurdecomposition(self, time_quadrule):
        Wt, Mt, nt = time_pencil(time_quadrule)
        Ut, D, g = arrowhead_factors(Wt, Mt)
        self._eigenvec_time = Ut
        self._eigenblocks_time = [D, g]
        self._nt = nt
        self._advection_time_corrector = [1.0] * max(1, self.nbDoFsPerNode)
        self._plans = {}

    def add_scalar_time_correctors(self, advection_corrector: Optional[List[float]] = None):
        """``time_arrow.py:99-109``."""
        if isinstance(advection_corrector, (list, tuple)) and len(advection_corrector) == self.nbDoFsPerNode:
            self._advection_time_corrector = list(advection_corrector)

    def invalidate(self):
        for p in self._plans.values():
            p.version = None

    @property
    def sptm_free_nodes(self) -> List[np.ndarray]:
        """``time_template.py:61-71``: (time 1..n_t-1) x (spatial free nodes), time slowest."""
        space = np.asarray(self.space_fd.space_free_nodes[0], dtype=np.int64)
        t = np.arange(1, self._nt, dtype=np.int64)
        return [(t[:, None] * self.space_fd.sp_nnz + space[None, :]).ravel()]

    def box(self):
        n, lo, ln = self.space_fd.box()
        return (np.append(n, self._nt), np.append(lo, 1), np.append(ln, self._nt - 1))

    def plan(self, dtype: str) -> STFDPlan:
        if self._nt == 0:
            raise RuntimeError("Space-time preconditioner not defined.")
        p = self._plans.get(dtype)
        if p is None:
            n, lo, ln = self.box()
            if len(n) != 4:
                raise NotImplementedError("the space-time FD apply is provided for three space directions")
            p = STFDPlan(n, lo, ln, self.space_fd.eigenvec_by_dir_space, self._eigenvec_time, dtype)
            self._plans[dtype] = p
        lam = self.space_fd.space_eigenvalues[0]
        adv = float(self._advection_time_corrector[0])
        version = (id(lam), adv)
        if p.version != version:
            D, g = self._eigenblocks_time
            p.set_blocks(lam, D, g, adv, version)
        return p

    def apply_device(self, x, out=None, nrhs: int = 1, stream=None, transpose: bool = False):
        t = dv.torch()
        dtype = "f32" if x.dtype == t.float32 else "f64"
        if out is None:
            out = t.empty_like(x)
        return self.plan(dtype).apply(x, out, nrhs, transpose, stream)

    def apply_transpose_device(self, x, out=None, nrhs: int = 1, stream=None):
        return self.apply_device(x, out, nrhs, stream, transpose=True)


class SpaceTimeOperator:
    """The space-time heat operator ``A x = u_v(c; time_ders=(0, 1)) x +
    gradu_gradv(C; time_ders=(0, 0)) x`` (``examples/performance_matvec_sptm.py:44-57``;
    the terms are ``bsplines_ops.py:290-355`` with the time direction last) on the
    device, coefficients resident in HBM.

    ``c`` is an ``(N_q,)`` array or a scalar; ``C`` a ``(d, d, N_q)`` array or a
    ``(d, d)`` matrix of constants (``d`` space directions).  Coefficient fields that
    are identically zero are left out of the sum (they contribute nothing)."""

    def __init__(self, quadrule_list, c=1.0, C=None, dtype: str = "f64"):
        from .operators import Term, terms_gradu_gradv, terms_u_v

        t = dv.torch()
        self.quadrule_list = list(quadrule_list)
        nd = len(self.quadrule_list)
        sd = nd - 1
        self.nq = int(np.prod([q.nbquadpts for q in self.quadrule_list], dtype=np.int64))
        self.ndof = int(np.prod([q.nbctrlpts for q in self.quadrule_list], dtype=np.int64))
        self.dtype = dtype
        tdt = t.float64 if dtype == "f64" else t.float32
        if C is None:
            C = np.eye(sd)

        def field(v):
            v = np.asarray(v, dtype=np.float64)
            if v.ndim == 0:
                if float(v) == 0.0:
                    return None
                return t.full((self.nq,), float(v), dtype=tdt, device=dv.device())
            if v.shape != (self.nq,):
                raise ValueError("coefficient field does not match the quadrature points")
            if not np.any(v != 0.0):
                return None
            return dv.to_device(v, dtype)

        C = np.asarray(C, dtype=np.float64)
        if C.shape[:2] != (sd, sd):
            raise ValueError("C must be (d, d) or (d, d, N_q) over the space directions")
        cfields = [[field(C[i, j]) for j in range(sd)] for i in range(sd)]
        terms = [tm for tm in terms_u_v(self.quadrule_list, field(c), (0, 1)) if tm.coef is not None]
        terms += [tm for tm in terms_gradu_gradv(self.quadrule_list, cfields, (0, 0)) if tm.coef is not None]
        self.terms: List[Term] = terms

    def apply_device(self, x, out=None, nrhs: int = 1, stream=None):
        from .operators import composed_apply

        return composed_apply(self.quadrule_list, self.terms, x, nrhs, out=out, stream=stream)

    apply = apply_device

    def algorithmic_bytes(self, nrhs: int = 1) -> int:
        """``b (2 N_dof r + c N_q)`` with ``c`` = the resident coefficient fields."""
        b = 8 if self.dtype == "f64" else 4
        return b * (2 * self.ndof * nrhs + len(self.terms) * self.nq)


class BoxMasked:
    """``out = mask(A mask(x))`` for a device operator, the free set a box (the
    cleandod wrappers of ``solvers/linear_solver.py:119-124``); the input is copied,
    not masked in place."""

    def __init__(self, op, box):
        self.op = op
        self.box = box

    def _mask(self, v, nrhs, stream=None):
        n, lo, ln = self.box
        na, npt = _lib.i64_array(n)
        la, lpt = _lib.i64_array(lo)
        lna, lnpt = _lib.i64_array(ln)
        _lib.call("iga_box_mask", dv.iga_dtype(v), len(n), npt, lpt, lnpt, int(nrhs), dv.ptr(v), dv.stream_ptr(stream))

    def apply_device(self, x, out=None, nrhs: int = 1, stream=None):
        xm = x.clone()
        self._mask(xm, nrhs, stream)
        out = self.op.apply_device(xm, out=out, nrhs=nrhs, stream=stream)
        self._mask(out, nrhs, stream)
        return out
