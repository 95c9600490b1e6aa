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
        version = (self.space_fd.eig_version, adv)
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


class STFDQZPlan(STFDPlan):
    """Device plan of the legacy QZ space-time FD (``iga_stfd_create_qz``)."""

    def __init__(self, n, lo, ln, space_eigvecs, Q, Z, dtype: str):
        self.dtype = dtype
        self.n, self.lo, self.ln = (np.asarray(v, dtype=np.int64) for v in (n, lo, ln))
        na, npt = _lib.i64_array(self.n)
        la, lpt = _lib.i64_array(self.lo)
        lna, lnpt = _lib.i64_array(self.ln)
        keep = [np.ascontiguousarray(U, dtype=np.float64) for U in space_eigvecs]
        arr = (_lib._f64p * 3)(*[k.ctypes.data_as(_lib._f64p) for k in keep])
        qr, qrp = _lib.f64_array(np.real(Q))
        qi, qip = _lib.f64_array(np.imag(Q))
        zr, zrp = _lib.f64_array(np.real(Z))
        zi, zip_ = _lib.f64_array(np.imag(Z))
        h = C.c_void_p()
        _lib.call("iga_stfd_create_qz", C.byref(h), npt, lpt, lnpt, arr, qrp, qip, zrp, zip_, dv.DTYPES[dtype])
        self._h = h
        self.version = None
        self._work = {}

    def set_qz(self, lam_space, S, T, adv: float, version):
        lam, lp = _lib.f64_array(lam_space)
        sr, srp = _lib.f64_array(np.real(S))
        si, sip = _lib.f64_array(np.imag(S))
        tr, trp = _lib.f64_array(np.real(T))
        ti, tip = _lib.f64_array(np.imag(T))
        _lib.call("iga_stfd_set_qz", self._h, lp, srp, sip, trp, tip, float(adv))
        self.version = version


class SpaceTimeQZFD(SpaceTimeFD):
    """The legacy QZ space-time FD (``SpTimeFD`` of ``fastdiag/utils/time_legacy.py:16-171``,
    ``SingleFD.compute_time_schurdecomposition(..., use_arrowhead=False)``) for one DOF per
    node: the complex QZ decomposition ``W_t = Q S Z^H``, ``M_t = Q T Z^H`` of the time
    pencil (host), applied on the device as the arrowhead form with ``Q^H`` / ``Z`` time
    transforms and one upper-triangular solve with ``adv S + lambda T`` per spatial
    eigenvalue in place of the arrowhead solve."""

    def compute_time_schurdecomposition(self, time_quadrule):
        Wt, Mt, nt = time_pencil(time_quadrule)
        S, T, Q, Z = sla.qz(Wt, Mt, output="complex")  # time_legacy.py:42-48
        self.schur_adv, self.schur_mass, self.schur_VSL, self.schur_VSR = S, T, Q, Z
        self._eigenvec_time = Q
        self._eigenblocks_time = []
        self._nt = nt
        self._advection_time_corrector = [1.0] * max(1, self.nbDoFsPerNode)
        self._plans = {}

    def plan(self, dtype: str) -> STFDPlan:
        if self._nt == 0:
            raise RuntimeError("Space-time preconditioner not defined.")
        p = self._plans.get(dtype)
        if p is None:
            n, lo, ln = self.box()
            if len(n) != 4:
                raise NotImplementedError("the space-time FD apply is provided for three space directions")
            p = STFDQZPlan(n, lo, ln, self.space_fd.eigenvec_by_dir_space, self.schur_VSL, self.schur_VSR, dtype)
            self._plans[dtype] = p
        lam = self.space_fd.space_eigenvalues[0]
        adv = float(self._advection_time_corrector[0])
        version = (self.space_fd.eig_version, adv)
        if p.version != version:
            p.set_qz(lam, self.schur_adv, self.schur_mass, adv, version)
        return p


class SpaceTimeOperator:
    """The space-time heat operator ``A x = u_v(c; time_ders=(0, 1)) x +
    gradu_gradv(C; time_ders=(0, 0)) x`` (``examples/performance_matvec_sptm.py:44-57``;
    the terms are ``bsplines_ops.py:290-355`` with the time direction last) on the
    device, coefficients resident in HBM.

    ``c`` is an ``(N_q,)`` array or a scalar; ``C`` a ``(d, d, N_q)`` array or a
    ``(d, d)`` matrix of constants (``d`` space directions).

    With three space directions whose rules suit the fused sweep, the operator is
    applied slice by slice in time: the time trial bases ``B0_t``, ``B1_t`` act first
    (one banded mode product), then, for every time quadrature point, the fused 3-D
    sweep with that point's coefficients (stiffness on the ``B0_t`` slice, mass on
    the ``B1_t`` slice; both terms test with ``W0_t``), and finally the time test
    weights ``W0_t``.  With a symmetric ``C`` the sweep reads both slices in one
    launch (seven coefficient fields); otherwise a general stiffness sweep and a mass
    sweep follow each other.  Otherwise (and with ``fused=False``) the composed device path
    evaluates the sum of terms; there, identically zero fields are left out."""

    def __init__(self, quadrule_list, c=1.0, C=None, dtype: str = "f64", fused: Optional[bool] = None):
        from .operators import fused_applicable

        self.quadrule_list = list(quadrule_list)
        nd = len(self.quadrule_list)
        self.sd = nd - 1
        self.nq = int(np.prod([q.nbquadpts for q in self.quadrule_list], dtype=np.int64))
        self.ndof = int(np.prod([q.nbctrlpts for q in self.quadrule_list], dtype=np.int64))
        self.dtype = dtype
        if C is None:
            C = np.eye(self.sd)
        C = C if isinstance(C, dv.torch().Tensor) else np.asarray(C, dtype=np.float64)
        if tuple(C.shape[:2]) != (self.sd, self.sd):
            raise ValueError("C must be (d, d) or (d, d, N_q) over the space directions")
        can_fuse = nd == 4 and fused_applicable(self.quadrule_list[:3])
        self.fused = can_fuse if fused is None else (bool(fused) and can_fuse)
        if fused and not can_fuse:
            raise NotImplementedError("the sliced fused path needs three space directions with fused-capable rules")
        if self.fused:
            self._build_fused(c, C)
        else:
            self._build_composed(c, C)

    # -- composed ---------------------------------------------------------------
    def _build_composed(self, c, C):
        from .operators import terms_gradu_gradv, terms_u_v

        t = dv.torch()
        tdt = dv.torch_dtype(self.dtype)

        def field(v):
            if isinstance(v, t.Tensor):
                return v.to(device=dv.device(), dtype=tdt).contiguous()
            v = np.asarray(v, dtype=np.float64)
            if v.ndim == 0:
                if float(v) == 0.0:
                    return None
                return t.full((self.nq,), float(v), dtype=tdt, device=dv.device())
            if v.shape != (self.nq,):
                raise ValueError("coefficient field does not match the quadrature points")
            if not np.any(v != 0.0):
                return None
            return dv.to_device(v, self.dtype)

        sd = self.sd
        cfields = [[field(C[i, j]) for j in range(sd)] for i in range(sd)]
        terms = [tm for tm in terms_u_v(self.quadrule_list, field(c), (0, 1)) if tm.coef is not None]
        terms += [tm for tm in terms_gradu_gradv(self.quadrule_list, cfields, (0, 0)) if tm.coef is not None]
        self.terms = terms
        self.ncoef_fields = len(terms)

    # -- fused, sliced in time -----------------------------------------------------
    def _build_fused(self, c, C):
        from .operators import FORM_GEN16, FORM_K6, FORM_KM7X, FORM_M1, band_of, fused_plan

        t = dv.torch()
        qs, qt = self.quadrule_list[:3], self.quadrule_list[3]
        self.nqs = int(np.prod([q.nbquadpts for q in qs], dtype=np.int64))
        self.ncs = int(np.prod([q.nbctrlpts for q in qs], dtype=np.int64))
        self.nqt, self.nct = int(qt.nbquadpts), int(qt.nbctrlpts)
        dev = dv.device()
        # coefficient sources: per-slice device arrays (components comp_stride apart)
        if isinstance(C, t.Tensor) or np.ndim(C) == 3:
            Cd = dv.to_device(C, "f64").reshape(3, 3, -1)
            if Cd.shape[2] != self.nq:
                raise ValueError("coefficient field does not match the quadrature points")
            c_stride, c_slice = self.nq, self.nqs
            sym = bool(t.equal(Cd, Cd.transpose(0, 1)))
        else:
            Cd = t.from_numpy(np.ascontiguousarray(np.broadcast_to(np.asarray(C, float)[:, :, None],
                                                                   (3, 3, self.nqs)))).to(dev)
            c_stride, c_slice = self.nqs, 0
            sym = bool(np.allclose(C, np.asarray(C).T, rtol=0, atol=0))
        if isinstance(c, t.Tensor) or np.ndim(c) == 1:
            md = dv.to_device(c, "f64").reshape(-1)
            if md.numel() != self.nq:
                raise ValueError("coefficient field does not match the quadrature points")
            m_slice = self.nqs
        else:
            md = t.full((self.nqs,), float(c), dtype=t.float64, device=dev)
            m_slice = 0
        nslab = -(-int(qs[2].nbquadpts) // 64)  # fewest slabs: many slices fill the device
        # symmetric C: one sweep with the stiffness on the time-interpolated vector and
        # the mass on the time-differentiated one (form KM7X); otherwise two sweeps
        self.dual = sym
        if sym:
            self.planK = fused_plan(qs, FORM_KM7X, self.dtype, nslab)
            self.planM = None
            self.packK = dv.empty(self.planK.coef_elems * self.nqt, self.dtype)
            self.packM = None
            for s in range(self.nqt):
                self.planK.pack_strided(dv.ptr(Cd) + 8 * s * c_slice, dv.ptr(md) + 8 * s * m_slice, c_stride,
                                        self.packK, s * self.planK.coef_elems)
        else:
            self.planK = fused_plan(qs, FORM_GEN16, self.dtype, nslab)
            self.planM = fused_plan(qs, FORM_M1, self.dtype, nslab)
            self.packK = dv.empty(self.planK.coef_elems * self.nqt, self.dtype)
            self.packM = dv.empty(self.planM.coef_elems * self.nqt, self.dtype)
            zero_m = t.zeros(self.nqs, dtype=t.float64, device=dev)
            for s in range(self.nqt):
                self.planK.pack_strided(dv.ptr(Cd) + 8 * s * c_slice, dv.ptr(zero_m), c_stride, self.packK,
                                        s * self.planK.coef_elems)
                self.planM.pack_strided(0, dv.ptr(md) + 8 * s * m_slice, self.nqs, self.packM,
                                        s * self.planM.coef_elems)
        self.time_trial = band_of(qt.basis[0], qt.basis[1])
        self.time_test = band_of(qt.weights[0])
        self.ncoef_fields = (6 if sym else 16) + 1
        self._work = {}

    def _buffers(self, nrhs: int):
        buf = self._work.get(nrhs)
        if buf is None:
            t = dv.torch()
            tdt = dv.torch_dtype(self.dtype)
            n = self.nqt * self.ncs * nrhs
            buf = (t.empty(n, dtype=tdt, device=dv.device()), t.empty(n, dtype=tdt, device=dv.device()),
                   t.empty(n, dtype=tdt, device=dv.device()), self.planK.workspace(self.nqt),
                   self.planM.workspace(self.nqt) if self.planM is not None else None)
            self._work[nrhs] = buf
        return buf

    def apply_device(self, x, out=None, nrhs: int = 1, stream=None):
        if not self.fused:
            from .operators import composed_apply

            return composed_apply(self.quadrule_list, self.terms, x, nrhs, out=out, stream=stream)
        t = dv.torch()
        X0, X1, Y, wK, wM = self._buffers(nrhs)
        # time trial bases: (n_c,t, N_s, r) -> (n_q,t, N_s, r) twice
        self.time_trial.apply(x, [X0, X1], 1, self.ncs * nrhs, stream=stream)
        es = X0.element_size()
        for r in range(nrhs):
            xd, xs = nrhs, self.ncs * nrhs
            off = r * es
            if self.dual:
                self.planK.apply_sliced2(self.packK, self.planK.coef_elems, dv.ptr(X0) + off, dv.ptr(X1) + off, xd,
                                         xs, dv.ptr(Y) + off, xd, xs, self.nqt, work=wK, stream=stream)
            else:
                self.planK.apply_sliced(self.packK, self.planK.coef_elems, dv.ptr(X0) + off, xd, xs,
                                        dv.ptr(Y) + off, xd, xs, self.nqt, work=wK, stream=stream)
                self.planM.apply_sliced(self.packM, self.planM.coef_elems, dv.ptr(X1) + off, xd, xs,
                                        dv.ptr(Y) + off, xd, xs, self.nqt, accumulate=True, work=wM, stream=stream)
        if out is None:
            out = t.empty(self.ndof * nrhs, dtype=x.dtype, device=x.device)
        self.time_test.apply(Y, [out], 1, self.ncs * nrhs, stream=stream)
        return out

    apply = apply_device

    def algorithmic_bytes(self, nrhs: int = 1) -> int:
        """``b (2 N_dof r + c N_q)``, c = 7 (6 symmetric stiffness + 1 mass fields)."""
        b = 8 if self.dtype == "f64" else 4
        return b * (2 * self.ndof * nrhs + 7 * self.nq)


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
