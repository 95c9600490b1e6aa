"""Matrix-free operators on the GPU, behind the reference's call signatures.

* ``MatrixFree.apply(matrix_list, array_in, is_transpose=False)``
  (``operators/utils/matrix_free.py:77-119``): d banded mode products on the device.
* ``BsplineOperations`` / ``NurbsOperations`` ``.compute_mf_scalar_{u_v, gradu_gradv,
  gradu_v, u_gradv}(quadrule_list, coefficients, array_in, time_ders=None,
  nurbs_weights=np.array([]))`` (``operators/utils/template.py:312-454``,
  ``operators/bsplines_ops.py:290-429``, ``operators/nurbs_ops.py:320-463``): same
  arguments and results; numpy inputs give numpy outputs (uploaded per call), torch
  device tensors stay on the device.
* ``StiffnessMassOperator``: a device-resident plan for ``y = a K x + b M x`` (the
  benchmark matvec), coefficients packed once; used by the solvers and the server.

Two execution paths, both HIP:
  - fused (``csrc/fused.hip``): 3-D, rules whose test weights depend only on the test
    derivative (Gauss, WQ type "2"), one pass over the coefficient stream;
  - composed (``csrc/lmode.hip``): any dimension / time derivatives / WQ type "1",
    mirroring the reference's sequence of L-mode products and point-wise products
    (terms with the same trial interpolation share it, terms with the same test
    weights share the test contraction).
"""
from __future__ import annotations

import copy
import ctypes as C
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
from scipy import sparse as sp

from . import _lib
from . import device as dv
from .bands import common_band, fused_direction_tables, test_weights_depend_on_alpha_only

# ---------------------------------------------------------------------------
# Device bands
# ---------------------------------------------------------------------------


class Band:
    """Device copy of ``nmat`` matrices sharing one band (``iga_band_t``)."""

    def __init__(self, mats: Sequence):
        first, width, vals = common_band([_as_csr(m) for m in mats])
        self.nrows, self.ncols = _as_csr(mats[0]).shape
        self.width = width
        self.nmat = len(mats)
        h = C.c_void_p()
        fa, fp = _lib.i32_array(first)
        va, vp = _lib.f64_array(vals)
        _lib.call("iga_band_create", C.byref(h), self.nrows, self.ncols, width, self.nmat, fp, vp)
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib.load().iga_band_destroy(h)
            self._h = None

    def apply(self, x, ys: Sequence, outer: int, inner: int, accumulate: bool = False, alpha: float = 1.0, stream=None):
        arr = (C.c_void_p * 4)(*[dv.ptr(y) for y in ys] + [0] * (4 - len(ys)))
        _lib.call(
            "iga_band_mode_product",
            self._h,
            dv.iga_dtype(x),
            int(outer),
            int(inner),
            dv.ptr(x),
            arr,
            int(bool(accumulate)),
            float(alpha),
            dv.stream_ptr(stream),
        )


def _as_csr(m):
    return sp.csr_array(m) if not isinstance(m, np.ndarray) else sp.csr_array(np.asarray(m))


_BAND_CACHE: Dict[Tuple, Tuple[list, Band]] = {}


def band_of(*mats, transpose: bool = False) -> Band:
    """Cached device band of the given matrices (kept alive with the band)."""
    key = (tuple(id(m) for m in mats), transpose)
    hit = _BAND_CACHE.get(key)
    if hit is not None:
        return hit[1]
    src = [(_as_csr(m).T if transpose else m) for m in mats]
    b = Band(src)
    _BAND_CACHE[key] = (list(mats), b)
    return b


def clear_caches():
    _BAND_CACHE.clear()
    _FUSED_CACHE.clear()


# ---------------------------------------------------------------------------
# MatrixFree.apply on the device
# ---------------------------------------------------------------------------


def mf_apply_device(bands: Sequence[Band], x, nrhs: int, out=None, accumulate=False, alpha=1.0, stream=None):
    """``(M_d (x) ... (x) M_1) x`` for ``bands = [M_1, .., M_d]`` on an ``(N, nrhs)``
    device array.  Directions are processed slowest first, like the reference."""
    t = dv.torch()
    cur = [b.ncols for b in bands]
    tensor = x
    nd = len(bands)
    for d in reversed(range(nd)):
        b = bands[d]
        outer = int(np.prod(cur[d + 1:], dtype=np.int64))
        inner = int(np.prod(cur[:d], dtype=np.int64)) * nrhs
        new = list(cur)
        new[d] = b.nrows
        last = d == 0
        n_out = int(np.prod(new, dtype=np.int64)) * nrhs
        if last and out is not None:
            y = out
        else:
            y = t.empty(n_out, dtype=x.dtype, device=x.device)
        b.apply(tensor, [y], outer, inner, accumulate=(accumulate and last), alpha=(alpha if last else 1.0), stream=stream)
        tensor = y
        cur = new
    return tensor


def _to_dev_2d(array_in, dtype="f64"):
    """Device (N*r,) view of an (N,) or (N, r) input; returns (tensor, r, was_numpy, shape)."""
    t = dv.torch()
    was_np = not isinstance(array_in, t.Tensor)
    shape = tuple(array_in.shape)
    r = int(np.prod(shape[1:], dtype=np.int64)) if len(shape) > 1 else 1
    x = dv.to_device(array_in, dtype).reshape(-1)
    return x, r, was_np, shape


def _from_dev(y, was_np, shape):
    y = y.reshape(shape)
    if was_np:
        return y.double().cpu().numpy()
    return y


class MatrixFree:
    """``operators/utils/matrix_free.py:7-119`` on the device."""

    @staticmethod
    def apply(matrix_list, array_in, is_transpose: bool = False):
        dv.require_gpu()
        t = dv.torch()
        if isinstance(array_in, t.Tensor):
            dtype = "f32" if array_in.dtype == t.float32 else "f64"
        else:
            if np.iscomplexobj(array_in) or any(np.iscomplexobj(m) for m in matrix_list if isinstance(m, np.ndarray)):
                raise NotImplementedError("complex MatrixFree.apply: use the real and imaginary parts")
            dtype = "f64"
        x, r, was_np, shape = _to_dev_2d(array_in, dtype)
        bands = [band_of(m, transpose=is_transpose) for m in matrix_list]
        if x.numel() != int(np.prod([b.ncols for b in bands])) * r:
            raise ValueError("array_in does not match the Kronecker product")
        y = mf_apply_device(bands, x, r)
        out_shape = (int(np.prod([b.nrows for b in bands])),) + tuple(shape[1:])
        return _from_dev(y, was_np, out_shape)


# ---------------------------------------------------------------------------
# Composed operator path
# ---------------------------------------------------------------------------


class Term:
    """``y += W^{test} diag(coef) B^{trial} x``: per-direction basis index ``trial``,
    weights index ``test``, coefficient device array ``coef`` (N_q,)."""

    __slots__ = ("trial", "test", "coef")

    def __init__(self, trial, test, coef):
        self.trial = tuple(int(v) for v in trial)
        self.test = tuple(int(v) for v in test)
        self.coef = coef


def composed_apply(quadrule_list, terms: List[Term], x, nrhs: int, out=None, stream=None):
    """Evaluate ``sum_terms`` with shared interpolations / test contractions."""
    t = dv.torch()
    nq = int(np.prod([q.nbquadpts for q in quadrule_list], dtype=np.int64))
    ncdof = int(np.prod([q.nbctrlpts for q in quadrule_list], dtype=np.int64))
    interp: Dict[Tuple, object] = {}
    acc: Dict[Tuple, object] = {}
    dt = dv.iga_dtype(x)
    for term in terms:
        if term.trial not in interp:
            bands = [band_of(q.basis[b]) for q, b in zip(quadrule_list, term.trial)]
            interp[term.trial] = mf_apply_device(bands, x, nrhs, stream=stream)
        u = interp[term.trial]
        first = term.test not in acc
        if first:
            acc[term.test] = t.empty(nq * nrhs, dtype=x.dtype, device=x.device)
        _lib.call(
            "iga_qp_scale",
            dt,
            nq,
            nrhs,
            dv.ptr(term.coef),
            dv.ptr(u),
            dv.ptr(acc[term.test]),
            0 if first else 1,
            1.0,
            dv.stream_ptr(stream),
        )
    y = out if out is not None else t.zeros(ncdof * nrhs, dtype=x.dtype, device=x.device)
    started = out is None
    for i, (test, a) in enumerate(acc.items()):
        bands = [band_of(q.weights[z]) for q, z in zip(quadrule_list, test)]
        mf_apply_device(bands, a, nrhs, out=y, accumulate=(i > 0 or started), stream=stream)
    if not acc and out is not None:
        y.zero_()
    return y


def _spacetime_dims(quadrule_list, time_ders):
    nd = len(quadrule_list)
    return nd, (nd - 1 if time_ders is not None else nd)


def terms_u_v(quadrule_list, c, time_ders=None):
    """``bsplines_ops.py:290-317``."""
    nd, _ = _spacetime_dims(quadrule_list, time_ders)
    tr = [0] * nd
    te = [0] * nd
    if time_ders is not None:
        tr[-1] = time_ders[1]
        te[-1] = time_ders[0]
    return [Term(tr, te, c)]


def terms_gradu_gradv(quadrule_list, C, time_ders=None):
    """``bsplines_ops.py:319-355`` (``C[i][j]``: test d_i, trial d_j)."""
    nd, sd = _spacetime_dims(quadrule_list, time_ders)
    out = []
    for j in range(sd):
        beta = [0] * nd
        beta[j] = 1
        if time_ders is not None:
            beta[-1] = time_ders[1]
        for i in range(sd):
            alpha = [0] * nd
            alpha[i] = 1
            if time_ders is not None:
                alpha[-1] = time_ders[0]
            out.append(Term(beta, [b + 2 * a for a, b in zip(alpha, beta)], C[i][j]))
    return out


def terms_gradu_v(quadrule_list, c, time_ders=None):
    """``bsplines_ops.py:357-390``: test d_i, trial value."""
    nd, sd = _spacetime_dims(quadrule_list, time_ders)
    beta = [0] * nd
    if time_ders is not None:
        beta[-1] = time_ders[1]
    out = []
    for i in range(sd):
        alpha = [0] * nd
        alpha[i] = 1
        if time_ders is not None:
            alpha[-1] = time_ders[0]
        out.append(Term(beta, [b + 2 * a for a, b in zip(alpha, beta)], c[i]))
    return out


def terms_u_gradv(quadrule_list, c, time_ders=None):
    """``bsplines_ops.py:392-429``: trial d_j, test weights index ``alpha`` (value)."""
    nd, sd = _spacetime_dims(quadrule_list, time_ders)
    alpha = [0] * nd
    if time_ders is not None:
        alpha[-1] = time_ders[0]
    out = []
    for i in range(sd):
        beta = [0] * nd
        beta[i] = 1
        if time_ders is not None:
            beta[-1] = time_ders[1]
        out.append(Term(beta, alpha, c[i]))
    return out


# ---------------------------------------------------------------------------
# NURBS weight projection (device)
# ---------------------------------------------------------------------------


def nurbs_weight_fields(quadrule_list, w_dev):
    """``W`` and ``grad W`` at the quadrature points (``nurbs_ops.py:14-62``), float64."""
    nd = len(quadrule_list)
    W = mf_apply_device([band_of(q.basis[0]) for q in quadrule_list], w_dev, 1)
    dW = []
    for j in range(nd):
        bands = [band_of(q.basis[1 if d == j else 0]) for d, q in enumerate(quadrule_list)]
        dW.append(mf_apply_device(bands, w_dev, 1))
    return W, dW


# ---------------------------------------------------------------------------
# Fused plans
# ---------------------------------------------------------------------------

FORM_M1, FORM_K6, FORM_KM7, FORM_SYM10, FORM_GEN16 = 0, 1, 2, 3, 4
FORM_KM7X = 5  # the KM7 fields, value (mass) term on a second input
FORM_GV4, FORM_UG4 = 6, 7  # first-order terms: gradient on the test / on the trial function
FORM_NCOEF = {0: 1, 1: 6, 2: 7, 3: 10, 4: 16, 5: 7, 6: 4, 7: 4}


class FusedPlan:
    """Device plan of ``csrc/fused.hip`` for three 1-D rules, one form and dtype."""

    def __init__(self, quadrule_list, form: int, dtype: str = "f64", nslab: int = 0):
        if len(quadrule_list) != 3:
            raise ValueError("the fused operator is three-dimensional")
        tabs = [fused_direction_tables(q) for q in quadrule_list]
        P = max(t[1] for t in tabs)
        tabs = [fused_direction_tables(q, P) if t[1] < P else t for q, t in zip(quadrule_list, tabs)]
        self.P = P
        self.form = form
        self.dtype = dtype
        self.nc = np.array([q.nbctrlpts for q in quadrule_list], dtype=np.int32)
        self.nq = np.array([q.nbquadpts for q in quadrule_list], dtype=np.int32)
        self._keep = []
        firsts = (_lib._i32p * 3)()
        tabps = (_lib._f64p * 3)()
        for d, (first, _, vals) in enumerate(tabs):
            fa, fp = _lib.i32_array(first)
            va, vp = _lib.f64_array(vals)
            self._keep += [fa, va]
            firsts[d] = fp
            tabps[d] = vp
        nca, ncp = _lib.i32_array(self.nc)
        nqa, nqp = _lib.i32_array(self.nq)
        h = C.c_void_p()
        _lib.call("iga_fused_create", C.byref(h), ncp, nqp, P, firsts, tabps, form, dv.DTYPES[dtype], int(nslab))
        self._h = h
        ce, pe = C.c_int64(), C.c_int64()
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _lib.call("iga_fused_info", h, C.byref(ce), C.byref(pe), C.byref(a), C.byref(b), C.byref(c))
        self.coef_elems = ce.value
        self.part_elems = pe.value
        self.tiles = (a.value, b.value)
        self.nslab = c.value
        self.rhs_block = 1
        self._work = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib.load().iga_fused_destroy(h)
            self._h = None

    def set_rhs_block(self, nr: int):
        """Right-hand sides per work-group when several share the coefficients (1, the
        default, or 2: each coefficient load and plane-phase barrier then serves two;
        the results agree to rounding)."""
        _lib.call("iga_fused_set_rhs_block", self._h, int(nr))
        self.rhs_block = int(nr)

    def workspace_bytes(self, nrhs: int) -> int:
        return int(_lib.load().iga_fused_workspace_bytes(self._h, int(nrhs)))

    def workspace(self, nrhs: int):
        w = self._work.get(nrhs)
        if w is None:
            w = dv.workspace(self.workspace_bytes(nrhs))
            self._work[nrhs] = w
        return w

    def pack(self, C=None, m=None, W=None, dW=None, stream=None):
        """Packed device coefficients from reference-layout float64 device arrays."""
        out = dv.empty(self.coef_elems, self.dtype)
        _lib.call(
            "iga_fused_pack",
            self._h,
            dv.ptr(C),
            dv.ptr(m),
            dv.ptr(W),
            dv.ptr(dW),
            dv.ptr(out),
            dv.stream_ptr(stream),
        )
        return out

    def pack_extruded(self, ext, field=None, with_mass=False, W=None, dW=None, stream=None):
        """Packed device coefficients of an extruded patch generated point by point
        (``iga_fused_pack_extruded``): ``ext`` from ``ExtrudedPatch.device_extrusion``,
        ``field`` from ``ExtrudedPatch.device_field_modes`` (``None``: ``kappa = 1``),
        ``W`` (n_plane) and ``dW`` (2 n_plane, concatenated) the NURBS weight function
        and its in-plane gradient on the cross-section grid."""
        out = dv.empty(self.coef_elems, self.dtype)
        H = Ez = None
        modes, sigma = 0, 0.0
        if field is not None:
            H, Ez, modes, sigma = field["H"], field["Ez"], field["modes"], field["sigma"]
        _lib.call("iga_fused_pack_extruded", self._h, dv.ptr(ext["plane"]), dv.ptr(ext["dz"]), dv.ptr(H), dv.ptr(Ez),
                  int(modes), float(sigma), int(bool(with_mass)), dv.ptr(W), dv.ptr(dW), dv.ptr(out),
                  dv.stream_ptr(stream))
        return out

    def apply_stage(self, stage, packed, x, y, nrhs=1, nurbs_w=None, alpha=1.0, accumulate=False, work=None, stream=None):
        """One of the two launches of ``apply`` (1: coefficient sweep, 2: gather)."""
        work = work if work is not None else self.workspace(nrhs)
        _lib.call("iga_fused_apply_stage", self._h, int(stage), dv.ptr(packed), dv.ptr(x), dv.ptr(y), int(nrhs),
                  dv.ptr(nurbs_w), float(alpha), int(bool(accumulate)), dv.ptr(work), dv.stream_ptr(stream))
        return y

    def pack_strided(self, C_ptr: int, m_ptr: int, comp_stride: int, out, out_offset: int = 0, stream=None):
        """Pack one slice of larger coefficient arrays (raw device addresses of its
        first component; components ``comp_stride`` elements apart) into
        ``out[out_offset:]``."""
        es = out.element_size()
        _lib.call("iga_fused_pack_strided", self._h, int(C_ptr), int(m_ptr), 0, 0, int(comp_stride),
                  dv.ptr(out) + int(out_offset) * es, dv.stream_ptr(stream))
        return out

    def apply_sliced(self, packed, coef_slice_stride: int, x_ptr: int, x_dstride: int, x_sstride: int, y_ptr: int,
                     y_dstride: int, y_sstride: int, nslices: int, alpha=1.0, accumulate=False, work=None,
                     stream=None):
        """``nslices`` 3-D applies with per-slice coefficients and strided vectors."""
        work = work if work is not None else self.workspace(nslices)
        _lib.call("iga_fused_apply_sliced", self._h, dv.ptr(packed), int(coef_slice_stride), int(x_ptr),
                  int(x_dstride), int(x_sstride), int(y_ptr), int(y_dstride), int(y_sstride), int(nslices), 0,
                  float(alpha), int(bool(accumulate)), dv.ptr(work), dv.stream_ptr(stream))

    def apply_sliced2(self, packed, coef_slice_stride: int, x_ptr: int, x1_ptr: int, x_dstride: int, x_sstride: int,
                      y_ptr: int, y_dstride: int, y_sstride: int, nslices: int, alpha=1.0, accumulate=False,
                      work=None, stream=None):
        """``apply_sliced`` of a FORM_KM7X plan: the stiffness on ``x``, the mass on
        ``x1`` (same strides), one sweep."""
        work = work if work is not None else self.workspace(nslices)
        _lib.call("iga_fused_apply_sliced2", self._h, dv.ptr(packed), int(coef_slice_stride), int(x_ptr),
                  int(x1_ptr), int(x_dstride), int(x_sstride), int(y_ptr), int(y_dstride), int(y_sstride),
                  int(nslices), 0, float(alpha), int(bool(accumulate)), dv.ptr(work), dv.stream_ptr(stream))

    def apply(self, packed, x, y, nrhs: int = 1, nurbs_w=None, alpha=1.0, accumulate=False, work=None, stream=None):
        work = work if work is not None else self.workspace(nrhs)
        _lib.call(
            "iga_fused_apply",
            self._h,
            dv.ptr(packed),
            dv.ptr(x),
            dv.ptr(y),
            int(nrhs),
            dv.ptr(nurbs_w),
            float(alpha),
            int(bool(accumulate)),
            dv.ptr(work),
            dv.stream_ptr(stream),
        )
        return y


_FUSED_CACHE: Dict[Tuple, Tuple[list, FusedPlan]] = {}

# right-hand sides from which StiffnessMassOperator.apply takes the fewest-slab plan
BATCH_PLAN_MIN_RHS = 4


def fused_plan(quadrule_list, form, dtype="f64", nslab=0) -> FusedPlan:
    key = (tuple(id(q) for q in quadrule_list), form, dtype, nslab)
    hit = _FUSED_CACHE.get(key)
    if hit is None:
        hit = (list(quadrule_list), FusedPlan(quadrule_list, form, dtype, nslab))
        _FUSED_CACHE[key] = hit
    return hit[1]


def fused_applicable(quadrule_list, time_ders=None) -> bool:
    return (
        len(quadrule_list) == 3
        and time_ders is None
        and all(test_weights_depend_on_alpha_only(q) for q in quadrule_list)
        and all(2 <= q.degree <= 7 for q in quadrule_list)
    )


# ---------------------------------------------------------------------------
# Device operator: y = a K x + b M x
# ---------------------------------------------------------------------------


def _is_symmetric(Cd) -> bool:
    t = dv.torch()
    return bool(t.equal(Cd, Cd.transpose(0, 1)))


class StiffnessMassOperator:
    """Device-resident ``y = K x + M x`` (either part may be omitted).

    ``C`` (3,3,N_q) stiffness coefficients (``C[i][j]``: test d_i, trial d_j), ``m``
    (N_q,) mass coefficients, ``nurbs_weights`` (N_dof,) for a NURBS patch.  Inputs may
    be numpy or device tensors; they are uploaded / packed once.  ``dtype`` is the
    arithmetic of the apply ("f64" or "f32")."""

    def __init__(self, quadrule_list, C=None, m=None, nurbs_weights=None, dtype="f64", force_composed=False, nslab=0,
                 symmetric=None, nurbs_fields=None):
        dv.require_gpu()
        t = dv.torch()
        self.quadrule_list = list(quadrule_list)
        self.dtype = dtype
        self.ndof = int(np.prod([q.nbctrlpts for q in quadrule_list]))
        self.nq = int(np.prod([q.nbquadpts for q in quadrule_list]))
        self.is_nurbs = nurbs_weights is not None and int(
            nurbs_weights.numel() if isinstance(nurbs_weights, t.Tensor) else np.size(nurbs_weights)) > 0
        Cd = None if C is None else dv.to_device(C, "f64").reshape(3, 3, -1)
        md = None if m is None else dv.to_device(m, "f64").reshape(-1)
        self.w64 = dv.to_device(nurbs_weights, "f64") if self.is_nurbs else None
        self.w = None if not self.is_nurbs else self.w64.to(dv.torch_dtype(dtype))
        self.fused = (not force_composed) and fused_applicable(self.quadrule_list)
        if self.fused:
            # `symmetric` skips the check (a device comparison the host waits for)
            sym = Cd is None or (bool(symmetric) if symmetric is not None else _is_symmetric(Cd))
            if self.is_nurbs:
                form = FORM_SYM10 if sym else FORM_GEN16
            elif Cd is None:
                form = FORM_M1
            elif md is None:
                form = FORM_K6 if sym else FORM_GEN16
            else:
                form = FORM_KM7 if sym else FORM_GEN16
            if form == FORM_GEN16 and md is None:
                md = t.zeros(self.nq, dtype=t.float64, device=Cd.device)
            if form == FORM_M1 and md is None:
                raise ValueError("no coefficients")
            self.plan = fused_plan(self.quadrule_list, form, dtype, nslab)
            self._auto_slabs = nslab <= 0
            W = dW = None
            if self.is_nurbs:
                Wf, dWf = nurbs_fields if nurbs_fields is not None else nurbs_weight_fields(self.quadrule_list,
                                                                                            self.w64)
                W, dW = Wf, t.cat(dWf)
                if Cd is None:
                    Cd = t.zeros((3, 3, self.nq), dtype=t.float64, device=W.device)
                if md is None:
                    md = t.zeros(self.nq, dtype=t.float64, device=W.device)
            self.packed = self.plan.pack(Cd, md, W, dW)
            self.form = form
        else:
            self.terms = self._composed_terms(Cd, md)
        self.coef_count = FORM_NCOEF[self.form] if self.fused else len(self.terms)
        self._work = {}

    @classmethod
    def from_extruded(cls, patch, field=None, dtype: str = "f64", mass: bool = False, nurbs_w64=None, nurbs_w=None,
                      stream=None):
        """``K (+ M)`` of an extruded patch (``geometry.ExtrudedPatch``) with the
        conductivity ``kappa detJ J^+ J^+T``, ``kappa`` the lognormal field of
        ``patch.device_field_modes`` (or 1 when ``field`` is None): the same operator as
        ``StiffnessMassOperator(qs, *patch.device_coefficients(...), nurbs_weights=...)``,
        with the coefficients generated and packed in one launch instead of going
        through a (3, 3, N_q) array.  For a NURBS cross-section the weight function and
        its gradient come from the cross-section grid (``patch.device_extrusion``), not
        from N_q-sized fields.  ``nurbs_w64`` (the control-point weights on the device,
        float64) and ``nurbs_w`` (the same in ``dtype``, shared by the operators of one
        shape) may be passed when cached by the caller."""
        dv.require_gpu()
        self = cls.__new__(cls)
        self.quadrule_list = list(patch.quadrule_list)
        self.dtype = dtype
        self.ndof = patch.nbctrlpts_total
        self.nq = patch.nbquadpts_total
        self.is_nurbs = patch.is_nurbs
        self.w64 = None
        if self.is_nurbs:
            self.w64 = nurbs_w64 if nurbs_w64 is not None else dv.upload(np.asarray(patch.nurbs_weights, dtype=float))
        self.w = None
        if self.is_nurbs:
            self.w = nurbs_w if nurbs_w is not None else self.w64.to(dv.torch_dtype(dtype))
        if not fused_applicable(self.quadrule_list):
            raise NotImplementedError("generated coefficients need the fused operator")
        self.fused = True
        if self.is_nurbs:
            self.form = FORM_SYM10
        else:
            self.form = FORM_KM7 if mass else FORM_K6
        self.plan = fused_plan(self.quadrule_list, self.form, dtype, 0)
        ext = patch.device_extrusion()
        W = dW = None
        if self.is_nurbs:
            npl = ext["nplane"]
            W, dW = ext["nurbs"][:npl], ext["nurbs"][npl:]
        self.packed = self.plan.pack_extruded(ext, field, mass, W, dW, stream=stream)
        self.coef_count = FORM_NCOEF[self.form]
        self._work = {}
        return self

    def workspace(self, nrhs: int):
        """This operator's own partial-block buffer, so that operators sharing a plan
        can run concurrently on different streams."""
        w = self._work.get(nrhs)
        if w is None:
            w = dv.workspace(self.plan.workspace_bytes(nrhs))
            self._work[nrhs] = w
        return w

    def batch_plan(self) -> "FusedPlan":
        """The fused plan of this operator with the fewest q3 slabs (one per 128 planes),
        for launches of several right-hand sides: their work-groups fill the device in
        many rounds, where each slab's fill and drain is overhead (p=4 n=64, 16 right-hand
        sides: FP64 1.58 against 1.65-1.70 ms, FP32 0.82 against 0.93-0.97 ms with the
        single-right-hand-side slab count).  It reads the same packed coefficients: their
        layout depends on the tiles, not on the slabs."""
        if not self.fused:
            raise ValueError("the composed operator has no fused plan")
        bp = getattr(self, "_batch", None)
        if bp is None:
            bp = self.plan
            if getattr(self, "_auto_slabs", True):
                bp = fused_plan(self.quadrule_list, self.form, self.dtype, 1)
                bp.set_rhs_block(self.plan.rhs_block)
            self._batch = bp
        return bp

    def batch_view(self) -> "StiffnessMassOperator":
        """This operator on :meth:`batch_plan` (same coefficients and weights; its own
        workspaces), for callers that fix the plan of all their launches and launch many
        right-hand sides at a time.  Launches of one or a few right-hand sides are slower
        on it (p=4 n=64 FP32, one right-hand side: 118 against 73 us), so solves whose
        group steps shrink (refinement phases desynchronise the runs of
        ``runs.run_batched``) keep the operator's own plan."""
        v = copy.copy(self)
        v.plan = self.batch_plan()
        v._batch = v.plan
        v._work = {}
        v._work_batch = {}
        return v

    def set_rhs_block(self, nr: int):
        """Right-hand sides per work-group of this operator's multi-right-hand-side
        launches (``FusedPlan.set_rhs_block`` on both of its plans)."""
        self.plan.set_rhs_block(nr)
        if self.fused:
            self.batch_plan().set_rhs_block(nr)

    def _batch_workspace(self, nrhs: int):
        cache = self.__dict__.setdefault("_work_batch", {})
        w = cache.get(nrhs)
        if w is None:
            w = dv.workspace(self.batch_plan().workspace_bytes(nrhs))
            cache[nrhs] = w
        return w

    def _composed_terms(self, Cd, md):
        t = dv.torch()
        qs = self.quadrule_list
        tdt = dv.torch_dtype(self.dtype)
        terms = []
        if not self.is_nurbs:
            if Cd is not None:
                Cc = [[Cd[i, j].to(tdt).contiguous() for j in range(3)] for i in range(3)]
                terms += terms_gradu_gradv(qs, Cc)
            if md is not None:
                terms += terms_u_v(qs, md.to(tdt).contiguous())
            return terms
        Wf, dWf = nurbs_weight_fields(qs, self.w64)
        wi = 1.0 / Wf
        z = [g * wi * wi for g in dWf]
        if Cd is not None:
            Cw = [[(Cd[i, j] * wi * wi).to(tdt).contiguous() for j in range(3)] for i in range(3)]
            terms += terms_gradu_gradv(qs, Cw)
            zcz = sum(Cd[i, j] * z[i] * z[j] for i in range(3) for j in range(3))
            terms += terms_u_v(qs, zcz.to(tdt).contiguous())
            ctz = [(-sum(Cd[i, j] * z[i] for i in range(3)) * wi).to(tdt).contiguous() for j in range(3)]
            terms += terms_u_gradv(qs, ctz)
            cz = [(-sum(Cd[i, j] * z[j] for j in range(3)) * wi).to(tdt).contiguous() for i in range(3)]
            terms += terms_gradu_v(qs, cz)
        if md is not None:
            terms += terms_u_v(qs, (md * wi * wi).to(tdt).contiguous())
        return terms

    def apply(self, x, out=None, nrhs: int = 1, alpha: float = 1.0, accumulate: bool = False, stream=None,
              work=None):
        """``out = alpha A x (+ out)`` on device arrays of the operator dtype; ``work``
        (``plan.workspace_bytes(nrhs)`` bytes) keeps the partial blocks of callers on
        different streams apart (the operator's own buffer otherwise).  Without ``work``,
        ``BATCH_PLAN_MIN_RHS`` or more right-hand sides run on :meth:`batch_plan`."""
        t = dv.torch()
        if out is None:
            out = t.empty_like(x)
        if self.fused:
            if work is None and nrhs >= BATCH_PLAN_MIN_RHS:
                # several right-hand sides: the plan with the fewest slabs, own workspace
                self.batch_plan().apply(self.packed, x, out, nrhs, self.w, alpha, accumulate,
                                        work=self._batch_workspace(nrhs), stream=stream)
                return out
            self.plan.apply(self.packed, x, out, nrhs, self.w, alpha, accumulate,
                            work=work if work is not None else self.workspace(nrhs), stream=stream)
            return out
        xin = x
        if self.is_nurbs:
            xin = (x.reshape(self.ndof, nrhs) * self.w[:, None]).reshape(-1)
        tmp = composed_apply(self.quadrule_list, self.terms, xin, nrhs, stream=stream)
        if self.is_nurbs:
            tmp = (tmp.reshape(self.ndof, nrhs) * self.w[:, None]).reshape(-1)
        if accumulate:
            out.add_(tmp, alpha=alpha)
        else:
            out.copy_(tmp).mul_(alpha) if alpha != 1.0 else out.copy_(tmp)
        return out

    __call__ = apply

    def apply_device(self, x, out=None, nrhs: int = 1, stream=None):
        return self.apply(x, out=out, nrhs=nrhs, stream=stream)

    def algorithmic_bytes(self, nrhs: int = 1) -> int:
        """``b (2 N_dof r + c N_q)`` with c = 7 (B-spline K+M) / 11 (NURBS K+M)."""
        b = 8 if self.dtype == "f64" else 4
        has_k = self.form != FORM_M1 if self.fused else True
        if self.is_nurbs:
            c = 11 if has_k else 1
        else:
            c = 7 if (has_k and self.form in (FORM_KM7, FORM_GEN16)) else (6 if has_k else 1)
        return b * (2 * self.ndof * nrhs + c * self.nq)


class FirstOrderOperator:
    """Device-resident first-order term in one fused sweep: ``kind`` "gradu_v",
    ``y_A = sum_i int d_i N_A c_i u`` (``compute_mf_scalar_gradu_v``,
    ``operators/bsplines_ops.py:357-390``), or "u_gradv", ``y_A = sum_j int N_A c_j
    d_j u`` (``bsplines_ops.py:392-429``); ``c`` is (3, N_q).  For a NURBS patch the
    rational correction of ``operators/nurbs_ops.py:392-430`` (``c W~^2`` on the
    first-order term, ``-(c . z) W~`` on the value term) is folded into the four
    packed fields per point, so the sweep reads ``4 N_q`` coefficients and no
    ``N_q``-sized intermediate reaches HBM."""

    def __init__(self, quadrule_list, c, kind: str, nurbs_weights=None, dtype: str = "f64"):
        dv.require_gpu()
        t = dv.torch()
        if kind not in ("gradu_v", "u_gradv"):
            raise ValueError("kind must be 'gradu_v' or 'u_gradv'")
        self.quadrule_list = list(quadrule_list)
        if not fused_applicable(self.quadrule_list):
            raise NotImplementedError("the fused first-order sweep takes three directions of WQ type 2 rules")
        self.dtype = dtype
        self.kind = kind
        self.ndof = int(np.prod([q.nbctrlpts for q in quadrule_list]))
        self.nq = int(np.prod([q.nbquadpts for q in quadrule_list]))
        self.is_nurbs = nurbs_weights is not None and int(
            nurbs_weights.numel() if isinstance(nurbs_weights, t.Tensor) else np.size(nurbs_weights)) > 0
        cd = dv.to_device(c, "f64").reshape(3, -1).contiguous()
        self.form = FORM_GV4 if kind == "gradu_v" else FORM_UG4
        self.plan = fused_plan(self.quadrule_list, self.form, dtype)
        W = dW = None
        self.w = None
        if self.is_nurbs:
            w64 = dv.to_device(nurbs_weights, "f64")
            self.w = w64.to(dv.torch_dtype(dtype))
            Wf, dWf = nurbs_weight_fields(self.quadrule_list, w64)
            W, dW = Wf, t.cat(dWf)
        self.packed = self.plan.pack(cd, None, W, dW)
        self.fused = True
        self._work = {}

    def workspace(self, nrhs: int):
        w = self._work.get(nrhs)
        if w is None:
            w = dv.workspace(self.plan.workspace_bytes(nrhs))
            self._work[nrhs] = w
        return w

    def apply(self, x, out=None, nrhs: int = 1, alpha: float = 1.0, accumulate: bool = False, stream=None):
        t = dv.torch()
        if out is None:
            out = t.empty_like(x)
        self.plan.apply(self.packed, x, out, nrhs, self.w, alpha, accumulate, work=self.workspace(nrhs), stream=stream)
        return out

    __call__ = apply

    def algorithmic_bytes(self, nrhs: int = 1) -> int:
        """``b (2 N_dof r + c N_q)``, c = 3 (B-spline) / 4 (NURBS: the value-value term)."""
        b = 8 if self.dtype == "f64" else 4
        return b * (2 * self.ndof * nrhs + (4 if self.is_nurbs else 3) * self.nq)


def apply_group(ops, X, outs, box=None, work=None, stream=None, member_major: bool = False):
    """The fused operators of several solves that share one plan (same rules, form
    and dtype: requests of one shape with their own coefficients, or one operator
    repeated for several right-hand sides) in one launch (``iga_fused_apply_group``):
    member k applies ``ops[k]`` to column k of ``X`` (``(N, g)`` C order; row k of a
    ``(g, N)`` array with ``member_major``) into the vector ``outs[k]``, bit for bit its
    own ``apply``.  ``box`` = ``(lo, len)`` (int64
    arrays) zeroes the outputs outside the Dirichlet box; ``work`` holds
    ``plan.workspace_bytes(g)`` bytes (the plan's own buffer if None)."""
    g = len(ops)
    if g < 1 or len(outs) != g:
        raise ValueError("one output per operator")
    plan = ops[0].plan if ops[0].fused else None
    if plan is None or any((not op.fused) or op.plan is not plan for op in ops):
        raise ValueError("a grouped apply takes fused operators of one plan")
    ptrs = (C.c_void_p * g)(*[dv.ptr(op.packed) for op in ops])
    yptrs = (C.c_void_p * g)(*[dv.ptr(y) for y in outs])
    work = work if work is not None else plan.workspace(g)
    nd, lop, lenp = 0, None, None
    if box is not None:
        lo = np.ascontiguousarray(box[0], dtype=np.int64)
        ln = np.ascontiguousarray(box[1], dtype=np.int64)
        nd, lop, lenp = 3, lo.ctypes.data_as(_lib._i64p), ln.ctypes.data_as(_lib._i64p)
    n = ops[0].ndof
    xd, xm = (1, n) if member_major else (g, 1)
    _lib.call("iga_fused_apply_group", plan._h, g, ptrs, dv.ptr(X), xd, xm, yptrs, dv.ptr(ops[0].w), dv.ptr(work),
              nd, lop, lenp, dv.stream_ptr(stream))
    return outs


# ---------------------------------------------------------------------------
# Reference-signature operator classes
# ---------------------------------------------------------------------------


def _operator_call(kind, quadrule_list, coefficients, array_in, time_ders, nurbs_weights):
    dv.require_gpu()
    t = dv.torch()
    qs = list(quadrule_list)
    nd = len(qs)
    is_nurbs = nurbs_weights is not None and np.size(nurbs_weights) > 0
    dtype = "f32" if isinstance(array_in, t.Tensor) and array_in.dtype == t.float32 else "f64"
    x, r, was_np, shape = _to_dev_2d(array_in, dtype)
    ndof = int(np.prod([q.nbctrlpts for q in qs]))
    if x.numel() != ndof * r:
        raise ValueError("array_in does not match the number of control points")
    coef = dv.to_device(coefficients, "f64")
    if (
        kind in ("u_v", "gradu_gradv")
        and fused_applicable(qs, time_ders)
        and (kind == "u_v" or coefficients.shape[:2] == (3, 3))
    ):
        if kind == "u_v":
            op = StiffnessMassOperator(qs, C=None, m=coef.reshape(-1), nurbs_weights=nurbs_weights, dtype=dtype)
        else:
            op = StiffnessMassOperator(qs, C=coef.reshape(3, 3, -1), m=None, nurbs_weights=nurbs_weights, dtype=dtype)
        y = op.apply(x, nrhs=r)
        return _from_dev(y, was_np, shape)
    if kind in ("gradu_v", "u_gradv") and fused_applicable(qs, time_ders) and tuple(coefficients.shape[:1]) == (3,):
        op = FirstOrderOperator(qs, coef.reshape(3, -1), kind, nurbs_weights=nurbs_weights, dtype=dtype)
        y = op.apply(x, nrhs=r)
        return _from_dev(y, was_np, shape)
    tdt = dv.torch_dtype(dtype)
    xin = x
    w = None
    if is_nurbs:
        w64 = dv.to_device(nurbs_weights, "f64")
        w = w64.to(tdt)
        xin = (x.reshape(ndof, r) * w[:, None]).reshape(-1)
        Wf, dWf = nurbs_weight_fields(qs, w64)
        wi = 1.0 / Wf
        sd = nd - 1 if time_ders is not None else nd
        z = [g * wi * wi for g in dWf][:sd]
    c = coef
    terms: List[Term] = []
    cast = lambda a: a.to(tdt).contiguous()
    if kind == "u_v":
        cc = c.reshape(-1)
        terms = terms_u_v(qs, cast(cc * wi * wi if is_nurbs else cc), time_ders)
    elif kind == "gradu_gradv":
        sd = c.shape[0]
        if is_nurbs:
            terms += terms_gradu_gradv(qs, [[cast(c[i, j] * wi * wi) for j in range(sd)] for i in range(sd)], time_ders)
            zcz = sum(c[i, j] * z[i] * z[j] for i in range(sd) for j in range(sd))
            terms += terms_u_v(qs, cast(zcz), time_ders)
            terms += terms_u_gradv(qs, [cast(-sum(c[i, j] * z[i] for i in range(sd)) * wi) for j in range(sd)], time_ders)
            terms += terms_gradu_v(qs, [cast(-sum(c[i, j] * z[j] for j in range(sd)) * wi) for i in range(sd)], time_ders)
        else:
            terms = terms_gradu_gradv(qs, [[cast(c[i, j]) for j in range(sd)] for i in range(sd)], time_ders)
    elif kind in ("gradu_v", "u_gradv"):
        sd = c.shape[0]
        build = terms_gradu_v if kind == "gradu_v" else terms_u_gradv
        if is_nurbs:
            terms += build(qs, [cast(c[i] * wi * wi) for i in range(sd)], time_ders)
            terms += terms_u_v(qs, cast(-sum(c[i] * z[i] for i in range(sd)) * wi), time_ders)
        else:
            terms = build(qs, [cast(c[i]) for i in range(sd)], time_ders)
    else:
        raise ValueError(kind)
    y = composed_apply(qs, terms, xin, r)
    if is_nurbs:
        y = (y.reshape(ndof, r) * w[:, None]).reshape(-1)
    return _from_dev(y, was_np, shape)


def _asm(kind: str, name: str):
    def call(quadrule_list, coefficients, *args, **kwargs):
        from . import assembly

        return getattr(getattr(assembly, kind), name)(quadrule_list, coefficients, *args, **kwargs)

    call.__name__ = name
    call.__doc__ = f"``{name}``: sparse assembly / load vector on the device (``igaops_amd.assembly``)."
    return staticmethod(call)


class BsplineOperations:
    """``operators/bsplines_ops.py:56-429`` on the device: matrix-free applies and the
    sum-factorised assembly (``assemble_scalar_*``)."""

    assemble_scalar_u_v = _asm("BsplineAssembly", "assemble_scalar_u_v")
    assemble_scalar_gradu_gradv = _asm("BsplineAssembly", "assemble_scalar_gradu_gradv")
    assemble_scalar_gradu_v = _asm("BsplineAssembly", "assemble_scalar_gradu_v")
    assemble_scalar_u_gradv = _asm("BsplineAssembly", "assemble_scalar_u_gradv")
    assemble_scalar_u_force = _asm("BsplineAssembly", "assemble_scalar_u_force")
    assemble_scalar_gradu_force = _asm("BsplineAssembly", "assemble_scalar_gradu_force")

    @staticmethod
    def compute_mf_scalar_u_v(quadrule_list, coefficients, array_in, time_ders=None, nurbs_weights=np.array([])):
        return _operator_call("u_v", quadrule_list, coefficients, array_in, time_ders, None)

    @staticmethod
    def compute_mf_scalar_gradu_gradv(quadrule_list, coefficients, array_in, time_ders=None, nurbs_weights=np.array([])):
        return _operator_call("gradu_gradv", quadrule_list, coefficients, array_in, time_ders, None)

    @staticmethod
    def compute_mf_scalar_gradu_v(quadrule_list, coefficients, array_in, time_ders=None, nurbs_weights=np.array([])):
        return _operator_call("gradu_v", quadrule_list, coefficients, array_in, time_ders, None)

    @staticmethod
    def compute_mf_scalar_u_gradv(quadrule_list, coefficients, array_in, time_ders=None, nurbs_weights=np.array([])):
        return _operator_call("u_gradv", quadrule_list, coefficients, array_in, time_ders, None)


class NurbsOperations:
    """``operators/nurbs_ops.py:80-463`` on the device: matrix-free applies and the
    sum-factorised assembly (``assemble_scalar_*``)."""

    assemble_scalar_u_v = _asm("NurbsAssembly", "assemble_scalar_u_v")
    assemble_scalar_gradu_gradv = _asm("NurbsAssembly", "assemble_scalar_gradu_gradv")
    assemble_scalar_gradu_v = _asm("NurbsAssembly", "assemble_scalar_gradu_v")
    assemble_scalar_u_gradv = _asm("NurbsAssembly", "assemble_scalar_u_gradv")
    assemble_scalar_u_force = _asm("NurbsAssembly", "assemble_scalar_u_force")
    assemble_scalar_gradu_force = _asm("NurbsAssembly", "assemble_scalar_gradu_force")

    @staticmethod
    def compute_mf_scalar_u_v(quadrule_list, coefficients, array_in, time_ders=None, nurbs_weights=np.array([])):
        return _operator_call("u_v", quadrule_list, coefficients, array_in, time_ders, nurbs_weights)

    @staticmethod
    def compute_mf_scalar_gradu_gradv(quadrule_list, coefficients, array_in, time_ders=None, nurbs_weights=np.array([])):
        return _operator_call("gradu_gradv", quadrule_list, coefficients, array_in, time_ders, nurbs_weights)

    @staticmethod
    def compute_mf_scalar_gradu_v(quadrule_list, coefficients, array_in, time_ders=None, nurbs_weights=np.array([])):
        return _operator_call("gradu_v", quadrule_list, coefficients, array_in, time_ders, nurbs_weights)

    @staticmethod
    def compute_mf_scalar_u_gradv(quadrule_list, coefficients, array_in, time_ders=None, nurbs_weights=np.array([])):
        return _operator_call("u_gradv", quadrule_list, coefficients, array_in, time_ders, nurbs_weights)


def operator_engine(nurbs_weights):
    """``geometry/mixin/template.py:74-77``: NURBS engine iff weights are given."""
    return NurbsOperations if nurbs_weights is not None and np.size(nurbs_weights) > 0 else BsplineOperations
