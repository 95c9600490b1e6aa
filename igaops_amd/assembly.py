"""Sparse assembly by sum factorisation (``operators/utils/operations.py:59-115``,
``operators/bsplines_ops.py:138-288``, ``operators/nurbs_ops.py:182-319``).

The reference assembles one column of the matrix at a time by contracting the
coefficient tensor with ``W_d diag(B_d[:, j_d])`` direction by direction.  Here
the values of every nonzero are computed on the device (``iga_sumfact_assemble``,
one wave per row): the 1-D products ``W_{t,d}[i, q] B_{t,d}[q, j]`` of every term
``t`` and direction ``d`` are tabulated on the host (1-D, small), the 3-D (or 4-D)
pattern is their Kronecker product, and

    A[i, j] = sum_t s_t sum_q prod_d W_{t,d}[i_d, q_d] B_{t,d}[q_d, j_d] c_t[q].

The result is a ``scipy.sparse.csr_array`` with sorted indices and the exact
zeros removed, like the reference's.  The load vectors (``assemble_scalar_*_force``)
are weighted-quadrature applies (``MatrixFree.apply`` on the device).
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np
from scipy import sparse as sp

from . import _lib
from . import device as dv


class AsmTerm:
    """One term ``s * W^{weights} diag(coef) B^{basis}`` (per-direction indices)."""

    __slots__ = ("weights", "basis", "coef", "scale")

    def __init__(self, weights, basis, coef, scale: float = 1.0):
        self.weights = [int(v) for v in weights]
        self.basis = [int(v) for v in basis]
        self.coef = coef
        self.scale = float(scale)


def _one_dim_products(W, B):
    """All nonzero ``W[i, q] B[q, j]`` as arrays (i, j, q, value), sorted by (i, j, q)."""
    W = sp.csr_array(W)
    B = sp.csr_array(B)
    Wc = W.tocoo()
    wi, wq, wv = Wc.row.astype(np.int64), Wc.col.astype(np.int64), Wc.data
    start = B.indptr[wq]
    cnt = (B.indptr[wq + 1] - start).astype(np.int64)
    rep = np.repeat(np.arange(wi.size), cnt)
    pos = np.repeat(start - np.concatenate(([0], np.cumsum(cnt)[:-1])), cnt) + np.arange(int(cnt.sum()))
    I, Q = wi[rep], wq[rep]
    J = B.indices[pos].astype(np.int64)
    V = wv[rep] * B.data[pos]
    keep = V != 0.0
    I, J, Q, V = I[keep], J[keep], Q[keep], V[keep]
    order = np.lexsort((Q, J, I))
    return I[order], J[order], Q[order], V[order]


def _tables(quadrule_list, terms: Sequence[AsmTerm]):
    """Per direction: the union pattern of (i, j) pairs over the terms, and per
    (term, direction) the offsets of each pair's (q, value) list."""
    ndim = len(quadrule_list)
    nr = [None] * ndim
    nc = [None] * ndim
    nq = [int(q.nbquadpts) for q in quadrule_list]
    pat_ptr, pat_col, lst_ptr, lst_q, lst_v = [], [], [[None] * ndim for _ in terms], [], []
    for d, q in enumerate(quadrule_list):
        prods = []
        for term in terms:
            W, B = q.weights[term.weights[d]], q.basis[term.basis[d]]
            nr[d], nc[d] = int(W.shape[0]), int(B.shape[1])
            prods.append(_one_dim_products(W, B))
        keys = np.unique(np.concatenate([p[0] * nc[d] + p[1] for p in prods]))
        rows = keys // nc[d]
        pat_ptr.append(np.concatenate(([0], np.cumsum(np.bincount(rows, minlength=nr[d])))).astype(np.int32))
        pat_col.append((keys % nc[d]).astype(np.int32))
        qs, vs, base = [], [], 0
        for t, (I, J, Q, V) in enumerate(prods):
            pair = np.searchsorted(keys, I * nc[d] + J)
            counts = np.bincount(pair, minlength=keys.size)
            lst_ptr[t][d] = (base + np.concatenate(([0], np.cumsum(counts)))).astype(np.int32)
            qs.append(Q.astype(np.int32))
            vs.append(V.astype(np.float64))
            base += Q.size
        lst_q.append(np.concatenate(qs) if qs else np.zeros(0, np.int32))
        lst_v.append(np.concatenate(vs) if vs else np.zeros(0))
    return nr, nq, nc, pat_ptr, pat_col, lst_ptr, lst_q, lst_v


def sumfact_assemble(quadrule_list, terms: Sequence[AsmTerm], w=None) -> sp.csr_array:
    """``sum_t`` of ``Operations.sumfact_assembling`` (``operations.py:59-115``) over
    ``terms``, optionally scaled to ``diag(w) A diag(w)``; float64."""
    dv.require_gpu()
    t = dv.torch()
    ndim = len(quadrule_list)
    if not terms:
        raise ValueError("no terms to assemble")
    nr, nq, nc, pat_ptr, pat_col, lst_ptr, lst_q, lst_v = _tables(quadrule_list, terms)
    n_rows, n_cols, n_q = int(np.prod(nr)), int(np.prod(nc)), int(np.prod(nq))
    i64 = lambda v: (C.c_int64 * len(v))(*[int(x) for x in v])
    ptrs_i32 = lambda arrs: (C.POINTER(C.c_int32) * len(arrs))(
        *[a.ctypes.data_as(C.POINTER(C.c_int32)) for a in arrs])
    ptrs_f64 = lambda arrs: (C.POINTER(C.c_double) * len(arrs))(
        *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])
    nr_a, nq_a, nc_a = i64(nr), i64(nq), i64(nc)
    pp = ptrs_i32(pat_ptr)
    lib = _lib.load()
    nnz = int(lib.iga_sumfact_nnz(ndim, nr_a, pp))
    if nnz < 0:
        raise ValueError("bad assembly pattern")
    coefs = []
    for term in terms:
        c = dv.to_device(term.coef, "f64").reshape(-1)
        if c.numel() != n_q:
            raise ValueError(f"coefficient field has {c.numel()} values, the rule has {n_q} points")
        coefs.append(c.contiguous())
    flat_lst_ptr = [lst_ptr[ti][d] for ti in range(len(terms)) for d in range(ndim)]
    wd = None
    if w is not None:
        wd = dv.to_device(np.asarray(w, dtype=np.float64) if not isinstance(w, t.Tensor) else w, "f64").reshape(-1)
        if wd.numel() != n_rows:
            raise ValueError("scaling vector does not match the matrix")
    row_ptr = t.empty(n_rows + 1, dtype=t.int64, device=dv.device())
    cols = t.empty(max(nnz, 1), dtype=t.int64, device=dv.device())
    vals = t.empty(max(nnz, 1), dtype=t.float64, device=dv.device())
    coef_ptrs = (C.c_void_p * len(coefs))(*[dv.ptr(c) for c in coefs])
    scales = (C.c_double * len(terms))(*[term.scale for term in terms])
    _lib.call("iga_sumfact_assemble", ndim, len(terms), nr_a, nq_a, nc_a, pp, ptrs_i32(pat_col),
              i64([a.size for a in pat_col]), ptrs_i32(flat_lst_ptr), ptrs_i32(lst_q), ptrs_f64(lst_v),
              i64([a.size for a in lst_q]), coef_ptrs, scales, dv.ptr(row_ptr), dv.ptr(wd), dv.ptr(cols),
              dv.ptr(vals), dv.stream_ptr())
    mat = sp.csr_array((vals[:nnz].cpu().numpy(), cols[:nnz].cpu().numpy(), row_ptr.cpu().numpy()),
                       shape=(n_rows, n_cols))
    mat.eliminate_zeros()
    return mat


# ---------------------------------------------------------------------------
# Term lists of the reference's assemble_* functions
# ---------------------------------------------------------------------------


def asm_terms_u_v(quadrule_list, c, time_ders=None) -> List[AsmTerm]:
    """``bsplines_ops.py:139-158``."""
    nd = len(quadrule_list)
    b, w = [0] * nd, [0] * nd
    if time_ders is not None:
        b[-1] = time_ders[1]
        w[-1] = time_ders[0]
    return [AsmTerm(w, b, c)]


def asm_terms_gradu_gradv(quadrule_list, C, time_ders=None) -> List[AsmTerm]:
    """``bsplines_ops.py:160-191``: every direction (the time direction too, with its
    derivative orders overwritten by ``time_ders``), weights ``beta + 2 alpha``."""
    nd = len(quadrule_list)
    out = []
    for j in range(nd):
        beta = [0] * nd
        beta[j] = 1
        if time_ders is not None:
            beta[-1] = time_ders[1]
        for i in range(nd):
            alpha = [0] * nd
            alpha[i] = 1
            if time_ders is not None:
                alpha[-1] = time_ders[0]
            out.append(AsmTerm([b + 2 * a for a, b in zip(alpha, beta)], beta, C[i][j]))
    return out


def asm_terms_gradu_v(quadrule_list, c, time_ders=None) -> List[AsmTerm]:
    """``bsplines_ops.py:193-217``."""
    nd = len(quadrule_list)
    beta = [0] * nd
    if time_ders is not None:
        beta[-1] = time_ders[1]
    out = []
    for i in range(nd):
        alpha = [0] * nd
        alpha[i] = 1
        if time_ders is not None:
            alpha[-1] = time_ders[0]
        out.append(AsmTerm([b + 2 * a for a, b in zip(alpha, beta)], beta, c[i]))
    return out


def asm_terms_u_gradv(quadrule_list, c, time_ders=None) -> List[AsmTerm]:
    """``bsplines_ops.py:219-246``."""
    nd = len(quadrule_list)
    out = []
    for j in range(nd):
        beta = [0] * nd
        beta[j] = 1
        if time_ders is not None:
            beta[-1] = time_ders[1]
        alpha = [0] * nd
        if time_ders is not None:
            alpha[-1] = time_ders[0]
        out.append(AsmTerm([b + 2 * a for a, b in zip(alpha, beta)], beta, c[j]))
    return out


def _negate(terms: List[AsmTerm]) -> List[AsmTerm]:
    for term in terms:
        term.scale = -term.scale
    return terms


def _nurbs_fields(quadrule_list, nurbs_weights):
    """``W~ = 1/W`` and ``z = grad W W~^2`` at the points (``nurbs_ops.py:14-62``)."""
    from .operators import nurbs_weight_fields

    w = dv.to_device(np.asarray(nurbs_weights, dtype=np.float64), "f64")
    W, dW = nurbs_weight_fields(quadrule_list, w)
    wt = 1.0 / W
    z = [g * wt * wt for g in dW]
    return w, wt, z


def _dev(c):
    return dv.to_device(c, "f64")


def _fields(coefficients, ndim: int):
    t = dv.torch()
    was_np = not isinstance(coefficients, t.Tensor)
    c = dv.to_device(coefficients, "f64")
    if c.dim() != ndim:
        raise ValueError(f"Expected {ndim}D array.")
    return c, was_np


def _weights_apply(quadrule_list, zeta, c):
    """``MatrixFree.apply([q.weights[zeta_d]], c.T).T`` for ``c`` of shape (nm, N_q)."""
    from .operators import MatrixFree

    return MatrixFree.apply([q.weights[z] for q, z in zip(quadrule_list, zeta)], c.T.contiguous()).T.contiguous()


def _gradu_force(quadrule_list, c):
    nd = len(quadrule_list)
    out = None
    for i in range(nd):
        part = _weights_apply(quadrule_list, [2 if d == i else 0 for d in range(nd)], c[:, i, :])
        out = part if out is None else out + part
    return out


class BsplineAssembly:
    """``BsplineOperations.assemble_scalar_*`` (``bsplines_ops.py:138-288``)."""

    @staticmethod
    def assemble_scalar_u_v(quadrule_list, coefficients, time_ders=None, nurbs_weights=np.array([])):
        return sumfact_assemble(quadrule_list, asm_terms_u_v(quadrule_list, coefficients, time_ders))

    @staticmethod
    def assemble_scalar_gradu_gradv(quadrule_list, coefficients, time_ders=None, nurbs_weights=np.array([])):
        return sumfact_assemble(quadrule_list, asm_terms_gradu_gradv(quadrule_list, coefficients, time_ders))

    @staticmethod
    def assemble_scalar_gradu_v(quadrule_list, coefficients, time_ders=None, nurbs_weights=np.array([])):
        return sumfact_assemble(quadrule_list, asm_terms_gradu_v(quadrule_list, coefficients, time_ders))

    @staticmethod
    def assemble_scalar_u_gradv(quadrule_list, coefficients, time_ders=None, nurbs_weights=np.array([])):
        return sumfact_assemble(quadrule_list, asm_terms_u_gradv(quadrule_list, coefficients, time_ders))

    @staticmethod
    def assemble_scalar_u_force(quadrule_list, coefficients, nurbs_weights=np.array([])):
        """``bsplines_ops.py:248-262``: ``(nm, N_q)`` fields -> ``(nm, N_dof)`` loads
        (numpy in, numpy out; device tensors stay on the device)."""
        c, was_np = _fields(coefficients, 2)
        y = _weights_apply(quadrule_list, [0] * len(quadrule_list), c)
        return y.cpu().numpy() if was_np else y

    @staticmethod
    def assemble_scalar_gradu_force(quadrule_list, coefficients, nurbs_weights=np.array([])):
        """``bsplines_ops.py:264-287``: ``(nm, d, N_q)`` fields, weights ``2 e_i``."""
        c, was_np = _fields(coefficients, 3)
        y = _gradu_force(quadrule_list, c)
        return y.cpu().numpy() if was_np else y


class NurbsAssembly:
    """``NurbsOperations.assemble_scalar_*`` (``nurbs_ops.py:182-319``): the B-spline
    assembly of the rational terms, then ``diag(w) A diag(w)`` in the same kernel."""

    @staticmethod
    def assemble_scalar_u_v(quadrule_list, coefficients, time_ders=None, nurbs_weights=np.array([])):
        w, wt, _ = _nurbs_fields(quadrule_list, nurbs_weights)
        c = _dev(coefficients).reshape(-1) * wt * wt
        return sumfact_assemble(quadrule_list, asm_terms_u_v(quadrule_list, c, time_ders), w=w)

    @staticmethod
    def assemble_scalar_gradu_gradv(quadrule_list, coefficients, time_ders=None, nurbs_weights=np.array([])):
        w, wt, z = _nurbs_fields(quadrule_list, nurbs_weights)
        Cd = _dev(coefficients)
        nd = len(z)
        Cq = Cd.reshape(nd, nd, -1)
        terms = asm_terms_gradu_gradv(quadrule_list, [[Cq[i, j] * wt * wt for j in range(nd)] for i in range(nd)],
                                      time_ders)
        zcz = sum(Cq[i, j] * z[i] * z[j] for i in range(nd) for j in range(nd))
        terms += asm_terms_u_v(quadrule_list, zcz, time_ders)
        czj = [sum(Cq[i, j] * z[i] for i in range(nd)) * wt for j in range(nd)]  # sum_i C_ij z_i W~
        terms += _negate(asm_terms_u_gradv(quadrule_list, czj, time_ders))
        czi = [sum(Cq[i, j] * z[j] for j in range(nd)) * wt for i in range(nd)]  # sum_j C_ij z_j W~
        terms += _negate(asm_terms_gradu_v(quadrule_list, czi, time_ders))
        return sumfact_assemble(quadrule_list, terms, w=w)

    @staticmethod
    def _first_order(quadrule_list, coefficients, time_ders, nurbs_weights, kind):
        w, wt, z = _nurbs_fields(quadrule_list, nurbs_weights)
        nd = len(z)
        cq = _dev(coefficients).reshape(nd, -1)
        build = asm_terms_gradu_v if kind == "gradu_v" else asm_terms_u_gradv
        terms = build(quadrule_list, [cq[i] * wt * wt for i in range(nd)], time_ders)
        cz = sum(cq[i] * z[i] for i in range(nd)) * wt
        terms += _negate(asm_terms_u_v(quadrule_list, cz, time_ders))
        return sumfact_assemble(quadrule_list, terms, w=w)

    @staticmethod
    def assemble_scalar_gradu_v(quadrule_list, coefficients, time_ders=None, nurbs_weights=np.array([])):
        return NurbsAssembly._first_order(quadrule_list, coefficients, time_ders, nurbs_weights, "gradu_v")

    @staticmethod
    def assemble_scalar_u_gradv(quadrule_list, coefficients, time_ders=None, nurbs_weights=np.array([])):
        return NurbsAssembly._first_order(quadrule_list, coefficients, time_ders, nurbs_weights, "u_gradv")

    @staticmethod
    def assemble_scalar_u_force(quadrule_list, coefficients, nurbs_weights=np.array([])):
        """``nurbs_ops.py:289-300``: ``w * u_force(c W~)``."""
        w, wt, _ = _nurbs_fields(quadrule_list, nurbs_weights)
        c, was_np = _fields(coefficients, 2)
        y = _weights_apply(quadrule_list, [0] * len(quadrule_list), c * wt[None, :]) * w[None, :]
        return y.cpu().numpy() if was_np else y

    @staticmethod
    def assemble_scalar_gradu_force(quadrule_list, coefficients, nurbs_weights=np.array([])):
        """``nurbs_ops.py:302-319``: ``w * (gradu_force(C W~) - u_force(sum_j C_j z_j))``."""
        w, wt, z = _nurbs_fields(quadrule_list, nurbs_weights)
        c, was_np = _fields(coefficients, 3)  # (nm, d, N_q)
        y = _gradu_force(quadrule_list, c * wt[None, None, :])
        cz = sum(c[:, j, :] * z[j][None, :] for j in range(len(z)))
        y = (y - _weights_apply(quadrule_list, [0] * len(quadrule_list), cz)) * w[None, :]
        return y.cpu().numpy() if was_np else y
