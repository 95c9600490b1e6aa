"""Oracle: matrix-free B-spline and NURBS operators.

Follows ``operators/bsplines_ops.py:290-429`` (``compute_mf_scalar_{u_v, gradu_gradv,
gradu_v, u_gradv}``) and ``operators/nurbs_ops.py:14-77,320-463`` (rational versions
built from the B-spline ones on ``w o x`` with the weight function ``W`` projected at
the quadrature points).  Index conventions kept from the reference:

* ``basis[beta]`` is ``(n_q x n_c)``; ``weights[zeta]`` is ``(n_c x n_q)`` with
  ``zeta = beta + 2 alpha`` (trial derivative beta, test derivative alpha).
* ``C[i, j]`` multiplies the trial derivative ``d_j`` and the test derivative ``d_i``.
* With ``time_ders=(a, b)`` the last direction is time: trial basis ``B_b``; the mass
  operator tests with ``weights[a]`` (``bsplines_ops.py:309-316``), the others with
  ``zeta = b + 2a`` and loop over the spatial directions only.

Pinned by (tests/test_oracle_operators.py): explicit sparse assembly of the same
integrals (``oracle.assembly``), constant-coefficient Kronecker sums on the identity
geometry (Gauss-exact), ``M 1 = int N_A``, ``K 1 = 0``, symmetry of K and M, NURBS
with unit weights == B-spline (all four operators), the quarter-annulus volume
``1^T M 1 = 3 pi / 4``, all four NURBS operators against direct assembly from the
rational basis ``R_A = w_A N_A / W`` and its gradient on the Gauss annulus (with planted
chain-rule errors that must fail), and the space-time operator
``u_v(time_ders=(0,1)) + gradu_gradv(time_ders=(0,0))`` against the closed form
``kron(W_t, M_s^3) + kron(M_t, K_s)`` from 1-D matrices built with
``scipy.interpolate.BSpline`` (plus the WQ type "1" time-weight convention of
``bsplines_ops.py:309-316``, where a wrong weight index fails).
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from . import matrix_free as mf


def _basis(quadrule_list, ders):
    return [q.basis[b] for q, b in zip(quadrule_list, ders)]


def _weights(quadrule_list, ders):
    return [q.weights[z] for q, z in zip(quadrule_list, ders)]


def _bcast(c: np.ndarray, arr: np.ndarray) -> np.ndarray:
    """``c_q * arr_q...`` with trailing batch dims on ``arr``."""
    return c.reshape(c.shape + (1,) * (arr.ndim - 1)) * arr


def bspline_u_v(quadrule_list, coefficients, array_in, time_ders: Optional[Sequence[int]] = None):
    """``bsplines_ops.py:290-317``."""
    nd = len(quadrule_list)
    ders = np.zeros(nd, dtype=int)
    if time_ders is not None:
        ders[-1] = time_ders[1]
    tmp = mf.apply(_basis(quadrule_list, ders), array_in)
    if time_ders is not None:
        ders[-1] = time_ders[0]
    return mf.apply(_weights(quadrule_list, ders), _bcast(np.asarray(coefficients), tmp))


def bspline_gradu_gradv(quadrule_list, coefficients, array_in, time_ders=None):
    """``bsplines_ops.py:319-355``."""
    nd = len(quadrule_list)
    sd = nd - 1 if time_ders is not None else nd
    out = np.zeros_like(array_in, dtype=float)
    for j in range(sd):
        beta = np.zeros(nd, dtype=int)
        beta[j] = 1
        if time_ders is not None:
            beta[-1] = time_ders[1]
        tmp = mf.apply(_basis(quadrule_list, beta), array_in)
        for i in range(sd):
            alpha = np.zeros(nd, dtype=int)
            alpha[i] = 1
            if time_ders is not None:
                alpha[-1] = time_ders[0]
            zeta = beta + 2 * alpha
            out += mf.apply(_weights(quadrule_list, zeta), _bcast(coefficients[i, j], tmp))
    return out


def bspline_gradu_v(quadrule_list, coefficients, array_in, time_ders=None):
    """``bsplines_ops.py:357-390``: ``sum_i int d_i N_A c_i u``."""
    nd = len(quadrule_list)
    sd = nd - 1 if time_ders is not None else nd
    out = np.zeros_like(array_in, dtype=float)
    beta = np.zeros(nd, dtype=int)
    if time_ders is not None:
        beta[-1] = time_ders[1]
    tmp = mf.apply(_basis(quadrule_list, beta), array_in)
    for i in range(sd):
        alpha = np.zeros(nd, dtype=int)
        alpha[i] = 1
        if time_ders is not None:
            alpha[-1] = time_ders[0]
        out += mf.apply(_weights(quadrule_list, beta + 2 * alpha), _bcast(coefficients[i], tmp))
    return out


def bspline_u_gradv(quadrule_list, coefficients, array_in, time_ders=None):
    """``bsplines_ops.py:392-429``: ``sum_j int N_A c_j d_j u``."""
    nd = len(quadrule_list)
    sd = nd - 1 if time_ders is not None else nd
    nq = int(np.prod([q.nbquadpts for q in quadrule_list]))
    acc = np.zeros((nq,) + array_in.shape[1:])
    for i in range(sd):
        beta = np.zeros(nd, dtype=int)
        beta[i] = 1
        if time_ders is not None:
            beta[-1] = time_ders[1]
        acc += _bcast(coefficients[i], mf.apply(_basis(quadrule_list, beta), array_in))
    alpha = np.zeros(nd, dtype=int)
    if time_ders is not None:
        alpha[-1] = time_ders[0]
    return mf.apply(_weights(quadrule_list, alpha), acc)


# --- NURBS (operators/nurbs_ops.py) ------------------------------------------


def project_nurbs_weights(quadrule_list, nurbs_weights, nders: int = 1):
    """``W~ = 1/W`` and ``z = grad W * W~^2`` at the quadrature points
    (``nurbs_ops.py:14-62``)."""
    nd = len(quadrule_list)
    W = mf.apply(_basis(quadrule_list, np.zeros(nd, dtype=int)), nurbs_weights)
    winv = 1.0 / W
    if nders == 0:
        return winv, None
    grad = []
    for j in range(nd):
        b = np.zeros(nd, dtype=int)
        b[j] = 1
        grad.append(mf.apply(_basis(quadrule_list, b), nurbs_weights))
    return winv, np.stack(grad) * winv ** 2


def _scale_rows(w, arr):
    return _bcast(np.asarray(w), arr)


def nurbs_u_v(quadrule_list, coefficients, array_in, time_ders=None, nurbs_weights=None):
    """``nurbs_ops.py:320-341``."""
    winv, _ = project_nurbs_weights(quadrule_list, nurbs_weights, nders=0)
    xs = _scale_rows(nurbs_weights, array_in)
    out = bspline_u_v(quadrule_list, coefficients * winv ** 2, xs, time_ders)
    return _scale_rows(nurbs_weights, out)


def nurbs_gradu_gradv(quadrule_list, coefficients, array_in, time_ders=None, nurbs_weights=None):
    """``nurbs_ops.py:343-390``: the chain rule of ``R_A = w_A N_A / W`` gives four
    B-spline operators on ``x~ = w o x``: ``C W~^2`` (stiffness), ``z^T C z`` (mass),
    ``-(C^T z W~)`` (trial-gradient / test-value) and ``-(C z W~)`` (trial-value /
    test-gradient)."""
    nd = len(quadrule_list)
    sd = nd - 1 if time_ders is not None else nd
    winv, z = project_nurbs_weights(quadrule_list, nurbs_weights, nders=1)
    z = z[:sd]
    xs = _scale_rows(nurbs_weights, array_in)
    C = np.asarray(coefficients)
    out = bspline_gradu_gradv(quadrule_list, C * winv[None, None] ** 2, xs, time_ders)
    out += bspline_u_v(quadrule_list, np.einsum("ij...,i...,j...->...", C, z, z), xs, time_ders)
    out -= bspline_u_gradv(quadrule_list, np.einsum("ij...,i...,...->j...", C, z, winv), xs, time_ders)
    out -= bspline_gradu_v(quadrule_list, np.einsum("ij...,j...,...->i...", C, z, winv), xs, time_ders)
    return _scale_rows(nurbs_weights, out)


def _nurbs_first_order(quadrule_list, coefficients, array_in, time_ders, nurbs_weights, kind):
    """``nurbs_ops.py:392-429`` (shared by gradu_v and u_gradv)."""
    nd = len(quadrule_list)
    sd = nd - 1 if time_ders is not None else nd
    winv, z = project_nurbs_weights(quadrule_list, nurbs_weights, nders=1)
    z = z[:sd]
    xs = _scale_rows(nurbs_weights, array_in)
    c = np.asarray(coefficients)
    op = bspline_gradu_v if kind == "gradu_v" else bspline_u_gradv
    out = op(quadrule_list, c * winv[None] ** 2, xs, time_ders)
    out -= bspline_u_v(quadrule_list, np.einsum("i...,i...,...->...", c, z, winv), xs, time_ders)
    return _scale_rows(nurbs_weights, out)


def nurbs_gradu_v(quadrule_list, coefficients, array_in, time_ders=None, nurbs_weights=None):
    return _nurbs_first_order(quadrule_list, coefficients, array_in, time_ders, nurbs_weights, "gradu_v")


def nurbs_u_gradv(quadrule_list, coefficients, array_in, time_ders=None, nurbs_weights=None):
    return _nurbs_first_order(quadrule_list, coefficients, array_in, time_ders, nurbs_weights, "u_gradv")


def stiffness_plus_mass(quadrule_list, C, m, x, nurbs_weights=None):
    """``y = K x + M x`` as two reference operator calls (the benchmark matvec)."""
    if nurbs_weights is not None and np.size(nurbs_weights) > 0:
        return nurbs_gradu_gradv(quadrule_list, C, x, nurbs_weights=nurbs_weights) + nurbs_u_v(
            quadrule_list, m, x, nurbs_weights=nurbs_weights
        )
    return bspline_gradu_gradv(quadrule_list, C, x) + bspline_u_v(quadrule_list, m, x)
