"""Oracle: explicitly assembled operator matrices, an independent pin of the
matrix-free operators.

Instead of L-mode products these build the Kronecker matrices
``B^beta = B_d^{beta_d} (x) ... (x) B_1^{beta_1}`` and ``W^zeta`` explicitly with
``scipy.sparse.kron`` and form ``sum W^zeta diag(c) B^beta`` (the sum-factorised
assembly of ``operators/utils/operations.py:59-115`` computes the same matrices).
For Gauss rules the NURBS operators are also assembled from the rational basis
``R_A = w_A N_A / W`` and its gradient ``dR_A = w_A (dN_A W - N_A dW) / W^2`` evaluated
directly, which checks the chain-rule algebra of ``operators/nurbs_ops.py:343-390``
without reusing it.  ``sumfact_terms`` is the same Kronecker formula for any list of
(weights, basis, coefficient, sign) terms (pinned by reproducing ``bspline_mass`` and
``bspline_stiffness``); the load vectors ``nurbs_*force_gauss`` integrate the rational
basis directly.  Test infrastructure only.
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np
from scipy import sparse as sp


def kron_list(mats: Sequence) -> sp.csr_array:
    """``M_d (x) ... (x) M_1`` for ``mats = [M_1, ..., M_d]``."""
    out = sp.csr_array(mats[0])
    for m in mats[1:]:
        out = sp.csr_array(sp.kron(sp.csr_array(m), out))
    return out


def basis_kron(quadrule_list, ders) -> sp.csr_array:
    return kron_list([q.basis[b] for q, b in zip(quadrule_list, ders)])


def weights_kron(quadrule_list, ders) -> sp.csr_array:
    return kron_list([q.weights[z] for q, z in zip(quadrule_list, ders)])


def bspline_stiffness(quadrule_list, C, time_ders=None) -> sp.csr_array:
    """``K = sum_ij W^{beta_j + 2 alpha_i} diag(C_ij) B^{beta_j}``."""
    nd = len(quadrule_list)
    sd = nd - 1 if time_ders is not None else nd
    K = None
    for i in range(sd):
        for j in range(sd):
            beta = np.zeros(nd, dtype=int)
            beta[j] = 1
            alpha = np.zeros(nd, dtype=int)
            alpha[i] = 1
            if time_ders is not None:
                beta[-1] = time_ders[1]
                alpha[-1] = time_ders[0]
            term = weights_kron(quadrule_list, beta + 2 * alpha) @ sp.diags_array(C[i, j]) @ basis_kron(
                quadrule_list, beta
            )
            K = term if K is None else K + term
    return sp.csr_array(K)


def bspline_mass(quadrule_list, m, time_ders=None) -> sp.csr_array:
    nd = len(quadrule_list)
    bd = np.zeros(nd, dtype=int)
    wd = np.zeros(nd, dtype=int)
    if time_ders is not None:
        bd[-1] = time_ders[1]
        wd[-1] = time_ders[0]
    return sp.csr_array(weights_kron(quadrule_list, wd) @ sp.diags_array(m) @ basis_kron(quadrule_list, bd))


def gauss_point_weights(quadrule_list) -> np.ndarray:
    w = np.asarray(quadrule_list[0].parametric_weights)
    for q in quadrule_list[1:]:
        w = np.kron(np.asarray(q.parametric_weights), w)
    return w


def rational_basis(quadrule_list, nurbs_weights) -> List[sp.csr_array]:
    """``[R, dR/dxi_1, ..., dR/dxi_d]`` as ``(N_q x N_dof)`` matrices."""
    nd = len(quadrule_list)
    Wd = sp.diags_array(np.asarray(nurbs_weights, dtype=float))
    N0 = basis_kron(quadrule_list, np.zeros(nd, dtype=int))
    W = N0 @ np.asarray(nurbs_weights, dtype=float)
    out = [sp.csr_array(sp.diags_array(1.0 / W) @ N0 @ Wd)]
    for j in range(nd):
        b = np.zeros(nd, dtype=int)
        b[j] = 1
        Nj = basis_kron(quadrule_list, b)
        dW = Nj @ np.asarray(nurbs_weights, dtype=float)
        dR = sp.diags_array(1.0 / W) @ Nj @ Wd - sp.diags_array(dW / W ** 2) @ N0 @ Wd
        out.append(sp.csr_array(dR))
    return out


def nurbs_stiffness_gauss(quadrule_list, C, nurbs_weights) -> sp.csr_array:
    """``K_AB = sum_q omega_q dR_A^T C dR_B`` for Gauss rules."""
    R = rational_basis(quadrule_list, nurbs_weights)
    om = gauss_point_weights(quadrule_list)
    nd = len(quadrule_list)
    K = None
    for i in range(nd):
        for j in range(nd):
            term = R[1 + i].T @ sp.diags_array(om * C[i, j]) @ R[1 + j]
            K = term if K is None else K + term
    return sp.csr_array(K)


def nurbs_mass_gauss(quadrule_list, m, nurbs_weights) -> sp.csr_array:
    R = rational_basis(quadrule_list, nurbs_weights)
    om = gauss_point_weights(quadrule_list)
    return sp.csr_array(R[0].T @ sp.diags_array(om * m) @ R[0])


def sumfact_terms(quadrule_list, terms) -> sp.csr_array:
    """``sum_t s_t W^{zeta_t} diag(c_t) B^{beta_t}`` for ``terms`` of
    ``(zeta, beta, c, s)`` tuples: ``Operations.sumfact_assembling``
    (``operators/utils/operations.py:59-115``) summed over terms, written as the
    explicit Kronecker products it factorises."""
    A = None
    for zeta, beta, c, s in terms:
        term = s * (weights_kron(quadrule_list, zeta) @ sp.diags_array(np.asarray(c, dtype=float)) @
                    basis_kron(quadrule_list, beta))
        A = term if A is None else A + term
    return sp.csr_array(A)


def nurbs_force_gauss(quadrule_list, c, nurbs_weights) -> np.ndarray:
    """``F_A = sum_q omega_q R_A c_q`` (Gauss rules), ``c`` of shape (nm, N_q)."""
    R = rational_basis(quadrule_list, nurbs_weights)
    om = gauss_point_weights(quadrule_list)
    return np.stack([R[0].T @ (om * ci) for ci in np.atleast_2d(c)])


def nurbs_grad_force_gauss(quadrule_list, c, nurbs_weights) -> np.ndarray:
    """``F_A = sum_q omega_q grad R_A . c_q`` (Gauss rules), ``c`` of shape (nm, d, N_q)."""
    R = rational_basis(quadrule_list, nurbs_weights)
    om = gauss_point_weights(quadrule_list)
    nd = len(quadrule_list)
    return np.stack([sum(R[1 + i].T @ (om * ci[i]) for i in range(nd)) for ci in c])


def nurbs_gradu_v_gauss(quadrule_list, c, nurbs_weights) -> sp.csr_array:
    """``G_AB = sum_q omega_q sum_i d_i R_A c_i R_B`` (test gradient, trial value;
    the integrand of ``nurbs_ops.py:392-429`` with ``mf_type='gradu_v'``), Gauss rules."""
    R = rational_basis(quadrule_list, nurbs_weights)
    om = gauss_point_weights(quadrule_list)
    nd = len(quadrule_list)
    return sp.csr_array(sum(R[1 + i].T @ sp.diags_array(om * c[i]) @ R[0] for i in range(nd)))


def nurbs_u_gradv_gauss(quadrule_list, c, nurbs_weights) -> sp.csr_array:
    """``G_AB = sum_q omega_q sum_j R_A c_j d_j R_B`` (test value, trial gradient),
    Gauss rules."""
    R = rational_basis(quadrule_list, nurbs_weights)
    om = gauss_point_weights(quadrule_list)
    nd = len(quadrule_list)
    return sp.csr_array(sum(R[0].T @ sp.diags_array(om * c[j]) @ R[1 + j] for j in range(nd)))
