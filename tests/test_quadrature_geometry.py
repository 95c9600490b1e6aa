"""Host set-up pins: basis evaluation, Gauss and WQ rules, geometry maps."""
import json
import os

import numpy as np
import pytest
from scipy.interpolate import BSpline
from scipy.sparse.linalg import lsqr

from igaops_amd.geometry import lognormal_conductivity, quarter_annulus, rational_arc_ctrlpts, unit_cube
from igaops_amd.quadrature import (
    StandardGauss,
    WeightedQuadrature,
    eval_ders_basis_sparse,
    eval_nbquadpts_per_element,
    lobatto,
    make_open_knotvector,
    solve_min_weighted_norm,
)

GOLD = os.path.join(os.path.dirname(__file__), "golden", "wq_point_counts.json")


@pytest.mark.parametrize("p", [1, 2, 3, 5])
def test_basis_matches_scipy_bspline(p):
    kv = make_open_knotvector(p, 6, multiplicity=1)
    x = np.linspace(0, 1, 53)[:-1]  # scipy is open at the right end
    n = len(kv) - p - 1
    B = eval_ders_basis_sparse(p, kv, x, nders=min(p, 2))
    for k in range(min(p, 2) + 1):
        ref = np.stack([np.nan_to_num(BSpline(kv, np.eye(n)[i], p, extrapolate=False)(x, nu=k)) for i in range(n)], 1)
        np.testing.assert_allclose(B[k].toarray(), ref, atol=1e-11 * max(1, np.abs(ref).max()))
    # partition of unity (values) and zero sum of derivatives, including x = 1
    B1 = eval_ders_basis_sparse(p, kv, np.array([0.0, 0.37, 1.0]), nders=1)
    np.testing.assert_allclose(B1[0].sum(axis=1), 1.0, rtol=1e-14)
    np.testing.assert_allclose(B1[1].sum(axis=1), 0.0, atol=1e-11)


def test_point_counts_golden():
    for p, nel, n2, n1, ng in json.load(open(GOLD))["cases"]:
        kv = make_open_knotvector(p, nel)
        assert int(eval_nbquadpts_per_element(p, kv, "2").sum()) + 1 == n2
        assert int(eval_nbquadpts_per_element(p, kv, "1").sum()) + 1 == n1
        assert StandardGauss(p, kv, "legendre").nbquadpts == ng


@pytest.mark.parametrize("position_rule", ["endpoint", "cellcenter"])
@pytest.mark.parametrize("quadtype", ["1", "2"])
def test_wq_integrals_match_gauss(position_rule, quadtype):
    """``tests/test_weighted_quadrature.py:13-64`` restated: ``W_kl B_l`` equals the
    Gauss-exact 1-D integrals to 1e-8 (relative Frobenius)."""
    errs = []
    for p in range(2, 5):
        for nel in (4, 8, 16):
            kv = make_open_knotvector(p, nel)
            wq = WeightedQuadrature(p, kv, quadtype=quadtype, position_rule=position_rule)
            gs = StandardGauss(p, kv, "legendre")
            for wi, bi in [(0, 0), (1, 1), (2, 0), (3, 1)]:
                ref = (gs.weights[wi] @ gs.basis[bi]).toarray()
                app = (wq.weights[wi] @ wq.basis[bi]).toarray()
                errs.append(np.linalg.norm(app - ref) / np.linalg.norm(ref))
    assert max(errs) < 1e-8


def test_min_norm_weights_closed_form():
    """The weights are ``z o pinv(A diag z) b`` (the minimum weighted-norm solution) on
    the full system, although each row is solved on the block of its nonzero columns and
    the rows of ``A`` touching them: rows with different supports, a zero inside a
    support, rows of ``A`` touching no column of a support, and a row of ``Z`` that is
    zero."""
    rng = np.random.default_rng(0)
    A = rng.standard_normal((7, 12))
    A[0, 6:] = 0.0  # row 0 touches only columns 0-5
    A[6, :9] = 0.0  # row 6 touches only columns 9-11
    Z = np.abs(rng.standard_normal((5, 12)))
    Z[:2, :3] = 0.0
    Z[2, 6:] = 0.0
    Z[3, 7] = 0.0
    Z[4] = 0.0
    Bm = rng.standard_normal((5, 7))
    got = solve_min_weighted_norm(Z, A, Bm)
    for i in range(5):
        ref = Z[i] * (np.linalg.pinv(A * Z[i]) @ Bm[i])
        np.testing.assert_allclose(got[i], ref, rtol=1e-10, atol=1e-12)


def test_min_norm_weights_match_lsqr_on_a_wq_system():
    """On the (well-conditioned) value-weight system of a cubic rule, the batched
    solution equals one ``lsqr`` per row on the full system, as the reference computes
    it (``quadrature/utils/operations.py:107-143``)."""
    from igaops_amd.common import Constants
    from igaops_amd.quadrature import _knot_support

    kv = make_open_knotvector(3, 8)
    wq = WeightedQuadrature(3, kv, quadtype="2")
    gs = StandardGauss(3, kv, "legendre")
    A = gs.eval_basis(wq.quadpts)[0].T.toarray()
    Z = (gs.eval_basis(wq.quadpts)[0].T.toarray() * _knot_support(wq.quadpts, gs.unique_kv))
    Bm = (gs.weights[0] @ gs.basis[0]).toarray()
    tol = Constants.SAFEGUARD
    ref = np.stack([z * lsqr(A * z, b, atol=tol, btol=tol)[0] for z, b in zip(Z, Bm)])
    np.testing.assert_allclose(solve_min_weighted_norm(Z, A, Bm), ref, rtol=0, atol=1e-12)


def test_gauss_lobatto():
    x, w = lobatto(5)
    xr = np.array([-1.0, -np.sqrt(3 / 7), 0.0, np.sqrt(3 / 7), 1.0])
    wr = np.array([1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10])
    np.testing.assert_allclose(x, xr, atol=1e-14)
    np.testing.assert_allclose(w, wr, atol=1e-14)


def test_unit_cube_identity_map():
    C, m = unit_cube(3, 4).conductivity_coefficients()
    np.testing.assert_allclose(m, 1.0, atol=1e-13)
    np.testing.assert_allclose(C, np.broadcast_to(np.eye(3)[:, :, None], C.shape), atol=1e-13)


def test_quarter_annulus_closed_forms():
    """Points lie on circles of radius 1 + xi_1; detJ = r (r_out - r_in) |c'(xi_2)| H;
    the arc control net reproduces the quadratic rational arc exactly."""
    patch = quarter_annulus(3, 5, quadclass="gauss", quadtype="legendre")
    q = patch.quadrule_list[0]
    pg = patch.plane_geometry()
    r = np.hypot(pg.x[0], pg.x[1])
    np.testing.assert_allclose(r, np.broadcast_to(1.0 + q.quadpts[None, :], r.shape), atol=1e-13)
    theta = np.arctan2(pg.x[1], pg.x[0])
    assert theta.min() >= -1e-14 and theta.max() <= np.pi / 2 + 1e-14
    # angular speed from the rational arc: c(t) = X^w(t) / W(t)
    t = q.quadpts
    wm = np.cos(np.pi / 4)
    W = (1 - t) ** 2 + 2 * t * (1 - t) * wm + t ** 2
    X = (1 - t) ** 2 + 2 * t * (1 - t) * wm
    Y = 2 * t * (1 - t) * wm + t ** 2
    dW = -2 * (1 - t) + 2 * (1 - 2 * t) * wm + 2 * t
    dX = -2 * (1 - t) + 2 * (1 - 2 * t) * wm
    dY = 2 * (1 - 2 * t) * wm + 2 * t
    speed = np.hypot((dX * W - X * dW) / W ** 2, (dY * W - Y * dW) / W ** 2)
    det_ref = (1.0 + q.quadpts[None, None, :]) * 1.0 * speed[None, :, None] * 1.0
    det = patch.det_jac().reshape(q.nbquadpts, q.nbquadpts, q.nbquadpts)
    np.testing.assert_allclose(det, np.broadcast_to(det_ref, det.shape), rtol=1e-12)


def test_lognormal_field_statistics():
    patch = unit_cube(2, 8)
    k1 = lognormal_conductivity(patch, 0.5, seed=3)
    k2 = lognormal_conductivity(patch, 0.5, seed=3)
    np.testing.assert_array_equal(k1, k2)
    g = np.log(k1) / 0.5
    assert np.all(k1 > 0) and 0.3 < g.std() < 2.0


def test_cross_section_nurbs_weight_function_closed_form():
    """The quarter annulus's NURBS weight function on the cross-section grid (what the
    generated coefficients fold in): W(xi_2) of the quadratic rational arc, whatever the
    degree elevation and refinement, constant along the radial xi_1; the xi_3 weights are
    the cross-section's repeated, so these are the weight function of the whole patch."""
    patch = quarter_annulus(4, 6, quadclass="gauss", quadtype="legendre")
    t = patch.quadrule_list[1].quadpts
    wm = np.cos(np.pi / 4)
    W = (1 - t) ** 2 + 2 * t * (1 - t) * wm + t ** 2
    dW = -2 * (1 - t) + 2 * (1 - 2 * t) * wm + 2 * t
    w = patch.plane_geometry().w
    assert w.shape == (3,) + patch.plane_geometry().x.shape[1:]
    np.testing.assert_allclose(w[0], np.broadcast_to(W[:, None], w[0].shape), rtol=1e-13)
    np.testing.assert_allclose(w[1], 0.0, atol=1e-12)
    np.testing.assert_allclose(w[2], np.broadcast_to(dW[:, None], w[2].shape), rtol=1e-12, atol=1e-13)
    n1, n2, n3 = patch.nbctrlpts
    np.testing.assert_array_equal(patch.nurbs_weights.reshape(n3, n1 * n2), np.tile(patch.weights2d, (n3, 1)))
    assert unit_cube(2, 3).plane_geometry().w is None


def test_alpha_only_rule_answer_follows_the_weights():
    """The fused-operator rule test is kept per rule and redone when the rule's weight
    matrices are replaced (a rebuild of the rule sets new ones)."""
    from igaops_amd.bands import test_weights_depend_on_alpha_only as alpha_only

    q = WeightedQuadrature(3, make_open_knotvector(3, 6), quadtype="2")
    assert alpha_only(q) and alpha_only(q)
    w = list(q.weights)
    q._weights = [w[0], 2.0 * w[1], w[2], w[3]]  # as a rebuild of the rule would
    assert not alpha_only(q)
    q._weights = w
    assert alpha_only(q)
