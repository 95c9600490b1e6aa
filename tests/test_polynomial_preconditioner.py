"""GMRES polynomial preconditioner: oracle pins (CPU) and the device class against
the oracle (GPU)."""
import numpy as np
import pytest

from oracle import polynomial as opl


def _spd(n, seed=0, cond=100.0):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.geomspace(1.0, cond, n)
    return (Q * lam) @ Q.T, lam


def test_full_degree_on_diagonal_is_the_inverse():
    lam = np.linspace(1.0, 7.0, 12)
    A = np.diag(lam)
    roots, m = opl.harmonic_ritz(lambda v: A @ v, 12, 12)
    assert m == 12
    np.testing.assert_allclose(np.sort(roots), lam, rtol=1e-8)
    x = np.random.default_rng(1).standard_normal(12)
    y = opl.apply(lambda v: A @ v, opl.leja_order(roots), x)
    np.testing.assert_allclose(y, x / lam, rtol=1e-7)


def test_harmonic_ritz_values_lie_in_the_spectral_interval():
    A, lam = _spd(60, seed=2)
    roots, m = opl.harmonic_ritz(lambda v: A @ v, 60, 10)
    assert m == 10
    assert np.all(roots >= lam.min() * (1 - 1e-10)) and np.all(roots <= lam.max() * (1 + 1e-10))


def test_residual_polynomial_is_the_gmres_residual():
    """b - A p(A) b = prod (I - A / theta) b is the m-step GMRES residual from x0 = 0,
    computed independently by least squares over the Krylov space."""
    A, _ = _spd(50, seed=3, cond=30.0)
    m = 8
    b = np.random.default_rng(0).standard_normal(50)  # the default b0
    roots, _ = opl.harmonic_ritz(lambda v: A @ v, 50, m)
    y = opl.apply(lambda v: A @ v, opl.leja_order(roots), b)
    K = np.stack([np.linalg.matrix_power(A, k) @ b for k in range(m)], axis=1)
    Q, _ = np.linalg.qr(K)
    c = np.linalg.lstsq(A @ Q, b, rcond=None)[0]
    r_gmres = b - A @ Q @ c
    np.testing.assert_allclose(b - A @ y, r_gmres, rtol=1e-6, atol=1e-9 * np.linalg.norm(b))


def test_leja_order_does_not_change_the_result():
    A, _ = _spd(40, seed=4, cond=20.0)
    roots, _ = opl.harmonic_ritz(lambda v: A @ v, 40, 6)
    x = np.random.default_rng(5).standard_normal(40)
    a = opl.apply(lambda v: A @ v, opl.leja_order(roots), x)
    b = opl.apply(lambda v: A @ v, list(roots), x)
    np.testing.assert_allclose(a, b, rtol=1e-10)


def test_planted_root_error_is_detected():
    """Dropping one root must break the GMRES-residual property."""
    A, _ = _spd(30, seed=6, cond=20.0)
    b = np.random.default_rng(0).standard_normal(30)
    roots, _ = opl.harmonic_ritz(lambda v: A @ v, 30, 6)
    full = b - A @ opl.apply(lambda v: A @ v, list(roots), b)
    short = b - A @ opl.apply(lambda v: A @ v, list(roots[:-1]), b)
    assert np.linalg.norm(full - short) > 1e-3 * np.linalg.norm(full)


@pytest.mark.gpu
def test_device_polynomial_matches_oracle():
    import torch

    from igaops_amd.polynomial import GMRESPolynomialPreconditioner

    A, _ = _spd(200, seed=7, cond=50.0)
    P = GMRESPolynomialPreconditioner(A, degree=10)
    roots, m = opl.harmonic_ritz(lambda v: A @ v, 200, 10)
    assert P.degree_actual == m
    np.testing.assert_allclose(np.sort(P.roots), np.sort(roots), rtol=1e-10)
    x = np.random.default_rng(8).standard_normal(200)
    ref = opl.apply(lambda v: A @ v, opl.leja_order(roots), x)
    np.testing.assert_allclose(P @ x, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    y = P.apply_device(torch.as_tensor(x, device="cuda"))
    np.testing.assert_allclose(y.cpu().numpy(), ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


@pytest.mark.gpu
def test_device_polynomial_on_the_stiffness_operator():
    """A device operator (the fused stiffness + mass matvec) as A: p(A) against the
    oracle's polynomial built from the oracle operator on the same start vector."""
    import torch

    from igaops_amd.geometry import lognormal_conductivity, unit_cube
    from igaops_amd.operators import StiffnessMassOperator
    from igaops_amd.polynomial import GMRESPolynomialPreconditioner
    from oracle import operators as oo

    patch = unit_cube(2, 6)
    C, m = patch.conductivity_coefficients(lognormal_conductivity(patch, 0.3, seed=1))
    qs = patch.quadrule_list
    n = patch.nbctrlpts_total
    op = StiffnessMassOperator(qs, C=C, m=m)
    P = GMRESPolynomialPreconditioner(op, degree=8, n=n)
    mv = lambda v: oo.bspline_gradu_gradv(qs, C, v) + oo.bspline_u_v(qs, m, v)
    roots, _ = opl.harmonic_ritz(mv, n, 8)
    np.testing.assert_allclose(np.sort(P.roots), np.sort(roots), rtol=1e-8)
    x = np.random.default_rng(9).standard_normal(n)
    ref = opl.apply(mv, opl.leja_order(roots), x)
    got = P.apply_device(torch.as_tensor(x, device="cuda")).cpu().numpy()
    np.testing.assert_allclose(got, ref, rtol=1e-7, atol=1e-10 * np.abs(ref).max())
