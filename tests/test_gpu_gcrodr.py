"""GPU GCRODR (recycled Krylov) against the oracle and the reference's own test case."""
import numpy as np
import pytest

from oracle import solvers as osv

pytestmark = pytest.mark.gpu

N, NTEST, M, K, TOL = 1000, 4, 60, 20, 1e-6


def tridiag_apply():
    import torch

    def A(x, out):
        out.copy_(2.05 * x)
        out[:-1] -= x[1:]
        out[1:] -= x[:-1]
        return out

    return A


def test_gcrodr_tridiagonal_sequence():
    """``tests/test_gcrodr.py`` restated on the device: convergence, agreement with
    GMRES, true residual <= 10 tol over a sequence of scaled right-hand sides, and the
    recycled-space invariants ``C^H C = I``, ``C = M A U``."""
    import torch

    from igaops_amd.gcrodr import GCRODR, ZBlas, complex_apply

    A = tridiag_apply()
    e0 = torch.zeros(N, dtype=torch.float64, device="cuda")
    e0[0] = 1.0
    b0 = torch.empty_like(e0)
    A(e0, b0)
    rhs = [float(s) * b0 for s in np.linspace(0, 1, NTEST + 1)[1:]]
    g = GCRODR(TOL, 500, krylov_size=M, recycled_size=K)
    out = g.solve_device(A, rhs[0])
    assert out.residual.min() <= TOL
    Ad = np.diag(np.full(N, 2.05)) - np.diag(np.ones(N - 1), 1) - np.diag(np.ones(N - 1), -1)
    ref = osv.gmres(lambda v: Ad @ v, rhs[0].cpu().numpy(), tol=TOL, maxiters=500).solution
    assert np.allclose(out.solution.cpu().numpy(), ref, atol=1e-4)
    for b in rhs:
        out = g.solve_device(A, b)
        x = out.solution.cpu().numpy()
        assert np.linalg.norm(b.cpu().numpy() - Ad @ x) <= 10 * TOL
    zb = ZBlas(N)
    Cm, U = g.last_C, g.last_U
    np.testing.assert_allclose(zb.gram(Cm, Cm), np.eye(len(Cm)), atol=1e-8)
    Ac = complex_apply(A)
    for u, c in zip(U, Cm):
        au = torch.empty_like(u)
        Ac(u, au)
        assert float(torch.linalg.vector_norm(au - c)) <= 1e-8 * max(1.0, float(torch.linalg.vector_norm(c)))


def test_gcrodr_matches_oracle_iterations():
    """Same left-preconditioned problem in the oracle: same number of matvecs (+-2)
    and solutions within the tolerance."""
    import torch

    from igaops_amd.gcrodr import GCRODR

    A = tridiag_apply()
    Ad = np.diag(np.full(N, 2.05)) - np.diag(np.ones(N - 1), 1) - np.diag(np.ones(N - 1), -1)
    rng = np.random.default_rng(0)
    b = rng.standard_normal(N)
    dg = GCRODR(1e-8, 2000, krylov_size=30, recycled_size=10)
    og = osv.GCRODR(1e-8, 2000, krylov_size=30, recycled_size=10)
    for s in (1.0, 1.05):
        bd = torch.as_tensor(s * b, device="cuda")
        o1 = dg.solve_device(A, bd)
        o2 = og.solve(lambda v: Ad @ v, s * b)
        assert o1.success and o2.success
        assert abs(o1.iterations - o2.iterations) <= 2
        x1, x2 = o1.solution.cpu().numpy(), o2.solution
        assert np.linalg.norm(x1 - x2) <= 1e-6 * np.linalg.norm(x2)


@pytest.mark.parametrize("m,k", [(80, 30), (100, 70)])
def test_gcrodr_beyond_64_basis_vectors_matches_oracle(m, k):
    """Krylov and recycled sizes past one block call (the tall products are split over
    32-row gram and 64-row combine calls): the oracle's matvec counts (+-2) and
    solutions over a sequence in which the recycled space is carried across solves."""
    import torch

    from igaops_amd.gcrodr import GCRODR

    n = 3000
    Ad = np.diag(np.full(n, 2.01)) - np.diag(np.ones(n - 1), 1) - np.diag(np.ones(n - 1), -1)

    def A(x, out):
        out.copy_(2.01 * x)
        out[:-1] -= x[1:]
        out[1:] -= x[:-1]
        return out

    rng = np.random.default_rng(3)
    b = rng.standard_normal(n)
    dg = GCRODR(1e-8, 5000, krylov_size=m, recycled_size=k)
    og = osv.GCRODR(1e-8, 5000, krylov_size=m, recycled_size=k)
    for s in (1.0, 1.02, 0.97):
        bd = torch.as_tensor(s * b, device="cuda")
        o1 = dg.solve_device(A, bd)
        o2 = og.solve(lambda v: Ad @ v, s * b)
        assert o1.success and o2.success
        if s == 1.0:
            assert o1.iterations > m  # more than one cycle: the recycled space is built and used
        assert abs(o1.iterations - o2.iterations) <= 2
        x1, x2 = o1.solution.cpu().numpy(), o2.solution
        assert np.linalg.norm(x1 - x2) <= 1e-6 * np.linalg.norm(x2)
        assert np.linalg.norm(s * b - Ad @ x1) <= 10 * 1e-8 * np.linalg.norm(s * b)
