"""Pins of the oracle's FD preconditioners (oracle/fastdiag.py)."""
import numpy as np
import pytest
import scipy.linalg as sla
from scipy.sparse.linalg import spsolve

from igaops_amd.quadrature import StandardGauss, make_open_knotvector
from oracle import assembly as asm
from oracle import fastdiag as ofd

VALS = [3.0, 2.0, 5.0]


def gauss(p=3, nel=4):
    return StandardGauss(p, make_open_knotvector(p, nel), "legendre")


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_fd_equals_dense_kron_reconstruction(ndim):
    """``tests/test_tensor_operations.py:130-156`` restated."""
    qs = [gauss()] * ndim
    table = np.ones((1, ndim, 2), dtype=bool)
    table[0, 0, 0] = False
    fd = ofd.SpaceFD(qs, table)
    free = fd.free_nodes()
    UU = fd.eigvecs[0]
    for U in fd.eigvecs[1:]:
        UU = np.kron(U, UU)
    lam = fd.eigenvalues()
    v = np.random.default_rng(0).random(int(np.prod(fd.nnz)))
    ref = np.zeros_like(v)
    ref[free] = UU @ ((UU.T @ v[free]) / lam)
    np.testing.assert_allclose(fd.apply(v), ref, rtol=1e-12, atol=1e-14)


def test_eigenvalues_and_vectors_against_independent_eigh():
    """``tests/test_fast_diagonalization.py:20-119`` restated: mixed Dirichlet table,
    stiffness correctors [3, 2, 5]."""
    q = gauss()
    W, B = q.weights, q.basis
    M = (W[0] @ B[0]).toarray()
    K = (W[-1] @ B[-1]).toarray()
    idx = [np.arange(1, q.nbctrlpts), np.arange(0, q.nbctrlpts), np.arange(0, q.nbctrlpts - 1)]
    ev, U = [], []
    for i in idx:
        Kd = K[np.ix_(i, i)] + (1e-8 * M[np.ix_(i, i)] if len(i) == q.nbctrlpts else 0)
        l, u = sla.eigh(Kd, M[np.ix_(i, i)])
        ev.append(l)
        U.append(u)
    ones = [np.ones_like(l) for l in ev]
    ref = (VALS[0] * np.kron(np.kron(ones[2], ones[1]), ev[0]) + VALS[1] * np.kron(np.kron(ones[2], ev[1]), ones[0])
           + VALS[2] * np.kron(np.kron(ev[2], ones[1]), ones[0]))
    table = np.zeros((1, 3, 2), dtype=bool)
    table[0, 0, 0] = True
    table[0, 2, 1] = True
    fd = ofd.SpaceFD([q] * 3, table)
    fd.stiff_corr = np.array(VALS)
    np.testing.assert_allclose(fd.eigenvalues(), ref, rtol=1e-10)
    for d in range(3):
        Md = M[np.ix_(idx[d], idx[d])]
        np.testing.assert_allclose(fd.eigvecs[d].T @ Md @ fd.eigvecs[d], np.eye(len(idx[d])), atol=1e-10)


def test_fd_is_exact_inverse_of_separable_stiffness():
    """``tests/test_tensor_operations.py:159-188``: on the Gauss cube with all faces
    Dirichlet the FD apply solves the trimmed stiffness exactly."""
    q = gauss(3, 4)
    qs = [q] * 3
    nq = q.nbquadpts ** 3
    C = np.zeros((3, 3, nq))
    for i in range(3):
        C[i, i] = 1.0
    K = asm.bspline_stiffness(qs, C)
    fd = ofd.SpaceFD(qs, np.ones((1, 3, 2), dtype=bool))
    free = fd.free_nodes()
    v = np.random.default_rng(1).random(K.shape[0])
    ref = np.zeros_like(v)
    ref[free] = spsolve(K[np.ix_(free, free)].tocsc(), v[free])
    np.testing.assert_allclose(fd.apply(v), ref, rtol=1e-8, atol=1e-10)


def test_arrowhead_decomposition_invariants():
    qt = StandardGauss(2, make_open_knotvector(2, 6), "legendre")
    Wt, Mt = ofd.time_pencil(qt)
    Ut, diag, g = ofd.arrowhead_decomposition(Wt, Mt)
    np.testing.assert_allclose(Ut.conj().T @ Mt @ Ut, np.eye(len(diag)), atol=1e-10)
    D = np.diag(diag).astype(complex)
    D[:-1, -1] = g
    D[-1, :-1] = -g.conj()
    np.testing.assert_allclose(Ut.conj().T @ Wt @ Ut, D, atol=1e-9)


def test_arrowhead_solve_against_dense():
    rng = np.random.default_rng(3)
    k, m = 5, 7
    H = rng.standard_normal((k + 1, m)) + 1j * rng.standard_normal((k + 1, m)) + 4.0
    b = rng.standard_normal(k) + 1j * rng.standard_normal(k)
    rhs = rng.standard_normal((k + 1) * m) + 1j * rng.standard_normal((k + 1) * m)
    x = ofd.solve_arrowhead(H, b, rhs)
    # dense system: for each spatial index j the (k+1) x (k+1) arrowhead block
    Xr = rhs.reshape(k + 1, m)
    X = x.reshape(k + 1, m)
    for j in range(m):
        A = np.diag(H[:, j]).astype(complex)
        A[:-1, -1] = b
        A[-1, :-1] = -b.conj()
        np.testing.assert_allclose(A @ X[:, j], Xr[:, j], rtol=1e-10, atol=1e-10)


def test_spacetime_fd_inverts_separable_spacetime_operator():
    """With Gauss rules the arrowhead FD is the exact inverse of the separable
    space-time operator ``W_t (x) M_s + M_t (x) K_s`` on the free nodes."""
    qs = gauss(2, 3)
    qt = StandardGauss(2, make_open_knotvector(2, 4), "legendre")
    sfd = ofd.SpaceFD([qs] * 3, np.ones((1, 3, 2), dtype=bool))
    st = ofd.SpaceTimeFD(sfd, qt)
    M = (qs.weights[0] @ qs.basis[0]).toarray()
    K = (qs.weights[-1] @ qs.basis[-1]).toarray()
    Mt = (qt.weights[0] @ qt.basis[0]).toarray()
    Wt = (qt.weights[1] @ qt.basis[1]).toarray()
    kron3 = lambda a, b, c: np.kron(c, np.kron(b, a))
    Ms = kron3(M, M, M)
    Ks = kron3(K, M, M) + kron3(M, K, M) + kron3(M, M, K)
    A = np.kron(Wt, Ms) + np.kron(Mt, Ks)
    free = st.free_nodes()
    v = np.random.default_rng(2).standard_normal(A.shape[0])
    ref = np.zeros_like(v)
    ref[free] = np.linalg.solve(A[np.ix_(free, free)], v[free])
    np.testing.assert_allclose(st.apply(v), ref, rtol=1e-7, atol=1e-9 * np.abs(ref).max())
    refT = np.zeros_like(v)
    refT[free] = np.linalg.solve(A[np.ix_(free, free)].T, v[free])
    np.testing.assert_allclose(st.apply_transpose(v), refT, rtol=1e-7, atol=1e-9 * np.abs(refT).max())


def test_qz_spacetime_fd_inverts_separable_spacetime_operator():
    """The legacy QZ form: ``Q S Z^H = W_t``, ``Q T Z^H = M_t`` with S, T upper
    triangular and Q, Z unitary, and, with Gauss rules, the same exact inverse of the
    separable space-time operator as the arrowhead form (apply and transpose).  A
    planted error (the transposed solve without the transpose) fails."""
    qs = gauss(2, 3)
    qt = StandardGauss(2, make_open_knotvector(2, 4), "legendre")
    sfd = ofd.SpaceFD([qs] * 3, np.ones((1, 3, 2), dtype=bool))
    st = ofd.SpaceTimeQZFD(sfd, qt)
    Wt, Mt = ofd.time_pencil(qt)
    np.testing.assert_allclose(st.Q @ st.S @ st.Z.conj().T, Wt, atol=1e-12)
    np.testing.assert_allclose(st.Q @ st.T @ st.Z.conj().T, Mt, atol=1e-12)
    assert np.allclose(np.tril(st.S, -1), 0) and np.allclose(np.tril(st.T, -1), 0)
    np.testing.assert_allclose(st.Q.conj().T @ st.Q, np.eye(len(Wt)), atol=1e-12)
    M = (qs.weights[0] @ qs.basis[0]).toarray()
    K = (qs.weights[-1] @ qs.basis[-1]).toarray()
    Mtf = (qt.weights[0] @ qt.basis[0]).toarray()
    Wtf = (qt.weights[1] @ qt.basis[1]).toarray()
    kron3 = lambda a, b, c: np.kron(c, np.kron(b, a))
    Ms = kron3(M, M, M)
    Ks = kron3(K, M, M) + kron3(M, K, M) + kron3(M, M, K)
    A = np.kron(Wtf, Ms) + np.kron(Mtf, Ks)
    free = st.free_nodes()
    v = np.random.default_rng(3).standard_normal(A.shape[0])
    ref = np.zeros_like(v)
    ref[free] = np.linalg.solve(A[np.ix_(free, free)], v[free])
    np.testing.assert_allclose(st.apply(v), ref, rtol=1e-7, atol=1e-9 * np.abs(ref).max())
    refT = np.zeros_like(v)
    refT[free] = np.linalg.solve(A[np.ix_(free, free)].T, v[free])
    np.testing.assert_allclose(st.apply_transpose(v), refT, rtol=1e-7, atol=1e-9 * np.abs(refT).max())
    arrow = ofd.SpaceTimeFD(sfd, qt)
    np.testing.assert_allclose(st.apply(v), arrow.apply(v), atol=1e-9 * np.abs(ref).max())

    class Planted(ofd.SpaceTimeQZFD):
        def _solve(self, t, transpose):
            return super()._solve(t, False)

    bad = Planted(sfd, qt)
    assert not np.allclose(bad.apply_transpose(v), refT, rtol=1e-7, atol=1e-9 * np.abs(refT).max())
