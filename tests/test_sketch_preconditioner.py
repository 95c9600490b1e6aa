"""Randomized low-rank preconditioners of ``mu I + S``: the oracle pinned by closed
forms (CPU), the library's device apply against the oracle (GPU)."""
import numpy as np
import pytest
from scipy.sparse.linalg import aslinearoperator

from oracle import sketch as osk


def low_rank(n, r, rng, spd):
    A = rng.standard_normal((n, r))
    if spd:
        return A @ np.diag(np.linspace(5.0, 1.0, r)) @ A.T
    return A @ rng.standard_normal((r, n))


def test_dense_form_is_the_inverse():
    rng = np.random.default_rng(0)
    n = 40
    S = low_rank(n, 40, rng, True)
    for cls in (osk.SPD, osk.General):
        P = cls(aslinearoperator(S), 0.3, rank=8, small_threshold=576, random_state=1).build()
        assert P.dense
        x = rng.standard_normal(n)
        np.testing.assert_allclose(P.apply(x), np.linalg.solve(S + 0.3 * np.eye(n), x), rtol=1e-10)


@pytest.mark.parametrize("rank", [12, 0])
def test_spd_form_on_an_exactly_low_rank_operator(rank):
    """S of rank 8 with the retained rank at least 8 (fixed 12, or adaptive): the
    basis spans range(S) and P^{-1} (mu I + S) = (lambda_r + mu) U U^T + mu (I - U U^T)."""
    rng = np.random.default_rng(1)
    n, mu = 700, 0.5
    S = low_rank(n, 8, rng, True)
    P = osk.SPD(aslinearoperator(S), mu, rank=rank, small_threshold=576, random_state=2).build()
    assert not P.dense and P.U.shape[1] >= 8
    # the range of S is captured
    np.testing.assert_allclose(P.U @ (P.U.T @ S), S, atol=1e-8 * np.abs(S).max())
    x = rng.standard_normal(n)
    lhs = P.apply(mu * x + S @ x)
    proj = P.U @ (P.U.T @ x)
    np.testing.assert_allclose(lhs, (P.Lam[-1] + mu) * proj + mu * (x - proj), rtol=1e-8, atol=1e-9)


@pytest.mark.parametrize("rank", [12, 0])
def test_general_form_is_exact_on_a_low_rank_operator(rank):
    """Non-symmetric S of rank 6: the Woodbury form inverts mu I + S; a planted sign
    error in its correction fails."""
    rng = np.random.default_rng(3)
    n, mu = 650, 2.0
    S = low_rank(n, 6, rng, False)
    P = osk.General(aslinearoperator(S), mu, rank=rank, small_threshold=576, random_state=4).build()
    x = rng.standard_normal(n)
    b = mu * x + S @ x
    Uk = P.U[:, P.Lam > 1e-8 * P.Lam[0]]
    np.testing.assert_allclose(Uk @ (Uk.T @ S), S, atol=1e-8 * np.abs(S).max())
    if np.all(P.Lam > 1e-8 * P.Lam[0]):
        np.testing.assert_allclose(P.apply(b), x, rtol=1e-8, atol=1e-8)
    bad = lambda v: v / mu + P.U @ osk.lu_solve(P.R, P.V.T @ (v / mu))
    assert not np.allclose(bad(b), x, rtol=1e-6, atol=1e-6)


def test_adaptive_range_finder_captures_a_low_rank_range():
    rng = np.random.default_rng(5)
    S = low_rank(300, 7, rng, False)
    Q = osk.adaptive_range_finder(aslinearoperator(S), 300, 10, 1e-2, np.random.default_rng(6))
    assert 7 <= Q.shape[1] <= 12
    np.testing.assert_allclose(Q @ (Q.T @ S), S, atol=1e-8 * np.abs(S).max())


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["spd", "general"])
@pytest.mark.parametrize("rank,n", [(16, 900), (0, 800), (8, 300)])
def test_device_apply_matches_oracle(kind, rank, n):
    """``RandomPreconditioner`` built on the same sketch (same seed) as the oracle and
    applied on the device (numpy and device vectors, one and several columns): equal to
    the oracle's apply to rounding; rank 8 at n = 300 is the dense form."""
    import torch

    from igaops_amd.sketch import RandomPreconditioner

    print("case", kind, rank, n, flush=True)
    rng = np.random.default_rng(7)
    S = low_rank(n, 10, rng, kind == "spd") + 1e-3 * np.eye(n)
    op = aslinearoperator(S)
    P = RandomPreconditioner(op, 0.7, kind, rank=rank, small_threshold=576, random_state=9)
    P.build()
    O = (osk.SPD if kind == "spd" else osk.General)(op, 0.7, rank=rank, small_threshold=576, random_state=9).build()
    assert P.is_dense == O.dense
    if not P.is_dense:
        assert P.low_rank_size == O.Lam.size
    X = rng.standard_normal((n, 3))
    ref = np.stack([O.apply(X[:, k]) for k in range(3)], axis=1)
    got = P.apply(X)
    assert np.linalg.norm(got - ref) <= 1e-10 * np.linalg.norm(ref)
    xd = torch.as_tensor(X[:, 1], device="cuda")
    yd = P.apply(xd)
    assert isinstance(yd, torch.Tensor) and yd.is_cuda
    assert np.linalg.norm(yd.cpu().numpy() - ref[:, 1]) <= 1e-10 * np.linalg.norm(ref[:, 1])
    assert "rank" in repr(P) or "dense" in repr(P)
