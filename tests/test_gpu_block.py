"""GPU block kernels of the recycled Krylov solver on complex row lists, against
float64 numpy: ``iga_zgram`` (``X^H Y``) and ``iga_zcombine``.

Bound for the Gram: each part of ``G_ab`` is a sum of 2n products, so its error in
any summation order is at most 2n eps sum_i |x_i| |y_i| (times a small constant)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rows(torch, rng, n, count, pool):
    """``count`` complex rows of length n, taken from ``pool`` separate allocations
    (rows of one 2-D tensor and of separate tensors are mixed)."""
    rows = []
    for p in range(pool):
        k = (count - len(rows)) if p == pool - 1 else count // pool
        if k <= 0:
            continue
        h = rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))
        d = torch.as_tensor(h, dtype=torch.complex128).cuda()
        rows.extend(d[i] for i in range(k))
    return rows


def _gram_raw(torch, X, Y, n):
    from igaops_amd import _lib
    from igaops_amd import device as dv

    px = torch.tensor([r.data_ptr() for r in X], dtype=torch.int64).cuda()
    py = torch.tensor([r.data_ptr() for r in Y], dtype=torch.int64).cuda()
    work = dv.workspace(_lib.load().iga_zgram_workspace_bytes(n, len(X), len(Y)))
    out = torch.full((2 * len(X) * len(Y),), np.nan, dtype=torch.float64, device="cuda")
    _lib.call("iga_zgram", n, len(X), dv.ptr(px), len(Y), dv.ptr(py), dv.ptr(out), dv.ptr(work), dv.stream_ptr())
    torch.cuda.synchronize()
    v = out.cpu().numpy()
    return (v[0::2] + 1j * v[1::2]).reshape(len(X), len(Y))


@pytest.mark.parametrize("n,na,nb", [(0, 2, 3), (1, 1, 1), (7, 3, 2), (8, 1, 1), (9, 16, 8), (1000, 20, 20),
                                     (12345, 20, 1), (3001, 17, 9), (4097, 33, 40), (20011, 64, 64),
                                     (600000, 1, 1), (300001, 32, 24), (100003, 2, 2), (50001, 1, 2),
                                     (50001, 2, 1), (3, 2, 2)])
def test_zgram_matches_numpy(n, na, nb):
    """Every tile shape of the matrix-core kernel (1..2 x 1..4 fragments of 16 x 16),
    the element-per-thread kernel for at most 2 x 2 rows, ragged n (partial last step
    of 8 elements), rows from several allocations, and repeat calls giving the same
    bits."""
    import torch

    print("case", n, na, nb, flush=True)
    rng = np.random.default_rng(n + 31 * na + nb)
    X = _rows(torch, rng, n, na, 3 if na >= 3 else 1)
    Y = _rows(torch, rng, n, nb, 2 if nb >= 2 else 1)
    if n == 0:
        X = [torch.zeros(1, dtype=torch.complex128, device="cuda") for _ in range(na)]
        Y = [torch.zeros(1, dtype=torch.complex128, device="cuda") for _ in range(nb)]
    got = _gram_raw(torch, X, Y, n)
    Xh = np.stack([r[:n].cpu().numpy() for r in X])
    Yh = np.stack([r[:n].cpu().numpy() for r in Y])
    exact = Xh.conj() @ Yh.T
    bound = 8 * max(n, 1) * np.finfo(np.float64).eps * (np.abs(Xh) @ np.abs(Yh).T) + 1e-300
    assert np.all(np.abs(got.real - exact.real) <= bound)
    assert np.all(np.abs(got.imag - exact.imag) <= bound)
    again = _gram_raw(torch, X, Y, n)
    np.testing.assert_array_equal(got, again)


def test_zgram_same_rows_both_sides():
    """``C^H C`` with the same list on both sides (the CholQR Gram): Hermitian to
    rounding, real positive diagonal equal to the squared norms."""
    import torch

    rng = np.random.default_rng(5)
    n = 50021
    C = _rows(torch, rng, n, 20, 2)
    G = _gram_raw(torch, C, C, n)
    Ch = np.stack([r.cpu().numpy() for r in C])
    np.testing.assert_allclose(G, Ch.conj() @ Ch.T, rtol=0, atol=1e-10 * n)
    assert np.all(np.abs(np.diag(G).imag) <= 16 * n * np.finfo(float).eps * np.diag(G).real)
    np.testing.assert_allclose(np.diag(G).real, np.sum(np.abs(Ch) ** 2, axis=1), rtol=1e-13)


def test_zblas_gram_and_combine():
    """The solver's row-list wrappers (row lists split in 32-row calls, coefficient
    blocks of 24 outputs) against numpy."""
    import torch

    from igaops_amd.gcrodr import ZBlas

    rng = np.random.default_rng(11)
    n = 77777
    X = _rows(torch, rng, n, 41, 3)
    Y = _rows(torch, rng, n, 40, 2)
    zb = ZBlas(n)
    G = zb.gram(X, Y)
    Xh = np.stack([r.cpu().numpy() for r in X])
    Yh = np.stack([r.cpu().numpy() for r in Y])
    np.testing.assert_allclose(G, Xh.conj() @ Yh.T, rtol=0, atol=1e-10 * np.sqrt(n))
    coef = rng.standard_normal((41, 30)) + 1j * rng.standard_normal((41, 30))
    out = [torch.empty(n, dtype=torch.complex128, device="cuda") for _ in range(30)]
    for o in out:
        o.fill_(1.0 + 2.0j)
    zb.combine(X, coef, out, beta=0.5)
    torch.cuda.synchronize()
    ref = 0.5 * (1.0 + 2.0j) + coef.T @ Xh
    got = np.stack([o.cpu().numpy() for o in out])
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * 41 * np.abs(coef).max() * np.abs(Xh).max())
    assert abs(zb.norm(X[0]) - np.linalg.norm(Xh[0])) <= 1e-13 * np.linalg.norm(Xh[0])
