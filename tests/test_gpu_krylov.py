"""GPU many-dot / many-axpy reductions against float64 numpy sums.

Bound: |computed - exact| <= 4 n eps sum_i |V_ji w_i| (a summation of n
products in float64 in any order stays within n eps of that sum)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup():
    import torch

    from igaops_amd import _lib
    from igaops_amd import device as dv

    return torch, _lib, dv


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("n,k,pad,woff", [(0, 3, 2, 0), (1, 1, 0, 0), (3, 17, 0, 0), (1001, 5, 1, 1),
                                          (1001, 1, 0, 0), (70001, 1, 2, 0), (4099, 1, 0, 1),
                                          (4099, 16, 1, 0), (70001, 64, 3, 0), (262147, 33, 0, 0),
                                          (314432, 130, 0, 0)])
def test_dot_many_matches_float64_sums(dtype, n, k, pad, woff):
    """Vector path (16-byte loads, with a scalar tail) where ldv and the pointers allow
    it (any ldv for a single row), scalar path otherwise (odd ldv with several rows,
    w at an odd element offset)."""
    torch, _lib, dv = _setup()
    print("case", dtype, n, k, pad, woff, flush=True)
    tdt = torch.float64 if dtype == "f64" else torch.float32
    code = _lib.IGA_F64 if dtype == "f64" else _lib.IGA_F32
    g = torch.Generator().manual_seed(n + k)
    ldv = n + pad
    Vh = torch.randn(k, ldv, generator=g, dtype=torch.float64).to(tdt)
    wh = torch.randn(ldv + 1, generator=g, dtype=torch.float64).to(tdt)
    V, w = Vh.cuda(), wh.cuda()
    wp = w[woff:]
    wref = wh[woff:][:n].double().numpy()
    work = dv.workspace(_lib.load().iga_reduce_workspace_bytes(n, k))
    outs = []
    for rep in range(3):
        kk = k if rep != 1 else max(1, k // 2)
        out = torch.full((k,), np.nan, dtype=torch.float64, device="cuda")
        _lib.call("iga_dot_many", code, n, kk, dv.ptr(V), ldv, wp.data_ptr(), dv.ptr(out), dv.ptr(work),
                  dv.stream_ptr())
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy()[:kk])
    Vn = Vh[:, :n].double().numpy()
    exact = Vn @ wref
    bound = 4 * max(n, 1) * np.finfo(np.float64).eps * (np.abs(Vn) @ np.abs(wref)) + 1e-300
    for got in outs:
        assert np.all(np.abs(got - exact[: got.size]) <= bound[: got.size])
    # deterministic: the same call gives the same bits
    np.testing.assert_array_equal(outs[0], outs[2])


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_axpy_many_matches_float64(dtype):
    torch, _lib, dv = _setup()
    n, k = 50003, 37
    tdt = torch.float64 if dtype == "f64" else torch.float32
    code = _lib.IGA_F64 if dtype == "f64" else _lib.IGA_F32
    g = torch.Generator().manual_seed(7)
    Vh = torch.randn(k, n, generator=g, dtype=torch.float64).to(tdt)
    wh = torch.randn(n, generator=g, dtype=torch.float64).to(tdt)
    h = torch.randn(k, generator=g, dtype=torch.float64)
    w = wh.cuda()
    _lib.call("iga_axpy_many", code, n, k, dv.ptr(Vh.cuda()), n, dv.ptr(h.cuda()), -0.5, dv.ptr(w), dv.stream_ptr())
    torch.cuda.synchronize()
    exact = wh.double().numpy() - 0.5 * (h.numpy() @ Vh.double().numpy())
    got = w.cpu().double().numpy()
    eps = np.finfo(np.float64 if dtype == "f64" else np.float32).eps
    scale = np.abs(wh.double().numpy()) + 0.5 * (np.abs(h.numpy()) @ np.abs(Vh.double().numpy()))
    assert np.all(np.abs(got - exact) <= (k + 2) * eps * scale)
