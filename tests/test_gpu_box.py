"""GPU free-box gather / scatter / mask (the Dirichlet box of the FD preconditioner and
of the solver mask) against numpy slicing, in 1-4 directions, with trailing
right-hand-side counts, empty and whole boxes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    ((7,), (2,), (3,)),
    ((68, 68, 68), (1, 1, 1), (66, 66, 66)),
    ((9, 5, 6), (0, 2, 1), (9, 3, 4)),
    ((4, 3, 5, 6), (1, 0, 2, 1), (2, 3, 3, 4)),
    ((5, 4), (1, 1), (0, 2)),
    ((6, 7), (0, 0), (6, 7)),
]


def _full(shape, r, dtype, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((int(np.prod(shape)), r)).astype(dtype)


def _as_nd(v, shape, r):
    # direction 1 fastest: numpy C order of the reversed shape
    return v.reshape(tuple(reversed(shape)) + (r,))


def _slices(lo, ln):
    return tuple(slice(a, a + b) for a, b in zip(reversed(lo), reversed(ln)))


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("r", [1, 3])
@pytest.mark.parametrize("shape,lo,ln", CASES)
def test_box_gather_scatter_mask(shape, lo, ln, r, dt):
    import torch

    from igaops_amd import _lib
    from igaops_amd import device as dv

    print("case", shape, lo, ln, r, dt, flush=True)
    npdt = np.float64 if dt == "f64" else np.float32
    code = _lib.IGA_F64 if dt == "f64" else _lib.IGA_F32
    nd = len(shape)
    (na, npt), (la, lpt), (lna, lnpt) = [_lib.i64_array(a) for a in (shape, lo, ln)]
    full = _full(shape, r, npdt, sum(shape) + r)
    nsub = int(np.prod(ln))
    expect_sub = _as_nd(full, shape, r)[_slices(lo, ln)].reshape(nsub, r)
    fd = torch.as_tensor(full).cuda()
    sub = torch.full((max(nsub, 1), r), np.nan, dtype=fd.dtype, device="cuda")
    _lib.call("iga_box_gather", code, nd, npt, lpt, lnpt, r, dv.ptr(fd), dv.ptr(sub), dv.stream_ptr())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(sub.cpu().numpy()[:nsub], expect_sub)

    src = _full(ln if nsub else (1,), r, npdt, 7)[: max(nsub, 1)]
    sd = torch.as_tensor(src).cuda()
    for zero_outside in (0, 1):
        out = fd.clone()
        _lib.call("iga_box_scatter", code, nd, npt, lpt, lnpt, r, dv.ptr(sd), dv.ptr(out), zero_outside,
                  dv.stream_ptr())
        torch.cuda.synchronize()
        ref = full.copy() if not zero_outside else np.zeros_like(full)
        refv = _as_nd(ref, shape, r)
        if nsub:
            refv[_slices(lo, ln)] = src[:nsub].reshape(refv[_slices(lo, ln)].shape)
        np.testing.assert_array_equal(out.cpu().numpy(), ref)

    out = fd.clone()
    _lib.call("iga_box_mask", code, nd, npt, lpt, lnpt, r, dv.ptr(out), dv.stream_ptr())
    torch.cuda.synchronize()
    ref = np.zeros_like(full)
    _as_nd(ref, shape, r)[_slices(lo, ln)] = _as_nd(full, shape, r)[_slices(lo, ln)]
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
