"""iga_fused_apply_noalloc: the fused stiffness + mass apply with its partial blocks in
the library's built-in device scratch (no allocation), checked against
iga_fused_apply with a caller workspace on two calls in a row, the second with a
smaller input, and its size limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def _operator(geometry, p, nel, seed, dtype):
    from igaops_amd.geometry import lognormal_conductivity, quarter_annulus, unit_cube
    from igaops_amd.operators import StiffnessMassOperator

    patch = (quarter_annulus if geometry == "annulus" else unit_cube)(p, nel)
    C, m = patch.conductivity_coefficients(lognormal_conductivity(patch, 0.5, seed=seed))
    w = patch.nurbs_weights if patch.is_nurbs else None
    op = StiffnessMassOperator(patch.quadrule_list, C=C, m=m, nurbs_weights=w, dtype=dtype)
    assert op.fused
    return op


def _noalloc(op, x, y, nrhs, alpha=1.0, accumulate=False):
    from igaops_amd import _lib
    from igaops_amd import device as dv

    _lib.call("iga_fused_apply_noalloc", op.plan._h, dv.ptr(op.packed), dv.ptr(x), dv.ptr(y), int(nrhs),
              dv.ptr(op.w), float(alpha), int(bool(accumulate)), dv.stream_ptr(None))


def _with_workspace(op, x, y, nrhs, alpha=1.0, accumulate=False):
    from igaops_amd import _lib
    from igaops_amd import device as dv

    work = dv.workspace(op.plan.workspace_bytes(nrhs))
    _lib.call("iga_fused_apply", op.plan._h, dv.ptr(op.packed), dv.ptr(x), dv.ptr(y), int(nrhs), dv.ptr(op.w),
              float(alpha), int(bool(accumulate)), dv.ptr(work), dv.stream_ptr(None))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_noalloc_apply_matches_workspace_apply_on_two_calls(dtype):
    import torch

    tdt = torch.float64 if dtype == "f64" else torch.float32
    big = _operator("annulus", 4, 16, 1, dtype)
    small = _operator("cube", 2, 6, 2, dtype)
    assert small.plan.workspace_bytes(1) < big.plan.workspace_bytes(3) <= SCRATCH_BYTES
    g = torch.Generator(device="cuda").manual_seed(3)
    # first call: the annulus patch (NURBS), three right-hand sides
    xb = torch.randn(big.ndof * 3, generator=g, device="cuda", dtype=tdt)
    yb = torch.empty_like(xb)
    rb = torch.empty_like(xb)
    print("noalloc", dtype, "annulus p=4 nel=16 nrhs=3", flush=True)
    _noalloc(big, xb, yb, 3)
    _with_workspace(big, xb, rb, 3)
    torch.cuda.synchronize()
    assert torch.equal(yb, rb)
    # second call, smaller input: the cube patch (B-spline), one right-hand side,
    # accumulated with a scale into an existing output; the scratch still holds the
    # first call's larger partial blocks
    xs = torch.randn(small.ndof, generator=g, device="cuda", dtype=tdt)
    ys = torch.randn(small.ndof, generator=g, device="cuda", dtype=tdt)
    rs = ys.clone()
    print("noalloc", dtype, "cube p=2 nel=6 nrhs=1 accumulate", flush=True)
    _noalloc(small, xs, ys, 1, alpha=0.5, accumulate=True)
    _with_workspace(small, xs, rs, 1, alpha=0.5, accumulate=True)
    torch.cuda.synchronize()
    assert torch.equal(ys, rs)
    # and the first problem again, after the smaller call
    _noalloc(big, xb, yb, 3)
    torch.cuda.synchronize()
    assert torch.equal(yb, rb)


def test_noalloc_apply_rejects_sizes_above_its_scratch():
    import torch

    op = _operator("cube", 3, 8, 4, "f64")
    per_rhs = op.plan.workspace_bytes(1)
    nrhs = SCRATCH_BYTES // per_rhs + 1
    assert op.plan.workspace_bytes(nrhs) > SCRATCH_BYTES
    x = torch.zeros(op.ndof * nrhs, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    with pytest.raises(ValueError, match="built-in scratch"):
        _noalloc(op, x, y, nrhs)
