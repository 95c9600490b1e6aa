"""GPU parity of the operator path against the oracle (float64: <= 1e-12 relative;
float32: <= 5e-6 relative, measured against the float64 oracle)."""
import numpy as np
import pytest

from igaops_amd.geometry import lognormal_conductivity, quarter_annulus, unit_cube
from igaops_amd.quadrature import StandardGauss, WeightedQuadrature, make_open_knotvector
from oracle import matrix_free as omf
from oracle import operators as oo

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
F32_TOL = 5e-6


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def spd_coefficients(nq, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((3, 3, nq))
    C = np.einsum("ikq,jkq->ijq", A, A) + 0.5 * np.eye(3)[:, :, None]
    m = rng.random(nq) + 0.5
    return C, m


def device_op(patch, C, m, dtype="f64", force_composed=False, nslab=0):
    from igaops_amd.operators import StiffnessMassOperator

    w = patch.nurbs_weights if patch.is_nurbs else None
    return StiffnessMassOperator(patch.quadrule_list, C=C, m=m, nurbs_weights=w, dtype=dtype,
                                 force_composed=force_composed, nslab=nslab)


def run_op(op, x, nrhs=1, dtype="f64"):
    import torch

    xd = torch.as_tensor(x, device="cuda").to(torch.float64 if dtype == "f64" else torch.float32).reshape(-1)
    y = op.apply(xd, nrhs=nrhs)
    torch.cuda.synchronize()
    return y.double().cpu().numpy().reshape(x.shape)


CASES = [
    ("cube_p3_n6_wq2", lambda: unit_cube(3, 6)),
    ("cube_p2_n12_wq2", lambda: unit_cube(2, 12)),
    ("annulus_p4_n5_wq2", lambda: quarter_annulus(4, 5)),
    ("annulus_p3_n7_gauss", lambda: quarter_annulus(3, 7, quadclass="gauss", quadtype="legendre")),
    ("cube_p6_n9_wq2", lambda: unit_cube(6, 9)),
    ("annulus_p7_n3_wq2", lambda: quarter_annulus(7, 3)),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("path", ["fused", "composed"])
def test_stiffness_plus_mass_f64(name, make, path):
    print(f"case {name} {path}", flush=True)
    patch = make()
    C, m = spd_coefficients(patch.nbquadpts_total, 1)
    x = np.random.default_rng(2).standard_normal(patch.nbctrlpts_total)
    w = patch.nurbs_weights if patch.is_nurbs else None
    ref = oo.stiffness_plus_mass(patch.quadrule_list, C, m, x, nurbs_weights=w)
    op = device_op(patch, C, m, force_composed=(path == "composed"))
    assert op.fused == (path == "fused")
    assert rel(run_op(op, x), ref) <= F64_TOL


@pytest.mark.parametrize("nslab", [1, 2, 5])
def test_fused_slabs_and_multi_rhs(nslab):
    patch = quarter_annulus(3, 6)
    C, m = spd_coefficients(patch.nbquadpts_total, 3)
    X = np.random.default_rng(4).standard_normal((patch.nbctrlpts_total, 3))
    ref = oo.stiffness_plus_mass(patch.quadrule_list, C, m, X, nurbs_weights=patch.nurbs_weights)
    op = device_op(patch, C, m, nslab=nslab)
    assert rel(run_op(op, X, nrhs=3), ref) <= F64_TOL


@pytest.mark.parametrize("name,make", CASES[:4], ids=[c[0] for c in CASES[:4]])
def test_stiffness_plus_mass_f32(name, make):
    patch = make()
    C, m = spd_coefficients(patch.nbquadpts_total, 5)
    x = np.random.default_rng(6).standard_normal(patch.nbctrlpts_total)
    w = patch.nurbs_weights if patch.is_nurbs else None
    ref = oo.stiffness_plus_mass(patch.quadrule_list, C, m, x, nurbs_weights=w)
    op = device_op(patch, C, m, dtype="f32")
    assert rel(run_op(op, x, dtype="f32"), ref) <= F32_TOL


@pytest.mark.parametrize("kind", ["u_v", "gradu_gradv", "gradu_v", "u_gradv"])
@pytest.mark.parametrize("geom", ["cube", "annulus"])
def test_reference_signature_operators(kind, geom):
    from igaops_amd.operators import BsplineOperations, NurbsOperations

    patch = unit_cube(3, 4) if geom == "cube" else quarter_annulus(3, 4)
    qs = patch.quadrule_list
    nq = patch.nbquadpts_total
    rng = np.random.default_rng(8)
    x = rng.standard_normal(patch.nbctrlpts_total)
    coef = {"u_v": rng.random(nq), "gradu_gradv": rng.standard_normal((3, 3, nq)),
            "gradu_v": rng.standard_normal((3, nq)), "u_gradv": rng.standard_normal((3, nq))}[kind]
    eng = BsplineOperations if geom == "cube" else NurbsOperations
    kw = {} if geom == "cube" else {"nurbs_weights": patch.nurbs_weights}
    got = getattr(eng, f"compute_mf_scalar_{kind}")(qs, coef, x, **kw)
    fn = getattr(oo, ("bspline_" if geom == "cube" else "nurbs_") + kind)
    ref = fn(qs, coef, x, **kw)
    assert isinstance(got, np.ndarray) and got.shape == x.shape
    assert rel(got, ref) <= F64_TOL


def test_wq_type1_and_spacetime_composed():
    """WQ type "1" (test weights depend on the trial derivative) and a 4-D space-time
    operator go through the composed path."""
    from igaops_amd.operators import BsplineOperations

    q1 = WeightedQuadrature(3, make_open_knotvector(3, 4), "1")
    qs = [q1] * 3
    nq = q1.nbquadpts ** 3
    rng = np.random.default_rng(9)
    C = rng.standard_normal((3, 3, nq))
    x = rng.standard_normal(q1.nbctrlpts ** 3)
    assert rel(BsplineOperations.compute_mf_scalar_gradu_gradv(qs, C, x), oo.bspline_gradu_gradv(qs, C, x)) <= F64_TOL
    qsp = WeightedQuadrature(2, make_open_knotvector(2, 4), "2")
    qt = WeightedQuadrature(2, make_open_knotvector(2, 5), "2")
    qst = [qsp, qsp, qsp, qt]
    nq = qsp.nbquadpts ** 3 * qt.nbquadpts
    x = rng.standard_normal(qsp.nbctrlpts ** 3 * qt.nbctrlpts)
    c = rng.random(nq)
    Cs = rng.standard_normal((3, 3, nq))
    got = BsplineOperations.compute_mf_scalar_u_v(qst, c, x, time_ders=(0, 1)) + \
        BsplineOperations.compute_mf_scalar_gradu_gradv(qst, Cs, x, time_ders=(0, 0))
    ref = oo.bspline_u_v(qst, c, x, time_ders=(0, 1)) + oo.bspline_gradu_gradv(qst, Cs, x, time_ders=(0, 0))
    assert rel(got, ref) <= F64_TOL


@pytest.mark.parametrize("transpose", [False, True])
def test_matrixfree_apply(transpose):
    from igaops_amd.operators import MatrixFree

    rng = np.random.default_rng(10)
    q = WeightedQuadrature(3, make_open_knotvector(3, 5), "2")
    mats = [q.basis[0], q.basis[1], q.weights[2].T.tocsr()]
    n_in = int(np.prod([m.shape[0 if transpose else 1] for m in mats]))
    X = rng.standard_normal((n_in, 2))
    got = MatrixFree.apply(mats, X, is_transpose=transpose)
    ref = omf.apply(mats, X, is_transpose=transpose)
    assert rel(got, ref) <= F64_TOL
    dense = [rng.standard_normal((4, 3)), rng.standard_normal((2, 5))]
    v = rng.standard_normal(15)
    assert rel(MatrixFree.apply(dense, v), omf.apply(dense, v)) <= F64_TOL


def test_lognormal_annulus_config_small():
    """The benchmark's coefficient field on a small curved patch."""
    patch = quarter_annulus(4, 8)
    kappa = lognormal_conductivity(patch, 0.5, seed=0)
    C, m = patch.conductivity_coefficients(kappa)
    x = np.random.default_rng(11).standard_normal(patch.nbctrlpts_total)
    ref = oo.stiffness_plus_mass(patch.quadrule_list, C, m, x, nurbs_weights=patch.nurbs_weights)
    assert rel(run_op(device_op(patch, C, m), x), ref) <= F64_TOL
    assert rel(run_op(device_op(patch, C, m, dtype="f32"), x, dtype="f32"), ref) <= F32_TOL


def test_zero_and_repeat_calls_are_deterministic():
    import torch

    patch = unit_cube(3, 6)
    C, m = spd_coefficients(patch.nbquadpts_total, 12)
    op = device_op(patch, C, m)
    z = torch.zeros(patch.nbctrlpts_total, dtype=torch.float64, device="cuda")
    assert float(op.apply(z).abs().max()) == 0.0
    x = torch.randn(patch.nbctrlpts_total, dtype=torch.float64, device="cuda")
    y1 = op.apply(x).clone()
    y2 = op.apply(x).clone()
    assert torch.equal(y1, y2)


def test_unit_cube_p3_n32_full_size():
    """The CPU-runnable configuration at its full size: unit cube, degree 3, 32
    elements per direction (42,875 DOF, 101^3 WQ points), unit conductivity: the
    float64 K+M matvec against the oracle, and FD-preconditioned GMRES on K x = b to
    1e-10 against the oracle's GMRES (the FD is the exact inverse of this separable
    operator, so both stop within two iterations)."""
    import torch

    from igaops_amd.fastdiag import SingleFD
    from igaops_amd.geometry import unit_cube
    from igaops_amd.operators import StiffnessMassOperator
    from igaops_amd.solvers import LinearSolver
    from oracle import fastdiag as ofd
    from oracle import solvers as osv
    from igaops_amd.quadrature import StandardGauss

    patch = unit_cube(3, 32)
    qs = patch.quadrule_list
    n = patch.nbctrlpts_total
    assert n == 42875 and patch.nbquadpts_total == 1030301
    C, m = patch.conductivity_coefficients()
    x = np.random.default_rng(0).standard_normal(n)
    op = StiffnessMassOperator(qs, C=C, m=m)
    assert op.fused
    y = op.apply(torch.as_tensor(x, device="cuda")).cpu().numpy()
    assert rel(y, oo.stiffness_plus_mass(qs, C, m, x)) <= 1e-12
    K = StiffnessMassOperator(qs, C=C)
    fd = SingleFD("spatial")
    fd.compute_space_eigendecomposition(qs, np.ones((1, 3, 2), bool))
    fd.update_space_eigenvalues([0.0, 1.0])
    free = np.asarray(fd.space_preconditioner.space_free_nodes[0])
    dod = np.setdiff1d(np.arange(n), free).tolist()
    b = np.random.default_rng(1).standard_normal(n)
    out = LinearSolver(1e-10, 50, cleandod=dod).solve(K, b, Pfun=fd)
    gq = [StandardGauss(q.degree, q.knotvector, "legendre") for q in qs]
    ref = osv.gmres(lambda v: oo.bspline_gradu_gradv(qs, C, v), b, ofd.SpaceFD(gq, np.ones((1, 3, 2), bool)).apply,
                    tol=1e-10, maxiters=50, cleandod=dod)
    assert out.success and ref.success and out.iterations <= 2 and abs(out.iterations - ref.iterations) <= 1
    assert rel(out.solution, ref.solution) <= 1e-9


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("geom", ["cube", "annulus"])
def test_grouped_apply_equals_each_members_apply(dtype, geom):
    """Three operators of one shape with their own lognormal fields in one grouped
    launch (iga_fused_apply_group): every column is that operator's own apply bit for
    bit, with and without the Dirichlet box, and the float64 columns match the oracle."""
    import torch

    from igaops_amd.operators import StiffnessMassOperator, apply_group

    print("case", geom, dtype, flush=True)
    patch = unit_cube(3, 6) if geom == "cube" else quarter_annulus(3, 6)
    qs = patch.quadrule_list
    w = patch.nurbs_weights if patch.is_nurbs else None
    coefs = [patch.conductivity_coefficients(lognormal_conductivity(patch, 0.3 + 0.2 * k, seed=k)) for k in range(3)]
    ops = [StiffnessMassOperator(qs, C=C, m=m, nurbs_weights=w, dtype=dtype) for C, m in coefs]
    assert all(op.fused and op.plan is ops[0].plan for op in ops)
    n = patch.nbctrlpts_total
    tdt = torch.float64 if dtype == "f64" else torch.float32
    X = torch.randn(n, 3, dtype=torch.float64, device="cuda").to(tdt)
    Y = [torch.empty(n, dtype=tdt, device="cuda") for _ in range(3)]
    apply_group(ops, X, Y)
    alone = [op.apply(X[:, k].contiguous()) for k, op in enumerate(ops)]
    for k in range(3):
        assert torch.equal(Y[k], alone[k])
    # member-major input
    W = [torch.empty(n, dtype=tdt, device="cuda") for _ in range(3)]
    apply_group(ops, X.t().contiguous(), W, member_major=True)
    for k in range(3):
        assert torch.equal(W[k], alone[k])
    # one operator for several right-hand sides: the multi-right-hand-side apply
    Z = [torch.empty(n, dtype=tdt, device="cuda") for _ in range(3)]
    apply_group([ops[1]] * 3, X, Z)
    multi = ops[1].apply(X.reshape(-1).contiguous(), nrhs=3).reshape(n, 3)
    for k in range(3):
        assert torch.equal(Z[k], multi[:, k])
    nc = [q.nbctrlpts for q in qs]
    lo, ln = np.array([1, 1, 1]), np.array([c - 2 for c in nc])
    apply_group(ops, X, Y, box=(lo, ln))
    inside = torch.zeros(nc[2], nc[1], nc[0], dtype=torch.bool, device="cuda")
    inside[1:-1, 1:-1, 1:-1] = True
    inside = inside.reshape(-1)
    for k in range(3):
        assert torch.equal(Y[k], torch.where(inside, alone[k], torch.zeros_like(alone[k])))
    Xh = X.double().cpu().numpy()
    for k in (0, 2):
        C, m = coefs[k]
        ref = oo.stiffness_plus_mass(qs, C, m, Xh[:, k], nurbs_weights=w)
        assert rel(alone[k].double().cpu().numpy(), ref) <= (F64_TOL if dtype == "f64" else F32_TOL)


@pytest.mark.parametrize("kind", ["gradu_v", "u_gradv"])
@pytest.mark.parametrize("geom", ["cube", "annulus"])
def test_first_order_terms_in_one_sweep(kind, geom):
    """The first-order terms through the fused sweep (four packed fields per point,
    the NURBS correction folded in): float64 and float32, two right-hand sides, against
    the oracle's composed restatement; the reference-signature call takes this path."""
    import torch

    from igaops_amd.operators import FirstOrderOperator

    print("case", kind, geom, flush=True)
    patch = unit_cube(3, 7) if geom == "cube" else quarter_annulus(4, 6)
    qs = patch.quadrule_list
    w = patch.nurbs_weights if patch.is_nurbs else None
    rng = np.random.default_rng(31)
    c = rng.standard_normal((3, patch.nbquadpts_total))
    n = patch.nbctrlpts_total
    X = rng.standard_normal((n, 2))
    fn = getattr(oo, ("bspline_" if geom == "cube" else "nurbs_") + kind)
    kw = {} if w is None else {"nurbs_weights": w}
    ref = np.stack([fn(qs, c, X[:, k], **kw) for k in range(2)], axis=1)
    for dtype, tol in (("f64", F64_TOL), ("f32", F32_TOL)):
        op = FirstOrderOperator(qs, c, kind, nurbs_weights=w, dtype=dtype)
        xd = torch.as_tensor(X, device="cuda").to(torch.float64 if dtype == "f64" else torch.float32)
        y = op.apply(xd.reshape(-1).contiguous(), nrhs=2).reshape(n, 2).double().cpu().numpy()
        assert rel(y, ref) <= tol, (dtype, rel(y, ref))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_fewest_slab_plan_of_several_right_hand_sides(dtype):
    """Several right-hand sides run on the operator's fewest-slab plan (the same packed
    coefficients): the same values as the single-right-hand-side plan to rounding (the
    partial blocks are summed over other slabs), each column against the oracle; the
    batch view applies one right-hand side on that plan too."""
    import torch

    from igaops_amd.operators import BATCH_PLAN_MIN_RHS, StiffnessMassOperator

    patch = quarter_annulus(4, 12)
    qs = patch.quadrule_list
    w = patch.nurbs_weights
    C, m = patch.conductivity_coefficients(lognormal_conductivity(patch, 0.5, seed=6))
    op = StiffnessMassOperator(qs, C=C, m=m, nurbs_weights=w, dtype=dtype)
    bp = op.batch_plan()
    print("case", dtype, "slabs", op.plan.nslab, bp.nslab, flush=True)
    assert bp.nslab <= op.plan.nslab and bp.nslab == -(-int(patch.nbquadpts[2]) // 128)
    n = patch.nbctrlpts_total
    r = BATCH_PLAN_MIN_RHS + 1
    tdt = torch.float64 if dtype == "f64" else torch.float32
    X = torch.randn(n, r, dtype=torch.float64, device="cuda").to(tdt)
    xf = X.reshape(-1).contiguous()
    y_batch = op.apply(xf, nrhs=r).reshape(n, r)
    y_plan = op.apply(xf, nrhs=r, work=op.workspace(r)).reshape(n, r)
    d = float((y_batch - y_plan).abs().max() / y_plan.abs().max())
    assert d <= (1e-14 if dtype == "f64" else 1e-6), d
    Xh = X.double().cpu().numpy()
    tol = F64_TOL if dtype == "f64" else F32_TOL
    for k in (0, r - 1):
        ref = oo.stiffness_plus_mass(qs, C, m, Xh[:, k], nurbs_weights=w)
        assert rel(y_batch[:, k].double().cpu().numpy(), ref) <= tol
    v = op.batch_view()
    assert v.plan is bp and v.packed is op.packed
    y1 = v.apply(X[:, 0].contiguous())
    assert float((y1 - y_plan[:, 0]).abs().max() / y_plan[:, 0].abs().max()) <= (1e-14 if dtype == "f64" else 1e-6)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("geom,mass", [("annulus", True), ("cube", True), ("cube", False)])
def test_two_right_hand_sides_per_workgroup_same_bits(dtype, geom, mass):
    """Right-hand sides sharing one coefficient set swept two per work-group (the
    coefficient loads and plane phases shared) agree with one per work-group to
    rounding (the compiler may contract the per-point sums differently), for the NURBS
    (10 fields), B-spline K+M (7) and K (6) forms, and match the oracle."""
    import torch

    from igaops_amd.operators import StiffnessMassOperator

    print("case", geom, mass, dtype, flush=True)
    patch = unit_cube(3, 7) if geom == "cube" else quarter_annulus(4, 6)
    qs = patch.quadrule_list
    w = patch.nurbs_weights if patch.is_nurbs else None
    C, m = patch.conductivity_coefficients(lognormal_conductivity(patch, 0.5, seed=4))
    op = StiffnessMassOperator(qs, C=C, m=m if mass else None, nurbs_weights=w, dtype=dtype)
    n = patch.nbctrlpts_total
    tdt = torch.float64 if dtype == "f64" else torch.float32
    X = torch.randn(n, 6, dtype=torch.float64, device="cuda").to(tdt)
    outs = {}
    for nr in (1, 2):
        op.set_rhs_block(nr)
        outs[nr] = op.apply(X.reshape(-1).contiguous(), nrhs=6).reshape(n, 6)
    op.set_rhs_block(1)
    d = float((outs[1] - outs[2]).abs().max() / outs[1].abs().max())
    print("two-per-work-group difference", d, flush=True)
    assert d <= (1e-14 if dtype == "f64" else 1e-6)
    Xh = X.double().cpu().numpy()
    for k in (0, 5):
        ref = (oo.stiffness_plus_mass(qs, C, m, Xh[:, k], nurbs_weights=w) if mass
               else (oo.nurbs_gradu_gradv(qs, C, Xh[:, k], nurbs_weights=w) if w is not None
                     else oo.bspline_gradu_gradv(qs, C, Xh[:, k])))
        assert rel(outs[2][:, k].double().cpu().numpy(), ref) <= (F64_TOL if dtype == "f64" else F32_TOL)
