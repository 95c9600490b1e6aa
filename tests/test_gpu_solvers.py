"""GPU parity of the FD apply and the Krylov drivers against the oracle."""
import numpy as np
import pytest

from igaops_amd.geometry import lognormal_conductivity, quarter_annulus, unit_cube
from igaops_amd.quadrature import StandardGauss, make_open_knotvector
from oracle import fastdiag as ofd
from oracle import operators as oo
from oracle import solvers as osv

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("table", ["all", "mixed", "none"])
@pytest.mark.parametrize("p,nel", [(3, 6), (2, 9)])
def test_fd_apply_matches_oracle(table, p, nel):
    from igaops_amd.fastdiag import SingleFD

    patch = unit_cube(p, nel)
    qs = patch.quadrule_list
    tab = {"all": np.ones((1, 3, 2), bool), "none": np.zeros((1, 3, 2), bool)}.get(table)
    if tab is None:
        tab = np.zeros((1, 3, 2), bool)
        tab[0, 0, 0] = True
        tab[0, 2, 1] = True
    fd = SingleFD("spatial")
    fd.compute_space_eigendecomposition(qs, tab)
    fd.update_space_eigenvalues([0.3, 1.0])
    gq = [StandardGauss(q.degree, q.knotvector, "legendre") for q in qs]
    ofd_ = ofd.SpaceFD(gq, tab)
    ofd_.scalar_coefs = (0.3, 1.0)
    X = np.random.default_rng(1).standard_normal((patch.nbctrlpts_total, 2))
    np.testing.assert_array_equal(np.asarray(fd.space_preconditioner.space_free_nodes[0]), ofd_.free_nodes())
    assert rel(fd.apply(X), ofd_.apply(X)) <= 1e-12
    assert rel(fd.apply(X[:, 0]), ofd_.apply(X[:, 0])) <= 1e-12
    import torch

    x32 = torch.as_tensor(X[:, 0], device="cuda").float()
    y32 = fd.apply(x32).double().cpu().numpy()
    assert rel(y32, ofd_.apply(X[:, 0])) <= 1e-4  # float32 GEMMs on a conditioned inverse


@pytest.mark.parametrize("dt,nels", [("f64", (42, 72, 97)), ("f64", (122, 112, 66)), ("f32", (102, 62, 122)),
                                     ("f32", (132, 52, 142)), ("bf16", (72, 97, 122))])
def test_fd_apply_every_staging_size(dt, nels):
    """Three-launch FD applies on boxes of unequal sides (degree 2, every face Dirichlet,
    box side nel), so that the line kernel runs every operand-staging size (flat with
    16 / 24 / 40 / 64 registers per thread for sides up to 64 / 78 / 101 / 128, row
    rounds past 128) and the plane kernels one to three tiles per wave; two right-hand
    sides, against the oracle's apply."""
    import torch

    from igaops_amd.fastdiag import SingleFD
    from igaops_amd.quadrature import WeightedQuadrature

    qs = [WeightedQuadrature(2, make_open_knotvector(2, n), "2") for n in nels]
    tab = np.ones((1, 3, 2), bool)
    fd = SingleFD("spatial")
    fd.compute_space_eigendecomposition(qs, tab)
    fd.update_space_eigenvalues([0.5, 1.0])
    sp = fd.space_preconditioner
    o = ofd.SpaceFD([StandardGauss(q.degree, q.knotvector, "legendre") for q in qs], tab)
    o.scalar_coefs = (0.5, 1.0)
    n = int(np.prod([q.nbctrlpts for q in qs]))
    X = np.random.default_rng(21).standard_normal((n, 2))
    ref = o.apply(X)
    Xd = torch.as_tensor(X, device="cuda")
    if dt != "f64":
        sp.float32_plan = dt
        Xd = Xd.float()
    yd = sp.apply_device(Xd.reshape(-1).contiguous(), nrhs=2)
    # the right-hand-side-major entry gives the same bits per right-hand side
    plan = sp.plan(dt)
    ym = torch.empty(2, n, dtype=Xd.dtype, device="cuda")
    plan.apply(Xd.t().contiguous(), ym, 2, rhs_major=True)
    assert torch.equal(ym.t(), yd.reshape(n, 2))
    y = yd.reshape(n, 2).double().cpu().numpy()
    sp.float32_plan = "f32"
    err = rel(y, ref)
    print(f"FD {dt} box {nels}: relative error {err:.2e}", flush=True)
    assert err <= {"f64": 1e-12, "f32": 1e-5, "bf16": 1e-2}[dt]
    free = np.zeros(n, bool)
    free[o.free_nodes()] = True
    assert np.all(y[~free] == 0)


def build_problem(geom, p, nel, sigma=0.5):
    patch = unit_cube(p, nel) if geom == "cube" else quarter_annulus(p, nel)
    kappa = lognormal_conductivity(patch, sigma, seed=1) if sigma > 0 else None
    C, m = patch.conductivity_coefficients(kappa)
    return patch, C, m


def test_gmres_device_matches_oracle_gmres():
    """Same preconditioned problem, same stopping rule: iteration counts within one,
    solutions within the attainable accuracy."""
    import torch

    from igaops_amd.fastdiag import SingleFD
    from igaops_amd.operators import StiffnessMassOperator
    from igaops_amd.solvers import LinearSolver

    patch, C, m = build_problem("annulus", 3, 6)
    qs = patch.quadrule_list
    w = patch.nurbs_weights
    n = patch.nbctrlpts_total
    fd = SingleFD("spatial")
    fd.compute_space_eigendecomposition(qs, np.ones((1, 3, 2), bool))
    fd.update_space_eigenvalues([0.0, 1.0])
    free = np.asarray(fd.space_preconditioner.space_free_nodes[0])
    dod = np.setdiff1d(np.arange(n), free).tolist()
    op = StiffnessMassOperator(qs, C=C, m=None, nurbs_weights=w)
    b = np.random.default_rng(3).standard_normal(n)
    solver = LinearSolver(1e-10, 200, cleandod=dod)
    out = solver.solve(op, b, Pfun=fd)
    gq = [StandardGauss(q.degree, q.knotvector, "legendre") for q in qs]
    ofd_ = ofd.SpaceFD(gq, np.ones((1, 3, 2), bool))
    ref = osv.gmres(lambda v: oo.nurbs_gradu_gradv(qs, C, v, nurbs_weights=w), b, ofd_.apply, tol=1e-10,
                    maxiters=200, cleandod=dod)
    assert out.success and ref.success
    assert abs(out.iterations - ref.iterations) <= 1
    assert rel(out.solution, ref.solution) <= 1e-8
    # true residual of the device solution, masked as the solver sees it
    Ax = oo.nurbs_gradu_gradv(qs, C, out.solution, nurbs_weights=w)
    r = b.copy()
    r[dod] = 0
    Ax[dod] = 0
    assert np.linalg.norm(r - Ax) / np.linalg.norm(r) <= 1e-9


def test_mixed_precision_refinement_reaches_1e10():
    import torch

    from igaops_amd.fastdiag import SingleFD
    from igaops_amd.operators import StiffnessMassOperator
    from igaops_amd.solvers import Mask, PrecisionAdapter, refinement_solve

    patch, C, m = build_problem("annulus", 4, 8)
    qs = patch.quadrule_list
    w = patch.nurbs_weights
    n = patch.nbctrlpts_total
    fd = SingleFD("spatial")
    fd.compute_space_eigendecomposition(qs, np.ones((1, 3, 2), bool))
    sp = fd.space_preconditioner
    A64 = StiffnessMassOperator(qs, C=C, m=m, nurbs_weights=w, dtype="f64")
    A32 = StiffnessMassOperator(qs, C=C, m=m, nurbs_weights=w, dtype="f32")
    fd.update_space_eigenvalues([1.0, 1.0])
    mask = Mask(box=sp.box())
    b = torch.as_tensor(np.random.default_rng(4).standard_normal(n), device="cuda")
    mask(b)
    P32 = PrecisionAdapter(fd, n, "f32")
    A32a = PrecisionAdapter(A32, n, "f32")
    out = refinement_solve(lambda x, o: A64.apply(x, out=o), lambda x, o: A32a.apply_device(x, out=o),
                           lambda x, o: P32.apply_device(x, out=o), b, tol=1e-10, inner_tol=1e-5, mask=mask)
    assert out.success
    x = out.solution.cpu().numpy()
    bh = b.cpu().numpy()
    Ax = oo.stiffness_plus_mass(qs, C, m, x, nurbs_weights=w)
    free = np.zeros(n, bool)
    free[np.asarray(sp.space_free_nodes[0])] = True
    Ax[~free] = 0
    assert np.linalg.norm(bh - Ax) / np.linalg.norm(bh) <= 1e-10


def _annulus_fd_problem(p=3, nel=6, mass=True):
    from igaops_amd.fastdiag import SingleFD
    from igaops_amd.operators import StiffnessMassOperator
    from igaops_amd.solvers import Mask

    patch, C, m = build_problem("annulus", p, nel)
    qs = patch.quadrule_list
    w = patch.nurbs_weights
    fd = SingleFD("spatial")
    fd.compute_space_eigendecomposition(qs, np.ones((1, 3, 2), bool))
    fd.update_space_eigenvalues([1.0 if mass else 0.0, 1.0])
    sp = fd.space_preconditioner
    ops = {dt: StiffnessMassOperator(qs, C=C, m=m if mass else None, nurbs_weights=w, dtype=dt)
           for dt in ("f64", "f32")}
    return patch, C, m if mass else None, fd, Mask(box=sp.box()), ops


def test_device_step_gmres_matches_host_step_gmres():
    """The device-side Arnoldi/Givens steps (runs.GmresRun) against the host-side
    Givens loop on the same problem: same iteration count, same residual history,
    solutions within the attainable accuracy; both with and without lookahead."""
    import torch

    from igaops_amd.runs import GmresRun
    from igaops_amd.solvers import gmres_device

    patch, C, m, fd, mask, ops = _annulus_fd_problem()
    n = patch.nbctrlpts_total
    A = lambda x, o: ops["f64"].apply(x, out=o)
    P = lambda x, o: fd.apply_device(x, out=o)
    b = torch.as_tensor(np.random.default_rng(11).standard_normal(n), device="cuda")
    mask(b)
    ref = gmres_device(A, b, P, 1e-10, 200, mask=mask, restart=400)  # host Givens path (one cycle)
    for la in (1, 3):
        out = GmresRun(A, P, b, 1e-10, 200, mask=mask, lookahead=la).run()
        assert out.success and ref.success
        assert out.iterations == ref.iterations
        np.testing.assert_allclose(out.residual[:len(ref.residual)], ref.residual[:len(out.residual)], rtol=1e-6,
                                   atol=1e-14)
        assert rel(out.solution.cpu().numpy(), ref.solution.cpu().numpy()) <= 1e-9
    # true residual
    r = b - ops["f64"].apply(out.solution)
    mask(r)
    assert float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b)) <= 1e-9


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_device_step_gmres_at_maxiters_and_orthogonality(dtype):
    """Stopped at the Krylov dimension (the last step finishes the last basis
    vector): the same minimiser as the host-step solve over the same space, and a
    basis orthonormal to working precision (delayed reorthogonalisation)."""
    import torch

    from igaops_amd.runs import GmresRun
    from igaops_amd.solvers import gmres_device

    patch, C, m, fd, mask, ops = _annulus_fd_problem()
    n = patch.nbctrlpts_total
    tdt = torch.float64 if dtype == "f64" else torch.float32
    A = lambda x, o: ops[dtype].apply(x, out=o)
    P = lambda x, o: fd.apply_device(x, out=o)
    b = torch.as_tensor(np.random.default_rng(12).standard_normal(n), device="cuda", dtype=tdt)
    mask(b)
    kmax = 9
    run = GmresRun(A, P, b, 1e-14, kmax, mask=mask)
    out = run.run()
    assert out.iterations == kmax and out.reason == "maxiters"
    Q = run.V[:kmax, : run.n].double()  # rows are padded to 16-byte multiples
    orth = float((Q @ Q.T - torch.eye(kmax, dtype=torch.float64, device="cuda")).abs().max())
    assert orth <= (1e-13 if dtype == "f64" else 1e-5)
    if dtype == "f64":
        ref = gmres_device(A, b, P, 1e-14, kmax, mask=mask, restart=400)
        assert ref.iterations == kmax
        np.testing.assert_allclose(out.residual, ref.residual, rtol=1e-8)
        assert rel(out.solution.cpu().numpy(), ref.solution.cpu().numpy()) <= 1e-10


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("precond", [True, False])
def test_one_call_arnoldi_step_matches_python_driven_steps(dtype, precond):
    """``iga_arnoldi_step`` queues the kernels of the Python-driven step with the same
    arguments (the FD leaves zeros outside its box, so the mask of ``z`` it skips
    changes nothing): the same residual history, iteration count and solution bits,
    with lookahead 1 and 3, with and without the FD."""
    import torch

    from igaops_amd.runs import GmresRun, NativeStep

    patch, C, m, fd, mask, ops = _annulus_fd_problem()
    n = patch.nbctrlpts_total
    tdt = torch.float64 if dtype == "f64" else torch.float32
    plan = fd.space_preconditioner.plan(dtype)
    work = plan.workspace(1)
    A = lambda x, o: ops[dtype].apply(x, out=o)
    P = (lambda x, o: plan.apply(x, o, 1, None, work=work)) if precond else None
    native = NativeStep(ops[dtype], plan if precond else None, work if precond else None, mask.box)
    b = torch.as_tensor(np.random.default_rng(13).standard_normal(n), device="cuda", dtype=tdt)
    mask(b)
    tol = 1e-10 if dtype == "f64" else 1e-6
    for la in (1, 3):
        ref = GmresRun(A, P, b, tol, 150, mask=mask, lookahead=la).run()
        out = GmresRun(A, P, b, tol, 150, mask=mask, lookahead=la, native=native).run()
        assert out.iterations == ref.iterations and out.reason == ref.reason
        assert out.iterations >= (10 if precond else 40)
        np.testing.assert_array_equal(out.residual, ref.residual)
        assert torch.equal(out.solution, ref.solution)


def test_one_call_arnoldi_step_inside_refinement():
    """Mixed-precision refinement with its inner solves stepped by ``iga_arnoldi_step``:
    the same refinements, inner iterations and solution bits as with Python-driven
    inner steps, and a true residual of 1e-10."""
    import torch

    from igaops_amd.runs import NativeStep, RefinementRun

    patch, C, m, fd, mask, ops = _annulus_fd_problem(4, 8)
    n = patch.nbctrlpts_total
    plan = fd.space_preconditioner.plan("f32")
    work = plan.workspace(1)
    A64 = lambda x, o: ops["f64"].apply(x, out=o)
    A32 = lambda x, o: ops["f32"].apply(x, out=o)
    P = lambda x, o: plan.apply(x, o, 1, None, work=work)
    b = torch.as_tensor(np.random.default_rng(14).standard_normal(n), device="cuda")
    mask(b)
    ref = RefinementRun(A64, A32, P, b, tol=1e-10, inner_tol=1e-5, mask=mask).run()
    out = RefinementRun(A64, A32, P, b, tol=1e-10, inner_tol=1e-5, mask=mask,
                        native_lo=NativeStep(ops["f32"], plan, work, mask.box)).run()
    assert out.success and ref.success
    assert out.history["inner_iterations"] == ref.history["inner_iterations"]
    assert len(out.history["inner_iterations"]) >= 2
    assert torch.equal(out.solution, ref.solution)
    r = b - ops["f64"].apply(out.solution)
    mask(r)
    assert float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b)) <= 1e-10


@pytest.mark.parametrize("lockstep", [True, False])
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_batched_steps_of_several_right_hand_sides_match_single_runs(mode, lockstep):
    """Five right-hand sides of one operator driven together (``run_batched``: the FD
    of the ready runs in one multi-right-hand-side launch, their operators in one
    grouped launch, then each run's orthogonalisation; in lock step or as they become
    ready): the same iteration counts and solution bits as each solve driven alone, and
    a true residual of 1e-10."""
    import torch

    from igaops_amd.runs import GmresRun, NativeStep, RefinementRun, run_batched

    print("case", mode, lockstep, flush=True)
    patch, C, m, fd, mask, ops = _annulus_fd_problem(4, 8)
    n = patch.nbctrlpts_total
    dt = "f64" if mode == "fp64" else "f32"
    plan = fd.space_preconditioner.plan(dt)
    work = plan.workspace(1)
    A64 = lambda x, o: ops["f64"].apply(x, out=o)
    A32 = lambda x, o: ops["f32"].apply(x, out=o)
    P = lambda x, o: plan.apply(x, o, 1, None, work=work)
    rng = np.random.default_rng(15)
    B = [torch.as_tensor(rng.standard_normal(n), device="cuda") for _ in range(5)]
    for b in B:
        mask(b)

    def make(b):
        if mode == "fp64":
            return GmresRun(A64, P, b, 1e-10, 200, mask=mask, native=NativeStep(ops["f64"], plan, work, mask.box))
        return RefinementRun(A64, A32, P, b, tol=1e-10, inner_tol=1e-5, mask=mask,
                             native_lo=NativeStep(ops["f32"], plan, work, mask.box))

    alone = [make(b).run() for b in B]
    together, done_at = run_batched([make(b) for b in B], lockstep=lockstep)
    assert all(d is not None for d in done_at)
    for b, a, t in zip(B, alone, together):
        assert a.success and t.success
        assert t.iterations == a.iterations
        assert torch.equal(t.solution, a.solution)
        r = b - ops["f64"].apply(t.solution)
        mask(r)
        assert float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b)) <= 1e-10


@pytest.mark.parametrize("tol", [1e-6, 1e-10])
def test_device_steps_with_the_increment_criterion_match_oracle(tol):
    """One-call device steps with the reference's third criterion, the relative
    increment of the iterate (linear_solver.py:107-117, 503-515), which is what stops
    the reference's GMRES first on this FD-preconditioned annulus problem: the same
    stop reason and iteration count as the oracle's GMRES, the increment history, and
    the solution within 1e-8."""
    import torch

    from igaops_amd.runs import GmresRun, NativeStep

    patch, C, m, fd, mask, ops = _annulus_fd_problem(3, 6)
    qs, w, n = patch.quadrule_list, patch.nurbs_weights, patch.nbctrlpts_total
    plan = fd.space_preconditioner.plan("f64")
    work = plan.workspace(1)
    A = lambda x, o: ops["f64"].apply(x, out=o)
    P = lambda x, o: plan.apply(x, o, 1, None, work=work)
    b = np.random.default_rng(3).standard_normal(n)
    o = ofd.SpaceFD([StandardGauss(q.degree, q.knotvector, "legendre") for q in qs], np.ones((1, 3, 2), bool))
    o.scalar_coefs = (1.0, 1.0)
    dod = np.setdiff1d(np.arange(n), o.free_nodes()).tolist()
    ref = osv.gmres(lambda v: oo.stiffness_plus_mass(qs, C, m, v, nurbs_weights=w), b, o.apply, tol=tol,
                    maxiters=200, cleandod=dod)
    bd = torch.as_tensor(b, device="cuda")
    mask(bd)
    run = GmresRun(A, P, bd, tol, 200, mask=mask, native=NativeStep(ops["f64"], plan, work, mask.box),
                   increment=True)
    out = run.run()
    print("oracle", ref.iterations, ref.reason, "device", out.iterations, out.reason, flush=True)
    assert ref.reason == "increment" and out.reason == "increment"
    assert out.success and out.iterations == ref.iterations
    assert rel(out.solution.cpu().numpy(), ref.solution) <= 1e-8
    assert min(run.increments) < tol <= min(run.increments[:-1])


def test_refinement_run_float32_vectors_reaches_1e10():
    import torch

    from igaops_amd.solvers import mixed_solve

    patch, C, m, fd, mask, ops = _annulus_fd_problem(4, 8)
    n = patch.nbctrlpts_total
    b = torch.as_tensor(np.random.default_rng(12).standard_normal(n), device="cuda")
    mask(b)
    out = mixed_solve(lambda x, o: ops["f64"].apply(x, out=o), lambda x, o: ops["f32"].apply(x, out=o),
                      lambda x, o: fd.apply_device(x, out=o), b, tol=1e-10, inner_tol=1e-5, mask=mask)
    assert out.success and len(out.history["inner_iterations"]) >= 2
    x = out.solution.cpu().numpy()
    Ax = oo.stiffness_plus_mass(patch.quadrule_list, C, m, x, nurbs_weights=patch.nurbs_weights)
    bh = b.cpu().numpy()
    free = np.zeros(n, bool)
    free[np.asarray(fd.space_preconditioner.space_free_nodes[0])] = True
    Ax[~free] = 0
    assert np.linalg.norm(bh - Ax) / np.linalg.norm(bh) <= 1e-10


def test_multi_patch_fd_matches_per_patch_oracle():
    """Block-diagonal FD over two patches of different sizes: each slice is the
    single-patch oracle apply of that patch (``multi_preconditioner.py:55-59``)."""
    from igaops_amd.fastdiag import MultiFD, SingleFD

    blocks, refs = [], []
    for patch, coefs in ((unit_cube(2, 4), (0.0, 1.0)), (quarter_annulus(3, 3), (0.5, 1.0))):
        qs = patch.quadrule_list
        tab = np.ones((1, 3, 2), bool)
        fd = SingleFD("spatial")
        fd.compute_space_eigendecomposition(qs, tab)
        fd.update_space_eigenvalues(list(coefs))
        blocks.append(fd)
        o = ofd.SpaceFD([StandardGauss(q.degree, q.knotvector, "legendre") for q in qs], tab)
        o.scalar_coefs = coefs
        refs.append((o, patch.nbctrlpts_total))
    multi = MultiFD(blocks)
    n = sum(nb for _, nb in refs)
    X = np.random.default_rng(9).standard_normal((n, 2))
    ref = np.concatenate([o.apply(X[a:a + nb]) for (o, nb), a in zip(refs, np.cumsum([0] + [nb for _, nb in refs]))])
    assert rel(multi.apply(X), ref) <= 1e-12
    assert rel(multi.apply_transpose(X[:, 0]), ref[:, 0]) <= 1e-12  # symmetric spatial FD
    with pytest.raises(ValueError):
        multi.apply(X[:-1])


def test_fd_eigenvalue_updates_reach_the_device_plan():
    """Changing the scalar coefficients or correctors after a first apply updates the
    device plan (``space_precond.py:195-236`` recomputes Lambda each time)."""
    from igaops_amd.fastdiag import SingleFD

    patch = unit_cube(2, 5)
    qs = patch.quadrule_list
    tab = np.ones((1, 3, 2), bool)
    fd = SingleFD("spatial")
    fd.compute_space_eigendecomposition(qs, tab)
    o = ofd.SpaceFD([StandardGauss(q.degree, q.knotvector, "legendre") for q in qs], tab)
    x = np.random.default_rng(10).standard_normal(patch.nbctrlpts_total)
    for coefs in ((0.0, 1.0), (2.0, 1.0), (0.5, 3.0)):
        fd.update_space_eigenvalues(list(coefs))
        o.scalar_coefs = coefs
        assert rel(fd.apply(x), o.apply(x)) <= 1e-12
    fd.add_scalar_space_time_correctors(stiffness_corrector=[np.array([1.0, 2.0, 0.5])])
    o.stiff_corr = np.array([1.0, 2.0, 0.5])
    assert rel(fd.apply(x), o.apply(x)) <= 1e-12


def test_device_step_gmres_with_initial_guess_and_exact_preconditioner():
    """The device-step GMRES from a nonzero x0 reaches the same solution; with the
    exact inverse as preconditioner it stops after one step (the next tentative
    basis vector vanishes: the breakdown path of the delayed reorthogonalisation)."""
    import torch

    from igaops_amd.runs import GmresRun

    rng = np.random.default_rng(21)
    n = 300
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    Ah = (Q * np.geomspace(1.0, 200.0, n)) @ Q.T
    A = torch.as_tensor(Ah, device="cuda")
    Ainv = torch.as_tensor(np.linalg.inv(Ah), device="cuda")
    mv = lambda x, o: o.copy_(A @ x)
    b = torch.as_tensor(rng.standard_normal(n), device="cuda")
    ref = np.linalg.solve(Ah, b.cpu().numpy())
    out = GmresRun(mv, None, b, 1e-12, 400).run()
    assert out.success
    np.testing.assert_allclose(out.solution.cpu().numpy(), ref, rtol=1e-8, atol=1e-10 * np.abs(ref).max())
    x0 = torch.as_tensor(ref + 1e-3 * rng.standard_normal(n), device="cuda")
    out0 = GmresRun(mv, None, b, 1e-12, 400, x0=x0).run()
    assert out0.success and out0.iterations <= out.iterations
    np.testing.assert_allclose(out0.solution.cpu().numpy(), ref, rtol=1e-8, atol=1e-10 * np.abs(ref).max())
    exact = GmresRun(mv, lambda x, o: o.copy_(Ainv @ x), b, 1e-12, 50).run()
    assert exact.success and exact.iterations <= 2
    np.testing.assert_allclose(exact.solution.cpu().numpy(), ref, rtol=1e-9, atol=1e-11 * np.abs(ref).max())


def _annulus_stiffness_gmres_problem():
    from igaops_amd.fastdiag import SingleFD
    from igaops_amd.operators import StiffnessMassOperator

    patch, C, m = build_problem("annulus", 3, 6)
    qs = patch.quadrule_list
    w = patch.nurbs_weights
    n = patch.nbctrlpts_total
    fd = SingleFD("spatial")
    fd.compute_space_eigendecomposition(qs, np.ones((1, 3, 2), bool))
    fd.update_space_eigenvalues([0.0, 1.0])
    free = np.asarray(fd.space_preconditioner.space_free_nodes[0])
    dod = np.setdiff1d(np.arange(n), free).tolist()
    op = StiffnessMassOperator(qs, C=C, m=None, nurbs_weights=w)
    gq = [StandardGauss(q.degree, q.knotvector, "legendre") for q in qs]
    ofd_ = ofd.SpaceFD(gq, np.ones((1, 3, 2), bool))
    oA = lambda v: oo.nurbs_gradu_gradv(qs, C, v, nurbs_weights=w)
    return patch, op, fd, dod, oA, ofd_


@pytest.mark.parametrize("guess,tol", [("near", 1e-10), ("near", 1e-6), ("random", 1e-10)])
def test_linear_solver_warm_start_matches_oracle(guess, tol):
    """``LinearSolver.solve(..., x0=x0)`` solves ``A d = b - A x0`` from zero, measures
    the relative criterion against ``||b - A x0||`` and returns ``x0 + d``
    (``linear_solver.py:228-236``): the oracle's iteration count within one and its
    solution."""
    from igaops_amd.solvers import LinearSolver

    patch, op, fd, dod, oA, ofd_ = _annulus_stiffness_gmres_problem()
    n = patch.nbctrlpts_total
    rng = np.random.default_rng(3)
    b = rng.standard_normal(n)
    exact = osv.gmres(oA, b, ofd_.apply, tol=1e-12, maxiters=200, cleandod=dod).solution
    x0 = exact + 1e-3 * rng.standard_normal(n) if guess == "near" else rng.standard_normal(n)
    x0_in = x0.copy()
    out = LinearSolver(tol, 200, cleandod=dod).solve(op, b, Pfun=fd, x0=x0)
    ref = osv.gmres(oA, b, ofd_.apply, tol=tol, maxiters=200, cleandod=dod, x0=x0)
    np.testing.assert_array_equal(x0, x0_in)  # x0 is copied, not masked in place
    assert out.success and ref.success
    assert abs(out.iterations - ref.iterations) <= 1
    assert rel(out.solution, ref.solution) <= (1e-8 if tol <= 1e-10 else 1e-4)
    # the true residual reduction is measured from the initial residual
    r0 = b.copy()
    r0[dod] = 0
    ax0 = oA(np.where(np.isin(np.arange(n), dod), 0.0, x0))
    ax0[dod] = 0
    r0 -= ax0
    Ax = oA(out.solution)
    Ax[dod] = 0
    r = b.copy()
    r[dod] = 0
    assert np.linalg.norm(r - Ax) <= 10 * tol * np.linalg.norm(r0)


def test_device_step_gmres_warm_start_matches_oracle():
    """``GmresRun(..., x0=x0)`` on the annulus: the oracle's iterations within one
    (the oracle stops on its relative criterion here) and its solution."""
    import torch

    from igaops_amd.runs import GmresRun
    from igaops_amd.solvers import Mask

    patch, op, fd, dod, oA, ofd_ = _annulus_stiffness_gmres_problem()
    n = patch.nbctrlpts_total
    rng = np.random.default_rng(3)
    b = rng.standard_normal(n)
    exact = osv.gmres(oA, b, ofd_.apply, tol=1e-12, maxiters=200, cleandod=dod).solution
    x0 = exact + 1e-3 * rng.standard_normal(n)
    x0[dod] = 0
    bm = b.copy()
    bm[dod] = 0
    ref = osv.gmres(oA, b, ofd_.apply, tol=1e-10, maxiters=200, cleandod=dod, x0=x0)
    assert ref.reason == "relative"
    mask = Mask(box=fd.space_preconditioner.box())
    A = lambda x, o: op.apply(x, out=o)
    P = lambda x, o: fd.apply_device(x, out=o)
    out = GmresRun(A, P, torch.as_tensor(bm, device="cuda"), 1e-10, 200, mask=mask,
                   x0=torch.as_tensor(x0, device="cuda")).run()
    assert out.success and out.reason == "relative"
    assert abs(out.iterations - ref.iterations) <= 1
    assert rel(out.solution.cpu().numpy(), ref.solution) <= 1e-8


@pytest.mark.parametrize("p,nel", [(3, 6), (2, 9), (4, 20)])
def test_fd_bf16_apply_against_oracle(p, nel):
    """BF16 matrix-core inputs with FP32 accumulation (eigenvectors rounded to bf16,
    vector operands split into bf16 high and low parts): within 1e-2 of the float64
    oracle apply (the eigenvector rounding, 2^-9 relative per entry), one and several
    right-hand sides; FP32 and FP64 plans of the same three-launch path to their own
    precision."""
    import torch

    from igaops_amd.fastdiag import SingleFD

    patch = unit_cube(p, nel)
    qs = patch.quadrule_list
    tab = np.ones((1, 3, 2), bool)
    fd = SingleFD("spatial")
    fd.compute_space_eigendecomposition(qs, tab)
    fd.update_space_eigenvalues([0.5, 1.0])
    sp = fd.space_preconditioner
    o = ofd.SpaceFD([StandardGauss(q.degree, q.knotvector, "legendre") for q in qs], tab)
    o.scalar_coefs = (0.5, 1.0)
    n = patch.nbctrlpts_total
    X = np.random.default_rng(13).standard_normal((n, 3))
    ref = o.apply(X)
    Xd = torch.as_tensor(X, device="cuda")
    sp.float32_plan = "bf16"
    y = sp.apply_device(Xd.float().reshape(-1).contiguous(), nrhs=3).reshape(n, 3).double().cpu().numpy()
    err = rel(y, ref)
    print("bf16 FD relative error", err, flush=True)
    assert err <= 1e-2
    y1 = sp.apply_device(Xd[:, 1].float().contiguous()).double().cpu().numpy()
    assert rel(y1, ref[:, 1]) <= 1e-2
    free = np.zeros(n, bool)
    free[o.free_nodes()] = True
    assert np.all(y[~free] == 0)
    sp.float32_plan = "f32"
    assert rel(sp.apply_device(Xd.float().reshape(-1).contiguous(), nrhs=3).reshape(n, 3).double().cpu().numpy(),
               ref) <= 1e-5
    assert rel(sp.apply_device(Xd.reshape(-1).contiguous(), nrhs=3).reshape(n, 3).cpu().numpy(), ref) <= 1e-12
