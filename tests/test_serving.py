"""Solve server: request stream and admission logic (CPU), results alone vs in the mix (GPU)."""
import numpy as np
import pytest

from igaops_amd.serving import SolveRequest, SolveResult, footprint_bytes, poisson_stream, summarize


def test_poisson_stream_is_seeded_and_in_range():
    a = poisson_stream(200, qps=10.0, seed=3)
    b = poisson_stream(200, qps=10.0, seed=3)
    assert [(r.geometry, r.degree, r.nel, r.sigma, r.seed, r.arrival) for r in a] == \
        [(r.geometry, r.degree, r.nel, r.sigma, r.seed, r.arrival) for r in b]
    arr = np.array([r.arrival for r in a])
    assert arr[0] == 0.0 and np.all(np.diff(arr) >= 0)
    # mean inter-arrival close to 1/qps
    assert abs(np.mean(np.diff(arr)) - 0.1) < 0.03
    assert {r.geometry for r in a} == {"cube", "annulus"}
    assert all(2 <= r.degree <= 5 and 16 <= r.nel <= 96 and 0.2 <= r.sigma <= 1.0 for r in a)


def test_footprint_grows_with_size_and_precision():
    small = footprint_bytes(20 ** 3, 60 ** 3, False, "fp64", 300, 200)
    big = footprint_bytes(60 ** 3, 180 ** 3, False, "fp64", 300, 200)
    assert big > 20 * small
    # float32 inner bases are half the size of float64 ones
    assert footprint_bytes(60 ** 3, 180 ** 3, True, "mixed", 300, 200) < footprint_bytes(60 ** 3, 180 ** 3, True,
                                                                                         "fp64", 300, 200)


def test_annulus_correctors_are_the_mean_geometric_conductivity():
    """The FD correctors of an annulus shape equal the per-direction means over the
    quadrature points of the diagonal of ``conductivity_coefficients`` (separable
    factors on the host against the full (3, 3, N_q) array); the angular direction is
    the smallest."""
    from igaops_amd.geometry import quarter_annulus
    from igaops_amd.serving import annulus_correctors, wq_rule

    c = annulus_correctors(3, 7)
    C, _ = quarter_annulus(3, 7, quadrule=wq_rule(3, 7)).conductivity_coefficients(None)
    np.testing.assert_allclose(c, [C[d, d].mean() for d in range(3)], rtol=1e-12)
    assert c[1] < 0.5 * min(c[0], c[2])


def test_summarize_counts_only_converged():
    rs = [SolveResult(i, ("cube", 2, 4), 100, 10, 1e-11 if i % 2 else 1e-8, True, 0.0, 0.0, 0.1 * (i + 1))
          for i in range(10)]
    s = summarize(rs, 1e-10, 2.0)
    assert s["requests"] == 10 and s["reached_tol"] == 5 and s["solves_per_s"] == 2.5
    assert abs(s["p50_ms"] - 550.0) < 1e-6


def _mix():
    shapes = [("cube", 2, 6), ("annulus", 3, 5), ("cube", 3, 4), ("annulus", 2, 6), ("cube", 2, 6), ("annulus", 4, 4)]
    return [SolveRequest(i, g, p, n, 0.3 + 0.1 * i, 100 + i) for i, (g, p, n) in enumerate(shapes)]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_results_same_alone_and_in_mix(mode):
    """Every request of a mixed batch (two of them of the same shape, overlapping on
    the device, their FD applications batched into shared launches) returns the
    same solution as when it is served alone."""
    from igaops_amd.serving import SolveServer

    print("case", mode, flush=True)
    srv = SolveServer(mode=mode, max_active=6, keep_solutions=True, maxiters=200)
    mix = {r.rid: r for r in srv.serve(_mix(), realtime=False)}
    assert srv.stats["max_active"] >= 4
    print("batched steps", srv.stats.get("batched_steps", 0), flush=True)
    for req in _mix():
        alone = SolveServer(mode=mode, max_active=1, keep_solutions=True, maxiters=200).serve([req],
                                                                                            realtime=False)[0]
        got = mix[req.rid]
        assert got.success and alone.success
        assert got.true_rel_residual <= 1e-10 and alone.true_rel_residual <= 1e-10
        assert got.iterations == alone.iterations
        a, b = alone.solution.cpu().numpy(), got.solution.cpu().numpy()
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(a)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_same_shape_requests_batch_fd_and_operator(mode):
    """Four requests of one shape (their own fields and right-hand sides) admitted
    together: their FD applications and their operators go out as batched launches
    (one multi-right-hand-side FD, one grouped operator launch per step of the group),
    and every solution equals the one served alone."""
    from igaops_amd.serving import SolveServer

    print("case", mode, flush=True)
    reqs = lambda: [SolveRequest(i, "annulus", 3, 8, 0.3 + 0.2 * i, 500 + i) for i in range(4)]
    srv = SolveServer(mode=mode, max_active=4, keep_solutions=True, maxiters=200)
    mix = {r.rid: r for r in srv.serve(reqs(), realtime=False)}
    print("batched steps", srv.stats.get("batched_steps", 0), flush=True)
    assert srv.stats.get("batched_steps", 0) > 0
    for req in reqs():
        alone = SolveServer(mode=mode, max_active=1, keep_solutions=True, maxiters=200).serve([req],
                                                                                            realtime=False)[0]
        got = mix[req.rid]
        assert got.success and alone.success
        assert got.true_rel_residual <= 1e-10 and alone.true_rel_residual <= 1e-10
        assert got.iterations == alone.iterations
        a, b = alone.solution.cpu().numpy(), got.solution.cpu().numpy()
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(a)


@pytest.mark.gpu
def test_device_coefficients_match_host():
    import torch

    from igaops_amd.geometry import lognormal_conductivity, quarter_annulus

    patch = quarter_annulus(3, 5)
    C, m = patch.conductivity_coefficients(lognormal_conductivity(patch, 0.7, seed=5))
    Cd, md = patch.device_coefficients(0.7, 5)
    np.testing.assert_allclose(Cd.cpu().numpy(), C, rtol=1e-12, atol=1e-14 * np.abs(C).max())
    np.testing.assert_allclose(md.cpu().numpy(), m, rtol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("geometry,mass", [("cube", False), ("cube", True), ("annulus", False)])
@pytest.mark.parametrize("sigma", [0.0, 0.6])
def test_generated_coefficients_match_the_oracle(geometry, mass, sigma):
    """The operator of generated coefficients (``from_extruded``: field modes, then
    one generating pack launch) against the oracle on the host coefficients
    (``lognormal_conductivity`` and ``conductivity_coefficients``): float64 within
    1e-12, float32 within 5e-6; its packed coefficients equal those packed from the
    device (3, 3, N_q) array to 1e-13."""
    import torch

    from igaops_amd.geometry import lognormal_conductivity, quarter_annulus, unit_cube
    from igaops_amd.operators import StiffnessMassOperator
    from oracle import operators as oo

    patch = (unit_cube if geometry == "cube" else quarter_annulus)(3, 6)
    kappa = lognormal_conductivity(patch, sigma, seed=9) if sigma > 0 else None
    C, m = patch.conductivity_coefficients(kappa)
    if not mass:
        m = np.zeros_like(m)
    w = patch.nurbs_weights if patch.is_nurbs else None
    n = patch.nbctrlpts_total
    x = np.random.default_rng(4).standard_normal(n)
    ref = oo.stiffness_plus_mass(patch.quadrule_list, C, m, x, nurbs_weights=w)
    xd = torch.as_tensor(x, device="cuda")
    for dt, tol in (("f64", 1e-12), ("f32", 5e-6)):
        field = patch.device_field_modes(sigma, 9)
        op = StiffnessMassOperator.from_extruded(patch, field, dtype=dt, mass=mass)
        y = op.apply(xd.to(op.packed.dtype)).double().cpu().numpy()
        assert np.linalg.norm(y - ref) <= tol * np.linalg.norm(ref)
        Cd, md = patch.device_coefficients(sigma, 9, mass=mass)
        old = StiffnessMassOperator(patch.quadrule_list, C=Cd, m=md, nurbs_weights=w, dtype=dt, symmetric=True)
        assert op.form == old.form
        a, b = op.packed.double(), old.packed.double()
        assert float((a - b).abs().max()) <= (1e-13 if dt == "f64" else 1e-6) * float(b.abs().max())


@pytest.mark.gpu
def test_admission_respects_memory_budget():
    from igaops_amd.serving import SolveServer

    reqs = _mix()
    srv = SolveServer(mode="fp64", max_active=6, maxiters=200)
    one = max(srv.footprint(r) for r in reqs)
    srv.budget = int(2.5 * one)
    res = srv.serve(reqs, realtime=False)
    assert len(res) == len(reqs) and all(r.true_rel_residual <= 1e-10 for r in res)
    assert srv.stats["max_bytes"] <= srv.budget
    assert srv.stats["deferred_for_memory"] > 0


def _span_mix():
    """Shapes at both ends of the served stream's range: nel 16 and 96, degree 2 and
    5, both geometries; two requests share the largest shape."""
    shapes = [("cube", 2, 16), ("annulus", 5, 96), ("annulus", 2, 16), ("cube", 5, 96), ("annulus", 5, 96),
              ("cube", 3, 40)]
    return [SolveRequest(i, g, p, n, 0.2 + 0.15 * i, 500 + i) for i, (g, p, n) in enumerate(shapes)]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_results_same_alone_and_in_mix_across_the_stream_range(mode):
    """Requests at the ends of the stream's size range (1.03 M DOF and 26 M points for
    degree 5 and 96 elements) give the same solution alone and in the mix, each to a
    true 1e-10."""
    from igaops_amd.serving import SolveServer

    print("case", mode, flush=True)
    srv = SolveServer(mode=mode, max_active=6, keep_solutions=True)
    mix = {r.rid: r for r in srv.serve(_span_mix(), realtime=False)}
    assert srv.stats["max_active"] >= 4
    for req in _span_mix():
        alone = SolveServer(mode=mode, max_active=1, keep_solutions=True).serve([req], realtime=False)[0]
        got = mix[req.rid]
        print(req.shape, got.iterations, got.true_rel_residual, flush=True)
        assert got.success and alone.success
        assert got.true_rel_residual <= 1e-10 and alone.true_rel_residual <= 1e-10
        assert got.iterations == alone.iterations
        a, b = alone.solution.cpu().numpy(), got.solution.cpu().numpy()
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(a)
        del alone, got
    mix.clear()
