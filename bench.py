"""Benchmark of the weighted-quadrature matrix-free operator path on one MI355X.

A step is one stiffness + mass matvec ``y = (K + M) x`` on one right-hand side,
inputs resident in HBM.  Default workload (the FP64 GPU drop-in case): quarter
thick-walled annulus (r in [1,2]) extruded to height 1 as a NURBS map, degree 4,
64 elements per direction (68^3 = 314,432 DOF), WQ type "2" rules (199^3 =
7,880,599 points), conductivity ``kappa detJ J^-1 J^-T`` with a lognormal ``kappa``
(sigma 0.5, seeded Fourier modes, 8 per axis), seeded N(0,1) vectors, float64.

Printed: one JSON line with GDOF/s (= N_dof / t_step / 1e9); the roofline of the
fused sweep kernel (algorithmic bytes ``8 (2 N_dof + 11 N_q)`` over its average
event-timed launch, and the HBM bytes it moves, read from two rocprofv3 counter
passes over a child run and scaled by the same counters on a 512 MiB device copy);
the 16 right-hand sides of the workload batched in one launch; the relative error
against the CPU oracle; the oracle's own single-thread throughput on the host
(cpu_baseline); and FD-preconditioned GMRES solves of the 16 right-hand sides to
1e-10 (plain FP64, and FP32 operators inside FP64 iterative refinement) with
solves/s and p50/p99 latency; and a stream of 2000 Poisson-arrival mixed-shape
requests through the solve server, per mode (server start reported apart).
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}  # dense MFMA peaks of the MI355X

WORKLOADS = {
    "annulus_p4_n64": dict(geometry="quarter_annulus", degree=4, nel=64, nrhs=16, sigma=0.5),
    "cube_p3_n32": dict(geometry="unit_cube", degree=3, nel=32, nrhs=1, sigma=0.0),
    "annulus_p6_n128": dict(geometry="quarter_annulus", degree=6, nel=128, nrhs=1, sigma=0.5),
    "annulus_p2_n128": dict(geometry="quarter_annulus", degree=2, nel=128, nrhs=1, sigma=0.5),
}


def build_patch(cfg):
    from igaops_amd.geometry import lognormal_conductivity, quarter_annulus, unit_cube

    patch = (quarter_annulus if cfg["geometry"] == "quarter_annulus" else unit_cube)(cfg["degree"], cfg["nel"])
    kappa = lognormal_conductivity(patch, cfg["sigma"], seed=0) if cfg["sigma"] > 0 else None
    C, m = patch.conductivity_coefficients(kappa)
    return patch, C, m


def time_region(fn, steps, warmup):
    """Milliseconds per call of ``fn`` over ``steps`` back-to-back calls, from one pair
    of events on the current stream around all of them (after ``warmup`` calls): the
    average launch duration over the timed region, without a per-launch event."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        fn()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps


def time_events(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    end = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    for i in range(steps):
        start[i].record()
        fn()
        end[i].record()
    torch.cuda.synchronize()
    return [s.elapsed_time(e) for s, e in zip(start, end)]


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(patch, C, m, x, budget_s=20.0):
    """The oracle's matvec (the reference algorithm in numpy/scipy) on the host: one
    core (median of repeats), and all cores of this process's share (one matvec per
    thread on independent right-hand sides, at most 16 threads; scipy's sparse
    products release the GIL)."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import operators as oo

    from threadpoolctl import threadpool_limits

    w = patch.nurbs_weights if patch.is_nurbs else None
    f = lambda v: oo.stiffness_plus_mass(patch.quadrule_list, C, m, v, nurbs_weights=w)
    with threadpool_limits(limits=1):  # one BLAS thread per matvec
        t0 = time.perf_counter()
        y = f(x)
        first = time.perf_counter() - t0
        reps = max(1, min(5, int(budget_s / max(first, 1e-3)) - 1))
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            f(x)
            times.append(time.perf_counter() - t0)
        cores = min(16, len(os.sched_getaffinity(0)))
        X = [np.roll(x, k) for k in range(cores)]
        with ThreadPoolExecutor(cores) as ex:
            t0 = time.perf_counter()
            list(ex.map(f, X))
            t_all = time.perf_counter() - t0
    return y, float(np.median(times)), reps, {"cores": cores, "seconds": t_all, "matvecs": cores}


def fd_error_vs_oracle(patch, dtype, x, reps=20):
    """The FD apply (Dirichlet on all faces, ``Lambda = M + K`` eigenvalues) in
    ``dtype``: relative error against the oracle's float64 apply on ``x``, device time
    per apply (HIP events, ``reps`` applies), its MFMA flops ``4 N_sub sum_d n_d`` (six
    mode products) against the dense peak, and the oracle's time per apply on one
    host core."""
    import torch

    from igaops_amd.fastdiag import SingleFD
    from igaops_amd.quadrature import StandardGauss
    from oracle import fastdiag as ofd

    qs = patch.quadrule_list
    table = np.ones((1, 3, 2), bool)
    fd = SingleFD("spatial")
    fd.compute_space_eigendecomposition(qs, table)
    fd.update_space_eigenvalues([1.0, 1.0])
    sp = fd.space_preconditioner

    def run(dt):
        sp.float32_plan = "bf16" if dt == "bf16" else "f32"
        xd = torch.as_tensor(x, device="cuda").to(torch.float64 if dt == "f64" else torch.float32)
        yv = fd.apply(xd).double().cpu().numpy()
        out = torch.empty_like(xd)
        for _ in range(3):
            fd.apply_device(xd, out)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fd.apply_device(xd, out)
        e1.record()
        torch.cuda.synchronize()
        sp.float32_plan = "f32"
        return yv, e0.elapsed_time(e1) / reps * 1e-3

    runs = {dt: run(dt) for dt in ("f64", "f32", "bf16")}
    y, t_dev = runs[dtype]
    ref_fd = ofd.SpaceFD([StandardGauss(q.degree, q.knotvector, "legendre") for q in qs], table)
    ref_fd.scalar_coefs = (1.0, 1.0)
    from threadpoolctl import threadpool_limits

    with threadpool_limits(limits=1):  # one core, as reported; median of 3 warm calls
        ref = ref_fd.apply(x)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            ref_fd.apply(x)
            times.append(time.perf_counter() - t0)
    t_cpu = float(np.median(times))
    lens = [len(ix) for ix in ref_fd.indices]
    nsub = int(np.prod(lens))
    flops = 4.0 * nsub * sum(lens)
    peak = MFMA_PEAK_TFLOPS[dtype]
    rel = lambda v: float(np.linalg.norm(v - ref) / np.linalg.norm(ref))
    return {
        "rel_err_vs_oracle": rel(y),
        "free_box": lens,
        "us_per_apply": t_dev * 1e6,
        "launches": 3,
        "mfma": {"achieved": flops / t_dev / 1e12, "peak": peak, "unit": "TFLOP/s",
                 "frac": flops / t_dev / 1e12 / peak, "flops": flops},
        "by_precision": {dt: {"us_per_apply": tt * 1e6, "rel_err_vs_oracle": rel(yy),
                              "mfma_frac": flops / tt / 1e12 / MFMA_PEAK_TFLOPS[dt]}
                         for dt, (yy, tt) in runs.items()},
        "cpu_baseline": {"value": t_cpu * 1e6, "unit": "us per apply", "cores": 1, "kind": "port",
                         "sample": "the oracle's FD apply (numpy / scipy dgemm, one BLAS thread) on the same vector, median of 3 after one warm-up"},
    }


def solve_section(patch, C, m, nrhs, tol=1e-10):
    """FD-preconditioned GMRES on ``(K + M) x = b`` for each right-hand side: plain
    FP64 and FP32 operators inside FP64 iterative refinement.  The FD carries the mean
    conductivity per direction and the mean mass coefficient as its correctors."""
    import torch

    from igaops_amd.fastdiag import SingleFD
    from igaops_amd.operators import StiffnessMassOperator
    from igaops_amd.runs import REFINEMENT_INNER_TOL, GmresRun, NativeStep, RefinementRun
    from igaops_amd.solvers import Mask

    qs = patch.quadrule_list
    w = patch.nurbs_weights if patch.is_nurbs else None
    n = patch.nbctrlpts_total
    fd = SingleFD("spatial")
    fd.compute_space_eigendecomposition(qs, np.ones((1, 3, 2), bool))
    # the reference's scalar correctors (fastdiag/preconditioner.py:110-144): mean
    # conductivity per direction and mean mass coefficient over the points
    fd.add_scalar_space_time_correctors(mass_corrector=[float(np.mean(m))],
                                        stiffness_corrector=[np.array([np.mean(C[d, d]) for d in range(3)])])
    fd.update_space_eigenvalues([1.0, 1.0])
    mask = Mask(box=fd.space_preconditioner.box())
    A64 = StiffnessMassOperator(qs, C=C, m=m, nurbs_weights=w, dtype="f64")
    A32 = StiffnessMassOperator(qs, C=C, m=m, nurbs_weights=w, dtype="f32")
    rng = np.random.default_rng(2)
    B = [torch.as_tensor(rng.standard_normal(n), device="cuda") for _ in range(nrhs)]
    for b in B:
        mask(b)
    sp = fd.space_preconditioner
    plans = {dt: sp.plan(dt) for dt in ("f64", "f32")}
    works = {dt: plans[dt].workspace(1) for dt in plans}
    P = {dt: (lambda x, o, dt=dt: plans[dt].apply(x, o, 1, None, work=works[dt])) for dt in plans}
    a64 = lambda x, o: A64.apply(x, out=o)
    a32 = lambda x, o: A32.apply(x, out=o)
    res = {}
    for mode in ("fp64", "mixed"):
        lat, its, ok, true_res = [], [], 0, []
        for b in [None] + B:  # an untimed first solve pays the one-off set-up of the mode
            warm = b is None
            b = B[0] if warm else b
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            # device-stepped runs, one library call per Arnoldi step (residual criteria)
            if mode == "fp64":
                out = GmresRun(a64, P["f64"], b, tol, 400, mask=mask,
                               native=NativeStep(A64, plans["f64"], works["f64"], mask.box)).run()
            else:
                out = RefinementRun(a64, a32, P["f32"], b, tol=tol, inner_tol=REFINEMENT_INNER_TOL, inner_maxiters=300,
                                    mask=mask,
                                    native_lo=NativeStep(A32, plans["f32"], works["f32"], mask.box)).run()
            torch.cuda.synchronize()
            if warm:
                continue
            lat.append(time.perf_counter() - t0)
            its.append(out.iterations)
            r = b - A64.apply(out.solution)
            mask(r)
            tr = float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b))
            true_res.append(tr)
            ok += int(tr <= tol * 1.0001)
        res[mode] = {
            "solves_per_s": len(B) / sum(lat),
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "iterations_mean": float(np.mean(its)),
            "max_true_rel_residual": float(max(true_res)),
            "reached_1e-10": ok,
        }
    res["speedup_mixed_vs_fp64"] = res["mixed"]["solves_per_s"] / res["fp64"]["solves_per_s"]
    # all right-hand sides at once: their Arnoldi steps batched (one multi-right-hand-side
    # FD launch and one operator launch reading the coefficient stream once per group step)
    from igaops_amd import device as dv
    from igaops_amd.runs import run_batched

    # the solves on four streams (their starts, refinement residuals and solution updates);
    # a group step's FD, operator and orthogonalisation launches go on a batch stream
    streams = [torch.cuda.Stream() for _ in range(4)]
    bstream = torch.cuda.Stream()
    # scratch per stream: the FD workspace and the operators' partial blocks of work on
    # different streams must not be shared
    fd_works = {dt: [plans[dt].workspace(1) for _ in streams] for dt in plans}
    op_works = [dv.workspace(A64.plan.workspace_bytes(1)) for _ in streams]

    def make(mode, i, b):
        j = i % len(streams)
        s = streams[j]
        Pi = {dt: (lambda x, o, dt=dt, s=s, j=j: plans[dt].apply(x, o, 1, s, work=fd_works[dt][j])) for dt in plans}
        a64s = lambda x, o, s=s, j=j: A64.apply(x, out=o, stream=s, work=op_works[j])
        with torch.cuda.stream(s):
            if mode == "fp64":
                return GmresRun(a64s, Pi["f64"], b, tol, 400, mask=mask, stream=s,
                                native=NativeStep(A64, plans["f64"], fd_works["f64"][j], mask.box))
            return RefinementRun(a64s, a32, Pi["f32"], b, tol=tol, inner_tol=REFINEMENT_INNER_TOL, inner_maxiters=300,
                                 mask=mask,
                                 stream=s, native_lo=NativeStep(A32, plans["f32"], fd_works["f32"][j], mask.box))

    for mode in ("fp64", "mixed"):
        # untimed: the same set of solves once, so that their device buffers, pinned host
        # buffers and the batch scratch come from the allocators' caches when timed
        run_batched([make(mode, i, b) for i, b in enumerate(B)], batch_stream=bstream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        outs, done_at = run_batched([make(mode, i, b) for i, b in enumerate(B)], batch_stream=bstream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        lat = [d - t0 for d in done_at]
        true_res = []
        for b, out in zip(B, outs):
            r = b - A64.apply(out.solution)
            mask(r)
            true_res.append(float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b)))
        res[mode + "_batched"] = {
            "solves_per_s": len(B) / wall,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "iterations_mean": float(np.mean([o.iterations for o in outs])),
            "max_true_rel_residual": float(max(true_res)),
            "reached_1e-10": int(sum(tr <= tol * 1.0001 for tr in true_res)),
        }
    res["speedup_mixed_vs_fp64_batched"] = (res["mixed_batched"]["solves_per_s"]
                                            / res["fp64_batched"]["solves_per_s"])
    return res


PMC_COPY_BYTES = 512 * 1024 * 1024


def serve_section(nreq=2000, qps=200.0, seed=0, max_active=16, budget_gb=200.0, workers=8):
    """Poisson-arrival stream of Poisson/diffusion solves (nel U{16..96}, p U{2..5},
    cube or annulus, sigma U[0.2,1.0]) served continuously, plain float64 and mixed
    precision on the same stream; latency from arrival to a checked 1e-10 residual.
    At server start the host set-up of every (p, nel) the server accepts is built (the
    whole declared range, not the stream's sizes) and its time is reported; the
    per-shape device set-up and the per-request set-up are inside the latencies.
    The default rate (200 requests/s) is above what either mode serves, so the stream
    backs up and solves/s is each mode's capacity on the same stream; p50/p99 then
    include the time requests wait for admission."""
    from igaops_amd.serving import SolveServer, poisson_stream, summarize

    out = {"requests": nreq, "qps": qps, "seed": seed, "max_active": max_active,
           "stream": "Poisson arrivals; nel U{16..96}, p U{2..5}, cube or quarter annulus, sigma U[0.2,1.0]; "
                     "homogeneous Dirichlet; FD-preconditioned GMRES (fp64) or float64 refinement around "
                     "float32 GMRES (mixed) to a true relative residual of 1e-10"}
    import torch

    for mode in ("fp64", "mixed"):
        torch.cuda.empty_cache()  # each mode starts from an empty allocator cache
        torch.cuda.reset_peak_memory_stats()
        srv = SolveServer(mode=mode, max_active=max_active, hbm_budget_gb=budget_gb)
        start = srv.prepare(range(2, 6), range(16, 97), workers=workers)
        t1 = time.perf_counter()
        res = srv.serve(poisson_stream(nreq, qps, seed=seed), progress_every=250)
        wall = time.perf_counter() - t1
        s = summarize(res, 1e-10, wall)
        s.update({"server_start": dict(start, note="all 324 (p, nel) of p 2..5, nel 16..96: rules and FD pencils "
                                                   "in a process pool (built once per process), device FD and "
                                                   "operator plans of this mode, per-shape device "
                                                   "resources of both geometries"),
                  "max_concurrent": srv.stats["max_active"], "peak_admitted_gb": srv.stats["max_bytes"] / 1e9,
                  "deferred_for_memory": srv.stats["deferred_for_memory"], "shapes": len(srv.shapes),
                  "host_admit_s": srv.stats["admit_s"], "host_advance_s": srv.stats["advance_s"],
                  "host_complete_s": srv.stats["complete_s"], "host_shape_setup_s": srv.stats["shape_setup_s"],
                  "batched_steps": srv.stats.get("batched_steps", 0),
                  "host_setup_s": {k[6:-2]: v for k, v in srv.stats.items() if k.startswith("setup_")},
                  "allocator": {k: srv.stats[k] for k in ("alloc_retries", "device_frees", "device_allocs",
                                                           "peak_reserved_gb")}})
        out[mode] = s
        print(f"serve {mode}: {json.dumps(s)}", file=sys.stderr, flush=True)
    out["speedup_mixed_vs_fp64"] = out["mixed"]["solves_per_s"] / max(out["fp64"]["solves_per_s"], 1e-30)
    out["same_shape"] = same_shape_section()
    return out


def same_shape_section(nreq=96, max_active=16):
    """A stream whose requests share three shapes (annulus p=3 nel=24, cube p=4
    nel=32, annulus p=2 nel=40; their own fields, sigma in [0.3, 0.8), and right-hand
    sides), all present at the start: served with same-shape batching (one
    multi-right-hand-side FD launch and one grouped operator launch per group step)
    and without, per mode; solves/s over the whole stream and the number of
    request steps that went out in groups."""
    from igaops_amd.serving import SolveRequest, SolveServer, summarize

    shapes = [("annulus", 3, 24), ("cube", 4, 32), ("annulus", 2, 40)]
    reqs = lambda: [SolveRequest(i, *shapes[i % 3], 0.3 + 0.5 * ((i * 37) % 100) / 100.0, 1000 + i)
                    for i in range(nreq)]
    out = {"requests": nreq, "shapes": [list(s) for s in shapes], "max_active": max_active}
    for mode in ("fp64", "mixed"):
        row = {}
        for batch in (False, True):
            srv = SolveServer(mode=mode, max_active=max_active, batch_steps=batch)
            srv.prepare((2, 3, 4), (24, 32, 40), workers=1)
            srv.serve(reqs()[:max_active], realtime=False)  # shape set-up and allocator warm-up
            t1 = time.perf_counter()
            res = srv.serve(reqs(), realtime=False)
            wall = time.perf_counter() - t1
            s = summarize(res, 1e-10, wall)
            row["batched" if batch else "unbatched"] = {
                "solves_per_s": s["solves_per_s"], "reached_tol": s["reached_tol"],
                "mean_iterations": s["mean_iterations"], "wall_s": wall,
                "batched_steps": srv.stats.get("batched_steps", 0)}
        row["speedup_batched"] = row["batched"]["solves_per_s"] / max(row["unbatched"]["solves_per_s"], 1e-30)
        out[mode] = row
        print(f"same-shape {mode}: {json.dumps(row)}", file=sys.stderr, flush=True)
    return out


def c3_section(degrees=(2, 4, 6), nel=128, tol=1e-10):
    """Mixed precision at 128 elements per direction on the quarter annulus (lognormal
    sigma 0.5): per degree, the FP64 and FP32 operators' errors against the oracle's
    FP64 matvec and their event-timed applies; the FP64, FP32 and BF16-input FD
    applies' errors against the oracle's FD; and one right-hand side solved to a true
    1e-10 by FP64 GMRES and by FP64 refinement around FP32 GMRES with the FP32 FD and
    with the BF16-input FD (one-call Arnoldi steps)."""
    import torch

    from igaops_amd.fastdiag import SingleFD
    from igaops_amd.operators import StiffnessMassOperator
    from igaops_amd.quadrature import StandardGauss
    from igaops_amd.runs import REFINEMENT_INNER_TOL, GmresRun, NativeStep, RefinementRun
    from igaops_amd.solvers import Mask
    from oracle import fastdiag as ofd
    from oracle import operators as oo

    out = {}
    for p in degrees:
        cfg = dict(geometry="quarter_annulus", degree=p, nel=nel, nrhs=1, sigma=0.5)
        patch, C, m = build_patch(cfg)
        qs, w, n = patch.quadrule_list, patch.nurbs_weights, patch.nbctrlpts_total
        A64 = StiffnessMassOperator(qs, C=C, m=m, nurbs_weights=w, dtype="f64")
        A32 = StiffnessMassOperator(qs, C=C, m=m, nurbs_weights=w, dtype="f32")
        x = np.random.default_rng(3).standard_normal(n)
        xd = torch.as_tensor(x, device="cuda")
        xf = xd.float()
        ref = oo.stiffness_plus_mass(qs, C, m, x, nurbs_weights=w)
        rel = lambda a: float(np.linalg.norm(a - ref) / np.linalg.norm(ref))
        row = {"ndof": n, "nq": patch.nbquadpts_total,
               "matvec_f64": {"rel_err_vs_oracle": rel(A64.apply(xd).cpu().numpy()),
                              "ms": float(np.mean(time_events(lambda: A64.apply(xd), 10, 2)))},
               "matvec_f32": {"rel_err_vs_oracle": rel(A32.apply(xf).double().cpu().numpy()),
                              "ms": float(np.mean(time_events(lambda: A32.apply(xf), 10, 2)))}}
        del ref
        fd = SingleFD("spatial")
        fd.compute_space_eigendecomposition(qs, np.ones((1, 3, 2), bool))
        fd.add_scalar_space_time_correctors(mass_corrector=[float(np.mean(m))],
                                            stiffness_corrector=[np.array([np.mean(C[d, d]) for d in range(3)])])
        fd.update_space_eigenvalues([1.0, 1.0])
        sp = fd.space_preconditioner
        oref = ofd.SpaceFD([StandardGauss(q.degree, q.knotvector, "legendre") for q in qs], np.ones((1, 3, 2), bool))
        oref.stiff_corr = np.array([np.mean(C[d, d]) for d in range(3)])
        oref.mass_corr = float(np.mean(m))
        oref.scalar_coefs = (1.0, 1.0)
        pref = oref.apply(x)
        mask = Mask(box=sp.box())
        b = torch.as_tensor(np.random.default_rng(4).standard_normal(n), device="cuda")
        mask(b)
        a64 = lambda v, o: A64.apply(v, out=o)
        a32 = lambda v, o: A32.apply(v, out=o)

        def true_res(xs):
            r = b - A64.apply(xs)
            mask(r)
            return float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b))

        fdr = {}
        for plan_name in ("f64", "f32", "bf16"):
            sp.float32_plan = "bf16" if plan_name == "bf16" else "f32"
            dt = "f64" if plan_name == "f64" else "f32"
            plan = sp.plan(plan_name)
            work = plan.workspace(1)
            v = xd if dt == "f64" else xf
            y = torch.empty_like(v)
            plan.apply(v, y, 1, None, work=work)
            fdr[plan_name] = {"rel_err_vs_oracle": float(np.linalg.norm(y.double().cpu().numpy() - pref)
                                                         / np.linalg.norm(pref)),
                              "us": float(np.mean(time_events(lambda: plan.apply(v, y, 1, None, work=work), 20, 2)))
                              * 1e3}
            P = lambda vv, o, plan=plan, work=work: plan.apply(vv, o, 1, None, work=work)
            for rep in range(2):  # the second solve is timed (the first pays one-off set-up)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                if plan_name == "f64":
                    res = GmresRun(a64, P, b, tol, 600, mask=mask, native=NativeStep(A64, plan, work, mask.box)).run()
                else:
                    res = RefinementRun(a64, a32, P, b, tol=tol, inner_tol=REFINEMENT_INNER_TOL, inner_maxiters=400,
                                        mask=mask,
                                        native_lo=NativeStep(A32, plan, work, mask.box)).run()
                torch.cuda.synchronize()
            fdr[plan_name]["solve"] = {"s": time.perf_counter() - t0, "iterations": int(res.iterations),
                                       "refinements": len(res.history.get("inner_iterations", [])) or None,
                                       "true_rel_residual": true_res(res.solution), "success": bool(res.success)}
        sp.float32_plan = "f32"
        row["fd"] = fdr
        row["speedup_mixed_f32_vs_fp64"] = fdr["f64"]["solve"]["s"] / fdr["f32"]["solve"]["s"]
        row["speedup_mixed_bf16_vs_fp64"] = fdr["f64"]["solve"]["s"] / fdr["bf16"]["solve"]["s"]
        out[f"p{p}"] = row
        print(f"c3 p={p}: {json.dumps(row)}", file=sys.stderr, flush=True)
        del A64, A32, C, m
        torch.cuda.empty_cache()
    return out


def fd_section(nels=(64, 128), reps=50):
    """The FD apply alone (degree 4, every face Dirichlet, ``Lambda = M + K``): free boxes
    66^3 and 130^3, FP64 / FP32 / BF16-input plans, event-timed on the launch stream;
    flops ``12 n^4`` per apply (two products of ``n x n`` factors per direction and
    way) against the dense matrix-core peak of the precision."""
    import torch

    from igaops_amd.fastdiag import SingleFD
    from igaops_amd.quadrature import WeightedQuadrature, make_open_knotvector

    out = {}
    for nel in nels:
        q = WeightedQuadrature(4, make_open_knotvector(4, nel), quadtype="2")
        fd = SingleFD("spatial")
        fd.compute_space_eigendecomposition([q, q, q], np.ones((1, 3, 2), bool))
        fd.update_space_eigenvalues([1.0, 1.0])
        sp = fd.space_preconditioner
        n = q.nbctrlpts ** 3
        box = q.nbctrlpts - 2
        flops = 12.0 * box ** 4
        row = {}
        for dt in ("f64", "f32", "bf16"):
            sp.float32_plan = "bf16" if dt == "bf16" else "f32"
            plan = sp.plan(dt)
            work = plan.workspace(1)
            x = torch.randn(n, device="cuda", dtype=torch.float64 if dt == "f64" else torch.float32)
            y = torch.empty_like(x)
            us = float(np.mean(time_events(lambda: plan.apply(x, y, 1, None, work=work), reps, 3))) * 1e3
            row[dt] = {"us_per_apply": us, "tflops": flops / us * 1e-6,
                       "mfma_frac": flops / us * 1e-6 / MFMA_PEAK_TFLOPS[dt]}
        sp.float32_plan = "f32"
        out[f"box{box}"] = row
        print(f"fd box {box}: {json.dumps(row)}", file=sys.stderr, flush=True)
    return out


def krylov_section(n=68 ** 3, m=200, sizes=(16, 64, 128), reps=20):
    """One device GMRES Arnoldi orthogonalisation step (``iga_gmres_step``: two-vector
    many-dot pass, small dense kernel, update pass) at the annulus p=4 n=64 vector
    length, on random bases of j vectors, HIP-event timed; microseconds per step."""
    import torch

    from igaops_amd import _lib
    from igaops_amd import device as dv

    lib = _lib.load()
    out = {}
    for name, tdt, code in (("f64", torch.float64, _lib.IGA_F64), ("f32", torch.float32, _lib.IGA_F32)):
        g = torch.Generator(device="cuda").manual_seed(5)
        V = torch.randn(m + 2, n, dtype=tdt, device="cuda", generator=g) / n ** 0.5
        w = torch.randn(n, dtype=tdt, device="cuda", generator=g) / n ** 0.5
        ls = torch.zeros(int(lib.iga_gmres_state_doubles(m)), dtype=torch.float64, device="cuda")
        work = dv.workspace(lib.iga_reduce_workspace_bytes(n, m + 1))
        sp = dv.stream_ptr()
        res = {}
        for j in sizes:
            def step():
                _lib.call("iga_gmres_step", code, n, j + 1, m, dv.ptr(V), n, dv.ptr(w), dv.ptr(ls), dv.ptr(work), sp)
            for _ in range(3):
                step()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                step()
            e1.record()
            torch.cuda.synchronize()
            res[str(j)] = e0.elapsed_time(e1) / reps * 1e3
        out[name] = res
    return {"n": n, "step_us_by_basis_size": out}


def assembly_section(p=2, nel=32, reps=3):
    """Sum-factorised sparse assembly of the stiffness + mass matrix of the unit cube
    (degree p, nel elements/dir, WQ type "2", lognormal conductivity) on the device:
    seconds per matrix (tables, kernel, copy to a scipy CSR), and the oracle's
    explicit Kronecker assembly of the same matrix on one host core for scale."""
    import torch

    from igaops_amd.geometry import lognormal_conductivity, unit_cube
    from igaops_amd.operators import BsplineOperations
    from oracle import assembly as asm

    patch = unit_cube(p, nel)
    C, m = patch.conductivity_coefficients(lognormal_conductivity(patch, 0.5, seed=1))
    qs = patch.quadrule_list
    A = None
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        A = BsplineOperations.assemble_scalar_gradu_gradv(qs, C) + BsplineOperations.assemble_scalar_u_v(qs, m)
        times.append(time.perf_counter() - t0)
    # CPU leg: the oracle as baseline and checker
    t0 = time.perf_counter()
    ref = asm.bspline_stiffness(qs, C) + asm.bspline_mass(qs, m)
    t_cpu = time.perf_counter() - t0
    err = float(abs(A - ref).max() / abs(ref).max())
    return {"workload": f"unit cube p={p} nel={nel}, stiffness + mass", "n": int(A.shape[0]), "nnz": int(A.nnz),
            "seconds_per_matrix": min(times), "max_rel_diff_vs_oracle": err,
            "cpu_baseline": {"value": t_cpu, "unit": "s per matrix", "cores": 1, "kind": "port",
                             "sample": "the oracle's explicit Kronecker assembly of the same matrix"}}


def krylov_roofline(n, j=64):
    """One FP64 GMRES orthogonalisation step at basis size ``j`` on vectors of the
    workload's length: two passes over ``j + 2`` vectors, against the HBM peak."""
    us = krylov_section(n=n, m=max(2 * j, 16), sizes=(j,))["step_us_by_basis_size"]["f64"][str(j)]
    nbytes = 2.0 * (j + 2) * n * 8
    return {"basis_size": j, "us_per_step": us,
            "hbm": {"achieved": nbytes / (us * 1e-6) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": nbytes / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, "bytes": nbytes}}


def spacetime_section(steps=5, nrhs_seq=10, tol=1e-10, space=(3, 32), time_=(2, 64)):
    """Space-time heat equation: 3-D unit cube (degree 3, 32 elements/dir) x time
    (degree 2, 64 elements), WQ type "2" rules, unit diffusivity,
    ``A = u_v(time_ders=(0,1)) + gradu_gradv(time_ders=(0,0))``, Dirichlet on the
    space boundary and the initial time; arrowhead space-time FD; GCRODR (m=40, k=20)
    over 10 right-hand sides ``b_{i+1} = b_i + 0.05 |b_i| / sqrt(N) xi_i``."""
    import torch

    from igaops_amd.fastdiag import SingleFD
    from igaops_amd.gcrodr import GCRODR
    from igaops_amd.quadrature import WeightedQuadrature, make_open_knotvector
    from igaops_amd.spacetime import BoxMasked, SpaceTimeOperator

    t0 = time.perf_counter()
    qs = [WeightedQuadrature(space[0], make_open_knotvector(*space), quadtype="2") for _ in range(3)]
    qt = WeightedQuadrature(time_[0], make_open_knotvector(*time_), quadtype="2")
    ql = qs + [qt]
    fd = SingleFD("spacetime")
    fd.compute_space_eigendecomposition(qs, np.ones((1, 3, 2), bool))
    fd.compute_time_schurdecomposition(qt)
    fd.update_space_eigenvalues([0.0, 1.0])
    op = SpaceTimeOperator(ql, 1.0, np.eye(3))
    A = BoxMasked(op, fd.sptm_preconditioner.box())
    n = op.ndof
    rng = np.random.default_rng(5)
    x = torch.as_tensor(rng.standard_normal(n), device="cuda")
    y = torch.empty_like(x)
    setup_s = time.perf_counter() - t0
    mv = float(np.mean(time_events(lambda: op.apply_device(x, out=y), steps, 1)))
    pc = float(np.mean(time_events(lambda: fd.sptm_preconditioner.apply_device(x, out=y), steps, 1)))
    b = torch.as_tensor(rng.standard_normal(n), device="cuda")
    A._mask(b, 1)
    g = GCRODR(tol, 2000, krylov_size=40, recycled_size=20)
    lat, its, true_res = [], [], []
    for i in range(nrhs_seq):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        out = g.solve_device(A, b, Pfun=fd.sptm_preconditioner)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t1)
        its.append(out.iterations)
        r = b - A.apply_device(out.solution)
        true_res.append(float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b)))
        xi = torch.as_tensor(rng.standard_normal(n), device="cuda")
        A._mask(xi, 1)
        b = b + 0.05 * torch.linalg.vector_norm(b) / n ** 0.5 * xi
    nq = op.nq
    return {
        "workload": f"space-time heat, space p={space[0]} nel={space[1]}, time p={time_[0]} nel={time_[1]}, "
                    f"WQ type 2, unit cube, unit diffusivity",
        "ndof": n,
        "nq": nq,
        "matvec_ms": mv,
        "matvec_gdofs": n / mv / 1e6,
        "matvec_path": "fused, sliced in time" if op.fused else "composed",
        "matvec_coefficient_fields": op.ncoef_fields,
        "matvec_pct_hbm_roofline": op.algorithmic_bytes(1) / (mv * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "fd_apply_ms": pc,
        "gcrodr_iterations": its,
        "solve_s": lat,
        "solves_per_s": len(lat) / sum(lat),
        "preconditioned_tol": tol,
        "max_true_rel_residual": max(true_res),
        "setup_s": setup_s,
    }


def pmc_child(args):
    """Body of the counter runs: three launches of the sweep kernel on one
    right-hand side, then three 512 MiB device copies (the counters' scale)."""
    import torch

    from igaops_amd.operators import BATCH_PLAN_MIN_RHS, StiffnessMassOperator

    cfg = dict(WORKLOADS[args.workload])
    patch, C, m = build_patch(cfg)
    w = patch.nurbs_weights if patch.is_nurbs else None
    op = StiffnessMassOperator(patch.quadrule_list, C=C, m=m, nurbs_weights=w, dtype=args.dtype)
    tdt = torch.float64 if args.dtype == "f64" else torch.float32
    x = torch.randn(patch.nbctrlpts_total, dtype=tdt, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        op.plan.apply_stage(1, op.packed, x, y, 1, op.w)
    a = torch.ones(PMC_COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("pmc child done", flush=True)


def _counter_values(csv_path, name_part):
    import csv

    with open(csv_path) as fh:
        return [float(r["Counter_Value"]) for r in csv.DictReader(fh) if name_part in r["Kernel_Name"]]


def pmc_traffic(args):
    """HBM bytes per launch of the sweep kernel from two rocprofv3 passes
    (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by the ratio of
    the bytes a 512 MiB copy moves to what the same counter reads for it."""
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None, "rocprofv3 not found"
    out = {}
    tmp = tempfile.mkdtemp(prefix="igaops_pmc_", dir="/tmp")
    env = dict(os.environ, TMPDIR="/tmp")
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(tmp, counter)
            cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--", sys.executable,
                   os.path.abspath(__file__), "--pmc-child", "--workload", args.workload, "--dtype", args.dtype]
            subprocess.run(cmd, cwd="/tmp", env=env, timeout=150, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL, check=False)
            csvs = [os.path.join(r, f) for r, _, fs in os.walk(d) for f in fs if f.endswith("counter_collection.csv")]
            if not csvs:
                return None, f"no {counter} counters collected"
            kv = _counter_values(csvs[0], "fused_apply_kernel")
            # the three large copies (set-up uploads also run the copy kernel, at small sizes)
            cv = _counter_values(csvs[0], "copyBuffer")
            kern = sum(kv) / len(kv) if kv else None
            copy = max(cv) if cv else None
            if not kern or not copy:
                return None, f"{counter}: kernel or copy dispatch missing"
            scale = PMC_COPY_BYTES / (copy * 1024.0)
            out[counter] = {"kib": kern, "scale": scale, "bytes": kern * 1024.0 * scale}
    except (subprocess.TimeoutExpired, OSError) as e:
        return None, f"counter pass failed: {type(e).__name__}"
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return out, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="annulus_p4_n64", choices=sorted(WORKLOADS))
    ap.add_argument("--dtype", default="f64", choices=["f64", "f32"])
    ap.add_argument("--nrhs", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-traffic", action="store_true")
    ap.add_argument("--spacetime", action="store_true", help="only the space-time GCRODR section")
    ap.add_argument("--krylov", action="store_true", help="only the GMRES orthogonalisation step timings")
    ap.add_argument("--fd", action="store_true", help="only the FD apply timings (66^3 and 130^3 boxes)")
    ap.add_argument("--c3", action="store_true", help="only the mixed-precision case at 128 elements, p 2, 4, 6")
    ap.add_argument("--assembly", action="store_true", help="only the sparse assembly timing")
    ap.add_argument("--serve", type=int, default=0, help="only the serving section, with this many requests")
    ap.add_argument("--same-shape", action="store_true", help="only the same-shape batching stream")
    ap.add_argument("--solve-only", action="store_true", help="only the solves of the workload's right-hand sides")
    ap.add_argument("--qps", type=float, default=200.0)
    ap.add_argument("--no-serve", action="store_true", help="skip the short serving stream of the default run")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.pmc_child:
        pmc_child(args)
        return
    if args.spacetime:
        print(json.dumps({"spacetime": spacetime_section()}))
        return
    if args.krylov:
        print(json.dumps({"krylov": krylov_section()}))
        return
    if args.fd:
        print(json.dumps({"fd": fd_section()}))
        return
    if args.c3:
        print(json.dumps({"c3": c3_section()}))
        return
    if args.assembly:
        print(json.dumps({"assembly": assembly_section()}))
        return
    if args.same_shape:
        print(json.dumps({"same_shape": same_shape_section()}))
        return
    if args.solve_only:
        cfg = dict(WORKLOADS[args.workload])
        patch, C, m = build_patch(cfg)
        print(json.dumps({"solve": solve_section(patch, C, m, cfg["nrhs"])}))
        return
    if args.serve:
        print(json.dumps({"serve": serve_section(args.serve, args.qps)}))
        return

    import torch

    from igaops_amd.operators import BATCH_PLAN_MIN_RHS, StiffnessMassOperator

    cfg = dict(WORKLOADS[args.workload])
    if args.nrhs:
        cfg["nrhs"] = args.nrhs
    r_batch = cfg["nrhs"]
    t_setup = time.perf_counter()
    patch, C, m = build_patch(cfg)
    w = patch.nurbs_weights if patch.is_nurbs else None
    op = StiffnessMassOperator(patch.quadrule_list, C=C, m=m, nurbs_weights=w, dtype=args.dtype)
    assert op.fused
    setup_s = time.perf_counter() - t_setup
    n, nq = patch.nbctrlpts_total, patch.nbquadpts_total
    tdt = torch.float64 if args.dtype == "f64" else torch.float32
    rng = np.random.default_rng(1)
    X = torch.as_tensor(rng.standard_normal((n, r_batch)), device="cuda").to(tdt).contiguous()
    x1 = X[:, 0].contiguous()
    y1 = torch.empty_like(x1)
    Y = torch.empty_like(X)

    # timed region: K single-right-hand-side matvecs bracketed by synchronisations
    def step():
        op.apply(x1, out=y1, nrhs=1)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed * 1e3 / args.steps
    value = n * args.steps / elapsed / 1e9

    # dominant kernel (the coefficient sweep) timed alone: K back-to-back launches between
    # two events on its stream (the launch stream is the current one)
    plan, packed = op.plan, op.packed
    main_ms = time_region(lambda: plan.apply_stage(1, packed, x1, y1, 1, op.w), args.steps, args.warmup)
    gather_ms = time_region(lambda: plan.apply_stage(2, packed, x1, y1, 1, op.w), args.steps, args.warmup)
    t_main = float(main_ms) * 1e-3
    alg_bytes = op.algorithmic_bytes(1)
    achieved = alg_bytes / t_main / 1e9

    # the batch of right-hand sides of the workload in one launch
    batched = None
    if r_batch > 1:
        # one or two right-hand sides per work-group (two: each coefficient load and plane
        # barrier serves both; equal to rounding); the faster is reported
        Xf, Yf = X.reshape(-1), Y.reshape(-1)
        by_nr, outs = {}, {}
        # (the operator's multi-right-hand-side launches run on its fewest-slab plan)
        for nr in (1, 2):
            op.set_rhs_block(nr)
            bms = time_events(lambda: op.apply(Xf, out=Yf, nrhs=r_batch), max(3, args.steps // 2), 1)
            by_nr[nr] = float(np.mean(bms))
            outs[nr] = Yf.clone()
        best = min(by_nr, key=by_nr.get)
        op.set_rhs_block(1)
        tb = by_nr[best] * 1e-3
        batched = {
            "nrhs": r_batch,
            "q3_slabs": op.batch_plan().nslab if r_batch >= BATCH_PLAN_MIN_RHS else plan.nslab,
            "ms": tb * 1e3,
            "gdofs": n * r_batch / tb / 1e9,
            "pct_hbm_roofline": op.algorithmic_bytes(r_batch) / tb / 1e9 / HBM_PEAK_GBS,
            "rhs_per_workgroup": best,
            "ms_by_rhs_per_workgroup": {str(k): v for k, v in by_nr.items()},
            "max_rel_diff_between_blockings": float((outs[1] - outs[2]).abs().max() / outs[1].abs().max()),
            "note": "the right-hand sides share one coefficient stream, so this launch is bound by the "
                    "per-point contractions, not by HBM; the fraction is of the HBM roofline all the same",
        }

    # parity on one right-hand side against the oracle, and the CPU baseline
    x0 = x1.double().cpu().numpy()
    y0 = op.apply(x1).double().cpu().numpy()
    cpu = None
    relerr = None
    if not args.no_cpu:
        yref, t_cpu, reps, allc = cpu_baseline(patch, C, m, x0)
        relerr = float(np.linalg.norm(y0 - yref) / np.linalg.norm(yref))
        cpu = {
            "value": n / t_cpu / 1e9,
            "unit": "GDOF/s",
            "cores": 1,
            "kind": "port",
            "sample": f"one right-hand side of the same workload, oracle matvec (numpy/scipy, one BLAS thread), "
                      f"median of {reps} after one warm-up call",
            "seconds_per_matvec": t_cpu,
            "cpu_model": cpu_model(),
            "all_cores": {
                "value": n * allc["matvecs"] / allc["seconds"] / 1e9,
                "unit": "GDOF/s",
                "cores": allc["cores"],
                "sample": f"{allc['matvecs']} matvecs of the same workload on {allc['cores']} threads at once "
                          f"(independent right-hand sides), wall time {allc['seconds']:.2f} s",
            },
        }
        fd_info = fd_error_vs_oracle(patch, args.dtype, x0)
        fd_err = fd_info["rel_err_vs_oracle"]
    out = {
        "metric": "matvec GDOF/s (% HBM roofline); solves/s + p50/p99 latency at 1e-10 residual",
        "value": value,
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seeded geometry, lognormal conductivity and N(0,1) right-hand sides)",
        "config": {
            "workload": f"{args.workload}: {cfg['geometry']} NURBS map, degree {cfg['degree']}, {cfg['nel']} "
            f"elements/dir, WQ type 2, stiffness+mass matvec, one right-hand side per step "
            f"({r_batch} right-hand sides of the workload batched in 'batched')",
            "ndof": n,
            "nq": nq,
            "nrhs_per_step": 1,
            "coef_form": op.form,
            "slabs": op.plan.nslab,
            "tiles": list(op.plan.tiles),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "fused_apply_kernel",
            "kernel_ms": t_main * 1e3,
            "gather_ms": float(np.mean(gather_ms)),
            "algorithmic_bytes": alg_bytes,
        },
        "pct_hbm_roofline_matvec": alg_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "batched": batched,
        "rel_err_vs_oracle": relerr,
        "fd_rel_err_vs_oracle": None if args.no_cpu else fd_err,
        "fd_apply": None if args.no_cpu else fd_info,
        "gmres_step": krylov_roofline(n),
        "cpu_baseline": cpu,
        "setup_s": setup_s,
    }
    if not args.no_traffic:
        pm, why = pmc_traffic(args)
        if pm is not None:
            t = pm["FETCH_SIZE"]["bytes"] + pm["WRITE_SIZE"]["bytes"]
            out["roofline"]["traffic"] = t
            out["roofline"]["traffic_detail"] = {
                "fetch_bytes": pm["FETCH_SIZE"]["bytes"],
                "write_bytes": pm["WRITE_SIZE"]["bytes"],
                "fetch_scale": pm["FETCH_SIZE"]["scale"],
                "write_scale": pm["WRITE_SIZE"]["scale"],
                "traffic_over_algorithmic": t / alg_bytes,
            }
        else:
            out["roofline"]["traffic_note"] = why
    if not args.no_solve:
        out["solve"] = solve_section(patch, C, m, r_batch)
    if not args.no_serve:
        # the serving stream: 2000 Poisson-arrival requests per mode at 200 requests/s
        out["serve"] = serve_section(2000, args.qps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
