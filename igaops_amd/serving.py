"""A persistent solve server on one GPU: continuous admission of Poisson/diffusion
solves of mixed degree, mesh size, geometry and coefficient field.

Each request is ``K(kappa) x = b`` on a unit cube or an extruded quarter annulus
(degree ``p``, ``nel`` elements per direction, WQ type "2" rules), homogeneous
Dirichlet on every face, ``kappa`` lognormal (``sigma``, ``seed``), solved to a true
relative residual ``tol`` by FD-preconditioned GMRES, either in float64
(``"fp64"``) or as float64 iterative refinement around float32 inner solves
(``"mixed"``).

How the loop works:

* Requests become visible at their arrival time.  One is admitted when a stream
  slot is free and its HBM footprint (packed coefficients, the set-up temporaries,
  the Krylov bases of its solver) fits in what the active requests leave of the
  budget; otherwise it waits.  Admission is in arrival order.
* Everything a request touches lives in HBM for its lifetime: its coefficients, its
  right-hand side and solution, its Krylov basis and least-squares state.  What
  depends only on the shape (geometry, quadrature, FD eigenvectors, fused-operator
  tables) is built once per shape and shared read-only.
* Every active request owns a stream and private scratch buffers.  The loop visits
  the active requests in turn and queues whatever each can do without waiting
  (the next Arnoldi steps, a refinement residual, the solution update), so the
  work of many requests is in flight on the device at once.  Same-shape requests
  share plans; when several of them are ready for an Arnoldi step in the same
  pass, their FD preconditioner applications go out as one multi-right-hand-side
  launch (the FD depends on the shape only: parametric pencils with per-direction
  correctors from the mean geometric conductivity) and their operators as one
  grouped launch of the shared plan with each request's coefficients, on a batch
  stream, ordered by events.
  Nothing a request computes depends on which others are present, so its result
  is the same alone and in the mix (to rounding: the batched FD sums in the same
  order per column, and each column of the grouped operator is that request's own
  apply bit for bit).
* Latency is measured from arrival to completion (the true residual checked).  The
  loop never waits for one request's stream: set-up uploads go through pinned
  memory, residual estimates and the final true residual come back through pinned
  memory behind events, and each Arnoldi step is one library call.
"""
from __future__ import annotations

import sys
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import device as dv
from .runs import REFINEMENT_INNER_TOL

_RULES: Dict[Tuple[int, int], object] = {}
_PENCILS: Dict[Tuple[int, int], tuple] = {}
_CORRECTORS: Dict[Tuple[int, int], np.ndarray] = {}


def wq_rule(degree: int, nel: int):
    """WQ type "2" rule on the open uniform knot vector, cached by ``(degree, nel)``."""
    from .quadrature import WeightedQuadrature, make_open_knotvector

    key = (int(degree), int(nel))
    q = _RULES.get(key)
    if q is None:
        q = WeightedQuadrature(degree, make_open_knotvector(degree, nel), quadtype="2")
        _RULES[key] = q
    return q


def fd_pencil(degree: int, nel: int):
    """Free indices and 1-D eigenpairs of the FD of ``(degree, nel)`` with both ends
    Dirichlet (the Gauss rewrite of the rule: the same for every direction and for
    both geometries), cached by ``(degree, nel)``."""
    from .fastdiag import pencil_eigenpairs, rewrite_quadrature

    key = (int(degree), int(nel))
    e = _PENCILS.get(key)
    if e is None:
        e = pencil_eigenpairs(rewrite_quadrature([wq_rule(degree, nel)])[0], True, True)
        _PENCILS[key] = e
    return e


def annulus_correctors(degree: int, nel: int) -> np.ndarray:
    """Per-direction FD stiffness correctors of the quarter annulus of ``(degree,
    nel)``: the mean over the quadrature points of the diagonal of its geometric
    conductivity ``detJ J^+ J^+T`` (the scalar correctors of
    ``fastdiag/preconditioner.py:110-144``, which the reference leaves to the caller).
    The angular direction is up to three times longer than the radial one, which a
    parametric FD ignores.  Cached by ``(degree, nel)``."""
    from .geometry import quarter_annulus

    key = (int(degree), int(nel))
    c = _CORRECTORS.get(key)
    if c is None:
        plane, dz, _, _ = quarter_annulus(degree, nel, quadrule=wq_rule(*key)).extrusion_factors()
        det2, i00, i11 = plane[0], plane[1], plane[2]
        c = np.array([np.mean(dz) * np.mean(det2 * i00), np.mean(dz) * np.mean(det2 * i11),
                      np.mean(det2) * np.mean(1.0 / dz)])
        _CORRECTORS[key] = c
    return c


def _host_setup(key):
    """Worker-process body of :func:`prepare_shape_table`."""
    return key, wq_rule(*key), fd_pencil(*key), annulus_correctors(*key)


def _worker_init():
    """One BLAS thread per worker: the pool already fills the cores, and the dense
    problems here are far too small for threaded BLAS.  The limit reaches only the
    BLAS libraries already loaded, so scipy's LAPACK is loaded first."""
    import scipy.linalg  # noqa: F401
    import scipy.sparse.linalg  # noqa: F401
    from threadpoolctl import threadpool_limits

    threadpool_limits(limits=1)


def prepare_shape_table(degrees, nels, workers: int = 8) -> dict:
    """Build the rule, FD pencil and annulus FD correctors of every ``(degree, nel)``
    of a declared range (the host set-up a request of that size needs) in a pool of
    ``workers`` processes,
    largest first.  Entries already cached are skipped.  Returns the number built and
    the wall time, which a server reports as its start-up cost."""
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor

    keys = sorted({(int(p), int(n)) for p in degrees for n in nels} - (_RULES.keys() & _PENCILS.keys()
                                                                         & _CORRECTORS.keys()),
                  key=lambda k: -(k[0] + 1) * k[1])
    t0 = time.perf_counter()
    if workers > 1 and len(keys) > 1:
        # fresh interpreters: the workers never touch the device
        ctx = multiprocessing.get_context("spawn")
        with ProcessPoolExecutor(max_workers=min(workers, len(keys)), mp_context=ctx,
                                 initializer=_worker_init) as ex:
            for key, q, e, c in ex.map(_host_setup, keys, chunksize=2):
                _RULES[key], _PENCILS[key], _CORRECTORS[key] = q, e, c
    else:
        for key in keys:
            _host_setup(key)
    return {"entries": len(keys), "workers": workers, "seconds": time.perf_counter() - t0}


@dataclass
class SolveRequest:
    rid: int
    geometry: str  # "cube" or "annulus"
    degree: int
    nel: int
    sigma: float
    seed: int
    arrival: float = 0.0  # seconds after the start of the stream

    @property
    def shape(self):
        return (self.geometry, self.degree, self.nel)


@dataclass
class SolveResult:
    rid: int
    shape: tuple
    ndof: int
    iterations: int
    true_rel_residual: float
    success: bool
    arrival: float
    admitted: float
    completed: float
    solution: object = None
    reason: str = ""

    @property
    def latency(self) -> float:
        return self.completed - self.arrival


class ShapeResources:
    """Everything a request of one (geometry, degree, nel) shares with its peers."""

    def __init__(self, geometry: str, degree: int, nel: int, fd: "FDResources"):
        from .geometry import quarter_annulus, unit_cube

        q = wq_rule(degree, nel)
        build = unit_cube if geometry == "cube" else quarter_annulus
        self.patch = build(degree, nel, quadrule=q)
        self.qs = self.patch.quadrule_list
        self.ndof = self.patch.nbctrlpts_total
        self.nq = self.patch.nbquadpts_total
        self.fd = fd
        self.space_fd, self.mask, self.fd_plans = fd.space_fd, fd.mask, fd.plans
        t = dv.torch()
        self.w64 = dv.to_device(self.patch.nurbs_weights, "f64") if self.patch.is_nurbs else None
        # the weights in each operator precision, shared by the requests of this shape
        # (a grouped operator launch takes one weight vector for all its members)
        self.w_by_dtype = {"f64": self.w64, "f32": None if self.w64 is None else self.w64.float()}
        # cross-section and extrusion factors (and the NURBS weight function on the
        # cross-section grid), on the device
        self.patch.device_extrusion()
        # requests read these from their own streams, which wait for this event
        self.ready = t.cuda.Event()
        self.ready.record(t.cuda.current_stream())


class FDResources:
    """The FD preconditioner of one ``(geometry, degree, nel)``: the parametric Gauss
    pencils of the rules (every face Dirichlet) with per-direction stiffness
    correctors, 1 for the unit cube and :func:`annulus_correctors` for the annulus."""

    def __init__(self, geometry: str, degree: int, nel: int, dtypes=("f64", "f32")):
        from .fastdiag import SpaceFD
        from .solvers import Mask

        q = wq_rule(degree, nel)
        pencil = fd_pencil(degree, nel)
        fd = SpaceFD()
        fd.set_space_eigenpairs(np.ones((1, 3, 2), bool), [q.nbctrlpts] * 3, [pencil] * 3)
        if geometry == "annulus":
            fd.add_scalar_space_correctors(stiffness_corrector=[annulus_correctors(degree, nel)])
        fd.update_space_eigenvalues([0.0, 1.0])
        self.space_fd = fd
        self.mask = Mask(box=fd.box())
        self.plans = {dt: fd.plan(dt) for dt in dtypes}
        dv.torch().cuda.current_stream().synchronize()


def footprint_bytes(ndof: int, nq: int, nurbs: bool, mode: str, maxiters: int, inner_maxiters: int) -> int:
    """Upper estimate of the HBM a request holds at its peak."""
    ncoef = 10 if nurbs else 6
    setup = 18 * nq * 8  # C (3,3,N_q), the field, NURBS weight fields: released after packing
    coef = ncoef * nq * 8 + (ncoef * nq * 4 if mode == "mixed" else 0)
    # one basis of m + 2 vectors and the dense GMRES state (two (m+1)^2 matrices)
    if mode == "mixed":
        krylov = (inner_maxiters + 2) * ndof * 4 + 2 * (inner_maxiters + 1) ** 2 * 8 + 6 * ndof * 8
    else:
        krylov = (maxiters + 2) * ndof * 8 + 2 * (maxiters + 1) ** 2 * 8 + 4 * ndof * 8
    scratch = 12 * ndof * 8
    return int(setup + coef + krylov + scratch)


class _Active:
    """One admitted request: its stream, operators, scratch and solver run."""

    def __init__(self, req: SolveRequest, shape: ShapeResources, mode: str, stream, tol: float, maxiters: int,
                 inner_tol: float, inner_maxiters: int, lookahead: int, admitted: float, host_buf=None):
        from .operators import StiffnessMassOperator
        from .runs import GmresRun, NativeStep, RefinementRun

        t = dv.torch()
        self.req, self.shape, self.mode, self.stream = req, shape, mode, stream
        self.admitted = admitted
        self.setup_times = {}
        tic = time.perf_counter()

        def lap(name):
            nonlocal tic
            now = time.perf_counter()
            self.setup_times[name] = now - tic
            tic = now

        stream.wait_event(shape.ready)
        with t.cuda.stream(stream):
            # the lognormal field as separable modes; the coefficients are generated
            # at the points and packed in one launch per precision
            field = shape.patch.device_field_modes(req.sigma, req.seed)
            lap("coefficients")
            kw = dict(nurbs_w64=shape.w64)
            self.A64 = StiffnessMassOperator.from_extruded(shape.patch, field, "f64", nurbs_w=shape.w_by_dtype["f64"],
                                                           **kw)
            self.A32 = StiffnessMassOperator.from_extruded(shape.patch, field, "f32", nurbs_w=shape.w_by_dtype["f32"],
                                                           **kw) if mode == "mixed" else None
            # the operators' partial blocks come from this request's stream pool (the
            # inner solves of the mixed mode are built outside any stream context)
            for op in (self.A64, self.A32):
                if op is not None:
                    op.workspace(1)
            lap("operators")
            del field
            gen = t.Generator(device=dv.device())
            gen.manual_seed(int(req.seed) + 7919)
            b = t.randn(shape.ndof, generator=gen, dtype=t.float64, device=dv.device())
            shape.mask(b)
            self.b = b
            dt = "f32" if mode == "mixed" else "f64"
            plan = shape.fd_plans[dt]
            fd_work = plan.workspace(1)
            self._fd_work = fd_work
            # the closures capture the operators, not this object, so that a finished
            # request holds no reference cycle and its HBM is released at once
            op64, op32 = self.A64, self.A32
            P = lambda x, o: plan.apply(x, o, 1, None, work=fd_work)
            A64 = lambda x, o: op64.apply(x, out=o)
            # each Arnoldi step is one library call (FD, operator, mask, orthogonalisation)
            native = NativeStep(op32 if mode == "mixed" else op64, plan, fd_work, shape.mask.box)
            if mode == "mixed":
                A32 = lambda x, o: op32.apply(x, out=o)
                self.run = RefinementRun(A64, A32, P, b, tol=tol, inner_tol=inner_tol,
                                         inner_maxiters=inner_maxiters, mask=shape.mask, stream=stream,
                                         lookahead=lookahead, host_buf=host_buf, native_lo=native)
            else:
                self.run = GmresRun(A64, P, b, tol, maxiters, mask=shape.mask, stream=stream, lookahead=lookahead,
                                    host_buf=host_buf, native=native)
            lap("solver")
        self.t0 = time.perf_counter()

    def advance(self) -> bool:
        return self.run.advance()

    @property
    def done(self) -> bool:
        return self.run.done

    def queue_check(self, host2):
        """Queue the true residual ``||b - A x||^2`` and ``||b||^2`` (float64, masked)
        and their copy into ``host2`` (two pinned doubles) behind an event, so that
        the loop never waits for a request's stream."""
        from . import _lib

        t = dv.torch()
        x = self.run.result.solution
        with t.cuda.stream(self.stream):
            r = self.b - self.A64.apply(x)
            self.shape.mask(r)
            nrm = t.empty(2, dtype=t.float64, device=r.device)
            work = dv.workspace(_lib.load().iga_reduce_workspace_bytes(r.numel(), 1))
            sp = dv.stream_ptr(self.stream)
            for i, v in enumerate((r, self.b)):
                _lib.call("iga_dot_many", _lib.IGA_F64, v.numel(), 1, dv.ptr(v), v.numel(), dv.ptr(v),
                          dv.ptr(nrm[i:]), dv.ptr(work), sp)
            host2.copy_(nrm, non_blocking=True)
            self._check = (host2, nrm, work, r)
            self.check_event = t.cuda.Event()
            self.check_event.record(self.stream)

    check_event = None

    def result(self, completed: float, keep_solution: bool) -> SolveResult:
        out = self.run.result
        host2 = self._check[0]
        rel = float(host2[0]) ** 0.5 / float(host2[1]) ** 0.5
        x = out.solution
        res = SolveResult(self.req.rid, self.req.shape, self.shape.ndof, int(out.iterations), rel,
                          bool(out.success), self.req.arrival, self.admitted, completed,
                          x.clone() if keep_solution else None, str(getattr(out, "reason", "")))
        # release everything this request holds on the device
        self.run = self.A64 = self.A32 = self.b = self._fd_work = self._check = None
        return res


class SolveServer:
    """Continuous admission over a stream of requests (see the module docstring)."""

    def __init__(self, mode: str = "mixed", hbm_budget_gb: float = 200.0, max_active: int = 16, tol: float = 1e-10,
                 maxiters: int = 800, inner_tol: float = REFINEMENT_INNER_TOL, inner_maxiters: int = 300,
                 lookahead: int = 2,
                 keep_solutions: bool = False, batch_steps: bool = True):
        if mode not in ("mixed", "fp64"):
            raise ValueError("mode must be 'mixed' or 'fp64'")
        self.mode = mode
        self.budget = int(hbm_budget_gb * 1e9)
        self.max_active = int(max_active)
        self.tol, self.maxiters = float(tol), int(maxiters)
        self.inner_tol, self.inner_maxiters = float(inner_tol), int(inner_maxiters)
        self.lookahead = int(lookahead)
        self.keep_solutions = keep_solutions
        self.batch_steps = bool(batch_steps)
        self._batch_stream = None
        self._batch_work: Dict[tuple, object] = {}
        self.shapes: Dict[tuple, ShapeResources] = {}
        self.fds: Dict[tuple, FDResources] = {}
        self.stats = {"max_active": 0, "max_bytes": 0, "deferred_for_memory": 0, "admit_s": 0.0, "advance_s": 0.0,
                      "complete_s": 0.0, "shape_setup_s": 0.0, "start": None}

    @property
    def dtypes(self):
        return ("f64", "f32") if self.mode == "mixed" else ("f64",)

    def prepare(self, degrees, nels, workers: int = 8, shapes: bool = True) -> dict:
        """Server start: build what depends on ``(degree, nel)`` only, for every size
        the server accepts, so that no request waits for it: the rules and FD pencils
        (host, in a process pool), then the device FD plans and the fused-operator
        plans of both geometries' coefficient forms in the mode's precisions (one-off
        uploads that would otherwise wait behind other requests' work).  Sizes outside
        the declared ranges are still served; their set-up is then part of the first
        such request's latency.  With ``shapes`` the per-shape device resources of both
        geometries are built too (patch, device extrusion factors and NURBS weights:
        a few MB per shape); their time is reported as ``shape_seconds``."""
        from .operators import FORM_K6, FORM_SYM10, fused_plan

        # the device is opened before the worker processes start (they never use it)
        dv.require_gpu()
        info = prepare_shape_table(degrees, nels, workers)
        t0 = time.perf_counter()
        for p in degrees:
            for nel in nels:
                key = (int(p), int(nel))
                for geo in ("cube", "annulus"):
                    if (geo,) + key not in self.fds:
                        self.fds[(geo,) + key] = FDResources(geo, *key, dtypes=self.dtypes)
                q = wq_rule(*key)
                for form in (FORM_K6, FORM_SYM10):  # unit cube; quarter annulus (NURBS)
                    for dt in self.dtypes:
                        fused_plan([q, q, q], form, dt)
        dv.torch().cuda.synchronize()
        info["device_plans_seconds"] = time.perf_counter() - t0
        info["seconds"] += info["device_plans_seconds"]
        if shapes:
            t0 = time.perf_counter()
            for p in degrees:
                for nel in nels:
                    for geo in ("cube", "annulus"):
                        self.shape((geo, int(p), int(nel)))
            dv.torch().cuda.synchronize()
            info["shape_seconds"] = time.perf_counter() - t0
            info["seconds"] += info["shape_seconds"]
            self.stats["shape_setup_s"] = 0.0  # counts set-up during serving only
        self.stats["start"] = info
        return info

    def _batched_steps(self, states) -> bool:
        """One Arnoldi step for every group of same-shape requests ready for it, as one
        library call per group (``iga_arnoldi_step_group``): their preconditioner
        applications in one multi-right-hand-side FD launch and their operators in one
        grouped launch (one plan, each request's coefficients) on the batch stream,
        then their orthogonalisations (one launch per kernel for the group); events order
        the batch after each request's previous work and each request after the batch."""
        from .runs import launch_group

        groups: Dict[tuple, list] = {}
        for st in states:
            run = st.run.step_run() if st.run is not None else None
            if run is not None and run._ctx is not None:
                groups.setdefault((st.req.shape, run.dtype), []).append(run)
        done = False
        for runs in groups.values():
            for j in range(0, len(runs), 16):
                grp = runs[j:j + 16]
                if len(grp) < 2:
                    continue
                launch_group(grp, self._batch_work, self._batch_stream)
                self.stats["batched_steps"] = self.stats.get("batched_steps", 0) + len(grp)
                done = True
        return done

    def shape(self, key) -> ShapeResources:
        s = self.shapes.get(key)
        if s is None:
            t0 = time.perf_counter()
            fk = tuple(key)
            fd = self.fds.get(fk)
            if fd is None:
                fd = FDResources(*fk, dtypes=self.dtypes)
                self.fds[fk] = fd
            s = ShapeResources(*key, fd)
            self.shapes[key] = s
            self.stats["shape_setup_s"] += time.perf_counter() - t0
        return s

    def footprint(self, req: SolveRequest) -> int:
        n = (req.nel + req.degree) ** 3
        q = wq_rule(req.degree, req.nel)
        nq = q.nbquadpts ** 3
        return footprint_bytes(n, nq, req.geometry == "annulus", self.mode, self.maxiters, self.inner_maxiters)

    def serve(self, requests: List[SolveRequest], realtime: bool = True, progress_every: int = 0) -> List[SolveResult]:
        """Run the stream to completion.  With ``realtime`` the requests arrive at their
        ``arrival`` times (wall clock from the call); otherwise all are present at once.
        ``progress_every > 0`` prints a line to stderr every that many completions."""
        t = dv.torch()
        dv.require_gpu()
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        streams = [t.cuda.Stream() for _ in range(self.max_active)]
        self._batch_stream = t.cuda.Stream()
        self._batch_work = {}  # group-step scratch, allocated on this batch stream
        # one pinned host buffer per stream slot, lent to the solver runs of that slot
        from .runs import GmresRun

        hd = (GmresRun.host_doubles(self.inner_maxiters) + 2 if self.mode == "mixed"
              else GmresRun.host_doubles(self.maxiters))
        host_bufs = [t.zeros(hd, dtype=t.float64, pin_memory=True) for _ in range(self.max_active)]
        check_bufs = [t.zeros(2, dtype=t.float64, pin_memory=True) for _ in range(self.max_active)]
        free_streams = list(range(self.max_active))
        active: List[Tuple[_Active, int, int]] = []  # (state, stream index, bytes)
        used = 0
        results: List[SolveResult] = []
        start = time.perf_counter()
        now = lambda: time.perf_counter() - start
        head = 0
        while head < len(pending) or active:
            clock = now()
            # admission in arrival order
            while head < len(pending) and free_streams:
                req = pending[head]
                if realtime and req.arrival > clock:
                    break
                fp = self.footprint(req)
                if active and used + fp > self.budget:
                    self.stats["deferred_for_memory"] += 1
                    break
                si = free_streams.pop()
                ta = time.perf_counter()
                shape = self.shape(req.shape)
                st = _Active(req, shape, self.mode, streams[si], self.tol, self.maxiters, self.inner_tol,
                             self.inner_maxiters, self.lookahead, admitted=now(), host_buf=host_bufs[si])
                self.stats["admit_s"] += time.perf_counter() - ta
                for k_, v_ in st.setup_times.items():
                    self.stats["setup_" + k_ + "_s"] = self.stats.get("setup_" + k_ + "_s", 0.0) + v_
                active.append((st, si, fp))
                used += fp
                head += 1
                if not realtime:
                    req.arrival = 0.0
            self.stats["max_active"] = max(self.stats["max_active"], len(active))
            self.stats["max_bytes"] = max(self.stats["max_bytes"], used)
            progressed = False
            ta = time.perf_counter()
            if self.batch_steps and len(active) > 1:
                # only requests whose shape another active request shares can batch their
                # FD steps: they advance without launching, then the batches go out
                seen, shared = set(), set()
                for st, _, _ in active:
                    (shared if st.req.shape in seen else seen).add(st.req.shape)
                if shared:
                    cand = [st for st, _, _ in active if st.req.shape in shared]
                    for st in cand:
                        progressed |= st.run.advance(launch=False)
                    progressed |= self._batched_steps(cand)
            for st, _, _ in active:
                progressed |= st.advance()
            tb = time.perf_counter()
            self.stats["advance_s"] += tb - ta
            still = []
            for st, si, fp in active:
                if st.done and st.check_event is None:
                    st.queue_check(check_bufs[si])
                    progressed = True
                if st.done and st.check_event.query():
                    results.append(st.result(now(), self.keep_solutions))
                    if progress_every > 0 and len(results) % progress_every == 0:
                        print(f"[serve {self.mode}] {len(results)}/{len(pending)} done at {now():.1f} s, "
                              f"{len(active)} active, {len(pending) - head} waiting", file=sys.stderr, flush=True)
                    self.stats["complete_s"] += time.perf_counter() - tb
                    tb = time.perf_counter()
                    free_streams.append(si)
                    used -= fp
                else:
                    still.append((st, si, fp))
            if len(still) != len(active):
                progressed = True
            active = still
            if not progressed:
                if active:
                    # nothing to queue: wait briefly for the device
                    time.sleep(20e-6)
                elif head < len(pending):
                    time.sleep(max(0.0, min(1e-3, pending[head].arrival - now())))
        t.cuda.synchronize()
        ms = t.cuda.memory_stats()
        self.stats["alloc_retries"] = int(ms.get("num_alloc_retries", 0))
        self.stats["device_frees"] = int(ms.get("num_device_free", 0))
        self.stats["device_allocs"] = int(ms.get("num_device_alloc", 0))
        self.stats["peak_reserved_gb"] = ms.get("reserved_bytes.all.peak", 0) / 1e9
        return results


def poisson_stream(n: int, qps: float, seed: int = 0, nel_range=(16, 96), degree_range=(2, 5),
                   sigma_range=(0.2, 1.0), geometries=("cube", "annulus")) -> List[SolveRequest]:
    """Seeded request stream with exponential inter-arrival times at ``qps``."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / qps, size=n) if qps > 0 else np.zeros(n)
    arr = np.cumsum(gaps) - gaps[0]
    out = []
    for i in range(n):
        out.append(SolveRequest(
            rid=i,
            geometry=str(geometries[int(rng.integers(len(geometries)))]),
            degree=int(rng.integers(degree_range[0], degree_range[1] + 1)),
            nel=int(rng.integers(nel_range[0], nel_range[1] + 1)),
            sigma=float(rng.uniform(*sigma_range)),
            seed=int(rng.integers(1 << 30)),
            arrival=float(arr[i]),
        ))
    return out


def summarize(results: List[SolveResult], tol: float, wall: float) -> dict:
    lat = np.array([r.latency for r in results]) if results else np.zeros(1)
    good = lambda r: r.success and r.true_rel_residual <= tol * 1.0001
    ok = sum(1 for r in results if good(r))
    failures = [{"rid": r.rid, "shape": list(r.shape), "iterations": r.iterations, "reason": r.reason,
                 "true_rel_residual": r.true_rel_residual} for r in results if not good(r)][:10]
    return {
        "failures": failures,
        "requests": len(results),
        "reached_tol": ok,
        "solves_per_s": ok / wall if wall > 0 else 0.0,
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_iterations": float(np.mean([r.iterations for r in results])) if results else 0.0,
        "max_true_rel_residual": float(max((r.true_rel_residual for r in results), default=0.0)),
        "wall_s": wall,
    }
