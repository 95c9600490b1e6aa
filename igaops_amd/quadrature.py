"""Host-side construction of the 1-D factors of the operators: B-spline basis
matrices and Gauss / weighted-quadrature (WQ) rules.

This is set-up code (it runs once per degree / knot vector, on the CPU) and keeps
the reference's public classes and attributes so that the operator, FD and solver
layers can be fed by objects built here or by the reference's own objects:

* ``StandardGauss`` (``quadrature/gauss_quad.py:14-126``): ``basis = [B0, B1]``
  (``n_q x n_c`` CSR), ``weights = [w0, w0, w1, w1]`` with ``w_k = B_k^T diag(omega)``.
* ``WeightedQuadrature`` (``quadrature/weighted_quad.py:238-363``): quadrature points
  from the S1-S4 layout + greedy/MILP point count (``quadrature/utils/wq_helpers.py``),
  weights by a minimum weighted-norm least-squares solve per test function
  (``quadrature/utils/operations.py:107-143``); ``weights = [W0, W0, W1, W1]`` for
  type "2" and ``[W00, W01, W10, W11]`` for type "1" (``weighted_quad.py:355``).

Basis functions and derivatives are evaluated with the Cox-de Boor recurrences of
The NURBS Book (Piegl & Tiller, algorithms A2.1 FindSpan and A2.3 DersBasisFuns),
vectorised over evaluation points; the reference delegates this to geomdl
(``quadrature/utils/operations.py:52-105``), which implements the same algorithms.
"""
from __future__ import annotations

import logging
from typing import Callable, List, Literal, Optional, Sequence, Tuple

import numpy as np
from scipy import sparse as sp
from scipy.optimize import Bounds, LinearConstraint, milp

from .common import validate_entry

logger = logging.getLogger(__name__)

# ---------------------------------------------------------------------------
# Knot vectors and basis evaluation
# ---------------------------------------------------------------------------


def make_open_knotvector(degree: int, nbel: int, multiplicity: int = 1) -> np.ndarray:
    """Open uniform knot vector on [0, 1] (``geometry/mixin/operations.py:8-30``)."""
    inner = np.linspace(0.0, 1.0, nbel + 1)[1:-1]
    return np.concatenate((np.zeros(degree + 1), np.repeat(inner, multiplicity), np.ones(degree + 1)))


def evaluate_greville(degree: int, knotvector: np.ndarray) -> np.ndarray:
    """Greville abscissae (``geometry/mixin/operations.py:54-71``)."""
    kv = np.asarray(knotvector, dtype=float)
    n = len(kv) - degree - 1
    if degree == 0:
        return 0.5 * (kv[:n] + kv[1:n + 1])
    return np.array([kv[i + 1:i + degree + 1].sum() / degree for i in range(n)])


def find_multiplicity(knot: float, knotvector: np.ndarray, tol: float = 1e-12) -> int:
    return int(np.sum(np.abs(np.asarray(knotvector, dtype=float) - knot) <= tol))


def find_spans(degree: int, knotvector: np.ndarray, nbctrlpts: int, knots: np.ndarray) -> np.ndarray:
    """Knot-span index of each evaluation point (A2.1, right-continuous).

    The last interval is closed, so ``knot == knotvector[-1]`` maps to ``nbctrlpts-1``.
    """
    kv = np.asarray(knotvector, dtype=float)
    span = np.searchsorted(kv, np.asarray(knots, dtype=float), side="right") - 1
    return np.clip(span, degree, nbctrlpts - 1)


def ders_basis_funs(
    degree: int, knotvector: np.ndarray, spans: np.ndarray, knots: np.ndarray, nders: int
) -> np.ndarray:
    """Non-zero basis functions and derivatives (A2.3), vectorised over points.

    Returns ``ders[k, j, pt]`` = k-th derivative of ``N_{span-p+j}`` at point ``pt``.
    """
    p = degree
    u = np.asarray(knots, dtype=float)
    kv = np.asarray(knotvector, dtype=float)
    npt = u.shape[0]
    left = np.zeros((p + 1, npt))
    right = np.zeros((p + 1, npt))
    ndu = np.zeros((p + 1, p + 1, npt))
    ndu[0, 0] = 1.0
    for j in range(1, p + 1):
        left[j] = u - kv[spans + 1 - j]
        right[j] = kv[spans + j] - u
        saved = np.zeros(npt)
        for r in range(j):
            ndu[j, r] = right[r + 1] + left[j - r]
            temp = ndu[r, j - 1] / ndu[j, r]
            ndu[r, j] = saved + right[r + 1] * temp
            saved = left[j - r] * temp
        ndu[j, j] = saved
    ders = np.zeros((nders + 1, p + 1, npt))
    ders[0] = ndu[:, p]
    a = np.zeros((2, p + 1, npt))
    for r in range(p + 1):
        s1, s2 = 0, 1
        a[:] = 0.0
        a[0, 0] = 1.0
        for k in range(1, nders + 1):
            d = np.zeros(npt)
            rk, pk = r - k, p - k
            if r >= k:
                a[s2, 0] = a[s1, 0] / ndu[pk + 1, rk]
                d = a[s2, 0] * ndu[rk, pk]
            j1 = 1 if rk >= -1 else -rk
            j2 = k - 1 if r - 1 <= pk else p - r
            for j in range(j1, j2 + 1):
                a[s2, j] = (a[s1, j] - a[s1, j - 1]) / ndu[pk + 1, rk + j]
                d = d + a[s2, j] * ndu[rk + j, pk]
            if r <= pk:
                a[s2, k] = -a[s1, k - 1] / ndu[pk + 1, r]
                d = d + a[s2, k] * ndu[r, pk]
            ders[k, r] = d
            s1, s2 = s2, s1
    fac = float(p)
    for k in range(1, nders + 1):
        ders[k] *= fac
        fac *= p - k
    return ders


def eval_ders_basis_sparse(
    degree: int, knotvector: np.ndarray, knots: np.ndarray, nders: int = 1
) -> List[sp.csr_array]:
    """Basis matrices ``[B_0, ..., B_nders]``, each ``(len(knots) x n_c)`` CSR with
    explicit zeros removed (``quadrature/utils/operations.py:52-105``)."""
    kv = np.asarray(knotvector, dtype=float)
    knots = np.asarray(knots, dtype=float)
    n = len(kv) - degree - 1
    spans = find_spans(degree, kv, n, knots)
    ders = ders_basis_funs(degree, kv, spans, knots, nders)
    rows = np.repeat(np.arange(len(knots)), degree + 1)
    cols = (spans[:, None] - degree + np.arange(degree + 1)[None, :]).ravel()
    out = []
    for k in range(nders + 1):
        vals = ders[k].T.ravel()
        b = sp.coo_array((vals, (rows, cols)), shape=(len(knots), n))
        b = sp.csr_array(b)
        b.sum_duplicates()
        b.eliminate_zeros()
        out.append(b)
    return out


def increase_multiplicity_to_knotvector(repeat: int, degree: int, knotvector: np.ndarray) -> np.ndarray:
    """Raise every interior knot multiplicity by ``repeat``
    (``quadrature/utils/operations.py:12-50``)."""
    kv = np.asarray(knotvector, dtype=float)
    uniq = np.unique(kv)
    out: List[float] = []
    for i, knot in enumerate(uniq):
        extra = 0 if i in (0, len(uniq) - 1) else repeat
        m = find_multiplicity(knot, kv) + extra
        if m > degree + 1:
            raise Warning("Introduces numerical instability.")
        out.extend([knot] * m)
    return np.sort(np.asarray(out))


def lobatto(n: int) -> Tuple[np.ndarray, np.ndarray]:
    """Gauss-Lobatto nodes/weights on [-1, 1] (``quadrature/gauss_quad.py:129-180``):
    end points plus the roots of P'_{n-1}, weights ``2 / (n (n-1) P_{n-1}(x)^2)``."""
    if n < 2:
        raise ValueError("n must be >= 2")
    if n == 2:
        return np.array([-1.0, 1.0]), np.array([1.0, 1.0])
    pn1 = np.polynomial.legendre.Legendre.basis(n - 1)
    roots = np.sort(np.real(pn1.deriv().roots()))
    x = np.concatenate(([-1.0], roots, [1.0]))
    w = 2.0 / (n * (n - 1) * pn1(x) ** 2)
    return x, w


# ---------------------------------------------------------------------------
# Quadrature classes
# ---------------------------------------------------------------------------


class _QuadratureBase:
    """Attributes shared by the quadrature rules (``quadrature/utils/template.py:11-155``)."""

    def __init__(self, degree: int, knotvector: np.ndarray):
        self._degree = int(degree)
        self._knotvector = np.asarray(knotvector, dtype=float)
        self._parametric_space = (0.0, 1.0)
        self._unique_kv = np.unique(np.clip(self._knotvector, 0.0, 1.0))
        self._nbctrlpts = len(self._knotvector) - self._degree - 1
        self._quadpts = np.array([])
        self._basis: List[sp.csr_array] = []
        self._weights: List[sp.csr_array] = []

    degree = property(lambda self: self._degree)
    knotvector = property(lambda self: self._knotvector)
    parametric_space = property(lambda self: self._parametric_space)
    unique_kv = property(lambda self: self._unique_kv)
    nbctrlpts = property(lambda self: self._nbctrlpts)
    quadpts = property(lambda self: self._quadpts)
    basis = property(lambda self: self._basis)
    weights = property(lambda self: self._weights)

    @property
    def nbquadpts(self) -> int:
        return len(self._quadpts)

    @property
    def max_h_size(self) -> float:
        return float(np.max(np.abs(np.diff(self._unique_kv))))

    def eval_minimal_continuity(self) -> int:
        cont = self._degree - 1
        for knot in self._unique_kv[1:-1]:
            cont = min(cont, self._degree - find_multiplicity(knot, self._knotvector))
        return cont

    def eval_basis(self, knots_to_interp: np.ndarray, nders: int = 1) -> List[sp.csr_array]:
        return eval_ders_basis_sparse(self._degree, self._knotvector, knots_to_interp, nders=nders)


class StandardGauss(_QuadratureBase):
    """Gauss-Legendre / Gauss-Lobatto rule on every knot span (``gauss_quad.py:14-126``)."""

    def __init__(
        self,
        degree: int,
        knotvector: np.ndarray,
        quadtype: Literal["legendre", "lobatto"],
        nbptsperel: Optional[int] = None,
        nders: Optional[int] = None,
    ):
        super().__init__(degree, knotvector)
        self._quadrature_type = validate_entry(quadtype, ["legendre", "lobatto"])
        self._nders = int(nders) if nders is not None else 1
        default = self._degree + (1 if self._quadrature_type == "legendre" else 2)
        n = int(nbptsperel) if nbptsperel is not None else default
        table = np.polynomial.legendre.leggauss if self._quadrature_type == "legendre" else lobatto
        self._iso_pos, self._iso_w = table(n)
        pts, wts = self.interpolate_points_and_weights(self._unique_kv)
        self._quadpts = pts
        self._parametric_weights = wts
        basis = self.eval_basis(self._quadpts, nders=self._nders)
        weights: List[sp.csr_array] = []
        for b in basis:
            w = sp.csr_array(b.T @ sp.diags_array(self._parametric_weights))
            weights.extend([w, w])
        self._basis = basis
        self._weights = weights

    quadrature_class = property(lambda self: "GAUSS")
    quadrature_type = property(lambda self: self._quadrature_type)
    parametric_weights = property(lambda self: self._parametric_weights)

    @property
    def nbptsperel(self) -> int:
        return len(self._iso_pos)

    def interpolate_points_and_weights(self, knots: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
        knots = np.unique(knots)
        a, b = knots[:-1, None], knots[1:, None]
        pts = (0.5 * ((b - a) * self._iso_pos[None, :] + a + b)).ravel()
        wts = (0.5 * (b - a) * self._iso_w[None, :]).ravel()
        return pts, wts


# --- WQ point layout (quadrature/utils/wq_helpers.py:11-259) ----------------


def _element_supports(kv: np.ndarray, degree: int, breaks: np.ndarray) -> List[Tuple[int, int]]:
    """Elements covered (with positive measure) by each B-spline (``wq_helpers.py:90-116``)."""
    n = len(kv) - degree - 1
    lo_el, hi_el = breaks[:-1], breaks[1:]
    out = []
    for i in range(n):
        a, b = kv[i], kv[i + degree + 1]
        cover = np.nonzero(np.minimum(b, hi_el) > np.maximum(a, lo_el))[0]
        out.append((int(cover.min()), int(cover.max())))
    return out


def compute_quadrature_layout(
    knot_vector_test: Sequence[float],
    knot_vector_target: Sequence[float],
    degree_test: int,
    degree_target: int,
) -> np.ndarray:
    """Upper-triangular table ``S4[r, s]`` of the minimum number of points needed in
    elements ``r..s`` (steps S1-S4 of ``wq_helpers.py:11-82``).

    For test function i, ``mu_i[r, s]`` counts the target functions whose common
    elements with it are exactly ``r..s``, and the number of targets confined to
    ``r..s`` is ``sum_{r <= k <= l <= s} mu_i[k, l]``: since ``mu_i`` vanishes below the
    diagonal this is a two-dimensional prefix sum (a reversed cumulative sum over k, a
    cumulative sum over l), computed for all test functions at once.  ``S4`` is its
    maximum over the test functions."""
    kv_test = np.asarray(knot_vector_test, dtype=float)
    kv_target = np.asarray(knot_vector_target, dtype=float)
    breaks = np.unique(kv_test)
    nel = len(breaks) - 1
    sup_test = np.asarray(_element_supports(kv_test, degree_test, breaks), dtype=np.int64)
    sup_target = np.asarray(_element_supports(kv_target, degree_target, breaks), dtype=np.int64)
    r = np.maximum(sup_test[:, None, 0], sup_target[None, :, 0])
    s_ = np.minimum(sup_test[:, None, 1], sup_target[None, :, 1])
    i_idx, j_idx = np.nonzero(r <= s_)
    mu = np.zeros((len(sup_test), nel, nel), dtype=np.int64)
    np.add.at(mu, (i_idx, r[i_idx, j_idx], s_[i_idx, j_idx]), 1)
    mb = np.flip(np.cumsum(np.flip(mu, axis=1), axis=1), axis=1).cumsum(axis=2)
    return mb.max(axis=0)


def convert_to_greedy_input(S4: np.ndarray):
    """Lower bounds and covering constraints of the point-count problem (``wq_helpers.py:200-218``)."""
    nel = S4.shape[0]
    lb = np.diag(S4).copy()
    rows, rhs = [], []
    for d in range(1, nel):
        for r in range(nel - d):
            s = r + d
            row = np.zeros(nel, dtype=np.int64)
            row[r:s + 1] = 1
            rows.append(row)
            rhs.append(S4[r, s])
    A = np.array(rows, dtype=np.int64).reshape(len(rows), nel)
    return lb, A, np.array(rhs, dtype=np.int64)


def greedy_min_norm1(lb: np.ndarray, A: np.ndarray, b: np.ndarray) -> np.ndarray:
    """min sum(x) s.t. x >= lb, A x >= b, x integer (``wq_helpers.py:221-259``):
    ``lb`` itself when feasible, otherwise scipy's MILP."""
    lb = np.asarray(lb, dtype=float)
    A = np.asarray(A, dtype=float)
    b = np.asarray(b, dtype=float)
    if A.shape[0] == 0 or np.all(A @ lb - b >= 0):
        return np.round(lb).astype(int)
    res = milp(
        c=np.ones(lb.shape[0]),
        constraints=LinearConstraint(A, lb=b, ub=np.inf),
        integrality=np.ones(lb.shape[0]),
        bounds=Bounds(lb=lb, ub=np.inf),
    )
    if res.success:
        return np.round(np.asarray(res.x)).astype(int)
    raise ValueError(f"Solver failed {res.message}")


def _target_degree_and_knots(degree: int, knotvector: np.ndarray, method: str):
    """Target space of the WQ exactness conditions (``weighted_quad.py:76-93``)."""
    if method == "1":
        return degree - 1, np.asarray(knotvector, dtype=float)[1:-1]
    if method == "2":
        return degree, increase_multiplicity_to_knotvector(1, degree, knotvector)
    raise ValueError("Unknown method for weighted quadrature.")


def eval_nbquadpts_per_element(degree: int, knotvector: np.ndarray, quadrature_type: str) -> np.ndarray:
    """Points per element (``weighted_quad.py:207-235``), including the reference's
    ``+1`` stability margin before the covering check.  The check ``A lb >= b`` of
    ``greedy_min_norm1`` is evaluated from cumulative sums (the sum of ``lb`` over
    elements ``r..s`` against ``S4[r, s]``); the constraint matrix is built only when
    it fails and the MILP is needed."""
    deg_t, kv_t = _target_degree_and_knots(degree, knotvector, quadrature_type)
    S4 = compute_quadrature_layout(knotvector, kv_t, degree, deg_t)
    lb = np.diag(S4).copy() + 1
    c = np.concatenate([[0], np.cumsum(lb)])
    span = c[None, 1:] - c[:-1, None]  # span[r, s] = sum(lb[r..s]) for r <= s
    if np.all(np.triu(span - S4, 1) >= 0):
        return lb.astype(int)
    _, A, b = convert_to_greedy_input(S4)
    return greedy_min_norm1(lb, A, b).astype(int)


def _points_on_spans(breaks: np.ndarray, n_list: np.ndarray, rule: Callable) -> np.ndarray:
    """Concatenate per-span point sets and drop duplicates (``weighted_quad.py:49-71``)."""
    pts: List[float] = []
    nsp = len(breaks) - 1
    for i in range(nsp):
        pts.extend(rule(breaks[i], breaks[i + 1], int(n_list[i])))
    return np.sort(np.unique(pts))


def endpoint_rule(knots: np.ndarray, n_list: np.ndarray) -> np.ndarray:
    """Equispaced points including both span ends (``weighted_quad.py:31-35``)."""
    return _points_on_spans(knots, np.asarray(n_list) + 1, np.linspace)


def cellcenter_rule(knots: np.ndarray, n_list: np.ndarray) -> np.ndarray:
    """Cell-centred points (``weighted_quad.py:37-47``)."""

    def rule(a, b, n):
        t = (2.0 * np.arange(1, n + 1) - 1.0) / (2.0 * n)
        return (1.0 - t) * a + t * b

    return _points_on_spans(knots, n_list, rule)


def _knot_support(quadpts: np.ndarray, unique_kv: np.ndarray) -> np.ndarray:
    """Length attributed to each point: distance between midpoints to its
    neighbours, with mirrored ghosts at both ends (``weighted_quad.py:133-138``)."""
    ext = np.concatenate([[2 * unique_kv[0] - quadpts[0]], quadpts, [2 * unique_kv[-1] - quadpts[-1]]])
    mid = 0.5 * (ext[:-1] + ext[1:])
    return np.diff(mid)


def solve_min_weighted_norm(Z: np.ndarray, A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """Row-wise ``min ||diag(z_i)^{-1} w_i||`` s.t. ``A w_i = b_i``, i.e.
    ``w_i = z_i * u_i`` with ``u_i`` the minimum-norm solution of ``A diag(z_i) u = b_i``
    (the reference runs ``lsqr`` on it, ``quadrature/utils/operations.py:107-143``).

    Only the columns where ``z_i`` is nonzero (a contiguous run: the support of test
    function i) and the rows of ``A`` that touch them (contiguous: the overlapping
    target functions) take part; the other rows are ``0 = b`` and the other entries of
    ``u_i`` are zero in the minimum-norm solution.  The small dense systems are solved
    with ``numpy.linalg.pinv``, batched over rows of equal block shape."""
    n, m = Z.shape
    out = np.zeros_like(Z)
    nzA = A != 0
    groups = {}
    for i in range(n):
        cols = np.flatnonzero(Z[i])
        if cols.size == 0:
            continue
        c0, c1 = int(cols[0]), int(cols[-1]) + 1
        touch = np.flatnonzero(nzA[:, c0:c1].any(axis=1))
        if touch.size == 0:
            continue
        r0, r1 = int(touch[0]), int(touch[-1]) + 1
        groups.setdefault((r1 - r0, c1 - c0), []).append((i, r0, c0))
    for (nr, nc), items in groups.items():
        idx = np.array([it[0] for it in items])
        r0 = np.array([it[1] for it in items])[:, None] + np.arange(nr)
        c0 = np.array([it[2] for it in items])[:, None] + np.arange(nc)
        z = Z[idx[:, None], c0]
        M = A[r0[:, :, None], c0[:, None, :]] * z[:, None, :]
        u = np.einsum("gij,gj->gi", np.linalg.pinv(M), B[idx[:, None], r0])
        out[idx[:, None], c0] = z * u
    return out


def derivative_weights(p: int, kv: np.ndarray, Q: np.ndarray) -> np.ndarray:
    """Test-derivative weights from degree ``p-1`` value weights via
    ``N'_{i,p} = p/(k_{i+p}-k_i) N_{i,p-1} - p/(k_{i+p+1}-k_{i+1}) N_{i+1,p-1}``
    (``weighted_quad.py:189-205``)."""
    n = len(kv) - p - 1
    out = np.zeros((n, Q.shape[1]))
    for i in range(n):
        d1 = kv[i + p] - kv[i]
        if abs(d1) > 0 and 0 <= i < len(Q):
            out[i] += (p / d1) * Q[i]
        d2 = kv[i + p + 1] - kv[i + 1]
        if abs(d2) > 0 and 0 <= i + 1 < len(Q):
            out[i] -= (p / d2) * Q[i + 1]
    return out


class WeightedQuadrature(_QuadratureBase):
    """Weighted quadrature rule (``quadrature/weighted_quad.py:238-363``)."""

    def __init__(
        self,
        degree: int,
        knotvector: np.ndarray,
        quadtype: Literal["1", "2"],
        position_rule: Literal["endpoint", "cellcenter"] = "endpoint",
        eval_nbpts_per_elem: Optional[Callable] = None,
    ):
        super().__init__(degree, knotvector)
        self._quadrature_type = validate_entry(quadtype, ["1", "2"])
        self._position_rule = validate_entry(position_rule, ["endpoint", "cellcenter"])
        self._use_fallback = self._degree <= 1 or len(self._unique_kv) <= 2
        if self._use_fallback:
            logger.warning("Weighted quadrature needs degree >= 2 and two elements; Gauss is used.")
        if self.eval_minimal_continuity() < 0:
            raise RuntimeError("The knotvector is not regular enough.")
        self._eval_nbpts = eval_nbpts_per_elem if callable(eval_nbpts_per_elem) else eval_nbquadpts_per_element
        self._generate()

    quadrature_class = property(lambda self: "WEIGHTED")
    quadrature_type = property(lambda self: self._quadrature_type)
    position_rule = property(lambda self: self._position_rule)

    def _generate(self):
        if self._use_fallback:
            q = StandardGauss(self._degree, self._knotvector, quadtype="legendre")
            self._quadpts, self._weights, self._basis = q.quadpts, q.weights, q.basis
            return
        n_list = self._eval_nbpts(self._degree, self._knotvector, self._quadrature_type)
        rule = endpoint_rule if self._position_rule == "endpoint" else cellcenter_rule
        self._quadpts = rule(self._unique_kv, n_list)
        self._set_basis_and_weights()

    def _w0_weights(self, test: StandardGauss, tg_b_gauss, tg_b_wq, te_b_gauss, te_b_wq):
        """Value weights of the test space ``test`` at the WQ points
        (``weighted_quad.py:140-187``)."""
        basis = test.eval_basis(self._quadpts, nders=1)
        support = _knot_support(self._quadpts, test.unique_kv)
        Z = (basis[0].T @ sp.diags_array(support)).toarray()
        w0_test = test.weights[0]
        out = []
        if self._quadrature_type == "1":
            out.append(solve_min_weighted_norm(Z, te_b_wq, (w0_test @ te_b_gauss).toarray()))
        out.append(solve_min_weighted_norm(Z, tg_b_wq, (w0_test @ tg_b_gauss).toarray()))
        return basis, out

    def _set_basis_and_weights(self):
        p, kv, qt = self._degree, self._knotvector, self._quadrature_type
        gauss_test = StandardGauss(p, kv, quadtype="legendre")
        deg_t, kv_t = _target_degree_and_knots(p, kv, qt)
        gauss_target = StandardGauss(deg_t, kv_t, quadtype="legendre", nbptsperel=gauss_test.nbptsperel)
        # target / test bases at the Gauss points of the test space and at the WQ points
        tg_b_gauss = gauss_target.eval_basis(gauss_test.quadpts)[0]
        tg_b_wq = gauss_target.eval_basis(self._quadpts)[0].T.toarray()
        te_b_gauss = gauss_test.basis[0]
        te_b_wq = gauss_test.eval_basis(self._quadpts)[0].T.toarray()

        basis, weights = self._w0_weights(gauss_test, tg_b_gauss, tg_b_wq, te_b_gauss, te_b_wq)
        # derivative weights are built from the value weights of S^{p-1}_{r-1}
        dm1, kvm1 = _target_degree_and_knots(p, kv, "1")
        gauss_m1 = StandardGauss(dm1, kvm1, quadtype="legendre", nbptsperel=gauss_test.nbptsperel)
        _, weights_m1 = self._w0_weights(gauss_m1, tg_b_gauss, tg_b_wq, te_b_gauss, te_b_wq)
        zeros = np.zeros((1, self.nbquadpts))
        for w in weights_m1:
            weights.append(derivative_weights(p, kv, np.vstack([zeros, w, zeros])))
        order = {"1": [0, 1, 2, 3], "2": [0, 0, 1, 1]}[qt]
        csr = []
        for idx in order:
            w = sp.csr_array(weights[idx])
            w.eliminate_zeros()
            csr.append(w)
        self._basis = basis
        self._weights = csr
