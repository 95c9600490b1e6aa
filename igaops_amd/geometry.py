"""Host-side geometry set-up: single-patch 3-D maps and the conductivity
coefficients that feed the stiffness / mass operators.

The reference builds patches through geomdl (``geometry/patch.py:25-92``), evaluates
``J``, ``detJ = sqrt(det J^T J)`` and ``J^+ = (J^T J)^{-1} J^T`` at the quadrature points
(``geometry/mixin/template.py:85-130``, ``geometry/mixin/operations.py:137-159``), and
leaves it to the caller to form ``C = kappa detJ J^+ J^+T`` and ``m = detJ``.

Here a patch is a 2-D B-spline/NURBS surface in ``(xi_1, xi_2)`` extruded linearly
along ``xi_3`` (``z = z0 + H xi_3``, Greville control points, weights constant in
``xi_3``).  That is exactly a 3-D NURBS map with control net ``P_ijk = (x_ij, y_ij, z_k)``
and weights ``w_ijk = w_ij``; it covers the unit cube and the extruded quarter
annulus, and lets ``J`` factor into a plane part on the ``(q_1, q_2)`` grid and a line
part on ``q_3``, so the set-up never materialises per-point 3x3 matrices it does not need.

The quarter-annulus arc is not produced by degree elevation + knot insertion (geomdl,
``tests/test_assemble.py:23-50``) but by interpolating, at the Greville points of the
target spline space, the numerator and denominator of the quadratic rational arc;
both are polynomials that lie in that space, so the interpolant reproduces exactly the
control net that elevation + insertion would give.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np
from scipy import sparse as sp

from .common import kron_vectors
from .quadrature import (
    StandardGauss,
    WeightedQuadrature,
    eval_ders_basis_sparse,
    evaluate_greville,
    make_open_knotvector,
)


def _apply2(Ba: sp.csr_array, Bb: sp.csr_array, X: np.ndarray) -> np.ndarray:
    """``(Bb (x) Ba) x`` for a field stored as ``(n_b, n_a)`` (direction a fastest)."""
    T = (Ba @ X.T).T  # (n_b, q_a)
    return Bb @ T  # (q_b, q_a)


@dataclass
class PlaneGeometry:
    """Map of the 2-D cross-section on the ``(q_2, q_1)`` grid."""

    x: np.ndarray  # (2, nq2, nq1) physical coordinates
    jac: np.ndarray  # (2, 2, nq2, nq1): jac[i, j] = d x_i / d xi_j
    # (3, nq2, nq1): the NURBS weight function W and dW/dxi_1, dW/dxi_2 (None for B-splines)
    w: Optional[np.ndarray] = None


class ExtrudedPatch:
    """A 3-D single patch: 2-D surface ``(xi_1, xi_2) -> (x, y)`` times ``z = z0 + H xi_3``."""

    def __init__(
        self,
        quadrule_list: Sequence,
        ctrlpts2d: np.ndarray,
        weights2d: Optional[np.ndarray],
        z0: float = 0.0,
        height: float = 1.0,
        name: str = "patch",
    ):
        if len(quadrule_list) != 3:
            raise ValueError("ExtrudedPatch is three-dimensional")
        self.quadrule_list = list(quadrule_list)
        self.ctrlpts2d = np.asarray(ctrlpts2d, dtype=float)  # (2, n1*n2), dir 1 fastest
        self.weights2d = None if weights2d is None else np.asarray(weights2d, dtype=float)
        self.z0 = float(z0)
        self.height = float(height)
        self.name = name
        self._plane: Optional[PlaneGeometry] = None

    # --- sizes -----------------------------------------------------------
    @property
    def nbctrlpts(self) -> np.ndarray:
        return np.array([q.nbctrlpts for q in self.quadrule_list], dtype=int)

    @property
    def nbquadpts(self) -> np.ndarray:
        return np.array([q.nbquadpts for q in self.quadrule_list], dtype=int)

    @property
    def nbctrlpts_total(self) -> int:
        return int(np.prod(self.nbctrlpts))

    @property
    def nbquadpts_total(self) -> int:
        return int(np.prod(self.nbquadpts))

    @property
    def is_nurbs(self) -> bool:
        return self.weights2d is not None

    @property
    def nurbs_weights(self) -> np.ndarray:
        """Per-control-point weights ``(N_dof,)`` (empty for a B-spline patch)."""
        if self.weights2d is None:
            return np.array([])
        return np.tile(self.weights2d, int(self.nbctrlpts[2]))

    @property
    def ctrlpts(self) -> np.ndarray:
        """Cartesian control points ``(3, N_dof)`` in direction-1-fastest order."""
        n3 = int(self.nbctrlpts[2])
        q3 = self.quadrule_list[2]
        z = self.z0 + self.height * evaluate_greville(q3.degree, q3.knotvector)
        xy = np.tile(self.ctrlpts2d, (1, n3))
        zz = np.repeat(z, self.ctrlpts2d.shape[1])
        return np.vstack([xy, zz[None, :]])

    # --- evaluation ------------------------------------------------------
    def plane_geometry(self) -> PlaneGeometry:
        """Rational map and its Jacobian on the cross-section quadrature grid
        (the ``NurbsOperations.eval_jacobien`` formula, ``operators/nurbs_ops.py:136-160``:
        ``dX = (d(wP) - X dW) / W``).  Evaluated once per patch: the control net and
        the rules do not change after construction, and callers only read the arrays."""
        if self._plane is not None:
            return self._plane
        q1, q2 = self.quadrule_list[0], self.quadrule_list[1]
        n1, n2 = q1.nbctrlpts, q2.nbctrlpts
        B1 = q1.basis[:2]
        B2 = q2.basis[:2]
        w = np.ones(n1 * n2) if self.weights2d is None else self.weights2d
        Wf = w.reshape(n2, n1)
        W = _apply2(B1[0], B2[0], Wf)
        dW = [_apply2(B1[1], B2[0], Wf), _apply2(B1[0], B2[1], Wf)]
        x = np.zeros((2, q2.nbquadpts, q1.nbquadpts))
        jac = np.zeros((2, 2) + x.shape[1:])
        for i in range(2):
            Pw = (self.ctrlpts2d[i] * w).reshape(n2, n1)
            X = _apply2(B1[0], B2[0], Pw) / W
            x[i] = X
            dX = [_apply2(B1[1], B2[0], Pw), _apply2(B1[0], B2[1], Pw)]
            for j in range(2):
                jac[i, j] = (dX[j] - X * dW[j]) / W
        x.flags.writeable = False
        jac.flags.writeable = False
        wf = None
        if self.weights2d is not None:
            wf = np.stack([W, dW[0], dW[1]])
            wf.flags.writeable = False
        self._plane = PlaneGeometry(x=x, jac=jac, w=wf)
        return self._plane

    def line_geometry(self) -> Tuple[np.ndarray, np.ndarray]:
        """``z(q_3)`` and ``dz/dxi_3`` on the extrusion direction."""
        q3 = self.quadrule_list[2]
        g = self.z0 + self.height * evaluate_greville(q3.degree, q3.knotvector)
        z = q3.basis[0] @ g
        dz = q3.basis[1] @ g
        return z, dz

    def det_jac(self) -> np.ndarray:
        """``detJ`` at every quadrature point ``(N_q,)``."""
        pg = self.plane_geometry()
        _, dz = self.line_geometry()
        det2 = pg.jac[0, 0] * pg.jac[1, 1] - pg.jac[0, 1] * pg.jac[1, 0]
        if np.any(det2 * dz[0] <= 0.0):
            raise ValueError("Potential geometry issue: quadrature points with detJ <= 0")
        return (dz[:, None, None] * det2[None, :, :]).ravel()

    def jacobian(self) -> np.ndarray:
        """Full ``J (3, 3, N_q)`` (``jac[i, j] = dx_i/dxi_j``); for small problems and tests."""
        pg = self.plane_geometry()
        _, dz = self.line_geometry()
        n3 = dz.shape[0]
        shp = (n3,) + pg.jac.shape[2:]
        J = np.zeros((3, 3) + shp)
        for i in range(2):
            for j in range(2):
                J[i, j] = np.broadcast_to(pg.jac[i, j][None], shp)
        J[2, 2] = np.broadcast_to(dz[:, None, None], shp)
        return J.reshape(3, 3, -1)

    def physical_points(self) -> np.ndarray:
        """Physical coordinates of the quadrature points ``(3, N_q)``."""
        pg = self.plane_geometry()
        z, _ = self.line_geometry()
        shp = (z.shape[0],) + pg.x.shape[1:]
        out = np.empty((3,) + shp)
        out[0] = pg.x[0][None]
        out[1] = pg.x[1][None]
        out[2] = z[:, None, None]
        return out.reshape(3, -1)

    def conductivity_coefficients(self, kappa: Optional[np.ndarray] = None) -> Tuple[np.ndarray, np.ndarray]:
        """Stiffness coefficients ``C = kappa detJ J^+ J^+T`` ``(3, 3, N_q)`` and mass
        coefficients ``m = detJ`` ``(N_q,)``.

        For the extruded map ``J^+ J^+T = (J^T J)^{-1} = diag((J_2^T J_2)^{-1}, 1/z'^2)``
        and ``detJ = det(J_2) z'``."""
        pg = self.plane_geometry()
        _, dz = self.line_geometry()
        j = pg.jac
        det2 = j[0, 0] * j[1, 1] - j[0, 1] * j[1, 0]
        if np.any(det2 * dz[0] <= 0.0):
            raise ValueError("Potential geometry issue: quadrature points with detJ <= 0")
        g00 = j[0, 0] ** 2 + j[1, 0] ** 2
        g11 = j[0, 1] ** 2 + j[1, 1] ** 2
        g01 = j[0, 0] * j[0, 1] + j[1, 0] * j[1, 1]
        dG = g00 * g11 - g01 ** 2
        inv = [[g11 / dG, -g01 / dG], [-g01 / dG, g00 / dG]]
        n3 = dz.shape[0]
        plane_shape = det2.shape
        det = dz[:, None, None] * det2[None]
        C = np.zeros((3, 3, n3) + plane_shape)
        for a in range(2):
            for b in range(2):
                C[a, b] = det * inv[a][b][None]
        C[2, 2] = det * (1.0 / dz[:, None, None] ** 2)
        C = C.reshape(3, 3, -1)
        m = det.ravel()
        if kappa is not None:
            C *= np.asarray(kappa).ravel()[None, None, :]
        return C, m


    def extrusion_factors(self):
        """Host factors of the extruded map that the coefficients are built from:
        ``plane (4, n_plane)`` = ``det J_2`` and the entries 11, 22, 12 of
        ``(J_2^T J_2)^{-1}`` on the cross-section grid; ``dz (n3)`` = ``dz/dxi_3``; and
        the coordinates scaled to the bounding box of the control net (``xy (2,
        n_plane)``, ``z (n3)``) at which the lognormal field is evaluated."""
        pg = self.plane_geometry()
        z, dz = self.line_geometry()
        j = pg.jac
        det2 = j[0, 0] * j[1, 1] - j[0, 1] * j[1, 0]
        if np.any(det2 * dz[0] <= 0.0):
            raise ValueError("Potential geometry issue: quadrature points with detJ <= 0")
        g00 = j[0, 0] ** 2 + j[1, 0] ** 2
        g11 = j[0, 1] ** 2 + j[1, 1] ** 2
        g01 = j[0, 0] * j[0, 1] + j[1, 0] * j[1, 1]
        dG = g00 * g11 - g01 ** 2
        plane = np.stack([det2.ravel(), (g11 / dG).ravel(), (g00 / dG).ravel(), (-g01 / dG).ravel()])
        cp = self.ctrlpts
        lo, hi = cp.min(axis=1), cp.max(axis=1)
        span = np.where(hi > lo, hi - lo, 1.0)
        xy = np.stack([((pg.x[0] - lo[0]) / span[0]).ravel(), ((pg.x[1] - lo[1]) / span[1]).ravel()])
        return plane, np.asarray(dz, dtype=float), xy, (z - lo[2]) / span[2]

    @staticmethod
    def lognormal_modes(sigma: float, seed: int, modes: int = 8):
        """Complex mode coefficients ``coef[a, b, c]`` (wave numbers ``a+1, b+1, c+1``
        along x, y, z) of the field of :func:`lognormal_conductivity`, from the same
        random numbers."""
        rng = np.random.default_rng(seed)
        k = np.arange(1, modes + 1)
        K1, K2, K3 = np.meshgrid(k, k, k, indexing="ij")
        amp = rng.standard_normal(K1.shape) / (1.0 + K1 ** 2 + K2 ** 2 + K3 ** 2)
        phase = rng.uniform(0.0, 2 * np.pi, K1.shape)
        scale = np.sqrt(np.sum(amp ** 2) / 2.0)
        return amp * np.exp(1j * phase) / scale

    def device_extrusion(self):
        """``extrusion_factors`` on the device (float64), computed once per patch:
        ``plane (4 n_plane)``, ``dz``, ``xy (2 n_plane)``, the host ``z``, and for a
        NURBS cross-section ``nurbs (3 n_plane)`` = ``W``, ``dW/dxi_1``, ``dW/dxi_2`` on
        the cross-section grid (the weights repeat along ``xi_3``, so these are the
        weight function of the whole patch; ``None`` for B-splines)."""
        cache = getattr(self, "_dev_extrusion", None)
        if cache is None:
            from . import device as dv

            plane, dz, xy, zh = self.extrusion_factors()
            wf = self.plane_geometry().w
            parts = [plane.ravel(), dz, xy.ravel()] + ([wf.ravel()] if wf is not None else [])
            allv = dv.upload(np.concatenate(parts))
            npl = plane.shape[1]
            o = 4 * npl + dz.size
            cache = {"plane": allv[:4 * npl], "dz": allv[4 * npl:o], "xy": allv[o:o + 2 * npl], "z": zh,
                     "nplane": npl, "nurbs": allv[o + 2 * npl:] if wf is not None else None}
            self._dev_extrusion = cache
        return cache

    def device_field_modes(self, sigma: float, seed: int, modes: int = 8, stream=None):
        """The separable factors of ``sigma g`` of :func:`lognormal_conductivity` on the
        device: ``H (n_plane, modes)`` (one launch of ``iga_lognormal_plane_modes``) and
        ``Ez (n3, modes)``, complex as interleaved float64; ``g = Re sum_c Ez[q3, c]
        H[plane, c]``.  ``None`` when ``sigma`` is 0 (``kappa = 1``)."""
        if sigma == 0.0:
            return None
        from . import _lib
        from . import device as dv

        ext = self.device_extrusion()
        coef = self.lognormal_modes(sigma, seed, modes)
        k = np.arange(1, modes + 1)
        Ez = np.exp(1j * np.pi * ext["z"][:, None] * k[None, :])
        host = np.concatenate([np.ascontiguousarray(coef).view(np.float64).ravel(),
                               np.ascontiguousarray(Ez).view(np.float64).ravel()])
        t = dv.torch()
        allv = dv.upload(host)
        nc = 2 * modes ** 3
        H = t.empty(2 * ext["nplane"] * modes, dtype=t.float64, device=allv.device)
        _lib.call("iga_lognormal_plane_modes", ext["nplane"], modes, dv.ptr(ext["xy"]), dv.ptr(allv[:nc]),
                  dv.ptr(H), dv.stream_ptr(stream))
        return {"H": H, "Ez": allv[nc:], "modes": modes, "sigma": float(sigma), "_coef": allv}

    def device_coefficients(self, sigma: float = 0.0, seed: int = 0, modes: int = 8, mass: bool = True):
        """``conductivity_coefficients(lognormal_conductivity(self, sigma, seed, modes))``
        formed on the device: the cross-section and extrusion factors are evaluated on
        the host (``n_q1 n_q2`` and ``n_q3`` points), the ``N_q`` products, the Fourier
        field (one complex matrix product) and its exponential on the device.
        Returns float64 device tensors ``C (3, 3, N_q)`` and ``m (N_q,)`` (``None``
        when ``mass`` is false)."""
        import torch as t

        from . import device as dv

        ext = self.device_extrusion()
        npl = ext["nplane"]
        det2_d = ext["plane"][:npl]
        invd = {k: ext["plane"][(i + 1) * npl:(i + 2) * npl] for i, k in enumerate(((0, 0), (1, 1), (0, 1)))}
        dz_d = ext["dz"]
        det = t.outer(dz_d, det2_d)  # (n3, nplane): q3 slowest
        if sigma > 0.0:
            kdet = det * self._device_lognormal(sigma, seed, modes)
        else:
            kdet = det
        nq = det.numel()
        C = t.zeros((3, 3, nq), dtype=t.float64, device=dv.device())
        for (a, b), v in invd.items():
            val = (kdet * v[None, :]).reshape(-1)
            C[a, b] = val
            if a != b:
                C[b, a] = val
        C[2, 2] = (kdet / (dz_d * dz_d)[:, None]).reshape(-1)
        m = det.reshape(-1).clone() if mass else None
        return C, m

    def _device_lognormal(self, sigma, seed, modes):
        """The field of ``lognormal_conductivity`` (same random numbers), on the device,
        as one complex matrix product ``(n3, modes^2) x (modes^2, n_plane)``."""
        import torch as t

        from . import device as dv

        ext = self.device_extrusion()
        coef = self.lognormal_modes(sigma, seed, modes)
        k = np.arange(1, modes + 1)
        Ez = np.exp(1j * np.pi * ext["z"][:, None] * k[None, :])
        Zc = np.einsum("abc,zc->abz", coef, Ez).reshape(-1, ext["z"].shape[0])  # (m*m, n3)
        npl = ext["nplane"]
        xd, yd = ext["xy"][:npl], ext["xy"][npl:]
        kd = dv.upload(k.astype(np.float64))
        Ex = t.exp(1j * np.pi * xd[:, None] * kd[None, :])
        Ey = t.exp(1j * np.pi * yd[:, None] * kd[None, :])
        Pxy = (Ex[:, :, None] * Ey[:, None, :]).reshape(npl, -1)
        g = t.real(dv.upload(Zc).T.contiguous() @ Pxy.T.contiguous())  # (n3, nplane)
        return t.exp(sigma * g)


# ---------------------------------------------------------------------------
# Patch builders
# ---------------------------------------------------------------------------


def make_quadrature(degree: int, nbel: int, quadclass: str = "weighted", quadtype: str = "2"):
    kv = make_open_knotvector(degree, nbel)
    if quadclass == "gauss":
        return StandardGauss(degree, kv, quadtype=quadtype if quadtype in ("legendre", "lobatto") else "legendre")
    return WeightedQuadrature(degree, kv, quadtype=quadtype)


def unit_cube(degree: int, nbel: int, quadclass: str = "weighted", quadtype: str = "2", quadrule=None) -> ExtrudedPatch:
    """Identity map of ``[0,1]^3`` (Greville control points, so ``J = I``)."""
    q = quadrule if quadrule is not None else make_quadrature(degree, nbel, quadclass, quadtype)
    g = evaluate_greville(q.degree, q.knotvector)
    n = q.nbctrlpts
    X = np.tile(g, n)
    Y = np.repeat(g, n)
    return ExtrudedPatch([q, q, q], np.vstack([X, Y]), None, 0.0, 1.0, name="unit_cube")


def rational_arc_ctrlpts(degree: int, knotvector: np.ndarray, a0: float = 0.0, a1: float = np.pi / 2):
    """Control points and weights of the circular arc ``[a0, a1]`` in the spline space
    of ``(degree, knotvector)``.  The quadratic rational arc (end points weight 1, middle
    point weight ``cos((a1-a0)/2)``) has polynomial numerator ``X^w(t)`` and denominator
    ``W(t)``; interpolating both at the Greville points of the space gives its exact
    representation there.  Returns Cartesian points ``(2, n)`` and weights ``(n,)``."""
    beta = 0.5 * (a0 + a1)
    wm = np.cos(0.5 * (a1 - a0))
    P = np.array([[np.cos(a0), np.sin(a0), 1.0], [np.cos(beta), np.sin(beta), wm], [np.cos(a1), np.sin(a1), 1.0]])
    # P rows hold (w x, w y, w) with the middle point already weighted, as in geomdl's ctrlptsw
    g = evaluate_greville(degree, knotvector)
    bern = np.stack([(1 - g) ** 2, 2 * g * (1 - g), g ** 2], axis=1)
    target = bern @ P  # (n, 3) values of (X^w, Y^w, W) at the Greville points
    N = eval_ders_basis_sparse(degree, knotvector, g, nders=0)[0].toarray()
    coef = np.linalg.solve(N, target)
    w = coef[:, 2]
    return (coef[:, :2] / w[:, None]).T, w


def quarter_annulus(
    degree: int,
    nbel: int,
    r_in: float = 1.0,
    r_out: float = 2.0,
    height: float = 1.0,
    quadclass: str = "weighted",
    quadtype: str = "2",
    quadrule=None,
) -> ExtrudedPatch:
    """Quarter of a thick-walled annulus (``r in [r_in, r_out]``, ``theta in [0, pi/2]``)
    extruded to ``z in [0, height]`` as a NURBS map: direction 1 radial, 2 angular, 3 axial
    (same construction as ``tests/test_assemble.py:53-90``, plus the extrusion)."""
    if degree < 2:
        raise ValueError("the circular arc needs degree >= 2")
    q = quadrule if quadrule is not None else make_quadrature(degree, nbel, quadclass, quadtype)
    kv = q.knotvector
    arc_xy, arc_w = rational_arc_ctrlpts(q.degree, kv)
    g = evaluate_greville(q.degree, kv)
    radius = r_in + g * (r_out - r_in)
    n = q.nbctrlpts
    X = (radius[None, :] * arc_xy[0][:, None]).ravel()  # index a1 + n a2 -> [a2, a1]
    Y = (radius[None, :] * arc_xy[1][:, None]).ravel()
    W = np.repeat(arc_w, n)
    return ExtrudedPatch([q, q, q], np.vstack([X, Y]), W, 0.0, height, name="quarter_annulus")


# ---------------------------------------------------------------------------
# Lognormal conductivity
# ---------------------------------------------------------------------------


def lognormal_conductivity(patch: ExtrudedPatch, sigma: float, seed: int, modes: int = 8) -> np.ndarray:
    """``kappa = exp(sigma g)`` at the quadrature points, ``g`` a smooth random Fourier
    field with ``modes`` wave numbers per axis:

    ``g(x) = Re sum_k a_k e^{i phi_k} exp(i pi k . xhat) / s``,  ``k in {1..modes}^3``,

    ``xhat`` the physical point scaled to the bounding box of the control net,
    ``a_k ~ N(0,1) / (1 + |k|^2)``, ``phi_k ~ U[0, 2 pi)`` from ``default_rng(seed)``, and
    ``s^2 = sum a_k^2 / 2`` (unit variance in expectation over the phases).  The sum is
    separable between the cross-section and the extrusion direction, so it is evaluated
    as one complex matrix product of size ``(n_q1 n_q2, modes^2) x (modes^2, n_q3)``."""
    k = np.arange(1, modes + 1)
    coef = ExtrudedPatch.lognormal_modes(sigma, seed, modes)  # (k1, k2, k3)
    cp = patch.ctrlpts
    lo, hi = cp.min(axis=1), cp.max(axis=1)
    span = np.where(hi > lo, hi - lo, 1.0)
    pg = patch.plane_geometry()
    z, _ = patch.line_geometry()
    xh = ((pg.x[0] - lo[0]) / span[0]).ravel()
    yh = ((pg.x[1] - lo[1]) / span[1]).ravel()
    zh = (z - lo[2]) / span[2]
    Ex = np.exp(1j * np.pi * xh[:, None] * k[None, :])  # (nplane, m)
    Ey = np.exp(1j * np.pi * yh[:, None] * k[None, :])
    Ez = np.exp(1j * np.pi * zh[:, None] * k[None, :])  # (n3, m)
    Pxy = (Ex[:, :, None] * Ey[:, None, :]).reshape(xh.shape[0], -1)  # (nplane, m*m)
    Zc = np.einsum("abc,zc->abz", coef, Ez).reshape(-1, zh.shape[0])  # (m*m, n3)
    g = np.real(Pxy @ Zc)  # (nplane, n3)
    return np.exp(sigma * g.T.ravel())
