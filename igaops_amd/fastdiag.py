"""Fast-diagonalization (FD) preconditioner: host LAPACK set-up, device apply.

Set-up follows ``fastdiag/utils/space_precond.py:124-236`` (Gauss rewrite of the
rules, Dirichlet trimming, ``+1e-8 M`` regularisation of untrimmed directions,
``scipy.linalg.eigh`` of each 1-D pencil, Kronecker-sum eigenvalues with correctors,
clip at ``1e-14 max``).  The apply ``y[free] = U (Lambda^{-1} o U^T x[free])``,
``y = 0`` elsewhere (``space_precond.py:273-320``) runs on the device: the free nodes
form a box (one interior range per direction), so gather/scatter are box copies, the
Kronecker transforms are dense mode products and ``Lambda^{-1}`` is a point-wise scale.
"""
from __future__ import annotations

from typing import List, Optional, Sequence

import numpy as np
import scipy.linalg as sla

from . import _lib
from . import device as dv
from .common import Constants, ParametricDirection, combine_arrays, kron_nonzero_indices, validate_entry
from .quadrature import StandardGauss


def rewrite_quadrature(quadrule_list) -> List[StandardGauss]:
    """Gauss-Legendre rules on the same spaces (``fastdiag/utils/template.py:28-45``)."""
    if not isinstance(quadrule_list, (list, tuple)):
        raise ValueError("quadrule_list should be iterable")
    out = []
    for q in quadrule_list:
        if getattr(q, "quadrature_class", None) == "GAUSS" and getattr(q, "quadrature_type", None) == "legendre":
            out.append(q)
        else:
            out.append(StandardGauss(q.degree, q.knotvector, quadtype="legendre"))
    return out


def pencil_eigenpairs(q, dirichlet_lo: bool, dirichlet_hi: bool):
    """Free indices and generalized eigenpairs ``K U = M U diag(lam)`` of one direction
    (``space_precond.py:124-193``): ``M``, ``K`` from the Gauss rule ``q``, rows and
    columns of Dirichlet ends removed, ``K + 1e-8 M`` when no end is removed,
    ``scipy.linalg.eigh`` (``U^T M U = I``)."""
    M = (q.weights[0] @ q.basis[0]).toarray()
    K = (q.weights[-1] @ q.basis[-1]).toarray()
    n = q.nbctrlpts
    idx = np.arange(1 if dirichlet_lo else 0, n - (1 if dirichlet_hi else 0))
    Mi = M[np.ix_(idx, idx)]
    Ki = K[np.ix_(idx, idx)]
    if len(idx) == n:
        Ki = Ki + Constants.TINY * Mi
    lam, U = sla.eigh(Ki, Mi)
    return idx, np.real(lam), np.real(U)


class SpaceFD:
    """Spatial FD for one DOF per node (``nbDoFsPerNode = 1``, the scalar problems of
    the path; the reference's vector-valued loop over DOFs per node is the same apply
    repeated on each component)."""

    def __init__(self):
        self.nbDoFsPerNode = 0
        self.space_scalar_coefs = (0.0, 1.0)
        self._eigvals_cache = None
        self._dev = {}
        # bumped whenever the eigenvalues change; device plans compare it
        self.eig_version = 0
        # plan of float32 inputs: "f32" or "bf16" (bf16 matrix-core inputs, FP32 accumulation)
        self.float32_plan = "f32"

    def compute_space_eigendecomposition(self, space_quadrule_list, space_table_dirichlet):
        table = np.asarray(space_table_dirichlet, dtype=bool)
        if table.ndim != 3:
            raise ValueError("space_table_dirichlet must have shape (ndof_per_node, ndim, 2)")
        if table.shape[0] != 1:
            raise NotImplementedError("one DOF per node is supported")
        qs = rewrite_quadrature(space_quadrule_list)
        nbdirs = min(table.shape[1], len(qs))
        solved = {}
        pairs = []
        for d in range(nbdirs):
            q = qs[d]
            # directions with the same space, rule and boundary flags share one pencil
            key = (q.degree, q.knotvector.tobytes(), q.nbquadpts, bool(table[0, d, 0]), bool(table[0, d, 1]))
            if key not in solved:
                solved[key] = pencil_eigenpairs(q, table[0, d, 0], table[0, d, 1])
            pairs.append(solved[key])
        self.set_space_eigenpairs(table, [q.nbctrlpts for q in qs[:nbdirs]], pairs)

    def set_space_eigenpairs(self, space_table_dirichlet, nnz, pairs):
        """Install per-direction ``(indices, eigenvalues, eigenvectors)`` computed by
        :func:`pencil_eigenpairs` (possibly elsewhere, e.g. in a worker process)."""
        table = np.asarray(space_table_dirichlet, dtype=bool)
        self.nbDoFsPerNode = table.shape[0]
        self.nnz = [int(n) for n in nnz]
        nbdirs = len(self.nnz)
        self.space_dirs = [ParametricDirection(i) for i in range(nbdirs)]
        self.indices = [np.asarray(p[0]) for p in pairs]
        self.eigenval_by_dir_space = [np.asarray(p[1]) for p in pairs]
        self.eigenvec_by_dir_space = [np.asarray(p[2]) for p in pairs]
        self.space_table_dirichlet = table
        self.mass_space_corrector = [1.0]
        self.stiff_space_corrector = [np.ones(nbdirs)]
        self._eigvals_cache = None
        self._dev = {}

    # -- reference-compatible accessors (fastdiag/utils/space_precond.py:40-102) --
    @property
    def sp_nnz(self) -> int:
        return int(np.prod(self.nnz))

    @property
    def space_free_nodes(self) -> List[List[int]]:
        nnz = list(self.nnz)[::-1]
        return [kron_nonzero_indices([idx.tolist() for idx in self.indices[::-1]], nnz)]

    def update_space_eigenvalues(self, scalar_coefs: Sequence[float]) -> List[np.ndarray]:
        """``space_precond.py:195-236``."""
        if len(scalar_coefs) < 2:
            return []
        mixed = combine_arrays(self.eigenval_by_dir_space, self.stiff_space_corrector[0])
        cur = scalar_coefs[0] * self.mass_space_corrector[0] * np.ones_like(mixed) + scalar_coefs[1] * mixed
        cur = np.maximum(cur, Constants.SAFEGUARD * np.max(cur))
        self.space_scalar_coefs = (float(scalar_coefs[0]), float(scalar_coefs[1]))
        self._eigvals_cache = [cur]
        self.eig_version += 1
        return [cur]

    @property
    def space_eigenvalues(self) -> List[np.ndarray]:
        if self._eigvals_cache is None:
            self.update_space_eigenvalues(self.space_scalar_coefs)
        return self._eigvals_cache

    def add_scalar_space_correctors(self, mass_corrector=None, stiffness_corrector=None):
        ok = lambda e: isinstance(e, (list, tuple)) and len(e) == self.nbDoFsPerNode
        if mass_corrector is not None and ok(mass_corrector):
            self.mass_space_corrector = list(mass_corrector)
            self._eigvals_cache = None
        if stiffness_corrector is not None and ok(stiffness_corrector):
            self.stiff_space_corrector = [np.asarray(s, dtype=float) for s in stiffness_corrector]
            self._eigvals_cache = None

    # -- device apply ---------------------------------------------------------
    def plan(self, dtype: str) -> "FDPlan":
        key = ("plan", dtype)
        p = self._dev.get(key)
        if p is None:
            n, lo, ln = self.box()
            p = FDPlan(n, lo, ln, self.eigenvec_by_dir_space, dtype)
            self._dev[key] = p
        lam = self.space_eigenvalues[0]
        if p.eig_version != self.eig_version:
            p.set_eigenvalues(lam, self.eig_version)
        return p

    def box(self):
        lo = np.array([int(idx[0]) if len(idx) else 0 for idx in self.indices], dtype=np.int64)
        ln = np.array([len(idx) for idx in self.indices], dtype=np.int64)
        return np.array(self.nnz, dtype=np.int64), lo, ln

    def apply_device(self, x, out=None, nrhs: int = 1, stream=None):
        """``out = P x`` for device arrays ((N, nrhs) C order), MFMA mode products;
        float32 vectors use the plan named by ``float32_plan`` ("f32", or "bf16": bf16
        matrix-core inputs with FP32 accumulation)."""
        t = dv.torch()
        dtype = self.float32_plan if x.dtype == t.float32 else "f64"
        if out is None:
            out = t.empty_like(x)
        return self.plan(dtype).apply(x, out, nrhs, stream)


class FDPlan:
    """Device FD plan (``iga_fd_t``): eigenvectors uploaded once, ``Lambda^{-1}`` on set."""

    def __init__(self, n, lo, ln, eigvecs, dtype: str):
        import ctypes as C

        self.dtype = dtype
        self.nsub = int(np.prod(ln))
        # the free box (grid sizes, first free index, free count per direction)
        self.box = tuple(np.array(a, dtype=np.int64) for a in (n, lo, ln))
        na, npt = _lib.i64_array(n)
        la, lpt = _lib.i64_array(lo)
        lna, lnpt = _lib.i64_array(ln)
        keep = [np.ascontiguousarray(U, dtype=np.float64) for U in eigvecs]
        arr = (_lib._f64p * len(keep))(*[k.ctypes.data_as(_lib._f64p) for k in keep])
        h = C.c_void_p()
        _lib.call("iga_fd_create", C.byref(h), len(n), npt, lpt, lnpt, arr, dv.DTYPES[dtype])
        self._h = h
        self.eig_version = None
        self._work = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib.load().iga_fd_destroy(h)
            self._h = None

    def set_eigenvalues(self, lam: np.ndarray, version=None):
        la, lp = _lib.f64_array(lam)
        if la.size != self.nsub:
            raise ValueError("eigenvalues do not match the free box")
        _lib.call("iga_fd_set_eigenvalues", self._h, lp)
        self.eig_version = version

    def workspace_bytes(self, nrhs: int = 1) -> int:
        return int(_lib.load().iga_fd_workspace_bytes(self._h, int(nrhs)))

    def workspace(self, nrhs: int = 1):
        return dv.workspace(self.workspace_bytes(nrhs))

    def apply(self, x, out, nrhs: int = 1, stream=None, work=None, rhs_major: bool = False):
        """``out = P x``; ``work`` (from ``workspace``) lets concurrent callers of one
        plan keep their scratch apart (the plan's own buffer is used otherwise).
        Several right-hand sides are ``(N, nrhs)`` C order, or ``(nrhs, N)`` with
        ``rhs_major`` (``iga_fd_apply_rhs_major``)."""
        w = work
        if w is None:
            w = self._work.get(nrhs)
        if w is None:
            w = self.workspace(nrhs)
            self._work[nrhs] = w
        _lib.call("iga_fd_apply_rhs_major" if rhs_major else "iga_fd_apply", self._h, dv.ptr(x), dv.ptr(out),
                  int(nrhs), dv.ptr(w), dv.stream_ptr(stream))
        return out


class SingleFastDiagonalization:
    """``fastdiag/preconditioner.py:15-197`` (spatial type; the space-time type is in
    ``igaops_amd.spacetime``)."""

    def __init__(self, preconditioner_type: str):
        self._preconditioner_type = validate_entry(preconditioner_type, ["spatial", "spacetime"])
        self._space_fd: Optional[SpaceFD] = None
        self._sptime_fd = None

    @property
    def preconditioner_type(self):
        return self._preconditioner_type

    @property
    def space_preconditioner(self) -> SpaceFD:
        if self._space_fd is None:
            raise RuntimeError("Space preconditioner not defined.")
        return self._space_fd

    @property
    def sptm_preconditioner(self):
        if self._preconditioner_type == "spatial" or self._sptime_fd is None:
            raise RuntimeError("Space-time preconditioner not defined.")
        return self._sptime_fd

    def compute_space_eigendecomposition(self, space_quadrule_list, space_table_dirichlet):
        fd = SpaceFD()
        fd.compute_space_eigendecomposition(space_quadrule_list, space_table_dirichlet)
        self._space_fd = fd

    def compute_time_schurdecomposition(self, time_quadrule, use_arrowhead: bool = True):
        if self._preconditioner_type == "spatial":
            raise RuntimeError("Space-time preconditioner not defined.")
        from .spacetime import SpaceTimeFD, SpaceTimeQZFD

        # the arrowhead form, or the legacy QZ form (fastdiag/preconditioner.py:65-88)
        st = (SpaceTimeFD if use_arrowhead else SpaceTimeQZFD)(self.space_preconditioner)
        st.compute_time_schurdecomposition(time_quadrule)
        self._sptime_fd = st

    def update_space_eigenvalues(self, scalar_coefs):
        self.space_preconditioner.update_space_eigenvalues(scalar_coefs)
        if self._sptime_fd is not None:
            self._sptime_fd.invalidate()

    def add_scalar_space_time_correctors(self, mass_corrector=None, stiffness_corrector=None, advection_corrector=None):
        if self._space_fd is not None:
            self._space_fd.add_scalar_space_correctors(mass_corrector, stiffness_corrector)
        if self._sptime_fd is not None:
            self._sptime_fd.add_scalar_time_correctors(advection_corrector)

    def apply_device(self, x, out=None, nrhs=1, stream=None):
        if self._preconditioner_type == "spatial":
            return self.space_preconditioner.apply_device(x, out, nrhs, stream)
        return self.sptm_preconditioner.apply_device(x, out, nrhs, stream)

    def apply_transpose_device(self, x, out=None, nrhs=1, stream=None):
        if self._preconditioner_type == "spatial":
            return self.space_preconditioner.apply_device(x, out, nrhs, stream)
        return self.sptm_preconditioner.apply_device(x, out, nrhs, stream, transpose=True)

    def _host_call(self, fn, array_in):
        dv.require_gpu()
        t = dv.torch()
        was_np = not isinstance(array_in, t.Tensor)
        shape = tuple(array_in.shape)
        r = int(np.prod(shape[1:])) if len(shape) > 1 else 1
        dtype = "f32" if (not was_np and array_in.dtype == t.float32) else "f64"
        x = dv.to_device(array_in, dtype).reshape(-1)
        y = fn(x, None, r)
        y = y.reshape(shape)
        return y.double().cpu().numpy() if was_np else y

    def apply(self, array_in):
        """``preconditioner.py:147-163``."""
        return self._host_call(self.apply_device, array_in)

    def apply_transpose(self, array_in):
        """``preconditioner.py:165-182``."""
        return self._host_call(self.apply_transpose_device, array_in)

    __call__ = apply


SingleFD = SingleFastDiagonalization


class MultiFastDiagonalization:
    """``fastdiag/multi_preconditioner.py:13-103``: block-diagonal FD over several
    patches, patch ``i`` owning the contiguous slice of its DOFs (space DOFs, times
    the time DOFs for the space-time type) in list order.  Each block runs the
    patch's device apply on its slice; numpy in, numpy out (device tensors stay on
    the device)."""

    def __init__(self, list_preconditioner):
        if len(list_preconditioner) < 1:
            raise RuntimeError("Empty list of preconditioners")
        if not all(isinstance(m, SingleFastDiagonalization) for m in list_preconditioner):
            raise RuntimeError("All elements should be preconditioners")
        self._list_fd = list(list_preconditioner)
        self._preconditioner_type = self._list_fd[0].preconditioner_type
        for fd in self._list_fd:
            if fd.preconditioner_type != self._preconditioner_type:
                raise RuntimeError("Preconditioners should be of the same type")
        self._slices = self._set_slices()

    @property
    def preconditioner_type(self):
        return self._preconditioner_type

    def _set_slices(self):
        sizes = [0]
        for fd in self._list_fd:
            sp = fd.space_preconditioner
            n = max(1, sp.nbDoFsPerNode) * sp.sp_nnz
            if self._preconditioner_type == "spacetime":
                n *= fd.sptm_preconditioner.box()[0][-1]
            sizes.append(int(n))
        c = np.cumsum(sizes)
        return [slice(int(c[i]), int(c[i + 1])) for i in range(len(c) - 1)]

    def apply_device(self, x, out=None, nrhs: int = 1, stream=None, transpose: bool = False):
        t = dv.torch()
        if out is None:
            out = t.empty_like(x)
        X = x.reshape(-1, nrhs)
        Y = out.reshape(-1, nrhs)
        if X.shape[0] != self._slices[-1].stop:
            raise ValueError("array does not match the patches")
        for fd, slc in zip(self._list_fd, self._slices):
            xi = X[slc].contiguous().reshape(-1)
            fn = fd.apply_transpose_device if transpose else fd.apply_device
            Y[slc] = fn(xi, None, nrhs, stream).reshape(-1, nrhs)
        return out

    def _host_call(self, array_in, transpose: bool):
        dv.require_gpu()
        t = dv.torch()
        was_np = not isinstance(array_in, t.Tensor)
        shape = tuple(array_in.shape)
        r = int(np.prod(shape[1:])) if len(shape) > 1 else 1
        dtype = "f32" if (not was_np and array_in.dtype == t.float32) else "f64"
        x = dv.to_device(array_in, dtype).reshape(-1)
        y = self.apply_device(x, None, r, None, transpose).reshape(shape)
        return y.double().cpu().numpy() if was_np else y

    def apply(self, array_in):
        """``multi_preconditioner.py:61-81``."""
        return self._host_call(array_in, False)

    def apply_transpose(self, array_in):
        """``multi_preconditioner.py:83-103``."""
        return self._host_call(array_in, True)


MultiFD = MultiFastDiagonalization
