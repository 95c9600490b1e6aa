"""ctypes binding of ``libigaops_hip.so`` (declared in ``include/igaops_hip.h``).

Return codes are turned into the exception families of the reference
(``ValueError`` for bad arguments, ``RuntimeError`` for runtime failures,
``MemoryError``, ``NotImplementedError``).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libigaops_hip.so"

IGA_F64 = 0
IGA_F32 = 1
IGA_BF16 = 2  # FD plans: float vectors, bf16 matrix-core inputs, FP32 accumulation

_i32, _i64, _f64, _vp = C.c_int, C.c_int64, C.c_double, C.c_void_p
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "iga_version": (_i32, []),
    "iga_last_error": (_i32, [C.c_char_p, _i32]),
    "iga_device_count": (_i32, [C.POINTER(C.c_int)]),
    "iga_band_create": (_i32, [C.POINTER(_vp), _i32, _i32, _i32, _i32, _i32p, _f64p]),
    "iga_band_destroy": (_i32, [_vp]),
    "iga_band_mode_product": (_i32, [_vp, _i32, _i64, _i64, _vp, C.POINTER(_vp), _i32, _f64, _vp]),
    "iga_qp_scale": (_i32, [_i32, _i64, _i64, _vp, _vp, _vp, _i32, _f64, _vp]),
    "iga_reduce_workspace_bytes": (_i64, [_i64, _i32]),
    "iga_dot_many": (_i32, [_i32, _i64, _i32, _vp, _i64, _vp, _vp, _vp, _vp]),
    "iga_axpy_many": (_i32, [_i32, _i64, _i32, _vp, _i64, _vp, _f64, _vp, _vp]),
    "iga_zgram_workspace_bytes": (_i64, [_i64, _i32, _i32]),
    "iga_zgram": (_i32, [_i64, _i32, _vp, _i32, _vp, _vp, _vp, _vp]),
    "iga_zcombine": (_i32, [_i64, _i32, _vp, _i32, _vp, _f64, _f64, _vp, _vp]),
    "iga_zdot_many": (_i32, [_i64, _i32, _vp, _i64, _vp, _vp, _vp, _vp]),
    "iga_zaxpy_many": (_i32, [_i64, _i32, _vp, _i64, _vp, _f64, _f64, _vp, _vp]),
    "iga_axpby": (_i32, [_i32, _i64, _f64, _vp, _i32, _vp, _f64, _vp, _vp]),
    "iga_cast": (_i32, [_i32, _i32, _i64, _vp, _vp, _vp]),
    "iga_box_gather": (_i32, [_i32, _i32, _i64p, _i64p, _i64p, _i64, _vp, _vp, _vp]),
    "iga_box_scatter": (_i32, [_i32, _i32, _i64p, _i64p, _i64p, _i64, _vp, _vp, _i32, _vp]),
    "iga_box_mask": (_i32, [_i32, _i32, _i64p, _i64p, _i64p, _i64, _vp, _vp]),
    "iga_index_zero": (_i32, [_i32, _i64, _vp, _i64, _vp, _vp]),
    "iga_fused_create": (
        _i32,
        [C.POINTER(_vp), _i32p, _i32p, _i32, C.POINTER(_i32p), C.POINTER(_f64p), _i32, _i32, _i32],
    ),
    "iga_fused_destroy": (_i32, [_vp]),
    "iga_fused_set_rhs_block": (_i32, [_vp, _i32]),
    "iga_fused_info": (_i32, [_vp, _i64p, _i64p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "iga_fused_workspace_bytes": (_i64, [_vp, _i32]),
    "iga_fused_pack": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "iga_fused_apply": (_i32, [_vp, _vp, _vp, _vp, _i32, _vp, _f64, _i32, _vp, _vp]),
    "iga_fused_apply_noalloc": (_i32, [_vp, _vp, _vp, _vp, _i32, _vp, _f64, _i32, _vp]),
    "iga_fused_apply_group": (_i32, [_vp, _i32, C.POINTER(_vp), _vp, _i64, _i64, C.POINTER(_vp), _vp, _vp, _i32,
                                     _i64p, _i64p, _vp]),
    "iga_fused_apply_stage": (_i32, [_vp, _i32, _vp, _vp, _vp, _i32, _vp, _f64, _i32, _vp, _vp]),
    "iga_fd_create": (_i32, [C.POINTER(_vp), _i32, _i64p, _i64p, _i64p, C.POINTER(_f64p), _i32]),
    "iga_fd_set_eigenvalues": (_i32, [_vp, _f64p]),
    "iga_fd_workspace_bytes": (_i64, [_vp, _i32]),
    "iga_fd_apply": (_i32, [_vp, _vp, _vp, _i32, _vp, _vp]),
    "iga_fd_apply_rhs_major": (_i32, [_vp, _vp, _vp, _i32, _vp, _vp]),
    "iga_fd_destroy": (_i32, [_vp]),
    "iga_fused_apply_sliced": (_i32, [_vp, _vp, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _i32, _vp, C.c_double, _i32,
                                      _vp, _vp]),
    "iga_fused_apply_sliced2": (_i32, [_vp, _vp, _i64, _vp, _vp, _i64, _i64, _vp, _i64, _i64, _i32, _vp, C.c_double,
                                       _i32, _vp, _vp]),
    "iga_fused_pack_strided": (_i32, [_vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp]),
    "iga_fused_pack_extruded": (_i32, [_vp, _vp, _vp, _vp, _vp, _i32, _f64, _i32, _vp, _vp, _vp, _vp]),
    "iga_lognormal_plane_modes": (_i32, [_i64, _i32, _vp, _vp, _vp, _vp]),
    "iga_gmres_state_doubles": (_i64, [_i32]),
    "iga_gmres_layout": (_i32, [_i32, _i64p]),
    "iga_sumfact_nnz": (_i64, [_i32, _i64p, C.POINTER(_i32p)]),
    "iga_sumfact_assemble": (
        _i32,
        [_i32, _i32, _i64p, _i64p, _i64p, C.POINTER(_i32p), C.POINTER(_i32p), _i64p, C.POINTER(_i32p),
         C.POINTER(_i32p), C.POINTER(_f64p), _i64p, C.POINTER(_vp), _f64p, _vp, _vp, _vp, _vp, _vp],
    ),
    "iga_gmres_start": (_i32, [_i32, _i64, _i32, _vp, _vp, _vp, _vp, _vp]),
    "iga_gmres_step": (_i32, [_i32, _i64, _i32, _i32, _vp, _i64, _vp, _vp, _vp, _vp]),
    "iga_stfd_create": (_i32, [C.POINTER(_vp), _i64p, _i64p, _i64p, C.POINTER(_f64p), _f64p, _f64p, _i32]),
    "iga_stfd_set_blocks": (_i32, [_vp, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_double]),
    "iga_stfd_workspace_bytes": (_i64, [_vp, _i32]),
    "iga_stfd_apply": (_i32, [_vp, _vp, _vp, _i32, _i32, _vp, _vp]),
    "iga_stfd_destroy": (_i32, [_vp]),
    "iga_stfd_create_qz": (_i32, [C.POINTER(_vp), _i64p, _i64p, _i64p, C.POINTER(_f64p), _f64p, _f64p, _f64p,
                                  _f64p, _i32]),
    "iga_stfd_set_qz": (_i32, [_vp, _f64p, _f64p, _f64p, _f64p, _f64p, _f64]),
    "iga_arnoldi_create": (_i32, [C.POINTER(_vp), _i32, _i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i64p, _i64p,
                                  _i64p, _vp, _vp, _vp, _vp, _vp, _i64]),
    "iga_arnoldi_start": (_i32, [_vp, _vp, _vp, _vp]),
    "iga_arnoldi_step": (_i32, [_vp, _i32, _vp, _i64, _i32, _vp]),
    "iga_arnoldi_group_workspace_bytes": (_i64, [_vp, _i32]),
    "iga_arnoldi_increment_doubles": (_i64, [_i32]),
    "iga_arnoldi_set_increment": (_i32, [_vp, _vp, _i64, _vp, _vp]),
    "iga_arnoldi_step_group": (_i32, [_i32, C.POINTER(_vp), C.POINTER(C.c_int), C.POINTER(_vp), _i64p,
                                      C.POINTER(_vp), _vp, _vp, _vp, _vp]),
    "iga_arnoldi_destroy": (_i32, [_vp]),
    "iga_host_alloc": (_i32, [_i64, C.POINTER(_vp)]),
    "iga_host_free": (_i32, [_vp]),
}

_ERRORS = {-1: ValueError, -2: RuntimeError, -3: MemoryError, -4: NotImplementedError, -5: RuntimeError}

_lib = None


def load():
    """Load the library (raises if it has not been built: there is no fallback path)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH.name} is not built; run `python -m igaops_amd.build` (the HIP path has no CPU fallback)"
            )
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int, what: str = ""):
    if rc == 0:
        return
    buf = C.create_string_buffer(2048)
    load().iga_last_error(buf, 2048)
    exc = _ERRORS.get(rc, RuntimeError)
    raise exc(f"{what}: {buf.value.decode(errors='replace')}" if what else buf.value.decode(errors="replace"))


def call(name: str, *args):
    """Call an int-returning entry point and raise on failure."""
    rc = getattr(load(), name)(*args)
    check(rc, name)


def i32_array(a):
    import numpy as np

    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_i32p)


def i64_array(a):
    import numpy as np

    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(_i64p)


def f64_array(a):
    import numpy as np

    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_f64p)
