"""The C-ABI library loads and exports every symbol declared in include/igaops_hip.h
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from igaops_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "igaops_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(iga_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    if not _lib.LIB_PATH.exists():
        pytest.fail("libigaops_hip.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    names = declared_symbols()
    assert set(names) <= set(_lib.SIGNATURES), sorted(set(names) - set(_lib.SIGNATURES))


def test_error_path_without_gpu_compute():
    """A bad argument is rejected on the host with a readable message and code -1."""
    lib = _lib.load()
    rc = lib.iga_band_create(None, 0, 0, 0, 0, None, None)
    assert rc == -1
    buf = ctypes.create_string_buffer(256)
    lib.iga_last_error(buf, 256)
    assert b"NULL" in buf.value or b"bad" in buf.value
    with pytest.raises(ValueError):
        _lib.check(rc, "iga_band_create")
    assert lib.iga_version() >= 100


def test_arnoldi_context_rejects_bad_arguments_on_the_host():
    """The one-call Arnoldi step checks its context on the host: missing pointers, a
    box whose grid is not the vector, a box that is not three-dimensional, and a tag
    whose low 32 bits are not zero are errors (-1); destroying no context is a no-op."""
    import numpy as np

    lib = _lib.load()
    h = ctypes.c_void_p()
    fake = ctypes.c_void_p(1 << 20)  # never dereferenced: each call fails its checks first
    tag = 5 << 32
    assert lib.iga_arnoldi_create(ctypes.byref(h), 0, 27, 5, None, fake, None, fake, None, None, 0, None, None, None,
                                  fake, fake, fake, fake, fake, tag) == -1
    n, lo, ln = (np.array(a, dtype=np.int64) for a in ((3, 3, 3), (1, 1, 1), (1, 1, 1)))
    ptrs = [a.ctypes.data_as(_lib._i64p) for a in (n, lo, ln)]
    assert lib.iga_arnoldi_create(ctypes.byref(h), 0, 28, 5, fake, fake, None, fake, None, None, 3, *ptrs,
                                  fake, fake, fake, fake, fake, tag) == -1
    buf = ctypes.create_string_buffer(256)
    lib.iga_last_error(buf, 256)
    assert b"box grid" in buf.value
    assert lib.iga_arnoldi_create(ctypes.byref(h), 0, 27, 5, fake, fake, None, fake, None, None, 2, *ptrs,
                                  fake, fake, fake, fake, fake, tag) == -1
    assert lib.iga_arnoldi_create(ctypes.byref(h), 0, 27, 5, fake, fake, None, fake, None, None, 3, *ptrs,
                                  fake, fake, fake, fake, fake, tag + 1) == -1
    lib.iga_last_error(buf, 256)
    assert b"tag" in buf.value
    assert lib.iga_arnoldi_destroy(None) == 0


def test_native_step_spec_checks():
    """NativeStep refuses a non-fused operator, an FD without workspace, an FD box
    other than the mask box and mismatched vector types."""
    import types

    import numpy as np

    from igaops_amd.runs import NativeStep

    box = (np.array([5, 5, 5]), np.array([1, 1, 1]), np.array([3, 3, 3]))
    op = types.SimpleNamespace(fused=True, dtype="f64")
    fd = types.SimpleNamespace(box=box, dtype="f64")
    NativeStep(op, fd, object(), box)
    NativeStep(op, None, None, box)
    with pytest.raises(ValueError, match="fused"):
        NativeStep(types.SimpleNamespace(fused=False, dtype="f64"))
    with pytest.raises(ValueError, match="workspace"):
        NativeStep(op, fd, None, box)
    with pytest.raises(ValueError, match="box"):
        NativeStep(op, fd, object(), (box[0], np.array([0, 1, 1]), box[2]))
    with pytest.raises(ValueError, match="vector types"):
        NativeStep(op, types.SimpleNamespace(box=box, dtype="f32"), object(), box)
