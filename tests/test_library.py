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
