"""Device plumbing: torch owns HBM buffers and streams; the library sees raw pointers."""
from __future__ import annotations

from typing import Union

import numpy as np

from . import _lib

DTYPES = {"f64": _lib.IGA_F64, "f32": _lib.IGA_F32, "bf16": _lib.IGA_BF16}


def torch():
    import torch as _t

    return _t


def require_gpu():
    """The HIP path is the only path: fail loudly without a device or the library."""
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("igaops_amd needs a HIP device (MI355X); no CPU fallback is provided")
    _lib.load()
    return t


def device():
    return torch().device("cuda", torch().cuda.current_device())


_DEV_INDEX = None


def stream_ptr(stream=None) -> int:
    """The hipStream_t of ``stream`` (a torch stream) or of the current stream."""
    global _DEV_INDEX
    if stream is not None:
        return int(stream.cuda_stream)
    t = torch()
    raw = getattr(t._C, "_cuda_getCurrentRawStream", None)
    if raw is not None:
        # the current device does not change within a process of this library
        if _DEV_INDEX is None:
            _DEV_INDEX = t.cuda.current_device()
        return int(raw(_DEV_INDEX))
    return int(t.cuda.current_stream().cuda_stream)


def torch_dtype(dtype: str):
    t = torch()
    # "bf16" names the FD arithmetic; its vectors are float32
    return {"f64": t.float64, "f32": t.float32, "bf16": t.float32}[dtype]


def iga_dtype(t) -> int:
    tt = torch()
    if t.dtype == tt.float64:
        return _lib.IGA_F64
    if t.dtype == tt.float32:
        return _lib.IGA_F32
    raise ValueError(f"unsupported dtype {t.dtype}")


def ptr(t) -> int:
    if t is None:
        return 0
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return int(t.data_ptr())


def to_device(a: Union[np.ndarray, "object"], dtype: str = "f64"):
    t = torch()
    if isinstance(a, t.Tensor):
        return a.to(device=device(), dtype=torch_dtype(dtype)).contiguous()
    return t.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(device=device(), dtype=torch_dtype(dtype))


def upload(a: np.ndarray):
    """Host array to the device without waiting for the copy: staged through pinned
    memory (torch's pinned-block cache keeps the block until the copy on the current
    stream has finished), so the host does not wait behind the device work queued
    ahead of the copy."""
    t = torch()
    return t.from_numpy(np.ascontiguousarray(a)).pin_memory().to(device(), non_blocking=True)


def empty(n, dtype: str = "f64"):
    t = torch()
    shape = (n,) if np.isscalar(n) else tuple(n)
    return t.empty(shape, dtype=torch_dtype(dtype), device=device())


def zeros(n, dtype: str = "f64"):
    t = torch()
    shape = (n,) if np.isscalar(n) else tuple(n)
    return t.zeros(shape, dtype=torch_dtype(dtype), device=device())


def workspace(nbytes: int):
    """Raw device scratch (uint8) of at least ``nbytes`` bytes, 256-byte aligned by torch."""
    t = torch()
    return t.empty((max(int(nbytes), 256),), dtype=t.uint8, device=device())
