"""ctypes binding of libsbh_amd.so (the C ABI declared in include/sbh_amd.h).

This is the host side of the drop-in boundary: it plays the role XLA's FFI
call frame plays for the reference (python/jax_sbh_ffi.py:44-59 loads the
library with ctypes.CDLL and hands handler symbols to jax.ffi).  Device
memory and streams come from PyTorch-ROCm; the arithmetic is all in the HIP
library.  There is no CPU fallback: without the built library every call
raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import torch

LIB_NAME = "libsbh_amd.so"
_HERE = os.path.dirname(os.path.abspath(__file__))

SBH_OK = 0
SBH_INVALID_ARGUMENT = 3
SBH_INTERNAL = 13

_DTYPE_CODES = {torch.uint8: 6, torch.int32: 4, torch.int64: 5, torch.float32: 11}


class SbhBuffer(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("dtype", ctypes.c_int32),
        ("rank", ctypes.c_int32),
        ("dims", ctypes.c_int64 * 4),
    ]


_lib: Optional[ctypes.CDLL] = None


def library_path() -> str:
    return os.path.join(_HERE, LIB_NAME)


def load_library(path: Optional[str] = None) -> ctypes.CDLL:
    """Load libsbh_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or library_path()
    if not os.path.exists(p):
        raise FileNotFoundError(
            f"{p} not found: build the HIP library first (python -c "
            "'import __graft_entry__ as g; g.build()')"
        )
    lib = ctypes.CDLL(p)
    lib.sbh_last_error.restype = ctypes.c_char_p
    for name in ("sbh_leaf3_size", "sbh_node3_size", "sbh_octree_level3_max_nodes",
                 "sbh_octree8_node_size", "sbh_leaf2_size", "sbh_node2_size"):
        getattr(lib, name).restype = ctypes.c_size_t
    if path is None:
        _lib = lib
    return lib


def buf(t: Optional[torch.Tensor]):
    """SbhBuffer view of a contiguous device tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("sbh_amd expects device (HIP) tensors")
    if not t.is_contiguous():
        raise ValueError("sbh_amd expects contiguous tensors")
    code = _DTYPE_CODES.get(t.dtype)
    if code is None:
        raise TypeError(f"unsupported dtype {t.dtype}")
    if t.dim() > 4:
        raise ValueError("rank > 4 not supported")
    b = SbhBuffer()
    b.data = t.data_ptr() if t.numel() > 0 else None
    b.dtype = code
    b.rank = t.dim()
    for i, d in enumerate(t.shape):
        b.dims[i] = int(d)
    return ctypes.byref(b)


def stream_handle(device: Optional[torch.device] = None) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def check(rc: int) -> None:
    if rc == SBH_OK:
        return
    msg = (load_library().sbh_last_error() or b"").decode()
    if rc == SBH_INVALID_ARGUMENT:
        raise ValueError(msg)
    raise RuntimeError(f"sbh_amd internal error: {msg}")


def f32(x) -> ctypes.c_float:
    return ctypes.c_float(float(x))


def i32(x) -> ctypes.c_int32:
    return ctypes.c_int32(int(x))


def call(name: str, *args) -> None:
    fn = getattr(load_library(), name)
    fn.restype = ctypes.c_int
    check(fn(*args))
