"""Device plumbing: the torch device, the current HIP stream and a per-(device, stream)
workspace buffer for the C ABI's scratch memory.

PyTorch supplies device memory and streams; all permanent arithmetic runs in the
HIP kernels of libtheboss_amd.so.
"""
import ctypes
from typing import Dict, Tuple

import numpy as np

from . import _lib

_workspaces: Dict[Tuple[int, int], "object"] = {}


def torch():
    import torch as _torch

    return _torch


_gpu_checked = False


def require_gpu():
    """torch, after checking once per process that a GPU is visible and the HIP library
    loads (both raise otherwise; the check is not repeated on every call, since
    torch.cuda.is_available() re-counts the devices each time)."""
    global _gpu_checked
    t = torch()
    if not _gpu_checked:
        if not t.cuda.is_available():
            raise RuntimeError(
                "theboss_amd needs a ROCm GPU (gfx950); torch.cuda.is_available() is False")
        _lib.load()
        _gpu_checked = True
    return t


def device(index: int = None):
    t = require_gpu()
    if index is None:
        index = t.cuda.current_device()
    return t.device("cuda", index)


def stream_ptr(dev=None) -> ctypes.c_void_p:
    t = torch()
    return ctypes.c_void_p(t.cuda.current_stream(dev).cuda_stream)


def workspace(nbytes: int, dev=None):
    """A uint8 device buffer of at least nbytes, reused per (device, stream). Calls that
    share it are ordered by the stream."""
    t = require_gpu()
    dev = dev if dev is not None else device()
    key = (dev.index, t.cuda.current_stream(dev).cuda_stream)
    buf = _workspaces.get(key)
    if buf is None or buf.numel() < nbytes:
        size = max(int(nbytes), 1 << 20)
        if buf is not None:
            size = max(size, 2 * buf.numel())
        buf = t.empty(size, dtype=t.uint8, device=dev)
        _workspaces[key] = buf
    return buf


def ptr(tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(tensor.data_ptr())


def to_device_c128(array, dev=None):
    t = require_gpu()
    a = np.ascontiguousarray(np.asarray(array, dtype=np.complex128))
    return t.from_numpy(a).to(dev if dev is not None else device(), non_blocking=False)


def to_device_i32(array, dev=None):
    t = require_gpu()
    a = np.ascontiguousarray(np.asarray(array, dtype=np.int32))
    return t.from_numpy(a).to(dev if dev is not None else device(), non_blocking=False)
