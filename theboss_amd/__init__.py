"""theboss_amd: MI355X (gfx950) permanent hot path for theboss-style boson sampling.

Permanent arithmetic runs in hand-written HIP kernels (libtheboss_amd.so, C ABI in
include/theboss_amd.h), reached through ctypes on PyTorch-ROCm tensors.
"""
from . import _lib

__version__ = "0.1.0"


def library_path() -> str:
    return _lib.LIB_PATH
