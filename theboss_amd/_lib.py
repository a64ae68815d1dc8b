"""ctypes binding of libtheboss_amd.so (the gfx950 C ABI declared in include/theboss_amd.h).

The library is loaded from this package directory. There is no fallback: if the
shared object is missing or no GPU is visible, every GPU entry point raises.
"""
import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("THEBOSS_AMD_LIB") or os.path.join(HERE, "libtheboss_amd.so")

TB_OK = 0
TB_EINVAL = -1
TB_ERANGE = -2
TB_EHIP = -3
TB_EWORKSPACE = -4

TB_PREC_FP64 = 0
TB_PREC_FP32_COLSUM = 1
TB_PREC_FP32_COLSUM_ANY_N = 2
TB_FP32_COLSUM_MAX_N = 23
TB_FORMULA_RYSER = 0
TB_FORMULA_CHIN_HUH = 1

TB_TUNE_SWEEP = 1
TB_TUNE_GLYNN = 2
TB_TUNE_GRAD = 3
TB_TUNE_MAX_SIZE = 48
TB_TUNE_BATCH_BUCKETS = 16

TB_GLYNN_MAX_N = 48
TB_GLYNN_NOALLOC_SCRATCH_BYTES = 16 << 20
TB_SWEEP_MAX_K = 40
TB_GUAN_MAX_ROWS = 48
TB_GUAN_MAX_COLS = 32
TB_HIST_MAX_M = 64
TB_HIST_MAX_N = 64


class TbError(RuntimeError):
    """A C-ABI call returned a negative status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"theboss_amd error {code}: {message}")
        self.code = code


class GccbConfig(ctypes.Structure):
    _fields_ = [
        ("m", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("samples", ctypes.c_int32),
        ("chunk_bits", ctypes.c_int32 * (TB_SWEEP_MAX_K + 1)),
        ("per_sample", ctypes.c_int32),
        ("u_stride", ctypes.c_int64),
    ]


_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_sz = ctypes.c_size_t

# name -> (restype, argtypes)
SIGNATURES = {
    "tb_version": (_i32, []),
    "tb_last_error": (ctypes.c_char_p, []),
    "tb_device_count": (_i32, []),
    "tb_tune_set": (_i32, [_i32, _i32, _i64, _i32]),
    "tb_tune_get": (_i32, [_i32, _i32, _i64]),
    "tb_glynn_workspace_size": (_sz, [_i32, _i32, _i32]),
    "tb_glynn_batch": (_i32, [_vp, _i32, _i32, _i64, _i32, _vp, _vp, _sz, _vp]),
    "tb_glynn_batch_noalloc": (_i32, [_vp, _i32, _i32, _i64, _i32, _vp, _vp]),
    "tb_guan_perm_batch": (_i32, [_vp, _i32, _vp, _vp, _i32, _vp, _vp, _i32, _i32, _i32, _vp, _vp]),
    "tb_guan_perm_outer": (
        _i32, [_vp, _i32, _vp, _vp, _i32, _i32, _vp, _vp, _i32, _i32, _i32, _vp, _vp]),
    "tb_cc_workspace_size": (_sz, [_i32, _i32, _i32, _i32]),
    "tb_cc_submatrices_batch": (
        _i32, [_vp, _i32, _vp, _vp, _i32, _vp, _vp, _i32, _i32, _i32, _i32, _vp, _vp, _sz, _vp]),
    "tb_sweep_workspace_size": (_sz, [_i32, _i32]),
    "tb_sweep_minors_batch": (_i32, [_vp, _i32, _i32, _i64, _i32, _vp, _vp, _sz, _vp]),
    "tb_perm_grad_workspace_size": (_sz, [_i32, _i32]),
    "tb_perm_grad_batch": (_i32, [_vp, _i32, _i32, _i64, _vp, _vp, _vp, _sz, _vp]),
    "tb_outcome_count": (_i64, [_i32, _i32, _i32]),
    "tb_outcome_histogram": (_i32, [_vp, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _vp]),
    "tb_gccb_workspace_size": (_sz, [ctypes.POINTER(GccbConfig)]),
    "tb_gccb_sample": (
        _i32,
        [ctypes.POINTER(GccbConfig), _vp, _i32, _vp, ctypes.c_uint64, ctypes.c_double, _vp, _vp,
         _vp, _vp, _vp, _vp, _sz, _vp]),
    "tb_gccb_status": (_i32, [_vp, _vp]),
    "tb_realtime_hz": (_i64, []),
    "tb_clock_probe": (_i32, [_vp, _i64, _vp]),
}

_lock = threading.Lock()
_lib = None


def load() -> ctypes.CDLL:
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} not found: build it with `python -m theboss_amd.build` "
                    "(or __graft_entry__.build())")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(rc: int) -> None:
    if rc != TB_OK:
        msg = load().tb_last_error().decode(errors="replace")
        raise TbError(rc, msg)


def exported_symbols():
    return list(SIGNATURES)
