"""tb_glynn_batch_noalloc (scratch in a static device array) against tb_glynn_batch (caller
workspace), through the C ABI: the same kernels with the same plan, so the permanents are
bit-identical, on two calls in a row (the second smaller, so it reuses the array's first
bytes after a larger call), in FP64 and with FP32 column sums (which also stage the FP32
copy of the matrices in the array), and against the oracle. A call that needs more scratch
than the array holds is refused before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from theboss_amd import _lib

    _lib.load()
    return torch


def _blocks(oracle, n, B, seed):
    return np.stack([oracle.haar_unitary(2 * n, seed + b)[:n, :n] for b in range(B)])


def _noalloc(T, A, precision):
    from theboss_amd import _device as D
    from theboss_amd import _lib as L

    B, n = A.shape[0], A.shape[1]
    out = T.empty(B, dtype=T.complex128, device=A.device)
    L.check(L.load().tb_glynn_batch_noalloc(D.ptr(A), n, B, n * n, precision, D.ptr(out),
                                            D.stream_ptr(A.device)))
    return out


def test_noalloc_equals_workspace_entry_point(T, oracle):
    from theboss_amd import _lib as L
    from theboss_amd import ops

    for n, B, prec in ((28, 16, L.TB_PREC_FP64), (14, 5, L.TB_PREC_FP64),
                       (20, 8, L.TB_PREC_FP32_COLSUM), (12, 3, L.TB_PREC_FP32_COLSUM)):
        A = _blocks(oracle, n, B, 900 + 10 * n)
        Ad = T.from_numpy(A).cuda()
        ref = ops.glynn_batch(Ad, prec)
        got = _noalloc(T, Ad, prec)
        T.cuda.synchronize()
        assert T.equal(got, ref), (n, B, prec)
        want = oracle.perm_glynn_batch(A, 1)
        err = np.max(np.abs(got.cpu().numpy() - want) / np.abs(want))
        assert err <= (1e-10 if prec == L.TB_PREC_FP64 else 1e-4), (n, B, err)


def test_noalloc_refuses_more_than_its_scratch(T):
    from theboss_amd import _lib as L

    lib = L.load()
    n, B = 40, 256  # 256 x 8192 waves x 16 B of partial sums = 32 MiB
    assert lib.tb_glynn_workspace_size(n, B, L.TB_PREC_FP64) > L.TB_GLYNN_NOALLOC_SCRATCH_BYTES
    rc = lib.tb_glynn_batch_noalloc(ctypes.c_void_p(8), n, B, n * n, L.TB_PREC_FP64,
                                    ctypes.c_void_p(8), None)
    assert rc == L.TB_ERANGE
    assert b"scratch" in lib.tb_last_error()
