"""CPU-side checks of the C ABI boundary: the library loads, exports every symbol the
header declares, and host-side argument validation rejects bad input before any launch
(no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "theboss_amd.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tb_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from theboss_amd import _lib

    return _lib.load()


def test_header_declares_expected_entry_points():
    names = header_functions()
    for required in ("tb_glynn_batch", "tb_guan_perm_batch", "tb_cc_submatrices_batch",
                     "tb_sweep_minors_batch", "tb_perm_grad_batch", "tb_gccb_sample",
                     "tb_last_error", "tb_version"):
        assert required in names


def test_library_exports_every_header_symbol(lib):
    for name in header_functions():
        assert hasattr(lib, name), f"{name} declared in the header but not exported"


def test_python_binding_covers_header(lib):
    from theboss_amd import _lib

    assert set(_lib.exported_symbols()) == set(header_functions())


def test_version_and_error_channel(lib):
    assert lib.tb_version() >= 100
    assert isinstance(lib.tb_last_error(), bytes)


def test_workspace_queries_are_host_only(lib):
    from theboss_amd import _lib, ops

    assert lib.tb_glynn_workspace_size(32, 16, 0) > 0
    assert lib.tb_glynn_workspace_size(8, 16, 0) == 0
    assert lib.tb_sweep_workspace_size(20, 4096) > 0
    assert lib.tb_perm_grad_workspace_size(20, 64) > 0
    cfg = ops.gccb_config(400, 20, 4096)
    assert lib.tb_gccb_workspace_size(ctypes.byref(cfg)) > 4096 * 20 * 20 * 16


@pytest.mark.parametrize("call", [
    lambda l: l.tb_glynn_batch(None, 49, 1, 49 * 49, 0, None, None, 0, None),
    lambda l: l.tb_glynn_batch(None, 4, 1, 16, 7, ctypes.c_void_p(8), None, 0, None),
    lambda l: l.tb_sweep_minors_batch(None, 41, 40, 41 * 40, 1, None, None, 0, None),
    lambda l: l.tb_perm_grad_batch(None, 50, 1, 2500, ctypes.c_void_p(8), None, None, 0, None),
    lambda l: l.tb_guan_perm_batch(None, 1, None, None, 1, None, None, 1, 1, 5, None, None),
    lambda l: l.tb_guan_perm_outer(None, 1, None, None, 1, 1, None, None, 1, 1, 5, None, None),
    # either side may be enumerated: a 33-slot output side exceeds the 32 input slots
    lambda l: l.tb_guan_perm_outer(*[ctypes.c_void_p(8), 40] + [ctypes.c_void_p(8)] * 2 +
                                   [33, 4] + [ctypes.c_void_p(8)] * 2 + [4, 4, 0] +
                                   [ctypes.c_void_p(8), None]),
    # more pairs than a launch indexes
    lambda l: l.tb_guan_perm_outer(*[ctypes.c_void_p(8), 40] + [ctypes.c_void_p(8)] * 2 +
                                   [4, 1 << 16] + [ctypes.c_void_p(8)] * 2 + [4, 1 << 16, 0] +
                                   [ctypes.c_void_p(8), None]),
    # more scratch than the static array of the no-workspace variant holds
    lambda l: l.tb_glynn_batch_noalloc(ctypes.c_void_p(8), 40, 256, 1600, 0, ctypes.c_void_p(8), None),
])
def test_invalid_arguments_fail_before_launch(lib, call):
    from theboss_amd import _lib

    rc = call(lib)
    assert rc in (_lib.TB_EINVAL, _lib.TB_ERANGE)
    assert lib.tb_last_error()


def test_zero_batch_is_a_no_op(lib):
    assert lib.tb_glynn_batch(None, 5, 0, 25, 0, None, None, 0, None) == 0
    assert lib.tb_sweep_minors_batch(None, 5, 4, 20, 0, None, None, 0, None) == 0
    assert lib.tb_guan_perm_outer(None, 1, None, None, 1, 0, None, None, 1, 7, 0, None, None) == 0


def test_tuning_table_roundtrip_and_errors(lib):
    from theboss_amd import _lib

    k = 20
    try:
        assert lib.tb_tune_get(_lib.TB_TUNE_SWEEP, k, 4096) == 0  # built-in plan by default
        base = lib.tb_sweep_workspace_size(k, 4096)
        assert lib.tb_tune_set(_lib.TB_TUNE_SWEEP, k, 4096, 8) == _lib.TB_OK
        assert lib.tb_tune_get(_lib.TB_TUNE_SWEEP, k, 4096) == 8
        # the bucket of 4096 (2049..4096) only: batch 4097 and batch 1 keep the built-in plan
        assert lib.tb_tune_get(_lib.TB_TUNE_SWEEP, k, 2049) == 8
        assert lib.tb_tune_get(_lib.TB_TUNE_SWEEP, k, 4097) == 0
        assert lib.tb_tune_get(_lib.TB_TUNE_SWEEP, k, 1) == 0
        # 2^(19-8-6) waves per item instead of 2^(19-11-6): 8x the partial sums
        assert lib.tb_sweep_workspace_size(k, 4096) == 8 * base
        assert lib.tb_tune_set(_lib.TB_TUNE_GLYNN, 32, 0, 10) == _lib.TB_OK  # every bucket
        assert all(lib.tb_tune_get(_lib.TB_TUNE_GLYNN, 32, 1 << b) == 10
                   for b in range(_lib.TB_TUNE_BATCH_BUCKETS))
        assert lib.tb_tune_get(_lib.TB_TUNE_GLYNN, 32, 1 << 40) == 10  # last bucket
        assert lib.tb_tune_set(_lib.TB_TUNE_GRAD, 20, 64, 9) == _lib.TB_OK
        assert lib.tb_tune_get(_lib.TB_TUNE_GRAD, 20, 33) == 9
        assert lib.tb_tune_get(_lib.TB_TUNE_GRAD, 20, 32) == 0
    finally:
        for kern, size in ((_lib.TB_TUNE_SWEEP, k), (_lib.TB_TUNE_GLYNN, 32), (_lib.TB_TUNE_GRAD, 20)):
            lib.tb_tune_set(kern, size, 0, 0)
    assert lib.tb_sweep_workspace_size(k, 4096) == base
    assert lib.tb_tune_set(7, 20, 1, 8) == _lib.TB_EINVAL
    assert lib.tb_tune_get(0, 20, 1) == _lib.TB_EINVAL
    assert lib.tb_tune_get(_lib.TB_TUNE_SWEEP, 20, 0) == _lib.TB_EINVAL
    assert lib.tb_tune_set(_lib.TB_TUNE_SWEEP, _lib.TB_TUNE_MAX_SIZE + 1, 1, 8) == _lib.TB_ERANGE
    assert lib.tb_tune_set(_lib.TB_TUNE_SWEEP, 20, 1, 17) == _lib.TB_EINVAL
    assert lib.tb_tune_set(_lib.TB_TUNE_SWEEP, 20, 1, -1) == _lib.TB_EINVAL


def test_sweep_plan_fills_the_gpu_for_small_batches(lib):
    """The built-in sweep plan depends on the batch: a single k = 20 item is cut into 2^5-step
    chunks (256 waves: 2^14 chunks x 20 minors of partial sums), 4096 items keep 2^11-step
    chunks (4 waves each)."""
    k = 20
    assert lib.tb_sweep_workspace_size(k, 1) == 256 * k * 16
    assert lib.tb_sweep_workspace_size(k, 4096) == 4096 * 4 * k * 16
    assert lib.tb_sweep_workspace_size(k, 16) == 16 * 256 * k * 16   # 4096 waves in total
    assert lib.tb_sweep_workspace_size(k, 256) == 256 * 16 * k * 16  # 4096 waves in total


def test_gradient_plan_matches_its_python_mirror(lib):
    """autotune.grad_builtin, the centre of the tuner's candidates, is the library's built-in
    gradient chunk: the workspace holds batch x wpi x n x (columns of the groups) partial
    sums with wpi = 2^min(n - 7 - c, 13), which fixes c. Shapes where the plan grows the
    chunk past the ~8192-wave choice (while the rounds stay full) and where it does not."""
    from theboss_amd import _lib, autotune

    for n, B in ((14, 1000), (16, 256), (17, 64), (20, 1), (20, 16), (20, 64), (22, 32),
                 (24, 16), (26, 4)):
        lib.tb_tune_set(_lib.TB_TUNE_GRAD, n, 0, 0)
        c = autotune.grad_builtin(n, B)
        g = autotune.grad_columns(n)
        cols = (n // 8) * 8 + 4 if g == 8 and 0 < n % 8 <= 4 else -(-n // g) * g
        want = B * 2 ** min(n - 7 - c, 13) * n * cols * 16
        assert lib.tb_perm_grad_workspace_size(n, B) == -(-want // 256) * 256, (n, B, c)
    assert autotune.grad_builtin(20, 64) == 9  # 2048 waves of 2^9 steps, not 8192 of 2^7


def test_sampler_workspace_follows_tuned_sweeps(lib):
    from theboss_amd import _lib, ops

    cfg = ops.gccb_config(400, 20, 4096)
    base = lib.tb_gccb_workspace_size(ctypes.byref(cfg))
    try:
        lib.tb_tune_set(_lib.TB_TUNE_SWEEP, 20, 4096, 7)
        assert lib.tb_gccb_workspace_size(ctypes.byref(cfg)) > base
        cfg.chunk_bits[20] = 11  # an explicit per-call value wins over the table
        assert lib.tb_gccb_workspace_size(ctypes.byref(cfg)) == base
    finally:
        lib.tb_tune_set(_lib.TB_TUNE_SWEEP, 20, 0, 0)


def test_glynn_fp32_workspace_follows_the_fp64_rule(lib):
    """FP32 column sums stop at n = 23: above, TB_PREC_FP32_COLSUM plans the FP64 launch
    (no FP32 matrix copy in the workspace), TB_PREC_FP32_COLSUM_ANY_N the FP32 one."""
    from theboss_amd import _lib

    for n in (16, 23, 24, 32):
        f64 = lib.tb_glynn_workspace_size(n, 16, _lib.TB_PREC_FP64)
        f32 = lib.tb_glynn_workspace_size(n, 16, _lib.TB_PREC_FP32_COLSUM)
        anyn = lib.tb_glynn_workspace_size(n, 16, _lib.TB_PREC_FP32_COLSUM_ANY_N)
        copy = 16 * n * n * 8  # FP32 complex copy of the batch
        assert anyn >= f64 + copy
        if n <= _lib.TB_FP32_COLSUM_MAX_N:
            assert f32 == anyn
        else:
            assert f32 == f64


def test_outcome_count_host_only(lib):
    from theboss_amd import _lib

    assert lib.tb_outcome_count(3, 4, 0) == 20   # C(6, 3)
    assert lib.tb_outcome_count(3, 4, 1) == 35   # C(7, 4): 0..3 particles
    assert lib.tb_outcome_count(0, 5, 0) == 1
    assert lib.tb_outcome_count(10, 20, 0) == 20030010
    assert lib.tb_outcome_count(-1, 4, 0) == _lib.TB_EINVAL
    assert lib.tb_outcome_count(3, 0, 0) == _lib.TB_EINVAL
    assert lib.tb_outcome_count(3, _lib.TB_HIST_MAX_M + 1, 0) == _lib.TB_ERANGE
    assert lib.tb_outcome_count(20, 40, 0) == _lib.TB_ERANGE  # more than 2^31 - 1
    assert lib.tb_outcome_histogram(None, 4, 0, 4, 3, 0, None, None, None) == _lib.TB_OK
    assert lib.tb_outcome_histogram(None, 4, 8, 4, 3, 0, None, None, None) == _lib.TB_EINVAL
    assert lib.tb_outcome_histogram(None, 2, -1, 4, 3, 0, None, None, None) == _lib.TB_EINVAL
