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
])
def test_invalid_arguments_fail_before_launch(lib, call):
    from theboss_amd import _lib

    rc = call(lib)
    assert rc in (_lib.TB_EINVAL, _lib.TB_ERANGE)
    assert lib.tb_last_error()


def test_zero_batch_is_a_no_op(lib):
    assert lib.tb_glynn_batch(None, 5, 0, 25, 0, None, None, 0, None) == 0
    assert lib.tb_sweep_minors_batch(None, 5, 4, 20, 0, None, None, 0, None) == 0
