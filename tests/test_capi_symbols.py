"""The C-ABI libraries load on the CPU and export every declared symbol."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:gkomi|gko_c)_[a-z0-9_]+)\s*\(", text)))


def test_kernel_library_exports_every_declared_symbol():
    from ginkgo_amd import _capi
    lib = _capi.lib()
    names = declared("gkomi.h")
    assert len(names) > 80
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the binding's signature table only names real entry points
    assert set(_capi._SIGS) <= set(names)


def test_kernel_library_reports_errors_without_gpu():
    from ginkgo_amd import _capi
    lib = _capi.lib()
    assert b"gfx950" in lib.gkomi_version()
    # workspace sizing is pure host arithmetic
    assert lib.gkomi_reduction_workspace_bytes(1000, 1) > 0
    assert lib.gkomi_csr_transpose_workspace_bytes(10, 10, 30) > 0
    assert lib.gkomi_csr_lb_chunk(8) > 0
    # a null descriptor is rejected with a typed error, before any device call
    with pytest.raises(_capi.InvalidArgument):
        _capi.call("gkomi_csr_spmv_f64", None, None, 1, None, 1, None, 1)
