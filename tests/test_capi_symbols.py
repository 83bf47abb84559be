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
    # load_balance::clac_size is host arithmetic too; it equals the oracle's
    # restatement (pinned in test_oracle.py) over the thresholds of both forms
    import oracle as ora
    for nnz in (0, 1, 63, 64, 65, 4, 199_999, 200_000, 999_999, 1_000_000, 1_999_999,
                2_000_000, 9_999_999, 10_000_000, 16_000_000, 20_000_000, 200_000_000):
        for nwarps, ws, cuda, intel in ((1024, 64, 0, 0), (2, 32, 1, 0), (8192, 32, 1, 0),
                                        (512, 32, 1, 1), (3, 16, 0, 0)):
            assert lib.gkomi_csr_lb_clac_size(nnz, nwarps, ws, cuda, intel) == \
                ora.lb_clac_size(nnz, nwarps, ws, cuda, intel), (nnz, nwarps, ws, cuda, intel)
    # a null descriptor is rejected with a typed error, before any device call
    with pytest.raises(_capi.InvalidArgument):
        _capi.call("gkomi_csr_spmv_f64", None, None, 1, None, 1, None, 1)
    # an unknown value-type code of a mixed product is rejected the same way
    with pytest.raises(_capi.InvalidArgument):
        _capi.call("gkomi_csr_spmv_mixed", None, 7, None, 1, 0, None, 1, 0, None, 1)


def test_host_library_has_no_cpu_compute_path():
    """The gko:: host library loads without a GPU; LinOps on the host executor
    hold memory only and refuse to compute (no CPU fallback)."""
    import numpy as np
    import ginkgo_amd.gko as gko
    host = gko.HostExecutor.create()
    A = gko.csr_from_host(host, np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32),
                          np.array([2.0, 3.0]))
    b = gko.matrix.Dense("f64").from_numpy(host, np.ones((2, 1)))
    x = gko.matrix.Dense("f64").create(host, 2, 1)
    with pytest.raises(gko.NotSupported):
        A.apply(b, x)
    # factories resolve their parameters without touching a device
    params = (gko.solver.Cg.build("f64")
              .with_criteria(gko.stop.Iteration.build().with_max_iters(5),
                             gko.stop.ResidualNorm.build("f64").with_reduction_factor(1e-8)))
    assert params is not None


def test_host_batch_api_has_no_cpu_compute_path():
    """batch::solver::Bicgstab on the host executor: storage and the launch
    plan are built, the solve refuses (no CPU fallback)."""
    import numpy as np
    import ginkgo_amd.gko as gko
    host = gko.HostExecutor.create()
    rp = np.array([0, 1, 2], np.int32)
    A = gko.batch.matrix.Csr("f64").from_numpy(host, 2, 2, rp, np.array([0, 1], np.int32),
                                               np.ones(4))
    s = gko.batch.solver.Bicgstab.build("f64").with_tolerance(1e-8).on(host).generate(A)
    assert s.block_sizes == [64]
    b = gko.batch.MultiVector("f64").from_numpy(host, 2, 2, np.ones(4))
    x = gko.batch.MultiVector("f64").create(host, 2, 2)
    with pytest.raises(gko.NotSupported):
        s.apply(b, x)
