"""The drop-in boundary on the GPU: the load_balance partition, the SELL-P and
ELL advanced products, the mixed-precision products and the operand checks
of the Python layer.

* The partition (gkomi_csr_lb_partition) is index work: bit-exact against
  the oracle's restatement of Csr::load_balance::process
  (include/ginkgo/core/matrix/csr.hpp:417-468), and the load-balance SpMV
  accepts the reference's own srow array unchanged.
* Advanced products: the reference's KATs (reference/test/matrix/
  sellp_kernels.cpp:103-171, 460-500; ell_kernels.cpp:233-339, 566-604),
  beta = 0 with NaN in c included, then the oracle on larger matrices.
* Mixed products compute in the highest precision of the operands
  (reference/matrix/csr_kernels.cpp:47-121): against the oracle's double
  product of the widened inputs, within 4 k eps (|A||x|)_i of the arithmetic
  type plus one rounding of the output type.
"""
import json
import os

import numpy as np
import pytest

import oracle as ora
from ginkgo_amd import generators as gen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
DTYPES = [np.float64, np.float32]
TD = {np.float64: torch.float64, np.float32: torch.float32}
OTHER = {np.float64: np.float32, np.float32: np.float64}


@pytest.fixture(scope="module")
def dev():
    from ginkgo_amd import _capi
    _capi.lib()
    return torch.device("cuda", 0)


def D():
    from ginkgo_amd import device
    return device


def to_dev(a, dev):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dev)


def kat_matrix(name, dtype):
    m = KATS["matrices"][name]
    return (np.array(m["row_ptrs"], np.int32), np.array(m["col_idxs"], np.int32),
            np.array(m["values"], dtype), m["n_cols"])


def arr(x, dtype):
    return np.array([[float(v) for v in row] for row in x], dtype=dtype)


def abs_bound(rp, ci, v, x, eps):
    """4 k eps (|A||x|)_i through the oracle's SpMV on |A| (no dense matrix)."""
    ax = ora.csr_spmv(rp, ci, np.abs(v).astype(np.float64), np.abs(x.astype(np.float64)))
    k = np.maximum(np.diff(rp), 1).astype(np.float64)[:, None]
    return 4 * k * eps * ax


def irregular(n, seed):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 12, n)
    lens[:9] = 0
    lens[rng.integers(0, n, 4)] = rng.integers(200, 900, 4)
    lens[n // 2] = min(n, 4096)
    lens[-5:] = 0
    rows_cols = [np.sort(rng.choice(n, l, replace=False)) for l in lens]
    rp = np.zeros(n + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    ci = np.concatenate(rows_cols).astype(np.int32)
    return rp, ci, rng.uniform(-1, 1, ci.shape[0])


# ------------------------------------------------------ row partition ----

@pytest.mark.parametrize("form", ["hip", "cuda32", "tiny"])
@pytest.mark.parametrize("maker", ["irregular", "powerlaw", "laplace", "one_row"])
def test_lb_partition_bit_exact(dev, maker, form):
    if maker == "irregular":
        rp, ci, v = irregular(6000, seed=4)
    elif maker == "powerlaw":
        rp, ci, v = gen.powerlaw(50000, seed=6)
    elif maker == "laplace":
        rp, ci, v = gen.laplace3d(20)
    else:
        rp = np.array([0, 5000], np.int32)
        ci = np.arange(5000, dtype=np.int32)
        v = np.ones(5000)
    n = len(rp) - 1
    nnz = int(rp[-1])
    nwarps, ws, cuda = {"hip": (D().default_lb_warps(dev), 64, False),
                        "cuda32": (1024, 32, True), "tiny": (3, 16, False)}[form]
    A = D().DeviceCsr(rp, ci, v, n_cols=max(n, int(ci.max()) + 1), strategy="load_balance",
                      device=dev, lb_nwarps=nwarps, lb_warp_size=ws, lb_cuda_strategy=cuda)
    expect_n = ora.lb_clac_size(nnz, nwarps, ws, cuda)
    assert A.num_srow == expect_n
    np.testing.assert_array_equal(A.srow[:A.num_srow].cpu().numpy(),
                                  ora.lb_process(rp, expect_n, ws))
    # and the product over that partition meets the bound
    x = np.random.default_rng(1).uniform(-1, 1, (A.n_cols, 1))
    y = A.spmv(to_dev(x, dev)).cpu().numpy()
    ref = ora.csr_spmv(rp, ci, v, x)
    assert np.all(np.abs(y - ref) <= 2 * abs_bound(rp, ci, v, x, np.finfo(np.float64).eps))


def test_lb_partition_full_size_powerlaw(dev):
    """The 1M-row power-law matrix (lognormal rows up to 4096): the device
    partition equals the reference's process() on the HipExecutor's warps."""
    rp, ci, v = gen.powerlaw(1_000_000, seed=3)
    nnz = int(rp[-1])
    A = D().DeviceCsr(rp, ci, v, strategy="load_balance", device=dev)
    nw = D().default_lb_warps(dev)
    expect = ora.lb_process(rp, ora.lb_clac_size(nnz, nw, 64, False), 64)
    np.testing.assert_array_equal(A.srow[:A.num_srow].cpu().numpy(), expect)


def test_lb_spmv_takes_the_reference_srow(dev):
    """A maintainer passes Csr::get_const_srow() straight through: an srow
    built on the host by the reference's rule, uploaded as is, drives the
    load-balance kernel to the right product.  Also the KAT partition of
    core/test/matrix/csr.cpp:127-131 (one element, 0)."""
    from ginkgo_amd import _capi as C
    k = KATS["lb_partition"]
    rp, ci, v, nc = kat_matrix(k["matrix"], np.float64)
    assert ora.lb_clac_size(int(rp[-1]), k["nwarps"], k["warp_size"], k["cuda_strategy"]) == \
        k["num_srow"]
    np.testing.assert_array_equal(ora.lb_process(rp, k["num_srow"], k["warp_size"]), k["srow"])
    for (rp, ci, v, nc), nwarps, ws in (((rp, ci, v, nc), 1, 32),
                                        ((*irregular(3000, 9), 3000), 777, 32),
                                        ((*irregular(3000, 10), 3000), 50, 64)):
        A = D().DeviceCsr(rp, ci, v, n_cols=nc, strategy="classical", device=dev)
        srow = to_dev(ora.lb_process(rp, nwarps, ws), dev)
        A.srow, A.num_srow, A.strategy = srow, nwarps, "load_balance"
        A.desc = A._desc()
        x = np.random.default_rng(2).uniform(-1, 1, (nc, 1))
        c0 = np.random.default_rng(3).uniform(-1, 1, (len(rp) - 1, 1))
        y = A.spmv(to_dev(x, dev)).cpu().numpy()
        ref = ora.csr_spmv(rp, ci, v, x)
        bound = abs_bound(rp, ci, v, x, np.finfo(np.float64).eps)
        assert np.all(np.abs(y - ref) <= 2 * bound)
        c = to_dev(c0, dev)
        A.advanced_spmv(-2.0, to_dev(x, dev), 0.5, c)
        ref_a = ora.csr_advanced_spmv(-2.0, rp, ci, v, x, 0.5, c0)
        assert np.all(np.abs(c.cpu().numpy() - ref_a) <= 4 * bound + 1e-15)
    assert C.CSR_LOAD_BALANCE == 1


def test_lb_partition_through_the_cpp_strategy(dev):
    """Csr with load_balance(nwarps, warp_size, cuda_strategy) and with the
    default (the executor's warps) builds the reference's srow."""
    import ginkgo_amd.gko as gko
    ex = gko.Mi355xExecutor.create(0, torch.cuda.current_stream().cuda_stream)
    rp, ci, v = irregular(4000, seed=12)
    nnz = int(rp[-1])
    for strat, (nw, ws, cuda) in (("load_balance:2:32", (2, 32, True)),
                                  ("load_balance:100:64", (100, 64, True)),
                                  ("load_balance", (D().default_lb_warps(dev), 64, False))):
        A = gko.csr_from_host(ex, rp, ci, v, 4000, strat)
        assert A.effective_strategy == "load_balance"
        n_s = ora.lb_clac_size(nnz, nw, ws, cuda)
        assert A.num_srow_elements == n_s
        np.testing.assert_array_equal(A.srow(), ora.lb_process(rp, n_s, ws))
        x = np.random.default_rng(4).uniform(-1, 1, (4000, 1))
        b = gko.matrix.Dense("f64").from_numpy(ex, x)
        out = gko.matrix.Dense("f64").create(ex, 4000, 1)
        A.apply(b, out)
        y = out.to_numpy()
        ref = ora.csr_spmv(rp, ci, v, x)
        assert np.all(np.abs(y - ref) <= 2 * abs_bound(rp, ci, v, x, np.finfo(np.float64).eps))


# ----------------------------------------------- SELL-P / ELL advanced ----

@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", KATS["sellp_advanced"], ids=lambda c: c["name"])
def test_sellp_advanced_kats(dev, case, dtype):
    rp, ci, v, nc = kat_matrix(case["matrix"], dtype)
    A = D().DeviceCsr(rp, ci, v, n_cols=nc, device=dev)
    S = D().DeviceSellp(A, case["slice_size"], case["stride_factor"])
    c = to_dev(arr(case["y"], dtype), dev)
    S.advanced_spmv(case["alpha"], to_dev(arr(case["x"], dtype), dev), case["beta"], c)
    np.testing.assert_array_equal(c.cpu().numpy(), arr(case["expected"], dtype))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", KATS["ell_advanced"], ids=lambda c: c["name"])
def test_ell_advanced_kats(dev, case, dtype):
    rp, ci, v, nc = kat_matrix(case["matrix"], dtype)
    A = D().DeviceCsr(rp, ci, v, n_cols=nc, device=dev)
    E = D().DeviceEll(A, stride=case["stride"])
    c = to_dev(arr(case["y"], dtype), dev)
    E.advanced_spmv(case["alpha"], to_dev(arr(case["x"], dtype), dev), case["beta"], c)
    np.testing.assert_array_equal(c.cpu().numpy(), arr(case["expected"], dtype))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("fmt", ["sellp", "ell"])
def test_format_advanced_against_oracle(dev, fmt, dtype):
    rp, ci, v = gen.powerlaw(5000, mean=9, seed=31)
    v = v.astype(dtype)
    n = 5000
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (n, 2)).astype(dtype)
    c0 = rng.uniform(-1, 1, (n, 2)).astype(dtype)
    A = D().DeviceCsr(rp, ci, v, device=dev)
    if fmt == "sellp":
        M = D().DeviceSellp(A, 64, 1)
        sets, lens, cols, vals = ora.csr_to_sellp(rp, ci, v, 64, 1)
        ref = ora.sellp_advanced_spmv(1.5, n, 64, sets, lens, cols, vals, x, -0.25, c0)
    else:
        M = D().DeviceEll(A)
        w, s, cols, vals = ora.csr_to_ell(rp, ci, v)
        ref = ora.ell_advanced_spmv(1.5, n, w, s, cols, vals, x, -0.25, c0)
    c = to_dev(c0, dev)
    M.advanced_spmv(1.5, to_dev(x, dev), -0.25, c)
    eps = np.finfo(dtype).eps
    bound = 1.5 * abs_bound(rp, ci, v, x, eps) + 4 * eps * np.abs(c0.astype(np.float64))
    assert np.all(np.abs(c.cpu().numpy().astype(np.float64) - ref) <= 2 * bound + 1e-30)
    # beta = 0: c is ignored, NaN included
    cn = torch.full((n, 2), float("nan"), dtype=TD[dtype], device=dev)
    M.advanced_spmv(1.0, to_dev(x, dev), 0.0, cn)
    assert not torch.isnan(cn).any()


# ------------------------------------------------------- mixed precision ----

@pytest.mark.parametrize("mdt", DTYPES)
@pytest.mark.parametrize("fmt", ["csr", "sellp", "ell"])
def test_mixed_apply_kats(dev, fmt, mdt):
    k = KATS["mixed_apply"]
    rp, ci, v, nc = kat_matrix(k["matrix"], mdt)
    A = D().DeviceCsr(rp, ci, v, n_cols=nc, device=dev)
    M = {"csr": A, "sellp": D().DeviceSellp(A) if fmt == "sellp" else None,
         "ell": D().DeviceEll(A) if fmt == "ell" else None}[fmt]
    for bt, ct in k["combos"]:
        bdt = mdt if bt == "same" else OTHER[mdt]
        cdt = mdt if ct == "same" else OTHER[mdt]
        out = torch.empty((2, 1), dtype=TD[cdt], device=dev)
        M.spmv(to_dev(arr(k["apply"]["x"], bdt), dev), out)
        np.testing.assert_array_equal(out.cpu().numpy(), arr(k["apply"]["expected"], cdt))
        a = k["advanced"]
        c = to_dev(arr(a["y"], cdt), dev)
        M.advanced_spmv(torch.full((1,), a["alpha"], dtype=TD[mdt], device=dev),
                        to_dev(arr(a["x"], bdt), dev),
                        torch.full((1,), a["beta"], dtype=TD[cdt], device=dev), c)
        np.testing.assert_array_equal(c.cpu().numpy(), arr(a["expected"], cdt))


@pytest.mark.parametrize("nrhs", [1, 3])
@pytest.mark.parametrize("types", [(np.float32, np.float64, np.float64),
                                   (np.float64, np.float32, np.float32),
                                   (np.float32, np.float32, np.float64),
                                   (np.float64, np.float64, np.float32),
                                   (np.float32, np.float64, np.float32),
                                   (np.float64, np.float32, np.float64)],
                         ids=lambda t: "".join("d" if x == np.float64 else "s" for x in t))
@pytest.mark.parametrize("strategy", ["classical", "load_balance"])
def test_mixed_csr_against_widened_oracle(dev, types, nrhs, strategy):
    """Csr<float> x Dense<double> computes in double: the exact widening of the
    float values, no rounding to float anywhere (a product formed in float
    would miss the double-precision bound by orders of magnitude)."""
    ta, tb, tc = types
    rp, ci, v = irregular(8000, seed=21)
    v = v.astype(ta)
    rng = np.random.default_rng(22)
    x = rng.uniform(-1, 1, (8000, nrhs)).astype(tb)
    A = D().DeviceCsr(rp, ci, v, strategy=strategy, device=dev)
    out = torch.empty((8000, nrhs), dtype=TD[tc], device=dev)
    A.spmv(to_dev(x, dev), out)
    wide = ora.csr_spmv(rp, ci, v.astype(np.float64), x.astype(np.float64))
    eps_acc = np.finfo(np.float64).eps
    bound = abs_bound(rp, ci, v, x, eps_acc) + np.finfo(tc).eps * np.abs(wide)
    assert np.all(np.abs(out.cpu().numpy().astype(np.float64) - wide) <= 2 * bound + 1e-300)
    c0 = rng.uniform(-1, 1, (8000, nrhs)).astype(tc)
    c = to_dev(c0, dev)
    A.advanced_spmv(torch.full((1,), 0.5, dtype=TD[ta], device=dev), to_dev(x, dev),
                    torch.full((1,), -3.0, dtype=TD[tc], device=dev), c)
    ref = 0.5 * wide - 3.0 * c0.astype(np.float64)
    bound_a = 0.5 * abs_bound(rp, ci, v, x, eps_acc) + 4 * eps_acc * 3.0 * np.abs(c0) + \
        np.finfo(tc).eps * np.abs(ref)
    assert np.all(np.abs(c.cpu().numpy().astype(np.float64) - ref) <= 2 * bound_a + 1e-300)


@pytest.mark.parametrize("fmt", ["csr", "sellp", "ell"])
def test_mixed_apply_through_the_cpp_api(dev, fmt):
    """Csr/Sellp/Ell<float>::apply on Dense<double> (and the reverse) through
    the gko:: surface: no conversion of the operands through float."""
    import ginkgo_amd.gko as gko
    ex = gko.Mi355xExecutor.create(0, torch.cuda.current_stream().cuda_stream)
    rp, ci, v = gen.poisson2d(50)
    n = 2500
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, (n, 1))
    for mdt, vdt in (("f32", "f64"), ("f64", "f32")):
        A = gko.csr_from_host(ex, rp, ci, v.astype(np.float32 if mdt == "f32" else np.float64),
                              n, "classical")
        M = {"csr": A, "sellp": A.to_sellp() if fmt == "sellp" else None,
             "ell": A.to_ell() if fmt == "ell" else None}[fmt]
        npv = np.float64 if vdt == "f64" else np.float32
        b = gko.matrix.Dense(vdt).from_numpy(ex, x.astype(npv))
        out = gko.matrix.Dense(vdt).create(ex, n, 1)
        M.apply(b, out)
        y = out.to_numpy()
        vm = v.astype(np.float32 if mdt == "f32" else np.float64).astype(np.float64)
        wide = ora.csr_spmv(rp, ci, vm, x.astype(npv).astype(np.float64))
        bound = abs_bound(rp, ci, vm, x, np.finfo(np.float64).eps) + \
            np.finfo(npv).eps * np.abs(wide)
        assert np.all(np.abs(y.astype(np.float64) - wide) <= 2 * bound + 1e-300)
        alpha = gko.matrix.Dense(mdt).from_numpy(ex, np.array([[2.0]]).astype(
            np.float32 if mdt == "f32" else np.float64))
        beta = gko.matrix.Dense(vdt).from_numpy(ex, np.array([[0.0]]).astype(npv))
        cn = gko.matrix.Dense(vdt).from_numpy(ex, np.full((n, 1), np.nan).astype(npv))
        M.apply(alpha, b, beta, cn)
        yc = cn.to_numpy()
        assert not np.isnan(yc).any()
        assert np.all(np.abs(yc.astype(np.float64) - 2 * wide) <= 4 * bound + 1e-300)


# ------------------------------------------------- Python operand checks ----

def test_operand_checks_reject_before_launch(dev):
    from ginkgo_amd import _capi as C
    rp, ci, v = gen.poisson2d(10)
    A = D().DeviceCsr(rp, ci, v, device=dev)
    good = torch.ones(100, dtype=torch.float64, device=dev)
    with pytest.raises(C.InvalidArgument):
        A.spmv(torch.ones(100, dtype=torch.float64))              # host tensor
    with pytest.raises(C.DimensionMismatch):
        A.spmv(torch.ones(99, dtype=torch.float64, device=dev))   # wrong rows
    with pytest.raises(C.NotSupported):
        A.spmv(torch.ones(100, dtype=torch.int32, device=dev))    # not a value type
    with pytest.raises(C.NotSupported):
        A.spmv(torch.ones(200, dtype=torch.float64, device=dev)[::2])  # strided vector
    with pytest.raises(C.NotSupported):
        A.spmv(torch.ones(3, 100, dtype=torch.float64, device=dev).t())  # column stride 3
    with pytest.raises(C.DimensionMismatch):
        A.spmv(good, torch.empty((100, 2), dtype=torch.float64, device=dev))
    with pytest.raises(C.NotSupported):                            # overlapping rows
        A.advanced_spmv(1.0, good, 0.0, torch.zeros(2, dtype=torch.float64, device=dev)
                        .as_strided((100, 2), (0, 1)))
    with pytest.raises(C.NotSupported):
        D().dot(good, torch.ones(100, dtype=torch.float32, device=dev))
    with pytest.raises(C.DimensionMismatch):
        D().dot(good, torch.ones(50, dtype=torch.float64, device=dev))
    with pytest.raises(C.NotSupported):
        D().norm2(torch.ones(200, dtype=torch.float64, device=dev)[::2])
    with pytest.raises(C.NotSupported):
        D().copy(good, torch.empty(100, dtype=torch.float32, device=dev))
    # the valid calls still work
    assert torch.allclose(D().norm2(good), torch.tensor([10.0], dtype=torch.float64, device=dev))
    y = A.spmv(good)
    assert y.shape == (100, 1)
