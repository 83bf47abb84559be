"""The merge_path Csr strategy (csr.hpp:283-301; merge-path SpMV after
Merrill & Garland, SC16): y = A x and the advanced product against the
oracle within 4 k eps (|A||x|)_i, on matrices whose rows cross thread and
workgroup shares of the merge (long rows, empty rows, no entries)."""
import numpy as np
import pytest

import oracle as ora
from ginkgo_amd import generators as gen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
DTYPES = [np.float64, np.float32]


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


def check(rp, ci, v, x, y, dtype):
    n = len(rp) - 1
    exact = gen.csr_to_dense(rp, ci, v.astype(np.float64), x.shape[0]) @ x.astype(np.float64)
    a = gen.csr_to_dense(rp, ci, np.abs(v).astype(np.float64), x.shape[0])
    k = np.maximum(np.diff(rp), 1).astype(np.float64)[:, None]
    bound = 4 * k * np.finfo(dtype).eps * (a @ np.abs(x.astype(np.float64))) + 1e-30
    assert y.shape[0] == n
    assert np.all(np.abs(y.astype(np.float64) - exact) <= bound)


def rows_of(lens, n_cols, rng):
    rp = np.zeros(len(lens) + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    ci = np.concatenate([np.sort(rng.choice(n_cols, l, replace=False)) for l in lens] +
                        [np.zeros(0, np.int64)]).astype(np.int32)
    return rp, ci, rng.uniform(-1, 1, ci.shape[0])


def make(maker):
    rng = np.random.default_rng(11)
    if maker == "powerlaw":
        return gen.powerlaw(30000, seed=7)
    if maker == "laplace":
        return gen.laplace3d(18)
    if maker == "long_rows":
        # rows longer than a workgroup's share (2048 merge steps) next to
        # short and empty rows
        lens = rng.integers(0, 6, 8000)
        lens[[5, 6, 4000, 7999]] = [7000, 2047, 5000, 3000]
        return rows_of(lens, 8000, rng)
    if maker == "empty_rows":
        lens = np.zeros(9000, np.int64)
        lens[rng.integers(0, 9000, 300)] = 3
        return rows_of(lens, 9000, rng)
    if maker == "single_row":
        return rows_of(np.array([6000]), 6000, rng)
    raise ValueError(maker)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("maker", ["powerlaw", "laplace", "long_rows", "empty_rows",
                                   "single_row"])
def test_merge_path_spmv(dev, dtype, maker):
    rp, ci, v = make(maker)
    v = v.astype(dtype)
    n_cols = int(ci.max()) + 1 if ci.size else len(rp) - 1
    n_cols = max(n_cols, len(rp) - 1)
    x = np.random.default_rng(3).uniform(-1, 1, (n_cols, 1)).astype(dtype)
    A = D().DeviceCsr(rp, ci, v, n_cols=n_cols, strategy="merge_path", device=dev)
    y = A.spmv(to_dev(x, dev)).cpu().numpy()
    check(rp, ci, v, x, y, dtype)
    # deterministic: a second product is bit-identical
    np.testing.assert_array_equal(A.spmv(to_dev(x, dev)).cpu().numpy(), y)


@pytest.mark.parametrize("beta", [0.0, 2.0])
def test_merge_path_advanced(dev, beta):
    rp, ci, v = make("long_rows")
    n = len(rp) - 1
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (n, 1))
    c0 = rng.uniform(-1, 1, (n, 1))
    if beta == 0.0:
        c0[::7] = np.nan        # beta == 0 ignores c, NaN included
    A = D().DeviceCsr(rp, ci, v, strategy="merge_path", device=dev)
    c = to_dev(c0, dev)
    A.advanced_spmv(-0.75, to_dev(x, dev), beta, c)
    ref = ora.csr_advanced_spmv(-0.75, rp, ci, v, x, beta, c0)
    got = c.cpu().numpy()
    assert np.all(np.isfinite(got))
    a = gen.csr_to_dense(rp, ci, np.abs(v))
    k = np.maximum(np.diff(rp), 1)[:, None]
    bound = 4 * k * np.finfo(np.float64).eps * 0.75 * (a @ np.abs(x)) + 1e-14
    assert np.all(np.abs(got - ref) <= bound)


def test_merge_path_no_entries_and_kat(dev):
    A = D().DeviceCsr(np.zeros(4001, np.int32), np.zeros(0, np.int32), np.zeros(0),
                      strategy="merge_path", device=dev)
    y = torch.full((4000,), 5.0, dtype=torch.float64, device=dev)
    A.spmv(torch.ones(4000, dtype=torch.float64, device=dev), y)
    assert torch.all(y == 0)
    # the reference's KAT AppliesToDenseVector (reference/test/matrix/
    # csr_kernels.cpp:353-364): [[1, 3, 2], [0, 5, 0]] [2, 1, 4]^T = [13, 5]
    import json
    import os
    kats = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                       "reference_kats.json")))
    case = kats["spmv"][0]
    m = kats["matrices"][case["matrix"]]
    K = D().DeviceCsr(m["row_ptrs"], m["col_idxs"], np.array(m["values"]), m["n_cols"],
                      strategy="merge_path", device=dev)
    yk = K.spmv(to_dev(np.array(case["x"]), dev)).cpu().numpy()
    np.testing.assert_array_equal(yk, np.array(case["expected"]))


def test_cpp_csr_merge_path_in_a_solve(dev):
    """Csr with the merge_path strategy drives the generic CG path."""
    import ginkgo_amd.gko as gko
    ex = gko.Mi355xExecutor.create(0, torch.cuda.current_stream().cuda_stream)
    rp, ci, v = gen.poisson2d(40)
    n = len(rp) - 1
    b = np.random.default_rng(1).uniform(-1, 1, (n, 1))
    A = gko.csr_from_host(ex, rp, ci, v, strategy="merge_path")
    assert A.effective_strategy == "merge_path"
    s = (gko.solver.Cg.build("f64")
         .with_criteria(gko.stop.Iteration.build().with_max_iters(1000),
                        gko.stop.ResidualNorm.build("f64").with_reduction_factor(1e-10))
         .on(ex).generate(A))
    x = gko.matrix.Dense("f64").create(ex, n, 1)
    x.fill(0.0)
    s.apply(gko.matrix.Dense("f64").from_numpy(ex, b), x)
    assert not s.last_used_fused_path
    d = gen.csr_to_dense(rp, ci, v)
    assert np.linalg.norm(b - d @ x.to_numpy()) <= 2e-10 * np.linalg.norm(b)
    # sparselib resolves as automatical does (classical here)
    S = gko.csr_from_host(ex, rp, ci, v, strategy="sparselib")
    assert S.effective_strategy == "classical"
