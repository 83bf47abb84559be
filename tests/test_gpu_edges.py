"""Edge cases the reference's tests exercise: empty matrices and vectors,
matrices without entries, a single very long row among short ones, zero
right-hand sides, one-row batch items.  Expected values follow from the
definitions (reference/matrix/csr_kernels.cpp:47-121: an empty row gives 0;
beta == 0 ignores c; core/solver/cg.cpp: b = 0 converges at iteration 0)."""
import numpy as np
import pytest

import oracle as ora
from ginkgo_amd import generators as gen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    from ginkgo_amd import _capi
    _capi.lib()
    return torch.device("cuda", 0)


def D():
    from ginkgo_amd import device
    return device


def test_spmv_on_zero_rows(dev):
    rp = np.zeros(1, np.int32)
    for strategy in ("classical", "load_balance"):
        A = D().DeviceCsr(rp, np.zeros(0, np.int32), np.zeros(0), n_cols=4,
                          strategy=strategy, device=dev)
        y = A.spmv(torch.ones(4, dtype=torch.float64, device=dev))
        assert y.numel() == 0


@pytest.mark.parametrize("strategy", ["classical", "load_balance"])
def test_matrix_without_entries(dev, strategy):
    n = 37
    A = D().DeviceCsr(np.zeros(n + 1, np.int32), np.zeros(0, np.int32), np.zeros(0),
                      strategy=strategy, device=dev)
    x = torch.ones(n, dtype=torch.float64, device=dev)
    y = torch.full((n,), 5.0, dtype=torch.float64, device=dev)
    A.spmv(x, y)
    assert torch.all(y == 0)
    c = torch.full((n, 1), 3.0, dtype=torch.float64, device=dev)
    A.advanced_spmv(1.0, x, 2.0, c)
    assert torch.all(c == 6.0)
    # conversions of an entry-free matrix
    S = D().DeviceSellp(A, 64, 1)
    y.fill_(5.0)
    S.spmv(x, y)
    assert torch.all(y == 0)
    E = D().DeviceEll(A)
    assert E.width == 0
    y.fill_(5.0)
    E.spmv(x, y)
    assert torch.all(y == 0)
    t_rp, t_ci, t_v = A.transpose()
    assert torch.all(t_rp == 0) and t_ci.numel() == 0


@pytest.mark.parametrize("strategy", ["classical", "load_balance"])
def test_one_long_row_among_short_ones(dev, strategy):
    rng = np.random.default_rng(4)
    n = 6000
    lens = rng.integers(0, 4, n)
    lens[1234] = 5000
    rows = np.repeat(np.arange(n), lens)
    cols = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in lens])
    vals = rng.uniform(-1, 1, cols.shape[0])
    rp = np.zeros(n + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    x = rng.uniform(-1, 1, (n, 1))
    A = D().DeviceCsr(rp, cols.astype(np.int32), vals, strategy=strategy, device=dev)
    y = A.spmv(torch.as_tensor(x).to(dev)).cpu().numpy()
    exact = gen.csr_to_dense(rp, cols, vals, n) @ x
    k = np.maximum(np.diff(rp), 1)[:, None]
    bound = 4 * k * np.finfo(np.float64).eps * (gen.csr_to_dense(rp, cols, np.abs(vals), n)
                                                @ np.abs(x)) + 1e-300
    assert np.all(np.abs(y - exact) <= bound)


def test_reductions_of_empty_vectors(dev):
    z = torch.zeros((0, 3), dtype=torch.float64, device=dev)
    assert torch.all(D().dot(z, z) == 0)
    assert torch.all(D().norm2(z) == 0)


def test_zero_rhs_converges_at_iteration_zero(dev):
    import ginkgo_amd.gko as gko
    rp, ci, v = gen.poisson2d(16)
    n = 256
    crit = [{"kind": "iteration", "max_iters": 100},
            {"kind": "residual_norm", "reduction_factor": 1e-10}]
    xo, so, io = ora.cg_solve(rp, ci, v, np.zeros((n, 1)), np.zeros((n, 1)), crit)
    assert io.iterations == 0
    A = D().DeviceCsr(rp, ci, v, strategy="classical", device=dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    solver = D().FusedCg(A, crit, chunk=4)
    assert solver.apply(torch.zeros(n, dtype=torch.float64, device=dev), x) == 0
    assert solver.status.item() == so[0]
    assert torch.all(x == 0)
    # the same through the C++ solver, two columns (fused block path)
    ex = gko.Mi355xExecutor.create(0, torch.cuda.current_stream().cuda_stream)
    G = gko.csr_from_host(ex, rp, ci, v)
    s = (gko.solver.Cg.build("f64")
         .with_criteria(gko.stop.Iteration.build().with_max_iters(100),
                        gko.stop.ResidualNorm.build("f64").with_reduction_factor(1e-10))
         .on(ex).generate(G))
    xb = gko.matrix.Dense("f64").create(ex, n, 2)
    xb.fill(0.0)
    s.apply(gko.matrix.Dense("f64").from_numpy(ex, np.zeros((n, 2))), xb)
    assert s.last_iterations == 0
    assert np.all(xb.to_numpy() == 0)


def test_batch_items_of_one_row(dev):
    # three 1 x 1 systems among larger ones: x = b / a
    row_off = np.array([0, 1, 4, 5, 6], np.int64)
    d = [np.array([[4.0]]), np.array([[4.0, -1, 0], [-1, 4, -1], [0, -1, 4]]),
         np.array([[2.0]]), np.array([[-8.0]])]
    rp, ci, vals = [0], [], []
    for m in d:
        r, c, v = gen.dense_to_csr(m)
        rp.extend((np.asarray(r[1:]) + rp[-1]).tolist())
        ci.extend(c.tolist())
        vals.extend(v.tolist())
    rp, ci, vals = np.array(rp, np.int32), np.array(ci, np.int32), np.array(vals)
    b = np.array([1.0, 1, 2, 3, 4, 4])
    A = D().DeviceBatchCsr(row_off, rp, ci, vals, device=dev)
    x = torch.zeros(6, dtype=torch.float64, device=dev)
    it, res = A.solve(torch.as_tensor(b).to(dev), x,
                      A.settings(tol=1e-12, relative=True, max_iters=50, jacobi=True))
    xh = x.cpu().numpy()
    np.testing.assert_allclose(xh[[0, 4, 5]], [0.25, 2.0, -0.5], rtol=1e-14)
    np.testing.assert_allclose(d[1] @ xh[1:4], b[1:4], rtol=0, atol=1e-11)
