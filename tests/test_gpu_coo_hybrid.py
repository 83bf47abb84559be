"""matrix::Coo and matrix::Hybrid on the device: the reference's
known-answer tests (reference/test/matrix/coo_kernels.cpp,
core/test/matrix/hybrid.cpp, reference/test/matrix/hybrid_kernels.cpp), the
index conversions and the Hybrid layout bit-exact against the oracle, and
the products within 4 k eps (|A||x|)_i."""
import json
import os

import numpy as np
import pytest

import oracle as ora
from ginkgo_amd import generators as gen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


@pytest.fixture(scope="module")
def gx():
    import ginkgo_amd.gko as gko
    from ginkgo_amd import _capi
    _capi.lib()
    return gko, gko.Mi355xExecutor.create(0, torch.cuda.current_stream().cuda_stream)


def dense(gko, ex, a, dt="f64"):
    a = np.asarray(a, dtype=np.float64 if dt == "f64" else np.float32)
    return gko.matrix.Dense(dt).from_numpy(ex, a.reshape(a.shape[0], -1))


def scalar(gko, ex, v, dt="f64"):
    return dense(gko, ex, np.array([[v]]), dt)


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("unsorted", [False, True])
def test_coo_reference_kats(gx, dt, unsorted):
    gko, ex = gx
    k = KATS["coo"]
    cols = k["unsorted_col_idxs"] if unsorted else k["col_idxs"]
    vals = k["unsorted_values"] if unsorted else k["values"]
    npdt = np.float64 if dt == "f64" else np.float32
    A = gko.matrix.Coo(dt).from_numpy(ex, k["n_rows"], k["n_cols"], np.array(k["row_idxs"]),
                                      np.array(cols), np.array(vals, npdt))
    for case in k["cases"]:
        x = dense(gko, ex, case["x"], dt)
        exp = np.array(case["expected"])
        if case["op"] == "apply":
            y = gko.matrix.Dense(dt).create(ex, k["n_rows"], exp.shape[1])
            A.apply(x, y)
        elif case["op"] == "advanced":
            y = dense(gko, ex, case["y"], dt)
            A.apply(scalar(gko, ex, case["alpha"], dt), x, scalar(gko, ex, case["beta"], dt), y)
        elif "alpha" in case:
            y = dense(gko, ex, case["y"], dt)
            A.apply2(scalar(gko, ex, case["alpha"], dt), x, y)
        else:
            y = dense(gko, ex, case["y"], dt)
            A.apply2(x, y)
        np.testing.assert_array_equal(y.to_numpy(), exp, err_msg=case["name"])
    csr = A.to_csr()
    assert csr.row_ptrs().tolist() == k["csr"]["row_ptrs"]
    assert csr.col_idxs().tolist() == cols


def test_coo_rejects_unsorted_rows(gx):
    gko, ex = gx
    with pytest.raises(gko.NotSupported):
        gko.matrix.Coo("f64").from_numpy(ex, 2, 2, np.array([1, 0]), np.array([0, 0]),
                                         np.ones(2))


@pytest.mark.parametrize("maker", ["powerlaw", "laplace", "empty_rows"])
def test_csr_coo_round_trip_and_products(gx, maker):
    """Csr -> Coo gives the row index of every entry (bit-exact); the Coo
    products match the oracle's entry-by-entry restatement within the FP
    bound, for every RHS count; Coo -> Csr restores the row pointers."""
    gko, ex = gx
    if maker == "powerlaw":
        rp, ci, v = gen.powerlaw(20000, seed=4)
    elif maker == "laplace":
        rp, ci, v = gen.laplace3d(16)
    else:
        rp = np.zeros(3001, np.int32)
        rp[2000:] = np.arange(1001)
        ci = np.arange(1000, dtype=np.int32)
        v = np.random.default_rng(1).uniform(-1, 1, 1000)
    n = len(rp) - 1
    rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(rp))
    A = gko.csr_from_host(ex, rp, ci, v)
    C = A.to_coo()
    np.testing.assert_array_equal(C.row_idxs(), rows)
    np.testing.assert_array_equal(C.to_csr().row_ptrs(), rp)
    rng = np.random.default_rng(2)
    a_abs = gen.csr_to_dense(rp, ci, np.abs(v), n)
    kk = np.maximum(np.diff(rp), 1)[:, None]
    for nrhs in (1, 3):
        x = rng.uniform(-1, 1, (n, nrhs))
        y0 = rng.uniform(-1, 1, (n, nrhs))
        y = dense(gko, ex, y0)
        C.apply2(scalar(gko, ex, -0.5), dense(gko, ex, x), y)
        ref = ora.coo_spmv2(rows, ci, v, x, y0, alpha=-0.5)
        bound = 4 * kk * np.finfo(np.float64).eps * 0.5 * (a_abs @ np.abs(x)) + 1e-14
        assert np.all(np.abs(y.to_numpy() - ref) <= bound + 4e-16 * np.abs(y0))
        z = gko.matrix.Dense("f64").create(ex, n, nrhs)
        C.apply(dense(gko, ex, x), z)
        exact = gen.csr_to_dense(rp, ci, v, n) @ x
        assert np.all(np.abs(z.to_numpy() - exact) <=
                      4 * kk * np.finfo(np.float64).eps * (a_abs @ np.abs(x)) + 1e-300)


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_hybrid_reference_kats(gx, dt):
    gko, ex = gx
    h = KATS["hybrid"]
    npdt = np.float64 if dt == "f64" else np.float32
    A = gko.csr_from_host(ex, np.array(h["csr"]["row_ptrs"], np.int32),
                          np.array(h["csr"]["col_idxs"], np.int32),
                          np.array(h["csr"]["values"], npdt), n_cols=h["n_cols"])
    for lay in h["layouts"]:
        H = A.to_hybrid(lay["strategy"], lay.get("p0", 0.0))
        assert H.ell_width == lay["ell_width"] and H.ell_stride == lay["ell_stride"]
        assert H.ell_col_idxs().tolist() == lay["ell_col_idxs"]
        assert H.ell_values().tolist() == lay["ell_values"]
        assert H.coo_row_idxs().tolist() == lay["coo_row_idxs"]
        assert H.coo_col_idxs().tolist() == lay["coo_col_idxs"]
        assert H.coo_values().tolist() == lay["coo_values"]
        # reference/test/matrix/hybrid_kernels.cpp:102-160 on mtx1 =
        # [[1, 3, 2], [0, 5, 0]], converted with the same strategy
        ap = h["apply"]
        mrp, mci, mv = gen.dense_to_csr(np.array(ap["matrix"]))
        H = gko.csr_from_host(ex, mrp, mci, mv.astype(npdt), n_cols=3).to_hybrid(
            lay["strategy"], lay.get("p0", 0.0))
        y = gko.matrix.Dense(dt).create(ex, 2, 1)
        H.apply(dense(gko, ex, ap["x"], dt), y)
        np.testing.assert_array_equal(y.to_numpy(), np.array(ap["expected"]))
        y = dense(gko, ex, ap["y"], dt)
        H.apply(scalar(gko, ex, ap["alpha"], dt), dense(gko, ex, ap["x"], dt),
                scalar(gko, ex, ap["beta"], dt), y)
        np.testing.assert_array_equal(y.to_numpy(), np.array(ap["expected_advanced"]))


@pytest.mark.parametrize("strategy,p0,p1", [("automatic", 0, 0), ("column_limit", 4, 0),
                                            ("imbalance_limit", 0.8, 0),
                                            ("imbalance_bounded_limit", 0.8, 0.0001),
                                            ("minimal_storage_limit", 0, 0)])
def test_hybrid_layout_bit_exact_and_product(gx, strategy, p0, p1):
    gko, ex = gx
    rp, ci, v = gen.powerlaw(30000, seed=9)
    n = len(rp) - 1
    A = gko.csr_from_host(ex, rp, ci, v)
    H = A.to_hybrid(strategy, p0, p1)
    w = ora.hybrid_ell_width(np.diff(rp), strategy, p0, p1)
    assert H.ell_width == w
    ec, ev, cr, cc, cv = ora.csr_to_hybrid(rp, ci, v, w, n)
    np.testing.assert_array_equal(H.ell_col_idxs(), ec)
    np.testing.assert_array_equal(H.ell_values(), ev)
    np.testing.assert_array_equal(H.coo_row_idxs(), cr)
    np.testing.assert_array_equal(H.coo_col_idxs(), cc)
    np.testing.assert_array_equal(H.coo_values(), cv)
    x = np.random.default_rng(5).uniform(-1, 1, (n, 1))
    y = gko.matrix.Dense("f64").create(ex, n, 1)
    H.apply(dense(gko, ex, x), y)
    exact = gen.csr_to_dense(rp, ci, v, n) @ x
    a_abs = gen.csr_to_dense(rp, ci, np.abs(v), n)
    kk = np.maximum(np.diff(rp), 1)[:, None]
    assert np.all(np.abs(y.to_numpy() - exact) <=
                  4 * kk * np.finfo(np.float64).eps * (a_abs @ np.abs(x)) + 1e-300)


def test_hybrid_drives_a_solve(gx):
    """A Hybrid as the system matrix of the generic CG path."""
    gko, ex = gx
    rp, ci, v = gen.poisson2d(30)
    n = len(rp) - 1
    b = np.random.default_rng(1).uniform(-1, 1, (n, 1))
    H = gko.csr_from_host(ex, rp, ci, v).to_hybrid("column_limit", 3)
    assert H.coo_nnz > 0
    s = (gko.solver.Cg.build("f64")
         .with_criteria(gko.stop.Iteration.build().with_max_iters(1000),
                        gko.stop.ResidualNorm.build("f64").with_reduction_factor(1e-10))
         .on(ex).generate(H))
    x = gko.matrix.Dense("f64").create(ex, n, 1)
    x.fill(0.0)
    s.apply(dense(gko, ex, b), x)
    d = gen.csr_to_dense(rp, ci, v)
    assert np.linalg.norm(b - d @ x.to_numpy()) <= 2e-10 * np.linalg.norm(b)
