"""The gko:: C++ API on the MI355X executor, driven through the pybind11
binding, against the reference's known answers and the oracle.

These tests read like the reference's reference/test/solver/*.cpp: build a
solver factory with criteria, generate it on a matrix, apply it, and compare
with the known solution at the reference's tolerance r<T> = 10 eps.
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
DTYPES = ["f64", "f32"]
NP = {"f64": np.float64, "f32": np.float32}
TD = {"f64": torch.float64, "f32": torch.float32}


@pytest.fixture(scope="module")
def gko():
    import ginkgo_amd.gko as g
    return g


@pytest.fixture(scope="module")
def exec_(gko):
    return gko.Mi355xExecutor.create(0, torch.cuda.current_stream().cuda_stream)


def r_tol(dt, spec):
    eps = np.finfo(NP[dt]).eps
    if spec == "r":
        return 10 * eps
    if isinstance(spec, str) and spec.startswith("r*"):
        return 10 * eps * float(spec[2:])
    return float(spec)


def rel_frob(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = max(np.sum(a * a), np.sum(b * b))
    return 0.0 if den == 0 else float(np.sqrt(np.sum((a - b) ** 2) / den))


def csr(gko, exec_, name, dt, strategy="automatical"):
    m = KATS["matrices"][name]
    if "dense" in m:
        rp, ci, v = gen.dense_to_csr(np.array(m["dense"]))
        nc = len(m["dense"][0])
    else:
        rp, ci, v = np.array(m["row_ptrs"]), np.array(m["col_idxs"]), np.array(m["values"])
        nc = m["n_cols"]
    return gko.csr_from_host(exec_, rp.astype(np.int32), ci.astype(np.int32), v.astype(NP[dt]),
                             nc, strategy)


def dense(gko, exec_, a, dt):
    return gko.matrix.Dense(dt).from_numpy(exec_, np.array(a, dtype=NP[dt]))


def crits(gko, spec, dt):
    out = []
    for c in spec:
        if c["kind"] == "iteration":
            out.append(gko.stop.Iteration.build().with_max_iters(c["max_iters"]))
        else:
            maker = (gko.stop.ResidualNorm if c["kind"] == "residual_norm"
                     else gko.stop.ImplicitResidualNorm)
            out.append(maker.build(dt).with_reduction_factor(r_tol(dt, c["reduction_factor"])))
    return out


@pytest.mark.parametrize("strategy", ["classical", "load_balance"])
@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("case", KATS["spmv"], ids=lambda c: c["name"])
def test_csr_apply_kats(gko, exec_, case, dt, strategy):
    A = csr(gko, exec_, case["matrix"], dt, strategy)
    assert A.effective_strategy == strategy
    x = dense(gko, exec_, case["x"], dt)
    if "alpha" in case:
        y = dense(gko, exec_, [[float(v) for v in r] for r in case["y"]], dt)
        A.apply(dense(gko, exec_, [[case["alpha"]]], dt), x,
                dense(gko, exec_, [[case["beta"]]], dt), y)
    else:
        y = gko.matrix.Dense(dt).create(exec_, A.size[0], x.size[1])
        A.apply(x, y)
    np.testing.assert_array_equal(y.to_numpy(), np.array(case["expected"], NP[dt]))


def test_conversion_kats(gko, exec_):
    A = csr(gko, exec_, "mtx_2x3", "f64")
    S = A.to_sellp()
    k = KATS["sellp_layout"]
    assert list(S.slice_sets()) == k["slice_sets"]
    cols, vals = S.col_idxs(), S.values()
    for p, c in k["cols_at"].items():
        assert cols[int(p)] == c
    for p, c in k["vals_at"].items():
        assert vals[int(p)] == c
    E = A.to_ell()
    assert list(E.col_idxs()) == KATS["ell_layout"]["cols"]
    assert list(E.values()) == KATS["ell_layout"]["vals"]
    for case in KATS["transpose"]:
        T = csr(gko, exec_, case["matrix"], "f64").transpose()
        d = gen.csr_to_dense(T.row_ptrs(), T.col_idxs(), T.values(), T.size[1])
        np.testing.assert_array_equal(d, np.array(case["expected_dense"]))
    # SpMV through the converted formats
    x = dense(gko, exec_, [[2.0], [1.0], [4.0]], "f64")
    for op in (S, E):
        y = gko.matrix.Dense("f64").create(exec_, 2, 1)
        op.apply(x, y)
        np.testing.assert_array_equal(y.to_numpy(), [[13.0], [5.0]])


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("case", KATS["solves"], ids=lambda c: c["name"])
def test_solver_kats(gko, exec_, case, dt):
    A = csr(gko, exec_, case["matrix"], dt)
    b = dense(gko, exec_, case["b"], dt)
    x = dense(gko, exec_, case["x0"], dt)
    c = crits(gko, case["criteria"], dt)
    if case["solver"] == "cg":
        factory = gko.solver.Cg.build(dt).with_criteria(*c).on(exec_)
    elif case["solver"] == "bicgstab":
        factory = gko.solver.Bicgstab.build(dt).with_criteria(*c).on(exec_)
    else:
        factory = gko.solver.Ir.build(dt).with_criteria(*c).on(exec_)
    s = factory.generate(A)
    if "alpha" in case:
        s.apply(dense(gko, exec_, [[case["alpha"]]], dt), b, dense(gko, exec_, [[case["beta"]]], dt),
                x)
    else:
        s.apply(b, x)
    assert rel_frob(x.to_numpy(), case["expected"]) <= r_tol(dt, case["tol"])


@pytest.mark.parametrize("case", KATS["jacobi"], ids=lambda c: c["name"])
def test_jacobi_kats(gko, exec_, case):
    A = csr(gko, exec_, case["matrix"], "f64")
    J = gko.preconditioner.Jacobi.build("f64").with_max_block_size(1).on(exec_).generate(A)
    b = dense(gko, exec_, case["b"], "f64")
    if "alpha" in case:
        x = dense(gko, exec_, case["x0"], "f64")
        J.apply(dense(gko, exec_, [[case["alpha"]]], "f64"), b,
                dense(gko, exec_, [[case["beta"]]], "f64"), x)
    else:
        x = gko.matrix.Dense("f64").create(exec_, *b.size)
        J.apply(b, x)
    assert rel_frob(x.to_numpy(), case["expected"]) <= r_tol("f64", case["tol"])


def test_block_jacobi_is_not_supported(gko, exec_):
    A = csr(gko, exec_, "jacobi_5x5", "f64")
    with pytest.raises(gko.NotSupported):
        gko.preconditioner.Jacobi.build("f64").with_max_block_size(4).on(exec_).generate(A)


def test_dimension_mismatch(gko, exec_):
    A = csr(gko, exec_, "mtx_2x3", "f64")
    x = dense(gko, exec_, [[1.0], [2.0]], "f64")
    y = gko.matrix.Dense("f64").create(exec_, 2, 1)
    with pytest.raises(gko.DimensionMismatch):
        A.apply(x, y)


@pytest.mark.parametrize("dt", DTYPES)
def test_config1_poisson_cg_jacobi(gko, exec_, dt):
    """256^2 5-pt Poisson, CG + scalar Jacobi to 1e-10 (<= 2000 iterations),
    b = A x_true with x_true ~ U[-1,1] (seed 1): the reference's plain-method
    configuration on the fused device path, against the oracle."""
    rp, ci, v = gen.poisson2d(256)
    v = v.astype(NP[dt])
    n = 65536
    xt = np.random.default_rng(1).uniform(-1, 1, (n, 1)).astype(NP[dt])
    b_h = ora.csr_spmv(rp, ci, v, xt)
    tol = 1e-10 if dt == "f64" else 1e-5
    spec = [{"kind": "iteration", "max_iters": 2000},
            {"kind": "residual_norm", "reduction_factor": tol}]
    inv = ora.jacobi_inverse_diagonal(rp, ci, v)
    xo, so, io = ora.cg_solve(rp, ci, v, b_h, np.zeros_like(b_h), spec, inv_diag=inv)
    A = gko.csr_from_host(exec_, rp, ci, v)
    solver = (gko.solver.Cg.build(dt).with_criteria(*crits(gko, spec, dt))
              .with_preconditioner(gko.preconditioner.Jacobi.build(dt).with_max_block_size(1))
              .on(exec_).generate(A))
    logger = gko.log.Convergence.create()
    solver.add_logger(logger)
    b = dense(gko, exec_, b_h, dt)
    x = gko.matrix.Dense(dt).create(exec_, n, 1)
    x.fill(0.0)
    solver.apply(b, x)
    assert solver.last_used_fused_path
    band = 3 if dt == "f64" else max(3, io.iterations // 10)
    assert abs(solver.last_iterations - io.iterations) <= band
    assert logger.num_iterations == solver.last_iterations
    assert logger.has_converged
    assert solver.last_status == [int(so[0])]
    xh = x.to_numpy().astype(np.float64)
    # true residual in FP64 through the oracle's O(nnz) SpMV
    r = b_h.astype(np.float64) - ora.csr_spmv(rp, ci, v.astype(np.float64), xh)
    res = np.linalg.norm(r) / np.linalg.norm(b_h.astype(np.float64))
    assert res <= tol * (2 if dt == "f64" else 20)
    err = np.max(np.abs(xh - xt.astype(np.float64)))
    print(f"C1 {dt}: {solver.last_iterations} iterations (oracle {io.iterations}), "
          f"relative residual {res:.2e}, max |x - x_true| {err:.2e}")


@pytest.mark.parametrize("k,jacobi,dt", [(2, False, "f64"), (5, True, "f64"), (64, False, "f64"),
                                          (70, True, "f64"), (7, False, "f32")])
def test_cg_many_rhs_block_path(gko, exec_, k, jacobi, dt):
    """k right-hand sides on the fused block path: per-column scalars and
    stopping status against the oracle's restatement of cg.cpp."""
    rp, ci, v = gen.poisson2d(40)
    v = v.astype(NP[dt])
    n = 1600
    rng = np.random.default_rng(k)
    b_h = rng.uniform(-1, 1, (n, k)).astype(NP[dt])
    tol = 1e-10 if dt == "f64" else 1e-5
    spec = [{"kind": "iteration", "max_iters": 1000},
            {"kind": "residual_norm", "reduction_factor": tol}]
    inv = ora.jacobi_inverse_diagonal(rp, ci, v) if jacobi else None
    xo, so, io = ora.cg_solve(rp, ci, v, b_h, np.zeros_like(b_h), spec, inv_diag=inv)
    A = gko.csr_from_host(exec_, rp, ci, v)
    params = gko.solver.Cg.build(dt).with_criteria(*crits(gko, spec, dt))
    if jacobi:
        params = params.with_preconditioner(
            gko.preconditioner.Jacobi.build(dt).with_max_block_size(1))
    s = params.on(exec_).generate(A)
    x = gko.matrix.Dense(dt).create(exec_, n, k)
    x.fill(0.0)
    s.apply(dense(gko, exec_, b_h, dt), x)
    assert s.last_used_fused_path
    band = 3 if dt == "f64" else max(3, io.iterations // 10)
    assert abs(s.last_iterations - io.iterations) <= band
    if dt == "f64":
        assert s.last_status == [int(c) for c in so]
        np.testing.assert_allclose(x.to_numpy(), xo, atol=1e-7)
    d = gen.csr_to_dense(rp, ci, v.astype(np.float64))
    xh = x.to_numpy().astype(np.float64)
    res = np.linalg.norm(b_h - d @ xh, axis=0) / np.linalg.norm(b_h, axis=0)
    assert np.all(res <= tol * (2 if dt == "f64" else 20))


def test_cg_many_rhs_block_path_long_rows(gko, exec_):
    """36 right-hand sides (one wave per row in the product) on a 27-point
    matrix: rows of 8..27 entries take several rounds of gathers per row.
    Iterations, per-column status and x against the oracle's cg.cpp."""
    rp, ci, v = gen.diffusion27(9)
    n, k = len(rp) - 1, 36
    rng = np.random.default_rng(36)
    b_h = rng.uniform(-1, 1, (n, k))
    spec = [{"kind": "iteration", "max_iters": 1000},
            {"kind": "residual_norm", "reduction_factor": 1e-10}]
    xo, so, io = ora.cg_solve(rp, ci, v, b_h, np.zeros_like(b_h), spec)
    A = gko.csr_from_host(exec_, rp, ci, v)
    s = gko.solver.Cg.build("f64").with_criteria(*crits(gko, spec, "f64")).on(exec_).generate(A)
    x = gko.matrix.Dense("f64").create(exec_, n, k)
    x.fill(0.0)
    s.apply(dense(gko, exec_, b_h, "f64"), x)
    assert s.last_used_fused_path
    assert abs(s.last_iterations - io.iterations) <= 3
    assert s.last_status == [int(c) for c in so]
    np.testing.assert_allclose(x.to_numpy(), xo, atol=1e-8)


def test_cg_many_rhs_block_path_4gib_vectors(gko, exec_):
    """k-RHS CG whose n x k blocks span 2^32 bytes (2^24 rows, 64 FP32
    columns), so the direction kernel indexes them with 64-bit offsets.
    Column j's right-hand side is 2^(j-32) u: every CG operation commutes
    with a power-of-two scale, so x[:, j] must be exactly 2^(j-32) x[:, 32],
    which fails if any column reads another's entries.  Column 32 matches
    the oracle's CG on u (cg.cpp) to FP32 rounding."""
    n, k, iters = 1 << 24, 64, 3
    rp = np.zeros(n + 1, np.int32)
    lens = np.full(n, 3, np.int32)
    lens[0] = lens[-1] = 2
    rp[1:] = np.cumsum(lens)
    rows = np.repeat(np.arange(n, dtype=np.int64), lens)
    ci = (rows + (np.arange(rp[-1]) - rp[rows]) - (rows > 0)).astype(np.int32)
    v = np.where(ci == rows, np.float32(3.0), np.float32(-1.0)).astype(np.float32)
    del rows
    u = np.random.default_rng(17).uniform(0.5, 1.5, n).astype(np.float32)
    scales = np.ldexp(np.float32(1), np.arange(k) - 32).astype(np.float32)
    b_h = u[:, None] * scales[None, :]
    spec = [{"kind": "iteration", "max_iters": iters},
            {"kind": "residual_norm", "reduction_factor": 1e-6}]
    A = gko.csr_from_host(exec_, rp, ci, v)
    s = gko.solver.Cg.build("f32").with_criteria(*crits(gko, spec, "f32")).on(exec_).generate(A)
    b = gko.matrix.Dense("f32").from_numpy(exec_, b_h)
    del b_h
    x = gko.matrix.Dense("f32").create(exec_, n, k)
    x.fill(0.0)
    s.apply(b, x)
    assert s.last_used_fused_path
    assert s.last_iterations == iters
    del b
    xh = x.to_numpy()
    del x
    ref = xh[:, 32].copy()
    for j in range(k):
        assert np.array_equal(xh[:, j], ref * scales[j]), j
    del xh
    xo, _, io = ora.cg_solve(rp, ci, v.astype(np.float64), u.astype(np.float64)[:, None],
                             np.zeros((n, 1)), spec)
    assert io.iterations == iters
    np.testing.assert_allclose(ref, xo[:, 0], rtol=1e-4, atol=1e-5 * np.abs(xo).max())


def test_cg_block_per_column_stopping(gko, exec_):
    """Columns that stop at different iterations keep their p, x and r
    (reference/solver/cg_kernels.cpp:55-95): a zero column stops at
    iteration 0, an eigenvector column at iteration 1, the rest later."""
    nx = 30
    rp, ci, v = gen.poisson2d(nx)
    n = nx * nx
    i = np.arange(nx) + 1
    eig = np.outer(np.sin(np.pi * i / (nx + 1)), np.sin(2 * np.pi * i / (nx + 1))).reshape(-1)
    rng = np.random.default_rng(5)
    b_h = np.stack([np.zeros(n), eig, rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)], axis=1)
    spec = [{"kind": "iteration", "max_iters": 500},
            {"kind": "residual_norm", "reduction_factor": 1e-10}]
    xo, so, io = ora.cg_solve(rp, ci, v, b_h, np.zeros_like(b_h), spec)
    A = gko.csr_from_host(exec_, rp, ci, v)
    s = gko.solver.Cg.build("f64").with_criteria(*crits(gko, spec, "f64")).on(exec_).generate(A)
    x = gko.matrix.Dense("f64").create(exec_, n, 4)
    x.fill(0.0)
    s.apply(dense(gko, exec_, b_h, "f64"), x)
    assert s.last_used_fused_path
    assert s.last_status == [int(c) for c in so]
    assert abs(s.last_iterations - io.iterations) <= 3
    xh = x.to_numpy()
    assert np.all(xh[:, 0] == 0.0)
    np.testing.assert_allclose(xh, xo, atol=1e-9)


def test_cg_many_rhs_generic_path(gko, exec_):
    """A non-Csr system matrix (Ell) takes the generic reference-order loop."""
    rp, ci, v = gen.poisson2d(48)
    n = 48 * 48
    rng = np.random.default_rng(3)
    b_h = rng.uniform(-1, 1, (n, 5))
    spec = [{"kind": "iteration", "max_iters": 1000},
            {"kind": "residual_norm", "reduction_factor": 1e-10}]
    xo, so, io = ora.cg_solve(rp, ci, v, b_h, np.zeros_like(b_h), spec)
    A = gko.csr_from_host(exec_, rp, ci, v).to_ell()
    s = gko.solver.Cg.build("f64").with_criteria(*crits(gko, spec, "f64")).on(exec_).generate(A)
    x = gko.matrix.Dense("f64").create(exec_, n, 5)
    x.fill(0.0)
    s.apply(dense(gko, exec_, b_h, "f64"), x)
    assert not s.last_used_fused_path
    assert abs(s.last_iterations - io.iterations) <= 3
    assert s.last_status == [int(v) for v in so]
    np.testing.assert_allclose(x.to_numpy(), xo, atol=1e-7)


def test_bicgstab_nonsymmetric_jacobi(gko, exec_):
    rng = np.random.default_rng(11)
    n = 400
    d = (rng.random((n, n)) < 0.02) * rng.uniform(-1, 1, (n, n))
    d += np.diag(np.abs(d).sum(1) + 1)
    rp, ci, v = gen.dense_to_csr(d)
    b_h = rng.uniform(-1, 1, (n, 2))
    spec = [{"kind": "iteration", "max_iters": 500},
            {"kind": "residual_norm", "reduction_factor": 1e-12}]
    inv = ora.jacobi_inverse_diagonal(rp, ci, v)
    xo, so, io = ora.bicgstab_solve(rp, ci, v, b_h, np.zeros_like(b_h), spec, inv_diag=inv)
    A = gko.csr_from_host(exec_, rp, ci, v)
    s = (gko.solver.Bicgstab.build("f64").with_criteria(*crits(gko, spec, "f64"))
         .with_preconditioner(gko.preconditioner.Jacobi.build("f64").with_max_block_size(1))
         .on(exec_).generate(A))
    x = gko.matrix.Dense("f64").create(exec_, n, 2)
    x.fill(0.0)
    s.apply(dense(gko, exec_, b_h, "f64"), x)
    assert abs(s.last_iterations - io.iterations) <= 2
    np.testing.assert_allclose(x.to_numpy(), np.linalg.solve(d, b_h), atol=1e-9)


def test_ir_mixed_precision(gko, exec_):
    """solver::Ir<double> around Cg<float> (the mixed-precision-ir example):
    the outer FP64 refinement reaches 1e-12 although the inner solve is FP32."""
    rp, ci, v = gen.diffusion27(14)
    n = 14 ** 3
    b_h = np.random.default_rng(4).uniform(-1, 1, (n, 1))
    outer = [{"kind": "iteration", "max_iters": 100},
             {"kind": "residual_norm", "reduction_factor": 1e-12}]
    inner = [{"kind": "iteration", "max_iters": 2000},
             {"kind": "residual_norm", "reduction_factor": 1e-3}]
    xo, so, io, inner_total = ora.ir_solve(rp, ci, v, b_h[:, 0], np.zeros(n), outer, inner)
    A = gko.csr_from_host(exec_, rp, ci, v)
    inner_f = gko.solver.Cg.build("f32").with_criteria(*crits(gko, inner, "f32"))
    s = (gko.solver.Ir.build("f64").with_criteria(*crits(gko, outer, "f64"))
         .with_solver(inner_f).on(exec_).generate(A))
    x = gko.matrix.Dense("f64").create(exec_, n, 1)
    x.fill(0.0)
    s.apply(dense(gko, exec_, b_h, "f64"), x)
    d = gen.csr_to_dense(rp, ci, v)
    xh = x.to_numpy()[:, 0]
    assert np.linalg.norm(b_h[:, 0] - d @ xh) <= 1.05e-12 * np.linalg.norm(b_h)
    assert abs(s.last_iterations - io.iterations) <= 1
    np.testing.assert_allclose(xh, np.linalg.solve(d, b_h[:, 0]), rtol=0, atol=1e-9)
