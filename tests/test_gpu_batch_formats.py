"""batch::matrix::Ell and Dense on the device: the batched product against
the reference's fixtures and the oracle (reference/test/matrix/
batch_{ell,dense}_kernels.cpp), and the batched BiCGSTAB and CG over both
formats against the oracle and the Csr batch (reference/test/solver/
batch_{bicgstab,cg}_kernels.cpp run the same solvers over Csr, Ell and
Dense)."""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle as ora
from ginkgo_amd import generators as gen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
KAT = GOLD["batch_formats"]
SFX = {np.float64: "f64", np.float32: "f32"}


@pytest.fixture(scope="module")
def gko():
    import ginkgo_amd.gko as g
    return g


@pytest.fixture(scope="module")
def ex(gko):
    return gko.Mi355xExecutor.create(0, torch.cuda.current_stream().cuda_stream)


def random_items(num, rows, cols, per_row, seed, spd=False, dtype=np.float64):
    """Items with one shared pattern (the diagonal plus per_row random
    columns per row), values per item; SPD items are symmetric and strictly
    diagonally dominant, the others diagonally dominant."""
    rng = np.random.default_rng(seed)
    mask = np.zeros((rows, cols), bool)
    for i in range(rows):
        mask[i, rng.choice(cols, size=min(per_row, cols), replace=False)] = True
        if i < cols:
            mask[i, i] = True
    if spd:
        mask = mask | mask.T
    vals = rng.uniform(-1.0, 1.0, size=(num, rows, cols)) * mask
    if spd:
        vals = 0.5 * (vals + vals.transpose(0, 2, 1))
    if rows == cols:
        idx = np.arange(rows)
        vals[:, idx, idx] = 0.0
        vals[:, idx, idx] = np.abs(vals).sum(axis=2) + 1.0 + rng.uniform(0, 1, size=(num, rows))
    return vals.astype(dtype), mask


def csr_pattern(mask):
    rows = mask.shape[0]
    rp = np.zeros(rows + 1, np.int32)
    ci = []
    for i in range(rows):
        cols = np.nonzero(mask[i])[0]
        ci.extend(cols)
        rp[i + 1] = len(ci)
    return rp, np.array(ci, np.int32)


def csr_values(dense, rp, ci):
    rows = len(rp) - 1
    rows_of = np.repeat(np.arange(rows), np.diff(rp))
    return dense[:, rows_of, ci].reshape(-1)


def bound(dense, b, eps):
    """4 k eps (|A| |b|)_i for the rows of every item."""
    k = np.count_nonzero(dense, axis=2).max()
    return 4 * max(k, 1) * eps * np.einsum("kij,kj->ki", np.abs(dense), np.abs(b)) + 1e-300


def make(gko, ex, fmt, dense, mask):
    dt = SFX[dense.dtype.type]
    num, rows, cols = dense.shape
    if fmt == "ell":
        return gko.batch.matrix.Ell(dt).from_dense(ex, dense)
    if fmt == "dense":
        return gko.batch.matrix.Dense(dt).from_numpy(ex, dense)
    rp, ci = csr_pattern(mask)
    return gko.batch.matrix.Csr(dt).from_numpy(ex, num, rows, rp, ci, csr_values(dense, rp, ci),
                                               cols=cols)


def mv(gko, ex, a):
    a = np.asarray(a)
    return gko.batch.MultiVector(SFX[a.dtype.type]).from_numpy(ex, a.shape[0], a.shape[1],
                                                              a.reshape(-1))


@pytest.mark.parametrize("fmt", ["ell", "dense"])
@pytest.mark.parametrize("advanced", [False, True])
def test_apply_reference_fixture(gko, ex, fmt, advanced):
    """The reference's 2 x (2 x 3) fixture, one column of b at a time:
    every value is a small binary fraction, so the product is exact."""
    a = KAT["apply"]
    mtx = np.array(a["mtx"])
    A = make(gko, ex, fmt, mtx, None)
    for col in range(3):
        b = np.array(a["b"])[:, :, col]
        x0 = np.array(a["x"])[:, :, col]
        expect = np.einsum("kij,kj->ki", mtx, b)
        x = mv(gko, ex, x0)
        if advanced:
            al, be = np.array(a["alpha"]), np.array(a["beta"])
            A.apply(mv(gko, ex, al[:, None]), mv(gko, ex, b), mv(gko, ex, be[:, None]), x)
            expect = al[:, None] * expect + be[:, None] * x0
        else:
            A.apply(mv(gko, ex, b), x)
        np.testing.assert_array_equal(x.to_numpy().reshape(2, 2), expect)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("fmt", ["csr", "ell", "dense"])
@pytest.mark.parametrize("shape", [(400, 40, 40, 5), (64, 200, 200, 9), (100, 30, 50, 4)],
                         ids=["40x40", "200x200", "30x50"])
def test_apply_random_against_oracle(gko, ex, dtype, fmt, shape):
    num, rows, cols, per_row = shape
    dense, mask = random_items(num, rows, cols, per_row, seed=rows + cols, dtype=dtype)
    rng = np.random.default_rng(1)
    b = rng.uniform(-1, 1, size=(num, cols)).astype(dtype)
    x0 = rng.uniform(-1, 1, size=(num, rows)).astype(dtype)
    al = rng.uniform(-2, 2, size=num).astype(dtype)
    be = rng.uniform(-2, 2, size=num).astype(dtype)
    A = make(gko, ex, fmt, dense, mask)
    eps = np.finfo(dtype).eps
    # simple
    x = mv(gko, ex, x0)
    A.apply(mv(gko, ex, b), x)
    if fmt == "ell":
        E = gko.batch.matrix.Ell(SFX[dtype]).from_dense(gko.HostExecutor.create(), dense)
        ref = ora.batch_ell_spmv(num, rows, cols, E.num_stored_elements_per_row, E.col_idxs(),
                                 E.values(), b.reshape(-1))
    elif fmt == "dense":
        ref = ora.batch_dense_spmv(num, rows, cols, dense.reshape(-1), b.reshape(-1))
    else:
        rp, ci = csr_pattern(mask)
        ref = ora.batch_csr_spmv(num, rows, cols, rp, ci, csr_values(dense, rp, ci),
                                 b.reshape(-1))
    got = x.to_numpy().reshape(num, rows)
    assert np.all(np.abs(got - ref.reshape(num, rows)) <= bound(dense, b, eps))
    # advanced: x = alpha A b + beta x
    x = mv(gko, ex, x0)
    A.apply(mv(gko, ex, al[:, None]), mv(gko, ex, b), mv(gko, ex, be[:, None]), x)
    expect = al[:, None] * ref.reshape(num, rows) + be[:, None] * x0
    tol = np.abs(al)[:, None] * bound(dense, b, eps) + 4 * eps * (
        np.abs(al[:, None] * ref.reshape(num, rows)) + np.abs(be[:, None] * x0))
    assert np.all(np.abs(x.to_numpy().reshape(num, rows) - expect) <= tol)


def test_apply_uneven_block_diagonal(gko, ex):
    """BlockDiagCsr items of different sizes through the same product."""
    row_off, rp, ci, v, _ = gen.uneven_batch(num_systems=300, n_min=16, n_max=120, seed=3)
    A = gko.batch.matrix.BlockDiagCsr("f64").from_numpy(ex, row_off, rp, ci, v)
    rng = np.random.default_rng(2)
    b = rng.uniform(-1, 1, row_off[-1])
    x = gko.batch.MultiVector("f64").uneven(ex, row_off, np.zeros(row_off[-1]))
    A.apply(gko.batch.MultiVector("f64").uneven(ex, row_off, b), x)
    got = x.to_numpy()
    for s in range(0, 300, 7):
        r0, r1 = row_off[s], row_off[s + 1]
        d = gen.csr_to_dense(rp[r0:r1 + 1] - rp[r0], ci[rp[r0]:rp[r1]], v[rp[r0]:rp[r1]],
                             r1 - r0)
        np.testing.assert_allclose(got[r0:r1], d @ b[r0:r1], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("method", ["bicgstab", "cg"])
@pytest.mark.parametrize("fmt", ["ell", "dense"])
@pytest.mark.parametrize("jacobi", [False, True])
def test_solvers_match_csr_batch_and_oracle(gko, ex, method, fmt, jacobi):
    """The same SPD batch as Csr, Ell and Dense (Dense adds the zeros of
    each row, 0 * x_j = 0): the three differ only in the rounding of the
    row sums and dot products (the Csr product may split a row over lanes),
    so iteration counts agree to one.  The oracle runs the reference's
    per-system loop on the Csr form."""
    num, n = 300, 48
    dense, mask = random_items(num, n, n, 4, seed=11, spd=True)
    rng = np.random.default_rng(5)
    b = rng.uniform(-1, 1, size=(num, n))
    solver_t = gko.batch.solver.Cg if method == "cg" else gko.batch.solver.Bicgstab

    def solve(A):
        p = (solver_t.build("f64").with_max_iterations(200).with_tolerance(1e-10)
             .with_tolerance_type(gko.batch.stop.tolerance_type.relative))
        if jacobi:
            p = p.with_preconditioner(gko.batch.preconditioner.Jacobi.build("f64"))
        s = p.on(ex).generate(A)
        log = gko.batch.log.BatchConvergence("f64").create()
        s.add_logger(log)
        x = mv(gko, ex, np.zeros((num, n)))
        s.apply(mv(gko, ex, b), x)
        return x.to_numpy().reshape(num, n), np.array(log.num_iterations)

    xf, itf = solve(make(gko, ex, fmt, dense, mask))
    xc, itc = solve(make(gko, ex, "csr", dense, mask))
    assert np.max(np.abs(itf - itc)) <= 1
    np.testing.assert_allclose(xf, xc, rtol=1e-6, atol=1e-8)
    rp, ci = csr_pattern(mask)
    row_off = np.arange(num + 1, dtype=np.int64) * n
    glob_rp = np.concatenate([rp[:-1] + k * rp[-1] for k in range(num)] + [[num * rp[-1]]])
    oracle_solve = ora.batch_cg_solve if method == "cg" else ora.batch_bicgstab_solve
    xo, io, ro = oracle_solve(row_off, glob_rp, np.tile(ci, num), csr_values(dense, rp, ci),
                              b.reshape(-1), np.zeros(num * n), use_jacobi=jacobi,
                              relative=True, tol=1e-10, max_iters=200)
    assert np.max(np.abs(itf - io)) <= 1
    res = np.linalg.norm(b - np.einsum("kij,kj->ki", dense, xf), axis=1) / np.linalg.norm(b, axis=1)
    # CG stops on the implicit norm sqrt|r.z|; BiCGSTAB on ||r||
    assert np.all(res <= (1e-8 if method == "cg" and jacobi else 1e-9))


@pytest.mark.parametrize("method", ["bicgstab", "cg"])
def test_dense_global_memory_path(gko, ex, method):
    """Dense items too large for LDS (160 x 160 FP64 = 200 KiB) run from
    global memory."""
    num, n = 40, 160
    dense, mask = random_items(num, n, n, 6, seed=21, spd=True)
    b = np.random.default_rng(6).uniform(-1, 1, size=(num, n))
    solver_t = gko.batch.solver.Cg if method == "cg" else gko.batch.solver.Bicgstab
    s = (solver_t.build("f64").with_max_iterations(300).with_tolerance(1e-10)
         .with_tolerance_type(gko.batch.stop.tolerance_type.relative)
         .with_preconditioner(gko.batch.preconditioner.Jacobi.build("f64"))
         .on(ex).generate(make(gko, ex, "dense", dense, mask)))
    # -1: resolved at the launch, where 160 * 160 * 8 bytes exceed the LDS
    assert s.matrix_in_lds == [-1] and n * n * 8 > 160 * 1024
    x = mv(gko, ex, np.zeros((num, n)))
    s.apply(mv(gko, ex, b), x)
    xh = x.to_numpy().reshape(num, n)
    res = np.linalg.norm(b - np.einsum("kij,kj->ki", dense, xh), axis=1) / np.linalg.norm(b, axis=1)
    assert np.all(res <= 1e-8)


def test_tuned_ell_batch(gko, ex):
    """The tuner times LDS and global residency for an Ell batch and keeps
    the faster; the solution does not depend on the choice."""
    num, n = 500, 64
    dense, mask = random_items(num, n, n, 5, seed=31, spd=True)
    b = np.random.default_rng(7).uniform(-1, 1, size=(num, n))
    A = make(gko, ex, "ell", dense, mask)
    s = (gko.batch.solver.Bicgstab.build("f64").with_max_iterations(200).with_tolerance(1e-10)
         .with_tolerance_type(gko.batch.stop.tolerance_type.relative).on(ex).generate(A))
    bb = mv(gko, ex, b)
    x = mv(gko, ex, np.zeros((num, n)))
    s.tune(bb, x)
    assert s.block_sizes[0] in (64, 128, 256, 512) and s.matrix_in_lds[0] in (0, 1)
    s.apply(bb, x)
    xh = x.to_numpy().reshape(num, n)
    res = np.linalg.norm(b - np.einsum("kij,kj->ki", dense, xh), axis=1) / np.linalg.norm(b, axis=1)
    assert np.all(res <= 1e-9)


def test_ell_jacobi_missing_diagonal_is_one(gko, ex):
    """An Ell row without a diagonal entry gets the Jacobi value 1, as in the
    reference's ScalarJacobi::generate for Ell."""
    # 2 x 2 items [[2, 1], [1, 0]] would have a missing (not stored) diagonal
    # in row 1; with b = [3, 1] the solution is [1, 1]
    ci = np.array([0, 0, 1, -1], np.int32)
    vals = np.array([2.0, 1.0, 1.0, 0.0] * 3)
    A = gko.batch.matrix.Ell("f64").from_numpy(ex, 3, 2, 2, 2, ci, vals)
    s = (gko.batch.solver.Bicgstab.build("f64").with_max_iterations(50).with_tolerance(1e-12)
         .with_preconditioner(gko.batch.preconditioner.Jacobi.build("f64")).on(ex).generate(A))
    x = mv(gko, ex, np.zeros((3, 2)))
    s.apply(mv(gko, ex, np.tile([3.0, 1.0], (3, 1))), x)
    np.testing.assert_allclose(x.to_numpy(), np.ones(6), rtol=1e-10)


def test_descriptor_errors(gko, ex):
    """The C ABI rejects an uneven Ell without item offsets, a rectangular
    solve and a half-given alpha/beta pair."""
    from ginkgo_amd import _capi as C, device as D
    x = torch.zeros(8, dtype=torch.float64, device="cuda")
    ctx = D.ctx(x.device)
    off = torch.tensor([0, 4, 8], dtype=torch.int64, device="cuda")
    ci = torch.zeros(8, dtype=torch.int32, device="cuda")
    d = C.BatchCsr(num_systems=2, max_rows=4, max_nnz=8, row_offsets=D.ptr(off),
                   col_idxs=D.ptr(ci), values=D.ptr(x), format=C.BATCH_ELL, ell_width=2)
    with pytest.raises(C.InvalidArgument):
        C.call("gkomi_batch_spmv_f64", ctypes.byref(ctx), ctypes.byref(d), None, D.ptr(x), None,
               D.ptr(x))
    d = C.BatchCsr(num_systems=2, max_rows=2, max_nnz=4, values=D.ptr(x), uniform_rows=2,
                   uniform_nnz=4, format=C.BATCH_DENSE, uniform_cols=2)
    with pytest.raises(C.InvalidArgument):
        C.call("gkomi_batch_spmv_f64", ctypes.byref(ctx), ctypes.byref(d), D.ptr(x), D.ptr(x),
               None, D.ptr(x))
    with pytest.raises(gko.DimensionMismatch):
        rect = gko.batch.matrix.Dense("f64").from_numpy(ex, np.ones((2, 2, 3)))
        gko.batch.solver.Cg.build("f64").on(ex).generate(rect)
    torch.cuda.synchronize()


GOLD_BATCH = {"bicgstab": GOLD["batch_bicgstab"], "cg": GOLD["batch_cg"]}


@pytest.mark.parametrize("fmt", ["ell", "dense"])
@pytest.mark.parametrize("method", ["bicgstab", "cg"])
def test_reference_stencil_kats_in_ell_and_dense(gko, ex, fmt, method):
    """The reference's batch solver KATs (SolvesStencilSystem,
    StencilSystemLoggerLogsIterations of reference/test/solver/
    batch_{bicgstab,cg}_kernels.cpp) with the 3-point stencil stored as
    batch::matrix::Ell and Dense (the reference's solver tests run over the
    same matrix types)."""
    k = GOLD_BATCH[method][0]
    n, nb = k["num_rows"], k["num_batch_items"]
    d = np.diag(np.full(n, k["diag"])) + np.diag(np.full(n - 1, k["off"]), 1) + \
        np.diag(np.full(n - 1, k["off"]), -1)
    dense = np.tile(d, (nb, 1, 1))
    xs = np.full(n, k["exact_sol_fill"])
    b = np.tile(d @ xs, (nb, 1))
    A = make(gko, ex, fmt, dense, d != 0)
    solver_t = gko.batch.solver.Cg if method == "cg" else gko.batch.solver.Bicgstab
    s = (solver_t.build("f64").with_max_iterations(k["max_iterations"])
         .with_tolerance(k["tolerance"])
         .with_tolerance_type(gko.batch.stop.tolerance_type.relative).on(ex).generate(A))
    x = mv(gko, ex, np.zeros((nb, n)))
    s.apply(mv(gko, ex, b), x)
    xh = x.to_numpy().reshape(nb, n)
    relres_factor = float(k["check_relres_le"].split("*")[0])
    for i in range(nb):
        rel = np.linalg.norm(b[i] - d @ xh[i]) / np.linalg.norm(b[i])
        assert rel <= relres_factor * k["tolerance"]
        assert np.max(np.abs(xh[i] - xs)) <= 10 * k["tolerance"]
    # the iteration KAT: tolerance 0 and the limit 5
    kit = GOLD_BATCH[method][-1]
    s5 = (solver_t.build("f64").with_max_iterations(kit["max_iterations"]).with_tolerance(0.0)
          .with_tolerance_type(gko.batch.stop.tolerance_type.relative).on(ex).generate(A))
    log = gko.batch.log.BatchConvergence("f64").create()
    s5.add_logger(log)
    x = mv(gko, ex, np.zeros((nb, n)))
    s5.apply(mv(gko, ex, b), x)
    its = np.array(log.num_iterations)
    if "check_iters_eq" in kit:
        assert np.all(its == kit["check_iters_eq"])
    else:
        assert np.all(its <= kit["check_iters_le"])


# ------------------------------------------- uneven Ell / Dense batches ----

def uneven_dense_blocks(num, n_min, n_max, seed, spd):
    """Dense blocks of different sizes (the diagonal plus 4 random columns
    per row; SPD ones symmetric), diagonally dominant."""
    rng = np.random.default_rng(seed)
    blocks = []
    for k in range(num):
        n = int(rng.integers(n_min, n_max + 1))
        d, _ = random_items(1, n, n, 4, seed=seed * 1000 + k, spd=spd)
        blocks.append(d[0])
    return blocks


def block_diag(blocks):
    row_off, rp, ci, v = [0], [0], [], []
    for d in blocks:
        for i in range(d.shape[0]):
            cols = np.nonzero(d[i])[0]
            ci.extend(cols)
            v.extend(d[i, cols])
            rp.append(len(ci))
        row_off.append(row_off[-1] + d.shape[0])
    return (np.array(row_off, np.int64), np.array(rp, np.int32), np.array(ci, np.int32),
            np.array(v))


def test_apply_uneven_ell_and_dense(gko, ex):
    """BlockDiagEll and BlockDiagDense products: every item against its dense
    block; the Ell product sums a row in the stored order, as the
    BlockDiagCsr product does, so the two agree bit for bit."""
    blocks = uneven_dense_blocks(150, 1, 70, seed=4, spd=False)
    row_off, rp, ci, v = block_diag(blocks)
    b = np.random.default_rng(8).uniform(-1, 1, row_off[-1])
    out = {}
    for name, A in (("csr", gko.batch.matrix.BlockDiagCsr("f64").from_numpy(ex, row_off, rp, ci, v)),
                    ("ell", gko.batch.matrix.BlockDiagEll("f64").from_numpy(ex, row_off, rp, ci, v)),
                    ("dense", gko.batch.matrix.BlockDiagDense("f64").from_block_diag_csr(
                        ex, row_off, rp, ci, v))):
        x = gko.batch.MultiVector("f64").uneven(ex, row_off, np.zeros(row_off[-1]))
        A.apply(gko.batch.MultiVector("f64").uneven(ex, row_off, b), x)
        out[name] = x.to_numpy()
    for k, d in enumerate(blocks):
        r0, r1 = row_off[k], row_off[k + 1]
        ref = d @ b[r0:r1]
        tol = 4 * d.shape[0] * np.finfo(np.float64).eps * (np.abs(d) @ np.abs(b[r0:r1])) + 1e-300
        for name in out:
            assert np.all(np.abs(out[name][r0:r1] - ref) <= tol), (name, k)
    np.testing.assert_array_equal(out["ell"], out["csr"])


@pytest.mark.parametrize("method", ["bicgstab", "cg"])
def test_solvers_uneven_ell_and_dense(gko, ex, method):
    """Batched BiCGSTAB and CG over BlockDiagEll and BlockDiagDense items of
    16 to 64 rows: iteration counts within one of the oracle's (the
    reference's per-system loop on the Csr form of each item), every
    system's true relative residual at the tolerance, and the Ell and Dense
    solves agreeing (Dense adds each row's zeros, 0 * x_j = 0)."""
    blocks = uneven_dense_blocks(200, 16, 64, seed=7, spd=(method == "cg"))
    row_off, rp, ci, v = block_diag(blocks)
    b = np.random.default_rng(9).uniform(-1, 1, row_off[-1])
    solver_t = gko.batch.solver.Cg if method == "cg" else gko.batch.solver.Bicgstab
    got = {}
    for name, A in (("ell", gko.batch.matrix.BlockDiagEll("f64").from_numpy(ex, row_off, rp, ci, v)),
                    ("dense", gko.batch.matrix.BlockDiagDense("f64").from_block_diag_csr(
                        ex, row_off, rp, ci, v))):
        s = (solver_t.build("f64").with_max_iterations(300).with_tolerance(1e-10)
             .with_tolerance_type(gko.batch.stop.tolerance_type.relative)
             .with_preconditioner(gko.batch.preconditioner.Jacobi.build("f64"))
             .on(ex).generate(A))
        log = gko.batch.log.BatchConvergence("f64").create()
        s.add_logger(log)
        x = gko.batch.MultiVector("f64").uneven(ex, row_off, np.zeros(row_off[-1]))
        s.apply(gko.batch.MultiVector("f64").uneven(ex, row_off, b), x)
        got[name] = (x.to_numpy(), np.array(log.num_iterations))
    oracle_solve = ora.batch_cg_solve if method == "cg" else ora.batch_bicgstab_solve
    xo, io, ro = oracle_solve(row_off, rp, ci, v, b, np.zeros_like(b), use_jacobi=True,
                              relative=True, tol=1e-10, max_iters=300)
    for name, (xh, it) in got.items():
        assert np.max(np.abs(it - io)) <= 1, name
        for k, d in enumerate(blocks):
            r0, r1 = row_off[k], row_off[k + 1]
            res = np.linalg.norm(b[r0:r1] - d @ xh[r0:r1]) / np.linalg.norm(b[r0:r1])
            assert res <= (1e-8 if method == "cg" else 1e-9), (name, k, res)
    assert np.max(np.abs(got["ell"][1] - got["dense"][1])) <= 1
    np.testing.assert_allclose(got["ell"][0], got["dense"][0], rtol=1e-7, atol=1e-9)
