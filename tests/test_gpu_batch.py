"""Batched BiCGSTAB on the device against the reference's batch tests and
the oracle (a per-system restatement of reference/solver/batch_bicgstab_kernels.hpp)."""
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
def dev():
    from ginkgo_amd import _capi
    _capi.lib()
    return torch.device("cuda", 0)


def D():
    from ginkgo_amd import device
    return device


def stencil_batch(k):
    row_off, rp, ci, v = gen.stencil3pt_batch(k["num_batch_items"], k["num_rows"], k["diag"],
                                              k["off"])
    n = k["num_rows"]
    d = gen.csr_to_dense(rp[: n + 1], ci[: rp[n]], v[: rp[n]])
    xs = np.full(n, k["exact_sol_fill"])
    b = np.tile(d @ xs, k["num_batch_items"])
    return row_off, rp, ci, v, b, xs, n


@pytest.mark.parametrize("uniform", [False, True])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reference_stencil_kat(dev, dtype, uniform):
    k = KATS["batch_bicgstab"][0]
    row_off, rp, ci, v, b, xs, n = stencil_batch(k)
    if uniform:
        A = D().DeviceBatchCsr.uniform(k["num_batch_items"], rp[: n + 1], ci[: rp[n]],
                                       v.astype(dtype), device=dev)
    else:
        A = D().DeviceBatchCsr(row_off, rp, ci, v.astype(dtype), device=dev)
    bd = torch.as_tensor(b.astype(dtype)).to(dev)
    x = torch.zeros_like(bd)
    st = A.settings(tol=k["tolerance"], relative=True, max_iters=k["max_iterations"],
                    jacobi=False)
    iters, res = A.solve(bd, x, st)
    xh = x.cpu().numpy().astype(np.float64)
    for s in range(k["num_batch_items"]):
        bs = b[s * n:(s + 1) * n]
        d = gen.csr_to_dense(rp[: n + 1], ci[: rp[n]], v[: rp[n]])
        true_res = np.linalg.norm(bs - d @ xh[s * n:(s + 1) * n])
        assert true_res / np.linalg.norm(bs) <= 2 * k["tolerance"]
        assert float(res[s]) / np.linalg.norm(bs) <= k["tolerance"] * 1.0001
        err = np.linalg.norm(xh[s * n:(s + 1) * n] - xs) / np.linalg.norm(xs)
        assert err <= 10 * k["tolerance"]


def test_reference_iteration_cap_kat(dev):
    k = KATS["batch_bicgstab"][1]
    row_off, rp, ci, v, b, xs, n = stencil_batch(k)
    A = D().DeviceBatchCsr(row_off, rp, ci, v, device=dev)
    bd = torch.as_tensor(b).to(dev)
    x = torch.zeros_like(bd)
    iters, _ = A.solve(bd, x, A.settings(tol=0.0, relative=True, max_iters=5, jacobi=False))
    assert int(iters.max()) <= k["check_iters_le"]
    xo, io, ro = ora.batch_bicgstab_solve(row_off, rp, ci, v, b, np.zeros_like(b),
                                          use_jacobi=False, tol=0.0, max_iters=5)
    assert iters.cpu().tolist() == io.tolist()


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 1e-5)])
def test_uneven_batch_against_oracle(dev, dtype, tol):
    row_off, rp, ci, v, b = gen.uneven_batch(300, seed=12)
    v = v.astype(dtype)
    b = b.astype(dtype)
    A = D().DeviceBatchCsr(row_off, rp, ci, v, device=dev)
    bd = torch.as_tensor(b).to(dev)
    x = torch.zeros_like(bd)
    st = A.settings(tol=tol, relative=True, max_iters=500, jacobi=True)
    plan = A.plan()
    iters, res = A.solve(bd, x, st, plan)
    xo, io, ro = ora.batch_bicgstab_solve(row_off, rp, ci, v, b, np.zeros_like(b),
                                          use_jacobi=True, relative=True, tol=tol,
                                          max_iters=500)
    it = iters.cpu().numpy()
    assert np.all(np.abs(it - io) <= (1 if dtype == np.float64 else 3))
    xh = x.cpu().numpy().astype(np.float64)
    for s in range(0, 300, 7):
        r0, r1 = row_off[s], row_off[s + 1]
        sub = rp[r0:r1 + 1] - rp[r0]
        d = gen.csr_to_dense(sub, ci[rp[r0]:rp[r1]], v[rp[r0]:rp[r1]].astype(np.float64))
        bs = b[r0:r1].astype(np.float64)
        assert np.linalg.norm(bs - d @ xh[r0:r1]) <= 3 * tol * np.linalg.norm(bs)


def test_lds_and_global_residency_agree(dev):
    row_off, rp, ci, v, b = gen.uneven_batch(120, seed=3)
    A = D().DeviceBatchCsr(row_off, rp, ci, v, device=dev)
    bd = torch.as_tensor(b).to(dev)
    st = A.settings(tol=1e-10, max_iters=300)
    outs = []
    for mat in (1, 0):
        for bs in (64, 256, 512):
            plan = A.plan(block_size=bs)
            for i in range(plan.num_bins):
                plan.matrix_in_lds[i] = mat
                plan.vectors[i] = 1    # the LDS work-vector kernel
            x = torch.zeros_like(bd)
            iters, res = A.solve(bd, x, st, plan)
            outs.append((x.cpu().numpy(), iters.cpu().numpy()))
    # identical arithmetic order within a block size: bit-identical
    np.testing.assert_array_equal(outs[0][0], outs[3][0])
    np.testing.assert_array_equal(outs[2][0], outs[5][0])
    for xo, it in outs[1:]:
        np.testing.assert_allclose(xo, outs[0][0], rtol=0, atol=1e-8)


@pytest.mark.parametrize("method", ["bicgstab", "cg"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("uniform", [False, True])
def test_register_and_lds_work_vectors_agree(dev, method, dtype, uniform):
    """The register work-vector kernel against the LDS work-vector kernel
    with one lane per row: both sum every row in stored order by one thread
    and reduce in the same order.  CG agrees bit for bit; BiCGSTAB in the
    iteration counts, with solutions and residual norms within a few units
    in the last place (the compiler fuses the BiCGSTAB updates' products
    differently in the two kernels).  One and two rows per thread."""
    if uniform:
        k = KATS["batch_bicgstab"][0]
        _, rp, ci, v, _, _, n = stencil_batch(dict(k, num_batch_items=64, num_rows=300))
        A = D().DeviceBatchCsr.uniform(64, rp[: n + 1], ci[: rp[n]],
                                       np.tile(v[: rp[n]], 64).astype(dtype), device=dev)
        b = np.random.default_rng(3).uniform(-1, 1, 64 * n)
    else:
        maker = gen.uneven_spd_batch if method == "cg" else gen.uneven_batch
        row_off, rp, ci, v, b = maker(400, n_min=16, n_max=512, seed=17)
        A = D().DeviceBatchCsr(row_off, rp, ci, v.astype(dtype), device=dev)
    bd = torch.as_tensor(b.astype(dtype)).to(dev)
    tol = 1e-10 if dtype == np.float64 else 1e-5
    st = A.settings(tol=tol, relative=True, max_iters=500, jacobi=True)
    # (vectors, matrix_in_lds): LDS work vectors; register work vectors with
    # the matrix in LDS; and (uneven) with the matrix rows in registers
    variants = [(1, 1), (2, 1)] + ([] if uniform else [(2, 0)])
    for bs in (256, 512):
        outs = []
        for vec, mat in variants:
            plan = A.plan(block_size=bs)
            for i in range(plan.num_bins):
                plan.vectors[i] = vec
                plan.matrix_in_lds[i] = mat
                plan.subwave[i] = 1
            x = torch.zeros_like(bd)
            it, res = A.solve(bd, x, st, plan, method=method)
            outs.append((x.cpu().numpy(), it.cpu().numpy(), res.cpu().numpy()))
        for o in outs[1:]:
            np.testing.assert_array_equal(outs[0][1], o[1])
            if method == "cg":
                np.testing.assert_array_equal(outs[0][0], o[0])
                np.testing.assert_array_equal(outs[0][2], o[2])
            else:
                eps = np.finfo(dtype).eps
                scale = float(np.abs(outs[0][0]).max())
                np.testing.assert_allclose(o[0], outs[0][0], rtol=16 * eps,
                                           atol=4 * eps * scale)
                # the final residual norms are small differences of O(1) terms
                np.testing.assert_allclose(o[2], outs[0][2], rtol=1e-2 if
                                           dtype == np.float32 else 1e-6, atol=0)
            assert int(o[1].max()) < 500


def test_register_vectors_refused_where_they_do_not_apply(dev):
    from ginkgo_amd import _capi as C
    row_off, rp, ci, v, b = gen.uneven_batch(20, n_min=300, n_max=512, seed=2)
    A = D().DeviceBatchCsr(row_off, rp, ci, v, device=dev)
    bd = torch.as_tensor(b).to(dev)
    plan = A.plan(block_size=64)           # 512 rows need more than two rows per thread
    for i in range(plan.num_bins):
        plan.vectors[i] = 2
    with pytest.raises(C.NotSupported):
        A.solve(bd, torch.zeros_like(bd), A.settings(), plan)
    for i in range(plan.num_bins):
        plan.vectors[i] = 0                # automatic: the LDS-vector kernel runs
    it, _ = A.solve(bd, torch.zeros_like(bd), A.settings(tol=1e-10, max_iters=300), plan)
    assert int(it.max()) < 300


def test_tuned_plan_and_padded_uniform_agree(dev):
    row_off, rp, ci, v, b = gen.uneven_batch(200, seed=8, n_max=96)
    A = D().DeviceBatchCsr(row_off, rp, ci, v, device=dev)
    bd = torch.as_tensor(b).to(dev)
    x0 = torch.zeros_like(bd)
    st = A.settings(tol=1e-10, max_iters=300)
    plan = A.tune(bd, x0, st, A.plan())
    assert all(plan.block_size[i] in (64, 128, 256, 512) for i in range(plan.num_bins))
    x = torch.zeros_like(bd)
    iters, res = A.solve(bd, x, st, plan)
    # the reference-compatible padded uniform batch solves the same systems
    prp, pci, pv, pb = gen.pad_to_uniform(row_off, rp, ci, v, b)
    npad = len(prp) - 1
    U = D().DeviceBatchCsr.uniform(200, prp, pci, pv, device=dev)
    ub = torch.as_tensor(pb).to(dev)
    ux = torch.zeros_like(ub)
    U.solve(ub, ux, U.settings(tol=1e-10, max_iters=300))
    uxh = ux.cpu().numpy().reshape(200, npad)
    xh = x.cpu().numpy()
    sizes = np.diff(row_off)
    for s in range(200):
        np.testing.assert_allclose(uxh[s, :sizes[s]], xh[row_off[s]:row_off[s + 1]], atol=1e-8)
        assert np.all(uxh[s, sizes[s]:] == 0)


# ------------------------------------------------- the gko:: batch API ----
# batch::solver::Bicgstab<V>::build().with_*(...).on(exec)->generate(A)->apply(b, x)
# through the C++ library, as reference/test/solver/batch_bicgstab_kernels.cpp
# drives it.

@pytest.fixture(scope="module")
def gko_exec(dev):
    import ginkgo_amd.gko as gko
    return gko, gko.Mi355xExecutor.create(0, torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("dt", ["f64", "f32"])
def test_api_reference_stencil_kat(gko_exec, dt):
    gko, ex = gko_exec
    k = KATS["batch_bicgstab"][0]
    row_off, rp, ci, v, b, xs, n = stencil_batch(k)
    npdt = np.float64 if dt == "f64" else np.float32
    nb = k["num_batch_items"]
    A = gko.batch.matrix.Csr(dt).from_numpy(ex, nb, n, rp[: n + 1], ci[: rp[n]],
                                            v.astype(npdt))
    solver = (gko.batch.solver.Bicgstab.build(dt)
              .with_max_iterations(k["max_iterations"])
              .with_tolerance(k["tolerance"])
              .with_tolerance_type(gko.batch.stop.tolerance_type.relative)
              .on(ex).generate(A))
    logger = gko.batch.log.BatchConvergence(dt).create()
    solver.add_logger(logger)
    bb = gko.batch.MultiVector(dt).from_numpy(ex, nb, n, b.astype(npdt))
    x = gko.batch.MultiVector(dt).create(ex, nb, n)
    x.fill(0.0)
    solver.apply(bb, x)
    xh = x.to_numpy().astype(np.float64)
    d = gen.csr_to_dense(rp[: n + 1], ci[: rp[n]], v[: rp[n]])
    for s in range(nb):
        bs = b[s * n:(s + 1) * n]
        assert np.linalg.norm(bs - d @ xh[s * n:(s + 1) * n]) / np.linalg.norm(bs) <= \
            2 * k["tolerance"]
        assert np.linalg.norm(xh[s * n:(s + 1) * n] - xs) / np.linalg.norm(xs) <= \
            10 * k["tolerance"]
    assert len(logger.num_iterations) == nb and np.all(logger.num_iterations > 0)
    assert np.all(logger.residual_norm / np.linalg.norm(b[:n]) <= k["tolerance"] * 1.0001)


def test_api_uneven_block_diagonal_against_oracle(gko_exec):
    gko, ex = gko_exec
    row_off, rp, ci, v, b = gen.uneven_batch(150, n_min=16, n_max=300, seed=9)
    A = gko.batch.matrix.BlockDiagCsr("f64").from_numpy(ex, row_off, rp, ci, v)
    solver = (gko.batch.solver.Bicgstab.build("f64").with_max_iterations(500)
              .with_tolerance(1e-10)
              .with_tolerance_type(gko.batch.stop.tolerance_type.relative)
              .with_preconditioner(gko.batch.preconditioner.Jacobi.build("f64"))
              .on(ex).generate(A))
    logger = gko.batch.log.BatchConvergence("f64").create()
    solver.add_logger(logger)
    bb = gko.batch.MultiVector("f64").uneven(ex, row_off, b)
    x = gko.batch.MultiVector("f64").uneven(ex, row_off, np.zeros_like(b))
    solver.tune(bb, x)
    assert all(bs in (64, 128, 256, 512) for bs in solver.block_sizes)
    solver.apply(bb, x)
    xo, io, ro = ora.batch_bicgstab_solve(row_off, rp, ci, v, b, np.zeros_like(b),
                                          use_jacobi=True, tol=1e-10, max_iters=500)
    xh = x.to_numpy()
    for s in range(150):
        r0, r1 = row_off[s], row_off[s + 1]
        bs = b[r0:r1]
        d = gen.csr_to_dense(rp[r0:r1 + 1] - rp[r0], ci[rp[r0]:rp[r1]], v[rp[r0]:rp[r1]])
        assert np.linalg.norm(bs - d @ xh[r0:r1]) / np.linalg.norm(bs) <= 2e-10
    assert np.max(np.abs(logger.num_iterations - io)) <= 2


@pytest.mark.parametrize("method", ["Bicgstab", "Cg"])
def test_api_graph_capture_replays(gko_exec, method):
    """batch::solver::{Bicgstab, Cg}::capture: the whole solve (x = 0 by a
    fill kernel, then every bin's launch) as one HIP graph.  Three replays,
    with eager work on the same buffers and host synchronisations between
    them, each equal bit for bit to the eager solve."""
    gko, ex = gko_exec
    row_off, rp, ci, v, b = gen.uneven_batch(150, n_min=16, n_max=300, seed=9)
    if method == "Cg":
        row_off, rp, ci, v, b = gen.uneven_spd_batch(150, n_min=16, n_max=300, seed=9)
    A = gko.batch.matrix.BlockDiagCsr("f64").from_numpy(ex, row_off, rp, ci, v)
    solver = (getattr(gko.batch.solver, method).build("f64").with_max_iterations(500)
              .with_tolerance(1e-10)
              .with_tolerance_type(gko.batch.stop.tolerance_type.relative)
              .with_preconditioner(gko.batch.preconditioner.Jacobi.build("f64"))
              .on(ex).generate(A))
    logger = gko.batch.log.BatchConvergence("f64").create()
    solver.add_logger(logger)
    bb = gko.batch.MultiVector("f64").uneven(ex, row_off, b)
    x = gko.batch.MultiVector("f64").uneven(ex, row_off, np.zeros_like(b))
    solver.apply(bb, x)
    eager = x.to_numpy()
    eager_iters = np.array(logger.num_iterations)
    graph = solver.capture(bb, x)
    for rep in range(3):
        x.fill(7.0)                       # eager work on the graph's buffers
        solver.apply(bb, x)               # and an eager solve in between
        ex.synchronize()
        x.fill(-1.0)
        graph.replay()
        ex.synchronize()
        np.testing.assert_array_equal(x.to_numpy(), eager, err_msg=f"replay {rep}")
        solver.record_logs()
        np.testing.assert_array_equal(np.array(logger.num_iterations), eager_iters)
    # without the zero fill a replay starts from x as it is: from the
    # solution itself the iterations stop at once
    g2 = solver.capture(bb, x, zero_initial_guess=False)
    g2.replay()
    ex.synchronize()
    solver.record_logs()
    assert np.max(np.array(logger.num_iterations)) <= 2
    del graph, g2


def test_api_errors(gko_exec):
    gko, ex = gko_exec
    with pytest.raises(gko.NotSupported):
        gko.batch.solver.Bicgstab.build("f64").with_preconditioner(
            gko.batch.preconditioner.Jacobi.build("f64").with_max_block_size(4))
    rp = np.array([0, 1, 3], dtype=np.int32)
    with pytest.raises(gko.DimensionMismatch):   # column 2 is outside a 2-row item
        gko.batch.matrix.BlockDiagCsr("f64").from_numpy(
            ex, np.array([0, 2]), rp, np.array([0, 0, 2], dtype=np.int32), np.ones(3))
    A = gko.batch.matrix.Csr("f64").from_numpy(ex, 3, 2, rp, np.array([0, 0, 1], np.int32),
                                               np.ones(9))
    s = gko.batch.solver.Bicgstab.build("f64").on(ex).generate(A)
    with pytest.raises(gko.DimensionMismatch):
        s.apply(gko.batch.MultiVector("f64").create(ex, 2, 2),
                gko.batch.MultiVector("f64").create(ex, 2, 2))


def test_graph_captured_batch_solve_replays(dev):
    """The whole uneven batched solve captured in one HIP graph: every replay
    (with eager work and synchronisation between replays) reproduces the
    eager solve bit for bit -- each system is solved by one workgroup in a
    fixed order, and the queue counters are reset by the kernels."""
    row_off, rp, ci, v, b = gen.uneven_batch(300, n_min=16, n_max=400, seed=21)
    A = D().DeviceBatchCsr(row_off, rp, ci, v, device=dev)
    bd = torch.as_tensor(b).to(dev)
    st = A.settings(tol=1e-10, relative=True, max_iters=500, jacobi=True)
    plan = A.plan()
    x_eager = torch.zeros_like(bd)
    it_eager, res_eager = A.solve(bd, x_eager, st, plan)
    x = torch.full_like(bd, 7.0)
    graph, it_g, res_g = A.capture(bd, x, st, plan)
    for rep in range(3):
        # unrelated eager work on the same buffers' neighbours between replays
        scratch = torch.randn(1 << 16, dtype=torch.float64, device=dev)
        scratch.mul_(2.0).sum()
        x.fill_(-3.0)
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(x, x_eager), rep
        assert torch.equal(it_g, it_eager), rep
        assert torch.equal(res_g, res_eager), rep
