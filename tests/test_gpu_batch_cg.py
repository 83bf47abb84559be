"""Batched CG on the device against the reference's batch CG tests
(reference/test/solver/batch_cg_kernels.cpp) and the oracle's restatement of
reference/solver/batch_cg_kernels.hpp:111-221."""
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
    return row_off, rp, ci, v, b, xs, n, d


@pytest.mark.parametrize("uniform", [False, True])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reference_stencil_kats(dev, dtype, uniform):
    """SolvesStencilSystem and StencilSystemLoggerLogsResidual
    (batch_cg_kernels.cpp:81-120)."""
    k = KATS["batch_cg"][0]
    row_off, rp, ci, v, b, xs, n, d = stencil_batch(k)
    nb = k["num_batch_items"]
    if uniform:
        A = D().DeviceBatchCsr.uniform(nb, rp[: n + 1], ci[: rp[n]], v.astype(dtype), device=dev)
    else:
        A = D().DeviceBatchCsr(row_off, rp, ci, v.astype(dtype), device=dev)
    bd = torch.as_tensor(b.astype(dtype)).to(dev)
    x = torch.zeros_like(bd)
    st = A.settings(tol=k["tolerance"], relative=True, max_iters=k["max_iterations"],
                    jacobi=False)
    iters, res = A.solve(bd, x, st, method="cg")
    xh = x.cpu().numpy().astype(np.float64)
    rh = res.cpu().numpy().astype(np.float64)
    for s in range(nb):
        bs = b[s * n:(s + 1) * n]
        true_res = np.linalg.norm(bs - d @ xh[s * n:(s + 1) * n])
        assert true_res / np.linalg.norm(bs) <= 6 * k["tolerance"]
        assert np.max(np.abs(xh[s * n:(s + 1) * n] - xs)) <= 10 * k["tolerance"]
        assert rh[s] / np.linalg.norm(bs) <= k["tolerance"] * 1.0001
        assert abs(rh[s] - true_res) <= 10 * k["tolerance"]


def test_reference_iteration_kat(dev):
    """StencilSystemLoggerLogsIterations (batch_cg_kernels.cpp:123-139):
    exactly 5 iterations with tolerance 0 and the limit 5."""
    k = KATS["batch_cg"][2]
    row_off, rp, ci, v, b, xs, n, d = stencil_batch(k)
    A = D().DeviceBatchCsr(row_off, rp, ci, v, device=dev)
    bd = torch.as_tensor(b).to(dev)
    x = torch.zeros_like(bd)
    iters, res = A.solve(bd, x, A.settings(tol=0.0, relative=True, max_iters=5, jacobi=False),
                         method="cg")
    assert iters.cpu().tolist() == [k["check_iters_eq"]] * k["num_batch_items"]
    xo, io, ro = ora.batch_cg_solve(row_off, rp, ci, v, b, np.zeros_like(b), use_jacobi=False,
                                    tol=0.0, max_iters=5)
    np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(res.cpu().numpy(), ro, rtol=1e-10)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 1e-5)])
def test_uneven_spd_batch_against_oracle(dev, dtype, tol):
    row_off, rp, ci, v, b = gen.uneven_spd_batch(300, seed=12)
    v = v.astype(dtype)
    b = b.astype(dtype)
    A = D().DeviceBatchCsr(row_off, rp, ci, v, device=dev)
    bd = torch.as_tensor(b).to(dev)
    x = torch.zeros_like(bd)
    st = A.settings(tol=tol, relative=True, max_iters=500, jacobi=True)
    iters, res = A.solve(bd, x, st, A.plan(), method="cg")
    xo, io, ro = ora.batch_cg_solve(row_off, rp, ci, v, b, np.zeros_like(b), use_jacobi=True,
                                    relative=True, tol=tol, max_iters=500)
    it = iters.cpu().numpy()
    assert np.all(np.abs(it - io) <= (1 if dtype == np.float64 else 3))
    xh = x.cpu().numpy().astype(np.float64)
    for s in range(0, 300, 7):
        r0, r1 = row_off[s], row_off[s + 1]
        d = gen.csr_to_dense(rp[r0:r1 + 1] - rp[r0], ci[rp[r0]:rp[r1]],
                             v[rp[r0]:rp[r1]].astype(np.float64))
        bs = b[r0:r1].astype(np.float64)
        exact = np.linalg.solve(d, bs)
        # the implicit (Jacobi-weighted) norm met tol * ||b||; the true
        # residual and the error follow within the conditioning of M^(1/2)
        assert np.linalg.norm(bs - d @ xh[r0:r1]) <= 10 * tol * np.linalg.norm(bs)
        assert np.linalg.norm(xh[r0:r1] - exact) <= 10 * tol * np.linalg.norm(exact)


def test_lds_and_global_residency_agree(dev):
    row_off, rp, ci, v, b = gen.uneven_spd_batch(120, seed=3)
    A = D().DeviceBatchCsr(row_off, rp, ci, v, device=dev)
    bd = torch.as_tensor(b).to(dev)
    st = A.settings(tol=1e-10, max_iters=300)
    outs = []
    for mat in (1, 0):
        for bs in (64, 512):
            plan = A.plan(block_size=bs)
            for i in range(plan.num_bins):
                plan.matrix_in_lds[i] = mat
            x = torch.zeros_like(bd)
            A.solve(bd, x, st, plan, method="cg")
            outs.append(x.cpu().numpy())
    np.testing.assert_array_equal(outs[0], outs[2])
    np.testing.assert_array_equal(outs[1], outs[3])
    np.testing.assert_allclose(outs[1], outs[0], rtol=0, atol=1e-8)


def test_api_batch_cg(dev):
    import ginkgo_amd.gko as gko
    ex = gko.Mi355xExecutor.create(0, torch.cuda.current_stream().cuda_stream)
    row_off, rp, ci, v, b = gen.uneven_spd_batch(150, n_min=16, n_max=300, seed=9)
    A = gko.batch.matrix.BlockDiagCsr("f64").from_numpy(ex, row_off, rp, ci, v)
    solver = (gko.batch.solver.Cg.build("f64").with_max_iterations(500)
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
    xo, io, ro = ora.batch_cg_solve(row_off, rp, ci, v, b, np.zeros_like(b), use_jacobi=True,
                                    tol=1e-10, max_iters=500)
    xh = x.to_numpy()
    for s in range(150):
        r0, r1 = row_off[s], row_off[s + 1]
        bs = b[r0:r1]
        d = gen.csr_to_dense(rp[r0:r1 + 1] - rp[r0], ci[rp[r0]:rp[r1]], v[rp[r0]:rp[r1]])
        assert np.linalg.norm(bs - d @ xh[r0:r1]) / np.linalg.norm(bs) <= 1e-9
    assert np.max(np.abs(logger.num_iterations - io)) <= 1
    # the stencil KAT through the uniform batch::matrix::Csr
    k = KATS["batch_cg"][0]
    srow_off, srp, sci, sv, sb, xs, n, d = stencil_batch(k)
    nb = k["num_batch_items"]
    S = gko.batch.matrix.Csr("f64").from_numpy(ex, nb, n, srp[: n + 1], sci[: srp[n]], sv)
    s2 = (gko.batch.solver.Cg.build("f64").with_max_iterations(k["max_iterations"])
          .with_tolerance(k["tolerance"])
          .with_tolerance_type(gko.batch.stop.tolerance_type.relative).on(ex).generate(S))
    xb = gko.batch.MultiVector("f64").create(ex, nb, n)
    xb.fill(0.0)
    s2.apply(gko.batch.MultiVector("f64").from_numpy(ex, nb, n, sb), xb)
    assert np.max(np.abs(xb.to_numpy() - np.tile(xs, nb))) <= 10 * k["tolerance"]


def test_graph_captured_batch_cg_replays(dev):
    row_off, rp, ci, v, b = gen.uneven_spd_batch(200, n_min=16, n_max=400, seed=21)
    A = D().DeviceBatchCsr(row_off, rp, ci, v, device=dev)
    bd = torch.as_tensor(b).to(dev)
    st = A.settings(tol=1e-10, relative=True, max_iters=500, jacobi=True)
    plan = A.plan()
    x_eager = torch.zeros_like(bd)
    it_eager, res_eager = A.solve(bd, x_eager, st, plan, method="cg")
    x = torch.full_like(bd, 7.0)
    graph, it_g, res_g = A.capture(bd, x, st, plan, method="cg")
    for rep in range(3):
        scratch = torch.randn(1 << 16, dtype=torch.float64, device=dev)
        scratch.mul_(2.0).sum()
        x.fill_(-3.0)
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(x, x_eager), rep
        assert torch.equal(it_g, it_eager), rep
        assert torch.equal(res_g, res_eager), rep
