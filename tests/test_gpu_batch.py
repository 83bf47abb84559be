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
            x = torch.zeros_like(bd)
            iters, res = A.solve(bd, x, st, plan)
            outs.append((x.cpu().numpy(), iters.cpu().numpy()))
    # identical arithmetic order within a block size: bit-identical
    np.testing.assert_array_equal(outs[0][0], outs[3][0])
    np.testing.assert_array_equal(outs[2][0], outs[5][0])
    for xo, it in outs[1:]:
        np.testing.assert_allclose(xo, outs[0][0], rtol=0, atol=1e-8)


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
