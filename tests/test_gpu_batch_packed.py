"""The one-launch batched solver (plan.single_launch = 1): the whole uneven
batch in one persistent launch, 256-thread workgroups taking work items from
one queue ordered largest first, small systems side by side in one
workgroup.  Checked against the per-bin register kernels (the same per-row
and per-wave arithmetic for systems above 64 rows), the oracle (a per-system
restatement of reference/solver/batch_bicgstab_kernels.hpp and
batch_cg_kernels.hpp) and the true residuals."""
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


def batch(method, num, seed):
    if method == "cg":
        # three off-diagonals per row before symmetrisation: rows of at most
        # 16 entries, as the one-launch kernel keeps them in registers
        row_off, rp, ci, v, b = gen.uneven_spd_batch(num, n_min=16, n_max=512, k=3, seed=seed)
    else:
        row_off, rp, ci, v, b = gen.uneven_batch(num, n_min=16, n_max=512, seed=seed)
    assert int(np.diff(rp).max()) <= 16
    return row_off, rp, ci, v, b


def true_relres(row_off, rp, ci, v, b, x):
    """Largest ||b - A x|| / ||b|| over the systems with b != 0 (the product
    by the oracle's CSR SpMV in FP64 over global column indices)."""
    first_row = np.repeat(row_off[:-1], np.diff(row_off))          # per row
    gci = (ci + first_row[np.repeat(np.arange(len(rp) - 1), np.diff(rp))]).astype(np.int32)
    ax = ora.csr_spmv(rp, gci, v.astype(np.float64), x.astype(np.float64).reshape(-1, 1))[:, 0]
    b = b.astype(np.float64)
    worst = 0.0
    for s in range(len(row_off) - 1):
        r0, r1 = row_off[s], row_off[s + 1]
        nb = np.linalg.norm(b[r0:r1])
        if nb > 0:
            worst = max(worst, np.linalg.norm(b[r0:r1] - ax[r0:r1]) / nb)
    return worst


@pytest.mark.parametrize("method", ["bicgstab", "cg"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_one_launch_against_per_bin_kernels_and_oracle(dev, method, dtype):
    row_off, rp, ci, v, b = batch(method, 700, seed=21)
    v = v.astype(dtype)
    b = b.astype(dtype)
    A = D().DeviceBatchCsr(row_off, rp, ci, v, device=dev)
    bd = torch.as_tensor(b).to(dev)
    tol = 1e-10 if dtype == np.float64 else 1e-5
    st = A.settings(tol=tol, relative=True, max_iters=500, jacobi=True)
    plan = A.plan()
    assert plan.num_bins == 4
    # per-bin register kernels with two rows per thread where the bin allows
    # it: 256 threads for 257..512 rows, 128 for 129..256, 64 for 65..128
    for i, bs in enumerate([256, 128, 64, 64]):
        plan.block_size[i] = bs
        plan.vectors[i] = 2
        plan.matrix_in_lds[i] = 0
        plan.subwave[i] = 1
    plan.single_launch = 0
    xb = torch.zeros_like(bd)
    itb, resb = A.solve(bd, xb, st, plan, method=method)
    plan.single_launch = 1
    x1 = torch.zeros_like(bd)
    it1, res1 = A.solve(bd, x1, st, plan, method=method)
    itb, it1 = itb.cpu().numpy(), it1.cpu().numpy()
    xb, x1 = xb.cpu().numpy(), x1.cpu().numpy()
    sizes = np.diff(row_off)
    big = sizes > 64
    rows_big = np.repeat(big, sizes)
    # systems above 64 rows: the same rows per thread, per-row sums and wave
    # order as the per-bin kernels
    np.testing.assert_array_equal(itb[big], it1[big])
    eps = np.finfo(dtype).eps
    scale = float(np.abs(xb).max())
    np.testing.assert_allclose(x1[rows_big], xb[rows_big], rtol=16 * eps, atol=4 * eps * scale)
    # every system against the oracle's iteration counts and its true residual
    solve = ora.batch_cg_solve if method == "cg" else ora.batch_bicgstab_solve
    xo, io, ro = solve(row_off, rp, ci, v, b, np.zeros_like(b), use_jacobi=True, relative=True,
                       tol=tol, max_iters=500)
    assert np.all(np.abs(it1 - io) <= (1 if dtype == np.float64 else 3))
    bound = 3 * tol if dtype == np.float64 else 30 * tol
    assert true_relres(row_off, rp, ci, v, b.astype(np.float64), x1.astype(np.float64)) <= bound
    assert np.all(res1.cpu().numpy() >= 0)


def test_one_launch_small_systems_side_by_side(dev):
    """Eight systems of at most 64 rows per workgroup, of different sizes and
    iteration counts (one stops at its first check, b = 0), with a partial
    last work item; the iteration cap is honoured per system."""
    row_off, rp, ci, v, b = gen.uneven_batch(37, n_min=16, n_max=64, seed=4)
    b = b.copy()
    b[row_off[5]:row_off[6]] = 0.0
    A = D().DeviceBatchCsr(row_off, rp, ci, v, device=dev)
    bd = torch.as_tensor(b).to(dev)
    plan = A.plan()
    plan.single_launch = 1
    for cap in (500, 3):
        st = A.settings(tol=1e-10, relative=False, max_iters=cap, jacobi=True)
        x = torch.zeros_like(bd)
        it, res = A.solve(bd, x, st, plan)
        xo, io, ro = ora.batch_bicgstab_solve(row_off, rp, ci, v, b, np.zeros_like(b),
                                              use_jacobi=True, relative=False, tol=1e-10,
                                              max_iters=cap)
        it = it.cpu().numpy()
        assert np.all(np.abs(it - io) <= (1 if cap == 500 else 0))
        assert it[5] == 0 and float(res[5]) == 0.0
        if cap == 3:
            assert int(it.max()) <= 3
            np.testing.assert_allclose(x.cpu().numpy(), xo, rtol=0, atol=1e-12)
        else:
            assert true_relres(row_off, rp, ci, v, b, x.cpu().numpy()) <= 1e-8


def test_one_launch_refused_where_it_does_not_apply(dev):
    from ginkgo_amd import _capi as C
    # rows longer than 16 entries (the SPD batch of eight off-diagonals)
    row_off, rp, ci, v, b = gen.uneven_spd_batch(30, n_min=100, n_max=200, seed=3)
    assert int(np.diff(rp).max()) > 16
    A = D().DeviceBatchCsr(row_off, rp, ci, v, device=dev)
    bd = torch.as_tensor(b).to(dev)
    plan = A.plan()
    plan.single_launch = 1
    with pytest.raises(C.NotSupported):
        A.solve(bd, torch.zeros_like(bd), A.settings(), plan, method="cg")
    plan.single_launch = 2
    with pytest.raises(C.InvalidArgument):
        A.solve(bd, torch.zeros_like(bd), A.settings(), plan, method="cg")
    # the tuner leaves it off where it does not apply
    plan.single_launch = 0
    plan = A.tune(bd, torch.zeros_like(bd), A.settings(tol=1e-8, max_iters=200), plan,
                  method="cg")
    assert plan.single_launch == 0


@pytest.mark.parametrize("method", ["bicgstab", "cg"])
def test_tuner_choice_and_graph_replay(dev, method):
    row_off, rp, ci, v, b = batch(method, 400, seed=10)
    A = D().DeviceBatchCsr(row_off, rp, ci, v, device=dev)
    bd = torch.as_tensor(b).to(dev)
    st = A.settings(tol=1e-10, relative=True, max_iters=500, jacobi=True)
    plan = A.tune(bd, torch.zeros_like(bd), st, A.plan(), method=method)
    assert plan.single_launch in (0, 1)
    for single in (0, 1):
        plan.single_launch = single
        x = torch.zeros_like(bd)
        it, res = A.solve(bd, x, st, plan, method=method)
        assert true_relres(row_off, rp, ci, v, b, x.cpu().numpy()) <= 3e-10
        graph, git, gres = A.capture(bd, torch.empty_like(bd), st, plan, method=method)
        for _ in range(3):
            graph.replay()
            torch.cuda.synchronize()
            np.testing.assert_array_equal(git.cpu().numpy(), it.cpu().numpy())
            np.testing.assert_array_equal(gres.cpu().numpy(), res.cpu().numpy())


def test_one_launch_allocation_free_entry(dev):
    import ctypes
    from ginkgo_amd import _capi as C
    row_off, rp, ci, v, b = gen.uneven_batch(300, seed=31)
    A = D().DeviceBatchCsr(row_off, rp, ci, v, device=dev)
    bd = torch.as_tensor(b).to(dev)
    st = A.settings(tol=1e-10, relative=True, max_iters=500, jacobi=True)
    plan = A.plan()
    plan.single_launch = 1
    x = torch.zeros_like(bd)
    it, res = A.solve(bd, x, st, plan)
    x2 = torch.zeros_like(bd)
    it2 = torch.empty_like(it)
    res2 = torch.empty_like(res)
    c = D().ctx(dev)
    C.call("gkomi_batch_bicgstab_solve_noalloc_f64", ctypes.byref(c), ctypes.byref(A.desc),
           ctypes.byref(st), ctypes.byref(plan), D().ptr(bd), D().ptr(x2), D().ptr(it2),
           D().ptr(res2))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(x2.cpu().numpy(), x.cpu().numpy())
    np.testing.assert_array_equal(it2.cpu().numpy(), it.cpu().numpy())
