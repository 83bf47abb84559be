"""The allocation-free batched BiCGSTAB (gkomi_batch_bicgstab_solve_noalloc_f64)
against the workspace entry point it varies (gkomi_batch_bicgstab_solve_f64):
the same kernels over the same plan, so the solutions, iteration counts and
residual norms agree bit for bit.  Two calls in a row, the second on a
smaller batch, then one without a plan; each checked against the workspace
entry point on the same input."""
import ctypes

import numpy as np
import pytest

from ginkgo_amd import generators as gen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dev():
    from ginkgo_amd import _capi
    _capi.lib()
    return torch.device("cuda", 0)


def _noalloc(A, b, x, st, plan):
    from ginkgo_amd import _capi as C
    from ginkgo_amd import device as D
    iters = torch.empty(A.num_systems, dtype=torch.int32, device=A.device)
    res = torch.empty(A.num_systems, dtype=torch.float64, device=A.device)
    c = D.ctx(A.device)
    C.call("gkomi_batch_bicgstab_solve_noalloc_f64", ctypes.byref(c), ctypes.byref(A.desc),
           ctypes.byref(st), ctypes.byref(plan) if plan is not None else None, D.ptr(b),
           D.ptr(x), D.ptr(iters), D.ptr(res))
    return iters, res


@pytest.mark.parametrize("num,seed", [(600, 31)])
def test_noalloc_batch_matches_workspace_entry(dev, num, seed):
    from ginkgo_amd import device as D
    row_off, rp, ci, v, b = gen.uneven_batch(num, seed=seed)
    cases = [(row_off, rp, ci, v, b)]
    # a smaller batch second: the first systems of the same set
    k = num // 5
    r1 = int(row_off[k])
    cases.append((row_off[:k + 1], rp[:r1 + 1], ci[:rp[r1]], v[:rp[r1]], b[:r1]))
    for i, (ro, p, c, vals, rhs) in enumerate(cases):
        A = D.DeviceBatchCsr(ro, p, c, vals, device=dev)
        bd = torch.as_tensor(rhs).to(dev)
        st = A.settings(tol=1e-10, relative=True, max_iters=500, jacobi=True)
        for plan in (A.plan(), None):
            x_ref = torch.zeros_like(bd)
            it_ref, res_ref = A.solve(bd, x_ref, st, plan)
            x = torch.zeros_like(bd)
            it, res = _noalloc(A, bd, x, st, plan)
            torch.cuda.synchronize()
            print(f"call {i} ({A.num_systems} systems, plan {'on' if plan else 'off'}): "
                  f"iterations max {int(it.max())}", flush=True)
            assert torch.equal(it, it_ref)
            assert torch.equal(res, res_ref)
            assert torch.equal(x, x_ref)
            assert int(it.max()) < 500
