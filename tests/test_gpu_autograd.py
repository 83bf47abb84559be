"""Differentiable solve: forward solve, adjoint backward on the device-built
transpose, sampled outer product -- checked against central finite
differences (h = 1e-4) and against a dense FP64 torch.linalg.solve gradient."""
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


def problem(nx, convection, nrhs, dev, seed=7):
    rp, ci, vals, theta, vmap, coef = gen.anisotropic2d(nx, seed=6, convection=convection)
    n = nx * nx
    nnz = len(ci)
    vmap_t = torch.as_tensor(vmap).to(dev)
    coef_t = torch.as_tensor(coef).to(dev)
    theta0 = torch.as_tensor(theta).to(dev)
    # constant (convection) part of the values
    diff0 = torch.zeros(nnz, dtype=torch.float64, device=dev).index_add_(
        0, vmap_t[:, 0], coef_t * theta0[vmap_t[:, 1]])
    const = torch.as_tensor(vals).to(dev) - diff0
    rng = np.random.default_rng(seed)
    b = torch.as_tensor(rng.uniform(-1, 1, (n, nrhs))).to(dev)
    xt = torch.as_tensor(rng.uniform(-1, 1, (n, nrhs))).to(dev)
    rp_t = torch.as_tensor(rp).to(dev)
    ci_t = torch.as_tensor(ci).to(dev)

    def values_of(th):
        return torch.zeros(nnz, dtype=torch.float64, device=dev).index_add(
            0, vmap_t[:, 0], coef_t * th[vmap_t[:, 1]]) + const

    return rp_t, ci_t, theta0, values_of, b, xt, n


def loss_fn(x, xt):
    return 0.5 * ((x - xt) ** 2).sum()


@pytest.mark.parametrize("method,convection,jacobi", [("cg", 0.0, False), ("bicgstab", 10.0, False),
                                                     ("cg", 0.0, True), ("bicgstab", 10.0, True)])
def test_gradients_match_finite_differences_and_dense(dev, method, convection, jacobi):
    """With jacobi the forward and the adjoint solve are both preconditioned
    (scalar Jacobi of A and of its transpose); the gradient is the same."""
    from ginkgo_amd.autograd import sparse_solve
    rp, ci, theta0, values_of, b, xt, n = problem(12, convection, 3, dev)
    tol = 1e-14
    theta = theta0.clone().requires_grad_(True)
    x = sparse_solve(values_of(theta), b, rp, ci, method=method, tol=tol, max_iters=5000,
                     jacobi=jacobi)
    loss = loss_fn(x, xt)
    loss.backward()
    g = theta.grad.detach().cpu().numpy()

    # dense FP64 reference gradient on the CPU
    th_c = theta0.detach().cpu().clone().requires_grad_(True)
    vals_c = values_of(th_c.to(dev)).cpu()
    rp_h, ci_h = rp.cpu().numpy(), ci.cpu().numpy()
    rows = torch.as_tensor(np.repeat(np.arange(n), np.diff(rp_h)))
    Ad = torch.zeros(n, n, dtype=torch.float64).index_put((rows, torch.as_tensor(ci_h).long()),
                                                          vals_c, accumulate=True)
    xd = torch.linalg.solve(Ad, b.cpu())
    loss_fn(xd, xt.cpu()).backward()
    gd = th_c.grad.numpy()
    np.testing.assert_allclose(x.detach().cpu().numpy(), xd.detach().numpy(), rtol=0, atol=1e-10)
    assert np.max(np.abs(g - gd)) <= 1e-8 * np.max(np.abs(gd))

    # central finite differences on random theta entries.  h = 1e-4 balances
    # the O(h^2) truncation error (about 3e-8 relative here, measured with a
    # dense solve) against the rounding noise of L divided by 2h, which at
    # h = 1e-6 already reaches 4e-6 relative with an exact dense solve.
    rng = np.random.default_rng(8)
    idx = rng.choice(theta0.numel(), 12, replace=False)
    h = 1e-4
    for i in idx:
        tp = theta0.clone()
        tm = theta0.clone()
        tp[i] += h
        tm[i] -= h
        lp = loss_fn(sparse_solve(values_of(tp), b, rp, ci, method=method, tol=tol,
                                  max_iters=5000, jacobi=jacobi), xt).item()
        lm = loss_fn(sparse_solve(values_of(tm), b, rp, ci, method=method, tol=tol,
                                  max_iters=5000, jacobi=jacobi), xt).item()
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(fd), 1e-3 * np.max(np.abs(gd))), (i, fd, g[i])


def test_rhs_gradient_is_adjoint_solution(dev):
    from ginkgo_amd.autograd import sparse_solve
    rp, ci, theta0, values_of, b, xt, n = problem(10, 0.0, 2, dev)
    bb = b.clone().requires_grad_(True)
    x = sparse_solve(values_of(theta0), bb, rp, ci, tol=1e-14)
    loss_fn(x, xt).backward()
    # A symmetric: dL/db = A^{-1} (x - xt)
    lam = sparse_solve(values_of(theta0), (x - xt).detach(), rp, ci, tol=1e-14)
    np.testing.assert_allclose(bb.grad.cpu().numpy(), lam.cpu().numpy(), rtol=0, atol=1e-12)


@pytest.mark.parametrize("nrhs", [1, 3, 64, 70])
def test_sampled_outer_product(dev, nrhs):
    from ginkgo_amd.autograd import sampled_outer_product
    rp, ci, v = gen.poisson2d(20)
    n = 400
    rng = np.random.default_rng(nrhs)
    lam = rng.uniform(-1, 1, (n, nrhs))
    x = rng.uniform(-1, 1, (n, nrhs))
    out = sampled_outer_product(torch.as_tensor(rp).to(dev), torch.as_tensor(ci).to(dev),
                                torch.as_tensor(v).to(dev), torch.as_tensor(lam).to(dev),
                                torch.as_tensor(x).to(dev), alpha=-1.0).cpu().numpy()
    rows = np.repeat(np.arange(n), np.diff(rp))
    ref = -np.einsum("kj,kj->k", lam[rows], x[ci])
    np.testing.assert_allclose(out, ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("method,convection", [("cg", 0.0), ("bicgstab", 10.0)])
def test_unconverged_solves_raise(dev, method, convection):
    """A forward solve stopped by the iteration limit raises; so does an
    adjoint solve that cannot converge while the forward one did (b = 0 gives
    x = 0 at iteration 0, the adjoint right-hand side x - x_t does not)."""
    from ginkgo_amd.autograd import SolveNotConverged, sparse_solve
    rp, ci, theta0, values_of, b, xt, n = problem(12, convection, 2, dev)
    theta = theta0.clone().requires_grad_(True)
    with pytest.raises(SolveNotConverged, match="forward"):
        sparse_solve(values_of(theta), b, rp, ci, method=method, tol=1e-12, max_iters=3)
    x = sparse_solve(values_of(theta), torch.zeros_like(b), rp, ci, method=method, tol=1e-12,
                     max_iters=3)
    assert float(x.detach().abs().max()) == 0.0
    with pytest.raises(SolveNotConverged, match="adjoint"):
        loss_fn(x, xt).backward()
    # the policy can downgrade the error to a warning
    with pytest.warns(RuntimeWarning, match="forward"):
        sparse_solve(values_of(theta0), b, rp, ci, method=method, tol=1e-12, max_iters=3,
                     on_not_converged="warn")
