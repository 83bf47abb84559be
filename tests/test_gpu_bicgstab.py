"""Fused single-RHS BiCGSTAB (five kernels per iteration, both convergence
checks on the device) against the oracle's restatement of
core/solver/bicgstab.cpp:95-255 and reference/solver/bicgstab_kernels.cpp,
and against direct solves."""
import numpy as np
import pytest

import oracle as ora
from ginkgo_amd import generators as gen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TD = {np.float64: torch.float64, np.float32: torch.float32}


@pytest.fixture(scope="module")
def dev():
    from ginkgo_amd import _capi
    _capi.lib()
    return torch.device("cuda", 0)


def D():
    from ginkgo_amd import device
    return device


def to_dev(a, dev):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dev)


def convection_diffusion(nx, convection=10.0):
    rp, ci, v, *_ = gen.anisotropic2d(nx, seed=6, convection=convection)
    return rp, ci, v


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 1e-5)])
@pytest.mark.parametrize("jacobi", [True, False])
def test_fused_bicgstab_convection_diffusion(dev, dtype, tol, jacobi):
    rp, ci, v = convection_diffusion(40)
    v = v.astype(dtype)
    n = len(rp) - 1
    b = np.random.default_rng(7).uniform(-1, 1, n).astype(dtype)
    crit = [{"kind": "iteration", "max_iters": 2000},
            {"kind": "residual_norm", "reduction_factor": tol}]
    inv = ora.jacobi_inverse_diagonal(rp, ci, v) if jacobi else None
    xo, so, io = ora.bicgstab_solve(rp, ci, v, b.reshape(-1, 1), np.zeros((n, 1), dtype), crit,
                                    inv_diag=inv)
    A = D().DeviceCsr(rp, ci, v, strategy="classical", device=dev)
    solver = D().FusedBicgstab(A, crit, jacobi=jacobi, chunk=8)
    x = torch.zeros(n, dtype=TD[dtype], device=dev)
    iters = solver.apply(to_dev(b, dev), x)
    band = 3 if dtype == np.float64 else max(3, int(0.1 * io.iterations))
    assert abs(iters - io.iterations) <= band, (iters, io.iterations)
    assert solver.status.item() == so[0]
    xh = x.cpu().numpy().astype(np.float64)
    dense = gen.csr_to_dense(rp, ci, v.astype(np.float64))
    # the recursive residual met the criterion; the true residual is close to it
    res = np.linalg.norm(b - dense @ xh) / np.linalg.norm(b)
    assert res <= tol * (2 if dtype == np.float64 else 20)
    if dtype == np.float64:
        exact = np.linalg.solve(dense, b.astype(np.float64))
        np.testing.assert_allclose(xh, exact, rtol=0, atol=1e-7 * np.abs(exact).max())


def test_fused_bicgstab_half_step_stop(dev):
    """A = c I: s = 0 after the first half step, so the half-step check stops
    the solve at iteration 0 and finalize adds alpha y (status converged by
    criterion 2, then finalized)."""
    n = 3000
    c = 4.0
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    v = np.full(n, c)
    b = np.random.default_rng(5).uniform(-1, 1, n)
    crit = [{"kind": "iteration", "max_iters": 100},
            {"kind": "residual_norm", "reduction_factor": 1e-12}]
    xo, so, io = ora.bicgstab_solve(rp, ci, v, b.reshape(-1, 1), np.zeros((n, 1)), crit)
    assert io.iterations == 0 and so[0] == (0x80 | 0x40 | 2)
    A = D().DeviceCsr(rp, ci, v, strategy="classical", device=dev)
    for jacobi in (False, True):
        solver = D().FusedBicgstab(A, crit, jacobi=jacobi, chunk=4)
        x = torch.zeros(n, dtype=torch.float64, device=dev)
        assert solver.apply(to_dev(b, dev), x) == 0
        assert solver.status.item() == so[0]
        np.testing.assert_allclose(x.cpu().numpy(), b / c, rtol=1e-15)


def test_fused_bicgstab_iteration_limit_matches_oracle(dev):
    """Stopping by the iteration limit: the same 6 iterations as the oracle,
    status 'criterion 1, finalized', and the same iterate to rounding."""
    rp, ci, v = convection_diffusion(24)
    n = len(rp) - 1
    b = np.ones(n)
    crit = [{"kind": "iteration", "max_iters": 6},
            {"kind": "residual_norm", "reduction_factor": 1e-14}]
    xo, so, io = ora.bicgstab_solve(rp, ci, v, b.reshape(-1, 1), np.zeros((n, 1)), crit)
    A = D().DeviceCsr(rp, ci, v, strategy="classical", device=dev)
    solver = D().FusedBicgstab(A, crit, chunk=16)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    assert solver.apply(to_dev(b, dev), x) == 6 == io.iterations
    assert solver.status.item() == so[0] == (1 | 0x40)
    np.testing.assert_allclose(x.cpu().numpy(), xo[:, 0], rtol=1e-9, atol=1e-12)
    # a second solve with the same workspace repeats the first bit for bit
    x2 = torch.zeros(n, dtype=torch.float64, device=dev)
    assert solver.apply(to_dev(b, dev), x2) == 6
    np.testing.assert_array_equal(x.cpu().numpy(), x2.cpu().numpy())


def test_fused_bicgstab_zero_rhs_and_initial_guess(dev):
    rp, ci, v = convection_diffusion(16)
    n = len(rp) - 1
    crit = [{"kind": "iteration", "max_iters": 100},
            {"kind": "residual_norm", "reduction_factor": 1e-10}]
    A = D().DeviceCsr(rp, ci, v, strategy="classical", device=dev)
    solver = D().FusedBicgstab(A, crit, chunk=4)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    assert solver.apply(torch.zeros(n, dtype=torch.float64, device=dev), x) == 0
    assert torch.all(x == 0)
    # an initial guess that is the solution: r0 = b - A x0 ~ 0 stops at once
    # with the initial-residual baseline 'absolute' criterion
    dense = gen.csr_to_dense(rp, ci, v)
    xs = np.linalg.solve(dense, np.ones(n))
    crit_abs = [{"kind": "iteration", "max_iters": 100},
                {"kind": "residual_norm", "reduction_factor": 1e-10, "baseline": "absolute"}]
    xo, so, io = ora.bicgstab_solve(rp, ci, v, np.ones((n, 1)), xs.reshape(-1, 1), crit_abs)
    solver = D().FusedBicgstab(A, crit_abs, chunk=4)
    x = to_dev(xs, dev)
    it = solver.apply(to_dev(np.ones(n), dev), x)
    assert it == io.iterations
    assert solver.status.item() == so[0]


@pytest.mark.parametrize("maker", ["convection", "d27"])
def test_fused_bicgstab_compact_offsets_bit_identical(dev, maker):
    rp, ci, v = convection_diffusion(32) if maker == "convection" else gen.diffusion27(12)
    n = len(rp) - 1
    b = to_dev(np.random.default_rng(2).uniform(-1, 1, n), dev)
    crit = [{"kind": "iteration", "max_iters": 2000},
            {"kind": "residual_norm", "reduction_factor": 1e-10}]
    out = []
    for compact in (False, True):
        A = D().DeviceCsr(rp, ci, v, strategy="classical", device=dev)
        if compact:
            assert A.compact()
        x = torch.zeros(n, dtype=torch.float64, device=dev)
        it = D().FusedBicgstab(A, crit, jacobi=True, chunk=8).apply(b, x)
        out.append((it, x.cpu().numpy()))
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_cpp_bicgstab_one_column_takes_fused_path(dev):
    import ginkgo_amd.gko as gko
    rp, ci, v = convection_diffusion(48)
    n = len(rp) - 1
    b_h = np.random.default_rng(3).uniform(-1, 1, (n, 1))
    spec = [{"kind": "iteration", "max_iters": 1000},
            {"kind": "residual_norm", "reduction_factor": 1e-12}]
    inv = ora.jacobi_inverse_diagonal(rp, ci, v)
    xo, so, io = ora.bicgstab_solve(rp, ci, v, b_h, np.zeros_like(b_h), spec, inv_diag=inv)
    ex = gko.Mi355xExecutor.create(0, torch.cuda.current_stream().cuda_stream)
    A = gko.csr_from_host(ex, rp, ci, v)
    s = (gko.solver.Bicgstab.build("f64")
         .with_criteria(gko.stop.Iteration.build().with_max_iters(1000),
                        gko.stop.ResidualNorm.build("f64").with_reduction_factor(1e-12))
         .with_preconditioner(gko.preconditioner.Jacobi.build("f64").with_max_block_size(1))
         .on(ex).generate(A))
    x = gko.matrix.Dense("f64").create(ex, n, 1)
    x.fill(0.0)
    s.apply(gko.matrix.Dense("f64").from_numpy(ex, b_h), x)
    assert s.last_used_fused_path
    assert abs(s.last_iterations - io.iterations) <= 3
    assert s.last_status == [int(t) for t in so]
    dense = gen.csr_to_dense(rp, ci, v)
    exact = np.linalg.solve(dense, b_h)
    np.testing.assert_allclose(x.to_numpy(), exact, rtol=0, atol=1e-9 * np.abs(exact).max())


def test_fused_bicgstab_implicit_residual_norm(dev):
    """ImplicitResidualNorm reads rho = rr.r as the squared norm
    (core/solver/bicgstab.cpp:169-175): same stop as the oracle."""
    rp, ci, v = convection_diffusion(32)
    n = len(rp) - 1
    b = np.random.default_rng(9).uniform(-1, 1, n)
    crit = [{"kind": "iteration", "max_iters": 1000},
            {"kind": "implicit_residual_norm", "reduction_factor": 1e-9}]
    xo, so, io = ora.bicgstab_solve(rp, ci, v, b.reshape(-1, 1), np.zeros((n, 1)), crit)
    A = D().DeviceCsr(rp, ci, v, strategy="classical", device=dev)
    solver = D().FusedBicgstab(A, crit, chunk=8)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    it = solver.apply(to_dev(b, dev), x)
    assert abs(it - io.iterations) <= 3, (it, io.iterations)
    assert solver.status.item() == so[0]
