"""The sliced copy of a CSR matrix (SELL-P layout with 64-row slices, no
reference counterpart): the classical single right-hand-side SpMV and the
fused CG read it when it is present.

* Products against the oracle's FP64 product within 4 k eps (|A||x|)_i, with
  int32 slot columns and with 2-byte offsets (bit-identical to each other),
  plain and advanced (beta = 0 ignoring NaN in c), strided b, empty rows and
  a row count that is not a multiple of 64.
* The fused CG over the sliced copy against the oracle's CG: iteration
  counts within the FP64 band, the true residual at the tolerance; through
  the Python layer and through Csr::build_slices in C++.
"""
import sys

import numpy as np
import pytest

import oracle as ora
from ginkgo_amd import generators as gen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

EPS = {np.float64: np.finfo(np.float64).eps, np.float32: np.finfo(np.float32).eps}
TD = {np.float64: torch.float64, np.float32: torch.float32}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    from ginkgo_amd import _capi
    _capi.lib()
    return torch.device("cuda", 0)


def D():
    from ginkgo_amd import device
    return device


def ragged(seed=13, n=3001, n_cols=2900):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 30, n)
    lens[rng.choice(n, 200, replace=False)] = 0
    lens[::211] = 90
    rp = np.zeros(n + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    ci = np.concatenate([np.sort(rng.choice(n_cols, size=int(L), replace=False))
                         if L else np.zeros(0, np.int64) for L in lens]).astype(np.int32)
    return rp, ci, rng.uniform(-1, 1, len(ci)), n_cols


def cases():
    out = []
    rp, ci, v = gen.laplace3d(20)
    out.append(("lap20", rp, ci, v, len(rp) - 1))
    rp, ci, v = gen.diffusion27(12)
    out.append(("d27_12", rp, ci, v, len(rp) - 1))
    rp, ci, v, nc = ragged()
    out.append(("ragged", rp, ci, v, nc))
    return out


CASES = cases()


def bound_check(rp, ci, v, x, y, dtype, label):
    vw = v.astype(dtype).astype(np.float64)
    xw = x.astype(dtype).astype(np.float64).reshape(-1, 1)
    exact = ora.csr_spmv(rp, ci, vw, xw)[:, 0]
    ax = ora.csr_spmv(rp, ci, np.abs(vw), np.abs(xw))[:, 0]
    k = np.maximum(np.diff(rp), 1).astype(np.float64)
    bound = 4 * k * (EPS[dtype] + EPS[np.float64]) * ax
    err = np.abs(y.astype(np.float64) - exact)
    say(f"{label}: max |y - y_oracle| {err.max() if err.size else 0:.3e}")
    assert np.all(err <= bound), label


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_sliced_spmv(dev, case, dtype):
    name, rp, ci, v, n_cols = case
    vd = v.astype(dtype)
    x = np.random.default_rng(3).uniform(-1, 1, (n_cols, 2)).astype(dtype)
    xd = torch.as_tensor(x).to(dev)
    A = D().DeviceCsr(rp, ci, vd, n_cols=n_cols, strategy="classical", device=dev)
    compact = A.build_slices(compact=True)
    assert compact
    y_off = A.spmv(xd[:, 0:1]).cpu().numpy()[:, 0]
    bound_check(rp, ci, v, x[:, 0], y_off, dtype, f"sliced {name} offsets")
    np.testing.assert_array_equal(A.spmv(xd[:, 0:1]).cpu().numpy()[:, 0], y_off)
    A.build_slices(compact=False)
    y_col = A.spmv(xd[:, 0:1]).cpu().numpy()[:, 0]
    np.testing.assert_array_equal(y_col, y_off)
    # strided b (the second column of a row-major 2-column block)
    y2 = A.spmv(xd[:, 1:2]).cpu().numpy()[:, 0]
    bound_check(rp, ci, v, x[:, 1], y2, dtype, f"sliced {name} strided")
    # c = alpha A b + beta c, and beta = 0 ignoring NaN in c
    n = len(rp) - 1
    c0 = np.random.default_rng(4).uniform(-1, 1, (n, 1)).astype(dtype)
    cd = torch.as_tensor(c0).to(dev)
    A.advanced_spmv(2.0, xd[:, 0:1], -0.5, cd)
    ref = 2.0 * y_off.astype(np.float64) - 0.5 * c0[:, 0].astype(np.float64)
    tol = 8 * EPS[dtype] * (2 * np.abs(y_off) + 0.5 * np.abs(c0[:, 0])) + 8 * EPS[dtype] * \
        ora.csr_spmv(rp, ci, np.abs(vd.astype(np.float64)),
                     np.abs(x[:, :1].astype(np.float64)))[:, 0]
    assert np.all(np.abs(cd.cpu().numpy()[:, 0] - ref) <= tol)
    cn = torch.full((n, 1), float("nan"), dtype=TD[dtype], device=dev)
    A.advanced_spmv(1.0, xd[:, 0:1], 0.0, cn)
    np.testing.assert_array_equal(cn.cpu().numpy()[:, 0], y_off)


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_fused_cg_over_slices(dev, dtype):
    """The fused CG reading the sliced copy: the oracle's iteration count
    within the band of the other fused CG tests, the true residual at the
    tolerance."""
    rp, ci, v = gen.diffusion27(24)
    n = len(rp) - 1
    xt = np.random.default_rng(4).uniform(-1, 1, (n, 1))
    b = ora.csr_spmv(rp, ci, v, xt)
    tol = 1e-10 if dtype == np.float64 else 1e-5
    spec = [{"kind": "iteration", "max_iters": 5000},
            {"kind": "residual_norm", "reduction_factor": tol}]
    xo, so, io = ora.cg_solve(rp, ci, v.astype(dtype).astype(np.float64), b,
                              np.zeros_like(b), spec)
    A = D().DeviceCsr(rp, ci, v.astype(dtype), strategy="classical", device=dev)
    A.compact()
    A.build_slices(compact=True)
    cg = D().FusedCg(A, spec)
    bd = torch.as_tensor(b[:, 0].astype(dtype)).to(dev)
    x = torch.zeros_like(bd)
    it = cg.apply(bd, x)
    xh = x.cpu().numpy().astype(np.float64)
    res = np.linalg.norm(b[:, 0] - ora.csr_spmv(rp, ci, v, xh.reshape(-1, 1))[:, 0]) / \
        np.linalg.norm(b)
    say(f"sliced CG {np.dtype(dtype).name}: {it} iterations (oracle {io.iterations}), "
        f"true relative residual {res:.3e}")
    band = 3 if dtype == np.float64 else max(3, io.iterations // 10)
    assert abs(it - io.iterations) <= band
    assert res <= (1.05 * tol if dtype == np.float64 else 20 * tol)


def test_cpp_build_slices(dev):
    """Csr::build_slices through the gko:: surface: Cg over the sliced copy
    matches the same Cg over the plain Csr to the tolerance, the classical
    apply matches within the bound, and a mutable accessor drops the copy."""
    import ginkgo_amd.gko as gko
    ex = gko.Mi355xExecutor.create(0, torch.cuda.current_stream().cuda_stream)
    rp, ci, v = gen.diffusion27(16)
    n = len(rp) - 1
    b_h = ora.csr_spmv(rp, ci, v, np.random.default_rng(5).uniform(-1, 1, (n, 1)))
    crit = (gko.stop.Iteration.build().with_max_iters(3000),
            gko.stop.ResidualNorm.build("f64").with_reduction_factor(1e-10))
    out = {}
    for sliced in (False, True):
        A = gko.csr_from_host(ex, rp, ci, v, strategy="classical")
        if sliced:
            A.build_slices(True)
            assert A.has_slices
        s = gko.solver.Cg.build("f64").with_criteria(*crit).on(ex).generate(A)
        x = gko.matrix.Dense("f64").create(ex, n, 1)
        x.fill(0.0)
        s.apply(gko.matrix.Dense("f64").from_numpy(ex, b_h), x)
        out[sliced] = (x.to_numpy(), s.last_iterations)
        y = gko.matrix.Dense("f64").create(ex, n, 1)
        A.apply(gko.matrix.Dense("f64").from_numpy(ex, b_h), y)
        bound_check(rp, ci, v, b_h[:, 0], y.to_numpy()[:, 0], np.float64,
                    f"C++ apply sliced={sliced}")
        if sliced:
            _ = A.values_data_ptr
            assert not A.has_slices
    assert abs(out[True][1] - out[False][1]) <= 3
    for sliced in (False, True):
        xh = out[sliced][0]
        res = np.linalg.norm(b_h - ora.csr_spmv(rp, ci, v, xh)) / np.linalg.norm(b_h)
        assert res <= 1.05e-10
