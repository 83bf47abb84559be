"""Parity of the gfx950 kernels (called through the C ABI) with the oracle.

Bars: index and layout work bit-exact; FP SpMV within
|y - y_ref|_i <= 4 k eps (|A||x|)_i for row length k; reductions within the
Higham bound; solver runs meet the same relative-residual criterion as the
oracle's with iteration counts within a small band (summation order differs).
"""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle as ora
from ginkgo_amd import generators as gen

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
DTYPES = [np.float64, np.float32]
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


def spmv_bound(rp, ci, v, x, dtype):
    a = gen.csr_to_dense(rp, ci, np.abs(v).astype(np.float64), x.shape[0])
    k = np.maximum(np.diff(rp), 1).astype(np.float64)[:, None]
    return 4 * k * np.finfo(dtype).eps * (a @ np.abs(x.astype(np.float64))) + 1e-30


def check_spmv(rp, ci, v, x, y, dtype):
    ref = ora.csr_spmv(rp, ci, v.astype(dtype), x.astype(dtype)).astype(np.float64)
    exact = gen.csr_to_dense(rp, ci, v.astype(np.float64), x.shape[0]) @ x.astype(np.float64)
    bound = spmv_bound(rp, ci, v, x, dtype)
    assert np.all(np.abs(y.astype(np.float64) - exact) <= bound + np.abs(ref - exact))


def irregular(n, seed):
    """rows of lengths 0..200 plus a few very long ones; some empty rows."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 12, n)
    lens[rng.integers(0, n, 5)] = rng.integers(100, 300, 5)
    lens[rng.integers(0, n, 2)] = 3000
    lens = np.minimum(lens, n)
    rows = np.repeat(np.arange(n), lens)
    cols = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in lens])
    vals = rng.uniform(-1, 1, rows.shape[0])
    rp = np.zeros(n + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    return rp, cols.astype(np.int32), vals


# ----------------------------------------------------------------- SpMV ----

@pytest.mark.parametrize("strategy", ["classical", "load_balance"])
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", KATS["spmv"], ids=lambda c: c["name"])
def test_spmv_kats(dev, case, dtype, strategy):
    m = KATS["matrices"][case["matrix"]]
    A = D().DeviceCsr(m["row_ptrs"], m["col_idxs"], np.array(m["values"], dtype), m["n_cols"],
                      strategy=strategy, device=dev)
    x = to_dev(np.array(case["x"], dtype), dev)
    exp = np.array(case["expected"], dtype)
    if "alpha" in case:
        y = to_dev(np.array([[float(v) for v in r] for r in case["y"]], dtype), dev)
        A.advanced_spmv(case["alpha"], x, case["beta"], y)
    else:
        y = A.spmv(x)
    np.testing.assert_array_equal(y.cpu().numpy(), exp)


@pytest.mark.parametrize("subwave", [1, 2, 4, 8, 16, 32, 64])
@pytest.mark.parametrize("dtype", DTYPES)
def test_spmv_classical_every_subwave(dev, dtype, subwave):
    rp, ci, v = irregular(3000, seed=subwave)
    v = v.astype(dtype)
    x = np.random.default_rng(1).uniform(-1, 1, (3000, 1)).astype(dtype)
    A = D().DeviceCsr(rp, ci, v, strategy="classical", subwave=subwave, device=dev)
    y = A.spmv(to_dev(x, dev)).cpu().numpy()
    check_spmv(rp, ci, v, x, y, dtype)
    # rows in flight per lane group (subwave bits 8..15) change the schedule,
    # never the per-row summation order: bit-identical results
    for rpg in (1, 2, 4, 8):
        A2 = D().DeviceCsr(rp, ci, v, strategy="classical", subwave=subwave | (rpg << 8),
                           device=dev)
        y2 = A2.spmv(to_dev(x, dev)).cpu().numpy()
        np.testing.assert_array_equal(y2, y)
        # 2-byte column offsets instead of the indices: the same bits
        assert A2.compact()
        y3 = A2.spmv(to_dev(x, dev)).cpu().numpy()
        np.testing.assert_array_equal(y3, y)


STAGED = 0x10000     # GKOMI_CSR_STAGED


def auto_staged_lanes(rp, dtype):
    n, nnz = len(rp) - 1, int(rp[-1])
    for g in ((1, 2, 4) if dtype == np.float64 else (1, 2, 4, 8)):
        if (64 // g) * nnz <= 448 * n:
            return g
    return 0


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", ["irregular", "stencil7", "stencil27", "tiny"])
def test_spmv_staged_every_group(dev, dtype, case):
    """The staged kernel (windows of 64 / G rows, their entries copied to LDS
    with contiguous loads, G lanes per row): windows whose entries exceed the
    448-entry staging chunk (rows of 100..3000 entries cross chunks), empty
    rows, a last partial window.  Sub-lane s of a row (lane r + s 64/G)
    adds entries s, s + G, ... and a butterfly with lane strides 64/G, ...
    adds the sub-lanes as the group tree pairs lanes, so the product equals
    the sub-wave shape with G lanes per row bit for bit, with the indices or
    the 2-byte offsets; the automatic classical choice for short rows is that
    kernel."""
    if case == "irregular":
        rp, ci, v = irregular(3000, seed=7)
    elif case == "stencil7":
        rp, ci, v = gen.laplace3d(21)
    elif case == "stencil27":
        rp, ci, v = gen.diffusion27(14)
    else:
        rp, ci, v = irregular(37, seed=2)
    n = len(rp) - 1
    v = v.astype(dtype)
    x = np.random.default_rng(1).uniform(-1, 1, (n, 1)).astype(dtype)
    xd = to_dev(x, dev)
    outs = {}
    for g in (1, 2, 4, 8, 16):
        sub = D().DeviceCsr(rp, ci, v, strategy="classical", subwave=g | (1 << 8), device=dev)
        y = sub.spmv(xd).cpu().numpy()
        check_spmv(rp, ci, v, x, y, dtype)
        A = D().DeviceCsr(rp, ci, v, strategy="classical", subwave=STAGED | g, device=dev)
        np.testing.assert_array_equal(A.spmv(xd).cpu().numpy(), y)
        if A.compact():
            np.testing.assert_array_equal(A.spmv(xd).cpu().numpy(), y)
        outs[g] = y
    g = auto_staged_lanes(rp, dtype)
    if g:
        auto = D().DeviceCsr(rp, ci, v, strategy="classical", device=dev)
        np.testing.assert_array_equal(auto.spmv(xd).cpu().numpy(), outs[g])
    # advanced: y = alpha A x + beta y; beta = 0 ignores NaN in y
    A = D().DeviceCsr(rp, ci, v, strategy="classical", subwave=STAGED | 4, device=dev)
    y0 = np.random.default_rng(3).uniform(-1, 1, (n, 1)).astype(dtype)
    yd = to_dev(y0, dev)
    A.advanced_spmv(-2.0, xd, 0.5, yd)
    ref = ora.csr_advanced_spmv(-2.0, rp, ci, v, x, 0.5, y0.copy()).astype(np.float64)
    tol = 2 * spmv_bound(rp, ci, v, x, dtype) + 4 * np.finfo(dtype).eps * np.abs(y0)
    assert np.all(np.abs(yd.cpu().numpy().astype(np.float64) - ref) <= tol)
    nan = to_dev(np.full((n, 1), np.nan, dtype), dev)
    A.advanced_spmv(1.0, xd, 0.0, nan)
    np.testing.assert_array_equal(nan.cpu().numpy(), outs[4])
    # a lane count outside 1..16 is refused
    from ginkgo_amd import _capi as C
    with pytest.raises(C.InvalidArgument):
        D().DeviceCsr(rp, ci, v, strategy="classical", subwave=STAGED | 3, device=dev).spmv(xd)
    # a strided right-hand side takes the sub-wave kernel
    X = np.random.default_rng(4).uniform(-1, 1, (n, 3)).astype(dtype)
    Xd = to_dev(X, dev)
    ys = A.spmv(Xd[:, 1:2]).cpu().numpy()
    check_spmv(rp, ci, v, X[:, 1:2].copy(), ys, dtype)


def test_compact_offsets_refuse_wide_rows(dev):
    n = 40000
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    ci[5] = 39000                        # offset 38995 does not fit int16
    A = D().DeviceCsr(rp, ci, np.ones(n), strategy="classical", device=dev)
    assert not A.compact()
    assert A.col_offsets is None
    ci[5] = 5 + 32767                    # the widest offset that fits
    B = D().DeviceCsr(rp, ci, np.ones(n), strategy="classical", device=dev)
    assert B.compact()
    x = torch.arange(n, dtype=torch.float64, device=dev)
    y = B.spmv(x).cpu().numpy()
    assert y[5] == 5 + 32767 and y[6] == 6


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 1000, 4099, 100003])
@pytest.mark.parametrize("dtype", DTYPES)
def test_dense_copy(dev, dtype, n):
    # packed (16-byte vector path plus element tail), offset by one element
    # (element path), and strided
    src = torch.as_tensor(np.random.default_rng(n).uniform(-1, 1, n + 3).astype(dtype)).to(dev)
    for off in (0, 1):
        dst = torch.full((n + 3,), 7.0, dtype=TD[dtype], device=dev)
        D().copy(src[off:off + n], dst[off:off + n])
        assert torch.equal(dst[off:off + n], src[off:off + n])
        assert torch.all(dst[:off] == 7.0) and torch.all(dst[off + n:] == 7.0)
    if n >= 2:
        a = src[: (n // 2) * 2].reshape(n // 2, 2)
        b = torch.zeros(n // 2, 3, dtype=TD[dtype], device=dev)
        D().copy(a, b[:, :2])
        assert torch.equal(b[:, :2], a) and torch.all(b[:, 2] == 0)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("maker", ["irregular", "powerlaw", "laplace", "empty_rows"])
def test_spmv_load_balance(dev, dtype, maker):
    if maker == "irregular":
        rp, ci, v = irregular(5000, seed=3)
    elif maker == "powerlaw":
        rp, ci, v = gen.powerlaw(20000, seed=5)
    elif maker == "laplace":
        rp, ci, v = gen.laplace3d(17)
    else:
        rp = np.zeros(3001, np.int32)
        rp[2000:] = np.arange(1001)  # 2000 empty rows, then one nnz per row
        ci = np.arange(1000, dtype=np.int32)
        v = np.ones(1000)
    v = v.astype(dtype)
    n = len(rp) - 1
    x = np.random.default_rng(2).uniform(-1, 1, (n, 1)).astype(dtype)
    A = D().DeviceCsr(rp, ci, v, strategy="load_balance", device=dev)
    y = A.spmv(to_dev(x, dev)).cpu().numpy()
    check_spmv(rp, ci, v, x, y, dtype)
    # deterministic: a second run is bit-identical
    y2 = A.spmv(to_dev(x, dev)).cpu().numpy()
    np.testing.assert_array_equal(y, y2)


@pytest.mark.parametrize("nrhs", [2, 3, 64, 70])
def test_spmv_many_rhs_and_advanced(dev, nrhs):
    rp, ci, v = gen.poisson2d(40)
    n = 1600
    rng = np.random.default_rng(nrhs)
    x = rng.uniform(-1, 1, (n, nrhs))
    A = D().DeviceCsr(rp, ci, v, device=dev)
    y = A.spmv(to_dev(x, dev)).cpu().numpy()
    check_spmv(rp, ci, v, x, y, np.float64)
    c0 = rng.uniform(-1, 1, (n, nrhs))
    c = to_dev(c0, dev)
    A.advanced_spmv(-0.5, to_dev(x, dev), 2.0, c)
    ref = ora.csr_advanced_spmv(-0.5, rp, ci, v, x, 2.0, c0)
    np.testing.assert_allclose(c.cpu().numpy(), ref, rtol=0, atol=1e-13)


def test_spmm_long_rows_and_strided_input(dev):
    """36 right-hand sides (one wave per row, lane = column) on rows of 0 to
    3000 entries, empty rows included: rows longer than one batch of eight
    gathers take several batches; the input block is a column slice of a
    wider array (row stride 40)."""
    rp, ci, v = irregular(3000, seed=11)
    n = len(rp) - 1
    rng = np.random.default_rng(36)
    xw = rng.uniform(-1, 1, (n, 40))
    xd = to_dev(xw, dev)[:, 2:38]
    x = xw[:, 2:38].copy()
    A = D().DeviceCsr(rp, ci, v, device=dev)
    y = A.spmv(xd).cpu().numpy()
    check_spmv(rp, ci, v, x, y, np.float64)
    c0 = rng.uniform(-1, 1, (n, 36))
    c = to_dev(c0, dev)
    A.advanced_spmv(1.5, xd, -1.0, c)
    ref = ora.csr_advanced_spmv(1.5, rp, ci, v, x, -1.0, c0)
    tol = 2 * spmv_bound(rp, ci, v, x, np.float64) * 1.5 + 4 * np.finfo(np.float64).eps * np.abs(c0)
    assert np.all(np.abs(c.cpu().numpy() - ref) <= tol)


def test_spmv_rectangular_and_strided(dev):
    rng = np.random.default_rng(8)
    d = rng.uniform(-1, 1, (300, 170)) * (rng.random((300, 170)) < 0.05)
    rp, ci, v = gen.dense_to_csr(d)
    A = D().DeviceCsr(rp, ci, v, n_cols=170, device=dev)
    xb = to_dev(rng.uniform(-1, 1, (170, 5)), dev)
    x = xb[:, 1:2]  # column view with stride 5
    out = torch.zeros((300, 3), dtype=torch.float64, device=dev)
    A.spmv(x, out[:, 2:3])
    np.testing.assert_allclose(out[:, 2].cpu().numpy(), d @ x.cpu().numpy()[:, 0], atol=1e-13)
    assert not out[:, :2].any()


# --------------------------------------------------------- SELL-P / ELL ----

@pytest.mark.parametrize("slice_size,stride_factor", [(64, 1), (2, 2), (8, 3), (32, 4)])
@pytest.mark.parametrize("dtype", DTYPES)
def test_sellp_conversion_bit_exact(dev, dtype, slice_size, stride_factor):
    rp, ci, v = irregular(2100, seed=slice_size + stride_factor)
    v = v.astype(dtype)
    A = D().DeviceCsr(rp, ci, v, device=dev)
    S = D().DeviceSellp(A, slice_size, stride_factor)
    sets, lens, cols, vals = ora.csr_to_sellp(rp, ci, v, slice_size, stride_factor)
    np.testing.assert_array_equal(S.slice_sets.cpu().numpy(), sets)
    np.testing.assert_array_equal(S.slice_lengths.cpu().numpy()[: len(lens)], lens)
    np.testing.assert_array_equal(S.col_idxs.cpu().numpy()[: S.total], cols)
    np.testing.assert_array_equal(S.values.cpu().numpy()[: S.total], vals)
    x = np.random.default_rng(0).uniform(-1, 1, (2100, 2)).astype(dtype)
    y = S.spmv(to_dev(x, dev)).cpu().numpy()
    check_spmv(rp, ci, v, x, y, dtype)


def test_sellp_kat_layout(dev):
    k = KATS["sellp_layout"]
    m = KATS["matrices"][k["matrix"]]
    A = D().DeviceCsr(m["row_ptrs"], m["col_idxs"], np.array(m["values"]), m["n_cols"],
                      device=dev)
    S = D().DeviceSellp(A, 64, 1)
    assert S.slice_sets.cpu().tolist() == k["slice_sets"]
    cols = S.col_idxs.cpu().numpy()
    vals = S.values.cpu().numpy()
    for p, c in k["cols_at"].items():
        assert cols[int(p)] == c
    for p, c in k["vals_at"].items():
        assert vals[int(p)] == c


@pytest.mark.parametrize("dtype", DTYPES)
def test_ell_conversion_bit_exact(dev, dtype):
    rp, ci, v = gen.poisson2d(37)
    v = v.astype(dtype)
    A = D().DeviceCsr(rp, ci, v, device=dev)
    E = D().DeviceEll(A)
    w, s, cols, vals = ora.csr_to_ell(rp, ci, v)
    assert (E.width, E.stride) == (w, s)
    np.testing.assert_array_equal(E.col_idxs.cpu().numpy(), cols)
    np.testing.assert_array_equal(E.values.cpu().numpy(), vals)
    x = np.random.default_rng(0).uniform(-1, 1, (37 * 37, 3)).astype(dtype)
    y = E.spmv(to_dev(x, dev)).cpu().numpy()
    check_spmv(rp, ci, v, x, y, dtype)
    k = KATS["ell_layout"]
    m = KATS["matrices"][k["matrix"]]
    A2 = D().DeviceCsr(m["row_ptrs"], m["col_idxs"], np.array(m["values"]), 3, device=dev)
    E2 = D().DeviceEll(A2)
    assert E2.col_idxs.cpu().tolist() == k["cols"]
    assert E2.values.cpu().tolist() == k["vals"]


# ------------------------------------------------------------ transpose ----

@pytest.mark.parametrize("maker", ["powerlaw", "dense_column", "rect", "kat"])
def test_transpose_bit_exact(dev, maker):
    if maker == "powerlaw":
        rp, ci, v = gen.powerlaw(20000, mean=9, seed=1)
        nr = nc = 20000
    elif maker == "dense_column":
        rng = np.random.default_rng(3)
        d = (rng.random((500, 400)) < 0.01) * rng.uniform(-1, 1, (500, 400))
        d[:, 7] = rng.uniform(1, 2, 500)    # one column with 500 entries (wave path)
        d[:40, 11] = 3.0                    # one with 40 entries
        rp, ci, v = gen.dense_to_csr(d)
        nr, nc = 500, 400
    elif maker == "rect":
        rng = np.random.default_rng(4)
        d = (rng.random((50, 3000)) < 0.02) * rng.uniform(-1, 1, (50, 3000))
        rp, ci, v = gen.dense_to_csr(d)
        nr, nc = 50, 3000
    else:
        m = KATS["matrices"]["mtx_2x3"]
        rp, ci, v = (np.array(m["row_ptrs"], np.int32), np.array(m["col_idxs"], np.int32),
                     np.array(m["values"]))
        nr, nc = 2, 3
    A = D().DeviceCsr(rp, ci, v, n_cols=nc, device=dev)
    t_rp, t_ci, t_v = A.transpose()
    o_rp, o_ci, o_v = ora.csr_transpose(nr, nc, rp, ci, v)
    np.testing.assert_array_equal(t_rp.cpu().numpy(), o_rp)
    np.testing.assert_array_equal(t_ci.cpu().numpy(), o_ci)
    np.testing.assert_array_equal(t_v.cpu().numpy(), o_v)


# ----------------------------------------------------------- reductions ----

@pytest.mark.parametrize("nrhs", [1, 5, 64, 100])
@pytest.mark.parametrize("dtype", DTYPES)
def test_dot_and_norm(dev, dtype, nrhs):
    rng = np.random.default_rng(nrhs)
    n = 100_003
    x = rng.uniform(-1, 1, (n, nrhs)).astype(dtype)
    y = rng.uniform(-1, 1, (n, nrhs)).astype(dtype)
    d = D().dot(to_dev(x, dev), to_dev(y, dev)).cpu().numpy().astype(np.float64)
    exact = np.sum(x.astype(np.float64) * y.astype(np.float64), axis=0)
    bound = 2 * n * np.finfo(dtype).eps * np.sum(np.abs(x.astype(np.float64) * y), axis=0)
    assert np.all(np.abs(d - exact) <= bound)
    nr = D().norm2(to_dev(x, dev)).cpu().numpy()
    # sum of n squares: relative error of the squared norm <= n eps, halved by the sqrt
    np.testing.assert_allclose(nr, np.linalg.norm(x.astype(np.float64), axis=0),
                               rtol=n * np.finfo(dtype).eps)


def test_dot_exact_integers_all_sizes(dev):
    # integer-valued data: every partial sum is exact, so the result is exact
    for n in (1, 2, 63, 64, 65, 255, 256, 257, 1000, 4097, 70000):
        x = torch.arange(1, n + 1, dtype=torch.float64, device=dev)
        got = D().dot(x, torch.ones_like(x)).item()
        assert got == n * (n + 1) / 2, n


# --------------------------------------------------------------- Jacobi ----

@pytest.mark.parametrize("case", KATS["jacobi"], ids=lambda c: c["name"])
def test_jacobi_kats(dev, case):
    from ginkgo_amd import _capi as C
    m = KATS["matrices"][case["matrix"]]
    A = D().DeviceCsr(m["row_ptrs"], m["col_idxs"], np.array(m["values"]), device=dev)
    inv = A.jacobi()
    b = to_dev(np.array(case["b"]), dev)
    n, k = b.shape
    if "alpha" in case:
        x = to_dev(np.array(case["x0"]), dev)
        a = D().scalar(case["alpha"], torch.float64, dev)
        bt = D().scalar(case["beta"], torch.float64, dev)
        C.call("gkomi_jacobi_advanced_apply_f64", ctypes.byref(D().ctx(dev)), n, k, D().ptr(inv),
               D().ptr(a), D().ptr(b), k, D().ptr(bt), D().ptr(x), k)
    else:
        x = torch.empty_like(b)
        C.call("gkomi_jacobi_apply_f64", ctypes.byref(D().ctx(dev)), n, k, D().ptr(inv),
               D().ptr(b), k, D().ptr(x), k)
    np.testing.assert_allclose(x.cpu().numpy(), np.array(case["expected"]), rtol=0, atol=1e-15)


# ---------------------------------------------------- CG / BiCGSTAB KATs ----

def _full(v, dev, dt):
    return torch.full((2, 2), float(v), dtype=dt, device=dev)


def _vec(v, dev, dt):
    return torch.tensor(v, dtype=dt, device=dev)


@pytest.mark.parametrize("dtype", DTYPES)
def test_cg_and_bicgstab_step_kats(dev, dtype):
    from ginkgo_amd import _capi as C
    dt = TD[dtype]
    s = "f64" if dtype == np.float64 else "f32"
    c = D().ctx(dev)
    keep = []  # every tensor handed to the C ABI stays alive until the call returns

    def P(t):
        keep.append(t)
        return D().ptr(t)

    def st(v):
        return torch.tensor(v, dtype=torch.uint8, device=dev)

    k = KATS["cg_kernels"]["step_1"]
    p, z = _full(k["p_fill"], dev, dt), _full(k["z_fill"], dev, dt)
    C.call(f"gkomi_cg_step_1_{s}", ctypes.byref(c), 2, 2, P(p), P(z), 2, P(_vec(k["rho"], dev, dt)),
           P(_vec(k["prev_rho"], dev, dt)), P(st(k["status"])))
    np.testing.assert_array_equal(p.cpu().numpy(), k["expect_p"])
    k = KATS["cg_kernels"]["step_2"]
    x, r = _full(k["x_fill"], dev, dt), _full(k["r_fill"], dev, dt)
    C.call(f"gkomi_cg_step_2_{s}", ctypes.byref(c), 2, 2, P(x), 2, P(r),
           P(_full(k["p_fill"], dev, dt)), P(_full(k["q_fill"], dev, dt)), 2,
           P(_vec(k["beta"], dev, dt)), P(_vec(k["rho"], dev, dt)), P(st(k["status"])))
    np.testing.assert_array_equal(x.cpu().numpy(), k["expect_x"])
    np.testing.assert_array_equal(r.cpu().numpy(), k["expect_r"])
    kb = KATS["bicgstab_kernels"]
    for name in ("step_1", "step_1_div_rho_zero", "step_1_div_omega_zero"):
        k = kb[name]
        p = _full(k["p_fill"], dev, dt)
        C.call(f"gkomi_bicgstab_step_1_{s}", ctypes.byref(c), 2, 2, P(_full(k["r_fill"], dev, dt)),
               P(p), P(_full(k["v_fill"], dev, dt)), 2, P(_vec(k["rho"], dev, dt)),
               P(_vec(k["prev_rho"], dev, dt)), P(_vec(k["alpha"], dev, dt)),
               P(_vec(k["omega"], dev, dt)), P(st(k["status"])))
        np.testing.assert_array_equal(p.cpu().numpy(), k["expect_p"], err_msg=name)
    for name in ("step_2", "step_2_div_zero"):
        k = kb[name]
        sv = _full(k["s_fill"], dev, dt)
        al = _vec(k["alpha"], dev, dt)
        C.call(f"gkomi_bicgstab_step_2_{s}", ctypes.byref(c), 2, 2, P(_full(k["r_fill"], dev, dt)),
               P(sv), P(_full(k["v_fill"], dev, dt)), 2, P(_vec(k["rho"], dev, dt)), P(al),
               P(_vec(k["beta"], dev, dt)), P(st(k["status"])))
        np.testing.assert_array_equal(sv.cpu().numpy(), k["expect_s"], err_msg=name)
        np.testing.assert_array_equal(al.cpu().numpy(), k["expect_alpha"], err_msg=name)
    for name in ("step_3", "step_3_div_zero"):
        k = kb[name]
        x, r = _full(k["x_fill"], dev, dt), _full(k["r_fill"], dev, dt)
        om = _vec(k["omega"], dev, dt)
        C.call(f"gkomi_bicgstab_step_3_{s}", ctypes.byref(c), 2, 2, P(x), 2, P(r),
               P(_full(k["s_fill"], dev, dt)), P(_full(k["t_fill"], dev, dt)),
               P(_full(k["y_fill"], dev, dt)), P(_full(k["z_fill"], dev, dt)), 2,
               P(_vec(k["alpha"], dev, dt)), P(_vec(k["beta"], dev, dt)),
               P(_vec(k["gamma"], dev, dt)), P(om), P(st(k["status"])))
        np.testing.assert_array_equal(x.cpu().numpy(), k["expect_x"], err_msg=name)
        np.testing.assert_array_equal(om.cpu().numpy(), k["expect_omega"], err_msg=name)
    k = kb["finalize"]
    x = _full(k["x_fill"], dev, dt)
    status = st(k["status"])
    C.call(f"gkomi_bicgstab_finalize_{s}", ctypes.byref(c), 2, 2, P(x), 2,
           P(_full(k["y_fill"], dev, dt)), 2, P(_vec(k["alpha"], dev, dt)), P(status))
    np.testing.assert_array_equal(x.cpu().numpy(), k["expect_x"])
    assert status.cpu().tolist() == k["expect_status"]


def test_criteria_check_status_bits(dev):
    from ginkgo_amd import _capi as C
    crit = C.criteria([{"kind": "iteration", "max_iters": 5},
                       {"kind": "residual_norm", "reduction_factor": 0.1}])
    tau = torch.tensor([0.05, 0.5], dtype=torch.float64, device=dev)
    orig = torch.tensor([0, 0, 1.0, 1.0], dtype=torch.float64, device=dev)
    status = torch.zeros(2, dtype=torch.uint8, device=dev)
    flags = torch.zeros(2, dtype=torch.int64, device=dev)
    flags[0] = 3
    c = D().ctx(dev)
    P = D().ptr
    C.call("gkomi_criteria_check_f64", ctypes.byref(c), ctypes.byref(crit), 2, P(tau), None,
           P(orig), 1, P(status), P(flags))
    # column 0 converged by criterion 2 (finalized), column 1 still running
    assert status.cpu().tolist() == [0x80 | 0x40 | 2, 0]
    assert int(flags[1].item()) & 0xffffffff == 0
    flags[0] = 5
    C.call("gkomi_criteria_check_f64", ctypes.byref(c), ctypes.byref(crit), 2, P(tau), None,
           P(orig), 0, P(status), P(flags))
    # the iteration limit stops column 1 with id 1, not finalized, not converged
    assert status.cpu().tolist() == [0x80 | 0x40 | 2, 1]
    assert int(flags[1].item()) & 0xffffffff == 1


# --------------------------------------------------------------- fused CG ----

@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 1e-5)])
@pytest.mark.parametrize("jacobi", [True, False])
def test_fused_cg_poisson(dev, dtype, tol, jacobi):
    rp, ci, v = gen.poisson2d(64)
    v = v.astype(dtype)
    n = 64 * 64
    rng = np.random.default_rng(1)
    xt = rng.uniform(-1, 1, n)
    b = ora.csr_spmv(rp, ci, v, xt.astype(dtype).reshape(-1, 1))[:, 0]
    crit = [{"kind": "iteration", "max_iters": 2000},
            {"kind": "residual_norm", "reduction_factor": tol}]
    inv = ora.jacobi_inverse_diagonal(rp, ci, v) if jacobi else None
    xo, so, io = ora.cg_solve(rp, ci, v, b.reshape(-1, 1), np.zeros((n, 1), dtype), crit, inv)
    A = D().DeviceCsr(rp, ci, v, device=dev)
    solver = D().FusedCg(A, crit, jacobi=jacobi, chunk=8)
    x = torch.zeros(n, dtype=TD[dtype], device=dev)
    bt = to_dev(b, dev)
    iters = solver.apply(bt, x)
    # same recurrence, different summation order: a few iterations apart in
    # FP64, up to 10% in FP32 where rounding differences grow faster
    band = 3 if dtype == np.float64 else max(3, int(0.1 * io.iterations))
    assert abs(iters - io.iterations) <= band, (iters, io.iterations)
    assert solver.status.item() == so[0]
    xh = x.cpu().numpy().astype(np.float64)
    dense = gen.csr_to_dense(rp, ci, v.astype(np.float64))
    # the recursive residual met the criterion; the true residual is close to it
    res = np.linalg.norm(b - dense @ xh) / np.linalg.norm(b)
    assert res <= tol * (2 if dtype == np.float64 else 20)
    # forward error against the exact solution, bounded by kappa * residual
    kappa = 8.0 / (4 - 4 * np.cos(np.pi / 65))
    err = np.linalg.norm(xh - xt) / np.linalg.norm(xt)
    assert err <= 2 * kappa * max(res, np.finfo(dtype).eps)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 1e-5)])
@pytest.mark.parametrize("jacobi", [True, False])
def test_fused_cg_split_form_27pt(dev, dtype, tol, jacobi):
    """Rows of 27 entries take the split form of kernel 1 (p update, then
    q = A p with one gather per entry): same recurrence as the oracle."""
    rp, ci, v = gen.diffusion27(16)
    v = v.astype(dtype)
    n = 16 ** 3
    b = np.random.default_rng(3).uniform(-1, 1, n).astype(dtype)
    crit = [{"kind": "iteration", "max_iters": 3000},
            {"kind": "residual_norm", "reduction_factor": tol}]
    inv = ora.jacobi_inverse_diagonal(rp, ci, v) if jacobi else None
    xo, so, io = ora.cg_solve(rp, ci, v, b.reshape(-1, 1), np.zeros((n, 1), dtype), crit, inv)
    A = D().DeviceCsr(rp, ci, v, strategy="classical", device=dev)
    solver = D().FusedCg(A, crit, jacobi=jacobi, chunk=8)
    x = torch.zeros(n, dtype=TD[dtype], device=dev)
    iters = solver.apply(to_dev(b, dev), x)
    band = 3 if dtype == np.float64 else max(3, int(0.1 * io.iterations))
    assert abs(iters - io.iterations) <= band, (iters, io.iterations)
    assert solver.status.item() == so[0]
    xh = x.cpu().numpy().astype(np.float64)
    dense = gen.csr_to_dense(rp, ci, v.astype(np.float64))
    res = np.linalg.norm(b - dense @ xh) / np.linalg.norm(b)
    assert res <= tol * (2 if dtype == np.float64 else 20)
    if dtype == np.float64:
        np.testing.assert_allclose(xh, xo[:, 0], rtol=0, atol=1e-8 * np.abs(xo).max())


@pytest.mark.parametrize("maker", ["poisson", "d27"])
def test_fused_cg_compact_offsets_bit_identical(dev, maker):
    """The fused CG over the 2-byte column offsets (both kernel-1 forms)
    reproduces the 4-byte-index run bit for bit."""
    rp, ci, v = gen.poisson2d(48) if maker == "poisson" else gen.diffusion27(14)
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
        it = D().FusedCg(A, crit, jacobi=True, chunk=8).apply(b, x)
        out.append((it, x.cpu().numpy()))
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_fused_cg_iteration_limit_and_restart(dev):
    rp, ci, v = gen.poisson2d(32)
    n = 1024
    b = np.ones(n)
    crit = [{"kind": "iteration", "max_iters": 7},
            {"kind": "residual_norm", "reduction_factor": 1e-14}]
    xo, so, io = ora.cg_solve(rp, ci, v, b.reshape(-1, 1), np.zeros((n, 1)), crit)
    A = D().DeviceCsr(rp, ci, v, device=dev)
    solver = D().FusedCg(A, crit, chunk=16)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    assert solver.apply(to_dev(b, dev), x) == 7 == io.iterations
    assert solver.status.item() == so[0] == (1 | 0x40)
    np.testing.assert_allclose(x.cpu().numpy(), xo[:, 0], rtol=1e-10, atol=1e-12)
    # a second apply reuses the workspace (tickets were reset by the kernels)
    x2 = torch.zeros(n, dtype=torch.float64, device=dev)
    assert solver.apply(to_dev(b, dev), x2) == 7
    np.testing.assert_array_equal(x.cpu().numpy(), x2.cpu().numpy())
