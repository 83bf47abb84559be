"""Column-panel CSR product (gkomi_csr_build_panels_*, no reference
counterpart; it replaces the gathers of csr::spmv / advanced_spmv,
common/cuda_hip/matrix/csr_kernels.template.cpp:150-290, for matrices whose
input vector does not fit one XCD's L2).

* The panel build is integer work: its row pointers, column indices, copied
  values and chunk bounds are compared bit for bit with a numpy statement of
  the documented layout.
* Rows of at most 32 entries per panel are summed sequentially from LDS and
  the panel sums added in panel order, so the product is compared bit for
  bit with that order restated in Python floats (and numpy float32).
* Every product is checked against the oracle's FP64 product within
  4 k eps (|A||x|)_i, and a repeated product is bit-identical.
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
CHUNK = 1536  # rows plus entries per chunk of an evenly filled panel


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


def a256(x):
    return (x + 255) // 256 * 256


def layout(n, nnz, P, vb):
    C = max(1, -(-(n + -(-nnz // P)) // CHUNK))
    off = {}
    o = 0
    off["rp"] = o
    o += a256(4 * (P * n + 1))
    off["ci"] = o
    o += a256(4 * nnz)
    off["v"] = o
    o += a256(vb * nnz)
    off["bounds"] = o
    o += a256(4 * P * (C + 1))
    off["partial"] = o
    return C, off


def host_panels(rp, ci, v, n_cols, P):
    """The documented panel layout restated on the host."""
    n = len(rp) - 1
    W = max(1, -(-n_cols // P))
    rows = np.repeat(np.arange(n), np.diff(rp))
    pan = np.minimum(ci.astype(np.int64) // W, P - 1)
    counts = np.zeros((P, n), np.int64)
    np.add.at(counts, (pan, rows), 1)
    prp = np.zeros(P * n + 1, np.int64)
    np.cumsum(counts.reshape(-1), out=prp[1:])
    # stable order: by panel, then row, then original position
    order = np.lexsort((np.arange(len(ci)), rows, pan))
    nnz = len(ci)
    C = max(1, -(-(n + -(-nnz // P)) // CHUNK))
    bounds = np.zeros((P, C + 1), np.int64)
    for p in range(P):
        r = prp[p * n:p * n + n + 1]
        cum = np.arange(1, n + 1) + (r[1:] - r[0])
        total = n + (r[n] - r[0])
        for k in range(C + 1):
            bounds[p, k] = np.searchsorted(cum, k * total // C, side="right")
    return prp, ci[order], v[order], bounds, pan, W


def read_blob(A, dtype):
    n, nnz, P = A.n_rows, A.nnz, A.num_panels
    vb = np.dtype(dtype).itemsize
    C, off = layout(n, nnz, P, vb)
    raw = A.panel_buf.cpu().numpy()
    prp = raw[off["rp"]:off["rp"] + 4 * (P * n + 1)].view(np.int32)
    pci = raw[off["ci"]:off["ci"] + 4 * nnz].view(np.int32)
    pv = raw[off["v"]:off["v"] + vb * nnz].view(dtype)
    bnd = raw[off["bounds"]:off["bounds"] + 4 * P * (C + 1)].view(np.int32).reshape(P, C + 1)
    return prp, pci, pv, bnd


def sequential_panel_product(rp, ci, v, x, n_cols, P, dtype):
    """Row sums panel by panel in row order, panel sums added in panel order."""
    W = max(1, -(-n_cols // P))
    f = (lambda a: float(a)) if dtype == np.float64 else np.float32
    y = np.zeros(len(rp) - 1, dtype)
    for i in range(len(rp) - 1):
        parts = [f(0.0)] * P
        for j in range(rp[i], rp[i + 1]):
            p = min(int(ci[j]) // W, P - 1)
            parts[p] = f(parts[p] + f(f(v[j]) * f(x[ci[j]])))
        s = parts[0]
        for p in range(1, P):
            s = f(s + parts[p])
        y[i] = s
    return y


def bound_check(rp, ci, v, x, y, dtype, label):
    vw = v.astype(dtype).astype(np.float64)
    xw = x.astype(dtype).astype(np.float64)
    exact = ora.csr_spmv(rp, ci, vw, xw.reshape(-1, 1), omp=True)[:, 0]
    ax = ora.csr_spmv(rp, ci, np.abs(vw), np.abs(xw).reshape(-1, 1), omp=True)[:, 0]
    k = np.maximum(np.diff(rp), 1).astype(np.float64)
    bound = 4 * k * (EPS[dtype] + EPS[np.float64]) * ax
    err = np.abs(y.astype(np.float64) - exact)
    say(f"{label}: max |y - y_oracle| {err.max() if err.size else 0:.3e}")
    assert np.all(err <= bound), label


def rand_csr(rng, n, n_cols, lens, cols_hi=None):
    cols_hi = n_cols if cols_hi is None else cols_hi
    rp = np.zeros(n + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    ci = np.concatenate([np.sort(rng.choice(cols_hi, size=int(L), replace=False))
                         if L else np.zeros(0, np.int64) for L in lens]).astype(np.int32)
    v = rng.uniform(-1, 1, len(ci))
    return rp, ci, v


def cases():
    rng = np.random.default_rng(11)
    out = []
    # rectangular, empty rows, rows of up to 40 entries
    lens = rng.integers(0, 41, 3000)
    lens[rng.choice(3000, 300, replace=False)] = 0
    out.append(("ragged", 2500) + rand_csr(rng, 3000, 2500, lens))
    # rows beyond one 64-entry step of the build, long rows per panel (> 32)
    lens = rng.integers(0, 20, 2000)
    lens[::97] = 700
    out.append(("long_rows", 4000) + rand_csr(rng, 2000, 4000, lens))
    # one row of 5000 entries: its chunk does not fit LDS
    lens = rng.integers(1, 8, 1500)
    lens[700] = 5000
    out.append(("huge_row", 6000) + rand_csr(rng, 1500, 6000, lens))
    # fewer columns than panels: the last panels are empty
    lens = rng.integers(0, 4, 500)
    out.append(("narrow", 3) + rand_csr(rng, 500, 3, lens))
    # every entry in the first panel
    lens = rng.integers(0, 10, 4000)
    out.append(("one_panel", 8000) + rand_csr(rng, 4000, 8000, lens, cols_hi=900))
    return out


CASES = cases()


@pytest.mark.parametrize("P", [2, 4, 8])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_panel_build_and_product(dev, case, P, dtype):
    name, n_cols, rp, ci, v = case
    say(f"panels {name} P={P} {np.dtype(dtype).name}: n {len(rp) - 1} x {n_cols}, "
        f"nnz {rp[-1]}")
    vd = v.astype(dtype)
    A = D().DeviceCsr(rp, ci, vd, n_cols=n_cols, strategy="load_balance", device=dev)
    A.build_panels(P)
    prp, pci, pv, bnd = read_blob(A, dtype)
    hp, hci, hv, hb, pan, W = host_panels(rp, ci, vd, n_cols, P)
    np.testing.assert_array_equal(prp, hp)
    np.testing.assert_array_equal(pci, hci)
    np.testing.assert_array_equal(pv, hv)
    np.testing.assert_array_equal(bnd, hb)
    x = np.random.default_rng(5).uniform(-1, 1, n_cols).astype(dtype)
    xd = torch.as_tensor(x).to(dev).reshape(-1, 1)
    y = A.spmv(xd).cpu().numpy()[:, 0]
    bound_check(rp, ci, v, x, y, dtype, f"panels {name} P={P}")
    np.testing.assert_array_equal(A.spmv(xd).cpu().numpy()[:, 0], y)
    # the panel rows of at most 32 entries: the exact summation order
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    per = np.zeros((P, len(rp) - 1), np.int64)
    np.add.at(per, (pan, rows), 1)
    # (a chunk beyond the LDS capacity, as every chunk of the first panel in
    # one_panel with P = 8, takes several LDS passes in the same order)
    if per.max(initial=0) <= 32:
        ys = sequential_panel_product(rp, ci, vd, x, n_cols, P, dtype)
        np.testing.assert_array_equal(y, ys)


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_panel_advanced_and_strided(dev, dtype):
    rng = np.random.default_rng(12)
    lens = rng.integers(0, 30, 2048)
    rp, ci, v = rand_csr(rng, 2048, 3000, lens)
    vd = v.astype(dtype)
    A = D().DeviceCsr(rp, ci, vd, n_cols=3000, strategy="classical", device=dev)
    A.build_panels(4)
    x = rng.uniform(-1, 1, (3000, 3)).astype(dtype)
    xd = torch.as_tensor(x).to(dev)
    # strided single column
    y = A.spmv(xd[:, 1:2]).cpu().numpy()[:, 0]
    bound_check(rp, ci, v, x[:, 1], y, dtype, "panels strided b")
    # c = alpha A b + beta c, and beta = 0 ignoring NaN in c
    c0 = rng.uniform(-1, 1, (2048, 1)).astype(dtype)
    cd = torch.as_tensor(c0).to(dev)
    A.advanced_spmv(2.0, xd[:, 0:1], -0.5, cd)
    ax = ora.csr_spmv(rp, ci, vd.astype(np.float64), x[:, :1].astype(np.float64), omp=True)
    ab = ora.csr_spmv(rp, ci, np.abs(vd.astype(np.float64)),
                      np.abs(x[:, :1].astype(np.float64)), omp=True)
    k = np.maximum(np.diff(rp), 1)[:, None]
    ref = 2.0 * ax - 0.5 * c0.astype(np.float64)
    tol = 8 * k * EPS[dtype] * (2 * ab + 0.5 * np.abs(c0))
    assert np.all(np.abs(cd.cpu().numpy() - ref) <= tol)
    cn = torch.full((2048, 1), float("nan"), dtype=TD[dtype], device=dev)
    A.advanced_spmv(1.0, xd[:, 0:1], 0.0, cn)
    assert not torch.isnan(cn).any()
    # several right-hand sides take the row-wave kernel: unchanged results
    Y = A.spmv(xd).cpu().numpy()
    bound_check(rp, ci, v, x[:, 2], Y[:, 2], dtype, "panels multi-rhs column")


def test_panel_empty_and_rejected(dev):
    # no entries: the product is zero
    rp = np.zeros(101, np.int32)
    A = D().DeviceCsr(rp, np.zeros(0, np.int32), np.zeros(0), n_cols=50,
                      strategy="load_balance", device=dev)
    A.build_panels(8)
    y = A.spmv(torch.ones((50, 1), dtype=torch.float64, device=dev))
    assert torch.all(y == 0)
    from ginkgo_amd import _capi as C
    with pytest.raises(C.InvalidArgument):
        A.build_panels(3)
    assert int(C.lib().gkomi_csr_panels_bytes(10, 10, 5, C.VALUE_F64)) == -1
    assert int(C.lib().gkomi_csr_panels_bytes(10, 10, 8, 7)) == -1
    # removing the panels restores the plain product
    A.build_panels(0)
    assert A.num_panels == 0


@pytest.mark.parametrize("P", [4, 8])
def test_panel_powerlaw_1m(dev, P):
    rp, ci, v = gen.powerlaw(1_000_000, seed=3)
    n = len(rp) - 1
    x = np.random.default_rng(3).uniform(-1, 1, n)
    say(f"panels power-law 1M P={P}: nnz {rp[-1]}")
    A = D().DeviceCsr(rp, ci, v, strategy="load_balance", device=dev)
    A.build_panels(P)
    xd = torch.as_tensor(x).to(dev).reshape(-1, 1)
    y = A.spmv(xd).cpu().numpy()[:, 0]
    bound_check(rp, ci, v, x, y, np.float64, f"panels power-law P={P}")
    np.testing.assert_array_equal(A.spmv(xd).cpu().numpy()[:, 0], y)
    prp, pci, pv, bnd = read_blob(A, np.float64)
    hp, hci, hv, hb, _, _ = host_panels(rp, ci, v, n, P)
    np.testing.assert_array_equal(prp, hp)
    np.testing.assert_array_equal(bnd, hb)


def test_panels_through_the_cpp_api(dev):
    """Csr::build_column_panels: apply and advanced apply through the gko::
    surface equal the Python path's panel product bit for bit; a mutable
    accessor drops the panels (the plain product again)."""
    import ginkgo_amd.gko as gko
    ex = gko.Mi355xExecutor.create(0, torch.cuda.current_stream().cuda_stream)
    name, n_cols, rp, ci, v = CASES[0]
    n = len(rp) - 1
    x = np.random.default_rng(9).uniform(-1, 1, (n_cols, 1))
    A = gko.csr_from_host(ex, rp, ci, v, n_cols, "load_balance")
    b = gko.matrix.Dense("f64").from_numpy(ex, x)
    out = gko.matrix.Dense("f64").create(ex, n, 1)
    A.apply(b, out)
    plain = out.to_numpy()
    A.build_column_panels(4)
    assert A.num_column_panels == 4
    A.apply(b, out)
    y = out.to_numpy()
    P4 = D().DeviceCsr(rp, ci, v, n_cols=n_cols, strategy="load_balance", device=dev)
    P4.build_panels(4)
    ref = P4.spmv(torch.as_tensor(x).to(dev)).cpu().numpy()
    np.testing.assert_array_equal(y, ref)
    np.testing.assert_array_equal(y[:, 0], sequential_panel_product(rp, ci, v, x[:, 0], n_cols,
                                                                    4, np.float64))
    alpha = gko.matrix.Dense("f64").from_numpy(ex, np.array([[1.0]]))
    beta = gko.matrix.Dense("f64").from_numpy(ex, np.array([[0.0]]))
    cn = gko.matrix.Dense("f64").from_numpy(ex, np.full((n, 1), np.nan))
    A.apply(alpha, b, beta, cn)
    np.testing.assert_array_equal(cn.to_numpy(), ref)
    _ = A.values_data_ptr
    assert A.num_column_panels == 0
    A.apply(b, out)
    np.testing.assert_array_equal(out.to_numpy(), plain)
    with pytest.raises(Exception):
        A.build_column_panels(5)
