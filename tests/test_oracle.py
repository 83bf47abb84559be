"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each oracle function is checked against (a) the known-answer tests
transcribed from the reference's own test-suite (tests/golden/reference_kats.json)
and (b) what the mathematics fixes: brute-force dense products and solves,
closed forms on stencils, invariants of the layouts.  The checks are chosen
so that a dropped term, a wrong index or a transposed operand fails.
"""
import json
import os

import numpy as np
import pytest

import oracle as ora
from ginkgo_amd import generators as gen

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")
KATS = json.load(open(GOLDEN))
DTYPES = [np.float64, np.float32]


def r_tol(dtype, spec):
    eps = np.finfo(dtype).eps
    if spec == 0:
        return 0.0
    if spec == "r":
        return 10 * eps
    if isinstance(spec, str) and spec.startswith("r*"):
        return 10 * eps * float(spec[2:])
    return float(spec)


def rel_frob(a, b):
    """GKO_ASSERT_MTX_NEAR metric (core/test/utils/assertions.hpp:290-306)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.sum(a * a), np.sum(b * b))
    if den == 0:
        return 0.0
    return float(np.sqrt(np.sum((a - b) ** 2) / den))


def mat(name, dtype):
    m = KATS["matrices"][name]
    if "dense" in m:
        return gen.dense_to_csr(np.array(m["dense"], dtype=dtype))
    return (np.array(m["row_ptrs"], np.int32), np.array(m["col_idxs"], np.int32),
            np.array(m["values"], dtype))


def arr(x, dtype):
    return np.array([[float(v) for v in row] for row in x], dtype=dtype)


# ------------------------------------------------------------------ SpMV ----

@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", KATS["spmv"], ids=lambda c: c["name"])
def test_spmv_kats(case, dtype):
    rp, ci, v = mat(case["matrix"], dtype)
    x = arr(case["x"], dtype)
    if "alpha" in case:
        y = ora.csr_advanced_spmv(case["alpha"], rp, ci, v, x, case["beta"], arr(case["y"], dtype))
    else:
        y = ora.csr_spmv(rp, ci, v, x)
    np.testing.assert_array_equal(y, arr(case["expected"], dtype))


@pytest.mark.parametrize("dtype", DTYPES)
def test_spmv_matches_dense_product(dtype):
    rng = np.random.default_rng(42)
    d = rng.uniform(-1, 1, (57, 41)) * (rng.random((57, 41)) < 0.2)
    rp, ci, v = gen.dense_to_csr(d.astype(dtype))
    x = rng.uniform(-1, 1, (41, 3)).astype(dtype)
    y = ora.csr_spmv(rp, ci, v, x)
    ref = d.astype(np.float64) @ x.astype(np.float64)
    bound = 4 * 41 * np.finfo(dtype).eps * (np.abs(d) @ np.abs(x.astype(np.float64))) + 1e-300
    assert np.all(np.abs(y - ref) <= bound)
    # transposed operand must not pass
    y_wrong = ora.csr_spmv(*gen.dense_to_csr(d.T[:41, :41].astype(dtype)), x[:41])
    assert y_wrong.shape != y.shape or not np.allclose(y_wrong, y)


def test_spmv_omp_equals_sequential():
    rp, ci, v = gen.powerlaw(4000, seed=11)
    x = np.random.default_rng(1).uniform(-1, 1, (4000, 1))
    np.testing.assert_array_equal(ora.csr_spmv(rp, ci, v, x), ora.csr_spmv(rp, ci, v, x, omp=True))


def test_laplacian_closed_form():
    # 7-pt Laplacian applied to the ones vector gives 6 - (number of neighbours)
    rp, ci, v = gen.laplace3d(6)
    y = ora.csr_spmv(rp, ci, v, np.ones((216, 1)))
    g = np.indices((6, 6, 6)).reshape(3, -1)
    nb = sum(((g[d] > 0).astype(int) + (g[d] < 5).astype(int)) for d in range(3))
    np.testing.assert_array_equal(y[:, 0], 6 - nb)


# ------------------------------------------------- load_balance partition ----

def test_lb_clac_size_known_values():
    # core/test/matrix/csr.cpp:127-131: load_balance(2) on the 2x3 matrix with
    # 4 entries has one srow element
    assert ora.lb_clac_size(4, 2) == 1
    # hand-computed: min(ceil(nnz / warp_size), nwarps * multiple)
    # HipExecutor form (1024 warps of 64, no cuda_strategy): multiples 8 / 16 / 64
    hip = dict(nwarps=1024, warp_size=64, cuda_strategy=False)
    assert ora.lb_clac_size(100_000, **hip) == 1563          # ceil(1e5/64)
    assert ora.lb_clac_size(999_999, **hip) == 8192          # 1024 * 8
    assert ora.lb_clac_size(1_000_000, **hip) == 15625       # ceil(1e6/64) < 1024*16
    assert ora.lb_clac_size(9_999_999, **hip) == 16384       # 1024 * 16
    assert ora.lb_clac_size(10_000_000, **hip) == 65536      # 1024 * 64
    assert ora.lb_clac_size(0, **hip) == 0
    # CUDA form (32-wide warps): multiples 8 / 32 / 128 / 512 / 2048
    assert ora.lb_clac_size(199_999, 2, 32) == 16            # 2 * 8
    assert ora.lb_clac_size(200_000, 2, 32) == 64            # 2 * 32
    assert ora.lb_clac_size(2_000_000, 2, 32) == 256         # 2 * 128
    assert ora.lb_clac_size(20_000_000, 2, 32) == 1024       # 2 * 512
    assert ora.lb_clac_size(200_000_000, 2, 32) == 4096      # 2 * 2048
    # the "intel" name: 8 / 32 / 256
    assert ora.lb_clac_size(20_000_000, 2, 32, intel=True) == 64
    assert ora.lb_clac_size(199_999, 2, 32, intel=True) == 16
    assert ora.lb_clac_size(100, 2, 0) == 0                  # warp size 0


def _lb_by_warp_start(rp, nwarps, ws):
    """srow from the reference GPU kernel's own warp ranges
    (common/cuda_hip/matrix/csr_kernels.template.cpp:199-205, 222-228): warp w
    starts at entry floor(w * ceil(nnz / ws) / nwarps) * ws and begins its
    walk at srow[w], the row that holds that entry (empty rows ending there
    skipped).  Brute force over every row, independent of process()."""
    nnz = int(rp[-1])
    lines = -(-nnz // ws)
    out = []
    for w in range(nwarps):
        start = (w * lines // nwarps) * ws
        out.append(next(i for i in range(len(rp) - 1) if rp[i + 1] > start))
    return np.array(out, np.int32)


@pytest.mark.parametrize("ws", [32, 64])
def test_lb_process_matches_warp_starts(ws):
    rng = np.random.default_rng(9)
    # empty rows at the start, inside and at the end, one very long row
    lens = rng.integers(0, 40, 700)
    lens[:5] = 0
    lens[100:130] = 0
    lens[-7:] = 0
    lens[300] = 5000
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    for nwarps in (1, 2, 7, 64, 333, ora.lb_clac_size(int(rp[-1]), 1024, ws, False)):
        s = ora.lb_process(rp, nwarps, ws)
        np.testing.assert_array_equal(s, _lb_by_warp_start(rp, nwarps, ws))
        assert np.all(np.diff(s) >= 0) and s[0] == 5 and s[-1] < 700


def test_lb_process_kat_and_dropped_term_fails():
    # core/test/matrix/csr.cpp:20-45, 127-131: 2x3 matrix, load_balance(2)
    np.testing.assert_array_equal(ora.lb_process(np.array([0, 3, 4], np.int32), 1), [0])
    rp = np.concatenate([[0], np.cumsum(np.arange(1, 200))]).astype(np.int32)
    s = ora.lb_process(rp, 50, 32)
    # a partition that forgot the ceil of the bucket (floor instead) differs
    nnz = int(rp[-1])
    div = -(-nnz // 32)
    floor_b = np.zeros(50, np.int64)
    for i in range(199):
        bkt = (-(-int(rp[i + 1]) // 32)) * 50 // div
        if bkt < 50:
            floor_b[bkt] += 1
    assert not np.array_equal(np.cumsum(floor_b), s)


# ---------------------------------------------------------- SELL-P / ELL ----

@pytest.mark.parametrize("dtype", DTYPES)
def test_sellp_layout_kat(dtype):
    k = KATS["sellp_layout"]
    rp, ci, v = mat(k["matrix"], dtype)
    sets, lens, cols, vals = ora.csr_to_sellp(rp, ci, v, k["slice_size"], k["stride_factor"])
    assert list(sets) == k["slice_sets"]
    assert list(lens) == k["slice_lengths"]
    for pos, c in k["cols_at"].items():
        assert cols[int(pos)] == c
    for pos, c in k["vals_at"].items():
        assert vals[int(pos)] == dtype(c)


@pytest.mark.parametrize("dtype", DTYPES)
def test_sellp_small_slices_kat(dtype):
    k = KATS["sellp_small_slices"]
    rp, ci, v = mat(k["matrix"], dtype)
    sets, lens, cols, vals = ora.csr_to_sellp(rp, ci, v, k["slice_size"], k["stride_factor"])
    y = ora.sellp_spmv(2, k["slice_size"], sets, lens, cols, vals, arr(k["x"], dtype))
    np.testing.assert_array_equal(y, arr(k["expected"], dtype))
    assert all(l % k["stride_factor"] == 0 for l in lens)


@pytest.mark.parametrize("slice_size,stride_factor", [(64, 1), (2, 2), (8, 3), (1, 1)])
def test_sellp_roundtrip_and_spmv(slice_size, stride_factor):
    rp, ci, v = gen.powerlaw(777, mean=6, seed=5)
    sets, lens, cols, vals = ora.csr_to_sellp(rp, ci, v, slice_size, stride_factor)
    # round trip: every CSR entry at its SELL-P slot, every other slot padding
    n = 777
    dense = np.zeros((n, n))
    for r in range(n):
        s, lr = divmod(r, slice_size)
        for k in range(lens[s]):
            idx = (sets[s] + k) * slice_size + lr
            if cols[idx] != -1:
                dense[r, cols[idx]] += vals[idx]
            else:
                assert vals[idx] == 0
    np.testing.assert_array_equal(dense, gen.csr_to_dense(rp, ci, v))
    x = np.random.default_rng(3).uniform(-1, 1, (n, 2))
    np.testing.assert_array_equal(ora.sellp_spmv(n, slice_size, sets, lens, cols, vals, x),
                                  ora.csr_spmv(rp, ci, v, x))
    # slice lengths are the padded maxima
    rl = np.diff(rp)
    for s in range(len(lens)):
        m = rl[s * slice_size:(s + 1) * slice_size].max()
        assert lens[s] == -(-m // stride_factor) * stride_factor


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", KATS["sellp_advanced"], ids=lambda c: c["name"])
def test_sellp_advanced_kats(case, dtype):
    rp, ci, v = mat(case["matrix"], dtype)
    sets, lens, cols, vals = ora.csr_to_sellp(rp, ci, v, case["slice_size"],
                                              case["stride_factor"])
    y = ora.sellp_advanced_spmv(case["alpha"], 2, case["slice_size"], sets, lens, cols, vals,
                                arr(case["x"], dtype), case["beta"], arr(case["y"], dtype))
    np.testing.assert_array_equal(y, arr(case["expected"], dtype))


@pytest.mark.parametrize("slice_size,stride_factor", [(64, 1), (8, 3), (1, 1)])
def test_sellp_advanced_and_omp_against_dense(slice_size, stride_factor):
    rng = np.random.default_rng(17)
    rp, ci, v = gen.powerlaw(901, mean=7, seed=21)
    sets, lens, cols, vals = ora.csr_to_sellp(rp, ci, v, slice_size, stride_factor)
    x = rng.uniform(-1, 1, (901, 3))
    c0 = rng.uniform(-1, 1, (901, 3))
    dense = gen.csr_to_dense(rp, ci, v)
    ref = -0.75 * (dense @ x) + 1.5 * c0
    ya = ora.sellp_advanced_spmv(-0.75, 901, slice_size, sets, lens, cols, vals, x, 1.5, c0)
    # within the bound of the k-term sums (k = the longest row), and never a
    # sign or operand swap: alpha and beta exchanged must fail
    k = int(np.diff(rp).max())
    bound = 4 * (k + 2) * np.finfo(np.float64).eps * (np.abs(dense) @ np.abs(x) * 0.75 +
                                                      1.5 * np.abs(c0))
    assert np.all(np.abs(ya - ref) <= bound)
    swapped = 1.5 * (dense @ x) - 0.75 * c0
    assert np.max(np.abs(ya - swapped)) > 1e-3
    # beta = 0 ignores c, NaN included
    nan_c = np.full((901, 3), np.nan)
    y0 = ora.sellp_advanced_spmv(2.0, 901, slice_size, sets, lens, cols, vals, x, 0.0, nan_c)
    np.testing.assert_array_equal(y0, 2.0 * ora.sellp_spmv(901, slice_size, sets, lens, cols,
                                                           vals, x))
    # the OpenMP form sums each row in the same order: bit-identical
    np.testing.assert_array_equal(ora.sellp_spmv(901, slice_size, sets, lens, cols, vals, x,
                                                 omp=True),
                                  ora.sellp_spmv(901, slice_size, sets, lens, cols, vals, x))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", KATS["ell_advanced"], ids=lambda c: c["name"])
def test_ell_advanced_kats(case, dtype):
    rp, ci, v = mat(case["matrix"], dtype)
    w, s, cols, vals = ora.csr_to_ell(rp, ci, v, stride=case["stride"])
    y = ora.ell_advanced_spmv(case["alpha"], 2, w, s, cols, vals, arr(case["x"], dtype),
                              case["beta"], arr(case["y"], dtype))
    np.testing.assert_array_equal(y, arr(case["expected"], dtype))


def test_mixed_apply_kats_through_the_widening_rule():
    """A mixed product computes in the highest precision of its operands
    (reference/matrix/csr_kernels.cpp:47-121): with float and double
    operands that is the double product of the exactly widened inputs,
    rounded to the output type once.  The reference's mixed KATs hold for
    every (matrix, input, output) combination this way."""
    k = KATS["mixed_apply"]
    for mdt in DTYPES:
        rp, ci, v = mat(k["matrix"], mdt)
        for bt, ct in k["combos"]:
            bdt = mdt if bt == "same" else (np.float32 if mdt == np.float64 else np.float64)
            cdt = mdt if ct == "same" else (np.float32 if mdt == np.float64 else np.float64)
            xb = arr(k["apply"]["x"], bdt)
            y = ora.csr_spmv(rp, ci, v.astype(np.float64), xb.astype(np.float64)).astype(cdt)
            np.testing.assert_array_equal(y, arr(k["apply"]["expected"], cdt))
            a = k["advanced"]
            ya = ora.csr_advanced_spmv(a["alpha"], rp, ci, v.astype(np.float64),
                                       arr(a["x"], bdt).astype(np.float64), a["beta"],
                                       arr(a["y"], cdt).astype(np.float64)).astype(cdt)
            np.testing.assert_array_equal(ya, arr(a["expected"], cdt))


@pytest.mark.parametrize("dtype", DTYPES)
def test_ell_layout_kat(dtype):
    k = KATS["ell_layout"]
    rp, ci, v = mat(k["matrix"], dtype)
    w, s, cols, vals = ora.csr_to_ell(rp, ci, v)
    assert (w, s) == (k["ell_width"], k["stride"])
    assert list(cols) == k["cols"]
    np.testing.assert_array_equal(vals, np.array(k["vals"], dtype))


def test_ell_spmv_equals_csr():
    rp, ci, v = gen.poisson2d(23)
    w, s, cols, vals = ora.csr_to_ell(rp, ci, v)
    x = np.random.default_rng(0).uniform(-1, 1, (529, 3))
    np.testing.assert_array_equal(ora.ell_spmv(529, w, s, cols, vals, x), ora.csr_spmv(rp, ci, v, x))
    c0 = np.random.default_rng(1).uniform(-1, 1, (529, 3))
    ya = ora.ell_advanced_spmv(-2.0, 529, w, s, cols, vals, x, 0.5, c0)
    ref = -2.0 * gen.csr_to_dense(rp, ci, v) @ x + 0.5 * c0
    np.testing.assert_allclose(ya, ref, rtol=0, atol=1e-13)


# ------------------------------------------------------------- transpose ----

@pytest.mark.parametrize("case", KATS["transpose"], ids=lambda c: c["name"])
def test_transpose_kats(case):
    m = KATS["matrices"][case["matrix"]]
    rp, ci, v = mat(case["matrix"], np.float64)
    nr, nc = (m.get("n_rows"), m.get("n_cols")) if "n_rows" in m else np.array(m["dense"]).shape
    t_rp, t_ci, t_v = ora.csr_transpose(nr, nc, rp, ci, v)
    np.testing.assert_array_equal(gen.csr_to_dense(t_rp, t_ci, t_v, nr),
                                  np.array(case["expected_dense"]))


def test_transpose_is_sorted_and_involutive():
    rp, ci, v = gen.powerlaw(3000, mean=9, seed=7)
    t = ora.csr_transpose(3000, 3000, rp, ci, v)
    # each transposed row sorted by column (original row order)
    for r in range(0, 3000, 97):
        seg = t[1][t[0][r]:t[0][r + 1]]
        assert np.all(np.diff(seg) > 0)
    tt = ora.csr_transpose(3000, 3000, *t)
    np.testing.assert_array_equal(tt[0], rp)
    np.testing.assert_array_equal(tt[1], ci)
    np.testing.assert_array_equal(tt[2], v)
    np.testing.assert_array_equal(gen.csr_to_dense(*t), gen.csr_to_dense(rp, ci, v).T)


# ------------------------------------------------------------ dot / norm ----

@pytest.mark.parametrize("dtype", DTYPES)
def test_dot_norm_higham(dtype):
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, (5000, 3)).astype(dtype)
    y = rng.uniform(-1, 1, (5000, 3)).astype(dtype)
    d = ora.dot(x, y)
    exact = np.sum(x.astype(np.longdouble) * y.astype(np.longdouble), axis=0)
    bound = 5000 * np.finfo(dtype).eps * np.sum(np.abs(x.astype(np.float64) * y), axis=0)
    assert np.all(np.abs(d.astype(np.longdouble) - exact) <= bound)
    nrm = ora.norm2(x)
    np.testing.assert_allclose(nrm, np.linalg.norm(x.astype(np.float64), axis=0),
                               rtol=100 * np.finfo(dtype).eps)


# --------------------------------------------------------------- Jacobi ----

@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", KATS["jacobi"], ids=lambda c: c["name"])
def test_jacobi_kats(case, dtype):
    rp, ci, v = mat(case["matrix"], dtype)
    inv = ora.jacobi_inverse_diagonal(rp, ci, v)
    b = arr(case["b"], dtype)
    z = ora.jacobi_apply(inv, b)
    if "alpha" in case:
        z = case["beta"] * arr(case["x0"], dtype) + case["alpha"] * z
    assert rel_frob(z, arr(case["expected"], dtype)) <= r_tol(dtype, case["tol"])


def test_jacobi_zero_diagonal_maps_to_one():
    rp = np.array([0, 1, 2], np.int32)
    ci = np.array([1, 1], np.int32)   # row 0 has no diagonal entry
    v = np.array([3.0, 4.0])
    inv = ora.jacobi_inverse_diagonal(rp, ci, v)
    np.testing.assert_array_equal(inv, [1.0, 0.25])


# ---------------------------------------------------------- CG kernels ----

def _stopped(s):
    return s


@pytest.mark.parametrize("dtype", DTYPES)
def test_cg_step_kats(dtype):
    import ctypes
    k = KATS["cg_kernels"]
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    n, m = 2, 2
    c = k["step_1"]
    p = np.full((n, m), c["p_fill"], dtype)
    z = np.full((n, m), c["z_fill"], dtype)
    st = np.array(c["status"], np.uint8)
    ora.cg_kernel("cg_step_1", dtype)(ctypes.c_int64(n), ctypes.c_int64(m), P(p), P(z),
                                      ctypes.c_int64(m), P(np.array(c["rho"], dtype)),
                                      P(np.array(c["prev_rho"], dtype)), P(st))
    np.testing.assert_array_equal(p, np.array(c["expect_p"], dtype))
    c = k["step_1_div_by_zero"]
    p = np.full((n, m), c["p_fill"], dtype)
    z = np.full((n, m), c["z_fill"], dtype)
    ora.cg_kernel("cg_step_1", dtype)(ctypes.c_int64(n), ctypes.c_int64(m), P(p), P(z),
                                      ctypes.c_int64(m), P(np.array(c["rho"], dtype)),
                                      P(np.array(c["prev_rho"], dtype)),
                                      P(np.array(c["status"], np.uint8)))
    np.testing.assert_array_equal(p, np.array(c["expect_p"], dtype))
    for name in ("step_2", "step_2_div_by_zero"):
        c = k[name]
        x = np.full((n, m), c["x_fill"], dtype)
        r = np.full((n, m), c["r_fill"], dtype)
        p = np.full((n, m), c["p_fill"], dtype)
        q = np.full((n, m), c["q_fill"], dtype)
        ora.cg_kernel("cg_step_2", dtype)(ctypes.c_int64(n), ctypes.c_int64(m), P(x),
                                          ctypes.c_int64(m), P(r), P(p), P(q), ctypes.c_int64(m),
                                          P(np.array(c["beta"], dtype)),
                                          P(np.array(c["rho"], dtype)),
                                          P(np.array(c["status"], np.uint8)))
        np.testing.assert_array_equal(x, np.array(c["expect_x"], dtype))
        np.testing.assert_array_equal(r, np.array(c["expect_r"], dtype))


@pytest.mark.parametrize("dtype", DTYPES)
def test_cg_initialize_kat(dtype):
    import ctypes
    c = KATS["cg_kernels"]["initialize"]
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    n = m = 2
    b = np.full((n, m), c["b_fill"], dtype)
    r, z, p, q = (np.full((n, m), v, dtype) for v in (0, 1, 1, 1))
    prev, rho = np.zeros(m, dtype), np.ones(m, dtype)
    st = np.array([1, 1], np.uint8)
    ora.cg_kernel("cg_initialize", dtype)(ctypes.c_int64(n), ctypes.c_int64(m), P(b),
                                          ctypes.c_int64(m), P(r), P(z), P(p), P(q),
                                          ctypes.c_int64(m), P(prev), P(rho), P(st))
    np.testing.assert_array_equal(r, b)
    assert not z.any() and not p.any() and not q.any()
    np.testing.assert_array_equal(rho, c["expect_rho"])
    np.testing.assert_array_equal(prev, c["expect_prev_rho"])
    np.testing.assert_array_equal(st, c["expect_status"])


@pytest.mark.parametrize("dtype", DTYPES)
def test_bicgstab_step_kats(dtype):
    import ctypes
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    I = ctypes.c_int64
    k = KATS["bicgstab_kernels"]
    n = m = 2
    full = lambda v: np.full((n, m), v, dtype)
    vec = lambda v: np.array(v, dtype)
    for name in ("step_1", "step_1_div_rho_zero", "step_1_div_omega_zero", "step_1_div_both_zero"):
        c = k[name]
        p = full(c["p_fill"])
        ora.cg_kernel("bicgstab_step_1", dtype)(I(n), I(m), P(full(c["r_fill"])), P(p),
                                                P(full(c["v_fill"])), I(m), P(vec(c["rho"])),
                                                P(vec(c["prev_rho"])), P(vec(c["alpha"])),
                                                P(vec(c["omega"])),
                                                P(np.array(c["status"], np.uint8)))
        np.testing.assert_array_equal(p, np.array(c["expect_p"], dtype), err_msg=name)
    for name in ("step_2", "step_2_div_zero"):
        c = k[name]
        s = full(c["s_fill"])
        alpha = vec(c["alpha"])
        ora.cg_kernel("bicgstab_step_2", dtype)(I(n), I(m), P(full(c["r_fill"])), P(s),
                                                P(full(c["v_fill"])), I(m), P(vec(c["rho"])),
                                                P(alpha), P(vec(c["beta"])),
                                                P(np.array(c["status"], np.uint8)))
        np.testing.assert_array_equal(s, np.array(c["expect_s"], dtype), err_msg=name)
        np.testing.assert_array_equal(alpha, np.array(c["expect_alpha"], dtype), err_msg=name)
    for name in ("step_3", "step_3_div_zero"):
        c = k[name]
        x, r = full(c["x_fill"]), full(c["r_fill"])
        omega = vec(c["omega"])
        ora.cg_kernel("bicgstab_step_3", dtype)(I(n), I(m), P(x), I(m), P(r),
                                                P(full(c["s_fill"])), P(full(c["t_fill"])),
                                                P(full(c["y_fill"])), P(full(c["z_fill"])), I(m),
                                                P(vec(c["alpha"])), P(vec(c["beta"])),
                                                P(vec(c["gamma"])), P(omega),
                                                P(np.array(c["status"], np.uint8)))
        np.testing.assert_array_equal(x, np.array(c["expect_x"], dtype), err_msg=name)
        np.testing.assert_array_equal(omega, np.array(c["expect_omega"], dtype), err_msg=name)
        if "expect_r" in c:
            np.testing.assert_array_equal(r, np.array(c["expect_r"], dtype), err_msg=name)
    c = k["finalize"]
    x = full(c["x_fill"])
    st = np.array(c["status"], np.uint8)
    ora.cg_kernel("bicgstab_finalize", dtype)(I(n), I(m), P(x), I(m), P(full(c["y_fill"])), I(m),
                                              P(vec(c["alpha"])), P(st))
    np.testing.assert_array_equal(x, np.array(c["expect_x"], dtype))
    np.testing.assert_array_equal(st, c["expect_status"])


# --------------------------------------------------------------- solves ----

def _crit(spec, dtype):
    out = []
    for c in spec:
        c = dict(c)
        if "reduction_factor" in c and isinstance(c["reduction_factor"], str):
            c["reduction_factor"] = r_tol(dtype, c["reduction_factor"])
        out.append(c)
    return out


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", KATS["solves"], ids=lambda c: c["name"])
def test_solver_kats(case, dtype):
    rp, ci, v = mat(case["matrix"], dtype)
    b = arr(case["b"], dtype)
    x0 = arr(case["x0"], dtype)
    crit = _crit(case["criteria"], dtype)
    if case["solver"] == "cg":
        x, st, info = ora.cg_solve(rp, ci, v, b, x0, crit)
    elif case["solver"] == "bicgstab":
        x, st, info = ora.bicgstab_solve(rp, ci, v, b, x0, crit)
    else:
        if dtype == np.float32:
            pytest.skip("solver::Ir oracle is the FP64 outer loop")
        x, st, info, _ = ora.ir_solve(rp, ci, v, b, x0, crit, inner_cg=False)
        x = x.reshape(-1, 1)
    if "alpha" in case:
        x = case["alpha"] * x + case["beta"] * x0
    assert rel_frob(x, arr(case["expected"], dtype)) <= r_tol(dtype, case["tol"])


def test_cg_against_direct_solve_and_criterion():
    rp, ci, v = gen.poisson2d(24)
    rng = np.random.default_rng(1)
    xt = rng.uniform(-1, 1, (576, 1))
    b = ora.csr_spmv(rp, ci, v, xt)
    inv = ora.jacobi_inverse_diagonal(rp, ci, v)
    crit = [{"kind": "iteration", "max_iters": 2000},
            {"kind": "residual_norm", "reduction_factor": 1e-10}]
    x, st, info = ora.cg_solve(rp, ci, v, b, np.zeros_like(b), crit, inv_diag=inv)
    assert st[0] & ora.STATUS_CONVERGED and (st[0] & ora.STATUS_ID_MASK) == 2
    res = np.linalg.norm(b - gen.csr_to_dense(rp, ci, v) @ x) / np.linalg.norm(b)
    assert res <= 1e-10
    np.testing.assert_allclose(x, np.linalg.solve(gen.csr_to_dense(rp, ci, v), b), atol=1e-8)
    # CG in exact arithmetic converges in <= n iterations; far fewer here
    assert 10 < info.iterations < 200
    # the OpenMP variant follows the same recurrence (different dot order)
    x2, st2, info2 = ora.cg_solve(rp, ci, v, b, np.zeros_like(b), crit, inv_diag=inv, omp=True)
    assert abs(info2.iterations - info.iterations) <= 2


def test_iteration_limit_sets_stop_not_converged():
    rp, ci, v = gen.poisson2d(16)
    b = np.ones((256, 1))
    x, st, info = ora.cg_solve(rp, ci, v, b, np.zeros_like(b),
                               [{"kind": "iteration", "max_iters": 3},
                                {"kind": "residual_norm", "reduction_factor": 1e-14}])
    assert info.iterations == 3
    assert st[0] == (1 | ora.STATUS_FINALIZED)


def test_bicgstab_nonsymmetric_against_direct():
    rng = np.random.default_rng(4)
    n = 60
    d = rng.uniform(-1, 1, (n, n)) * (rng.random((n, n)) < 0.1)
    d += np.diag(np.abs(d).sum(1) + 1)
    rp, ci, v = gen.dense_to_csr(d)
    b = rng.uniform(-1, 1, (n, 2))
    crit = [{"kind": "iteration", "max_iters": 500},
            {"kind": "residual_norm", "reduction_factor": 1e-12}]
    x, st, info = ora.bicgstab_solve(rp, ci, v, b, np.zeros_like(b), crit)
    np.testing.assert_allclose(x, np.linalg.solve(d, b), atol=1e-10)
    assert all(s & ora.STATUS_CONVERGED for s in st)


def test_ir_mixed_precision_reaches_fp64_accuracy():
    rp, ci, v = gen.diffusion27(8)
    n = 512
    rng = np.random.default_rng(2)
    b = rng.uniform(-1, 1, n)
    outer = [{"kind": "iteration", "max_iters": 100},
             {"kind": "residual_norm", "reduction_factor": 1e-12}]
    inner = [{"kind": "iteration", "max_iters": 1000},
             {"kind": "residual_norm", "reduction_factor": 1e-3}]
    x, st, info, inner_total = ora.ir_solve(rp, ci, v, b, np.zeros(n), outer, inner)
    d = gen.csr_to_dense(rp, ci, v)
    assert np.linalg.norm(b - d @ x) <= 1e-12 * np.linalg.norm(b)
    # the FP32 inner solver alone cannot reach 1e-12: several refinement sweeps
    assert info.iterations >= 3 and inner_total > info.iterations


def test_ir_inner_iterations_per_sweep():
    """Why mixed-precision IR needs more inner FP32 iterations than plain FP64
    CG needs in total: every sweep restarts CG with the residual as initial
    guess (core/solver/ir.cpp:242-250, examples/mixed-precision-ir/
    mixed-precision-ir.cpp:124-130), so the Krylov space is rebuilt and the
    first inner residual b - A b exceeds the baseline ||b||.  27-point 32^3
    diffusion with b = A x_true: the per-sweep counts add up to the total, the
    first sweep is short, and the inner total exceeds plain CG's count."""
    rp, ci, v = gen.diffusion27(32)
    n = len(rp) - 1
    xt = np.random.default_rng(4).uniform(-1, 1, (n, 1))
    b = ora.csr_spmv(rp, ci, v, xt, omp=True)
    cg_crit = [{"kind": "iteration", "max_iters": 20000},
               {"kind": "residual_norm", "reduction_factor": 1e-12}]
    _, st, info = ora.cg_solve(rp, ci, v, b, np.zeros((n, 1)), cg_crit, omp=True)
    outer = [{"kind": "iteration", "max_iters": 200},
             {"kind": "residual_norm", "reduction_factor": 1e-12}]
    inner = [{"kind": "iteration", "max_iters": 20000},
             {"kind": "residual_norm", "reduction_factor": 1e-3}]
    x, sti, ii, total, sweeps = ora.ir_solve(rp, ci, v, b[:, 0], np.zeros(n), outer, inner,
                                             omp=True, per_sweep=True)
    print(f"32^3: plain CG {info.iterations} iterations; IR {ii.iterations} sweeps, "
          f"inner {[int(s) for s in sweeps]} = {total}")
    assert sti[0] & ora.STATUS_CONVERGED and st[0] & ora.STATUS_CONVERGED
    assert len(sweeps) == ii.iterations and int(sweeps.sum()) == total
    assert sweeps[0] < min(sweeps[1:])
    assert total > info.iterations
    relres = np.linalg.norm(b[:, 0] - ora.csr_spmv(rp, ci, v, x.reshape(-1, 1))[:, 0])
    assert relres <= 1e-12 * np.linalg.norm(b)


def test_batch_bicgstab_reference_stencil_kat():
    k = KATS["batch_bicgstab"][0]
    row_off, rp, ci, v = gen.stencil3pt_batch(k["num_batch_items"], k["num_rows"])
    n = k["num_rows"]
    d = gen.csr_to_dense(rp[: n + 1], ci[:rp[n]], v[:rp[n]])
    xs = np.full(n, k["exact_sol_fill"])
    b = np.tile(d @ xs, k["num_batch_items"])
    x, iters, res = ora.batch_bicgstab_solve(row_off, rp, ci, v, b, np.zeros_like(b),
                                             use_jacobi=False, relative=True,
                                             tol=k["tolerance"], max_iters=k["max_iterations"])
    for s in range(k["num_batch_items"]):
        bs = b[s * n:(s + 1) * n]
        assert res[s] / np.linalg.norm(bs) <= 2 * k["tolerance"]
        assert np.max(np.abs(x[s * n:(s + 1) * n] - xs)) <= 10 * k["tolerance"] * 2.5
    k2 = KATS["batch_bicgstab"][1]
    x, iters, res = ora.batch_bicgstab_solve(row_off, rp, ci, v, b, np.zeros_like(b),
                                             use_jacobi=False, relative=True,
                                             tol=k2["tolerance"], max_iters=k2["max_iterations"])
    assert np.all(iters <= k2["check_iters_le"])


def test_batch_bicgstab_uneven_against_dense_lu():
    row_off, rp, ci, v, b = gen.uneven_batch(30, seed=12)
    x, iters, res = ora.batch_bicgstab_solve(row_off, rp, ci, v, b, np.zeros_like(b),
                                             use_jacobi=True, relative=True, tol=1e-10,
                                             max_iters=500)
    for s in range(30):
        r0, r1 = row_off[s], row_off[s + 1]
        sub_rp = rp[r0:r1 + 1] - rp[r0]
        d = gen.csr_to_dense(sub_rp, ci[rp[r0]:rp[r1]], v[rp[r0]:rp[r1]])
        np.testing.assert_allclose(x[r0:r1], np.linalg.solve(d, b[r0:r1]), atol=1e-8)
        assert np.linalg.norm(b[r0:r1] - d @ x[r0:r1]) <= 1e-10 * np.linalg.norm(b[r0:r1]) * 1.01


def _stencil_batch_system(k):
    row_off, rp, ci, v = gen.stencil3pt_batch(k["num_batch_items"], k["num_rows"],
                                               k["diag"], k["off"])
    n = k["num_rows"]
    d = gen.csr_to_dense(rp[: n + 1], ci[:rp[n]], v[:rp[n]])
    xs = np.full(n, k["exact_sol_fill"])
    b = np.tile(d @ xs, k["num_batch_items"])
    return row_off, rp, ci, v, d, xs, b


def test_batch_cg_reference_stencil_kats():
    """reference/test/solver/batch_cg_kernels.cpp:81-139 restated."""
    k = KATS["batch_cg"][0]
    row_off, rp, ci, v, d, xs, b = _stencil_batch_system(k)
    n = k["num_rows"]
    x, iters, res = ora.batch_cg_solve(row_off, rp, ci, v, b, np.zeros_like(b), use_jacobi=False,
                                       relative=True, tol=k["tolerance"],
                                       max_iters=k["max_iterations"])
    for s in range(k["num_batch_items"]):
        bs, xk = b[s * n:(s + 1) * n], x[s * n:(s + 1) * n]
        true_res = np.linalg.norm(bs - d @ xk)
        assert true_res / np.linalg.norm(bs) <= 6 * k["tolerance"]
        assert np.max(np.abs(xk - xs)) <= 10 * k["tolerance"]
        # StencilSystemLoggerLogsResidual: the logged norm meets the tolerance
        # and stays close to the true residual norm
        assert res[s] / np.linalg.norm(bs) <= k["tolerance"]
        assert abs(res[s] - true_res) <= 10 * k["tolerance"]
    k3 = KATS["batch_cg"][2]
    x, iters, res = ora.batch_cg_solve(row_off, rp, ci, v, b, np.zeros_like(b), use_jacobi=False,
                                       relative=True, tol=k3["tolerance"],
                                       max_iters=k3["max_iterations"])
    assert np.all(iters == k3["check_iters_eq"])


def test_batch_cg_spd_systems_against_dense_solve():
    """Uneven SPD systems: the solution of a dense solve, the logged norm
    equal to sqrt(r^T M r) of the final iterate's residual (the implicit norm
    with Jacobi), and at most n + 1 iterations -- in exact arithmetic CG
    ends within n steps, and the count starts at 1."""
    row_off, rp, ci, v, b = gen.uneven_spd_batch(25, n_min=16, n_max=48, seed=9)
    x, iters, res = ora.batch_cg_solve(row_off, rp, ci, v, b, np.zeros_like(b), use_jacobi=True,
                                       relative=True, tol=1e-12, max_iters=500)
    for s in range(25):
        r0, r1 = row_off[s], row_off[s + 1]
        d = gen.csr_to_dense(rp[r0:r1 + 1] - rp[r0], ci[rp[r0]:rp[r1]], v[rp[r0]:rp[r1]])
        exact = np.linalg.solve(d, b[r0:r1])
        np.testing.assert_allclose(x[r0:r1], exact, atol=1e-10 * np.abs(exact).max())
        r = b[r0:r1] - d @ x[r0:r1]
        implicit = np.sqrt(abs(r @ (r / np.diag(d))))
        assert res[s] <= 1e-12 * np.linalg.norm(b[r0:r1])
        assert abs(res[s] - implicit) <= 1e-12 * np.linalg.norm(b[r0:r1])
        assert 1 <= iters[s] <= (r1 - r0) + 1


def test_batch_cg_restart_and_zero_rhs():
    """A zero right-hand side stops at the first check (iteration 1, norm 0);
    an identity matrix converges after one step (x = b)."""
    n = 6
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.array([0, 1, 2, 0, 1, 2], dtype=np.int32)  # local to each system
    v = np.ones(n)
    row_off = np.array([0, 3, 6], dtype=np.int64)
    b = np.array([0.0, 0, 0, 1, 2, 3])
    x, iters, res = ora.batch_cg_solve(row_off, rp, ci, v, b, np.zeros(n), use_jacobi=False,
                                       relative=False, tol=1e-14, max_iters=10)
    assert iters[0] == 1 and res[0] == 0
    assert iters[1] == 2 and res[1] == 0
    np.testing.assert_array_equal(x, b)


def _kat_mat(a):
    return np.array(a, dtype=np.float64)


def test_coo_reference_kats():
    """reference/test/matrix/coo_kernels.cpp: apply, advanced apply (beta
    scales c, then c += alpha A b: coo_kernels.cpp:45-56), apply2 and the
    advanced apply2, on the sorted and the column-unsorted matrix."""
    k = KATS["coo"]
    for cols, vals in ((k["col_idxs"], k["values"]), (k["unsorted_col_idxs"], k["unsorted_values"])):
        rows = np.array(k["row_idxs"], np.int32)
        cols = np.array(cols, np.int32)
        vals = np.array(vals)
        for case in k["cases"]:
            x = _kat_mat(case["x"])
            if case["op"] == "apply":
                got = ora.coo_spmv2(rows, cols, vals, x, np.zeros((k["n_rows"], x.shape[1])))
            elif case["op"] == "advanced":
                c = case["beta"] * _kat_mat(case["y"])
                got = ora.coo_spmv2(rows, cols, vals, x, c, alpha=case["alpha"])
            else:
                got = ora.coo_spmv2(rows, cols, vals, x, _kat_mat(case["y"]),
                                    alpha=case.get("alpha"))
            np.testing.assert_array_equal(got, _kat_mat(case["expected"]), err_msg=case["name"])


def test_hybrid_reference_kats():
    """core/test/matrix/hybrid.cpp:168-240: the ELL width each strategy picks
    for row lengths [3, 1] and the layout of csr::convert_to_hybrid (the
    explicit zero entry is kept)."""
    h = KATS["hybrid"]
    rp = np.array(h["csr"]["row_ptrs"], np.int32)
    ci = np.array(h["csr"]["col_idxs"], np.int32)
    v = np.array(h["csr"]["values"])
    for lay in h["layouts"]:
        w = ora.hybrid_ell_width(np.diff(rp), lay["strategy"], lay.get("p0", 0.0))
        assert w == lay["ell_width"], lay["strategy"]
        ec, ev, cr, cc, cv = ora.csr_to_hybrid(rp, ci, v, w, lay["ell_stride"])
        assert ec.tolist() == lay["ell_col_idxs"] and ev.tolist() == lay["ell_values"]
        assert cr.tolist() == lay["coo_row_idxs"] and cc.tolist() == lay["coo_col_idxs"]
        assert cv.tolist() == lay["coo_values"]


def test_hybrid_parts_sum_to_the_product():
    """ELL part + COO part of any width reproduce A x (brute force), and the
    COO part holds exactly the entries beyond the width."""
    rng = np.random.default_rng(3)
    d = rng.uniform(-1, 1, (60, 50)) * (rng.random((60, 50)) < 0.15)
    d[7, :] = rng.uniform(-1, 1, 50)            # one full row
    rp, ci, v = gen.dense_to_csr(d)
    x = rng.uniform(-1, 1, 50)
    for w in (0, 1, 3, 8, 50):
        ec, ev, cr, cc, cv = ora.csr_to_hybrid(rp, ci, v, w, 60)
        y = np.zeros(60)
        for k in range(w):
            for r in range(60):
                if ec[r + k * 60] >= 0:
                    y[r] += ev[r + k * 60] * x[ec[r + k * 60]]
        y = ora.coo_spmv2(cr, cc, cv, x.reshape(-1, 1), y.reshape(-1, 1))[:, 0]
        np.testing.assert_allclose(y, d @ x, rtol=0, atol=1e-13)
        assert len(cv) == int(np.maximum(np.diff(rp) - w, 0).sum())
