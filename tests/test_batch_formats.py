"""batch::matrix::Ell and Dense without a GPU: the oracle's batched products
against the reference's fixtures (reference/test/matrix/batch_{ell,dense}_
kernels.cpp) and the dense product, and the Ell layout built on the host
executor against core/test/matrix/batch_ell.cpp."""
import json
import os

import numpy as np
import pytest

import oracle as ora

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
KAT = GOLD["batch_formats"]


def fixture_cases():
    """The reference's fixture split into single right-hand sides: items of
    2 x 3, b with three columns, x with three columns."""
    a = KAT["apply"]
    mtx = np.array(a["mtx"])
    b = np.array(a["b"])
    x = np.array(a["x"])
    for col in range(3):
        yield mtx, b[:, :, col], x[:, :, col]


def ell_of(mtx):
    """Shared column-major Ell pattern of the items' union of nonzeros."""
    num, rows, cols = mtx.shape
    pattern = [[j for j in range(cols) if np.any(mtx[:, i, j] != 0)] for i in range(rows)]
    width = max(len(p) for p in pattern)
    ci = -np.ones(rows * width, np.int32)
    vals = np.zeros((num, rows * width))
    for i, p in enumerate(pattern):
        for s, j in enumerate(p):
            ci[i + s * rows] = j
            vals[:, i + s * rows] = mtx[:, i, j]
    return width, ci, vals.reshape(-1)


def csr_of(mtx):
    num, rows, cols = mtx.shape
    rp, ci = [0], []
    for i in range(rows):
        for j in range(cols):
            if np.any(mtx[:, i, j] != 0):
                ci.append(j)
        rp.append(len(ci))
    vals = np.concatenate([[mtx[k, i, j] for i in range(rows) for j in ci[rp[i]:rp[i + 1]]]
                           for k in range(num)])
    return np.array(rp, np.int32), np.array(ci, np.int32), vals


@pytest.mark.parametrize("fmt", ["csr", "ell", "dense"])
@pytest.mark.parametrize("advanced", [False, True])
def test_oracle_batch_apply_matches_dense_product(fmt, advanced):
    """Every value of the fixture is a small binary fraction, so the dense
    product is exact and the restatement must reproduce it bit for bit."""
    a = KAT["apply"]
    alpha = np.array(a["alpha"]) if advanced else None
    beta = np.array(a["beta"]) if advanced else None
    for mtx, b, x0 in fixture_cases():
        num, rows, cols = mtx.shape
        expect = np.einsum("kij,kj->ki", mtx, b)
        if advanced:
            expect = alpha[:, None] * expect + beta[:, None] * x0
        kw = dict(x=x0.reshape(-1), alpha=alpha, beta=beta) if advanced else {}
        if fmt == "dense":
            got = ora.batch_dense_spmv(num, rows, cols, mtx.reshape(-1), b.reshape(-1), **kw)
        elif fmt == "ell":
            width, ci, vals = ell_of(mtx)
            got = ora.batch_ell_spmv(num, rows, cols, width, ci, vals, b.reshape(-1), **kw)
        else:
            rp, ci, vals = csr_of(mtx)
            got = ora.batch_csr_spmv(num, rows, cols, rp, ci, vals, b.reshape(-1), **kw)
        np.testing.assert_array_equal(got.reshape(num, rows), expect)


def test_oracle_ell_row_ends_at_first_padding():
    """A row ends at its first -1: a value stored after it is not read."""
    ci = np.array([0, -1, 1, 1], np.int32)   # rows 2, width 2: row 0 = {0, 1}, row 1 = {}
    vals = np.array([2.0, 7.0, 3.0, 5.0])
    got = ora.batch_ell_spmv(1, 2, 2, 2, ci, vals, np.array([1.0, 10.0]))
    np.testing.assert_array_equal(got, [32.0, 0.0])


def test_oracle_advanced_beta_zero_propagates_nan():
    """beta * x is formed also for beta = 0 (unlike the unbatched Csr)."""
    got = ora.batch_dense_spmv(1, 1, 1, np.array([1.0]), np.array([1.0]),
                               x=np.array([np.nan]), alpha=np.array([1.0]),
                               beta=np.array([0.0]))
    assert np.isnan(got[0])


@pytest.mark.parametrize("case", KAT["ell_layout"], ids=["sparse", "full"])
def test_ell_layout_from_dense_items(case):
    """batch::matrix::Ell from per-item dense data on the host executor: the
    reference's expected column-major values and shared column indices."""
    import ginkgo_amd.gko as gko
    host = gko.HostExecutor.create()
    d = np.array(case["dense"])
    m = gko.batch.matrix.Ell("f64").from_dense(host, d)
    assert m.num_stored_elements_per_row == case["width"]
    np.testing.assert_array_equal(m.values(), case["values"])
    if "col_idxs" in case:
        np.testing.assert_array_equal(m.col_idxs(), case["col_idxs"])
    assert m.size == (2, 2, 3)


def test_ell_and_dense_reject_bad_input():
    import ginkgo_amd.gko as gko
    host = gko.HostExecutor.create()
    with pytest.raises(gko.DimensionMismatch):
        gko.batch.matrix.Ell("f64").from_numpy(host, 2, 2, 2, 1, np.array([0, 2], np.int32),
                                              np.ones(4))
    with pytest.raises(gko.DimensionMismatch):
        gko.batch.matrix.Ell("f64").from_numpy(host, 2, 2, 2, 1, np.array([0, 1], np.int32),
                                              np.ones(3))
    with pytest.raises(gko.DimensionMismatch):
        gko.batch.matrix.Dense("f64").from_numpy(host, np.ones((2, 3)))


def test_batch_csr_shapes_and_bad_input():
    """batch::matrix::Csr items of any shape (include/ginkgo/core/matrix/
    batch_csr.hpp: a batch_dim<2>); patterns that do not fit the item raise."""
    import ginkgo_amd.gko as gko
    host = gko.HostExecutor.create()
    rp = np.array([0, 2, 3], np.int32)
    ci = np.array([0, 4, 2], np.int32)
    m = gko.batch.matrix.Csr("f64").from_numpy(host, 3, 2, rp, ci, np.ones(9), cols=5)
    assert m.size == (3, 2, 5)
    with pytest.raises(gko.DimensionMismatch):  # column 4 outside a 2 x 2 item
        gko.batch.matrix.Csr("f64").from_numpy(host, 3, 2, rp, ci, np.ones(9))
    with pytest.raises(gko.DimensionMismatch):  # values of 2 items, 3 declared
        gko.batch.matrix.Csr("f64").from_numpy(host, 3, 2, rp, ci, np.ones(6), cols=5)
    with pytest.raises(gko.DimensionMismatch):  # decreasing row pointers
        gko.batch.matrix.Csr("f64").from_numpy(host, 1, 2, np.array([0, 3, 2], np.int32),
                                               np.array([0, 1], np.int32), np.ones(2), cols=5)


def block_diag_csr_of(blocks):
    """Block-diagonal CSR (global row pointers, local columns) of dense
    blocks of different sizes."""
    row_off, rp, ci, v = [0], [0], [], []
    for d in blocks:
        for i in range(d.shape[0]):
            cols = np.nonzero(d[i])[0]
            ci.extend(cols)
            v.extend(d[i, cols])
            rp.append(len(ci))
        row_off.append(row_off[-1] + d.shape[0])
    return (np.array(row_off, np.int64), np.array(rp, np.int32), np.array(ci, np.int32),
            np.array(v))


def uneven_blocks(seed=3, num=7):
    rng = np.random.default_rng(seed)
    blocks = []
    for k in range(num):
        n = int(rng.integers(1, 9))
        d = rng.uniform(-1, 1, (n, n)) * (rng.random((n, n)) < 0.4)
        d[np.arange(n), np.arange(n)] = 4.0 + k
        blocks.append(d)
    return blocks


def test_block_diag_ell_and_dense_layout():
    """The uneven Ell / Dense batches (extension: items of different sizes)
    built on the host executor: per item n_k x w_k column-major slots with
    the item's own local columns and -1 after a row's end, w_k the item's
    longest row; Dense n_k x n_k row-major; item offsets their prefix sums."""
    import ginkgo_amd.gko as gko
    host = gko.HostExecutor.create()
    blocks = uneven_blocks()
    row_off, rp, ci, v = block_diag_csr_of(blocks)
    E = gko.batch.matrix.BlockDiagEll("f64").from_numpy(host, row_off, rp, ci, v)
    widths = [max(int(np.count_nonzero(d[i])) for i in range(d.shape[0])) for d in blocks]
    offs = np.concatenate([[0], np.cumsum([d.shape[0] * w for d, w in zip(blocks, widths)])])
    np.testing.assert_array_equal(E.item_widths, widths)
    np.testing.assert_array_equal(E.item_offsets, offs)
    eci, ev = E.col_idxs(), E.values()
    for k, d in enumerate(blocks):
        n, w = d.shape[0], widths[k]
        for i in range(n):
            cols = np.nonzero(d[i])[0]
            for s in range(w):
                slot = offs[k] + i + s * n
                if s < len(cols):
                    assert eci[slot] == cols[s] and ev[slot] == d[i, cols[s]]
                else:
                    assert eci[slot] == -1 and ev[slot] == 0.0
    D1 = gko.batch.matrix.BlockDiagDense("f64").from_block_diag_csr(host, row_off, rp, ci, v)
    D2 = gko.batch.matrix.BlockDiagDense("f64").from_numpy(
        host, row_off, np.concatenate([d.reshape(-1) for d in blocks]))
    want = np.concatenate([d.reshape(-1) for d in blocks])
    np.testing.assert_array_equal(D1.values(), want)
    np.testing.assert_array_equal(D2.values(), want)
    np.testing.assert_array_equal(D1.item_offsets,
                                  np.concatenate([[0], np.cumsum([d.size for d in blocks])]))
    with pytest.raises(gko.DimensionMismatch):
        gko.batch.matrix.BlockDiagDense("f64").from_numpy(host, row_off, np.ones(3))
    with pytest.raises(gko.DimensionMismatch):   # a column outside its item
        gko.batch.matrix.BlockDiagEll("f64").from_numpy(
            host, np.array([0, 1]), np.array([0, 1], np.int32), np.array([1], np.int32),
            np.ones(1))
