"""Seeded synthetic problems for the sparse solver path (numpy, host side).

Every generator returns CSR arrays (int32 row pointers and column indices,
rows sorted by column) plus whatever vectors the workload needs.  The
stencils follow benchmark/utils/stencil_matrix.hpp of the reference
(Dirichlet boundaries: neighbours outside the grid are dropped).

Workloads and seeds:
  * poisson2d(256)           : 5-pt, diag 4 / off -1; x_true ~ U[-1,1] (seed 1)
  * laplace3d(128)           : 7-pt, diag 6 / off -1; x ~ U[-1,1] (seed 2)
  * powerlaw(10**6)          : lognormal row lengths (mean 16, sigma 1.2, <= 4096),
                               uniform columns without duplicates, U[-1,1] (seed 3)
  * diffusion27(96)          : variable-coefficient 27-pt SPD M-matrix (seed 4)
  * uneven_batch(20000)      : block-diagonal batch, n_i ~ U{16..512} (seed 5)
  * anisotropic2d(512)       : per-edge conductivities theta = exp(N(0, .5^2)) (seed 6)
"""
import numpy as np


def _csr_from_coo(n_rows, rows, cols, vals, n_cols=None):
    n_cols = n_rows if n_cols is None else n_cols
    key = rows.astype(np.int64) * n_cols + cols.astype(np.int64)
    order = np.argsort(key, kind="stable")
    rows = rows[order]
    cols = cols[order].astype(np.int32)
    vals = vals[order]
    counts = np.bincount(rows, minlength=n_rows)
    row_ptrs = np.zeros(n_rows + 1, dtype=np.int64)
    np.cumsum(counts, out=row_ptrs[1:])
    assert row_ptrs[-1] < 2**31, "int32 index range exceeded"
    return row_ptrs.astype(np.int32), cols, vals


def stencil(dims, offsets_weights, diag, dtype=np.float64):
    """Constant-coefficient stencil on a box grid with Dirichlet boundaries."""
    dims = tuple(dims)
    n = int(np.prod(dims))
    grid = np.indices(dims).reshape(len(dims), -1)
    rows_l, cols_l, vals_l = [np.arange(n)], [np.arange(n)], [np.full(n, diag)]
    strides = np.array([int(np.prod(dims[i + 1:])) for i in range(len(dims))])
    for off, w in offsets_weights:
        off = np.asarray(off)
        nb = grid + off[:, None]
        ok = np.all((nb >= 0) & (nb < np.asarray(dims)[:, None]), axis=0)
        src = np.nonzero(ok)[0]
        dst = (nb[:, ok] * strides[:, None]).sum(axis=0)
        rows_l.append(src)
        cols_l.append(dst)
        vals_l.append(np.full(src.shape[0], w))
    rows = np.concatenate(rows_l)
    cols = np.concatenate(cols_l)
    vals = np.concatenate(vals_l).astype(dtype)
    return _csr_from_coo(n, rows, cols, vals)


def box_grid_spd(n, nnz, shift=1e-3, dtype=np.float64):
    """SPD matrix of about n rows and nnz entries: a 2-D grid of
    ceil(sqrt(n)) columns and enough rows for n points, coupled to every
    point of a (2a+1) x (2b+1) box around it with weight -1, the box chosen
    so that its size is closest to nnz / n; the diagonal is the number of
    neighbours plus `shift` (weakly dominant, so CG converges slowly)."""
    k = max(1.0, nnz / n)
    best = None
    for a in range(0, 40):
        for b in range(a, 40):
            d = abs((2 * a + 1) * (2 * b + 1) - k)
            if best is None or d < best[0]:
                best = (d, a, b)
    _, a, b = best
    nx = int(np.ceil(np.sqrt(n)))
    ny = int(np.ceil(n / nx))
    offs = [((i, j), -1.0) for i in range(-a, a + 1) for j in range(-b, b + 1) if (i, j) != (0, 0)]
    rp, ci, v = stencil((ny, nx), offs, 0.0, dtype)
    counts = np.diff(rp.astype(np.int64)) - 1
    diag_pos = np.repeat(np.arange(len(rp) - 1), np.diff(rp)) == ci
    v[diag_pos] = (counts + shift).astype(dtype)
    return rp, ci, v


def poisson2d(nx, dtype=np.float64):
    offs = [((-1, 0), -1.0), ((1, 0), -1.0), ((0, -1), -1.0), ((0, 1), -1.0)]
    return stencil((nx, nx), offs, 4.0, dtype)


def laplace3d(nx, dtype=np.float64):
    offs = []
    for d in range(3):
        for s in (-1, 1):
            o = [0, 0, 0]
            o[d] = s
            offs.append((tuple(o), -1.0))
    return stencil((nx, nx, nx), offs, 6.0, dtype)


def uniform(rng, shape, dtype=np.float64):
    return rng.uniform(-1.0, 1.0, size=shape).astype(dtype)


def powerlaw(n, mean=16.0, sigma=1.2, cap=4096, seed=3, dtype=np.float64):
    """Rows with lognormal lengths (arithmetic mean `mean`), columns uniform
    over [0, n) without duplicates inside a row, values U[-1,1]."""
    rng = np.random.default_rng(seed)
    mu = np.log(mean) - sigma * sigma / 2.0
    lengths = np.rint(rng.lognormal(mu, sigma, size=n)).astype(np.int64)
    lengths = np.clip(lengths, 1, min(cap, n))
    nnz = int(lengths.sum())
    rows = np.repeat(np.arange(n, dtype=np.int64), lengths)
    cols = rng.integers(0, n, size=nnz, dtype=np.int64)
    # resample duplicates inside a row until none remain
    for _ in range(64):
        key = rows * n + cols
        order = np.argsort(key, kind="stable")
        ks = key[order]
        dup = np.zeros(nnz, dtype=bool)
        dup[order[1:][ks[1:] == ks[:-1]]] = True
        nd = int(dup.sum())
        if nd == 0:
            break
        cols[dup] = rng.integers(0, n, size=nd, dtype=np.int64)
    vals = uniform(rng, nnz, dtype)
    return _csr_from_coo(n, rows, cols, vals)


def diffusion27(nx, seed=4, dtype=np.float64):
    """27-point variable-coefficient diffusion on an nx^3 cell grid.

    Cell coefficients c ~ lognormal(0, 1).  The coupling between two
    neighbouring cells is -w * harmonic_mean(c_i, c_j) with w = 1 for faces,
    1/2 for edges and 1/4 for corners (symmetric).  The diagonal is the row's
    absolute off-diagonal sum plus 1e-3, so A is a symmetric, strictly
    diagonally dominant M-matrix (SPD)."""
    rng = np.random.default_rng(seed)
    dims = (nx, nx, nx)
    n = nx ** 3
    c = rng.lognormal(0.0, 1.0, size=n)
    grid = np.indices(dims).reshape(3, -1)
    strides = np.array([nx * nx, nx, 1])
    rows_l, cols_l, vals_l = [], [], []
    absum = np.zeros(n)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                k = abs(dz) + abs(dy) + abs(dx)
                if k == 0:
                    continue
                w = {1: 1.0, 2: 0.5, 3: 0.25}[k]
                off = np.array([dz, dy, dx])
                nb = grid + off[:, None]
                ok = np.all((nb >= 0) & (nb < nx), axis=0)
                src = np.nonzero(ok)[0]
                dst = (nb[:, ok] * strides[:, None]).sum(axis=0)
                h = 2.0 / (1.0 / c[src] + 1.0 / c[dst])
                v = -w * h
                absum += np.bincount(src, weights=-v, minlength=n)
                rows_l.append(src)
                cols_l.append(dst)
                vals_l.append(v)
    rows_l.append(np.arange(n))
    cols_l.append(np.arange(n))
    vals_l.append(absum + 1e-3)
    rows = np.concatenate(rows_l)
    cols = np.concatenate(cols_l)
    vals = np.concatenate(vals_l).astype(dtype)
    return _csr_from_coo(n, rows, cols, vals)


def uneven_batch(num_systems=20000, n_min=16, n_max=512, k=8, band=32, seed=5,
                 dtype=np.float64):
    """Block-diagonal batch of independent nonsymmetric systems.

    System s has n_s ~ U{n_min..n_max} rows.  Each row i gets k distinct
    off-diagonal columns drawn from [i-band, i+band] minus {i}, clipped to the
    system; values U[-1,1]; diagonal = row abs sum + 1.  Returns
    (row_off[int64, S+1], row_ptrs[int32], col_idxs[int32, local], vals, b)."""
    if n_min <= k:
        raise ValueError("uneven_batch: every system needs more than k rows")
    rng = np.random.default_rng(seed)
    sizes = rng.integers(n_min, n_max + 1, size=num_systems).astype(np.int64)
    row_off = np.zeros(num_systems + 1, dtype=np.int64)
    np.cumsum(sizes, out=row_off[1:])
    total = int(row_off[-1])
    sys_of_row = np.repeat(np.arange(num_systems), sizes)
    local = np.arange(total, dtype=np.int64) - row_off[sys_of_row]
    nsz = sizes[sys_of_row]
    offsets = np.concatenate([np.arange(-band, 0), np.arange(1, band + 1)])
    picks = np.empty((total, k), dtype=np.int64)
    chunk = 1 << 20
    for s0 in range(0, total, chunk):
        s1 = min(total, s0 + chunk)
        cand = local[s0:s1, None] + offsets[None, :]
        valid = (cand >= 0) & (cand < nsz[s0:s1, None])
        keys = rng.random(size=cand.shape, dtype=np.float32)
        keys[~valid] = 2.0
        sel = np.argpartition(keys, k - 1, axis=1)[:, :k]
        picks[s0:s1] = np.take_along_axis(cand, sel, axis=1)
    off_vals = uniform(rng, (total, k), np.float64)
    diag = np.abs(off_vals).sum(axis=1) + 1.0
    rows = np.concatenate([np.repeat(np.arange(total, dtype=np.int64), k),
                           np.arange(total, dtype=np.int64)])
    cols = np.concatenate([picks.reshape(-1), local])
    vals = np.concatenate([off_vals.reshape(-1), diag]).astype(dtype)
    # sort by global (row, local col); local col < n_max so the key is unique
    key = rows * (n_max + 1) + cols
    order = np.argsort(key, kind="stable")
    rows, cols, vals = rows[order], cols[order].astype(np.int32), vals[order]
    counts = np.bincount(rows, minlength=total)
    row_ptrs = np.zeros(total + 1, dtype=np.int64)
    np.cumsum(counts, out=row_ptrs[1:])
    b = uniform(rng, total, dtype)
    return row_off, row_ptrs.astype(np.int32), cols, vals, b


def uneven_spd_batch(num_systems=20000, n_min=16, n_max=512, k=8, band=32, seed=5,
                     dtype=np.float64):
    """The systems of uneven_batch made symmetric positive definite (for batch
    CG): the off-diagonal part O becomes O + O^T (entries where both exist are
    added), and the diagonal is again the row abs sum + 1, so every system is
    symmetric and strictly diagonally dominant with a positive diagonal.
    Returns (row_off, row_ptrs, col_idxs, vals, b) as uneven_batch."""
    row_off, rp, ci, v, b = uneven_batch(num_systems, n_min, n_max, k, band, seed, np.float64)
    total = int(row_off[-1])
    sizes = np.diff(row_off)
    rows = np.repeat(np.arange(total, dtype=np.int64), np.diff(rp))
    sys_of_row = np.repeat(np.arange(num_systems), sizes)
    local = rows - row_off[sys_of_row[rows]]
    off = ci != local
    r_o, c_o, v_o = rows[off], ci[off].astype(np.int64), v[off]
    g_c = c_o + row_off[sys_of_row[r_o]]          # global column of each entry
    lr = r_o - row_off[sys_of_row[r_o]]           # local row
    rows2 = np.concatenate([r_o, g_c])
    cols2 = np.concatenate([c_o, lr])
    vals2 = np.concatenate([v_o, v_o])
    key = rows2 * (n_max + 1) + cols2
    uniq, inv = np.unique(key, return_inverse=True)
    sv = np.bincount(inv, weights=vals2, minlength=uniq.shape[0])
    ur, uc = uniq // (n_max + 1), uniq % (n_max + 1)
    diag = np.bincount(ur, weights=np.abs(sv), minlength=total) + 1.0
    all_r = np.concatenate([ur, np.arange(total, dtype=np.int64)])
    all_c = np.concatenate([uc, np.arange(total, dtype=np.int64) - row_off[sys_of_row]])
    all_v = np.concatenate([sv, diag])
    order = np.argsort(all_r * (n_max + 1) + all_c, kind="stable")
    all_r, all_c, all_v = all_r[order], all_c[order].astype(np.int32), all_v[order]
    row_ptrs = np.zeros(total + 1, dtype=np.int64)
    np.cumsum(np.bincount(all_r, minlength=total), out=row_ptrs[1:])
    return row_off, row_ptrs.astype(np.int32), all_c, all_v.astype(dtype), b.astype(dtype)


def stencil3pt_batch(num_items, n, diag=6.0, off=-1.0, dtype=np.float64):
    """The reference's generate_3pt_stencil_batch_matrix
    (core/test/utils/batch_helpers.hpp:95-120) as a block-diagonal batch."""
    rp, ci, v = stencil((n,), [((-1,), off), ((1,), off)], diag, dtype)
    row_off = np.arange(num_items + 1, dtype=np.int64) * n
    nnz = int(rp[-1])
    row_ptrs = np.concatenate([rp[:-1] + s * nnz for s in range(num_items)] +
                              [np.array([num_items * nnz])]).astype(np.int32)
    col_idxs = np.tile(ci, num_items)
    vals = np.tile(v, num_items)
    return row_off, row_ptrs, col_idxs, vals


def anisotropic2d(nx, seed=6, eps_y=0.1, convection=0.0, dtype=np.float64):
    """2-D diffusion on an nx x nx grid with per-edge conductivities.

    theta_e = exp(g_e), g_e ~ N(0, 0.5^2), one per grid edge, including the
    edges that connect boundary cells to the Dirichlet exterior (which only
    reach the diagonal).  y-edges are scaled by eps_y (anisotropy).  With
    convection > 0 an upwind term in +x is added (a_ii += c, a_i,i-x -= c),
    which makes A nonsymmetric.

    Returns (row_ptrs, col_idxs, vals, theta, value_map, coef): the diffusion
    part of the values is the sum over t of
    coef[t] * theta[value_map[t, 1]] into vals[value_map[t, 0]], so vals is
    linear in theta (the convection part is constant)."""
    rng = np.random.default_rng(seed)
    n = nx * nx
    idx = np.arange(n)
    # edges as (a, b) cell pairs, -1 being the Dirichlet exterior: x-edges run
    # along each grid row and y-edges along each grid column, both ends included
    ea, eb, es = [], [], []
    for i in range(nx):
        row = i * nx + np.arange(nx)
        ea.append(np.concatenate([[-1], row]))
        eb.append(np.concatenate([row, [-1]]))
        es.append(np.ones(nx + 1))
    for j in range(nx):
        col = np.arange(nx) * nx + j
        ea.append(np.concatenate([[-1], col]))
        eb.append(np.concatenate([col, [-1]]))
        es.append(np.full(nx + 1, eps_y))
    ea = np.concatenate(ea)
    eb = np.concatenate(eb)
    es = np.concatenate(es)
    ne = ea.shape[0]
    theta = np.exp(rng.normal(0.0, 0.5, size=ne))
    # contributions: (row, col, edge, coeff): a_aa += s*theta, a_bb += s*theta,
    # a_ab = a_ba -= s*theta (only for interior pairs)
    r_l, c_l, e_l, k_l = [], [], [], []
    eidx = np.arange(ne)
    for (p, q) in ((ea, eb), (eb, ea)):
        m = p >= 0
        r_l.append(p[m]); c_l.append(p[m]); e_l.append(eidx[m]); k_l.append(es[m])
        m2 = (p >= 0) & (q >= 0)
        r_l.append(p[m2]); c_l.append(q[m2]); e_l.append(eidx[m2]); k_l.append(-es[m2])
    rows = np.concatenate(r_l)
    cols = np.concatenate(c_l)
    eds = np.concatenate(e_l)
    coef = np.concatenate(k_l)
    key = rows.astype(np.int64) * n + cols
    uniq, inv = np.unique(key, return_inverse=True)
    vals = np.zeros(uniq.shape[0])
    np.add.at(vals, inv, coef * theta[eds])
    if convection > 0.0:
        conv = np.zeros(uniq.shape[0])
        # upwind in +x: a_ii += c, a_{i, i-1} -= c (x-neighbour inside the grid)
        diag_pos = np.searchsorted(uniq, idx.astype(np.int64) * n + idx)
        conv[diag_pos] += convection
        has_left = (idx % nx) > 0
        left_key = idx[has_left].astype(np.int64) * n + (idx[has_left] - 1)
        conv[np.searchsorted(uniq, left_key)] -= convection
        vals = vals + conv
    urow = (uniq // n).astype(np.int64)
    ucol = (uniq % n).astype(np.int32)
    counts = np.bincount(urow, minlength=n)
    row_ptrs = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=row_ptrs[1:])
    value_map = np.stack([inv.astype(np.int64), eds.astype(np.int64)], axis=1)
    return (row_ptrs.astype(np.int32), ucol, vals.astype(dtype), theta, value_map,
            coef)


def csr_to_dense(row_ptrs, col_idxs, vals, n_cols=None):
    n = len(row_ptrs) - 1
    n_cols = n if n_cols is None else n_cols
    d = np.zeros((n, n_cols), dtype=vals.dtype)
    rows = np.repeat(np.arange(n), np.diff(row_ptrs))
    np.add.at(d, (rows, col_idxs), vals)
    return d


def dense_to_csr(d, keep_zeros=False):
    d = np.asarray(d)
    mask = np.ones_like(d, dtype=bool) if keep_zeros else d != 0
    rows, cols = np.nonzero(mask)
    counts = np.bincount(rows, minlength=d.shape[0])
    row_ptrs = np.zeros(d.shape[0] + 1, dtype=np.int32)
    np.cumsum(counts, out=row_ptrs[1:])
    return row_ptrs, cols.astype(np.int32), d[rows, cols]


def pad_to_uniform(row_off, row_ptrs, col_idxs, vals, b, n_pad=None):
    """The reference-compatible uniform batch of an uneven one: every system
    padded to n_pad rows with one shared pattern, the union of all systems'
    patterns plus the diagonal (batch::matrix::Csr needs one pattern,
    include/ginkgo/core/matrix/batch_csr.hpp:29-31).  Padded rows carry a
    unit diagonal and a zero right-hand side, so the padded unknowns are 0.
    Returns (row_ptrs[n_pad+1], col_idxs, values[S * nnz], b[S * n_pad])."""
    row_off = np.asarray(row_off, dtype=np.int64)
    sizes = np.diff(row_off)
    S = len(sizes)
    n_pad = int(sizes.max()) if n_pad is None else int(n_pad)
    total = int(row_off[-1])
    sys_of_row = np.repeat(np.arange(S), sizes)
    local_row = np.arange(total) - row_off[sys_of_row]
    counts = np.diff(row_ptrs)
    e_row = np.repeat(local_row, counts)
    e_sys = np.repeat(sys_of_row, counts)
    mask = np.zeros((n_pad, n_pad), dtype=bool)
    mask[e_row, col_idxs] = True
    mask[np.arange(n_pad), np.arange(n_pad)] = True
    pr, pc = np.nonzero(mask)
    nnz = pr.shape[0]
    index = np.full((n_pad, n_pad), -1, dtype=np.int64)
    index[pr, pc] = np.arange(nnz)
    rp = np.zeros(n_pad + 1, dtype=np.int32)
    np.cumsum(np.bincount(pr, minlength=n_pad), out=rp[1:])
    values = np.zeros((S, nnz), dtype=vals.dtype)
    values[e_sys, index[e_row, col_idxs]] = vals
    # unit diagonal on the padded rows of each system
    diag_pos = index[np.arange(n_pad), np.arange(n_pad)]
    pad_mask = np.arange(n_pad)[None, :] >= sizes[:, None]
    rows_s, rows_i = np.nonzero(pad_mask)
    values[rows_s, diag_pos[rows_i]] = 1.0
    bp = np.zeros((S, n_pad), dtype=b.dtype)
    bp[sys_of_row, local_row] = b
    return rp, pc.astype(np.int32), values.reshape(-1), bp.reshape(-1)
