"""ctypes wrapper of the CPU oracle (test infrastructure only).

Every function takes and returns numpy arrays.  Dense blocks are row-major
``(n, k)`` arrays, the layout of ``matrix::view::dense``
(include/ginkgo/core/matrix/device_views.hpp:24-45).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

STATUS_CONVERGED = 0x80
STATUS_FINALIZED = 0x40
STATUS_ID_MASK = 0x3F

CRIT_ITERATION = 1
CRIT_RESIDUAL_NORM = 2
CRIT_IMPLICIT_RESNORM = 3
BASELINE = {"rhs_norm": 0, "initial_resnorm": 1, "absolute": 2}


class Criterion(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("baseline", ctypes.c_int32),
        ("max_iters", ctypes.c_int64),
        ("reduction_factor", ctypes.c_double),
    ]


class SolveInfo(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int64),
        ("spmv_count", ctypes.c_int64),
        ("final_resnorm", ctypes.c_double * 64),
    ]


def build():
    """Compile liboracle.so with its Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _suffix(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float64:
        return "d"
    if dtype == np.float32:
        return "s"
    raise TypeError(f"unsupported value type {dtype}")


def _scalar(dtype, v):
    return ctypes.c_double(v) if np.dtype(dtype) == np.float64 else ctypes.c_float(v)


def criteria(spec, dtype=np.float64):
    """Build a Criterion array from [{'kind': 'iteration', 'max_iters': N},
    {'kind': 'residual_norm', 'reduction_factor': f, 'baseline': 'rhs_norm'}]."""
    arr = (Criterion * max(1, len(spec)))()
    for i, c in enumerate(spec):
        kind = c["kind"]
        if kind == "iteration":
            arr[i].kind = CRIT_ITERATION
            arr[i].max_iters = int(c["max_iters"])
        else:
            arr[i].kind = CRIT_RESIDUAL_NORM if kind == "residual_norm" else CRIT_IMPLICIT_RESNORM
            arr[i].baseline = BASELINE[c.get("baseline", "rhs_norm")]
            arr[i].reduction_factor = float(c["reduction_factor"])
    return arr, len(spec)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def csr_spmv(row_ptrs, col_idxs, vals, b, omp=False):
    dt = vals.dtype
    b = _c(b, dt).reshape(b.shape[0], -1)
    n = len(row_ptrs) - 1
    c = np.zeros((n, b.shape[1]), dtype=dt)
    fn = getattr(lib(), f"ora_csr_spmv{'_omp' if omp else ''}_{_suffix(dt)}")
    fn(ctypes.c_int64(n), ctypes.c_int64(b.shape[1]), _p(_c(row_ptrs, np.int32)),
       _p(_c(col_idxs, np.int32)), _p(_c(vals, dt)), _p(b), ctypes.c_int64(b.shape[1]),
       _p(c), ctypes.c_int64(c.shape[1]))
    return c


def csr_advanced_spmv(alpha, row_ptrs, col_idxs, vals, b, beta, c):
    dt = vals.dtype
    b = _c(b, dt).reshape(b.shape[0], -1)
    c = _c(c, dt).reshape(c.shape[0], -1).copy()
    n = len(row_ptrs) - 1
    getattr(lib(), f"ora_csr_advanced_spmv_{_suffix(dt)}")(
        _scalar(dt, alpha), ctypes.c_int64(n), ctypes.c_int64(b.shape[1]),
        _p(_c(row_ptrs, np.int32)), _p(_c(col_idxs, np.int32)), _p(_c(vals, dt)), _p(b),
        ctypes.c_int64(b.shape[1]), _scalar(dt, beta), _p(c), ctypes.c_int64(c.shape[1]))
    return c


def sellp_slice_sets(row_ptrs, slice_size=64, stride_factor=1):
    n = len(row_ptrs) - 1
    ns = (n + slice_size - 1) // slice_size
    sets = np.zeros(ns + 1, dtype=np.int64)
    lens = np.zeros(max(ns, 1), dtype=np.int64)
    lib().ora_sellp_compute_slice_sets(ctypes.c_int64(n), _p(_c(row_ptrs, np.int32)),
                                       ctypes.c_int64(slice_size), ctypes.c_int64(stride_factor),
                                       _p(sets), _p(lens))
    return sets, lens[:ns]


def csr_to_sellp(row_ptrs, col_idxs, vals, slice_size=64, stride_factor=1):
    """Returns (slice_sets, slice_lengths, cols, vals).  Storage for rows past
    n in the last slice is pre-filled with (-1, 0)."""
    dt = vals.dtype
    n = len(row_ptrs) - 1
    sets, lens = sellp_slice_sets(row_ptrs, slice_size, stride_factor)
    total = int(sets[-1]) * slice_size
    s_cols = np.full(max(total, 1), -1, dtype=np.int32)
    s_vals = np.zeros(max(total, 1), dtype=dt)
    lens_buf = lens if len(lens) else np.zeros(1, np.int64)
    getattr(lib(), f"ora_csr_to_sellp_{_suffix(dt)}")(
        ctypes.c_int64(n), _p(_c(row_ptrs, np.int32)), _p(_c(col_idxs, np.int32)),
        _p(_c(vals, dt)), ctypes.c_int64(slice_size), _p(sets), _p(lens_buf), _p(s_cols),
        _p(s_vals))
    return sets, lens, s_cols[:total], s_vals[:total]


def sellp_spmv(n, slice_size, sets, lens, s_cols, s_vals, b, omp=False):
    dt = s_vals.dtype
    b = _c(b, dt).reshape(b.shape[0], -1)
    c = np.zeros((n, b.shape[1]), dtype=dt)
    lens_buf = lens if len(lens) else np.zeros(1, np.int64)
    getattr(lib(), f"ora_sellp_spmv{'_omp' if omp else ''}_{_suffix(dt)}")(
        ctypes.c_int64(n), ctypes.c_int64(b.shape[1]), ctypes.c_int64(slice_size), _p(sets),
        _p(lens_buf), _p(_c(s_cols, np.int32)), _p(_c(s_vals, dt)), _p(b),
        ctypes.c_int64(b.shape[1]), _p(c), ctypes.c_int64(c.shape[1]))
    return c


def sellp_advanced_spmv(alpha, n, slice_size, sets, lens, s_cols, s_vals, b, beta, c):
    dt = s_vals.dtype
    b = _c(b, dt).reshape(b.shape[0], -1)
    c = _c(c, dt).reshape(c.shape[0], -1).copy()
    lens_buf = lens if len(lens) else np.zeros(1, np.int64)
    getattr(lib(), f"ora_sellp_advanced_spmv_{_suffix(dt)}")(
        _scalar(dt, alpha), ctypes.c_int64(n), ctypes.c_int64(b.shape[1]),
        ctypes.c_int64(slice_size), _p(sets), _p(lens_buf), _p(_c(s_cols, np.int32)),
        _p(_c(s_vals, dt)), _p(b), ctypes.c_int64(b.shape[1]), _scalar(dt, beta), _p(c),
        ctypes.c_int64(c.shape[1]))
    return c


def max_row_nnz(row_ptrs):
    f = lib().ora_csr_max_row_nnz
    f.restype = ctypes.c_int64
    return int(f(ctypes.c_int64(len(row_ptrs) - 1), _p(_c(row_ptrs, np.int32))))


def lb_clac_size(nnz, nwarps, warp_size=32, cuda_strategy=True, intel=False):
    """Csr::load_balance::clac_size (HIP-platform build): srow elements."""
    f = lib().ora_csr_lb_clac_size
    f.restype = ctypes.c_int64
    return int(f(ctypes.c_int64(nnz), ctypes.c_int64(nwarps), ctypes.c_int32(warp_size),
                 ctypes.c_int32(int(bool(cuda_strategy))), ctypes.c_int32(int(bool(intel)))))


def lb_process(row_ptrs, nwarps, warp_size=32):
    """Csr::load_balance::process: the srow partition of nwarps entries."""
    srow = np.zeros(max(1, nwarps), dtype=np.int32)
    lib().ora_csr_lb_process(ctypes.c_int64(len(row_ptrs) - 1), _p(_c(row_ptrs, np.int32)),
                             ctypes.c_int64(warp_size), _p(srow), ctypes.c_int64(nwarps))
    return srow[:nwarps]


def csr_to_ell(row_ptrs, col_idxs, vals, width=None, stride=None):
    dt = vals.dtype
    n = len(row_ptrs) - 1
    width = max_row_nnz(row_ptrs) if width is None else width
    stride = n if stride is None else stride
    e_cols = np.zeros(max(width * stride, 1), dtype=np.int32)
    e_vals = np.zeros(max(width * stride, 1), dtype=dt)
    getattr(lib(), f"ora_csr_to_ell_{_suffix(dt)}")(
        ctypes.c_int64(n), _p(_c(row_ptrs, np.int32)), _p(_c(col_idxs, np.int32)),
        _p(_c(vals, dt)), ctypes.c_int64(width), ctypes.c_int64(stride), _p(e_cols), _p(e_vals))
    return width, stride, e_cols[: width * stride], e_vals[: width * stride]


def ell_spmv(n, width, stride, e_cols, e_vals, b):
    dt = e_vals.dtype
    b = _c(b, dt).reshape(b.shape[0], -1)
    c = np.zeros((n, b.shape[1]), dtype=dt)
    ec = e_cols if len(e_cols) else np.zeros(1, np.int32)
    ev = e_vals if len(e_vals) else np.zeros(1, dt)
    getattr(lib(), f"ora_ell_spmv_{_suffix(dt)}")(
        ctypes.c_int64(n), ctypes.c_int64(b.shape[1]), ctypes.c_int64(width),
        ctypes.c_int64(stride), _p(_c(ec, np.int32)), _p(_c(ev, dt)), _p(b),
        ctypes.c_int64(b.shape[1]), _p(c), ctypes.c_int64(c.shape[1]))
    return c


def ell_advanced_spmv(alpha, n, width, stride, e_cols, e_vals, b, beta, c):
    dt = e_vals.dtype
    b = _c(b, dt).reshape(b.shape[0], -1)
    c = _c(c, dt).reshape(c.shape[0], -1).copy()
    getattr(lib(), f"ora_ell_advanced_spmv_{_suffix(dt)}")(
        _scalar(dt, alpha), ctypes.c_int64(n), ctypes.c_int64(b.shape[1]), ctypes.c_int64(width),
        ctypes.c_int64(stride), _p(_c(e_cols, np.int32)), _p(_c(e_vals, dt)), _p(b),
        ctypes.c_int64(b.shape[1]), _scalar(dt, beta), _p(c), ctypes.c_int64(c.shape[1]))
    return c


def csr_transpose(n_rows, n_cols, row_ptrs, col_idxs, vals):
    dt = vals.dtype
    nnz = int(row_ptrs[-1])
    t_rp = np.zeros(n_cols + 1, dtype=np.int32)
    t_ci = np.zeros(max(nnz, 1), dtype=np.int32)
    t_v = np.zeros(max(nnz, 1), dtype=dt)
    getattr(lib(), f"ora_csr_transpose_{_suffix(dt)}")(
        ctypes.c_int64(n_rows), ctypes.c_int64(n_cols), _p(_c(row_ptrs, np.int32)),
        _p(_c(col_idxs, np.int32) if nnz else np.zeros(1, np.int32)),
        _p(_c(vals, dt) if nnz else np.zeros(1, dt)), _p(t_rp), _p(t_ci), _p(t_v))
    return t_rp, t_ci[:nnz], t_v[:nnz]


def dot(x, y):
    dt = x.dtype
    x = _c(x, dt).reshape(x.shape[0], -1)
    y = _c(y, dt).reshape(y.shape[0], -1)
    out = np.zeros(x.shape[1], dtype=dt)
    getattr(lib(), f"ora_dense_dot_{_suffix(dt)}")(
        ctypes.c_int64(x.shape[0]), ctypes.c_int64(x.shape[1]), _p(x), ctypes.c_int64(x.shape[1]),
        _p(y), ctypes.c_int64(y.shape[1]), _p(out))
    return out


def norm2(x):
    dt = x.dtype
    x = _c(x, dt).reshape(x.shape[0], -1)
    out = np.zeros(x.shape[1], dtype=dt)
    getattr(lib(), f"ora_dense_norm2_{_suffix(dt)}")(
        ctypes.c_int64(x.shape[0]), ctypes.c_int64(x.shape[1]), _p(x), ctypes.c_int64(x.shape[1]),
        _p(out))
    return out


def jacobi_inverse_diagonal(row_ptrs, col_idxs, vals):
    dt = vals.dtype
    n = len(row_ptrs) - 1
    diag = np.zeros(n, dtype=dt)
    inv = np.zeros(n, dtype=dt)
    s = _suffix(dt)
    getattr(lib(), f"ora_csr_extract_diagonal_{s}")(
        ctypes.c_int64(n), _p(_c(row_ptrs, np.int32)), _p(_c(col_idxs, np.int32)),
        _p(_c(vals, dt)), _p(diag))
    getattr(lib(), f"ora_jacobi_invert_diagonal_{s}")(ctypes.c_int64(n), _p(diag), _p(inv))
    return inv


def jacobi_apply(inv_diag, b):
    dt = inv_diag.dtype
    b = _c(b, dt).reshape(b.shape[0], -1)
    x = np.zeros_like(b)
    getattr(lib(), f"ora_jacobi_simple_scalar_apply_{_suffix(dt)}")(
        ctypes.c_int64(b.shape[0]), ctypes.c_int64(b.shape[1]), _p(inv_diag), _p(b),
        ctypes.c_int64(b.shape[1]), _p(x), ctypes.c_int64(x.shape[1]))
    return x


def _solve(kind, row_ptrs, col_idxs, vals, b, x0, crit, inv_diag=None, omp=False):
    dt = vals.dtype
    b = _c(b, dt).reshape(b.shape[0], -1)
    x = _c(x0, dt).reshape(b.shape).copy()
    n, k = b.shape
    carr, nc = criteria(crit, dt)
    status = np.zeros(k, dtype=np.uint8)
    info = SolveInfo()
    fn = getattr(lib(), f"ora_{kind}_solve_{_suffix(dt)}")
    rc = fn(ctypes.c_int64(n), ctypes.c_int64(k), _p(_c(row_ptrs, np.int32)),
            _p(_c(col_idxs, np.int32)), _p(_c(vals, dt)),
            _p(_c(inv_diag, dt)) if inv_diag is not None else None, _p(b), _p(x), carr,
            ctypes.c_int32(nc), ctypes.c_int(1 if omp else 0), _p(status), ctypes.byref(info))
    if rc != 0:
        raise MemoryError("oracle solve failed to allocate")
    return x, status, info


def cg_solve(row_ptrs, col_idxs, vals, b, x0, crit, inv_diag=None, omp=False):
    return _solve("cg", row_ptrs, col_idxs, vals, b, x0, crit, inv_diag, omp)


def bicgstab_solve(row_ptrs, col_idxs, vals, b, x0, crit, inv_diag=None, omp=False):
    return _solve("bicgstab", row_ptrs, col_idxs, vals, b, x0, crit, inv_diag, omp)


def ir_solve(row_ptrs, col_idxs, vals, b, x0, outer, inner=None, relaxation=1.0,
             inner_cg=True, omp=False, per_sweep=False):
    """(x, status, info, inner iterations in total); with per_sweep also the
    inner iterations of every sweep (numpy int64 array)."""
    vals = _c(vals, np.float64)
    b = _c(b, np.float64).reshape(-1)
    x = _c(x0, np.float64).reshape(-1).copy()
    n = b.shape[0]
    oarr, no = criteria(outer)
    iarr, ni = criteria(inner or [{"kind": "iteration", "max_iters": 1}])
    status = np.zeros(1, dtype=np.uint8)
    info = SolveInfo()
    inner_total = ctypes.c_int64(0)
    cap = max(1024, max((int(c.get("max_iters", 0)) for c in outer), default=0) + 1)
    sweeps = np.full(cap, -1, dtype=np.int64)
    rc = lib().ora_ir_solve(ctypes.c_int64(n), _p(_c(row_ptrs, np.int32)),
                            _p(_c(col_idxs, np.int32)), _p(vals), _p(b), _p(x), oarr,
                            ctypes.c_int32(no), iarr, ctypes.c_int32(ni),
                            ctypes.c_double(relaxation), ctypes.c_int(1 if inner_cg else 0),
                            ctypes.c_int(1 if omp else 0), _p(status), ctypes.byref(info),
                            ctypes.byref(inner_total), _p(sweeps), ctypes.c_int64(cap))
    if rc != 0:
        raise MemoryError("oracle IR failed to allocate")
    if per_sweep:
        return x, status, info, inner_total.value, sweeps[:max(0, int(info.iterations))]
    return x, status, info, inner_total.value


def batch_bicgstab_solve(row_off, row_ptrs, col_idxs, vals, b, x0, use_jacobi=True,
                         relative=True, tol=1e-10, max_iters=100, omp=False):
    dt = vals.dtype
    row_off = _c(row_off, np.int64)
    ns = len(row_off) - 1
    x = _c(x0, dt).copy()
    iters = np.zeros(ns, dtype=np.int32)
    res = np.zeros(ns, dtype=dt)
    rc = getattr(lib(), f"ora_batch_bicgstab_solve_{_suffix(dt)}")(
        ctypes.c_int64(ns), _p(row_off), _p(_c(row_ptrs, np.int32)), _p(_c(col_idxs, np.int32)),
        _p(_c(vals, dt)), _p(_c(b, dt)), _p(x), ctypes.c_int(1 if use_jacobi else 0),
        ctypes.c_int(1 if relative else 0), ctypes.c_double(tol), ctypes.c_int32(max_iters),
        ctypes.c_int(1 if omp else 0), _p(iters), _p(res))
    if rc != 0:
        raise MemoryError("oracle batch solve failed to allocate")
    return x, iters, res


def coo_spmv2(rows, cols, vals, b, c, alpha=None):
    """c += (alpha *) A b entry by entry (reference/matrix/coo_kernels.cpp:61-101)."""
    dt = vals.dtype
    b = _c(b, dt).reshape(b.shape[0], -1)
    c = _c(c, dt).reshape(c.shape[0], -1).copy()
    a = np.array([alpha], dtype=dt) if alpha is not None else None
    getattr(lib(), f"ora_coo_spmv2_{_suffix(dt)}")(
        ctypes.c_int64(len(vals)), _p(_c(rows, np.int32)), _p(_c(cols, np.int32)), _p(_c(vals, dt)),
        _p(a) if a is not None else None, ctypes.c_int64(b.shape[1]), _p(b),
        ctypes.c_int64(b.shape[1]), _p(c), ctypes.c_int64(c.shape[1]))
    return c


def csr_to_hybrid(row_ptrs, col_idxs, vals, width, stride=None):
    """(ell_cols, ell_vals, coo_rows, coo_cols, coo_vals) of
    reference/matrix/csr_kernels.cpp:911-957 for ELL width `width`."""
    dt = vals.dtype
    n = len(row_ptrs) - 1
    stride = n if stride is None else stride
    lens = np.diff(np.asarray(row_ptrs, np.int64))
    coo_nnz = int(np.maximum(lens - width, 0).sum())
    ec = np.zeros(width * stride, np.int32)
    ev = np.zeros(width * stride, dt)
    cr = np.zeros(coo_nnz, np.int32)
    cc = np.zeros(coo_nnz, np.int32)
    cv = np.zeros(coo_nnz, dt)
    getattr(lib(), f"ora_csr_to_hybrid_{_suffix(dt)}")(
        ctypes.c_int64(n), _p(_c(row_ptrs, np.int32)), _p(_c(col_idxs, np.int32)), _p(_c(vals, dt)),
        ctypes.c_int64(width), ctypes.c_int64(stride), _p(ec), _p(ev), _p(cr), _p(cc), _p(cv))
    return ec, ev, cr, cc, cv


def hybrid_ell_width(row_nnz, strategy, p0=0.0, p1=0.0, value_bytes=8, index_bytes=4):
    """ELL width of the Hybrid strategies (include/ginkgo/core/matrix/hybrid.hpp:
    column_limit, imbalance_limit, imbalance_bounded_limit,
    minimal_storage_limit, automatic), from the row lengths."""
    lens = np.sort(np.asarray(row_nnz, np.int64))
    n = len(lens)

    def imbalance(percent):
        percent = min(max(percent, 0.0), 1.0)
        if n == 0:
            return 0
        return int(lens[int(n * percent)]) if percent < 1 else int(lens[n - 1])

    if strategy == "column_limit":
        return int(p0)
    if strategy == "imbalance_limit":
        return imbalance(p0)
    if strategy == "imbalance_bounded_limit":
        return min(imbalance(p0), int(n * p1))
    if strategy == "minimal_storage_limit":
        return imbalance(index_bytes / (value_bytes + 2 * index_bytes))
    if strategy == "automatic":
        return min(imbalance(1.0 / 3.0), int(n * 0.001))
    raise ValueError(strategy)


def batch_cg_solve(row_off, row_ptrs, col_idxs, vals, b, x0, use_jacobi=True, relative=True,
                   tol=1e-10, max_iters=100, omp=False):
    """Batch CG (reference/solver/batch_cg_kernels.hpp:31-221) over the
    block-diagonal batch format; returns (x, iterations, implicit residual norms)."""
    dt = vals.dtype
    row_off = _c(row_off, np.int64)
    ns = len(row_off) - 1
    x = _c(x0, dt).copy()
    iters = np.zeros(ns, dtype=np.int32)
    res = np.zeros(ns, dtype=dt)
    rc = getattr(lib(), f"ora_batch_cg_solve_{_suffix(dt)}")(
        ctypes.c_int64(ns), _p(row_off), _p(_c(row_ptrs, np.int32)), _p(_c(col_idxs, np.int32)),
        _p(_c(vals, dt)), _p(_c(b, dt)), _p(x), ctypes.c_int(1 if use_jacobi else 0),
        ctypes.c_int(1 if relative else 0), ctypes.c_double(tol), ctypes.c_int32(max_iters),
        ctypes.c_int(1 if omp else 0), _p(iters), _p(res))
    if rc != 0:
        raise MemoryError("oracle batch solve failed to allocate")
    return x, iters, res


def _raw(name):
    return getattr(lib(), name)


def cg_kernel(name, dtype):
    """Direct handle of a CG/BiCGSTAB step kernel (for the kernel-level KATs)."""
    return _raw(f"ora_{name}_{_suffix(dtype)}")


def omp_threads():
    return int(lib().ora_omp_threads())


def _batch_scalars(alpha, beta, dt):
    if (alpha is None) != (beta is None):
        raise ValueError("alpha and beta must both be given or both be None")
    if alpha is None:
        return None, None
    return _c(np.asarray(alpha), dt), _c(np.asarray(beta), dt)


def batch_csr_spmv(num, rows, cols, row_ptrs, col_idxs, vals, b, x=None, alpha=None, beta=None):
    """batch::matrix::Csr apply per system (common/cuda_hip/matrix/
    batch_csr_kernels.hpp:37-86); alpha/beta one value per system."""
    dt = vals.dtype
    a, bt = _batch_scalars(alpha, beta, dt)
    out = np.zeros(num * rows, dt) if x is None else _c(x, dt).copy()
    getattr(lib(), f"ora_batch_csr_spmv_{_suffix(dt)}")(
        ctypes.c_int64(num), ctypes.c_int64(rows), ctypes.c_int64(cols),
        _p(_c(row_ptrs, np.int32)), _p(_c(col_idxs, np.int32)), _p(_c(vals, dt)),
        _p(a) if a is not None else None, _p(_c(b, dt)), _p(bt) if bt is not None else None,
        _p(out))
    return out


def batch_ell_spmv(num, rows, cols, width, col_idxs, vals, b, x=None, alpha=None, beta=None):
    """batch::matrix::Ell apply per system (common/cuda_hip/matrix/
    batch_ell_kernels.hpp:30-100): a row ends at its first -1."""
    dt = vals.dtype
    a, bt = _batch_scalars(alpha, beta, dt)
    out = np.zeros(num * rows, dt) if x is None else _c(x, dt).copy()
    getattr(lib(), f"ora_batch_ell_spmv_{_suffix(dt)}")(
        ctypes.c_int64(num), ctypes.c_int64(rows), ctypes.c_int64(cols), ctypes.c_int64(width),
        _p(_c(col_idxs, np.int32)), _p(_c(vals, dt)), _p(a) if a is not None else None,
        _p(_c(b, dt)), _p(bt) if bt is not None else None, _p(out))
    return out


def batch_dense_spmv(num, rows, cols, vals, b, x=None, alpha=None, beta=None):
    """batch::matrix::Dense apply per system (common/cuda_hip/matrix/
    batch_dense_kernels.hpp:30-110), summed in column order."""
    dt = vals.dtype
    a, bt = _batch_scalars(alpha, beta, dt)
    out = np.zeros(num * rows, dt) if x is None else _c(x, dt).copy()
    getattr(lib(), f"ora_batch_dense_spmv_{_suffix(dt)}")(
        ctypes.c_int64(num), ctypes.c_int64(rows), ctypes.c_int64(cols), _p(_c(vals, dt)),
        _p(a) if a is not None else None, _p(_c(b, dt)), _p(bt) if bt is not None else None,
        _p(out))
    return out
