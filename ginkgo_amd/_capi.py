"""ctypes binding of the C ABI in include/gkomi.h (libgkomi.so).

The library is loaded from ``ginkgo_amd/lib``; if it is missing the import
fails loudly -- there is no CPU fallback.  Every wrapper raises a typed
exception (mirroring the reference's exception classes) on a non-zero
status.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgkomi.so")

SUCCESS = 0
INVALID_ARGUMENT = 1
NOT_SUPPORTED = 2
ALLOCATION_ERROR = 3
HIP_ERROR = 4
DIMENSION_MISMATCH = 5

VALUE_F64 = 0
VALUE_F32 = 1

CSR_CLASSICAL = 0
CSR_LOAD_BALANCE = 1
CSR_MERGE_PATH = 2

STATUS_CONVERGED = 0x80
STATUS_FINALIZED = 0x40
STATUS_ID_MASK = 0x3F

CRIT_ITERATION = 1
CRIT_RESIDUAL_NORM = 2
CRIT_IMPLICIT_RESNORM = 3
BASELINE_RHS_NORM = 0
BASELINE_INITIAL_RESNORM = 1
BASELINE_ABSOLUTE = 2
MAX_CRITERIA = 4


class GkoError(RuntimeError):
    """Base class of errors reported by the kernel library."""


class InvalidArgument(GkoError, ValueError):
    pass


class NotSupported(GkoError):
    pass


class AllocationError(GkoError, MemoryError):
    pass


class HipError(GkoError):
    pass


class DimensionMismatch(GkoError, ValueError):
    pass


_ERRORS = {
    INVALID_ARGUMENT: InvalidArgument,
    NOT_SUPPORTED: NotSupported,
    ALLOCATION_ERROR: AllocationError,
    HIP_ERROR: HipError,
    DIMENSION_MISMATCH: DimensionMismatch,
}


class Ctx(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("stream", ctypes.c_void_p)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("arch", ctypes.c_char * 32),
                ("num_cus", ctypes.c_int32), ("warp_size", ctypes.c_int32),
                ("max_lds_per_block", ctypes.c_int64), ("total_global_mem", ctypes.c_int64),
                ("clock_khz", ctypes.c_int32), ("memory_clock_khz", ctypes.c_int32),
                ("memory_bus_width", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class Csr(ctypes.Structure):
    _fields_ = [("n_rows", ctypes.c_int64), ("n_cols", ctypes.c_int64), ("nnz", ctypes.c_int64),
                ("row_ptrs", ctypes.c_void_p), ("col_idxs", ctypes.c_void_p),
                ("values", ctypes.c_void_p), ("strategy", ctypes.c_int32),
                ("subwave", ctypes.c_int32), ("srow", ctypes.c_void_p),
                ("num_srow", ctypes.c_int64), ("col_offsets", ctypes.c_void_p),
                ("work", ctypes.c_void_p), ("work_bytes", ctypes.c_int64),
                ("panels", ctypes.c_void_p), ("panels_bytes", ctypes.c_int64),
                ("num_panels", ctypes.c_int32), ("slice_sets", ctypes.c_void_p),
                ("slice_cols", ctypes.c_void_p), ("slice_offsets", ctypes.c_void_p),
                ("slice_values", ctypes.c_void_p)]


class Sellp(ctypes.Structure):
    _fields_ = [("n_rows", ctypes.c_int64), ("n_cols", ctypes.c_int64),
                ("slice_size", ctypes.c_int64), ("stride_factor", ctypes.c_int64),
                ("slice_sets", ctypes.c_void_p), ("slice_lengths", ctypes.c_void_p),
                ("col_idxs", ctypes.c_void_p), ("values", ctypes.c_void_p)]


class Ell(ctypes.Structure):
    _fields_ = [("n_rows", ctypes.c_int64), ("n_cols", ctypes.c_int64),
                ("width", ctypes.c_int64), ("stride", ctypes.c_int64),
                ("col_idxs", ctypes.c_void_p), ("values", ctypes.c_void_p)]


class Criterion(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("baseline", ctypes.c_int32),
                ("max_iters", ctypes.c_int64), ("reduction_factor", ctypes.c_double)]


class Criteria(ctypes.Structure):
    _fields_ = [("num", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("c", Criterion * MAX_CRITERIA)]


class StopFlags(ctypes.Structure):
    _fields_ = [("iteration", ctypes.c_int64), ("all_stopped", ctypes.c_int32),
                ("one_changed", ctypes.c_int32)]


class CgState(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("x", ctypes.c_void_p), ("b", ctypes.c_void_p),
                ("r", ctypes.c_void_p), ("z", ctypes.c_void_p), ("p_a", ctypes.c_void_p),
                ("p_b", ctypes.c_void_p), ("q", ctypes.c_void_p),
                ("inv_diag", ctypes.c_void_p), ("scalars", ctypes.c_void_p),
                ("orig_tau", ctypes.c_void_p), ("status", ctypes.c_void_p),
                ("flags", ctypes.c_void_p), ("work", ctypes.c_void_p),
                ("issued", ctypes.c_int64)]


class BlockCgState(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("nrhs", ctypes.c_int64),
                ("x", ctypes.c_void_p), ("x_stride", ctypes.c_int64),
                ("b", ctypes.c_void_p), ("b_stride", ctypes.c_int64),
                ("r", ctypes.c_void_p), ("z", ctypes.c_void_p), ("p_a", ctypes.c_void_p),
                ("p_b", ctypes.c_void_p), ("q", ctypes.c_void_p),
                ("inv_diag", ctypes.c_void_p), ("scalars", ctypes.c_void_p),
                ("orig_tau", ctypes.c_void_p), ("status", ctypes.c_void_p),
                ("flags", ctypes.c_void_p), ("work", ctypes.c_void_p),
                ("issued", ctypes.c_int64)]


class BicgstabState(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("x", ctypes.c_void_p), ("b", ctypes.c_void_p),
                ("r", ctypes.c_void_p), ("rr", ctypes.c_void_p), ("p", ctypes.c_void_p),
                ("v", ctypes.c_void_p), ("s", ctypes.c_void_p), ("t", ctypes.c_void_p),
                ("y", ctypes.c_void_p), ("z", ctypes.c_void_p),
                ("inv_diag", ctypes.c_void_p), ("scalars", ctypes.c_void_p),
                ("orig_tau", ctypes.c_void_p), ("status", ctypes.c_void_p),
                ("flags", ctypes.c_void_p), ("work", ctypes.c_void_p),
                ("issued", ctypes.c_int64)]


BATCH_MAX_BINS = 8


class BatchCsr(ctypes.Structure):
    _fields_ = [("num_systems", ctypes.c_int64), ("max_rows", ctypes.c_int64),
                ("max_nnz", ctypes.c_int64), ("row_offsets", ctypes.c_void_p),
                ("row_ptrs", ctypes.c_void_p), ("col_idxs", ctypes.c_void_p),
                ("values", ctypes.c_void_p), ("uniform_rows", ctypes.c_int64),
                ("uniform_nnz", ctypes.c_int64), ("format", ctypes.c_int32),
                ("ell_width", ctypes.c_int32), ("uniform_cols", ctypes.c_int64),
                ("item_offsets", ctypes.c_void_p)]


BATCH_CSR, BATCH_ELL, BATCH_DENSE = 0, 1, 2


class BatchSettings(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int32), ("tolerance_type", ctypes.c_int32),
                ("tolerance", ctypes.c_double), ("preconditioner", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class BatchPlan(ctypes.Structure):
    _fields_ = [("num_bins", ctypes.c_int32), ("single_launch", ctypes.c_int32),
                ("order", ctypes.c_void_p),
                ("bin_begin", ctypes.c_int64 * (BATCH_MAX_BINS + 1)),
                ("bin_max_rows", ctypes.c_int64 * BATCH_MAX_BINS),
                ("bin_max_nnz", ctypes.c_int64 * BATCH_MAX_BINS),
                ("block_size", ctypes.c_int32 * BATCH_MAX_BINS),
                ("matrix_in_lds", ctypes.c_int32 * BATCH_MAX_BINS),
                ("blocks_per_cu", ctypes.c_int32 * BATCH_MAX_BINS),
                ("subwave", ctypes.c_int32 * BATCH_MAX_BINS),
                ("vectors", ctypes.c_int32 * BATCH_MAX_BINS),
                ("bin_max_row_nnz", ctypes.c_int64 * BATCH_MAX_BINS)]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_SZ = ctypes.c_size_t
_D = ctypes.c_double
_F = ctypes.c_float
_CTX = ctypes.POINTER(Ctx)

# name -> (restype, [argtypes]); value-typed pointers are void*
_SIGS = {
    "gkomi_last_error": (ctypes.c_char_p, []),
    "gkomi_version": (ctypes.c_char_p, []),
    "gkomi_get_device_count": (ctypes.c_int, [ctypes.POINTER(_I32)]),
    "gkomi_get_device_info": (ctypes.c_int, [_I32, ctypes.POINTER(DeviceInfo)]),
    "gkomi_stream_create": (ctypes.c_int, [_I32, ctypes.POINTER(_P)]),
    "gkomi_stream_destroy": (ctypes.c_int, [_I32, _P]),
    "gkomi_synchronize": (ctypes.c_int, [_CTX]),
    "gkomi_malloc": (ctypes.c_int, [_CTX, _SZ, ctypes.POINTER(_P)]),
    "gkomi_free": (ctypes.c_int, [_CTX, _P]),
    "gkomi_host_malloc": (ctypes.c_int, [_SZ, ctypes.POINTER(_P)]),
    "gkomi_host_free": (ctypes.c_int, [_P]),
    "gkomi_memcpy_async": (ctypes.c_int, [_CTX, _P, _P, _SZ]),
    "gkomi_memcpy_to_host": (ctypes.c_int, [_CTX, _P, _P, _SZ]),
    "gkomi_event_create": (ctypes.c_int, [_I32, ctypes.POINTER(_P)]),
    "gkomi_event_destroy": (ctypes.c_int, [_I32, _P]),
    "gkomi_event_record": (ctypes.c_int, [_CTX, _P]),
    "gkomi_event_synchronize": (ctypes.c_int, [_P]),
    "gkomi_event_elapsed_ms": (ctypes.c_int, [_P, _P, ctypes.POINTER(_F)]),
    "gkomi_fill_zero": (ctypes.c_int, [_CTX, _P, _SZ]),
    "gkomi_fill_f64": (ctypes.c_int, [_CTX, _I64, _D, _P]),
    "gkomi_fill_f32": (ctypes.c_int, [_CTX, _I64, _F, _P]),
    "gkomi_convert_f64_to_f32": (ctypes.c_int, [_CTX, _I64, _P, _P]),
    "gkomi_convert_f32_to_f64": (ctypes.c_int, [_CTX, _I64, _P, _P]),
    "gkomi_reduction_workspace_bytes": (_SZ, [_I64, _I64]),
    "gkomi_csr_lb_clac_size": (_I64, [_I64, _I64, _I32, _I32, _I32]),
    "gkomi_csr_lb_partition": (ctypes.c_int, [_CTX, _I64, _P, _I64, _I32, _P, _I64]),
    "gkomi_csr_max_row_nnz": (ctypes.c_int, [_CTX, _I64, _P, _P]),
    "gkomi_csr_strategy_work_bytes": (_I64, [_I32, _I64, _I64]),
    "gkomi_csr_col_offsets": (ctypes.c_int, [_CTX, ctypes.POINTER(Csr), _P, _P]),
    "gkomi_csr_panels_bytes": (_I64, [_I64, _I64, _I32, _I32]),
    "gkomi_sellp_col_offsets": (ctypes.c_int, [_CTX, ctypes.POINTER(Sellp), _P, _P]),
    "gkomi_csr_build_panels_f64": (ctypes.c_int, [_CTX, ctypes.POINTER(Csr), _I32, _P, _I64]),
    "gkomi_csr_build_panels_f32": (ctypes.c_int, [_CTX, ctypes.POINTER(Csr), _I32, _P, _I64]),
    "gkomi_csr_transpose_workspace_bytes": (_SZ, [_I64, _I64, _I64]),
    "gkomi_scan_workspace_bytes": (_SZ, [_I64]),
    "gkomi_sellp_compute_slice_sets": (ctypes.c_int, [_CTX, _I64, _P, _I64, _I64, _P, _P, _P,
                                                      _SZ]),
    "gkomi_cg_work_bytes": (_SZ, [_I64]),
    "gkomi_cg_current_p": (_P, [ctypes.POINTER(CgState)]),
    "gkomi_block_cg_work_bytes": (_SZ, [_I64, _I64]),
    "gkomi_block_cg_current_p": (_P, [ctypes.POINTER(BlockCgState)]),
    "gkomi_bicgstab_work_bytes": (_SZ, [_I64]),
    "gkomi_batch_workspace_bytes": (_SZ, [ctypes.POINTER(BatchCsr)]),
    "gkomi_batch_lds_bytes": (_SZ, [_I32, _I64, _I64, _I32]),
    "gkomi_batch_cg_lds_bytes": (_SZ, [_I32, _I64, _I64, _I32]),
}

_SIGS["gkomi_batch_bicgstab_solve_noalloc_f64"] = (
    ctypes.c_int, [_CTX, ctypes.POINTER(BatchCsr), ctypes.POINTER(BatchSettings),
                   ctypes.POINTER(BatchPlan), _P, _P, _P, _P])

for _f, _d in (("csr", Csr), ("sellp", Sellp), ("ell", Ell)):
    _SIGS[f"gkomi_{_f}_spmv_mixed"] = (ctypes.c_int, [_CTX, _I32, ctypes.POINTER(_d), _I64,
                                                      _I32, _P, _I64, _I32, _P, _I64])
    _SIGS[f"gkomi_{_f}_advanced_spmv_mixed"] = (ctypes.c_int, [_CTX, _P, _I32,
                                                               ctypes.POINTER(_d), _I64, _I32,
                                                               _P, _I64, _P, _I32, _P, _I64])

for _s in ("f64", "f32"):
    _SIGS.update({
        f"gkomi_dense_dot_{_s}": (ctypes.c_int, [_CTX, _I64, _I64, _P, _I64, _P, _I64, _P, _P,
                                                 _SZ]),
        f"gkomi_dense_norm2_{_s}": (ctypes.c_int, [_CTX, _I64, _I64, _P, _I64, _P, _P, _SZ]),
        f"gkomi_dense_add_scaled_{_s}": (ctypes.c_int, [_CTX, _I64, _I64, _P, _I64, _P, _I64,
                                                        _P, _I64]),
        f"gkomi_dense_scale_{_s}": (ctypes.c_int, [_CTX, _I64, _I64, _P, _I64, _P, _I64]),
        f"gkomi_dense_copy_{_s}": (ctypes.c_int, [_CTX, _I64, _I64, _P, _I64, _P, _I64]),
        f"gkomi_csr_spmv_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(Csr), _I64, _P, _I64, _P,
                                                _I64]),
        f"gkomi_csr_advanced_spmv_{_s}": (ctypes.c_int, [_CTX, _P, ctypes.POINTER(Csr), _I64,
                                                         _P, _I64, _P, _P, _I64]),
        f"gkomi_csr_transpose_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(Csr), _P, _P, _P, _P,
                                                     _SZ]),
        f"gkomi_csr_to_sellp_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(Csr), _I64, _P, _P, _P,
                                                    _P]),
        f"gkomi_sellp_spmv_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(Sellp), _I64, _P, _I64,
                                                  _P, _I64]),
        f"gkomi_sellp_advanced_spmv_{_s}": (ctypes.c_int, [_CTX, _P, ctypes.POINTER(Sellp),
                                                           _I64, _P, _I64, _P, _P, _I64]),
        f"gkomi_csr_to_ell_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(Csr), _I64, _I64, _P,
                                                  _P]),
        f"gkomi_ell_spmv_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(Ell), _I64, _P, _I64, _P,
                                                _I64]),
        f"gkomi_ell_advanced_spmv_{_s}": (ctypes.c_int, [_CTX, _P, ctypes.POINTER(Ell), _I64,
                                                         _P, _I64, _P, _P, _I64]),
        f"gkomi_jacobi_generate_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(Csr), _P]),
        f"gkomi_jacobi_apply_{_s}": (ctypes.c_int, [_CTX, _I64, _I64, _P, _P, _I64, _P, _I64]),
        f"gkomi_jacobi_advanced_apply_{_s}": (ctypes.c_int, [_CTX, _I64, _I64, _P, _P, _P, _I64,
                                                             _P, _P, _I64]),
        f"gkomi_criteria_check_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(Criteria), _I64, _P,
                                                      _P, _P, _I32, _P, _P]),
        f"gkomi_cg_initialize_{_s}": (ctypes.c_int, [_CTX, _I64, _I64, _P, _I64, _P, _P, _P, _P,
                                                     _I64, _P, _P, _P]),
        f"gkomi_cg_step_1_{_s}": (ctypes.c_int, [_CTX, _I64, _I64, _P, _P, _I64, _P, _P, _P]),
        f"gkomi_cg_step_2_{_s}": (ctypes.c_int, [_CTX, _I64, _I64, _P, _I64, _P, _P, _P, _I64,
                                                 _P, _P, _P]),
        f"gkomi_cg_fused_start_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(Csr),
                                                      ctypes.POINTER(Criteria),
                                                      ctypes.POINTER(CgState)]),
        f"gkomi_cg_fused_iterate_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(Csr),
                                                        ctypes.POINTER(Criteria),
                                                        ctypes.POINTER(CgState), _I32]),
        f"gkomi_block_cg_start_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(Csr),
                                                      ctypes.POINTER(Criteria),
                                                      ctypes.POINTER(BlockCgState)]),
        f"gkomi_block_cg_iterate_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(Csr),
                                                        ctypes.POINTER(Criteria),
                                                        ctypes.POINTER(BlockCgState), _I32]),
        f"gkomi_bicgstab_fused_start_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(Csr),
                                                            ctypes.POINTER(Criteria),
                                                            ctypes.POINTER(BicgstabState)]),
        f"gkomi_bicgstab_fused_iterate_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(Csr),
                                                              ctypes.POINTER(Criteria),
                                                              ctypes.POINTER(BicgstabState),
                                                              _I32]),
        f"gkomi_bicgstab_initialize_{_s}": (ctypes.c_int, [_CTX, _I64, _I64, _P, _I64] +
                                            [_P] * 8 + [_I64] + [_P] * 7),
        f"gkomi_bicgstab_step_1_{_s}": (ctypes.c_int, [_CTX, _I64, _I64, _P, _P, _P, _I64, _P,
                                                       _P, _P, _P, _P]),
        f"gkomi_bicgstab_step_2_{_s}": (ctypes.c_int, [_CTX, _I64, _I64, _P, _P, _P, _I64, _P,
                                                       _P, _P, _P]),
        f"gkomi_bicgstab_step_3_{_s}": (ctypes.c_int, [_CTX, _I64, _I64, _P, _I64, _P, _P, _P,
                                                       _P, _P, _I64, _P, _P, _P, _P, _P]),
        f"gkomi_bicgstab_finalize_{_s}": (ctypes.c_int, [_CTX, _I64, _I64, _P, _I64, _P, _I64,
                                                         _P, _P]),
        f"gkomi_csr_sampled_outer_product_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(Csr),
                                                                _I64, _D if _s == "f64" else _F,
                                                                _P, _I64, _P, _I64, _P]),
        f"gkomi_batch_bicgstab_solve_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(BatchCsr),
                                                           ctypes.POINTER(BatchSettings),
                                                           ctypes.POINTER(BatchPlan), _P, _P, _P,
                                                           _P, _P, _SZ]),
        f"gkomi_batch_bicgstab_tune_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(BatchCsr),
                                                          ctypes.POINTER(BatchSettings),
                                                          ctypes.POINTER(BatchPlan), _P, _P, _P,
                                                          _P, _P, _P, _SZ]),
        f"gkomi_batch_spmv_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(BatchCsr), _P, _P, _P,
                                                   _P]),
        f"gkomi_batch_cg_solve_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(BatchCsr),
                                                     ctypes.POINTER(BatchSettings),
                                                     ctypes.POINTER(BatchPlan), _P, _P, _P, _P,
                                                     _P, _SZ]),
        f"gkomi_batch_cg_tune_{_s}": (ctypes.c_int, [_CTX, ctypes.POINTER(BatchCsr),
                                                    ctypes.POINTER(BatchSettings),
                                                    ctypes.POINTER(BatchPlan), _P, _P, _P, _P,
                                                    _P, _P, _SZ]),
    })

_lib = None


def lib():
    """The loaded kernel library.  Raises ImportError if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m ginkgo_amd.build`; "
                "this package has no CPU fallback")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def declared_symbols():
    """Names of every function declared in include/gkomi.h."""
    import re
    hdr = os.path.join(os.path.dirname(_HERE), "include", "gkomi.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"\b(gkomi_[a-z0-9_]+)\s*\(", text)))


def check(rc, what=""):
    if rc != SUCCESS:
        msg = lib().gkomi_last_error().decode(errors="replace")
        raise _ERRORS.get(rc, GkoError)(f"{what}: {msg}" if what else msg)
    return rc


def call(name, *args):
    """Call an int-returning entry point and raise on failure."""
    return check(getattr(lib(), name)(*args), name)


def criteria(spec):
    """Criteria from a list like [{'kind': 'iteration', 'max_iters': 100},
    {'kind': 'residual_norm', 'reduction_factor': 1e-10, 'baseline': 'rhs_norm'}]."""
    kinds = {"iteration": CRIT_ITERATION, "residual_norm": CRIT_RESIDUAL_NORM,
             "implicit_residual_norm": CRIT_IMPLICIT_RESNORM}
    bases = {"rhs_norm": BASELINE_RHS_NORM, "initial_resnorm": BASELINE_INITIAL_RESNORM,
             "absolute": BASELINE_ABSOLUTE}
    if not 1 <= len(spec) <= MAX_CRITERIA:
        raise NotSupported("between 1 and 4 stopping criteria are supported")
    c = Criteria()
    c.num = len(spec)
    for i, s in enumerate(spec):
        c.c[i].kind = kinds[s["kind"]]
        c.c[i].max_iters = int(s.get("max_iters", 0))
        c.c[i].baseline = bases[s.get("baseline", "rhs_norm")]
        c.c[i].reduction_factor = float(s.get("reduction_factor", 0.0))
    return c
