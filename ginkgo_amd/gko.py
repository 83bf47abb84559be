"""The gko:: host library from Python, with the reference's names.

    import ginkgo_amd.gko as gko
    exec = gko.Mi355xExecutor.create(0, torch.cuda.current_stream().cuda_stream)
    A = gko.csr_view(exec, row_ptrs, col_idxs, values)          # torch device tensors
    solver = (gko.solver.Cg.build("f64")
              .with_criteria(gko.stop.Iteration.build().with_max_iters(400),
                             gko.stop.ResidualNorm.build("f64").with_reduction_factor(1e-10))
              .with_preconditioner(gko.preconditioner.Jacobi.build("f64").with_max_block_size(1))
              .on(exec).generate(A))
    solver.apply(gko.dense_view(exec, b), gko.dense_view(exec, x))

Every call runs in the C++ library (libginkgo_amd.so) and its kernels in
libgkomi.so; this module only names things.
"""
import types

from . import _gko as _g

Mi355xExecutor = _g.Mi355xExecutor
HostExecutor = _g.HostExecutor
LinOp = _g.LinOp
NotSupported = _g.NotSupported
DimensionMismatch = _g.DimensionMismatch
NotImplemented_ = _g.NotImplemented
AllocationError = _g.AllocationError
HipError = _g.HipError
GkoError = _g.GkoError

_SFX = {"f64": "_f64", "f32": "_f32", "float64": "_f64", "float32": "_f32",
        "double": "_f64", "float": "_f32"}


def _sfx(dtype):
    key = str(dtype).replace("torch.", "").replace("<class 'numpy.", "").replace("'>", "")
    try:
        return _SFX[key]
    except KeyError:
        raise NotSupported(f"value type {dtype} is not supported (float64 / float32)")


def _typed(name):
    return lambda dtype="f64": getattr(_g, name + _sfx(dtype))


matrix = types.SimpleNamespace(
    Dense=_typed("Dense"), Csr=_typed("Csr"), Sellp=_typed("Sellp"), Ell=_typed("Ell"),
    Coo=_typed("Coo"), Hybrid=_typed("Hybrid"))

stop = types.SimpleNamespace(
    Iteration=types.SimpleNamespace(build=lambda: _g.Iteration()),
    ResidualNorm=types.SimpleNamespace(build=lambda dtype="f64": _typed("ResidualNorm")(dtype)()),
    ImplicitResidualNorm=types.SimpleNamespace(
        build=lambda dtype="f64": _typed("ImplicitResidualNorm")(dtype)()),
    mode=_g.mode,
)

preconditioner = types.SimpleNamespace(
    Jacobi=types.SimpleNamespace(build=lambda dtype="f64": _typed("JacobiParams")(dtype)()))

solver = types.SimpleNamespace(
    Cg=types.SimpleNamespace(build=lambda dtype="f64": _typed("CgParams")(dtype)()),
    Bicgstab=types.SimpleNamespace(build=lambda dtype="f64": _typed("BicgstabParams")(dtype)()),
    Ir=types.SimpleNamespace(build=lambda dtype="f64": _typed("IrParams")(dtype)()),
)

log = types.SimpleNamespace(
    Convergence=_g.Convergence, ProfilerHook=_g.ProfilerHook,
    TableSummaryWriter=_g.TableSummaryWriter, profile_event_category=_g.profile_event_category,
    summary_entry=_g.summary_entry, nested_summary_entry=_g.nested_summary_entry)
Timer = _g.Timer
CpuTimer = _g.CpuTimer
Mi355xTimer = _g.Mi355xTimer

# batch module: batch.matrix.Csr (shared pattern), Ell, Dense, BlockDiagCsr /
# BlockDiagEll / BlockDiagDense
# (items of different sizes), batch.MultiVector, batch.solver.Bicgstab and Cg,
# batch.preconditioner.Jacobi, batch.log.BatchConvergence, batch.stop.tolerance_type
batch = types.SimpleNamespace(
    MultiVector=_typed("BatchMultiVector"),
    matrix=types.SimpleNamespace(Csr=_typed("BatchCsr"), Ell=_typed("BatchEll"),
                                 Dense=_typed("BatchDense"),
                                 BlockDiagCsr=_typed("BatchBlockDiagCsr"),
                                 BlockDiagEll=_typed("BatchBlockDiagEll"),
                                 BlockDiagDense=_typed("BatchBlockDiagDense")),
    stop=types.SimpleNamespace(tolerance_type=_g.batch_tolerance_type),
    preconditioner=types.SimpleNamespace(
        Jacobi=types.SimpleNamespace(build=lambda dtype="f64": _typed("BatchJacobiParams")(dtype)())),
    solver=types.SimpleNamespace(
        Bicgstab=types.SimpleNamespace(build=lambda dtype="f64": _typed("BatchBicgstabParams")(dtype)()),
        Cg=types.SimpleNamespace(build=lambda dtype="f64": _typed("BatchCgParams")(dtype)())),
    log=types.SimpleNamespace(BatchConvergence=_typed("BatchConvergence")),
)


def dense_view(exec, t):
    """Non-owning Dense view of a 1-D or 2-D device tensor (kept alive by the caller)."""
    t2 = t.reshape(t.shape[0], 1) if t.dim() == 1 else t
    if t2.stride(-1) != 1:
        raise NotSupported("Dense views need unit column stride")
    cls = getattr(_g, "Dense" + _sfx(t.dtype))
    return cls.create_view(exec, t2.shape[0], t2.shape[1], t2.data_ptr(), t2.stride(0))


def csr_view(exec, row_ptrs, col_idxs, values, n_cols=None, strategy="automatical", subwave=0):
    """Non-owning Csr view of device tensors (int32 indices)."""
    n = row_ptrs.numel() - 1
    cls = getattr(_g, "Csr" + _sfx(values.dtype))
    return cls.create_view(exec, n, n if n_cols is None else n_cols, values.numel(),
                           values.data_ptr(), col_idxs.data_ptr(), row_ptrs.data_ptr(),
                           strategy, subwave)


def csr_from_host(exec, row_ptrs, col_idxs, values, n_cols=None, strategy="automatical"):
    """Csr copied from host (numpy) arrays."""
    n = len(row_ptrs) - 1
    cls = getattr(_g, "Csr" + _sfx(values.dtype))
    return cls.create_from_host(exec, n, n if n_cols is None else n_cols, row_ptrs, col_idxs,
                                values, strategy)
