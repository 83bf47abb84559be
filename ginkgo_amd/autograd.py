"""Differentiable sparse solve (no counterpart in the reference).

    x = sparse_solve(values, b, row_ptrs, col_idxs, method="cg", tol=1e-12)

solves A(values) x = b on the MI355X with the gko:: solvers and makes the
result differentiable with respect to the matrix values and b:

  forward : x = A^{-1} b                      (Cg for SPD A, Bicgstab otherwise)
  backward: A^T lam = dL/dx                   (same solver; A^T built on the
                                               device by Csr::transpose)
            dL/db = lam
            dL/da_ij = -sum_k lam_ik x_jk     (sampled outer product on the
                                               pattern of A, a gfx950 kernel)

values is the CSR value array (nnz,), b is (n,) or (n, k).  Gradients flow
on to whatever produced `values` (e.g. per-edge conductivities scattered
into the matrix with torch ops).

Both solves must converge: a column whose stopping status lacks the
converged bit (include/ginkgo/core/stop/stopping_status.hpp:21-100), e.g.
one stopped by the iteration limit, raises SolveNotConverged (or warns, or
passes, as `on_not_converged` says), since its x and every gradient through
it would be wrong.
"""
import ctypes
import warnings

import torch

from . import _capi as C
from . import device as D
from .device import as_2d

_EXECS = {}

# iteration counts of the most recent forward and adjoint solves
last_solve_info = {}


class SolveNotConverged(C.GkoError):
    """A forward or adjoint solve stopped without meeting its residual
    criterion (for instance on the iteration limit)."""


def _check_converged(solver, phase, tol, max_iters, policy):
    status = list(solver.last_status)
    bad = [j for j, st in enumerate(status) if not (st & C.STATUS_CONVERGED)]
    if not bad or policy == "ignore":
        return
    msg = (f"{phase} solve did not converge to relative residual {tol:g} within "
           f"{max_iters} iterations ({solver.last_iterations} run) in column(s) "
           f"{bad[:8]}{'...' if len(bad) > 8 else ''} of {len(status)}")
    if policy == "warn":
        warnings.warn(msg, RuntimeWarning, stacklevel=4)
    else:
        raise SolveNotConverged(msg)


def executor(device=None):
    """One gko::Mi355xExecutor per (device, torch stream), sharing torch's
    current stream so torch work and solver work stay ordered."""
    from . import gko
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else device
    s = torch.cuda.current_stream(dev)
    key = (dev.index, s.cuda_stream)
    if key not in _EXECS:
        _EXECS[key] = gko.Mi355xExecutor.create(dev.index or 0, s.cuda_stream)
    return _EXECS[key]


def _criteria(gko, dt, tol, max_iters):
    return (gko.stop.Iteration.build().with_max_iters(int(max_iters)),
            gko.stop.ResidualNorm.build(dt).with_reduction_factor(float(tol)))


def _solve(gko, exec_, A, rhs, method, tol, max_iters, jacobi):
    dt = D.suffix(rhs.dtype)
    build = gko.solver.Cg.build if method == "cg" else gko.solver.Bicgstab.build
    params = build(dt).with_criteria(*_criteria(gko, dt, tol, max_iters))
    if jacobi:
        params = params.with_preconditioner(
            gko.preconditioner.Jacobi.build(dt).with_max_block_size(1))
    solver = params.on(exec_).generate(A)
    x = torch.zeros_like(rhs)
    solver.apply(gko.dense_view(exec_, rhs), gko.dense_view(exec_, x))
    return x, solver


def sampled_outer_product(row_ptrs, col_idxs, values, lam, x, alpha=-1.0):
    """out[k] = alpha * sum_j lam[row(k), j] * x[col(k), j] on the CSR pattern."""
    lam2 = as_2d(lam).contiguous()
    x2 = as_2d(x).contiguous()
    out = torch.empty_like(values)
    d = C.Csr()
    d.n_rows, d.n_cols, d.nnz = row_ptrs.numel() - 1, x2.shape[0], values.numel()
    d.row_ptrs, d.col_idxs, d.values = row_ptrs.data_ptr(), col_idxs.data_ptr(), values.data_ptr()
    d.strategy = C.CSR_CLASSICAL
    c = D.ctx(values.device)
    C.call(f"gkomi_csr_sampled_outer_product_{D.suffix(values.dtype)}", ctypes.byref(c),
           ctypes.byref(d), lam2.shape[1], alpha, D.ptr(lam2), lam2.stride(0), D.ptr(x2),
           x2.stride(0), D.ptr(out))
    return out


class SparseSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, row_ptrs, col_idxs, method, tol, max_iters, jacobi, policy):
        from . import gko
        if method not in ("cg", "bicgstab"):
            raise C.NotSupported("method must be 'cg' (SPD) or 'bicgstab'")
        if policy not in ("raise", "warn", "ignore"):
            raise C.InvalidArgument("on_not_converged must be 'raise', 'warn' or 'ignore'")
        exec_ = executor(values.device)
        vals = values.detach().contiguous()
        rhs = b.detach().reshape(b.shape[0], -1).contiguous()
        rp = row_ptrs.to(torch.int32).contiguous()
        ci = col_idxs.to(torch.int32).contiguous()
        A = gko.csr_view(exec_, rp, ci, vals, strategy="classical")
        x, solver = _solve(gko, exec_, A, rhs, method, tol, max_iters, jacobi)
        _check_converged(solver, "forward", tol, max_iters, policy)
        last_solve_info["forward_iterations"] = int(solver.last_iterations)
        ctx.save_for_backward(vals, x, rp, ci)
        ctx.settings = (method, tol, max_iters, jacobi, b.dim(), policy)
        ctx.A = A
        ctx.iterations = solver.last_iterations
        return x.reshape(b.shape)

    @staticmethod
    def backward(ctx, grad_x):
        from . import gko
        vals, x, rp, ci = ctx.saved_tensors
        method, tol, max_iters, jacobi, bdim, policy = ctx.settings
        exec_ = executor(vals.device)
        g = grad_x.reshape(x.shape).contiguous()
        # adjoint system: A is symmetric for CG; otherwise transpose on the device
        At = ctx.A if method == "cg" else ctx.A.transpose()
        lam, adjoint = _solve(gko, exec_, At, g, method, tol, max_iters, jacobi)
        _check_converged(adjoint, "adjoint", tol, max_iters, policy)
        last_solve_info["adjoint_iterations"] = int(adjoint.last_iterations)
        grad_vals = None
        if ctx.needs_input_grad[0]:
            grad_vals = sampled_outer_product(rp, ci, vals, lam, x, alpha=-1.0)
        grad_b = lam.reshape(grad_x.shape) if ctx.needs_input_grad[1] else None
        return grad_vals, grad_b, None, None, None, None, None, None, None


def sparse_solve(values, b, row_ptrs, col_idxs, method="cg", tol=1e-12, max_iters=10000,
                 jacobi=False, on_not_converged="raise"):
    """Differentiable x = A(values)^{-1} b (see the module docstring).
    on_not_converged: 'raise' (SolveNotConverged), 'warn' or 'ignore' when a
    forward or adjoint column stops without meeting the criterion."""
    return SparseSolve.apply(values, b, row_ptrs, col_idxs, method, tol, max_iters, jacobi,
                             on_not_converged)
