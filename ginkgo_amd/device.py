"""Thin PyTorch-ROCm plumbing over the C ABI (device memory and streams).

torch provides device allocations and the stream; every computation is a
call into libgkomi.so.  Matrices are held as torch tensors plus the plain C
descriptor the C ABI takes.
"""
import ctypes

import numpy as np
import torch

from . import _capi as C

_VT = {torch.float64: "f64", torch.float32: "f32"}


def value_code(dtype):
    """GKOMI_VALUE_F64 / GKOMI_VALUE_F32 of a torch dtype."""
    suffix(dtype)
    return C.VALUE_F64 if dtype == torch.float64 else C.VALUE_F32


def suffix(dtype):
    try:
        return _VT[dtype]
    except KeyError:
        raise C.NotSupported(f"value type {dtype} is not supported (float64/float32 only)")


def ctx(device=None, stream=None):
    """gkomi_ctx for torch's current (or given) device and stream."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    s = torch.cuda.current_stream(dev) if stream is None else stream
    c = C.Ctx()
    c.device = dev.index if dev.index is not None else torch.cuda.current_device()
    c.stream = ctypes.c_void_p(s.cuda_stream)
    return c


def as_2d(t):
    """A 1-D vector as one column, a 2-D block unchanged (also when empty)."""
    return t.reshape(t.shape[0], 1) if t.dim() == 1 else t.reshape(t.shape[0], t.shape[1])


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else ctypes.c_void_p(0)


def workspace(nbytes, device):
    """Zero-filled device scratch of at least one byte."""
    return torch.zeros(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def scalar(value, dtype, device):
    return torch.full((1,), float(value), dtype=dtype, device=device)


def check_operand(t, name, dtype=None, device=None, rows=None, cols=None):
    """Rejects, before any launch, a dense operand the C ABI cannot take as
    (pointer, row stride): not a CUDA tensor of `device`, another value type
    than `dtype`, more than two dimensions, columns that are not unit-stride,
    overlapping rows, or rows/columns other than expected."""
    if not torch.is_tensor(t):
        raise C.InvalidArgument(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.device.type != "cuda" or (device is not None and t.device != torch.device(device)):
        raise C.InvalidArgument(f"{name}: tensor on {t.device}, expected {device}")
    if dtype is not None and t.dtype != dtype:
        raise C.NotSupported(f"{name}: value type {t.dtype}, expected {dtype}")
    if t.dim() not in (1, 2):
        raise C.DimensionMismatch(f"{name}: expected a vector or a 2-D block, got {t.dim()}-D")
    if t.dim() == 1 and t.numel() > 1 and t.stride(0) != 1:
        raise C.NotSupported(f"{name}: a vector must be contiguous (stride {t.stride(0)})")
    if t.dim() == 2:
        if t.shape[1] > 1 and t.stride(1) != 1:
            raise C.NotSupported(f"{name}: columns must be unit-stride (stride {t.stride(1)})")
        if t.shape[0] > 1 and t.stride(0) < t.shape[1]:
            raise C.NotSupported(f"{name}: rows overlap (row stride {t.stride(0)})")
    if rows is not None and t.shape[0] != rows:
        raise C.DimensionMismatch(f"{name}: {t.shape[0]} rows, expected {rows}")
    if cols is not None and (t.shape[1] if t.dim() == 2 else 1) != cols:
        raise C.DimensionMismatch(f"{name}: wrong number of columns, expected {cols}")
    return t


def _cols(t):
    return t.shape[1] if t.dim() == 2 else 1


def _check_scalar(s, name, dtype, device):
    check_operand(s, name, dtype, device)
    if s.numel() < 1:
        raise C.DimensionMismatch(f"{name}: empty scalar")
    return s


def default_lb_warps(device):
    """The HipExecutor's warp count, compute units x 4 (hip/base/executor.hip.cpp:234-238)."""
    dev = torch.device(device)
    info = C.DeviceInfo()
    C.call("gkomi_get_device_info", dev.index if dev.index is not None else 0, ctypes.byref(info))
    return int(info.num_cus) * 4


class DeviceCsr:
    """Csr<V, int32> on the device with its SpMV strategy.

    strategy: 'classical', 'load_balance', 'merge_path' or 'automatical'
    (picks classical for short, uniform rows and load_balance otherwise, as
    the reference's automatical::process does from row statistics,
    csr.hpp:606-673)."""

    def __init__(self, row_ptrs, col_idxs, values, n_cols=None, strategy="automatical",
                 subwave=0, device="cuda", lb_nwarps=None, lb_warp_size=64,
                 lb_cuda_strategy=False):
        dev = torch.device(device)
        self.row_ptrs = torch.as_tensor(np.asarray(row_ptrs, dtype=np.int32)).to(dev) \
            if not torch.is_tensor(row_ptrs) else row_ptrs.to(dev, torch.int32)
        self.col_idxs = torch.as_tensor(np.asarray(col_idxs, dtype=np.int32)).to(dev) \
            if not torch.is_tensor(col_idxs) else col_idxs.to(dev, torch.int32)
        self.values = torch.as_tensor(np.asarray(values)).to(dev) \
            if not torch.is_tensor(values) else values.to(dev)
        self.dtype = self.values.dtype
        suffix(self.dtype)
        self.n_rows = self.row_ptrs.numel() - 1
        self.n_cols = self.n_rows if n_cols is None else int(n_cols)
        self.nnz = self.col_idxs.numel()
        self.device = self.values.device
        dev = self.device
        rp = self.row_ptrs.cpu().numpy() if strategy == "automatical" else None
        if strategy == "automatical":
            # the C++ Csr's automatical rule: classical up to the reference's
            # AMD row limit of 768 entries when the longest row is within 8x
            # the mean plus 64, load_balance otherwise
            lens = np.diff(rp) if self.n_rows > 0 else np.zeros(1)
            mx = int(lens.max()) if lens.size else 0
            mean = float(lens.mean()) if lens.size else 0.0
            strategy = "classical" if (mx <= 768 and mx <= 8 * mean + 64) else "load_balance"
        self.strategy = strategy
        self.subwave = int(subwave)
        self.srow = None
        self.num_srow = 0
        if strategy == "load_balance":
            # Csr::load_balance with the executor's warps (HipExecutor form:
            # warp size 64, no cuda_strategy) unless given
            nw = default_lb_warps(dev) if lb_nwarps is None else int(lb_nwarps)
            self.lb = (nw, int(lb_warp_size), bool(lb_cuda_strategy))
            self.num_srow = int(C.lib().gkomi_csr_lb_clac_size(
                self.nnz, nw, int(lb_warp_size), int(bool(lb_cuda_strategy)), 0))
            self.srow = torch.empty(max(self.num_srow, 1), dtype=torch.int32, device=dev)
            c = ctx(dev)
            C.call("gkomi_csr_lb_partition", ctypes.byref(c), self.n_rows, ptr(self.row_ptrs),
                   self.nnz, int(lb_warp_size), ptr(self.srow), self.num_srow)
        elif strategy == "merge_path":
            wb = int(C.lib().gkomi_csr_strategy_work_bytes(C.CSR_MERGE_PATH, self.n_rows,
                                                           self.nnz))
            self.work = workspace(wb, dev)
        elif strategy != "classical":
            raise C.InvalidArgument(f"unknown strategy {strategy}")
        self.col_offsets = None
        self.desc = self._desc()

    def compact(self):
        """Builds the 2-byte column offsets (col - row) the classical SpMV and
        the fused CG read instead of the 4-byte indices; returns whether every
        offset fits int16 (else nothing changes).  Results are bit-identical."""
        off = torch.empty(max(self.nnz, 1), dtype=torch.int16, device=self.device)
        fits = torch.zeros(1, dtype=torch.int32, device=self.device)
        c = ctx(self.device)
        d = self._desc()
        d.col_offsets = None
        C.call("gkomi_csr_col_offsets", ctypes.byref(c), ctypes.byref(d), ptr(off), ptr(fits))
        if self.nnz > 0 and int(fits.item()) == 1:
            self.col_offsets = off
            self.desc = self._desc()
            return True
        return False

    def build_panels(self, num_panels=8):
        """Builds the column panels (gkomi_csr_build_panels_*): a copy of the
        matrix in num_panels (2, 4 or 8) column panels, one per XCD group, so
        that the gathers from b of a single right-hand-side SpMV stay within
        one XCD's L2 window; every such SpMV then uses them.  The copy is
        taken now: call again after changing the values.  Results differ
        from the unpanelled product only in the order of a row's sum (panel by
        panel), and are deterministic.  num_panels 0 removes them."""
        num_panels = int(num_panels)
        self.panel_buf, self.num_panels = None, 0
        if num_panels == 0:
            self.desc = self._desc()
            return
        nb = int(C.lib().gkomi_csr_panels_bytes(self.n_rows, self.nnz, num_panels,
                                                value_code(self.dtype)))
        if nb < 0:
            raise C.InvalidArgument(f"column panels: num_panels {num_panels} not 2, 4 or 8")
        buf = workspace(nb, self.device)
        c = ctx(self.device)
        d = self._desc()
        C.call(f"gkomi_csr_build_panels_{suffix(self.dtype)}", ctypes.byref(c), ctypes.byref(d),
               num_panels, ptr(buf), nb)
        self.panel_buf, self.num_panels = buf, num_panels
        self.desc = self._desc()

    def build_slices(self, compact=True):
        """Builds the sliced copy (SELL-P layout, 64-row slices; no reference
        counterpart) that the classical single right-hand-side SpMV and the
        fused CG then read: one lane per row, a row's entries summed in their
        stored order.  With compact, the slot columns are kept as 2-byte
        offsets col - row when they fit.  Returns whether they did.  The copy
        is taken now: call again after changing the values."""
        self.slices = None
        self.desc = self._desc()
        S = DeviceSellp(self, 64, 1)
        offs = None
        if compact and S.total > 0:
            o = torch.empty(S.total, dtype=torch.int16, device=self.device)
            fits = torch.zeros(1, dtype=torch.int32, device=self.device)
            c = ctx(self.device)
            C.call("gkomi_sellp_col_offsets", ctypes.byref(c), ctypes.byref(S.desc), ptr(o),
                   ptr(fits))
            if int(fits.item()) == 1:
                offs = o
        self.slices = (S, offs)
        self.desc = self._desc()
        return offs is not None

    def _desc(self):
        d = C.Csr()
        d.n_rows, d.n_cols, d.nnz = self.n_rows, self.n_cols, self.nnz
        d.row_ptrs = ptr(self.row_ptrs).value
        d.col_idxs = ptr(self.col_idxs).value
        d.values = ptr(self.values).value
        d.strategy = {"classical": C.CSR_CLASSICAL, "load_balance": C.CSR_LOAD_BALANCE,
                      "merge_path": C.CSR_MERGE_PATH}[self.strategy]
        d.subwave = self.subwave
        d.srow = ptr(self.srow).value if self.srow is not None else None
        d.num_srow = self.num_srow
        d.col_offsets = ptr(self.col_offsets).value if getattr(self, "col_offsets", None) \
            is not None else None
        work = getattr(self, "work", None)
        d.work = ptr(work).value if work is not None else None
        d.work_bytes = work.numel() if work is not None else 0
        buf = getattr(self, "panel_buf", None)
        d.panels = ptr(buf).value if buf is not None else None
        d.panels_bytes = buf.numel() if buf is not None else 0
        d.num_panels = getattr(self, "num_panels", 0) if buf is not None else 0
        sl = getattr(self, "slices", None)
        if sl is not None:
            S, offs = sl
            d.slice_sets = ptr(S.slice_sets).value
            d.slice_values = ptr(S.values).value
            d.slice_cols = ptr(S.col_idxs).value
            d.slice_offsets = ptr(offs).value if offs is not None else None
        return d

    def spmv(self, b, out=None, stream=None):
        """out = A b.  b and out may be float64 or float32 independently of
        the matrix: a mixed product computes in the highest precision
        (gkomi_csr_spmv_mixed), as the reference's Csr::apply does."""
        check_operand(b, "b", None, self.device, rows=self.n_cols)
        suffix(b.dtype)
        if out is None:
            out = torch.empty((self.n_rows, _cols(b)), dtype=b.dtype, device=self.device)
        check_operand(out, "out", None, self.device, rows=self.n_rows, cols=_cols(b))
        suffix(out.dtype)
        b2, o2 = as_2d(b), as_2d(out)
        c = ctx(self.device, stream)
        if b.dtype == self.dtype and out.dtype == self.dtype:
            C.call(f"gkomi_csr_spmv_{suffix(self.dtype)}", ctypes.byref(c),
                   ctypes.byref(self.desc), b2.shape[1], ptr(b2), b2.stride(0), ptr(o2),
                   o2.stride(0))
        else:
            C.call("gkomi_csr_spmv_mixed", ctypes.byref(c), value_code(self.dtype),
                   ctypes.byref(self.desc), b2.shape[1], value_code(b.dtype), ptr(b2),
                   b2.stride(0), value_code(out.dtype), ptr(o2), o2.stride(0))
        return out

    def advanced_spmv(self, alpha, b, beta, c_vec, stream=None):
        """c = alpha A b + beta c; alpha in the matrix's value type, beta in
        c's (mixed types as in spmv)."""
        check_operand(b, "b", None, self.device, rows=self.n_cols)
        check_operand(c_vec, "c", None, self.device, rows=self.n_rows, cols=_cols(b))
        suffix(b.dtype)
        suffix(c_vec.dtype)
        a = alpha if torch.is_tensor(alpha) else scalar(alpha, self.dtype, self.device)
        bt = beta if torch.is_tensor(beta) else scalar(beta, c_vec.dtype, self.device)
        _check_scalar(a, "alpha", self.dtype, self.device)
        _check_scalar(bt, "beta", c_vec.dtype, self.device)
        b2, c2 = as_2d(b), as_2d(c_vec)
        c = ctx(self.device, stream)
        if b.dtype == self.dtype and c_vec.dtype == self.dtype:
            C.call(f"gkomi_csr_advanced_spmv_{suffix(self.dtype)}", ctypes.byref(c), ptr(a),
                   ctypes.byref(self.desc), b2.shape[1], ptr(b2), b2.stride(0), ptr(bt),
                   ptr(c2), c2.stride(0))
        else:
            C.call("gkomi_csr_advanced_spmv_mixed", ctypes.byref(c), ptr(a),
                   value_code(self.dtype), ctypes.byref(self.desc), b2.shape[1],
                   value_code(b.dtype), ptr(b2), b2.stride(0), ptr(bt),
                   value_code(c_vec.dtype), ptr(c2), c2.stride(0))
        return c_vec

    def max_row_nnz(self):
        out = torch.zeros(1, dtype=torch.int64, device=self.device)
        c = ctx(self.device)
        C.call("gkomi_csr_max_row_nnz", ctypes.byref(c), self.n_rows, ptr(self.row_ptrs),
               ptr(out))
        return int(out.item())

    def transpose(self):
        t_rp = torch.empty(self.n_cols + 1, dtype=torch.int32, device=self.device)
        t_ci = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=self.device)
        t_v = torch.empty(max(self.nnz, 1), dtype=self.dtype, device=self.device)
        wsb = C.lib().gkomi_csr_transpose_workspace_bytes(self.n_rows, self.n_cols, self.nnz)
        ws = workspace(wsb, self.device)
        c = ctx(self.device)
        C.call(f"gkomi_csr_transpose_{suffix(self.dtype)}", ctypes.byref(c),
               ctypes.byref(self.desc), ptr(t_rp), ptr(t_ci), ptr(t_v), ptr(ws), wsb)
        return t_rp, t_ci[: self.nnz], t_v[: self.nnz]

    def jacobi(self):
        inv = torch.empty(self.n_rows, dtype=self.dtype, device=self.device)
        c = ctx(self.device)
        C.call(f"gkomi_jacobi_generate_{suffix(self.dtype)}", ctypes.byref(c),
               ctypes.byref(self.desc), ptr(inv))
        return inv


class DeviceSellp:
    """Sellp<V, int32> built on the device from a DeviceCsr."""

    def __init__(self, csr, slice_size=64, stride_factor=1):
        self.n_rows, self.n_cols = csr.n_rows, csr.n_cols
        self.dtype, self.device = csr.dtype, csr.device
        self.slice_size, self.stride_factor = int(slice_size), int(stride_factor)
        ns = -(-self.n_rows // self.slice_size)
        self.num_slices = ns
        self.slice_sets = torch.empty(ns + 1, dtype=torch.int64, device=self.device)
        self.slice_lengths = torch.empty(max(ns, 1), dtype=torch.int64, device=self.device)
        wsb = C.lib().gkomi_scan_workspace_bytes(ns)
        ws = workspace(wsb, self.device)
        c = ctx(self.device)
        C.call("gkomi_sellp_compute_slice_sets", ctypes.byref(c), self.n_rows,
               ptr(csr.row_ptrs), self.slice_size, self.stride_factor, ptr(self.slice_sets),
               ptr(self.slice_lengths), ptr(ws), wsb)
        total = int(self.slice_sets[-1].item()) * self.slice_size
        self.col_idxs = torch.empty(max(total, 1), dtype=torch.int32, device=self.device)
        self.values = torch.empty(max(total, 1), dtype=self.dtype, device=self.device)
        C.call(f"gkomi_csr_to_sellp_{suffix(self.dtype)}", ctypes.byref(c),
               ctypes.byref(csr.desc), self.slice_size, ptr(self.slice_sets),
               ptr(self.slice_lengths), ptr(self.col_idxs), ptr(self.values))
        self.total = total
        d = C.Sellp()
        d.n_rows, d.n_cols = self.n_rows, self.n_cols
        d.slice_size, d.stride_factor = self.slice_size, self.stride_factor
        d.slice_sets = ptr(self.slice_sets).value
        d.slice_lengths = ptr(self.slice_lengths).value
        d.col_idxs = ptr(self.col_idxs).value
        d.values = ptr(self.values).value
        self.desc = d

    def spmv(self, b, out=None, stream=None):
        """out = A b, mixed value types as DeviceCsr.spmv."""
        return _format_spmv(self, "sellp", b, out, stream)

    def advanced_spmv(self, alpha, b, beta, c_vec, stream=None):
        """c = alpha A b + beta c (beta == 0 ignores c, NaN included)."""
        return _format_advanced_spmv(self, "sellp", alpha, b, beta, c_vec, stream)


class DeviceEll:
    """Ell<V, int32> (column-major, stride = n_rows) built on the device."""

    def __init__(self, csr, width=None, stride=None):
        self.n_rows, self.n_cols = csr.n_rows, csr.n_cols
        self.dtype, self.device = csr.dtype, csr.device
        self.width = csr.max_row_nnz() if width is None else int(width)
        self.stride = self.n_rows if stride is None else int(stride)
        n = max(self.width * self.stride, 1)
        self.col_idxs = torch.empty(n, dtype=torch.int32, device=self.device)
        self.values = torch.empty(n, dtype=self.dtype, device=self.device)
        c = ctx(self.device)
        C.call(f"gkomi_csr_to_ell_{suffix(self.dtype)}", ctypes.byref(c), ctypes.byref(csr.desc),
               self.width, self.stride, ptr(self.col_idxs), ptr(self.values))
        d = C.Ell()
        d.n_rows, d.n_cols, d.width, d.stride = self.n_rows, self.n_cols, self.width, self.stride
        d.col_idxs = ptr(self.col_idxs).value
        d.values = ptr(self.values).value
        self.desc = d

    def spmv(self, b, out=None, stream=None):
        """out = A b, mixed value types as DeviceCsr.spmv."""
        return _format_spmv(self, "ell", b, out, stream)

    def advanced_spmv(self, alpha, b, beta, c_vec, stream=None):
        """c = alpha A b + beta c (beta == 0 ignores c, NaN included)."""
        return _format_advanced_spmv(self, "ell", alpha, b, beta, c_vec, stream)



def _format_spmv(m, fmt, b, out, stream):
    check_operand(b, "b", None, m.device, rows=m.n_cols)
    suffix(b.dtype)
    if out is None:
        out = torch.empty((m.n_rows, _cols(b)), dtype=b.dtype, device=m.device)
    check_operand(out, "out", None, m.device, rows=m.n_rows, cols=_cols(b))
    suffix(out.dtype)
    b2, o2 = as_2d(b), as_2d(out)
    c = ctx(m.device, stream)
    if b.dtype == m.dtype and out.dtype == m.dtype:
        C.call(f"gkomi_{fmt}_spmv_{suffix(m.dtype)}", ctypes.byref(c), ctypes.byref(m.desc),
               b2.shape[1], ptr(b2), b2.stride(0), ptr(o2), o2.stride(0))
    else:
        C.call(f"gkomi_{fmt}_spmv_mixed", ctypes.byref(c), value_code(m.dtype),
               ctypes.byref(m.desc), b2.shape[1], value_code(b.dtype), ptr(b2), b2.stride(0),
               value_code(out.dtype), ptr(o2), o2.stride(0))
    return out


def _format_advanced_spmv(m, fmt, alpha, b, beta, c_vec, stream):
    check_operand(b, "b", None, m.device, rows=m.n_cols)
    check_operand(c_vec, "c", None, m.device, rows=m.n_rows, cols=_cols(b))
    suffix(b.dtype)
    suffix(c_vec.dtype)
    a = alpha if torch.is_tensor(alpha) else scalar(alpha, m.dtype, m.device)
    bt = beta if torch.is_tensor(beta) else scalar(beta, c_vec.dtype, m.device)
    _check_scalar(a, "alpha", m.dtype, m.device)
    _check_scalar(bt, "beta", c_vec.dtype, m.device)
    b2, c2 = as_2d(b), as_2d(c_vec)
    c = ctx(m.device, stream)
    if b.dtype == m.dtype and c_vec.dtype == m.dtype:
        C.call(f"gkomi_{fmt}_advanced_spmv_{suffix(m.dtype)}", ctypes.byref(c), ptr(a),
               ctypes.byref(m.desc), b2.shape[1], ptr(b2), b2.stride(0), ptr(bt), ptr(c2),
               c2.stride(0))
    else:
        C.call(f"gkomi_{fmt}_advanced_spmv_mixed", ctypes.byref(c), ptr(a),
               value_code(m.dtype), ctypes.byref(m.desc), b2.shape[1], value_code(b.dtype),
               ptr(b2), b2.stride(0), ptr(bt), value_code(c_vec.dtype), ptr(c2),
               c2.stride(0))
    return c_vec


def dot(x, y):
    check_operand(x, "x")
    suffix(x.dtype)
    check_operand(y, "y", x.dtype, x.device, rows=x.shape[0], cols=_cols(x))
    x2 = as_2d(x)
    y2 = as_2d(y)
    out = torch.empty(x2.shape[1], dtype=x.dtype, device=x.device)
    wsb = C.lib().gkomi_reduction_workspace_bytes(x2.shape[0], x2.shape[1])
    ws = workspace(wsb, x.device)
    c = ctx(x.device)
    C.call(f"gkomi_dense_dot_{suffix(x.dtype)}", ctypes.byref(c), x2.shape[0], x2.shape[1],
           ptr(x2), x2.stride(0), ptr(y2), y2.stride(0), ptr(out), ptr(ws), wsb)
    return out


def copy(x, y):
    """y = x for device tensors of equal shape (Dense::copy_from)."""
    check_operand(x, "x")
    suffix(x.dtype)
    check_operand(y, "y", x.dtype, x.device, rows=x.shape[0], cols=_cols(x))
    x2 = as_2d(x)
    y2 = as_2d(y)
    c = ctx(x.device)
    C.call(f"gkomi_dense_copy_{suffix(x.dtype)}", ctypes.byref(c), x2.shape[0], x2.shape[1],
           ptr(x2), x2.stride(0), ptr(y2), y2.stride(0))
    return y


def norm2(x):
    check_operand(x, "x")
    suffix(x.dtype)
    x2 = as_2d(x)
    out = torch.empty(x2.shape[1], dtype=x.dtype, device=x.device)
    wsb = C.lib().gkomi_reduction_workspace_bytes(x2.shape[0], x2.shape[1])
    ws = workspace(wsb, x.device)
    c = ctx(x.device)
    C.call(f"gkomi_dense_norm2_{suffix(x.dtype)}", ctypes.byref(c), x2.shape[0], x2.shape[1],
           ptr(x2), x2.stride(0), ptr(out), ptr(ws), wsb)
    return out


class FusedCg:
    """Single-RHS CG on the fused gfx950 kernels with the convergence test
    on the device; the host enqueues `chunk` iterations ahead and polls the
    stop flags once per chunk."""

    def __init__(self, csr, criteria_spec, jacobi=False, chunk=16):
        if csr.strategy != "classical":
            raise C.NotSupported("fused CG needs the classical CSR strategy")
        self.A = csr
        self.crit = C.criteria(criteria_spec)
        self.inv_diag = csr.jacobi() if jacobi else None
        self.chunk = int(chunk)
        dev, dt, n = csr.device, csr.dtype, csr.n_rows
        self.r = torch.empty(n, dtype=dt, device=dev)
        self.z = torch.empty(n, dtype=dt, device=dev)
        self.pa = torch.empty(n, dtype=dt, device=dev)
        self.pb = torch.empty(n, dtype=dt, device=dev)
        self.q = torch.empty(n, dtype=dt, device=dev)
        self.scalars = torch.zeros(8, dtype=dt, device=dev)
        self.orig = torch.zeros(C.MAX_CRITERIA, dtype=dt, device=dev)
        self.status = torch.zeros(1, dtype=torch.uint8, device=dev)
        self.flags = torch.zeros(2, dtype=torch.int64, device=dev)
        wsb = C.lib().gkomi_cg_work_bytes(n)
        self.work = workspace(wsb, dev)
        self.flags_host = torch.zeros(2, dtype=torch.int64).pin_memory()

    def _state(self, b, x):
        st = C.CgState()
        st.n = self.A.n_rows
        st.x, st.b = ptr(x).value, ptr(b).value
        st.r, st.z = ptr(self.r).value, ptr(self.z).value
        st.p_a, st.p_b, st.q = ptr(self.pa).value, ptr(self.pb).value, ptr(self.q).value
        st.inv_diag = ptr(self.inv_diag).value if self.inv_diag is not None else None
        st.scalars, st.orig_tau = ptr(self.scalars).value, ptr(self.orig).value
        st.status, st.flags, st.work = (ptr(self.status).value, ptr(self.flags).value,
                                        ptr(self.work).value)
        return st

    def apply(self, b, x, max_chunks=1 << 30):
        """Solve A x = b in place (x is the initial guess).  Returns the
        iteration count of the reference's Convergence logger."""
        st = self._state(b, x)
        c = ctx(self.A.device)
        sfx = suffix(self.A.dtype)
        C.call(f"gkomi_cg_fused_start_{sfx}", ctypes.byref(c), ctypes.byref(self.A.desc),
               ctypes.byref(self.crit), ctypes.byref(st))
        done = False
        chunks = 0
        while not done and chunks < max_chunks:
            C.call(f"gkomi_cg_fused_iterate_{sfx}", ctypes.byref(c), ctypes.byref(self.A.desc),
                   ctypes.byref(self.crit), ctypes.byref(st), self.chunk)
            self.flags_host.copy_(self.flags, non_blocking=True)
            torch.cuda.current_stream(self.A.device).synchronize()
            done = (int(self.flags_host[1].item()) & 0xffffffff) != 0
            chunks += 1
        return int(self.flags_host[0].item())

    @property
    def iteration(self):
        return int(self.flags[0].item())


class FusedBicgstab:
    """Single-RHS BiCGSTAB on the fused gfx950 kernels (five per iteration,
    both convergence checks on the device), polled once per chunk."""

    def __init__(self, csr, criteria_spec, jacobi=False, chunk=16):
        if csr.strategy != "classical":
            raise C.NotSupported("fused BiCGSTAB needs the classical CSR strategy")
        self.A = csr
        self.crit = C.criteria(criteria_spec)
        self.inv_diag = csr.jacobi() if jacobi else None
        self.chunk = int(chunk)
        dev, dt, n = csr.device, csr.dtype, csr.n_rows
        names = ("r", "rr", "p", "v", "s", "t") + (("y", "z") if jacobi else ())
        self.vec = {k: torch.empty(n, dtype=dt, device=dev) for k in names}
        self.scalars = torch.zeros(8, dtype=dt, device=dev)
        self.orig = torch.zeros(C.MAX_CRITERIA, dtype=dt, device=dev)
        self.status = torch.zeros(1, dtype=torch.uint8, device=dev)
        self.flags = torch.zeros(2, dtype=torch.int64, device=dev)
        self.work = workspace(C.lib().gkomi_bicgstab_work_bytes(n), dev)
        self.flags_host = torch.zeros(2, dtype=torch.int64).pin_memory()

    def _state(self, b, x):
        st = C.BicgstabState()
        st.n = self.A.n_rows
        st.x, st.b = ptr(x).value, ptr(b).value
        for k in ("r", "rr", "p", "v", "s", "t", "y", "z"):
            setattr(st, k, ptr(self.vec[k]).value if k in self.vec else None)
        st.inv_diag = ptr(self.inv_diag).value if self.inv_diag is not None else None
        st.scalars, st.orig_tau = ptr(self.scalars).value, ptr(self.orig).value
        st.status, st.flags, st.work = (ptr(self.status).value, ptr(self.flags).value,
                                        ptr(self.work).value)
        return st

    def apply(self, b, x, max_chunks=1 << 30):
        """Solve A x = b in place (x is the initial guess).  Returns the
        iteration count of the reference's loop (core/solver/bicgstab.cpp)."""
        st = self._state(b, x)
        c = ctx(self.A.device)
        sfx = suffix(self.A.dtype)
        C.call(f"gkomi_bicgstab_fused_start_{sfx}", ctypes.byref(c), ctypes.byref(self.A.desc),
               ctypes.byref(self.crit), ctypes.byref(st))
        done = False
        chunks = 0
        while not done and chunks < max_chunks:
            C.call(f"gkomi_bicgstab_fused_iterate_{sfx}", ctypes.byref(c),
                   ctypes.byref(self.A.desc), ctypes.byref(self.crit), ctypes.byref(st),
                   self.chunk)
            self.flags_host.copy_(self.flags, non_blocking=True)
            torch.cuda.current_stream(self.A.device).synchronize()
            done = (int(self.flags_host[1].item()) & 0xffffffff) != 0
            chunks += 1
        return int(self.flags_host[0].item())


class DeviceBatchCsr:
    """A batch of independent systems on the device.

    Uneven (row_off given): block-diagonal CSR with column indices local to
    each system.  Uniform (`DeviceBatchCsr.uniform`): the reference's
    batch::matrix::Csr with one shared pattern and values per system."""

    def __init__(self, row_off, row_ptrs, col_idxs, values, device="cuda"):
        dev = torch.device(device)
        self.device = dev
        row_off = np.asarray(row_off, dtype=np.int64)
        rp = np.asarray(row_ptrs, dtype=np.int32)
        sizes = np.diff(row_off)
        nnzs = rp[row_off[1:]] - rp[row_off[:-1]]
        self.sizes, self.nnzs = sizes, nnzs
        # longest row of each system
        lens = np.diff(rp).astype(np.int64)
        self.max_row_nnz = np.zeros(len(sizes), dtype=np.int64)
        nonempty = sizes > 0
        if lens.size and nonempty.any():
            self.max_row_nnz[nonempty] = np.maximum.reduceat(lens, row_off[:-1][nonempty])
        self.num_systems = len(sizes)
        self.total_rows = int(row_off[-1])
        self.row_off = torch.as_tensor(row_off).to(dev)
        self.row_ptrs = torch.as_tensor(rp).to(dev)
        self.col_idxs = torch.as_tensor(np.asarray(col_idxs, dtype=np.int32)).to(dev)
        self.values = torch.as_tensor(np.asarray(values)).to(dev) \
            if not torch.is_tensor(values) else values.to(dev)
        self.dtype = self.values.dtype
        self.uniform = False
        self.desc = self._desc()

    @classmethod
    def uniform(cls, num_systems, row_ptrs, col_idxs, values, device="cuda"):
        self = cls.__new__(cls)
        dev = torch.device(device)
        self.device = dev
        rp = np.asarray(row_ptrs, dtype=np.int32)
        n = len(rp) - 1
        nnz = int(rp[-1])
        self.sizes = np.full(num_systems, n, dtype=np.int64)
        self.nnzs = np.full(num_systems, nnz, dtype=np.int64)
        self.max_row_nnz = np.full(num_systems, int(np.diff(rp).max()) if n else 0,
                                   dtype=np.int64)
        self.num_systems = num_systems
        self.total_rows = num_systems * n
        self.row_off = None
        self.row_ptrs = torch.as_tensor(rp).to(dev)
        self.col_idxs = torch.as_tensor(np.asarray(col_idxs, dtype=np.int32)).to(dev)
        self.values = torch.as_tensor(np.asarray(values)).to(dev) \
            if not torch.is_tensor(values) else values.to(dev)
        self.dtype = self.values.dtype
        self.uniform = True
        self.desc = self._desc()
        return self

    def _desc(self):
        d = C.BatchCsr()
        d.num_systems = self.num_systems
        d.max_rows = int(self.sizes.max()) if self.num_systems else 0
        d.max_nnz = int(self.nnzs.max()) if self.num_systems else 0
        d.row_offsets = ptr(self.row_off).value if self.row_off is not None else None
        d.row_ptrs = ptr(self.row_ptrs).value
        d.col_idxs = ptr(self.col_idxs).value
        d.values = ptr(self.values).value
        if self.uniform:
            d.uniform_rows = int(self.sizes[0]) if self.num_systems else 0
            d.uniform_nnz = int(self.nnzs[0]) if self.num_systems else 0
        return d

    def plan(self, edges=(1 << 16, 512, 256, 128, 64, 0), block_size=None):
        """Size-ordered work queue split into bins (n in (edges[i+1], edges[i]])."""
        order = np.argsort(-self.sizes, kind="stable").astype(np.int32)
        ss = self.sizes[order]
        nz = self.nnzs[order]
        self._order = torch.as_tensor(order).to(self.device)
        p = C.BatchPlan()
        p.order = ptr(self._order).value
        begin = 0
        nb = 0
        for hi, lo in zip(edges[:-1], edges[1:]):
            end = int(np.searchsorted(-ss, -lo, side="left"))
            if end > begin:
                p.bin_begin[nb] = begin
                p.bin_max_rows[nb] = int(ss[begin:end].max())
                p.bin_max_nnz[nb] = int(nz[begin:end].max())
                p.bin_max_row_nnz[nb] = int(self.max_row_nnz[order[begin:end]].max())
                r = p.bin_max_rows[nb]
                p.block_size[nb] = block_size or (256 if r > 256 else (128 if r > 64 else 64))
                p.matrix_in_lds[nb] = -1
                p.blocks_per_cu[nb] = 0
                nb += 1
                begin = end
        p.num_bins = max(nb, 1)
        p.bin_begin[p.num_bins] = self.num_systems
        if nb == 0:
            p.bin_max_rows[0] = 0
            p.bin_max_nnz[0] = 0
            p.block_size[0] = 64
        return p

    def workspace(self):
        if getattr(self, "_ws", None) is None:
            self._ws = workspace(C.lib().gkomi_batch_workspace_bytes(ctypes.byref(self.desc)),
                                 self.device)
        return self._ws

    def settings(self, tol=1e-10, relative=True, max_iters=100, jacobi=True):
        s = C.BatchSettings()
        s.max_iterations = int(max_iters)
        s.tolerance_type = 1 if relative else 0
        s.tolerance = float(tol)
        s.preconditioner = 1 if jacobi else 0
        return s

    @staticmethod
    def _method(method):
        if method not in ("bicgstab", "cg"):
            raise C.NotSupported("batch method must be 'bicgstab' or 'cg'")
        return method

    def solve(self, b, x, settings, plan=None, stream=None, method="bicgstab"):
        """BiCGSTAB (or CG, method="cg") on every system; x holds the initial
        guess and gets the solution.  Returns (iterations, residual norms) as
        device tensors."""
        m = self._method(method)
        iters = torch.empty(self.num_systems, dtype=torch.int32, device=self.device)
        res = torch.empty(self.num_systems, dtype=self.dtype, device=self.device)
        ws = self.workspace()
        c = ctx(self.device, stream)
        C.call(f"gkomi_batch_{m}_solve_{suffix(self.dtype)}", ctypes.byref(c),
               ctypes.byref(self.desc), ctypes.byref(settings),
               ctypes.byref(plan) if plan is not None else None, ptr(b), ptr(x), ptr(iters),
               ptr(res), ptr(ws), ws.numel())
        return iters, res

    def capture(self, b, x, settings, plan=None, x0=None, method="bicgstab"):
        """The whole batched solve -- x = x0 (zero when x0 is None), then every
        bin's persistent launch -- captured in one HIP graph on a side
        stream.  Returns (graph, iterations, residual norms); graph.replay()
        re-runs it on the same buffers.  The work-queue counters are reset by
        the kernels themselves, and x is set by a fill/copy kernel, so the
        graph holds no memset node."""
        m = self._method(method)
        iters = torch.empty(self.num_systems, dtype=torch.int32, device=self.device)
        res = torch.empty(self.num_systems, dtype=self.dtype, device=self.device)
        ws = self.workspace()
        sfx = suffix(self.dtype)
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=side):
            if x0 is None:
                x.zero_()
            else:
                x.copy_(x0)
            c = ctx(self.device, side)
            C.call(f"gkomi_batch_{m}_solve_{sfx}", ctypes.byref(c),
                   ctypes.byref(self.desc), ctypes.byref(settings),
                   ctypes.byref(plan) if plan is not None else None, ptr(b), ptr(x),
                   ptr(iters), ptr(res), ptr(ws), ws.numel())
        torch.cuda.current_stream(self.device).wait_stream(side)
        return graph, iters, res

    def tune(self, b, x, settings, plan, method="bicgstab"):
        """Autotune block size and LDS residency per bin (writes into plan)."""
        m = self._method(method)
        xs = torch.empty_like(x)
        iters = torch.empty(self.num_systems, dtype=torch.int32, device=self.device)
        res = torch.empty(self.num_systems, dtype=self.dtype, device=self.device)
        ws = self.workspace()
        c = ctx(self.device)
        C.call(f"gkomi_batch_{m}_tune_{suffix(self.dtype)}", ctypes.byref(c),
               ctypes.byref(self.desc), ctypes.byref(settings), ctypes.byref(plan), ptr(b),
               ptr(x), ptr(xs), ptr(iters), ptr(res), ptr(ws), ws.numel())
        return plan
