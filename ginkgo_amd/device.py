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


class DeviceCsr:
    """Csr<V, int32> on the device with its SpMV strategy.

    strategy: 'classical', 'load_balance', 'merge_path' or 'automatical'
    (picks classical for short, uniform rows and load_balance otherwise, as
    the reference's automatical::process does from row statistics,
    csr.hpp:606-673)."""

    def __init__(self, row_ptrs, col_idxs, values, n_cols=None, strategy="automatical",
                 subwave=0, device="cuda"):
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
        self.device = dev
        rp = self.row_ptrs.cpu().numpy() if strategy == "automatical" else None
        if strategy == "automatical":
            lens = np.diff(rp) if self.n_rows > 0 else np.zeros(1)
            mx = int(lens.max()) if lens.size else 0
            mean = float(lens.mean()) if lens.size else 0.0
            strategy = "classical" if (mx <= 64 and mx <= 4 * mean + 4) else "load_balance"
        self.strategy = strategy
        self.subwave = int(subwave)
        self.srow = None
        self.srow_blocks = 0
        if strategy == "load_balance":
            chunk = int(C.lib().gkomi_csr_lb_chunk(self.values.element_size()))
            self.srow_blocks = max(1, -(-self.nnz // chunk))
            self.srow = torch.empty(self.srow_blocks + 1, dtype=torch.int32, device=dev)
            c = ctx(dev)
            C.call("gkomi_csr_lb_partition", ctypes.byref(c), self.n_rows, ptr(self.row_ptrs),
                   self.nnz, chunk, ptr(self.srow), self.srow_blocks)
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
        d.srow_blocks = self.srow_blocks
        d.col_offsets = ptr(self.col_offsets).value if getattr(self, "col_offsets", None) \
            is not None else None
        work = getattr(self, "work", None)
        d.work = ptr(work).value if work is not None else None
        d.work_bytes = work.numel() if work is not None else 0
        return d

    def spmv(self, b, out=None, stream=None):
        b2 = as_2d(b)
        if out is None:
            out = torch.empty((self.n_rows, b2.shape[1]), dtype=self.dtype, device=self.device)
        o2 = as_2d(out)
        if b2.shape[0] != self.n_cols or o2.shape[0] != self.n_rows or o2.shape[1] != b2.shape[1]:
            raise C.DimensionMismatch("spmv: operand sizes do not conform")
        c = ctx(self.device, stream)
        C.call(f"gkomi_csr_spmv_{suffix(self.dtype)}", ctypes.byref(c), ctypes.byref(self.desc),
               b2.shape[1], ptr(b2), b2.stride(0), ptr(o2), o2.stride(0))
        return out

    def advanced_spmv(self, alpha, b, beta, c_vec, stream=None):
        b2 = as_2d(b)
        c2 = as_2d(c_vec)
        a = alpha if torch.is_tensor(alpha) else scalar(alpha, self.dtype, self.device)
        bt = beta if torch.is_tensor(beta) else scalar(beta, self.dtype, self.device)
        c = ctx(self.device, stream)
        C.call(f"gkomi_csr_advanced_spmv_{suffix(self.dtype)}", ctypes.byref(c), ptr(a),
               ctypes.byref(self.desc), b2.shape[1], ptr(b2), b2.stride(0), ptr(bt), ptr(c2),
               c2.stride(0))
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
        b2 = as_2d(b)
        if out is None:
            out = torch.empty((self.n_rows, b2.shape[1]), dtype=self.dtype, device=self.device)
        o2 = as_2d(out)
        c = ctx(self.device, stream)
        C.call(f"gkomi_sellp_spmv_{suffix(self.dtype)}", ctypes.byref(c), ctypes.byref(self.desc),
               b2.shape[1], ptr(b2), b2.stride(0), ptr(o2), o2.stride(0))
        return out

    def advanced_spmv(self, alpha, b, beta, c_vec):
        b2 = as_2d(b)
        c2 = as_2d(c_vec)
        a = scalar(alpha, self.dtype, self.device)
        bt = scalar(beta, self.dtype, self.device)
        c = ctx(self.device)
        C.call(f"gkomi_sellp_advanced_spmv_{suffix(self.dtype)}", ctypes.byref(c), ptr(a),
               ctypes.byref(self.desc), b2.shape[1], ptr(b2), b2.stride(0), ptr(bt), ptr(c2),
               c2.stride(0))
        return c_vec


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
        b2 = as_2d(b)
        if out is None:
            out = torch.empty((self.n_rows, b2.shape[1]), dtype=self.dtype, device=self.device)
        o2 = as_2d(out)
        c = ctx(self.device, stream)
        C.call(f"gkomi_ell_spmv_{suffix(self.dtype)}", ctypes.byref(c), ctypes.byref(self.desc),
               b2.shape[1], ptr(b2), b2.stride(0), ptr(o2), o2.stride(0))
        return out

    def advanced_spmv(self, alpha, b, beta, c_vec):
        b2 = as_2d(b)
        c2 = as_2d(c_vec)
        a = scalar(alpha, self.dtype, self.device)
        bt = scalar(beta, self.dtype, self.device)
        c = ctx(self.device)
        C.call(f"gkomi_ell_advanced_spmv_{suffix(self.dtype)}", ctypes.byref(c), ptr(a),
               ctypes.byref(self.desc), b2.shape[1], ptr(b2), b2.stride(0), ptr(bt), ptr(c2),
               c2.stride(0))
        return c_vec


def dot(x, y):
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
    x2 = as_2d(x)
    y2 = as_2d(y)
    c = ctx(x.device)
    C.call(f"gkomi_dense_copy_{suffix(x.dtype)}", ctypes.byref(c), x2.shape[0], x2.shape[1],
           ptr(x2), x2.stride(0), ptr(y2), y2.stride(0))
    return y


def norm2(x):
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
