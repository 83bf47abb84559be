"""ginkgo_amd: the sparse iterative-solver hot path of Ginkgo, MI355X-native.

Native parts (built in-tree by ``python -m ginkgo_amd.build``):
  * lib/libgkomi.so       -- hand-written gfx950 HIP kernels behind the C ABI in include/gkomi.h
  * lib/libginkgo_amd.so  -- the C++ host library (gko:: executors, LinOps, solvers)

Python modules:
  * ginkgo_amd._capi      -- ctypes binding of the C ABI
  * ginkgo_amd.device     -- torch-tensor plumbing over the C ABI
  * ginkgo_amd.generators -- seeded synthetic problems
"""
__version__ = "0.1.0"
