"""CPU oracle (test infrastructure only).

Python access to ``liboracle.so``, the plain-C restatement of Ginkgo's
ReferenceExecutor kernels and solver loops.  Only the test-suite, the smoke
check and the CPU-baseline leg of ``bench.py`` import this package; the
product (``ginkgo_amd``) never does.
"""
from .oracle import *  # noqa: F401,F403
