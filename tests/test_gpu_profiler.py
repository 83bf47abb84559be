"""Logger events and log::ProfilerHook on the MI355X executor: the range
sequence of an iterative solver's applies (core/test/log/profiler_hook.cpp:
214-249), propagation from the executor to every object on it, summaries
timed with HIP events on the executor's stream, and roctx ranges."""
import gc
import json
import os
import time

import numpy as np
import pytest

from ginkgo_amd import generators as gen

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
EVENTS = GOLD["profiler_events"]


@pytest.fixture(scope="module")
def gko():
    import ginkgo_amd.gko as g
    return g


@pytest.fixture(scope="module")
def g():
    import ginkgo_amd._gko as g
    return g


@pytest.fixture()
def exec_(gko):
    # a fresh executor per test: the executor's loggers are its own state
    return gko.Mi355xExecutor.create(0, torch.cuda.current_stream().cuda_stream)


def poisson(gko, exec_, m):
    rp, ci, v = gen.poisson2d(m)
    return gko.csr_from_host(exec_, rp.astype(np.int32), ci.astype(np.int32),
                             v.astype(np.float64), m * m, "automatical")


def vec(gko, exec_, a):
    return gko.matrix.Dense("f64").from_numpy(exec_, np.asarray(a, np.float64).reshape(-1, 1))


def recorder():
    out = []
    return out, (lambda name, cat: out.append("begin:" + name)), \
        (lambda name, cat: out.append("end:" + name))


def balanced(events):
    """Every end closes the innermost open range of the same name."""
    stack = []
    for e in events:
        kind, name = e.split(":", 1)
        if kind == "begin":
            stack.append(name)
        else:
            assert stack and stack[-1] == name, (e, stack)
            stack.pop()
    return not stack


def test_iterative_solver_apply_ranges_match_reference(gko, g, exec_):
    """Ir with one iteration: an "iteration" range inside the apply and the
    advanced apply, objects named by set_object_name, the unnamed scalars
    by their type (core/test/log/profiler_hook.cpp:214-249)."""
    out, begin, end = recorder()
    hook = g.ProfilerHook.create_custom(begin, end)
    A = poisson(gko, exec_, 4)
    solver = (gko.solver.Ir.build("f64")
              .with_criteria(gko.stop.Iteration.build().with_max_iters(1))
              .on(exec_).generate(A))
    b = vec(gko, exec_, np.ones(16))
    x = vec(gko, exec_, np.zeros(16))
    alpha = gko.matrix.Dense("f64").from_numpy(exec_, np.ones((1, 1)))
    hook.set_object_name(solver, "solver")
    hook.set_object_name(b, "mtx")
    hook.set_object_name(x, "mtx")
    solver.add_logger(hook)
    solver.apply(b, x)
    solver.apply(alpha, b, alpha, x)
    solver.remove_logger(hook)
    solver.apply(b, x)
    assert out == EVENTS["iteration"]


@pytest.mark.parametrize("method", ["cg", "bicgstab"])
def test_fused_krylov_apply_ranges(gko, g, exec_, method):
    """The fused device loops report one "iteration" range per apply (the
    iterations run without a host round trip); Convergence still sees the
    final state."""
    out, begin, end = recorder()
    hook = g.ProfilerHook.create_custom(begin, end)
    conv = gko.log.Convergence.create()
    A = poisson(gko, exec_, 16)
    solver_t = gko.solver.Cg if method == "cg" else gko.solver.Bicgstab
    solver = (solver_t.build("f64")
              .with_criteria(gko.stop.Iteration.build().with_max_iters(500),
                             gko.stop.ResidualNorm.build("f64").with_reduction_factor(1e-10))
              .on(exec_).generate(A))
    b = vec(gko, exec_, np.ones(256))
    x = vec(gko, exec_, np.zeros(256))
    hook.set_object_name(solver, "solver")
    hook.set_object_name(b, "b")
    hook.set_object_name(x, "x")
    solver.add_logger(hook)
    solver.add_logger(conv)
    solver.apply(b, x)
    assert out[0] == "begin:apply(solver * b = x)"
    assert out[1] == "begin:iteration"
    assert out[-2] == "end:iteration"
    assert out[-1] == "end:apply(solver * b = x)"
    assert balanced(out)
    assert conv.has_converged and 0 < conv.num_iterations < 500


def test_executor_propagation(gko, g, exec_):
    """A hook on the executor sees the executor's memory events and, since
    it needs propagation, the generate and apply of every object on it."""
    out, begin, end = recorder()
    hook = g.ProfilerHook.create_custom(begin, end)
    exec_.add_logger(hook)
    A = poisson(gko, exec_, 8)
    factory = (gko.solver.Cg.build("f64")
               .with_criteria(gko.stop.Iteration.build().with_max_iters(100),
                              gko.stop.ResidualNorm.build("f64").with_reduction_factor(1e-8))
               .on(exec_))
    hook.set_object_name(factory, "cg_factory")
    solver = factory.generate(A)
    hook.set_object_name(solver, "solver")
    b = vec(gko, exec_, np.ones(64))
    x = vec(gko, exec_, np.zeros(64))
    hook.set_object_name(b, "b")
    hook.set_object_name(x, "x")
    n_before = len(out)
    solver.apply(b, x)
    applied = out[n_before:]
    exec_.remove_logger(hook)
    n_after = len(out)
    solver.apply(b, x)
    assert len(out) == n_after
    assert balanced(out)
    assert "begin:generate(cg_factory)" in out and "end:generate(cg_factory)" in out
    assert "begin:allocate" in out and "begin:copy" in out and "begin:free" in out
    assert applied[0] == "begin:apply(solver * b = x)"
    assert applied[1] == "begin:iteration"
    assert applied[-1] == "end:apply(solver * b = x)"


def test_unnamed_objects_are_named_by_type(gko, g, exec_):
    """Without set_object_name a range names the object's demangled type."""
    out, begin, end = recorder()
    hook = g.ProfilerHook.create_custom(begin, end)
    A = poisson(gko, exec_, 4)
    b = vec(gko, exec_, np.ones(16))
    x = vec(gko, exec_, np.zeros(16))
    A.add_logger(hook)
    A.apply(b, x)
    A.remove_logger(hook)
    name = "apply(gko::matrix::Csr<double, int> * gko::matrix::Dense<double> = " \
           "gko::matrix::Dense<double>)"
    assert out == ["begin:" + name, "end:" + name]


def test_propagation_switched_off(gko, g, exec_):
    """set_log_propagation(False): the executor's hook keeps its own memory
    events but no longer sees the objects' applies."""
    out, begin, end = recorder()
    hook = g.ProfilerHook.create_custom(begin, end)
    A = poisson(gko, exec_, 4)
    b = vec(gko, exec_, np.ones(16))
    x = vec(gko, exec_, np.zeros(16))
    exec_.add_logger(hook)
    exec_.set_log_propagation(False)
    A.apply(b, x)
    tmp = vec(gko, exec_, np.ones(16))
    del tmp
    gc.collect()
    exec_.remove_logger(hook)
    assert not any(e.startswith("begin:apply") for e in out)
    assert "begin:allocate" in out and "begin:free" in out


class _Capture:
    def __init__(self):
        self.entries = None
        self.root = None

    def write(self, entries, overhead_ns):
        self.entries = {e.name: (e.inclusive_ns, e.exclusive_ns, e.count) for e in entries}

    def write_nested(self, root, overhead_ns):
        self.root = root


def test_summary_with_device_timer(gko, g, exec_):
    """create_summary timed by HIP events on the executor's stream: the
    solve's apply range holds the device time of the solve, within the wall
    time around it."""
    cap = _Capture()
    hook = g.ProfilerHook.create_summary(g.Mi355xTimer(exec_), cap)
    A = poisson(gko, exec_, 128)
    solver = (gko.solver.Cg.build("f64")
              .with_criteria(gko.stop.Iteration.build().with_max_iters(50))
              .on(exec_).generate(A))
    hook.set_object_name(solver, "solver")
    b = vec(gko, exec_, np.ones(128 * 128))
    x = vec(gko, exec_, np.zeros(128 * 128))
    hook.set_object_name(b, "b")
    hook.set_object_name(x, "x")
    exec_.add_logger(hook)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.apply(b, x)
    exec_.synchronize()
    wall_ns = (time.perf_counter() - t0) * 1e9
    exec_.remove_logger(hook)
    del hook
    gc.collect()
    e = cap.entries
    inc, exc, cnt = e["apply(solver * b = x)"]
    assert cnt == 1 and inc > 0 and 0 <= exc <= inc
    assert e["iteration"][2] == 1
    assert inc <= wall_ns * 1.05 + 50_000
    assert e["total"][0] >= inc
    assert e["iteration"][0] <= inc


def test_nested_summary_with_device_timer(gko, g, exec_):
    cap = _Capture()
    hook = g.ProfilerHook.create_nested_summary(g.Timer.create_for_executor(exec_), cap)
    A = poisson(gko, exec_, 16)
    exec_.add_logger(hook)
    factory = (gko.solver.Cg.build("f64")
               .with_criteria(gko.stop.Iteration.build().with_max_iters(20))
               .on(exec_))
    hook.set_object_name(factory, "f")
    solver = factory.generate(A)
    hook.set_object_name(solver, "solver")
    b = vec(gko, exec_, np.ones(256))
    x = vec(gko, exec_, np.zeros(256))
    hook.set_object_name(b, "b")
    hook.set_object_name(x, "x")
    solver.apply(b, x)
    exec_.remove_logger(hook)
    del hook
    gc.collect()
    root = cap.root
    assert root.name == "total"
    top = {c.name: c for c in root.children}
    assert "generate(f)" in top and "apply(solver * b = x)" in top
    apply_node = top["apply(solver * b = x)"]
    assert [c.name for c in apply_node.children][:1] == ["iteration"]
    assert apply_node.elapsed_ns >= sum(c.elapsed_ns for c in apply_node.children)


def test_synchronization_ranges(gko, g, exec_):
    """set_synchronization(True) puts a "synchronize" range around each
    executor wait before a memory range opens or closes."""
    cap = _Capture()
    hook = g.ProfilerHook.create_summary(g.CpuTimer(), cap)
    hook.set_synchronization(True)
    exec_.add_logger(hook)
    tmp = vec(gko, exec_, np.ones(1000))
    del tmp
    gc.collect()
    exec_.remove_logger(hook)
    del hook
    gc.collect()
    assert cap.entries["synchronize"][2] >= 4
    assert cap.entries["allocate"][2] >= 1 and cap.entries["free"][2] >= 1


def test_roctx_hook(gko, g, exec_):
    """create_roctx / create_for_executor load roctx at run time and run a
    solve under it."""
    hook = g.ProfilerHook.create_for_executor(exec_)
    g.ProfilerHook.create_roctx()
    A = poisson(gko, exec_, 8)
    exec_.add_logger(hook)
    solver = (gko.solver.Cg.build("f64")
              .with_criteria(gko.stop.Iteration.build().with_max_iters(100),
                             gko.stop.ResidualNorm.build("f64").with_reduction_factor(1e-8))
              .on(exec_).generate(A))
    b = vec(gko, exec_, np.ones(64))
    x = vec(gko, exec_, np.zeros(64))
    with hook.user_range("solve"):
        solver.apply(b, x)
    exec_.remove_logger(hook)
    r = np.ones(64) - gen.csr_to_dense(A.row_ptrs(), A.col_idxs(), A.values(), 64) @ x.to_numpy()[:, 0]
    assert np.linalg.norm(r) <= 1e-8 * 8 * 1.01
