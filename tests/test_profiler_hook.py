"""log::ProfilerHook and the logger events on the host: the Markdown tables
against the reference's own expected output, the flat and nested summaries
over user ranges, the scope guard, and the executor's allocate / copy / free
events (core/test/log/profiler_hook.cpp)."""
import gc
import io
import json
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))
EVENTS = GOLD["profiler_events"]


@pytest.fixture(scope="module")
def g():
    import ginkgo_amd._gko as g
    return g


def recorder():
    out = []
    return out, (lambda name, cat: out.append("begin:" + name)), \
        (lambda name, cat: out.append("end:" + name))


def test_table_writer_flat_matches_reference(g):
    """TableSummaryWriter.write: sorted by total, one decimal in the largest
    unit, average by count (division by zero gives 0)
    (core/test/log/profiler_hook.cpp:437-467)."""
    kat = GOLD["profiler_table"]
    stream = io.StringIO()
    w = g.TableSummaryWriter(stream, kat["header"])
    entries = [g.summary_entry(*e) for e in kat["flat_entries"]]
    w.write(entries, kat["overhead_ns"])
    assert stream.getvalue() == kat["flat_expected"]


def test_table_writer_nested_matches_reference(g):
    """TableSummaryWriter.write_nested: children longest first, "(self)"
    before the first shorter child (core/test/log/profiler_hook.cpp:470-500)."""
    kat = GOLD["profiler_table"]

    def node(d):
        return g.nested_summary_entry(d["name"], d["elapsed_ns"], d["count"],
                                      [node(c) for c in d["children"]])

    stream = io.StringIO()
    g.TableSummaryWriter(stream, kat["header"]).write_nested(node(kat["nested_root"]),
                                                             kat["overhead_ns"])
    assert stream.getvalue() == kat["nested_expected"]


def test_table_writer_units_and_padding(g):
    """Unit boundaries of the duration format: 999 ns stays in ns, 1000 ns
    is 1.0 us, 59.999 s stays in s, 60 s is 1.0 m."""
    stream = io.StringIO()
    entries = [g.summary_entry("a", 999, 999, 1), g.summary_entry("b", 1000, 0, 1),
               g.summary_entry("c", 59_999_999_999, 0, 1), g.summary_entry("d", 60_000_000_000, 0, 1)]
    g.TableSummaryWriter(stream, "h").write(entries, 0)
    rows = stream.getvalue().splitlines()[4:]
    cells = {r.split("|")[1].strip(): r.split("|")[2].strip() for r in rows}
    assert cells == {"a": "999.0 ns", "b": "1.0 us", "c": "60.0 s", "d": "1.0 m"}


class _Capture:
    def __init__(self):
        self.entries = None
        self.root = None
        self.overhead = None

    def write(self, entries, overhead_ns):
        self.entries = [(e.name, e.inclusive_ns, e.exclusive_ns, e.count) for e in entries]
        self.overhead = overhead_ns

    def write_nested(self, root, overhead_ns):
        self.root = root
        self.overhead = overhead_ns


def _ranges_unique(hook):
    with hook.user_range("foo"):
        with hook.user_range("bar"):
            pass
        with hook.user_range("bar"):
            pass
        with hook.user_range("baz"):
            with hook.user_range("bazz"):
                pass
            with hook.user_range("bazzz"):
                pass
        with hook.user_range("bazzzz"):
            pass


def test_summary_counts_and_exclusive_times(g):
    """create_summary: one entry per name in first-opened order with
    inclusive = exclusive + the inclusive time of the direct children
    (core/test/log/profiler_hook.cpp:284-359)."""
    cap = _Capture()
    hook = g.ProfilerHook.create_summary(g.CpuTimer(), cap)
    _ranges_unique(hook)
    del hook
    gc.collect()
    kat = EVENTS["summary_ranges_unique"]
    assert [e[0] for e in cap.entries] == kat["names"]
    assert [e[3] for e in cap.entries] == kat["counts"]
    inc = {e[0]: e[1] for e in cap.entries}
    exc = {e[0]: e[2] for e in cap.entries}
    assert inc["total"] == exc["total"] + inc["foo"]
    assert inc["foo"] == exc["foo"] + inc["bar"] + inc["baz"] + inc["bazzzz"]
    assert inc["baz"] == exc["baz"] + inc["bazz"] + inc["bazzz"]
    for leaf in ("bar", "bazz", "bazzz", "bazzzz"):
        assert inc[leaf] == exc[leaf]
    assert all(v >= 0 for v in exc.values())
    assert cap.overhead >= 0


def test_nested_summary_tree(g):
    """create_nested_summary: one node per path of names, children in the
    order first opened (core/test/log/profiler_hook.cpp:362-434)."""
    cap = _Capture()
    hook = g.ProfilerHook.create_nested_summary(g.CpuTimer(), cap)
    with hook.user_range("foo"):
        with hook.user_range("foo"):
            pass
        with hook.user_range("foo"):
            pass
        with hook.user_range("bar"):
            with hook.user_range("baz"):
                pass
            with hook.user_range("bazz"):
                pass
        with hook.user_range("baz"):
            pass
    del hook
    gc.collect()

    def shape(n):
        return [n.name, n.count, [shape(c) for c in n.children]]

    assert shape(cap.root) == EVENTS["nested_ranges"]["tree"]

    def check(n):
        assert n.elapsed_ns >= sum(c.elapsed_ns for c in n.children)
        for c in n.children:
            check(c)

    check(cap.root)


def test_summary_closes_open_ranges_at_destruction(g):
    """Ranges still open when the hook goes away are closed before the
    table is written."""
    cap = _Capture()
    hook = g.ProfilerHook.create_summary(g.CpuTimer(), cap)
    r = hook.user_range("open")
    r.__enter__()
    del hook, r
    gc.collect()
    names = [e[0] for e in cap.entries]
    assert names == ["total", "open"]
    assert [e[3] for e in cap.entries] == [1, 1]


def test_summary_default_writer_prints_table(g, capfd):
    """Without a writer the table goes to stderr under "Runtime summary"."""
    hook = g.ProfilerHook.create_summary()
    with hook.user_range("work"):
        pass
    del hook
    gc.collect()
    err = capfd.readouterr().err
    assert err.startswith("Runtime summary\nOverhead estimate ")
    assert "| work " in err and "| total " in err


def test_scope_guard_order(g):
    """user_range opens on enter and closes on exit, nested ranges in
    order (core/test/log/profiler_hook.cpp:252-281)."""
    out = []
    hook = g.ProfilerHook.create_custom(lambda n, c: out.append(n), lambda n, c: out.append(n))
    with hook.user_range("foo"):
        with hook.user_range("bar"):
            pass
        with hook.user_range("baz"):
            pass
    assert out == EVENTS["scope_guard"]


def test_custom_hook_categories(g):
    cats = []
    hook = g.ProfilerHook.create_custom(lambda n, c: cats.append(c), lambda n, c: None)
    with hook.user_range("x"):
        pass
    assert cats == [g.profile_event_category.user]


def test_executor_allocate_copy_free_events(g):
    """A hook on the executor sees the allocation, the copy in and the free
    of an array (core/test/log/profiler_hook.cpp:59-78)."""
    import ginkgo_amd.gko as gko
    out, begin, end = recorder()
    host = gko.HostExecutor.create()
    hook = g.ProfilerHook.create_custom(begin, end)
    host.add_logger(hook)
    assert host.num_loggers == 1 and host.should_propagate_log()
    d = gko.matrix.Dense("f64").from_numpy(host, np.ones((3, 1)))
    del d
    gc.collect()
    host.remove_logger(hook)
    assert out == EVENTS["allocate_copy_free"]
    assert host.num_loggers == 0 and not host.should_propagate_log()
    with pytest.raises(gko.GkoError):
        host.remove_logger(hook)


def test_executor_events_with_propagation_switched_off(g):
    """set_log_propagation(False) keeps the executor's own events."""
    import ginkgo_amd.gko as gko
    out, begin, end = recorder()
    host = gko.HostExecutor.create()
    hook = g.ProfilerHook.create_custom(begin, end)
    host.add_logger(hook)
    host.set_log_propagation(False)
    assert not host.should_propagate_log()
    d = gko.matrix.Dense("f64").from_numpy(host, np.ones((2, 1)))
    del d
    gc.collect()
    assert out == EVENTS["allocate_copy_free"]


def test_create_for_host_executor_is_not_supported(g):
    import ginkgo_amd.gko as gko
    with pytest.raises(gko.NotSupported):
        g.ProfilerHook.create_for_executor(gko.HostExecutor.create())
    with pytest.raises(gko.NotSupported):
        g.Mi355xTimer(gko.HostExecutor.create())
