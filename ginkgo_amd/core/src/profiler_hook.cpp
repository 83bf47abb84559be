// log::ProfilerHook: range names from the logger events, roctx ranges, the
// flat and nested summaries and their Markdown tables (core/log/
// profiler_hook.cpp, profiler_hook_summary.cpp,
// profiler_hook_summary_writer.cpp).
#include <cxxabi.h>
#include <dlfcn.h>

#include <algorithm>
#include <array>
#include <cstdlib>
#include <iomanip>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <sstream>
#include <typeinfo>

#include "ginkgo_amd/profiler_hook.hpp"

namespace gko {
namespace log {

namespace {

std::string demangle(const std::type_info& t)
{
    int status = 0;
    char* s = abi::__cxa_demangle(t.name(), nullptr, nullptr, &status);
    std::string out = (status == 0 && s) ? s : t.name();
    std::free(s);
    return out;
}

bool is_iterative(const LinOp* op)
{
    return dynamic_cast<const solver::IterativeBase*>(op) != nullptr;
}

}  // namespace

// ------------------------------------------------------------ the events ----

ProfilerHook::ProfilerHook(hook_function begin, hook_function end)
    : begin_hook_{std::move(begin)}, end_hook_{std::move(end)}
{}

void ProfilerHook::maybe_synchronize(const Executor* exec) const
{
    if (!synchronize_ || !exec) return;
    profiling_scope_guard guard{"synchronize", profile_event_category::internal, begin_hook_,
                                end_hook_};
    exec->synchronize();
}

std::string ProfilerHook::stringify(const LinOp* obj) const
{
    if (!obj) return "nullptr";
    auto it = name_map_.find(obj);
    return it != name_map_.end() ? it->second : demangle(typeid(*obj));
}

std::string ProfilerHook::stringify(const LinOpFactory* obj) const
{
    if (!obj) return "nullptr";
    auto it = name_map_.find(obj);
    return it != name_map_.end() ? it->second : demangle(typeid(*obj));
}

void ProfilerHook::set_object_name(const LinOp* obj, std::string name)
{
    name_map_[obj] = std::move(name);
}

void ProfilerHook::set_object_name(const LinOpFactory* obj, std::string name)
{
    name_map_[obj] = std::move(name);
}

void ProfilerHook::on_allocation_started(const Executor* exec, const size_type&) const
{
    maybe_synchronize(exec);
    begin_hook_("allocate", profile_event_category::memory);
}

void ProfilerHook::on_allocation_completed(const Executor* exec, const size_type&,
                                           const uintptr&) const
{
    maybe_synchronize(exec);
    end_hook_("allocate", profile_event_category::memory);
}

void ProfilerHook::on_free_started(const Executor* exec, const uintptr&) const
{
    maybe_synchronize(exec);
    begin_hook_("free", profile_event_category::memory);
}

void ProfilerHook::on_free_completed(const Executor* exec, const uintptr&) const
{
    maybe_synchronize(exec);
    end_hook_("free", profile_event_category::memory);
}

void ProfilerHook::on_copy_started(const Executor* from, const Executor* to, const uintptr&,
                                   const uintptr&, const size_type&) const
{
    maybe_synchronize(from);
    maybe_synchronize(to);
    begin_hook_("copy", profile_event_category::operation);
}

void ProfilerHook::on_copy_completed(const Executor* from, const Executor* to, const uintptr&,
                                     const uintptr&, const size_type&) const
{
    maybe_synchronize(from);
    maybe_synchronize(to);
    end_hook_("copy", profile_event_category::operation);
}

void ProfilerHook::on_linop_apply_started(const LinOp* A, const LinOp* b, const LinOp* x) const
{
    const auto name = "apply(" + stringify(A) + " * " + stringify(b) + " = " + stringify(x) + ")";
    begin_hook_(name.c_str(), profile_event_category::linop);
    if (is_iterative(A)) begin_hook_("iteration", profile_event_category::solver);
}

void ProfilerHook::on_linop_apply_completed(const LinOp* A, const LinOp* b,
                                            const LinOp* x) const
{
    const auto name = "apply(" + stringify(A) + " * " + stringify(b) + " = " + stringify(x) + ")";
    if (is_iterative(A)) end_hook_("iteration", profile_event_category::solver);
    end_hook_(name.c_str(), profile_event_category::linop);
}

namespace {
template <typename S>
std::string advanced_name(const S& str, const LinOp* A, const LinOp* alpha, const LinOp* b,
                          const LinOp* beta, const LinOp* x)
{
    return "advanced_apply(" + str(alpha) + " * " + str(A) + " * " + str(b) + " + " + str(beta) +
           " * " + str(x) + ")";
}
}  // namespace

void ProfilerHook::on_linop_advanced_apply_started(const LinOp* A, const LinOp* alpha,
                                                   const LinOp* b, const LinOp* beta,
                                                   const LinOp* x) const
{
    const auto name =
        advanced_name([this](const LinOp* o) { return stringify(o); }, A, alpha, b, beta, x);
    begin_hook_(name.c_str(), profile_event_category::linop);
    if (is_iterative(A)) begin_hook_("iteration", profile_event_category::solver);
}

void ProfilerHook::on_linop_advanced_apply_completed(const LinOp* A, const LinOp* alpha,
                                                     const LinOp* b, const LinOp* beta,
                                                     const LinOp* x) const
{
    const auto name =
        advanced_name([this](const LinOp* o) { return stringify(o); }, A, alpha, b, beta, x);
    if (is_iterative(A)) end_hook_("iteration", profile_event_category::solver);
    end_hook_(name.c_str(), profile_event_category::linop);
}

void ProfilerHook::on_linop_factory_generate_started(const LinOpFactory* factory,
                                                     const LinOp*) const
{
    const auto name = "generate(" + stringify(factory) + ")";
    begin_hook_(name.c_str(), profile_event_category::factory);
}

void ProfilerHook::on_linop_factory_generate_completed(const LinOpFactory* factory,
                                                       const LinOp*, const LinOp*) const
{
    const auto name = "generate(" + stringify(factory) + ")";
    end_hook_(name.c_str(), profile_event_category::factory);
}

// a solver that reports an unfinished iteration starts the next
// "iteration" range (profiler_hook.cpp: on_iteration_complete)
void ProfilerHook::on_iteration_complete(const LinOp* solver, int64 num_iterations,
                                         const std::vector<double>&,
                                         const std::vector<stopping_status>&,
                                         bool all_stopped) const
{
    if (num_iterations > 0 && is_iterative(solver) && !all_stopped) {
        end_hook_("iteration", profile_event_category::solver);
        begin_hook_("iteration", profile_event_category::solver);
    }
}

profiling_scope_guard ProfilerHook::user_range(const char* name) const
{
    return profiling_scope_guard{name, profile_event_category::user, begin_hook_, end_hook_};
}

// ----------------------------------------------------- the scope guard ----

profiling_scope_guard::profiling_scope_guard()
    : empty_{true}, category_{profile_event_category::internal}
{}

profiling_scope_guard::profiling_scope_guard(const char* name,
                                             profile_event_category category,
                                             ProfilerHook::hook_function begin,
                                             ProfilerHook::hook_function end)
    : empty_{false}, name_{name}, category_{category}, end_{std::move(end)}
{
    begin(name, category);
}

profiling_scope_guard::~profiling_scope_guard()
{
    if (empty_) return;
    // a destructor must not throw: a failing end function is reported
    try {
        end_(name_.c_str(), category_);
    } catch (const std::exception& e) {
        std::cerr << "WARNING: closing profiler range \"" << name_ << "\" failed: " << e.what()
                  << '\n';
    } catch (...) {
        std::cerr << "WARNING: closing profiler range \"" << name_ << "\" failed\n";
    }
}

profiling_scope_guard::profiling_scope_guard(profiling_scope_guard&& other) noexcept
    : empty_{other.empty_},
      name_{std::move(other.name_)},
      category_{other.category_},
      end_{std::move(other.end_)}
{
    other.empty_ = true;
}

profiling_scope_guard& profiling_scope_guard::operator=(profiling_scope_guard&& other) noexcept
{
    if (this != &other) {
        if (!empty_) end_(name_.c_str(), category_);
        empty_ = other.empty_;
        name_ = std::move(other.name_);
        category_ = other.category_;
        end_ = std::move(other.end_);
        other.empty_ = true;
    }
    return *this;
}

// ---------------------------------------------------------------- roctx ----

namespace {

// roctx is loaded at run time, so that the library does not depend on the
// profiler's libraries; rocprofv3 --marker-trace records these ranges
struct roctx_api {
    int (*push)(const char*) = nullptr;
    int (*pop)() = nullptr;
};

const roctx_api* load_roctx()
{
    static const roctx_api api = [] {
        roctx_api a;
        for (const char* lib : {"librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so",
                                "libroctx64.so.4", "libroctx64.so"}) {
            void* h = dlopen(lib, RTLD_NOW | RTLD_GLOBAL);
            if (!h) continue;
            a.push = reinterpret_cast<int (*)(const char*)>(dlsym(h, "roctxRangePushA"));
            a.pop = reinterpret_cast<int (*)()>(dlsym(h, "roctxRangePop"));
            if (a.push && a.pop) return a;
            a = roctx_api{};
        }
        return a;
    }();
    return api.push ? &api : nullptr;
}

}  // namespace

std::shared_ptr<ProfilerHook> ProfilerHook::create_custom(hook_function begin, hook_function end)
{
    if (!begin || !end) GKO_NOT_SUPPORTED_MSG("ProfilerHook: empty hook function");
    return std::shared_ptr<ProfilerHook>(new ProfilerHook(std::move(begin), std::move(end)));
}

std::shared_ptr<ProfilerHook> ProfilerHook::create_roctx()
{
    const roctx_api* api = load_roctx();
    if (!api) {
        GKO_NOT_SUPPORTED_MSG(
            "ProfilerHook::create_roctx: no roctx library (librocprofiler-sdk-roctx.so, "
            "libroctx64.so) could be loaded");
    }
    return create_custom([api](const char* name, profile_event_category) { api->push(name); },
                         [api](const char*, profile_event_category) { api->pop(); });
}

std::shared_ptr<ProfilerHook> ProfilerHook::create_for_executor(
    std::shared_ptr<const Executor> exec)
{
    if (!exec || exec->is_host()) {
        GKO_NOT_SUPPORTED_MSG(
            "ProfilerHook::create_for_executor: the host executor has no profiler here "
            "(TAU and VTune are not part of this library)");
    }
    return create_roctx();
}

// ------------------------------------------------------------ summaries ----

namespace {

using cpu_clock = std::chrono::steady_clock;

// What the flat and the nested summary share: the timer and a pool of its
// time points, the overhead of the bookkeeping, and the checks of a pop
// against the stack of open ranges.
struct summary_state {
    explicit summary_state(std::shared_ptr<Timer> t) : timer{std::move(t)}
    {
        for (int i = 0; i < 8; ++i) spare.push_back(timer->create_time_point());
    }
    time_point stamp()
    {
        time_point t = [this] {
            if (spare.empty()) return timer->create_time_point();
            time_point p = std::move(spare.back());
            spare.pop_back();
            return p;
        }();
        timer->record(t);
        return t;
    }
    void recycle(time_point&& t) { spare.push_back(std::move(t)); }
    // false (and broken from now on) if the pop does not match the stack
    bool pop_allowed(size_type depth, const std::string& top, const char* name,
                     bool allow_root)
    {
        if (broken) return false;
        const char* problem = nullptr;
        if (depth == 0) {
            problem = "a range was closed that was never opened";
        } else if (depth == 1 && !allow_root) {
            problem = "the root range was closed while the program ran";
        } else if (check_nesting && top != name) {
            problem = "ranges are not properly nested";
        }
        if (!problem) return true;
        std::cerr << "WARNING: ProfilerHook summary: " << problem << " (closing \"" << name
                  << "\"); the summary is incomplete.  Create the hook in the outermost "
                     "scope that uses the library.\n";
        broken = true;
        return false;
    }

    std::shared_ptr<Timer> timer;
    std::vector<time_point> spare;
    std::chrono::nanoseconds overhead{0};
    bool broken{false};
    bool check_nesting{false};
    std::mutex mutex;
};

struct flat_summary : summary_state {
    struct open_range {
        int64 id;
        time_point start;
    };
    explicit flat_summary(std::shared_ptr<Timer> t) : summary_state{std::move(t)}
    {
        push("total");
    }
    int64 id_of(const char* name)
    {
        auto it = ids.find(name);
        if (it != ids.end()) return it->second;
        const auto id = static_cast<int64>(entries.size());
        ids.emplace(name, id);
        entries.push_back(ProfilerHook::summary_entry{name, {}, {}, 0});
        return id;
    }
    void push(const char* name)
    {
        const auto t0 = cpu_clock::now();
        std::lock_guard<std::mutex> g{mutex};
        if (broken) return;
        const auto id = id_of(name);
        stack.push_back(open_range{id, stamp()});
        overhead += cpu_clock::now() - t0;
    }
    void pop(const char* name, bool allow_root = false)
    {
        const auto t0 = cpu_clock::now();
        std::lock_guard<std::mutex> g{mutex};
        const std::string top = stack.empty() ? std::string{} : entries[stack.back().id].name;
        if (!pop_allowed(stack.size(), top, name, allow_root)) return;
        time_point now = stamp();
        open_range r = std::move(stack.back());
        stack.pop_back();
        const auto t1 = cpu_clock::now();
        timer->wait(now);  // waiting for the device is not overhead
        const auto t2 = cpu_clock::now();
        const auto elapsed = timer->difference_async(r.start, now);
        recycle(std::move(r.start));
        recycle(std::move(now));
        auto& e = entries[r.id];
        e.count++;
        e.inclusive += elapsed;
        e.exclusive += elapsed;
        if (!stack.empty()) entries[stack.back().id].exclusive -= elapsed;
        overhead += (t1 - t0) + (cpu_clock::now() - t2);
    }
    void close_all()
    {
        while (!stack.empty() && !broken) {
            const std::string top = entries[stack.back().id].name;
            pop(top.c_str(), true);
        }
    }

    std::vector<open_range> stack;
    std::unordered_map<std::string, int64> ids;
    std::vector<ProfilerHook::summary_entry> entries;
};

struct tree_summary : summary_state {
    struct node {
        int64 name;
        int64 parent;
        std::chrono::nanoseconds elapsed{0};
        int64 count{};
    };
    struct open_range {
        int64 node;
        time_point start;
    };
    explicit tree_summary(std::shared_ptr<Timer> t) : summary_state{std::move(t)}
    {
        push("total");
    }
    int64 name_of(const char* name)
    {
        auto it = name_ids.find(name);
        if (it != name_ids.end()) return it->second;
        const auto id = static_cast<int64>(names.size());
        name_ids.emplace(name, id);
        names.emplace_back(name);
        return id;
    }
    int64 parent() const { return stack.empty() ? -1 : stack.back().node; }
    void push(const char* name)
    {
        const auto t0 = cpu_clock::now();
        std::lock_guard<std::mutex> g{mutex};
        if (broken) return;
        const auto key = std::make_pair(name_of(name), parent());
        auto it = node_ids.find(key);
        if (it == node_ids.end()) {
            it = node_ids.emplace(key, static_cast<int64>(nodes.size())).first;
            nodes.push_back(node{key.first, key.second});
        }
        stack.push_back(open_range{it->second, stamp()});
        overhead += cpu_clock::now() - t0;
    }
    void pop(const char* name, bool allow_root = false)
    {
        const auto t0 = cpu_clock::now();
        std::lock_guard<std::mutex> g{mutex};
        const std::string top =
            stack.empty() ? std::string{} : names[nodes[stack.back().node].name];
        if (!pop_allowed(stack.size(), top, name, allow_root)) return;
        time_point now = stamp();
        open_range r = std::move(stack.back());
        stack.pop_back();
        const auto t1 = cpu_clock::now();
        timer->wait(now);
        const auto t2 = cpu_clock::now();
        const auto elapsed = timer->difference_async(r.start, now);
        recycle(std::move(r.start));
        recycle(std::move(now));
        nodes[r.node].count++;
        nodes[r.node].elapsed += elapsed;
        overhead += (t1 - t0) + (cpu_clock::now() - t2);
    }
    void close_all()
    {
        while (!stack.empty() && !broken) {
            const std::string top = names[nodes[stack.back().node].name];
            pop(top.c_str(), true);
        }
    }
    // the tree under node 0 ("total"); children in the order they were
    // first opened under their parent
    ProfilerHook::nested_summary_entry tree() const
    {
        std::vector<std::vector<int64>> children(nodes.size());
        for (size_type i = 0; i < nodes.size(); ++i) {
            if (nodes[i].parent >= 0) children[nodes[i].parent].push_back(static_cast<int64>(i));
        }
        auto build = [&](auto& self, int64 id) -> ProfilerHook::nested_summary_entry {
            ProfilerHook::nested_summary_entry e{names[nodes[id].name], nodes[id].elapsed,
                                                 nodes[id].count, {}};
            for (auto c : children[id]) e.children.push_back(self(self, c));
            return e;
        };
        if (nodes.empty()) return {};
        return build(build, 0);
    }

    std::vector<open_range> stack;
    std::unordered_map<std::string, int64> name_ids;
    std::vector<std::string> names;
    std::map<std::pair<int64, int64>, int64> node_ids;
    std::vector<node> nodes;
};

}  // namespace

std::shared_ptr<ProfilerHook> ProfilerHook::create_summary(std::shared_ptr<Timer> timer,
                                                           std::unique_ptr<SummaryWriter> writer,
                                                           bool debug_check_nesting)
{
    if (!timer || !writer) GKO_NOT_SUPPORTED_MSG("ProfilerHook::create_summary: null argument");
    // the state lives as long as the hook's functions; its deleter closes
    // what is still open and hands the table to the writer
    std::shared_ptr<SummaryWriter> w{std::move(writer)};
    std::shared_ptr<flat_summary> data{new flat_summary{std::move(timer)}, [w](flat_summary* s) {
                                           s->close_all();
                                           w->write(s->entries, s->overhead);
                                           delete s;
                                       }};
    data->check_nesting = debug_check_nesting;
    return create_custom([data](const char* name, profile_event_category) { data->push(name); },
                         [data](const char* name, profile_event_category) { data->pop(name); });
}

std::shared_ptr<ProfilerHook> ProfilerHook::create_nested_summary(
    std::shared_ptr<Timer> timer, std::unique_ptr<NestedSummaryWriter> writer,
    bool debug_check_nesting)
{
    if (!timer || !writer) {
        GKO_NOT_SUPPORTED_MSG("ProfilerHook::create_nested_summary: null argument");
    }
    std::shared_ptr<NestedSummaryWriter> w{std::move(writer)};
    std::shared_ptr<tree_summary> data{new tree_summary{std::move(timer)}, [w](tree_summary* s) {
                                           s->close_all();
                                           w->write_nested(s->tree(), s->overhead);
                                           delete s;
                                       }};
    data->check_nesting = debug_check_nesting;
    return create_custom([data](const char* name, profile_event_category) { data->push(name); },
                         [data](const char* name, profile_event_category) { data->pop(name); });
}

// ---------------------------------------------------------------- tables ----

namespace {

// One decimal in the largest unit the value reaches: ns below 1 us, us
// below 1 ms, ms below 1 s, then s, minutes, hours and days (the last three
// padded to two characters so the columns align).
std::string format_duration(std::chrono::nanoseconds t)
{
    struct unit {
        int64 max_ns;
        double ns_per_unit;
        const char* name;
    };
    static const unit units[] = {
        {999, 1.0, "ns"},
        {999'999, 1e3, "us"},
        {999'999'999, 1e6, "ms"},
        {59'999'999'999, 1e9, "s "},
        {3'599'999'999'999, 60e9, "m "},
        {86'399'999'999'999, 3600e9, "h "},
        {std::numeric_limits<int64>::max(), 86400e9, "d "},
    };
    const int64 ns = t.count();
    const unit* u = &units[0];
    for (const auto& c : units) {
        u = &c;
        if (ns <= c.max_ns) break;
    }
    std::ostringstream ss;
    ss << std::fixed << std::setprecision(1) << double(ns) / u->ns_per_unit << ' ' << u->name;
    return ss.str();
}

std::string format_avg(std::chrono::nanoseconds t, int64 count)
{
    return format_duration(std::chrono::nanoseconds{t.count() / std::max<int64>(count, 1)});
}

std::string format_fraction(std::chrono::nanoseconds part, std::chrono::nanoseconds whole)
{
    std::ostringstream ss;
    ss << std::fixed << std::setprecision(1)
       << 100.0 * double(part.count()) / double(whole.count()) << " %";
    return ss.str();
}

// Markdown table: headers centred, the first column left-aligned and the
// others right-aligned (":" in the separator row)
template <size_type N>
void print_table(std::ostream& out, const std::array<std::string, N>& headers,
                 const std::vector<std::array<std::string, N>>& rows)
{
    std::array<size_type, N> w;
    for (size_type i = 0; i < N; ++i) {
        w[i] = headers[i].size();
        for (const auto& r : rows) w[i] = std::max(w[i], r[i].size());
    }
    for (size_type i = 0; i < N; ++i) {
        const auto pad = w[i] - headers[i].size();
        out << '|' << std::string(pad / 2, ' ') << headers[i] << std::string(pad - pad / 2, ' ');
    }
    out << "|\n";
    for (size_type i = 0; i < N; ++i) {
        out << '|' << std::string(w[i] - 1, '-') << (i == 0 ? '-' : ':');
    }
    out << "|\n";
    for (const auto& r : rows) {
        for (size_type i = 0; i < N; ++i) {
            const auto pad = std::string(w[i] - r[i].size(), ' ');
            out << '|' << (i == 0 ? r[i] + pad : pad + r[i]);
        }
        out << "|\n";
    }
}

}  // namespace

ProfilerHook::TableSummaryWriter::TableSummaryWriter(std::ostream& output, std::string header)
    : output_{&output}, header_{std::move(header)}
{}

void ProfilerHook::TableSummaryWriter::write(const std::vector<summary_entry>& entries,
                                             std::chrono::nanoseconds overhead)
{
    *output_ << header_ << '\n' << "Overhead estimate " << format_duration(overhead) << '\n';
    auto sorted = entries;
    std::stable_sort(sorted.begin(), sorted.end(),
                     [](const summary_entry& a, const summary_entry& b) {
                         return a.inclusive > b.inclusive;
                     });
    std::vector<std::array<std::string, 6>> rows;
    for (const auto& e : sorted) {
        rows.push_back({" " + e.name + " ", " " + format_duration(e.inclusive) + " ",
                        " " + format_duration(e.exclusive) + " ",
                        " " + std::to_string(e.count) + " ",
                        " " + format_avg(e.inclusive, e.count) + " ",
                        " " + format_avg(e.exclusive, e.count) + " "});
    }
    print_table<6>(*output_,
                   {" name ", " total ", " total (self) ", " count ", " avg ", " avg (self) "},
                   rows);
}

// Each node is followed by its children, longest first, indented by two
// spaces per level; a node with children also gets a "(self)" row for the
// time outside them, placed before the first child that took less.
void ProfilerHook::TableSummaryWriter::write_nested(const nested_summary_entry& root,
                                                    std::chrono::nanoseconds overhead)
{
    *output_ << header_ << '\n' << "Overhead estimate " << format_duration(overhead) << '\n';
    std::vector<std::array<std::string, 5>> rows;
    auto visit = [&rows](auto& self, const nested_summary_entry& n,
                         std::chrono::nanoseconds parent, size_type depth) -> void {
        rows.push_back({std::string(2 * depth + 1, ' ') + n.name + " ",
                        " " + format_duration(n.elapsed) + " ",
                        format_fraction(n.elapsed, parent) + " ",
                        " " + std::to_string(n.count) + " ", " " + format_avg(n.elapsed, n.count) + " "});
        std::chrono::nanoseconds in_children{0};
        for (const auto& c : n.children) in_children += c.elapsed;
        const nested_summary_entry own{"(self)", n.elapsed - in_children, n.count, {}};
        std::vector<const nested_summary_entry*> order;
        for (const auto& c : n.children) order.push_back(&c);
        std::stable_sort(order.begin(), order.end(),
                         [](const nested_summary_entry* a, const nested_summary_entry* b) {
                             return a->elapsed > b->elapsed;
                         });
        bool own_done = n.children.empty();
        for (const auto* c : order) {
            if (!own_done && c->elapsed < own.elapsed) {
                self(self, own, n.elapsed, depth + 1);
                own_done = true;
            }
            self(self, *c, n.elapsed, depth + 1);
        }
        if (!own_done) self(self, own, n.elapsed, depth + 1);
    };
    visit(visit, root, root.elapsed, 0);
    print_table<5>(*output_, {" name ", " total ", " fraction ", " count ", " avg "}, rows);
}

}  // namespace log
}  // namespace gko
