// log::ProfilerHook (include/ginkgo/core/log/profiler_hook.hpp:40-430,
// core/log/profiler_hook.cpp, profiler_hook_summary.cpp,
// profiler_hook_summary_writer.cpp): turns the logger events into nested
// named ranges, handed to a pair of begin / end functions.
//   create_custom          any pair of functions
//   create_roctx           roctx ranges (roctxRangePushA / roctxRangePop),
//                          visible to rocprofv3 --marker-trace
//   create_for_executor    roctx for a device executor
//   create_summary         flat table of inclusive / exclusive time and
//                          count per range name
//   create_nested_summary  the same as a tree of range paths
// The range names follow the reference: "allocate", "free", "copy",
// "apply(A * b = x)", "advanced_apply(alpha * A * b + beta * x)",
// "generate(factory)", and "iteration" inside the apply of an iterative
// solver.  An object is named by set_object_name, else by its demangled
// dynamic type.
#pragma once

#include <chrono>
#include <functional>
#include <iostream>
#include <string>
#include <unordered_map>

#include "ginkgo_amd/lin_op.hpp"
#include "ginkgo_amd/timer.hpp"

namespace gko {
namespace log {

enum class profile_event_category {
    memory,
    operation,
    object,
    linop,
    factory,
    solver,
    criterion,
    user,
    internal,
};

class profiling_scope_guard;

class ProfilerHook : public Logger {
public:
    using hook_function = std::function<void(const char*, profile_event_category)>;

    void on_allocation_started(const Executor* exec, const size_type& num_bytes) const override;
    void on_allocation_completed(const Executor* exec, const size_type& num_bytes,
                                 const uintptr& location) const override;
    void on_free_started(const Executor* exec, const uintptr& location) const override;
    void on_free_completed(const Executor* exec, const uintptr& location) const override;
    void on_copy_started(const Executor* from, const Executor* to, const uintptr& location_from,
                         const uintptr& location_to, const size_type& num_bytes) const override;
    void on_copy_completed(const Executor* from, const Executor* to,
                           const uintptr& location_from, const uintptr& location_to,
                           const size_type& num_bytes) const override;
    void on_linop_apply_started(const LinOp* A, const LinOp* b, const LinOp* x) const override;
    void on_linop_apply_completed(const LinOp* A, const LinOp* b, const LinOp* x) const override;
    void on_linop_advanced_apply_started(const LinOp* A, const LinOp* alpha, const LinOp* b,
                                         const LinOp* beta, const LinOp* x) const override;
    void on_linop_advanced_apply_completed(const LinOp* A, const LinOp* alpha, const LinOp* b,
                                           const LinOp* beta, const LinOp* x) const override;
    void on_linop_factory_generate_started(const LinOpFactory* factory,
                                           const LinOp* input) const override;
    void on_linop_factory_generate_completed(const LinOpFactory* factory, const LinOp* input,
                                             const LinOp* output) const override;
    void on_iteration_complete(const LinOp* solver, int64 num_iterations,
                               const std::vector<double>& residual_norm,
                               const std::vector<stopping_status>& status,
                               bool all_stopped) const override;
    // the hook sees every object on the executors it is attached to
    bool needs_propagation() const override { return true; }

    // names an object in the range names instead of its type
    void set_object_name(const LinOp* obj, std::string name);
    void set_object_name(const LinOpFactory* obj, std::string name);
    template <typename T>
    void set_object_name(const std::shared_ptr<T>& obj, std::string name)
    {
        set_object_name(obj.get(), std::move(name));
    }
    // synchronize the executor before each memory / copy range opens and
    // closes, so that the ranges time device work, not its enqueueing
    void set_synchronization(bool synchronize) { synchronize_ = synchronize; }
    // a range of the user's own, closed when the guard goes out of scope
    profiling_scope_guard user_range(const char* name) const;

    // one row of create_summary: time with (inclusive) and without
    // (exclusive) the nested ranges, summed over `count` occurrences
    struct summary_entry {
        std::string name;
        std::chrono::nanoseconds inclusive{0};
        std::chrono::nanoseconds exclusive{0};
        int64 count{};
    };
    // a node of create_nested_summary: a range name under its parents
    struct nested_summary_entry {
        std::string name;
        std::chrono::nanoseconds elapsed{0};
        int64 count{};
        std::vector<nested_summary_entry> children;
    };
    class SummaryWriter {
    public:
        virtual ~SummaryWriter() = default;
        virtual void write(const std::vector<summary_entry>& entries,
                           std::chrono::nanoseconds overhead) = 0;
    };
    class NestedSummaryWriter {
    public:
        virtual ~NestedSummaryWriter() = default;
        virtual void write_nested(const nested_summary_entry& root,
                                  std::chrono::nanoseconds overhead) = 0;
    };
    // Markdown table (profiler_hook_summary_writer.cpp)
    class TableSummaryWriter : public SummaryWriter, public NestedSummaryWriter {
    public:
        explicit TableSummaryWriter(std::ostream& output = std::cerr,
                                    std::string header = "Runtime summary");
        void write(const std::vector<summary_entry>& entries,
                   std::chrono::nanoseconds overhead) override;
        void write_nested(const nested_summary_entry& root,
                          std::chrono::nanoseconds overhead) override;

    private:
        std::ostream* output_;
        std::string header_;
    };

    static std::shared_ptr<ProfilerHook> create_custom(hook_function begin, hook_function end);
    static std::shared_ptr<ProfilerHook> create_roctx();
    static std::shared_ptr<ProfilerHook> create_for_executor(std::shared_ptr<const Executor> exec);
    // The writer receives the table when the hook (and every copy of its
    // functions) is destroyed; ranges still open then are closed first.
    static std::shared_ptr<ProfilerHook> create_summary(
        std::shared_ptr<Timer> timer = std::make_shared<CpuTimer>(),
        std::unique_ptr<SummaryWriter> writer = std::make_unique<TableSummaryWriter>(),
        bool debug_check_nesting = false);
    static std::shared_ptr<ProfilerHook> create_nested_summary(
        std::shared_ptr<Timer> timer = std::make_shared<CpuTimer>(),
        std::unique_ptr<NestedSummaryWriter> writer = std::make_unique<TableSummaryWriter>(),
        bool debug_check_nesting = false);

private:
    ProfilerHook(hook_function begin, hook_function end);
    void maybe_synchronize(const Executor* exec) const;
    std::string stringify(const LinOp* obj) const;
    std::string stringify(const LinOpFactory* obj) const;

    std::unordered_map<const void*, std::string> name_map_;
    bool synchronize_{false};
    hook_function begin_hook_;
    hook_function end_hook_;
};

// Calls end(name) when it goes out of scope (profiler_hook.hpp:430-470).
class profiling_scope_guard {
public:
    profiling_scope_guard();
    profiling_scope_guard(const char* name, profile_event_category category,
                          ProfilerHook::hook_function begin, ProfilerHook::hook_function end);
    ~profiling_scope_guard();
    profiling_scope_guard(profiling_scope_guard&& other) noexcept;
    profiling_scope_guard& operator=(profiling_scope_guard&& other) noexcept;
    profiling_scope_guard(const profiling_scope_guard&) = delete;
    profiling_scope_guard& operator=(const profiling_scope_guard&) = delete;

private:
    bool empty_;
    std::string name_;
    profile_event_category category_;
    ProfilerHook::hook_function end_;
};

}  // namespace log
}  // namespace gko
