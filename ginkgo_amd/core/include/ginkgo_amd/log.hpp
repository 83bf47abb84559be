// Loggers: the event interface of include/ginkgo/core/log/logger.hpp:132-540
// (the events this library emits, each a no-op by default) and
// log::Convergence (include/ginkgo/core/log/convergence.hpp:37,
// core/log/convergence.cpp:52-110).
//
// Events and where they come from:
//   allocation / free / copy      Executor::alloc, free and copy_from
//   linop (advanced) apply        LinOp::apply
//   linop factory generate        LinOpFactory::generate
//   iteration complete            the Krylov solvers and Ir, once per apply
//                                 after the device loop has stopped (the
//                                 iterations run without a host round trip,
//                                 so there is no per-iteration host event)
// A logger attached to an executor receives that executor's events; if its
// needs_propagation() is true it also receives the events of every object
// on the executor (logger.hpp:826-862, executor.hpp should_propagate_log).
#pragma once

#include <cstdint>
#include <vector>

#include "ginkgo_amd/base.hpp"

namespace gko {

class LinOp;
class LinOpFactory;

namespace log {

class Logger {
public:
    virtual ~Logger() = default;

    // executor events (logger.hpp on_allocation_* / on_free_* / on_copy_*)
    virtual void on_allocation_started(const Executor* exec, const size_type& num_bytes) const {}
    virtual void on_allocation_completed(const Executor* exec, const size_type& num_bytes,
                                         const uintptr& location) const
    {}
    virtual void on_free_started(const Executor* exec, const uintptr& location) const {}
    virtual void on_free_completed(const Executor* exec, const uintptr& location) const {}
    virtual void on_copy_started(const Executor* from, const Executor* to,
                                 const uintptr& location_from, const uintptr& location_to,
                                 const size_type& num_bytes) const
    {}
    virtual void on_copy_completed(const Executor* from, const Executor* to,
                                   const uintptr& location_from, const uintptr& location_to,
                                   const size_type& num_bytes) const
    {}

    // linop events (logger.hpp on_linop_apply_* / on_linop_advanced_apply_*)
    virtual void on_linop_apply_started(const LinOp* A, const LinOp* b, const LinOp* x) const {}
    virtual void on_linop_apply_completed(const LinOp* A, const LinOp* b, const LinOp* x) const
    {}
    virtual void on_linop_advanced_apply_started(const LinOp* A, const LinOp* alpha,
                                                 const LinOp* b, const LinOp* beta,
                                                 const LinOp* x) const
    {}
    virtual void on_linop_advanced_apply_completed(const LinOp* A, const LinOp* alpha,
                                                   const LinOp* b, const LinOp* beta,
                                                   const LinOp* x) const
    {}

    // factory events (logger.hpp on_linop_factory_generate_*)
    virtual void on_linop_factory_generate_started(const LinOpFactory* factory,
                                                   const LinOp* input) const
    {}
    virtual void on_linop_factory_generate_completed(const LinOpFactory* factory,
                                                     const LinOp* input,
                                                     const LinOp* output) const
    {}

    // solver event: the per-column residual norms and stopping status after
    // `num_iterations` iterations; all_stopped when every column has stopped
    virtual void on_iteration_complete(const LinOp* solver, int64 num_iterations,
                                       const std::vector<double>& residual_norm,
                                       const std::vector<stopping_status>& status,
                                       bool all_stopped) const
    {}

    // logger.hpp:712: forward the events of all objects on an executor
    // this logger is attached to
    virtual bool needs_propagation() const { return false; }
};

// log::Convergence: iteration count, convergence flag and the final
// residual norm of each right-hand side.
class Convergence : public Logger {
public:
    static std::shared_ptr<Convergence> create() { return std::make_shared<Convergence>(); }
    void on_iteration_complete(const LinOp*, int64 num_iterations,
                               const std::vector<double>& residual_norm,
                               const std::vector<stopping_status>& status,
                               bool all_stopped) const override
    {
        if (!all_stopped) return;
        num_iterations_ = num_iterations;
        residual_norm_ = residual_norm;
        converged_ = !status.empty();
        for (auto s : status) converged_ = converged_ && s.has_converged();
        status_ = status;
    }
    int64 get_num_iterations() const { return num_iterations_; }
    bool has_converged() const { return converged_; }
    const std::vector<double>& get_residual_norm() const { return residual_norm_; }
    const std::vector<stopping_status>& get_status() const { return status_; }

private:
    mutable int64 num_iterations_{0};
    mutable bool converged_{false};
    mutable std::vector<double> residual_norm_;
    mutable std::vector<stopping_status> status_;
};

}  // namespace log

// Calls f(logger) for the loggers of an object on `exec` that must see an
// event: the propagating loggers of the executor, then the object's own
// (EnableLogging::log, logger.hpp:856-862).
template <typename F>
void log_event(const Executor* exec,
               const std::vector<std::shared_ptr<const log::Logger>>& own, F&& f)
{
    if (exec && exec->should_propagate_log()) {
        for (auto& l : exec->get_loggers()) {
            if (l->needs_propagation()) f(*l);
        }
    }
    for (auto& l : own) f(*l);
}

}  // namespace gko
