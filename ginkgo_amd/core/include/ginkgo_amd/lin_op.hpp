// LinOp, LinOpFactory, Dense and Identity: include/ginkgo/core/base/lin_op.hpp
// (apply / apply_impl, conformance checks :300-323, generate :384-408) and
// include/ginkgo/core/matrix/{dense,identity}.hpp, re-designed over the
// gkomi C ABI.
#pragma once

#include <functional>
#include <limits>
#include <type_traits>

#include "ginkgo_amd/base.hpp"
#include "ginkgo_amd/log.hpp"

namespace gko {

class LinOp : public std::enable_shared_from_this<LinOp> {
public:
    virtual ~LinOp() = default;

    // x = op(b), between the linop_apply events (lin_op.hpp:60-80)
    LinOp* apply(const LinOp* b, LinOp* x)
    {
        static_cast<const LinOp*>(this)->apply(b, x);
        return this;
    }
    const LinOp* apply(const LinOp* b, LinOp* x) const
    {
        check_apply(b, x);
        log_event(exec_.get(), loggers_,
                  [&](const log::Logger& l) { l.on_linop_apply_started(this, b, x); });
        apply_impl(b, x);
        log_event(exec_.get(), loggers_,
                  [&](const log::Logger& l) { l.on_linop_apply_completed(this, b, x); });
        return this;
    }
    // x = alpha op(b) + beta x, between the linop_advanced_apply events
    LinOp* apply(const LinOp* alpha, const LinOp* b, const LinOp* beta, LinOp* x)
    {
        static_cast<const LinOp*>(this)->apply(alpha, b, beta, x);
        return this;
    }
    const LinOp* apply(const LinOp* alpha, const LinOp* b, const LinOp* beta, LinOp* x) const
    {
        check_apply(b, x);
        check_scalar(alpha, b);
        check_scalar(beta, x);
        log_event(exec_.get(), loggers_, [&](const log::Logger& l) {
            l.on_linop_advanced_apply_started(this, alpha, b, beta, x);
        });
        apply_impl(alpha, b, beta, x);
        log_event(exec_.get(), loggers_, [&](const log::Logger& l) {
            l.on_linop_advanced_apply_completed(this, alpha, b, beta, x);
        });
        return this;
    }
    template <typename B, typename X>
    auto apply(const std::shared_ptr<B>& b, const std::shared_ptr<X>& x) const
    {
        return apply(b.get(), x.get());
    }
    template <typename B, typename X>
    auto apply(const std::unique_ptr<B>& b, const std::unique_ptr<X>& x) const
    {
        return apply(b.get(), x.get());
    }

    const dim<2>& get_size() const noexcept { return size_; }
    std::shared_ptr<const Executor> get_executor() const noexcept { return exec_; }

    // log::EnableLogging (include/ginkgo/core/log/logger.hpp:786-817)
    void add_logger(std::shared_ptr<const log::Logger> logger) const
    {
        loggers_.push_back(std::move(logger));
    }
    void remove_logger(const log::Logger* logger) const
    {
        for (auto it = loggers_.begin(); it != loggers_.end(); ++it) {
            if (it->get() == logger) {
                loggers_.erase(it);
                return;
            }
        }
        throw Error(__FILE__, __LINE__, "remove_logger: the logger is not attached");
    }
    void clear_loggers() const { loggers_.clear(); }
    const std::vector<std::shared_ptr<const log::Logger>>& get_loggers() const
    {
        return loggers_;
    }

    // A solver that reads x as its initial guess (solver_base.hpp).
    virtual bool apply_uses_initial_guess() const { return false; }

protected:
    LinOp(std::shared_ptr<const Executor> exec, dim<2> size = {})
        : exec_{std::move(exec)}, size_{size}
    {}
    void set_size(dim<2> s) { size_ = s; }
    virtual void apply_impl(const LinOp* b, LinOp* x) const = 0;
    virtual void apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                            LinOp* x) const = 0;

    // GKO_ASSERT_CONFORMANT / EQUAL_ROWS / EQUAL_COLS (lin_op.hpp:300-323)
    void check_apply(const LinOp* b, LinOp* x) const
    {
        if (!b || !x) {
            throw NotSupported(__FILE__, __LINE__, "apply: null operand");
        }
        if (size_.cols != b->get_size().rows) {
            throw DimensionMismatch(__FILE__, __LINE__,
                                    "apply: operator columns != rows of b");
        }
        if (size_.rows != x->get_size().rows) {
            throw DimensionMismatch(__FILE__, __LINE__, "apply: operator rows != rows of x");
        }
        if (b->get_size().cols != x->get_size().cols) {
            throw DimensionMismatch(__FILE__, __LINE__, "apply: columns of b != columns of x");
        }
    }
    static void check_scalar(const LinOp* s, const LinOp* v)
    {
        if (!s || s->get_size().rows != 1 ||
            (s->get_size().cols != 1 && s->get_size().cols != v->get_size().cols)) {
            throw DimensionMismatch(__FILE__, __LINE__, "apply: alpha/beta must be 1x1 or 1xk");
        }
    }

private:
    std::shared_ptr<const Executor> exec_;
    dim<2> size_;
    mutable std::vector<std::shared_ptr<const log::Logger>> loggers_;
};

class LinOpFactory {
public:
    virtual ~LinOpFactory() = default;
    // LinOpFactory::generate (lin_op.hpp:384-408), between the
    // linop_factory_generate events
    std::unique_ptr<LinOp> generate(std::shared_ptr<const LinOp> input) const
    {
        const LinOp* in = input.get();
        log_event(exec_.get(), loggers_, [&](const log::Logger& l) {
            l.on_linop_factory_generate_started(this, in);
        });
        auto out = generate_impl(std::move(input));
        log_event(exec_.get(), loggers_, [&](const log::Logger& l) {
            l.on_linop_factory_generate_completed(this, in, out.get());
        });
        return out;
    }
    std::shared_ptr<const Executor> get_executor() const { return exec_; }
    void add_logger(std::shared_ptr<const log::Logger> logger) const
    {
        loggers_.push_back(std::move(logger));
    }
    void clear_loggers() const { loggers_.clear(); }
    const std::vector<std::shared_ptr<const log::Logger>>& get_loggers() const
    {
        return loggers_;
    }

protected:
    explicit LinOpFactory(std::shared_ptr<const Executor> exec) : exec_{std::move(exec)} {}
    virtual std::unique_ptr<LinOp> generate_impl(std::shared_ptr<const LinOp> input) const = 0;

private:
    std::shared_ptr<const Executor> exec_;
    mutable std::vector<std::shared_ptr<const log::Logger>> loggers_;
};

namespace solver {

// Marks the iterative solvers (solver::IterativeBase, include/ginkgo/core/
// solver/solver_base.hpp): a profiler opens an "iteration" range inside
// their applies.
class IterativeBase {
public:
    virtual ~IterativeBase() = default;
};

}  // namespace solver

// A factory parameter that is resolved on the executor passed to .on(exec)
// (deferred_factory_parameter, include/ginkgo/core/base/abstract_factory.hpp:521-604).
template <typename FactoryType>
class deferred_factory_parameter {
public:
    deferred_factory_parameter() = default;
    deferred_factory_parameter(std::nullptr_t) {}
    template <typename F, typename = std::enable_if_t<std::is_convertible<
                              std::shared_ptr<F>, std::shared_ptr<const FactoryType>>::value>>
    deferred_factory_parameter(std::shared_ptr<F> f)
        : gen_{[f](std::shared_ptr<const Executor>) {
              return std::shared_ptr<const FactoryType>(f);
          }}
    {}
    template <typename F, typename = std::enable_if_t<std::is_convertible<
                              std::shared_ptr<F>, std::shared_ptr<const FactoryType>>::value>>
    deferred_factory_parameter(std::unique_ptr<F> f)
        : deferred_factory_parameter(std::shared_ptr<F>(std::move(f)))
    {}
    // parameters objects with an .on(exec) method
    template <typename P, typename = decltype(std::declval<const P&>().on(
                              std::declval<std::shared_ptr<const Executor>>()))>
    deferred_factory_parameter(P params)
        : gen_{[params](std::shared_ptr<const Executor> exec) {
              return std::shared_ptr<const FactoryType>(params.on(exec));
          }}
    {}
    std::shared_ptr<const FactoryType> on(std::shared_ptr<const Executor> exec) const
    {
        return gen_ ? gen_(std::move(exec)) : nullptr;
    }
    explicit operator bool() const { return static_cast<bool>(gen_); }

private:
    std::function<std::shared_ptr<const FactoryType>(std::shared_ptr<const Executor>)> gen_;
};

namespace matrix {

// Dense<V>: row-major n x k block with a stride (matrix::view::dense,
// device_views.hpp:24-45; dense.hpp:89).
template <typename ValueType = double>
class Dense : public LinOp {
public:
    using value_type = ValueType;
    static std::unique_ptr<Dense> create(std::shared_ptr<const Executor> exec, dim<2> size = {},
                                         size_type stride = 0)
    {
        return std::unique_ptr<Dense>(
            new Dense(std::move(exec), size, stride ? stride : size.cols));
    }
    // non-owning view over existing memory of `exec`
    static std::unique_ptr<Dense> create_view(std::shared_ptr<const Executor> exec, dim<2> size,
                                              ValueType* values, size_type stride)
    {
        auto d = std::unique_ptr<Dense>(new Dense(exec, {}, 0));
        d->set_size(size);
        d->stride_ = stride;
        d->values_ = array<ValueType>::view(exec, size.rows ? (size.rows - 1) * stride + size.cols
                                                            : 0,
                                            values);
        return d;
    }
    // host data (row-major) copied to exec
    static std::unique_ptr<Dense> create_from_host(std::shared_ptr<const Executor> exec,
                                                   dim<2> size,
                                                   const std::vector<ValueType>& host)
    {
        if (host.size() != size.rows * size.cols) {
            throw DimensionMismatch(__FILE__, __LINE__, "host data size");
        }
        // one allocation, filled by one copy
        auto d = std::unique_ptr<Dense>(new Dense(exec, {}, 0));
        d->set_size(size);
        d->stride_ = size.cols;
        d->values_ = array<ValueType>(exec, host);
        return d;
    }
    static std::unique_ptr<Dense> create_with_config_of(const Dense* other)
    {
        return create(other->get_executor(), other->get_size(), other->get_stride());
    }

    ValueType* get_values() noexcept { return values_.get_data(); }
    const ValueType* get_const_values() const noexcept { return values_.get_const_data(); }
    size_type get_stride() const noexcept { return stride_; }
    size_type get_num_stored_elements() const noexcept { return values_.get_size(); }
    std::vector<ValueType> to_host() const;   // row-major, stride removed
    std::unique_ptr<Dense> clone() const;

    void fill(ValueType value);
    // Dense::scale / add_scaled / compute_dot / compute_norm2 (dense.cpp:400-466)
    void scale(const Dense* alpha);
    void add_scaled(const Dense* alpha, const Dense* b);
    void compute_dot(const Dense* b, Dense* result) const;
    void compute_conj_dot(const Dense* b, Dense* result) const { compute_dot(b, result); }
    void compute_norm2(Dense* result) const;
    // copy with precision conversion (Dense::convert_to, dense.cpp)
    template <typename Other>
    void copy_from(const Dense<Other>* other);
    void copy_from(const Dense* other);

protected:
    Dense(std::shared_ptr<const Executor> exec, dim<2> size, size_type stride)
        : LinOp(exec, size),
          values_(exec, size.rows ? (size.rows - 1) * stride + size.cols : 0),
          stride_{stride}
    {}
    void apply_impl(const LinOp* b, LinOp* x) const override;
    void apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                    LinOp* x) const override;

private:
    array<ValueType> values_;
    size_type stride_{};
    mutable array<char> workspace_;
    template <typename>
    friend class Dense;

public:
    char* reduction_workspace(size_type bytes) const;
};

// matrix::Identity (include/ginkgo/core/matrix/identity.hpp): x = b.
template <typename ValueType = double>
class Identity : public LinOp {
public:
    static std::unique_ptr<Identity> create(std::shared_ptr<const Executor> exec, size_type n)
    {
        return std::unique_ptr<Identity>(new Identity(std::move(exec), n));
    }

protected:
    Identity(std::shared_ptr<const Executor> exec, size_type n) : LinOp(exec, dim<2>{n, n}) {}
    void apply_impl(const LinOp* b, LinOp* x) const override;
    void apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                    LinOp* x) const override;
};

}  // namespace matrix

// gko::initialize<Dense<V>>({...}, exec) for column vectors and scalars
template <typename DenseType>
std::unique_ptr<DenseType> initialize(std::initializer_list<typename DenseType::value_type> vals,
                                      std::shared_ptr<const Executor> exec)
{
    std::vector<typename DenseType::value_type> v(vals);
    return DenseType::create_from_host(exec, dim<2>{v.size(), 1}, v);
}

template <typename DenseType>
std::unique_ptr<DenseType> initialize(
    std::initializer_list<std::initializer_list<typename DenseType::value_type>> rows,
    std::shared_ptr<const Executor> exec)
{
    std::vector<typename DenseType::value_type> v;
    size_type cols = rows.size() ? rows.begin()->size() : 0;
    for (auto& r : rows) {
        if (r.size() != cols) throw DimensionMismatch(__FILE__, __LINE__, "ragged initializer");
        v.insert(v.end(), r.begin(), r.end());
    }
    return DenseType::create_from_host(exec, dim<2>{rows.size(), cols}, v);
}

// Casts a LinOp to Dense<V>, or throws NotSupported (gko::as).
template <typename ValueType>
const matrix::Dense<ValueType>* as_dense(const LinOp* op)
{
    auto d = dynamic_cast<const matrix::Dense<ValueType>*>(op);
    if (!d) GKO_NOT_SUPPORTED_MSG("operand is not a Dense of the operator's value type");
    return d;
}
template <typename ValueType>
matrix::Dense<ValueType>* as_dense(LinOp* op)
{
    auto d = dynamic_cast<matrix::Dense<ValueType>*>(op);
    if (!d) GKO_NOT_SUPPORTED_MSG("operand is not a Dense of the operator's value type");
    return d;
}

}  // namespace gko
