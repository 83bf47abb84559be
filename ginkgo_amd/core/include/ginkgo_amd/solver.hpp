// Preconditioner and solvers:
//   preconditioner::Jacobi (core/preconditioner/jacobi.cpp:338-370, scalar case),
//   solver::Cg (core/solver/cg.cpp), solver::Bicgstab (core/solver/bicgstab.cpp),
//   solver::Ir (core/solver/ir.cpp).
// The builder API follows abstract_factory.hpp: Cg<V>::build()
// .with_criteria(...).with_preconditioner(...).on(exec)->generate(A)->apply(b, x).
#pragma once

#include "ginkgo_amd/matrix.hpp"
#include "ginkgo_amd/stop.hpp"

namespace gko {

namespace preconditioner {

// Scalar Jacobi: inv_diag = 1 / diag(A) with 0 -> 1.  max_block_size must be
// 1 (the block-Jacobi variants of the reference are outside this path).
template <typename ValueType = double, typename IndexType = int32>
class Jacobi : public LinOp {
public:
    using value_type = ValueType;
    class Factory;
    struct parameters_type {
        uint32_t max_block_size{32};  // reference default (jacobi.hpp:315)
        bool skip_sorting{false};
        parameters_type& with_max_block_size(uint32_t b)
        {
            max_block_size = b;
            return *this;
        }
        parameters_type& with_skip_sorting(bool s)
        {
            skip_sorting = s;
            return *this;
        }
        std::unique_ptr<Factory> on(std::shared_ptr<const Executor> exec) const
        {
            return std::unique_ptr<Factory>(new Factory(std::move(exec), *this));
        }
    };
    class Factory : public LinOpFactory {
    public:
        Factory(std::shared_ptr<const Executor> exec, parameters_type p)
            : LinOpFactory(std::move(exec)), params_{p}
        {}
        const parameters_type& get_parameters() const { return params_; }

    protected:
        std::unique_ptr<LinOp> generate_impl(std::shared_ptr<const LinOp> a) const override
        {
            return std::unique_ptr<LinOp>(new Jacobi(get_executor(), params_, a.get()));
        }

    private:
        parameters_type params_;
    };
    static parameters_type build() { return {}; }
    const ValueType* get_inverse_diagonal() const { return inv_diag_.get_const_data(); }
    std::vector<ValueType> inverse_diagonal_to_host() const { return inv_diag_.to_host(); }

protected:
    Jacobi(std::shared_ptr<const Executor> exec, parameters_type p, const LinOp* a);
    void apply_impl(const LinOp* b, LinOp* x) const override;
    void apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                    LinOp* x) const override;

private:
    array<ValueType> inv_diag_;
};

}  // namespace preconditioner

namespace solver {

// Shared parameters of the Krylov solvers (solver_base.hpp:818-870).
struct iterative_parameters {
    criteria_list criteria;
    deferred_factory_parameter<const LinOpFactory> preconditioner;
    std::shared_ptr<const LinOp> generated_preconditioner;
};

template <typename Derived, typename Params>
struct krylov_builder : Params {
    template <typename... C>
    Derived& with_criteria(C&&... c)
    {
        this->criteria.list = {deferred_factory_parameter<const stop::CriterionFactory>(
            std::forward<C>(c))...};
        return static_cast<Derived&>(*this);
    }
    Derived& with_criteria_list(
        std::vector<deferred_factory_parameter<const stop::CriterionFactory>> l)
    {
        this->criteria.list = std::move(l);
        return static_cast<Derived&>(*this);
    }
    Derived& with_preconditioner(deferred_factory_parameter<const LinOpFactory> p)
    {
        this->preconditioner = std::move(p);
        return static_cast<Derived&>(*this);
    }
    Derived& with_generated_preconditioner(std::shared_ptr<const LinOp> p)
    {
        this->generated_preconditioner = std::move(p);
        return static_cast<Derived&>(*this);
    }
};

// Device-side statistics of the last apply.
struct solve_stats {
    int64 iterations{};
    std::vector<stopping_status> status;
    bool used_fused_path{};
    // Ir: the inner solver's iterations in every sweep (Krylov inner solvers)
    std::vector<int64> inner_iterations;
};

template <typename ValueType>
class KrylovBase : public LinOp, public IterativeBase {
public:
    using value_type = ValueType;
    bool apply_uses_initial_guess() const override { return true; }
    std::shared_ptr<const LinOp> get_system_matrix() const { return system_matrix_; }
    std::shared_ptr<const LinOp> get_preconditioner() const { return precond_; }
    const solve_stats& get_last_stats() const { return stats_; }
    // iterations the host enqueues ahead of each convergence poll
    void set_poll_chunk(int chunk) { chunk_ = chunk > 0 ? chunk : 1; }

protected:
    KrylovBase(std::shared_ptr<const Executor> exec, const iterative_parameters& p,
               std::shared_ptr<const LinOp> a);
    using LinOp::apply_impl;
    void apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                    LinOp* x) const override;
    void log_result(const matrix::Dense<ValueType>* b, const matrix::Dense<ValueType>* x) const;

    std::shared_ptr<const LinOp> system_matrix_;
    std::shared_ptr<const LinOp> precond_;  // nullptr = identity
    const ValueType* inv_diag_{};           // set when precond_ is scalar Jacobi
    gkomi_criteria crit_{};
    mutable solve_stats stats_;
    int chunk_{8};
};

template <typename ValueType = double>
class Cg : public KrylovBase<ValueType> {
public:
    class Factory;
    struct parameters_type : krylov_builder<parameters_type, iterative_parameters> {
        std::unique_ptr<Factory> on(std::shared_ptr<const Executor> exec) const
        {
            return std::unique_ptr<Factory>(new Factory(std::move(exec), *this));
        }
    };
    class Factory : public LinOpFactory {
    public:
        Factory(std::shared_ptr<const Executor> exec, parameters_type p)
            : LinOpFactory(std::move(exec)), params_{std::move(p)}
        {}
        const parameters_type& get_parameters() const { return params_; }

    protected:
        std::unique_ptr<LinOp> generate_impl(std::shared_ptr<const LinOp> a) const override
        {
            return std::unique_ptr<LinOp>(new Cg(get_executor(), params_, std::move(a)));
        }

    private:
        parameters_type params_;
    };
    static parameters_type build() { return {}; }

protected:
    Cg(std::shared_ptr<const Executor> exec, const parameters_type& p,
       std::shared_ptr<const LinOp> a)
        : KrylovBase<ValueType>(std::move(exec), p, std::move(a))
    {}
    void apply_impl(const LinOp* b, LinOp* x) const override;
    using KrylovBase<ValueType>::apply_impl;

private:
    void apply_dense(const matrix::Dense<ValueType>* b, matrix::Dense<ValueType>* x) const;
    bool try_fused(const matrix::Dense<ValueType>* b, matrix::Dense<ValueType>* x) const;
    bool try_fused_block(const matrix::Dense<ValueType>* b, matrix::Dense<ValueType>* x) const;
};

template <typename ValueType = double>
class Bicgstab : public KrylovBase<ValueType> {
public:
    class Factory;
    struct parameters_type : krylov_builder<parameters_type, iterative_parameters> {
        std::unique_ptr<Factory> on(std::shared_ptr<const Executor> exec) const
        {
            return std::unique_ptr<Factory>(new Factory(std::move(exec), *this));
        }
    };
    class Factory : public LinOpFactory {
    public:
        Factory(std::shared_ptr<const Executor> exec, parameters_type p)
            : LinOpFactory(std::move(exec)), params_{std::move(p)}
        {}
        const parameters_type& get_parameters() const { return params_; }

    protected:
        std::unique_ptr<LinOp> generate_impl(std::shared_ptr<const LinOp> a) const override
        {
            return std::unique_ptr<LinOp>(new Bicgstab(get_executor(), params_, std::move(a)));
        }

    private:
        parameters_type params_;
    };
    static parameters_type build() { return {}; }

protected:
    Bicgstab(std::shared_ptr<const Executor> exec, const parameters_type& p,
             std::shared_ptr<const LinOp> a)
        : KrylovBase<ValueType>(std::move(exec), p, std::move(a))
    {}
    void apply_impl(const LinOp* b, LinOp* x) const override;
    using KrylovBase<ValueType>::apply_impl;

private:
    void apply_dense(const matrix::Dense<ValueType>* b, matrix::Dense<ValueType>* x) const;
    bool try_fused(const matrix::Dense<ValueType>* b, matrix::Dense<ValueType>* x) const;
};

// solver::Ir (core/solver/ir.cpp:193-257): x += w * inner_solve(b - A x).
// The inner solver may run in a lower precision; the system matrix is then
// converted once at generate time (the explicit conversion of the
// mixed-precision-ir example, examples/mixed-precision-ir:85-94).
template <typename ValueType = double>
class Ir : public LinOp, public IterativeBase {
public:
    using value_type = ValueType;
    class Factory;
    struct parameters_type {
        criteria_list criteria;
        deferred_factory_parameter<const LinOpFactory> solver;
        std::shared_ptr<const LinOp> generated_solver;
        double relaxation_factor{1.0};
        template <typename... C>
        parameters_type& with_criteria(C&&... c)
        {
            criteria.list = {deferred_factory_parameter<const stop::CriterionFactory>(
                std::forward<C>(c))...};
            return *this;
        }
        parameters_type& with_solver(deferred_factory_parameter<const LinOpFactory> s)
        {
            solver = std::move(s);
            return *this;
        }
        parameters_type& with_generated_solver(std::shared_ptr<const LinOp> s)
        {
            generated_solver = std::move(s);
            return *this;
        }
        parameters_type& with_relaxation_factor(double w)
        {
            relaxation_factor = w;
            return *this;
        }
        std::unique_ptr<Factory> on(std::shared_ptr<const Executor> exec) const
        {
            return std::unique_ptr<Factory>(new Factory(std::move(exec), *this));
        }
    };
    class Factory : public LinOpFactory {
    public:
        Factory(std::shared_ptr<const Executor> exec, parameters_type p)
            : LinOpFactory(std::move(exec)), params_{std::move(p)}
        {}
        const parameters_type& get_parameters() const { return params_; }

    protected:
        std::unique_ptr<LinOp> generate_impl(std::shared_ptr<const LinOp> a) const override
        {
            return std::unique_ptr<LinOp>(new Ir(get_executor(), params_, std::move(a)));
        }

    private:
        parameters_type params_;
    };
    static parameters_type build() { return {}; }
    bool apply_uses_initial_guess() const override { return true; }
    std::shared_ptr<const LinOp> get_solver() const { return inner_; }
    const solve_stats& get_last_stats() const { return stats_; }

protected:
    Ir(std::shared_ptr<const Executor> exec, const parameters_type& p,
       std::shared_ptr<const LinOp> a);
    void apply_impl(const LinOp* b, LinOp* x) const override;
    void apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                    LinOp* x) const override;

private:
    std::shared_ptr<const LinOp> system_matrix_;
    std::shared_ptr<const LinOp> inner_;  // nullptr = Richardson (Identity)
    gkomi_criteria crit_{};
    double relaxation_;
    mutable solve_stats stats_;
};

}  // namespace solver
}  // namespace gko
