// Stopping criteria (include/ginkgo/core/stop/{criterion,iteration,residual_norm,
// combined}.hpp).  Criterion factories keep the reference's parameters and
// builder API; their evaluation happens on the device (gkomi_criteria),
// so a solver never has to copy a residual norm to the host to decide.
#pragma once

#include "ginkgo_amd/lin_op.hpp"

namespace gko {
namespace stop {

// stop::mode (residual_norm.hpp:37)
enum class mode { absolute, initial_resnorm, rhs_norm };

class CriterionFactory {
public:
    virtual ~CriterionFactory() = default;
    // appends this criterion (or, for Combined, its members) to `out`
    virtual void append_to(gkomi_criteria& out) const = 0;
    std::shared_ptr<const Executor> get_executor() const { return exec_; }

protected:
    explicit CriterionFactory(std::shared_ptr<const Executor> exec) : exec_{std::move(exec)} {}

private:
    std::shared_ptr<const Executor> exec_;
};

inline void push_criterion(gkomi_criteria& out, const gkomi_criterion& c)
{
    if (out.num >= GKOMI_MAX_CRITERIA) {
        GKO_NOT_SUPPORTED_MSG("at most 4 stopping criteria can be combined");
    }
    out.c[out.num++] = c;
}

// stop::Iteration (core/stop/iteration.cpp:14-24)
class Iteration {
public:
    class Factory : public CriterionFactory {
    public:
        Factory(std::shared_ptr<const Executor> exec, size_type max_iters)
            : CriterionFactory(std::move(exec)), max_iters_{max_iters}
        {}
        void append_to(gkomi_criteria& out) const override
        {
            gkomi_criterion c{};
            c.kind = GKOMI_CRIT_ITERATION;
            c.max_iters = static_cast<int64>(max_iters_);
            push_criterion(out, c);
        }
        size_type get_max_iters() const { return max_iters_; }

    private:
        size_type max_iters_;
    };
    struct parameters_type {
        size_type max_iters{0};
        parameters_type& with_max_iters(size_type n)
        {
            max_iters = n;
            return *this;
        }
        std::unique_ptr<Factory> on(std::shared_ptr<const Executor> exec) const
        {
            return std::make_unique<Factory>(std::move(exec), max_iters);
        }
    };
    static parameters_type build() { return {}; }
};

namespace detail {
template <typename ValueType, int Kind>
class ResidualNormCriterion {
public:
    class Factory : public CriterionFactory {
    public:
        Factory(std::shared_ptr<const Executor> exec, double reduction, mode baseline)
            : CriterionFactory(std::move(exec)), reduction_{reduction}, baseline_{baseline}
        {}
        void append_to(gkomi_criteria& out) const override
        {
            gkomi_criterion c{};
            c.kind = Kind;
            c.baseline = baseline_ == mode::rhs_norm
                             ? GKOMI_BASELINE_RHS_NORM
                             : (baseline_ == mode::initial_resnorm ? GKOMI_BASELINE_INITIAL_RESNORM
                                                                   : GKOMI_BASELINE_ABSOLUTE);
            c.reduction_factor = reduction_;
            push_criterion(out, c);
        }
        double get_reduction_factor() const { return reduction_; }
        mode get_baseline() const { return baseline_; }

    private:
        double reduction_;
        mode baseline_;
    };
    struct parameters_type {
        // default 5 * eps of the value type (residual_norm.hpp:124-128)
        double reduction_factor{5 * std::numeric_limits<ValueType>::epsilon()};
        mode baseline{mode::rhs_norm};
        parameters_type& with_reduction_factor(double r)
        {
            reduction_factor = r;
            return *this;
        }
        parameters_type& with_baseline(mode m)
        {
            baseline = m;
            return *this;
        }
        std::unique_ptr<Factory> on(std::shared_ptr<const Executor> exec) const
        {
            return std::make_unique<Factory>(std::move(exec), reduction_factor, baseline);
        }
    };
    static parameters_type build() { return {}; }
};
}  // namespace detail

// stop::ResidualNorm<V> (core/stop/residual_norm.cpp:165-207)
template <typename ValueType = double>
using ResidualNorm = detail::ResidualNormCriterion<ValueType, GKOMI_CRIT_RESIDUAL_NORM>;
// stop::ImplicitResidualNorm<V> (core/stop/residual_norm.cpp:210-226)
template <typename ValueType = double>
using ImplicitResidualNorm = detail::ResidualNormCriterion<ValueType, GKOMI_CRIT_IMPLICIT_RESNORM>;

// stop::Combined (core/stop/combined.cpp)
class Combined {
public:
    class Factory : public CriterionFactory {
    public:
        Factory(std::shared_ptr<const Executor> exec,
                std::vector<std::shared_ptr<const CriterionFactory>> list)
            : CriterionFactory(std::move(exec)), list_{std::move(list)}
        {
            if (list_.empty()) GKO_NOT_SUPPORTED_MSG("Combined needs at least one criterion");
        }
        void append_to(gkomi_criteria& out) const override
        {
            for (auto& c : list_) {
                if (c) c->append_to(out);
            }
        }

    private:
        std::vector<std::shared_ptr<const CriterionFactory>> list_;
    };
};

// gko::stop::max_iters / gko::stop::combine helpers
inline Iteration::parameters_type max_iters(size_type n)
{
    return Iteration::build().with_max_iters(n);
}

}  // namespace stop

// A list of deferred criteria, as solver parameters hold them.
struct criteria_list {
    std::vector<deferred_factory_parameter<const stop::CriterionFactory>> list;
    gkomi_criteria resolve(std::shared_ptr<const Executor> exec) const
    {
        gkomi_criteria out{};
        for (auto& d : list) {
            auto f = d.on(exec);
            if (f) f->append_to(out);
        }
        if (out.num == 0) {
            GKO_NOT_SUPPORTED_MSG("a solver needs at least one stopping criterion");
        }
        return out;
    }
};

}  // namespace gko
