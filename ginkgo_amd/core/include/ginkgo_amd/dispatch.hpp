// precision_dispatch (include/ginkgo/core/base/precision_dispatch.hpp:43-125):
// runs a function on Dense<V> views of its operands, converting Dense
// operands of the other floating-point precision through temporaries
// (outputs are converted back afterwards).
#pragma once

#include "ginkgo_amd/lin_op.hpp"

namespace gko {
namespace detail {

template <typename V, typename Other>
std::unique_ptr<matrix::Dense<V>> convert_dense(const matrix::Dense<Other>* d)
{
    auto t = matrix::Dense<V>::create(d->get_executor(), d->get_size());
    if (d->get_stride() == d->get_size().cols) {
        t->copy_from(d);
    } else {
        auto c = d->clone();
        t->copy_from(c.get());
    }
    return t;
}

template <typename V>
using other_t = std::conditional_t<std::is_same<V, double>::value, float, double>;

// a Dense<V> view of op, possibly a converted temporary
template <typename V>
struct dense_in {
    const matrix::Dense<V>* ptr{};
    std::unique_ptr<matrix::Dense<V>> tmp;
    explicit dense_in(const LinOp* op)
    {
        if ((ptr = dynamic_cast<const matrix::Dense<V>*>(op))) return;
        if (auto o = dynamic_cast<const matrix::Dense<other_t<V>>*>(op)) {
            tmp = convert_dense<V>(o);
            ptr = tmp.get();
            return;
        }
        GKO_NOT_SUPPORTED_MSG("operand is not a Dense matrix");
    }
};

template <typename V>
struct dense_out {
    matrix::Dense<V>* ptr{};
    matrix::Dense<other_t<V>>* orig{};
    std::unique_ptr<matrix::Dense<V>> tmp;
    explicit dense_out(LinOp* op)
    {
        if ((ptr = dynamic_cast<matrix::Dense<V>*>(op))) return;
        if ((orig = dynamic_cast<matrix::Dense<other_t<V>>*>(op))) {
            tmp = convert_dense<V>(orig);
            ptr = tmp.get();
            return;
        }
        GKO_NOT_SUPPORTED_MSG("operand is not a Dense matrix");
    }
    void write_back()
    {
        if (orig) orig->copy_from(ptr);
    }
};

}  // namespace detail

template <typename V, typename Fn>
inline void precision_dispatch(Fn fn, const LinOp* b, LinOp* x)
{
    detail::dense_in<V> db(b);
    detail::dense_out<V> dx(x);
    fn(db.ptr, dx.ptr);
    dx.write_back();
}

// alpha, b -> x (the apply2 of Coo: x += alpha A b)
template <typename V, typename Fn>
inline void precision_dispatch(Fn fn, const LinOp* alpha, const LinOp* b, LinOp* x)
{
    detail::dense_in<V> da(alpha);
    detail::dense_in<V> db(b);
    detail::dense_out<V> dx(x);
    fn(da.ptr, db.ptr, dx.ptr);
    dx.write_back();
}

template <typename V, typename Fn>
inline void precision_dispatch(Fn fn, const LinOp* alpha, const LinOp* b, const LinOp* beta, LinOp* x)
{
    detail::dense_in<V> da(alpha);
    detail::dense_in<V> db(b);
    detail::dense_in<V> dbeta(beta);
    detail::dense_out<V> dx(x);
    fn(da.ptr, db.ptr, dbeta.ptr, dx.ptr);
    dx.write_back();
}

// mixed_precision_dispatch (include/ginkgo/core/base/precision_dispatch.hpp:
// 228-262, the GINKGO_MIXED_PRECISION form): runs fn on the operands as
// Dense of their own value types, each of V or the other precision, with
// nothing converted; the kernel computes in the highest precision.
template <typename V, typename Fn>
inline void mixed_precision_dispatch(Fn fn, const LinOp* b, LinOp* x)
{
    using O = detail::other_t<V>;
    auto with_out = [&](auto db) {
        if (auto dx = dynamic_cast<matrix::Dense<V>*>(x)) {
            fn(db, dx);
        } else if (auto dx = dynamic_cast<matrix::Dense<O>*>(x)) {
            fn(db, dx);
        } else {
            GKO_NOT_SUPPORTED_MSG("output operand is not a Dense matrix");
        }
    };
    if (auto db = dynamic_cast<const matrix::Dense<V>*>(b)) {
        with_out(db);
    } else if (auto db = dynamic_cast<const matrix::Dense<O>*>(b)) {
        with_out(db);
    } else {
        GKO_NOT_SUPPORTED_MSG("input operand is not a Dense matrix");
    }
}

// The C ABI's value-type code of T.
template <typename T>
constexpr int32_t value_code()
{
    return std::is_same<T, double>::value ? GKOMI_VALUE_F64 : GKOMI_VALUE_F32;
}

}  // namespace gko
