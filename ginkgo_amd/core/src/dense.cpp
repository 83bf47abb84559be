// matrix::Dense, matrix::Identity and precision dispatch over the C ABI.
#include "ginkgo_amd/dispatch.hpp"

namespace gko {
namespace matrix {

namespace {

template <typename T>
struct ops;

template <>
struct ops<double> {
    static constexpr auto fill = gkomi_fill_f64;
    static constexpr auto scale = gkomi_dense_scale_f64;
    static constexpr auto add_scaled = gkomi_dense_add_scaled_f64;
    static constexpr auto dot = gkomi_dense_dot_f64;
    static constexpr auto norm2 = gkomi_dense_norm2_f64;
    static constexpr auto copy = gkomi_dense_copy_f64;
};

template <>
struct ops<float> {
    static constexpr auto fill = gkomi_fill_f32;
    static constexpr auto scale = gkomi_dense_scale_f32;
    static constexpr auto add_scaled = gkomi_dense_add_scaled_f32;
    static constexpr auto dot = gkomi_dense_dot_f32;
    static constexpr auto norm2 = gkomi_dense_norm2_f32;
    static constexpr auto copy = gkomi_dense_copy_f32;
};

}  // namespace

template <typename V>
char* Dense<V>::reduction_workspace(size_type bytes) const
{
    // partials of the reduction kernels; kept per vector to avoid reallocation
    if (workspace_.get_size() < bytes) {
        workspace_ = array<char>(this->get_executor(), bytes);
    }
    return workspace_.get_data();
}

template <typename V>
std::vector<V> Dense<V>::to_host() const
{
    auto raw = values_.to_host();
    const auto s = this->get_size();
    std::vector<V> out(s.rows * s.cols);
    for (size_type i = 0; i < s.rows; ++i) {
        for (size_type j = 0; j < s.cols; ++j) out[i * s.cols + j] = raw[i * stride_ + j];
    }
    return out;
}

template <typename V>
std::unique_ptr<Dense<V>> Dense<V>::clone() const
{
    auto d = create(this->get_executor(), this->get_size(), this->get_size().cols);
    d->copy_from(this);
    return d;
}

template <typename V>
void Dense<V>::fill(V value)
{
    const auto s = this->get_size();
    if (s.rows * s.cols == 0) return;
    if (stride_ == s.cols) {
        GKO_CALL(ops<V>::fill(ctx_of(this->get_executor()), static_cast<int64_t>(s.rows * s.cols),
                              value, values_.get_data()));
    } else {
        for (size_type i = 0; i < s.rows; ++i) {
            GKO_CALL(ops<V>::fill(ctx_of(this->get_executor()), static_cast<int64_t>(s.cols),
                                  value, values_.get_data() + i * stride_));
        }
    }
}

template <typename V>
void Dense<V>::scale(const Dense* alpha)
{
    const auto s = this->get_size();
    GKO_CALL(ops<V>::scale(ctx_of(this->get_executor()), s.rows, s.cols,
                           alpha->get_const_values(), alpha->get_size().cols,
                           values_.get_data(), stride_));
}

template <typename V>
void Dense<V>::add_scaled(const Dense* alpha, const Dense* b)
{
    const auto s = this->get_size();
    if (b->get_size() != s) GKO_DIMENSION_MISMATCH_MSG("add_scaled: sizes differ");
    GKO_CALL(ops<V>::add_scaled(ctx_of(this->get_executor()), s.rows, s.cols,
                                alpha->get_const_values(), alpha->get_size().cols,
                                b->get_const_values(), b->get_stride(), values_.get_data(),
                                stride_));
}

template <typename V>
void Dense<V>::compute_dot(const Dense* b, Dense* result) const
{
    const auto s = this->get_size();
    if (b->get_size() != s || result->get_size() != dim<2>{1, s.cols}) {
        GKO_DIMENSION_MISMATCH_MSG("compute_dot: sizes do not conform");
    }
    const size_type wsb = gkomi_reduction_workspace_bytes(s.rows, s.cols);
    GKO_CALL(ops<V>::dot(ctx_of(this->get_executor()), s.rows, s.cols, get_const_values(),
                         stride_, b->get_const_values(), b->get_stride(), result->get_values(),
                         reduction_workspace(wsb), wsb));
}

template <typename V>
void Dense<V>::compute_norm2(Dense* result) const
{
    const auto s = this->get_size();
    if (result->get_size() != dim<2>{1, s.cols}) {
        GKO_DIMENSION_MISMATCH_MSG("compute_norm2: result must be 1 x cols");
    }
    const size_type wsb = gkomi_reduction_workspace_bytes(s.rows, s.cols);
    GKO_CALL(ops<V>::norm2(ctx_of(this->get_executor()), s.rows, s.cols, get_const_values(),
                           stride_, result->get_values(), reduction_workspace(wsb), wsb));
}

template <typename V>
void Dense<V>::copy_from(const Dense* other)
{
    const auto s = other->get_size();
    if (this->get_size() != s) GKO_DIMENSION_MISMATCH_MSG("copy_from: sizes differ");
    if (other->get_executor()->is_host() || this->get_executor()->is_host() ||
        other->get_executor() != this->get_executor()) {
        if (stride_ != s.cols || other->stride_ != s.cols) {
            GKO_NOT_SUPPORTED_MSG("cross-executor copies need contiguous Dense operands");
        }
        this->get_executor()->copy_from(other->get_executor().get(), s.rows * s.cols,
                                        other->get_const_values(), values_.get_data());
        return;
    }
    GKO_CALL(ops<V>::copy(ctx_of(this->get_executor()), s.rows, s.cols, other->get_const_values(),
                          other->get_stride(), values_.get_data(), stride_));
}

template <typename V>
template <typename Other>
void Dense<V>::copy_from(const Dense<Other>* other)
{
    const auto s = other->get_size();
    if (this->get_size() != s) GKO_DIMENSION_MISMATCH_MSG("copy_from: sizes differ");
    if (stride_ != s.cols || other->get_stride() != s.cols) {
        GKO_NOT_SUPPORTED_MSG("precision conversion needs contiguous Dense operands");
    }
    const int64_t n = static_cast<int64_t>(s.rows * s.cols);
    if constexpr (std::is_same<V, float>::value && std::is_same<Other, double>::value) {
        GKO_CALL(gkomi_convert_f64_to_f32(ctx_of(this->get_executor()), n,
                                          other->get_const_values(), values_.get_data()));
    } else if constexpr (std::is_same<V, double>::value && std::is_same<Other, float>::value) {
        GKO_CALL(gkomi_convert_f32_to_f64(ctx_of(this->get_executor()), n,
                                          other->get_const_values(), values_.get_data()));
    } else {
        copy_from(static_cast<const Dense*>(other));
    }
}

template <typename V>
void Dense<V>::apply_impl(const LinOp*, LinOp*) const
{
    throw NotImplemented(__FILE__, __LINE__,
                         "Dense::apply (a dense product) is not part of the sparse path; "
                         "convert the operator to Csr");
}

template <typename V>
void Dense<V>::apply_impl(const LinOp*, const LinOp*, const LinOp*, LinOp*) const
{
    throw NotImplemented(__FILE__, __LINE__,
                         "Dense::apply (a dense product) is not part of the sparse path; "
                         "convert the operator to Csr");
}

template class Dense<double>;
template class Dense<float>;
template void Dense<double>::copy_from<float>(const Dense<float>*);
template void Dense<float>::copy_from<double>(const Dense<double>*);

// Identity::apply_impl (core/matrix/identity.cpp): x = b, or x = alpha b + beta x
template <typename V>
void Identity<V>::apply_impl(const LinOp* b, LinOp* x) const
{
    precision_dispatch<V>([](auto db, auto dx) { dx->copy_from(db); }, b, x);
}

template <typename V>
void Identity<V>::apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                             LinOp* x) const
{
    precision_dispatch<V>(
        [](auto da, auto db, auto dbeta, auto dx) {
            dx->scale(dbeta);
            dx->add_scaled(da, db);
        },
        alpha, b, beta, x);
}

template class Identity<double>;
template class Identity<float>;

}  // namespace matrix

}  // namespace gko

