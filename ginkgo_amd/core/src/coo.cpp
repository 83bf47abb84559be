// matrix::Coo and matrix::Hybrid host classes (core/matrix/coo.cpp,
// core/matrix/hybrid.cpp): storage, conversions from and to Csr, and the
// products, dispatched to the C ABI.
#include "ginkgo_amd/dispatch.hpp"
#include "ginkgo_amd/matrix.hpp"

namespace gko {
namespace matrix {

namespace {

template <typename T>
struct fk;

template <>
struct fk<double> {
    static constexpr auto spmv2 = gkomi_coo_spmv2_f64;
    static constexpr auto to_hybrid = gkomi_csr_to_hybrid_f64;
    static constexpr auto ell_spmv = gkomi_ell_spmv_f64;
    static constexpr auto ell_adv = gkomi_ell_advanced_spmv_f64;
};

template <>
struct fk<float> {
    static constexpr auto spmv2 = gkomi_coo_spmv2_f32;
    static constexpr auto to_hybrid = gkomi_csr_to_hybrid_f32;
    static constexpr auto ell_spmv = gkomi_ell_spmv_f32;
    static constexpr auto ell_adv = gkomi_ell_advanced_spmv_f32;
};

}  // namespace

// ---------------------------------------------------------------- Coo ----

template <typename V, typename I>
std::unique_ptr<Coo<V, I>> Coo<V, I>::create(std::shared_ptr<const Executor> exec, dim<2> size,
                                             size_type nnz)
{
    auto m = std::unique_ptr<Coo>(new Coo(exec, size, nnz));
    if (nnz == 0) m->build_index(false);
    return m;
}

template <typename V, typename I>
std::unique_ptr<Coo<V, I>> Coo<V, I>::create_from_host(std::shared_ptr<const Executor> exec,
                                                       dim<2> size,
                                                       const std::vector<I>& row_idxs,
                                                       const std::vector<I>& col_idxs,
                                                       const std::vector<V>& values)
{
    const auto nnz = values.size();
    if (row_idxs.size() != nnz || col_idxs.size() != nnz) {
        throw DimensionMismatch(__FILE__, __LINE__, "Coo: array sizes do not conform");
    }
    for (size_type k = 0; k < nnz; ++k) {
        if (row_idxs[k] < 0 || static_cast<size_type>(row_idxs[k]) >= size.rows ||
            col_idxs[k] < 0 || static_cast<size_type>(col_idxs[k]) >= size.cols) {
            throw DimensionMismatch(__FILE__, __LINE__, "Coo: index out of range");
        }
        if (k > 0 && row_idxs[k] < row_idxs[k - 1]) {
            GKO_NOT_SUPPORTED_MSG("Coo: the entries must be sorted by row");
        }
    }
    auto m = std::unique_ptr<Coo>(new Coo(exec, size, 0));
    m->values_ = array<V>(exec, values);
    m->col_idxs_ = array<I>(exec, col_idxs);
    m->row_idxs_ = array<I>(exec, row_idxs);
    m->build_index(false);
    return m;
}

// The row-pointer index of the entries and the Csr view over it; the view's
// automatical strategy picks the sub-wave or the load-balanced kernels from
// the row lengths, as for any Csr.
template <typename V, typename I>
void Coo<V, I>::build_index(bool ptrs_ready)
{
    auto exec = this->get_executor();
    const auto n = this->get_size().rows;
    const auto nnz = values_.get_size();
    if (row_ptrs_.get_size() != n + 1) row_ptrs_ = array<I>(exec, n + 1);
    if (!ptrs_ready) {
        GKO_CALL(gkomi_idxs_to_ptrs(ctx_of(exec), static_cast<int64>(nnz),
                                    row_idxs_.get_const_data(), static_cast<int64>(n),
                                    row_ptrs_.get_data()));
    }
    index_ = Csr<V, I>::create_view(exec, this->get_size(), nnz, values_.get_data(),
                                    col_idxs_.get_data(), row_ptrs_.get_data());
}

template <typename V, typename I>
void Coo<V, I>::apply_impl(const LinOp* b, LinOp* x) const
{
    index_->apply(b, x);
}

template <typename V, typename I>
void Coo<V, I>::apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                           LinOp* x) const
{
    index_->apply(alpha, b, beta, x);
}

template <typename V, typename I>
void Coo<V, I>::apply2(const LinOp* b, LinOp* x) const
{
    this->check_apply(b, x);
    precision_dispatch<V>(
        [this](auto db, auto dx) {
            const auto d = index_->descriptor();
            GKO_CALL(fk<V>::spmv2(ctx_of(this->get_executor()), nullptr, &d,
                                  db->get_size().cols, db->get_const_values(),
                                  db->get_stride(), dx->get_values(), dx->get_stride()));
        },
        b, x);
}

template <typename V, typename I>
void Coo<V, I>::apply2(const LinOp* alpha, const LinOp* b, LinOp* x) const
{
    this->check_apply(b, x);
    if (!alpha || alpha->get_size().rows != 1 || alpha->get_size().cols != 1) {
        throw DimensionMismatch(__FILE__, __LINE__, "apply2: alpha must be 1x1");
    }
    precision_dispatch<V>(
        [this](auto da, auto db, auto dx) {
            const auto d = index_->descriptor();
            GKO_CALL(fk<V>::spmv2(ctx_of(this->get_executor()), da->get_const_values(), &d,
                                  db->get_size().cols, db->get_const_values(),
                                  db->get_stride(), dx->get_values(), dx->get_stride()));
        },
        alpha, b, x);
}

template <typename V, typename I>
std::unique_ptr<Csr<V, I>> Coo<V, I>::to_csr() const
{
    auto exec = this->get_executor();
    return Csr<V, I>::create(exec, this->get_size(), array<V>(exec, values_),
                             array<I>(exec, col_idxs_), array<I>(exec, row_ptrs_));
}

template <typename V, typename I>
std::unique_ptr<Coo<V, I>> Csr<V, I>::to_coo() const
{
    auto exec = this->get_executor();
    const auto nnz = values_.get_size();
    const auto n = this->get_size().rows;
    auto m = std::unique_ptr<Coo<V, I>>(new Coo<V, I>(exec, this->get_size(), 0));
    m->values_ = array<V>(exec, values_);
    m->col_idxs_ = array<I>(exec, col_idxs_);
    m->row_idxs_ = array<I>(exec, nnz);
    m->row_ptrs_ = array<I>(exec, row_ptrs_);
    GKO_CALL(gkomi_ptrs_to_idxs(ctx_of(exec), static_cast<int64>(n), row_ptrs_.get_const_data(),
                                static_cast<int64>(nnz), m->row_idxs_.get_data()));
    m->build_index(true);
    return m;
}

// ------------------------------------------------------------- Hybrid ----

template <typename V, typename I>
std::unique_ptr<Hybrid<V, I>> Hybrid<V, I>::create_from_csr(const Csr<V, I>* csr,
                                                            std::shared_ptr<strategy_type> s)
{
    if (!csr) GKO_NOT_SUPPORTED_MSG("Hybrid: null source");
    auto exec = csr->get_executor();
    const auto size = csr->get_size();
    const auto n = size.rows;
    auto m = std::unique_ptr<Hybrid>(new Hybrid(exec, size));
    m->strategy_ = s ? std::move(s) : std::make_shared<automatic>();
    // the strategy reads the row lengths on the host, as the reference's
    // compute_hybrid_config does
    const auto rp = csr->row_ptrs_to_host();
    std::vector<size_type> row_nnz(n);
    for (size_type r = 0; r < n; ++r) row_nnz[r] = static_cast<size_type>(rp[r + 1] - rp[r]);
    auto lens = row_nnz;
    const size_type width = m->strategy_->compute_ell_num_stored_elements_per_row(lens);
    size_type coo_nnz = 0;
    for (auto l : row_nnz) coo_nnz += l > width ? l - width : 0;
    m->ell_ = std::unique_ptr<Ell<V, I>>(new Ell<V, I>(exec, size, width, n));
    m->coo_ = std::unique_ptr<Coo<V, I>>(new Coo<V, I>(exec, size, coo_nnz));
    const auto d = csr->descriptor();
    const gkomi_ctx* ctx = ctx_of(exec);
    array<char> ws(exec, gkomi_csr_hybrid_workspace_bytes(static_cast<int64>(n)));
    GKO_CALL(gkomi_csr_hybrid_coo_row_ptrs(ctx, &d, static_cast<int64>(width),
                                           m->coo_->row_ptrs_.get_data(), ws.get_data(),
                                           ws.get_size()));
    GKO_CALL(fk<V>::to_hybrid(ctx, &d, static_cast<int64>(width), static_cast<int64>(n),
                              m->ell_->col_idxs_.get_data(), m->ell_->values_.get_data(),
                              m->coo_->row_ptrs_.get_const_data(),
                              m->coo_->row_idxs_.get_data(), m->coo_->col_idxs_.get_data(),
                              m->coo_->values_.get_data()));
    m->coo_->build_index(true);
    return m;
}

template <typename V, typename I>
void Hybrid<V, I>::apply_impl(const LinOp* b, LinOp* x) const
{
    ell_->apply(b, x);
    if (coo_->get_num_stored_elements()) coo_->apply2(b, x);
}

template <typename V, typename I>
void Hybrid<V, I>::apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                              LinOp* x) const
{
    ell_->apply(alpha, b, beta, x);
    if (coo_->get_num_stored_elements()) coo_->apply2(alpha, b, x);
}

template class Coo<double, int32>;
template class Coo<float, int32>;
template class Hybrid<double, int32>;
template class Hybrid<float, int32>;
template std::unique_ptr<Coo<double, int32>> Csr<double, int32>::to_coo() const;
template std::unique_ptr<Coo<float, int32>> Csr<float, int32>::to_coo() const;

}  // namespace matrix
}  // namespace gko
