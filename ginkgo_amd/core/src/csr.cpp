// matrix::Csr / Sellp / Ell host classes: storage, strategy resolution,
// conversions and SpMV dispatch to the C ABI.
#include <cmath>

#include "ginkgo_amd/dispatch.hpp"
#include "ginkgo_amd/matrix.hpp"

namespace gko {
namespace matrix {

namespace {

template <typename T>
struct sp;

template <>
struct sp<double> {
    static constexpr auto transpose = gkomi_csr_transpose_f64;
    static constexpr auto to_sellp = gkomi_csr_to_sellp_f64;
    static constexpr auto to_ell = gkomi_csr_to_ell_f64;
};

template <>
struct sp<float> {
    static constexpr auto transpose = gkomi_csr_transpose_f32;
    static constexpr auto to_sellp = gkomi_csr_to_sellp_f32;
    static constexpr auto to_ell = gkomi_csr_to_ell_f32;
};

}  // namespace

template <typename V, typename I>
std::unique_ptr<Csr<V, I>> Csr<V, I>::create(std::shared_ptr<const Executor> exec, dim<2> size,
                                             array<V> values, array<I> col_idxs,
                                             array<I> row_ptrs,
                                             std::shared_ptr<strategy_type> strategy)
{
    if (values.get_size() != col_idxs.get_size() || row_ptrs.get_size() != size.rows + 1) {
        throw DimensionMismatch(__FILE__, __LINE__, "Csr: array sizes do not conform");
    }
    auto m = std::unique_ptr<Csr>(new Csr(exec, size, 0, std::move(strategy)));
    values.set_executor(exec);
    col_idxs.set_executor(exec);
    row_ptrs.set_executor(exec);
    m->values_ = std::move(values);
    m->col_idxs_ = std::move(col_idxs);
    m->row_ptrs_ = std::move(row_ptrs);
    m->make_srow();
    return m;
}

template <typename V, typename I>
std::unique_ptr<Csr<V, I>> Csr<V, I>::create_from_host(std::shared_ptr<const Executor> exec,
                                                       dim<2> size,
                                                       const std::vector<I>& row_ptrs,
                                                       const std::vector<I>& col_idxs,
                                                       const std::vector<V>& values,
                                                       std::shared_ptr<strategy_type> strategy)
{
    return create(exec, size, array<V>(exec, values), array<I>(exec, col_idxs),
                  array<I>(exec, row_ptrs), std::move(strategy));
}

template <typename V, typename I>
std::unique_ptr<Csr<V, I>> Csr<V, I>::create_view(std::shared_ptr<const Executor> exec,
                                                  dim<2> size, size_type nnz, V* values,
                                                  I* col_idxs, I* row_ptrs,
                                                  std::shared_ptr<strategy_type> strategy)
{
    return create(exec, size, array<V>::view(exec, nnz, values),
                  array<I>::view(exec, nnz, col_idxs), array<I>::view(exec, size.rows + 1, row_ptrs),
                  std::move(strategy));
}

template <typename V, typename I>
std::unique_ptr<Csr<V, I>> Csr<V, I>::create_from_data(std::shared_ptr<const Executor> exec,
                                                       matrix_data<V, I> data,
                                                       std::shared_ptr<strategy_type> strategy)
{
    data.sort_row_major();
    std::vector<I> rp(data.size.rows + 1, 0), ci;
    std::vector<V> v;
    ci.reserve(data.nonzeros.size());
    v.reserve(data.nonzeros.size());
    for (size_type k = 0; k < data.nonzeros.size(); ++k) {
        const auto& nz = data.nonzeros[k];
        if (nz.row < 0 || static_cast<size_type>(nz.row) >= data.size.rows || nz.column < 0 ||
            static_cast<size_type>(nz.column) >= data.size.cols) {
            throw Error(__FILE__, __LINE__, "matrix_data entry out of range");
        }
        if (!ci.empty() && k > 0 && data.nonzeros[k - 1].row == nz.row &&
            data.nonzeros[k - 1].column == nz.column) {
            v.back() += nz.value;  // duplicates are summed
            continue;
        }
        ci.push_back(nz.column);
        v.push_back(nz.value);
        rp[nz.row + 1]++;
    }
    for (size_type r = 0; r < data.size.rows; ++r) rp[r + 1] += rp[r];
    return create_from_host(exec, data.size, rp, ci, v, std::move(strategy));
}

template <typename V, typename I>
std::unique_ptr<Csr<V, I>> Csr<V, I>::create_from_dense(
    std::shared_ptr<const Executor> exec, const std::vector<std::vector<V>>& rows,
    std::shared_ptr<strategy_type> strategy)
{
    matrix_data<V, I> d;
    d.size = dim<2>{rows.size(), rows.empty() ? 0 : rows[0].size()};
    for (size_type i = 0; i < rows.size(); ++i) {
        for (size_type j = 0; j < rows[i].size(); ++j) {
            if (rows[i][j] != V{0}) {
                d.nonzeros.push_back({static_cast<I>(i), static_cast<I>(j), rows[i][j]});
            }
        }
    }
    return create_from_data(exec, d, std::move(strategy));
}

template <typename V, typename I>
void Csr<V, I>::set_strategy(std::shared_ptr<strategy_type> s)
{
    strategy_ = std::move(s);
    make_srow();
}

// Resolves the strategy and builds the load-balance partition on the
// device (the role of make_srow, csr.hpp:1651-1655).  automatical (and
// sparselib) decide from the longest and the mean row length: classical
// (sub-wave per row) when no row is longer than the reference's AMD limit
// of 768 entries (automatical::process, csr.hpp:524-535) and the longest row
// is within 8x the mean plus 64 (the lanes of a sub-wave then stay busy),
// load_balance otherwise.
// merge_path keeps a workspace for the workgroups' carries.
template <typename V, typename I>
void Csr<V, I>::make_srow()
{
    if (this->get_executor()->is_host()) {
        effective_ = "classical";
        return;
    }
    std::string name = strategy_->get_name();
    subwave_ = 0;
    if (auto c = std::dynamic_pointer_cast<classical>(strategy_)) subwave_ = c->subwave();
    max_row_nnz_ = -1;
    if (name == "automatical" || name == "sparselib") {
        const double mean = this->get_size().rows
                                ? double(values_.get_size()) / double(this->get_size().rows)
                                : 0.0;
        const auto mx = static_cast<double>(max_row_nnz());
        name = (mx <= 768 && mx <= 8 * mean + 64) ? "classical" : "load_balance";
    }
    effective_ = name;
    if (name == "load_balance") {
        // an explicit load_balance keeps its warps; automatical resolving to
        // it uses the executor's, as the reference's automatical does for a
        // HipExecutor (csr.hpp:606-673)
        auto lb = std::dynamic_pointer_cast<load_balance>(strategy_);
        const load_balance strat = lb ? lb->resolved(*this->get_executor())
                                      : load_balance().resolved(*this->get_executor());
        const int64 nnz = static_cast<int64>(values_.get_size());
        srow_ = array<I>(this->get_executor(), static_cast<size_type>(strat.clac_size(nnz)));
        strat.process(this->get_executor(), this->get_size().rows, row_ptrs_.get_const_data(),
                      nnz, &srow_);
    } else {
        srow_.clear();
    }
    if (name == "merge_path") {
        const int64 wb = gkomi_csr_strategy_work_bytes(
            GKOMI_CSR_MERGE_PATH, static_cast<int64>(this->get_size().rows),
            static_cast<int64>(values_.get_size()));
        strategy_work_ = array<char>(this->get_executor(), static_cast<size_type>(wb));
    } else {
        strategy_work_.clear();
    }
}

template <typename V, typename I>
size_type Csr<V, I>::max_row_nnz() const
{
    if (max_row_nnz_ < 0) {
        array<int64> out(this->get_executor(), 1);
        GKO_CALL(gkomi_csr_max_row_nnz(ctx_of(this->get_executor()), this->get_size().rows,
                                       row_ptrs_.get_const_data(), out.get_data()));
        max_row_nnz_ = out.to_host()[0];
    }
    return static_cast<size_type>(max_row_nnz_);
}

template <typename V, typename I>
gkomi_csr Csr<V, I>::descriptor() const
{
    gkomi_csr d{};
    d.n_rows = this->get_size().rows;
    d.n_cols = this->get_size().cols;
    d.nnz = values_.get_size();
    d.row_ptrs = row_ptrs_.get_const_data();
    d.col_idxs = col_idxs_.get_const_data();
    d.values = values_.get_const_data();
    d.strategy = effective_ == "load_balance"
                     ? GKOMI_CSR_LOAD_BALANCE
                     : (effective_ == "merge_path" ? GKOMI_CSR_MERGE_PATH : GKOMI_CSR_CLASSICAL);
    d.subwave = subwave_;
    d.srow = srow_.get_const_data();
    d.num_srow = static_cast<int64>(srow_.get_size());
    d.col_offsets = compact_ ? col_offsets_.get_const_data() : nullptr;
    d.work = strategy_work_.get_data();
    d.work_bytes = static_cast<int64>(strategy_work_.get_size());
    if (panels_valid_ && num_panels_ > 0) {
        d.panels = panels_.get_data();
        d.panels_bytes = static_cast<int64>(panels_.get_size());
        d.num_panels = num_panels_;
    }
    if (slices_valid_ && slices_) {
        d.slice_sets = slices_->get_const_slice_sets();
        d.slice_cols = slices_->get_const_col_idxs();
        d.slice_values = slices_->get_const_values();
        d.slice_offsets = slice_offsets_valid_ ? slice_offsets_.get_const_data() : nullptr;
    }
    return d;
}

template <typename V, typename I>
void Csr<V, I>::build_slices(bool compact)
{
    slices_valid_ = false;
    slice_offsets_valid_ = false;
    slices_.reset();
    slice_offsets_.clear();
    auto exec = this->get_executor();
    if (exec->is_host()) {
        GKO_NOT_SUPPORTED_MSG("the sliced copy lives on a device executor");
    }
    std::shared_ptr<const Sellp<V, I>> s(to_sellp(64, 1));
    if (compact && s->get_num_stored_elements() > 0) {
        array<int16> off(exec, s->get_num_stored_elements());
        array<int32> fits(exec, 1);
        const auto sd = s->descriptor();
        GKO_CALL(gkomi_sellp_col_offsets(ctx_of(exec), &sd, off.get_data(), fits.get_data()));
        if (exec->copy_val_to_host(fits.get_const_data()) == 1) {
            slice_offsets_ = std::move(off);
            slice_offsets_valid_ = true;
        }
    }
    slices_ = std::move(s);
    slices_valid_ = true;
}

template <typename V, typename I>
void Csr<V, I>::build_column_panels(int num_panels)
{
    panels_valid_ = false;
    num_panels_ = 0;
    if (num_panels == 0) {
        panels_.clear();
        return;
    }
    const int64 nb = gkomi_csr_panels_bytes(static_cast<int64>(this->get_size().rows),
                                            static_cast<int64>(values_.get_size()), num_panels,
                                            value_code<V>());
    if (nb < 0) {
        throw Error(__FILE__, __LINE__, "column panels: num_panels must be 2, 4 or 8");
    }
    auto exec = this->get_executor();
    if (exec->is_host()) {
        GKO_NOT_SUPPORTED_MSG("column panels live on a device executor");
    }
    panels_ = array<char>(exec, static_cast<size_type>(nb));
    auto d = descriptor();
    const auto build = std::is_same<V, double>::value ? gkomi_csr_build_panels_f64
                                                      : gkomi_csr_build_panels_f32;
    GKO_CALL(build(ctx_of(exec), &d, num_panels, panels_.get_data(), nb));
    num_panels_ = num_panels;
    panels_valid_ = true;
}

template <typename V, typename I>
bool Csr<V, I>::compact_indices()
{
    auto exec = this->get_executor();
    const auto ctx = ctx_of(exec);
    const size_type nnz = values_.get_size();
    if (nnz == 0) return false;
    array<int16> off(exec, nnz);
    array<int32> fits(exec, 1);
    auto d = descriptor();
    d.col_offsets = nullptr;
    GKO_CALL(gkomi_csr_col_offsets(ctx, &d, off.get_data(), fits.get_data()));
    if (exec->copy_val_to_host(fits.get_const_data()) == 0) return false;
    col_offsets_ = std::move(off);
    compact_ = true;
    return true;
}

// Csr::apply_impl (core/matrix/csr.cpp:257-315) -> csr::spmv / advanced_spmv
// over mixed_precision_dispatch: b and x may each be Dense<double> or
// Dense<float>; alpha is taken in the matrix's precision, beta in x's
// (make_temporary_conversion, csr.cpp:297-303).
template <typename V, typename I>
void Csr<V, I>::apply_impl(const LinOp* b, LinOp* x) const
{
    mixed_precision_dispatch<V>(
        [this](auto db, auto dx) {
            using TB = typename std::decay_t<decltype(*db)>::value_type;
            using TC = typename std::decay_t<decltype(*dx)>::value_type;
            const auto d = descriptor();
            GKO_CALL(gkomi_csr_spmv_mixed(ctx_of(this->get_executor()), value_code<V>(), &d,
                                          db->get_size().cols, value_code<TB>(),
                                          db->get_const_values(), db->get_stride(),
                                          value_code<TC>(), dx->get_values(), dx->get_stride()));
        },
        b, x);
}

template <typename V, typename I>
void Csr<V, I>::apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                           LinOp* x) const
{
    mixed_precision_dispatch<V>(
        [this, alpha, beta](auto db, auto dx) {
            using TB = typename std::decay_t<decltype(*db)>::value_type;
            using TC = typename std::decay_t<decltype(*dx)>::value_type;
            detail::dense_in<V> da(alpha);
            detail::dense_in<TC> dbeta(beta);
            const auto d = descriptor();
            GKO_CALL(gkomi_csr_advanced_spmv_mixed(
                ctx_of(this->get_executor()), da.ptr->get_const_values(), value_code<V>(), &d,
                db->get_size().cols, value_code<TB>(), db->get_const_values(), db->get_stride(),
                dbeta.ptr->get_const_values(), value_code<TC>(), dx->get_values(),
                dx->get_stride()));
        },
        b, x);
}

template <typename V, typename I>
std::unique_ptr<Csr<V, I>> Csr<V, I>::transpose() const
{
    auto exec = this->get_executor();
    const auto s = this->get_size();
    const auto nnz = values_.get_size();
    array<I> t_rp(exec, s.cols + 1);
    array<I> t_ci(exec, nnz);
    array<V> t_v(exec, nnz);
    const size_type wsb = gkomi_csr_transpose_workspace_bytes(s.rows, s.cols, nnz);
    array<char> ws(exec, wsb);
    const auto d = descriptor();
    GKO_CALL(sp<V>::transpose(ctx_of(exec), &d, t_rp.get_data(), t_ci.get_data(),
                              t_v.get_data(), ws.get_data(), wsb));
    return create(exec, dim<2>{s.cols, s.rows}, std::move(t_v), std::move(t_ci),
                  std::move(t_rp), strategy_);
}

template <typename V, typename I>
std::unique_ptr<Sellp<V, I>> Csr<V, I>::to_sellp(size_type slice_size,
                                                 size_type stride_factor) const
{
    auto exec = this->get_executor();
    const auto n = this->get_size().rows;
    const size_type num_slices = (n + slice_size - 1) / slice_size;
    auto m = std::unique_ptr<Sellp<V, I>>(
        new Sellp<V, I>(exec, this->get_size(), slice_size, stride_factor));
    m->slice_sets_ = array<int64>(exec, num_slices + 1);
    m->slice_lengths_ = array<int64>(exec, std::max<size_type>(num_slices, 1));
    const size_type wsb = gkomi_scan_workspace_bytes(num_slices);
    array<char> ws(exec, wsb);
    GKO_CALL(gkomi_sellp_compute_slice_sets(ctx_of(exec), n, row_ptrs_.get_const_data(),
                                            slice_size, stride_factor,
                                            m->slice_sets_.get_data(),
                                            m->slice_lengths_.get_data(), ws.get_data(), wsb));
    // the one host read of the conversion: the total (core/matrix/csr.cpp:473-474)
    m->total_cols_ = static_cast<size_type>(
        exec->copy_val_to_host(m->slice_sets_.get_const_data() + num_slices));
    m->values_ = array<V>(exec, m->total_cols_ * slice_size);
    m->col_idxs_ = array<I>(exec, m->total_cols_ * slice_size);
    const auto d = descriptor();
    GKO_CALL(sp<V>::to_sellp(ctx_of(exec), &d, slice_size, m->slice_sets_.get_const_data(),
                             m->slice_lengths_.get_const_data(), m->col_idxs_.get_data(),
                             m->values_.get_data()));
    return m;
}

template <typename V, typename I>
std::unique_ptr<Ell<V, I>> Csr<V, I>::to_ell() const
{
    auto exec = this->get_executor();
    const auto width = max_row_nnz();
    const auto n = this->get_size().rows;
    auto m = std::unique_ptr<Ell<V, I>>(new Ell<V, I>(exec, this->get_size(), width, n));
    const auto d = descriptor();
    GKO_CALL(sp<V>::to_ell(ctx_of(exec), &d, width, n, m->col_idxs_.get_data(),
                           m->values_.get_data()));
    return m;
}

template <typename V, typename I>
template <typename Other>
std::unique_ptr<Csr<Other, I>> Csr<V, I>::convert_precision() const
{
    auto exec = this->get_executor();
    const auto nnz = values_.get_size();
    array<Other> v(exec, nnz);
    if (nnz) {
        if constexpr (std::is_same<V, double>::value && std::is_same<Other, float>::value) {
            GKO_CALL(gkomi_convert_f64_to_f32(ctx_of(exec), nnz, values_.get_const_data(),
                                              v.get_data()));
        } else if constexpr (std::is_same<V, float>::value && std::is_same<Other, double>::value) {
            GKO_CALL(gkomi_convert_f32_to_f64(ctx_of(exec), nnz, values_.get_const_data(),
                                              v.get_data()));
        } else {
            v = values_;
        }
    }
    std::shared_ptr<typename Csr<Other, I>::strategy_type> st;
    if (effective_ == "load_balance") {
        st = std::make_shared<typename Csr<Other, I>::load_balance>();
    } else if (effective_ == "merge_path") {
        st = std::make_shared<typename Csr<Other, I>::merge_path>();
    } else {
        st = std::make_shared<typename Csr<Other, I>::classical>(subwave_);
    }
    return Csr<Other, I>::create(exec, this->get_size(), std::move(v), array<I>(exec, col_idxs_),
                                 array<I>(exec, row_ptrs_), st);
}

template <typename V, typename I>
std::unique_ptr<Csr<V, I>> Csr<V, I>::clone() const
{
    auto exec = this->get_executor();
    return create(exec, this->get_size(), array<V>(exec, values_), array<I>(exec, col_idxs_),
                  array<I>(exec, row_ptrs_), strategy_);
}

// -------------------------------------------------------------- Sellp ----

template <typename V, typename I>
gkomi_sellp Sellp<V, I>::descriptor() const
{
    gkomi_sellp d{};
    d.n_rows = this->get_size().rows;
    d.n_cols = this->get_size().cols;
    d.slice_size = slice_size_;
    d.stride_factor = stride_factor_;
    d.slice_sets = slice_sets_.get_const_data();
    d.slice_lengths = slice_lengths_.get_const_data();
    d.col_idxs = col_idxs_.get_const_data();
    d.values = values_.get_const_data();
    return d;
}

// Sellp::apply_impl (core/matrix/sellp.cpp) over mixed_precision_dispatch,
// as Csr.
template <typename V, typename I>
void Sellp<V, I>::apply_impl(const LinOp* b, LinOp* x) const
{
    mixed_precision_dispatch<V>(
        [this](auto db, auto dx) {
            using TB = typename std::decay_t<decltype(*db)>::value_type;
            using TC = typename std::decay_t<decltype(*dx)>::value_type;
            const auto d = descriptor();
            GKO_CALL(gkomi_sellp_spmv_mixed(ctx_of(this->get_executor()), value_code<V>(), &d,
                                            db->get_size().cols, value_code<TB>(),
                                            db->get_const_values(), db->get_stride(),
                                            value_code<TC>(), dx->get_values(),
                                            dx->get_stride()));
        },
        b, x);
}

template <typename V, typename I>
void Sellp<V, I>::apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                             LinOp* x) const
{
    mixed_precision_dispatch<V>(
        [this, alpha, beta](auto db, auto dx) {
            using TB = typename std::decay_t<decltype(*db)>::value_type;
            using TC = typename std::decay_t<decltype(*dx)>::value_type;
            detail::dense_in<V> da(alpha);
            detail::dense_in<TC> dbeta(beta);
            const auto d = descriptor();
            GKO_CALL(gkomi_sellp_advanced_spmv_mixed(
                ctx_of(this->get_executor()), da.ptr->get_const_values(), value_code<V>(), &d,
                db->get_size().cols, value_code<TB>(), db->get_const_values(), db->get_stride(),
                dbeta.ptr->get_const_values(), value_code<TC>(), dx->get_values(),
                dx->get_stride()));
        },
        b, x);
}

// ---------------------------------------------------------------- Ell ----

template <typename V, typename I>
gkomi_ell Ell<V, I>::descriptor() const
{
    gkomi_ell d{};
    d.n_rows = this->get_size().rows;
    d.n_cols = this->get_size().cols;
    d.width = width_;
    d.stride = stride_;
    d.col_idxs = col_idxs_.get_const_data();
    d.values = values_.get_const_data();
    return d;
}

// Ell::apply_impl (core/matrix/ell.cpp) over mixed_precision_dispatch, as Csr.
template <typename V, typename I>
void Ell<V, I>::apply_impl(const LinOp* b, LinOp* x) const
{
    mixed_precision_dispatch<V>(
        [this](auto db, auto dx) {
            using TB = typename std::decay_t<decltype(*db)>::value_type;
            using TC = typename std::decay_t<decltype(*dx)>::value_type;
            const auto d = descriptor();
            GKO_CALL(gkomi_ell_spmv_mixed(ctx_of(this->get_executor()), value_code<V>(), &d,
                                          db->get_size().cols, value_code<TB>(),
                                          db->get_const_values(), db->get_stride(),
                                          value_code<TC>(), dx->get_values(), dx->get_stride()));
        },
        b, x);
}

template <typename V, typename I>
void Ell<V, I>::apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                           LinOp* x) const
{
    mixed_precision_dispatch<V>(
        [this, alpha, beta](auto db, auto dx) {
            using TB = typename std::decay_t<decltype(*db)>::value_type;
            using TC = typename std::decay_t<decltype(*dx)>::value_type;
            detail::dense_in<V> da(alpha);
            detail::dense_in<TC> dbeta(beta);
            const auto d = descriptor();
            GKO_CALL(gkomi_ell_advanced_spmv_mixed(
                ctx_of(this->get_executor()), da.ptr->get_const_values(), value_code<V>(), &d,
                db->get_size().cols, value_code<TB>(), db->get_const_values(), db->get_stride(),
                dbeta.ptr->get_const_values(), value_code<TC>(), dx->get_values(),
                dx->get_stride()));
        },
        b, x);
}

template class Csr<double, int32>;
template class Csr<float, int32>;
template std::unique_ptr<Csr<float, int32>> Csr<double, int32>::convert_precision<float>() const;
template std::unique_ptr<Csr<double, int32>> Csr<float, int32>::convert_precision<double>() const;
template std::unique_ptr<Csr<double, int32>> Csr<double, int32>::convert_precision<double>() const;
template std::unique_ptr<Csr<float, int32>> Csr<float, int32>::convert_precision<float>() const;
template class Sellp<double, int32>;
template class Sellp<float, int32>;
template class Ell<double, int32>;
template class Ell<float, int32>;

}  // namespace matrix
}  // namespace gko
