// The batch module on the Mi355xExecutor: storage, the size-binned launch
// plan, and batch::solver::{Bicgstab, Cg} over gkomi_batch_{bicgstab,cg}_{solve,tune}.
#include "ginkgo_amd/batch.hpp"

#include <algorithm>
#include <numeric>
#include <string>

namespace gko {
namespace batch {

// ---------------------------------------------------------- MultiVector ----

template <typename V>
std::unique_ptr<MultiVector<V>> MultiVector<V>::create(std::shared_ptr<const Executor> exec,
                                                       batch_dim<2> size)
{
    if (size.common_size.cols != 1) {
        GKO_NOT_SUPPORTED_MSG("batch::MultiVector: one right-hand side per item");
    }
    const size_type total = size.num_batch_items * size.common_size.rows;
    return std::unique_ptr<MultiVector>(new MultiVector(exec, size, total));
}

template <typename V>
std::unique_ptr<MultiVector<V>> MultiVector<V>::create_from_host(
    std::shared_ptr<const Executor> exec, size_type num_items, size_type n,
    const std::vector<V>& values)
{
    if (values.size() != num_items * n) {
        GKO_DIMENSION_MISMATCH_MSG("batch::MultiVector: values do not match num_items * n");
    }
    auto mv = create(exec, batch_dim<2>{num_items, dim<2>{n, 1}});
    if (!values.empty()) {
        exec->copy_from(exec->get_master().get(), values.size(), values.data(), mv->get_values());
    }
    return mv;
}

template <typename V>
std::unique_ptr<MultiVector<V>> MultiVector<V>::create_uneven(
    std::shared_ptr<const Executor> exec, const std::vector<int64>& row_offsets,
    const std::vector<V>* values)
{
    if (row_offsets.empty() || row_offsets.front() != 0) {
        GKO_DIMENSION_MISMATCH_MSG("batch::MultiVector: row offsets must start at 0");
    }
    size_type max_rows = 0;
    for (size_type k = 0; k + 1 < row_offsets.size(); ++k) {
        if (row_offsets[k + 1] < row_offsets[k]) {
            GKO_DIMENSION_MISMATCH_MSG("batch::MultiVector: decreasing row offsets");
        }
        max_rows = std::max<size_type>(max_rows, row_offsets[k + 1] - row_offsets[k]);
    }
    const size_type total = static_cast<size_type>(row_offsets.back());
    std::unique_ptr<MultiVector> mv(new MultiVector(
        exec, batch_dim<2>{row_offsets.size() - 1, dim<2>{max_rows, 1}}, total));
    if (values) {
        if (values->size() != total) {
            GKO_DIMENSION_MISMATCH_MSG("batch::MultiVector: values do not match the offsets");
        }
        if (total) {
            exec->copy_from(exec->get_master().get(), total, values->data(), mv->get_values());
        }
    }
    return mv;
}

namespace matrix {

// ------------------------------------------------------------------ Csr ----

template <typename V, typename I>
std::unique_ptr<Csr<V, I>> Csr<V, I>::create_from_host(std::shared_ptr<const Executor> exec,
                                                       size_type num_items, size_type n,
                                                       const std::vector<I>& row_ptrs,
                                                       const std::vector<I>& col_idxs,
                                                       const std::vector<V>& values)
{
    return create_from_host(std::move(exec), num_items, dim<2>{n, n}, row_ptrs, col_idxs,
                            values);
}

template <typename V, typename I>
std::unique_ptr<Csr<V, I>> Csr<V, I>::create_from_host(std::shared_ptr<const Executor> exec,
                                                       size_type num_items, dim<2> size,
                                                       const std::vector<I>& row_ptrs,
                                                       const std::vector<I>& col_idxs,
                                                       const std::vector<V>& values)
{
    if (row_ptrs.size() != size.rows + 1 || row_ptrs.front() != 0 ||
        static_cast<size_type>(row_ptrs.back()) != col_idxs.size() ||
        values.size() != num_items * col_idxs.size()) {
        GKO_DIMENSION_MISMATCH_MSG("batch::matrix::Csr: inconsistent pattern or values");
    }
    for (size_type i = 0; i < size.rows; ++i) {
        if (row_ptrs[i + 1] < row_ptrs[i]) {
            GKO_DIMENSION_MISMATCH_MSG("batch::matrix::Csr: row pointers must not decrease");
        }
    }
    for (const auto c : col_idxs) {
        if (c < 0 || static_cast<size_type>(c) >= size.cols) {
            GKO_DIMENSION_MISMATCH_MSG("batch::matrix::Csr: column index outside the item");
        }
    }
    std::unique_ptr<Csr> m(new Csr(exec, batch_dim<2>{num_items, size}));
    m->row_ptrs_ = array<I>(exec, row_ptrs);
    m->col_idxs_ = array<I>(exec, col_idxs);
    m->values_ = array<V>(exec, values);
    return m;
}

template <typename V, typename I>
gkomi_batch_csr Csr<V, I>::descriptor() const
{
    gkomi_batch_csr d{};
    d.num_systems = static_cast<int64_t>(this->size_.num_batch_items);
    d.max_rows = static_cast<int64_t>(this->size_.common_size.rows);
    d.max_nnz = static_cast<int64_t>(col_idxs_.get_size());
    d.row_offsets = nullptr;
    d.row_ptrs = row_ptrs_.get_const_data();
    d.col_idxs = col_idxs_.get_const_data();
    d.values = values_.get_const_data();
    d.uniform_rows = d.max_rows;
    d.uniform_nnz = d.max_nnz;
    d.uniform_cols = static_cast<int64_t>(this->size_.common_size.cols);
    return d;
}

// ---------------------------------------------------- BatchMatrix apply ----

namespace {
template <typename V>
struct bm;
template <>
struct bm<double> {
    static constexpr auto spmv = gkomi_batch_spmv_f64;
};
template <>
struct bm<float> {
    static constexpr auto spmv = gkomi_batch_spmv_f32;
};
}  // namespace

template <typename V>
void BatchMatrix<V>::check_apply(const MultiVector<V>* b, const MultiVector<V>* x) const
{
    if (!b || !x) GKO_NOT_SUPPORTED_MSG("batch apply: null vector");
    const auto num = this->get_num_batch_items();
    if (b->get_num_batch_items() != num || x->get_num_batch_items() != num ||
        b->get_total_rows() != get_total_cols() || x->get_total_rows() != get_total_rows()) {
        GKO_DIMENSION_MISMATCH_MSG("batch apply: vectors do not match the matrix");
    }
}

template <typename V>
void BatchMatrix<V>::apply(const MultiVector<V>* b, MultiVector<V>* x) const
{
    check_apply(b, x);
    const auto d = descriptor();
    GKO_CALL(bm<V>::spmv(ctx_of(this->get_executor()), &d, nullptr, b->get_const_values(),
                         nullptr, x->get_values()));
}

template <typename V>
void BatchMatrix<V>::apply(const MultiVector<V>* alpha, const MultiVector<V>* b,
                           const MultiVector<V>* beta, MultiVector<V>* x) const
{
    check_apply(b, x);
    const auto num = this->get_num_batch_items();
    if (!alpha || !beta || alpha->get_num_batch_items() != num ||
        beta->get_num_batch_items() != num || alpha->get_total_rows() != num ||
        beta->get_total_rows() != num) {
        GKO_DIMENSION_MISMATCH_MSG("batch apply: alpha and beta need one value per item");
    }
    const auto d = descriptor();
    GKO_CALL(bm<V>::spmv(ctx_of(this->get_executor()), &d, alpha->get_const_values(),
                         b->get_const_values(), beta->get_const_values(), x->get_values()));
}

// ------------------------------------------------------------------ Ell ----

template <typename V, typename I>
std::unique_ptr<Ell<V, I>> Ell<V, I>::create_from_host(std::shared_ptr<const Executor> exec,
                                                       size_type num_items, dim<2> size,
                                                       size_type num_elems_per_row,
                                                       const std::vector<I>& col_idxs,
                                                       const std::vector<V>& values)
{
    const size_type per_item = size.rows * num_elems_per_row;
    if (col_idxs.size() != per_item || values.size() != num_items * per_item) {
        GKO_DIMENSION_MISMATCH_MSG(
            "batch::matrix::Ell: col_idxs must hold rows * num_elems_per_row entries and "
            "values num_items times that");
    }
    for (auto c : col_idxs) {
        if (c < -1 || (c >= 0 && static_cast<size_type>(c) >= size.cols)) {
            GKO_DIMENSION_MISMATCH_MSG("batch::matrix::Ell: column index out of range");
        }
    }
    std::unique_ptr<Ell> m(new Ell(exec, batch_dim<2>{num_items, size}));
    m->width_ = num_elems_per_row;
    m->col_idxs_ = array<I>(exec, col_idxs);
    m->values_ = array<V>(exec, values);
    return m;
}

template <typename V, typename I>
std::unique_ptr<Ell<V, I>> Ell<V, I>::create_from_dense_host(std::shared_ptr<const Executor> exec,
                                                             size_type num_items, dim<2> size,
                                                             const std::vector<V>& dense)
{
    const size_type rows = size.rows, cols = size.cols;
    if (dense.size() != num_items * rows * cols) {
        GKO_DIMENSION_MISMATCH_MSG("batch::matrix::Ell: dense data must be items * rows * cols");
    }
    // the union pattern, row by row
    std::vector<std::vector<I>> pattern(rows);
    size_type width = 0;
    for (size_type i = 0; i < rows; ++i) {
        for (size_type j = 0; j < cols; ++j) {
            for (size_type k = 0; k < num_items; ++k) {
                if (dense[(k * rows + i) * cols + j] != V{}) {
                    pattern[i].push_back(static_cast<I>(j));
                    break;
                }
            }
        }
        width = std::max(width, pattern[i].size());
    }
    std::vector<I> ci(rows * width, I{-1});
    std::vector<V> vals(num_items * rows * width, V{});
    for (size_type i = 0; i < rows; ++i) {
        for (size_type s = 0; s < pattern[i].size(); ++s) {
            const size_type j = static_cast<size_type>(pattern[i][s]);
            ci[i + s * rows] = pattern[i][s];
            for (size_type k = 0; k < num_items; ++k) {
                vals[k * rows * width + i + s * rows] = dense[(k * rows + i) * cols + j];
            }
        }
    }
    return create_from_host(std::move(exec), num_items, size, width, ci, vals);
}

template <typename V, typename I>
gkomi_batch_csr Ell<V, I>::descriptor() const
{
    gkomi_batch_csr d{};
    d.num_systems = static_cast<int64_t>(this->size_.num_batch_items);
    d.max_rows = static_cast<int64_t>(this->size_.common_size.rows);
    d.max_nnz = static_cast<int64_t>(col_idxs_.get_size());
    d.col_idxs = col_idxs_.get_const_data();
    d.values = values_.get_const_data();
    d.uniform_rows = d.max_rows;
    d.uniform_nnz = d.max_nnz;
    d.format = GKOMI_BATCH_ELL;
    d.ell_width = static_cast<int32_t>(width_);
    d.uniform_cols = static_cast<int64_t>(this->size_.common_size.cols);
    return d;
}

// ---------------------------------------------------------------- Dense ----

template <typename V>
std::unique_ptr<Dense<V>> Dense<V>::create_from_host(std::shared_ptr<const Executor> exec,
                                                     size_type num_items, dim<2> size,
                                                     const std::vector<V>& values)
{
    if (values.size() != num_items * size.rows * size.cols) {
        GKO_DIMENSION_MISMATCH_MSG("batch::matrix::Dense: values must be items * rows * cols");
    }
    std::unique_ptr<Dense> m(new Dense(exec, batch_dim<2>{num_items, size}));
    m->values_ = array<V>(exec, values);
    return m;
}

template <typename V>
gkomi_batch_csr Dense<V>::descriptor() const
{
    gkomi_batch_csr d{};
    d.num_systems = static_cast<int64_t>(this->size_.num_batch_items);
    d.max_rows = static_cast<int64_t>(this->size_.common_size.rows);
    d.max_nnz = static_cast<int64_t>(this->size_.common_size.rows * this->size_.common_size.cols);
    d.values = values_.get_const_data();
    d.uniform_rows = d.max_rows;
    d.uniform_nnz = d.max_nnz;
    d.format = GKOMI_BATCH_DENSE;
    d.uniform_cols = static_cast<int64_t>(this->size_.common_size.cols);
    return d;
}

// --------------------------------------------------------- BlockDiagCsr ----

template <typename V, typename I>
std::unique_ptr<BlockDiagCsr<V, I>> BlockDiagCsr<V, I>::create_from_host(
    std::shared_ptr<const Executor> exec, const std::vector<int64>& row_offsets,
    const std::vector<I>& row_ptrs, const std::vector<I>& col_idxs, const std::vector<V>& values)
{
    if (row_offsets.empty() || row_offsets.front() != 0 ||
        row_ptrs.size() != static_cast<size_type>(row_offsets.back()) + 1 ||
        static_cast<size_type>(row_ptrs.back()) != col_idxs.size() ||
        values.size() != col_idxs.size()) {
        GKO_DIMENSION_MISMATCH_MSG("batch::matrix::BlockDiagCsr: inconsistent arrays");
    }
    const size_type num = row_offsets.size() - 1;
    std::vector<int64> rows(num), nnz(num), max_row(num, 0);
    int64 max_rows = 0;
    for (size_type k = 0; k < num; ++k) {
        rows[k] = row_offsets[k + 1] - row_offsets[k];
        if (rows[k] < 0) GKO_DIMENSION_MISMATCH_MSG("BlockDiagCsr: decreasing row offsets");
        nnz[k] = row_ptrs[row_offsets[k + 1]] - row_ptrs[row_offsets[k]];
        for (int64 r = row_offsets[k]; r < row_offsets[k + 1]; ++r) {
            max_row[k] = std::max<int64>(max_row[k], row_ptrs[r + 1] - row_ptrs[r]);
            for (I p = row_ptrs[r]; p < row_ptrs[r + 1]; ++p) {
                if (col_idxs[p] < 0 || col_idxs[p] >= rows[k]) {
                    GKO_DIMENSION_MISMATCH_MSG(
                        "BlockDiagCsr: column index outside its item (indices are local)");
                }
            }
        }
        max_rows = std::max(max_rows, rows[k]);
    }
    std::unique_ptr<BlockDiagCsr> m(new BlockDiagCsr(
        exec, batch_dim<2>{num, dim<2>{static_cast<size_type>(max_rows),
                                       static_cast<size_type>(max_rows)}}));
    m->row_offsets_ = array<int64>(exec, row_offsets);
    m->row_ptrs_ = array<I>(exec, row_ptrs);
    m->col_idxs_ = array<I>(exec, col_idxs);
    m->values_ = array<V>(exec, values);
    m->row_offsets_host_ = row_offsets;
    m->item_rows_ = std::move(rows);
    m->item_nnz_ = std::move(nnz);
    m->item_max_row_nnz_ = std::move(max_row);
    return m;
}

template <typename V, typename I>
gkomi_batch_csr BlockDiagCsr<V, I>::descriptor() const
{
    gkomi_batch_csr d{};
    d.num_systems = static_cast<int64_t>(this->size_.num_batch_items);
    d.max_rows = static_cast<int64_t>(this->size_.common_size.rows);
    d.max_nnz = item_nnz_.empty() ? 0 : *std::max_element(item_nnz_.begin(), item_nnz_.end());
    d.row_offsets = row_offsets_.get_const_data();
    d.row_ptrs = row_ptrs_.get_const_data();
    d.col_idxs = col_idxs_.get_const_data();
    d.values = values_.get_const_data();
    d.uniform_rows = 0;
    d.uniform_nnz = 0;
    return d;
}

// ------------------------------------------- BlockDiagEll / BlockDiagDense ----

namespace {

// item k's rows from row offsets, checked
std::vector<int64> item_rows_of(const std::vector<int64>& row_offsets, const char* who)
{
    if (row_offsets.empty() || row_offsets.front() != 0) {
        GKO_DIMENSION_MISMATCH_MSG(std::string(who) + ": row offsets must start at 0");
    }
    std::vector<int64> rows(row_offsets.size() - 1);
    for (size_type k = 0; k + 1 < row_offsets.size(); ++k) {
        rows[k] = row_offsets[k + 1] - row_offsets[k];
        if (rows[k] < 0) GKO_DIMENSION_MISMATCH_MSG(std::string(who) + ": decreasing row offsets");
    }
    return rows;
}

template <typename I>
void check_block_diag_csr(const std::vector<int64>& row_offsets, const std::vector<I>& row_ptrs,
                          const std::vector<I>& col_idxs, size_type num_values, const char* who)
{
    if (row_offsets.empty() || row_ptrs.size() != static_cast<size_type>(row_offsets.back()) + 1 ||
        static_cast<size_type>(row_ptrs.back()) != col_idxs.size() ||
        num_values != col_idxs.size()) {
        GKO_DIMENSION_MISMATCH_MSG(std::string(who) + ": inconsistent arrays");
    }
    for (size_type k = 0; k + 1 < row_offsets.size(); ++k) {
        const int64 n = row_offsets[k + 1] - row_offsets[k];
        for (int64 r = row_offsets[k]; r < row_offsets[k + 1]; ++r) {
            if (row_ptrs[r + 1] < row_ptrs[r]) {
                GKO_DIMENSION_MISMATCH_MSG(std::string(who) + ": decreasing row pointers");
            }
            for (I p = row_ptrs[r]; p < row_ptrs[r + 1]; ++p) {
                if (col_idxs[p] < 0 || col_idxs[p] >= n) {
                    GKO_DIMENSION_MISMATCH_MSG(std::string(who) +
                                               ": column index outside its item (local indices)");
                }
            }
        }
    }
}

}  // namespace

template <typename V>
BlockDiagBase<V>::BlockDiagBase(std::shared_ptr<const Executor> exec,
                                const std::vector<int64>& row_offsets,
                                const std::vector<int64>& item_offsets,
                                const std::vector<V>& values)
    : BatchMatrix<V>(exec, batch_dim<2>{}),
      row_offsets_(exec, row_offsets),
      item_offsets_(exec, item_offsets),
      values_(exec, values),
      row_offsets_host_(row_offsets),
      item_offsets_host_(item_offsets)
{
    item_rows_ = item_rows_of(row_offsets, "batch BlockDiag matrix");
    const size_type num = item_rows_.size();
    item_nnz_.resize(num);
    int64 max_rows = 0;
    for (size_type k = 0; k < num; ++k) {
        item_nnz_[k] = item_offsets[k + 1] - item_offsets[k];
        max_rows = std::max(max_rows, item_rows_[k]);
    }
    this->size_ = batch_dim<2>{num, dim<2>{static_cast<size_type>(max_rows),
                                           static_cast<size_type>(max_rows)}};
}

template <typename V>
gkomi_batch_csr BlockDiagBase<V>::base_descriptor(int32_t format) const
{
    gkomi_batch_csr d{};
    d.num_systems = static_cast<int64_t>(this->size_.num_batch_items);
    d.max_rows = static_cast<int64_t>(this->size_.common_size.rows);
    d.max_nnz = item_nnz_.empty() ? 0 : *std::max_element(item_nnz_.begin(), item_nnz_.end());
    d.row_offsets = row_offsets_.get_const_data();
    d.item_offsets = item_offsets_.get_const_data();
    d.values = values_.get_const_data();
    d.format = format;
    return d;
}

template <typename V, typename I>
BlockDiagEll<V, I>::BlockDiagEll(std::shared_ptr<const Executor> exec,
                                 const std::vector<int64>& row_offsets,
                                 const std::vector<int64>& item_offsets,
                                 const std::vector<I>& col_idxs, const std::vector<V>& values,
                                 std::vector<int64> widths)
    : BlockDiagBase<V>(exec, row_offsets, item_offsets, values),
      col_idxs_(exec, col_idxs),
      item_width_(std::move(widths))
{}

template <typename V, typename I>
std::unique_ptr<BlockDiagEll<V, I>> BlockDiagEll<V, I>::create_from_block_diag_csr_host(
    std::shared_ptr<const Executor> exec, const std::vector<int64>& row_offsets,
    const std::vector<I>& row_ptrs, const std::vector<I>& col_idxs, const std::vector<V>& values)
{
    const char* who = "batch::matrix::BlockDiagEll";
    const auto rows = item_rows_of(row_offsets, who);
    check_block_diag_csr(row_offsets, row_ptrs, col_idxs, values.size(), who);
    const size_type num = rows.size();
    std::vector<int64> widths(num, 0), offsets(num + 1, 0);
    for (size_type k = 0; k < num; ++k) {
        for (int64 r = row_offsets[k]; r < row_offsets[k + 1]; ++r) {
            widths[k] = std::max<int64>(widths[k], row_ptrs[r + 1] - row_ptrs[r]);
        }
        offsets[k + 1] = offsets[k] + rows[k] * widths[k];
    }
    if (offsets.back() >= (int64{1} << 31)) {
        GKO_NOT_SUPPORTED_MSG("batch::matrix::BlockDiagEll: more than 2^31 - 1 slots");
    }
    std::vector<I> ci(static_cast<size_type>(offsets.back()), I{-1});
    std::vector<V> v(static_cast<size_type>(offsets.back()), V{0});
    for (size_type k = 0; k < num; ++k) {
        for (int64 i = 0; i < rows[k]; ++i) {
            const int64 r = row_offsets[k] + i;
            for (I p = row_ptrs[r]; p < row_ptrs[r + 1]; ++p) {
                const int64 slot = offsets[k] + i + (p - row_ptrs[r]) * rows[k];
                ci[slot] = col_idxs[p];
                v[slot] = values[p];
            }
        }
    }
    return std::unique_ptr<BlockDiagEll>(
        new BlockDiagEll(exec, row_offsets, offsets, ci, v, std::move(widths)));
}

template <typename V, typename I>
gkomi_batch_csr BlockDiagEll<V, I>::descriptor() const
{
    auto d = this->base_descriptor(GKOMI_BATCH_ELL);
    d.col_idxs = col_idxs_.get_const_data();
    return d;
}

template <typename V>
std::unique_ptr<BlockDiagDense<V>> BlockDiagDense<V>::create_from_host(
    std::shared_ptr<const Executor> exec, const std::vector<int64>& row_offsets,
    const std::vector<V>& values)
{
    const auto rows = item_rows_of(row_offsets, "batch::matrix::BlockDiagDense");
    std::vector<int64> offsets(rows.size() + 1, 0);
    for (size_type k = 0; k < rows.size(); ++k) offsets[k + 1] = offsets[k] + rows[k] * rows[k];
    if (static_cast<size_type>(offsets.back()) != values.size()) {
        GKO_DIMENSION_MISMATCH_MSG(
            "batch::matrix::BlockDiagDense: values must hold n_k * n_k entries per item");
    }
    return std::unique_ptr<BlockDiagDense>(new BlockDiagDense(exec, row_offsets, offsets, values));
}

template <typename V>
template <typename I>
std::unique_ptr<BlockDiagDense<V>> BlockDiagDense<V>::create_from_block_diag_csr_host(
    std::shared_ptr<const Executor> exec, const std::vector<int64>& row_offsets,
    const std::vector<I>& row_ptrs, const std::vector<I>& col_idxs, const std::vector<V>& values)
{
    const char* who = "batch::matrix::BlockDiagDense";
    const auto rows = item_rows_of(row_offsets, who);
    check_block_diag_csr(row_offsets, row_ptrs, col_idxs, values.size(), who);
    std::vector<int64> offsets(rows.size() + 1, 0);
    for (size_type k = 0; k < rows.size(); ++k) offsets[k + 1] = offsets[k] + rows[k] * rows[k];
    std::vector<V> d(static_cast<size_type>(offsets.back()), V{0});
    for (size_type k = 0; k < rows.size(); ++k) {
        for (int64 i = 0; i < rows[k]; ++i) {
            const int64 r = row_offsets[k] + i;
            for (I p = row_ptrs[r]; p < row_ptrs[r + 1]; ++p) {
                d[static_cast<size_type>(offsets[k] + i * rows[k] + col_idxs[p])] += values[p];
            }
        }
    }
    return std::unique_ptr<BlockDiagDense>(new BlockDiagDense(exec, row_offsets, offsets, d));
}

template <typename V>
gkomi_batch_csr BlockDiagDense<V>::descriptor() const
{
    return this->base_descriptor(GKOMI_BATCH_DENSE);
}

template class BlockDiagBase<double>;
template class BlockDiagBase<float>;
template class BlockDiagEll<double, int32>;
template class BlockDiagEll<float, int32>;
template class BlockDiagDense<double>;
template class BlockDiagDense<float>;
template std::unique_ptr<BlockDiagDense<double>>
BlockDiagDense<double>::create_from_block_diag_csr_host<int32>(
    std::shared_ptr<const Executor>, const std::vector<int64>&, const std::vector<int32>&,
    const std::vector<int32>&, const std::vector<double>&);
template std::unique_ptr<BlockDiagDense<float>>
BlockDiagDense<float>::create_from_block_diag_csr_host<int32>(
    std::shared_ptr<const Executor>, const std::vector<int64>&, const std::vector<int32>&,
    const std::vector<int32>&, const std::vector<float>&);

}  // namespace matrix

// ------------------------------------------------------ Bicgstab and Cg ----

namespace solver {

namespace {

template <typename V, batch_method M>
struct bk;
template <>
struct bk<double, batch_method::bicgstab> {
    static constexpr auto solve = gkomi_batch_bicgstab_solve_f64;
    static constexpr auto tune = gkomi_batch_bicgstab_tune_f64;
    static constexpr const char* name = "batch::solver::Bicgstab";
};
template <>
struct bk<float, batch_method::bicgstab> {
    static constexpr auto solve = gkomi_batch_bicgstab_solve_f32;
    static constexpr auto tune = gkomi_batch_bicgstab_tune_f32;
    static constexpr const char* name = "batch::solver::Bicgstab";
};
template <>
struct bk<double, batch_method::cg> {
    static constexpr auto solve = gkomi_batch_cg_solve_f64;
    static constexpr auto tune = gkomi_batch_cg_tune_f64;
    static constexpr const char* name = "batch::solver::Cg";
};
template <>
struct bk<float, batch_method::cg> {
    static constexpr auto solve = gkomi_batch_cg_solve_f32;
    static constexpr auto tune = gkomi_batch_cg_tune_f32;
    static constexpr const char* name = "batch::solver::Cg";
};

std::string what_of(const char* who, const char* what)
{
    return std::string(who) + ": " + what;
}

// Size bins (rows in (lo, hi]) of the launch plan, largest first.
constexpr int64 kBinEdges[] = {int64{1} << 31, 512, 256, 128, 64, 0};

}  // namespace

template <typename V, batch_method M>
std::unique_ptr<BatchKrylov<V, M>> BatchKrylov<V, M>::Factory::generate(
    std::shared_ptr<const BatchLinOp> a) const
{
    return std::unique_ptr<BatchKrylov>(new BatchKrylov(exec_, params_, std::move(a)));
}

template <typename V, batch_method M>
BatchKrylov<V, M>::BatchKrylov(std::shared_ptr<const Executor> exec, parameters_type params,
                               std::shared_ptr<const BatchLinOp> a)
    : exec_{std::move(exec)},
      params_{std::move(params)},
      system_{std::move(a)},
      order_(exec_),
      workspace_(exec_),
      iterations_(exec_),
      residuals_(exec_)
{
    const char* who = bk<V, M>::name;
    if (!system_) GKO_NOT_SUPPORTED_MSG(what_of(who, "no system matrix"));
    if (params_.max_iterations < 0 || params_.tolerance < 0) {
        throw InvalidStateError(__FILE__, __LINE__,
                                what_of(who, "negative tolerance or iterations"));
    }
    if (system_->get_common_size().rows != system_->get_common_size().cols) {
        GKO_DIMENSION_MISMATCH_MSG(what_of(who, "items must be square"));
    }
    std::vector<int64> rows, nnz, max_row;
    const size_type num = system_->get_num_batch_items();
    auto mat = dynamic_cast<const matrix::BatchMatrix<V>*>(system_.get());
    if (!mat) {
        GKO_NOT_SUPPORTED_MSG(
            what_of(who, "system must be a batch::matrix (Csr, Ell, Dense, BlockDiagCsr, "
                         "BlockDiagEll or BlockDiagDense)"));
    }
    desc_ = mat->descriptor();
    rows = mat->get_item_rows();
    nnz = mat->get_item_nnz();
    max_row = mat->get_item_max_row_nnz();
    total_rows_ = mat->get_total_rows();
    // plan: items sorted by size (largest first), split into bins
    std::vector<int32> order(num);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32 x, int32 y) { return rows[x] > rows[y]; });
    order_ = array<int32>(exec_, order);
    plan_ = gkomi_batch_plan{};
    plan_.order = order_.get_const_data();
    size_type begin = 0;
    int nb = 0;
    for (int e = 0; e + 1 < static_cast<int>(sizeof(kBinEdges) / sizeof(kBinEdges[0])); ++e) {
        size_type end = begin;
        while (end < num && rows[order[end]] > kBinEdges[e + 1]) ++end;
        if (end > begin) {
            int64 mr = 0, mz = 0, ml = 0;
            bool known = true;
            for (size_type k = begin; k < end; ++k) {
                mr = std::max(mr, rows[order[k]]);
                mz = std::max(mz, nnz[order[k]]);
                ml = std::max(ml, max_row[order[k]]);
                known = known && (max_row[order[k]] > 0 || nnz[order[k]] == 0);
            }
            plan_.bin_begin[nb] = static_cast<int64_t>(begin);
            plan_.bin_max_rows[nb] = mr;
            plan_.bin_max_nnz[nb] = mz;
            plan_.bin_max_row_nnz[nb] = known ? ml : 0;
            plan_.block_size[nb] = mr > 256 ? 256 : (mr > 64 ? 128 : 64);
            plan_.matrix_in_lds[nb] = -1;
            plan_.blocks_per_cu[nb] = 0;
            ++nb;
            begin = end;
        }
    }
    if (nb == 0) {
        plan_.bin_max_rows[0] = 0;
        plan_.bin_max_nnz[0] = 0;
        plan_.block_size[0] = 64;
        nb = 1;
    }
    plan_.num_bins = nb;
    plan_.bin_begin[nb] = static_cast<int64_t>(num);
    workspace_ = array<char>(exec_, gkomi_batch_workspace_bytes(&desc_));
    workspace_.fill(0);
    iterations_ = array<int32>(exec_, num);
    residuals_ = array<V>(exec_, num);
}

template <typename V, batch_method M>
gkomi_batch_settings BatchKrylov<V, M>::settings() const
{
    gkomi_batch_settings s{};
    s.max_iterations = params_.max_iterations;
    s.tolerance_type = params_.tolerance_type == stop::tolerance_type::relative ? 1 : 0;
    s.tolerance = params_.tolerance;
    s.preconditioner = params_.preconditioner ? 1 : 0;
    return s;
}

template <typename V, batch_method M>
void BatchKrylov<V, M>::check_vectors(const MultiVector<V>* b, const MultiVector<V>* x) const
{
    const char* who = bk<V, M>::name;
    if (!b || !x) GKO_NOT_SUPPORTED_MSG(what_of(who, "null vector"));
    if (b->get_num_batch_items() != system_->get_num_batch_items() ||
        x->get_num_batch_items() != system_->get_num_batch_items() ||
        b->get_total_rows() != total_rows_ || x->get_total_rows() != total_rows_) {
        GKO_DIMENSION_MISMATCH_MSG(what_of(who, "vectors do not match the system"));
    }
}

template <typename V, batch_method M>
void BatchKrylov<V, M>::apply(const MultiVector<V>* b, MultiVector<V>* x) const
{
    check_vectors(b, x);
    const auto s = settings();
    using K = bk<V, M>;
    GKO_CALL(K::solve(ctx_of(exec_), &desc_, &s, &plan_, b->get_const_values(), x->get_values(),
                      iterations_.get_data(), residuals_.get_data(), workspace_.get_data(),
                      workspace_.get_size()));
    record_logs();
}

template <typename V, batch_method M>
void BatchKrylov<V, M>::record_logs() const
{
    if (!loggers_.empty()) {
        auto it = iterations_.to_host();
        auto res = residuals_.to_host();
        for (auto& l : loggers_) l->record(it, res);
    }
}

template <typename V, batch_method M>
std::unique_ptr<Graph> BatchKrylov<V, M>::capture(const MultiVector<V>* b, MultiVector<V>* x,
                                                  bool zero_initial_guess) const
{
    check_vectors(b, x);
    const auto s = settings();
    using K = bk<V, M>;
    gkomi_ctx cap{};
    GKO_CALL(gkomi_graph_begin(ctx_of(exec_), &cap));
    int rc = GKOMI_SUCCESS;
    if (zero_initial_guess) {
        if constexpr (std::is_same<V, double>::value) {
            rc = gkomi_fill_f64(&cap, static_cast<int64>(total_rows_), 0.0, x->get_values());
        } else {
            rc = gkomi_fill_f32(&cap, static_cast<int64>(total_rows_), 0.0f, x->get_values());
        }
    }
    if (rc == GKOMI_SUCCESS) {
        rc = K::solve(&cap, &desc_, &s, &plan_, b->get_const_values(), x->get_values(),
                      iterations_.get_data(), residuals_.get_data(), workspace_.get_data(),
                      workspace_.get_size());
    }
    // the capture ends in every case (its stream is destroyed there)
    gkomi_graph* g = nullptr;
    const int end_rc = gkomi_graph_end(&cap, &g);
    if (rc != GKOMI_SUCCESS) {
        gkomi_graph_destroy(g);
        GKO_CALL(rc);
    }
    GKO_CALL(end_rc);
    return std::unique_ptr<Graph>(new Graph(exec_, g));
}

template <typename V, batch_method M>
void BatchKrylov<V, M>::tune(const MultiVector<V>* b, const MultiVector<V>* x)
{
    check_vectors(b, x);
    const auto s = settings();
    array<V> scratch(exec_, total_rows_);
    using K = bk<V, M>;
    GKO_CALL(K::tune(ctx_of(exec_), &desc_, &s, &plan_, b->get_const_values(),
                     x->get_const_values(), scratch.get_data(), iterations_.get_data(),
                     residuals_.get_data(), workspace_.get_data(), workspace_.get_size()));
}

template <typename V, batch_method M>
std::vector<int> BatchKrylov<V, M>::get_block_sizes() const
{
    return std::vector<int>(plan_.block_size, plan_.block_size + plan_.num_bins);
}

template <typename V, batch_method M>
std::vector<int> BatchKrylov<V, M>::get_matrix_in_lds() const
{
    return std::vector<int>(plan_.matrix_in_lds, plan_.matrix_in_lds + plan_.num_bins);
}

template class BatchKrylov<double, batch_method::bicgstab>;
template class BatchKrylov<float, batch_method::bicgstab>;
template class BatchKrylov<double, batch_method::cg>;
template class BatchKrylov<float, batch_method::cg>;

}  // namespace solver

template class MultiVector<double>;
template class MultiVector<float>;
template class matrix::Csr<double, int32>;
template class matrix::Csr<float, int32>;
template class matrix::BlockDiagCsr<double, int32>;
template class matrix::BlockDiagCsr<float, int32>;
template class matrix::BatchMatrix<double>;
template class matrix::BatchMatrix<float>;
template class matrix::Ell<double, int32>;
template class matrix::Ell<float, int32>;
template class matrix::Dense<double>;
template class matrix::Dense<float>;

}  // namespace batch
}  // namespace gko
