// Sparse formats: Csr (include/ginkgo/core/matrix/csr.hpp:104-1690, with the
// classical / load_balance / merge_path / automatical strategies), Sellp
// (sellp.hpp), Ell (ell.hpp), Coo (coo.hpp) and Hybrid (hybrid.hpp), their
// conversions and the device transpose; plus matrix_data for assembling from
// host triplets.
#pragma once

#include <algorithm>
#include <string>
#include <tuple>
#include <vector>

#include "ginkgo_amd/lin_op.hpp"

namespace gko {

// include/ginkgo/core/base/matrix_data.hpp: host triplets.
template <typename ValueType = double, typename IndexType = int32>
struct matrix_data {
    struct nonzero_type {
        IndexType row;
        IndexType column;
        ValueType value;
    };
    dim<2> size;
    std::vector<nonzero_type> nonzeros;
    void sort_row_major()
    {
        std::stable_sort(nonzeros.begin(), nonzeros.end(), [](const auto& a, const auto& b) {
            return std::tie(a.row, a.column) < std::tie(b.row, b.column);
        });
    }
};

namespace matrix {

template <typename ValueType, typename IndexType>
class Sellp;
template <typename ValueType, typename IndexType>
class Ell;
template <typename ValueType, typename IndexType>
class Coo;
template <typename ValueType, typename IndexType>
class Hybrid;

template <typename ValueType = double, typename IndexType = int32>
class Csr : public LinOp {
    static_assert(std::is_same<IndexType, int32>::value, "the MI355X path uses int32 indices");

public:
    using value_type = ValueType;
    using index_type = IndexType;

    // csr::strategy_type (csr.hpp:176-699)
    class strategy_type {
    public:
        virtual ~strategy_type() = default;
        virtual std::string get_name() const = 0;
    };
    // Sub-wavefront per row.  `subwave` = lanes per row (bits 0..7, a power
    // of two <= 64) | rows in flight per lane group << 8 (1, 2, 4 or 8);
    // 0 in either field = chosen from the mean row length.
    class classical : public strategy_type {
    public:
        classical(int subwave = 0) : subwave_{subwave} {}
        std::string get_name() const override { return "classical"; }
        int subwave() const { return subwave_; }

    private:
        int subwave_;
    };
    // Csr::load_balance (csr.hpp:354-520): the srow partition of
    // clac_size(nnz) warps, built by process().  The default constructor
    // takes the executor's warps when the matrix builds its partition (the
    // reference's default asks CUDA device 0 for them, csr.hpp:359-364).
    class load_balance : public strategy_type {
    public:
        load_balance() = default;
        explicit load_balance(std::shared_ptr<const Mi355xExecutor> exec)
            : load_balance(exec->get_num_warps(), exec->get_warp_size(), false)
        {}
        load_balance(int64 nwarps, int warp_size = 32, bool cuda_strategy = true,
                     std::string strategy_name = "none")
            : nwarps_{nwarps},
              warp_size_{warp_size},
              cuda_strategy_{cuda_strategy},
              strategy_name_{std::move(strategy_name)}
        {}
        std::string get_name() const override { return "load_balance"; }
        // csr.hpp:475-505 (HIP-platform build)
        int64 clac_size(int64 nnz) const
        {
            return gkomi_csr_lb_clac_size(nnz, nwarps_, warp_size_, cuda_strategy_ ? 1 : 0,
                                          strategy_name_ == "intel" ? 1 : 0);
        }
        // csr.hpp:417-468, on the device (bit-identical integers)
        void process(std::shared_ptr<const Executor> exec, size_type n_rows,
                     const IndexType* row_ptrs, int64 nnz, array<IndexType>* srow) const
        {
            if (srow->get_size() == 0) return;
            GKO_CALL(gkomi_csr_lb_partition(ctx_of(exec), static_cast<int64>(n_rows), row_ptrs,
                                            nnz, warp_size_, srow->get_data(),
                                            static_cast<int64>(srow->get_size())));
        }
        bool has_warps() const { return nwarps_ > 0; }
        // the executor's warps for a default-constructed strategy
        load_balance resolved(const Executor& exec) const
        {
            if (has_warps()) return *this;
            auto m = dynamic_cast<const Mi355xExecutor*>(&exec);
            if (!m) return load_balance(1, 64, false);
            return load_balance(m->get_num_warps(), m->get_warp_size(), false);
        }

    private:
        int64 nwarps_{0};
        int warp_size_{64};
        bool cuda_strategy_{false};
        std::string strategy_name_{"none"};
    };
    // merge-path partition of rows and entries (csr.hpp:283-301): every
    // workgroup takes an equal share of n + nnz merge steps
    class merge_path : public strategy_type {
    public:
        merge_path() = default;
        std::string get_name() const override { return "merge_path"; }
    };
    // picks classical for short uniform rows, load_balance otherwise
    // (automatical::process, csr.hpp:606-673, decides from row statistics)
    class automatical : public strategy_type {
    public:
        automatical() = default;
        std::string get_name() const override { return "automatical"; }
    };
    // "use the sparse library's own kernel" (csr.hpp:332-349: cuSPARSE /
    // hipSPARSE there): on this executor the library is this one, so the
    // choice is made as automatical makes it.
    class sparselib : public strategy_type {
    public:
        sparselib() = default;
        std::string get_name() const override { return "sparselib"; }
    };

    static std::unique_ptr<Csr> create(std::shared_ptr<const Executor> exec, dim<2> size = {},
                                       size_type nnz = 0,
                                       std::shared_ptr<strategy_type> strategy = nullptr)
    {
        return std::unique_ptr<Csr>(new Csr(std::move(exec), size, nnz, std::move(strategy)));
    }
    // from arrays (moved in or copied to exec)
    static std::unique_ptr<Csr> create(std::shared_ptr<const Executor> exec, dim<2> size,
                                       array<ValueType> values, array<IndexType> col_idxs,
                                       array<IndexType> row_ptrs,
                                       std::shared_ptr<strategy_type> strategy = nullptr);
    // host CSR arrays copied to exec
    static std::unique_ptr<Csr> create_from_host(std::shared_ptr<const Executor> exec,
                                                 dim<2> size,
                                                 const std::vector<IndexType>& row_ptrs,
                                                 const std::vector<IndexType>& col_idxs,
                                                 const std::vector<ValueType>& values,
                                                 std::shared_ptr<strategy_type> strategy = nullptr);
    // non-owning view over device arrays
    static std::unique_ptr<Csr> create_view(std::shared_ptr<const Executor> exec, dim<2> size,
                                            size_type nnz, ValueType* values,
                                            IndexType* col_idxs, IndexType* row_ptrs,
                                            std::shared_ptr<strategy_type> strategy = nullptr);
    // Csr::read (core/matrix/csr.cpp:563-634): duplicates are summed
    static std::unique_ptr<Csr> create_from_data(std::shared_ptr<const Executor> exec,
                                                 matrix_data<ValueType, IndexType> data,
                                                 std::shared_ptr<strategy_type> strategy = nullptr);
    static std::unique_ptr<Csr> create_from_dense(std::shared_ptr<const Executor> exec,
                                                  const std::vector<std::vector<ValueType>>& rows,
                                                  std::shared_ptr<strategy_type> strategy = nullptr);

    // the mutable accessors drop the column panels and the sliced copy
    // (their copies of the matrix could go stale)
    ValueType* get_values() noexcept
    {
        panels_valid_ = false;
        slices_valid_ = false;
        return values_.get_data();
    }
    const ValueType* get_const_values() const noexcept { return values_.get_const_data(); }
    IndexType* get_col_idxs() noexcept
    {
        panels_valid_ = false;
        slices_valid_ = false;
        return col_idxs_.get_data();
    }
    const IndexType* get_const_col_idxs() const noexcept { return col_idxs_.get_const_data(); }
    IndexType* get_row_ptrs() noexcept
    {
        panels_valid_ = false;
        slices_valid_ = false;
        return row_ptrs_.get_data();
    }
    const IndexType* get_const_row_ptrs() const noexcept { return row_ptrs_.get_const_data(); }
    const IndexType* get_const_srow() const noexcept { return srow_.get_const_data(); }
    size_type get_num_srow_elements() const noexcept { return srow_.get_size(); }
    size_type get_num_stored_elements() const noexcept { return values_.get_size(); }
    std::shared_ptr<strategy_type> get_strategy() const { return strategy_; }
    // the strategy in effect after `automatical` was resolved
    std::string get_effective_strategy() const { return effective_; }
    void set_strategy(std::shared_ptr<strategy_type> s);

    // Csr::transpose (core/matrix/csr.cpp:1106-1116), built on the device
    std::unique_ptr<Csr> transpose() const;
    // Csr::convert_to(Sellp/Ell/Csr<other precision>) (csr.cpp:318-335, 457-529)
    std::unique_ptr<Sellp<ValueType, IndexType>> to_sellp(size_type slice_size = 64,
                                                         size_type stride_factor = 1) const;
    std::unique_ptr<Ell<ValueType, IndexType>> to_ell() const;
    // Csr::convert_to(Coo) (core/matrix/csr.cpp): row indices from the row pointers
    std::unique_ptr<Coo<ValueType, IndexType>> to_coo() const;
    template <typename Other>
    std::unique_ptr<Csr<Other, IndexType>> convert_precision() const;
    std::unique_ptr<Csr> clone() const;
    // host copies of the three arrays
    std::vector<IndexType> row_ptrs_to_host() const { return row_ptrs_.to_host(); }
    std::vector<IndexType> srow_to_host() const { return srow_.to_host(); }
    std::vector<IndexType> col_idxs_to_host() const { return col_idxs_.to_host(); }
    std::vector<ValueType> values_to_host() const { return values_.to_host(); }

    // 2-byte column offsets (col - row), read instead of the 4-byte indices
    // by the classical single right-hand-side SpMV and the fused CG; built
    // only when every offset fits int16 (returns whether it did).  Results
    // are bit-identical either way.  No reference counterpart.  Values or
    // indices changed later through a view require calling it again.
    bool compact_indices();
    bool has_compact_indices() const noexcept { return compact_; }

    // Column panels (gkomi_csr_build_panels_*, no reference counterpart): a
    // copy of the matrix split into num_panels (2, 4 or 8) column panels, so
    // that each XCD gathers the input vector from one panel's window; every
    // single right-hand-side apply then uses them, with any strategy.  Pays
    // off when the input vector exceeds one XCD's 4 MB L2 and the columns
    // are scattered (the 1M-row power-law matrix: 1.8x); for banded matrices
    // the extra row visits cost more than they save.  The copy is taken now:
    // get_values() / get_col_idxs() / get_row_ptrs() drop it, and writes
    // through a pointer obtained earlier need another call.  0 removes them.
    void build_column_panels(int num_panels = 4);
    int get_num_column_panels() const noexcept { return panels_valid_ ? num_panels_ : 0; }

    // Sliced copy (no reference counterpart): the matrix in SELL-P layout
    // (64-row slices), which the classical single right-hand-side apply and
    // the fused CG then read, one lane per row summing the row in its stored
    // order; with compact the slot columns are kept as 2-byte offsets col -
    // row when they fit.  Faster for rows of nearly equal length (27-point
    // FP32: 6.6 against 4.5 TB/s), slower when a slice holds one long row.
    // Dropped by the mutable accessors, like the column panels.
    void build_slices(bool compact = true);
    bool has_slices() const noexcept { return slices_valid_; }

    // the plain descriptor the C ABI takes
    gkomi_csr descriptor() const;
    // longest row (cached after the first query)
    size_type max_row_nnz() const;

protected:
    Csr(std::shared_ptr<const Executor> exec, dim<2> size, size_type nnz,
        std::shared_ptr<strategy_type> strategy)
        : LinOp(exec, size),
          values_(exec, nnz),
          col_idxs_(exec, nnz),
          row_ptrs_(exec, size.rows + 1),
          srow_(exec),
          strategy_work_(exec),
          col_offsets_(exec),
          panels_(exec),
          slice_offsets_(exec),
          strategy_{strategy ? std::move(strategy) : std::make_shared<automatical>()}
    {
        if (size.rows == 0 && nnz == 0) {
            row_ptrs_.fill(0);
        }
    }
    void apply_impl(const LinOp* b, LinOp* x) const override;
    void apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                    LinOp* x) const override;
    void make_srow();

private:
    array<ValueType> values_;
    array<IndexType> col_idxs_;
    array<IndexType> row_ptrs_;
    array<IndexType> srow_;
    mutable array<char> strategy_work_;  // merge_path carries
    array<int16> col_offsets_;
    bool compact_{false};
    mutable array<char> panels_;  // column panels (also written: partial sums)
    int num_panels_{0};
    bool panels_valid_{false};
    std::shared_ptr<const Sellp<ValueType, IndexType>> slices_;
    array<int16> slice_offsets_;
    bool slices_valid_{false};
    bool slice_offsets_valid_{false};
    std::shared_ptr<strategy_type> strategy_;
    std::string effective_{"classical"};
    int subwave_{0};
    mutable int64 max_row_nnz_{-1};
    template <typename, typename>
    friend class Csr;
    friend class Coo<ValueType, IndexType>;
};

template <typename ValueType = double, typename IndexType = int32>
class Sellp : public LinOp {
public:
    using value_type = ValueType;
    ValueType* get_values() noexcept { return values_.get_data(); }
    const ValueType* get_const_values() const noexcept { return values_.get_const_data(); }
    const IndexType* get_const_col_idxs() const noexcept { return col_idxs_.get_const_data(); }
    const int64* get_const_slice_sets() const noexcept { return slice_sets_.get_const_data(); }
    const int64* get_const_slice_lengths() const noexcept
    {
        return slice_lengths_.get_const_data();
    }
    size_type get_slice_size() const noexcept { return slice_size_; }
    size_type get_stride_factor() const noexcept { return stride_factor_; }
    size_type get_total_cols() const noexcept { return total_cols_; }
    size_type get_num_stored_elements() const noexcept { return values_.get_size(); }
    gkomi_sellp descriptor() const;
    std::vector<int64> slice_sets_to_host() const { return slice_sets_.to_host(); }
    std::vector<IndexType> col_idxs_to_host() const { return col_idxs_.to_host(); }
    std::vector<ValueType> values_to_host() const { return values_.to_host(); }

protected:
    Sellp(std::shared_ptr<const Executor> exec, dim<2> size, size_type slice_size,
          size_type stride_factor)
        : LinOp(exec, size),
          values_(exec),
          col_idxs_(exec),
          slice_sets_(exec),
          slice_lengths_(exec),
          slice_size_{slice_size},
          stride_factor_{stride_factor}
    {}
    void apply_impl(const LinOp* b, LinOp* x) const override;
    void apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                    LinOp* x) const override;

private:
    array<ValueType> values_;
    array<IndexType> col_idxs_;
    array<int64> slice_sets_;
    array<int64> slice_lengths_;
    size_type slice_size_;
    size_type stride_factor_;
    size_type total_cols_{};
    friend class Csr<ValueType, IndexType>;
};

template <typename ValueType = double, typename IndexType = int32>
class Ell : public LinOp {
public:
    using value_type = ValueType;
    const ValueType* get_const_values() const noexcept { return values_.get_const_data(); }
    const IndexType* get_const_col_idxs() const noexcept { return col_idxs_.get_const_data(); }
    size_type get_num_stored_elements_per_row() const noexcept { return width_; }
    size_type get_stride() const noexcept { return stride_; }
    gkomi_ell descriptor() const;
    std::vector<IndexType> col_idxs_to_host() const { return col_idxs_.to_host(); }
    std::vector<ValueType> values_to_host() const { return values_.to_host(); }

protected:
    Ell(std::shared_ptr<const Executor> exec, dim<2> size, size_type width, size_type stride)
        : LinOp(exec, size),
          values_(exec, width * stride),
          col_idxs_(exec, width * stride),
          width_{width},
          stride_{stride}
    {}
    void apply_impl(const LinOp* b, LinOp* x) const override;
    void apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                    LinOp* x) const override;

private:
    array<ValueType> values_;
    array<IndexType> col_idxs_;
    size_type width_;
    size_type stride_;
    friend class Csr<ValueType, IndexType>;
    friend class Hybrid<ValueType, IndexType>;
};

// matrix::Coo (include/ginkgo/core/matrix/coo.hpp): one (row, column, value)
// triplet per entry, rows sorted (the order within a row is free, as the
// reference's unsorted tests use).  The products read the entries through a
// Csr view over their row-pointer index, built once (no reference
// counterpart): 4 (n + 1) bytes of row pointers per product instead of
// 4 nnz bytes of row indices, and the Csr strategy's kernels.
template <typename ValueType = double, typename IndexType = int32>
class Coo : public LinOp {
public:
    using value_type = ValueType;
    using index_type = IndexType;
    static std::unique_ptr<Coo> create(std::shared_ptr<const Executor> exec, dim<2> size = {},
                                       size_type nnz = 0);
    // host triplets copied to exec; throws NotSupported when rows are not sorted
    static std::unique_ptr<Coo> create_from_host(std::shared_ptr<const Executor> exec,
                                                 dim<2> size,
                                                 const std::vector<IndexType>& row_idxs,
                                                 const std::vector<IndexType>& col_idxs,
                                                 const std::vector<ValueType>& values);
    ValueType* get_values() noexcept { return values_.get_data(); }
    const ValueType* get_const_values() const noexcept { return values_.get_const_data(); }
    const IndexType* get_const_col_idxs() const noexcept { return col_idxs_.get_const_data(); }
    const IndexType* get_const_row_idxs() const noexcept { return row_idxs_.get_const_data(); }
    size_type get_num_stored_elements() const noexcept { return values_.get_size(); }
    std::vector<IndexType> row_idxs_to_host() const { return row_idxs_.to_host(); }
    std::vector<IndexType> col_idxs_to_host() const { return col_idxs_.to_host(); }
    std::vector<ValueType> values_to_host() const { return values_.to_host(); }
    // Coo::apply2 (coo.hpp): x += A b and x += alpha A b
    void apply2(const LinOp* b, LinOp* x) const;
    void apply2(const LinOp* alpha, const LinOp* b, LinOp* x) const;
    // Coo::convert_to(Csr)
    std::unique_ptr<Csr<ValueType, IndexType>> to_csr() const;
    // the Csr view the products run on (shares this matrix's arrays)
    const Csr<ValueType, IndexType>* get_index() const { return index_.get(); }

protected:
    Coo(std::shared_ptr<const Executor> exec, dim<2> size, size_type nnz)
        : LinOp(exec, size),
          values_(exec, nnz),
          col_idxs_(exec, nnz),
          row_idxs_(exec, nnz),
          row_ptrs_(exec, size.rows + 1)
    {}
    void apply_impl(const LinOp* b, LinOp* x) const override;
    void apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                    LinOp* x) const override;
    // row_ptrs_ from row_idxs_ (or already filled), then the Csr view
    void build_index(bool ptrs_ready);

private:
    array<ValueType> values_;
    array<IndexType> col_idxs_;
    array<IndexType> row_idxs_;
    array<IndexType> row_ptrs_;
    std::unique_ptr<Csr<ValueType, IndexType>> index_;
    friend class Csr<ValueType, IndexType>;
    friend class Hybrid<ValueType, IndexType>;
};

// matrix::Hybrid (include/ginkgo/core/matrix/hybrid.hpp): an Ell part holding
// the first K entries of every row and a Coo part with the rest; K comes from
// the strategy and the row lengths (hybrid.hpp strategy classes, restated).
// apply = Ell product, then the Coo part added (core/matrix/hybrid.cpp).
template <typename ValueType = double, typename IndexType = int32>
class Hybrid : public LinOp {
public:
    using value_type = ValueType;
    using index_type = IndexType;

    class strategy_type {
    public:
        virtual ~strategy_type() = default;
        virtual std::string get_name() const = 0;
        // K from the row lengths (the vector may be reordered)
        virtual size_type compute_ell_num_stored_elements_per_row(
            std::vector<size_type>& row_nnz) const = 0;
    };
    // K = num_columns
    class column_limit : public strategy_type {
    public:
        explicit column_limit(size_type num_column = 0) : num_columns_{num_column} {}
        std::string get_name() const override { return "column_limit"; }
        size_type compute_ell_num_stored_elements_per_row(
            std::vector<size_type>&) const override
        {
            return num_columns_;
        }
        size_type get_num_columns() const { return num_columns_; }

    private:
        size_type num_columns_;
    };
    // K = the row length at position floor(n * percent) of the sorted
    // lengths (the longest when percent >= 1); percent clamped to [0, 1]
    class imbalance_limit : public strategy_type {
    public:
        explicit imbalance_limit(double percent = 0.8)
            : percent_{std::max(0.0, std::min(percent, 1.0))}
        {}
        std::string get_name() const override { return "imbalance_limit"; }
        size_type compute_ell_num_stored_elements_per_row(
            std::vector<size_type>& row_nnz) const override
        {
            const auto n = row_nnz.size();
            if (n == 0) return 0;
            std::sort(row_nnz.begin(), row_nnz.end());
            if (percent_ < 1) return row_nnz[static_cast<size_type>(n * percent_)];
            return row_nnz[n - 1];
        }
        double get_percentage() const { return percent_; }

    private:
        double percent_;
    };
    // imbalance_limit bounded by floor(n * ratio)
    class imbalance_bounded_limit : public strategy_type {
    public:
        imbalance_bounded_limit(double percent = 0.8, double ratio = 0.0001)
            : inner_{percent}, ratio_{ratio}
        {}
        std::string get_name() const override { return "imbalance_bounded_limit"; }
        size_type compute_ell_num_stored_elements_per_row(
            std::vector<size_type>& row_nnz) const override
        {
            const auto n = row_nnz.size();
            const auto k = inner_.compute_ell_num_stored_elements_per_row(row_nnz);
            return std::min(k, static_cast<size_type>(n * ratio_));
        }
        double get_percentage() const { return inner_.get_percentage(); }
        double get_ratio() const { return ratio_; }

    private:
        imbalance_limit inner_;
        double ratio_;
    };
    // imbalance_limit at sizeof(I) / (sizeof(V) + 2 sizeof(I))
    class minimal_storage_limit : public strategy_type {
    public:
        minimal_storage_limit()
            : inner_{static_cast<double>(sizeof(IndexType)) /
                     static_cast<double>(sizeof(ValueType) + 2 * sizeof(IndexType))}
        {}
        std::string get_name() const override { return "minimal_storage_limit"; }
        size_type compute_ell_num_stored_elements_per_row(
            std::vector<size_type>& row_nnz) const override
        {
            return inner_.compute_ell_num_stored_elements_per_row(row_nnz);
        }
        double get_percentage() const { return inner_.get_percentage(); }

    private:
        imbalance_limit inner_;
    };
    // imbalance_bounded_limit(1/3, 0.001)
    class automatic : public strategy_type {
    public:
        automatic() : inner_{1.0 / 3.0, 0.001} {}
        std::string get_name() const override { return "automatic"; }
        size_type compute_ell_num_stored_elements_per_row(
            std::vector<size_type>& row_nnz) const override
        {
            return inner_.compute_ell_num_stored_elements_per_row(row_nnz);
        }

    private:
        imbalance_bounded_limit inner_;
    };

    // Csr::convert_to(Hybrid) (reference/matrix/csr_kernels.cpp:911-957)
    static std::unique_ptr<Hybrid> create_from_csr(const Csr<ValueType, IndexType>* csr,
                                                   std::shared_ptr<strategy_type> strategy =
                                                       nullptr);
    const Ell<ValueType, IndexType>* get_ell() const { return ell_.get(); }
    const Coo<ValueType, IndexType>* get_coo() const { return coo_.get(); }
    size_type get_ell_num_stored_elements_per_row() const
    {
        return ell_->get_num_stored_elements_per_row();
    }
    size_type get_ell_stride() const { return ell_->get_stride(); }
    size_type get_coo_num_stored_elements() const { return coo_->get_num_stored_elements(); }
    std::shared_ptr<strategy_type> get_strategy() const { return strategy_; }

protected:
    Hybrid(std::shared_ptr<const Executor> exec, dim<2> size) : LinOp(exec, size) {}
    void apply_impl(const LinOp* b, LinOp* x) const override;
    void apply_impl(const LinOp* alpha, const LinOp* b, const LinOp* beta,
                    LinOp* x) const override;

private:
    std::unique_ptr<Ell<ValueType, IndexType>> ell_;
    std::unique_ptr<Coo<ValueType, IndexType>> coo_;
    std::shared_ptr<strategy_type> strategy_;
};

}  // namespace matrix
}  // namespace gko
