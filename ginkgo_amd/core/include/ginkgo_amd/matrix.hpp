// Sparse formats: Csr (include/ginkgo/core/matrix/csr.hpp:104-1690, with the
// classical / load_balance / automatical strategies), Sellp (sellp.hpp) and
// Ell (ell.hpp), their conversions and the device transpose; plus
// matrix_data for assembling from host triplets.
#pragma once

#include <algorithm>
#include <tuple>

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
    class load_balance : public strategy_type {
    public:
        load_balance() = default;
        explicit load_balance(int64) {}
        std::string get_name() const override { return "load_balance"; }
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

    ValueType* get_values() noexcept { return values_.get_data(); }
    const ValueType* get_const_values() const noexcept { return values_.get_const_data(); }
    IndexType* get_col_idxs() noexcept { return col_idxs_.get_data(); }
    const IndexType* get_const_col_idxs() const noexcept { return col_idxs_.get_const_data(); }
    IndexType* get_row_ptrs() noexcept { return row_ptrs_.get_data(); }
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
    template <typename Other>
    std::unique_ptr<Csr<Other, IndexType>> convert_precision() const;
    std::unique_ptr<Csr> clone() const;
    // host copies of the three arrays
    std::vector<IndexType> row_ptrs_to_host() const { return row_ptrs_.to_host(); }
    std::vector<IndexType> col_idxs_to_host() const { return col_idxs_.to_host(); }
    std::vector<ValueType> values_to_host() const { return values_.to_host(); }

    // 2-byte column offsets (col - row), read instead of the 4-byte indices
    // by the classical single right-hand-side SpMV and the fused CG; built
    // only when every offset fits int16 (returns whether it did).  Results
    // are bit-identical either way.  No reference counterpart.  Values or
    // indices changed later through a view require calling it again.
    bool compact_indices();
    bool has_compact_indices() const noexcept { return compact_; }

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
    std::shared_ptr<strategy_type> strategy_;
    std::string effective_{"classical"};
    int subwave_{0};
    mutable int64 max_row_nnz_{-1};
    template <typename, typename>
    friend class Csr;
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
};

}  // namespace matrix
}  // namespace gko
