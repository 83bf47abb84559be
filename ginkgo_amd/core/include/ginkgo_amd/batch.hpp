// The batch module: many small independent systems solved together.
//
// Mirrors the reference's batch surface:
//   batch::MultiVector          include/ginkgo/core/base/batch_multi_vector.hpp
//   batch::matrix::Csr          include/ginkgo/core/matrix/batch_csr.hpp (one shared pattern)
//   batch::matrix::Ell          include/ginkgo/core/matrix/batch_ell.hpp (shared column indices)
//   batch::matrix::Dense        include/ginkgo/core/matrix/batch_dense.hpp
//   BatchLinOp::apply of the matrices (simple and advanced)
//   batch::stop::tolerance_type include/ginkgo/core/stop/batch_stop_enum.hpp
//   batch::preconditioner::Jacobi (scalar) include/ginkgo/core/preconditioner/batch_jacobi.hpp
//   batch::solver::Bicgstab     include/ginkgo/core/solver/batch_bicgstab.hpp,
//   batch::solver::Cg           include/ginkgo/core/solver/batch_cg.hpp,
//                               both with batch_solver_base.hpp:161-178
//                               (defaults: 100 iterations, tolerance 1e-11,
//                               absolute)
//   batch::log::BatchConvergence include/ginkgo/core/log/batch_logger.hpp
// plus one extension with no reference counterpart: batch::matrix::BlockDiagCsr,
// systems of different sizes in one block-diagonal CSR (the reference's batch
// formats require one common size), solved in size-ordered bins with an
// autotuned launch per bin.
#pragma once

#include <memory>
#include <vector>

#include "ginkgo_amd/base.hpp"

namespace gko {

// include/ginkgo/core/base/batch_dim.hpp
template <int Dim = 2>
struct batch_dim {
    size_type num_batch_items{};
    dim<2> common_size{};
    size_type get_num_batch_items() const { return num_batch_items; }
    dim<2> get_common_size() const { return common_size; }
};

namespace batch {

// Base of the batched operators (include/ginkgo/core/base/batch_lin_op.hpp).
class BatchLinOp {
public:
    virtual ~BatchLinOp() = default;
    std::shared_ptr<const Executor> get_executor() const { return exec_; }
    size_type get_num_batch_items() const { return size_.num_batch_items; }
    dim<2> get_common_size() const { return size_.common_size; }
    batch_dim<2> get_size() const { return size_; }

protected:
    BatchLinOp(std::shared_ptr<const Executor> exec, batch_dim<2> size)
        : exec_{std::move(exec)}, size_{size}
    {}
    std::shared_ptr<const Executor> exec_;
    batch_dim<2> size_;
};

// One column per item.  Uniform items hold n rows each (item k at
// values[k * n]); the uneven form holds item k in rows
// [row_offsets[k], row_offsets[k + 1]).
template <typename ValueType>
class MultiVector : public BatchLinOp {
public:
    using value_type = ValueType;
    static std::unique_ptr<MultiVector> create(std::shared_ptr<const Executor> exec,
                                               batch_dim<2> size);
    static std::unique_ptr<MultiVector> create_from_host(std::shared_ptr<const Executor> exec,
                                                         size_type num_items, size_type n,
                                                         const std::vector<ValueType>& values);
    // the layout of an uneven batch (item k has row_offsets[k+1] - row_offsets[k] rows)
    static std::unique_ptr<MultiVector> create_uneven(std::shared_ptr<const Executor> exec,
                                                      const std::vector<int64>& row_offsets,
                                                      const std::vector<ValueType>* values =
                                                          nullptr);
    ValueType* get_values() { return values_.get_data(); }
    const ValueType* get_const_values() const { return values_.get_const_data(); }
    size_type get_total_rows() const { return values_.get_size(); }
    std::vector<ValueType> to_host() const { return values_.to_host(); }
    void fill(ValueType v) { values_.fill(v); }

private:
    MultiVector(std::shared_ptr<const Executor> exec, batch_dim<2> size, size_type total)
        : BatchLinOp(exec, size), values_(exec, total)
    {}
    array<ValueType> values_;
};

namespace stop {
// include/ginkgo/core/stop/batch_stop_enum.hpp
enum class tolerance_type { absolute, relative };
}  // namespace stop

namespace matrix {

// What the batched matrices share: the descriptor the kernels read, the
// per-item sizes the solvers plan with, and BatchLinOp::apply
// (include/ginkgo/core/base/batch_lin_op.hpp) over gkomi_batch_spmv_*.
template <typename ValueType>
class BatchMatrix : public BatchLinOp {
public:
    using value_type = ValueType;
    virtual gkomi_batch_csr descriptor() const = 0;
    // rows and stored values of each item
    virtual std::vector<int64> get_item_rows() const
    {
        return std::vector<int64>(size_.num_batch_items,
                                  static_cast<int64>(size_.common_size.rows));
    }
    virtual std::vector<int64> get_item_nnz() const
    {
        return std::vector<int64>(size_.num_batch_items, descriptor().uniform_nnz);
    }
    // longest row (stored entries) of each item, 0 where not known
    virtual std::vector<int64> get_item_max_row_nnz() const
    {
        return std::vector<int64>(size_.num_batch_items, 0);
    }
    virtual size_type get_total_rows() const
    {
        return size_.num_batch_items * size_.common_size.rows;
    }
    virtual size_type get_total_cols() const
    {
        return size_.num_batch_items * size_.common_size.cols;
    }
    // x_k = A_k b_k for every item k
    void apply(const MultiVector<ValueType>* b, MultiVector<ValueType>* x) const;
    // x_k = alpha_k A_k b_k + beta_k x_k; alpha and beta hold one value per item
    void apply(const MultiVector<ValueType>* alpha, const MultiVector<ValueType>* b,
               const MultiVector<ValueType>* beta, MultiVector<ValueType>* x) const;

protected:
    using BatchLinOp::BatchLinOp;

private:
    void check_apply(const MultiVector<ValueType>* b, const MultiVector<ValueType>* x) const;
};

// Reference batch::matrix::Csr: every item has the same n x n pattern
// (row_ptrs, col_idxs shared), values stored item after item.
template <typename ValueType, typename IndexType = int32>
class Csr : public BatchMatrix<ValueType> {
public:
    using value_type = ValueType;
    using index_type = IndexType;
    static std::unique_ptr<Csr> create_from_host(std::shared_ptr<const Executor> exec,
                                                 size_type num_items, size_type n,
                                                 const std::vector<IndexType>& row_ptrs,
                                                 const std::vector<IndexType>& col_idxs,
                                                 const std::vector<ValueType>& values);
    // items of size.rows x size.cols (include/ginkgo/core/matrix/batch_csr.hpp:
    // batch_dim<2> of any shape); the column indices must be below size.cols
    static std::unique_ptr<Csr> create_from_host(std::shared_ptr<const Executor> exec,
                                                 size_type num_items, dim<2> size,
                                                 const std::vector<IndexType>& row_ptrs,
                                                 const std::vector<IndexType>& col_idxs,
                                                 const std::vector<ValueType>& values);
    size_type get_num_elements_per_item() const { return col_idxs_.get_size(); }
    const IndexType* get_const_row_ptrs() const { return row_ptrs_.get_const_data(); }
    const IndexType* get_const_col_idxs() const { return col_idxs_.get_const_data(); }
    const ValueType* get_const_values() const { return values_.get_const_data(); }
    ValueType* get_values() { return values_.get_data(); }
    gkomi_batch_csr descriptor() const override;

private:
    Csr(std::shared_ptr<const Executor> exec, batch_dim<2> size)
        : BatchMatrix<ValueType>(exec, size), row_ptrs_(exec), col_idxs_(exec), values_(exec)
    {}
    array<IndexType> row_ptrs_;
    array<IndexType> col_idxs_;
    array<ValueType> values_;
};

// Reference batch::matrix::Ell: the column indices of one item,
// column-major (slot j of row i at i + j * rows), are shared by all items;
// -1 marks the slots after a row's last entry.  The values of item k use
// the same layout and start at k * rows * num_stored_elements_per_row.
template <typename ValueType, typename IndexType = int32>
class Ell : public BatchMatrix<ValueType> {
public:
    using value_type = ValueType;
    using index_type = IndexType;
    static std::unique_ptr<Ell> create_from_host(std::shared_ptr<const Executor> exec,
                                                 size_type num_items, dim<2> size,
                                                 size_type num_elems_per_row,
                                                 const std::vector<IndexType>& col_idxs,
                                                 const std::vector<ValueType>& values);
    // From row-major dense items (num_items * rows * cols values).  The
    // stored pattern is the union of the items' nonzeros, each row's columns
    // ascending; an item without one of those entries stores a zero there.
    static std::unique_ptr<Ell> create_from_dense_host(std::shared_ptr<const Executor> exec,
                                                       size_type num_items, dim<2> size,
                                                       const std::vector<ValueType>& dense);
    size_type get_num_stored_elements_per_row() const { return width_; }
    const IndexType* get_const_col_idxs() const { return col_idxs_.get_const_data(); }
    const ValueType* get_const_values() const { return values_.get_const_data(); }
    std::vector<IndexType> col_idxs_to_host() const { return col_idxs_.to_host(); }
    std::vector<ValueType> values_to_host() const { return values_.to_host(); }
    gkomi_batch_csr descriptor() const override;

private:
    Ell(std::shared_ptr<const Executor> exec, batch_dim<2> size)
        : BatchMatrix<ValueType>(exec, size), col_idxs_(exec), values_(exec)
    {}
    size_type width_{};
    array<IndexType> col_idxs_;
    array<ValueType> values_;
};

// Reference batch::matrix::Dense: item k's values, row-major, start at
// k * rows * cols.
template <typename ValueType>
class Dense : public BatchMatrix<ValueType> {
public:
    using value_type = ValueType;
    static std::unique_ptr<Dense> create_from_host(std::shared_ptr<const Executor> exec,
                                                   size_type num_items, dim<2> size,
                                                   const std::vector<ValueType>& values);
    const ValueType* get_const_values() const { return values_.get_const_data(); }
    std::vector<ValueType> values_to_host() const { return values_.to_host(); }
    gkomi_batch_csr descriptor() const override;

private:
    Dense(std::shared_ptr<const Executor> exec, batch_dim<2> size)
        : BatchMatrix<ValueType>(exec, size), values_(exec)
    {}
    array<ValueType> values_;
};

// Extension (no reference counterpart): items of different sizes.  Item k
// owns global rows [row_offsets[k], row_offsets[k+1]); row_ptrs are global
// (size total_rows + 1), column indices local to the item.
template <typename ValueType, typename IndexType = int32>
class BlockDiagCsr : public BatchMatrix<ValueType> {
public:
    using value_type = ValueType;
    using index_type = IndexType;
    static std::unique_ptr<BlockDiagCsr> create_from_host(
        std::shared_ptr<const Executor> exec, const std::vector<int64>& row_offsets,
        const std::vector<IndexType>& row_ptrs, const std::vector<IndexType>& col_idxs,
        const std::vector<ValueType>& values);
    std::vector<int64> get_item_rows() const override { return item_rows_; }
    std::vector<int64> get_item_nnz() const override { return item_nnz_; }
    std::vector<int64> get_item_max_row_nnz() const override { return item_max_row_nnz_; }
    size_type get_total_rows() const override
    {
        return row_offsets_host_.empty() ? 0 : static_cast<size_type>(row_offsets_host_.back());
    }
    size_type get_total_cols() const override { return get_total_rows(); }
    const std::vector<int64>& get_row_offsets() const { return row_offsets_host_; }
    gkomi_batch_csr descriptor() const override;

private:
    BlockDiagCsr(std::shared_ptr<const Executor> exec, batch_dim<2> size)
        : BatchMatrix<ValueType>(exec, size),
          row_offsets_(exec),
          row_ptrs_(exec),
          col_idxs_(exec),
          values_(exec)
    {}
    array<int64> row_offsets_;
    array<IndexType> row_ptrs_;
    array<IndexType> col_idxs_;
    array<ValueType> values_;
    std::vector<int64> row_offsets_host_, item_rows_, item_nnz_, item_max_row_nnz_;
};

// Extension (no reference counterpart; the reference's batch Ell and Dense
// have one size): Ell and Dense items of different sizes.  Item k owns rows
// [row_offsets[k], row_offsets[k+1]) of the vectors (n_k rows, square) and
// its stored entries from item_offsets[k] on.  BlockDiagEll: n_k * w_k slots
// column-major (slot j of row i at i + j * n_k), w_k the item's longest row,
// the item's own local column indices at the same positions, -1 after a
// row's last entry (the reference's Ell layout per item).  BlockDiagDense:
// n_k * n_k values, row-major.
template <typename ValueType>
class BlockDiagBase : public BatchMatrix<ValueType> {
public:
    std::vector<int64> get_item_rows() const override { return item_rows_; }
    std::vector<int64> get_item_nnz() const override { return item_nnz_; }
    size_type get_total_rows() const override
    {
        return row_offsets_host_.empty() ? 0 : static_cast<size_type>(row_offsets_host_.back());
    }
    size_type get_total_cols() const override { return get_total_rows(); }
    const std::vector<int64>& get_row_offsets() const { return row_offsets_host_; }
    const std::vector<int64>& get_item_offsets() const { return item_offsets_host_; }
    std::vector<ValueType> values_to_host() const { return values_.to_host(); }

protected:
    BlockDiagBase(std::shared_ptr<const Executor> exec, const std::vector<int64>& row_offsets,
                  const std::vector<int64>& item_offsets, const std::vector<ValueType>& values);
    gkomi_batch_csr base_descriptor(int32_t format) const;
    array<int64> row_offsets_;
    array<int64> item_offsets_;
    array<ValueType> values_;
    std::vector<int64> row_offsets_host_, item_offsets_host_, item_rows_, item_nnz_;
};

template <typename ValueType, typename IndexType = int32>
class BlockDiagEll : public BlockDiagBase<ValueType> {
public:
    using value_type = ValueType;
    using index_type = IndexType;
    // from a block-diagonal CSR on the host (global row pointers, local
    // column indices); each row keeps the order of its entries
    static std::unique_ptr<BlockDiagEll> create_from_block_diag_csr_host(
        std::shared_ptr<const Executor> exec, const std::vector<int64>& row_offsets,
        const std::vector<IndexType>& row_ptrs, const std::vector<IndexType>& col_idxs,
        const std::vector<ValueType>& values);
    std::vector<int64> get_item_max_row_nnz() const override { return item_width_; }
    const std::vector<int64>& get_item_widths() const { return item_width_; }
    std::vector<IndexType> col_idxs_to_host() const { return col_idxs_.to_host(); }
    gkomi_batch_csr descriptor() const override;

private:
    BlockDiagEll(std::shared_ptr<const Executor> exec, const std::vector<int64>& row_offsets,
                 const std::vector<int64>& item_offsets, const std::vector<IndexType>& col_idxs,
                 const std::vector<ValueType>& values, std::vector<int64> widths);
    array<IndexType> col_idxs_;
    std::vector<int64> item_width_;
};

template <typename ValueType>
class BlockDiagDense : public BlockDiagBase<ValueType> {
public:
    using value_type = ValueType;
    // item k's n_k x n_k values, row-major, one item after another
    static std::unique_ptr<BlockDiagDense> create_from_host(
        std::shared_ptr<const Executor> exec, const std::vector<int64>& row_offsets,
        const std::vector<ValueType>& values);
    // from a block-diagonal CSR on the host (duplicate entries are summed)
    template <typename IndexType>
    static std::unique_ptr<BlockDiagDense> create_from_block_diag_csr_host(
        std::shared_ptr<const Executor> exec, const std::vector<int64>& row_offsets,
        const std::vector<IndexType>& row_ptrs, const std::vector<IndexType>& col_idxs,
        const std::vector<ValueType>& values);
    gkomi_batch_csr descriptor() const override;

private:
    using BlockDiagBase<ValueType>::BlockDiagBase;
};

}  // namespace matrix

namespace preconditioner {

// Scalar batch Jacobi (reference/preconditioner/batch_scalar_jacobi.hpp):
// max_block_size must be 1.
template <typename ValueType>
class Jacobi {
public:
    struct parameters_type {
        unsigned max_block_size{1};
        parameters_type& with_max_block_size(unsigned b)
        {
            max_block_size = b;
            return *this;
        }
        std::shared_ptr<const parameters_type> on(std::shared_ptr<const Executor>) const
        {
            if (max_block_size != 1) {
                throw NotSupported(__FILE__, __LINE__,
                                   "batch::preconditioner::Jacobi: only the scalar "
                                   "(max_block_size = 1) batch Jacobi");
            }
            return std::make_shared<parameters_type>(*this);
        }
    };
    static parameters_type build() { return {}; }
};

}  // namespace preconditioner

namespace log {

// Per-item iteration counts and final residual norms of the last apply
// (include/ginkgo/core/log/batch_logger.hpp).
template <typename ValueType>
class BatchConvergence {
public:
    static std::shared_ptr<BatchConvergence> create()
    {
        return std::shared_ptr<BatchConvergence>(new BatchConvergence());
    }
    const std::vector<int32>& get_num_iterations() const { return iterations_; }
    const std::vector<ValueType>& get_residual_norm() const { return residual_norm_; }
    void record(std::vector<int32> it, std::vector<ValueType> res) const
    {
        iterations_ = std::move(it);
        residual_norm_ = std::move(res);
    }

private:
    BatchConvergence() = default;
    mutable std::vector<int32> iterations_;
    mutable std::vector<ValueType> residual_norm_;
};

}  // namespace log

namespace solver {

// The iterative batch solvers share their parameters, factory, plan and
// launch machinery; they differ in the per-system iteration.
enum class batch_method { bicgstab, cg };

// batch::solver::Bicgstab (include/ginkgo/core/solver/batch_bicgstab.hpp):
// an independent BiCGSTAB per item, no breakdown guards, the half-step exit
// (reference/solver/batch_bicgstab_kernels.hpp:180-354).
// batch::solver::Cg (include/ginkgo/core/solver/batch_cg.hpp): an
// independent preconditioned CG per item, stopped on the implicit norm
// sqrt|r.z| (reference/solver/batch_cg_kernels.hpp:111-221).
template <typename ValueType, batch_method Method>
class BatchKrylov {
public:
    using value_type = ValueType;
    using precond_params = typename preconditioner::Jacobi<ValueType>::parameters_type;

    class Factory;
    struct parameters_type {
        int max_iterations{100};
        double tolerance{1e-11};
        stop::tolerance_type tolerance_type{stop::tolerance_type::absolute};
        std::shared_ptr<const precond_params> preconditioner;
        parameters_type& with_max_iterations(int m)
        {
            max_iterations = m;
            return *this;
        }
        parameters_type& with_tolerance(double t)
        {
            tolerance = t;
            return *this;
        }
        parameters_type& with_tolerance_type(stop::tolerance_type t)
        {
            tolerance_type = t;
            return *this;
        }
        parameters_type& with_preconditioner(const precond_params& p)
        {
            preconditioner = p.on(nullptr);
            return *this;
        }
        std::shared_ptr<Factory> on(std::shared_ptr<const Executor> exec) const
        {
            return std::shared_ptr<Factory>(new Factory(std::move(exec), *this));
        }
    };

    class Factory {
    public:
        std::unique_ptr<BatchKrylov> generate(std::shared_ptr<const BatchLinOp> a) const;
        const parameters_type& get_parameters() const { return params_; }

    private:
        friend struct parameters_type;
        Factory(std::shared_ptr<const Executor> exec, parameters_type p)
            : exec_{std::move(exec)}, params_{std::move(p)}
        {}
        std::shared_ptr<const Executor> exec_;
        parameters_type params_;
    };

    static parameters_type build() { return {}; }

    // x holds the initial guess and receives the solution
    void apply(const MultiVector<ValueType>* b, MultiVector<ValueType>* x) const;
    // apply(b, x) recorded as one HIP graph (extension, no reference
    // counterpart): with zero_initial_guess, x is first set to zero by a fill
    // kernel inside the graph; otherwise every replay starts from x as it
    // is.  b, x and this solver must outlive the graph.  Replays write the
    // per-item iterations and residual norms but call no logger;
    // record_logs() hands the last ones to the loggers.
    std::unique_ptr<Graph> capture(const MultiVector<ValueType>* b, MultiVector<ValueType>* x,
                                   bool zero_initial_guess = true) const;
    void record_logs() const;
    // times the launch candidates of every size bin on (b, x) and keeps the
    // fastest (x is not modified); extension, no reference counterpart
    void tune(const MultiVector<ValueType>* b, const MultiVector<ValueType>* x);
    void add_logger(std::shared_ptr<const log::BatchConvergence<ValueType>> l)
    {
        loggers_.push_back(std::move(l));
    }
    // the plan in effect: workgroup size and LDS residency per size bin
    std::vector<int> get_block_sizes() const;
    std::vector<int> get_matrix_in_lds() const;
    const parameters_type& get_parameters() const { return params_; }

private:
    BatchKrylov(std::shared_ptr<const Executor> exec, parameters_type params,
                std::shared_ptr<const BatchLinOp> a);
    gkomi_batch_settings settings() const;
    void check_vectors(const MultiVector<ValueType>* b, const MultiVector<ValueType>* x) const;

    std::shared_ptr<const Executor> exec_;
    parameters_type params_;
    std::shared_ptr<const BatchLinOp> system_;
    gkomi_batch_csr desc_{};
    size_type total_rows_{};
    array<int32> order_;
    gkomi_batch_plan plan_{};
    mutable array<char> workspace_;
    mutable array<int32> iterations_;
    mutable array<ValueType> residuals_;
    std::vector<std::shared_ptr<const log::BatchConvergence<ValueType>>> loggers_;
};

template <typename ValueType>
using Bicgstab = BatchKrylov<ValueType, batch_method::bicgstab>;
template <typename ValueType>
using Cg = BatchKrylov<ValueType, batch_method::cg>;

}  // namespace solver
}  // namespace batch
}  // namespace gko
