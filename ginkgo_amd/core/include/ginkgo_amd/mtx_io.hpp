// MatrixMarket input/output (include/ginkgo/core/base/mtx_io.hpp:160,
// core/base/mtx_io.cpp): coordinate format, real / integer / pattern fields,
// general / symmetric / skew-symmetric symmetry; array format for dense.
#pragma once

#include <iosfwd>

#include "ginkgo_amd/matrix.hpp"

namespace gko {

// read_raw (mtx_io.hpp): the triplets of a MatrixMarket stream
template <typename ValueType = double, typename IndexType = int32>
matrix_data<ValueType, IndexType> read_raw(std::istream& is);

// write_raw: coordinate real general
template <typename ValueType = double, typename IndexType = int32>
void write_raw(std::ostream& os, const matrix_data<ValueType, IndexType>& data);

// gko::read<Csr<V>>(stream, exec)
template <typename ValueType = double>
std::unique_ptr<matrix::Csr<ValueType, int32>> read_csr(
    std::istream& is, std::shared_ptr<const Executor> exec,
    std::shared_ptr<typename matrix::Csr<ValueType, int32>::strategy_type> strategy = nullptr);

}  // namespace gko
