// MatrixMarket reader / writer for the host library.
#include "ginkgo_amd/mtx_io.hpp"

#include <algorithm>
#include <cctype>
#include <istream>
#include <ostream>
#include <sstream>

namespace gko {

namespace {

std::string lower(std::string s)
{
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return std::tolower(c); });
    return s;
}

}  // namespace

template <typename V, typename I>
matrix_data<V, I> read_raw(std::istream& is)
{
    std::string line;
    if (!std::getline(is, line)) throw Error(__FILE__, __LINE__, "mtx: empty stream");
    std::istringstream hdr(line);
    std::string banner, object, format, field, symmetry;
    hdr >> banner >> object >> format >> field >> symmetry;
    if (banner != "%%MatrixMarket" || lower(object) != "matrix") {
        throw Error(__FILE__, __LINE__, "mtx: missing %%MatrixMarket matrix header");
    }
    format = lower(format);
    field = lower(field);
    symmetry = lower(symmetry);
    if (field == "complex") GKO_NOT_SUPPORTED_MSG("mtx: complex matrices are not supported");
    const bool pattern = field == "pattern";
    const bool symmetric = symmetry == "symmetric" || symmetry == "hermitian";
    const bool skew = symmetry == "skew-symmetric";
    while (std::getline(is, line)) {
        if (!line.empty() && line[0] != '%') break;
    }
    std::istringstream sz(line);
    matrix_data<V, I> data;
    size_type rows = 0, cols = 0, nnz = 0;
    if (format == "coordinate") {
        sz >> rows >> cols >> nnz;
        data.size = dim<2>{rows, cols};
        data.nonzeros.reserve(symmetric || skew ? 2 * nnz : nnz);
        for (size_type k = 0; k < nnz; ++k) {
            int64 r = 0, c = 0;
            double v = 1.0;
            if (!(is >> r >> c)) throw Error(__FILE__, __LINE__, "mtx: truncated entries");
            if (!pattern && !(is >> v)) throw Error(__FILE__, __LINE__, "mtx: missing value");
            data.nonzeros.push_back({static_cast<I>(r - 1), static_cast<I>(c - 1),
                                     static_cast<V>(v)});
            if ((symmetric || skew) && r != c) {
                data.nonzeros.push_back({static_cast<I>(c - 1), static_cast<I>(r - 1),
                                         static_cast<V>(skew ? -v : v)});
            }
        }
    } else if (format == "array") {
        sz >> rows >> cols;
        data.size = dim<2>{rows, cols};
        for (size_type c = 0; c < cols; ++c) {
            for (size_type r = 0; r < rows; ++r) {
                double v;
                if (!(is >> v)) throw Error(__FILE__, __LINE__, "mtx: truncated array");
                if (v != 0.0) data.nonzeros.push_back({static_cast<I>(r), static_cast<I>(c),
                                                       static_cast<V>(v)});
            }
        }
    } else {
        GKO_NOT_SUPPORTED_MSG("mtx: unknown storage format");
    }
    data.sort_row_major();
    return data;
}

template <typename V, typename I>
void write_raw(std::ostream& os, const matrix_data<V, I>& data)
{
    os << "%%MatrixMarket matrix coordinate real general\n";
    os << data.size.rows << ' ' << data.size.cols << ' ' << data.nonzeros.size() << '\n';
    os.precision(17);
    for (const auto& nz : data.nonzeros) {
        os << nz.row + 1 << ' ' << nz.column + 1 << ' ' << static_cast<double>(nz.value) << '\n';
    }
}

template <typename V>
std::unique_ptr<matrix::Csr<V, int32>> read_csr(
    std::istream& is, std::shared_ptr<const Executor> exec,
    std::shared_ptr<typename matrix::Csr<V, int32>::strategy_type> strategy)
{
    return matrix::Csr<V, int32>::create_from_data(exec, read_raw<V, int32>(is),
                                                   std::move(strategy));
}

template matrix_data<double, int32> read_raw<double, int32>(std::istream&);
template matrix_data<float, int32> read_raw<float, int32>(std::istream&);
template void write_raw<double, int32>(std::ostream&, const matrix_data<double, int32>&);
template void write_raw<float, int32>(std::ostream&, const matrix_data<float, int32>&);
template std::unique_ptr<matrix::Csr<double, int32>> read_csr<double>(
    std::istream&, std::shared_ptr<const Executor>,
    std::shared_ptr<matrix::Csr<double, int32>::strategy_type>);
template std::unique_ptr<matrix::Csr<float, int32>> read_csr<float>(
    std::istream&, std::shared_ptr<const Executor>,
    std::shared_ptr<matrix::Csr<float, int32>::strategy_type>);

}  // namespace gko
