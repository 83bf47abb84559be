// CSR products with mixed value types: matrix values, the input b and the
// output c each double or float.  The reference forms every product and sum
// in the highest precision of the three (reference/matrix/csr_kernels.cpp:
// 47-121: arithmetic_type = highest_precision<MatrixValueType,
// InputValueType, OutputValueType>) and rounds into c once; so do these
// kernels, converting on load.  One right-hand side runs the classical
// sub-wave product (two rows in flight per lane group, lanes per row from the
// mean row length), several a wave per row with lanes over the right-hand
// sides.
#include <algorithm>

#include "common.hpp"
#include "spmv_common.hpp"

namespace gkomi {

template <typename TA, typename TB, typename TC, int SW, int Mode>
__global__ void __launch_bounds__(kBlock)
    csr_spmv_mixed_kernel(int64_t n_rows, const int32_t* __restrict__ rp,
                          const int32_t* __restrict__ ci, const TA* __restrict__ v,
                          const TB* __restrict__ b, int64_t bs, TC* __restrict__ c, int64_t cs,
                          const TA* __restrict__ alpha, const TC* __restrict__ beta)
{
    using Acc = acc_t<TA, TB, TC>;
    constexpr int kRpg = 2;
    const int64_t nb = static_cast<int64_t>(gridDim.x);
    const int64_t tiles = ceildiv(n_rows, subwave_tile_rows<SW, kRpg>);
    for (int64_t t = xcd_block(blockIdx.x, gridDim.x); t < tiles; t += nb) {
        subwave_tile<Acc, SW, kRpg, false>(
            n_rows, t, rp, ci, nullptr, v,
            [&](int32_t col) { return static_cast<Acc>(b[static_cast<int64_t>(col) * bs]); },
            [&](int64_t row, Acc sum) { store_result<Mode>(c + row * cs, sum, alpha, beta); });
    }
}

template <typename TA, typename TB, typename TC, int Mode>
__global__ void __launch_bounds__(kBlock)
    csr_spmm_mixed_kernel(int64_t n_rows, int64_t nrhs, const int32_t* __restrict__ rp,
                          const int32_t* __restrict__ ci, const TA* __restrict__ v,
                          const TB* __restrict__ b, int64_t bs, TC* __restrict__ c, int64_t cs,
                          const TA* __restrict__ alpha, const TC* __restrict__ beta)
{
    using Acc = acc_t<TA, TB, TC>;
    const int lane = threadIdx.x % kWave;
    const int64_t wave_id =
        static_cast<int64_t>(xcd_block(blockIdx.x, gridDim.x)) * kWavesPerBlock +
        threadIdx.x / kWave;
    const int64_t num_waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
    for (int64_t row = wave_id; row < n_rows; row += num_waves) {
        const int32_t s = rp[row];
        const int32_t e = rp[row + 1];
        for (int64_t j0 = 0; j0 < nrhs; j0 += kWave) {
            const int64_t j = j0 + lane;
            if (j < nrhs) {
                Acc sum = 0;
                for (int32_t k = s; k < e; ++k) {
                    sum += static_cast<Acc>(v[k]) *
                           static_cast<Acc>(b[static_cast<int64_t>(ci[k]) * bs + j]);
                }
                store_result<Mode>(c + row * cs + j, sum, alpha, beta);
            }
        }
    }
}

template <typename TA, typename TB, typename TC, int Mode>
int csr_mixed_launch(hipStream_t stream, const gkomi_csr* a, int64_t nrhs, const TB* b,
                     int64_t bs, TC* c, int64_t cs, const TA* alpha, const TC* beta)
{
    const TA* v = static_cast<const TA*>(a->values);
    if (nrhs > 1) {
        const unsigned g = grid_for(a->n_rows, kWavesPerBlock, 32);
        csr_spmm_mixed_kernel<TA, TB, TC, Mode><<<g, kBlock, 0, stream>>>(
            a->n_rows, nrhs, a->row_ptrs, a->col_idxs, v, b, bs, c, cs, alpha, beta);
        GKOMI_CHECK_LAUNCH("csr_spmm_mixed_kernel");
        return GKOMI_SUCCESS;
    }
    const int sw = pick_subwave(a, sizeof(acc_t<TA, TB, TC>)).sw;
#define GKOMI_MIXED_CASE(S)                                                                    \
    case S: {                                                                                  \
        const unsigned g = grid_for(a->n_rows, subwave_tile_rows<S, 2>, 64);                   \
        csr_spmv_mixed_kernel<TA, TB, TC, S, Mode><<<g, kBlock, 0, stream>>>(                  \
            a->n_rows, a->row_ptrs, a->col_idxs, v, b, bs, c, cs, alpha, beta);                \
        break;                                                                                 \
    }
    switch (sw) {
        GKOMI_MIXED_CASE(1)
        GKOMI_MIXED_CASE(2)
        GKOMI_MIXED_CASE(4)
        GKOMI_MIXED_CASE(8)
        GKOMI_MIXED_CASE(16)
        GKOMI_MIXED_CASE(32)
        GKOMI_MIXED_CASE(64)
    default:
        return fail(GKOMI_INVALID_ARGUMENT, "CSR: subwave must be a power of two <= 64");
    }
#undef GKOMI_MIXED_CASE
    GKOMI_CHECK_LAUNCH("csr_spmv_mixed_kernel");
    return GKOMI_SUCCESS;
}

template <typename TA, typename TB, typename TC, int Mode>
int csr_mixed_impl(const gkomi_ctx* ctx, const TA* alpha, const gkomi_csr* a, int64_t nrhs,
                   const TB* b, int64_t bs, const TC* beta, TC* c, int64_t cs)
{
    if constexpr (std::is_same<TA, TB>::value && std::is_same<TA, TC>::value) {
        // one value type: the typed entry points with the matrix's strategy
        if constexpr (Mode == kAdvanced) {
            if constexpr (std::is_same<TA, double>::value) {
                return gkomi_csr_advanced_spmv_f64(ctx, alpha, a, nrhs, b, bs, beta, c, cs);
            } else {
                return gkomi_csr_advanced_spmv_f32(ctx, alpha, a, nrhs, b, bs, beta, c, cs);
            }
        } else {
            if constexpr (std::is_same<TA, double>::value) {
                return gkomi_csr_spmv_f64(ctx, a, nrhs, b, bs, c, cs);
            } else {
                return gkomi_csr_spmv_f32(ctx, a, nrhs, b, bs, c, cs);
            }
        }
    } else {
        if (int rc = enter(ctx)) return rc;
        if (int rc = check_csr(a)) return rc;
        GKOMI_REQUIRE(nrhs >= 0, GKOMI_INVALID_ARGUMENT, "spmv: negative number of columns");
        if (a->n_rows == 0 || nrhs == 0) return GKOMI_SUCCESS;
        GKOMI_REQUIRE(c != nullptr && (a->nnz == 0 || b != nullptr), GKOMI_INVALID_ARGUMENT,
                      "spmv: null vector");
        GKOMI_REQUIRE(bs >= nrhs && cs >= nrhs, GKOMI_INVALID_ARGUMENT,
                      "spmv: stride smaller than the number of columns");
        GKOMI_REQUIRE(Mode != kAdvanced || (alpha != nullptr && beta != nullptr),
                      GKOMI_INVALID_ARGUMENT, "advanced spmv: null alpha or beta");
        return csr_mixed_launch<TA, TB, TC, Mode>(stream_of(ctx), a, nrhs, b, bs, c, cs, alpha,
                                                  beta);
    }
}

}  // namespace gkomi

using namespace gkomi;

extern "C" {

int gkomi_csr_spmv_mixed(const gkomi_ctx* ctx, int32_t a_type, const gkomi_csr* a, int64_t nrhs,
                         int32_t b_type, const void* b, int64_t b_stride, int32_t c_type,
                         void* c, int64_t c_stride)
{
    return dispatch_value_types(a_type, b_type, c_type, [&](auto ta, auto tb, auto tc) {
        using TA = decltype(ta);
        using TB = decltype(tb);
        using TC = decltype(tc);
        return csr_mixed_impl<TA, TB, TC, kPlain>(ctx, nullptr, a, nrhs,
                                                  static_cast<const TB*>(b), b_stride, nullptr,
                                                  static_cast<TC*>(c), c_stride);
    });
}

int gkomi_csr_advanced_spmv_mixed(const gkomi_ctx* ctx, const void* alpha, int32_t a_type,
                                  const gkomi_csr* a, int64_t nrhs, int32_t b_type,
                                  const void* b, int64_t b_stride, const void* beta,
                                  int32_t c_type, void* c, int64_t c_stride)
{
    return dispatch_value_types(a_type, b_type, c_type, [&](auto ta, auto tb, auto tc) {
        using TA = decltype(ta);
        using TB = decltype(tb);
        using TC = decltype(tc);
        return csr_mixed_impl<TA, TB, TC, kAdvanced>(
            ctx, static_cast<const TA*>(alpha), a, nrhs, static_cast<const TB*>(b), b_stride,
            static_cast<const TC*>(beta), static_cast<TC*>(c), c_stride);
    });
}

}  // extern "C"
