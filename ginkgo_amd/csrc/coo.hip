// Coo on gfx950: index conversions and the accumulating product.
//
// A matrix::Coo (include/ginkgo/core/matrix/coo.hpp) stores one row index per
// entry.  The products here do not read those: they run the CSR kernels
// (sub-wave, load_balance or merge_path) over the row-pointer index of the
// same entries, built once per matrix by gkomi_idxs_to_ptrs.  That reads
// 4 (n + 1) bytes of row pointers instead of 4 nnz bytes of row indices per
// product, and every row's sum is formed in one place, so the result is
// deterministic.  Rows must be sorted (the order within a row is free).
#include "common.hpp"
#include "spmv_common.hpp"

namespace gkomi {

namespace {

// ptrs[r] = number of entries whose row is below r (r = 0..n): a lower
// bound in the sorted row indices (components::convert_idxs_to_ptrs,
// core/components/format_conversion_kernels.hpp).
__global__ void __launch_bounds__(kBlock)
    idxs_to_ptrs_kernel(int64_t nnz, const int32_t* __restrict__ idxs, int64_t n,
                        int32_t* __restrict__ ptrs)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; r <= n;
         r += stride) {
        int64_t lo = 0, hi = nnz;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (idxs[mid] < r) {
                lo = mid + 1;
            } else {
                hi = mid;
            }
        }
        ptrs[r] = static_cast<int32_t>(lo);
    }
}

// idxs[e] = the row r with ptrs[r] <= e < ptrs[r + 1]: the last row start at
// or before e (components::convert_ptrs_to_idxs).
__global__ void __launch_bounds__(kBlock)
    ptrs_to_idxs_kernel(int64_t n, const int32_t* __restrict__ ptrs, int64_t nnz,
                        int32_t* __restrict__ idxs)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; e < nnz;
         e += stride) {
        int64_t lo = 0, hi = n - 1;  // answer in [0, n - 1]
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (ptrs[mid] <= e) {
                lo = mid;
            } else {
                hi = mid - 1;
            }
        }
        idxs[e] = static_cast<int32_t>(lo);
    }
}

template <typename T>
int coo_spmv2_impl(const gkomi_ctx* ctx, const T* alpha, const gkomi_csr* a, int64_t nrhs,
                   const T* b, int64_t bs, T* c, int64_t cs)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_csr(a)) return rc;
    GKOMI_REQUIRE(nrhs >= 0, GKOMI_INVALID_ARGUMENT, "coo spmv2: negative number of columns");
    if (a->n_rows == 0 || nrhs == 0 || a->nnz == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(b != nullptr && c != nullptr, GKOMI_INVALID_ARGUMENT,
                  "coo spmv2: null vector");
    GKOMI_REQUIRE(bs >= nrhs && cs >= nrhs, GKOMI_INVALID_ARGUMENT,
                  "coo spmv2: stride smaller than the number of columns");
    return csr_spmv_launch<T, kAccumulate>(stream_of(ctx), a, nrhs, b, bs, c, cs, alpha,
                                           nullptr);
}

}  // namespace

}  // namespace gkomi

using namespace gkomi;

extern "C" {

int gkomi_idxs_to_ptrs(const gkomi_ctx* ctx, int64_t nnz, const int32_t* idxs, int64_t n,
                       int32_t* ptrs)
{
    if (int rc = enter(ctx)) return rc;
    GKOMI_REQUIRE(nnz >= 0 && n >= 0, GKOMI_INVALID_ARGUMENT, "idxs_to_ptrs: negative size");
    GKOMI_REQUIRE(ptrs != nullptr && (nnz == 0 || idxs != nullptr), GKOMI_INVALID_ARGUMENT,
                  "idxs_to_ptrs: null pointer");
    idxs_to_ptrs_kernel<<<grid_for(n + 1, kBlock), kBlock, 0, stream_of(ctx)>>>(nnz, idxs, n,
                                                                                ptrs);
    GKOMI_CHECK_LAUNCH("idxs_to_ptrs_kernel");
    return GKOMI_SUCCESS;
}

int gkomi_ptrs_to_idxs(const gkomi_ctx* ctx, int64_t n, const int32_t* ptrs, int64_t nnz,
                       int32_t* idxs)
{
    if (int rc = enter(ctx)) return rc;
    GKOMI_REQUIRE(nnz >= 0 && n >= 0, GKOMI_INVALID_ARGUMENT, "ptrs_to_idxs: negative size");
    if (nnz == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(n > 0 && ptrs != nullptr && idxs != nullptr, GKOMI_INVALID_ARGUMENT,
                  "ptrs_to_idxs: entries without rows or null pointer");
    ptrs_to_idxs_kernel<<<grid_for(nnz, kBlock), kBlock, 0, stream_of(ctx)>>>(n, ptrs, nnz,
                                                                              idxs);
    GKOMI_CHECK_LAUNCH("ptrs_to_idxs_kernel");
    return GKOMI_SUCCESS;
}

int gkomi_coo_spmv2_f64(const gkomi_ctx* ctx, const double* alpha, const gkomi_csr* a,
                        int64_t nrhs, const double* b, int64_t b_stride, double* c,
                        int64_t c_stride)
{
    return coo_spmv2_impl(ctx, alpha, a, nrhs, b, b_stride, c, c_stride);
}

int gkomi_coo_spmv2_f32(const gkomi_ctx* ctx, const float* alpha, const gkomi_csr* a,
                        int64_t nrhs, const float* b, int64_t b_stride, float* c,
                        int64_t c_stride)
{
    return coo_spmv2_impl(ctx, alpha, a, nrhs, b, b_stride, c, c_stride);
}

}  // extern "C"
