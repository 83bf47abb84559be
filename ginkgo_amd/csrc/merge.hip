// CSR SpMV by merge-path partitioning (Matrix::Csr::merge_path, include/
// ginkgo/core/matrix/csr.hpp:283-301), written for gfx950 after the
// published algorithm (D. Merrill, M. Garland, "Merge-based parallel sparse
// matrix-vector multiplication", SC16).
//
// The work of y = A x is the merge of two sorted lists: the n row ends
// rp[1..n] and the nnz entry indices 0..nnz-1 (a row end goes first on a
// tie: row i is finished once all its entries are consumed).  Every
// workgroup takes an equal share of n + nnz steps of that merge, so empty
// rows, short rows and rows longer than a workgroup share all balance.
//
//  1. Two binary searches give the workgroup's first and last (row, entry)
//     coordinates.  The workgroup stages its rows' ends and the products
//     a_ij * x_j of its entries in LDS with coalesced loads.
//  2. Each thread searches its own share of kMpItems steps in LDS and walks
//     it: consuming an entry adds its product, finishing a row emits the
//     sum.  A thread's second and later rows are complete and stored at
//     once; its first row continues rows from earlier threads, whose
//     unfinished tails (the carries) are added by an inclusive segmented
//     scan over the threads in LDS.
//  3. The workgroup's last unfinished row leaves as a carry (row, value) in
//     the strategy workspace; a fix-up kernel adds each run of equal carry
//     rows, in workgroup order, to its row after the main kernel.
// No atomics: the result is deterministic.  Rows split across threads or
// workgroups are summed in another grouping than the reference's sequential
// loop (within the FP bound of the SpMV tests).
#include "common.hpp"
#include "spmv_common.hpp"

namespace gkomi {

namespace {

constexpr int kMpThreads = 256;
constexpr int kMpItems = 8;
constexpr int kMpSegment = kMpThreads * kMpItems;

// Rows finished after `diag` steps of the merge of the row ends a[i] =
// row_end(i) (i < n_rows) with the entry indices 0..nnz-1.
template <typename RowEnd>
__device__ __forceinline__ int64_t merge_search(int64_t diag, int64_t n_rows, int64_t nnz,
                                                RowEnd row_end)
{
    int64_t lo = diag > nnz ? diag - nnz : 0;
    int64_t hi = diag < n_rows ? diag : n_rows;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (row_end(mid) <= diag - mid - 1) {
            lo = mid + 1;
        } else {
            hi = mid;
        }
    }
    return lo;
}

template <typename T, int Mode>
__global__ void __launch_bounds__(kMpThreads)
    csr_spmv_merge_kernel(int64_t n_rows, int64_t nnz, const int32_t* __restrict__ rp,
                          const int32_t* __restrict__ ci, const T* __restrict__ v,
                          const T* __restrict__ b, int64_t bs, T* __restrict__ c, int64_t cs,
                          const T* __restrict__ alpha, const T* __restrict__ beta,
                          int32_t* __restrict__ carry_row, T* __restrict__ carry_val)
{
    __shared__ T prod[kMpSegment];
    __shared__ int32_t lend[kMpSegment + 1];
    __shared__ int64_t bounds[4];
    __shared__ int32_t crow[kMpThreads];
    __shared__ T cval[kMpThreads];
    const int tid = threadIdx.x;
    const int64_t total = n_rows + nnz;
    const int64_t d0 = static_cast<int64_t>(blockIdx.x) * kMpSegment;
    const int64_t d1 = d0 + kMpSegment < total ? d0 + kMpSegment : total;
    auto global_end = [&](int64_t i) { return static_cast<int64_t>(rp[i + 1]); };
    if (tid < 2) {
        const int64_t d = tid == 0 ? d0 : d1;
        const int64_t r = merge_search(d, n_rows, nnz, global_end);
        bounds[2 * tid] = r;
        bounds[2 * tid + 1] = d - r;
    }
    __syncthreads();
    const int64_t row0 = bounds[0], nz0 = bounds[1];
    const int64_t row1 = bounds[2], nz1 = bounds[3];
    const int rows_local = static_cast<int>(row1 - row0);  // row ends inside the segment
    const int nnz_local = static_cast<int>(nz1 - nz0);
    // stage the ends of rows row0 .. row1 (row1's end only when it exists)
    // relative to nz0, and the products of the segment's entries
    const int ends = rows_local + (row1 < n_rows ? 1 : 0);
    for (int i = tid; i < ends; i += kMpThreads) {
        lend[i] = static_cast<int32_t>(rp[row0 + i + 1] - nz0);
    }
    for (int j = tid; j < nnz_local; j += kMpThreads) {
        const int64_t e = nz0 + j;
        prod[j] = v[e] * b[static_cast<int64_t>(ci[e]) * bs];
    }
    __syncthreads();
    // this thread's share of the segment
    const int total_local = rows_local + nnz_local;
    const int dl = tid * kMpItems < total_local ? tid * kMpItems : total_local;
    const int de = dl + kMpItems < total_local ? dl + kMpItems : total_local;
    int i = static_cast<int>(
        merge_search(dl, rows_local, nnz_local, [&](int64_t r) { return lend[r]; }));
    int j = dl - i;
    T sum = T(0);
    bool first = true;
    int first_row = -1;
    T first_sum = T(0);
    for (int step = dl; step < de; ++step) {
        if (i < rows_local && lend[i] <= j) {
            if (first) {
                first = false;
                first_row = i;
                first_sum = sum;
            } else {
                store_result<Mode>(c + (row0 + i) * cs, sum, alpha, beta);
            }
            sum = T(0);
            ++i;
        } else {
            sum += prod[j];
            ++j;
        }
    }
    // carries: inclusive segmented scan of the unfinished tails over the
    // threads (rows are non-decreasing in the thread index, so equal rows
    // form contiguous runs)
    crow[tid] = i;
    cval[tid] = sum;
    __syncthreads();
    for (int off = 1; off < kMpThreads; off <<= 1) {
        T add = T(0);
        if (tid >= off && crow[tid - off] == crow[tid]) add = cval[tid - off];
        __syncthreads();
        cval[tid] += add;
        __syncthreads();
    }
    if (first_row >= 0) {
        // the first row finished here continues the carries of the threads
        // before (thread 0's continues earlier workgroups: fix-up kernel)
        T s = first_sum;
        if (tid > 0 && crow[tid - 1] == first_row) s = cval[tid - 1] + s;
        store_result<Mode>(c + (row0 + first_row) * cs, s, alpha, beta);
    }
    if (tid == kMpThreads - 1) {
        carry_row[blockIdx.x] = static_cast<int32_t>(row0 + crow[tid]);
        carry_val[blockIdx.x] = cval[tid];
    }
}

// Adds the workgroup carries to their rows: the first workgroup of each run
// of equal carry rows sums the run in order.
template <typename T, int Mode>
__global__ void __launch_bounds__(kBlock)
    csr_spmv_merge_fixup(int64_t groups, int64_t n_rows, const int32_t* __restrict__ carry_row,
                         const T* __restrict__ carry_val, T* __restrict__ c, int64_t cs,
                         const T* __restrict__ alpha)
{
    const int64_t w = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (w >= groups) return;
    const int32_t r = carry_row[w];
    if (r >= n_rows || (w > 0 && carry_row[w - 1] == r)) return;
    T s = carry_val[w];
    for (int64_t k = w + 1; k < groups && carry_row[k] == r; ++k) s += carry_val[k];
    if constexpr (Mode == kAdvanced) {
        s = alpha[0] * s;
    } else if constexpr (Mode == kAccumulate) {
        s = alpha ? alpha[0] * s : s;
    }
    c[static_cast<int64_t>(r) * cs] += s;
}

}  // namespace

int64_t merge_groups(int64_t n_rows, int64_t nnz)
{
    return ceildiv(n_rows + nnz, kMpSegment);
}

size_t merge_work_bytes(int64_t n_rows, int64_t nnz)
{
    const int64_t g = merge_groups(n_rows, nnz);
    return static_cast<size_t>(g) * (sizeof(int32_t) + sizeof(double)) + 16;
}

template <typename T, int Mode>
int csr_spmv_merge_launch(hipStream_t stream, const gkomi_csr* a, const T* b, int64_t bs, T* c,
                          int64_t cs, const T* alpha, const T* beta)
{
    const int64_t groups = merge_groups(a->n_rows, a->nnz);
    GKOMI_REQUIRE(groups < (int64_t{1} << 31), GKOMI_INVALID_ARGUMENT,
                  "CSR merge_path: matrix too large");
    GKOMI_REQUIRE(a->work != nullptr && a->work_bytes >= static_cast<int64_t>(merge_work_bytes(
                                                             a->n_rows, a->nnz)),
                  GKOMI_INVALID_ARGUMENT,
                  "CSR merge_path: strategy workspace missing or too small "
                  "(gkomi_csr_strategy_work_bytes)");
    // carry values first (8-byte aligned), then the carry rows
    T* carry_val = static_cast<T*>(a->work);
    int32_t* carry_row = reinterpret_cast<int32_t*>(static_cast<char*>(a->work) +
                                                    groups * static_cast<int64_t>(sizeof(double)));
    csr_spmv_merge_kernel<T, Mode><<<static_cast<unsigned>(groups), kMpThreads, 0, stream>>>(
        a->n_rows, a->nnz, a->row_ptrs, a->col_idxs, static_cast<const T*>(a->values), b, bs, c,
        cs, alpha, beta, carry_row, carry_val);
    GKOMI_CHECK_LAUNCH("csr_spmv_merge_kernel");
    csr_spmv_merge_fixup<T, Mode><<<grid_for(groups, kBlock), kBlock, 0, stream>>>(
        groups, a->n_rows, carry_row, carry_val, c, cs, alpha);
    GKOMI_CHECK_LAUNCH("csr_spmv_merge_fixup");
    return GKOMI_SUCCESS;
}

template int csr_spmv_merge_launch<double, kPlain>(hipStream_t, const gkomi_csr*, const double*,
                                                 int64_t, double*, int64_t, const double*,
                                                 const double*);
template int csr_spmv_merge_launch<double, kAdvanced>(hipStream_t, const gkomi_csr*, const double*,
                                                 int64_t, double*, int64_t, const double*,
                                                 const double*);
template int csr_spmv_merge_launch<double, kAccumulate>(hipStream_t, const gkomi_csr*, const double*,
                                                 int64_t, double*, int64_t, const double*,
                                                 const double*);
template int csr_spmv_merge_launch<float, kPlain>(hipStream_t, const gkomi_csr*, const float*,
                                                 int64_t, float*, int64_t, const float*,
                                                 const float*);
template int csr_spmv_merge_launch<float, kAdvanced>(hipStream_t, const gkomi_csr*, const float*,
                                                 int64_t, float*, int64_t, const float*,
                                                 const float*);
template int csr_spmv_merge_launch<float, kAccumulate>(hipStream_t, const gkomi_csr*, const float*,
                                                 int64_t, float*, int64_t, const float*,
                                                 const float*);

}  // namespace gkomi
