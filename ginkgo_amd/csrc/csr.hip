// CSR SpMV for gfx950 plus the CSR-side index work (load-balance partition,
// column panels, longest row) and the scalar Jacobi generate / apply.
//
// Kernels:
//  * csr_spmv_subwave<SW, RPG>: SW lanes of a wave64 per row, each lane group
//    holding RPG rows in flight, DPP / ds_swizzle group reduction; for rows
//    of (nearly) uniform short length such as stencils.  Rows are walked in
//    XCD-contiguous ranges.
//  * csr_spmv_balanced: workgroups own contiguous nnz ranges of the
//    partition `srow`; the block's row pointers and the products a_ij * x_j
//    are staged in LDS, then rows are reduced thread-per-row (short) or
//    wave-per-row (long).  Ranges that do not fit LDS take several passes;
//    a row beyond the LDS capacity is summed by the whole workgroup from
//    memory.  Fixed summation order: deterministic.
//  * csr_spmv_panels + panel_combine_kernel: the same row reduction over a
//    column-panel copy of the matrix, each XCD on one panel, then the panel
//    sums of a row added in panel order.
//  * csr_spmm_rowwave: one wave per row, lanes over right-hand sides (the
//    many-RHS case).
#include <algorithm>
#include <type_traits>
#include <cmath>

#include "common.hpp"
#include "scan.hpp"
#include "spmv_common.hpp"

namespace gkomi {

template <typename T, int SW, int RPG, int Mode, bool Strided, bool Compact>
__device__ __forceinline__ void csr_subwave_body(int64_t n_rows, const int32_t* __restrict__ rp,
                                                 const int32_t* __restrict__ ci,
                                                 const int16_t* __restrict__ co,
                                                 const T* __restrict__ v,
                                                 const T* __restrict__ b, int64_t bs,
                                                 T* __restrict__ c, int64_t cs,
                                                 const T* __restrict__ alpha,
                                                 const T* __restrict__ beta)
{
    const int64_t nb = static_cast<int64_t>(gridDim.x);
    const int64_t tiles = ceildiv(n_rows, subwave_tile_rows<SW, RPG>);
    for (int64_t t = xcd_block(blockIdx.x, gridDim.x); t < tiles; t += nb) {
        subwave_tile<T, SW, RPG, Compact>(
            n_rows, t, rp, ci, co, v,
            [&](int32_t col) { return Strided ? b[static_cast<int64_t>(col) * bs] : b[col]; },
            [&](int64_t row, T sum) { store_result<Mode>(c + row * cs, sum, alpha, beta); });
    }
}

// G lanes per row with staged, coalesced matrix loads (rows of nearly equal
// length: stencils): every wave walks windows of 64 / G consecutive rows with
// staged_window (spmv_common.hpp), which copies a window's entries to the
// wave's LDS in contiguous 64-entry runs and then sums each row in the order
// of the sub-wave kernel with G lanes per row.  The lanes of one sub-lane
// index hold consecutive rows, so a stencil's gathers read consecutive words.
// One window per wave and a grid of all windows: waves that end make room
// for new ones, which measured faster than a resident grid walking several
// windows per wave, also with the next window's loads issued ahead (7-point
// 128^3 FP64: 30.8 us against 36.5 and 38.4 us).
template <typename T, int G, int C, int Mode, bool Compact>
__global__ void __launch_bounds__(kBlock)
    csr_spmv_staged(int64_t n_rows, const int32_t* __restrict__ rp,
                    const int32_t* __restrict__ ci, const int16_t* __restrict__ co,
                    const T* __restrict__ v, const T* __restrict__ b, T* __restrict__ c,
                    int64_t cs, const T* __restrict__ alpha, const T* __restrict__ beta)
{
    using Idx = std::conditional_t<Compact, int16_t, int32_t>;
    __shared__ T sv[kWavesPerBlock][C];
    __shared__ Idx sc[kWavesPerBlock][C];
    const int wave = threadIdx.x / kWave;
    const int64_t windows = ceildiv(n_rows, kWave / G);
    const int64_t nw = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
    for (int64_t w = static_cast<int64_t>(xcd_block(blockIdx.x, gridDim.x)) * kWavesPerBlock +
                     wave;
         w < windows; w += nw) {
        staged_window<T, G, C, Compact>(
            n_rows, w, rp, ci, co, v, sv[wave], sc[wave], [&](int32_t col) { return b[col]; },
            [&](int64_t row, T sum) { store_result<Mode>(c + row * cs, sum, alpha, beta); });
    }
}

template <typename T, int SW, int RPG, int Mode>
__global__ void __launch_bounds__(kBlock)
    csr_spmv_subwave(int64_t n_rows, const int32_t* __restrict__ rp,
                     const int32_t* __restrict__ ci, const int16_t* __restrict__ co,
                     const T* __restrict__ v, const T* __restrict__ b, int64_t bs,
                     T* __restrict__ c, int64_t cs, const T* __restrict__ alpha,
                     const T* __restrict__ beta)
{
    // the compact offsets serve the unit-stride case; strided vectors read
    // the full indices
    if (bs == 1 && co != nullptr) {
        csr_subwave_body<T, SW, RPG, Mode, false, true>(n_rows, rp, ci, co, v, b, bs, c, cs,
                                                       alpha, beta);
    } else if (bs == 1) {
        csr_subwave_body<T, SW, RPG, Mode, false, false>(n_rows, rp, ci, co, v, b, bs, c, cs,
                                                        alpha, beta);
    } else {
        csr_subwave_body<T, SW, RPG, Mode, true, false>(n_rows, rp, ci, co, v, b, bs, c, cs,
                                                       alpha, beta);
    }
}

// One LDS pass over the rows [rs, re) of a CSR product, which hold at most
// kLbLdsRows rows and kLbLdsNnz entries: the row pointers and the products
// a_ij * x_j are staged in LDS, then rows up to kLbShortRow entries are
// summed thread-per-row (sequentially) and longer ones wave-per-row.
template <typename T, int Mode>
__device__ __forceinline__ void csr_rows_lds_pass(int32_t rs, int32_t re,
                                                  const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ ci,
                                                  const T* __restrict__ v,
                                                  const T* __restrict__ b, int64_t bs,
                                                  T* __restrict__ c, int64_t cs,
                                                  const T* __restrict__ alpha,
                                                  const T* __restrict__ beta, T* prod,
                                                  int32_t* lrp, int32_t* long_rows,
                                                  int32_t* num_long)
{
    const int32_t ps = rp[rs];
    const int32_t nrows = re - rs;
    const int32_t nnzb = rp[re] - ps;
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    if (threadIdx.x == 0) *num_long = 0;
    for (int i = threadIdx.x; i <= nrows; i += kBlock) {
        lrp[i] = rp[rs + i] - ps;
    }
    for (int i = threadIdx.x; i < nnzb; i += kBlock) {
        prod[i] = v[ps + i] * b[static_cast<int64_t>(ci[ps + i]) * bs];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nrows; i += kBlock) {
        const int32_t s = lrp[i];
        const int32_t e = lrp[i + 1];
        if (e - s <= kLbShortRow) {
            T sum = 0;
            for (int32_t j = s; j < e; ++j) sum += prod[j];
            store_result<Mode>(c + static_cast<int64_t>(rs + i) * cs, sum, alpha, beta);
        } else {
            long_rows[atomicAdd(num_long, 1)] = i;
        }
    }
    __syncthreads();
    const int nl = *num_long;
    for (int t = wave; t < nl; t += kWavesPerBlock) {
        const int i = long_rows[t];
        const int32_t s = lrp[i];
        const int32_t e = lrp[i + 1];
        T sum = 0;
        for (int32_t j = s + lane; j < e; j += kWave) sum += prod[j];
        sum = wave_sum(sum);
        if (lane == 0) {
            store_result<Mode>(c + static_cast<int64_t>(rs + i) * cs, sum, alpha, beta);
        }
    }
}

// The rows [rs, re) of a CSR product, reduced by one workgroup.  rp holds
// absolute positions into ci / v.  A range that fits LDS takes one pass;
// a larger one is cut into consecutive passes that fit, and a row longer
// than the LDS capacity is summed by the whole workgroup from memory.  The
// summation order depends only on the range: deterministic.
template <typename T, int Mode>
__device__ __forceinline__ void csr_rows_staged(int32_t rs, int32_t re,
                                                const int32_t* __restrict__ rp,
                                                const int32_t* __restrict__ ci,
                                                const T* __restrict__ v,
                                                const T* __restrict__ b, int64_t bs,
                                                T* __restrict__ c, int64_t cs,
                                                const T* __restrict__ alpha,
                                                const T* __restrict__ beta)
{
    __shared__ T prod[kLbLdsNnz];
    __shared__ int32_t lrp[kLbLdsRows + 1];
    __shared__ int32_t long_rows[kLbLdsRows];
    __shared__ int32_t num_long;
    __shared__ T red[kWavesPerBlock];
    if (re - rs <= kLbLdsRows && rp[re] - rp[rs] <= kLbLdsNnz) {
        csr_rows_lds_pass<T, Mode>(rs, re, rp, ci, v, b, bs, c, cs, alpha, beta, prod, lrp,
                                   long_rows, &num_long);
        return;
    }
    // every thread takes the same steps (uniform control flow)
    int32_t cur = rs;
    while (cur < re) {
        const int32_t base = rp[cur];
        // the longest run [cur, end) with at most kLbLdsRows rows and
        // kLbLdsNnz entries
        int32_t lo = cur;
        int32_t hi = min(re, cur + kLbLdsRows);
        while (lo < hi) {
            const int32_t mid = lo + (hi - lo + 1) / 2;
            if (rp[mid] - base <= kLbLdsNnz) {
                lo = mid;
            } else {
                hi = mid - 1;
            }
        }
        if (lo == cur) {
            const int32_t e = rp[cur + 1];
            T sum = 0;
            for (int32_t j = base + threadIdx.x; j < e; j += kBlock) {
                sum += v[j] * b[static_cast<int64_t>(ci[j]) * bs];
            }
            sum = block_sum(sum, red);
            if (threadIdx.x == 0) {
                store_result<Mode>(c + static_cast<int64_t>(cur) * cs, sum, alpha, beta);
            }
            cur += 1;
        } else {
            csr_rows_lds_pass<T, Mode>(cur, lo, rp, ci, v, b, bs, c, cs, alpha, beta, prod, lrp,
                                       long_rows, &num_long);
            __syncthreads();
            cur = lo;
        }
    }
}

template <typename T, int Mode>
__global__ void __launch_bounds__(kBlock)
    csr_spmv_balanced(int64_t n_rows, const int32_t* __restrict__ rp,
                      const int32_t* __restrict__ ci, const T* __restrict__ v,
                      const T* __restrict__ b, int64_t bs, T* __restrict__ c, int64_t cs,
                      const int32_t* __restrict__ srow, int64_t num_srow, int64_t group,
                      const T* __restrict__ alpha, const T* __restrict__ beta)
{
    // workgroup k takes the warps [k * group, (k + 1) * group) of the
    // partition and owns the rows [srow[k * group], srow[(k + 1) * group]):
    // row 0 onwards for the first, up to n_rows for the last.  Any
    // non-decreasing srow gives every row exactly one owner; the clamps keep
    // a malformed one in bounds.
    const int64_t k = xcd_block(blockIdx.x, gridDim.x);
    const int64_t w0 = k * group;
    const int64_t w1 = w0 + group;
    int64_t rs64 = k == 0 ? 0 : srow[w0];
    int64_t re64 = w1 >= num_srow ? n_rows : srow[w1];
    rs64 = rs64 < 0 ? 0 : (rs64 > n_rows ? n_rows : rs64);
    re64 = re64 < 0 ? 0 : (re64 > n_rows ? n_rows : re64);
    if (rs64 >= re64) {
        return;
    }
    csr_rows_staged<T, Mode>(static_cast<int32_t>(rs64), static_cast<int32_t>(re64), rp, ci, v,
                             b, bs, c, cs, alpha, beta);
}

// The CLASSICAL single right-hand-side product over the sliced copy: one
// wave per 64-row slice, slices walked in XCD-contiguous ranges.
template <typename T, int Mode, bool Compact>
__global__ void __launch_bounds__(kBlock)
    csr_spmv_sliced(int64_t n_rows, const int32_t* __restrict__ rp,
                    const int64_t* __restrict__ sets, const int32_t* __restrict__ cols,
                    const int16_t* __restrict__ offs, const T* __restrict__ vals,
                    const T* __restrict__ b, int64_t bs, T* __restrict__ c, int64_t cs,
                    const T* __restrict__ alpha, const T* __restrict__ beta)
{
    const int64_t slices = ceildiv(n_rows, kWave);
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
    for (int64_t sl = static_cast<int64_t>(xcd_block(blockIdx.x, gridDim.x)) * kWavesPerBlock +
                      threadIdx.x / kWave;
         sl < slices; sl += nwaves) {
        sliced_slice<T, Compact>(
            n_rows, sl, rp, sets, cols, offs, vals,
            [&](int32_t col) { return b[static_cast<int64_t>(col) * bs]; },
            [&](int64_t row, T sum) { store_result<Mode>(c + row * cs, sum, alpha, beta); });
    }
}

template <typename T, int Mode>
__global__ void __launch_bounds__(kBlock)
    csr_spmm_rowwave(int64_t n_rows, int64_t nrhs, const int32_t* __restrict__ rp,
                     const int32_t* __restrict__ ci, const T* __restrict__ v,
                     const T* __restrict__ b, int64_t bs, T* __restrict__ c, int64_t cs,
                     const T* __restrict__ alpha, const T* __restrict__ beta)
{
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
                // entries in batches of eight, every load of a batch issued
                // before its first product (the batch bounds are the same in
                // the whole wave); the products are added in stored order
                constexpr int B = 8;
                T sum = 0;
                for (int32_t k0 = s; k0 < e; k0 += B) {
                    T vv[B], bv[B];
#pragma unroll
                    for (int u = 0; u < B; ++u) {
                        if (k0 + u < e) {
                            vv[u] = v[k0 + u];
                            bv[u] = b[static_cast<int64_t>(ci[k0 + u]) * bs + j];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < B; ++u) {
                        if (k0 + u < e) sum += vv[u] * bv[u];
                    }
                }
                store_result<Mode>(c + row * cs + j, sum, alpha, beta);
            }
        }
    }
}

// The load_balance partition (Csr::load_balance::process,
// include/ginkgo/core/matrix/csr.hpp:417-468): the reference drops row i into
// the bucket ceil(L_i * nwarps / D), L_i = ceil(row_ptrs[i + 1] / warp_size),
// D = ceil(nnz / warp_size) (1 without entries), and prefix-counts the buckets.
// Since ceil(a / D) <= w exactly when a <= w * D and L_i never decreases,
// srow[w] is the first row i with L_i * nwarps > w * D (n_rows if none): one
// binary search per entry, the same integers as the reference's loops.
__global__ void __launch_bounds__(kBlock)
    lb_partition_kernel(int64_t n_rows, const int32_t* __restrict__ rp, int64_t nnz,
                        int64_t warp_size, int32_t* __restrict__ srow, int64_t nwarps)
{
    const int64_t divider = nnz > 0 ? ceildiv(nnz, warp_size) : 1;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t w = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; w < nwarps;
         w += stride) {
        const int64_t target = w * divider;
        int64_t lo = 0;
        int64_t hi = n_rows;
        while (lo < hi) {
            const int64_t mid = (lo + hi) / 2;
            if (ceildiv(rp[mid + 1], warp_size) * nwarps > target) {
                hi = mid;
            } else {
                lo = mid + 1;
            }
        }
        srow[w] = static_cast<int32_t>(lo);
    }
}

// ---------------------------------------------------------- column panels ----
// The columns split into P equal panels (P = 2, 4 or 8, width W =
// ceil(n_cols / P)); panel p keeps its own CSR copy of the entries with
// columns in [p W, (p + 1) W), rows in the same order and the entries of a row
// in their original order.  Workgroup b of a product works on panel b % P;
// the hardware places workgroup b on XCD b % 8, so each XCD gathers from one
// window of x of n_cols / P values, which for P = 8 and an 8 MB FP64 x stays
// in that XCD's 4 MB L2.  Panel p writes its row sums to partial[p][row];
// a second pass adds the P partial sums of a row in panel order.
//
// Panel storage in one device block (each part 256-byte aligned):
//   row_ptrs  int32 [P n_rows + 1]: panel p's row i starts at
//             row_ptrs[p n_rows + i]; one exclusive scan over the (panel,
//             row) counts, so panel p ends where panel p + 1 starts
//   col_idxs  int32 [nnz], values T [nnz]
//   bounds    int32 [P (C + 1)]: chunk k of panel p owns the rows
//             [bounds[p (C + 1) + k], bounds[p (C + 1) + k + 1])
//   partial   T [P n_rows] (also the scan's workspace while building)
constexpr int kMaxPanels = 8;
// rows plus entries per chunk of an evenly filled panel (a chunk beyond the
// LDS capacity takes several passes); on the 1M-row power-law matrix in 4
// panels 768 is 3% slower, 384 35% slower
constexpr int64_t kPanelChunk = 1536;

struct panel_layout {
    int64_t P, C, n, nnz;
    size_t row_ptrs, col_idxs, values, bounds, partial, total;
};

inline size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// Chunks per panel: at most kPanelChunk rows plus entries each for an
// evenly filled panel.
inline int64_t panel_chunks(int64_t n_rows, int64_t nnz, int64_t P)
{
    return std::max<int64_t>(1, ceildiv(n_rows + ceildiv(nnz, P), kPanelChunk));
}

inline panel_layout make_panel_layout(int64_t n_rows, int64_t nnz, int64_t P, size_t vb)
{
    panel_layout L{};
    L.P = P;
    L.n = n_rows;
    L.nnz = nnz;
    L.C = panel_chunks(n_rows, nnz, P);
    size_t off = 0;
    L.row_ptrs = off;
    off += align256(sizeof(int32_t) * static_cast<size_t>(P * n_rows + 1));
    L.col_idxs = off;
    off += align256(sizeof(int32_t) * static_cast<size_t>(nnz));
    L.values = off;
    off += align256(vb * static_cast<size_t>(nnz));
    L.bounds = off;
    off += align256(sizeof(int32_t) * static_cast<size_t>(P * (L.C + 1)));
    L.partial = off;
    off += align256(std::max(vb * static_cast<size_t>(P * n_rows),
                             scan_ws_bytes(P * n_rows, sizeof(int32_t))));
    L.total = off;
    return L;
}

__device__ __forceinline__ int panel_of(int32_t col, int64_t width, int P)
{
    const int64_t q = col < 0 ? 0 : col / width;
    return q >= P ? P - 1 : static_cast<int>(q);
}

// entries of each row per panel: counts[p n_rows + row], one wave per row
__global__ void __launch_bounds__(kBlock)
    panel_count_kernel(int64_t n_rows, const int32_t* __restrict__ rp,
                       const int32_t* __restrict__ ci, int64_t width, int P,
                       int32_t* __restrict__ counts)
{
    const int lane = threadIdx.x % kWave;
    const int64_t num_waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
    for (int64_t row = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
         row < n_rows; row += num_waves) {
        const int32_t e = rp[row + 1];
        int32_t mine = 0;  // lane q < P counts panel q
        for (int32_t j0 = rp[row]; j0 < e; j0 += kWave) {
            const int32_t j = j0 + lane;
            const int p = j < e ? panel_of(ci[j], width, P) : -1;
#pragma unroll
            for (int q = 0; q < kMaxPanels; ++q) {
                const uint64_t m = __ballot(p == q);
                if (lane == q) mine += __popcll(m);
            }
        }
        if (lane < P) counts[lane * n_rows + row] = mine;
    }
}

// copies every entry to its panel, keeping the order inside a row: the
// position is the panel row's start plus the entries of that row and panel
// in earlier 64-entry steps plus those of lower lanes in this step
template <typename T>
__global__ void __launch_bounds__(kBlock)
    panel_scatter_kernel(int64_t n_rows, const int32_t* __restrict__ rp,
                         const int32_t* __restrict__ ci, const T* __restrict__ v,
                         int64_t width, int P, const int32_t* __restrict__ prp,
                         int32_t* __restrict__ pci, T* __restrict__ pv)
{
    const int lane = threadIdx.x % kWave;
    const uint64_t below = (uint64_t{1} << lane) - 1;
    const int64_t num_waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
    for (int64_t row = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
         row < n_rows; row += num_waves) {
        int32_t next[kMaxPanels];
#pragma unroll
        for (int q = 0; q < kMaxPanels; ++q) next[q] = q < P ? prp[q * n_rows + row] : 0;
        const int32_t e = rp[row + 1];
        for (int32_t j0 = rp[row]; j0 < e; j0 += kWave) {
            const int32_t j = j0 + lane;
            const int32_t col = j < e ? ci[j] : 0;
            const int p = j < e ? panel_of(col, width, P) : -1;
            int32_t pos = 0;
#pragma unroll
            for (int q = 0; q < kMaxPanels; ++q) {
                const uint64_t m = __ballot(p == q);
                if (p == q) pos = next[q] + __popcll(m & below);
                next[q] += __popcll(m);
            }
            if (p >= 0) {
                pci[pos] = col;
                pv[pos] = v[j];
            }
        }
    }
}

// chunk bounds of every panel: chunk k of panel p starts at the first row i
// whose rows-plus-entries count through row i, (i + 1) + (entries of panel
// p in rows 0..i), exceeds k (n_rows + nnz_p) / C; chunk 0 starts at row 0,
// chunk C at n_rows
__global__ void __launch_bounds__(kBlock)
    panel_bounds_kernel(int64_t n_rows, int P, int64_t C, const int32_t* __restrict__ prp,
                        int32_t* __restrict__ bounds)
{
    const int64_t total_items = P * (C + 1);
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t idx = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
         idx < total_items; idx += stride) {
        const int64_t p = idx / (C + 1);
        const int64_t k = idx - p * (C + 1);
        const int32_t* rp = prp + p * n_rows;
        const int64_t base = rp[0];
        const int64_t total = n_rows + (static_cast<int64_t>(rp[n_rows]) - base);
        const int64_t target = k * total / C;
        int64_t lo = 0;
        int64_t hi = n_rows;
        while (lo < hi) {
            const int64_t mid = (lo + hi) / 2;
            if (mid + 1 + (static_cast<int64_t>(rp[mid + 1]) - base) > target) {
                hi = mid;
            } else {
                lo = mid + 1;
            }
        }
        bounds[idx] = static_cast<int32_t>(lo);
    }
}

// partial[p][row] for every panel: workgroup b takes chunk b / P of panel
// b % P (grid P C)
template <typename T>
__global__ void __launch_bounds__(kBlock)
    csr_spmv_panels(int64_t n_rows, int P, int64_t C, const int32_t* __restrict__ prp,
                    const int32_t* __restrict__ pci, const T* __restrict__ pv,
                    const int32_t* __restrict__ bounds, const T* __restrict__ b, int64_t bs,
                    T* __restrict__ partial)
{
    const int64_t p = blockIdx.x % P;
    const int64_t k = blockIdx.x / P;
    const int32_t* bnd = bounds + p * (C + 1) + k;
    int32_t rs = bnd[0];
    int32_t re = bnd[1];
    rs = rs < 0 ? 0 : (rs > n_rows ? static_cast<int32_t>(n_rows) : rs);
    re = re < rs ? rs : (re > n_rows ? static_cast<int32_t>(n_rows) : re);
    if (rs >= re) {
        return;
    }
    csr_rows_staged<T, kPlain>(rs, re, prp + p * n_rows, pci, pv, b, bs, partial + p * n_rows,
                               1, static_cast<const T*>(nullptr), static_cast<const T*>(nullptr));
}

// c[row] <- sum over panels of partial[p][row], in panel order
template <typename T, int Mode>
__global__ void __launch_bounds__(kBlock)
    panel_combine_kernel(int64_t n_rows, int P, const T* __restrict__ partial,
                         T* __restrict__ c, int64_t cs, const T* __restrict__ alpha,
                         const T* __restrict__ beta)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t row = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; row < n_rows;
         row += stride) {
        T sum = partial[row];
        for (int p = 1; p < P; ++p) sum += partial[p * n_rows + row];
        store_result<Mode>(c + row * cs, sum, alpha, beta);
    }
}

__global__ void __launch_bounds__(kBlock)
    max_row_nnz_kernel(int64_t n_rows, const int32_t* __restrict__ rp,
                       unsigned long long* __restrict__ result)
{
    __shared__ int64_t lds[kWavesPerBlock];
    int64_t m = 0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n_rows;
         i += stride) {
        m = max(m, static_cast<int64_t>(rp[i + 1] - rp[i]));
    }
    for (int off = 32; off >= 1; off /= 2) m = max(m, __shfl_xor(m, off, 64));
    if (threadIdx.x % kWave == 0) lds[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kWavesPerBlock; ++w) m = max(m, lds[w]);
        atomicMax(result, static_cast<unsigned long long>(m));
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    jacobi_generate_kernel(int64_t n_rows, const int32_t* __restrict__ rp,
                           const int32_t* __restrict__ ci, const T* __restrict__ v,
                           T* __restrict__ inv_diag)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t row = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; row < n_rows;
         row += stride) {
        T d = 0;
        for (int32_t k = rp[row]; k < rp[row + 1]; ++k) {
            if (ci[k] == row) {
                d = v[k];
                break;
            }
        }
        inv_diag[row] = (d == T(0)) ? T(1) : T(1) / d;
    }
}

template <typename T, bool Adv>
__global__ void __launch_bounds__(kBlock)
    jacobi_apply_kernel(int64_t n, int64_t nrhs, const T* __restrict__ inv_diag,
                        const T* __restrict__ alpha, const T* __restrict__ b, int64_t bs,
                        const T* __restrict__ beta, T* __restrict__ x, int64_t xs)
{
    const int64_t total = n * nrhs;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; e < total;
         e += stride) {
        const int64_t i = e / nrhs;
        const int64_t j = e - i * nrhs;
        if constexpr (Adv) {
            x[i * xs + j] = beta[0] * x[i * xs + j] + alpha[0] * b[i * bs + j] * inv_diag[i];
        } else {
            x[i * xs + j] = b[i * bs + j] * inv_diag[i];
        }
    }
}

// -------------------------------------------------------------- drivers ----

// col - row per entry as int16; a row that needs a wider offset clears *fits
__global__ void __launch_bounds__(kBlock)
    col_offsets_kernel(int64_t n_rows, const int32_t* __restrict__ rp,
                       const int32_t* __restrict__ ci, int16_t* __restrict__ out,
                       int32_t* __restrict__ fits)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    bool ok = true;
    for (int64_t row = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; row < n_rows;
         row += stride) {
        const int32_t e = rp[row + 1];
        for (int32_t p = rp[row]; p < e; ++p) {
            const int64_t d = static_cast<int64_t>(ci[p]) - row;
            ok = ok && d >= -32768 && d <= 32767;
            out[p] = static_cast<int16_t>(d);
        }
    }
    if (!ok) atomicMin(fits, 0);
}

int check_csr(const gkomi_csr* a)
{
    GKOMI_REQUIRE(a != nullptr, GKOMI_INVALID_ARGUMENT, "null CSR descriptor");
    GKOMI_REQUIRE(a->n_rows >= 0 && a->n_cols >= 0 && a->nnz >= 0, GKOMI_INVALID_ARGUMENT,
                  "CSR: negative size");
    GKOMI_REQUIRE(a->nnz < (int64_t{1} << 31) && a->n_rows < (int64_t{1} << 31),
                  GKOMI_INVALID_ARGUMENT, "CSR: sizes exceed the int32 index range");
    GKOMI_REQUIRE(a->n_rows == 0 || a->row_ptrs != nullptr, GKOMI_INVALID_ARGUMENT,
                  "CSR: null row pointers");
    GKOMI_REQUIRE(a->nnz == 0 || (a->col_idxs != nullptr && a->values != nullptr),
                  GKOMI_INVALID_ARGUMENT, "CSR: null column indices or values");
    GKOMI_REQUIRE(a->strategy == GKOMI_CSR_CLASSICAL || a->strategy == GKOMI_CSR_LOAD_BALANCE ||
                      a->strategy == GKOMI_CSR_MERGE_PATH,
                  GKOMI_INVALID_ARGUMENT, "CSR: unknown strategy");
    GKOMI_REQUIRE(a->strategy != GKOMI_CSR_LOAD_BALANCE || a->num_srow == 0 ||
                      (a->srow != nullptr && a->num_srow > 0),
                  GKOMI_INVALID_ARGUMENT, "CSR: load_balance strategy with a null partition");
    return GKOMI_SUCCESS;
}

// Lanes per row of the staged kernel (csr_spmv_staged), 0 = the sub-wave
// kernel, -1 = an invalid request: GKOMI_CSR_STAGED | G asks for it;
// automatic (subwave 0) takes the smallest power of two G with
// (64 / G) * mean row <= kStagedChunk (a window of rows holds on average at
// most one staging chunk), if G <= 4, or G = 8 in FP32.  Measured against the
// sub-wave kernel's automatic shape (us per product): 128^3 7-point FP64 G = 1
// 30.7 / 37.7, FP32 22.3 / 22.8; 1448^2 5-point FP64 25.8 / 31.0; 27-point
// 96^3 FP64 G = 4 61.5 / 70.0, FP32 31.1 / 45.3; a box stencil of 34.6
// entries per row (505,521 rows) G = 8 FP64 43.2 / 38.7 (slower: not taken),
// FP32 29.5 / 32.1.
int staged_lanes(const gkomi_csr* a, int value_bytes)
{
    if ((a->subwave & ~0xff) == GKOMI_CSR_STAGED) {
        const int G = a->subwave & 0xff;
        return (G == 1 || G == 2 || G == 4 || G == 8 || G == 16) ? G : -1;
    }
    if (a->subwave != 0 || a->n_rows <= 0) return 0;
    const int limit = value_bytes >= 8 ? 4 : 8;
    for (int G = 1; G <= limit; G *= 2) {
        if ((kWave / G) * a->nnz <= int64_t{kStagedChunk} * a->n_rows) return G;
    }
    return 0;
}

subwave_shape pick_subwave(const gkomi_csr* a, int value_bytes)
{
    subwave_shape sh{a->subwave & 0xff, (a->subwave >> 8) & 0xff};
    const double mean = a->n_rows > 0 ? static_cast<double>(a->nnz) / a->n_rows : 1.0;
    if (sh.sw == 0) {
        // about two entries per lane and row up to 8 lanes, three beyond
        // (rows of 25 and 35 entries: 16 lanes read 4 and 23% faster than 32)
        sh.sw = 1;
        while (sh.sw < 64 && (sh.sw < 8 ? 2 * sh.sw : 3 * sh.sw) < mean) sh.sw *= 2;
    }
    if (sh.rpg == 0) {
        // entries in flight per lane: eight, except four for FP64 rows that
        // take at most four lanes (measured over sw 1..32 and rpg 1..8 on the
        // 5-point 1448^2, 7-point 128^3 and 27-point 96^3 matrices, FP64 and
        // FP32; on the 7-point FP64 case four also reads the least: 1.01x the
        // algorithmic bytes against 1.06x with eight)
        const int target = (value_bytes >= 8 && sh.sw <= 4) ? 4 : 8;
        const int iters = std::max(1, static_cast<int>(std::ceil(mean / sh.sw)));
        sh.rpg = 1;
        while (sh.rpg < 8 && 2 * sh.rpg * iters <= target) sh.rpg *= 2;
    }
    return sh;
}

int check_panels(const gkomi_csr* a, size_t vb, panel_layout* L)
{
    GKOMI_REQUIRE(a->num_panels == 2 || a->num_panels == 4 || a->num_panels == 8,
                  GKOMI_INVALID_ARGUMENT, "CSR: column panels: num_panels must be 2, 4 or 8");
    *L = make_panel_layout(a->n_rows, a->nnz, a->num_panels, vb);
    GKOMI_REQUIRE(a->panels_bytes >= static_cast<int64_t>(L->total), GKOMI_INVALID_ARGUMENT,
                  "CSR: column panels smaller than gkomi_csr_panels_bytes");
    GKOMI_REQUIRE(L->P * L->C < (int64_t{1} << 31) && L->P * a->n_rows < (int64_t{1} << 31),
                  GKOMI_INVALID_ARGUMENT, "CSR: column panels too large");
    return GKOMI_SUCCESS;
}

template <typename T, int Mode>
int csr_spmv_panels_launch(hipStream_t stream, const gkomi_csr* a, const T* b, int64_t bs, T* c,
                           int64_t cs, const T* alpha, const T* beta)
{
    panel_layout L;
    if (int rc = check_panels(a, sizeof(T), &L)) return rc;
    char* base = static_cast<char*>(a->panels);
    T* partial = reinterpret_cast<T*>(base + L.partial);
    csr_spmv_panels<T><<<static_cast<unsigned>(L.P * L.C), kBlock, 0, stream>>>(
        a->n_rows, static_cast<int>(L.P), L.C, reinterpret_cast<const int32_t*>(base + L.row_ptrs),
        reinterpret_cast<const int32_t*>(base + L.col_idxs),
        reinterpret_cast<const T*>(base + L.values),
        reinterpret_cast<const int32_t*>(base + L.bounds), b, bs, partial);
    GKOMI_CHECK_LAUNCH("csr_spmv_panels");
    panel_combine_kernel<T, Mode><<<grid_for(a->n_rows, kBlock, 16), kBlock, 0, stream>>>(
        a->n_rows, static_cast<int>(L.P), partial, c, cs, alpha, beta);
    GKOMI_CHECK_LAUNCH("panel_combine_kernel");
    return GKOMI_SUCCESS;
}

template <typename T>
int csr_build_panels_impl(const gkomi_ctx* ctx, const gkomi_csr* a, int32_t num_panels,
                          void* panels, int64_t panels_bytes)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_csr(a)) return rc;
    gkomi_csr d = *a;
    d.panels = panels;
    d.panels_bytes = panels_bytes;
    d.num_panels = num_panels;
    panel_layout L;
    if (int rc = check_panels(&d, sizeof(T), &L)) return rc;
    if (a->n_rows == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(panels != nullptr, GKOMI_INVALID_ARGUMENT, "CSR: null column panels");
    const hipStream_t stream = stream_of(ctx);
    char* base = static_cast<char*>(panels);
    int32_t* prp = reinterpret_cast<int32_t*>(base + L.row_ptrs);
    const int P = num_panels;
    const int64_t width = std::max<int64_t>(1, ceildiv(a->n_cols, P));
    const unsigned g = grid_for(a->n_rows, kWavesPerBlock, 32);
    panel_count_kernel<<<g, kBlock, 0, stream>>>(a->n_rows, a->row_ptrs, a->col_idxs, width, P,
                                                 prp);
    GKOMI_CHECK_LAUNCH("panel_count_kernel");
    if (int rc = exclusive_scan<int32_t>(stream, L.P * a->n_rows, prp, prp, base + L.partial)) {
        return rc;
    }
    panel_scatter_kernel<T><<<g, kBlock, 0, stream>>>(
        a->n_rows, a->row_ptrs, a->col_idxs, static_cast<const T*>(a->values), width, P, prp,
        reinterpret_cast<int32_t*>(base + L.col_idxs), reinterpret_cast<T*>(base + L.values));
    GKOMI_CHECK_LAUNCH("panel_scatter_kernel");
    panel_bounds_kernel<<<grid_for(L.P * (L.C + 1), kBlock), kBlock, 0, stream>>>(
        a->n_rows, P, L.C, prp, reinterpret_cast<int32_t*>(base + L.bounds));
    GKOMI_CHECK_LAUNCH("panel_bounds_kernel");
    return GKOMI_SUCCESS;
}

template <typename T, int Mode>
int csr_spmv_launch(hipStream_t stream, const gkomi_csr* a, int64_t nrhs, const T* b, int64_t bs,
                    T* c, int64_t cs, const T* alpha, const T* beta)
{
    const T* v = static_cast<const T*>(a->values);
    if (nrhs > 1) {
        const unsigned g = grid_for(a->n_rows, kWavesPerBlock, 32);
        csr_spmm_rowwave<T, Mode><<<g, kBlock, 0, stream>>>(a->n_rows, nrhs, a->row_ptrs,
                                                          a->col_idxs, v, b, bs, c, cs, alpha,
                                                          beta);
        GKOMI_CHECK_LAUNCH("csr_spmm_rowwave");
        return GKOMI_SUCCESS;
    }
    if (a->panels != nullptr && a->num_panels > 0) {
        return csr_spmv_panels_launch<T, Mode>(stream, a, b, bs, c, cs, alpha, beta);
    }
    if (a->strategy == GKOMI_CSR_MERGE_PATH) {
        return csr_spmv_merge_launch<T, Mode>(stream, a, b, bs, c, cs, alpha, beta);
    }
    if (a->strategy == GKOMI_CSR_LOAD_BALANCE && a->num_srow > 0) {
        // consecutive warps of the partition per workgroup, about kLbChunk
        // entries each (a partition without elements has no entries: the
        // classical kernel below then writes the empty rows)
        const int64_t per_warp = std::max<int64_t>(1, ceildiv(a->nnz, a->num_srow));
        const int64_t group = std::max<int64_t>(1, kLbChunk / per_warp);
        const int64_t blocks = ceildiv(a->num_srow, group);
        GKOMI_REQUIRE(blocks < (int64_t{1} << 31), GKOMI_INVALID_ARGUMENT,
                      "CSR: partition too large");
        csr_spmv_balanced<T, Mode><<<static_cast<unsigned>(blocks), kBlock, 0, stream>>>(
            a->n_rows, a->row_ptrs, a->col_idxs, v, b, bs, c, cs, a->srow, a->num_srow, group,
            alpha, beta);
        GKOMI_CHECK_LAUNCH("csr_spmv_balanced");
        return GKOMI_SUCCESS;
    }
    if (a->slice_sets != nullptr && a->slice_values != nullptr &&
        (a->slice_offsets != nullptr || a->slice_cols != nullptr)) {
        const unsigned g = grid_for(ceildiv(a->n_rows, kWave), kWavesPerBlock, 64);
        const T* sv = static_cast<const T*>(a->slice_values);
        if (a->slice_offsets) {
            csr_spmv_sliced<T, Mode, true><<<g, kBlock, 0, stream>>>(
                a->n_rows, a->row_ptrs, a->slice_sets, a->slice_cols, a->slice_offsets, sv, b, bs,
                c, cs, alpha, beta);
        } else {
            csr_spmv_sliced<T, Mode, false><<<g, kBlock, 0, stream>>>(
                a->n_rows, a->row_ptrs, a->slice_sets, a->slice_cols, a->slice_offsets, sv, b, bs,
                c, cs, alpha, beta);
        }
        GKOMI_CHECK_LAUNCH("csr_spmv_sliced");
        return GKOMI_SUCCESS;
    }
    if (bs == 1) {
        const int G = staged_lanes(a, sizeof(T));
        GKOMI_REQUIRE(G >= 0, GKOMI_INVALID_ARGUMENT,
                      "CSR: GKOMI_CSR_STAGED needs 1, 2, 4, 8 or 16 lanes per row");
        if (G > 0) {
            const unsigned g =
                grid_for(ceildiv(a->n_rows, int64_t{kWave} / G), kWavesPerBlock, 64);
#define GKOMI_STAGED_CASE(GV)                                                                \
    if (G == GV) {                                                                          \
        if (a->col_offsets != nullptr) {                                                    \
            csr_spmv_staged<T, GV, kStagedChunk, Mode, true><<<g, kBlock, 0, stream>>>(      \
                a->n_rows, a->row_ptrs, a->col_idxs, a->col_offsets, v, b, c, cs, alpha,     \
                beta);                                                                      \
        } else {                                                                            \
            csr_spmv_staged<T, GV, kStagedChunk, Mode, false><<<g, kBlock, 0, stream>>>(     \
                a->n_rows, a->row_ptrs, a->col_idxs, a->col_offsets, v, b, c, cs, alpha,     \
                beta);                                                                      \
        }                                                                                   \
        GKOMI_CHECK_LAUNCH("csr_spmv_staged");                                              \
        return GKOMI_SUCCESS;                                                               \
    }
            GKOMI_STAGED_CASE(1)
            GKOMI_STAGED_CASE(2)
            GKOMI_STAGED_CASE(4)
            GKOMI_STAGED_CASE(8)
            GKOMI_STAGED_CASE(16)
#undef GKOMI_STAGED_CASE
        }
    }
    const subwave_shape sh = pick_subwave(a, sizeof(T));
    const int64_t tile_rows = int64_t{kBlock} / sh.sw * sh.rpg;
    const unsigned g = grid_for(a->n_rows, tile_rows, 64);
#define GKOMI_SPMV_CASE(S, R)                                                                 \
    csr_spmv_subwave<T, S, R, Mode><<<g, kBlock, 0, stream>>>(                                 \
        a->n_rows, a->row_ptrs, a->col_idxs, a->col_offsets, v, b, bs, c, cs, alpha, beta);   \
    break;
    GKOMI_SUBWAVE_SWITCH(sh, GKOMI_SPMV_CASE)
#undef GKOMI_SPMV_CASE
    GKOMI_CHECK_LAUNCH("csr_spmv_subwave");
    return GKOMI_SUCCESS;
}

template int csr_spmv_launch<double, kPlain>(hipStream_t, const gkomi_csr*, int64_t,
                                           const double*, int64_t, double*, int64_t,
                                           const double*, const double*);
template int csr_spmv_launch<double, kAdvanced>(hipStream_t, const gkomi_csr*, int64_t,
                                           const double*, int64_t, double*, int64_t,
                                           const double*, const double*);
template int csr_spmv_launch<double, kAccumulate>(hipStream_t, const gkomi_csr*, int64_t,
                                           const double*, int64_t, double*, int64_t,
                                           const double*, const double*);
template int csr_spmv_launch<float, kPlain>(hipStream_t, const gkomi_csr*, int64_t,
                                           const float*, int64_t, float*, int64_t,
                                           const float*, const float*);
template int csr_spmv_launch<float, kAdvanced>(hipStream_t, const gkomi_csr*, int64_t,
                                           const float*, int64_t, float*, int64_t,
                                           const float*, const float*);
template int csr_spmv_launch<float, kAccumulate>(hipStream_t, const gkomi_csr*, int64_t,
                                           const float*, int64_t, float*, int64_t,
                                           const float*, const float*);

template <typename T, int Mode>
int csr_spmv_impl(const gkomi_ctx* ctx, const T* alpha, const gkomi_csr* a, int64_t nrhs,
                  const T* b, int64_t bs, const T* beta, T* c, int64_t cs)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_csr(a)) return rc;
    GKOMI_REQUIRE(nrhs >= 0, GKOMI_INVALID_ARGUMENT, "spmv: negative number of columns");
    if (a->n_rows == 0 || nrhs == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(c != nullptr && (a->nnz == 0 || b != nullptr), GKOMI_INVALID_ARGUMENT,
                  "spmv: null vector");
    GKOMI_REQUIRE(bs >= nrhs && cs >= nrhs, GKOMI_INVALID_ARGUMENT,
                  "spmv: stride smaller than the number of columns");
    GKOMI_REQUIRE(Mode != kAdvanced || (alpha != nullptr && beta != nullptr),
                  GKOMI_INVALID_ARGUMENT,
                  "advanced spmv: null alpha or beta");
    return csr_spmv_launch<T, Mode>(stream_of(ctx), a, nrhs, b, bs, c, cs, alpha, beta);
}

template <typename T>
int jacobi_generate_impl(const gkomi_ctx* ctx, const gkomi_csr* a, T* inv_diag)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_csr(a)) return rc;
    GKOMI_REQUIRE(a->n_rows == a->n_cols, GKOMI_DIMENSION_MISMATCH,
                  "Jacobi: system matrix is not square");
    if (a->n_rows == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(inv_diag != nullptr, GKOMI_INVALID_ARGUMENT, "Jacobi: null output");
    jacobi_generate_kernel<T><<<grid_for(a->n_rows, kBlock), kBlock, 0, stream_of(ctx)>>>(
        a->n_rows, a->row_ptrs, a->col_idxs, static_cast<const T*>(a->values), inv_diag);
    GKOMI_CHECK_LAUNCH("jacobi_generate_kernel");
    return GKOMI_SUCCESS;
}

template <typename T, bool Adv>
int jacobi_apply_impl(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const T* inv_diag,
                      const T* alpha, const T* b, int64_t bs, const T* beta, T* x, int64_t xs)
{
    if (int rc = enter(ctx)) return rc;
    GKOMI_REQUIRE(n >= 0 && nrhs >= 0, GKOMI_INVALID_ARGUMENT, "Jacobi apply: negative size");
    if (n == 0 || nrhs == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(inv_diag && b && x, GKOMI_INVALID_ARGUMENT, "Jacobi apply: null pointer");
    GKOMI_REQUIRE(bs >= nrhs && xs >= nrhs, GKOMI_INVALID_ARGUMENT,
                  "Jacobi apply: stride smaller than the number of columns");
    GKOMI_REQUIRE(!Adv || (alpha && beta), GKOMI_INVALID_ARGUMENT,
                  "Jacobi apply: null alpha or beta");
    jacobi_apply_kernel<T, Adv><<<grid_for(n * nrhs, kBlock * 4), kBlock, 0, stream_of(ctx)>>>(
        n, nrhs, inv_diag, alpha, b, bs, beta, x, xs);
    GKOMI_CHECK_LAUNCH("jacobi_apply_kernel");
    return GKOMI_SUCCESS;
}

}  // namespace gkomi

using namespace gkomi;

extern "C" {

int64_t gkomi_csr_strategy_work_bytes(int32_t strategy, int64_t n_rows, int64_t nnz)
{
    if (strategy != GKOMI_CSR_MERGE_PATH || n_rows < 0 || nnz < 0) return 0;
    return static_cast<int64_t>(merge_work_bytes(n_rows, nnz));
}

// Csr::load_balance::clac_size (include/ginkgo/core/matrix/csr.hpp:475-505),
// as the reference's HIP-platform build compiles it.
int64_t gkomi_csr_lb_clac_size(int64_t nnz, int64_t nwarps, int32_t warp_size,
                               int32_t cuda_strategy, int32_t intel_strategy)
{
    if (warp_size <= 0 || nnz < 0 || nwarps < 0) return 0;
    int64_t multiple = 8;
    if (nnz >= 200000000) {
        multiple = 2048;
    } else if (nnz >= 20000000) {
        multiple = 512;
    } else if (nnz >= 2000000) {
        multiple = 128;
    } else if (nnz >= 200000) {
        multiple = 32;
    }
    if (intel_strategy) {
        multiple = nnz >= 200000000 ? 256 : (nnz >= 20000000 ? 32 : 8);
    }
    if (!cuda_strategy) {
        multiple = nnz >= 10000000 ? 64 : (nnz >= 1000000 ? 16 : 8);
    }
    return std::min(ceildiv(nnz, warp_size), nwarps * multiple);
}

int64_t gkomi_csr_panels_bytes(int64_t n_rows, int64_t nnz, int32_t num_panels,
                               int32_t value_type)
{
    if (n_rows < 0 || nnz < 0 || (num_panels != 2 && num_panels != 4 && num_panels != 8) ||
        (value_type != GKOMI_VALUE_F64 && value_type != GKOMI_VALUE_F32)) {
        return -1;
    }
    const size_t vb = value_type == GKOMI_VALUE_F64 ? sizeof(double) : sizeof(float);
    return static_cast<int64_t>(make_panel_layout(n_rows, nnz, num_panels, vb).total);
}

int gkomi_csr_build_panels_f64(const gkomi_ctx* ctx, const gkomi_csr* a, int32_t num_panels,
                               void* panels, int64_t panels_bytes)
{
    return csr_build_panels_impl<double>(ctx, a, num_panels, panels, panels_bytes);
}

int gkomi_csr_build_panels_f32(const gkomi_ctx* ctx, const gkomi_csr* a, int32_t num_panels,
                               void* panels, int64_t panels_bytes)
{
    return csr_build_panels_impl<float>(ctx, a, num_panels, panels, panels_bytes);
}

int gkomi_csr_lb_partition(const gkomi_ctx* ctx, int64_t n_rows, const int32_t* row_ptrs,
                           int64_t nnz, int32_t warp_size, int32_t* srow, int64_t nwarps)
{
    if (int rc = enter(ctx)) return rc;
    GKOMI_REQUIRE(n_rows >= 0 && nnz >= 0 && nwarps >= 0 && warp_size > 0,
                  GKOMI_INVALID_ARGUMENT, "partition: bad size");
    GKOMI_REQUIRE(n_rows < (int64_t{1} << 31) && nnz < (int64_t{1} << 31),
                  GKOMI_INVALID_ARGUMENT, "partition: sizes exceed the int32 index range");
    if (nwarps == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(srow != nullptr && row_ptrs != nullptr, GKOMI_INVALID_ARGUMENT,
                  "partition: null pointer");
    lb_partition_kernel<<<grid_for(nwarps, kBlock), kBlock, 0, stream_of(ctx)>>>(
        n_rows, row_ptrs, nnz, warp_size, srow, nwarps);
    GKOMI_CHECK_LAUNCH("lb_partition_kernel");
    return GKOMI_SUCCESS;
}

int gkomi_csr_max_row_nnz(const gkomi_ctx* ctx, int64_t n_rows, const int32_t* row_ptrs,
                          int64_t* result)
{
    if (int rc = enter(ctx)) return rc;
    GKOMI_REQUIRE(result != nullptr, GKOMI_INVALID_ARGUMENT, "max_row_nnz: null result");
    // zero the result with a kernel (graph-capture safe), then reduce into it
    if (int rc = gkomi_fill_zero(ctx, result, sizeof(int64_t))) return rc;
    if (n_rows == 0) return GKOMI_SUCCESS;
    max_row_nnz_kernel<<<grid_for(n_rows, kBlock * 8, 4), kBlock, 0, stream_of(ctx)>>>(
        n_rows, row_ptrs, reinterpret_cast<unsigned long long*>(result));
    GKOMI_CHECK_LAUNCH("max_row_nnz_kernel");
    return GKOMI_SUCCESS;
}

int gkomi_csr_col_offsets(const gkomi_ctx* ctx, const gkomi_csr* a, int16_t* out,
                          int32_t* fits)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_csr(a)) return rc;
    GKOMI_REQUIRE(fits != nullptr && (a->nnz == 0 || out != nullptr), GKOMI_INVALID_ARGUMENT,
                  "col_offsets: null output");
    // *fits = 1 by a kernel (graph-capture safe), cleared by rows that overflow
    static_assert(sizeof(int32_t) == sizeof(float), "fill by bit pattern");
    if (int rc = gkomi_fill_f32(ctx, 1, __builtin_bit_cast(float, int32_t{1}),
                                reinterpret_cast<float*>(fits))) {
        return rc;
    }
    if (a->n_rows == 0) return GKOMI_SUCCESS;
    col_offsets_kernel<<<grid_for(a->n_rows, kBlock * 4, 16), kBlock, 0, stream_of(ctx)>>>(
        a->n_rows, a->row_ptrs, a->col_idxs, out, fits);
    GKOMI_CHECK_LAUNCH("col_offsets_kernel");
    return GKOMI_SUCCESS;
}

int gkomi_csr_spmv_f64(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t nrhs, const double* b,
                       int64_t bs, double* c, int64_t cs)
{
    return csr_spmv_impl<double, kPlain>(ctx, nullptr, a, nrhs, b, bs, nullptr, c, cs);
}

int gkomi_csr_spmv_f32(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t nrhs, const float* b,
                       int64_t bs, float* c, int64_t cs)
{
    return csr_spmv_impl<float, kPlain>(ctx, nullptr, a, nrhs, b, bs, nullptr, c, cs);
}

int gkomi_csr_advanced_spmv_f64(const gkomi_ctx* ctx, const double* alpha, const gkomi_csr* a,
                                int64_t nrhs, const double* b, int64_t bs, const double* beta,
                                double* c, int64_t cs)
{
    return csr_spmv_impl<double, kAdvanced>(ctx, alpha, a, nrhs, b, bs, beta, c, cs);
}

int gkomi_csr_advanced_spmv_f32(const gkomi_ctx* ctx, const float* alpha, const gkomi_csr* a,
                                int64_t nrhs, const float* b, int64_t bs, const float* beta,
                                float* c, int64_t cs)
{
    return csr_spmv_impl<float, kAdvanced>(ctx, alpha, a, nrhs, b, bs, beta, c, cs);
}

int gkomi_jacobi_generate_f64(const gkomi_ctx* ctx, const gkomi_csr* a, double* inv_diag)
{
    return jacobi_generate_impl(ctx, a, inv_diag);
}

int gkomi_jacobi_generate_f32(const gkomi_ctx* ctx, const gkomi_csr* a, float* inv_diag)
{
    return jacobi_generate_impl(ctx, a, inv_diag);
}

int gkomi_jacobi_apply_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const double* inv_diag,
                           const double* b, int64_t bs, double* x, int64_t xs)
{
    return jacobi_apply_impl<double, false>(ctx, n, nrhs, inv_diag, nullptr, b, bs, nullptr, x,
                                            xs);
}

int gkomi_jacobi_apply_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const float* inv_diag,
                           const float* b, int64_t bs, float* x, int64_t xs)
{
    return jacobi_apply_impl<float, false>(ctx, n, nrhs, inv_diag, nullptr, b, bs, nullptr, x,
                                           xs);
}

int gkomi_jacobi_advanced_apply_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs,
                                    const double* inv_diag, const double* alpha,
                                    const double* b, int64_t bs, const double* beta, double* x,
                                    int64_t xs)
{
    return jacobi_apply_impl<double, true>(ctx, n, nrhs, inv_diag, alpha, b, bs, beta, x, xs);
}

int gkomi_jacobi_advanced_apply_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs,
                                    const float* inv_diag, const float* alpha, const float* b,
                                    int64_t bs, const float* beta, float* x, int64_t xs)
{
    return jacobi_apply_impl<float, true>(ctx, n, nrhs, inv_diag, alpha, b, bs, beta, x, xs);
}

}  // extern "C"

// ------------------------------------------------ sampled outer product ----
// out[k] = alpha * sum_j L[row(k), j] * X[col(k), j] for every stored entry k
// of the pattern (the gradient of a solve with respect to the matrix values
// is this product with L = adjoint solution, X = forward solution,
// alpha = -1).  One wave per row: the row's L entries are held in registers
// (lane = right-hand side), each stored entry gathers one row of X with a
// coalesced load and reduces across the wave.  nrhs == 1 uses one thread
// per stored entry.
namespace gkomi {

template <typename T>
__global__ void __launch_bounds__(kBlock)
    sampled_outer_rowwave(int64_t n_rows, int64_t nrhs, const int32_t* __restrict__ rp,
                          const int32_t* __restrict__ ci, const T* __restrict__ l, int64_t ls,
                          const T* __restrict__ x, int64_t xs, T alpha, T* __restrict__ out)
{
    const int lane = threadIdx.x % kWave;
    const int64_t wave_id =
        static_cast<int64_t>(xcd_block(blockIdx.x, gridDim.x)) * kWavesPerBlock +
        threadIdx.x / kWave;
    const int64_t num_waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
    for (int64_t row = wave_id; row < n_rows; row += num_waves) {
        const int32_t s = rp[row];
        const int32_t e = rp[row + 1];
        for (int32_t k = s; k < e; ++k) {
            const int64_t c = ci[k];
            T acc = 0;
            for (int64_t j = lane; j < nrhs; j += kWave) {
                acc += l[row * ls + j] * x[c * xs + j];
            }
            acc = wave_sum(acc);
            if (lane == 0) out[k] = alpha * acc;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    sampled_outer_single(int64_t n_rows, const int32_t* __restrict__ rp,
                         const int32_t* __restrict__ ci, const T* __restrict__ l, int64_t ls,
                         const T* __restrict__ x, int64_t xs, T alpha, T* __restrict__ out)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t row = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; row < n_rows;
         row += stride) {
        const T li = alpha * l[row * ls];
        for (int32_t k = rp[row]; k < rp[row + 1]; ++k) {
            out[k] = li * x[static_cast<int64_t>(ci[k]) * xs];
        }
    }
}

template <typename T>
int sampled_outer_impl(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t nrhs, T alpha,
                       const T* l, int64_t ls, const T* x, int64_t xs, T* out)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_csr(a)) return rc;
    GKOMI_REQUIRE(nrhs >= 1, GKOMI_INVALID_ARGUMENT, "sampled outer product: nrhs must be >= 1");
    if (a->nnz == 0 || a->n_rows == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(l && x && out, GKOMI_INVALID_ARGUMENT, "sampled outer product: null pointer");
    GKOMI_REQUIRE(ls >= nrhs && xs >= nrhs, GKOMI_INVALID_ARGUMENT,
                  "sampled outer product: stride smaller than the columns");
    if (nrhs == 1) {
        sampled_outer_single<T><<<grid_for(a->n_rows, kBlock), kBlock, 0, stream_of(ctx)>>>(
            a->n_rows, a->row_ptrs, a->col_idxs, l, ls, x, xs, alpha, out);
    } else {
        sampled_outer_rowwave<T><<<grid_for(a->n_rows, kWavesPerBlock, 32), kBlock, 0,
                                   stream_of(ctx)>>>(a->n_rows, nrhs, a->row_ptrs, a->col_idxs,
                                                     l, ls, x, xs, alpha, out);
    }
    GKOMI_CHECK_LAUNCH("sampled_outer_product");
    return GKOMI_SUCCESS;
}

}  // namespace gkomi

extern "C" {

int gkomi_csr_sampled_outer_product_f64(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t nrhs,
                                        double alpha, const double* l, int64_t ls,
                                        const double* x, int64_t xs, double* out)
{
    return sampled_outer_impl(ctx, a, nrhs, alpha, l, ls, x, xs, out);
}

int gkomi_csr_sampled_outer_product_f32(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t nrhs,
                                        float alpha, const float* l, int64_t ls, const float* x,
                                        int64_t xs, float* out)
{
    return sampled_outer_impl(ctx, a, nrhs, alpha, l, ls, x, xs, out);
}

}  // extern "C"
