// Device-wide exclusive prefix sum of integers (shared by the format
// conversions, the transpose and the CSR column panels).
#pragma once

#include <algorithm>

#include "common.hpp"

namespace gkomi {

// ------------------------------------------------------- exclusive scan ----
// out[0..n] = exclusive prefix sum of in[0..n) with out[n] = total.
// Three phases: per-tile sums, scan of tile sums in one workgroup, per-tile
// scan plus offset.  Integer only, hence exact.
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

template <typename I>
__device__ I block_exclusive_scan(I v, I* lds, I* total)
{
    // Hillis-Steele over the workgroup in LDS (integers)
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off *= 2) {
        I t = threadIdx.x >= off ? lds[threadIdx.x - off] : I(0);
        __syncthreads();
        lds[threadIdx.x] += t;
        __syncthreads();
    }
    const I incl = lds[threadIdx.x];
    *total = lds[kBlock - 1];
    __syncthreads();
    return incl - v;
}

template <typename I>
__global__ void __launch_bounds__(kBlock)
    scan_tile_sums(int64_t n, const I* __restrict__ in, I* __restrict__ tile_sums)
{
    __shared__ I lds[kBlock];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
    I acc = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + static_cast<int64_t>(k) * kBlock + threadIdx.x;
        if (i < n) acc += in[i];
    }
    I total;
    block_exclusive_scan(acc, lds, &total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

template <typename I>
__global__ void __launch_bounds__(kBlock) scan_tile_offsets(int64_t num_tiles, I* __restrict__ tile_sums)
{
    __shared__ I lds[kBlock];
    I carry = 0;
    for (int64_t t0 = 0; t0 < num_tiles; t0 += kBlock) {
        const int64_t t = t0 + threadIdx.x;
        const I v = t < num_tiles ? tile_sums[t] : I(0);
        I total;
        const I ex = block_exclusive_scan(v, lds, &total);
        if (t < num_tiles) tile_sums[t] = carry + ex;
        carry += total;
    }
}

template <typename I>
__global__ void __launch_bounds__(kBlock)
    scan_apply(int64_t n, const I* __restrict__ in, const I* __restrict__ tile_offsets,
               I* __restrict__ out)
{
    __shared__ I lds[kBlock];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
    // each thread owns kScanItems consecutive elements
    const int64_t first = base + static_cast<int64_t>(threadIdx.x) * kScanItems;
    I vals[kScanItems];
    I local = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = first + k;
        vals[k] = i < n ? in[i] : I(0);
        local += vals[k];
    }
    I total;
    I run = block_exclusive_scan(local, lds, &total) + tile_offsets[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = first + k;
        if (i < n) out[i] = run;
        run += vals[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) {
        out[n] = run;
    }
}

inline size_t scan_ws_bytes(int64_t n, size_t elem)
{
    return elem * static_cast<size_t>(std::max<int64_t>(1, ceildiv(n, kScanTile))) + 256;
}

// `in` and `out` may alias (in-place scan): phase 1 reads every input before
// phase 3 starts, and each phase-3 tile reads only its own inputs before
// writing them.
template <typename I>
int exclusive_scan(hipStream_t stream, int64_t n, const I* in, I* out, void* ws)
{
    const int64_t tiles = std::max<int64_t>(1, ceildiv(n, kScanTile));
    I* tile = static_cast<I*>(ws);
    scan_tile_sums<I><<<static_cast<unsigned>(tiles), kBlock, 0, stream>>>(n, in, tile);
    scan_tile_offsets<I><<<1, kBlock, 0, stream>>>(tiles, tile);
    scan_apply<I><<<static_cast<unsigned>(tiles), kBlock, 0, stream>>>(n, in, tile, out);
    GKOMI_CHECK_LAUNCH("exclusive_scan");
    return GKOMI_SUCCESS;
}

}  // namespace gkomi
