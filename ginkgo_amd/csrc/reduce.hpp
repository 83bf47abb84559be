// Deterministic column reductions (dot / squared norm) in two kernels.
//
// Per-thread and per-wave accumulation happens in the value type (FP32 for
// float operands), the per-workgroup partials and the final combine in FP64.
// A second one-workgroup kernel adds the partials in workgroup order, so
// results do not depend on scheduling.  No device-scope fence is used: on
// gfx950 a release fence writes back the XCD's L2, and thousands of them
// cost more than the reduction; the kernel boundary orders the two steps.
#pragma once

#include <algorithm>

#include "common.hpp"

namespace gkomi {

// bytes at the start of a reduction workspace that are not used any more
// (kept so that workspace sizes stay as documented)
constexpr size_t kTicketBytes = 256;
constexpr int kRedMaxBlocks = 2048;

template <typename T, bool Norm>
__global__ void __launch_bounds__(kBlock)
    reduce_one_column_kernel(int64_t n, const T* __restrict__ x, int64_t xs,
                             const T* __restrict__ y, int64_t ys, double* __restrict__ partials)
{
    __shared__ T lds_t[kWavesPerBlock];
    T acc = 0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += stride) {
        const T a = x[i * xs];
        acc += a * (Norm ? a : y[i * ys]);
    }
    const T bsum = block_sum(acc, lds_t);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = static_cast<double>(bsum);
    }
}

template <typename T, bool Norm>
__global__ void __launch_bounds__(kBlock)
    reduce_columns_kernel(int64_t n, int64_t nrhs, const T* __restrict__ x, int64_t xs,
                          const T* __restrict__ y, int64_t ys, double* __restrict__ partials)
{
    __shared__ double lds[kWavesPerBlock][kWave];
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int64_t rstride = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
    for (int64_t c0 = 0; c0 < nrhs; c0 += kWave) {
        const int64_t c = c0 + lane;
        T acc = 0;
        if (c < nrhs) {
            for (int64_t i = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wave; i < n;
                 i += rstride) {
                const T a = x[i * xs + c];
                acc += a * (Norm ? a : y[i * ys + c]);
            }
        }
        lds[wave][lane] = static_cast<double>(acc);
        __syncthreads();
        if (wave == 0 && c < nrhs) {
            double s = 0;
#pragma unroll
            for (int w = 0; w < kWavesPerBlock; ++w) s += lds[w][lane];
            partials[static_cast<int64_t>(blockIdx.x) * nrhs + c] = s;
        }
        __syncthreads();
    }
}

// Column sums of partials P[g][C] (g < blocks) spread over workgroups of
// kColSumThreads threads: workgroup b sums the columns [8 b, 8 b + 8), its
// thread t column 8 b + t % 8 over the rows g = t / 8 + 32 (4 m + u) into
// accumulator u, (a0 + a1) + (a2 + a3), then the 32 row groups in order
// through LDS: a fixed order.  fn(c, sum) runs once per column c < C.  One
// workgroup reading all the partials is bound by that CU's bandwidth (the
// partials come from every XCD: 1024 x 64 took 42 us); here each workgroup
// reads 1/(C/8) of them.  Every thread of the workgroup must call it.
constexpr int kColSumThreads = 256;
constexpr int kColSumCols = 8;

template <typename Fn>
__device__ __forceinline__ void column_sums_block(const double* __restrict__ P, int64_t blocks,
                                                  int64_t C, Fn fn)
{
    constexpr int kRowGroups = kColSumThreads / kColSumCols;
    __shared__ double lds[kColSumThreads];
    const int cl = threadIdx.x % kColSumCols;
    const int g0 = threadIdx.x / kColSumCols;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * kColSumCols + cl;
    double acc = 0;
    if (c < C) {
        double a[4] = {0, 0, 0, 0};
        for (int64_t g = g0; g < blocks; g += 4 * kRowGroups) {
            // all four loads issued before the adds; a slot past the last
            // row rereads row g and adds +0 (the accumulators start at +0,
            // so they never hold -0 and +0 leaves them be)
            double t[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t gg = g + u * kRowGroups;
                t[u] = P[(gg < blocks ? gg : g) * C + c];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] += g + u * kRowGroups < blocks ? t[u] : 0.0;
        }
        acc = (a[0] + a[1]) + (a[2] + a[3]);
    }
    lds[threadIdx.x] = acc;
    __syncthreads();
    if (g0 == 0 && c < C) {
        double sum = lds[cl];
        for (int r = 1; r < kRowGroups; ++r) sum += lds[r * kColSumCols + cl];
        fn(c, sum);
    }
}

inline unsigned colsum_grid(int64_t C)
{
    return static_cast<unsigned>(std::max<int64_t>(1, ceildiv(C, kColSumCols)));
}

// result[c] = sum of column c of the partials (sqrt for norms), several
// columns: column_sums_block.
template <typename T, bool Norm>
__global__ void __launch_bounds__(kColSumThreads)
    reduce_colsum_kernel(int64_t blocks, int64_t nrhs, const double* __restrict__ partials,
                         T* __restrict__ result)
{
    column_sums_block(partials, blocks, nrhs, [&](int64_t c, double s) {
        result[c] = static_cast<T>(Norm ? sqrt(s) : s);
    });
}

// One workgroup: result[c] = sum over `blocks` partials of column c, in
// workgroup order per wave, then the waves in order (sqrt for norms); the
// single-column reductions.
template <typename T, bool Norm>
__global__ void __launch_bounds__(kBlock)
    reduce_combine_kernel(int64_t blocks, int64_t nrhs, const double* __restrict__ partials,
                          T* __restrict__ result)
{
    __shared__ double lds[kWavesPerBlock][kWave];
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    if (nrhs == 1) {
        double acc = 0;
        for (int64_t k = threadIdx.x; k < blocks; k += kBlock) acc += partials[k];
        acc = wave_sum(acc);
        if (lane == 0) lds[wave][0] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = lds[0][0];
            for (int w = 1; w < kWavesPerBlock; ++w) s += lds[w][0];
            result[0] = static_cast<T>(Norm ? sqrt(s) : s);
        }
        return;
    }
    for (int64_t c0 = 0; c0 < nrhs; c0 += kWave) {
        const int64_t c = c0 + lane;
        double acc = 0;
        if (c < nrhs) {
            for (int64_t k = wave; k < blocks; k += kWavesPerBlock) acc += partials[k * nrhs + c];
        }
        lds[wave][lane] = acc;
        __syncthreads();
        if (wave == 0 && c < nrhs) {
            double s = 0;
#pragma unroll
            for (int w = 0; w < kWavesPerBlock; ++w) s += lds[w][lane];
            result[c] = static_cast<T>(Norm ? sqrt(s) : s);
        }
        __syncthreads();
    }
}

template <typename T, bool Norm>
int launch_column_reduction(hipStream_t stream, int64_t n, int64_t nrhs, const T* x,
                            int64_t xs, const T* y, int64_t ys, T* result, double* partials)
{
    int64_t g;
    if (nrhs == 1) {
        g = ceildiv(n, kBlock * 8);
        g = std::max<int64_t>(1, std::min<int64_t>(g, std::min<int64_t>(kRedMaxBlocks,
                                                                         num_cus() * 4)));
        reduce_one_column_kernel<T, Norm><<<static_cast<unsigned>(g), kBlock, 0, stream>>>(
            n, x, xs, y, ys, partials);
    } else {
        g = ceildiv(n, kWavesPerBlock * 16);
        g = std::max<int64_t>(1, std::min<int64_t>(g, std::min<int64_t>(kRedMaxBlocks,
                                                                         num_cus() * 4)));
        reduce_columns_kernel<T, Norm><<<static_cast<unsigned>(g), kBlock, 0, stream>>>(
            n, nrhs, x, xs, y, ys, partials);
    }
    GKOMI_CHECK_LAUNCH("column reduction");
    if (nrhs == 1) {
        reduce_combine_kernel<T, Norm><<<1, kBlock, 0, stream>>>(g, nrhs, partials, result);
    } else {
        reduce_colsum_kernel<T, Norm><<<colsum_grid(nrhs), kColSumThreads, 0, stream>>>(
            g, nrhs, partials, result);
    }
    GKOMI_CHECK_LAUNCH("column reduction combine");
    return GKOMI_SUCCESS;
}

}  // namespace gkomi
