// Deterministic single-pass column reductions (dot / squared norm).
//
// Per-thread and per-wave accumulation happens in the value type (FP32 for
// float operands), the per-workgroup partials and the final combine in FP64.
// The last workgroup to finish (ticket counter) adds the partials in
// workgroup order, so results do not depend on scheduling.
#pragma once

#include "common.hpp"

namespace gkomi {

constexpr size_t kTicketBytes = 256;
constexpr int kRedMaxBlocks = 2048;

template <typename T, bool Norm>
__global__ void __launch_bounds__(kBlock)
    reduce_one_column_kernel(int64_t n, const T* __restrict__ x, int64_t xs,
                             const T* __restrict__ y, int64_t ys, T* __restrict__ result,
                             unsigned* __restrict__ ticket, double* __restrict__ partials)
{
    __shared__ T lds_t[kWavesPerBlock];
    __shared__ double lds_d[kWavesPerBlock];
    __shared__ int flag;
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
    if (last_block_ticket(ticket, &flag)) {
        double a = 0;
        for (unsigned k = threadIdx.x; k < gridDim.x; k += kBlock) {
            a += partials[k];
        }
        const double total = block_sum(a, lds_d);
        if (threadIdx.x == 0) {
            result[0] = static_cast<T>(Norm ? sqrt(total) : total);
            *ticket = 0;
        }
    }
}

template <typename T, bool Norm>
__global__ void __launch_bounds__(kBlock)
    reduce_columns_kernel(int64_t n, int64_t nrhs, const T* __restrict__ x, int64_t xs,
                          const T* __restrict__ y, int64_t ys, T* __restrict__ result,
                          unsigned* __restrict__ ticket, double* __restrict__ partials)
{
    __shared__ double lds[kWavesPerBlock][kWave];
    __shared__ int flag;
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
    if (last_block_ticket(ticket, &flag)) {
        for (int64_t c0 = 0; c0 < nrhs; c0 += kWave) {
            const int64_t c = c0 + lane;
            double acc = 0;
            if (c < nrhs) {
                for (unsigned k = wave; k < gridDim.x; k += kWavesPerBlock) {
                    acc += partials[static_cast<int64_t>(k) * nrhs + c];
                }
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
        if (threadIdx.x == 0) {
            *ticket = 0;
        }
    }
}

template <typename T, bool Norm>
int launch_column_reduction(hipStream_t stream, int64_t n, int64_t nrhs, const T* x,
                            int64_t xs, const T* y, int64_t ys, T* result, unsigned* ticket,
                            double* partials)
{
    if (nrhs == 1) {
        int64_t g = ceildiv(n, kBlock * 8);
        g = std::max<int64_t>(1, std::min<int64_t>(g, std::min<int64_t>(kRedMaxBlocks,
                                                                         num_cus() * 4)));
        reduce_one_column_kernel<T, Norm><<<static_cast<unsigned>(g), kBlock, 0, stream>>>(
            n, x, xs, y, ys, result, ticket, partials);
    } else {
        int64_t g = ceildiv(n, kWavesPerBlock * 16);
        g = std::max<int64_t>(1, std::min<int64_t>(g, std::min<int64_t>(kRedMaxBlocks,
                                                                         num_cus() * 4)));
        reduce_columns_kernel<T, Norm><<<static_cast<unsigned>(g), kBlock, 0, stream>>>(
            n, nrhs, x, xs, y, ys, result, ticket, partials);
    }
    GKOMI_CHECK_LAUNCH("column reduction");
    return GKOMI_SUCCESS;
}

}  // namespace gkomi
