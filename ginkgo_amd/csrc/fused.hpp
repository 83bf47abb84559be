// Pieces shared by the fused Krylov iterations (CG, BiCGSTAB): FP64
// per-workgroup partial sums and their fixed-order combine, launch shapes
// and the criteria validation.
//
// Cross-workgroup sums never use a device-scope fence (on gfx950 a release
// fence writes back the XCD's whole L2, which at thousands of workgroups
// costs more than an iteration): each producer kernel writes one FP64
// partial per workgroup, and every workgroup of the consuming kernel adds
// all partials in the same fixed order, so all workgroups reach
// bit-identical scalars and the same stop decision.  The kernel boundary
// orders producer and consumer.
#pragma once

#include <algorithm>

#include "common.hpp"
#include "reduce.hpp"

namespace gkomi {

// Most workgroups of a fused SpMV kernel (the partial arrays hold this many).
constexpr int kCgMaxGrid = 2048;

// Workgroup sums of Count values in the value type, converted to FP64 and
// stored as this workgroup's partials (NT threads per workgroup).
template <typename T, int Count, int NT = kBlock>
__device__ __forceinline__ void store_partials(const T (&vals)[Count], double* partials)
{
    constexpr int kW = NT / kWave;
    __shared__ T lds[Count][kW];
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int c = 0; c < Count; ++c) {
        const T s = wave_sum(vals[c]);
        if (lane == 0) lds[c][wave] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < Count; ++c) {
            T s = lds[c][0];
            for (int w = 1; w < kW; ++w) s += lds[c][w];
            partials[static_cast<int64_t>(blockIdx.x) * Count + c] = static_cast<double>(s);
        }
    }
}

// Sum of `count` workgroup partials (Count values each) in a fixed order:
// thread t adds partials t, t + NT, ...; then the wave tree and the waves in
// order.  Every workgroup obtains the same bits.
template <int Count, int NT = kBlock>
__device__ __forceinline__ void combine_partials(const double* __restrict__ partials, int count,
                                                 double (&out)[Count])
{
    constexpr int kW = NT / kWave;
    __shared__ double lds[Count][kW];
    double acc[Count];
#pragma unroll
    for (int c = 0; c < Count; ++c) acc[c] = 0;
    for (int k = threadIdx.x; k < count; k += NT) {
#pragma unroll
        for (int c = 0; c < Count; ++c) acc[c] += partials[static_cast<int64_t>(k) * Count + c];
    }
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int c = 0; c < Count; ++c) {
        const double s = wave_sum(acc[c]);
        if (lane == 0) lds[c][wave] = s;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < Count; ++c) {
        double s = lds[c][0];
        for (int w = 1; w < kW; ++w) s += lds[c][w];
        out[c] = s;
    }
    __syncthreads();
}

// combine_partials that also hands one int from thread 0 to the whole
// workgroup through the same barrier: a kernel reads its stop flag in
// thread 0 at its start and learns the workgroup's decision together with
// the sums, instead of paying a barrier and a memory round trip for each.
template <int Count, int NT = kBlock>
__device__ __forceinline__ int combine_partials_bcast(const double* __restrict__ partials,
                                                      int count, double (&out)[Count],
                                                      int value0)
{
    constexpr int kW = NT / kWave;
    __shared__ double lds[Count][kW];
    __shared__ int bc;
    double acc[Count];
#pragma unroll
    for (int c = 0; c < Count; ++c) acc[c] = 0;
    for (int k = threadIdx.x; k < count; k += NT) {
#pragma unroll
        for (int c = 0; c < Count; ++c) acc[c] += partials[static_cast<int64_t>(k) * Count + c];
    }
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
#pragma unroll
    for (int c = 0; c < Count; ++c) {
        const double s = wave_sum(acc[c]);
        if (lane == 0) lds[c][wave] = s;
    }
    if (threadIdx.x == 0) bc = value0;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < Count; ++c) {
        double s = lds[c][0];
        for (int w = 1; w < kW; ++w) s += lds[c][w];
        out[c] = s;
    }
    const int v = bc;
    __syncthreads();
    return v;
}

// The elementwise kernels (update, split p) run 1024-thread workgroups on a
// grid of at most kCgUpdGrid: few partials to combine in their consumers,
// full occupancy for the streaming work.
constexpr int kUpdBlock = 1024;
constexpr int kCgUpdGrid = 512;

inline int fused_grid(int64_t tiles)
{
    int64_t g = std::min<int64_t>(tiles, static_cast<int64_t>(num_cus()) * 8);
    g = std::min<int64_t>(g, kCgMaxGrid);
    return static_cast<int>(std::max<int64_t>(g, 1));
}

inline int check_crit(const gkomi_criteria* crit)
{
    GKOMI_REQUIRE(crit != nullptr && crit->num >= 1 && crit->num <= GKOMI_MAX_CRITERIA,
                  GKOMI_NOT_SUPPORTED, "criteria: between 1 and 4 criteria are supported");
    for (int k = 0; k < crit->num; ++k) {
        const int kind = crit->c[k].kind;
        GKOMI_REQUIRE(kind == GKOMI_CRIT_ITERATION || kind == GKOMI_CRIT_RESIDUAL_NORM ||
                          kind == GKOMI_CRIT_IMPLICIT_RESNORM,
                      GKOMI_NOT_SUPPORTED, "criteria: unknown kind");
        GKOMI_REQUIRE(kind == GKOMI_CRIT_ITERATION ||
                          (crit->c[k].baseline >= 0 && crit->c[k].baseline <= 2),
                      GKOMI_NOT_SUPPORTED, "criteria: unknown baseline");
    }
    return GKOMI_SUCCESS;
}

}  // namespace gkomi
