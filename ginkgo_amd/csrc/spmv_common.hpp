// Constants and the result store shared by the SpMV kernels.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace gkomi {

// Load-balanced CSR: nominal nnz per workgroup of the partition, and the LDS
// capacity (products and row pointers) of the staged path.
constexpr int64_t kLbChunk = 1024;
constexpr int kLbLdsNnz = 2048;
constexpr int kLbLdsRows = 1024;
// rows up to this length are reduced by one thread from LDS, longer by a wave
constexpr int kLbShortRow = 32;

// Epilogue of a product row:
//   kPlain      c = sum;
//   kAdvanced   c = alpha * sum + beta * c, beta == 0 ignoring the old c
//               (reference/matrix/csr_kernels.cpp:96-117: NaN in c does not
//               propagate);
//   kAccumulate c += alpha * sum, alpha == NULL meaning 1 (the apply2 of a
//               matrix summed into another product, reference/matrix/
//               coo_kernels.cpp:61-101).
enum spmv_mode : int { kPlain = 0, kAdvanced = 1, kAccumulate = 2 };

// Mixed value types (the reference's arithmetic_type rule,
// reference/matrix/csr_kernels.cpp:47-121): `sum` is in the highest
// precision Acc of the operands, alpha has the matrix's type, beta and the
// destination the output's; the result is rounded to the output type once.
// With one type throughout this is the plain T arithmetic.
template <int Mode, typename Acc, typename TC, typename TS>
__device__ __forceinline__ void store_result(TC* dst, Acc sum, const TS* alpha, const TC* beta)
{
    if constexpr (Mode == kAdvanced) {
        const Acc bt = static_cast<Acc>(beta[0]);
        const Acc r = static_cast<Acc>(alpha[0]) * sum;
        *dst = static_cast<TC>((bt == Acc(0)) ? r : r + bt * static_cast<Acc>(*dst));
    } else if constexpr (Mode == kAccumulate) {
        *dst = static_cast<TC>(static_cast<Acc>(*dst) +
                               (alpha ? static_cast<Acc>(alpha[0]) * sum : sum));
    } else {
        *dst = static_cast<TC>(sum);
    }
}

// The arithmetic type of a product over matrix values TA, input TB and
// output TC: double when any of them is double.
template <typename TA, typename TB, typename TC>
using acc_t = std::conditional_t<std::is_same<TA, double>::value ||
                                     std::is_same<TB, double>::value ||
                                     std::is_same<TC, double>::value,
                                 double, float>;

// Rows per workgroup tile of the classical (sub-wave) kernels.
template <int SW, int RPG>
constexpr int64_t subwave_tile_rows = int64_t{kWave / SW} * RPG * kWavesPerBlock;

// One workgroup tile of the classical CSR product: rows
// [tile * subwave_tile_rows, ...).  Each group of SW lanes owns RPG rows at
// once (rows g, g + 64/SW, ... of the wave's contiguous range of 64/SW * RPG
// rows).  The body is branch-free: lanes past the end of a row load entry 0
// (valid whenever a row is active) and add a zero product, so the compiler
// issues the loads of all RPG rows back to back and waits once -- RPG times
// the bytes in flight per wave, which is what a short-row SpMV needs to cover
// HBM latency.  Per row the summation order is fixed (lane l adds entries l,
// l + SW, ... by fused multiply-add; then the group tree), independent of
// RPG and of the launch.  `gather(col)` supplies the vector entry of a
// column; `row_op(row, sum)` runs in lane 0 of the group for every row below
// n_rows.  With Compact the columns come from the 2-byte offsets `co`
// (column = row + co[p]) instead of `ci`.
template <typename T, int SW, int RPG, bool Compact = false, typename Gather, typename RowOp,
          typename TV>
__device__ __forceinline__ void subwave_tile(int64_t n_rows, int64_t tile,
                                             const int32_t* __restrict__ rp,
                                             const int32_t* __restrict__ ci,
                                             const int16_t* __restrict__ co,
                                             const TV* __restrict__ v, Gather gather,
                                             RowOp row_op)
{
    constexpr int kGroups = kWave / SW;
    constexpr int kRowsPerWave = kGroups * RPG;
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int g = lane / SW;
    const int l = lane % SW;
    const int64_t row0 = tile * subwave_tile_rows<SW, RPG> + wave * kRowsPerWave + g;
    int32_t pos[RPG];
    int32_t end[RPG];
    T sum[RPG];
    int32_t iters = 0;
    // rows past the end read rp[n_rows] twice: empty
#pragma unroll
    for (int k = 0; k < RPG; ++k) {
        const int64_t row = row0 + k * kGroups;
        pos[k] = rp[row < n_rows ? row : n_rows];
        end[k] = rp[row < n_rows ? row + 1 : n_rows];
    }
#pragma unroll
    for (int k = 0; k < RPG; ++k) {
        pos[k] += l;
        sum[k] = 0;
        const int32_t len = end[k] - pos[k];
        const int32_t it = len > 0 ? (len + SW - 1) / SW : 0;
        iters = it > iters ? it : iters;
    }
    for (int32_t j = 0; j < iters; ++j) {
        int32_t col[RPG];
        T val[RPG];
        bool act[RPG];
#pragma unroll
        for (int k = 0; k < RPG; ++k) {
            act[k] = pos[k] < end[k];
            const int32_t p = act[k] ? pos[k] : 0;
            if constexpr (Compact) {
                // an inactive slot gathers column 0 (always valid)
                const int32_t c = static_cast<int32_t>(row0 + k * kGroups) + co[p];
                col[k] = act[k] ? c : 0;
            } else {
                col[k] = ci[p];
            }
            val[k] = static_cast<T>(v[p]);
            pos[k] = act[k] ? pos[k] + SW : pos[k];
        }
        T xv[RPG];
#pragma unroll
        for (int k = 0; k < RPG; ++k) xv[k] = gather(col[k]);
#pragma unroll
        for (int k = 0; k < RPG; ++k) {
            sum[k] = fma(act[k] ? val[k] : T(0), act[k] ? xv[k] : T(0), sum[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < RPG; ++k) {
        sum[k] = group_sum<SW>(sum[k]);
        const int64_t row = row0 + k * kGroups;
        if (row < n_rows && l == 0) row_op(row, sum[k]);
    }
}

// One window of the staged CSR product, run by one wave: the R = 64 / G
// consecutive rows [R w, R w + R).  The window's entries are copied into the
// wave's LDS staging area (lv, lc: C values and C column indices, or with
// Compact the 2-byte offsets column - row) in runs of 64 consecutive entries
// (every load one contiguous run, all issued before the first LDS write).
// Row r of the window belongs to lanes r, r + R, ..., r + (G - 1) R: its
// sub-lane s (lane r + s R) adds the row's entries s, s + G, ... from LDS by
// fused multiply-add with gather(col), and the strided butterfly adds the
// sub-lanes: the per-row order of subwave_tile with SW = G, bit for bit.
// Lanes of one sub-lane index hold R consecutive rows, so each gather
// instruction reads R consecutive rows at G stencil offsets (R contiguous
// values per offset) rather than 64 / G rows at G interleaved offsets.  A
// window of more than C entries is staged in chunks; a lane's entries
// continue in the next chunk in the same order.  `row_op(row, sum)` runs in
// sub-lane 0 (lanes 0 .. R - 1) for every row below n_rows.  Only the wave
// itself reads its staging area: LDS accesses of one wave complete in
// order, and the wave barriers keep the compiler from moving them across
// the staging.
// Copies the entries [c0, c1) (c0 < c1 <= nnz, c1 - c0 <= C) into the
// wave's staging area: every load one contiguous run of 64 entries, all
// issued before the first LDS write.
template <int C, bool Compact, typename Idx, typename TV>
__device__ __forceinline__ void staged_copy(int32_t c0, int32_t c1,
                                            const int32_t* __restrict__ ci,
                                            const int16_t* __restrict__ co,
                                            const TV* __restrict__ v, TV* lv, Idx* lc)
{
    constexpr int kSteps = C / kWave;
    const int lane = threadIdx.x % kWave;
    TV tv[kSteps];
    Idx tc[kSteps];
#pragma unroll
    for (int u = 0; u < kSteps; ++u) {
        const int32_t e = c0 + u * kWave + lane;
        const int32_t ee = e < c1 ? e : c0;  // c0 < c1 <= nnz: a valid entry
        tv[u] = v[ee];
        if constexpr (Compact) {
            tc[u] = co[ee];
        } else {
            tc[u] = ci[ee];
        }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < kSteps; ++u) {
        lv[u * kWave + lane] = tv[u];
        lc[u * kWave + lane] = tc[u];
    }
    __builtin_amdgcn_wave_barrier();
}

// Adds the lane's entries k, k + G, ... below z (the row's end or the
// chunk's end) from the chunk staged at c0 to `sum`, in order, and advances
// k past them (at most to the first one not taken, which may lie in the
// next chunk; k never falls below c0: earlier chunks took those).
template <typename T, int G, bool Compact, typename Idx, typename TV, typename Gather>
__device__ __forceinline__ void staged_accumulate(int64_t row, int32_t c0, int32_t z,
                                                  int32_t& k, T& sum, const TV* lv,
                                                  const Idx* lc, Gather gather)
{
    constexpr int U = 4;
    while (__ballot(k < z) != 0) {
        int32_t col[U];
        T val[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int32_t kk = k + u * G;
            ok[u] = kk < z;
            const int32_t idx = ok[u] ? kk - c0 : 0;
            col[u] = Compact ? static_cast<int32_t>(row) + lc[idx] : lc[idx];
            val[u] = static_cast<T>(lv[idx]);
        }
        T xv[U];
        // a slot past the row's end gathers column 0 (always valid)
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = gather(ok[u] ? col[u] : 0);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            sum = fma(ok[u] ? val[u] : T(0), ok[u] ? xv[u] : T(0), sum);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) k = ok[u] ? k + G : k;
    }
}

template <typename T, int G, int C, bool Compact, typename Idx, typename Gather,
          typename RowOp, typename TV>
__device__ __forceinline__ void staged_window(int64_t n_rows, int64_t w,
                                              const int32_t* __restrict__ rp,
                                              const int32_t* __restrict__ ci,
                                              const int16_t* __restrict__ co,
                                              const TV* __restrict__ v, TV* lv, Idx* lc,
                                              Gather gather, RowOp row_op)
{
    constexpr int kRows = kWave / G;
    const int lane = threadIdx.x % kWave;
    const int l = lane / kRows;  // sub-lane
    const int64_t row = w * kRows + lane % kRows;
    const int32_t rs = rp[row < n_rows ? row : n_rows];
    const int32_t re = rp[row < n_rows ? row + 1 : n_rows];
    const int32_t ws = __shfl(rs, 0, kWave);
    const int32_t we = __shfl(re, kWave - 1, kWave);
    T sum = 0;
    int32_t k = rs + l;  // the lane's next entry (absolute position)
    for (int32_t c0 = ws; c0 < we; c0 += C) {
        const int32_t c1 = c0 + C < we ? c0 + C : we;
        staged_copy<C, Compact>(c0, c1, ci, co, v, lv, lc);
        staged_accumulate<T, G, Compact>(row, c0, re < c1 ? re : c1, k, sum, lv, lc, gather);
        __builtin_amdgcn_wave_barrier();
    }
    sum = strided_sum<G>(sum);
    if (row < n_rows && l == 0) row_op(row, sum);
}

// staged_window split in two around a wait for other data (the partial sums
// of a fused solver), for a wave's first window.  staged_prefetch loads the
// window's row pointers and, when the window fits one chunk, stages it and
// gathers both operands of the lane's first UB entries with load_a(col) and
// load_b(col); staged_finish adds fin(a, b) for those entries and then the
// rest as staged_window does, in the same order: fin(load_a(c), load_b(c))
// must equal gather(c), and the sums have the bits of staged_window's.
template <typename T, int UB>
struct staged_pf {
    int64_t row;
    int32_t re, ws, we, k;
    bool one;
    bool ok[UB];
    T val[UB], a[UB], b[UB];
};

template <typename T, int G, int C, bool Compact, int UB, typename Idx, typename TV,
          typename LoadA, typename LoadB>
__device__ __forceinline__ void staged_prefetch(int64_t n_rows, int64_t w,
                                                const int32_t* __restrict__ rp,
                                                const int32_t* __restrict__ ci,
                                                const int16_t* __restrict__ co,
                                                const TV* __restrict__ v, TV* lv, Idx* lc,
                                                LoadA load_a, LoadB load_b, staged_pf<T, UB>& pf)
{
    constexpr int kRows = kWave / G;
    const int lane = threadIdx.x % kWave;
    const int l = lane / kRows;
    pf.row = w * kRows + lane % kRows;
    const int32_t rs = rp[pf.row < n_rows ? pf.row : n_rows];
    pf.re = rp[pf.row < n_rows ? pf.row + 1 : n_rows];
    pf.ws = __shfl(rs, 0, kWave);
    pf.we = __shfl(pf.re, kWave - 1, kWave);
    pf.k = rs + l;
    pf.one = pf.we - pf.ws <= C;
#pragma unroll
    for (int u = 0; u < UB; ++u) {
        pf.ok[u] = false;
        pf.val[u] = pf.a[u] = pf.b[u] = T(0);
    }
    if (pf.one && pf.ws < pf.we) {
        staged_copy<C, Compact>(pf.ws, pf.we, ci, co, v, lv, lc);
        int32_t col[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int32_t kk = pf.k + u * G;
            pf.ok[u] = kk < pf.re;
            const int32_t idx = pf.ok[u] ? kk - pf.ws : 0;
            col[u] = Compact ? static_cast<int32_t>(pf.row) + lc[idx] : lc[idx];
            pf.val[u] = static_cast<T>(lv[idx]);
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int32_t c = pf.ok[u] ? col[u] : 0;
            pf.a[u] = load_a(c);
            pf.b[u] = load_b(c);
        }
    }
}

template <typename T, int G, int C, bool Compact, int UB, typename Idx, typename TV,
          typename Fin, typename Gather, typename RowOp>
__device__ __forceinline__ void staged_finish(int64_t n_rows, const int32_t* __restrict__ ci,
                                              const int16_t* __restrict__ co,
                                              const TV* __restrict__ v, TV* lv, Idx* lc,
                                              staged_pf<T, UB>& pf, Fin fin, Gather gather,
                                              RowOp row_op)
{
    constexpr int kRows = kWave / G;
    const int l = (threadIdx.x % kWave) / kRows;
    T sum = 0;
    int32_t k = pf.k;
    if (pf.one) {
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            sum = fma(pf.ok[u] ? pf.val[u] : T(0), pf.ok[u] ? fin(pf.a[u], pf.b[u]) : T(0), sum);
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) k = pf.ok[u] ? k + G : k;
        // entries past the first UB, from the staged window
        staged_accumulate<T, G, Compact>(pf.row, pf.ws, pf.re, k, sum, lv, lc, gather);
        __builtin_amdgcn_wave_barrier();
    } else {
        for (int32_t c0 = pf.ws; c0 < pf.we; c0 += C) {
            const int32_t c1 = c0 + C < pf.we ? c0 + C : pf.we;
            staged_copy<C, Compact>(c0, c1, ci, co, v, lv, lc);
            staged_accumulate<T, G, Compact>(pf.row, c0, pf.re < c1 ? pf.re : c1, k, sum, lv, lc,
                                             gather);
            __builtin_amdgcn_wave_barrier();
        }
    }
    sum = strided_sum<G>(sum);
    if (pf.row < n_rows && l == 0) row_op(pf.row, sum);
}

// Rows of one workgroup tile of the staged product (one window per wave).
template <int G>
constexpr int64_t staged_tile_rows = int64_t{kWave / G} * kWavesPerBlock;

// subwave_tile split in two, for kernels that have other loads to wait for
// first (the partial sums of a fused solver): subwave_prefetch issues the
// tile's row-pointer loads and the first step's column and value loads;
// subwave_finish gathers, sums and calls row_op as subwave_tile does, in the
// same order (a row's sum has the same bits either way).
template <typename T, int RPG>
struct subwave_pf {
    int64_t row0;
    int32_t pos[RPG];
    int32_t end[RPG];
    int32_t col[RPG];
    T val[RPG];
    bool act[RPG];
    int32_t iters;
};

template <typename T, int SW, int RPG, bool Compact>
__device__ __forceinline__ void subwave_prefetch(int64_t n_rows, int64_t tile,
                                                 const int32_t* __restrict__ rp,
                                                 const int32_t* __restrict__ ci,
                                                 const int16_t* __restrict__ co,
                                                 const T* __restrict__ v, subwave_pf<T, RPG>& pf)
{
    constexpr int kGroups = kWave / SW;
    constexpr int kRowsPerWave = kGroups * RPG;
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    const int g = lane / SW;
    const int l = lane % SW;
    pf.row0 = tile * subwave_tile_rows<SW, RPG> + wave * kRowsPerWave + g;
    pf.iters = 0;
#pragma unroll
    for (int k = 0; k < RPG; ++k) {
        const int64_t row = pf.row0 + k * kGroups;
        pf.pos[k] = rp[row < n_rows ? row : n_rows];
        pf.end[k] = rp[row < n_rows ? row + 1 : n_rows];
    }
#pragma unroll
    for (int k = 0; k < RPG; ++k) {
        pf.pos[k] += l;
        const int32_t len = pf.end[k] - pf.pos[k];
        const int32_t it = len > 0 ? (len + SW - 1) / SW : 0;
        pf.iters = it > pf.iters ? it : pf.iters;
    }
    // a lane with no entry in its rows loads nothing (the matrix may have no
    // entries at all); the others read entry 0 for their inactive slots, as
    // subwave_tile does
    const bool any = pf.iters > 0;
#pragma unroll
    for (int k = 0; k < RPG; ++k) {
        pf.act[k] = pf.pos[k] < pf.end[k];
        const int32_t p = pf.act[k] ? pf.pos[k] : 0;
        pf.col[k] = 0;
        pf.val[k] = T(0);
        if (any) {
            if constexpr (Compact) {
                const int32_t c = static_cast<int32_t>(pf.row0 + k * kGroups) + co[p];
                pf.col[k] = pf.act[k] ? c : 0;
            } else {
                pf.col[k] = ci[p];
            }
            pf.val[k] = v[p];
        }
        pf.pos[k] = pf.act[k] ? pf.pos[k] + SW : pf.pos[k];
    }
}

template <typename T, int SW, int RPG, bool Compact, typename Gather, typename RowOp>
__device__ __forceinline__ void subwave_finish(int64_t n_rows, const int32_t* __restrict__ ci,
                                               const int16_t* __restrict__ co,
                                               const T* __restrict__ v, subwave_pf<T, RPG>& pf,
                                               Gather gather, RowOp row_op)
{
    constexpr int kGroups = kWave / SW;
    const int l = (threadIdx.x % kWave) % SW;
    T sum[RPG];
    {
        T xv[RPG];
#pragma unroll
        for (int k = 0; k < RPG; ++k) xv[k] = gather(pf.col[k]);
#pragma unroll
        for (int k = 0; k < RPG; ++k) {
            sum[k] = fma(pf.act[k] ? pf.val[k] : T(0), pf.act[k] ? xv[k] : T(0), T(0));
        }
    }
    for (int32_t j = 1; j < pf.iters; ++j) {
        int32_t col[RPG];
        T val[RPG];
        bool act[RPG];
#pragma unroll
        for (int k = 0; k < RPG; ++k) {
            act[k] = pf.pos[k] < pf.end[k];
            const int32_t p = act[k] ? pf.pos[k] : 0;
            if constexpr (Compact) {
                const int32_t c = static_cast<int32_t>(pf.row0 + k * kGroups) + co[p];
                col[k] = act[k] ? c : 0;
            } else {
                col[k] = ci[p];
            }
            val[k] = v[p];
            pf.pos[k] = act[k] ? pf.pos[k] + SW : pf.pos[k];
        }
        T xv[RPG];
#pragma unroll
        for (int k = 0; k < RPG; ++k) xv[k] = gather(col[k]);
#pragma unroll
        for (int k = 0; k < RPG; ++k) {
            sum[k] = fma(act[k] ? val[k] : T(0), act[k] ? xv[k] : T(0), sum[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < RPG; ++k) {
        sum[k] = group_sum<SW>(sum[k]);
        const int64_t row = pf.row0 + k * kGroups;
        if (row < n_rows && l == 0) row_op(row, sum[k]);
    }
}

// One 64-row slice of a CSR matrix's sliced copy (SELL-P layout, slice 64,
// stride factor 1: slot k of row r at (sets[r / 64] + k) * 64 + r % 64), run
// by one wave with lane = row.  Each lane walks its own row's slots in order
// and stops at the row's length (rp[r + 1] - rp[r]), so the padding of a
// slice is never read; four slots are loaded before they are used.  With
// Compact the column is row + offs[slot], else cols[slot].  `gather(col)`
// supplies the vector entry; `row_op(row, sum)` runs for every row below
// n_rows.  The products and the row sum are formed in double for either
// value type (a product of two floats is exact in double), so an FP32 row
// sum is rounded once, when the caller stores it.
template <typename T, bool Compact, typename TV, typename Gather, typename RowOp>
__device__ __forceinline__ void sliced_slice(int64_t n_rows, int64_t s,
                                             const int32_t* __restrict__ rp,
                                             const int64_t* __restrict__ sets,
                                             const int32_t* __restrict__ cols,
                                             const int16_t* __restrict__ offs,
                                             const TV* __restrict__ vals, Gather gather,
                                             RowOp row_op)
{
    const int lane = threadIdx.x % kWave;
    const int64_t row = s * kWave + lane;
    const bool own = row < n_rows;
    const int32_t len = own ? rp[row + 1] - rp[row] : 0;
    const int64_t base = sets[s] * kWave + lane;
    const int32_t r32 = static_cast<int32_t>(row);
    auto col_of = [&](int64_t idx) -> int32_t {
        if constexpr (Compact) {
            return r32 + offs[idx];
        } else {
            return cols[idx];
        }
    };
    double sum = 0;
    int32_t k = 0;
    for (; k + 4 <= len; k += 4) {
        const int64_t i0 = base + static_cast<int64_t>(k) * kWave;
        int32_t c[4];
        TV v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            c[u] = col_of(i0 + u * kWave);
            v[u] = vals[i0 + u * kWave];
        }
        T x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = gather(c[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            sum = fma(static_cast<double>(v[u]), static_cast<double>(x[u]), sum);
        }
    }
    for (; k < len; ++k) {
        const int64_t i = base + static_cast<int64_t>(k) * kWave;
        sum = fma(static_cast<double>(vals[i]), static_cast<double>(gather(col_of(i))), sum);
    }
    if (own) row_op(row, static_cast<T>(sum));
}

// Dispatch over the classical shapes: sw in {1..64}, rpg in {1, 2, 4, 8};
// CASE(S, R) is expanded for each and must end in `break` or `return`.
#define GKOMI_SUBWAVE_SWITCH(sh, CASE)                                                  \
    switch ((sh).sw * 16 + (sh).rpg) {                                                  \
        GKOMI_SUBWAVE_ROW_(1, CASE) GKOMI_SUBWAVE_ROW_(2, CASE) GKOMI_SUBWAVE_ROW_(4, CASE) \
        GKOMI_SUBWAVE_ROW_(8, CASE) GKOMI_SUBWAVE_ROW_(16, CASE)                        \
        GKOMI_SUBWAVE_ROW_(32, CASE) GKOMI_SUBWAVE_ROW_(64, CASE)                       \
        default:                                                                        \
            return fail(GKOMI_INVALID_ARGUMENT,                                         \
                        "CSR: subwave must be a power of two <= 64 and rows per group "  \
                        "1, 2, 4 or 8");                                                \
    }
#define GKOMI_SUBWAVE_ROW_(S, CASE) \
    case S * 16 + 1: CASE(S, 1) case S * 16 + 2: CASE(S, 2) case S * 16 + 4: CASE(S, 4) \
    case S * 16 + 8: CASE(S, 8)

int check_csr(const gkomi_csr* a);
// Shape of the classical kernel: `sw` lanes per row, `rpg` rows in flight per
// lane group.  gkomi_csr::subwave holds sw in bits 0..7 and rpg in bits 8..15
// (0 = automatic).
struct subwave_shape {
    int sw;
    int rpg;
};

subwave_shape pick_subwave(const gkomi_csr* a, int value_bytes);

// Entries of a wave's LDS staging area in the staged kernel
// (csr_spmv_staged): 64 / G rows of up to 8 G entries in one chunk.
constexpr int kStagedChunk = 448;
int staged_lanes(const gkomi_csr* a, int value_bytes);


template <typename T, int Mode>
int csr_spmv_launch(hipStream_t stream, const gkomi_csr* a, int64_t nrhs, const T* b, int64_t bs,
                    T* c, int64_t cs, const T* alpha, const T* beta);

// merge_path strategy (merge.hip): single right-hand side.
size_t merge_work_bytes(int64_t n_rows, int64_t nnz);
template <typename T, int Mode>
int csr_spmv_merge_launch(hipStream_t stream, const gkomi_csr* a, const T* b, int64_t bs, T* c,
                          int64_t cs, const T* alpha, const T* beta);

}  // namespace gkomi
