// Shared device/host helpers of the gfx950 kernel library (libgkomi.so).
//
// Every kernel in this library is written for CDNA4 (gfx950): wave64,
// 256-thread workgroups (4 waves) unless stated otherwise, and DPP /
// ds_swizzle cross-lane reductions.  Cross-workgroup sums never use a
// device-scope fence or a ticket counter: each workgroup writes one FP64
// partial, and the consumer (a one-workgroup kernel, or every workgroup of
// the next kernel of a fused solver) adds the partials in one fixed order,
// so every reduction is deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "gkomi.h"

namespace gkomi {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

// ---------------------------------------------------------------- errors ----

void set_error(const char* fmt, ...);
int fail(int code, const char* what);
int hip_fail(hipError_t err, const char* what);

#define GKOMI_CHECK_HIP(expr)                                   \
    do {                                                        \
        hipError_t gkomi_err_ = (expr);                         \
        if (gkomi_err_ != hipSuccess) {                         \
            return ::gkomi::hip_fail(gkomi_err_, #expr);        \
        }                                                       \
    } while (0)

#define GKOMI_CHECK_LAUNCH(what)                                \
    do {                                                        \
        hipError_t gkomi_err_ = hipGetLastError();              \
        if (gkomi_err_ != hipSuccess) {                         \
            return ::gkomi::hip_fail(gkomi_err_, what);         \
        }                                                       \
    } while (0)

#define GKOMI_REQUIRE(cond, code, msg)                          \
    do {                                                        \
        if (!(cond)) {                                          \
            return ::gkomi::fail(code, msg);                    \
        }                                                       \
    } while (0)

inline hipStream_t stream_of(const gkomi_ctx* ctx)
{
    return ctx ? static_cast<hipStream_t>(ctx->stream) : nullptr;
}

// Sets the context's device current; returns non-zero on failure.
int enter(const gkomi_ctx* ctx);

// Number of compute units of the current device (cached per device).
int num_cus();

__host__ __device__ inline int64_t ceildiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Grid for `items` work items of `per_block` each, capped so that a
// grid-stride loop covers the rest (>= 64 workgroups per CU is plenty).
inline unsigned grid_for(int64_t items, int64_t per_block, int64_t cap_per_cu = 64)
{
    int64_t g = ceildiv(items, per_block);
    int64_t cap = static_cast<int64_t>(num_cus()) * cap_per_cu;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return static_cast<unsigned>(g);
}

// Calls fn(TA{}, TB{}, TC{}) with the value types (double / float) named by
// the codes GKOMI_VALUE_F64 / GKOMI_VALUE_F32 of a product's matrix, input
// and output operands.
template <typename Fn>
int dispatch_value_types(int32_t a_type, int32_t b_type, int32_t c_type, Fn fn)
{
    const auto valid = [](int32_t t) { return t == GKOMI_VALUE_F64 || t == GKOMI_VALUE_F32; };
    GKOMI_REQUIRE(valid(a_type) && valid(b_type) && valid(c_type), GKOMI_INVALID_ARGUMENT,
                  "unknown value type code (GKOMI_VALUE_F64 or GKOMI_VALUE_F32)");
    const auto with = [](int32_t t, auto f) {
        return t == GKOMI_VALUE_F64 ? f(double{}) : f(float{});
    };
    return with(a_type, [&](auto a) {
        return with(b_type, [&](auto b) { return with(c_type, [&](auto c) { return fn(a, b, c); }); });
    });
}

// --------------------------------------------------------- value traits ----

template <typename T>
struct vt;
template <>
struct vt<double> {
    static constexpr int bytes = 8;
};
template <>
struct vt<float> {
    static constexpr int bytes = 4;
};

// ------------------------------------------------ cross-lane reductions ----
// DPP controls (GFX9 encoding): quad_perm [1,0,3,2] = 0xB1, [2,3,0,1] = 0x4E,
// row_half_mirror = 0x141, row_mirror = 0x140.  ds_swizzle bit-mask mode
// with xor_mask 0x10 (offset 0x401F) swaps the 16-lane halves of each
// 32-lane group.  Each step leaves every lane of the group holding the
// group's sum, so the reductions below are "all-reduce" within the group.

template <int Ctrl>
__device__ __forceinline__ int dpp_i32(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, Ctrl, 0xF, 0xF, false);
}

template <int Ctrl>
__device__ __forceinline__ float dpp(float v)
{
    return __int_as_float(dpp_i32<Ctrl>(__float_as_int(v)));
}

template <int Ctrl>
__device__ __forceinline__ double dpp(double v)
{
    const long long bits = __double_as_longlong(v);
    const int lo = dpp_i32<Ctrl>(static_cast<int>(bits & 0xffffffffll));
    const int hi = dpp_i32<Ctrl>(static_cast<int>(bits >> 32));
    return __longlong_as_double((static_cast<long long>(hi) << 32) |
                                (static_cast<unsigned int>(lo)));
}

__device__ __forceinline__ float swz16(float v)
{
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x401F));
}

__device__ __forceinline__ double swz16(double v)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_swizzle(static_cast<int>(bits & 0xffffffffll), 0x401F);
    const int hi = __builtin_amdgcn_ds_swizzle(static_cast<int>(bits >> 32), 0x401F);
    return __longlong_as_double((static_cast<long long>(hi) << 32) |
                                (static_cast<unsigned int>(lo)));
}

// Sum over aligned groups of G consecutive lanes (G a power of two <= 64);
// every lane of a group receives the group's sum.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v)
{
    if constexpr (G >= 2) v += dpp<0xB1>(v);
    if constexpr (G >= 4) v += dpp<0x4E>(v);
    if constexpr (G >= 8) v += dpp<0x141>(v);
    if constexpr (G >= 16) v += dpp<0x140>(v);
    if constexpr (G >= 32) v += swz16(v);
    if constexpr (G >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
    return group_sum<64>(v);
}

// DPP move restricted to the rows (16-lane groups) in RowMask; the other rows
// receive 0.
template <int Ctrl, int RowMask>
__device__ __forceinline__ float dpp_rows(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), Ctrl, RowMask, 0xF,
                                                      false));
}

template <int Ctrl, int RowMask>
__device__ __forceinline__ double dpp_rows(double v)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(bits & 0xffffffffll), Ctrl,
                                               RowMask, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(bits >> 32), Ctrl, RowMask,
                                               0xF, false);
    return __longlong_as_double((static_cast<long long>(hi) << 32) |
                                (static_cast<unsigned int>(lo)));
}

// Sum of the 64 lanes of a wave, valid in lane 63 only, without LDS traffic.
// Each row of 16 lanes is reduced by the butterfly of group_sum; DPP
// row_bcast:15 (0x142) then adds lane 15's row total into row 1 and lane
// 47's into row 3, and row_bcast:31 (0x143) adds lane 31 (rows 0 + 1) into
// row 3.  Lane 63 ends with (r3 + r2) + (r1 + r0): the same pairs as
// wave_sum's (r0 + r1) + (r2 + r3), so the same bits.
template <typename T>
__device__ __forceinline__ T wave_total_last(T v)
{
    v = group_sum<16>(v);
    v += dpp_rows<0x142, 0xA>(v);
    v += dpp_rows<0x143, 0xC>(v);
    return v;
}

// v[lane] + v[lane ^ M] in every lane, for M = 4, 8, 16 or 32.  M = 32 and
// 16 use the gfx950 lane swaps with the value as both operands: one result
// holds the lower partner of every pair in both lanes, the other the upper
// one, whichever way the instruction swaps, so their sum is the pair's sum.
// M = 8 is DPP row_ror:8 (lane i of a row reads lane (i + 8) mod 16), M = 4
// a ds_swizzle in xor mode.
template <int M>
__device__ __forceinline__ void xor_pair(uint32_t v, uint32_t* a, uint32_t* b)
{
    static_assert(M == 4 || M == 8 || M == 16 || M == 32, "xor_pair: M");
    if constexpr (M == 32) {
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        *a = r[0];
        *b = r[1];
    } else if constexpr (M == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        *a = r[0];
        *b = r[1];
    } else if constexpr (M == 8) {
        *a = v;
        *b = static_cast<uint32_t>(dpp_i32<0x128>(static_cast<int>(v)));
    } else {
        *a = v;
        *b = static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(static_cast<int>(v), 0x101F));
    }
}

template <int M>
__device__ __forceinline__ float xor_pair_sum(float v)
{
    uint32_t a, b;
    xor_pair<M>(__float_as_uint(v), &a, &b);
    return __uint_as_float(a) + __uint_as_float(b);
}

template <int M>
__device__ __forceinline__ double xor_pair_sum(double v)
{
    const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(v));
    uint32_t alo, blo, ahi, bhi;
    xor_pair<M>(static_cast<uint32_t>(bits), &alo, &blo);
    xor_pair<M>(static_cast<uint32_t>(bits >> 32), &ahi, &bhi);
    const double a = __longlong_as_double(static_cast<long long>((uint64_t{ahi} << 32) | alo));
    const double b = __longlong_as_double(static_cast<long long>((uint64_t{bhi} << 32) | blo));
    return a + b;
}

// Sum over the G lanes r, r + 64/G, r + 2*64/G, ... (G a power of two <= 16),
// in every one of them: the butterfly of group_sum<G> with lane strides
// 64/G, 2*64/G, ..., so sub-lane s of a row (lane r + s*64/G) takes the place
// of lane s of group_sum's group and the sum has the same bits.
template <int G, typename T>
__device__ __forceinline__ T strided_sum(T v)
{
    constexpr int R = kWave / G;
    if constexpr (G >= 16) v = xor_pair_sum<R>(v);
    if constexpr (G >= 8) v = xor_pair_sum<R * (G >= 16 ? 2 : 1)>(v);
    if constexpr (G >= 4) v = xor_pair_sum<16>(v);
    if constexpr (G >= 2) v = xor_pair_sum<32>(v);
    return v;
}

// Workgroup sum over blockDim.x == 256 threads; result valid in all threads.
// `lds` must hold kWavesPerBlock values.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* lds)
{
    v = wave_sum(v);
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    T total = lds[0];
#pragma unroll
    for (int w = 1; w < kWavesPerBlock; ++w) total += lds[w];
    __syncthreads();
    return total;
}

// XCD-aware workgroup index: consecutive workgroups are dispatched round-robin
// over the 8 XCDs; this remaps them so that each XCD walks one contiguous
// range of the logical index space (better L2 reuse of gathered vectors).
__device__ __forceinline__ unsigned xcd_block(unsigned bid, unsigned nblocks)
{
    constexpr unsigned kXcd = 8;
    const unsigned per = nblocks / kXcd;
    const unsigned rem = nblocks % kXcd;
    const unsigned xcd = bid % kXcd;
    const unsigned pos = bid / kXcd;
    // the first `rem` XCDs own one extra block each
    const unsigned base = xcd * per + (xcd < rem ? xcd : rem);
    return base + pos;
}

// Contiguous share [begin, end) of `total` items for this workgroup, with the
// XCD-aware workgroup order (neighbouring shares on one XCD).
__device__ __forceinline__ void block_range(int64_t total, int64_t* begin, int64_t* end)
{
    const int64_t g = gridDim.x;
    const int64_t lb = xcd_block(blockIdx.x, gridDim.x);
    const int64_t per = ceildiv(total, g);
    *begin = min(total, lb * per);
    *end = min(total, *begin + per);
}


}  // namespace gkomi
