// Fused CG for k right-hand sides on the MI355X: k independent CG recurrences
// that share the matrix, with the reference's per-column semantics
// (core/solver/cg.cpp:93-196 with nrhs = k; reference/solver/cg_kernels.cpp:
// 23-101: a column whose stopping status is set keeps its p, x and r).
//
// Vectors are n x k, row-major with stride k.  A row of all k columns is
// handled by L lanes (L = min(64, next power of two >= k), RHS chunks of L
// when k > 64), so every gather x[col, :] is one coalesced segment of k
// values.  Per iteration four kernels:
//   check      (1 workgroup): rho_j = r_j.z_j, tau_j = ||r_j|| from the update
//              partials, the Combined check over all columns, the step_1
//              factor rho_j / prev_rho_j;
//   direction  p = z + f p (p = z on restart, p kept for stopped columns),
//              recomputed at gathered columns, q = A p, partials of p_j.q_j;
//   beta       (1 workgroup): beta_j from the direction partials;
//   update     x += rho/beta p, r -= rho/beta q (columns not stopped,
//              beta != 0), z = M r, partials of r.z and r.r.
// Sums are FP64 partials per workgroup added in fixed order by a
// one-workgroup kernel: deterministic and free of device-scope fences.
#include "common.hpp"
#include "reduce.hpp"
#include "spmv_common.hpp"
#include "stop.hpp"

namespace gkomi {

constexpr int kBcgMaxGrid = 1024;

// scalars (value type, 5k): rho | prev_used | prev_next | beta | tau
template <typename T>
struct bcg_scalars {
    T* rho;
    T* prev_used;
    T* prev_next;
    T* beta;
    T* tau;
};

template <typename T>
__host__ __device__ inline bcg_scalars<T> bcg_scalar_layout(void* base, int64_t k)
{
    T* s = static_cast<T*>(base);
    return {s, s + k, s + 2 * k, s + 3 * k, s + 4 * k};
}

struct bcg_work {
    double* part_a;  // [kBcgMaxGrid][2k]: (b.b, r.r) of the start, then (r.z, r.r)
    double* part_b;  // [kBcgMaxGrid][k]: p.q
    double* sums;    // [2k]: column sums of the last combine
};

__host__ __device__ inline bcg_work bcg_work_layout(void* base, int64_t k)
{
    double* d = static_cast<double*>(base);
    const int64_t g = kBcgMaxGrid;
    return {d, d + 2 * g * k, d + 3 * g * k};
}

// Column partials of one workgroup: thread t holds the sums of column
// chunk + t % L for its rows; the groups (blockDim / L of them) are added in
// order through LDS.
template <int L, int Count>
__device__ __forceinline__ void bcg_store_partials(const double (&acc)[Count], int64_t k,
                                                   int64_t c0, double* partials,
                                                   double (*lds)[kBlock])
{
    constexpr int kGroups = kBlock / L;
    const int lane = threadIdx.x % L;
#pragma unroll
    for (int c = 0; c < Count; ++c) lds[c][threadIdx.x] = acc[c];
    __syncthreads();
    if (threadIdx.x < L && c0 + lane < k) {
#pragma unroll
        for (int c = 0; c < Count; ++c) {
            double s = lds[c][lane];
            for (int g = 1; g < kGroups; ++g) s += lds[c][g * L + lane];
            partials[static_cast<int64_t>(blockIdx.x) * Count * k + c * k + c0 + lane] = s;
        }
    }
    __syncthreads();
}

// Single-workgroup column sums of the partials P[g][C] (g < blocks), for the
// one-workgroup kernels below (kCombineThreads threads): columns in chunks of
// 64 lanes, the 16 waves take g = w, w + 16, ... and their sums are added in
// wave order.  Each lane keeps four accumulators (g = w + 64 m + 16 u for
// u = 0..3), added as (a0 + a1) + (a2 + a3) at the end: a fixed order, and
// four independent loads in flight per lane instead of one dependent chain.
constexpr int kCombineThreads = 1024;
constexpr int kCombineWaves = kCombineThreads / kWave;

__device__ __forceinline__ void combine_columns(const double* __restrict__ P, int blocks,
                                                int64_t C, double* __restrict__ out,
                                                double (*lds)[kWave])
{
    const int lane = threadIdx.x % kWave;
    const int wave = threadIdx.x / kWave;
    for (int64_t c0 = 0; c0 < C; c0 += kWave) {
        const int64_t c = c0 + lane;
        double acc = 0;
        if (c < C) {
            double a[4] = {0, 0, 0, 0};
            int g = wave;
            for (; g + 3 * kCombineWaves < blocks; g += 4 * kCombineWaves) {
#pragma unroll
                for (int u = 0; u < 4; ++u) a[u] += P[(g + u * kCombineWaves) * C + c];
            }
            for (int u = 0; g < blocks; g += kCombineWaves, ++u) a[u] += P[g * C + c];
            acc = (a[0] + a[1]) + (a[2] + a[3]);
        }
        lds[wave][lane] = acc;
        __syncthreads();
        if (wave == 0 && c < C) {
            double s = lds[0][lane];
            for (int w = 1; w < kCombineWaves; ++w) s += lds[w][lane];
            out[c] = s;
        }
        __syncthreads();
    }
}

// Column sums of the partials (column_sums_block, reduce.hpp) for the
// one-workgroup check: a fixed order, one workgroup per 8 columns.  With
// `beta` the sums are also the step_1 factors (p.q per column).
template <typename T>
__global__ void __launch_bounds__(kColSumThreads)
    bcg_colsum_kernel(const double* __restrict__ P, int blocks, int64_t C,
                      double* __restrict__ out, T* __restrict__ beta,
                      const gkomi_stop_flags* __restrict__ flags)
{
    if (flags->all_stopped) return;
    column_sums_block(P, blocks, C, [&](int64_t c, double sum) {
        out[c] = sum;
        if (beta != nullptr) beta[c] = static_cast<T>(sum);
    });
}

// Row i of A times a gathered vector, kRowBatch entries at a time with all
// loads of a batch issued before their use (lanes past the row end reuse the
// row's first entry and add nothing).  Entry order within the row is the
// CSR order, one fused multiply-add each.
constexpr int kRowBatch = 6;

template <typename T, typename Gather>
__device__ __forceinline__ T row_product(const int32_t* __restrict__ ci,
                                         const T* __restrict__ v, int32_t beg, int32_t end,
                                         Gather gather)
{
    T sum = 0;
    for (int32_t p0 = beg; p0 < end; p0 += kRowBatch) {
        int32_t cc[kRowBatch];
        T vv[kRowBatch];
        bool act[kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) {
            act[u] = p0 + u < end;
            const int32_t p = act[u] ? p0 + u : p0;
            cc[u] = ci[p];
            vv[u] = v[p];
        }
        T gv[kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) gv[u] = gather(cc[u]);
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) {
            sum = fma(act[u] ? vv[u] : T(0), act[u] ? gv[u] : T(0), sum);
        }
    }
    return sum;
}

// Value of `x` in lane `src` of the wave (all lanes of the group active).
template <int L, typename V>
__device__ __forceinline__ V group_bcast(V x, int src)
{
    if constexpr (L == kWave) {
        // one group per wave, src a compile-time lane: a scalar read
        if constexpr (sizeof(V) == 8) {
            const uint64_t u = __builtin_bit_cast(uint64_t, x);
            const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(u), src);
            const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(u >> 32), src);
            return __builtin_bit_cast(V, (static_cast<uint64_t>(hi) << 32) | lo);
        } else {
            return __builtin_bit_cast(
                V, __builtin_amdgcn_readlane(__builtin_bit_cast(uint32_t, x), src));
        }
    } else {
        return __shfl(x, src, kWave);
    }
}

// R rows at once (the same scheme, the loads of all R rows issued together):
// sum[r] = row (beg[r], end[r]) times the gathered vector.  With at least
// kRowBatch lanes per group, lane u of the group loads entry u of the batch
// (one coalesced load for the group) and the entries are broadcast inside the
// group; narrower groups load every entry in every lane.  All lanes of a
// group must execute this together.
template <typename T, int R, int L, typename Gather>
__device__ __forceinline__ void rows_product(const int32_t* __restrict__ ci,
                                             const T* __restrict__ v, const int32_t (&beg)[R],
                                             const int32_t (&end)[R], Gather gather,
                                             T (&sum)[R])
{
    const int lane = threadIdx.x % L;
    const int base = threadIdx.x % kWave - lane;
    int32_t iters = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        sum[r] = 0;
        const int32_t len = end[r] - beg[r];
        const int32_t it = len > 0 ? (len + kRowBatch - 1) / kRowBatch : 0;
        iters = it > iters ? it : iters;
    }
    for (int32_t t = 0; t < iters; ++t) {
        int32_t cc[R][kRowBatch];
        T vv[R][kRowBatch];
        bool act[R][kRowBatch];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int32_t p0 = beg[r] + t * kRowBatch;
            if constexpr (L >= kRowBatch) {
                // rows of a tile are contiguous: beg[0] is a valid entry
                // whenever some row of the tile has one left
                const int32_t pl = p0 + lane;
                const int32_t pidx = (lane < kRowBatch && pl < end[r]) ? pl : beg[0];
                const int32_t cl = ci[pidx];
                const T vl = v[pidx];
#pragma unroll
                for (int u = 0; u < kRowBatch; ++u) {
                    act[r][u] = p0 + u < end[r];
                    cc[r][u] = group_bcast<L>(cl, base + u);
                    vv[r][u] = group_bcast<L>(vl, base + u);
                }
            } else {
#pragma unroll
                for (int u = 0; u < kRowBatch; ++u) {
                    act[r][u] = p0 + u < end[r];
                    const int32_t p = act[r][u] ? p0 + u : beg[0];
                    cc[r][u] = ci[p];
                    vv[r][u] = v[p];
                }
            }
        }
        T gv[R][kRowBatch];
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int u = 0; u < kRowBatch; ++u) gv[r][u] = gather(cc[r][u]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int u = 0; u < kRowBatch; ++u) {
                sum[r] = fma(act[r][u] ? vv[r][u] : T(0), act[r][u] ? gv[r][u] : T(0), sum[r]);
            }
        }
    }
}

// rows per lane group in flight in the product kernel: 3 rows of 6 slots
// (18 gathers in flight per wave on 5-point rows at 115 VGPRs, against 10
// with 2 rows of 8 at 122: 512^2, 64 columns 119.6 -> 110.0 us)
constexpr int kBcgRows = 3;

// ---------------------------------------------------------------- start ----
// r = b - A x; p_a = p_b = q = 0; partials of b_j.b_j and r_j.r_j.
template <typename T, int L>
__global__ void __launch_bounds__(kBlock)
    bcg_init_kernel(int64_t n, int64_t k, const int32_t* __restrict__ rp,
                    const int32_t* __restrict__ ci, const T* __restrict__ v,
                    const T* __restrict__ b, int64_t bs, const T* __restrict__ x, int64_t xs,
                    T* __restrict__ r, T* __restrict__ pa, T* __restrict__ pb,
                    T* __restrict__ q, bcg_work w)
{
    __shared__ double lds[2][kBlock];
    constexpr int kRowsPerBlock = kBlock / L;
    const int lane = threadIdx.x % L;
    const int g = threadIdx.x / L;
    int64_t rb0, rb1;
    block_range(ceildiv(n, kRowsPerBlock), &rb0, &rb1);
    for (int64_t c0 = 0; c0 < k; c0 += L) {
        const int64_t j = c0 + lane;
        const bool col_ok = j < k;
        double acc[2] = {0, 0};
        for (int64_t blk = rb0; blk < rb1; ++blk) {
            const int64_t i = blk * kRowsPerBlock + g;
            if (i >= n || !col_ok) continue;
            const T sum = row_product<T>(ci, v, rp[i], rp[i + 1], [&](int32_t c) {
                return x[static_cast<int64_t>(c) * xs + j];
            });
            const T bi = b[i * bs + j];
            const T ri = bi - sum;
            r[i * k + j] = ri;
            pa[i * k + j] = T(0);
            pb[i * k + j] = T(0);
            q[i * k + j] = T(0);
            acc[0] += static_cast<double>(bi * bi);
            acc[1] += static_cast<double>(ri * ri);
        }
        bcg_store_partials<L, 2>(acc, k, c0, w.part_a, lds);
    }
}

// One workgroup: the criteria baselines from (b.b, r.r), status and flags
// reset, prev_rho = 1 (reference/solver/cg_kernels.cpp:26-50).
template <typename T>
__global__ void __launch_bounds__(kCombineThreads)
    bcg_start_combine_kernel(int64_t k, int blocks, gkomi_criteria crit, bcg_work w,
                             T* __restrict__ scalars, T* __restrict__ orig,
                             uint8_t* __restrict__ status, gkomi_stop_flags* __restrict__ flags)
{
    __shared__ double lds[kCombineWaves][kWave];
    const bcg_scalars<T> sc = bcg_scalar_layout<T>(scalars, k);
    combine_columns(w.part_a, blocks, 2 * k, w.sums, lds);
    for (int64_t j = threadIdx.x; j < k; j += kCombineThreads) {
        const double bb = w.sums[j], rr = w.sums[k + j];
        for (int c = 0; c < crit.num; ++c) {
            const int base = crit.c[c].baseline;
            orig[c * k + j] = base == GKOMI_BASELINE_RHS_NORM
                                  ? static_cast<T>(sqrt(bb))
                                  : (base == GKOMI_BASELINE_INITIAL_RESNORM
                                         ? static_cast<T>(sqrt(rr))
                                         : T(1));
        }
        sc.rho[j] = T(0);
        sc.prev_used[j] = T(1);
        sc.prev_next[j] = T(1);
        sc.beta[j] = T(0);
        sc.tau[j] = T(0);
        status[j] = 0;
    }
    if (threadIdx.x == 0) {
        flags->iteration = -1;
        flags->all_stopped = 0;
        flags->one_changed = 0;
    }
}

// ----------------------------------------------------------- iteration ----
// One workgroup: the check of iteration `iter` from the column sums of the
// update partials (bcg_colsum_kernel has formed them), and the step_1
// factor.
template <typename T>
__global__ void __launch_bounds__(kCombineThreads)
    bcg_check_kernel(int64_t k, gkomi_criteria crit, bcg_work w, T* __restrict__ scalars,
                     const T* __restrict__ orig, uint8_t* __restrict__ status,
                     gkomi_stop_flags* __restrict__ flags, int64_t iter)
{
    if (flags->all_stopped) return;
    const bcg_scalars<T> sc = bcg_scalar_layout<T>(scalars, k);
    for (int64_t j = threadIdx.x; j < k; j += kCombineThreads) {
        sc.rho[j] = static_cast<T>(w.sums[j]);
        sc.tau[j] = static_cast<T>(sqrt(w.sums[k + j]));
    }
    __syncthreads();
    int all = 0, changed = 0;
    combined_check_block<kCombineThreads, T>(crit, iter, k, sc.tau, sc.rho, orig, true, status,
                                             &all, &changed);
    if (threadIdx.x == 0) {
        flags->iteration = iter;
        flags->all_stopped = all;
        flags->one_changed = changed;
    }
    __syncthreads();
    // rho of this iteration becomes prev_rho of the next; the direction
    // kernel reads the value it replaces
    for (int64_t j = threadIdx.x; j < k; j += kCombineThreads) {
        sc.prev_used[j] = sc.prev_next[j];
        sc.prev_next[j] = sc.rho[j];
    }
}

// Rows per staged chunk and LDS capacity (entries) of the direction kernel.
constexpr int kBcgChunkRows = 256;
constexpr int kBcgChunkNnz = 2048;

// q = A p and the partials of p.q, p being the new directions
// (bcg_direction_kernel has written them).  The workgroup stages a chunk of
// up to kBcgChunkRows rows -- row pointers and, when they fit, the chunk's
// column indices and values -- in LDS with coalesced loads, then each lane
// group takes R rows at a time: entries from LDS, the gathers of all R rows
// in flight.  The only dependent global round trip per row batch is the
// gather itself.  With Off32 (n k sizeof(T) below 4 GiB) the gathers use
// 32-bit byte offsets from the uniform base pointer.
template <typename T, int L, bool Off32>
__global__ void __launch_bounds__(kBlock)
    bcg_product_kernel(int64_t n, int64_t k, const int32_t* __restrict__ rp,
                       const int32_t* __restrict__ ci, const T* __restrict__ v,
                       const T* __restrict__ p, T* __restrict__ q,
                       const gkomi_stop_flags* __restrict__ flags, bcg_work w)
{
    if (flags->all_stopped) return;
    __shared__ double lds[1][kBlock];
    __shared__ int32_t lrp[kBcgChunkRows + 1];
    __shared__ int32_t lci[kBcgChunkNnz];
    __shared__ T lv[kBcgChunkNnz];
    constexpr int R = kBcgRows;
    constexpr int kGroups = kBlock / L;
    const int lane = threadIdx.x % L;
    const int g = threadIdx.x / L;
    int64_t r0, r1;
    block_range(n, &r0, &r1);
    for (int64_t c0 = 0; c0 < k; c0 += L) {
        const bool col_ok = c0 + lane < k;
        // lanes past the last column follow column c0 and store nothing
        const int64_t j = col_ok ? c0 + lane : c0;
        auto at = [&](int64_t row) { return p[row * k + j]; };
        // p of row `col`, column j
        auto gather = [&](int32_t col) {
            if constexpr (Off32) {
                const uint32_t b = (static_cast<uint32_t>(col) * static_cast<uint32_t>(k) +
                                    static_cast<uint32_t>(j)) *
                                   static_cast<uint32_t>(sizeof(T));
                return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + b);
            } else {
                return p[static_cast<int64_t>(col) * k + j];
            }
        };
        double acc[1] = {0};
        for (int64_t cs = r0; cs < r1; cs += kBcgChunkRows) {
            const int rows = static_cast<int>(min(int64_t{kBcgChunkRows}, r1 - cs));
            __syncthreads();
            for (int t = threadIdx.x; t <= rows; t += kBlock) lrp[t] = rp[cs + t];
            __syncthreads();
            const int32_t e0 = lrp[0];
            const int32_t ne = lrp[rows] - e0;
            const bool staged = ne <= kBcgChunkNnz;
            if (staged) {
                for (int t = threadIdx.x; t < ne; t += kBlock) {
                    lci[t] = ci[e0 + t];
                    lv[t] = v[e0 + t];
                }
            }
            __syncthreads();
            for (int rr = g * R; rr < rows; rr += kGroups * R) {
                // the rows' own p, loaded ahead of the gathers
                T own_p[R];
#pragma unroll
                for (int r = 0; r < R; ++r) own_p[r] = at(cs + (rr + r < rows ? rr + r : 0));
                T sum[R];
                if (staged) {
                    // entries from LDS, kRowBatch at a time for all R rows
                    int32_t beg[R], end[R];
                    int32_t iters = 0;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int lr = rr + r < rows ? rr + r : rows;
                        beg[r] = lrp[lr] - e0;
                        end[r] = lrp[rr + r < rows ? lr + 1 : rows] - e0;
                        sum[r] = 0;
                        const int32_t len = end[r] - beg[r];
                        const int32_t it = len > 0 ? (len + kRowBatch - 1) / kRowBatch : 0;
                        iters = it > iters ? it : iters;
                    }
                    for (int32_t t = 0; t < iters; ++t) {
                        // three phases -- entry slots, all gathers, arithmetic --
                        // so every gather of the batch is in flight at once.  A
                        // slot past its row's end holds -1 and loads nothing:
                        // a wave of 64 lanes owns one row, so such a slot
                        // would be a whole wave-wide load, and those cost
                        // 12% of the solve time on a 5-point stencil.
                        int32_t slot[R][kRowBatch];
#pragma unroll
                        for (int r = 0; r < R; ++r) {
#pragma unroll
                            for (int u = 0; u < kRowBatch; ++u) {
                                const int32_t e = beg[r] + t * kRowBatch + u;
                                slot[r][u] = e < end[r] ? e : -1;
                            }
                        }
                        T pv[R][kRowBatch];
#pragma unroll
                        for (int r = 0; r < R; ++r) {
#pragma unroll
                            for (int u = 0; u < kRowBatch; ++u) {
                                pv[r][u] = slot[r][u] >= 0 ? gather(lci[slot[r][u]]) : T(0);
                            }
                        }
#pragma unroll
                        for (int r = 0; r < R; ++r) {
#pragma unroll
                            for (int u = 0; u < kRowBatch; ++u) {
                                const bool act = slot[r][u] >= 0;
                                const T vv = lv[act ? slot[r][u] : 0];
                                sum[r] = fma(act ? vv : T(0), pv[r][u], sum[r]);
                            }
                        }
                    }
                } else {
                    int32_t beg[R], end[R];
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int lr = rr + r < rows ? rr + r : rows;
                        beg[r] = lrp[lr];
                        end[r] = lrp[rr + r < rows ? lr + 1 : rows];
                    }
                    rows_product<T, R, L>(ci, v, beg, end, at, sum);
                }
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int64_t i = cs + rr + r;
                    if (rr + r < rows && col_ok) {
                        q[i * k + j] = sum[r];
                        acc[0] += static_cast<double>(own_p[r] * sum[r]);
                    }
                }
            }
        }
        bcg_store_partials<L, 1>(acc, k, c0, w.part_b, lds);
    }
}

// The direction update: p_out = p_in (stopped column), z (restart, the
// previous rho is zero) or z + (rho / prev_rho) p_in, every row.  Written out
// here, so the product kernel gathers p only -- one load per entry instead
// of the two (p_in and z) a product that formed the direction at every
// gathered entry would need; 3% of the solve on a 5-point stencil with 64
// columns.
template <typename T, int L>
__global__ void __launch_bounds__(kBlock)
    bcg_direction_kernel(int64_t n, int64_t k, const T* __restrict__ z,
                         const T* __restrict__ p_in, T* __restrict__ p_out,
                         const T* __restrict__ scalars, const uint8_t* __restrict__ status,
                         const gkomi_stop_flags* __restrict__ flags)
{
    if (flags->all_stopped) return;
    const bcg_scalars<T> sc = bcg_scalar_layout<T>(const_cast<T*>(scalars), k);
    constexpr int kRowsPerBlock = kBlock / L;
    const int lane = threadIdx.x % L;
    const int g = threadIdx.x / L;
    int64_t rb0, rb1;
    block_range(ceildiv(n, kRowsPerBlock), &rb0, &rb1);
    for (int64_t c0 = 0; c0 < k; c0 += L) {
        const int64_t j = c0 + lane;
        if (j >= k) continue;
        int mode = 0;  // 0 keep, 1 z, 2 z + f p
        T f = 0;
        if (!has_stopped(status[j])) {
            const T prev = sc.prev_used[j];
            mode = prev == T(0) ? 1 : 2;
            f = mode == 2 ? sc.rho[j] / prev : T(0);
        }
        // four rows per step, all eight loads issued first (a row past the
        // end reads row 0 and stores nothing)
        constexpr int U = 4;
        for (int64_t blk = rb0; blk < rb1; blk += U) {
            int64_t o[U];
            bool ok[U];
            T pv[U], zv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t i = (blk + u) * kRowsPerBlock + g;
                ok[u] = blk + u < rb1 && i < n;
                o[u] = ok[u] ? i * k + j : j;
                pv[u] = __builtin_nontemporal_load(p_in + o[u]);
                zv[u] = __builtin_nontemporal_load(z + o[u]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const T fm = fma(f, pv[u], zv[u]);
                const T r1 = mode == 1 ? zv[u] : fm;
                if (ok[u]) __builtin_nontemporal_store(mode == 0 ? pv[u] : r1, p_out + o[u]);
            }
        }
    }
}

// Grid of the direction kernel (no partials, so not bound by their
// capacity): up to 8 workgroups per CU.
inline int bcg_direction_grid(int64_t n, int lanes)
{
    const int64_t tiles = ceildiv(std::max<int64_t>(n, 1), kBlock / lanes);
    const int64_t g = std::min<int64_t>(tiles, static_cast<int64_t>(num_cus()) * 8);
    return static_cast<int>(std::max<int64_t>(g, 1));
}

// Update (and, with Update = false, the tail of the start): x, r update of
// the active columns, z = M r, partials of r.z and r.r.  Without a
// preconditioner z is r itself (the caller passes the same buffer).
template <typename T, int L, bool Update, bool Jacobi>
__global__ void __launch_bounds__(kBlock)
    bcg_update_kernel(int64_t n, int64_t k, T* __restrict__ x, int64_t xs, T* __restrict__ r,
                      T* __restrict__ z, const T* __restrict__ p, const T* __restrict__ q,
                      const T* __restrict__ inv_diag, const T* __restrict__ scalars,
                      const uint8_t* __restrict__ status,
                      const gkomi_stop_flags* __restrict__ flags, bcg_work w)
{
    if (Update && flags->all_stopped) return;
    __shared__ double lds[2][kBlock];
    const bcg_scalars<T> sc = bcg_scalar_layout<T>(const_cast<T*>(scalars), k);
    constexpr int kRowsPerBlock = kBlock / L;
    const int lane = threadIdx.x % L;
    const int g = threadIdx.x / L;
    for (int64_t c0 = 0; c0 < k; c0 += L) {
        const int64_t j = c0 + lane;
        const bool col_ok = j < k;
        bool upd = false;
        T alpha = 0;
        if (Update && col_ok && !has_stopped(status[j])) {
            const T beta = sc.beta[j];
            upd = beta != T(0);
            alpha = upd ? sc.rho[j] / beta : T(0);
        }
        double acc[2] = {0, 0};
        // four rows per step, their loads issued first (a row past the end
        // reads entry 0 and stores nothing; the partials add the rows in
        // order).  x and q are not read again before the next update:
        // streamed with non-temporal loads and stores (2.8% of the adjoint
        // solve on a 512^2 stencil with 64 columns); r and z stay cached for
        // the direction kernel (non-temporal stores of them cost it 29%).
        constexpr int U = 4;
        // row tiles in grid-stride order (XCD-aware start), U at a time:
        // 125.5 -> 120.2 us on the 512^2 adjoint against a contiguous range
        // of tiles per workgroup (the direction kernel measured no gain)
        const int64_t tiles = ceildiv(n, kRowsPerBlock);
        const int64_t nb = gridDim.x;
        for (int64_t blk = xcd_block(blockIdx.x, gridDim.x); blk < tiles; blk += U * nb) {
            int64_t o[U], ox[U], id[U];
            bool ok[U];
            T rv[U], pv[U], xv[U], qv[U], dv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t i = (blk + u * nb) * kRowsPerBlock + g;
                ok[u] = blk + u * nb < tiles && i < n && col_ok;
                o[u] = ok[u] ? i * k + j : 0;
                ox[u] = ok[u] ? i * xs + j : 0;
                id[u] = ok[u] ? i : 0;
                rv[u] = r[o[u]];
                if (upd) {
                    pv[u] = p[o[u]];
                    xv[u] = __builtin_nontemporal_load(x + ox[u]);
                    qv[u] = __builtin_nontemporal_load(q + o[u]);
                }
                if constexpr (Jacobi) dv[u] = inv_diag[id[u]];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (!ok[u]) continue;
                T ri = rv[u];
                if (upd) {
                    __builtin_nontemporal_store(fma(alpha, pv[u], xv[u]), x + ox[u]);
                    ri = fma(-alpha, qv[u], ri);
                    r[o[u]] = ri;
                }
                T zi = ri;
                if constexpr (Jacobi) {
                    zi = ri * dv[u];
                    z[o[u]] = zi;
                }
                acc[0] += static_cast<double>(ri * zi);
                acc[1] += static_cast<double>(ri * ri);
            }
        }
        bcg_store_partials<L, 2>(acc, k, c0, w.part_a, lds);
    }
}

// ------------------------------------------------------------- drivers ----

int check_bcg_state(const gkomi_block_cg_state* st)
{
    GKOMI_REQUIRE(st != nullptr, GKOMI_INVALID_ARGUMENT, "null block CG state");
    GKOMI_REQUIRE(st->n >= 0 && st->nrhs >= 1, GKOMI_INVALID_ARGUMENT,
                  "block CG state: negative size or no right-hand side");
    GKOMI_REQUIRE(st->x_stride >= st->nrhs && st->b_stride >= st->nrhs, GKOMI_INVALID_ARGUMENT,
                  "block CG state: stride smaller than the number of columns");
    GKOMI_REQUIRE(st->x && st->b && st->r && st->z && st->p_a && st->p_b && st->q &&
                      st->scalars && st->orig_tau && st->status && st->flags && st->work,
                  GKOMI_INVALID_ARGUMENT, "block CG state: null pointer");
    GKOMI_REQUIRE(st->inv_diag != nullptr || st->z == st->r, GKOMI_INVALID_ARGUMENT,
                  "block CG state: without a preconditioner z must be r");
    return GKOMI_SUCCESS;
}

int check_bcg_crit(const gkomi_criteria* crit)
{
    GKOMI_REQUIRE(crit != nullptr && crit->num >= 1 && crit->num <= GKOMI_MAX_CRITERIA,
                  GKOMI_NOT_SUPPORTED, "criteria: between 1 and 4 criteria are supported");
    for (int c = 0; c < crit->num; ++c) {
        const int kind = crit->c[c].kind;
        GKOMI_REQUIRE(kind == GKOMI_CRIT_ITERATION || kind == GKOMI_CRIT_RESIDUAL_NORM ||
                          kind == GKOMI_CRIT_IMPLICIT_RESNORM,
                      GKOMI_NOT_SUPPORTED, "criteria: unknown kind");
    }
    return GKOMI_SUCCESS;
}

int bcg_lanes(int64_t k)
{
    int l = 1;
    while (l < 64 && l < k) l *= 2;
    return l;
}

int bcg_grid(int64_t n, int lanes)
{
    const int64_t tiles = ceildiv(std::max<int64_t>(n, 1), kBlock / lanes);
    int64_t g = std::min<int64_t>(tiles, static_cast<int64_t>(num_cus()) * 4);
    g = std::min<int64_t>(g, kBcgMaxGrid);
    return static_cast<int>(std::max<int64_t>(g, 1));
}

// Grid of the product kernel: at most the workgroups resident on the chip
// at once (a larger grid would run a partial second wave of workgroups).
template <typename T, int L, bool Off32>
int bcg_product_grid(int grid)
{
    static const int per_cu = [] {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, bcg_product_kernel<T, L, Off32>,
                                                         kBlock, 0) != hipSuccess ||
            nb < 1) {
            (void)hipGetLastError();
            nb = 1;
        }
        return nb;
    }();
    return static_cast<int>(std::min<int64_t>(grid, static_cast<int64_t>(num_cus()) * per_cu));
}

#define GKOMI_BCG_LANES(l, CASE)                                                       \
    switch (l) {                                                                       \
        case 1: CASE(1) case 2: CASE(2) case 4: CASE(4) case 8: CASE(8) case 16: CASE(16) \
        case 32: CASE(32) case 64: CASE(64)                                            \
        default: return fail(GKOMI_INVALID_ARGUMENT, "block CG: bad lane count");      \
    }

template <typename T>
int bcg_start_impl(const gkomi_ctx* ctx, const gkomi_csr* a, const gkomi_criteria* crit,
                   gkomi_block_cg_state* st)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_csr(a)) return rc;
    if (int rc = check_bcg_crit(crit)) return rc;
    if (int rc = check_bcg_state(st)) return rc;
    GKOMI_REQUIRE(a->n_rows == a->n_cols && a->n_rows == st->n, GKOMI_DIMENSION_MISMATCH,
                  "block CG: matrix and vectors do not conform");
    hipStream_t s = stream_of(ctx);
    const int64_t n = st->n, k = st->nrhs;
    const bcg_work w = bcg_work_layout(st->work, k);
    const int lanes = bcg_lanes(k);
    const int grid = bcg_grid(n, lanes);
    const T* v = static_cast<const T*>(a->values);
    T* r = static_cast<T*>(st->r);
#define GKOMI_BCG_INIT(L)                                                                  \
    bcg_init_kernel<T, L><<<grid, kBlock, 0, s>>>(                                         \
        n, k, a->row_ptrs, a->col_idxs, v, static_cast<const T*>(st->b), st->b_stride,     \
        static_cast<const T*>(st->x), st->x_stride, r, static_cast<T*>(st->p_a),           \
        static_cast<T*>(st->p_b), static_cast<T*>(st->q), w);                              \
    break;
    GKOMI_BCG_LANES(lanes, GKOMI_BCG_INIT)
#undef GKOMI_BCG_INIT
    GKOMI_CHECK_LAUNCH("bcg_init_kernel");
    bcg_start_combine_kernel<T><<<1, kCombineThreads, 0, s>>>(k, grid, *crit, w,
                                                     static_cast<T*>(st->scalars),
                                                     static_cast<T*>(st->orig_tau), st->status,
                                                     st->flags);
    GKOMI_CHECK_LAUNCH("bcg_start_combine_kernel");
    const T* inv = static_cast<const T*>(st->inv_diag);
    T* z = static_cast<T*>(st->z);
#define GKOMI_BCG_UPD0(L)                                                                  \
    if (inv) {                                                                             \
        bcg_update_kernel<T, L, false, true><<<grid, kBlock, 0, s>>>(                      \
            n, k, static_cast<T*>(st->x), st->x_stride, r, z, nullptr, nullptr, inv,       \
            static_cast<const T*>(st->scalars), st->status, st->flags, w);                 \
    } else {                                                                               \
        bcg_update_kernel<T, L, false, false><<<grid, kBlock, 0, s>>>(                     \
            n, k, static_cast<T*>(st->x), st->x_stride, r, z, nullptr, nullptr, nullptr,   \
            static_cast<const T*>(st->scalars), st->status, st->flags, w);                 \
    }                                                                                      \
    break;
    GKOMI_BCG_LANES(lanes, GKOMI_BCG_UPD0)
#undef GKOMI_BCG_UPD0
    GKOMI_CHECK_LAUNCH("bcg_update_kernel");
    st->issued = 0;
    return GKOMI_SUCCESS;
}

template <typename T>
int bcg_iterate_impl(const gkomi_ctx* ctx, const gkomi_csr* a, const gkomi_criteria* crit,
                     gkomi_block_cg_state* st, int32_t count)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_csr(a)) return rc;
    if (int rc = check_bcg_crit(crit)) return rc;
    if (int rc = check_bcg_state(st)) return rc;
    GKOMI_REQUIRE(a->n_rows == a->n_cols && a->n_rows == st->n, GKOMI_DIMENSION_MISMATCH,
                  "block CG: matrix and vectors do not conform");
    GKOMI_REQUIRE(count >= 0 && st->issued >= 0, GKOMI_INVALID_ARGUMENT,
                  "block CG: negative count or state not started");
    hipStream_t s = stream_of(ctx);
    const int64_t n = st->n, k = st->nrhs;
    const bcg_work w = bcg_work_layout(st->work, k);
    const int lanes = bcg_lanes(k);
    const int grid = bcg_grid(n, lanes);
    const T* v = static_cast<const T*>(a->values);
    T* sc = static_cast<T*>(st->scalars);
    T* r = static_cast<T*>(st->r);
    T* z = static_cast<T*>(st->z);
    const T* inv = static_cast<const T*>(st->inv_diag);
    // byte offsets of p and z entries fit 32 bits
    const bool off32 = static_cast<uint64_t>(n) * static_cast<uint64_t>(k) * sizeof(T) <
                       (uint64_t{1} << 32);
    for (int32_t it = 0; it < count; ++it) {
        T* pin = static_cast<T*>(st->p_a);
        T* pout = static_cast<T*>(st->p_b);
        bcg_colsum_kernel<T><<<colsum_grid(2 * k), kColSumThreads, 0, s>>>(
            w.part_a, grid, 2 * k, w.sums, nullptr, st->flags);
        bcg_check_kernel<T><<<1, kCombineThreads, 0, s>>>(k, *crit, w, sc,
                                                 static_cast<const T*>(st->orig_tau),
                                                 st->status, st->flags, st->issued);
        GKOMI_CHECK_LAUNCH("bcg_colsum_kernel, bcg_check_kernel");
#define GKOMI_BCG_PROD(L, OFF)                                                             \
    dgrid = bcg_product_grid<T, L, OFF>(grid);                                             \
    bcg_product_kernel<T, L, OFF><<<dgrid, kBlock, 0, s>>>(                                \
        n, k, a->row_ptrs, a->col_idxs, v, pout, static_cast<T*>(st->q), st->flags, w);
#define GKOMI_BCG_DIR(L)                                                                   \
    bcg_direction_kernel<T, L><<<bcg_direction_grid(n, L), kBlock, 0, s>>>(                 \
        n, k, z, pin, pout, sc, st->status, st->flags);                                    \
    if (off32) {                                                                           \
        GKOMI_BCG_PROD(L, true)                                                            \
    } else {                                                                               \
        GKOMI_BCG_PROD(L, false)                                                           \
    }                                                                                      \
    break;
        int dgrid = grid;
        GKOMI_BCG_LANES(lanes, GKOMI_BCG_DIR)
#undef GKOMI_BCG_DIR
#undef GKOMI_BCG_PROD
        GKOMI_CHECK_LAUNCH("bcg_direction_kernel, bcg_product_kernel");
        bcg_colsum_kernel<T><<<colsum_grid(k), kColSumThreads, 0, s>>>(
            w.part_b, dgrid, k, w.sums, bcg_scalar_layout<T>(sc, k).beta, st->flags);
        GKOMI_CHECK_LAUNCH("bcg_colsum_kernel (beta)");
#define GKOMI_BCG_UPD(L)                                                                   \
    if (inv) {                                                                             \
        bcg_update_kernel<T, L, true, true><<<grid, kBlock, 0, s>>>(                       \
            n, k, static_cast<T*>(st->x), st->x_stride, r, z, pout,                        \
            static_cast<const T*>(st->q), inv, sc, st->status, st->flags, w);              \
    } else {                                                                               \
        bcg_update_kernel<T, L, true, false><<<grid, kBlock, 0, s>>>(                      \
            n, k, static_cast<T*>(st->x), st->x_stride, r, z, pout,                        \
            static_cast<const T*>(st->q), nullptr, sc, st->status, st->flags, w);          \
    }                                                                                      \
    break;
        GKOMI_BCG_LANES(lanes, GKOMI_BCG_UPD)
#undef GKOMI_BCG_UPD
        GKOMI_CHECK_LAUNCH("bcg_update_kernel");
        std::swap(st->p_a, st->p_b);
        ++st->issued;
    }
    return GKOMI_SUCCESS;
}

}  // namespace gkomi

using namespace gkomi;

extern "C" {

size_t gkomi_block_cg_work_bytes(int64_t n, int64_t nrhs)
{
    (void)n;
    return sizeof(double) * (3 * static_cast<size_t>(kBcgMaxGrid) + 2) *
           static_cast<size_t>(std::max<int64_t>(nrhs, 1));
}

int gkomi_block_cg_start_f64(const gkomi_ctx* ctx, const gkomi_csr* a,
                             const gkomi_criteria* crit, gkomi_block_cg_state* st)
{
    return bcg_start_impl<double>(ctx, a, crit, st);
}

int gkomi_block_cg_start_f32(const gkomi_ctx* ctx, const gkomi_csr* a,
                             const gkomi_criteria* crit, gkomi_block_cg_state* st)
{
    return bcg_start_impl<float>(ctx, a, crit, st);
}

int gkomi_block_cg_iterate_f64(const gkomi_ctx* ctx, const gkomi_csr* a,
                               const gkomi_criteria* crit, gkomi_block_cg_state* st,
                               int32_t count)
{
    return bcg_iterate_impl<double>(ctx, a, crit, st, count);
}

int gkomi_block_cg_iterate_f32(const gkomi_ctx* ctx, const gkomi_csr* a,
                               const gkomi_criteria* crit, gkomi_block_cg_state* st,
                               int32_t count)
{
    return bcg_iterate_impl<float>(ctx, a, crit, st, count);
}

void* gkomi_block_cg_current_p(const gkomi_block_cg_state* st)
{
    return st ? st->p_a : nullptr;
}

}  // extern "C"
