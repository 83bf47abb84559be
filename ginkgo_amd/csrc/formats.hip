// SELL-P and ELL storage (conversion from CSR, SpMV) and the CSR transpose.
//
// SELL-P with slice_size 64 maps one slice onto one wave: lane = row in the
// slice, so every load of the column-major slice storage is a fully
// coalesced 64-wide access.  ELL is column-major with stride n_rows, one lane
// per row.  The transpose is a counting sort: column histogram, exclusive
// scan, scatter, then a per-column rank sort by original position that makes
// the order identical to the reference's sequential stable scatter.
#include "common.hpp"
#include "scan.hpp"
#include "spmv_common.hpp"

namespace gkomi {

// -------------------------------------------------------------- SELL-P ----

__global__ void __launch_bounds__(kBlock)
    slice_lengths_kernel(int64_t n_rows, const int32_t* __restrict__ rp, int64_t slice_size,
                         int64_t stride_factor, int64_t num_slices,
                         int64_t* __restrict__ slice_lengths)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t s = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; s < num_slices;
         s += stride) {
        int64_t len = 0;
        const int64_t r0 = s * slice_size;
        const int64_t r1 = std::min(n_rows, r0 + slice_size);
        for (int64_t r = r0; r < r1; ++r) {
            const int64_t l = rp[r + 1] - rp[r];
            len = max(len, ceildiv(l, stride_factor) * stride_factor);
        }
        slice_lengths[s] = len;
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    csr_to_sellp_kernel(int64_t n_rows, const int32_t* __restrict__ rp,
                        const int32_t* __restrict__ ci, const T* __restrict__ v,
                        int64_t slice_size, int64_t num_slices,
                        const int64_t* __restrict__ slice_sets,
                        const int64_t* __restrict__ slice_lengths, int32_t* __restrict__ s_cols,
                        T* __restrict__ s_vals)
{
    const int64_t total_rows = num_slices * slice_size;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; r < total_rows;
         r += stride) {
        const int64_t s = r / slice_size;
        const int64_t lr = r - s * slice_size;
        const int64_t base = slice_sets[s] * slice_size + lr;
        const int64_t slen = slice_lengths[s];
        int64_t k = 0;
        if (r < n_rows) {
            const int32_t b = rp[r];
            const int32_t len = rp[r + 1] - b;
            for (; k < len; ++k) {
                s_cols[base + k * slice_size] = ci[b + k];
                s_vals[base + k * slice_size] = v[b + k];
            }
        }
        for (; k < slen; ++k) {
            s_cols[base + k * slice_size] = -1;
            s_vals[base + k * slice_size] = T(0);
        }
    }
}

// Matrix values TA, input TB, output TC; the sum is formed in acc_t (the
// highest precision of the three), as the reference's mixed kernels do.
template <typename TA, typename TB, typename TC, bool Adv>
__global__ void __launch_bounds__(kBlock)
    sellp_spmv_kernel(int64_t n_rows, int64_t nrhs, int64_t slice_size,
                      const int64_t* __restrict__ slice_sets,
                      const int64_t* __restrict__ slice_lengths,
                      const int32_t* __restrict__ s_cols, const TA* __restrict__ s_vals,
                      const TB* __restrict__ b, int64_t bs, TC* __restrict__ c, int64_t cs,
                      const TA* __restrict__ alpha, const TC* __restrict__ beta)
{
    using Acc = acc_t<TA, TB, TC>;
    const int64_t nb = gridDim.x;
    const int64_t row_blocks = ceildiv(n_rows, kBlock);
    for (int64_t blk = xcd_block(blockIdx.x, gridDim.x); blk < row_blocks; blk += nb) {
        const int64_t r = blk * kBlock + threadIdx.x;
        if (r >= n_rows) continue;
        const int64_t s = r / slice_size;
        const int64_t base = slice_sets[s] * slice_size + (r - s * slice_size);
        const int64_t slen = slice_lengths[s];
        for (int64_t j = 0; j < nrhs; ++j) {
            Acc sum = 0;
            for (int64_t k = 0; k < slen; ++k) {
                const int64_t idx = base + k * slice_size;
                const int32_t col = s_cols[idx];
                if (col != -1) {
                    sum += static_cast<Acc>(s_vals[idx]) *
                           static_cast<Acc>(b[static_cast<int64_t>(col) * bs + j]);
                }
            }
            store_result<Adv>(c + r * cs + j, sum, alpha, beta);
        }
    }
}

// SELL-P product for slices of 64 rows and one right-hand side: one wave per
// slice, lane = row.  A row's stored entries precede its padding (the
// layout the conversions produce), so a lane stops at its row's first
// padding slot and never reads the padded tail of a slice -- which on
// matrices with a few long rows is most of the storage.
//  * Phase 1: each lane walks its own row over the first kSellpLanePhase
//    slots, four slots in flight at once.
//  * Phase 2: rows not finished by then (found by a ballot) are spread over
//    the wave, G = 64 / pow2ceil(count) lanes each, striding along the row
//    four slots at a time; a lane group sum joins the row's phase-1 sum.
// Every row sums in one fixed order: deterministic.
constexpr int kSellpLanePhase = 32;

template <typename TA, typename TB, typename TC, bool Adv>
__global__ void __launch_bounds__(kBlock)
    sellp_spmv_wave_kernel(int64_t n_rows, int64_t num_slices,
                           const int64_t* __restrict__ slice_sets,
                           const int64_t* __restrict__ slice_lengths,
                           const int32_t* __restrict__ s_cols, const TA* __restrict__ s_vals,
                           const TB* __restrict__ b, int64_t bs, TC* __restrict__ c, int64_t cs,
                           const TA* __restrict__ alpha, const TC* __restrict__ beta)
{
    using Acc = acc_t<TA, TB, TC>;
    constexpr int U = 4;
    const int lane = threadIdx.x % kWave;
    const int64_t num_waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
    for (int64_t s = static_cast<int64_t>(xcd_block(blockIdx.x, gridDim.x)) * kWavesPerBlock +
                     threadIdx.x / kWave;
         s < num_slices; s += num_waves) {
        const int64_t set = slice_sets[s];
        const int64_t len = slice_lengths[s];
        const int64_t row = s * kWave + lane;
        const int64_t base = set * kWave + lane;
        const int64_t k1 = len < kSellpLanePhase ? len : kSellpLanePhase;
        Acc sum = 0;
        bool done = row >= n_rows;
        for (int64_t k = 0; k < k1; k += U) {
            if (__ballot(!done) == 0) break;
            int32_t col[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                col[u] = (!done && k + u < k1) ? s_cols[base + (k + u) * kWave] : -2;
            }
            bool ok[U];
            bool go = !done;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                ok[u] = go && col[u] >= 0;
                go = ok[u];
                // padding ends the row; the end of phase 1 (-2) does not
                done = done || col[u] == -1;
            }
            TA val[U];
            TB xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                val[u] = s_vals[ok[u] ? base + (k + u) * kWave : base];
                xv[u] = b[static_cast<int64_t>(ok[u] ? col[u] : col[0] >= 0 ? col[0] : 0) * bs];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                sum = fma(ok[u] ? static_cast<Acc>(val[u]) : Acc(0),
                          ok[u] ? static_cast<Acc>(xv[u]) : Acc(0), sum);
            }
        }
        const bool pending = !done && len > kSellpLanePhase;
        const unsigned long long mask = __ballot(pending);
        if (mask != 0ull) {
            const int cnt = __popcll(mask);
            int gsz = kWave;
            while (gsz > 1 && gsz * cnt > kWave) gsz /= 2;
            const int g = lane / gsz;
            const int li = lane % gsz;
            Acc part = 0;
            if (g < cnt) {
                unsigned long long m = mask;
                for (int i = 0; i < g; ++i) m &= m - 1;
                const int prow = __ffsll(static_cast<long long>(m)) - 1;
                const int64_t pbase = set * kWave + prow;
                bool pdone = false;
                for (int64_t k = kSellpLanePhase + li; k < len && !pdone;
                     k += static_cast<int64_t>(U) * gsz) {
                    int32_t col[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int64_t kk = k + static_cast<int64_t>(u) * gsz;
                        col[u] = kk < len ? s_cols[pbase + kk * kWave] : -1;
                    }
                    bool ok[U];
                    bool go = true;
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        ok[u] = go && col[u] >= 0;
                        go = ok[u];
                    }
                    pdone = !go;
                    TA val[U];
                    TB xv[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int64_t kk = k + static_cast<int64_t>(u) * gsz;
                        val[u] = s_vals[ok[u] ? pbase + kk * kWave : pbase];
                        xv[u] = b[static_cast<int64_t>(ok[u] ? col[u] : 0) * bs];
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        part = fma(ok[u] ? static_cast<Acc>(val[u]) : Acc(0),
                                   ok[u] ? static_cast<Acc>(xv[u]) : Acc(0), part);
                    }
                }
            }
            for (int off = gsz / 2; off >= 1; off /= 2) part += __shfl_xor(part, off, kWave);
            const int rank = __popcll(mask & ((1ull << lane) - 1ull));
            const Acc add = __shfl(part, (rank * gsz) & (kWave - 1), kWave);
            if (pending) sum += add;
        }
        if (row < n_rows) store_result<Adv>(c + row * cs, sum, alpha, beta);
    }
}

// ------------------------------------------------------------------ ELL ----

template <typename T>
__global__ void __launch_bounds__(kBlock)
    csr_to_ell_kernel(int64_t n_rows, const int32_t* __restrict__ rp,
                      const int32_t* __restrict__ ci, const T* __restrict__ v, int64_t width,
                      int64_t stride_e, int32_t* __restrict__ e_cols, T* __restrict__ e_vals)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; r < stride_e;
         r += stride) {
        int64_t k = 0;
        if (r < n_rows) {
            const int32_t b = rp[r];
            const int64_t len = std::min<int64_t>(rp[r + 1] - b, width);
            for (; k < len; ++k) {
                e_cols[r + k * stride_e] = ci[b + k];
                e_vals[r + k * stride_e] = v[b + k];
            }
        }
        for (; k < width; ++k) {
            e_cols[r + k * stride_e] = -1;
            e_vals[r + k * stride_e] = T(0);
        }
    }
}

// Hybrid: counts[r] = entries of row r beyond the ELL width (the COO part's
// row lengths).
__global__ void __launch_bounds__(kBlock)
    hybrid_coo_counts_kernel(int64_t n_rows, const int32_t* __restrict__ rp, int64_t width,
                             int32_t* __restrict__ counts)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; r < n_rows;
         r += stride) {
        const int64_t len = rp[r + 1] - rp[r];
        counts[r] = static_cast<int32_t>(len > width ? len - width : 0);
    }
}

// csr::convert_to_hybrid (reference/matrix/csr_kernels.cpp:911-957): the
// first `width` entries of every row into the column-major ELL part
// (padding -1 / 0), the rest into the COO part at coo_rp[r] onwards, in row
// order.  One thread per row: the ELL writes of a column k are coalesced.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    csr_to_hybrid_kernel(int64_t n_rows, const int32_t* __restrict__ rp,
                         const int32_t* __restrict__ ci, const T* __restrict__ v, int64_t width,
                         int64_t stride_e, int32_t* __restrict__ e_cols, T* __restrict__ e_vals,
                         const int32_t* __restrict__ coo_rp, int32_t* __restrict__ coo_rows,
                         int32_t* __restrict__ coo_cols, T* __restrict__ coo_vals)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; r < stride_e;
         r += stride) {
        int64_t k = 0;
        if (r < n_rows) {
            const int32_t b = rp[r];
            const int64_t len = rp[r + 1] - b;
            const int64_t in_ell = std::min<int64_t>(len, width);
            for (; k < in_ell; ++k) {
                e_cols[r + k * stride_e] = ci[b + k];
                e_vals[r + k * stride_e] = v[b + k];
            }
            const int64_t p0 = coo_rp[r];
            for (int64_t q = width; q < len; ++q) {
                coo_rows[p0 + q - width] = static_cast<int32_t>(r);
                coo_cols[p0 + q - width] = ci[b + q];
                coo_vals[p0 + q - width] = v[b + q];
            }
        }
        for (; k < width; ++k) {
            e_cols[r + k * stride_e] = -1;
            e_vals[r + k * stride_e] = T(0);
        }
    }
}

template <typename TA, typename TB, typename TC, bool Adv>
__global__ void __launch_bounds__(kBlock)
    ell_spmv_kernel(int64_t n_rows, int64_t nrhs, int64_t width, int64_t stride_e,
                    const int32_t* __restrict__ e_cols, const TA* __restrict__ e_vals,
                    const TB* __restrict__ b, int64_t bs, TC* __restrict__ c, int64_t cs,
                    const TA* __restrict__ alpha, const TC* __restrict__ beta)
{
    using Acc = acc_t<TA, TB, TC>;
    const int64_t nb = gridDim.x;
    const int64_t row_blocks = ceildiv(n_rows, kBlock);
    for (int64_t blk = xcd_block(blockIdx.x, gridDim.x); blk < row_blocks; blk += nb) {
        const int64_t r = blk * kBlock + threadIdx.x;
        if (r >= n_rows) continue;
        for (int64_t j = 0; j < nrhs; ++j) {
            Acc sum = 0;
            for (int64_t k = 0; k < width; ++k) {
                const int32_t col = e_cols[r + k * stride_e];
                if (col != -1) {
                    sum += static_cast<Acc>(e_vals[r + k * stride_e]) *
                           static_cast<Acc>(b[static_cast<int64_t>(col) * bs + j]);
                }
            }
            store_result<Adv>(c + r * cs + j, sum, alpha, beta);
        }
    }
}

// ------------------------------------------------------------ transpose ----

__global__ void __launch_bounds__(kBlock)
    column_histogram(int64_t nnz, const int32_t* __restrict__ ci, int32_t* __restrict__ counts)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < nnz;
         i += stride) {
        atomicAdd(&counts[ci[i]], 1);
    }
}

// scatter (original index, row) into the column segments; order inside a
// segment is arbitrary here and fixed by the rank sort below
__global__ void __launch_bounds__(kBlock)
    transpose_scatter(int64_t n_rows, const int32_t* __restrict__ rp,
                      const int32_t* __restrict__ ci, int32_t* __restrict__ cursor,
                      int32_t* __restrict__ key, int32_t* __restrict__ row_of)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
    for (int64_t r = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; r < n_rows;
         r += stride) {
        for (int32_t i = rp[r]; i < rp[r + 1]; ++i) {
            const int32_t slot = atomicAdd(&cursor[ci[i]], 1);
            key[slot] = i;
            row_of[slot] = static_cast<int32_t>(r);
        }
    }
}

// rank sort of each column segment by original position (keys are unique):
// short segments by one lane, long ones by the whole wave
constexpr int kShortSegment = 16;

template <typename T>
__global__ void __launch_bounds__(kBlock)
    transpose_finish(int64_t n_cols, const int32_t* __restrict__ t_rp,
                     const int32_t* __restrict__ key, const int32_t* __restrict__ row_of,
                     const T* __restrict__ v, int32_t* __restrict__ t_ci, T* __restrict__ t_v)
{
    const int lane = threadIdx.x % kWave;
    const int64_t wave_id = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
    const int64_t num_waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
    for (int64_t c0 = wave_id * kWave; c0 < n_cols; c0 += num_waves * kWave) {
        const int64_t c = c0 + lane;
        int32_t s = 0, e = 0;
        if (c < n_cols) {
            s = t_rp[c];
            e = t_rp[c + 1];
        }
        const bool is_long = (e - s) > kShortSegment;
        if (c < n_cols && !is_long) {
            for (int32_t a = s; a < e; ++a) {
                const int32_t ka = key[a];
                int32_t rank = 0;
                for (int32_t q = s; q < e; ++q) rank += key[q] < ka ? 1 : 0;
                t_ci[s + rank] = row_of[a];
                t_v[s + rank] = v[ka];
            }
        }
        uint64_t long_mask = __ballot(is_long);
        while (long_mask) {
            const int l = __ffsll(static_cast<long long>(long_mask)) - 1;
            long_mask &= long_mask - 1;
            const int64_t cc = c0 + l;
            const int32_t ls = t_rp[cc];
            const int32_t le = t_rp[cc + 1];
            for (int32_t a = ls + lane; a < le; a += kWave) {
                const int32_t ka = key[a];
                int32_t rank = 0;
                for (int32_t q = ls; q < le; ++q) rank += key[q] < ka ? 1 : 0;
                t_ci[ls + rank] = row_of[a];
                t_v[ls + rank] = v[ka];
            }
        }
    }
}

struct transpose_ws {
    int32_t* cursor;  // n_cols + 1
    int32_t* key;     // nnz
    int32_t* row_of;  // nnz
    void* scan;
};

size_t transpose_ws_layout(int64_t n_rows, int64_t n_cols, int64_t nnz, void* base,
                           transpose_ws* out)
{
    (void)n_rows;
    auto align = [](size_t x) { return (x + 255) / 256 * 256; };
    size_t off = 0;
    const size_t cur = align(sizeof(int32_t) * static_cast<size_t>(n_cols + 1));
    const size_t kb = align(sizeof(int32_t) * static_cast<size_t>(std::max<int64_t>(nnz, 1)));
    const size_t sc = align(scan_ws_bytes(n_cols + 1, sizeof(int32_t)));
    if (out) {
        char* p = static_cast<char*>(base);
        out->cursor = reinterpret_cast<int32_t*>(p + off);
        out->key = reinterpret_cast<int32_t*>(p + off + cur);
        out->row_of = reinterpret_cast<int32_t*>(p + off + cur + kb);
        out->scan = p + off + cur + 2 * kb;
    }
    return cur + 2 * kb + sc;
}

// --------------------------------------------------------------- drivers ----

// col - row of every slot of a SELL-P matrix with 64-row slices (0 for
// padding), one wave per slice; a slot that needs a wider offset clears *fits
__global__ void __launch_bounds__(kBlock)
    sellp_col_offsets_kernel(int64_t n_rows, int64_t num_slices, const int64_t* __restrict__ sets,
                             const int32_t* __restrict__ cols, int16_t* __restrict__ out,
                             int32_t* __restrict__ fits)
{
    const int lane = threadIdx.x % kWave;
    const int64_t nwaves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
    bool ok = true;
    for (int64_t sl = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
         sl < num_slices; sl += nwaves) {
        const int64_t row = sl * kWave + lane;
        const int64_t e = sets[sl + 1];
        for (int64_t k = sets[sl]; k < e; ++k) {
            const int64_t idx = k * kWave + lane;
            const int32_t c = cols[idx];
            const int64_t d = c < 0 ? 0 : static_cast<int64_t>(c) - row;
            ok = ok && d >= -32768 && d <= 32767;
            out[idx] = static_cast<int16_t>(d);
        }
    }
    if (!ok) atomicMin(fits, 0);
}

int check_sellp(const gkomi_sellp* a)
{
    GKOMI_REQUIRE(a != nullptr, GKOMI_INVALID_ARGUMENT, "null SELL-P descriptor");
    GKOMI_REQUIRE(a->n_rows >= 0 && a->slice_size > 0 && a->stride_factor > 0,
                  GKOMI_INVALID_ARGUMENT, "SELL-P: bad sizes");
    GKOMI_REQUIRE(a->n_rows == 0 || (a->slice_sets && a->slice_lengths), GKOMI_INVALID_ARGUMENT,
                  "SELL-P: null slice arrays");
    return GKOMI_SUCCESS;
}

template <typename TA, typename TB, typename TC, bool Adv>
int sellp_spmv_impl(const gkomi_ctx* ctx, const TA* alpha, const gkomi_sellp* a, int64_t nrhs,
                    const TB* b, int64_t bs, const TC* beta, TC* c, int64_t cs)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_sellp(a)) return rc;
    if (a->n_rows == 0 || nrhs == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(nrhs > 0 && bs >= nrhs && cs >= nrhs, GKOMI_INVALID_ARGUMENT,
                  "SELL-P spmv: bad strides");
    GKOMI_REQUIRE(b && c, GKOMI_INVALID_ARGUMENT, "SELL-P spmv: null vector");
    GKOMI_REQUIRE(!Adv || (alpha && beta), GKOMI_INVALID_ARGUMENT, "SELL-P: null alpha/beta");
    if (a->slice_size == kWave && nrhs == 1) {
        const int64_t num_slices = ceildiv(a->n_rows, kWave);
        sellp_spmv_wave_kernel<TA, TB, TC, Adv>
            <<<grid_for(num_slices, kWavesPerBlock, 32), kBlock, 0, stream_of(ctx)>>>(
                a->n_rows, num_slices, a->slice_sets, a->slice_lengths, a->col_idxs,
                static_cast<const TA*>(a->values), b, bs, c, cs, alpha, beta);
        GKOMI_CHECK_LAUNCH("sellp_spmv_wave_kernel");
        return GKOMI_SUCCESS;
    }
    sellp_spmv_kernel<TA, TB, TC, Adv>
        <<<grid_for(a->n_rows, kBlock, 32), kBlock, 0, stream_of(ctx)>>>(
            a->n_rows, nrhs, a->slice_size, a->slice_sets, a->slice_lengths, a->col_idxs,
            static_cast<const TA*>(a->values), b, bs, c, cs, alpha, beta);
    GKOMI_CHECK_LAUNCH("sellp_spmv_kernel");
    return GKOMI_SUCCESS;
}

template <typename T>
int csr_to_sellp_impl(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t slice_size,
                      const int64_t* slice_sets, const int64_t* slice_lengths, int32_t* s_cols,
                      T* s_vals)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_csr(a)) return rc;
    GKOMI_REQUIRE(slice_size > 0, GKOMI_INVALID_ARGUMENT, "SELL-P: slice size must be > 0");
    const int64_t num_slices = ceildiv(a->n_rows, slice_size);
    if (num_slices == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(slice_sets && slice_lengths, GKOMI_INVALID_ARGUMENT,
                  "SELL-P: null slice arrays");
    csr_to_sellp_kernel<T><<<grid_for(num_slices * slice_size, kBlock), kBlock, 0,
                             stream_of(ctx)>>>(a->n_rows, a->row_ptrs, a->col_idxs,
                                               static_cast<const T*>(a->values), slice_size,
                                               num_slices, slice_sets, slice_lengths, s_cols,
                                               s_vals);
    GKOMI_CHECK_LAUNCH("csr_to_sellp_kernel");
    return GKOMI_SUCCESS;
}

template <typename T>
int csr_to_ell_impl(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t width, int64_t stride,
                    int32_t* e_cols, T* e_vals)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_csr(a)) return rc;
    GKOMI_REQUIRE(width >= 0 && stride >= a->n_rows, GKOMI_INVALID_ARGUMENT,
                  "ELL: stride must be >= number of rows");
    if (stride == 0 || width == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(e_cols && e_vals, GKOMI_INVALID_ARGUMENT, "ELL: null output");
    csr_to_ell_kernel<T><<<grid_for(stride, kBlock), kBlock, 0, stream_of(ctx)>>>(
        a->n_rows, a->row_ptrs, a->col_idxs, static_cast<const T*>(a->values), width, stride,
        e_cols, e_vals);
    GKOMI_CHECK_LAUNCH("csr_to_ell_kernel");
    return GKOMI_SUCCESS;
}

size_t hybrid_counts_bytes(int64_t n_rows)
{
    return (static_cast<size_t>(std::max<int64_t>(n_rows, 1)) * sizeof(int32_t) + 255) / 256 *
           256;
}

size_t hybrid_ws_bytes(int64_t n_rows)
{
    return hybrid_counts_bytes(n_rows) + scan_ws_bytes(n_rows, sizeof(int64_t));
}

int hybrid_coo_row_ptrs_impl(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t width,
                             int32_t* coo_rp, void* ws, size_t wsb)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_csr(a)) return rc;
    GKOMI_REQUIRE(width >= 0 && coo_rp != nullptr, GKOMI_INVALID_ARGUMENT,
                  "hybrid: negative width or null output");
    GKOMI_REQUIRE(ws != nullptr && wsb >= hybrid_ws_bytes(a->n_rows), GKOMI_INVALID_ARGUMENT,
                  "hybrid: workspace too small (gkomi_csr_hybrid_workspace_bytes)");
    hipStream_t s = stream_of(ctx);
    // the row counts, then the scan's tile sums, in the workspace
    int32_t* counts = static_cast<int32_t*>(ws);
    void* scan_ws = static_cast<char*>(ws) + hybrid_counts_bytes(a->n_rows);
    if (a->n_rows > 0) {
        hybrid_coo_counts_kernel<<<grid_for(a->n_rows, kBlock), kBlock, 0, s>>>(
            a->n_rows, a->row_ptrs, width, counts);
        GKOMI_CHECK_LAUNCH("hybrid_coo_counts_kernel");
    }
    return exclusive_scan<int32_t>(s, a->n_rows, counts, coo_rp, scan_ws);
}

template <typename T>
int csr_to_hybrid_impl(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t width, int64_t stride,
                       int32_t* e_cols, T* e_vals, const int32_t* coo_rp, int32_t* coo_rows,
                       int32_t* coo_cols, T* coo_vals)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_csr(a)) return rc;
    GKOMI_REQUIRE(width >= 0 && stride >= a->n_rows, GKOMI_INVALID_ARGUMENT,
                  "hybrid: ELL stride must be >= number of rows");
    if (stride == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(coo_rp != nullptr && (width == 0 || (e_cols && e_vals)),
                  GKOMI_INVALID_ARGUMENT, "hybrid: null output");
    csr_to_hybrid_kernel<T><<<grid_for(stride, kBlock), kBlock, 0, stream_of(ctx)>>>(
        a->n_rows, a->row_ptrs, a->col_idxs, static_cast<const T*>(a->values), width, stride,
        e_cols, e_vals, coo_rp, coo_rows, coo_cols, coo_vals);
    GKOMI_CHECK_LAUNCH("csr_to_hybrid_kernel");
    return GKOMI_SUCCESS;
}

template <typename TA, typename TB, typename TC, bool Adv>
int ell_spmv_impl(const gkomi_ctx* ctx, const TA* alpha, const gkomi_ell* a, int64_t nrhs,
                  const TB* b, int64_t bs, const TC* beta, TC* c, int64_t cs)
{
    if (int rc = enter(ctx)) return rc;
    GKOMI_REQUIRE(a != nullptr && a->n_rows >= 0 && a->width >= 0 && a->stride >= a->n_rows,
                  GKOMI_INVALID_ARGUMENT, "ELL: bad descriptor");
    if (a->n_rows == 0 || nrhs == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(nrhs > 0 && bs >= nrhs && cs >= nrhs, GKOMI_INVALID_ARGUMENT,
                  "ELL spmv: bad strides");
    GKOMI_REQUIRE(c && (a->width == 0 || (b && a->col_idxs && a->values)),
                  GKOMI_INVALID_ARGUMENT, "ELL spmv: null pointer");
    GKOMI_REQUIRE(!Adv || (alpha && beta), GKOMI_INVALID_ARGUMENT, "ELL: null alpha/beta");
    ell_spmv_kernel<TA, TB, TC, Adv>
        <<<grid_for(a->n_rows, kBlock, 32), kBlock, 0, stream_of(ctx)>>>(
            a->n_rows, nrhs, a->width, a->stride, a->col_idxs,
            static_cast<const TA*>(a->values), b, bs, c, cs, alpha, beta);
    GKOMI_CHECK_LAUNCH("ell_spmv_kernel");
    return GKOMI_SUCCESS;
}

template <typename T>
int transpose_impl(const gkomi_ctx* ctx, const gkomi_csr* a, int32_t* t_rp, int32_t* t_ci,
                   T* t_v, void* ws, size_t wsb)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_csr(a)) return rc;
    GKOMI_REQUIRE(t_rp != nullptr, GKOMI_INVALID_ARGUMENT, "transpose: null row pointers");
    GKOMI_REQUIRE(a->nnz == 0 || (t_ci && t_v), GKOMI_INVALID_ARGUMENT,
                  "transpose: null output");
    const size_t need = transpose_ws_layout(a->n_rows, a->n_cols, a->nnz, nullptr, nullptr);
    GKOMI_REQUIRE(ws != nullptr && wsb >= need, GKOMI_INVALID_ARGUMENT,
                  "transpose: workspace too small");
    transpose_ws w;
    transpose_ws_layout(a->n_rows, a->n_cols, a->nnz, ws, &w);
    hipStream_t st = stream_of(ctx);
    // counts -> exclusive scan -> t_rp; cursor = copy of t_rp
    if (int rc = gkomi_fill_zero(ctx, w.cursor, sizeof(int32_t) * (a->n_cols + 1))) return rc;
    if (a->nnz > 0) {
        column_histogram<<<grid_for(a->nnz, kBlock * 4), kBlock, 0, st>>>(a->nnz, a->col_idxs,
                                                                          w.cursor);
        GKOMI_CHECK_LAUNCH("column_histogram");
    }
    if (int rc = exclusive_scan<int32_t>(st, a->n_cols, w.cursor, t_rp, w.scan)) return rc;
    if (a->nnz == 0) return GKOMI_SUCCESS;
    GKOMI_CHECK_HIP(hipMemcpyAsync(w.cursor, t_rp, sizeof(int32_t) * (a->n_cols + 1),
                                   hipMemcpyDeviceToDevice, st));
    transpose_scatter<<<grid_for(a->n_rows, kBlock), kBlock, 0, st>>>(
        a->n_rows, a->row_ptrs, a->col_idxs, w.cursor, w.key, w.row_of);
    GKOMI_CHECK_LAUNCH("transpose_scatter");
    transpose_finish<T><<<grid_for(ceildiv(a->n_cols, kWave), kWavesPerBlock), kBlock, 0, st>>>(
        a->n_cols, t_rp, w.key, w.row_of, static_cast<const T*>(a->values), t_ci, t_v);
    GKOMI_CHECK_LAUNCH("transpose_finish");
    return GKOMI_SUCCESS;
}

}  // namespace gkomi

using namespace gkomi;

extern "C" {

size_t gkomi_scan_workspace_bytes(int64_t n) { return scan_ws_bytes(n, sizeof(int64_t)); }

size_t gkomi_csr_hybrid_workspace_bytes(int64_t n_rows) { return hybrid_ws_bytes(n_rows); }

int gkomi_csr_hybrid_coo_row_ptrs(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t ell_width,
                                  int32_t* coo_row_ptrs, void* workspace,
                                  size_t workspace_bytes)
{
    return hybrid_coo_row_ptrs_impl(ctx, a, ell_width, coo_row_ptrs, workspace,
                                    workspace_bytes);
}

int gkomi_csr_to_hybrid_f64(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t ell_width,
                            int64_t ell_stride, int32_t* ell_cols, double* ell_vals,
                            const int32_t* coo_row_ptrs, int32_t* coo_row_idxs,
                            int32_t* coo_col_idxs, double* coo_vals)
{
    return csr_to_hybrid_impl(ctx, a, ell_width, ell_stride, ell_cols, ell_vals, coo_row_ptrs,
                              coo_row_idxs, coo_col_idxs, coo_vals);
}

int gkomi_csr_to_hybrid_f32(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t ell_width,
                            int64_t ell_stride, int32_t* ell_cols, float* ell_vals,
                            const int32_t* coo_row_ptrs, int32_t* coo_row_idxs,
                            int32_t* coo_col_idxs, float* coo_vals)
{
    return csr_to_hybrid_impl(ctx, a, ell_width, ell_stride, ell_cols, ell_vals, coo_row_ptrs,
                              coo_row_idxs, coo_col_idxs, coo_vals);
}

int gkomi_sellp_compute_slice_sets(const gkomi_ctx* ctx, int64_t n_rows, const int32_t* row_ptrs,
                                   int64_t slice_size, int64_t stride_factor,
                                   int64_t* slice_sets, int64_t* slice_lengths, void* ws,
                                   size_t wsb)
{
    if (int rc = enter(ctx)) return rc;
    GKOMI_REQUIRE(n_rows >= 0 && slice_size > 0 && stride_factor > 0, GKOMI_INVALID_ARGUMENT,
                  "slice sets: bad sizes");
    const int64_t num_slices = ceildiv(n_rows, slice_size);
    GKOMI_REQUIRE(slice_sets != nullptr, GKOMI_INVALID_ARGUMENT, "slice sets: null output");
    GKOMI_REQUIRE(ws != nullptr && wsb >= scan_ws_bytes(num_slices, sizeof(int64_t)),
                  GKOMI_INVALID_ARGUMENT, "slice sets: workspace too small");
    if (num_slices > 0) {
        GKOMI_REQUIRE(slice_lengths && row_ptrs, GKOMI_INVALID_ARGUMENT,
                      "slice sets: null pointer");
        slice_lengths_kernel<<<grid_for(num_slices, kBlock), kBlock, 0, stream_of(ctx)>>>(
            n_rows, row_ptrs, slice_size, stride_factor, num_slices, slice_lengths);
        GKOMI_CHECK_LAUNCH("slice_lengths_kernel");
    }
    return exclusive_scan<int64_t>(stream_of(ctx), num_slices, slice_lengths, slice_sets, ws);
}

int gkomi_csr_to_sellp_f64(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t slice_size,
                           const int64_t* slice_sets, const int64_t* slice_lengths,
                           int32_t* s_cols, double* s_vals)
{
    return csr_to_sellp_impl(ctx, a, slice_size, slice_sets, slice_lengths, s_cols, s_vals);
}

int gkomi_csr_to_sellp_f32(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t slice_size,
                           const int64_t* slice_sets, const int64_t* slice_lengths,
                           int32_t* s_cols, float* s_vals)
{
    return csr_to_sellp_impl(ctx, a, slice_size, slice_sets, slice_lengths, s_cols, s_vals);
}

int gkomi_sellp_col_offsets(const gkomi_ctx* ctx, const gkomi_sellp* a, int16_t* out,
                            int32_t* fits)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_sellp(a)) return rc;
    GKOMI_REQUIRE(a->slice_size == kWave && a->stride_factor == 1, GKOMI_NOT_SUPPORTED,
                  "sellp col_offsets: needs slice size 64 and stride factor 1");
    GKOMI_REQUIRE(fits != nullptr, GKOMI_INVALID_ARGUMENT, "sellp col_offsets: null output");
    static_assert(sizeof(int32_t) == sizeof(float), "fill by bit pattern");
    if (int rc = gkomi_fill_f32(ctx, 1, __builtin_bit_cast(float, int32_t{1}),
                                reinterpret_cast<float*>(fits))) {
        return rc;
    }
    if (a->n_rows == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(out != nullptr && a->col_idxs != nullptr, GKOMI_INVALID_ARGUMENT,
                  "sellp col_offsets: null pointer");
    const int64_t num_slices = ceildiv(a->n_rows, kWave);
    sellp_col_offsets_kernel<<<grid_for(num_slices, kWavesPerBlock, 16), kBlock, 0,
                               stream_of(ctx)>>>(a->n_rows, num_slices, a->slice_sets,
                                                 a->col_idxs, out, fits);
    GKOMI_CHECK_LAUNCH("sellp_col_offsets_kernel");
    return GKOMI_SUCCESS;
}

int gkomi_sellp_spmv_f64(const gkomi_ctx* ctx, const gkomi_sellp* a, int64_t nrhs,
                         const double* b, int64_t bs, double* c, int64_t cs)
{
    return sellp_spmv_impl<double, double, double, false>(ctx, nullptr, a, nrhs, b, bs, nullptr, c, cs);
}

int gkomi_sellp_spmv_f32(const gkomi_ctx* ctx, const gkomi_sellp* a, int64_t nrhs,
                         const float* b, int64_t bs, float* c, int64_t cs)
{
    return sellp_spmv_impl<float, float, float, false>(ctx, nullptr, a, nrhs, b, bs, nullptr, c, cs);
}

int gkomi_sellp_advanced_spmv_f64(const gkomi_ctx* ctx, const double* alpha,
                                  const gkomi_sellp* a, int64_t nrhs, const double* b,
                                  int64_t bs, const double* beta, double* c, int64_t cs)
{
    return sellp_spmv_impl<double, double, double, true>(ctx, alpha, a, nrhs, b, bs, beta, c, cs);
}

int gkomi_sellp_advanced_spmv_f32(const gkomi_ctx* ctx, const float* alpha,
                                  const gkomi_sellp* a, int64_t nrhs, const float* b,
                                  int64_t bs, const float* beta, float* c, int64_t cs)
{
    return sellp_spmv_impl<float, float, float, true>(ctx, alpha, a, nrhs, b, bs, beta, c, cs);
}

int gkomi_csr_to_ell_f64(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t width,
                         int64_t stride, int32_t* e_cols, double* e_vals)
{
    return csr_to_ell_impl(ctx, a, width, stride, e_cols, e_vals);
}

int gkomi_csr_to_ell_f32(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t width,
                         int64_t stride, int32_t* e_cols, float* e_vals)
{
    return csr_to_ell_impl(ctx, a, width, stride, e_cols, e_vals);
}

int gkomi_ell_spmv_f64(const gkomi_ctx* ctx, const gkomi_ell* a, int64_t nrhs, const double* b,
                       int64_t bs, double* c, int64_t cs)
{
    return ell_spmv_impl<double, double, double, false>(ctx, nullptr, a, nrhs, b, bs, nullptr, c, cs);
}

int gkomi_ell_spmv_f32(const gkomi_ctx* ctx, const gkomi_ell* a, int64_t nrhs, const float* b,
                       int64_t bs, float* c, int64_t cs)
{
    return ell_spmv_impl<float, float, float, false>(ctx, nullptr, a, nrhs, b, bs, nullptr, c, cs);
}

int gkomi_ell_advanced_spmv_f64(const gkomi_ctx* ctx, const double* alpha, const gkomi_ell* a,
                                int64_t nrhs, const double* b, int64_t bs, const double* beta,
                                double* c, int64_t cs)
{
    return ell_spmv_impl<double, double, double, true>(ctx, alpha, a, nrhs, b, bs, beta, c, cs);
}

int gkomi_ell_advanced_spmv_f32(const gkomi_ctx* ctx, const float* alpha, const gkomi_ell* a,
                                int64_t nrhs, const float* b, int64_t bs, const float* beta,
                                float* c, int64_t cs)
{
    return ell_spmv_impl<float, float, float, true>(ctx, alpha, a, nrhs, b, bs, beta, c, cs);
}

size_t gkomi_csr_transpose_workspace_bytes(int64_t n_rows, int64_t n_cols, int64_t nnz)
{
    return transpose_ws_layout(n_rows, n_cols, nnz, nullptr, nullptr);
}

int gkomi_csr_transpose_f64(const gkomi_ctx* ctx, const gkomi_csr* a, int32_t* t_rp,
                            int32_t* t_ci, double* t_v, void* ws, size_t wsb)
{
    return transpose_impl(ctx, a, t_rp, t_ci, t_v, ws, wsb);
}

int gkomi_csr_transpose_f32(const gkomi_ctx* ctx, const gkomi_csr* a, int32_t* t_rp,
                            int32_t* t_ci, float* t_v, void* ws, size_t wsb)
{
    return transpose_impl(ctx, a, t_rp, t_ci, t_v, ws, wsb);
}


// Products with mixed value types (reference/matrix/sellp_kernels.cpp:26-95,
// reference/matrix/ell_kernels.cpp:29-110, instantiated for every mixed
// value type): a_type, b_type, c_type are GKOMI_VALUE_F64 / _F32.
int gkomi_sellp_spmv_mixed(const gkomi_ctx* ctx, int32_t a_type, const gkomi_sellp* a,
                           int64_t nrhs, int32_t b_type, const void* b, int64_t b_stride,
                           int32_t c_type, void* c, int64_t c_stride)
{
    return dispatch_value_types(a_type, b_type, c_type, [&](auto ta, auto tb, auto tc) {
        using TA = decltype(ta);
        using TB = decltype(tb);
        using TC = decltype(tc);
        return sellp_spmv_impl<TA, TB, TC, false>(ctx, nullptr, a, nrhs,
                                                  static_cast<const TB*>(b), b_stride, nullptr,
                                                  static_cast<TC*>(c), c_stride);
    });
}

int gkomi_sellp_advanced_spmv_mixed(const gkomi_ctx* ctx, const void* alpha, int32_t a_type,
                                    const gkomi_sellp* a, int64_t nrhs, int32_t b_type,
                                    const void* b, int64_t b_stride, const void* beta,
                                    int32_t c_type, void* c, int64_t c_stride)
{
    return dispatch_value_types(a_type, b_type, c_type, [&](auto ta, auto tb, auto tc) {
        using TA = decltype(ta);
        using TB = decltype(tb);
        using TC = decltype(tc);
        return sellp_spmv_impl<TA, TB, TC, true>(
            ctx, static_cast<const TA*>(alpha), a, nrhs, static_cast<const TB*>(b), b_stride,
            static_cast<const TC*>(beta), static_cast<TC*>(c), c_stride);
    });
}

int gkomi_ell_spmv_mixed(const gkomi_ctx* ctx, int32_t a_type, const gkomi_ell* a, int64_t nrhs,
                         int32_t b_type, const void* b, int64_t b_stride, int32_t c_type,
                         void* c, int64_t c_stride)
{
    return dispatch_value_types(a_type, b_type, c_type, [&](auto ta, auto tb, auto tc) {
        using TA = decltype(ta);
        using TB = decltype(tb);
        using TC = decltype(tc);
        return ell_spmv_impl<TA, TB, TC, false>(ctx, nullptr, a, nrhs,
                                                static_cast<const TB*>(b), b_stride, nullptr,
                                                static_cast<TC*>(c), c_stride);
    });
}

int gkomi_ell_advanced_spmv_mixed(const gkomi_ctx* ctx, const void* alpha, int32_t a_type,
                                  const gkomi_ell* a, int64_t nrhs, int32_t b_type,
                                  const void* b, int64_t b_stride, const void* beta,
                                  int32_t c_type, void* c, int64_t c_stride)
{
    return dispatch_value_types(a_type, b_type, c_type, [&](auto ta, auto tb, auto tc) {
        using TA = decltype(ta);
        using TB = decltype(tb);
        using TC = decltype(tc);
        return ell_spmv_impl<TA, TB, TC, true>(
            ctx, static_cast<const TA*>(alpha), a, nrhs, static_cast<const TB*>(b), b_stride,
            static_cast<const TC*>(beta), static_cast<TC*>(c), c_stride);
    });
}

}  // extern "C"
