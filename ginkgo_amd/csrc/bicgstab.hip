// Fused BiCGSTAB for one right-hand side on gfx950.
//
// Bicgstab::apply_dense_impl (core/solver/bicgstab.cpp:95-255) with the
// reference step semantics (reference/solver/bicgstab_kernels.cpp:26-185),
// five kernels per iteration and no host round trip inside the loop:
//   K1 p      rho = rr.r and ||r|| from the previous partials, the Combined
//             check of the iteration (finalizing), then
//             p = r + rho/prev_rho * alpha/omega * (p - omega v) and y = M p;
//   K2 v      v = A y and the partials of rr.v;
//   K3 s      beta = rr.v, alpha = rho / beta, s = r - alpha v, z = M s and
//             the partials of s.s;
//   K4 t      ||s|| and the half-step check (not finalizing); unless it
//             stopped, t = A z and the partials of s.t and t.t;
//   K5 update a half-step stop finalizes x += alpha y; otherwise
//             omega = s.t / t.t, x += alpha y + omega z, r = s - omega t
//             and the partials of rr.r and r.r for the next K1.
// Per iteration the kernels move 2 nnz (V + 4) + 8 (n + 1) + 25 n V bytes
// with scalar Jacobi (21 n V without), against the reference's
// 2 (matrix + n V) + 31 n V (core/solver/bicgstab.cpp:154-164).
//
// Scalars (value type): [0], [1] rho of even / odd iterations (slot 1 starts
// as prev_rho = 1), [2] alpha, [3] omega, [4], [5] ||r|| of even / odd
// iterations, [6] ||s|| of the half step that stopped the solve.  No
// workgroup reads a value that another workgroup of the same kernel writes.
#include <type_traits>

#include "common.hpp"
#include "fused.hpp"
#include "spmv_common.hpp"
#include "stop.hpp"

namespace gkomi {

namespace {

struct bcs_work {
    double* part_init;   // [kCgMaxGrid]: b.b of the start
    double* part_r;      // [kCgMaxGrid][2]: (rr.r, r.r) of the start / last update
    double* part_v;      // [kCgMaxGrid]: rr.v
    double* part_s;      // [kCgMaxGrid]: s.s
    double* part_t;      // [kCgMaxGrid][2]: (s.t, t.t)
    int64_t* half_stop;  // iteration whose half-step check stopped the solve, or -1
    uint8_t* status;     // [3]: after the check of even / odd iterations, after the half step
};

__host__ __device__ inline bcs_work bcs_work_layout(void* base)
{
    bcs_work w;
    double* d = static_cast<double*>(base);
    w.part_init = d;
    w.part_r = d + kCgMaxGrid;
    w.part_v = d + 3 * kCgMaxGrid;
    w.part_s = d + 4 * kCgMaxGrid;
    w.part_t = d + 5 * kCgMaxGrid;
    w.half_stop = reinterpret_cast<int64_t*>(d + 7 * kCgMaxGrid);
    w.status = reinterpret_cast<uint8_t*>(d + 7 * kCgMaxGrid + 1);
    return w;
}

constexpr size_t kBcsWorkBytes = (7 * kCgMaxGrid + 2) * sizeof(double);

// ------------------------------------------------------------- start ----
// r = b - A x, rr = r, p = v = 0 (reference/solver/bicgstab_kernels.cpp:
// 26-58 and core/solver/bicgstab.cpp:139-146); partials of b.b and of
// (rr.r, r.r) = (r.r, r.r); scalars, status and flags reset.
template <typename T, int SW, int RPG, bool Compact>
__global__ void __launch_bounds__(kBlock)
    bcs_init_kernel(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                    const int16_t* __restrict__ co, const T* __restrict__ av,
                    const T* __restrict__ b, const T* __restrict__ x, T* __restrict__ r,
                    T* __restrict__ rr, T* __restrict__ p, T* __restrict__ v,
                    T* __restrict__ scalars, uint8_t* __restrict__ status,
                    gkomi_stop_flags* __restrict__ flags, bcs_work w)
{
    int64_t rb0, rb1;
    block_range(ceildiv(n, subwave_tile_rows<SW, RPG>), &rb0, &rb1);
    T bb[1] = {T(0)};
    T rs[2] = {T(0), T(0)};
    for (int64_t blk = rb0; blk < rb1; ++blk) {
        subwave_tile<T, SW, RPG, Compact>(
            n, blk, rp, ci, co, av, [&](int32_t c) { return x[c]; },
            [&](int64_t row, T sum) {
                const T bi = b[row];
                const T ri = bi - sum;
                r[row] = ri;
                rr[row] = ri;
                p[row] = T(0);
                v[row] = T(0);
                bb[0] += bi * bi;
                rs[0] += ri * ri;
            });
    }
    rs[1] = rs[0];
    store_partials<T, 1>(bb, w.part_init);
    store_partials<T, 2>(rs, w.part_r);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scalars[0] = T(1);
        scalars[1] = T(1);  // prev_rho of iteration 0
        scalars[2] = T(1);  // alpha
        scalars[3] = T(1);  // omega
        scalars[4] = T(0);
        scalars[5] = T(0);
        scalars[6] = T(0);
        status[0] = 0;
        w.status[0] = 0;
        w.status[1] = 0;
        w.status[2] = 0;
        *w.half_stop = -1;
        flags->iteration = -1;
        flags->all_stopped = 0;
        flags->one_changed = 0;
    }
}

// One workgroup: the criteria baselines (core/stop/residual_norm.cpp:93-162)
// from ||b|| and ||r0||.
template <typename T>
__global__ void __launch_bounds__(kUpdBlock)
    bcs_baseline_kernel(gkomi_criteria crit, T* __restrict__ orig, bcs_work w, int count)
{
    double bb[1], rs[2];
    combine_partials<1, kUpdBlock>(w.part_init, count, bb);
    combine_partials<2, kUpdBlock>(w.part_r, count, rs);
    if (threadIdx.x == 0) {
        for (int k = 0; k < crit.num; ++k) {
            const int base = crit.c[k].baseline;
            orig[k] = base == GKOMI_BASELINE_RHS_NORM
                          ? static_cast<T>(sqrt(bb[0]))
                          : (base == GKOMI_BASELINE_INITIAL_RESNORM ? static_cast<T>(sqrt(rs[1]))
                                                                    : T(1));
        }
    }
}

// ---------------------------------------------------------- iteration ----
// K1 (core/solver/bicgstab.cpp:166-188; step_1,
// reference/solver/bicgstab_kernels.cpp:63-88).
template <typename T, bool Jacobi>
__global__ void __launch_bounds__(kUpdBlock)
    bcs_p_kernel(int64_t n, const T* __restrict__ r, T* __restrict__ p, const T* __restrict__ v,
                 T* __restrict__ y, const T* __restrict__ inv_diag, gkomi_criteria crit,
                 T* __restrict__ scalars, const T* __restrict__ orig,
                 uint8_t* __restrict__ status, gkomi_stop_flags* __restrict__ flags, bcs_work w,
                 int count_r, int64_t iter, int par)
{
    // the stop flag (thread 0; workgroup 0 writes it below), the scalars,
    // the status and this thread's first element are loaded together with
    // the partials, and the flag reaches the workgroup through the
    // combine's barrier
    int64_t i0, i1;
    block_range(ceildiv(n, kUpdBlock), &i0, &i1);
    const int64_t ia = i0 * kUpdBlock + threadIdx.x;
    const bool has_a = i0 < i1 && ia < n;
    T ra = T(0), pa = T(0), va = T(0), da = T(1);
    if (has_a) {
        ra = r[ia];
        pa = p[ia];
        va = v[ia];
        if constexpr (Jacobi) da = inv_diag[ia];
    }
    const int flag0 = threadIdx.x == 0 ? flags->all_stopped : 0;
    uint8_t st = w.status[par ^ 1];
    const T prev = scalars[par ^ 1];
    const T alpha = scalars[2];
    const T omega = scalars[3];
    double s[2];
    if (combine_partials_bcast<2, kUpdBlock>(w.part_r, count_r, s, flag0)) return;
    const T rho = static_cast<T>(s[0]);
    const T tau = static_cast<T>(sqrt(s[1]));
    int all = 0, changed = 0;
    combined_check<T>(crit, iter, 1, &tau, &rho, orig, true, &st, &all, &changed);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scalars[par] = rho;
        scalars[4 + par] = tau;
        w.status[par] = st;
        status[0] = st;
        flags->iteration = iter;
        flags->all_stopped = all;
        flags->one_changed = changed;
    }
    if (all) return;
    const bool go = prev * omega != T(0);
    const T tmp = go ? rho / prev * alpha / omega : T(0);
    for (int64_t blk = i0; blk < i1; ++blk) {
        const int64_t i = blk * kUpdBlock + threadIdx.x;
        if (i >= n) continue;
        const bool first = blk == i0;
        const T ri = first ? ra : r[i];
        const T pi = go ? ri + tmp * ((first ? pa : p[i]) - omega * (first ? va : v[i])) : ri;
        p[i] = pi;
        if (Jacobi) y[i] = pi * (first ? da : inv_diag[i]);
    }
}

// K2: v = A y and the partials of beta = rr.v (bicgstab.cpp:190-193).
template <typename T, int SW, int RPG, bool Compact>
__global__ void __launch_bounds__(kBlock)
    bcs_v_kernel(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                 const int16_t* __restrict__ co, const T* __restrict__ av,
                 const T* __restrict__ y, const T* __restrict__ rr, T* __restrict__ v,
                 const gkomi_stop_flags* __restrict__ flags, bcs_work w)
{
    // the first tile's loads go out before the stop flag is known
    int64_t rb0, rb1;
    block_range(ceildiv(n, subwave_tile_rows<SW, RPG>), &rb0, &rb1);
    subwave_pf<T, RPG> pf;
    if (rb0 < rb1) subwave_prefetch<T, SW, RPG, Compact>(n, rb0, rp, ci, co, av, pf);
    if (flags->all_stopped) return;
    T acc[1] = {T(0)};
    auto gather = [&](int32_t c) { return y[c]; };
    auto row_op = [&](int64_t row, T sum) {
        v[row] = sum;
        acc[0] += rr[row] * sum;
    };
    if (rb0 < rb1) subwave_finish<T, SW, RPG, Compact>(n, ci, co, av, pf, gather, row_op);
    for (int64_t blk = rb0 + 1; blk < rb1; ++blk) {
        subwave_tile<T, SW, RPG, Compact>(n, blk, rp, ci, co, av, gather, row_op);
    }
    store_partials<T, 1>(acc, w.part_v);
}

// K3 (step_2, reference/solver/bicgstab_kernels.cpp:93-117): alpha = rho /
// beta (0 when beta == 0, s = r then), s = r - alpha v, z = M s, partials
// of s.s.
template <typename T, bool Jacobi>
__global__ void __launch_bounds__(kUpdBlock)
    bcs_s_kernel(int64_t n, const T* __restrict__ r, const T* __restrict__ v,
                 T* __restrict__ s, T* __restrict__ z, const T* __restrict__ inv_diag,
                 T* __restrict__ scalars, const gkomi_stop_flags* __restrict__ flags,
                 bcs_work w, int count_v, int par)
{
    // no kernel writes the flag while this one runs: every thread reads it,
    // together with rho and its first element, before the combine
    int64_t i0, i1;
    block_range(ceildiv(n, kUpdBlock), &i0, &i1);
    const int64_t ia = i0 * kUpdBlock + threadIdx.x;
    const bool has_a = i0 < i1 && ia < n;
    T ra = T(0), va = T(0), da = T(1);
    if (has_a) {
        ra = r[ia];
        va = v[ia];
        if constexpr (Jacobi) da = inv_diag[ia];
    }
    const int stopped = flags->all_stopped;
    const T rho = scalars[par];
    double bsum[1];
    combine_partials<1, kUpdBlock>(w.part_v, count_v, bsum);
    if (stopped) return;
    const T beta = static_cast<T>(bsum[0]);
    const bool go = beta != T(0);
    const T alpha = go ? rho / beta : T(0);
    if (blockIdx.x == 0 && threadIdx.x == 0) scalars[2] = alpha;
    T acc[1] = {T(0)};
    for (int64_t blk = i0; blk < i1; ++blk) {
        const int64_t i = blk * kUpdBlock + threadIdx.x;
        if (i >= n) continue;
        const bool first = blk == i0;
        const T ri = first ? ra : r[i];
        const T si = go ? ri - alpha * (first ? va : v[i]) : ri;
        s[i] = si;
        if (Jacobi) z[i] = si * (first ? da : inv_diag[i]);
        acc[0] += si * si;
    }
    store_partials<T, 1, kUpdBlock>(acc, w.part_s);
}

// K4: the half-step check on ||s|| (bicgstab.cpp:204-212, not finalizing);
// unless it stopped, t = A z and the partials of gamma = s.t and t.t
// (bicgstab.cpp:223-230).
template <typename T, int SW, int RPG, bool Compact>
__global__ void __launch_bounds__(kBlock)
    bcs_t_kernel(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                 const int16_t* __restrict__ co, const T* __restrict__ av,
                 const T* __restrict__ z, const T* __restrict__ s, T* __restrict__ t,
                 gkomi_criteria crit, T* __restrict__ scalars, const T* __restrict__ orig,
                 uint8_t* __restrict__ status, gkomi_stop_flags* __restrict__ flags,
                 bcs_work w, int count_s, int64_t iter, int par)
{
    int64_t rb0, rb1;
    block_range(ceildiv(n, subwave_tile_rows<SW, RPG>), &rb0, &rb1);
    subwave_pf<T, RPG> pf;
    if (rb0 < rb1) subwave_prefetch<T, SW, RPG, Compact>(n, rb0, rp, ci, co, av, pf);
    const int flag0 = threadIdx.x == 0 ? flags->all_stopped : 0;
    const T rho = scalars[par];
    uint8_t st = w.status[par];
    double ss[1];
    if (combine_partials_bcast<1>(w.part_s, count_s, ss, flag0)) return;
    const T tau = static_cast<T>(sqrt(ss[0]));
    int all = 0, changed = 0;
    combined_check<T>(crit, iter, 1, &tau, &rho, orig, false, &st, &all, &changed);
    if (all) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            scalars[6] = tau;
            w.status[2] = st;
            status[0] = st;
            *w.half_stop = iter;
            flags->iteration = iter;
            flags->one_changed = changed;
            flags->all_stopped = 1;
        }
        return;
    }
    T acc[2] = {T(0), T(0)};
    auto gather = [&](int32_t c) { return z[c]; };
    auto row_op = [&](int64_t row, T sum) {
        t[row] = sum;
        acc[0] += s[row] * sum;
        acc[1] += sum * sum;
    };
    if (rb0 < rb1) subwave_finish<T, SW, RPG, Compact>(n, ci, co, av, pf, gather, row_op);
    for (int64_t blk = rb0 + 1; blk < rb1; ++blk) {
        subwave_tile<T, SW, RPG, Compact>(n, blk, rp, ci, co, av, gather, row_op);
    }
    store_partials<T, 2>(acc, w.part_t);
}

// K2 and K4 over the staged product (staged_window, G lanes per row; rows
// of stencil length), one window per wave: the same sums, checks and
// partials as bcs_v_kernel and bcs_t_kernel.
template <typename T, int G, bool Compact>
__global__ void __launch_bounds__(kBlock)
    bcs_v_staged_kernel(int64_t n, const int32_t* __restrict__ rp,
                        const int32_t* __restrict__ ci, const int16_t* __restrict__ co,
                        const T* __restrict__ av, const T* __restrict__ y,
                        const T* __restrict__ rr, T* __restrict__ v,
                        const gkomi_stop_flags* __restrict__ flags, bcs_work w)
{
    using Idx = std::conditional_t<Compact, int16_t, int32_t>;
    __shared__ T sv[kWavesPerBlock][kStagedChunk];
    __shared__ Idx sc[kWavesPerBlock][kStagedChunk];
    if (flags->all_stopped) return;
    T acc[1] = {T(0)};
    const int wave = threadIdx.x / kWave;
    int64_t rb0, rb1;
    block_range(ceildiv(n, staged_tile_rows<G>), &rb0, &rb1);
    for (int64_t blk = rb0; blk < rb1; ++blk) {
        staged_window<T, G, kStagedChunk, Compact>(
            n, blk * kWavesPerBlock + wave, rp, ci, co, av, sv[wave], sc[wave],
            [&](int32_t c) { return y[c]; },
            [&](int64_t row, T sum) {
                v[row] = sum;
                acc[0] += rr[row] * sum;
            });
    }
    store_partials<T, 1>(acc, w.part_v);
}

template <typename T, int G, bool Compact>
__global__ void __launch_bounds__(kBlock)
    bcs_t_staged_kernel(int64_t n, const int32_t* __restrict__ rp,
                        const int32_t* __restrict__ ci, const int16_t* __restrict__ co,
                        const T* __restrict__ av, const T* __restrict__ z,
                        const T* __restrict__ s, T* __restrict__ t, gkomi_criteria crit,
                        T* __restrict__ scalars, const T* __restrict__ orig,
                        uint8_t* __restrict__ status, gkomi_stop_flags* __restrict__ flags,
                        bcs_work w, int count_s, int64_t iter, int par)
{
    using Idx = std::conditional_t<Compact, int16_t, int32_t>;
    __shared__ T sv[kWavesPerBlock][kStagedChunk];
    __shared__ Idx sc[kWavesPerBlock][kStagedChunk];
    const int flag0 = threadIdx.x == 0 ? flags->all_stopped : 0;
    const T rho = scalars[par];
    uint8_t st = w.status[par];
    double ss[1];
    if (combine_partials_bcast<1>(w.part_s, count_s, ss, flag0)) return;
    const T tau = static_cast<T>(sqrt(ss[0]));
    int all = 0, changed = 0;
    combined_check<T>(crit, iter, 1, &tau, &rho, orig, false, &st, &all, &changed);
    if (all) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            scalars[6] = tau;
            w.status[2] = st;
            status[0] = st;
            *w.half_stop = iter;
            flags->iteration = iter;
            flags->one_changed = changed;
            flags->all_stopped = 1;
        }
        return;
    }
    T acc[2] = {T(0), T(0)};
    const int wave = threadIdx.x / kWave;
    int64_t rb0, rb1;
    block_range(ceildiv(n, staged_tile_rows<G>), &rb0, &rb1);
    for (int64_t blk = rb0; blk < rb1; ++blk) {
        staged_window<T, G, kStagedChunk, Compact>(
            n, blk * kWavesPerBlock + wave, rp, ci, co, av, sv[wave], sc[wave],
            [&](int32_t c) { return z[c]; },
            [&](int64_t row, T sum) {
                t[row] = sum;
                acc[0] += s[row] * sum;
                acc[1] += sum * sum;
            });
    }
    store_partials<T, 2>(acc, w.part_t);
}

// K5 (step_3, reference/solver/bicgstab_kernels.cpp:122-152; finalize,
// :157-172, after a half-step stop).
template <typename T>
__global__ void __launch_bounds__(kUpdBlock)
    bcs_update_kernel(int64_t n, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ rr,
                      const T* __restrict__ s, const T* __restrict__ t,
                      const T* __restrict__ y, const T* __restrict__ z,
                      T* __restrict__ scalars, uint8_t* __restrict__ status,
                      const gkomi_stop_flags* __restrict__ flags, bcs_work w, int count_t,
                      int64_t iter)
{
    // mode 0 update, 1 nothing to do, 2 finalize: decided by thread 0 and
    // handed over by the combine's barrier, with alpha and this thread's
    // first element loaded before it
    int64_t i0, i1;
    block_range(ceildiv(n, kUpdBlock), &i0, &i1);
    const int64_t ia = i0 * kUpdBlock + threadIdx.x;
    const bool has_a = i0 < i1 && ia < n;
    T ya = T(0), za = T(0), sa = T(0), ta = T(0), rra = T(0);
    if (has_a) {
        ya = y[ia];
        za = z[ia];
        sa = s[ia];
        ta = t[ia];
        rra = rr[ia];
    }
    int mode0 = 0;
    if (threadIdx.x == 0) mode0 = flags->all_stopped ? (*w.half_stop == iter ? 2 : 1) : 0;
    const T alpha = scalars[2];
    double ts[2];
    const int mode = combine_partials_bcast<2, kUpdBlock>(w.part_t, count_t, ts, mode0);
    if (mode == 1) return;
    if (mode == 2) {
        for (int64_t blk = i0; blk < i1; ++blk) {
            const int64_t i = blk * kUpdBlock + threadIdx.x;
            if (i < n) x[i] += alpha * y[i];
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            status[0] = static_cast<uint8_t>(w.status[2] | GKOMI_STATUS_FINALIZED);
        }
        return;
    }
    const T gamma = static_cast<T>(ts[0]);
    const T beta = static_cast<T>(ts[1]);
    const T omega = beta != T(0) ? gamma / beta : T(0);
    if (blockIdx.x == 0 && threadIdx.x == 0) scalars[3] = omega;
    T acc[2] = {T(0), T(0)};
    for (int64_t blk = i0; blk < i1; ++blk) {
        const int64_t i = blk * kUpdBlock + threadIdx.x;
        if (i >= n) continue;
        const bool first = blk == i0;
        x[i] += alpha * (first ? ya : y[i]) + omega * (first ? za : z[i]);
        const T ri = (first ? sa : s[i]) - omega * (first ? ta : t[i]);
        r[i] = ri;
        acc[0] += (first ? rra : rr[i]) * ri;
        acc[1] += ri * ri;
    }
    store_partials<T, 2, kUpdBlock>(acc, w.part_r);
}

// ------------------------------------------------------------ drivers ----

int check_bcs_state(const gkomi_bicgstab_state* st)
{
    GKOMI_REQUIRE(st != nullptr, GKOMI_INVALID_ARGUMENT, "null BiCGSTAB state");
    GKOMI_REQUIRE(st->n >= 0, GKOMI_INVALID_ARGUMENT, "BiCGSTAB state: negative size");
    GKOMI_REQUIRE(st->x && st->b && st->r && st->rr && st->p && st->v && st->s && st->t &&
                      st->scalars && st->orig_tau && st->status && st->flags && st->work,
                  GKOMI_INVALID_ARGUMENT, "BiCGSTAB state: null pointer");
    GKOMI_REQUIRE(!st->inv_diag || (st->y && st->z), GKOMI_INVALID_ARGUMENT,
                  "BiCGSTAB state: a preconditioner needs the y and z buffers");
    return GKOMI_SUCCESS;
}

int check_bcs_call(const gkomi_ctx* ctx, const gkomi_csr* a, const gkomi_criteria* crit,
                   const gkomi_bicgstab_state* st)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_csr(a)) return rc;
    if (int rc = check_crit(crit)) return rc;
    if (int rc = check_bcs_state(st)) return rc;
    GKOMI_REQUIRE(a->n_rows == a->n_cols && a->n_rows == st->n, GKOMI_DIMENSION_MISMATCH,
                  "fused BiCGSTAB: matrix and vectors do not conform");
    GKOMI_REQUIRE(a->strategy == GKOMI_CSR_CLASSICAL, GKOMI_NOT_SUPPORTED,
                  "fused BiCGSTAB: needs the classical (sub-wave) CSR strategy");
    return GKOMI_SUCCESS;
}

struct bcs_grids {
    subwave_shape sh;
    int staged;  // lanes per row of the staged K2 / K4, 0 = sub-wave
    int spmv;    // init, K2, K4
    int update;  // K1, K3, K5
};

template <typename T>
bcs_grids bcs_grids_for(const gkomi_csr* a, int64_t n)
{
    bcs_grids g;
    g.sh = pick_subwave(a, sizeof(T));
    // the staged product where the classical SpMV takes it (rows of
    // stencil length)
    g.staged = std::max(0, staged_lanes(a, sizeof(T)));
    const int64_t tile_rows = g.staged > 0 ? int64_t{kWave} / g.staged * kWavesPerBlock
                                           : int64_t{kBlock} / g.sh.sw * g.sh.rpg;
    g.spmv = fused_grid(ceildiv(std::max<int64_t>(n, 1), tile_rows));
    if (g.staged > 0) {
        // no more workgroups than are resident at once (four staging areas
        // per workgroup), as the fused CG
        const int64_t per_wg =
            int64_t{kWavesPerBlock} * kStagedChunk *
                (static_cast<int64_t>(sizeof(T)) + (a->col_offsets != nullptr ? 2 : 4)) +
            2048;
        const int64_t resident = std::max<int64_t>(1, std::min<int64_t>(8, (160 * 1024) / per_wg));
        g.spmv = static_cast<int>(std::min<int64_t>(g.spmv, resident * num_cus()));
    }
    g.update = static_cast<int>(std::max<int64_t>(
        1, std::min<int64_t>(ceildiv(std::max<int64_t>(n, 1), kUpdBlock), kCgUpdGrid)));
    return g;
}

template <typename T>
int bcs_start_impl(const gkomi_ctx* ctx, const gkomi_csr* a, const gkomi_criteria* crit,
                   gkomi_bicgstab_state* st)
{
    if (int rc = check_bcs_call(ctx, a, crit, st)) return rc;
    hipStream_t s = stream_of(ctx);
    const bcs_work w = bcs_work_layout(st->work);
    const T* av = static_cast<const T*>(a->values);
    T* sc = static_cast<T*>(st->scalars);
    const int64_t n = st->n;
    const bcs_grids g = bcs_grids_for<T>(a, n);
#define GKOMI_BCS_INIT_C(S, R, C)                                                            \
    bcs_init_kernel<T, S, R, C><<<g.spmv, kBlock, 0, s>>>(                                   \
        n, a->row_ptrs, a->col_idxs, a->col_offsets, av, static_cast<const T*>(st->b),       \
        static_cast<const T*>(st->x), static_cast<T*>(st->r), static_cast<T*>(st->rr),       \
        static_cast<T*>(st->p), static_cast<T*>(st->v), sc, st->status, st->flags, w);       \
    break;
#define GKOMI_BCS_INIT(S, R) GKOMI_BCS_INIT_C(S, R, false)
#define GKOMI_BCS_INIT_COMPACT(S, R) GKOMI_BCS_INIT_C(S, R, true)
    if (a->col_offsets) {
        GKOMI_SUBWAVE_SWITCH(g.sh, GKOMI_BCS_INIT_COMPACT)
    } else {
        GKOMI_SUBWAVE_SWITCH(g.sh, GKOMI_BCS_INIT)
    }
#undef GKOMI_BCS_INIT
#undef GKOMI_BCS_INIT_COMPACT
#undef GKOMI_BCS_INIT_C
    GKOMI_CHECK_LAUNCH("bcs_init_kernel");
    bcs_baseline_kernel<T><<<1, kUpdBlock, 0, s>>>(*crit, static_cast<T*>(st->orig_tau), w,
                                                   g.spmv);
    GKOMI_CHECK_LAUNCH("bcs_baseline_kernel");
    st->issued = 0;
    return GKOMI_SUCCESS;
}

template <typename T>
int bcs_iterate_impl(const gkomi_ctx* ctx, const gkomi_csr* a, const gkomi_criteria* crit,
                     gkomi_bicgstab_state* st, int32_t count)
{
    if (int rc = check_bcs_call(ctx, a, crit, st)) return rc;
    GKOMI_REQUIRE(count >= 0, GKOMI_INVALID_ARGUMENT, "fused BiCGSTAB: negative count");
    GKOMI_REQUIRE(st->issued >= 0, GKOMI_INVALID_ARGUMENT,
                  "fused BiCGSTAB: state not started (issued < 0)");
    hipStream_t s = stream_of(ctx);
    const bcs_work w = bcs_work_layout(st->work);
    const T* av = static_cast<const T*>(a->values);
    T* sc = static_cast<T*>(st->scalars);
    const T* orig = static_cast<const T*>(st->orig_tau);
    const T* dinv = static_cast<const T*>(st->inv_diag);
    const int64_t n = st->n;
    const bcs_grids g = bcs_grids_for<T>(a, n);
    const bool jac = dinv != nullptr;
    T* x = static_cast<T*>(st->x);
    T* r = static_cast<T*>(st->r);
    const T* rr = static_cast<const T*>(st->rr);
    T* p = static_cast<T*>(st->p);
    T* v = static_cast<T*>(st->v);
    T* sv = static_cast<T*>(st->s);
    T* t = static_cast<T*>(st->t);
    // without a preconditioner y = p and z = s (reference: a copy)
    T* y = jac ? static_cast<T*>(st->y) : p;
    T* z = jac ? static_cast<T*>(st->z) : sv;
    for (int32_t it = 0; it < count; ++it) {
        const int64_t iter = st->issued;
        const int par = static_cast<int>(iter & 1);
        const int count_r = iter == 0 ? g.spmv : g.update;
        if (jac) {
            bcs_p_kernel<T, true><<<g.update, kUpdBlock, 0, s>>>(
                n, r, p, v, y, dinv, *crit, sc, orig, st->status, st->flags, w, count_r, iter,
                par);
        } else {
            bcs_p_kernel<T, false><<<g.update, kUpdBlock, 0, s>>>(
                n, r, p, v, y, nullptr, *crit, sc, orig, st->status, st->flags, w, count_r,
                iter, par);
        }
        GKOMI_CHECK_LAUNCH("bcs_p_kernel");
        if (g.staged > 0) {
#define GKOMI_BCS_V_STAGED(GV, C)                                                            \
    if (g.staged == GV) {                                                                    \
        bcs_v_staged_kernel<T, GV, C><<<g.spmv, kBlock, 0, s>>>(                             \
            n, a->row_ptrs, a->col_idxs, a->col_offsets, av, y, rr, v, st->flags, w);        \
    }
            if (a->col_offsets) {
                GKOMI_BCS_V_STAGED(1, true) GKOMI_BCS_V_STAGED(2, true)
                GKOMI_BCS_V_STAGED(4, true) GKOMI_BCS_V_STAGED(8, true)
                GKOMI_BCS_V_STAGED(16, true)
            } else {
                GKOMI_BCS_V_STAGED(1, false) GKOMI_BCS_V_STAGED(2, false)
                GKOMI_BCS_V_STAGED(4, false) GKOMI_BCS_V_STAGED(8, false)
                GKOMI_BCS_V_STAGED(16, false)
            }
#undef GKOMI_BCS_V_STAGED
        } else {
#define GKOMI_BCS_V_C(S, R, C)                                                               \
    bcs_v_kernel<T, S, R, C><<<g.spmv, kBlock, 0, s>>>(n, a->row_ptrs, a->col_idxs,          \
                                                       a->col_offsets, av, y, rr, v,         \
                                                       st->flags, w);                        \
    break;
#define GKOMI_BCS_V(S, R) GKOMI_BCS_V_C(S, R, false)
#define GKOMI_BCS_V_COMPACT(S, R) GKOMI_BCS_V_C(S, R, true)
        if (a->col_offsets) {
            GKOMI_SUBWAVE_SWITCH(g.sh, GKOMI_BCS_V_COMPACT)
        } else {
            GKOMI_SUBWAVE_SWITCH(g.sh, GKOMI_BCS_V)
        }
#undef GKOMI_BCS_V
#undef GKOMI_BCS_V_COMPACT
#undef GKOMI_BCS_V_C
        }
        GKOMI_CHECK_LAUNCH("bcs_v_kernel");
        if (jac) {
            bcs_s_kernel<T, true><<<g.update, kUpdBlock, 0, s>>>(n, r, v, sv, z, dinv, sc,
                                                                st->flags, w, g.spmv, par);
        } else {
            bcs_s_kernel<T, false><<<g.update, kUpdBlock, 0, s>>>(n, r, v, sv, z, nullptr, sc,
                                                                 st->flags, w, g.spmv, par);
        }
        GKOMI_CHECK_LAUNCH("bcs_s_kernel");
        if (g.staged > 0) {
#define GKOMI_BCS_T_STAGED(GV, C)                                                            \
    if (g.staged == GV) {                                                                    \
        bcs_t_staged_kernel<T, GV, C><<<g.spmv, kBlock, 0, s>>>(                             \
            n, a->row_ptrs, a->col_idxs, a->col_offsets, av, z, sv, t, *crit, sc, orig,      \
            st->status, st->flags, w, g.update, iter, par);                                  \
    }
            if (a->col_offsets) {
                GKOMI_BCS_T_STAGED(1, true) GKOMI_BCS_T_STAGED(2, true)
                GKOMI_BCS_T_STAGED(4, true) GKOMI_BCS_T_STAGED(8, true)
                GKOMI_BCS_T_STAGED(16, true)
            } else {
                GKOMI_BCS_T_STAGED(1, false) GKOMI_BCS_T_STAGED(2, false)
                GKOMI_BCS_T_STAGED(4, false) GKOMI_BCS_T_STAGED(8, false)
                GKOMI_BCS_T_STAGED(16, false)
            }
#undef GKOMI_BCS_T_STAGED
        } else {
#define GKOMI_BCS_T_C(S, R, C)                                                               \
    bcs_t_kernel<T, S, R, C><<<g.spmv, kBlock, 0, s>>>(                                      \
        n, a->row_ptrs, a->col_idxs, a->col_offsets, av, z, sv, t, *crit, sc, orig,          \
        st->status, st->flags, w, g.update, iter, par);                                      \
    break;
#define GKOMI_BCS_T(S, R) GKOMI_BCS_T_C(S, R, false)
#define GKOMI_BCS_T_COMPACT(S, R) GKOMI_BCS_T_C(S, R, true)
        if (a->col_offsets) {
            GKOMI_SUBWAVE_SWITCH(g.sh, GKOMI_BCS_T_COMPACT)
        } else {
            GKOMI_SUBWAVE_SWITCH(g.sh, GKOMI_BCS_T)
        }
#undef GKOMI_BCS_T
#undef GKOMI_BCS_T_COMPACT
#undef GKOMI_BCS_T_C
        }
        GKOMI_CHECK_LAUNCH("bcs_t_kernel");
        bcs_update_kernel<T><<<g.update, kUpdBlock, 0, s>>>(n, x, r, rr, sv, t, y, z, sc,
                                                            st->status, st->flags, w, g.spmv,
                                                            iter);
        GKOMI_CHECK_LAUNCH("bcs_update_kernel");
        ++st->issued;
    }
    return GKOMI_SUCCESS;
}

}  // namespace

}  // namespace gkomi

using namespace gkomi;

extern "C" {

size_t gkomi_bicgstab_work_bytes(int64_t n)
{
    (void)n;
    return kBcsWorkBytes;
}

#define GKOMI_BCS_ENTRY(T, S)                                                                  \
    int gkomi_bicgstab_fused_start_##S(const gkomi_ctx* ctx, const gkomi_csr* a,               \
                                       const gkomi_criteria* crit, gkomi_bicgstab_state* st)   \
    {                                                                                          \
        return bcs_start_impl<T>(ctx, a, crit, st);                                            \
    }                                                                                          \
    int gkomi_bicgstab_fused_iterate_##S(const gkomi_ctx* ctx, const gkomi_csr* a,             \
                                         const gkomi_criteria* crit, gkomi_bicgstab_state* st, \
                                         int32_t count)                                        \
    {                                                                                          \
        return bcs_iterate_impl<T>(ctx, a, crit, st, count);                                   \
    }

GKOMI_BCS_ENTRY(double, f64)
GKOMI_BCS_ENTRY(float, f32)
#undef GKOMI_BCS_ENTRY

}  // extern "C"
