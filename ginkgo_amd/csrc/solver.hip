// CG and BiCGSTAB kernels for gfx950.
//
// Two families:
//  * reference-semantics step kernels (cg::initialize/step_1/step_2,
//    bicgstab::initialize/step_1/2/3/finalize) over n x nrhs blocks with
//    per-column stopping status, restated from
//    reference/solver/{cg,bicgstab}_kernels.cpp;
//  * the fused single-RHS CG iteration: two kernels per iteration that fold
//    the axpy updates, the Jacobi scaling, both dots and ||r|| into the SpMV
//    and update passes, and evaluate stop::Combined on the device so the host
//    never waits inside the loop.
#include <cmath>
#include <type_traits>

#include "common.hpp"
#include "reduce.hpp"
#include "fused.hpp"
#include "spmv_common.hpp"
#include "stop.hpp"

namespace gkomi {

// ------------------------------------------------ reference-semantics ----

#define GKOMI_ELEMENTS(total)                                                                \
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; e < (total);    \
         e += static_cast<int64_t>(gridDim.x) * kBlock)

template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg_initialize_kernel(int64_t n, int64_t nrhs, const T* __restrict__ b, int64_t bs,
                         T* __restrict__ r, T* __restrict__ z, T* __restrict__ p,
                         T* __restrict__ q, int64_t vs, T* __restrict__ prev_rho,
                         T* __restrict__ rho, uint8_t* __restrict__ status)
{
    GKOMI_ELEMENTS(n * nrhs)
    {
        const int64_t i = e / nrhs;
        const int64_t j = e - i * nrhs;
        r[i * vs + j] = b[i * bs + j];
        z[i * vs + j] = T(0);
        p[i * vs + j] = T(0);
        q[i * vs + j] = T(0);
        if (i == 0) {
            rho[j] = T(0);
            prev_rho[j] = T(1);
            status[j] = 0;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg_step_1_kernel(int64_t n, int64_t nrhs, T* __restrict__ p, const T* __restrict__ z,
                     int64_t vs, const T* __restrict__ rho, const T* __restrict__ prev_rho,
                     const uint8_t* __restrict__ status)
{
    GKOMI_ELEMENTS(n * nrhs)
    {
        const int64_t i = e / nrhs;
        const int64_t j = e - i * nrhs;
        if (has_stopped(status[j])) continue;
        if (prev_rho[j] == T(0)) {
            p[i * vs + j] = z[i * vs + j];
        } else {
            const T tmp = rho[j] / prev_rho[j];
            p[i * vs + j] = z[i * vs + j] + tmp * p[i * vs + j];
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    cg_step_2_kernel(int64_t n, int64_t nrhs, T* __restrict__ x, int64_t xs, T* __restrict__ r,
                     const T* __restrict__ p, const T* __restrict__ q, int64_t vs,
                     const T* __restrict__ beta, const T* __restrict__ rho,
                     const uint8_t* __restrict__ status)
{
    GKOMI_ELEMENTS(n * nrhs)
    {
        const int64_t i = e / nrhs;
        const int64_t j = e - i * nrhs;
        if (has_stopped(status[j])) continue;
        if (beta[j] != T(0)) {
            const T tmp = rho[j] / beta[j];
            x[i * xs + j] += tmp * p[i * vs + j];
            r[i * vs + j] -= tmp * q[i * vs + j];
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    bicgstab_initialize_kernel(int64_t n, int64_t nrhs, const T* __restrict__ b, int64_t bs,
                               T* r, T* rr, T* y, T* s, T* t, T* z, T* v, T* p, int64_t vs,
                               T* prev_rho, T* rho, T* alpha, T* beta, T* gamma, T* omega,
                               uint8_t* status)
{
    GKOMI_ELEMENTS(n * nrhs)
    {
        const int64_t i = e / nrhs;
        const int64_t j = e - i * nrhs;
        const int64_t o = i * vs + j;
        r[o] = b[i * bs + j];
        rr[o] = z[o] = v[o] = s[o] = t[o] = y[o] = p[o] = T(0);
        if (i == 0) {
            rho[j] = prev_rho[j] = alpha[j] = beta[j] = gamma[j] = omega[j] = T(1);
            status[j] = 0;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    bicgstab_step_1_kernel(int64_t n, int64_t nrhs, const T* __restrict__ r, T* __restrict__ p,
                           const T* __restrict__ v, int64_t vs, const T* __restrict__ rho,
                           const T* __restrict__ prev_rho, const T* __restrict__ alpha,
                           const T* __restrict__ omega, const uint8_t* __restrict__ status)
{
    GKOMI_ELEMENTS(n * nrhs)
    {
        const int64_t i = e / nrhs;
        const int64_t j = e - i * nrhs;
        if (has_stopped(status[j])) continue;
        const int64_t o = i * vs + j;
        if (prev_rho[j] * omega[j] != T(0)) {
            const T tmp = rho[j] / prev_rho[j] * alpha[j] / omega[j];
            p[o] = r[o] + tmp * (p[o] - omega[j] * v[o]);
        } else {
            p[o] = r[o];
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    bicgstab_step_2_kernel(int64_t n, int64_t nrhs, const T* __restrict__ r, T* __restrict__ s,
                           const T* __restrict__ v, int64_t vs, const T* __restrict__ rho,
                           T* __restrict__ alpha, const T* __restrict__ beta,
                           const uint8_t* __restrict__ status)
{
    GKOMI_ELEMENTS(n * nrhs)
    {
        const int64_t i = e / nrhs;
        const int64_t j = e - i * nrhs;
        if (has_stopped(status[j])) continue;
        const int64_t o = i * vs + j;
        if (beta[j] != T(0)) {
            const T a = rho[j] / beta[j];
            s[o] = r[o] - a * v[o];
            if (i == 0) alpha[j] = a;
        } else {
            s[o] = r[o];
            if (i == 0) alpha[j] = T(0);
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    bicgstab_step_3_kernel(int64_t n, int64_t nrhs, T* __restrict__ x, int64_t xs,
                           T* __restrict__ r, const T* __restrict__ s, const T* __restrict__ t,
                           const T* __restrict__ y, const T* __restrict__ z, int64_t vs,
                           const T* __restrict__ alpha, const T* __restrict__ beta,
                           const T* __restrict__ gamma, T* __restrict__ omega,
                           const uint8_t* __restrict__ status)
{
    GKOMI_ELEMENTS(n * nrhs)
    {
        const int64_t i = e / nrhs;
        const int64_t j = e - i * nrhs;
        if (has_stopped(status[j])) continue;
        const T w = beta[j] != T(0) ? gamma[j] / beta[j] : T(0);
        const int64_t o = i * vs + j;
        x[i * xs + j] += alpha[j] * y[o] + w * z[o];
        r[o] = s[o] - w * t[o];
    }
}

// omega is written by a separate single-pass kernel so that every element
// kernel above reads the old values consistently
template <typename T>
__global__ void bicgstab_omega_kernel(int64_t nrhs, const T* __restrict__ beta,
                                      const T* __restrict__ gamma, T* __restrict__ omega,
                                      const uint8_t* __restrict__ status)
{
    for (int64_t j = threadIdx.x; j < nrhs; j += blockDim.x) {
        if (has_stopped(status[j])) continue;
        omega[j] = beta[j] != T(0) ? gamma[j] / beta[j] : T(0);
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    bicgstab_finalize_kernel(int64_t n, int64_t nrhs, T* __restrict__ x, int64_t xs,
                             const T* __restrict__ y, int64_t vs, const T* __restrict__ alpha,
                             const uint8_t* __restrict__ status)
{
    GKOMI_ELEMENTS(n * nrhs)
    {
        const int64_t i = e / nrhs;
        const int64_t j = e - i * nrhs;
        if (has_stopped(status[j]) && !(status[j] & GKOMI_STATUS_FINALIZED)) {
            x[i * xs + j] += alpha[j] * y[i * vs + j];
        }
    }
}

__global__ void status_finalize_kernel(int64_t nrhs, uint8_t* status)
{
    for (int64_t j = threadIdx.x; j < nrhs; j += blockDim.x) {
        if (has_stopped(status[j])) status[j] |= GKOMI_STATUS_FINALIZED;
    }
}

template <typename T>
__global__ void criteria_check_kernel(gkomi_criteria crit, int64_t nrhs, const T* tau,
                                      const T* implicit_sq, const T* orig, int set_finalized,
                                      uint8_t* status, gkomi_stop_flags* flags)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        int all = 0, changed = 0;
        combined_check<T>(crit, flags->iteration, nrhs, tau, implicit_sq, orig,
                          set_finalized != 0, status, &all, &changed);
        flags->all_stopped = all;
        flags->one_changed = changed;
    }
}

// ---------------------------------------------------------- fused CG ----
// Two kernels per iteration; sums through FP64 per-workgroup partials that
// every workgroup of the consumer adds in the same order (fused.hpp).
// Values that a kernel both reads and publishes (rho, tau, the stopping status) live in two slots
// selected by the iteration's parity, so no workgroup reads a slot that
// another workgroup of the same kernel writes.
//
// scalars (value type): [0], [1] rho of even / odd iterations (slot 1 starts
// as prev_rho = 1), [2] beta, [3], [4] tau of even / odd iterations.

struct cg_work {
    double* part_init;  // [kCgMaxGrid][2]: (b.b, r0.r0) of the start
    double* part_upd;   // [kCgMaxGrid][2]: (r.z, r.r) of the last update kernel
    double* part_dir;   // [kCgMaxGrid]: p.q of the last direction kernel
    uint8_t* status;    // [2]: stopping status of even / odd iterations
};

__host__ __device__ inline cg_work cg_work_layout(void* base)
{
    cg_work w;
    double* d = static_cast<double*>(base);
    w.part_init = d;
    w.part_upd = d + 2 * kCgMaxGrid;
    w.part_dir = d + 4 * kCgMaxGrid;
    w.status = reinterpret_cast<uint8_t*>(d + 5 * kCgMaxGrid);
    return w;
}

// Start, kernel 1: r = b - A x (sub-wave SpMV), p_a = p_b = q = 0, and
// partials of b.b and r.r; resets the flags, the status and prev_rho = 1
// (reference/solver/cg_kernels.cpp:26-50).
template <typename T, int SW, int RPG, bool Compact>
__global__ void __launch_bounds__(kBlock)
    cg_fused_init_kernel(int64_t n, const int32_t* __restrict__ rp,
                         const int32_t* __restrict__ ci, const int16_t* __restrict__ co,
                         const T* __restrict__ v,
                         const T* __restrict__ b, const T* __restrict__ x, T* __restrict__ r,
                         T* __restrict__ pa, T* __restrict__ pb, T* __restrict__ q,
                         T* __restrict__ scalars, uint8_t* __restrict__ status,
                         gkomi_stop_flags* __restrict__ flags, cg_work w)
{
    int64_t rb0, rb1;
    block_range(ceildiv(n, subwave_tile_rows<SW, RPG>), &rb0, &rb1);
    T acc[2] = {T(0), T(0)};
    for (int64_t blk = rb0; blk < rb1; ++blk) {
        subwave_tile<T, SW, RPG, Compact>(
            n, blk, rp, ci, co, v, [&](int32_t c) { return x[c]; },
            [&](int64_t row, T sum) {
                const T bi = b[row];
                const T ri = bi - sum;
                r[row] = ri;
                pa[row] = T(0);
                pb[row] = T(0);
                q[row] = T(0);
                acc[0] += bi * bi;
                acc[1] += ri * ri;
            });
    }
    store_partials<T, 2>(acc, w.part_init);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scalars[0] = T(0);
        scalars[1] = T(1);  // prev_rho of iteration 0
        scalars[2] = T(0);
        scalars[3] = T(0);
        scalars[4] = T(0);
        status[0] = 0;
        w.status[0] = 0;
        w.status[1] = 0;
        flags->iteration = -1;
        flags->all_stopped = 0;
        flags->one_changed = 0;
    }
}

// The start's kernel 1 over the matrix's sliced copy: one wave per 64-row
// slice, a workgroup's tiles of four slices in a contiguous range.
constexpr int64_t kSlicedTileRows = int64_t{kWave} * kWavesPerBlock;

template <typename T, bool Compact>
__global__ void __launch_bounds__(kBlock)
    cg_fused_init_sliced_kernel(int64_t n, const int32_t* __restrict__ rp,
                                const int64_t* __restrict__ sets,
                                const int32_t* __restrict__ cols,
                                const int16_t* __restrict__ offs, const T* __restrict__ v,
                                const T* __restrict__ b, const T* __restrict__ x,
                                T* __restrict__ r, T* __restrict__ pa, T* __restrict__ pb,
                                T* __restrict__ q, T* __restrict__ scalars,
                                uint8_t* __restrict__ status, gkomi_stop_flags* __restrict__ flags,
                                cg_work w)
{
    int64_t rb0, rb1;
    block_range(ceildiv(n, kSlicedTileRows), &rb0, &rb1);
    const int64_t slices = ceildiv(n, kWave);
    T acc[2] = {T(0), T(0)};
    for (int64_t blk = rb0; blk < rb1; ++blk) {
        const int64_t sl = blk * kWavesPerBlock + threadIdx.x / kWave;
        if (sl >= slices) continue;
        sliced_slice<T, Compact>(
            n, sl, rp, sets, cols, offs, v, [&](int32_t c) { return x[c]; },
            [&](int64_t row, T sum) {
                const T bi = b[row];
                const T ri = bi - sum;
                r[row] = ri;
                pa[row] = T(0);
                pb[row] = T(0);
                q[row] = T(0);
                acc[0] += bi * bi;
                acc[1] += ri * ri;
            });
    }
    store_partials<T, 2>(acc, w.part_init);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scalars[0] = T(0);
        scalars[1] = T(1);  // prev_rho of iteration 0
        scalars[2] = T(0);
        scalars[3] = T(0);
        scalars[4] = T(0);
        status[0] = 0;
        w.status[0] = 0;
        w.status[1] = 0;
        flags->iteration = -1;
        flags->all_stopped = 0;
        flags->one_changed = 0;
    }
}

// Kernel 2 of an iteration (Update) and of the start (!Update):
//   Update: beta = p.q from the direction partials; x += rho/beta p,
//           r -= rho/beta q (skipped when beta == 0,
//           reference/solver/cg_kernels.cpp:89-95);
//   start:  the criteria baselines ||b|| / ||r0|| from the init partials;
// then z = M r and partials of r.z and r.r for the next direction kernel.
template <typename T, bool Update, bool Jacobi>
__global__ void __launch_bounds__(kUpdBlock)
    cg_fused_update_kernel(int64_t n, T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                           const T* __restrict__ p, const T* __restrict__ q,
                           const T* __restrict__ inv_diag, gkomi_criteria crit,
                           T* __restrict__ scalars, T* __restrict__ orig,
                           const gkomi_stop_flags* __restrict__ flags, cg_work w, int count_in,
                           int par)
{
    T alpha = T(0);
    bool do_update = false;
    int64_t i0, i1;
    block_range(ceildiv(n, kUpdBlock), &i0, &i1);
    // the loads of this thread's first element, the stop flag and rho go
    // out before the combine's barrier, so that their latencies overlap
    // with the partials' (no kernel writes the flag while this one runs)
    const int64_t ia = i0 * kUpdBlock + threadIdx.x;
    const bool has_a = i0 < i1 && ia < n;
    T ra = T(0), xa = T(0), pa = T(0), qa = T(0), da = T(1);
    if (has_a) {
        ra = r[ia];
        if constexpr (Update) {
            xa = x[ia];
            pa = p[ia];
            qa = q[ia];
        }
        if constexpr (Jacobi) da = inv_diag[ia];
    }
    if constexpr (Update) {
        const int stopped = flags->all_stopped;
        const T rho = scalars[par];
        double pq[1];
        combine_partials<1, kUpdBlock>(w.part_dir, count_in, pq);
        if (stopped) return;
        const T beta = static_cast<T>(pq[0]);
        do_update = beta != T(0);
        alpha = do_update ? rho / beta : T(0);
        if (blockIdx.x == 0 && threadIdx.x == 0) scalars[2] = beta;
    } else {
        double s[2];
        combine_partials<2, kUpdBlock>(w.part_init, count_in, s);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            for (int k = 0; k < crit.num; ++k) {
                const int base = crit.c[k].baseline;
                orig[k] = base == GKOMI_BASELINE_RHS_NORM
                              ? static_cast<T>(sqrt(s[0]))
                              : (base == GKOMI_BASELINE_INITIAL_RESNORM
                                     ? static_cast<T>(sqrt(s[1]))
                                     : T(1));
            }
        }
    }
    T acc[2] = {T(0), T(0)};
    for (int64_t blk = i0; blk < i1; ++blk) {
        const int64_t i = blk * kUpdBlock + threadIdx.x;
        if (i >= n) continue;
        const bool first = blk == i0;
        T ri = first ? ra : r[i];
        if (do_update) {
            x[i] = fma(alpha, first ? pa : p[i], first ? xa : x[i]);
            ri = fma(-alpha, first ? qa : q[i], ri);
            r[i] = ri;
        }
        const T zi = Jacobi ? ri * (first ? da : inv_diag[i]) : ri;
        z[i] = zi;
        acc[0] += ri * zi;
        acc[1] += ri * ri;
    }
    store_partials<T, 2, kUpdBlock>(acc, w.part_upd);
}

// Kernel 1 of iteration `iter` (parity par = iter % 2): rho = r.z and
// tau = ||r|| from the update partials, the Combined check of this
// iteration (core/solver/cg.cpp:145-158) in every workgroup; unless stopped,
// p_out = z + rho/prev_rho p_in (p_out = z when prev_rho == 0,
// reference/solver/cg_kernels.cpp:64-70), q = A p_out with p_out recomputed
// at the gathered columns, and partials of p.q.
template <typename T, int SW, int RPG, bool Compact>
__global__ void __launch_bounds__(kBlock)
    cg_fused_direction_kernel(int64_t n, const int32_t* __restrict__ rp,
                              const int32_t* __restrict__ ci, const int16_t* __restrict__ co,
                              const T* __restrict__ v,
                              const T* __restrict__ z, const T* __restrict__ p_in,
                              T* __restrict__ p_out, T* __restrict__ q, gkomi_criteria crit,
                              T* __restrict__ scalars, const T* __restrict__ orig,
                              uint8_t* __restrict__ status, gkomi_stop_flags* __restrict__ flags,
                              cg_work w, int count_in, int64_t iter, int par)
{
    // workgroup 0 may set the flag during this kernel (when this iteration
    // stops): thread 0 reads it once and the combine's barrier hands it to
    // the workgroup, so the early exit is uniform; the flag, rho_prev, the
    // status and the first tile's row pointers, columns and values are
    // loaded together with the partials
    int64_t rb0, rb1;
    block_range(ceildiv(n, subwave_tile_rows<SW, RPG>), &rb0, &rb1);
    subwave_pf<T, RPG> pf;
    if (rb0 < rb1) subwave_prefetch<T, SW, RPG, Compact>(n, rb0, rp, ci, co, v, pf);
    const int flag0 = threadIdx.x == 0 ? flags->all_stopped : 0;
    const T prev = scalars[par ^ 1];
    uint8_t st = w.status[par ^ 1];
    double s[2];
    if (combine_partials_bcast<2>(w.part_upd, count_in, s, flag0)) return;
    const T rho = static_cast<T>(s[0]);
    const T tau = static_cast<T>(sqrt(s[1]));
    int all = 0, changed = 0;
    combined_check<T>(crit, iter, 1, &tau, &rho, orig, true, &st, &all, &changed);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scalars[par] = rho;
        scalars[3 + par] = tau;
        w.status[par] = st;
        status[0] = st;
        flags->iteration = iter;
        flags->all_stopped = all;
        flags->one_changed = changed;
    }
    if (all) return;
    const bool restart = prev == T(0);
    const T tmp = restart ? T(0) : rho / prev;
    T acc[1] = {T(0)};
    auto gather = [&](int32_t c) { return restart ? z[c] : fma(tmp, p_in[c], z[c]); };
    auto row_op = [&](int64_t row, T sum) {
        const T pi = restart ? z[row] : fma(tmp, p_in[row], z[row]);
        p_out[row] = pi;
        q[row] = sum;
        acc[0] += pi * sum;
    };
    if (rb0 < rb1) subwave_finish<T, SW, RPG, Compact>(n, ci, co, v, pf, gather, row_op);
    for (int64_t blk = rb0 + 1; blk < rb1; ++blk) {
        subwave_tile<T, SW, RPG, Compact>(n, blk, rp, ci, co, v, gather, row_op);
    }
    store_partials<T, 1>(acc, w.part_dir);
}

// Split form of kernel 1 for rows with many entries, where recomputing p at
// every gathered column (two gathers per entry) costs more than one extra
// pass over p.  (a) the check of iteration `iter` exactly as in the fused
// direction kernel, then p_out = z + rho/prev_rho p_in elementwise;
template <typename T>
__global__ void __launch_bounds__(kUpdBlock)
    cg_split_p_kernel(int64_t n, const T* __restrict__ z, const T* __restrict__ p_in,
                      T* __restrict__ p_out, gkomi_criteria crit, T* __restrict__ scalars,
                      const T* __restrict__ orig, uint8_t* __restrict__ status,
                      gkomi_stop_flags* __restrict__ flags, cg_work w, int count_in,
                      int64_t iter, int par)
{
    // as in the direction kernel: flag, rho_prev and status ride with the
    // partials through one barrier, and so do the loads of this thread's
    // first z and p_in (p_in is never used on a restart, whatever it holds)
    const int flag0 = threadIdx.x == 0 ? flags->all_stopped : 0;
    const T prev = scalars[par ^ 1];
    uint8_t st = w.status[par ^ 1];
    int64_t i0, i1;
    block_range(ceildiv(n, kUpdBlock), &i0, &i1);
    const int64_t ia = i0 * kUpdBlock + threadIdx.x;
    T za = T(0), pa = T(0);
    if (i0 < i1 && ia < n) {
        za = z[ia];
        pa = p_in[ia];
    }
    double s[2];
    if (combine_partials_bcast<2, kUpdBlock>(w.part_upd, count_in, s, flag0)) return;
    const T rho = static_cast<T>(s[0]);
    const T tau = static_cast<T>(sqrt(s[1]));
    int all = 0, changed = 0;
    combined_check<T>(crit, iter, 1, &tau, &rho, orig, true, &st, &all, &changed);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scalars[par] = rho;
        scalars[3 + par] = tau;
        w.status[par] = st;
        status[0] = st;
        flags->iteration = iter;
        flags->all_stopped = all;
        flags->one_changed = changed;
    }
    if (all) return;
    const bool restart = prev == T(0);
    const T tmp = restart ? T(0) : rho / prev;
    for (int64_t blk = i0; blk < i1; ++blk) {
        const int64_t i = blk * kUpdBlock + threadIdx.x;
        if (i >= n) continue;
        const bool first = blk == i0;
        const T zi = first ? za : z[i];
        p_out[i] = restart ? zi : fma(tmp, first ? pa : p_in[i], zi);
    }
}

// (b) q = A p (one gather per entry) and the partials of p.q.
template <typename T, int SW, int RPG, bool Compact>
__global__ void __launch_bounds__(kBlock)
    cg_split_spmv_kernel(int64_t n, const int32_t* __restrict__ rp,
                         const int32_t* __restrict__ ci, const int16_t* __restrict__ co,
                         const T* __restrict__ v,
                         const T* __restrict__ p, T* __restrict__ q,
                         const gkomi_stop_flags* __restrict__ flags, cg_work w)
{
    // the first tile's loads go out before the stop flag is known
    int64_t rb0, rb1;
    block_range(ceildiv(n, subwave_tile_rows<SW, RPG>), &rb0, &rb1);
    subwave_pf<T, RPG> pf;
    if (rb0 < rb1) subwave_prefetch<T, SW, RPG, Compact>(n, rb0, rp, ci, co, v, pf);
    if (flags->all_stopped) return;
    T acc[1] = {T(0)};
    auto gather = [&](int32_t c) { return p[c]; };
    auto row_op = [&](int64_t row, T sum) {
        q[row] = sum;
        acc[0] += p[row] * sum;
    };
    if (rb0 < rb1) subwave_finish<T, SW, RPG, Compact>(n, ci, co, v, pf, gather, row_op);
    for (int64_t blk = rb0 + 1; blk < rb1; ++blk) {
        subwave_tile<T, SW, RPG, Compact>(n, blk, rp, ci, co, v, gather, row_op);
    }
    store_partials<T, 1>(acc, w.part_dir);
}

// Kernel 1 over the staged product (staged_window, G lanes per row; rows
// of stencil length): the check and p_out = z + rho/prev_rho p_in as in
// cg_fused_direction_kernel, q = A p_out with p_out recomputed at the
// gathered columns, and partials of p.q.  Workgroup tiles of one window
// per wave.  With Pre (small systems: the launch gives every workgroup at
// most one tile) the wave's window is staged and the lane's first entries'
// z and p_in gathered, and its row's z and p_in loaded, before the wait for
// the partials, so that only the arithmetic and the stores follow it; the
// sums are the same.
template <typename T, int G, bool Compact, bool Pre>
__global__ void __launch_bounds__(kBlock)
    cg_fused_direction_staged_kernel(int64_t n, const int32_t* __restrict__ rp,
                                     const int32_t* __restrict__ ci,
                                     const int16_t* __restrict__ co, const T* __restrict__ v,
                                     const T* __restrict__ z, const T* __restrict__ p_in,
                                     T* __restrict__ p_out, T* __restrict__ q,
                                     gkomi_criteria crit, T* __restrict__ scalars,
                                     const T* __restrict__ orig, uint8_t* __restrict__ status,
                                     gkomi_stop_flags* __restrict__ flags, cg_work w,
                                     int count_in, int64_t iter, int par)
{
    using Idx = std::conditional_t<Compact, int16_t, int32_t>;
    constexpr int UB = 8;
    __shared__ T sv[kWavesPerBlock][kStagedChunk];
    __shared__ Idx sc[kWavesPerBlock][kStagedChunk];
    const int flag0 = threadIdx.x == 0 ? flags->all_stopped : 0;
    const T prev = scalars[par ^ 1];
    uint8_t st = w.status[par ^ 1];
    const int wave = threadIdx.x / kWave;
    int64_t rb0, rb1;
    block_range(ceildiv(n, staged_tile_rows<G>), &rb0, &rb1);
    staged_pf<T, UB> pf;
    T z_row = T(0), p_row = T(0);
    if constexpr (Pre) {
        if (rb0 < rb1) {
            staged_prefetch<T, G, kStagedChunk, Compact, UB>(
                n, rb0 * kWavesPerBlock + wave, rp, ci, co, v, sv[wave], sc[wave],
                [&](int32_t c) { return z[c]; }, [&](int32_t c) { return p_in[c]; }, pf);
            if (pf.row < n) {
                z_row = z[pf.row];
                p_row = p_in[pf.row];
            }
        }
    }
    double s[2];
    if (combine_partials_bcast<2>(w.part_upd, count_in, s, flag0)) return;
    const T rho = static_cast<T>(s[0]);
    const T tau = static_cast<T>(sqrt(s[1]));
    int all = 0, changed = 0;
    combined_check<T>(crit, iter, 1, &tau, &rho, orig, true, &st, &all, &changed);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scalars[par] = rho;
        scalars[3 + par] = tau;
        w.status[par] = st;
        status[0] = st;
        flags->iteration = iter;
        flags->all_stopped = all;
        flags->one_changed = changed;
    }
    if (all) return;
    const bool restart = prev == T(0);
    const T tmp = restart ? T(0) : rho / prev;
    T acc[1] = {T(0)};
    const auto dir = [&](T zc, T pc) { return restart ? zc : fma(tmp, pc, zc); };
    auto gather = [&](int32_t c) { return dir(z[c], p_in[c]); };
    auto row_op = [&](int64_t row, T sum) {
        const T pi = dir(z[row], p_in[row]);
        p_out[row] = pi;
        q[row] = sum;
        acc[0] += pi * sum;
    };
    int64_t blk = rb0;
    if constexpr (Pre) {
        if (rb0 < rb1) {
            staged_finish<T, G, kStagedChunk, Compact, UB>(
                n, ci, co, v, sv[wave], sc[wave], pf, dir, gather,
                [&](int64_t row, T sum) {
                    const T pi = dir(z_row, p_row);
                    p_out[row] = pi;
                    q[row] = sum;
                    acc[0] += pi * sum;
                });
            ++blk;
        }
    }
    for (; blk < rb1; ++blk) {
        staged_window<T, G, kStagedChunk, Compact>(n, blk * kWavesPerBlock + wave, rp, ci, co,
                                                   v, sv[wave], sc[wave], gather, row_op);
    }
    store_partials<T, 1>(acc, w.part_dir);
}

// (b) over the staged product: q = A p and the partials of p.q.
template <typename T, int G, bool Compact>
__global__ void __launch_bounds__(kBlock)
    cg_split_spmv_staged_kernel(int64_t n, const int32_t* __restrict__ rp,
                                const int32_t* __restrict__ ci, const int16_t* __restrict__ co,
                                const T* __restrict__ v, const T* __restrict__ p,
                                T* __restrict__ q, const gkomi_stop_flags* __restrict__ flags,
                                cg_work w)
{
    using Idx = std::conditional_t<Compact, int16_t, int32_t>;
    __shared__ T sv[kWavesPerBlock][kStagedChunk];
    __shared__ Idx sc[kWavesPerBlock][kStagedChunk];
    if (flags->all_stopped) return;
    T acc[1] = {T(0)};
    const int wave = threadIdx.x / kWave;
    // tiles in grid-stride order (XCD-aware start): the resident waves move
    // through the matrix together
    const int64_t tiles = ceildiv(n, staged_tile_rows<G>);
    for (int64_t blk = xcd_block(blockIdx.x, gridDim.x); blk < tiles; blk += gridDim.x) {
        staged_window<T, G, kStagedChunk, Compact>(
            n, blk * kWavesPerBlock + wave, rp, ci, co, v, sv[wave], sc[wave],
            [&](int32_t c) { return p[c]; },
            [&](int64_t row, T sum) {
                q[row] = sum;
                acc[0] += p[row] * sum;
            });
    }
    store_partials<T, 1>(acc, w.part_dir);
}

// (b) over the matrix's sliced copy: q = A p and the partials of p.q.
template <typename T, bool Compact>
__global__ void __launch_bounds__(kBlock)
    cg_split_spmv_sliced_kernel(int64_t n, const int32_t* __restrict__ rp,
                                const int64_t* __restrict__ sets,
                                const int32_t* __restrict__ cols,
                                const int16_t* __restrict__ offs, const T* __restrict__ v,
                                const T* __restrict__ p, T* __restrict__ q,
                                const gkomi_stop_flags* __restrict__ flags, cg_work w)
{
    if (flags->all_stopped) return;
    int64_t rb0, rb1;
    block_range(ceildiv(n, kSlicedTileRows), &rb0, &rb1);
    const int64_t slices = ceildiv(n, kWave);
    T acc[1] = {T(0)};
    for (int64_t blk = rb0; blk < rb1; ++blk) {
        const int64_t sl = blk * kWavesPerBlock + threadIdx.x / kWave;
        if (sl >= slices) continue;
        sliced_slice<T, Compact>(n, sl, rp, sets, cols, offs, v,
                                 [&](int32_t c) { return p[c]; },
                                 [&](int64_t row, T sum) {
                                     q[row] = sum;
                                     acc[0] += p[row] * sum;
                                 });
    }
    store_partials<T, 1>(acc, w.part_dir);
}

bool has_slices(const gkomi_csr* a)
{
    return a->slice_sets != nullptr && a->slice_values != nullptr &&
           (a->slice_offsets != nullptr || a->slice_cols != nullptr);
}

// Rows of at least this mean length take the split form of kernel 1.
constexpr double kCgSplitMeanRow = 16.0;

int check_cg_state(const gkomi_cg_state* st)
{
    GKOMI_REQUIRE(st != nullptr, GKOMI_INVALID_ARGUMENT, "null CG state");
    GKOMI_REQUIRE(st->n >= 0, GKOMI_INVALID_ARGUMENT, "CG state: negative size");
    GKOMI_REQUIRE(st->x && st->b && st->r && st->z && st->p_a && st->p_b && st->q &&
                      st->scalars && st->orig_tau && st->status && st->flags && st->work,
                  GKOMI_INVALID_ARGUMENT, "CG state: null pointer");
    return GKOMI_SUCCESS;
}

// Launch shapes shared by start and iterate (the consumer of a partial
// array must know how many workgroups produced it).
struct cg_grids {
    subwave_shape sh;
    int staged;  // lanes per row of the staged product kernels, 0 = sub-wave
    int spmv;    // init and direction kernels
    int update;  // update kernels
};

// The split form's automatic shape: about 16 bytes of matrix values per
// lane and row (lanes per row: the power of two nearest to mean row length
// x value bytes / 16) and two rows in flight per lane group.  On the
// 27-point 96^3 matrix (mean 26.4) that is 16 lanes for FP64 and 8 for FP32,
// the fastest CG iterations of sw 1..32 x rpg 1..8 in both precisions
// (90.5 and 61.8 us against 92.3 and 67.6 us with the SpMV's own (16, 4)).
subwave_shape split_shape(double mean, int value_bytes)
{
    const double target = mean * value_bytes / 16.0;
    int sw = 1;
    while (sw < 64 && std::abs(2 * sw - target) < std::abs(sw - target)) sw *= 2;
    return subwave_shape{sw, 2};
}

template <typename T>
cg_grids cg_grids_for(const gkomi_csr* a, int64_t n)
{
    cg_grids g;
    g.sh = pick_subwave(a, sizeof(T));
    const double mean = n > 0 ? static_cast<double>(a->nnz) / static_cast<double>(n) : 0.0;
    if (a->subwave == 0 && mean >= kCgSplitMeanRow) {
        g.sh = split_shape(mean, sizeof(T));
    }
    // the staged product where the classical SpMV takes it (rows of
    // stencil length), unless the matrix carries its sliced copy
    g.staged = has_slices(a) ? 0 : std::max(0, staged_lanes(a, sizeof(T)));
    const int64_t tile_rows = has_slices(a) ? kSlicedTileRows
                              : g.staged > 0 ? int64_t{kWave} / g.staged * kWavesPerBlock
                                             : int64_t{kBlock} / g.sh.sw * g.sh.rpg;
    g.spmv = fused_grid(ceildiv(std::max<int64_t>(n, 1), tile_rows));
    if (g.staged > 0) {
        // no more workgroups than are resident at once: each holds four
        // staging areas (24 KB in FP64 with 4-byte indices, so 6 per CU of
        // 160 KB), and a partial second round of workgroups left the CUs of
        // the first idle at the end (measured on 7-entry rows of 0.5-1.2M
        // rows: 8 per CU 37.8 us, resident only 31-32 us per iteration)
        const int64_t per_wg =
            int64_t{kWavesPerBlock} * kStagedChunk *
                (static_cast<int64_t>(sizeof(T)) + (a->col_offsets != nullptr ? 2 : 4)) +
            2048;
        const int64_t resident = std::max<int64_t>(1, std::min<int64_t>(8, (160 * 1024) / per_wg));
        g.spmv = static_cast<int>(std::min<int64_t>(g.spmv, resident * num_cus()));
    }
    g.update = static_cast<int>(
        std::max<int64_t>(1, std::min<int64_t>(ceildiv(std::max<int64_t>(n, 1), kUpdBlock),
                                               kCgUpdGrid)));
    return g;
}

template <typename T>
int cg_fused_start_impl(const gkomi_ctx* ctx, const gkomi_csr* a, const gkomi_criteria* crit,
                        gkomi_cg_state* st)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_csr(a)) return rc;
    if (int rc = check_crit(crit)) return rc;
    if (int rc = check_cg_state(st)) return rc;
    GKOMI_REQUIRE(a->n_rows == a->n_cols && a->n_rows == st->n, GKOMI_DIMENSION_MISMATCH,
                  "fused CG: matrix and vectors do not conform");
    GKOMI_REQUIRE(a->strategy == GKOMI_CSR_CLASSICAL, GKOMI_NOT_SUPPORTED,
                  "fused CG: needs the classical (sub-wave) CSR strategy");
    hipStream_t s = stream_of(ctx);
    const cg_work w = cg_work_layout(st->work);
    const T* v = static_cast<const T*>(a->values);
    T* sc = static_cast<T*>(st->scalars);
    const int64_t n = st->n;
    const cg_grids g = cg_grids_for<T>(a, n);
    if (has_slices(a)) {
        const T* sv = static_cast<const T*>(a->slice_values);
        if (a->slice_offsets) {
            cg_fused_init_sliced_kernel<T, true><<<g.spmv, kBlock, 0, s>>>(
                n, a->row_ptrs, a->slice_sets, a->slice_cols, a->slice_offsets, sv,
                static_cast<const T*>(st->b), static_cast<const T*>(st->x),
                static_cast<T*>(st->r), static_cast<T*>(st->p_a), static_cast<T*>(st->p_b),
                static_cast<T*>(st->q), sc, st->status, st->flags, w);
        } else {
            cg_fused_init_sliced_kernel<T, false><<<g.spmv, kBlock, 0, s>>>(
                n, a->row_ptrs, a->slice_sets, a->slice_cols, a->slice_offsets, sv,
                static_cast<const T*>(st->b), static_cast<const T*>(st->x),
                static_cast<T*>(st->r), static_cast<T*>(st->p_a), static_cast<T*>(st->p_b),
                static_cast<T*>(st->q), sc, st->status, st->flags, w);
        }
    } else {
#define GKOMI_INIT_CASE_C(S, R, C)                                                           \
    cg_fused_init_kernel<T, S, R, C><<<g.spmv, kBlock, 0, s>>>(                              \
        n, a->row_ptrs, a->col_idxs, a->col_offsets, v, static_cast<const T*>(st->b),        \
        static_cast<const T*>(st->x), static_cast<T*>(st->r), static_cast<T*>(st->p_a),      \
        static_cast<T*>(st->p_b), static_cast<T*>(st->q), sc, st->status, st->flags, w);     \
    break;
#define GKOMI_INIT_CASE(S, R) GKOMI_INIT_CASE_C(S, R, false)
#define GKOMI_INIT_CASE_COMPACT(S, R) GKOMI_INIT_CASE_C(S, R, true)
    if (a->col_offsets) {
        GKOMI_SUBWAVE_SWITCH(g.sh, GKOMI_INIT_CASE_COMPACT)
    } else {
        GKOMI_SUBWAVE_SWITCH(g.sh, GKOMI_INIT_CASE)
    }
#undef GKOMI_INIT_CASE
#undef GKOMI_INIT_CASE_COMPACT
#undef GKOMI_INIT_CASE_C
    }
    GKOMI_CHECK_LAUNCH("cg_fused_init_kernel");
    if (st->inv_diag) {
        cg_fused_update_kernel<T, false, true><<<g.update, kUpdBlock, 0, s>>>(
            n, static_cast<T*>(st->x), static_cast<T*>(st->r), static_cast<T*>(st->z),
            static_cast<const T*>(st->p_a), static_cast<const T*>(st->q),
            static_cast<const T*>(st->inv_diag), *crit, sc, static_cast<T*>(st->orig_tau),
            st->flags, w, g.spmv, 0);
    } else {
        cg_fused_update_kernel<T, false, false><<<g.update, kUpdBlock, 0, s>>>(
            n, static_cast<T*>(st->x), static_cast<T*>(st->r), static_cast<T*>(st->z),
            static_cast<const T*>(st->p_a), static_cast<const T*>(st->q), nullptr, *crit, sc,
            static_cast<T*>(st->orig_tau), st->flags, w, g.spmv, 0);
    }
    GKOMI_CHECK_LAUNCH("cg_fused_update_kernel");
    st->issued = 0;
    return GKOMI_SUCCESS;
}

template <typename T>
int cg_fused_iterate_impl(const gkomi_ctx* ctx, const gkomi_csr* a, const gkomi_criteria* crit,
                          gkomi_cg_state* st, int32_t count)
{
    if (int rc = enter(ctx)) return rc;
    if (int rc = check_csr(a)) return rc;
    if (int rc = check_crit(crit)) return rc;
    if (int rc = check_cg_state(st)) return rc;
    GKOMI_REQUIRE(a->n_rows == a->n_cols && a->n_rows == st->n, GKOMI_DIMENSION_MISMATCH,
                  "fused CG: matrix and vectors do not conform");
    GKOMI_REQUIRE(a->strategy == GKOMI_CSR_CLASSICAL, GKOMI_NOT_SUPPORTED,
                  "fused CG: needs the classical (sub-wave) CSR strategy");
    GKOMI_REQUIRE(count >= 0, GKOMI_INVALID_ARGUMENT, "fused CG: negative count");
    GKOMI_REQUIRE(st->issued >= 0, GKOMI_INVALID_ARGUMENT,
                  "fused CG: state not started (issued < 0)");
    hipStream_t s = stream_of(ctx);
    const cg_work w = cg_work_layout(st->work);
    const T* v = static_cast<const T*>(a->values);
    T* sc = static_cast<T*>(st->scalars);
    const int64_t n = st->n;
    const cg_grids g = cg_grids_for<T>(a, n);
    const bool sliced = has_slices(a);
    const bool split = sliced || (n > 0 && static_cast<double>(a->nnz) / static_cast<double>(n) >=
                                               kCgSplitMeanRow);
    for (int32_t it = 0; it < count; ++it) {
        T* pin = static_cast<T*>(st->p_a);
        T* pout = static_cast<T*>(st->p_b);
        const int64_t iter = st->issued;
        const int par = static_cast<int>(iter & 1);
        if (split) {
            cg_split_p_kernel<T><<<g.update, kUpdBlock, 0, s>>>(
                n, static_cast<const T*>(st->z), pin, pout, *crit, sc,
                static_cast<const T*>(st->orig_tau), st->status, st->flags, w, g.update, iter,
                par);
            GKOMI_CHECK_LAUNCH("cg_split_p_kernel");
            if (sliced) {
                const T* sv = static_cast<const T*>(a->slice_values);
                if (a->slice_offsets) {
                    cg_split_spmv_sliced_kernel<T, true><<<g.spmv, kBlock, 0, s>>>(
                        n, a->row_ptrs, a->slice_sets, a->slice_cols, a->slice_offsets, sv, pout,
                        static_cast<T*>(st->q), st->flags, w);
                } else {
                    cg_split_spmv_sliced_kernel<T, false><<<g.spmv, kBlock, 0, s>>>(
                        n, a->row_ptrs, a->slice_sets, a->slice_cols, a->slice_offsets, sv, pout,
                        static_cast<T*>(st->q), st->flags, w);
                }
                GKOMI_CHECK_LAUNCH("cg_split_spmv_sliced_kernel");
            } else if (g.staged > 0) {
#define GKOMI_STAGED_SPLIT(GV, C)                                                          \
    if (g.staged == GV) {                                                                  \
        cg_split_spmv_staged_kernel<T, GV, C><<<g.spmv, kBlock, 0, s>>>(                   \
            n, a->row_ptrs, a->col_idxs, a->col_offsets, v, pout, static_cast<T*>(st->q),  \
            st->flags, w);                                                                 \
    }
                if (a->col_offsets) {
                    GKOMI_STAGED_SPLIT(1, true) GKOMI_STAGED_SPLIT(2, true)
                    GKOMI_STAGED_SPLIT(4, true) GKOMI_STAGED_SPLIT(8, true)
                    GKOMI_STAGED_SPLIT(16, true)
                } else {
                    GKOMI_STAGED_SPLIT(1, false) GKOMI_STAGED_SPLIT(2, false)
                    GKOMI_STAGED_SPLIT(4, false) GKOMI_STAGED_SPLIT(8, false)
                    GKOMI_STAGED_SPLIT(16, false)
                }
#undef GKOMI_STAGED_SPLIT
                GKOMI_CHECK_LAUNCH("cg_split_spmv_staged_kernel");
            } else {
#define GKOMI_SPLIT_CASE_C(S, R, C)                                                        \
    cg_split_spmv_kernel<T, S, R, C><<<g.spmv, kBlock, 0, s>>>(                            \
        n, a->row_ptrs, a->col_idxs, a->col_offsets, v, pout, static_cast<T*>(st->q),      \
        st->flags, w);                                                                     \
    break;
#define GKOMI_SPLIT_CASE(S, R) GKOMI_SPLIT_CASE_C(S, R, false)
#define GKOMI_SPLIT_CASE_COMPACT(S, R) GKOMI_SPLIT_CASE_C(S, R, true)
            if (a->col_offsets) {
                GKOMI_SUBWAVE_SWITCH(g.sh, GKOMI_SPLIT_CASE_COMPACT)
            } else {
                GKOMI_SUBWAVE_SWITCH(g.sh, GKOMI_SPLIT_CASE)
            }
#undef GKOMI_SPLIT_CASE
#undef GKOMI_SPLIT_CASE_COMPACT
#undef GKOMI_SPLIT_CASE_C
            GKOMI_CHECK_LAUNCH("cg_split_spmv_kernel");
            }
        } else if (g.staged > 0) {
            // at most one tile per workgroup: the variant that stages its
            // window before waiting for the partials
            const bool pre = ceildiv(n, int64_t{kWave} / g.staged * kWavesPerBlock) <= g.spmv;
#define GKOMI_STAGED_DIR_P(GV, C, P)                                                       \
    cg_fused_direction_staged_kernel<T, GV, C, P><<<g.spmv, kBlock, 0, s>>>(               \
        n, a->row_ptrs, a->col_idxs, a->col_offsets, v, static_cast<const T*>(st->z), pin, \
        pout, static_cast<T*>(st->q), *crit, sc, static_cast<const T*>(st->orig_tau),      \
        st->status, st->flags, w, g.update, iter, par);
#define GKOMI_STAGED_DIR(GV, C)                                                            \
    if (g.staged == GV) {                                                                  \
        if (pre) {                                                                         \
            GKOMI_STAGED_DIR_P(GV, C, true)                                                \
        } else {                                                                           \
            GKOMI_STAGED_DIR_P(GV, C, false)                                               \
        }                                                                                  \
    }
            if (a->col_offsets) {
                GKOMI_STAGED_DIR(1, true) GKOMI_STAGED_DIR(2, true) GKOMI_STAGED_DIR(4, true)
                GKOMI_STAGED_DIR(8, true) GKOMI_STAGED_DIR(16, true)
            } else {
                GKOMI_STAGED_DIR(1, false) GKOMI_STAGED_DIR(2, false) GKOMI_STAGED_DIR(4, false)
                GKOMI_STAGED_DIR(8, false) GKOMI_STAGED_DIR(16, false)
            }
#undef GKOMI_STAGED_DIR
#undef GKOMI_STAGED_DIR_P
            GKOMI_CHECK_LAUNCH("cg_fused_direction_staged_kernel");
        } else {
#define GKOMI_DIR_CASE_C(S, R, C)                                                          \
    cg_fused_direction_kernel<T, S, R, C><<<g.spmv, kBlock, 0, s>>>(                       \
        n, a->row_ptrs, a->col_idxs, a->col_offsets, v, static_cast<const T*>(st->z), pin, \
        pout, static_cast<T*>(st->q), *crit, sc, static_cast<const T*>(st->orig_tau),      \
        st->status, st->flags, w, g.update, iter, par);                                    \
    break;
#define GKOMI_DIR_CASE(S, R) GKOMI_DIR_CASE_C(S, R, false)
#define GKOMI_DIR_CASE_COMPACT(S, R) GKOMI_DIR_CASE_C(S, R, true)
            if (a->col_offsets) {
                GKOMI_SUBWAVE_SWITCH(g.sh, GKOMI_DIR_CASE_COMPACT)
            } else {
                GKOMI_SUBWAVE_SWITCH(g.sh, GKOMI_DIR_CASE)
            }
#undef GKOMI_DIR_CASE
#undef GKOMI_DIR_CASE_COMPACT
#undef GKOMI_DIR_CASE_C
            GKOMI_CHECK_LAUNCH("cg_fused_direction_kernel");
        }
        if (st->inv_diag) {
            cg_fused_update_kernel<T, true, true><<<g.update, kUpdBlock, 0, s>>>(
                n, static_cast<T*>(st->x), static_cast<T*>(st->r), static_cast<T*>(st->z), pout,
                static_cast<const T*>(st->q), static_cast<const T*>(st->inv_diag), *crit, sc,
                static_cast<T*>(st->orig_tau), st->flags, w, g.spmv, par);
        } else {
            cg_fused_update_kernel<T, true, false><<<g.update, kUpdBlock, 0, s>>>(
                n, static_cast<T*>(st->x), static_cast<T*>(st->r), static_cast<T*>(st->z), pout,
                static_cast<const T*>(st->q), nullptr, *crit, sc, static_cast<T*>(st->orig_tau),
                st->flags, w, g.spmv, par);
        }
        GKOMI_CHECK_LAUNCH("cg_fused_update_kernel");
        std::swap(st->p_a, st->p_b);
        ++st->issued;
    }
    return GKOMI_SUCCESS;
}

// ------------------------------------------------- step-kernel drivers ----

inline unsigned egrid(int64_t total) { return grid_for(total, kBlock * 4); }

int check_block(int64_t n, int64_t nrhs, int64_t vs)
{
    GKOMI_REQUIRE(n >= 0 && nrhs >= 0, GKOMI_INVALID_ARGUMENT, "negative size");
    GKOMI_REQUIRE(vs >= nrhs, GKOMI_INVALID_ARGUMENT, "stride smaller than the columns");
    return GKOMI_SUCCESS;
}

}  // namespace gkomi

using namespace gkomi;

#define GKOMI_BEGIN(n, nrhs, vs)                            \
    if (int rc_ = enter(ctx)) return rc_;                   \
    if (int rc_ = check_block(n, nrhs, vs)) return rc_;     \
    if ((n) * (nrhs) == 0) return GKOMI_SUCCESS;            \
    hipStream_t s_ = stream_of(ctx);

extern "C" {

#define GKOMI_CG_INIT(S, T)                                                                    \
    int gkomi_cg_initialize_##S(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const T* b,     \
                                int64_t bs, T* r, T* z, T* p, T* q, int64_t vs, T* prev_rho,   \
                                T* rho, uint8_t* status)                                       \
    {                                                                                          \
        GKOMI_BEGIN(n, nrhs, vs)                                                               \
        cg_initialize_kernel<T><<<egrid(n * nrhs), kBlock, 0, s_>>>(n, nrhs, b, bs, r, z, p, q, \
                                                                   vs, prev_rho, rho, status); \
        GKOMI_CHECK_LAUNCH("cg_initialize");                                                   \
        return GKOMI_SUCCESS;                                                                  \
    }                                                                                          \
    int gkomi_cg_step_1_##S(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, T* p, const T* z,   \
                            int64_t vs, const T* rho, const T* prev_rho,                       \
                            const uint8_t* status)                                             \
    {                                                                                          \
        GKOMI_BEGIN(n, nrhs, vs)                                                               \
        cg_step_1_kernel<T><<<egrid(n * nrhs), kBlock, 0, s_>>>(n, nrhs, p, z, vs, rho,        \
                                                               prev_rho, status);              \
        GKOMI_CHECK_LAUNCH("cg_step_1");                                                       \
        return GKOMI_SUCCESS;                                                                  \
    }                                                                                          \
    int gkomi_cg_step_2_##S(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, T* x, int64_t xs,   \
                            T* r, const T* p, const T* q, int64_t vs, const T* beta,           \
                            const T* rho, const uint8_t* status)                               \
    {                                                                                          \
        GKOMI_BEGIN(n, nrhs, vs)                                                               \
        GKOMI_REQUIRE(xs >= nrhs, GKOMI_INVALID_ARGUMENT, "x stride");                         \
        cg_step_2_kernel<T><<<egrid(n * nrhs), kBlock, 0, s_>>>(n, nrhs, x, xs, r, p, q, vs,   \
                                                               beta, rho, status);             \
        GKOMI_CHECK_LAUNCH("cg_step_2");                                                       \
        return GKOMI_SUCCESS;                                                                  \
    }                                                                                          \
    int gkomi_bicgstab_initialize_##S(const gkomi_ctx* ctx, int64_t n, int64_t nrhs,          \
                                      const T* b, int64_t bs, T* r, T* rr, T* y, T* s, T* t,   \
                                      T* z, T* v, T* p, int64_t vs, T* prev_rho, T* rho,       \
                                      T* alpha, T* beta, T* gamma, T* omega, uint8_t* status)  \
    {                                                                                          \
        GKOMI_BEGIN(n, nrhs, vs)                                                               \
        bicgstab_initialize_kernel<T><<<egrid(n * nrhs), kBlock, 0, s_>>>(                     \
            n, nrhs, b, bs, r, rr, y, s, t, z, v, p, vs, prev_rho, rho, alpha, beta, gamma,    \
            omega, status);                                                                    \
        GKOMI_CHECK_LAUNCH("bicgstab_initialize");                                             \
        return GKOMI_SUCCESS;                                                                  \
    }                                                                                          \
    int gkomi_bicgstab_step_1_##S(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const T* r,  \
                                  T* p, const T* v, int64_t vs, const T* rho,                  \
                                  const T* prev_rho, const T* alpha, const T* omega,           \
                                  const uint8_t* status)                                       \
    {                                                                                          \
        GKOMI_BEGIN(n, nrhs, vs)                                                               \
        bicgstab_step_1_kernel<T><<<egrid(n * nrhs), kBlock, 0, s_>>>(                         \
            n, nrhs, r, p, v, vs, rho, prev_rho, alpha, omega, status);                        \
        GKOMI_CHECK_LAUNCH("bicgstab_step_1");                                                 \
        return GKOMI_SUCCESS;                                                                  \
    }                                                                                          \
    int gkomi_bicgstab_step_2_##S(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const T* r,  \
                                  T* s, const T* v, int64_t vs, const T* rho, T* alpha,        \
                                  const T* beta, const uint8_t* status)                        \
    {                                                                                          \
        GKOMI_BEGIN(n, nrhs, vs)                                                               \
        bicgstab_step_2_kernel<T><<<egrid(n * nrhs), kBlock, 0, s_>>>(n, nrhs, r, s, v, vs,   \
                                                                     rho, alpha, beta, status); \
        GKOMI_CHECK_LAUNCH("bicgstab_step_2");                                                 \
        return GKOMI_SUCCESS;                                                                  \
    }                                                                                          \
    int gkomi_bicgstab_step_3_##S(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, T* x,        \
                                  int64_t xs, T* r, const T* s, const T* t, const T* y,        \
                                  const T* z, int64_t vs, const T* alpha, const T* beta,       \
                                  const T* gamma, T* omega, const uint8_t* status)             \
    {                                                                                          \
        GKOMI_BEGIN(n, nrhs, vs)                                                               \
        GKOMI_REQUIRE(xs >= nrhs, GKOMI_INVALID_ARGUMENT, "x stride");                         \
        bicgstab_step_3_kernel<T><<<egrid(n * nrhs), kBlock, 0, s_>>>(                         \
            n, nrhs, x, xs, r, s, t, y, z, vs, alpha, beta, gamma, omega, status);             \
        bicgstab_omega_kernel<T><<<1, kBlock, 0, s_>>>(nrhs, beta, gamma, omega, status);      \
        GKOMI_CHECK_LAUNCH("bicgstab_step_3");                                                 \
        return GKOMI_SUCCESS;                                                                  \
    }                                                                                          \
    int gkomi_bicgstab_finalize_##S(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, T* x,      \
                                    int64_t xs, const T* y, int64_t vs, const T* alpha,        \
                                    uint8_t* status)                                           \
    {                                                                                          \
        if (int rc_ = enter(ctx)) return rc_;                                                  \
        if (int rc_ = check_block(n, nrhs, vs)) return rc_;                                    \
        hipStream_t s_ = stream_of(ctx);                                                       \
        if (n * nrhs > 0) {                                                                    \
            bicgstab_finalize_kernel<T><<<egrid(n * nrhs), kBlock, 0, s_>>>(n, nrhs, x, xs, y, \
                                                                           vs, alpha, status); \
        }                                                                                      \
        if (n * nrhs > 0) status_finalize_kernel<<<1, kBlock, 0, s_>>>(nrhs, status);          \
        GKOMI_CHECK_LAUNCH("bicgstab_finalize");                                               \
        return GKOMI_SUCCESS;                                                                  \
    }                                                                                          \
    int gkomi_criteria_check_##S(const gkomi_ctx* ctx, const gkomi_criteria* crit,             \
                                 int64_t nrhs, const T* tau, const T* implicit_sq,             \
                                 const T* orig_tau, int32_t set_finalized, uint8_t* status,    \
                                 gkomi_stop_flags* flags)                                      \
    {                                                                                          \
        if (int rc_ = enter(ctx)) return rc_;                                                  \
        if (int rc_ = check_crit(crit)) return rc_;                                            \
        GKOMI_REQUIRE(status && flags && nrhs >= 0, GKOMI_INVALID_ARGUMENT,                    \
                      "criteria check: null status/flags");                                   \
        for (int k = 0; k < crit->num; ++k) {                                                  \
            if (crit->c[k].kind == GKOMI_CRIT_IMPLICIT_RESNORM) {                              \
                GKOMI_REQUIRE(implicit_sq && orig_tau, GKOMI_NOT_SUPPORTED,                    \
                              "ImplicitResidualNorm needs the implicit squared norm");        \
            }                                                                                  \
            if (crit->c[k].kind == GKOMI_CRIT_RESIDUAL_NORM) {                                 \
                GKOMI_REQUIRE(orig_tau, GKOMI_INVALID_ARGUMENT, "null baselines");            \
            }                                                                                  \
        }                                                                                      \
        criteria_check_kernel<T><<<1, 64, 0, stream_of(ctx)>>>(                                \
            *crit, nrhs, tau, implicit_sq, orig_tau, set_finalized, status, flags);            \
        GKOMI_CHECK_LAUNCH("criteria_check");                                                  \
        return GKOMI_SUCCESS;                                                                  \
    }                                                                                          \
    int gkomi_cg_fused_start_##S(const gkomi_ctx* ctx, const gkomi_csr* a,                     \
                                 const gkomi_criteria* crit, gkomi_cg_state* st)               \
    {                                                                                          \
        return cg_fused_start_impl<T>(ctx, a, crit, st);                                       \
    }                                                                                          \
    int gkomi_cg_fused_iterate_##S(const gkomi_ctx* ctx, const gkomi_csr* a,                   \
                                   const gkomi_criteria* crit, gkomi_cg_state* st,             \
                                   int32_t count)                                              \
    {                                                                                          \
        return cg_fused_iterate_impl<T>(ctx, a, crit, st, count);                              \
    }

GKOMI_CG_INIT(f64, double)
GKOMI_CG_INIT(f32, float)

size_t gkomi_cg_work_bytes(int64_t n)
{
    (void)n;
    return sizeof(double) * 5 * static_cast<size_t>(kCgMaxGrid) + 64;
}

void* gkomi_cg_current_p(const gkomi_cg_state* st) { return st ? st->p_a : nullptr; }

}  // extern "C"
