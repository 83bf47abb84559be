// Device-side evaluation of stop::Combined over Iteration / ResidualNorm /
// ImplicitResidualNorm (core/stop/combined.cpp:40-60, core/stop/iteration.cpp:
// 14-24, core/stop/residual_norm.cpp:165-226,
// reference/stop/residual_norm_kernels.cpp:26-90).
#pragma once

#include "common.hpp"

namespace gkomi {

__device__ __forceinline__ bool has_stopped(uint8_t s) { return (s & GKOMI_STATUS_ID_MASK) != 0; }

// One thread evaluates the whole list for all columns.  `tau` (explicit
// residual norms) may be null when no ResidualNorm criterion is listed;
// `implicit_sq` may be null when no ImplicitResidualNorm is listed.
template <typename T>
__device__ void combined_check(const gkomi_criteria& crit, int64_t iter, int64_t nrhs,
                               const T* tau, const T* implicit_sq, const T* orig,
                               bool set_finalized, uint8_t* status, int* all_stopped,
                               int* one_changed)
{
    bool one_converged = false;
    bool changed = false;
    for (int k = 0; k < crit.num; ++k) {
        const uint8_t id = static_cast<uint8_t>(k + 1);
        const gkomi_criterion& c = crit.c[k];
        bool result = false;
        if (c.kind == GKOMI_CRIT_ITERATION) {
            result = iter >= c.max_iters;
            if (result) {
                for (int64_t j = 0; j < nrhs; ++j) {
                    if (!has_stopped(status[j])) {
                        uint8_t s = status[j] | (id & GKOMI_STATUS_ID_MASK);
                        if (set_finalized) s |= GKOMI_STATUS_FINALIZED;
                        status[j] = s;
                    }
                }
                changed = true;
            }
        } else {
            const T goal = static_cast<T>(c.reduction_factor);
            const T* o = orig + static_cast<int64_t>(k) * nrhs;
            const T* t = (c.kind == GKOMI_CRIT_RESIDUAL_NORM) ? tau : nullptr;
            if (c.kind == GKOMI_CRIT_RESIDUAL_NORM && tau == nullptr) {
                continue;  // residual check ignored (update_residual's first check)
            }
            bool all = true;
            for (int64_t j = 0; j < nrhs; ++j) {
                T value;
                if (t != nullptr) {
                    value = t[j];
                } else {
                    const T a = implicit_sq[j];
                    value = sqrt(a < T(0) ? -a : a);
                }
                if (value <= goal * o[j]) {
                    if (!has_stopped(status[j])) {
                        uint8_t s = status[j] | GKOMI_STATUS_CONVERGED |
                                    (id & GKOMI_STATUS_ID_MASK);
                        if (set_finalized) s |= GKOMI_STATUS_FINALIZED;
                        status[j] = s;
                    }
                    changed = true;
                }
                if (!has_stopped(status[j])) all = false;
            }
            result = all;
        }
        one_converged |= result;
        if (one_converged) break;
    }
    *all_stopped = one_converged ? 1 : 0;
    *one_changed = changed ? 1 : 0;
}

// combined_check evaluated by a whole workgroup of NT threads, column j in
// thread j % NT: per column the same status updates in the same criterion
// order, and the same flags; "every column stopped" and "a status changed"
// of each criterion are workgroup votes.  Every thread of the workgroup must
// call it; *all_stopped and *one_changed are valid in every thread.
template <int NT, typename T>
__device__ void combined_check_block(const gkomi_criteria& crit, int64_t iter, int64_t nrhs,
                                     const T* tau, const T* implicit_sq, const T* orig,
                                     bool set_finalized, uint8_t* status, int* all_stopped,
                                     int* one_changed)
{
    bool one_converged = false;
    bool changed = false;
    for (int k = 0; k < crit.num; ++k) {
        const uint8_t id = static_cast<uint8_t>(k + 1);
        const gkomi_criterion& c = crit.c[k];
        bool result = false;
        if (c.kind == GKOMI_CRIT_ITERATION) {
            result = iter >= c.max_iters;
            if (result) {
                for (int64_t j = threadIdx.x; j < nrhs; j += NT) {
                    if (!has_stopped(status[j])) {
                        uint8_t st = status[j] | (id & GKOMI_STATUS_ID_MASK);
                        if (set_finalized) st |= GKOMI_STATUS_FINALIZED;
                        status[j] = st;
                    }
                }
                changed = true;
            }
        } else {
            const T goal = static_cast<T>(c.reduction_factor);
            const T* o = orig + static_cast<int64_t>(k) * nrhs;
            const T* t = (c.kind == GKOMI_CRIT_RESIDUAL_NORM) ? tau : nullptr;
            if (c.kind == GKOMI_CRIT_RESIDUAL_NORM && tau == nullptr) {
                continue;  // residual check ignored (update_residual's first check)
            }
            int all = 1, hit = 0;
            for (int64_t j = threadIdx.x; j < nrhs; j += NT) {
                T value;
                if (t != nullptr) {
                    value = t[j];
                } else {
                    const T a = implicit_sq[j];
                    value = sqrt(a < T(0) ? -a : a);
                }
                if (value <= goal * o[j]) {
                    if (!has_stopped(status[j])) {
                        uint8_t st = status[j] | GKOMI_STATUS_CONVERGED |
                                     (id & GKOMI_STATUS_ID_MASK);
                        if (set_finalized) st |= GKOMI_STATUS_FINALIZED;
                        status[j] = st;
                    }
                    hit = 1;
                }
                if (!has_stopped(status[j])) all = 0;
            }
            result = __syncthreads_and(all) != 0;
            changed |= __syncthreads_or(hit) != 0;
        }
        one_converged |= result;
        if (one_converged) break;
    }
    *all_stopped = one_converged ? 1 : 0;
    *one_changed = changed ? 1 : 0;
}

}  // namespace gkomi
