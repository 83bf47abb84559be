/*
 * CPU oracle for the sparse solver hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of Ginkgo's ReferenceExecutor kernels and of the
 * host-side solver loops (core/solver/{cg,bicgstab,ir}.cpp).  It is the
 * checker for the HIP library: the test-suite, the smoke check and the
 * CPU-baseline leg of bench.py load it; nothing in the product does.
 *
 * Parity pinning of each function is listed in tests/test_oracle.py, which
 * checks it against the known-answer tests transcribed from the reference's
 * own test-suite (tests/golden/reference_kats.json), against brute-force
 * dense linear algebra on small inputs, and against closed forms.
 */
#include "gko_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define T double
#define S d
#define SQRT sqrt
#include "oracle_impl.inc"
#undef T
#undef S
#undef SQRT

#define T float
#define S s
#define SQRT sqrtf
#include "oracle_impl.inc"
#undef T
#undef S
#undef SQRT

int ora_omp_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* reference/matrix/sellp_kernels.cpp:109-132 (compute_slice_sets): slice
 * length = max over the slice of ceil(len / stride_factor) * stride_factor,
 * slice_sets = exclusive prefix sum (num_slices + 1 entries). */
void ora_sellp_compute_slice_sets(int64_t n_rows, const int32_t* row_ptrs,
                                  int64_t slice_size, int64_t stride_factor,
                                  int64_t* slice_sets, int64_t* slice_lengths)
{
    const int64_t num_slices = (n_rows + slice_size - 1) / slice_size;
    for (int64_t slice = 0; slice < num_slices; slice++) {
        int64_t len = 0;
        for (int64_t local = 0; local < slice_size; local++) {
            const int64_t row = slice * slice_size + local;
            const int64_t rl = row < n_rows ? row_ptrs[row + 1] - row_ptrs[row] : 0;
            const int64_t padded = (rl + stride_factor - 1) / stride_factor * stride_factor;
            len = padded > len ? padded : len;
        }
        slice_lengths[slice] = len;
    }
    int64_t run = 0;
    for (int64_t slice = 0; slice < num_slices; slice++) {
        slice_sets[slice] = run;
        run += slice_lengths[slice];
    }
    slice_sets[num_slices] = run;
}

/* Csr::load_balance::clac_size (include/ginkgo/core/matrix/csr.hpp:475-505)
 * as the HIP-platform build compiles it (GINKGO_HIP_PLATFORM_HCC set): the
 * CUDA multiples, overridden for the "intel" strategy name, overridden again
 * for a strategy built without cuda_strategy. */
int64_t ora_csr_lb_clac_size(int64_t nnz, int64_t nwarps, int32_t warp_size,
                             int32_t cuda_strategy, int32_t intel)
{
    if (warp_size <= 0) {
        return 0;
    }
    int64_t multiple = 8;
    if (nnz >= (int64_t)2e8) {
        multiple = 2048;
    } else if (nnz >= (int64_t)2e7) {
        multiple = 512;
    } else if (nnz >= (int64_t)2e6) {
        multiple = 128;
    } else if (nnz >= (int64_t)2e5) {
        multiple = 32;
    }
    if (intel) {
        multiple = 8;
        if (nnz >= (int64_t)2e8) {
            multiple = 256;
        } else if (nnz >= (int64_t)2e7) {
            multiple = 32;
        }
    }
    if (!cuda_strategy) {
        multiple = 8;
        if (nnz >= (int64_t)1e7) {
            multiple = 64;
        } else if (nnz >= (int64_t)1e6) {
            multiple = 16;
        }
    }
    const int64_t total = nwarps * multiple;
    const int64_t lines = (nnz + warp_size - 1) / warp_size;
    return lines < total ? lines : total;
}

/* Csr::load_balance::process (include/ginkgo/core/matrix/csr.hpp:417-468):
 * every row drops into the bucket ceil(ceil(row_end / warp_size) * nwarps /
 * ceil(nnz / warp_size)); srow is the inclusive prefix count of the buckets
 * below nwarps.  The loop order is the reference's. */
void ora_csr_lb_process(int64_t n_rows, const int32_t* row_ptrs, int64_t warp_size,
                        int32_t* srow, int64_t nwarps)
{
    if (nwarps <= 0) {
        return;
    }
    for (int64_t i = 0; i < nwarps; i++) {
        srow[i] = 0;
    }
    const int64_t num_elems = row_ptrs[n_rows];
    const int64_t bucket_divider =
        num_elems > 0 ? (num_elems + warp_size - 1) / warp_size : 1;
    for (int64_t i = 0; i < n_rows; i++) {
        const int64_t lines = ((int64_t)row_ptrs[i + 1] + warp_size - 1) / warp_size;
        const int64_t bucket = (lines * nwarps + bucket_divider - 1) / bucket_divider;
        if (bucket < nwarps) {
            srow[bucket]++;
        }
    }
    for (int64_t i = 1; i < nwarps; i++) {
        srow[i] += srow[i - 1];
    }
}

/* common/unified/matrix/csr_kernels.cpp (compute_max_row_nnz), used by
 * Csr::convert_to(Ell*) in core/matrix/csr.cpp:512-529. */
int64_t ora_csr_max_row_nnz(int64_t n_rows, const int32_t* row_ptrs)
{
    int64_t m = 0;
    for (int64_t i = 0; i < n_rows; ++i) {
        const int64_t l = row_ptrs[i + 1] - row_ptrs[i];
        m = l > m ? l : m;
    }
    return m;
}

/* solver::Ir<double>::apply_dense_impl (core/solver/ir.cpp:193-257) with the
 * residual update of core/solver/update_residual.hpp, where the inner solver
 * is Cg<float> on the float copy of A (the mixed-precision-ir example,
 * examples/mixed-precision-ir/mixed-precision-ir.cpp:85-143).  The inner
 * solve starts from the residual itself, as both reference forms do, and
 * its operands are converted through float temporaries
 * (include/ginkgo/core/base/precision_dispatch.hpp:43-125). */
int ora_ir_solve(int64_t n, const int32_t* rp, const int32_t* ci, const double* vals,
                 const double* b, double* x, const ora_criterion* outer,
                 int32_t n_outer, const ora_criterion* inner, int32_t n_inner,
                 double relaxation, int inner_cg, int use_omp, uint8_t* status,
                 ora_solve_info* info, int64_t* inner_iterations_total,
                 int64_t* inner_per_sweep, int64_t sweep_capacity)
{
    const int64_t nnz = rp[n];
    float* vf = (float*)malloc(sizeof(float) * (size_t)(nnz + 1));
    double* residual = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    float* bf = (float*)malloc(sizeof(float) * (size_t)(n + 1));
    float* xf = (float*)malloc(sizeof(float) * (size_t)(n + 1));
    double* orig = (double*)malloc(sizeof(double) * (size_t)(n_outer + 1));
    double tau[1];
    uint8_t inner_status[1];
    if (!vf || !residual || !bf || !xf || !orig) {
        return -1;
    }
    for (int64_t k = 0; k < nnz; ++k) {
        vf[k] = (float)vals[k];
    }
    status[0] = 0;
    memcpy(residual, b, sizeof(double) * (size_t)n);
    residual_any_d(use_omp, n, 1, rp, ci, vals, x, residual);
    criteria_baselines_d(use_omp, outer, n_outer, n, 1, b, residual, orig);
    int64_t iter = -1;
    int64_t inner_total = 0;
    int64_t spmvs = 1;
    int one_changed = 0;
    for (;;) {
        ++iter;
        int all_stopped;
        if (iter == 0) {
            all_stopped = combined_check_d(use_omp, outer, n_outer, iter, n, 1, residual,
                                           0, NULL, orig, 1, status, &one_changed, tau);
        } else {
            all_stopped = combined_check_d(use_omp, outer, n_outer, iter, n, 1, NULL, 1,
                                           NULL, orig, 0, status, &one_changed, tau);
            if (!all_stopped) {
                memcpy(residual, b, sizeof(double) * (size_t)n);
                residual_any_d(use_omp, n, 1, rp, ci, vals, x, residual);
                spmvs++;
                all_stopped = combined_check_d(use_omp, outer, n_outer, iter, n, 1,
                                               residual, 0, NULL, orig, 1, status,
                                               &one_changed, tau);
            }
        }
        if (all_stopped) {
            break;
        }
        if (!inner_cg) {
            /* default inner solver: Identity, i.e. Richardson x += w * r */
            for (int64_t i = 0; i < n; ++i) {
                x[i] += relaxation * residual[i];
            }
            continue;
        }
        for (int64_t i = 0; i < n; ++i) {
            bf[i] = (float)residual[i];
            xf[i] = bf[i];
        }
        ora_solve_info inner_info;
        ora_cg_solve_s(n, 1, rp, ci, vf, NULL, bf, xf, inner, n_inner, use_omp,
                       inner_status, &inner_info);
        if (inner_per_sweep && iter < sweep_capacity) {
            inner_per_sweep[iter] = inner_info.iterations;
        }
        inner_total += inner_info.iterations;
        spmvs += inner_info.spmv_count;
        for (int64_t i = 0; i < n; ++i) {
            x[i] += relaxation * (double)xf[i];
        }
    }
    if (info) {
        info->iterations = iter;
        info->spmv_count = spmvs;
        final_resnorm_d(use_omp, n, 1, rp, ci, vals, b, x, info);
    }
    if (inner_iterations_total) {
        *inner_iterations_total = inner_total;
    }
    free(vf);
    free(residual);
    free(bf);
    free(xf);
    free(orig);
    return 0;
}
