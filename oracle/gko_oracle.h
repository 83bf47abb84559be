/*
 * CPU oracle for the sparse solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a sequential (plus OpenMP for
 * the CPU-baseline timings) restatement of Ginkgo's ReferenceExecutor
 * kernels and solver loops.  It is loaded only by the test-suite, by the
 * smoke check and by the CPU-baseline leg of bench.py, where it serves as
 * the checker.  No product code path links, loads or calls it.
 *
 * Every function names the reference file:line it restates.  Value types
 * come in two flavours: suffix _d (double) and _s (float).  Indices are
 * int32 (the reference's default index type for these configurations);
 * SELL-P slice sets/lengths are int64 (the reference's size_type).
 *
 * Dense operands follow matrix::view::dense (device_views.hpp:24-45):
 * element (i, j) of an n x k block lives at ptr[i * stride + j].
 */
#ifndef GKO_ORACLE_H_
#define GKO_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* stopping_status bit layout, include/ginkgo/core/stop/stopping_status.hpp:115-118 */
#define ORA_STATUS_CONVERGED 0x80u
#define ORA_STATUS_FINALIZED 0x40u
#define ORA_STATUS_ID_MASK 0x3fu

/* One entry of a stop::Combined list (core/stop/combined.cpp:40-60). */
enum {
    ORA_CRIT_ITERATION = 1,         /* stop::Iteration, core/stop/iteration.cpp:14-24 */
    ORA_CRIT_RESIDUAL_NORM = 2,     /* stop::ResidualNorm, core/stop/residual_norm.cpp:165-207 */
    ORA_CRIT_IMPLICIT_RESNORM = 3   /* stop::ImplicitResidualNorm, residual_norm.cpp:210-226 */
};
/* stop::mode, include/ginkgo/core/stop/residual_norm.hpp:37 */
enum { ORA_BASELINE_RHS_NORM = 0, ORA_BASELINE_INITIAL_RESNORM = 1, ORA_BASELINE_ABSOLUTE = 2 };

typedef struct {
    int32_t kind;
    int32_t baseline;
    int64_t max_iters;
    double reduction_factor;
} ora_criterion;

/* Result record of one solve. */
typedef struct {
    int64_t iterations;      /* value of `iter` at loop exit (the Convergence logger's count) */
    int64_t spmv_count;      /* number of SpMV applications issued */
    double final_resnorm[64];/* ||b - A x||_2 recomputed in the value type, first 64 columns */
} ora_solve_info;

/* ---- SpMV family -------------------------------------------------------- */
#define ORA_DECL_SPMV(T, S)                                                        \
    void ora_csr_spmv_##S(int64_t n_rows, int64_t nrhs, const int32_t* row_ptrs,   \
                          const int32_t* col_idxs, const T* vals, const T* b,      \
                          int64_t b_stride, T* c, int64_t c_stride);               \
    void ora_csr_advanced_spmv_##S(T alpha, int64_t n_rows, int64_t nrhs,          \
                                   const int32_t* row_ptrs, const int32_t* col_idxs,\
                                   const T* vals, const T* b, int64_t b_stride,    \
                                   T beta, T* c, int64_t c_stride);                \
    void ora_csr_spmv_omp_##S(int64_t n_rows, int64_t nrhs, const int32_t* row_ptrs,\
                              const int32_t* col_idxs, const T* vals, const T* b,  \
                              int64_t b_stride, T* c, int64_t c_stride);           \
    void ora_csr_to_sellp_##S(int64_t n_rows, const int32_t* row_ptrs,             \
                              const int32_t* col_idxs, const T* vals,              \
                              int64_t slice_size, const int64_t* slice_sets,       \
                              const int64_t* slice_lengths, int32_t* s_cols,       \
                              T* s_vals);                                          \
    void ora_sellp_spmv_##S(int64_t n_rows, int64_t nrhs, int64_t slice_size,      \
                            const int64_t* slice_sets, const int64_t* slice_lengths,\
                            const int32_t* s_cols, const T* s_vals, const T* b,    \
                            int64_t b_stride, T* c, int64_t c_stride);             \
    void ora_sellp_advanced_spmv_##S(T alpha, int64_t n_rows, int64_t nrhs,        \
                                     int64_t slice_size, const int64_t* slice_sets,\
                                     const int64_t* slice_lengths,                 \
                                     const int32_t* s_cols, const T* s_vals,       \
                                     const T* b, int64_t b_stride, T beta, T* c,   \
                                     int64_t c_stride);                            \
    void ora_sellp_spmv_omp_##S(int64_t n_rows, int64_t nrhs, int64_t slice_size,  \
                                const int64_t* slice_sets,                         \
                                const int64_t* slice_lengths,                      \
                                const int32_t* s_cols, const T* s_vals, const T* b,\
                                int64_t b_stride, T* c, int64_t c_stride);         \
    void ora_csr_to_ell_##S(int64_t n_rows, const int32_t* row_ptrs,               \
                            const int32_t* col_idxs, const T* vals,                \
                            int64_t ell_width, int64_t stride, int32_t* e_cols,    \
                            T* e_vals);                                            \
    void ora_ell_spmv_##S(int64_t n_rows, int64_t nrhs, int64_t ell_width,         \
                          int64_t stride, const int32_t* e_cols, const T* e_vals,  \
                          const T* b, int64_t b_stride, T* c, int64_t c_stride);   \
    void ora_ell_advanced_spmv_##S(T alpha, int64_t n_rows, int64_t nrhs,          \
                                   int64_t ell_width, int64_t stride,              \
                                   const int32_t* e_cols, const T* e_vals,         \
                                   const T* b, int64_t b_stride, T beta, T* c,     \
                                   int64_t c_stride);                              \
    void ora_csr_transpose_##S(int64_t n_rows, int64_t n_cols,                     \
                               const int32_t* row_ptrs, const int32_t* col_idxs,   \
                               const T* vals, int32_t* t_row_ptrs,                 \
                               int32_t* t_col_idxs, T* t_vals);                    \
    void ora_dense_dot_##S(int64_t n, int64_t nrhs, const T* x, int64_t xs,        \
                           const T* y, int64_t ys, T* result);                     \
    void ora_dense_norm2_##S(int64_t n, int64_t nrhs, const T* x, int64_t xs,      \
                             T* result);                                           \
    void ora_csr_extract_diagonal_##S(int64_t n_rows, const int32_t* row_ptrs,     \
                                      const int32_t* col_idxs, const T* vals,      \
                                      T* diag);                                    \
    void ora_jacobi_invert_diagonal_##S(int64_t n, const T* diag, T* inv_diag);    \
    void ora_jacobi_simple_scalar_apply_##S(int64_t n, int64_t nrhs,               \
                                            const T* inv_diag, const T* b,         \
                                            int64_t bs, T* x, int64_t xs);         \
    void ora_cg_initialize_##S(int64_t n, int64_t nrhs, const T* b, int64_t bs,    \
                               T* r, T* z, T* p, T* q, int64_t vs, T* prev_rho,    \
                               T* rho, uint8_t* status);                           \
    void ora_cg_step_1_##S(int64_t n, int64_t nrhs, T* p, const T* z, int64_t vs,  \
                           const T* rho, const T* prev_rho, const uint8_t* status);\
    void ora_cg_step_2_##S(int64_t n, int64_t nrhs, T* x, int64_t xs, T* r,        \
                           const T* p, const T* q, int64_t vs, const T* beta,      \
                           const T* rho, const uint8_t* status);                   \
    void ora_bicgstab_step_1_##S(int64_t n, int64_t nrhs, const T* r, T* p,        \
                                 const T* v, int64_t vs, const T* rho,             \
                                 const T* prev_rho, const T* alpha,                \
                                 const T* omega, const uint8_t* status);           \
    void ora_bicgstab_step_2_##S(int64_t n, int64_t nrhs, const T* r, T* s,        \
                                 const T* v, int64_t vs, const T* rho, T* alpha,   \
                                 const T* beta, const uint8_t* status);            \
    void ora_bicgstab_step_3_##S(int64_t n, int64_t nrhs, T* x, int64_t xs, T* r,  \
                                 const T* s, const T* t, const T* y, const T* z,   \
                                 int64_t vs, const T* alpha, const T* beta,        \
                                 const T* gamma, T* omega, const uint8_t* status); \
    void ora_bicgstab_finalize_##S(int64_t n, int64_t nrhs, T* x, int64_t xs,      \
                                   const T* y, int64_t vs, const T* alpha,         \
                                   uint8_t* status);                               \
    int ora_cg_solve_##S(int64_t n, int64_t nrhs, const int32_t* row_ptrs,         \
                         const int32_t* col_idxs, const T* vals,                   \
                         const T* inv_diag, const T* b, T* x,                      \
                         const ora_criterion* crit, int32_t ncrit, int use_omp,    \
                         uint8_t* status, ora_solve_info* info);                   \
    int ora_bicgstab_solve_##S(int64_t n, int64_t nrhs, const int32_t* row_ptrs,   \
                               const int32_t* col_idxs, const T* vals,             \
                               const T* inv_diag, const T* b, T* x,                \
                               const ora_criterion* crit, int32_t ncrit,           \
                               int use_omp, uint8_t* status,                       \
                               ora_solve_info* info);                              \
    int ora_batch_bicgstab_solve_##S(int64_t num_systems, const int64_t* row_off,  \
                                     const int32_t* row_ptrs,                      \
                                     const int32_t* col_idxs, const T* vals,       \
                                     const T* b, T* x, int use_jacobi,             \
                                     int relative, double tol,                     \
                                     int32_t max_iters, int use_omp,               \
                                     int32_t* iters, T* res_norms);                \
    void ora_coo_spmv2_##S(int64_t nnz, const int32_t* rows, const int32_t* cols,  \
                           const T* vals, const T* alpha, int64_t nrhs, const T* b, \
                           int64_t bs, T* c, int64_t cs);                         \
    void ora_csr_to_hybrid_##S(int64_t n_rows, const int32_t* rp, const int32_t* ci,\
                               const T* v, int64_t width, int64_t stride,          \
                               int32_t* ell_cols, T* ell_vals, int32_t* coo_rows,  \
                               int32_t* coo_cols, T* coo_vals);                    \
    int ora_batch_cg_solve_##S(int64_t num_systems, const int64_t* row_off,        \
                               const int32_t* row_ptrs, const int32_t* col_idxs,   \
                               const T* vals, const T* b, T* x, int use_jacobi,    \
                               int relative, double tol, int32_t max_iters,        \
                               int use_omp, int32_t* iters, T* res_norms);        \
    void ora_batch_csr_spmv_##S(int64_t num, int64_t rows, int64_t cols,           \
                                const int32_t* rp, const int32_t* ci, const T* vals, \
                                const T* alpha, const T* b, const T* beta, T* x);  \
    void ora_batch_ell_spmv_##S(int64_t num, int64_t rows, int64_t cols,           \
                                int64_t width, const int32_t* ci, const T* vals,   \
                                const T* alpha, const T* b, const T* beta, T* x);  \
    void ora_batch_dense_spmv_##S(int64_t num, int64_t rows, int64_t cols,         \
                                  const T* vals, const T* alpha, const T* b,       \
                                  const T* beta, T* x);

ORA_DECL_SPMV(double, d)
ORA_DECL_SPMV(float, s)

/* Index-only work. */
void ora_sellp_compute_slice_sets(int64_t n_rows, const int32_t* row_ptrs,
                                  int64_t slice_size, int64_t stride_factor,
                                  int64_t* slice_sets, int64_t* slice_lengths);
int64_t ora_csr_max_row_nnz(int64_t n_rows, const int32_t* row_ptrs);
/* Csr::load_balance::clac_size and ::process (the srow row partition). */
int64_t ora_csr_lb_clac_size(int64_t nnz, int64_t nwarps, int32_t warp_size,
                             int32_t cuda_strategy, int32_t intel);
void ora_csr_lb_process(int64_t n_rows, const int32_t* row_ptrs, int64_t warp_size,
                        int32_t* srow, int64_t nwarps);

/* Mixed-precision iterative refinement: solver::Ir<double> around an inner
 * Cg<float> on the float copy of A (core/solver/ir.cpp:193-257,
 * examples/mixed-precision-ir/mixed-precision-ir.cpp:96-143).  With
 * inner_cg == 0 the inner solver is the Identity (Richardson), the
 * reference's default when no inner solver is given.  inner_per_sweep
 * (NULL = not wanted) receives the inner iterations of the first
 * sweep_capacity sweeps. */
int ora_ir_solve(int64_t n, const int32_t* row_ptrs, const int32_t* col_idxs,
                 const double* vals, const double* b, double* x,
                 const ora_criterion* outer, int32_t n_outer,
                 const ora_criterion* inner, int32_t n_inner, double relaxation,
                 int inner_cg, int use_omp, uint8_t* status, ora_solve_info* info,
                 int64_t* inner_iterations_total, int64_t* inner_per_sweep,
                 int64_t sweep_capacity);

/* Number of OpenMP threads the _omp variants use. */
int ora_omp_threads(void);

#ifdef __cplusplus
}
#endif

#endif
