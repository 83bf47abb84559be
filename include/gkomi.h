/*
 * gkomi.h -- C ABI of the MI355X (gfx950) kernel library libgkomi.so.
 *
 * This is the drop-in boundary of the sparse solver hot path: the kernel
 * namespace that Ginkgo reaches through GKO_REGISTER_OPERATION
 * (include/ginkgo/core/base/executor.hpp:419-474), one `extern "C"` symbol
 * per (operation, value type), with index type int32.  Arguments are plain
 * device pointers, sizes and strides; dense operands follow
 * matrix::view::dense (include/ginkgo/core/matrix/device_views.hpp:24-45):
 * element (i, j) of an n x k block is at ptr[i * stride + j].  Scalars that
 * the reference passes as 1x1 / 1xk Dense views (alpha, beta, rho, ...) are
 * device pointers here too.
 *
 * Every entry point returns a gkomi_status (0 = success).  On failure
 * gkomi_last_error() describes the problem.  Work is enqueued on the context's
 * stream and runs asynchronously, as the reference's HIP executor does; the
 * library never frees caller memory and never allocates inside a kernel
 * entry point: scratch comes from caller-provided workspace.
 *
 * Value-type suffixes: _f64 = double, _f32 = float.
 */
#ifndef GKOMI_H_
#define GKOMI_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ basics */

/* Error codes; the C++ layer maps them to the reference's exception types
 * (include/ginkgo/core/base/exception.hpp: HipError, NotSupported,
 * DimensionMismatch, AllocationError). */
typedef enum {
    GKOMI_SUCCESS = 0,
    GKOMI_INVALID_ARGUMENT = 1,
    GKOMI_NOT_SUPPORTED = 2,
    GKOMI_ALLOCATION_ERROR = 3,
    GKOMI_HIP_ERROR = 4,
    GKOMI_DIMENSION_MISMATCH = 5
} gkomi_status;

/* Execution context: a device ordinal and the HIP stream (as an opaque
 * pointer; NULL = the device's null stream) that orders all work.  Mirrors
 * HipExecutor's device id + stream (include/ginkgo/core/base/executor.hpp:
 * 1786-1988, hip/base/executor.hip.cpp:215-245). */
typedef struct {
    int32_t device;
    int32_t reserved;
    void* stream;
} gkomi_ctx;

/* HIP graphs (no reference counterpart; HipExecutor has none): work enqueued
 * on `capture` between gkomi_graph_begin and gkomi_graph_end is recorded,
 * not run, then instantiated; gkomi_graph_launch replays it on ctx->stream.
 * gkomi_graph_begin fills *capture with ctx's device and a new stream of its
 * own in relaxed capture mode; gkomi_graph_end destroys that stream.  A
 * replay runs on the buffers the recorded calls named.  The library's
 * kernels that may be recorded clear memory with fill kernels, never with
 * memset nodes. */
typedef struct gkomi_graph gkomi_graph;
int gkomi_graph_begin(const gkomi_ctx* ctx, gkomi_ctx* capture);
int gkomi_graph_end(gkomi_ctx* capture, gkomi_graph** graph);
int gkomi_graph_launch(const gkomi_ctx* ctx, const gkomi_graph* graph);
int gkomi_graph_destroy(gkomi_graph* graph);

/* Thread-local description of the last failure. */
const char* gkomi_last_error(void);
/* Library version string. */
const char* gkomi_version(void);

/* ------------------------------------------- executor / memory (A1 row) */

typedef struct {
    char name[64];
    char arch[32];              /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    int32_t num_cus;            /* multiProcessorCount */
    int32_t warp_size;          /* 64 on CDNA */
    int64_t max_lds_per_block;  /* sharedMemPerBlock */
    int64_t total_global_mem;
    int32_t clock_khz;
    int32_t memory_clock_khz;
    int32_t memory_bus_width;
    int32_t reserved;
} gkomi_device_info;

/* Executor::get_num_devices / HipExecutor::get_device_id & properties
 * (hip/base/executor.hip.cpp:215-245, include/ginkgo/core/base/executor.hpp:1786-1988). */
int gkomi_get_device_count(int32_t* count);
int gkomi_get_device_info(int32_t device, gkomi_device_info* info);
/* hip/base/stream.hip.cpp (hip_stream): create / destroy a default-priority stream. */
int gkomi_stream_create(int32_t device, void** stream);
int gkomi_stream_destroy(int32_t device, void* stream);
/* Executor::synchronize (include/ginkgo/core/base/executor.hpp:824-831). */
int gkomi_synchronize(const gkomi_ctx* ctx);
/* HipAllocator::allocate / deallocate (hip/base/memory.hip.cpp:56-111). */
int gkomi_malloc(const gkomi_ctx* ctx, size_t bytes, void** ptr);
int gkomi_free(const gkomi_ctx* ctx, void* ptr);
/* Pinned host memory (HipHostAllocator, hip/base/memory.hip.cpp:113-175). */
int gkomi_host_malloc(size_t bytes, void** ptr);
int gkomi_host_free(void* ptr);
/* Executor::copy_from / raw_copy_to (executor.hpp:745-815): stream-ordered copy
 * between any two of host, pinned host and device memory. */
int gkomi_memcpy_async(const gkomi_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Executor::copy_val_to_host (executor.hpp:809-815): blocking copy to host. */
int gkomi_memcpy_to_host(const gkomi_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Timer (hip/base/timer.hip.cpp:25-57): HIP events on the context stream. */
int gkomi_event_create(int32_t device, void** event);
int gkomi_event_destroy(int32_t device, void* event);
int gkomi_event_record(const gkomi_ctx* ctx, void* event);
int gkomi_event_synchronize(void* event);
int gkomi_event_elapsed_ms(void* start, void* stop, float* ms);

/* ------------------------------------------------------------- components */

/* components::fill_array (core/components/fill_array_kernels.hpp); a kernel,
 * safe under graph capture. */
int gkomi_fill_zero(const gkomi_ctx* ctx, void* ptr, size_t bytes);
int gkomi_fill_f64(const gkomi_ctx* ctx, int64_t n, double value, double* ptr);
int gkomi_fill_f32(const gkomi_ctx* ctx, int64_t n, float value, float* ptr);
/* components::convert_precision (core/components/precision_conversion_kernels.hpp). */
int gkomi_convert_f64_to_f32(const gkomi_ctx* ctx, int64_t n, const double* in, float* out);
int gkomi_convert_f32_to_f64(const gkomi_ctx* ctx, int64_t n, const float* in, double* out);

/* Bytes of scratch the reductions below need for an n x nrhs operand. */
size_t gkomi_reduction_workspace_bytes(int64_t n, int64_t nrhs);

/* Dense::compute_conj_dot / compute_norm2 (core/matrix/dense.cpp:400-466;
 * reference/matrix/dense_kernels.cpp:299-352): result[j] per column, on the
 * device.  Per-workgroup FP64 partials, then a one-workgroup combine in
 * fixed order (deterministic).  The workspace needs no initialisation. */
int gkomi_dense_dot_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const double* x,
                        int64_t x_stride, const double* y, int64_t y_stride,
                        double* result, void* workspace, size_t workspace_bytes);
int gkomi_dense_dot_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const float* x,
                        int64_t x_stride, const float* y, int64_t y_stride, float* result,
                        void* workspace, size_t workspace_bytes);
int gkomi_dense_norm2_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const double* x,
                          int64_t x_stride, double* result, void* workspace,
                          size_t workspace_bytes);
int gkomi_dense_norm2_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const float* x,
                          int64_t x_stride, float* result, void* workspace,
                          size_t workspace_bytes);
/* Dense::add_scaled (common/unified/matrix/dense_kernels.template.cpp:225-260):
 * y += alpha[0 or j] * x; alpha has 1 or nrhs entries (alpha_len). */
int gkomi_dense_add_scaled_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs,
                               const double* alpha, int64_t alpha_len, const double* x,
                               int64_t x_stride, double* y, int64_t y_stride);
int gkomi_dense_add_scaled_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs,
                               const float* alpha, int64_t alpha_len, const float* x,
                               int64_t x_stride, float* y, int64_t y_stride);
/* Dense::scale: y *= alpha[0 or j]. */
int gkomi_dense_scale_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const double* alpha,
                          int64_t alpha_len, double* y, int64_t y_stride);
int gkomi_dense_scale_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const float* alpha,
                          int64_t alpha_len, float* y, int64_t y_stride);
/* Dense::copy_from with strides (dense_kernels.template.cpp: copy). */
int gkomi_dense_copy_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const double* x,
                         int64_t x_stride, double* y, int64_t y_stride);
int gkomi_dense_copy_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const float* x,
                         int64_t x_stride, float* y, int64_t y_stride);

/* Value-type codes of the mixed-precision products (*_mixed). */
enum { GKOMI_VALUE_F64 = 0, GKOMI_VALUE_F32 = 1 };

/* ------------------------------------------------------------------ CSR */

/* Csr strategies (include/ginkgo/core/matrix/csr.hpp:176-699).
 * CLASSICAL: a sub-wavefront of S lanes per row, each lane group keeping R
 * rows in flight; `subwave` = S | (R << 8), S a power of two <= 64, R one of
 * 1, 2, 4, 8, either 0 = chosen from the mean row length; or `subwave` =
 * GKOMI_CSR_LANE_ROWS (below).
 * LOAD_BALANCE: the reference's load_balance partition `srow` (built by
 * gkomi_csr_lb_partition, or by the reference's own load_balance::process);
 * each workgroup takes consecutive warps of it and owns whole rows.
 * MERGE_PATH: merge-path partition of rows and entries (csr.hpp:283-301). */
enum { GKOMI_CSR_CLASSICAL = 0, GKOMI_CSR_LOAD_BALANCE = 1, GKOMI_CSR_MERGE_PATH = 2 };
/* CLASSICAL `subwave` = GKOMI_CSR_STAGED | G (G = 1, 2, 4, 8 or 16): G lanes
 * per row, each wave staging its 64 / G rows' entries in LDS with contiguous
 * loads (the per-row summation order of S = G, bit for bit).  `subwave` = 0
 * takes it for a single right-hand side with unit stride with the smallest G
 * for which 64 / G mean rows hold at most 448 entries (the staging chunk), if
 * G <= 4 (FP64) or G <= 8 (FP32): rows of up to 28 / 56 entries on average.  Other
 * right-hand-side shapes use the sub-wavefront kernel; a G outside the list
 * is GKOMI_INVALID_ARGUMENT. */
enum { GKOMI_CSR_STAGED = 0x10000 };

/* Matrix::Csr<V, int32> storage (csr.hpp:104-1690) plus its strategy data.
 * `values` has the value type of the entry point it is passed to. */
typedef struct {
    int64_t n_rows;
    int64_t n_cols;
    int64_t nnz;
    const int32_t* row_ptrs;
    const int32_t* col_idxs;
    const void* values;
    int32_t strategy;
    int32_t subwave;
    /* LOAD_BALANCE: the partition Csr::get_const_srow() with
     * Csr::get_num_srow_elements() entries (csr.hpp:1651-1655), any
     * non-decreasing row indices in [0, n_rows]; num_srow 0 = no partition
     * (the matrix then has no entries, clac_size(0) = 0). */
    const int32_t* srow;
    int64_t num_srow;
    /* Optional (NULL = absent, no reference counterpart): col_idxs[p] - row
     * as int16 for every entry, built by gkomi_csr_col_offsets.  col_idxs
     * stays required; the CLASSICAL single right-hand-side SpMV and the fused
     * CG read these 2-byte offsets instead of the 4-byte indices. */
    const int16_t* col_offsets;
    /* Strategy scratch of MERGE_PATH (the workgroups' carries):
     * gkomi_csr_strategy_work_bytes bytes of device memory kept with the
     * matrix, like srow; no initialisation needed.  Every SpMV through the
     * descriptor writes it, so SpMVs of one matrix must not overlap. */
    void* work;
    int64_t work_bytes;
    /* Optional column panels (NULL = absent, no reference counterpart),
     * built by gkomi_csr_build_panels_*: a copy of the matrix split into
     * num_panels (2, 4 or 8) column panels, so that each XCD gathers from one
     * panel's window of b.  When present, every single right-hand-side SpMV
     * through the descriptor (any strategy, same value type as the build)
     * uses them; they also hold the panels' partial sums, so SpMVs of one
     * matrix must not overlap.  Rebuild after changing the matrix. */
    void* panels;
    int64_t panels_bytes;
    int32_t num_panels;
    /* Optional sliced copy (NULL = absent, no reference counterpart): the
     * same matrix in SELL-P form, 64-row slices and stride factor 1, as
     * gkomi_sellp_compute_slice_sets / gkomi_csr_to_sellp_* build it:
     * slice_sets (num_slices + 1), slice_values, and each slot's column as
     * int32 (slice_cols) or, when slice_offsets is set, as the 2-byte offset
     * col - row (gkomi_sellp_col_offsets).  One lane per row walks its row's
     * slots and stops at the row's length (from row_ptrs), so padding is
     * never read; a row sums its entries in their stored order, in double
     * for either value type (rounded once to the value type).  When
     * present, the single right-hand-side CLASSICAL SpMV and the fused CG
     * read it (same value type as the matrix).  Rebuild after changing the
     * matrix. */
    const int64_t* slice_sets;
    const int32_t* slice_cols;
    const int16_t* slice_offsets;
    const void* slice_values;
} gkomi_csr;

/* Bytes of strategy scratch (gkomi_csr::work) an SpMV with `strategy` needs
 * for an n_rows x ? matrix with nnz entries (0 for CLASSICAL and LOAD_BALANCE). */
int64_t gkomi_csr_strategy_work_bytes(int32_t strategy, int64_t n_rows, int64_t nnz);

/* Csr::load_balance::clac_size (csr.hpp:475-505) as the reference's
 * HIP-platform build compiles it: the number of srow elements for nnz
 * entries.  A HipExecutor-built strategy has nwarps = compute units x 4 and
 * warp_size 64, cuda_strategy 0 (hip/base/executor.hip.cpp:230-238,
 * csr.hpp:379-382); intel_strategy is the "intel" strategy name. */
int64_t gkomi_csr_lb_clac_size(int64_t nnz, int64_t nwarps, int32_t warp_size,
                               int32_t cuda_strategy, int32_t intel_strategy);
/* Csr::load_balance::process (csr.hpp:417-468) on the device: srow[w] for
 * w < nwarps, bit-identical to the reference's host loops (srow[w] is the
 * first row i with ceil(row_ptrs[i+1] / warp_size) * nwarps >
 * w * ceil(nnz / warp_size)).  Integer work, deterministic. */
int gkomi_csr_lb_partition(const gkomi_ctx* ctx, int64_t n_rows, const int32_t* row_ptrs,
                           int64_t nnz, int32_t warp_size, int32_t* srow, int64_t nwarps);
/* Column panels (no reference counterpart): bytes of device memory that
 * gkomi_csr_build_panels_* needs for an n_rows-row matrix with nnz entries
 * in num_panels panels (2, 4 or 8) of value_type (GKOMI_VALUE_*); -1 for
 * invalid arguments.  Panel p holds the entries with columns in
 * [p W, (p + 1) W), W = ceil(n_cols / num_panels), in their row order; one
 * SpMV then reads the matrix once, writes and reads num_panels partial
 * sums per row, and gathers b within one panel per XCD.  The build copies
 * a's row pointers, indices and values (integer work, deterministic); the
 * product sums each row panel by panel, in panel order (deterministic). */
int64_t gkomi_csr_panels_bytes(int64_t n_rows, int64_t nnz, int32_t num_panels,
                               int32_t value_type);
int gkomi_csr_build_panels_f64(const gkomi_ctx* ctx, const gkomi_csr* a, int32_t num_panels,
                               void* panels, int64_t panels_bytes);
int gkomi_csr_build_panels_f32(const gkomi_ctx* ctx, const gkomi_csr* a, int32_t num_panels,
                               void* panels, int64_t panels_bytes);
/* Compact column indices (no reference counterpart): out[p] = col_idxs[p] -
 * row(p) as int16 for every entry.  *fits (device int32) becomes 1 when every
 * offset lies in [-32768, 32767], else 0 (then `out` must not be used). */
int gkomi_csr_col_offsets(const gkomi_ctx* ctx, const gkomi_csr* a, int16_t* out,
                          int32_t* fits);
/* Longest row (csr::compute_max_row_nnz, common/unified/matrix/csr_kernels.cpp);
 * result is written to device memory. */
int gkomi_csr_max_row_nnz(const gkomi_ctx* ctx, int64_t n_rows, const int32_t* row_ptrs,
                          int64_t* result);

/* csr::spmv (core/matrix/csr_kernels.hpp:29-36; reference/matrix/csr_kernels.cpp:47-78):
 * c = A b. */
int gkomi_csr_spmv_f64(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t nrhs,
                       const double* b, int64_t b_stride, double* c, int64_t c_stride);
int gkomi_csr_spmv_f32(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t nrhs,
                       const float* b, int64_t b_stride, float* c, int64_t c_stride);
/* csr::advanced_spmv (csr_kernels.hpp:38-43; reference/matrix/csr_kernels.cpp:84-121):
 * c = alpha A b + beta c; beta == 0 ignores c (NaN included). */
int gkomi_csr_advanced_spmv_f64(const gkomi_ctx* ctx, const double* alpha, const gkomi_csr* a,
                                int64_t nrhs, const double* b, int64_t b_stride,
                                const double* beta, double* c, int64_t c_stride);
int gkomi_csr_advanced_spmv_f32(const gkomi_ctx* ctx, const float* alpha, const gkomi_csr* a,
                                int64_t nrhs, const float* b, int64_t b_stride,
                                const float* beta, float* c, int64_t c_stride);

/* csr::spmv / advanced_spmv with mixed value types (reference/matrix/
 * csr_kernels.cpp:47-121, instantiated by GKO_INSTANTIATE_FOR_EACH_MIXED_VALUE_
 * AND_INDEX_TYPE; reached from Csr::apply_impl through
 * mixed_precision_dispatch_real_complex, core/matrix/csr.cpp:257-315): the
 * matrix values (a_type), b (b_type) and c (c_type) are each f64 or f32
 * (GKOMI_VALUE_*); every product and sum is formed in the highest of the three
 * precisions and rounded into c once.  alpha has the matrix's type, beta
 * c's.  Nothing is converted through temporaries.  With one type throughout
 * this is gkomi_csr_(advanced_)spmv_f64/_f32 with the matrix's strategy; a
 * mixed product uses the sub-wave kernel for every strategy. */
int gkomi_csr_spmv_mixed(const gkomi_ctx* ctx, int32_t a_type, const gkomi_csr* a, int64_t nrhs,
                         int32_t b_type, const void* b, int64_t b_stride, int32_t c_type,
                         void* c, int64_t c_stride);
int gkomi_csr_advanced_spmv_mixed(const gkomi_ctx* ctx, const void* alpha, int32_t a_type,
                                  const gkomi_csr* a, int64_t nrhs, int32_t b_type,
                                  const void* b, int64_t b_stride, const void* beta,
                                  int32_t c_type, void* c, int64_t c_stride);

/* csr::transpose (reference/matrix/csr_kernels.cpp:676-731): the CSR of A^T
 * with each row sorted by column, entries in the reference's stable order.
 * Bit-exact.  t_row_ptrs has n_cols + 1 entries. */
size_t gkomi_csr_transpose_workspace_bytes(int64_t n_rows, int64_t n_cols, int64_t nnz);
int gkomi_csr_transpose_f64(const gkomi_ctx* ctx, const gkomi_csr* a, int32_t* t_row_ptrs,
                            int32_t* t_col_idxs, double* t_values, void* workspace,
                            size_t workspace_bytes);
int gkomi_csr_transpose_f32(const gkomi_ctx* ctx, const gkomi_csr* a, int32_t* t_row_ptrs,
                            int32_t* t_col_idxs, float* t_values, void* workspace,
                            size_t workspace_bytes);

/* --------------------------------------------------------- SELL-P, ELL */

/* matrix::Sellp (include/ginkgo/core/matrix/sellp.hpp:17-18, 379-383): entry k of
 * row r lives at (slice_sets[r / slice_size] + k) * slice_size + r % slice_size. */
typedef struct {
    int64_t n_rows;
    int64_t n_cols;
    int64_t slice_size;
    int64_t stride_factor;
    const int64_t* slice_sets;    /* num_slices + 1 */
    const int64_t* slice_lengths; /* num_slices */
    const int32_t* col_idxs;      /* padding = -1 (invalid_index) */
    const void* values;           /* padding = 0 */
} gkomi_sellp;

/* matrix::Ell (include/ginkgo/core/matrix/ell.hpp:52, 385-388): entry k of row r
 * at r + k * stride. */
typedef struct {
    int64_t n_rows;
    int64_t n_cols;
    int64_t width;
    int64_t stride;
    const int32_t* col_idxs;
    const void* values;
} gkomi_ell;

/* sellp::compute_slice_sets (reference/matrix/sellp_kernels.cpp:109-132):
 * slice_lengths[s] = max over the slice of ceil(len / stride_factor) * stride_factor,
 * slice_sets = exclusive prefix sum with num_slices + 1 entries.  Bit-exact. */
int gkomi_sellp_compute_slice_sets(const gkomi_ctx* ctx, int64_t n_rows,
                                   const int32_t* row_ptrs, int64_t slice_size,
                                   int64_t stride_factor, int64_t* slice_sets,
                                   int64_t* slice_lengths, void* workspace,
                                   size_t workspace_bytes);
size_t gkomi_scan_workspace_bytes(int64_t n);
/* csr::convert_to_sellp (reference/matrix/csr_kernels.cpp:528-569).  Storage past
 * the last row of the last slice is written as padding too. */
int gkomi_csr_to_sellp_f64(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t slice_size,
                           const int64_t* slice_sets, const int64_t* slice_lengths,
                           int32_t* s_cols, double* s_vals);
int gkomi_csr_to_sellp_f32(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t slice_size,
                           const int64_t* slice_sets, const int64_t* slice_lengths,
                           int32_t* s_cols, float* s_vals);
/* Compact slot columns of a SELL-P matrix with 64-row slices (no reference
 * counterpart): out[slot] = col - row as int16 for every slot, 0 for padding.
 * *fits (device int32) becomes 1 when every offset lies in [-32768, 32767],
 * else 0 (then `out` must not be used). */
int gkomi_sellp_col_offsets(const gkomi_ctx* ctx, const gkomi_sellp* a, int16_t* out,
                            int32_t* fits);
/* sellp::spmv / advanced_spmv (reference/matrix/sellp_kernels.cpp:26-103). */
int gkomi_sellp_spmv_f64(const gkomi_ctx* ctx, const gkomi_sellp* a, int64_t nrhs,
                         const double* b, int64_t b_stride, double* c, int64_t c_stride);
int gkomi_sellp_spmv_f32(const gkomi_ctx* ctx, const gkomi_sellp* a, int64_t nrhs,
                         const float* b, int64_t b_stride, float* c, int64_t c_stride);
int gkomi_sellp_advanced_spmv_f64(const gkomi_ctx* ctx, const double* alpha,
                                  const gkomi_sellp* a, int64_t nrhs, const double* b,
                                  int64_t b_stride, const double* beta, double* c,
                                  int64_t c_stride);
int gkomi_sellp_advanced_spmv_f32(const gkomi_ctx* ctx, const float* alpha,
                                  const gkomi_sellp* a, int64_t nrhs, const float* b,
                                  int64_t b_stride, const float* beta, float* c,
                                  int64_t c_stride);
/* csr::convert_to_ell (reference/matrix/csr_kernels.cpp:575-600). */
int gkomi_csr_to_ell_f64(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t width,
                         int64_t stride, int32_t* e_cols, double* e_vals);
int gkomi_csr_to_ell_f32(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t width,
                         int64_t stride, int32_t* e_cols, float* e_vals);
/* ell::spmv / advanced_spmv (reference/matrix/ell_kernels.cpp:29-110). */
int gkomi_ell_spmv_f64(const gkomi_ctx* ctx, const gkomi_ell* a, int64_t nrhs, const double* b,
                       int64_t b_stride, double* c, int64_t c_stride);
int gkomi_ell_spmv_f32(const gkomi_ctx* ctx, const gkomi_ell* a, int64_t nrhs, const float* b,
                       int64_t b_stride, float* c, int64_t c_stride);
int gkomi_ell_advanced_spmv_f64(const gkomi_ctx* ctx, const double* alpha, const gkomi_ell* a,
                                int64_t nrhs, const double* b, int64_t b_stride,
                                const double* beta, double* c, int64_t c_stride);
int gkomi_ell_advanced_spmv_f32(const gkomi_ctx* ctx, const float* alpha, const gkomi_ell* a,
                                int64_t nrhs, const float* b, int64_t b_stride,
                                const float* beta, float* c, int64_t c_stride);
/* sellp::spmv / advanced_spmv (reference/matrix/sellp_kernels.cpp:26-95) and
 * ell::spmv / advanced_spmv (reference/matrix/ell_kernels.cpp:29-110) with
 * mixed value types, as gkomi_csr_spmv_mixed: computed in the highest
 * precision of the matrix, b and c (KATs reference/test/matrix/
 * sellp_kernels.cpp:68-79, 132-145). */
int gkomi_sellp_spmv_mixed(const gkomi_ctx* ctx, int32_t a_type, const gkomi_sellp* a,
                           int64_t nrhs, int32_t b_type, const void* b, int64_t b_stride,
                           int32_t c_type, void* c, int64_t c_stride);
int gkomi_sellp_advanced_spmv_mixed(const gkomi_ctx* ctx, const void* alpha, int32_t a_type,
                                    const gkomi_sellp* a, int64_t nrhs, int32_t b_type,
                                    const void* b, int64_t b_stride, const void* beta,
                                    int32_t c_type, void* c, int64_t c_stride);
int gkomi_ell_spmv_mixed(const gkomi_ctx* ctx, int32_t a_type, const gkomi_ell* a, int64_t nrhs,
                         int32_t b_type, const void* b, int64_t b_stride, int32_t c_type,
                         void* c, int64_t c_stride);
int gkomi_ell_advanced_spmv_mixed(const gkomi_ctx* ctx, const void* alpha, int32_t a_type,
                                  const gkomi_ell* a, int64_t nrhs, int32_t b_type,
                                  const void* b, int64_t b_stride, const void* beta,
                                  int32_t c_type, void* c, int64_t c_stride);

/* ---------------------------------------------------------- COO, Hybrid */

/* components::convert_idxs_to_ptrs / convert_ptrs_to_idxs
 * (core/components/format_conversion_kernels.hpp): ptrs[r] = number of
 * entries whose row is below r (r = 0..n, idxs sorted); idxs[e] = the row of
 * entry e.  Integer work, bit-exact. */
int gkomi_idxs_to_ptrs(const gkomi_ctx* ctx, int64_t nnz, const int32_t* idxs, int64_t n,
                       int32_t* ptrs);
int gkomi_ptrs_to_idxs(const gkomi_ctx* ctx, int64_t n, const int32_t* ptrs, int64_t nnz,
                       int32_t* idxs);
/* coo::spmv2 / advanced_spmv2 (core/matrix/coo_kernels.hpp;
 * reference/matrix/coo_kernels.cpp:61-101): c += alpha A b, alpha NULL
 * meaning 1.  A matrix::Coo (include/ginkgo/core/matrix/coo.hpp, rows sorted)
 * is passed as the gkomi_csr of its entries over their row-pointer index
 * (gkomi_idxs_to_ptrs), so the product reads 4 (n + 1) bytes of row pointers
 * instead of 4 nnz bytes of row indices and runs the CSR kernels of the
 * descriptor's strategy.  coo::spmv and advanced_spmv are gkomi_csr_spmv_* and
 * gkomi_csr_advanced_spmv_* on the same descriptor. */
int gkomi_coo_spmv2_f64(const gkomi_ctx* ctx, const double* alpha, const gkomi_csr* a,
                        int64_t nrhs, const double* b, int64_t b_stride, double* c,
                        int64_t c_stride);
int gkomi_coo_spmv2_f32(const gkomi_ctx* ctx, const float* alpha, const gkomi_csr* a,
                        int64_t nrhs, const float* b, int64_t b_stride, float* c,
                        int64_t c_stride);
/* csr::convert_to_hybrid (reference/matrix/csr_kernels.cpp:911-957) in two
 * steps.  (1) coo_row_ptrs[r] = number of entries of rows < r beyond the ELL
 * width (n_rows + 1 values: the row-pointer index of the COO part), with
 * gkomi_csr_hybrid_workspace_bytes(n_rows) bytes of scratch.  (2) The first
 * ell_width entries of every row into the column-major ELL part (padding
 * -1 / 0), the rest into the COO part in row order.  Bit-exact. */
size_t gkomi_csr_hybrid_workspace_bytes(int64_t n_rows);
int gkomi_csr_hybrid_coo_row_ptrs(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t ell_width,
                                  int32_t* coo_row_ptrs, void* workspace,
                                  size_t workspace_bytes);
int gkomi_csr_to_hybrid_f64(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t ell_width,
                            int64_t ell_stride, int32_t* ell_cols, double* ell_vals,
                            const int32_t* coo_row_ptrs, int32_t* coo_row_idxs,
                            int32_t* coo_col_idxs, double* coo_vals);
int gkomi_csr_to_hybrid_f32(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t ell_width,
                            int64_t ell_stride, int32_t* ell_cols, float* ell_vals,
                            const int32_t* coo_row_ptrs, int32_t* coo_row_idxs,
                            int32_t* coo_col_idxs, float* coo_vals);

/* ------------------------------------------------------- scalar Jacobi */

/* Scalar Jacobi generate (core/preconditioner/jacobi.cpp:338-370): extract the
 * diagonal (reference/matrix/csr_kernels.cpp:1318-1336, 0 when absent) and
 * invert it (reference/preconditioner/jacobi_kernels.cpp:582-591, 0 -> 1). */
int gkomi_jacobi_generate_f64(const gkomi_ctx* ctx, const gkomi_csr* a, double* inv_diag);
int gkomi_jacobi_generate_f32(const gkomi_ctx* ctx, const gkomi_csr* a, float* inv_diag);
/* jacobi::simple_scalar_apply (jacobi_kernels.cpp:554-566): x = inv_diag .* b. */
int gkomi_jacobi_apply_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const double* inv_diag,
                           const double* b, int64_t b_stride, double* x, int64_t x_stride);
int gkomi_jacobi_apply_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const float* inv_diag,
                           const float* b, int64_t b_stride, float* x, int64_t x_stride);
/* jacobi::scalar_apply (jacobi_kernels.cpp:534-550): x = beta x + alpha inv_diag .* b. */
int gkomi_jacobi_advanced_apply_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs,
                                    const double* inv_diag, const double* alpha,
                                    const double* b, int64_t b_stride, const double* beta,
                                    double* x, int64_t x_stride);
int gkomi_jacobi_advanced_apply_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs,
                                    const float* inv_diag, const float* alpha, const float* b,
                                    int64_t b_stride, const float* beta, float* x,
                                    int64_t x_stride);

/* ------------------------------------------------ stopping criteria */

/* stopping_status bits (include/ginkgo/core/stop/stopping_status.hpp:115-118). */
#define GKOMI_STATUS_CONVERGED 0x80u
#define GKOMI_STATUS_FINALIZED 0x40u
#define GKOMI_STATUS_ID_MASK 0x3fu

enum {
    GKOMI_CRIT_ITERATION = 1,       /* stop::Iteration (core/stop/iteration.cpp:14-24) */
    GKOMI_CRIT_RESIDUAL_NORM = 2,   /* stop::ResidualNorm (core/stop/residual_norm.cpp:165-207) */
    GKOMI_CRIT_IMPLICIT_RESNORM = 3 /* stop::ImplicitResidualNorm (residual_norm.cpp:210-226) */
};
/* stop::mode (include/ginkgo/core/stop/residual_norm.hpp:37). */
enum { GKOMI_BASELINE_RHS_NORM = 0, GKOMI_BASELINE_INITIAL_RESNORM = 1, GKOMI_BASELINE_ABSOLUTE = 2 };

typedef struct {
    int32_t kind;
    int32_t baseline;
    int64_t max_iters;
    double reduction_factor;
} gkomi_criterion;

/* stop::Combined (core/stop/combined.cpp:40-60): criteria in order, the k-th has
 * stopping id k + 1. */
#define GKOMI_MAX_CRITERIA 4
typedef struct {
    int32_t num;
    int32_t reserved;
    gkomi_criterion c[GKOMI_MAX_CRITERIA];
} gkomi_criteria;

/* Device-resident loop state of a Krylov solve.  `iteration` is the
 * reference's `iter` (core/solver/cg.cpp:145-158); all_stopped / one_changed
 * are the results of the last Combined::check. */
typedef struct {
    int64_t iteration;
    int32_t all_stopped;
    int32_t one_changed;
} gkomi_stop_flags;

/* residual_norm::residual_norm (reference/stop/residual_norm_kernels.cpp:26-56)
 * generalised to the whole Combined list and evaluated on the device:
 * `tau` = explicit residual norms (NULL if not available), `implicit_sq` =
 * implicit squared norms (NULL if not available), orig_tau[k * nrhs + j] =
 * baseline of criterion k.  Reads flags->iteration, updates status and flags. */
int gkomi_criteria_check_f64(const gkomi_ctx* ctx, const gkomi_criteria* crit, int64_t nrhs,
                             const double* tau, const double* implicit_sq,
                             const double* orig_tau, int32_t set_finalized,
                             uint8_t* status, gkomi_stop_flags* flags);
int gkomi_criteria_check_f32(const gkomi_ctx* ctx, const gkomi_criteria* crit, int64_t nrhs,
                             const float* tau, const float* implicit_sq,
                             const float* orig_tau, int32_t set_finalized, uint8_t* status,
                             gkomi_stop_flags* flags);

/* -------------------------------------------- CG (reference semantics) */

/* cg::initialize / step_1 / step_2 (core/solver/cg_kernels.hpp:20-68;
 * reference/solver/cg_kernels.cpp:23-101).  All vectors share stride vs except
 * x (x_stride) and b (b_stride). */
int gkomi_cg_initialize_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const double* b,
                            int64_t b_stride, double* r, double* z, double* p, double* q,
                            int64_t vs, double* prev_rho, double* rho, uint8_t* status);
int gkomi_cg_initialize_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const float* b,
                            int64_t b_stride, float* r, float* z, float* p, float* q,
                            int64_t vs, float* prev_rho, float* rho, uint8_t* status);
int gkomi_cg_step_1_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, double* p,
                        const double* z, int64_t vs, const double* rho,
                        const double* prev_rho, const uint8_t* status);
int gkomi_cg_step_1_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, float* p,
                        const float* z, int64_t vs, const float* rho, const float* prev_rho,
                        const uint8_t* status);
int gkomi_cg_step_2_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, double* x,
                        int64_t x_stride, double* r, const double* p, const double* q,
                        int64_t vs, const double* beta, const double* rho,
                        const uint8_t* status);
int gkomi_cg_step_2_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, float* x,
                        int64_t x_stride, float* r, const float* p, const float* q,
                        int64_t vs, const float* beta, const float* rho,
                        const uint8_t* status);

/* ------------------------------------------------ CG (fused, gfx950) */

/* Device state of a fused CG solve with one right-hand side.  All pointers
 * are device memory of the entry point's value type (V); `work` holds the
 * per-workgroup reduction partials (gkomi_cg_work_bytes bytes, no
 * initialisation needed).  p_a / p_b are used alternately.  `issued` counts
 * the iterations enqueued since the start (set by gkomi_cg_fused_start,
 * advanced by gkomi_cg_fused_iterate); the state is host memory that both
 * calls update. */
typedef struct {
    int64_t n;
    void* x;        /* solution (in: initial guess) */
    const void* b;  /* right-hand side */
    void* r;
    void* z;
    void* p_a;
    void* p_b;
    void* q;
    const void* inv_diag; /* scalar Jacobi, or NULL for the identity */
    void* scalars;  /* 8 values (library-owned layout) */
    void* orig_tau; /* GKOMI_MAX_CRITERIA values */
    uint8_t* status;
    gkomi_stop_flags* flags;
    void* work;
    int64_t issued;
} gkomi_cg_state;

size_t gkomi_cg_work_bytes(int64_t n);
/* Start of Cg::apply_dense_impl (core/solver/cg.cpp:93-144): initialize,
 * r = b - A x, the criteria baselines, z = M r and the partial sums of
 * rho = r.z -- all on the device; the check of iteration 0 runs in the first
 * iterate kernel. */
int gkomi_cg_fused_start_f64(const gkomi_ctx* ctx, const gkomi_csr* a,
                             const gkomi_criteria* crit, gkomi_cg_state* st);
int gkomi_cg_fused_start_f32(const gkomi_ctx* ctx, const gkomi_csr* a,
                             const gkomi_criteria* crit, gkomi_cg_state* st);
/* `count` CG iterations (core/solver/cg.cpp:145-192): per iteration two
 * kernels -- (1) rho = r.z, ||r|| and the Combined check of the iteration,
 * then p = z + rho/prev_rho p fused into q = A p and beta = p.q, (2)
 * x += rho/beta p, r -= rho/beta q, z = M r.  For rows of 16 or more entries
 * on average kernel 1 is split in two (the p update, then q = A p with one
 * gather per entry).  Once flags->all_stopped is set the remaining kernels
 * return immediately, so the host may enqueue ahead. */
int gkomi_cg_fused_iterate_f64(const gkomi_ctx* ctx, const gkomi_csr* a,
                               const gkomi_criteria* crit, gkomi_cg_state* st, int32_t count);
int gkomi_cg_fused_iterate_f32(const gkomi_ctx* ctx, const gkomi_csr* a,
                               const gkomi_criteria* crit, gkomi_cg_state* st, int32_t count);
/* Pointer to the p buffer that holds the current search direction. */
void* gkomi_cg_current_p(const gkomi_cg_state* st);

/* ------------------------------------------ CG, k right-hand sides (fused) */

/* Device state of a fused CG over k = nrhs right-hand sides sharing one
 * matrix: Cg::apply_dense_impl (core/solver/cg.cpp:93-196) with per-column
 * scalars and stopping status, a stopped column keeping its p, x and r
 * (reference/solver/cg_kernels.cpp:55-95).  x and b are n x nrhs with their
 * own strides; r, z, p_a, p_b, q are n x nrhs with stride nrhs.  Without a
 * preconditioner (inv_diag NULL) z must be the same buffer as r.
 * `scalars` holds 5 * nrhs values, `orig_tau` GKOMI_MAX_CRITERIA * nrhs,
 * `status` nrhs bytes; `work` gkomi_block_cg_work_bytes(n, nrhs) bytes (no
 * initialisation needed); `issued` as in gkomi_cg_state. */
typedef struct {
    int64_t n;
    int64_t nrhs;
    void* x;
    int64_t x_stride;
    const void* b;
    int64_t b_stride;
    void* r;
    void* z;
    void* p_a;
    void* p_b;
    void* q;
    const void* inv_diag;
    void* scalars;
    void* orig_tau;
    uint8_t* status;
    gkomi_stop_flags* flags;
    void* work;
    int64_t issued;
} gkomi_block_cg_state;

size_t gkomi_block_cg_work_bytes(int64_t n, int64_t nrhs);
/* Start: r = b - A x, the baselines per column, z = M r and the partials of
 * r.z (cg.cpp:93-144 with nrhs columns; reference/solver/cg_kernels.cpp:23-50). */
int gkomi_block_cg_start_f64(const gkomi_ctx* ctx, const gkomi_csr* a,
                             const gkomi_criteria* crit, gkomi_block_cg_state* st);
int gkomi_block_cg_start_f32(const gkomi_ctx* ctx, const gkomi_csr* a,
                             const gkomi_criteria* crit, gkomi_block_cg_state* st);
/* `count` iterations (cg.cpp:145-192), four kernels each: the Combined check
 * over all columns, p = z + rho/prev_rho p with q = A p, beta = p.q, and the
 * x / r update with z = M r.  Once flags->all_stopped is set the remaining
 * kernels return immediately. */
int gkomi_block_cg_iterate_f64(const gkomi_ctx* ctx, const gkomi_csr* a,
                               const gkomi_criteria* crit, gkomi_block_cg_state* st,
                               int32_t count);
int gkomi_block_cg_iterate_f32(const gkomi_ctx* ctx, const gkomi_csr* a,
                               const gkomi_criteria* crit, gkomi_block_cg_state* st,
                               int32_t count);
/* Pointer to the p buffer that holds the current search directions. */
void* gkomi_block_cg_current_p(const gkomi_block_cg_state* st);

/* ------------------------------------------ BiCGSTAB (fused, gfx950) */

/* Device state of a fused BiCGSTAB solve with one right-hand side
 * (core/solver/bicgstab.cpp:95-255).  All vectors are device memory of n
 * values of the entry point's type.  y = M p and z = M s are only needed
 * with a preconditioner (inv_diag non-NULL); without one they may be NULL
 * (p and s are used, as the reference's identity copy would give).
 * `scalars` holds 8 values, `orig_tau` GKOMI_MAX_CRITERIA; `work`
 * gkomi_bicgstab_work_bytes(n) bytes (no initialisation needed); `issued`
 * as in gkomi_cg_state. */
typedef struct {
    int64_t n;
    void* x;        /* solution (in: initial guess) */
    const void* b;  /* right-hand side */
    void* r;
    void* rr;       /* the shadow residual r0 */
    void* p;
    void* v;
    void* s;
    void* t;
    void* y;
    void* z;
    const void* inv_diag;
    void* scalars;
    void* orig_tau;
    uint8_t* status;
    gkomi_stop_flags* flags;
    void* work;
    int64_t issued;
} gkomi_bicgstab_state;

size_t gkomi_bicgstab_work_bytes(int64_t n);
/* Start (bicgstab.cpp:103-146; initialize,
 * reference/solver/bicgstab_kernels.cpp:26-58): r = b - A x, rr = r,
 * p = v = 0, scalars = 1, the criteria baselines -- on the device. */
int gkomi_bicgstab_fused_start_f64(const gkomi_ctx* ctx, const gkomi_csr* a,
                                   const gkomi_criteria* crit, gkomi_bicgstab_state* st);
int gkomi_bicgstab_fused_start_f32(const gkomi_ctx* ctx, const gkomi_csr* a,
                                   const gkomi_criteria* crit, gkomi_bicgstab_state* st);
/* `count` iterations (bicgstab.cpp:158-253), five kernels each: rho = rr.r
 * with the Combined check and step_1 (p, y = M p); v = A y with rr.v;
 * step_2 (alpha, s, z = M s) with s.s; the half-step check with t = A z,
 * s.t and t.t; step_3 (omega, x, r) or, after a half-step stop, finalize.
 * Once flags->all_stopped is set the remaining kernels return immediately,
 * so the host may enqueue ahead. */
int gkomi_bicgstab_fused_iterate_f64(const gkomi_ctx* ctx, const gkomi_csr* a,
                                     const gkomi_criteria* crit, gkomi_bicgstab_state* st,
                                     int32_t count);
int gkomi_bicgstab_fused_iterate_f32(const gkomi_ctx* ctx, const gkomi_csr* a,
                                     const gkomi_criteria* crit, gkomi_bicgstab_state* st,
                                     int32_t count);

/* ------------------------------------------- BiCGSTAB (reference semantics) */

/* bicgstab::initialize / step_1 / step_2 / step_3 / finalize
 * (core/solver/bicgstab_kernels.hpp:42-104; reference/solver/bicgstab_kernels.cpp:26-185). */
int gkomi_bicgstab_initialize_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs,
                                  const double* b, int64_t b_stride, double* r, double* rr,
                                  double* y, double* s, double* t, double* z, double* v,
                                  double* p, int64_t vs, double* prev_rho, double* rho,
                                  double* alpha, double* beta, double* gamma, double* omega,
                                  uint8_t* status);
int gkomi_bicgstab_initialize_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const float* b,
                                  int64_t b_stride, float* r, float* rr, float* y, float* s,
                                  float* t, float* z, float* v, float* p, int64_t vs,
                                  float* prev_rho, float* rho, float* alpha, float* beta,
                                  float* gamma, float* omega, uint8_t* status);
int gkomi_bicgstab_step_1_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const double* r,
                              double* p, const double* v, int64_t vs, const double* rho,
                              const double* prev_rho, const double* alpha,
                              const double* omega, const uint8_t* status);
int gkomi_bicgstab_step_1_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const float* r,
                              float* p, const float* v, int64_t vs, const float* rho,
                              const float* prev_rho, const float* alpha, const float* omega,
                              const uint8_t* status);
int gkomi_bicgstab_step_2_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const double* r,
                              double* s, const double* v, int64_t vs, const double* rho,
                              double* alpha, const double* beta, const uint8_t* status);
int gkomi_bicgstab_step_2_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const float* r,
                              float* s, const float* v, int64_t vs, const float* rho,
                              float* alpha, const float* beta, const uint8_t* status);
int gkomi_bicgstab_step_3_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, double* x,
                              int64_t x_stride, double* r, const double* s, const double* t,
                              const double* y, const double* z, int64_t vs,
                              const double* alpha, const double* beta, const double* gamma,
                              double* omega, const uint8_t* status);
int gkomi_bicgstab_step_3_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, float* x,
                              int64_t x_stride, float* r, const float* s, const float* t,
                              const float* y, const float* z, int64_t vs, const float* alpha,
                              const float* beta, const float* gamma, float* omega,
                              const uint8_t* status);
int gkomi_bicgstab_finalize_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, double* x,
                                int64_t x_stride, const double* y, int64_t vs,
                                const double* alpha, uint8_t* status);
int gkomi_bicgstab_finalize_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, float* x,
                                int64_t x_stride, const float* y, int64_t vs,
                                const float* alpha, uint8_t* status);


/* ----------------------------------------------------- batched BiCGSTAB */

/* A batch of independent systems: batch::matrix::Csr, Ell and Dense
 * (include/ginkgo/core/matrix/batch_{csr,ell,dense}.hpp), the Csr form
 * generalised to uneven sizes.  Vectors are concatenated per system.
 *  format GKOMI_BATCH_CSR:
 *   uneven (row_offsets != NULL): system k owns global rows
 *   [row_offsets[k], row_offsets[k+1]) of one block-diagonal CSR whose column
 *   indices are local to the system (square systems);
 *   uniform (row_offsets == NULL): every system has uniform_rows rows and the
 *   shared pattern row_ptrs[uniform_rows + 1] / col_idxs[uniform_nnz];
 *   system k's values start at k * uniform_nnz.
 *  format GKOMI_BATCH_ELL (batch_ell.hpp), uniform (row_offsets == NULL):
 *   col_idxs[uniform_rows * ell_width], shared, column-major (slot j of row i
 *   at i + j * uniform_rows), -1 after the last entry of a row; system k's
 *   values, in the same layout, start at k * uniform_nnz with uniform_nnz =
 *   uniform_rows * ell_width.
 *  format GKOMI_BATCH_DENSE (batch_dense.hpp), uniform: system k's values,
 *   row-major with stride uniform_cols, start at k * uniform_nnz with
 *   uniform_nnz = uniform_rows * uniform_cols; row_ptrs and col_idxs unused.
 *  ELL and DENSE, uneven (row_offsets != NULL; no reference counterpart,
 *   the reference's Ell and Dense batches have one size): system k has
 *   n_k = row_offsets[k+1] - row_offsets[k] rows (square) and its stored
 *   entries at [item_offsets[k], item_offsets[k+1]) -- ELL: n_k * w_k slots,
 *   column-major within the system (slot j of row i at i + j * n_k), its
 *   own local column indices at the same positions of col_idxs, -1 after a
 *   row's last entry; DENSE: n_k * n_k values, row-major.
 * uniform_cols == 0 means uniform_cols = uniform_rows; the solvers need square
 * systems, gkomi_batch_spmv_* takes rectangular ones.  A descriptor with
 * format, ell_width and uniform_cols zero is the Csr form above. */
#define GKOMI_BATCH_MAX_BINS 8
#define GKOMI_BATCH_CSR 0
#define GKOMI_BATCH_ELL 1
#define GKOMI_BATCH_DENSE 2
typedef struct {
    int64_t num_systems;
    int64_t max_rows;
    int64_t max_nnz;
    const int64_t* row_offsets;
    const int32_t* row_ptrs;
    const int32_t* col_idxs;
    const void* values;
    int64_t uniform_rows;
    int64_t uniform_nnz;
    int32_t format;
    int32_t ell_width;
    int64_t uniform_cols;
    const int64_t* item_offsets; /* uneven ELL / DENSE only, num_systems + 1 */
} gkomi_batch_csr;

/* BatchLinOp::apply of batch::matrix::{Csr, Ell, Dense}
 * (common/cuda_hip/matrix/batch_{csr,ell,dense}_kernels.hpp: simple_apply,
 * advanced_apply): per system k, x_k = A_k b_k, or with alpha and beta (one
 * value per system, device arrays) x_k = sum_j (alpha_k a_ij) b_j +
 * beta_k x_k, the scaled products summed as the reference sums them; beta_k
 * multiplies x_k also when it is zero.  alpha and beta are both NULL or both
 * set.  b has uniform_cols (uneven: the system's) rows per system, x
 * uniform_rows. */
int gkomi_batch_spmv_f64(const gkomi_ctx* ctx, const gkomi_batch_csr* a, const double* alpha,
                         const double* b, const double* beta, double* x);
int gkomi_batch_spmv_f32(const gkomi_ctx* ctx, const gkomi_batch_csr* a, const float* alpha,
                         const float* b, const float* beta, float* x);

/* kernels::batch_bicgstab::settings (core/solver/batch_bicgstab_kernels.hpp):
 * tolerance_type 0 = absolute, 1 = relative to ||b||
 * (reference/stop/batch_criteria.hpp); preconditioner 0 = identity,
 * 1 = scalar Jacobi (reference/preconditioner/batch_scalar_jacobi.hpp). */
typedef struct {
    int32_t max_iterations;
    int32_t tolerance_type;
    double tolerance;
    int32_t preconditioner;
    int32_t reserved;
} gkomi_batch_settings;

/* Launch plan: systems in `order` (device, largest first) split into bins of
 * similar size; each bin is one persistent launch with its own workgroup
 * size, LDS residency of the matrix (1 = LDS, 0 = global, -1 = LDS if it
 * fits), workgroups per CU (0 = from the LDS footprint), for the Csr
 * formats lanes per row of the product (1, 2, 4, 8 or 16; 0 = from the
 * bin's entries per row), and where the work vectors live: `vectors`
 * 2 = registers (Csr formats, up to two rows per thread; only the vectors
 * the products gather go to LDS, beside the matrix), 1 = LDS, 0 = registers
 * where that applies and fits, LDS otherwise.  No reference counterpart: the
 * reference launches one uniform kernel per batch
 * (hip/solver/batch_bicgstab_kernels.hip.cpp). */
typedef struct {
    int32_t num_bins;
    /* 1: the whole batch in one persistent launch (uneven Csr, systems of at
     * most 512 rows, every bin's longest row known and at most 16 entries):
     * one queue of work items ordered largest first, each item one system of
     * more than 256 rows or up to eight smaller ones side by side in one
     * 256-thread workgroup; block_size, matrix_in_lds, blocks_per_cu, subwave
     * and vectors are then unused.  0: one launch per bin.  The autotuner
     * sets it. */
    int32_t single_launch;
    const int32_t* order;
    int64_t bin_begin[GKOMI_BATCH_MAX_BINS + 1];
    int64_t bin_max_rows[GKOMI_BATCH_MAX_BINS];
    int64_t bin_max_nnz[GKOMI_BATCH_MAX_BINS];
    int32_t block_size[GKOMI_BATCH_MAX_BINS];
    int32_t matrix_in_lds[GKOMI_BATCH_MAX_BINS];
    int32_t blocks_per_cu[GKOMI_BATCH_MAX_BINS];
    int32_t subwave[GKOMI_BATCH_MAX_BINS];
    int32_t vectors[GKOMI_BATCH_MAX_BINS];
    /* longest row (stored entries) of any system in the bin, 0 = unknown;
     * with register work vectors on an uneven batch, rows of up to 32
     * entries then stay in registers too (matrix_in_lds -1 or 0) */
    int64_t bin_max_row_nnz[GKOMI_BATCH_MAX_BINS];
} gkomi_batch_plan;

/* Zero-filled scratch for the work queues; left zero-filled after every call. */
size_t gkomi_batch_workspace_bytes(const gkomi_batch_csr* a);
/* LDS bytes one workgroup needs for a system of `rows` rows and `nnz` entries
 * stored as Csr; for Ell and Dense (fewer index bytes) the value is an upper
 * bound. */
size_t gkomi_batch_lds_bytes(int32_t value_bytes, int64_t rows, int64_t nnz,
                             int32_t matrix_in_lds);
/* batch_bicgstab::apply (core/solver/batch_bicgstab_kernels.hpp:177-198;
 * reference/solver/batch_bicgstab_kernels.hpp:180-354): x (initial guess in,
 * solution out), per-system iteration counts and final residual norms (the
 * batch log_data).  plan == NULL: one bin, natural order. */
int gkomi_batch_bicgstab_solve_f64(const gkomi_ctx* ctx, const gkomi_batch_csr* a,
                                   const gkomi_batch_settings* settings,
                                   const gkomi_batch_plan* plan, const double* b, double* x,
                                   int32_t* iterations, double* residual_norms,
                                   void* workspace, size_t workspace_bytes);
int gkomi_batch_bicgstab_solve_f32(const gkomi_ctx* ctx, const gkomi_batch_csr* a,
                                   const gkomi_batch_settings* settings,
                                   const gkomi_batch_plan* plan, const float* b, float* x,
                                   int32_t* iterations, float* residual_norms, void* workspace,
                                   size_t workspace_bytes);
/* No counterpart in the reference: gkomi_batch_bicgstab_solve_f64 without the
 * workspace argument, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It allocates nothing: the work-queue counters live in a 4 KiB
 * device array of the library, one per GPU, which the call clears with a
 * kernel first.  A batch whose queues would need more than 4 KiB is
 * rejected with GKOMI_NOT_SUPPORTED.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (on different streams) race. */
int gkomi_batch_bicgstab_solve_noalloc_f64(const gkomi_ctx* ctx, const gkomi_batch_csr* a,
                                           const gkomi_batch_settings* settings,
                                           const gkomi_batch_plan* plan, const double* b,
                                           double* x, int32_t* iterations,
                                           double* residual_norms);
/* Autotuner: for every bin of `plan`, times the candidate workgroup sizes
 * (64..512) and LDS / global matrix residency on that bin's systems (solving
 * into x_scratch from the initial guess x, which is not modified) and
 * writes the fastest choice into the plan.  No reference counterpart. */
int gkomi_batch_bicgstab_tune_f64(const gkomi_ctx* ctx, const gkomi_batch_csr* a,
                                  const gkomi_batch_settings* settings, gkomi_batch_plan* plan,
                                  const double* b, const double* x, double* x_scratch,
                                  int32_t* iterations, double* residual_norms, void* workspace,
                                  size_t workspace_bytes);
int gkomi_batch_bicgstab_tune_f32(const gkomi_ctx* ctx, const gkomi_batch_csr* a,
                                  const gkomi_batch_settings* settings, gkomi_batch_plan* plan,
                                  const float* b, const float* x, float* x_scratch,
                                  int32_t* iterations, float* residual_norms, void* workspace,
                                  size_t workspace_bytes);

/* -------------------------------------------------------------- batched CG */

/* batch_cg::apply (core/solver/batch_cg_kernels.hpp:165-181;
 * reference/solver/batch_cg_kernels.hpp:31-221, batch_entry_cg_impl): an
 * independent preconditioned CG per system over the same batch formats,
 * settings, plans and workspace as the batched BiCGSTAB.  Per the reference:
 * the stopping test reads the implicit norm sqrt|r.z|, iterations are
 * counted from 1, p = z when rho_old == 0.  Six work vectors per system
 * (r, z, p, Ap, x, prec) stay in LDS, with the matrix when it fits
 * (gkomi_batch_cg_lds_bytes). */
size_t gkomi_batch_cg_lds_bytes(int32_t value_bytes, int64_t rows, int64_t nnz,
                                int32_t matrix_in_lds);
int gkomi_batch_cg_solve_f64(const gkomi_ctx* ctx, const gkomi_batch_csr* a,
                             const gkomi_batch_settings* settings, const gkomi_batch_plan* plan,
                             const double* b, double* x, int32_t* iterations,
                             double* residual_norms, void* workspace, size_t workspace_bytes);
int gkomi_batch_cg_solve_f32(const gkomi_ctx* ctx, const gkomi_batch_csr* a,
                             const gkomi_batch_settings* settings, const gkomi_batch_plan* plan,
                             const float* b, float* x, int32_t* iterations,
                             float* residual_norms, void* workspace, size_t workspace_bytes);
/* Autotuner of the batched CG (as gkomi_batch_bicgstab_tune). */
int gkomi_batch_cg_tune_f64(const gkomi_ctx* ctx, const gkomi_batch_csr* a,
                            const gkomi_batch_settings* settings, gkomi_batch_plan* plan,
                            const double* b, const double* x, double* x_scratch,
                            int32_t* iterations, double* residual_norms, void* workspace,
                            size_t workspace_bytes);
int gkomi_batch_cg_tune_f32(const gkomi_ctx* ctx, const gkomi_batch_csr* a,
                            const gkomi_batch_settings* settings, gkomi_batch_plan* plan,
                            const float* b, const float* x, float* x_scratch,
                            int32_t* iterations, float* residual_norms, void* workspace,
                            size_t workspace_bytes);

/* ------------------------------------------------------ adjoint support */

/* Sampled outer product on the pattern of A: out[k] = alpha * sum_j
 * L(row(k), j) * X(col(k), j) for every stored entry k.  With L the adjoint
 * solution of A^T L = dJ/dX and X the forward solution, alpha = -1 gives
 * dJ/da_ij.  No reference counterpart (the reference has no differentiable
 * solve); `a->values` is not read. */
int gkomi_csr_sampled_outer_product_f64(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t nrhs,
                                        double alpha, const double* l, int64_t l_stride,
                                        const double* x, int64_t x_stride, double* out);
int gkomi_csr_sampled_outer_product_f32(const gkomi_ctx* ctx, const gkomi_csr* a, int64_t nrhs,
                                        float alpha, const float* l, int64_t l_stride,
                                        const float* x, int64_t x_stride, float* out);

#ifdef __cplusplus
}
#endif

#endif /* GKOMI_H_ */
