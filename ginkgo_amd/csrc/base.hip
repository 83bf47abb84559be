// Executor/memory entry points, elementwise components and dot/norm
// reductions of libgkomi.so (gfx950).
#include <cstdarg>
#include <mutex>
#include <vector>

#include "common.hpp"
#include "reduce.hpp"

namespace gkomi {

namespace {
thread_local std::string g_last_error;
}

void set_error(const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

int fail(int code, const char* what)
{
    set_error("%s", what);
    return code;
}

int hip_fail(hipError_t err, const char* what)
{
    set_error("HIP error %d (%s) in %s", static_cast<int>(err), hipGetErrorString(err), what);
    if (err == hipErrorOutOfMemory || err == hipErrorMemoryAllocation) {
        return GKOMI_ALLOCATION_ERROR;
    }
    return GKOMI_HIP_ERROR;
}

int enter(const gkomi_ctx* ctx)
{
    if (ctx == nullptr) {
        return fail(GKOMI_INVALID_ARGUMENT, "null context");
    }
    int cur = -1;
    GKOMI_CHECK_HIP(hipGetDevice(&cur));
    if (cur != ctx->device) {
        GKOMI_CHECK_HIP(hipSetDevice(ctx->device));
    }
    return GKOMI_SUCCESS;
}

int num_cus()
{
    static std::mutex mtx;
    static std::vector<int> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        return 256;
    }
    std::lock_guard<std::mutex> guard(mtx);
    if (static_cast<int>(cache.size()) <= dev) {
        cache.resize(dev + 1, 0);
    }
    if (cache[dev] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                hipSuccess ||
            cus <= 0) {
            cus = 256;
        }
        cache[dev] = cus;
    }
    return cache[dev];
}

// ------------------------------------------------------------ kernels ----

__global__ void __launch_bounds__(kBlock) fill_bytes_kernel(uint8_t* ptr, size_t bytes)
{
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    // 16-byte stores over the aligned body, bytes at the ragged ends
    const uintptr_t base = reinterpret_cast<uintptr_t>(ptr);
    const size_t head = std::min<size_t>(bytes, (16 - (base & 15)) & 15);
    const size_t body = (bytes - head) / 16;
    const size_t tail_start = head + body * 16;
    const size_t tid = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    uint4* b16 = reinterpret_cast<uint4*>(ptr + head);
    for (size_t i = tid; i < body; i += stride) {
        b16[i] = make_uint4(0, 0, 0, 0);
    }
    for (size_t i = tid; i < head; i += stride) {
        ptr[i] = 0;
    }
    for (size_t i = tail_start + tid; i < bytes; i += stride) {
        ptr[i] = 0;
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) fill_kernel(int64_t n, T value, T* ptr)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += stride) {
        ptr[i] = value;
    }
}

template <typename In, typename Out>
__global__ void __launch_bounds__(kBlock) convert_kernel(int64_t n, const In* in, Out* out)
{
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += stride) {
        out[i] = static_cast<Out>(in[i]);
    }
}

// y(i, j) = op(x(i, j), y(i, j), alpha[j or 0]) over an n x nrhs block
enum class ew { add_scaled, scale, copy };

template <ew Op, typename T>
__global__ void __launch_bounds__(kBlock)
    dense_ew_kernel(int64_t n, int64_t nrhs, const T* __restrict__ alpha, int64_t alpha_len,
                    const T* __restrict__ x, int64_t xs, T* __restrict__ y, int64_t ys)
{
    const int64_t total = n * nrhs;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
         e += stride) {
        const int64_t i = e / nrhs;
        const int64_t j = e - i * nrhs;
        if constexpr (Op == ew::copy) {
            y[i * ys + j] = x[i * xs + j];
        } else {
            const T a = alpha[alpha_len > 1 ? j : 0];
            if constexpr (Op == ew::add_scaled) {
                y[i * ys + j] += a * x[i * xs + j];
            } else {
                y[i * ys + j] *= a;
            }
        }
    }
}

// Contiguous copy in 16-byte units (the Dense::copy_from case with packed
// rows, and the stream-copy bandwidth probe): one unit per thread and a
// workgroup per 256 units, no grid-stride loop.  Copies of at least
// kCopyStreamBytes use non-temporal loads and stores, which do not keep the
// streamed data in the caches.  Measured on 1 GiB: 6.2 TB/s plain and 6.5
// TB/s non-temporal, against 5.3 TB/s for a capped grid-stride loop of four
// units per thread.  The tail below 16 bytes is copied by the caller's
// element kernel.
constexpr int64_t kCopyStreamBytes = int64_t{256} << 20;

template <bool NonTemporal>
__global__ void __launch_bounds__(kBlock)
    copy16_kernel(int64_t n16, const uint4* __restrict__ x, uint4* __restrict__ y)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n16) return;
    if constexpr (NonTemporal) {
        uint4 v;
        v.x = __builtin_nontemporal_load(&x[i].x);
        v.y = __builtin_nontemporal_load(&x[i].y);
        v.z = __builtin_nontemporal_load(&x[i].z);
        v.w = __builtin_nontemporal_load(&x[i].w);
        __builtin_nontemporal_store(v.x, &y[i].x);
        __builtin_nontemporal_store(v.y, &y[i].y);
        __builtin_nontemporal_store(v.z, &y[i].z);
        __builtin_nontemporal_store(v.w, &y[i].w);
    } else {
        y[i] = x[i];
    }
}

template <typename T>
int fill_impl(const gkomi_ctx* ctx, int64_t n, T value, T* ptr)
{
    if (int rc = enter(ctx)) return rc;
    GKOMI_REQUIRE(n >= 0, GKOMI_INVALID_ARGUMENT, "fill: negative size");
    if (n == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(ptr != nullptr, GKOMI_INVALID_ARGUMENT, "fill: null pointer");
    fill_kernel<T><<<grid_for(n, kBlock * 4), kBlock, 0, stream_of(ctx)>>>(n, value, ptr);
    GKOMI_CHECK_LAUNCH("fill_kernel");
    return GKOMI_SUCCESS;
}

template <typename In, typename Out>
int convert_impl(const gkomi_ctx* ctx, int64_t n, const In* in, Out* out)
{
    if (int rc = enter(ctx)) return rc;
    GKOMI_REQUIRE(n >= 0, GKOMI_INVALID_ARGUMENT, "convert: negative size");
    if (n == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(in && out, GKOMI_INVALID_ARGUMENT, "convert: null pointer");
    convert_kernel<In, Out><<<grid_for(n, kBlock * 4), kBlock, 0, stream_of(ctx)>>>(n, in, out);
    GKOMI_CHECK_LAUNCH("convert_kernel");
    return GKOMI_SUCCESS;
}

template <ew Op, typename T>
int dense_ew_impl(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const T* alpha,
                  int64_t alpha_len, const T* x, int64_t xs, T* y, int64_t ys)
{
    if (int rc = enter(ctx)) return rc;
    GKOMI_REQUIRE(n >= 0 && nrhs >= 0, GKOMI_INVALID_ARGUMENT, "dense op: negative size");
    if (n == 0 || nrhs == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(ys >= nrhs && (Op == ew::scale || xs >= nrhs), GKOMI_INVALID_ARGUMENT,
                  "dense op: stride smaller than the number of columns");
    GKOMI_REQUIRE(y != nullptr && (Op == ew::scale || x != nullptr), GKOMI_INVALID_ARGUMENT,
                  "dense op: null pointer");
    GKOMI_REQUIRE(Op == ew::copy || (alpha != nullptr && (alpha_len == 1 || alpha_len == nrhs)),
                  GKOMI_DIMENSION_MISMATCH, "dense op: alpha must have 1 or nrhs entries");
    if constexpr (Op == ew::copy) {
        const int64_t bytes = n * nrhs * static_cast<int64_t>(sizeof(T));
        const bool packed = xs == nrhs && ys == nrhs;
        const bool aligned = (reinterpret_cast<uintptr_t>(x) % 16 == 0) &&
                             (reinterpret_cast<uintptr_t>(y) % 16 == 0);
        if (packed && aligned && bytes >= 16) {
            const int64_t n16 = bytes / 16;
            const int64_t blocks = ceildiv(n16, kBlock);
            GKOMI_REQUIRE(blocks < (int64_t{1} << 31), GKOMI_INVALID_ARGUMENT,
                          "copy: more than 2^31 workgroups");
            const auto* src = reinterpret_cast<const uint4*>(x);
            auto* dst = reinterpret_cast<uint4*>(y);
            if (bytes >= kCopyStreamBytes) {
                copy16_kernel<true><<<static_cast<unsigned>(blocks), kBlock, 0, stream_of(ctx)>>>(
                    n16, src, dst);
            } else {
                copy16_kernel<false><<<static_cast<unsigned>(blocks), kBlock, 0, stream_of(ctx)>>>(
                    n16, src, dst);
            }
            GKOMI_CHECK_LAUNCH("copy16_kernel");
            const int64_t done = n16 * 16 / static_cast<int64_t>(sizeof(T));
            if (done == n * nrhs) return GKOMI_SUCCESS;
            // remaining elements (fewer than 16 bytes) as one packed row
            return dense_ew_impl<Op, T>(ctx, n * nrhs - done, 1, nullptr, 0, x + done, 1,
                                        y + done, 1);
        }
    }
    dense_ew_kernel<Op, T><<<grid_for(n * nrhs, kBlock * 4), kBlock, 0, stream_of(ctx)>>>(
        n, nrhs, alpha, alpha_len, x, xs, y, ys);
    GKOMI_CHECK_LAUNCH("dense_ew_kernel");
    return GKOMI_SUCCESS;
}

template <typename T, bool Norm>
int reduce_impl(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const T* x, int64_t xs,
                const T* y, int64_t ys, T* result, void* ws, size_t ws_bytes)
{
    if (int rc = enter(ctx)) return rc;
    GKOMI_REQUIRE(n >= 0 && nrhs >= 0, GKOMI_INVALID_ARGUMENT, "reduction: negative size");
    if (nrhs == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(result != nullptr, GKOMI_INVALID_ARGUMENT, "reduction: null result");
    GKOMI_REQUIRE(n == 0 || (x != nullptr && (Norm || y != nullptr)), GKOMI_INVALID_ARGUMENT,
                  "reduction: null operand");
    GKOMI_REQUIRE(xs >= nrhs && (Norm || ys >= nrhs), GKOMI_INVALID_ARGUMENT,
                  "reduction: stride smaller than the number of columns");
    GKOMI_REQUIRE(ws != nullptr && ws_bytes >= gkomi_reduction_workspace_bytes(n, nrhs),
                  GKOMI_INVALID_ARGUMENT, "reduction: workspace too small");
    return launch_column_reduction<T, Norm>(
        stream_of(ctx), n, nrhs, x, xs, y, ys, result,
        reinterpret_cast<double*>(static_cast<char*>(ws) + kTicketBytes));
}

}  // namespace gkomi

using namespace gkomi;

// A captured and instantiated HIP graph (the opaque gkomi_graph of the C ABI).
struct gkomi_graph {
    int32_t device;
    hipGraph_t graph;
    hipGraphExec_t exec;
};

extern "C" {

const char* gkomi_last_error(void) { return g_last_error.c_str(); }

const char* gkomi_version(void) { return "ginkgo_amd-gkomi 0.1 (gfx950)"; }

int gkomi_get_device_count(int32_t* count)
{
    GKOMI_REQUIRE(count != nullptr, GKOMI_INVALID_ARGUMENT, "null count");
    int n = 0;
    GKOMI_CHECK_HIP(hipGetDeviceCount(&n));
    *count = n;
    return GKOMI_SUCCESS;
}

int gkomi_get_device_info(int32_t device, gkomi_device_info* info)
{
    GKOMI_REQUIRE(info != nullptr, GKOMI_INVALID_ARGUMENT, "null info");
    hipDeviceProp_t prop;
    GKOMI_CHECK_HIP(hipGetDeviceProperties(&prop, device));
    std::memset(info, 0, sizeof(*info));
    std::snprintf(info->name, sizeof(info->name), "%s", prop.name);
    std::snprintf(info->arch, sizeof(info->arch), "%s", prop.gcnArchName);
    info->num_cus = prop.multiProcessorCount;
    info->warp_size = prop.warpSize;
    info->max_lds_per_block = static_cast<int64_t>(prop.sharedMemPerBlock);
    info->total_global_mem = static_cast<int64_t>(prop.totalGlobalMem);
    info->clock_khz = prop.clockRate;
    info->memory_clock_khz = prop.memoryClockRate;
    info->memory_bus_width = prop.memoryBusWidth;
    return GKOMI_SUCCESS;
}

int gkomi_stream_create(int32_t device, void** stream)
{
    GKOMI_REQUIRE(stream != nullptr, GKOMI_INVALID_ARGUMENT, "null stream out");
    GKOMI_CHECK_HIP(hipSetDevice(device));
    hipStream_t s = nullptr;
    // default priority on purpose: every stream of this library has the same one
    GKOMI_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return GKOMI_SUCCESS;
}

int gkomi_stream_destroy(int32_t device, void* stream)
{
    if (stream == nullptr) return GKOMI_SUCCESS;
    GKOMI_CHECK_HIP(hipSetDevice(device));
    GKOMI_CHECK_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
    return GKOMI_SUCCESS;
}

int gkomi_synchronize(const gkomi_ctx* ctx)
{
    if (int rc = enter(ctx)) return rc;
    GKOMI_CHECK_HIP(hipStreamSynchronize(stream_of(ctx)));
    return GKOMI_SUCCESS;
}

int gkomi_malloc(const gkomi_ctx* ctx, size_t bytes, void** ptr)
{
    if (int rc = enter(ctx)) return rc;
    GKOMI_REQUIRE(ptr != nullptr, GKOMI_INVALID_ARGUMENT, "null pointer out");
    *ptr = nullptr;
    if (bytes == 0) return GKOMI_SUCCESS;
    void* p = nullptr;
    hipError_t err = hipMalloc(&p, bytes);
    if (err != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(err));
        return GKOMI_ALLOCATION_ERROR;
    }
    *ptr = p;
    return GKOMI_SUCCESS;
}

int gkomi_free(const gkomi_ctx* ctx, void* ptr)
{
    if (ptr == nullptr) return GKOMI_SUCCESS;
    if (int rc = enter(ctx)) return rc;
    GKOMI_CHECK_HIP(hipFree(ptr));
    return GKOMI_SUCCESS;
}

int gkomi_host_malloc(size_t bytes, void** ptr)
{
    GKOMI_REQUIRE(ptr != nullptr, GKOMI_INVALID_ARGUMENT, "null pointer out");
    *ptr = nullptr;
    if (bytes == 0) return GKOMI_SUCCESS;
    void* p = nullptr;
    hipError_t err = hipHostMalloc(&p, bytes, hipHostMallocDefault);
    if (err != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipHostMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(err));
        return GKOMI_ALLOCATION_ERROR;
    }
    *ptr = p;
    return GKOMI_SUCCESS;
}

int gkomi_host_free(void* ptr)
{
    if (ptr == nullptr) return GKOMI_SUCCESS;
    GKOMI_CHECK_HIP(hipHostFree(ptr));
    return GKOMI_SUCCESS;
}

int gkomi_memcpy_async(const gkomi_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (int rc = enter(ctx)) return rc;
    if (bytes == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(dst && src, GKOMI_INVALID_ARGUMENT, "memcpy: null pointer");
    GKOMI_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, stream_of(ctx)));
    return GKOMI_SUCCESS;
}

int gkomi_memcpy_to_host(const gkomi_ctx* ctx, void* dst, const void* src, size_t bytes)
{
    if (int rc = gkomi_memcpy_async(ctx, dst, src, bytes)) return rc;
    GKOMI_CHECK_HIP(hipStreamSynchronize(stream_of(ctx)));
    return GKOMI_SUCCESS;
}

int gkomi_event_create(int32_t device, void** event)
{
    GKOMI_REQUIRE(event != nullptr, GKOMI_INVALID_ARGUMENT, "null event out");
    GKOMI_CHECK_HIP(hipSetDevice(device));
    hipEvent_t e = nullptr;
    GKOMI_CHECK_HIP(hipEventCreate(&e));
    *event = e;
    return GKOMI_SUCCESS;
}

int gkomi_event_destroy(int32_t device, void* event)
{
    if (event == nullptr) return GKOMI_SUCCESS;
    GKOMI_CHECK_HIP(hipSetDevice(device));
    GKOMI_CHECK_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return GKOMI_SUCCESS;
}

int gkomi_event_record(const gkomi_ctx* ctx, void* event)
{
    if (int rc = enter(ctx)) return rc;
    GKOMI_CHECK_HIP(hipEventRecord(static_cast<hipEvent_t>(event), stream_of(ctx)));
    return GKOMI_SUCCESS;
}

int gkomi_event_synchronize(void* event)
{
    GKOMI_CHECK_HIP(hipEventSynchronize(static_cast<hipEvent_t>(event)));
    return GKOMI_SUCCESS;
}

int gkomi_event_elapsed_ms(void* start, void* stop, float* ms)
{
    GKOMI_REQUIRE(ms != nullptr, GKOMI_INVALID_ARGUMENT, "null output");
    GKOMI_CHECK_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                                        static_cast<hipEvent_t>(stop)));
    return GKOMI_SUCCESS;
}

int gkomi_graph_begin(const gkomi_ctx* ctx, gkomi_ctx* capture)
{
    if (int rc = enter(ctx)) return rc;
    GKOMI_REQUIRE(capture != nullptr, GKOMI_INVALID_ARGUMENT, "graph: null capture context");
    hipStream_t s = nullptr;
    // default priority, as every stream of this library
    GKOMI_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    const hipError_t err = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
    if (err != hipSuccess) {
        (void)hipStreamDestroy(s);
        set_error("hipStreamBeginCapture: %s", hipGetErrorString(err));
        return GKOMI_HIP_ERROR;
    }
    capture->device = ctx->device;
    capture->reserved = 0;
    capture->stream = s;
    return GKOMI_SUCCESS;
}

int gkomi_graph_end(gkomi_ctx* capture, gkomi_graph** graph)
{
    GKOMI_REQUIRE(capture != nullptr && capture->stream != nullptr && graph != nullptr,
                  GKOMI_INVALID_ARGUMENT, "graph: no capture in progress");
    if (int rc = enter(capture)) return rc;
    hipStream_t s = static_cast<hipStream_t>(capture->stream);
    hipGraph_t g = nullptr;
    const hipError_t err = hipStreamEndCapture(s, &g);
    (void)hipStreamDestroy(s);
    capture->stream = nullptr;
    if (err != hipSuccess || g == nullptr) {
        set_error("hipStreamEndCapture: %s", hipGetErrorString(err));
        return GKOMI_HIP_ERROR;
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ierr = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (ierr != hipSuccess) {
        (void)hipGraphDestroy(g);
        set_error("hipGraphInstantiate: %s", hipGetErrorString(ierr));
        return GKOMI_HIP_ERROR;
    }
    *graph = new gkomi_graph{capture->device, g, ex};
    return GKOMI_SUCCESS;
}

int gkomi_graph_launch(const gkomi_ctx* ctx, const gkomi_graph* graph)
{
    if (int rc = enter(ctx)) return rc;
    GKOMI_REQUIRE(graph != nullptr, GKOMI_INVALID_ARGUMENT, "graph: null graph");
    GKOMI_REQUIRE(ctx->device == graph->device, GKOMI_INVALID_ARGUMENT,
                  "graph: launched on another device than it was captured on");
    GKOMI_CHECK_HIP(hipGraphLaunch(graph->exec, stream_of(ctx)));
    return GKOMI_SUCCESS;
}

int gkomi_graph_destroy(gkomi_graph* graph)
{
    if (graph == nullptr) return GKOMI_SUCCESS;
    GKOMI_CHECK_HIP(hipSetDevice(graph->device));
    const hipError_t e1 = hipGraphExecDestroy(graph->exec);
    const hipError_t e2 = hipGraphDestroy(graph->graph);
    delete graph;
    GKOMI_CHECK_HIP(e1);
    GKOMI_CHECK_HIP(e2);
    return GKOMI_SUCCESS;
}

int gkomi_fill_zero(const gkomi_ctx* ctx, void* ptr, size_t bytes)
{
    if (int rc = enter(ctx)) return rc;
    if (bytes == 0) return GKOMI_SUCCESS;
    GKOMI_REQUIRE(ptr != nullptr, GKOMI_INVALID_ARGUMENT, "fill_zero: null pointer");
    fill_bytes_kernel<<<grid_for(static_cast<int64_t>(bytes / 16 + 1), kBlock * 4), kBlock, 0,
                        stream_of(ctx)>>>(static_cast<uint8_t*>(ptr), bytes);
    GKOMI_CHECK_LAUNCH("fill_bytes_kernel");
    return GKOMI_SUCCESS;
}

int gkomi_fill_f64(const gkomi_ctx* ctx, int64_t n, double value, double* ptr)
{
    return fill_impl(ctx, n, value, ptr);
}

int gkomi_fill_f32(const gkomi_ctx* ctx, int64_t n, float value, float* ptr)
{
    return fill_impl(ctx, n, value, ptr);
}

int gkomi_convert_f64_to_f32(const gkomi_ctx* ctx, int64_t n, const double* in, float* out)
{
    return convert_impl(ctx, n, in, out);
}

int gkomi_convert_f32_to_f64(const gkomi_ctx* ctx, int64_t n, const float* in, double* out)
{
    return convert_impl(ctx, n, in, out);
}

size_t gkomi_reduction_workspace_bytes(int64_t n, int64_t nrhs)
{
    (void)n;
    return kTicketBytes + sizeof(double) * static_cast<size_t>(kRedMaxBlocks) *
                              static_cast<size_t>(nrhs > 0 ? nrhs : 1);
}

int gkomi_dense_dot_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const double* x,
                        int64_t xs, const double* y, int64_t ys, double* result, void* ws,
                        size_t wsb)
{
    return reduce_impl<double, false>(ctx, n, nrhs, x, xs, y, ys, result, ws, wsb);
}

int gkomi_dense_dot_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const float* x,
                        int64_t xs, const float* y, int64_t ys, float* result, void* ws,
                        size_t wsb)
{
    return reduce_impl<float, false>(ctx, n, nrhs, x, xs, y, ys, result, ws, wsb);
}

int gkomi_dense_norm2_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const double* x,
                          int64_t xs, double* result, void* ws, size_t wsb)
{
    return reduce_impl<double, true>(ctx, n, nrhs, x, xs, nullptr, xs, result, ws, wsb);
}

int gkomi_dense_norm2_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const float* x,
                          int64_t xs, float* result, void* ws, size_t wsb)
{
    return reduce_impl<float, true>(ctx, n, nrhs, x, xs, nullptr, xs, result, ws, wsb);
}

int gkomi_dense_add_scaled_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs,
                               const double* alpha, int64_t alpha_len, const double* x,
                               int64_t xs, double* y, int64_t ys)
{
    return dense_ew_impl<ew::add_scaled>(ctx, n, nrhs, alpha, alpha_len, x, xs, y, ys);
}

int gkomi_dense_add_scaled_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs,
                               const float* alpha, int64_t alpha_len, const float* x,
                               int64_t xs, float* y, int64_t ys)
{
    return dense_ew_impl<ew::add_scaled>(ctx, n, nrhs, alpha, alpha_len, x, xs, y, ys);
}

int gkomi_dense_scale_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const double* alpha,
                          int64_t alpha_len, double* y, int64_t ys)
{
    return dense_ew_impl<ew::scale, double>(ctx, n, nrhs, alpha, alpha_len, nullptr, nrhs, y,
                                            ys);
}

int gkomi_dense_scale_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const float* alpha,
                          int64_t alpha_len, float* y, int64_t ys)
{
    return dense_ew_impl<ew::scale, float>(ctx, n, nrhs, alpha, alpha_len, nullptr, nrhs, y,
                                           ys);
}

int gkomi_dense_copy_f64(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const double* x,
                         int64_t xs, double* y, int64_t ys)
{
    return dense_ew_impl<ew::copy, double>(ctx, n, nrhs, nullptr, 0, x, xs, y, ys);
}

int gkomi_dense_copy_f32(const gkomi_ctx* ctx, int64_t n, int64_t nrhs, const float* x,
                         int64_t xs, float* y, int64_t ys)
{
    return dense_ew_impl<ew::copy, float>(ctx, n, nrhs, nullptr, 0, x, xs, y, ys);
}

}  // extern "C"
