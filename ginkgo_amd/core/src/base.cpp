// Executors, the caching device allocator and error mapping of the gko::
// host library.  The device side is reached only through gkomi.h.
#include <cstdlib>
#include <sstream>

#include "ginkgo_amd/base.hpp"
#include "ginkgo_amd/log.hpp"

namespace gko {

void throw_on_status(int status, const char* file, int line, const char* call)
{
    if (status == GKOMI_SUCCESS) return;
    std::string msg = std::string(call) + ": " + gkomi_last_error();
    switch (status) {
    case GKOMI_NOT_SUPPORTED:
        throw NotSupported(file, line, msg);
    case GKOMI_DIMENSION_MISMATCH:
        throw DimensionMismatch(file, line, msg);
    case GKOMI_ALLOCATION_ERROR:
        throw AllocationError(file, line, msg);
    case GKOMI_HIP_ERROR:
        throw HipError(file, line, msg);
    case GKOMI_INVALID_ARGUMENT:
    default:
        throw Error(file, line, msg);
    }
}

// -------------------------------------------------------- Executor logging ----

void Executor::add_logger(std::shared_ptr<const log::Logger> logger) const
{
    loggers_.push_back(std::move(logger));
}

void Executor::remove_logger(const log::Logger* logger) const
{
    for (auto it = loggers_.begin(); it != loggers_.end(); ++it) {
        if (it->get() == logger) {
            loggers_.erase(it);
            return;
        }
    }
    throw Error(__FILE__, __LINE__, "remove_logger: the logger is not attached");
}

void Executor::clear_loggers() const { loggers_.clear(); }

void* Executor::logged_alloc(size_type bytes) const
{
    for (auto& l : loggers_) l->on_allocation_started(this, bytes);
    void* p = raw_alloc(bytes);
    const auto loc = reinterpret_cast<uintptr>(p);
    for (auto& l : loggers_) l->on_allocation_completed(this, bytes, loc);
    return p;
}

void Executor::logged_free(void* ptr) const noexcept
{
    const auto loc = reinterpret_cast<uintptr>(ptr);
    // a logger that throws must not turn a free into a leak or a terminate
    try {
        for (auto& l : loggers_) l->on_free_started(this, loc);
    } catch (...) {
    }
    raw_free(ptr);
    try {
        for (auto& l : loggers_) l->on_free_completed(this, loc);
    } catch (...) {
    }
}

// the copy is logged to the destination's loggers and, when it is another
// executor, to the source's (executor.hpp:745-781)
void Executor::logged_copy_from(const Executor* src, size_type bytes, const void* src_ptr,
                                void* dst_ptr) const
{
    const auto from = reinterpret_cast<uintptr>(src_ptr);
    const auto to = reinterpret_cast<uintptr>(dst_ptr);
    for (auto& l : loggers_) l->on_copy_started(src, this, from, to, bytes);
    if (src != this) {
        for (auto& l : src->loggers_) l->on_copy_started(src, this, from, to, bytes);
    }
    raw_copy_from(src, bytes, src_ptr, dst_ptr);
    for (auto& l : loggers_) l->on_copy_completed(src, this, from, to, bytes);
    if (src != this) {
        for (auto& l : src->loggers_) l->on_copy_completed(src, this, from, to, bytes);
    }
}

// ----------------------------------------------------------- HostExecutor ----

std::shared_ptr<HostExecutor> HostExecutor::create()
{
    return std::shared_ptr<HostExecutor>(new HostExecutor());
}

void* HostExecutor::raw_alloc(size_type bytes) const
{
    if (bytes == 0) return nullptr;
    void* p = std::malloc(bytes);
    if (!p) throw AllocationError(__FILE__, __LINE__, "host allocation failed");
    return p;
}

void HostExecutor::raw_free(void* ptr) const noexcept { std::free(ptr); }

void HostExecutor::raw_copy_from(const Executor* src, size_type bytes, const void* src_ptr,
                                 void* dst_ptr) const
{
    if (bytes == 0) return;
    if (src->is_host()) {
        std::memcpy(dst_ptr, src_ptr, bytes);
        return;
    }
    // device -> host: blocking copy on the device executor's stream
    GKO_CALL(gkomi_memcpy_to_host(src->get_ctx(), dst_ptr, src_ptr, bytes));
}

std::shared_ptr<const HostExecutor> HostExecutor::get_master() const
{
    return std::static_pointer_cast<const HostExecutor>(shared_from_this());
}

// -------------------------------------------------------- Mi355xExecutor ----

int Mi355xExecutor::get_num_devices()
{
    int32_t n = 0;
    if (gkomi_get_device_count(&n) != GKOMI_SUCCESS) return 0;
    return n;
}

std::shared_ptr<Mi355xExecutor> Mi355xExecutor::create(int device_id, void* stream, bool caching)
{
    return std::shared_ptr<Mi355xExecutor>(new Mi355xExecutor(device_id, stream, caching, true));
}

std::shared_ptr<Mi355xExecutor> Mi355xExecutor::create_on_stream(int device_id, void* stream,
                                                                 bool caching)
{
    return std::shared_ptr<Mi355xExecutor>(new Mi355xExecutor(device_id, stream, caching, false));
}

Mi355xExecutor::Mi355xExecutor(int device_id, void* stream, bool caching, bool own_if_null)
    : caching_{caching}, master_{HostExecutor::create()}
{
    ctx_.device = device_id;
    GKO_CALL(gkomi_get_device_info(device_id, &info_));
    if (stream == nullptr && own_if_null) {
        GKO_CALL(gkomi_stream_create(device_id, &ctx_.stream));
        own_stream_ = true;
    } else {
        ctx_.stream = stream;
    }
}

Mi355xExecutor::~Mi355xExecutor()
{
    gkomi_synchronize(&ctx_);
    release_pool();
    if (own_stream_) gkomi_stream_destroy(ctx_.device, ctx_.stream);
}

namespace {
size_type bucket(size_type bytes)
{
    // 256 B granularity up to 1 MiB, then powers of two: bounded waste
    if (bytes <= (size_type{1} << 20)) return (bytes + 255) / 256 * 256;
    size_type b = size_type{1} << 20;
    while (b < bytes) b <<= 1;
    return b;
}
}  // namespace

void* Mi355xExecutor::raw_alloc(size_type bytes) const
{
    if (bytes == 0) return nullptr;
    const size_type b = caching_ ? bucket(bytes) : bytes;
    if (caching_) {
        std::lock_guard<std::mutex> g(pool_mutex_);
        auto it = free_blocks_.find(b);
        if (it != free_blocks_.end()) {
            void* p = it->second;
            free_blocks_.erase(it);
            pooled_bytes_ -= b;
            live_blocks_[p] = b;
            return p;
        }
    }
    void* p = nullptr;
    int rc = gkomi_malloc(&ctx_, b, &p);
    if (rc == GKOMI_ALLOCATION_ERROR && caching_) {
        // return cached blocks to the driver and retry once
        release_pool();
        rc = gkomi_malloc(&ctx_, b, &p);
    }
    if (rc != GKOMI_SUCCESS) {
        throw AllocationError(__FILE__, __LINE__,
                              "device allocation of " + std::to_string(b) + " bytes failed: " +
                                  gkomi_last_error());
    }
    if (caching_) {
        std::lock_guard<std::mutex> g(pool_mutex_);
        live_blocks_[p] = b;
    }
    return p;
}

void Mi355xExecutor::raw_free(void* ptr) const noexcept
{
    if (!ptr) return;
    if (!caching_) {
        gkomi_free(&ctx_, ptr);
        return;
    }
    // blocks are reused in stream order: work already enqueued on this
    // executor's stream that reads the block completes before any later
    // kernel on the same stream can overwrite it
    std::lock_guard<std::mutex> g(pool_mutex_);
    auto it = live_blocks_.find(ptr);
    if (it == live_blocks_.end()) {
        gkomi_free(&ctx_, ptr);
        return;
    }
    free_blocks_.emplace(it->second, ptr);
    pooled_bytes_ += it->second;
    live_blocks_.erase(it);
}

void Mi355xExecutor::raw_copy_from(const Executor* src, size_type bytes, const void* src_ptr,
                                   void* dst_ptr) const
{
    if (bytes == 0) return;
    if (src->is_host()) {
        // host -> device: stream-ordered, then wait so the host buffer may go
        GKO_CALL(gkomi_memcpy_async(&ctx_, dst_ptr, src_ptr, bytes));
        GKO_CALL(gkomi_synchronize(&ctx_));
    } else if (src == this) {
        GKO_CALL(gkomi_memcpy_async(&ctx_, dst_ptr, src_ptr, bytes));
    } else {
        // another device executor: order after its pending work
        src->synchronize();
        GKO_CALL(gkomi_memcpy_async(&ctx_, dst_ptr, src_ptr, bytes));
    }
}

void Mi355xExecutor::synchronize() const { GKO_CALL(gkomi_synchronize(&ctx_)); }

size_type Mi355xExecutor::get_pool_bytes() const
{
    std::lock_guard<std::mutex> g(pool_mutex_);
    return pooled_bytes_;
}

void Mi355xExecutor::release_pool() const
{
    std::multimap<size_type, void*> blocks;
    {
        std::lock_guard<std::mutex> g(pool_mutex_);
        blocks.swap(free_blocks_);
        pooled_bytes_ = 0;
    }
    if (blocks.empty()) return;
    gkomi_synchronize(&ctx_);
    for (auto& kv : blocks) gkomi_free(&ctx_, kv.second);
}

// ------------------------------------------------------------ array::fill ----

template <typename T>
void array<T>::fill(T value)
{
    if (size_ == 0) return;
    if (exec_->is_host()) {
        std::fill(data_, data_ + size_, value);
        return;
    }
    if constexpr (std::is_same<T, double>::value) {
        GKO_CALL(gkomi_fill_f64(exec_->get_ctx(), static_cast<int64_t>(size_), value, data_));
    } else if constexpr (std::is_same<T, float>::value) {
        GKO_CALL(gkomi_fill_f32(exec_->get_ctx(), static_cast<int64_t>(size_), value, data_));
    } else {
        if (value == T{}) {
            GKO_CALL(gkomi_fill_zero(exec_->get_ctx(), data_, size_ * sizeof(T)));
        } else {
            std::vector<T> host(size_, value);
            exec_->copy_from(exec_->get_master().get(), size_, host.data(), data_);
        }
    }
}

template class array<double>;
template class array<float>;
template class array<int16>;
template class array<int32>;
template class array<int64>;
template class array<uint8>;
template class array<char>;
template class array<stopping_status>;

}  // namespace gko
