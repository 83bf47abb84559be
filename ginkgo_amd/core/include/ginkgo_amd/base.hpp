// Core types of the gko:: host library: dimensions, exceptions, executors,
// arrays and the stopping status.  The public surface keeps the names and
// meaning of include/ginkgo/core/base/{types,exception,executor,array}.hpp;
// device work is delegated to libgkomi.so through its C ABI (gkomi.h).
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "gkomi.h"

namespace gko {

using size_type = std::size_t;
using int16 = std::int16_t;
using int32 = std::int32_t;
using int64 = std::int64_t;
using uint8 = std::uint8_t;
using uintptr = std::uintptr_t;

// include/ginkgo/core/base/dim.hpp
template <size_type Dim = 2>
struct dim {
    size_type rows{};
    size_type cols{};
    constexpr dim() = default;
    constexpr explicit dim(size_type n) : rows{n}, cols{n} {}
    constexpr dim(size_type r, size_type c) : rows{r}, cols{c} {}
    constexpr size_type operator[](size_type i) const { return i == 0 ? rows : cols; }
    friend bool operator==(const dim& a, const dim& b)
    {
        return a.rows == b.rows && a.cols == b.cols;
    }
    friend bool operator!=(const dim& a, const dim& b) { return !(a == b); }
};

// ------------------------------------------------------------ exceptions ----
// include/ginkgo/core/base/exception.hpp

class Error : public std::exception {
public:
    Error(const std::string& file, int line, const std::string& what)
        : what_{file + ":" + std::to_string(line) + ": " + what}
    {}
    const char* what() const noexcept override { return what_.c_str(); }

private:
    std::string what_;
};

class NotImplemented : public Error {
public:
    using Error::Error;
};
class NotSupported : public Error {
public:
    using Error::Error;
};
class DimensionMismatch : public Error {
public:
    using Error::Error;
};
class BadDimension : public Error {
public:
    using Error::Error;
};
class AllocationError : public Error {
public:
    using Error::Error;
};
class HipError : public Error {
public:
    using Error::Error;
};
class InvalidStateError : public Error {
public:
    using Error::Error;
};

// Converts a gkomi status into the reference's exception types.
void throw_on_status(int status, const char* file, int line, const char* call);

#define GKO_CALL(expr) ::gko::throw_on_status((expr), __FILE__, __LINE__, #expr)
#define GKO_NOT_SUPPORTED_MSG(msg) throw ::gko::NotSupported(__FILE__, __LINE__, msg)
#define GKO_DIMENSION_MISMATCH_MSG(msg) throw ::gko::DimensionMismatch(__FILE__, __LINE__, msg)

// ------------------------------------------------------- stopping_status ----
// include/ginkgo/core/stop/stopping_status.hpp:21-118 (same bit layout)

class stopping_status {
public:
    bool has_stopped() const noexcept { return get_id(); }
    bool has_converged() const noexcept { return data_ & converged_mask_; }
    bool is_finalized() const noexcept { return data_ & finalized_mask_; }
    uint8 get_id() const noexcept { return data_ & id_mask_; }
    void reset() noexcept { data_ = 0; }
    void stop(uint8 id, bool set_finalized = true) noexcept
    {
        if (!has_stopped()) {
            data_ |= (id & id_mask_);
            if (set_finalized) data_ |= finalized_mask_;
        }
    }
    void converge(uint8 id, bool set_finalized = true) noexcept
    {
        if (!has_stopped()) {
            data_ |= converged_mask_ | (id & id_mask_);
            if (set_finalized) data_ |= finalized_mask_;
        }
    }
    void finalize() noexcept
    {
        if (has_stopped()) data_ |= finalized_mask_;
    }
    uint8 raw() const noexcept { return data_; }
    friend bool operator==(const stopping_status& a, const stopping_status& b)
    {
        return a.data_ == b.data_;
    }

private:
    static constexpr uint8 converged_mask_ = uint8{1} << 7;
    static constexpr uint8 finalized_mask_ = uint8{1} << 6;
    static constexpr uint8 id_mask_ = (uint8{1} << 6) - uint8{1};
    uint8 data_{};
};
static_assert(sizeof(stopping_status) == 1, "stopping_status must be one byte");

// ------------------------------------------------------------- executors ----
// include/ginkgo/core/base/executor.hpp:616-1988.  Two executors exist here:
// HostExecutor (host memory only; the master of every device executor) and
// Mi355xExecutor (one gfx950 device and one stream).  LinOps compute on the
// Mi355xExecutor only; a LinOp asked to compute on the host throws
// NotSupported -- there is no CPU fallback in this library.

class HostExecutor;

namespace log {
class Logger;
}

class Executor : public std::enable_shared_from_this<Executor> {
public:
    virtual ~Executor() = default;
    // the allocation itself, without logging (Executor::raw_alloc /
    // raw_free, executor.hpp:705-740)
    virtual void* raw_alloc(size_type bytes) const = 0;
    virtual void raw_free(void* ptr) const noexcept = 0;
    // Executor::alloc / free: raw_alloc / raw_free between the allocation_*
    // and free_* events of this executor's loggers
    template <typename T>
    T* alloc(size_type num) const
    {
        const size_type bytes = num * sizeof(T);
        if (loggers_.empty()) return static_cast<T*>(raw_alloc(bytes));
        return static_cast<T*>(logged_alloc(bytes));
    }
    void free(void* ptr) const noexcept
    {
        if (loggers_.empty()) {
            raw_free(ptr);
        } else {
            logged_free(ptr);
        }
    }
    // the copy into this executor's memory, without logging
    // (executor.hpp:745-815)
    virtual void raw_copy_from(const Executor* src, size_type bytes, const void* src_ptr,
                               void* dst_ptr) const = 0;
    // Executor::copy_from: raw_copy_from between the copy_* events
    template <typename T>
    void copy_from(const Executor* src, size_type num, const T* src_ptr, T* dst_ptr) const
    {
        const size_type bytes = num * sizeof(T);
        if (loggers_.empty() && src->loggers_.empty()) {
            raw_copy_from(src, bytes, src_ptr, dst_ptr);
        } else {
            logged_copy_from(src, bytes, src_ptr, dst_ptr);
        }
    }
    template <typename T>
    T copy_val_to_host(const T* ptr) const
    {
        T out{};
        get_master()->copy_from(this, 1, ptr, &out);
        return out;
    }
    virtual void synchronize() const = 0;
    virtual std::shared_ptr<const HostExecutor> get_master() const = 0;
    virtual bool is_host() const = 0;
    // the gkomi context (device + stream) used for every kernel
    virtual const gkomi_ctx* get_ctx() const { return nullptr; }
    bool memory_accessible(const std::shared_ptr<const Executor>& other) const
    {
        return other.get() == this || (is_host() && other->is_host());
    }

    // log::EnableLogging of the executor (logger.hpp:786-817) and the
    // propagation switch (executor.hpp set_log_propagation /
    // should_propagate_log)
    void add_logger(std::shared_ptr<const log::Logger> logger) const;
    void remove_logger(const log::Logger* logger) const;
    void clear_loggers() const;
    const std::vector<std::shared_ptr<const log::Logger>>& get_loggers() const
    {
        return loggers_;
    }
    void set_log_propagation(bool propagate) const { propagate_log_ = propagate; }
    bool should_propagate_log() const { return propagate_log_ && !loggers_.empty(); }

private:
    void* logged_alloc(size_type bytes) const;
    void logged_free(void* ptr) const noexcept;
    void logged_copy_from(const Executor* src, size_type bytes, const void* src_ptr,
                          void* dst_ptr) const;

    mutable std::vector<std::shared_ptr<const log::Logger>> loggers_;
    mutable bool propagate_log_{true};
};

class HostExecutor : public Executor {
public:
    static std::shared_ptr<HostExecutor> create();
    void* raw_alloc(size_type bytes) const override;
    void raw_free(void* ptr) const noexcept override;
    void raw_copy_from(const Executor* src, size_type bytes, const void* src_ptr,
                       void* dst_ptr) const override;
    void synchronize() const override {}
    std::shared_ptr<const HostExecutor> get_master() const override;
    bool is_host() const override { return true; }

protected:
    HostExecutor() = default;
};

// Mi355xExecutor: the drop-in device executor (the role of HipExecutor,
// executor.hpp:1786-1988 / hip/base/executor.hip.cpp).  Allocation goes
// through a size-bucketed caching pool over gkomi_malloc so that solver
// workspaces do not hit the driver on every apply.
class Mi355xExecutor : public Executor {
public:
    // stream == nullptr: the executor creates (and owns) a default-priority
    // stream; otherwise it orders work on the given hipStream_t.
    static std::shared_ptr<Mi355xExecutor> create(int device_id = 0, void* stream = nullptr,
                                                  bool caching = true);
    // orders work on exactly `stream`, nullptr meaning the device's null
    // stream (the stream a framework such as PyTorch runs on by default)
    static std::shared_ptr<Mi355xExecutor> create_on_stream(int device_id, void* stream,
                                                            bool caching = true);
    ~Mi355xExecutor() override;
    void* raw_alloc(size_type bytes) const override;
    void raw_free(void* ptr) const noexcept override;
    void raw_copy_from(const Executor* src, size_type bytes, const void* src_ptr,
                       void* dst_ptr) const override;
    void synchronize() const override;
    std::shared_ptr<const HostExecutor> get_master() const override { return master_; }
    bool is_host() const override { return false; }
    const gkomi_ctx* get_ctx() const override { return &ctx_; }
    int get_device_id() const { return ctx_.device; }
    void* get_stream() const { return ctx_.stream; }
    int get_num_multiprocessor() const { return info_.num_cus; }
    int get_warp_size() const { return info_.warp_size; }
    // HipExecutor::get_num_warps_per_sm / get_num_warps (executor.hpp:1846-1882):
    // four SIMDs per compute unit on CDNA (hip/base/executor.hip.cpp:234-238)
    int get_num_warps_per_sm() const { return 4; }
    int get_num_warps() const { return info_.num_cus * get_num_warps_per_sm(); }
    const gkomi_device_info& get_device_info() const { return info_; }
    static int get_num_devices();
    // bytes currently held in the pool / handed out
    size_type get_pool_bytes() const;
    void release_pool() const;

protected:
    Mi355xExecutor(int device_id, void* stream, bool caching, bool own_if_null);

private:
    gkomi_ctx ctx_{};
    bool own_stream_{};
    bool caching_{};
    gkomi_device_info info_{};
    std::shared_ptr<HostExecutor> master_;
    mutable std::mutex pool_mutex_;
    mutable std::multimap<size_type, void*> free_blocks_;
    mutable std::map<void*, size_type> live_blocks_;
    mutable size_type pooled_bytes_{};
};

// ------------------------------------------------------------------ array ----
// include/ginkgo/core/base/array.hpp: a typed buffer owned by (or viewing
// memory of) an executor.

template <typename T>
class array {
public:
    array() = default;
    explicit array(std::shared_ptr<const Executor> exec) : exec_{std::move(exec)} {}
    array(std::shared_ptr<const Executor> exec, size_type n) : exec_{std::move(exec)}
    {
        resize_and_reset(n);
    }
    array(std::shared_ptr<const Executor> exec, std::initializer_list<T> init)
        : array(exec, init.size())
    {
        std::vector<T> tmp(init);
        exec_->copy_from(exec_->get_master().get(), tmp.size(), tmp.data(), data_);
    }
    array(std::shared_ptr<const Executor> exec, const std::vector<T>& host)
        : array(exec, host.size())
    {
        if (!host.empty()) {
            exec_->copy_from(exec_->get_master().get(), host.size(), host.data(), data_);
        }
    }
    // copy to another executor
    array(std::shared_ptr<const Executor> exec, const array& other) : array(exec, other.size_)
    {
        if (size_) exec_->copy_from(other.exec_.get(), size_, other.data_, data_);
    }
    array(const array& other) : array(other.exec_, other) {}
    array(array&& other) noexcept { swap(other); }
    array& operator=(const array& other)
    {
        if (this != &other) {
            if (!exec_) exec_ = other.exec_;
            if (owning_ || !data_) {
                resize_and_reset(other.size_);
            } else if (size_ != other.size_) {
                throw DimensionMismatch(__FILE__, __LINE__, "view size mismatch in assignment");
            }
            if (size_) exec_->copy_from(other.exec_.get(), size_, other.data_, data_);
        }
        return *this;
    }
    array& operator=(array&& other) noexcept
    {
        swap(other);
        return *this;
    }
    ~array() { clear(); }

    // array::view (include/ginkgo/core/base/array.hpp): non-owning
    static array view(std::shared_ptr<const Executor> exec, size_type n, T* data)
    {
        array a(std::move(exec));
        a.size_ = n;
        a.data_ = data;
        a.owning_ = false;
        return a;
    }

    void resize_and_reset(size_type n)
    {
        clear();
        size_ = n;
        owning_ = true;
        data_ = n ? exec_->template alloc<T>(n) : nullptr;
    }
    void clear() noexcept
    {
        if (owning_ && data_) exec_->free(data_);
        data_ = nullptr;
        size_ = 0;
        owning_ = true;
    }
    void fill(T value);
    std::vector<T> to_host() const
    {
        std::vector<T> out(size_);
        if (size_) {
            exec_->get_master()->copy_from(exec_.get(), size_, data_, out.data());
        }
        return out;
    }
    void set_executor(std::shared_ptr<const Executor> exec)
    {
        if (exec == exec_) return;
        array tmp(exec, *this);
        swap(tmp);
    }
    void swap(array& o) noexcept
    {
        std::swap(exec_, o.exec_);
        std::swap(data_, o.data_);
        std::swap(size_, o.size_);
        std::swap(owning_, o.owning_);
    }

    size_type get_size() const noexcept { return size_; }
    T* get_data() noexcept { return data_; }
    const T* get_const_data() const noexcept { return data_; }
    std::shared_ptr<const Executor> get_executor() const noexcept { return exec_; }
    bool is_owning() const noexcept { return owning_; }

private:
    std::shared_ptr<const Executor> exec_;
    T* data_{};
    size_type size_{};
    bool owning_{true};
};

// device kernels take this context
inline const gkomi_ctx* ctx_of(const std::shared_ptr<const Executor>& exec)
{
    if (!exec || exec->is_host()) {
        GKO_NOT_SUPPORTED_MSG(
            "this operation runs on the Mi355xExecutor; the host executor only holds memory");
    }
    return exec->get_ctx();
}

// A recorded sequence of device work (a HIP graph; no reference
// counterpart), replayed on its executor's stream.  The buffers the recorded
// calls named must outlive it.
class Graph {
public:
    Graph(std::shared_ptr<const Executor> exec, gkomi_graph* g)
        : exec_{std::move(exec)}, graph_{g}
    {}
    Graph(const Graph&) = delete;
    Graph& operator=(const Graph&) = delete;
    ~Graph() { gkomi_graph_destroy(graph_); }
    void replay() const { GKO_CALL(gkomi_graph_launch(ctx_of(exec_), graph_)); }
    std::shared_ptr<const Executor> get_executor() const { return exec_; }

private:
    std::shared_ptr<const Executor> exec_;
    gkomi_graph* graph_;
};

}  // namespace gko

namespace gko {

// gko::share (include/ginkgo/core/base/utils_helper.hpp): unique -> shared
template <typename T>
std::shared_ptr<T> share(std::unique_ptr<T>&& p)
{
    return std::shared_ptr<T>(std::move(p));
}

}  // namespace gko
