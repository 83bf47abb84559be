// Timers (include/ginkgo/core/base/timer.hpp): time points recorded on the
// host clock or in a device stream, and the elapsed time between two.
//   CpuTimer       std::chrono::steady_clock at the moment of record()
//   Mi355xTimer    a HIP event recorded on the executor's stream: the time
//                  at which the work enqueued before it has run (the role
//                  of HipTimer, hip/base/timer.hip.cpp)
#pragma once

#include <chrono>

#include "ginkgo_amd/base.hpp"

namespace gko {

class time_point {
public:
    ~time_point();
    time_point(time_point&& other) noexcept;
    time_point& operator=(time_point&& other) noexcept;
    time_point(const time_point&) = delete;
    time_point& operator=(const time_point&) = delete;

private:
    time_point() = default;
    friend class Timer;
    friend class CpuTimer;
    friend class Mi355xTimer;

    enum class type { cpu, device };
    type type_{type::cpu};
    std::chrono::steady_clock::time_point chrono_{};
    void* event_{};   // hipEvent_t of a device time point
    int device_{-1};  // the device the event belongs to
};

class Timer {
public:
    virtual ~Timer() = default;
    time_point create_time_point();
    // records the current time (the host clock, or the stream position)
    virtual void record(time_point& time) = 0;
    // waits until the time point has been reached
    virtual void wait(time_point& time) = 0;
    // stop - start, waiting for stop first
    std::chrono::nanoseconds difference(time_point& start, time_point& stop);
    // stop - start for time points that have been waited for
    virtual std::chrono::nanoseconds difference_async(const time_point& start,
                                                      const time_point& stop) = 0;
    // a CpuTimer for the host executor, an Mi355xTimer for a device one
    static std::unique_ptr<Timer> create_for_executor(std::shared_ptr<const Executor> exec);

protected:
    virtual void init_time_point(time_point& time) = 0;
};

class CpuTimer : public Timer {
public:
    void record(time_point& time) override;
    void wait(time_point& time) override;
    std::chrono::nanoseconds difference_async(const time_point& start,
                                              const time_point& stop) override;

protected:
    void init_time_point(time_point& time) override;
};

class Mi355xTimer : public Timer {
public:
    explicit Mi355xTimer(std::shared_ptr<const Executor> exec);
    void record(time_point& time) override;
    void wait(time_point& time) override;
    std::chrono::nanoseconds difference_async(const time_point& start,
                                              const time_point& stop) override;

protected:
    void init_time_point(time_point& time) override;

private:
    // the executor's device and stream, not the executor itself: a timer
    // held by a logger of the executor must not keep it alive
    gkomi_ctx ctx_{};
};

}  // namespace gko
