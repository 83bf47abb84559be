// Timers: host clock and HIP events on the executor's stream
// (core/base/timer.cpp, hip/base/timer.hip.cpp).
#include <cmath>

#include "ginkgo_amd/timer.hpp"

namespace gko {

time_point::~time_point()
{
    if (type_ == type::device && event_) gkomi_event_destroy(device_, event_);
}

time_point::time_point(time_point&& other) noexcept
    : type_{other.type_}, chrono_{other.chrono_}, event_{other.event_}, device_{other.device_}
{
    other.type_ = type::cpu;
    other.event_ = nullptr;
}

time_point& time_point::operator=(time_point&& other) noexcept
{
    if (this != &other) {
        std::swap(type_, other.type_);
        std::swap(chrono_, other.chrono_);
        std::swap(event_, other.event_);
        std::swap(device_, other.device_);
    }
    return *this;
}

time_point Timer::create_time_point()
{
    time_point t;
    init_time_point(t);
    return t;
}

std::chrono::nanoseconds Timer::difference(time_point& start, time_point& stop)
{
    wait(stop);
    return difference_async(start, stop);
}

std::unique_ptr<Timer> Timer::create_for_executor(std::shared_ptr<const Executor> exec)
{
    if (!exec || exec->is_host()) return std::unique_ptr<Timer>(new CpuTimer());
    return std::unique_ptr<Timer>(new Mi355xTimer(std::move(exec)));
}

void CpuTimer::init_time_point(time_point& time) { time.type_ = time_point::type::cpu; }

void CpuTimer::record(time_point& time)
{
    if (time.type_ != time_point::type::cpu) {
        throw Error(__FILE__, __LINE__, "CpuTimer: the time point belongs to another timer");
    }
    time.chrono_ = std::chrono::steady_clock::now();
}

void CpuTimer::wait(time_point&) {}

std::chrono::nanoseconds CpuTimer::difference_async(const time_point& start,
                                                    const time_point& stop)
{
    return std::chrono::duration_cast<std::chrono::nanoseconds>(stop.chrono_ - start.chrono_);
}

Mi355xTimer::Mi355xTimer(std::shared_ptr<const Executor> exec)
{
    ctx_ = *ctx_of(exec);  // rejects the host executor
}

void Mi355xTimer::init_time_point(time_point& time)
{
    time.type_ = time_point::type::device;
    time.device_ = ctx_.device;
    GKO_CALL(gkomi_event_create(time.device_, &time.event_));
}

void Mi355xTimer::record(time_point& time)
{
    if (time.type_ != time_point::type::device || !time.event_) {
        throw Error(__FILE__, __LINE__, "Mi355xTimer: the time point belongs to another timer");
    }
    GKO_CALL(gkomi_event_record(&ctx_, time.event_));
}

void Mi355xTimer::wait(time_point& time)
{
    GKO_CALL(gkomi_event_synchronize(time.event_));
}

std::chrono::nanoseconds Mi355xTimer::difference_async(const time_point& start,
                                                       const time_point& stop)
{
    float ms = 0.f;
    GKO_CALL(gkomi_event_elapsed_ms(start.event_, stop.event_, &ms));
    return std::chrono::nanoseconds{static_cast<int64>(std::llround(double(ms) * 1e6))};
}

}  // namespace gko
