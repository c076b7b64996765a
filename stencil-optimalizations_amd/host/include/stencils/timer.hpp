// Wall-clock timing of plug-in phases (reference: src/infrastructure/timer.hpp:10-52).
// The reference brackets each phase with cudaDeviceSynchronize; the HIP plug-ins
// here finish every phase with the device work done (gol_sync), so the host
// clock alone measures the phase.
#pragma once

#include <chrono>
#include <cstddef>
#include <limits>

namespace infrastructure {

class Timer {
   public:
    using clock_type = std::chrono::high_resolution_clock;

    template <typename Func>
    static double measure_ms(Func&& func) {
        auto start = clock_type::now();
        func();
        auto end = clock_type::now();
        return std::chrono::duration<double, std::milli>(end - start).count();
    }
};

class StopWatch {
   public:
    using clock_type = std::chrono::high_resolution_clock;
    static constexpr long MAX_TIME = std::numeric_limits<long>::max();

    explicit StopWatch(std::size_t count_down_seconds) : count_down_seconds_(static_cast<long>(count_down_seconds)) {
        restart();
    }
    void restart() { start_ = clock_type::now(); }
    bool time_is_up() const {
        const auto secs = std::chrono::duration_cast<std::chrono::seconds>(clock_type::now() - start_).count();
        return secs >= count_down_seconds_;
    }

   private:
    long count_down_seconds_;
    clock_type::time_point start_;
};

}  // namespace infrastructure
