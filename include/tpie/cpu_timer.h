// tpie::cpu_timer subset: wall-clock stage timers (nni.cpp:552-560,
// terranni.cpp:559-561).
#ifndef TERRANNI_TPIE_CPU_TIMER_H
#define TERRANNI_TPIE_CPU_TIMER_H

#include <chrono>

namespace tpie {

class cpu_timer {
public:
    void start()
    {
        if (!running_) {
            t0_ = std::chrono::steady_clock::now();
            running_ = true;
        }
    }
    void stop()
    {
        if (running_) {
            acc_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
            running_ = false;
        }
    }
    void reset()
    {
        acc_ = 0.0;
        running_ = false;
    }
    double wall_time() const
    {
        double a = acc_;
        if (running_) a += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
        return a;
    }

private:
    std::chrono::steady_clock::time_point t0_{};
    double acc_ = 0.0;
    bool running_ = false;
};

}  // namespace tpie

#endif
