// Exceptions must not escape an OpenMP region (that terminates the process):
// parallel loops run their bodies through ParallelErrors, which keeps the
// first exception and rethrows it on the calling thread after the loop.
#pragma once
#include <exception>
#include <mutex>

namespace hdd {

class ParallelErrors {
public:
    template <class F>
    void run(F&& body) {
        try {
            body();
        } catch (...) {
            std::lock_guard<std::mutex> g(m_);
            if (!first_) first_ = std::current_exception();
        }
    }
    void rethrow() {
        if (first_) std::rethrow_exception(first_);
    }

private:
    std::mutex m_;
    std::exception_ptr first_;
};

}  // namespace hdd
