// GPU-time accounting for the solver phases.
#pragma once

#include "common.hpp"

#include <utility>
#include <vector>

namespace aspher {

// Accumulates GPU time of (start, stop) event pairs without synchronising:
// completed pairs are drained whenever the host has synchronised anyway.
// With stride k only every k-th pair records events (each record leaves the
// device idle for a few microseconds) and the total is extrapolated from
// the timed ones.
class EventTimer {
public:
    explicit EventTimer(int stride = 1) : stride_(stride < 1 ? 1 : stride) {}
    ~EventTimer() {
        for (auto& pr : pending_) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
        for (auto e : pool_) (void)hipEventDestroy(e);
    }
    void start(hipStream_t s) {
        on_ = (calls_++ % stride_) == 0;
        if (!on_) return;
        cur_ = get();
        (void)hipEventRecord(cur_, s);
    }
    void stop(hipStream_t s) {
        if (!on_) return;
        hipEvent_t e = get();
        (void)hipEventRecord(e, s);
        pending_.emplace_back(cur_, e);
    }
    // a (start, stop) pair that ran without events (inside a replayed graph):
    // counted, and covered by the extrapolation from the timed pairs
    void skip() { ++calls_; }
    // takes the pairs that have completed (all of them after a stream
    // synchronisation); pairs still in flight stay pending
    void drain() {
        std::size_t keep = 0;
        for (std::size_t i = 0; i < pending_.size(); ++i) {
            auto& pr = pending_[i];
            if (hipEventQuery(pr.second) != hipSuccess) {
                pending_[keep++] = pr;
                continue;
            }
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
                total_ += ms * 1e-3;
                ++timed_;
            }
            pool_.push_back(pr.first);
            pool_.push_back(pr.second);
        }
        pending_.resize(keep);
    }
    // all calls: the mean of the timed pairs taken so far times the number
    // of (start, stop) calls (exact with stride 1 once every pair is drained)
    double seconds() const {
        return timed_ ? total_ * static_cast<double>(calls_) / static_cast<double>(timed_) : 0.0;
    }

private:
    hipEvent_t get() {
        if (!pool_.empty()) {
            hipEvent_t e = pool_.back();
            pool_.pop_back();
            return e;
        }
        // elapsed times only: no system-scope fence at record (the host
        // never reads device memory through these events), which would
        // stall the stream for several microseconds per record
        hipEvent_t e;
        ASPHER_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        return e;
    }
    std::vector<hipEvent_t> pool_;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending_;
    hipEvent_t cur_ = nullptr;
    double total_ = 0.0;
    int stride_ = 1;
    bool on_ = false;
    long long calls_ = 0, timed_ = 0;
};

} // namespace aspher
