// server_core.h -- device-independent parts of the serving loop: the
// first-fit arena that admission reserves request memory from, the FIFO
// admission rule and the reach of a failed batch.  Header-only so the host tests
// compile it without a GPU.
#pragma once
#include <cstddef>
#include <cstdint>
#include <deque>
#include <iterator>
#include <map>

namespace qtx {
namespace srv {

// Offsets into a region of `capacity` bytes; blocks are `align`-aligned.
// Free space is kept coalesced, so a freed region is reusable at once.
class Arena {
  public:
    explicit Arena(size_t capacity = 0, size_t align = 256) { reset(capacity, align); }

    void reset(size_t capacity, size_t align = 256) {
        align_ = align;
        cap_ = capacity - capacity % align;
        used_ = peak_ = 0;
        free_.clear();
        if (cap_) free_[0] = cap_;
    }

    size_t round(size_t n) const { return n == 0 ? align_ : (n + align_ - 1) / align_ * align_; }

    // First fit; false if no free run is large enough.
    bool alloc(size_t n, size_t &off) {
        n = round(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            off = it->first;
            const size_t rest = it->second - n;
            free_.erase(it);
            if (rest) free_[off + n] = rest;
            used_ += n;
            if (used_ > peak_) peak_ = used_;
            return true;
        }
        return false;
    }

    void release(size_t off, size_t n) {
        n = round(n);
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }

    size_t capacity() const { return cap_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    size_t largest_free() const {
        size_t m = 0;
        for (auto &kv : free_) m = kv.second > m ? kv.second : m;
        return m;
    }
    size_t free_runs() const { return free_.size(); }

  private:
    size_t align_ = 256, cap_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;  // offset -> length
};

// FIFO admission: take requests from the front of `waiting` while each one's
// reservation succeeds and fewer than max_take were taken; stop at the first
// that does not fit, so a large request is never overtaken (no starvation).
// reserve(req) makes the reservation and returns false if it cannot.
template <class Req, class Reserve>
size_t admit_fifo(std::deque<Req> &waiting, size_t max_take, Reserve &&reserve, std::deque<Req> &taken) {
    size_t n = 0;
    while (!waiting.empty() && n < max_take) {
        if (!reserve(waiting.front())) break;
        taken.push_back(waiting.front());
        waiting.pop_front();
        n++;
    }
    return n;
}

// How far a HIP error that a batch ran into reaches.  Errors raised by work on the
// device -- a memory fault, an assert or exception in a kernel, a time-out -- and a
// lost or uninitialised device leave the device's state unknown: the server fails
// the batch, everything in flight and everything waiting, and refuses new
// requests.  Every other error is a call the runtime refused before any of its
// work ran (an invalid argument, no memory for a launch, a capture conflict): it
// fails that batch's requests only, and the server goes on with the next batch.
// Codes are hipError_t values (hip_runtime_api.h).
enum class FaultScope { Batch, Server };
inline FaultScope fault_scope(int hip_error) {
    switch (hip_error) {
        case 3:    // hipErrorNotInitialized
        case 4:    // hipErrorDeinitialized
        case 53:   // hipErrorPriorLaunchFailure
        case 100:  // hipErrorNoDevice
        case 101:  // hipErrorInvalidDevice
        case 201:  // hipErrorInvalidContext
        case 214:  // hipErrorECCNotCorrectable
        case 700:  // hipErrorIllegalAddress
        case 702:  // hipErrorLaunchTimeOut
        case 709:  // hipErrorContextIsDestroyed
        case 710:  // hipErrorAssert
        case 719:  // hipErrorLaunchFailure
        case 999:  // hipErrorUnknown
            return FaultScope::Server;
        default:
            return FaultScope::Batch;
    }
}

// Removes the first element of `waiting` equal to `r` (a request still waiting
// for admission); true if it was there.
template <class Req> bool remove_waiting(std::deque<Req> &waiting, const Req &r) {
    for (auto it = waiting.begin(); it != waiting.end(); ++it)
        if (*it == r) {
            waiting.erase(it);
            return true;
        }
    return false;
}

}  // namespace srv
}  // namespace qtx
