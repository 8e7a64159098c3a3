// First-fit sub-allocator of one fixed device buffer (offsets only; no HIP).
//
// The patch server carves every admitted patch's state, tolerances and
// outputs out of one arena allocated at start-up, so admission is a host-side
// bookkeeping decision that never calls hipMalloc while patches run.  Blocks
// are aligned to kAlign bytes; freed blocks coalesce with their neighbours.
#pragma once
#include <stddef.h>

#include <deque>
#include <iterator>
#include <map>

namespace dengo {

class Arena {
public:
    static constexpr size_t kAlign = 256;

    explicit Arena(size_t bytes = 0) { reset(bytes); }

    void reset(size_t bytes)
    {
        size_ = bytes / kAlign * kAlign;
        free_.clear();
        used_ = 0;
        peak_ = 0;
        if (size_) free_[0] = size_;
    }

    static size_t round_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

    // Offset of a block of at least n bytes, or -1 when no free run is long
    // enough (the lowest-addressed run that fits is taken).
    long long alloc(size_t n)
    {
        n = round_up(n ? n : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            const size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > n) free_[off + n] = len - n;
            used_ += n;
            if (used_ > peak_) peak_ = used_;
            return (long long)off;
        }
        return -1;
    }

    // Returns a block from alloc (its rounded size must be given back).
    void release(size_t off, size_t n)
    {
        n = round_up(n ? n : 1);
        used_ -= n;
        auto next = free_.lower_bound(off);
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                off = prev->first;
                n += prev->second;
                free_.erase(prev);
            }
        }
        if (next != free_.end() && off + n == next->first) {
            n += next->second;
            free_.erase(next);
        }
        free_[off] = n;
    }

    bool fits(size_t n) const
    {
        n = round_up(n ? n : 1);
        for (const auto &kv : free_)
            if (kv.second >= n) return true;
        return false;
    }

    size_t size() const { return size_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    size_t largest_free() const
    {
        size_t m = 0;
        for (const auto &kv : free_)
            if (kv.second > m) m = kv.second;
        return m;
    }
    size_t free_runs() const { return free_.size(); }

private:
    size_t size_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;   // offset -> length of each free run
};

// Admission order of the patch server.  waiting holds requests in arrival
// order, each with a count of the times a later request was admitted ahead
// of it.  Returns the index of the first request that fits(), or -1.  A
// request that has been passed max_bypass times blocks every later one, so a
// large patch waits for memory at most max_bypass admissions.  The passed
// requests' counts are incremented.
template <class Req, class Fits>
int pick_admission(std::deque<Req> &waiting, Fits fits, int max_bypass)
{
    for (size_t i = 0; i < waiting.size(); ++i) {
        if (fits(waiting[i])) {
            for (size_t j = 0; j < i; ++j) waiting[j].bypassed++;
            return (int)i;
        }
        if (waiting[i].bypassed >= max_bypass) return -1;
    }
    return -1;
}

}  // namespace dengo
