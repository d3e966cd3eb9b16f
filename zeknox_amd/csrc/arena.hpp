// First-fit allocator over one pre-allocated device range (offsets only; the
// commit server owns the memory).  Free ranges are kept sorted and coalesced.
// Host code, not thread-safe (the server serialises calls).
#pragma once
#include <stdint.h>

#include <iterator>
#include <map>

namespace zk {

class Arena {
public:
    explicit Arena(uint64_t size = 0, uint64_t align = 256) : size_(size), align_(align)
    {
        if (size_) free_[0] = size_;
    }

    uint64_t size() const { return size_; }
    uint64_t used() const { return used_; }
    uint64_t round(uint64_t bytes) const { return (bytes + align_ - 1) / align_ * align_; }

    // Lowest free offset with room for `bytes`; false when nothing fits.
    bool alloc(uint64_t bytes, uint64_t &off)
    {
        const uint64_t need = round(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < need) continue;
            off = it->first;
            const uint64_t rest = it->second - need;
            free_.erase(it);
            if (rest) free_[off + need] = rest;
            used_ += need;
            return true;
        }
        return false;
    }

    // Returns a range obtained from alloc(bytes) (same byte count).
    void release(uint64_t off, uint64_t bytes)
    {
        const uint64_t len = round(bytes ? bytes : 1);
        used_ -= len;
        auto next = free_.lower_bound(off);
        uint64_t start = off, end = off + len;
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                start = prev->first;
                free_.erase(prev);
            }
        }
        if (next != free_.end() && next->first == end) {
            end += next->second;
            free_.erase(next);
        }
        free_[start] = end - start;
    }

    uint64_t largest_free() const
    {
        uint64_t m = 0;
        for (auto &f : free_) m = f.second > m ? f.second : m;
        return m;
    }
    size_t free_ranges() const { return free_.size(); }

private:
    uint64_t size_, align_, used_ = 0;
    std::map<uint64_t, uint64_t> free_;  // offset -> length
};

}  // namespace zk
