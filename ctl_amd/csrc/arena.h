// Best-fit allocator of the serving loop's HBM arena: host-side bookkeeping of offsets into one
// device allocation (no device code, so that it can be built and checked on its own).
#pragma once

#include <cstddef>
#include <iterator>
#include <map>

namespace ctlsrv {

constexpr size_t kAlign = 256;

inline size_t alignUp(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

// Best-fit allocator over [0, size): blocks are kAlign-aligned; neighbouring free blocks merge.
class Arena {
public:
    void init(size_t size)
    {
        size_ = size;
        free_.clear();
        free_[0] = size;
    }
    bool alloc(size_t bytes, size_t* off)
    {
        bytes = alignUp(bytes);
        auto best = free_.end();
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes && (best == free_.end() || it->second < best->second)) best = it;
        if (best == free_.end()) return false;
        *off = best->first;
        const size_t rest = best->second - bytes;
        const size_t start = best->first;
        free_.erase(best);
        if (rest) free_[start + bytes] = rest;
        used_ += bytes;
        return true;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = alignUp(bytes);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
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
    size_t size() const { return size_; }

    size_t used() const { return used_; }
    // free blocks (offset -> size), for checks
    const std::map<size_t, size_t>& freeBlocks() const { return free_; }

private:
    size_t size_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;
};

} // namespace ctlsrv
