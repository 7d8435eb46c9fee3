// First-fit sub-allocator over one device allocation (byte offsets only, no
// device calls), used by the commit server for HBM admission.  Holes are kept
// sorted by offset and merged with both neighbours on release, so a released
// range is immediately reusable by a job of up to the merged size.
#pragma once

#include <cstddef>
#include <iterator>
#include <map>

namespace tvh {

class FirstFitArena {
public:
    void reset(size_t size) {
        free_.clear();
        if (size) free_[0] = size;
        size_ = size;
        used_ = 0;
    }
    bool alloc(size_t bytes, size_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                size_t rest = it->second - bytes, noff = it->first + bytes;
                free_.erase(it);
                if (rest) free_[noff] = rest;
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }
    size_t size() const { return size_; }
    size_t holes() const { return free_.size(); }
    size_t largest_hole() const {
        size_t b = 0;
        for (auto& h : free_) b = h.second > b ? h.second : b;
        return b;
    }

private:
    std::map<size_t, size_t> free_;  // offset -> length
    size_t size_ = 0, used_ = 0;
};

}  // namespace tvh
