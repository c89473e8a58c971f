// arena.hpp -- first-fit sub-allocation of one device allocation: the HBM
// budget by which the request servers admit work.
#pragma once
#include <hip/hip_runtime.h>

#include <iterator>
#include <map>

namespace px {

constexpr size_t GRAIN = 256;  // arena allocation granule (bytes)

inline size_t round_up(size_t v, size_t g) { return (v + g - 1) / g * g; }

// First-fit allocator over one device allocation (offsets in bytes).
class Arena {
  public:
    ~Arena() {
        if (base_) (void)hipFree(base_);
    }
    hipError_t init(size_t bytes) {
        cap_ = round_up(bytes, GRAIN);
        hipError_t e = hipMalloc(&base_, cap_);
        if (e == hipSuccess) free_[0] = cap_;
        return e;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = round_up(bytes, GRAIN);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes, at = it->first + bytes;
                free_.erase(it);
                if (rest) free_[at] = rest;
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = round_up(bytes, GRAIN);
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
    char* base() const { return (char*)base_; }
    size_t capacity() const { return cap_; }
    size_t used() const { return used_; }

  private:
    void* base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;
};

}  // namespace px
