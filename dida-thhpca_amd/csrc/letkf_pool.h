// letkf_pool.h -- first-fit allocation inside one device region, with the
// bookkeeping on the host.  The serving loop allocates its whole HBM budget
// once and hands each admitted job a region of it; a job's buffers are then
// carved from its region, so no hipMalloc or hipFree (which synchronises the
// device) runs while other jobs' kernels are in flight.
#pragma once
#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>

namespace dida {

// Offsets and lengths are multiples of kGranule.  Free blocks are kept by
// offset and merged with their neighbours on release.
class FitAlloc {
 public:
  static constexpr size_t kGranule = 256;
  static constexpr size_t npos = ~size_t(0);

  void reset(size_t size) {
    free_.clear();
    size_ = size / kGranule * kGranule;
    used_ = 0;
    if (size_ > 0) free_[0] = size_;
  }
  static size_t round(size_t n) { return (n + kGranule - 1) / kGranule * kGranule; }

  // offset of a block of at least n bytes (the lowest that fits), or npos
  size_t alloc(size_t n) {
    n = round(n == 0 ? 1 : n);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      const size_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > n) free_[off + n] = len - n;
      used_ += n;
      return off;
    }
    return npos;
  }

  // returns a block from alloc (n as asked there); false if it overlaps
  // free space or lies outside the region (the bookkeeping is left unchanged)
  bool release(size_t off, size_t n) {
    n = round(n == 0 ? 1 : n);
    if (off % kGranule || off + n > size_ || off + n < off) return false;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && next->first < off + n) return false;
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second > off) return false;
    }
    size_t start = off, len = n;
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        start = prev->first;
        len += prev->second;
        free_.erase(prev);
      }
    }
    if (next != free_.end() && next->first == off + n) {
      len += next->second;
      free_.erase(next);
    }
    free_[start] = len;
    used_ -= n;
    return true;
  }

  size_t size() const { return size_; }
  size_t used() const { return used_; }
  size_t largest() const {
    size_t m = 0;
    for (const auto& b : free_) m = b.second > m ? b.second : m;
    return m;
  }
  size_t n_free_blocks() const { return free_.size(); }

 private:
  std::map<size_t, size_t> free_;   // offset -> length
  size_t size_ = 0, used_ = 0;
};

// A device region and its allocator.  Buffers of the engine made while a
// pool is bound to the calling host thread come from it.
struct DevicePool {
  char* base = nullptr;
  FitAlloc fit;
};

// binds pool (or none) to the calling host thread; returns the previous one
DevicePool* bind_thread_pool(DevicePool* pool);
DevicePool* thread_pool();

}  // namespace dida
