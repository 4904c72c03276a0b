// First-fit allocator of the serving scratch pool: offsets into one device
// buffer, handed out in whole 256-byte granules.  Header-only and free of
// HIP so that host tests can build it alone.
#pragma once

#include <cstdint>
#include <iterator>
#include <map>

namespace eithot {

struct ScratchPool {
  int64_t cap = 0;                   // usable bytes: whole granules of the buffer
  std::map<int64_t, int64_t> free_;  // offset -> length, coalesced

  static int64_t granule(int64_t n) { return (n + 255) & ~(int64_t)255; }

  void init(int64_t c) {
    cap = c > 0 ? c & ~(int64_t)255 : 0;
    free_.clear();
    if (cap > 0) free_[0] = cap;
  }
  // Whether a request of n bytes can ever be allocated (when the pool is empty).
  bool fits(int64_t n) const { return n <= 0 || granule(n) <= cap; }
  // Offset of n bytes, or -1 while no free run is long enough.
  int64_t alloc(int64_t n) {
    if (n <= 0) return 0;
    n = granule(n);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        const int64_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[off + n] = len - n;
        return off;
      }
    }
    return -1;
  }
  void release(int64_t off, int64_t n) {
    if (n <= 0 || off < 0) return;
    n = granule(n);
    auto it = free_.emplace(off, n).first;
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
  // Bytes free in total (for tests).
  int64_t free_bytes() const {
    int64_t n = 0;
    for (const auto& kv : free_) n += kv.second;
    return n;
  }
};

}  // namespace eithot
