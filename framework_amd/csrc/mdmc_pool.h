// First-fit allocator over one block of device memory, in offsets only (no
// HIP calls), so that its bookkeeping can be tested on the host.  Free blocks
// are kept in address order and merged with their neighbours on release.
#pragma once

#include <stdint.h>

#include <iterator>
#include <map>

namespace mdmc {

class Pool {
 public:
  explicit Pool(uint64_t size = 0, uint64_t align = 256) { reset(size, align); }

  void reset(uint64_t size, uint64_t align = 256) {
    align_ = align ? align : 1;
    size_ = size / align_ * align_;
    free_.clear();
    used_.clear();
    in_use_ = 0;
    if (size_) free_[0] = size_;
  }

  // Offset of a new block of at least `bytes` (rounded up to the alignment),
  // or false when no free block is large enough.
  bool alloc(uint64_t bytes, uint64_t *off) {
    const uint64_t need = (bytes + align_ - 1) / align_ * align_;
    if (need == 0 || need > size_) return false;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < need) continue;
      const uint64_t o = it->first, rest = it->second - need;
      free_.erase(it);
      if (rest) free_[o + need] = rest;
      used_[o] = need;
      in_use_ += need;
      *off = o;
      return true;
    }
    return false;
  }

  // Releases a block returned by alloc; false for an unknown offset.
  bool release(uint64_t off) {
    auto u = used_.find(off);
    if (u == used_.end()) return false;
    uint64_t o = u->first, len = u->second;
    in_use_ -= len;
    used_.erase(u);
    auto next = free_.lower_bound(o);
    if (next != free_.end() && o + len == next->first) {
      len += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == o) {
        prev->second += len;
        return true;
      }
    }
    free_[o] = len;
    return true;
  }

  // Cuts a block returned by alloc down to its first `bytes` (rounded up to
  // the alignment) and frees the rest; a block cut to nothing is released.
  // False for an unknown offset or a size above the block's.
  bool shrink(uint64_t off, uint64_t bytes) {
    auto u = used_.find(off);
    if (u == used_.end()) return false;
    const uint64_t keep = (bytes + align_ - 1) / align_ * align_;
    if (keep > u->second) return false;
    if (keep == 0) return release(off);
    if (keep == u->second) return true;
    const uint64_t tail = off + keep, tail_len = u->second - keep;
    u->second = keep;
    used_[tail] = tail_len;  // in_use_ still counts it: release() takes it off
    return release(tail);
  }

  uint64_t size() const { return size_; }
  uint64_t in_use() const { return in_use_; }
  uint64_t largest_free() const {
    uint64_t m = 0;
    for (const auto &f : free_) m = f.second > m ? f.second : m;
    return m;
  }
  size_t free_blocks() const { return free_.size(); }
  size_t used_blocks() const { return used_.size(); }

 private:
  uint64_t size_ = 0, align_ = 256, in_use_ = 0;
  std::map<uint64_t, uint64_t> free_;  // offset -> length
  std::map<uint64_t, uint64_t> used_;  // offset -> length
};

}  // namespace mdmc
