// Host check of the device-memory pool bookkeeping (framework_amd/csrc/mdmc_pool.h):
// random allocations and releases against a simple model of the address space.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "mdmc_pool.h"

static int fail(const char *what, unsigned long long a = 0, unsigned long long b = 0) {
  std::printf("FAIL %s %llu %llu\n", what, a, b);
  return 1;
}

int main() {
  const uint64_t size = 1ull << 24;
  mdmc::Pool pool(size, 256);
  std::vector<unsigned char> owner(size / 256, 0);  // 256-byte granules in use
  std::map<uint64_t, uint64_t> live;                // offset -> rounded length
  std::mt19937_64 rng(7);
  for (int step = 0; step < 200000; ++step) {
    if (live.empty() || (rng() % 100) < 55) {
      const uint64_t want = 1 + rng() % (1ull << (8 + rng() % 12));
      uint64_t off = 0;
      if (pool.alloc(want, &off)) {
        const uint64_t len = (want + 255) / 256 * 256;
        if (off % 256 || off + len > size) return fail("range", off, len);
        for (uint64_t g = off / 256; g < (off + len) / 256; ++g) {
          if (owner[g]) return fail("overlap", off, len);
          owner[g] = 1;
        }
        live[off] = len;
      } else {
        // a failed allocation means no free run of that length exists
        const uint64_t len = (want + 255) / 256;
        uint64_t run = 0, best = 0;
        for (unsigned char o : owner) {
          run = o ? 0 : run + 1;
          best = run > best ? run : best;
        }
        if (best >= len) return fail("missed fit", best, len);
      }
    } else if ((rng() % 100) < 60) {
      auto it = live.begin();
      std::advance(it, rng() % live.size());
      if (!pool.release(it->first)) return fail("release", it->first);
      for (uint64_t g = it->first / 256; g < (it->first + it->second) / 256; ++g) owner[g] = 0;
      live.erase(it);
    } else {
      // cut a block down to a random prefix (possibly all or nothing of it)
      auto it = live.begin();
      std::advance(it, rng() % live.size());
      const uint64_t off = it->first, len = it->second;
      const uint64_t want = rng() % (len + 1);
      const uint64_t keep = (want + 255) / 256 * 256;
      if (!pool.shrink(off, want)) return fail("shrink", off, want);
      for (uint64_t g = (off + keep) / 256; g < (off + len) / 256; ++g) owner[g] = 0;
      if (keep == 0) live.erase(it); else it->second = keep;
      if (pool.shrink(off, keep + 256)) return fail("grow accepted", off, keep);
    }
    uint64_t used = 0;
    for (auto &kv : live) used += kv.second;
    if (used != pool.in_use()) return fail("in_use", used, pool.in_use());
  }
  for (auto &kv : live) pool.release(kv.first);
  if (pool.in_use() != 0 || pool.free_blocks() != 1 || pool.largest_free() != size) return fail("coalesce");
  if (pool.release(12345)) return fail("unknown offset accepted");
  if (pool.shrink(12345, 0)) return fail("unknown offset shrunk");
  std::printf("OK\n");
  return 0;
}
