// Host check of the plane-tile classify's workgroup -> tile map (pt_tile_of):
// over random grid shapes, the grid the library launches (8 * B * ceil(NKB *
// groups / 8) workgroups) covers every tile (x-tile, row block, plane block)
// exactly once, sends the rest past the last plane block, and with single
// columns per group gives every XCD (workgroups b with b % 8 equal) the same
// number of tiles within one.  Plain C++ (g++); prints OK.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "mdmc_tilemap.h"

int main() {
  std::mt19937 rng(7);
  for (int it = 0; it < 3000; ++it) {
    const uint32_t XT = 1 + rng() % 5, NRB = 1 + rng() % 160, NKB = 1 + rng() % 300;
    const uint32_t Bs[] = {1, 1, 2, 3, 4, 8};
    const uint32_t B = Bs[rng() % 6];
    const uint32_t cols = XT * NRB, ng = (cols + B - 1) / B;
    const uint64_t blocks = 8ull * B * (((uint64_t)NKB * ng + 7) / 8);
    std::vector<uint8_t> seen((size_t)cols * NKB, 0);
    uint64_t per_xcd[8] = {0};
    for (uint64_t b = 0; b < blocks; ++b) {
      const mdmc::PtTile t = mdmc::pt_tile_of((uint32_t)b, XT, NRB, NKB, B);
      if (t.kb == NKB) continue;
      if (t.xt >= XT || t.rb >= NRB || t.kb > NKB) {
        std::printf("out of range: b %llu XT %u NRB %u NKB %u B %u\n", (unsigned long long)b, XT, NRB, NKB, B);
        return 1;
      }
      uint8_t &s = seen[((size_t)t.kb * NRB + t.rb) * XT + t.xt];
      if (s) {
        std::printf("tile twice: XT %u NRB %u NKB %u B %u\n", XT, NRB, NKB, B);
        return 1;
      }
      s = 1;
      per_xcd[b & 7] += 1;
    }
    for (uint8_t s : seen)
      if (!s) {
        std::printf("tile missed: XT %u NRB %u NKB %u B %u\n", XT, NRB, NKB, B);
        return 1;
      }
    uint64_t lo = per_xcd[0], hi = per_xcd[0];
    for (int x = 1; x < 8; ++x) {
      lo = per_xcd[x] < lo ? per_xcd[x] : lo;
      hi = per_xcd[x] > hi ? per_xcd[x] : hi;
    }
    if (B == 1 && hi - lo > 1) {
      std::printf("XCD imbalance %llu..%llu: XT %u NRB %u NKB %u\n", (unsigned long long)lo,
                  (unsigned long long)hi, XT, NRB, NKB);
      return 1;
    }
  }
  std::printf("OK\n");
  return 0;
}
