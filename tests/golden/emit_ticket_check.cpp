// Host check of the persistent emit's chunk tickets (emit_tickets and
// emit_ticket_chunk, mdmc_tilemap.h), following k_emit's control flow: a
// workgroup starts at its slot's ticket, asks counter x for the next ticket
// at the start of each iteration (when the range needs it), processes one
// chunk per wave of the current ticket (chunks below hi), and stops at the
// first ticket that starts at or past hi.  Over random chunk counts, grids
// and wave counts, with the workgroups advanced in random order (any order in
// which the counter's atomics can land), every chunk in [0, n) is processed
// exactly once, no chunk at or past n is touched, every workgroup stops, and
// each counter hands out at most (tickets needed + workgroups) values.
// Plain C++ (g++); prints OK.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "mdmc_tilemap.h"

int main() {
  std::mt19937_64 rng(11);
  for (int it = 0; it < 4000; ++it) {
    const uint32_t n = (it % 5 == 0) ? (uint32_t)(rng() % 40) : (uint32_t)(rng() % 60000);
    const uint32_t G = 8u * (1u + (uint32_t)(rng() % (it % 3 == 0 ? 4 : 300)));
    const uint32_t W = 1u + (uint32_t)(rng() % 8);
    std::vector<uint8_t> seen(n, 0);
    uint64_t ctr[8] = {0};
    struct Wg {
      mdmc::EmitTickets et;
      uint64_t first;
      bool done;
      uint32_t x;
    };
    std::vector<Wg> wgs(G);
    std::vector<uint32_t> live;
    for (uint32_t b = 0; b < G; ++b) {
      Wg &w = wgs[b];
      w.x = b & 7u;
      w.et = mdmc::emit_tickets(n, w.x, G, W);
      w.first = mdmc::emit_ticket_chunk(w.et, b >> 3, W);
      w.done = w.first >= w.et.hi;
      if (!w.done) live.push_back(b);
    }
    uint64_t steps = 0;
    while (!live.empty()) {
      if (++steps > 50000000ull) {
        std::printf("no progress: n %u G %u W %u\n", n, G, W);
        return 1;
      }
      const size_t k = (size_t)(rng() % live.size());
      Wg &w = wgs[live[k]];
      uint64_t next = 0;
      if (w.et.dynamic) next = ctr[w.x]++;
      for (uint32_t v = 0; v < W; ++v) {
        const uint64_t ch = w.first + v;
        if (ch >= w.et.hi) continue;
        if (ch >= n || seen[ch]) {
          std::printf("chunk %llu %s: n %u G %u W %u\n", (unsigned long long)ch, ch >= n ? "out of range" : "twice",
                      n, G, W);
          return 1;
        }
        seen[ch] = 1;
      }
      bool stop = !w.et.dynamic;
      if (!stop) {
        w.first = mdmc::emit_ticket_chunk(w.et, (uint64_t)w.et.slots + next, W);
        stop = w.first >= w.et.hi;
      }
      if (stop) {
        w.done = true;
        live[k] = live.back();
        live.pop_back();
      }
    }
    for (uint32_t c = 0; c < n; ++c)
      if (!seen[c]) {
        std::printf("chunk %u missed: n %u G %u W %u\n", c, n, G, W);
        return 1;
      }
    for (uint32_t x = 0; x < 8; ++x) {
      const mdmc::EmitTickets e = mdmc::emit_tickets(n, x, G, W);
      const uint64_t need = (e.hi - e.lo + W - 1) / W;
      if (ctr[x] > need + G / 8) {
        std::printf("counter %u handed out %llu: n %u G %u W %u\n", x, (unsigned long long)ctr[x], n, G, W);
        return 1;
      }
    }
  }
  std::printf("OK\n");
  return 0;
}
