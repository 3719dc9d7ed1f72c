// CPU emulation of the decoupled look-back protocol of k_scan
// (framework_amd/csrc/mdmc_kernels.hpp): many tiny tiles, taken by ticket in
// order by several threads that run at random speeds, so tiles publish and
// resolve in random order.  Each tile's four channel words are
// (status << 32) | value with status 1 = aggregate, 2 = inclusive, published
// one channel at a time; a tile resolves each channel on its own by reading
// its predecessors 64 at a time (one wave's lanes: lane l reads tile
// look - l, spinning while that word is unpublished) and summing back to the
// nearest inclusive word of that channel.  Every tile's exclusive prefixes
// must equal a sequential scan, and no spin may reach the limit.
//
// Usage: lookback_emulation [TILES THREADS SEED]; prints OK on success.
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

namespace {

constexpr int kWave = 64;
constexpr uint32_t kSpinLimit = 1u << 26;
constexpr uint64_t AGG = 1ull << 32, INC = 2ull << 32;

struct Shared {
  uint32_t ntiles;
  std::vector<uint32_t> agg;                  // 4 channels per tile
  std::vector<std::atomic<uint64_t>> state;   // 4 words per tile
  std::vector<uint32_t> excl;                 // result: 4 per tile
  std::atomic<uint32_t> ticket{0};
  std::atomic<uint32_t> timeouts{0};
  explicit Shared(uint32_t n) : ntiles(n), agg(4 * (size_t)n), state(4 * (size_t)n), excl(4 * (size_t)n) {
    for (auto &w : state) w.store(0, std::memory_order_relaxed);
  }
};

void jitter(std::mt19937 &rng) {
  const uint32_t r = rng() % 64;
  if (r < 2) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 50));
  else if (r < 24) std::this_thread::yield();
}

void tile_work(Shared &S, uint32_t t, std::mt19937 &rng) {
  std::atomic<uint64_t> *my = &S.state[4 * (size_t)t];
  const uint32_t *a = &S.agg[4 * (size_t)t];
  uint32_t ex[4] = {0, 0, 0, 0};
  jitter(rng);
  if (t == 0) {
    for (int ch = 0; ch < 4; ++ch) {
      my[ch].store(INC | a[ch], std::memory_order_relaxed);
      jitter(rng);
    }
  } else {
    for (int ch = 0; ch < 4; ++ch) {
      my[ch].store(AGG | a[ch], std::memory_order_relaxed);
      jitter(rng);
    }
    int64_t look = (int64_t)t - 1;
    uint32_t done = 0;
    while (done != 0xFu) {
      uint64_t w[kWave][4];
      // each lane reads its predecessor's words (not-yet-resolved channels),
      // spinning on the ones not published yet
      for (int lane = 0; lane < kWave; ++lane) {
        const int64_t idx = look - lane;
        for (int ch = 0; ch < 4; ++ch) w[lane][ch] = INC;  // before tile 0: inclusive zeros
        if (idx < 0) continue;
        const std::atomic<uint64_t> *p = &S.state[4 * (size_t)idx];
        uint32_t spins = 0;
        for (int ch = 0; ch < 4; ++ch) {
          if ((done >> ch) & 1u) continue;
          w[lane][ch] = p[ch].load(std::memory_order_relaxed);
          while ((w[lane][ch] >> 32) == 0 && ++spins < kSpinLimit) {
            if ((spins & 255u) == 0) std::this_thread::yield();
            w[lane][ch] = p[ch].load(std::memory_order_relaxed);
          }
        }
        if (spins >= kSpinLimit) S.timeouts.fetch_add(1);
      }
      for (int ch = 0; ch < 4; ++ch) {
        if ((done >> ch) & 1u) continue;
        // the nearest predecessor whose word is inclusive (ballot + ffs)
        int stop = kWave;
        for (int lane = 0; lane < kWave; ++lane)
          if ((w[lane][ch] >> 32) == 2) {
            stop = lane;
            break;
          }
        uint32_t add = 0;
        for (int lane = 0; lane < kWave && lane <= stop; ++lane)
          if (look - lane >= 0) add += (uint32_t)w[lane][ch];
        ex[ch] += add;
        if (stop < kWave) done |= 1u << ch;
      }
      look -= kWave;
      jitter(rng);
    }
    for (int ch = 0; ch < 4; ++ch) {
      my[ch].store(INC | (uint32_t)(ex[ch] + a[ch]), std::memory_order_relaxed);
      jitter(rng);
    }
  }
  for (int ch = 0; ch < 4; ++ch) S.excl[4 * (size_t)t + ch] = ex[ch];
}

}  // namespace

int main(int argc, char **argv) {
  const uint32_t ntiles = argc > 1 ? (uint32_t)std::atoi(argv[1]) : 20000u;
  const int nthreads = argc > 2 ? std::atoi(argv[2]) : 8;
  const uint32_t seed = argc > 3 ? (uint32_t)std::atoi(argv[3]) : 1u;
  Shared S(ntiles);
  std::mt19937 gen(seed);
  for (uint32_t t = 0; t < ntiles; ++t) {
    // tiny tiles, many of them empty in some channels
    S.agg[4 * (size_t)t + 0] = gen() % 3;
    S.agg[4 * (size_t)t + 1] = gen() % 8 == 0 ? gen() % 5000 : 0;
    S.agg[4 * (size_t)t + 2] = gen() % 100;
    S.agg[4 * (size_t)t + 3] = gen() % 2;
  }
  std::vector<std::thread> th;
  for (int k = 0; k < nthreads; ++k)
    th.emplace_back([&S, k, seed] {
      std::mt19937 rng(seed * 7919u + (uint32_t)k);
      for (;;) {
        const uint32_t t = S.ticket.fetch_add(1);
        if (t >= S.ntiles) break;
        tile_work(S, t, rng);
      }
    });
  for (auto &x : th) x.join();
  uint32_t run[4] = {0, 0, 0, 0};
  for (uint32_t t = 0; t < ntiles; ++t)
    for (int ch = 0; ch < 4; ++ch) {
      if (S.excl[4 * (size_t)t + ch] != run[ch]) {
        std::printf("FAIL tile %u channel %d: %u != %u\n", t, ch, S.excl[4 * (size_t)t + ch], run[ch]);
        return 1;
      }
      run[ch] += S.agg[4 * (size_t)t + ch];
    }
  if (S.timeouts.load()) {
    std::printf("FAIL %u spin timeouts\n", S.timeouts.load());
    return 1;
  }
  std::printf("OK\n");
  return 0;
}
