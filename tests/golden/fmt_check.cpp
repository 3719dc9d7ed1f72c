// Host check of the OBJ writer's number text (mdmc_fmt.h) against snprintf:
// fmt_g against "%g" of the float's double over every 997th 32-bit pattern
// (both signs, zeros, subnormals, NaNs, infinities), over 4000 floats either
// side of every power of ten and of every value where rounding to six
// digits carries into a new decade (x.999995e+k), and over random floats;
// fmt_u32 against "%u".  Plain C++ (g++); prints OK.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

#include "mdmc_fmt.h"

static bool same_g(float f) {
  char a[64], b[64];
  const size_t n = mdmc::fmt_g(a, f);
  const int m = std::snprintf(b, sizeof b, "%g", (double)f);
  if ((size_t)m != n || std::memcmp(a, b, n) != 0) {
    uint32_t bits;
    std::memcpy(&bits, &f, 4);
    std::printf("fmt_g 0x%08x: '%.*s' vs '%s'\n", bits, (int)n, a, b);
    return false;
  }
  return true;
}

static float from_bits(uint32_t b) {
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}

static bool around(float c, int span) {
  uint32_t b;
  std::memcpy(&b, &c, 4);
  for (int d = -span; d <= span; ++d) {
    const int64_t x = (int64_t)b + d;
    if (x < 0 || x > 0xFFFFFFFFll) continue;
    if (!same_g(from_bits((uint32_t)x)) || !same_g(-from_bits((uint32_t)x))) return false;
  }
  return true;
}

int main() {
  for (uint64_t b = 0; b <= 0xFFFFFFFFull; b += 997)
    if (!same_g(from_bits((uint32_t)b))) return 1;
  for (int k = -7; k <= 16; ++k) {
    if (!around((float)std::pow(10.0, k), 2000)) return 1;
    if (!around((float)(9.999995 * std::pow(10.0, k)), 2000)) return 1;
    if (!around((float)(9.99999 * std::pow(10.0, k)), 200)) return 1;
  }
  std::mt19937 rng(5);
  std::uniform_real_distribution<float> u(-600.0f, 600.0f);
  for (int i = 0; i < 2000000; ++i)
    if (!same_g(u(rng)) || !same_g(from_bits((uint32_t)rng()))) return 1;
  const uint32_t edge[] = {0u, 1u, 9u, 10u, 99u, 100u, 4294967295u, 2147483648u, 1000000000u, 999999999u};
  for (uint32_t v : edge) {
    char a[32], b[32];
    const size_t n = mdmc::fmt_u32(a, v);
    const int m = std::snprintf(b, sizeof b, "%u", v);
    if ((size_t)m != n || std::memcmp(a, b, n) != 0) {
      std::printf("fmt_u32 %u\n", v);
      return 1;
    }
  }
  for (int i = 0; i < 1000000; ++i) {
    const uint32_t v = (uint32_t)rng() >> (rng() % 32);
    char a[32], b[32];
    const size_t n = mdmc::fmt_u32(a, v);
    const int m = std::snprintf(b, sizeof b, "%u", v);
    if ((size_t)m != n || std::memcmp(a, b, n) != 0) {
      std::printf("fmt_u32 %u\n", v);
      return 1;
    }
  }
  std::printf("OK\n");
  return 0;
}
