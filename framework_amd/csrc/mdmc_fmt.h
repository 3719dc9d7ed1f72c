// Text of the OBJ writer's numbers, byte for byte what printf writes.
//
// fmt_g: a float as printf's "%g" (precision 6) writes the double it
// converts to.  The six significant digits are the value's exact decimal
// expansion rounded half to even, computed with 128-bit integers from the
// float's significand and binary exponent: for 1e-5 <= |v| < 1e15 the
// significand times a power of two and a power of ten fits in 128 bits.
// Then the C rule for %g: with X the decimal exponent of the rounded value,
// style f with 5 - X decimals when -4 <= X < 6, else style e with 5
// decimals; trailing zeros of the fraction dropped, and the point with them
// when no fraction is left.  Zero, subnormal, larger or smaller values, NaN
// and infinities go through snprintf.
//
// fmt_u32: an unsigned integer in decimal ("%u").
//
// Plain C++ (no HIP types) so that a host test can check it against snprintf.
#ifndef MDMC_FMT_H
#define MDMC_FMT_H

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

namespace mdmc {

// writes at most 16 bytes, no terminating zero; returns the length
inline size_t fmt_u32(char *o, uint32_t v) {
  char t[10];
  size_t n = 0;
  do {
    t[n++] = (char)('0' + v % 10u);
    v /= 10u;
  } while (v);
  for (size_t i = 0; i < n; ++i) o[i] = t[n - 1 - i];
  return n;
}

// writes at most 32 bytes, no terminating zero; returns the length
inline size_t fmt_g(char *o, float f) {
  typedef unsigned __int128 u128;
  const double a = std::fabs((double)f);
  if (!(a >= 1e-5 && a < 1e15)) {
    char t[48];
    const int n = std::snprintf(t, sizeof t, "%g", (double)f);
    std::memcpy(o, t, (size_t)n);
    return (size_t)n;
  }
  static const uint64_t kPow10[20] = {1ull,
                                      10ull,
                                      100ull,
                                      1000ull,
                                      10000ull,
                                      100000ull,
                                      1000000ull,
                                      10000000ull,
                                      100000000ull,
                                      1000000000ull,
                                      10000000000ull,
                                      100000000000ull,
                                      1000000000000ull,
                                      10000000000000ull,
                                      100000000000000ull,
                                      1000000000000000ull,
                                      10000000000000000ull,
                                      100000000000000000ull,
                                      1000000000000000000ull,
                                      10000000000000000000ull};
  uint32_t bits;
  std::memcpy(&bits, &f, sizeof bits);
  const uint64_t m = (bits & 0x7FFFFFu) | 0x800000u;  // normal: |f| >= 1e-5
  const int e = (int)((bits >> 23) & 0xFFu) - 150;    // |f| = m 2^e, -40 <= e <= 26
  // the fraction |f| / 10^k as num / den
  auto frac = [&](int k, u128 &num, u128 &den) {
    num = m;
    den = 1;
    if (e >= 0) num <<= e; else den <<= -e;
    if (k >= 0) den *= kPow10[k]; else num *= kPow10[-k];
  };
  // X = floor(log10 |f|), made exact: 10^X <= |f| < 10^(X+1)
  int X = (int)std::floor(std::log10(a));
  u128 num, den;
  frac(X, num, den);
  if (num < den) {
    --X;
  } else {
    frac(X + 1, num, den);
    if (num >= den) ++X;
  }
  // six significant digits, rounded half to even
  frac(X - 5, num, den);
  uint64_t q = (uint64_t)(num / den);
  const u128 r2 = (num - (u128)q * den) * 2;
  if (r2 > den || (r2 == den && (q & 1u))) ++q;
  if (q == 1000000u) {
    q = 100000u;
    ++X;
  }
  char dg[6];
  for (int i = 5; i >= 0; --i) {
    dg[i] = (char)('0' + q % 10u);
    q /= 10u;
  }
  int nd = 6;  // significant digits left after dropping trailing zeros
  while (nd > 1 && dg[nd - 1] == '0') --nd;
  size_t n = 0;
  if (bits >> 31) o[n++] = '-';
  if (X >= -4 && X < 6) {
    if (X >= 0) {
      for (int i = 0; i <= X; ++i) o[n++] = dg[i];
      if (nd > X + 1) {
        o[n++] = '.';
        for (int i = X + 1; i < nd; ++i) o[n++] = dg[i];
      }
    } else {
      o[n++] = '0';
      o[n++] = '.';
      for (int i = 0; i < -X - 1; ++i) o[n++] = '0';
      for (int i = 0; i < nd; ++i) o[n++] = dg[i];
    }
  } else {
    o[n++] = dg[0];
    if (nd > 1) {
      o[n++] = '.';
      for (int i = 1; i < nd; ++i) o[n++] = dg[i];
    }
    o[n++] = 'e';
    o[n++] = X < 0 ? '-' : '+';
    const int ax = X < 0 ? -X : X;
    if (ax < 10) o[n++] = '0';
    n += fmt_u32(o + n, (uint32_t)ax);
  }
  return n;
}

}  // namespace mdmc

#endif
