/*
 * Synthetic scalar fields used by the benchmark and test workloads.
 *
 * The reference ships no datasets; its only harness builds analytic fields on
 * the host (dmc.cu:2211-2287, a sphere and a gyroid).  These generators define
 * the workloads exactly so that the GPU and the CPU oracle see identical
 * float bits:
 *
 *   - random numbers are SplitMix64 outputs addressed by (seed, index), so
 *     every value is reproducible without sequential state;
 *   - Gaussian blobs are evaluated in separable form
 *       f(x,y,z) = sum_b ((a_b * ex_b[x]) * ey_b[y]) * ez_b[z]
 *     with 1-D factors ex_b[x] = exp(-(x-c_b)^2 / (2 sigma_b^2)) computed once
 *     on the host in double precision; blob by blob, in order, the product
 *     (a_b * ex_b[x]) * ey_b[y] is rounded and then multiplied by ez_b[z] and
 *     added to the sum in one fused multiply-add (C99 fma, the GPU's FMA:
 *     both correctly rounded), then the sum is rounded to float;
 *   - the gyroid is sx[x]*cy[y] + sy[y]*cz[z] + sz[z]*cx[x] plus hash noise,
 *     again from host-computed 1-D sin/cos tables and fixed-order double math.
 *
 * Plain C99: included by the HIP library's host code and by the oracle.
 */
#ifndef MDMC_FIELDS_H
#define MDMC_FIELDS_H

#include <math.h>
#include <stdint.h>

#define MDMC_GOLDEN_GAMMA 0x9E3779B97F4A7C15ull

static inline uint64_t mdmc_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Output number i (0-based) of a SplitMix64 generator started at `seed`. */
static inline uint64_t mdmc_rng_u64(uint64_t seed, uint64_t i) {
  return mdmc_mix64(seed + (i + 1) * MDMC_GOLDEN_GAMMA);
}

/* Uniform double in [0, 1) with 53 random bits. */
static inline double mdmc_rng_unit(uint64_t seed, uint64_t i) {
  return (double)(mdmc_rng_u64(seed, i) >> 11) * (1.0 / 9007199254740992.0);
}

/* Parameters of blob b: centre uniform in the grid box [0, dim) per axis,
 * sigma uniform in [sigma_lo, sigma_hi], amplitude uniform in [amp_lo, amp_hi].
 * Random stream slots 6b+0 .. 6b+4 are used (slot 6b+5 is unused). */
typedef struct mdmc_blob {
  double cx, cy, cz, sigma, amp;
} mdmc_blob;

static inline mdmc_blob mdmc_blob_params(uint64_t seed, uint32_t b, const uint32_t dims[3],
                                         double sigma_lo, double sigma_hi, double amp_lo,
                                         double amp_hi) {
  mdmc_blob p;
  uint64_t s = 6ull * b;
  p.cx = mdmc_rng_unit(seed, s + 0) * (double)dims[0];
  p.cy = mdmc_rng_unit(seed, s + 1) * (double)dims[1];
  p.cz = mdmc_rng_unit(seed, s + 2) * (double)dims[2];
  p.sigma = sigma_lo + (sigma_hi - sigma_lo) * mdmc_rng_unit(seed, s + 3);
  p.amp = amp_lo + (amp_hi - amp_lo) * mdmc_rng_unit(seed, s + 4);
  return p;
}

/* 1-D Gaussian factor exp(-(i - c)^2 / (2 sigma^2)) for i in [0, n). */
static inline void mdmc_gauss_1d(double *out, uint32_t n, double c, double sigma) {
  uint32_t i;
  double denom = 2.0 * sigma * sigma;
  for (i = 0; i < n; ++i) {
    double d = (double)i - c;
    out[i] = exp(-(d * d) / denom);
  }
}

/* Hash noise in [-amp, amp) for point index `idx`. */
static inline double mdmc_noise(uint64_t seed, uint64_t idx, double amp) {
  return amp * (2.0 * mdmc_rng_unit(seed, idx) - 1.0);
}

/* 1-D sin/cos tables of w*i, w = 2*pi/period. */
static inline void mdmc_trig_1d(double *s, double *c, uint32_t n, double period) {
  uint32_t i;
  double w = 6.283185307179586476925286766559 / period;
  for (i = 0; i < n; ++i) {
    s[i] = sin(w * (double)i);
    c[i] = cos(w * (double)i);
  }
}

#endif /* MDMC_FIELDS_H */
