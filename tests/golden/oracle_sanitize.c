/* Driver for a sanitizer build of the oracle (oracle/dmc_oracle.c, built with
 * -fsanitize=address,undefined): extractions over the cases the tests use --
 * sphere, blobs, gyroid with noise, integer ties, NaN and infinite samples,
 * degenerate and tiny grids, a world box, smoothing, the FP64 twin and the
 * window mode -- each checked for the oracle's own invariant counter and for
 * quad indices inside the vertex array.  Prints OK and exits 0 when every run
 * is clean (a sanitizer report aborts the program). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dmc_oracle.h"

static int failures = 0;

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static double urand(void) {
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static void check(const char *name, const float *f, const uint32_t d[3], float iso, uint32_t smooth, int fp64,
                  const uint32_t *grid_cells, const uint32_t *offset) {
  oracle_result r;
  const float mn[3] = {-1.5f, 0.0f, 2.0f}, mx[3] = {3.0f, 7.0f, 5.5f};
  int rc = oracle_dmc_run_window(f, d, grid_cells, offset, mn, mx, iso, fp64, smooth, &r);
  uint64_t q;
  if (rc != 0) {
    printf("FAIL %s: rc %d\n", name, rc);
    ++failures;
    return;
  }
  if (r.num_anomalies) {
    printf("FAIL %s: %u anomalies\n", name, r.num_anomalies);
    ++failures;
  }
  for (q = 0; q < 4ull * r.num_quads; ++q)
    if (r.quads[q] < 0 || (uint32_t)r.quads[q] >= r.num_vertices) {
      printf("FAIL %s: quad entry %lld out of range\n", name, (long long)r.quads[q]);
      ++failures;
      break;
    }
  oracle_result_free(&r);
}

int main(void) {
  uint32_t d[3] = {40, 33, 29};
  size_t n = (size_t)d[0] * d[1] * d[2], i;
  float *f = (float *)malloc(n * sizeof(float));
  double c[3] = {19.5, 16.0, 14.2};
  uint32_t gc[3] = {90, 80, 70}, off[3] = {11, 7, 3};
  uint32_t tiny[][3] = {{2, 2, 2}, {3, 2, 2}, {2, 5, 3}, {1, 4, 4}, {4, 1, 4}, {5, 5, 1}, {65, 2, 2}};
  oracle_set_threads(2);
  oracle_field_sphere(f, d, c, 11.0);
  check("sphere", f, d, 0.0f, 0, 1, NULL, NULL);
  check("sphere smoothed", f, d, 0.0f, 3, 0, NULL, NULL);
  check("sphere window", f, d, 0.0f, 1, 1, gc, off);
  oracle_field_blobs(f, d, 3, 12, 3.0, 10.0, -1.0, 1.0);
  check("blobs", f, d, 0.1f, 0, 1, NULL, NULL);
  oracle_field_gyroid(f, d, 8.0, 5, 0.8);
  check("gyroid", f, d, 0.0f, 2, 0, NULL, NULL);
  for (i = 0; i < n; ++i) f[i] = (float)((int)(urand() * 5.0) - 2);
  check("ties", f, d, 0.0f, 1, 0, NULL, NULL);
  for (i = 0; i < n; ++i) {
    const double u = urand();
    f[i] = u < 0.02 ? NAN : (u < 0.03 ? INFINITY : (u < 0.04 ? -INFINITY : (float)(urand() * 2.0 - 1.0)));
  }
  check("nan and inf", f, d, 0.1f, 1, 1, NULL, NULL);
  for (i = 0; i < sizeof(tiny) / sizeof(tiny[0]); ++i) {
    size_t m = (size_t)tiny[i][0] * tiny[i][1] * tiny[i][2], k;
    float *g = (float *)malloc((m ? m : 1) * sizeof(float));
    for (k = 0; k < m; ++k) g[k] = (float)(urand() * 2.0 - 1.0);
    check("tiny", g, tiny[i], 0.0f, 1, 1, NULL, NULL);
    free(g);
  }
  free(f);
  if (failures) return 1;
  printf("OK\n");
  return 0;
}
