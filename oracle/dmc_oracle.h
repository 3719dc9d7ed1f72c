/*
 * CPU oracle for the Dual Marching Cubes extraction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (DMC_CUDA/DMC_CUDA/src/dmc.cu, dmc::run_dmc at :2069-2208), kept
 * pass-for-pass so that the HIP library can be checked against it.  Only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it; the product never calls it.
 */
#ifndef MDMC_ORACLE_H
#define MDMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_result {
  uint32_t num_active;   /* active cells (0 < cfg < 255) */
  uint32_t num_vertices; /* iso vertices + owned edge points */
  uint32_t num_quads;
  uint32_t num_tie_edges;      /* owned edges bipolar under the reference
                                  predicate but not under the sign rule */
  uint32_t num_unset_vertices; /* iso vertices with no in-grid edge */
  uint32_t num_anomalies;      /* broken invariants (should stay 0) */
  float *vertices;             /* 3 * num_vertices, x y z */
  double *vertices64;          /* FP64 twin, or NULL */
  int32_t *quads;              /* 4 * num_quads */
  uint32_t *cell_index;        /* num_active: cell index1D, ascending */
  uint8_t *cfg;                /* num_active */
  uint8_t *info;               /* num_active: bits 0-2 e6/e9/e10 bipolar,
                                  3-5 their ccw bit, 7 use variant-1 table */
  uint32_t *vertex_begin;      /* num_active */
  uint32_t *quad_begin;        /* num_active */
} oracle_result;

/* dims are point counts (x fastest); cells are dims - 1 per axis.
 * want_fp64 != 0 also fills vertices64.  Returns 0 on success. */
int oracle_dmc_run(const float *field, const uint32_t dims[3], const float xyz_min[3],
                   const float xyz_max[3], float iso, int want_fp64, oracle_result *out);
/* The same with num_smooth smoothing iterations (the reference's num_smooth,
 * dmc.cu:2157-2169); vertices64 is not smoothed. */
int oracle_dmc_run_smooth(const float *field, const uint32_t dims[3], const float xyz_min[3],
                          const float xyz_max[3], float iso, int want_fp64, uint32_t num_smooth,
                          oracle_result *out);
/* A window of a larger grid: `field`/`dims` hold the window, which starts at
 * point `offset` of a grid of `grid_cells` cells per axis; coordinates are
 * those of the whole grid, ijk_to_xyz(offset + i, grid_cells, range, min).
 * Configurations, variants and vertices of cells off the window's faces equal
 * the whole grid's (cells on a face may see fewer neighbours).  NULL
 * grid_cells / offset: the window is the grid. */
int oracle_dmc_run_window(const float *field, const uint32_t dims[3], const uint32_t grid_cells[3],
                          const uint32_t offset[3], const float xyz_min[3], const float xyz_max[3], float iso,
                          int want_fp64, uint32_t num_smooth, oracle_result *out);
void oracle_result_free(oracle_result *r);
/* Threads for the per-cell loops that are independent (the compaction and
 * prefix sums stay in cell order, so results do not depend on it); 1 by
 * default. */
void oracle_set_threads(int n);

/* The reference tables rebuilt by the shared builder, for hashing in tests:
 * edge_label[2][256][12], num_vertex[2][256], amb_face[256] (7168 bytes). */
void oracle_luts_bytes(uint8_t *out7168);

/* Field generators (same definitions as the library's device generators). */
void oracle_field_sphere(float *out, const uint32_t dims[3], const double centre[3],
                         double radius);
void oracle_field_blobs(float *out, const uint32_t dims[3], uint64_t seed, uint32_t nblobs,
                        double sigma_lo, double sigma_hi, double amp_lo, double amp_hi);
void oracle_field_gyroid(float *out, const uint32_t dims[3], double period, uint64_t seed,
                         double noise_amp);

#ifdef __cplusplus
}
#endif
#endif
