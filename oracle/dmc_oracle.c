/*
 * CPU oracle for the Dual Marching Cubes extraction path.
 *
 * TEST INFRASTRUCTURE ONLY (see dmc_oracle.h).  Each function restates one
 * pass of the reference dmc::run_dmc (DMC_CUDA/DMC_CUDA/src/dmc.cu) and cites
 * the lines it follows.  Passes are kept separate and in the reference order
 * (dense flags -> scan -> compaction -> per-active passes through a dense
 * cell->compact map) so that the oracle shares no structure with the HIP
 * kernels it checks.
 *
 * Build: plain C99, -ffp-contract=off, no fast-math, so every float + - * /
 * is one correctly rounded IEEE operation (the sequence the kernels follow).
 *
 * Defined behaviour where the reference is undefined:
 *   ties     an owned edge is bipolar iff (v0 < iso) != (v1 < iso).  The
 *            reference test (dmc.cu:1101-1106) also fires when one endpoint
 *            equals iso and the other is above it; those edges then index the
 *            edge table at a non-bipolar entry (0xff).  They are counted in
 *            num_tie_edges and treated as not bipolar.
 *   NaN      a NaN corner is "not < iso" in the configuration and in the
 *            bipolar rule above.
 *   empty    no active cell gives an empty mesh (the reference launches a
 *            zero-block grid and exits, dmc.cu:2113-2114).
 *   unset    an iso vertex whose edges all have owners outside the grid keeps
 *            uninitialised memory in the reference (dmc.cu:1588-1591); here it
 *            is set to a quiet NaN (0x7fc00000) in each component.
 */
#include "dmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../framework_amd/csrc/dmc_luts.h"
#include "../framework_amd/csrc/mdmc_fields.h"

#define INVALID_U32 0xFFFFFFFFu

static mdmc_luts g_luts;
static int g_luts_ready = 0;
static int g_threads = 1; /* OpenMP threads of the independent per-cell loops */

void oracle_set_threads(int n) { g_threads = n > 0 ? n : 1; }

static const mdmc_luts *luts(void) {
  if (!g_luts_ready) {
    mdmc_build_luts(&g_luts);
    g_luts_ready = 1;
  }
  return &g_luts;
}

void oracle_luts_bytes(uint8_t *out) {
  const mdmc_luts *L = luts();
  memcpy(out, L->edge_label, sizeof(L->edge_label));
  memcpy(out + sizeof(L->edge_label), L->num_vertex, sizeof(L->num_vertex));
  memcpy(out + sizeof(L->edge_label) + sizeof(L->num_vertex), L->amb_face, sizeof(L->amb_face));
}

/* Owner of each cell edge: offset to the owning cell and the owner's local
 * edge (6, 9 or 10); local edges are owned by the cell itself
 * (edge_belonged_voxel_lut, dmc.cu:634-649). */
static const int EDGE_OWNER_D[12][3] = {
    {0, -1, -1}, {0, 0, -1}, {0, 0, -1}, {-1, 0, -1}, {-1, -1, 0}, {0, -1, 0},
    {0, 0, 0},   {-1, 0, 0}, {0, -1, 0}, {0, 0, 0},   {0, 0, 0},   {-1, 0, 0}};
static const int EDGE_OWNER_E[12] = {10, 9, 10, 9, 6, 6, 6, 6, 10, 9, 10, 9};

/* Four cells around an owned edge, counter-clockwise: ring position p holds
 * the cell at offset RING_D[l][p] and that cell's edge RING_E[l][p], for l =
 * 0,1,2 <-> edge 6, 9, 10 (circular_edge_lut dmc.cu:652-658 with the negated
 * owner offsets of CircularEdgeIterator::retrieve, dmc.cu:1631-1666). */
static const int RING_E[3][4] = {{6, 7, 4, 5}, {9, 1, 3, 11}, {10, 8, 0, 2}};
static const int LOCAL_EDGES[3] = {6, 9, 10};

typedef struct cellrec {
  uint32_t index1d;
  uint32_t vertex_begin;
  uint8_t cfg;
  uint8_t info;
  uint8_t num_vertices;
} cellrec;

/* utils::ijk_to_xyz (include/utils.hpp:118-122). */
static float ijk_to_xyz(uint32_t i, uint32_t size, float range, float mn) {
  return (float)i / (float)size * range + mn;
}

static double ijk_to_xyz64(uint32_t i, uint32_t size, double range, double mn) {
  return (double)i / (double)size * range + mn;
}

/* dmc.cu:1090-1098 lerp_float3, one component. */
static float lerp1(float p0, float p1, float t, float omt) { return p0 * omt + p1 * t; }

static double lerp1d(double p0, double p1, double t, double omt) { return p0 * omt + p1 * t; }

/* Field value at point (x, y, z). */
static float F(const float *f, uint32_t X, uint32_t Y, uint32_t x, uint32_t y, uint32_t z) {
  return f[((size_t)z * Y + y) * X + x];
}

/* voxel_config_mask (dmc.cu:1018-1027) over the 8 corners of cell (i,j,k). */
static void cell_values(const float *f, uint32_t X, uint32_t Y, uint32_t i, uint32_t j,
                        uint32_t k, float v[8]) {
  int c;
  for (c = 0; c < 8; ++c)
    v[c] = F(f, X, Y, i + MDMC_CORNER_D[c][0], j + MDMC_CORNER_D[c][1], k + MDMC_CORNER_D[c][2]);
}

static uint8_t config_of(const float v[8], float iso) {
  uint8_t m = 0;
  int c;
  for (c = 0; c < 8; ++c) m |= (uint8_t)((v[c] < iso) << c);
  return m;
}

/* Reference bipolar predicate, is_edge_bipolar (dmc.cu:1101-1106). */
static int ref_bipolar(float a, float b, float iso) {
  float lo, hi;
  if (a == b) return 0;
  if (a > b) { lo = b; hi = a; } else { lo = a; hi = b; }
  return !((lo < iso && hi < iso) || (lo > iso && hi > iso));
}

static int edge_shift(int e) { return e == 6 ? 0 : (e == 9 ? 1 : 2); }

/* _VoxelInfo::edge_vertex_index (dmc.cu:911-921): the owned edge points follow
 * the iso vertices, in e6, e9, e10 order. */
static uint32_t edge_vertex_index(const cellrec *r, const mdmc_luts *L, int e) {
  int s = edge_shift(e), b, off = 0;
  int lut2 = (r->info >> 7) & 1;
  for (b = 0; b < s; ++b) off += (r->info >> b) & 1;
  return r->vertex_begin + L->num_vertex[lut2][r->cfg] + (uint32_t)off;
}

/* Pass 6, calc_iso_vertices_kern (dmc.cu:1517-1593): each iso vertex is the
 * mean of the edge points of its group, read back from the owners' slots in
 * the vertex array (the stored points: after smoothing, the moved ones);
 * edges whose owner lies outside the grid are skipped. */
static void reaverage(oracle_result *out, const cellrec *rec, const uint32_t *map, uint32_t A, uint32_t cx,
                      uint32_t cy, uint32_t cz, const mdmc_luts *L, int want_fp64, int first) {
  int64_t a;
  uint32_t anomalies = 0, unset = 0;
  (void)cz;
  /* independent per cell: reads owners' edge-point slots, writes its own
   * iso-vertex slots */
#pragma omp parallel for schedule(static) num_threads(g_threads) reduction(+ : anomalies, unset)
  for (a = 0; a < (int64_t)A; ++a) {
    uint32_t id = rec[a].index1d;
    int64_t i = id % cx, j = (id / cx) % cy, k = id / (cx * cy);
    int lut2 = (rec[a].info >> 7) & 1, e, m;
    uint8_t cnt[4] = {0, 0, 0, 0};
    int nv = L->num_vertex[lut2][rec[a].cfg];
    for (e = 0; e < 12; ++e) {
      uint8_t lab = L->edge_label[lut2][rec[a].cfg][e];
      int64_t oi, oj, ok;
      uint32_t om, src, dst;
      if (lab == MDMC_NO_VERTEX) continue;
      oi = i + EDGE_OWNER_D[e][0]; oj = j + EDGE_OWNER_D[e][1]; ok = k + EDGE_OWNER_D[e][2];
      if (oi < 0 || oj < 0 || ok < 0) continue;
      om = map[((size_t)ok * cy + (size_t)oj) * cx + (size_t)oi];
      if (om == INVALID_U32 || !((rec[om].info >> edge_shift(EDGE_OWNER_E[e])) & 1)) {
        if (first) anomalies++;
        continue;
      }
      src = edge_vertex_index(&rec[om], L, EDGE_OWNER_E[e]);
      dst = rec[a].vertex_begin + lab;
      if (cnt[lab] == 0) {
        out->vertices[3 * (size_t)dst + 0] = out->vertices[3 * (size_t)src + 0];
        out->vertices[3 * (size_t)dst + 1] = out->vertices[3 * (size_t)src + 1];
        out->vertices[3 * (size_t)dst + 2] = out->vertices[3 * (size_t)src + 2];
        if (want_fp64) {
          out->vertices64[3 * (size_t)dst + 0] = out->vertices64[3 * (size_t)src + 0];
          out->vertices64[3 * (size_t)dst + 1] = out->vertices64[3 * (size_t)src + 1];
          out->vertices64[3 * (size_t)dst + 2] = out->vertices64[3 * (size_t)src + 2];
        }
      } else {
        out->vertices[3 * (size_t)dst + 0] += out->vertices[3 * (size_t)src + 0];
        out->vertices[3 * (size_t)dst + 1] += out->vertices[3 * (size_t)src + 1];
        out->vertices[3 * (size_t)dst + 2] += out->vertices[3 * (size_t)src + 2];
        if (want_fp64) {
          out->vertices64[3 * (size_t)dst + 0] += out->vertices64[3 * (size_t)src + 0];
          out->vertices64[3 * (size_t)dst + 1] += out->vertices64[3 * (size_t)src + 1];
          out->vertices64[3 * (size_t)dst + 2] += out->vertices64[3 * (size_t)src + 2];
        }
      }
      ++cnt[lab];
    }
    for (m = 0; m < nv; ++m) {
      uint32_t dst = rec[a].vertex_begin + (uint32_t)m;
      int c;
      if (cnt[m]) {
        for (c = 0; c < 3; ++c) {
          out->vertices[3 * (size_t)dst + c] /= (float)cnt[m];
          if (want_fp64) out->vertices64[3 * (size_t)dst + c] /= (double)cnt[m];
        }
      } else {
        union { uint32_t u; float f; } qnan;
        qnan.u = 0x7FC00000u;
        if (first) unset++;
        for (c = 0; c < 3; ++c) {
          out->vertices[3 * (size_t)dst + c] = qnan.f;
          if (want_fp64) out->vertices64[3 * (size_t)dst + c] = (double)qnan.f;
        }
      }
    }
  }
  out->num_anomalies += anomalies;
  out->num_unset_vertices += unset;
}

/* ---- smoothing (num_smooth > 0): smooth_edge_vertices (dmc.cu:1868-1981)
 * followed by calc_iso_vertices_kern again, num_smooth times
 * (dmc.cu:2157-2169).  For each owned bipolar edge whose ring lies in the
 * grid, the ring's four iso vertices (the quad, in its ccw/cw order) are
 * projected on the plane across the edge (e6: x,y; e9: x,z; e10: y,z); if
 * the quadrilateral is not self-crossing, it is split at a diagonal and the
 * edge point moves to the barycentric blend of the triangle that contains
 * the edge's trace, when one does.
 *
 * Defined here where the reference depends on CUDA approximations: the split
 * of a convex quadrilateral is at the corner of the largest angle, found as
 * the smallest cosine (a.b) / (sqrt(a.a) sqrt(b.b)) with IEEE sqrt and
 * division (first smallest wins), instead of acosf of rsqrtf-normalised
 * vectors (dmc.cu:1818-1860); every other operation is the reference's, one
 * correctly rounded float operation each, no contraction. */

typedef struct pt2 {
  float x, y;
} pt2;

/* calc_cross_z_sign (dmc.cu:1741-1747): +1 when the z of
 * (right - mid) x (left - mid) is >= 0. */
static int cross_sign(pt2 l, pt2 m, pt2 r) {
  float dx1 = r.x - m.x, dy1 = r.y - m.y;
  float dx2 = l.x - m.x, dy2 = l.y - m.y;
  float cz = dx1 * dy2 - dx2 * dy1;
  return cz >= 0.0f ? 1 : -1;
}

/* cosine of the corner angle at m */
static float corner_cos(pt2 l, pt2 m, pt2 r) {
  float ax = l.x - m.x, ay = l.y - m.y, bx = r.x - m.x, by = r.y - m.y;
  float num = ax * bx + ay * by;
  float la = sqrtf(ax * ax + ay * ay), lb = sqrtf(bx * bx + by * by);
  return num / (la * lb);
}

/* is_inside_triangle (dmc.cu:1894-1912): barycentric (alpha, beta, gamma)
 * of t in (p0, p1, p2), inside with a 1e-4 margin (double comparisons). */
static int inside_tri(pt2 p0, pt2 p1, pt2 p2, pt2 t, float *al, float *be, float *ga) {
  float v0x = p1.x - p0.x, v0y = p1.y - p0.y;
  float v1x = p2.x - p0.x, v1y = p2.y - p0.y;
  float v2x = t.x - p0.x, v2y = t.y - p0.y;
  float d00 = v0x * v0x + v0y * v0y;
  float d10 = v1x * v0x + v1y * v0y;
  float d11 = v1x * v1x + v1y * v1y;
  float d20 = v2x * v0x + v2y * v0y;
  float d21 = v2x * v1x + v2y * v1y;
  float inv = d00 * d11 - d10 * d10;
  float beta, gamma;
  inv = 1.0f / inv;
  beta = (d11 * d20 - d10 * d21) * inv;
  gamma = (d00 * d21 - d10 * d20) * inv;
  *be = beta;
  *ga = gamma;
  *al = 1.0f - beta - gamma;
  return (-1e-4 < (double)beta) && (-1e-4 < (double)gamma) && ((double)(beta + gamma) < 1.0 + 1e-4);
}

static void smooth_pass(oracle_result *out, const cellrec *rec, uint32_t A, uint32_t cx, uint32_t cy,
                        uint32_t cz, const mdmc_luts *L, const float xyz_min[3], const float range[3],
                        const uint32_t div[3], const uint32_t off[3]) {
  uint32_t a;
  float *V = out->vertices;
  for (a = 0; a < A; ++a) {
    uint32_t id = rec[a].index1d;
    uint32_t i = id % cx, j = (id / cx) % cy, k = id / (cx * cy);
    uint32_t q = out->quad_begin[a];
    int l;
    for (l = 0; l < 3; ++l) {
      int e = LOCAL_EDGES[l], c0 = l == 2 ? 1 : 0, c1 = l == 0 ? 1 : 2, p, npos = 0, nneg = 0, ipos = 0, ineg = 0;
      int s0, s1;
      const int32_t *iv;
      pt2 P[4], o;
      float al, be, ga;
      float *E;
      if (!((rec[a].info >> l) & 1)) continue;
      if (e == 6 && !(i + 1 < cx && j + 1 < cy)) continue;
      if (e == 9 && !(i + 1 < cx && k + 1 < cz)) continue;
      if (e == 10 && !(j + 1 < cy && k + 1 < cz)) continue;
      iv = out->quads + 4 * (size_t)q++;
      for (p = 0; p < 4; ++p) {
        P[p].x = V[3 * (size_t)iv[p] + c0];
        P[p].y = V[3 * (size_t)iv[p] + c1];
      }
      /* calc_quadrilateral_signs (dmc.cu:1749-1780) */
      for (p = 0; p < 4; ++p) {
        if (cross_sign(P[(p + 3) % 4], P[p], P[(p + 1) % 4]) > 0) { ++npos; ipos = p; }
        else { ++nneg; ineg = p; }
      }
      if (npos == nneg) continue; /* self-crossing (dmc.cu:1783-1786) */
      if (npos == 0 || nneg == 0) {
        float best = corner_cos(P[3], P[0], P[1]);
        s0 = 0;
        for (p = 1; p < 4; ++p) {
          float c = corner_cos(P[(p + 3) % 4], P[p], P[(p + 1) % 4]);
          if (c < best) { best = c; s0 = p; }
        }
      } else {
        s0 = npos < nneg ? ipos : ineg; /* the concave corner */
      }
      s1 = (s0 + 2) % 4;
      o.x = l == 2 ? ijk_to_xyz(off[1] + j + 1, div[1], range[1], xyz_min[1])
                   : ijk_to_xyz(off[0] + i + 1, div[0], range[0], xyz_min[0]);
      o.y = l == 0 ? ijk_to_xyz(off[1] + j + 1, div[1], range[1], xyz_min[1])
                   : ijk_to_xyz(off[2] + k + 1, div[2], range[2], xyz_min[2]);
      E = V + 3 * (size_t)edge_vertex_index(&rec[a], L, e);
      if (inside_tri(P[s0], P[(s0 + 1) % 4], P[s1], o, &al, &be, &ga)) {
        const float *va = V + 3 * (size_t)iv[s0], *vb = V + 3 * (size_t)iv[(s0 + 1) % 4], *vc = V + 3 * (size_t)iv[s1];
        int c;
        for (c = 0; c < 3; ++c) E[c] = al * va[c] + be * vb[c] + ga * vc[c];
      } else if (inside_tri(P[s1], P[(s1 + 1) % 4], P[s0], o, &al, &be, &ga)) {
        const float *va = V + 3 * (size_t)iv[s1], *vb = V + 3 * (size_t)iv[(s1 + 1) % 4], *vc = V + 3 * (size_t)iv[s0];
        int c;
        for (c = 0; c < 3; ++c) E[c] = al * va[c] + be * vb[c] + ga * vc[c];
      }
    }
  }
}

int oracle_dmc_run_smooth(const float *field, const uint32_t dims[3], const float xyz_min[3],
                          const float xyz_max[3], float iso, int want_fp64, uint32_t num_smooth,
                          oracle_result *out) {
  return oracle_dmc_run_window(field, dims, NULL, NULL, xyz_min, xyz_max, iso, want_fp64, num_smooth, out);
}

int oracle_dmc_run_window(const float *field, const uint32_t dims[3], const uint32_t grid_cells[3],
                          const uint32_t offset[3], const float xyz_min[3], const float xyz_max[3], float iso,
                          int want_fp64, uint32_t num_smooth, oracle_result *out) {
  const mdmc_luts *L = luts();
  uint32_t X = dims[0], Y = dims[1], Z = dims[2];
  uint32_t cx, cy, cz;
  size_t C, idx;
  uint32_t *map = NULL, *plane_base = NULL;
  cellrec *rec = NULL;
  uint32_t A = 0, V = 0, Q = 0, a, ties = 0, anomalies = 0;
  int64_t kk, aa;
  float range[3];
  double range64[3];
  uint32_t div[3], off[3];

  memset(out, 0, sizeof(*out));
  if (X < 2 || Y < 2 || Z < 2) return 0; /* no cells: empty mesh */
  cx = X - 1; cy = Y - 1; cz = Z - 1;
  C = (size_t)cx * cy * cz;
  /* coordinates: index + offset over the grid's cells per axis (a window of
   * a larger grid takes that grid's divisors and its own position in it) */
  div[0] = grid_cells ? grid_cells[0] : cx;
  div[1] = grid_cells ? grid_cells[1] : cy;
  div[2] = grid_cells ? grid_cells[2] : cz;
  off[0] = offset ? offset[0] : 0;
  off[1] = offset ? offset[1] : 0;
  off[2] = offset ? offset[2] : 0;
  /* xyz_range = xyz_max - xyz_min in float (dmc.cu:2149) */
  range[0] = xyz_max[0] - xyz_min[0];
  range[1] = xyz_max[1] - xyz_min[1];
  range[2] = xyz_max[2] - xyz_min[2];
  range64[0] = (double)xyz_max[0] - (double)xyz_min[0];
  range64[1] = (double)xyz_max[1] - (double)xyz_min[1];
  range64[2] = (double)xyz_max[2] - (double)xyz_min[2];

  /* Pass 1: flag_active_voxels_kern (dmc.cu:1035-1061), launch_thrust_count +
   * exclusive scan (dmc.cu:2092-2098) and compact_voxel_flags_kern
   * (dmc.cu:1134-1158): a dense cell -> compact index map. */
  map = (uint32_t *)malloc(C * sizeof(uint32_t));
  plane_base = (uint32_t *)malloc(((size_t)cz + 1) * sizeof(uint32_t));
  if (!map || !plane_base) { free(map); free(plane_base); return -1; }
  /* flags per cell and active counts per cell plane (planes in parallel),
   * then the planes' exclusive prefix, then the compact numbers in cell
   * order: the same map as one sequential count */
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (kk = 0; kk < (int64_t)cz; ++kk) {
    uint32_t n = 0, i, j, k = (uint32_t)kk;
    for (j = 0; j < cy; ++j)
      for (i = 0; i < cx; ++i) {
        float v[8];
        uint8_t cfg;
        size_t id = ((size_t)k * cy + j) * cx + i;
        cell_values(field, X, Y, i, j, k, v);
        cfg = config_of(v, iso);
        if (cfg != 0 && cfg != 0xFF) { map[id] = 0; ++n; }
        else map[id] = INVALID_U32;
      }
    plane_base[k] = n;
  }
  for (idx = 0; idx < cz; ++idx) {
    uint32_t n = plane_base[idx];
    plane_base[idx] = A;
    A += n;
  }
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (kk = 0; kk < (int64_t)cz; ++kk) {
    uint32_t next = plane_base[kk];
    size_t id = (size_t)kk * cy * cx, end = id + (size_t)cy * cx;
    for (; id < end; ++id)
      if (map[id] != INVALID_U32) map[id] = next++;
  }
  free(plane_base);
  out->num_active = A;
  if (A == 0) { free(map); return 0; }
  rec = (cellrec *)calloc(A, sizeof(cellrec));
  out->cell_index = (uint32_t *)malloc(A * sizeof(uint32_t));
  out->cfg = (uint8_t *)malloc(A);
  out->info = (uint8_t *)malloc(A);
  out->vertex_begin = (uint32_t *)malloc(A * sizeof(uint32_t));
  out->quad_begin = (uint32_t *)malloc(A * sizeof(uint32_t));
  if (!rec || !out->cell_index || !out->cfg || !out->info || !out->vertex_begin || !out->quad_begin) {
    free(map); free(rec); oracle_result_free(out); return -1;
  }
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (kk = 0; kk < (int64_t)C; ++kk)
    if (map[kk] != INVALID_U32) rec[map[kk]].index1d = (uint32_t)kk;

  /* Pass 2: init_voxels_info_kern (dmc.cu:1173-1221): config plus bipolar and
   * ccw bits of the owned edges 6 (p2,p6), 9 (p5,p6), 10 (p7,p6). */
#pragma omp parallel for schedule(static) num_threads(g_threads) reduction(+ : ties)
  for (aa = 0; aa < (int64_t)A; ++aa) {
    uint32_t a = (uint32_t)aa;
    uint32_t id = rec[a].index1d;
    uint32_t i = id % cx, j = (id / cx) % cy, k = id / (cx * cy);
    float v[8];
    static const int first[3] = {2, 5, 7};
    int l;
    cell_values(field, X, Y, i, j, k, v);
    rec[a].cfg = config_of(v, iso);
    rec[a].info = 0;
    for (l = 0; l < 3; ++l) {
      float v0 = v[first[l]], v1 = v[6];
      int bip = (v0 < iso) != (v1 < iso);
      if (ref_bipolar(v0, v1, iso) && !bip) ties++;
      if (bip) {
        rec[a].info |= (uint8_t)(1u << l);
        if (v0 <= iso) rec[a].info |= (uint8_t)(1u << (l + 3));
      }
    }
  }
  out->num_tie_edges += ties;

  /* Pass 3: correct_voxels_info_kern (dmc.cu:1345-1371) with
   * is_adjacent_ambiguous_config (dmc.cu:1301-1340): a 2B/3B cell whose
   * in-grid neighbour across its ambiguous face is also 2B/3B uses the
   * variant-1 table.  Only this cell is marked (dmc.cu:1368-1369). */
#pragma omp parallel for schedule(static) num_threads(g_threads) reduction(+ : anomalies)
  for (aa = 0; aa < (int64_t)A; ++aa) {
    uint32_t a = (uint32_t)aa;
    uint8_t face = L->amb_face[rec[a].cfg];
    uint32_t id = rec[a].index1d;
    int64_t i = id % cx, j = (id / cx) % cy, k = id / (cx * cy);
    int dir;
    uint32_t nmap;
    if (face == MDMC_NOT_AMBIGUOUS) continue;
    dir = MDMC_FACE_DIR[face];
    /* will_exceed_boundary / get_index3D_by_dir (dmc.cu:1250-1296) */
    switch (dir) {
      case 0: i += 1; break;
      case 1: i -= 1; break;
      case 2: j += 1; break;
      case 3: j -= 1; break;
      case 4: k += 1; break;
      default: k -= 1; break;
    }
    if (i < 0 || j < 0 || k < 0 || i >= cx || j >= cy || k >= cz) continue;
    nmap = map[((size_t)k * cy + (size_t)j) * cx + (size_t)i];
    if (nmap == INVALID_U32) { anomalies++; continue; }
    if (L->amb_face[rec[nmap].cfg] != MDMC_NOT_AMBIGUOUS) rec[a].info |= 0x80;
  }

  /* Pass 4: calc_num_vertices_per_voxel_kern (dmc.cu:1382-1406) and the
   * vertex_begin exclusive scan (dmc.cu:2133-2143). */
  for (a = 0; a < A; ++a) {
    int lut2 = (rec[a].info >> 7) & 1;
    uint32_t nv = L->num_vertex[lut2][rec[a].cfg] + (rec[a].info & 1) + ((rec[a].info >> 1) & 1) +
                  ((rec[a].info >> 2) & 1);
    rec[a].num_vertices = (uint8_t)nv;
    rec[a].vertex_begin = V;
    V += nv;
  }
  out->num_vertices = V;
  out->vertices = (float *)malloc((size_t)V * 3 * sizeof(float));
  if (want_fp64) out->vertices64 = (double *)malloc((size_t)V * 3 * sizeof(double));
  if (!out->vertices || (want_fp64 && !out->vertices64)) { free(map); free(rec); oracle_result_free(out); return -1; }

  /* Pass 5: sample_edge_intersection_vertices_kern (dmc.cu:1468-1514).  e6
   * runs p2 -> p6, e9 p5 -> p6, e10 p6 -> p7. */
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (aa = 0; aa < (int64_t)A; ++aa) {
    uint32_t a = (uint32_t)aa;
    uint32_t id = rec[a].index1d;
    uint32_t i = id % cx, j = (id / cx) % cy, k = id / (cx * cy);
    uint32_t w = rec[a].vertex_begin + L->num_vertex[(rec[a].info >> 7) & 1][rec[a].cfg];
    float x1 = ijk_to_xyz(off[0] + i + 1, div[0], range[0], xyz_min[0]);
    float y1 = ijk_to_xyz(off[1] + j + 1, div[1], range[1], xyz_min[1]);
    float z1 = ijk_to_xyz(off[2] + k + 1, div[2], range[2], xyz_min[2]);
    double X1 = ijk_to_xyz64(off[0] + i + 1, div[0], range64[0], xyz_min[0]);
    double Y1 = ijk_to_xyz64(off[1] + j + 1, div[1], range64[1], xyz_min[1]);
    double Z1 = ijk_to_xyz64(off[2] + k + 1, div[2], range64[2], xyz_min[2]);
    float v6 = F(field, X, Y, i + 1, j + 1, k + 1);
    int l;
    for (l = 0; l < 3; ++l) {
      float p0[3], p1[3], va, vb, t, omt;
      double q0[3], q1[3], T, OMT;
      if (!((rec[a].info >> l) & 1)) continue;
      if (l == 0) { /* edge 6, pt 2 -> 6 */
        float zc = ijk_to_xyz(off[2] + k, div[2], range[2], xyz_min[2]);
        p0[0] = x1; p0[1] = y1; p0[2] = zc; p1[0] = x1; p1[1] = y1; p1[2] = z1;
        q0[0] = X1; q0[1] = Y1; q0[2] = ijk_to_xyz64(off[2] + k, div[2], range64[2], xyz_min[2]);
        q1[0] = X1; q1[1] = Y1; q1[2] = Z1;
        va = F(field, X, Y, i + 1, j + 1, k); vb = v6;
      } else if (l == 1) { /* edge 9, pt 5 -> 6 */
        float yc = ijk_to_xyz(off[1] + j, div[1], range[1], xyz_min[1]);
        p0[0] = x1; p0[1] = yc; p0[2] = z1; p1[0] = x1; p1[1] = y1; p1[2] = z1;
        q0[0] = X1; q0[1] = ijk_to_xyz64(off[1] + j, div[1], range64[1], xyz_min[1]); q0[2] = Z1;
        q1[0] = X1; q1[1] = Y1; q1[2] = Z1;
        va = F(field, X, Y, i + 1, j, k + 1); vb = v6;
      } else { /* edge 10, pt 6 -> 7 */
        float xc = ijk_to_xyz(off[0] + i, div[0], range[0], xyz_min[0]);
        p0[0] = x1; p0[1] = y1; p0[2] = z1; p1[0] = xc; p1[1] = y1; p1[2] = z1;
        q0[0] = X1; q0[1] = Y1; q0[2] = Z1;
        q1[0] = ijk_to_xyz64(off[0] + i, div[0], range64[0], xyz_min[0]); q1[1] = Y1; q1[2] = Z1;
        va = v6; vb = F(field, X, Y, i, j + 1, k + 1);
      }
      t = (iso - va) / (vb - va);
      omt = 1.0f - t;
      out->vertices[3 * (size_t)w + 0] = lerp1(p0[0], p1[0], t, omt);
      out->vertices[3 * (size_t)w + 1] = lerp1(p0[1], p1[1], t, omt);
      out->vertices[3 * (size_t)w + 2] = lerp1(p0[2], p1[2], t, omt);
      if (want_fp64) {
        T = ((double)iso - (double)va) / ((double)vb - (double)va);
        OMT = 1.0 - T;
        out->vertices64[3 * (size_t)w + 0] = lerp1d(q0[0], q1[0], T, OMT);
        out->vertices64[3 * (size_t)w + 1] = lerp1d(q0[1], q1[1], T, OMT);
        out->vertices64[3 * (size_t)w + 2] = lerp1d(q0[2], q1[2], T, OMT);
      }
      ++w;
    }
  }

  /* Pass 6 */
  reaverage(out, rec, map, A, cx, cy, cz, L, want_fp64, 1);

  /* Pass 7: calc_num_triangles_per_voxel_kern (dmc.cu:1983-2008) with
   * circular_edge_exceed_boundary (dmc.cu:1696-1708), then the offsets scan
   * (dmc.cu:2178-2183).  Counted here in quads (two triangles each). */
  for (a = 0; a < A; ++a) {
    uint32_t id = rec[a].index1d;
    uint32_t i = id % cx, j = (id / cx) % cy, k = id / (cx * cy);
    uint32_t n = 0;
    out->quad_begin[a] = Q;
    if ((rec[a].info & 1) && i + 1 < cx && j + 1 < cy) ++n;
    if ((rec[a].info & 2) && i + 1 < cx && k + 1 < cz) ++n;
    if ((rec[a].info & 4) && j + 1 < cy && k + 1 < cz) ++n;
    Q += n;
  }
  out->num_quads = Q;
  out->quads = (int32_t *)malloc((size_t)(Q ? Q : 1) * 4 * sizeof(int32_t));
  if (!out->quads) { free(map); free(rec); oracle_result_free(out); return -1; }

  /* Pass 8: generate_triangles_kern (dmc.cu:2011-2047) with
   * get_circular_vertices_by_edge (dmc.cu:1868-1892): the four ring cells'
   * iso vertices, ccw order 0,1,2,3 or cw order 0,3,2,1. */
#pragma omp parallel for schedule(static) num_threads(g_threads) reduction(+ : anomalies)
  for (aa = 0; aa < (int64_t)A; ++aa) {
    uint32_t a = (uint32_t)aa;
    uint32_t id = rec[a].index1d;
    int64_t i = id % cx, j = (id / cx) % cy, k = id / (cx * cy);
    uint32_t q = out->quad_begin[a];
    int l;
    for (l = 0; l < 3; ++l) {
      int e = LOCAL_EDGES[l], p, ccw;
      int64_t members[4];
      int32_t vidx[4];
      if (!((rec[a].info >> l) & 1)) continue;
      if (e == 6 && !(i + 1 < cx && j + 1 < cy)) continue;
      if (e == 9 && !(i + 1 < cx && k + 1 < cz)) continue;
      if (e == 10 && !(j + 1 < cy && k + 1 < cz)) continue;
      ccw = (rec[a].info >> (l + 3)) & 1;
      for (p = 0; p < 4; ++p) {
        int re = RING_E[l][p];
        int64_t ri = i - EDGE_OWNER_D[re][0], rj = j - EDGE_OWNER_D[re][1], rk = k - EDGE_OWNER_D[re][2];
        uint32_t rm = map[((size_t)rk * cy + (size_t)rj) * cx + (size_t)ri];
        uint8_t lab;
        members[p] = rm;
        if (rm == INVALID_U32) { anomalies++; vidx[p] = -1; continue; }
        lab = L->edge_label[(rec[rm].info >> 7) & 1][rec[rm].cfg][re];
        if (lab == MDMC_NO_VERTEX) { anomalies++; vidx[p] = -1; continue; }
        vidx[p] = (int32_t)(rec[rm].vertex_begin + lab);
      }
      (void)members;
      out->quads[4 * (size_t)q + 0] = vidx[0];
      out->quads[4 * (size_t)q + 1] = ccw ? vidx[1] : vidx[3];
      out->quads[4 * (size_t)q + 2] = vidx[2];
      out->quads[4 * (size_t)q + 3] = ccw ? vidx[3] : vidx[1];
      ++q;
    }
  }
  out->num_anomalies += anomalies;

  /* smoothing iterations; the FP64 twin is left unsmoothed */
  {
    uint32_t it;
    for (it = 0; it < num_smooth; ++it) {
      smooth_pass(out, rec, A, cx, cy, cz, L, xyz_min, range, div, off);
      reaverage(out, rec, map, A, cx, cy, cz, L, 0, 0);
    }
  }

  for (a = 0; a < A; ++a) {
    out->cell_index[a] = rec[a].index1d;
    out->cfg[a] = rec[a].cfg;
    out->info[a] = rec[a].info;
    out->vertex_begin[a] = rec[a].vertex_begin;
  }
  free(map);
  free(rec);
  return 0;
}

int oracle_dmc_run(const float *field, const uint32_t dims[3], const float xyz_min[3],
                   const float xyz_max[3], float iso, int want_fp64, oracle_result *out) {
  return oracle_dmc_run_smooth(field, dims, xyz_min, xyz_max, iso, want_fp64, 0, out);
}

void oracle_result_free(oracle_result *r) {
  free(r->vertices);
  free(r->vertices64);
  free(r->quads);
  free(r->cell_index);
  free(r->cfg);
  free(r->info);
  free(r->vertex_begin);
  free(r->quad_begin);
  memset(r, 0, sizeof(*r));
}

/* ---- field generators (definitions in mdmc_fields.h) ---- */

void oracle_field_sphere(float *out, const uint32_t dims[3], const double c[3], double radius) {
  uint32_t x, y, z;
  for (z = 0; z < dims[2]; ++z)
    for (y = 0; y < dims[1]; ++y)
      for (x = 0; x < dims[0]; ++x) {
        double dx = (double)x - c[0], dy = (double)y - c[1], dz = (double)z - c[2];
        out[((size_t)z * dims[1] + y) * dims[0] + x] = (float)(sqrt(dx * dx + dy * dy + dz * dz) - radius);
      }
}

void oracle_field_blobs(float *out, const uint32_t dims[3], uint64_t seed, uint32_t nblobs,
                        double sigma_lo, double sigma_hi, double amp_lo, double amp_hi) {
  double *ex = (double *)malloc(sizeof(double) * nblobs * dims[0]);
  double *ey = (double *)malloc(sizeof(double) * nblobs * dims[1]);
  double *ez = (double *)malloc(sizeof(double) * nblobs * dims[2]);
  double *amp = (double *)malloc(sizeof(double) * (nblobs ? nblobs : 1));
  uint32_t b, x, y, z;
  for (b = 0; b < nblobs; ++b) {
    mdmc_blob p = mdmc_blob_params(seed, b, dims, sigma_lo, sigma_hi, amp_lo, amp_hi);
    mdmc_gauss_1d(ex + (size_t)b * dims[0], dims[0], p.cx, p.sigma);
    mdmc_gauss_1d(ey + (size_t)b * dims[1], dims[1], p.cy, p.sigma);
    mdmc_gauss_1d(ez + (size_t)b * dims[2], dims[2], p.cz, p.sigma);
    amp[b] = p.amp;
  }
  for (z = 0; z < dims[2]; ++z)
    for (y = 0; y < dims[1]; ++y)
      for (x = 0; x < dims[0]; ++x) {
        double acc = 0.0;
        for (b = 0; b < nblobs; ++b)
          acc = fma((amp[b] * ex[(size_t)b * dims[0] + x]) * ey[(size_t)b * dims[1] + y],
                    ez[(size_t)b * dims[2] + z], acc);
        out[((size_t)z * dims[1] + y) * dims[0] + x] = (float)acc;
      }
  free(ex); free(ey); free(ez); free(amp);
}

void oracle_field_gyroid(float *out, const uint32_t dims[3], double period, uint64_t seed,
                         double noise_amp) {
  double *sx = (double *)malloc(sizeof(double) * dims[0]), *cxv = (double *)malloc(sizeof(double) * dims[0]);
  double *sy = (double *)malloc(sizeof(double) * dims[1]), *cyv = (double *)malloc(sizeof(double) * dims[1]);
  double *sz = (double *)malloc(sizeof(double) * dims[2]), *czv = (double *)malloc(sizeof(double) * dims[2]);
  uint32_t x, y, z;
  mdmc_trig_1d(sx, cxv, dims[0], period);
  mdmc_trig_1d(sy, cyv, dims[1], period);
  mdmc_trig_1d(sz, czv, dims[2], period);
  for (z = 0; z < dims[2]; ++z)
    for (y = 0; y < dims[1]; ++y)
      for (x = 0; x < dims[0]; ++x) {
        size_t id = ((size_t)z * dims[1] + y) * dims[0] + x;
        double g = sx[x] * cyv[y] + sy[y] * czv[z] + sz[z] * cxv[x];
        out[id] = (float)(g + mdmc_noise(seed, id, noise_amp));
      }
  free(sx); free(cxv); free(sy); free(cyv); free(sz); free(czv);
}
