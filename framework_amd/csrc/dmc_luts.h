/*
 * Dual Marching Cubes case tables, generated from a connectivity rule plus a
 * small label table.
 *
 * The reference keeps these tables as literal arrays in
 * DMC_CUDA/DMC_CUDA/src/dmc.cu:45-620 (config_edge_lut1/2, num_vertex_lut1/2,
 * config_2B_3B_lut, config_2B_3B_ambiguous_face).  Here they are rebuilt:
 *
 *   - a cell configuration `cfg` has bit b set when corner b is inside
 *     (value < iso); an edge is bipolar when its two corner bits differ;
 *   - iso-vertex groups are the connected components of the bipolar edges,
 *     where two bipolar edges are linked when they bound the same face segment:
 *     on a face with two bipolar edges they link directly; on an ambiguous face
 *     (four bipolar edges, diagonal corners alike) variant 0 joins the inside
 *     corners, i.e. each outside corner's two face edges are linked, and
 *     variant 1 cuts the other way on that config's listed ambiguous face;
 *   - the 2B/3B configs (dmc.cu:607-620) are exactly the configs with two or
 *     three inside corners and exactly one ambiguous face; that face is listed;
 *   - group numbering follows the reference: by default groups are numbered in
 *     order of their lowest edge index, and LABEL_FIX_* lists the rows whose
 *     reference numbering differs from that order.
 *
 * The resulting byte arrays are pinned by SHA-256 against the reference arrays
 * in the test suite.  This header is plain C99 so that the CPU oracle (gcc)
 * and the HIP library (hipcc) build the same tables.
 */
#ifndef MDMC_DMC_LUTS_H
#define MDMC_DMC_LUTS_H

#include <stdint.h>
#include <string.h>

#define MDMC_NO_VERTEX 0xFFu
#define MDMC_NOT_AMBIGUOUS 0xFFu

/* Corner c sits at offset (CORNER_D[c][0], CORNER_D[c][1], CORNER_D[c][2]) of
 * the cell's lowest corner (dmc.cu:1044-1054). */
static const uint8_t MDMC_CORNER_D[8][3] = {
    {0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
    {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

/* Edge e joins corners EDGE_C[e][0] and EDGE_C[e][1] (dmc.cu:1109-1129). */
static const uint8_t MDMC_EDGE_C[12][2] = {
    {0, 1}, {1, 2}, {2, 3}, {0, 3}, {0, 4}, {1, 5},
    {2, 6}, {3, 7}, {4, 5}, {5, 6}, {6, 7}, {4, 7}};

/* Face f has corners FACE_C[f][0..3] in cyclic order.  Faces are, in order,
 * -z, -y, +x, +y, -x, +z (dmc.cu:622-632). */
static const uint8_t MDMC_FACE_C[6][4] = {
    {0, 1, 2, 3}, {0, 1, 5, 4}, {1, 2, 6, 5},
    {2, 3, 7, 6}, {0, 3, 7, 4}, {4, 5, 6, 7}};

/* Direction of the neighbour across face f, in the reference's encoding
 * POS_X=0, NEG_X=1, POS_Y=2, NEG_Y=3, POS_Z=4, NEG_Z=5 (dmc.cu:624-632). */
static const uint8_t MDMC_FACE_DIR[6] = {5, 3, 0, 2, 1, 4};

/* Rows whose reference group numbering is not "lowest edge first".  Each entry
 * is {cfg, code}; the label of the g-th group (groups ordered by lowest edge)
 * is (code >> 2g) & 3.  Variant 0 rows apply to both variants outside the
 * 2B/3B configs, where the two tables agree. */
static const uint8_t MDMC_LABEL_FIX_V0[][2] = {
    {0x1a, 0x01}, {0x25, 0x01}, {0x3d, 0x01}, {0x3e, 0x01}, {0x53, 0x01},
    {0x56, 0x01}, {0x59, 0x01}, {0x5c, 0x01}, {0x5e, 0x18}, {0x5f, 0x01},
    {0x69, 0x01}, {0x6a, 0x01}, {0x6b, 0x01}, {0x6d, 0x01}, {0x75, 0x01},
    {0x78, 0x01}, {0x7a, 0x09}, {0x7e, 0x01}, {0x95, 0x01}, {0x97, 0x01},
    {0x9e, 0x01}, {0xa3, 0x01}, {0xa5, 0xb4}, {0xa6, 0x01}, {0xa7, 0x18},
    {0xa9, 0x01}, {0xac, 0x01}, {0xaf, 0x01}, {0xb4, 0x01}, {0xb5, 0x06},
    {0xba, 0x01}, {0xbd, 0x01}, {0xc3, 0x01}, {0xc5, 0x01}, {0xc7, 0x01},
    {0xca, 0x01}, {0xcb, 0x01}, {0xd2, 0x01}, {0xda, 0x21}, {0xde, 0x01},
    {0xe1, 0x01}, {0xe5, 0x12}, {0xe7, 0x01}, {0xfa, 0x01}};

/* Variant 1 numbering for the 2B/3B configs (the only rows where the variant
 * 1 partition differs from variant 0). */
static const uint8_t MDMC_LABEL_FIX_V1[][2] = {
    {0x05, 0x01}, {0x18, 0x01}, {0x21, 0x01}, {0x34, 0x01}, {0x38, 0x01},
    {0x42, 0x01}, {0x45, 0x01}, {0x50, 0x01}, {0x61, 0x01}, {0x68, 0x01},
    {0x81, 0x01}, {0x8a, 0x01}, {0x92, 0x01}, {0x94, 0x01}, {0xa0, 0x01},
    {0xc1, 0x01}, {0xc2, 0x01}};

typedef struct mdmc_luts {
  /* [variant][cfg][edge] -> iso-vertex label, MDMC_NO_VERTEX when the edge is
   * not bipolar (dmc.cu:45-303 for variant 0, :327-585 for variant 1). */
  uint8_t edge_label[2][256][12];
  /* [variant][cfg] -> number of iso vertices (dmc.cu:307-325, :587-605). */
  uint8_t num_vertex[2][256];
  /* [cfg] -> ambiguous face of a 2B/3B config, MDMC_NOT_AMBIGUOUS otherwise
   * (dmc.cu:607-620). */
  uint8_t amb_face[256];
} mdmc_luts;

static inline int mdmc_edge_of(int a, int b) {
  int e;
  for (e = 0; e < 12; ++e)
    if ((MDMC_EDGE_C[e][0] == a && MDMC_EDGE_C[e][1] == b) ||
        (MDMC_EDGE_C[e][0] == b && MDMC_EDGE_C[e][1] == a))
      return e;
  return -1;
}

static inline int mdmc_uf_find(int *parent, int x) {
  while (parent[x] != x) x = parent[x];
  return x;
}

static inline void mdmc_uf_union(int *parent, int a, int b) {
  int ra = mdmc_uf_find(parent, a), rb = mdmc_uf_find(parent, b);
  if (ra != rb) parent[ra] = rb;
}

/* A face is ambiguous when its diagonal corners agree and adjacent corners
 * differ (four bipolar edges). */
static inline int mdmc_face_is_ambiguous(int cfg, int f) {
  int b0 = (cfg >> MDMC_FACE_C[f][0]) & 1, b1 = (cfg >> MDMC_FACE_C[f][1]) & 1;
  int b2 = (cfg >> MDMC_FACE_C[f][2]) & 1, b3 = (cfg >> MDMC_FACE_C[f][3]) & 1;
  return b0 == b2 && b1 == b3 && b0 != b1;
}

static inline int mdmc_popcount8(int v) {
  int n = 0;
  while (v) { n += v & 1; v >>= 1; }
  return n;
}

static inline uint8_t mdmc_label_code(const uint8_t (*tab)[2], int n, int cfg) {
  int i;
  for (i = 0; i < n; ++i)
    if (tab[i][0] == cfg) return tab[i][1];
  return 0xE4; /* identity: labels 0,1,2,3 */
}

static inline void mdmc_build_luts(mdmc_luts *L) {
  int cfg, f, v, e, q;
  memset(L, 0xFF, sizeof(*L));
  for (cfg = 0; cfg < 256; ++cfg) {
    int n_amb = 0, amb = -1, pc = mdmc_popcount8(cfg);
    for (f = 0; f < 6; ++f)
      if (mdmc_face_is_ambiguous(cfg, f)) { ++n_amb; amb = f; }
    L->amb_face[cfg] = (n_amb == 1 && (pc == 2 || pc == 3)) ? (uint8_t)amb : MDMC_NOT_AMBIGUOUS;
  }
  for (v = 0; v < 2; ++v) {
    for (cfg = 0; cfg < 256; ++cfg) {
      int parent[12], bip[12], root_label[12];
      int ngroups = 0, flip = -1;
      uint8_t code;
      if (v == 1 && L->amb_face[cfg] != MDMC_NOT_AMBIGUOUS) flip = L->amb_face[cfg];
      for (e = 0; e < 12; ++e) {
        parent[e] = e;
        bip[e] = ((cfg >> MDMC_EDGE_C[e][0]) & 1) != ((cfg >> MDMC_EDGE_C[e][1]) & 1);
      }
      for (f = 0; f < 6; ++f) {
        int fe[4], nb = 0, list[4];
        for (q = 0; q < 4; ++q) {
          fe[q] = mdmc_edge_of(MDMC_FACE_C[f][q], MDMC_FACE_C[f][(q + 1) & 3]);
          if (bip[fe[q]]) list[nb++] = fe[q];
        }
        if (nb == 2) {
          mdmc_uf_union(parent, list[0], list[1]);
        } else if (nb == 4) {
          /* link the two face edges meeting at each corner whose inside bit
           * equals `want`: outside corners (want 0) normally, inside corners on
           * the flipped face */
          int want = (f == flip) ? 1 : 0;
          for (q = 0; q < 4; ++q) {
            int c = MDMC_FACE_C[f][q];
            if (((cfg >> c) & 1) == want)
              mdmc_uf_union(parent, fe[q], fe[(q + 3) & 3]);
          }
        }
      }
      for (e = 0; e < 12; ++e) root_label[e] = -1;
      for (e = 0; e < 12; ++e) {
        int r;
        if (!bip[e]) continue;
        r = mdmc_uf_find(parent, e);
        if (root_label[r] < 0) root_label[r] = ngroups++;
      }
      if (v == 1 && flip >= 0)
        code = mdmc_label_code(MDMC_LABEL_FIX_V1, (int)(sizeof(MDMC_LABEL_FIX_V1) / 2), cfg);
      else
        code = mdmc_label_code(MDMC_LABEL_FIX_V0, (int)(sizeof(MDMC_LABEL_FIX_V0) / 2), cfg);
      for (e = 0; e < 12; ++e) {
        if (!bip[e]) { L->edge_label[v][cfg][e] = MDMC_NO_VERTEX; continue; }
        L->edge_label[v][cfg][e] = (uint8_t)((code >> (2 * root_label[mdmc_uf_find(parent, e)])) & 3);
      }
      L->num_vertex[v][cfg] = (uint8_t)ngroups;
    }
  }
}

/* Kernel-side forms of the tables.
 *   packed_label[v][cfg]: 12 nibbles, nibble e = label of edge e or 0xF.
 *   cfg_info[cfg]: bits 0-2 num_vertex variant 0, bits 3-5 num_vertex variant
 *   1, bit 6 = 2B/3B config, bits 7-9 = check direction (reference encoding,
 *   POS_X=0 ... NEG_Z=5). */
static inline void mdmc_pack_luts(const mdmc_luts *L, uint64_t packed_label[2][256],
                                  uint16_t cfg_info[256]) {
  int v, cfg, e;
  for (v = 0; v < 2; ++v)
    for (cfg = 0; cfg < 256; ++cfg) {
      uint64_t w = 0;
      for (e = 0; e < 12; ++e) {
        uint64_t lab = L->edge_label[v][cfg][e];
        w |= (lab == MDMC_NO_VERTEX ? 0xFull : lab) << (4 * e);
      }
      packed_label[v][cfg] = w;
    }
  for (cfg = 0; cfg < 256; ++cfg) {
    uint16_t w = (uint16_t)(L->num_vertex[0][cfg] | (L->num_vertex[1][cfg] << 3));
    if (L->amb_face[cfg] != MDMC_NOT_AMBIGUOUS)
      w = (uint16_t)(w | (1u << 6) | ((unsigned)MDMC_FACE_DIR[L->amb_face[cfg]] << 7));
    cfg_info[cfg] = w;
  }
}

#endif /* MDMC_DMC_LUTS_H */
