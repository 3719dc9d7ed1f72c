// Dual Marching Cubes kernels for gfx950 (MI355X).
//
// Pipeline for one volume (stream-ordered launches, no host sync):
//
//   k_classify_pt  dense pass over the field in plane tiles (7 cell rows x
//               one x-tile x 1-4 cell planes per 256-thread workgroup): point
//               rows staged in LDS as sign bits, per-cell configuration keys
//               by multiply-shift, uniform 256-cell chunks skipped with one
//               ballot; active cells get their 2B/3B table variant, vertex and
//               quad counts, a DPP-scan rank and a 64-bit record in their
//               segment's slot range (segment = one cell row of one x-tile in
//               one plane), plus a per-64-cell group map (active mask, first
//               rank).  Workgroups map to tiles XCD by XCD (mdmc_tilemap.h).
//               k_classify is the z-marching form for any row length.
//   k_scan      single-pass decoupled look-back scan (ticketed tiles) over the
//               segment totals: vertex/quad bases per segment, the non-empty
//               segment list and where each 64-cell chunk starts.
//   k_emit      one wave per chunk of 64 active cells: edge crossing points
//               in the owner's orientation, dual vertices (FP32, reference
//               operation order) and quads; ring cells are found through the
//               group map in O(1).
//   k_smooth, k_reaverage   optional smoothing iterations over the mesh.
//
// Batched forms (k_classify_batch, k_batch_counts, k_emit_batch) run the
// same per-tile and per-cell code over many volumes; k_slab_info reports an
// out-of-core slab in one record; k_field_* generate the synthetic benchmark
// fields.
//
// Reference semantics: DMC_CUDA/DMC_CUDA/src/dmc.cu:1018-2208 (dmc::run_dmc).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mdmc_tilemap.h"

#pragma clang fp contract(off)

namespace mdmc {

constexpr int kWave = 64;
constexpr uint32_t kStatusVertexOverflow = 1, kStatusQuadOverflow = 2, kStatusLookup = 4,
                   kStatusScanTimeout = 8, kStatusInternal = 16, kStatusIndexRange = 32;

// Case tables in constant memory (uploaded by mdmc_setup):
//   c_label[v][cfg]   12 nibbles, edge e -> iso-vertex label or 0xF
//   c_cfg_info[cfg]   bits 0-2 NV variant 0, 3-5 NV variant 1, 6 = 2B/3B,
//                     7-9 check direction (POS_X=0 .. NEG_Z=5)
__constant__ uint64_t c_label[2][256];
__constant__ uint16_t c_cfg_info[256];
// c_cfg_count[cfg]: bits 0-3 NV variant 0 + bipolar owned edges, 4-7 the same
// for variant 1, 8-10 bipolar owned edges 6/9/10, 11 = 2B/3B, 12-14 check dir.
__constant__ uint16_t c_cfg_count[256];

struct Totals {
  uint32_t num_active, num_vertices, num_quads, num_nonempty;
  uint32_t num_chunks, pad0, pad1, pad2;
};

struct FieldDesc {
  const float *f;
  uint32_t X, Y, Z;     // points
  uint32_t cx, cy, cz;  // cells
  float iso;
};

// Record of an active cell (64 bits):
//   bits  0-9   x offset inside its x-tile
//   bits 10-17  configuration
//   bit  18     uses the variant-1 table (2B/3B correction)
//   bits 19-31  first vertex, relative to the segment
//   bits 32-43  first quad, relative to the segment
__device__ __forceinline__ uint64_t make_record(uint32_t xl, uint32_t cfg, uint32_t lut2,
                                                uint32_t lvb, uint32_t lqb) {
  return (uint64_t)(xl | (cfg << 10) | (lut2 << 18) | (lvb << 19)) | ((uint64_t)lqb << 32);
}
__device__ __forceinline__ uint32_t rec_x(uint64_t r) { return (uint32_t)r & 0x3FFu; }
__device__ __forceinline__ uint32_t rec_cfg(uint64_t r) { return ((uint32_t)r >> 10) & 0xFFu; }
__device__ __forceinline__ uint32_t rec_lut2(uint64_t r) { return ((uint32_t)r >> 18) & 1u; }
__device__ __forceinline__ uint32_t rec_lvb(uint64_t r) { return ((uint32_t)r >> 19) & 0x1FFFu; }
__device__ __forceinline__ uint32_t rec_lqb(uint64_t r) { return (uint32_t)(r >> 32) & 0xFFFu; }

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Index of the calling wave in its workgroup, as a wave-uniform (scalar)
// value so that per-wave bookkeeping compiles to scalar instructions.
__device__ __forceinline__ uint32_t wave_id() {
  return __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b -> XCD b%8).
// Remap so that each XCD works on a contiguous range of tiles: neighbouring
// tiles (which share halo rows) then hit the same L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
  uint32_t q = n >> 3, r = n & 7, x = b & 7;
  return x * q + min(x, r) + (b >> 3);
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = (int)lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint32_t t = __shfl_up(v, d, kWave);
    if (lane >= d) v += t;
  }
  return v;
}

// DPP controls (GFX9 encoding).
constexpr int kDppRowShr = 0x110;  // + n: lane i reads lane i-n of its 16-lane row
constexpr int kDppWaveShl1 = 0x130;  // lane i reads lane i+1 across the wave
constexpr int kDppRowBcast15 = 0x142;
constexpr int kDppRowBcast31 = 0x143;

// Value of lane+1 (0 in lane 63).
__device__ __forceinline__ uint32_t lane_next(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kDppWaveShl1, 0xF, 0xF, true);
}

// Inclusive prefix sum over the 64 lanes with DPP (no LDS traffic).
__device__ __forceinline__ uint32_t wave_incl_scan_dpp(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kDppRowShr + 1, 0xF, 0xF, true);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kDppRowShr + 2, 0xF, 0xF, true);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kDppRowShr + 4, 0xF, 0xF, true);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kDppRowShr + 8, 0xF, 0xF, true);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kDppRowBcast15, 0xA, 0xF, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kDppRowBcast31, 0xC, 0xF, false);
  return x;
}

__device__ __forceinline__ uint32_t mbcnt64(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Configuration of cell (x, y, z) read straight from the field (used for the
// rare 2B/3B neighbour checks whose neighbour lies outside the tile window).
__device__ __forceinline__ uint32_t cell_cfg_global(const FieldDesc &g, uint32_t x, uint32_t y,
                                                    uint32_t z) {
  const size_t sy = g.X, sz = (size_t)g.X * g.Y;
  const float *p = g.f + (size_t)z * sz + (size_t)y * sy + x;
  const float iso = g.iso;
  uint32_t c = 0;
  c |= (uint32_t)(p[0] < iso);
  c |= (uint32_t)(p[1] < iso) << 1;
  c |= (uint32_t)(p[sy + 1] < iso) << 2;
  c |= (uint32_t)(p[sy] < iso) << 3;
  c |= (uint32_t)(p[sz] < iso) << 4;
  c |= (uint32_t)(p[sz + 1] < iso) << 5;
  c |= (uint32_t)(p[sz + sy + 1] < iso) << 6;
  c |= (uint32_t)(p[sz + sy] < iso) << 7;
  return c;
}

// 2B/3B correction (dmc.cu:1345-1371): a cell whose configuration is 2B/3B
// uses the variant-1 table when its in-grid neighbour across the listed face
// is also 2B/3B.
__device__ __forceinline__ uint32_t use_variant1(const FieldDesc &g, const uint16_t *info_tab,
                                                 uint32_t info, uint32_t x, uint32_t y,
                                                 uint32_t z) {
  if (!(info & 64u)) return 0;
  uint32_t dir = (info >> 7) & 7u;
  int nx = (int)x, ny = (int)y, nz = (int)z;
  switch (dir) {
    case 0: nx += 1; break;
    case 1: nx -= 1; break;
    case 2: ny += 1; break;
    case 3: ny -= 1; break;
    case 4: nz += 1; break;
    default: nz -= 1; break;
  }
  if (nx < 0 || ny < 0 || nz < 0 || nx >= (int)g.cx || ny >= (int)g.cy || nz >= (int)g.cz) return 0;
  uint32_t ncfg = cell_cfg_global(g, (uint32_t)nx, (uint32_t)ny, (uint32_t)nz);
  return (info_tab[ncfg] >> 6) & 1u;
}

// The scan's ticket word opens a region of the scratch of kTicketBytes: word 0
// is the scan's tile ticket, and word kEmitCtr * (x + 1) the emit's chunk
// ticket counter of XCD x (k_emit), each on a 256-byte line of its own (the
// atomics on one line serialise).  The classify zeroes all nine.
constexpr uint32_t kEmitCtr = 64;
constexpr uint32_t kTicketBytes = 9 * kEmitCtr * 4;

__device__ __forceinline__ void zero_tickets(uint32_t *t, uint32_t tid) {
  if (tid < 9u) t[tid * kEmitCtr] = 0u;
}

struct ClassifyArgs {
  FieldDesc g;
  uint32_t XT, NRB, NKC, Kc;
  uint2 *seg_counts;      // per segment: {nact | nquad << 16, nvert}
  uint64_t *slots;        // per segment: 64*W record slots
  uint4 *gmap;            // per segment and 64-cell group: {active mask lo, hi, rank of
                          // the group's first active cell, 0}; written for non-empty groups
  unsigned long long *scan_state;  // zeroed here for k_scan
  uint32_t scan_state_words;
  uint2 *seg_pad;         // segment totals past the last segment, up to a whole scan tile
  uint32_t seg_pad_n;     // (zeroed here, read by the scan)
  uint32_t *scan_ticket;  // the ticket region, zeroed here
  Totals *totals;
  uint32_t *counts_out;   // user mdmc_counts (4 words), zeroed here
};

// One wave per tile; see the file header.
template <int W, int R>
__global__ __launch_bounds__(64) void k_classify(ClassifyArgs a) {
  static_assert(R >= 1 && R <= 15, "rows per tile must fit the 16-bit plane fields");
  static_assert(W >= 1 && W <= 16, "x-tile is at most 1024 cells");
  __shared__ uint16_t s_info[256];
  const uint32_t lane = lane_id();
  for (uint32_t i = lane; i < 256; i += kWave) s_info[i] = c_cfg_info[i];
  if (blockIdx.x == 0) {
    for (uint32_t i = lane; i < a.scan_state_words; i += kWave) a.scan_state[i] = 0ull;
    for (uint32_t i = lane; i < a.seg_pad_n; i += kWave) a.seg_pad[i] = make_uint2(0u, 0u);
    if (lane < 8) reinterpret_cast<uint32_t *>(a.totals)[lane] = 0u;
    zero_tickets(a.scan_ticket, lane);
    if (a.counts_out != nullptr && lane < 4) a.counts_out[lane] = 0u;
  }
  __syncthreads();

  const FieldDesc g = a.g;
  const float iso = g.iso;
  constexpr uint32_t CAP = 64u * W;
  const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t xt = tile % a.XT;
  const uint32_t rest = tile / a.XT;
  const uint32_t rb = rest % a.NRB;
  const uint32_t kc = rest / a.NRB;
  const uint32_t x0 = xt * CAP;
  const uint32_t j0 = rb * R;
  const uint32_t k0 = kc * a.Kc;
  const uint32_t k1 = min(k0 + a.Kc, g.cz);
  if (k0 >= g.cz) return;
  const size_t plane = (size_t)g.X * g.Y;
  const bool has_extra = x0 + CAP < g.X;  // first point of the next x-tile
  const float qnan = __int_as_float(0x7FC00000);

  float vals[R + 1][W];
  float vext[R + 1];

  auto load_plane = [&](uint32_t z) {
#pragma unroll
    for (int r = 0; r <= R; ++r) {
      const uint32_t y = j0 + r;
      const bool row_ok = (z < g.Z) && (y < g.Y);
      const float *rowp = g.f + (size_t)z * plane + (size_t)y * g.X;
#pragma unroll
      for (int c = 0; c < W; ++c) {
        const uint32_t x = x0 + 64u * c + lane;
        vals[r][c] = (row_ok && x < g.X) ? rowp[x] : qnan;
      }
      vext[r] = (row_ok && has_extra) ? rowp[x0 + CAP] : qnan;
    }
  };
  auto plane_bits = [&](int c) {
    uint32_t b = 0;
#pragma unroll
    for (int r = 0; r <= R; ++r) b |= (uint32_t)(vals[r][c] < iso) << r;
    return b;
  };
  auto ext_bits = [&]() {
    uint32_t b = 0;
#pragma unroll
    for (int r = 0; r <= R; ++r) b |= (uint32_t)(vext[r] < iso) << r;
    return b;
  };

  uint32_t B[W];
  uint32_t E;
  load_plane(k0);
#pragma unroll
  for (int c = 0; c < W; ++c) B[c] = plane_bits(c) << 16;
  E = ext_bits() << 16;
  load_plane(k0 + 1);

  constexpr uint32_t rowmask = (1u << (R + 1)) - 1u;
  constexpr uint32_t fullmask = rowmask | (rowmask << 16);

  for (uint32_t k = k0; k < k1; ++k) {
#pragma unroll
    for (int c = 0; c < W; ++c) B[c] = (B[c] >> 16) | (plane_bits(c) << 16);
    E = (E >> 16) | (ext_bits() << 16);
    if (k + 1 < k1) load_plane(k + 2);  // prefetch the plane the next step needs

    // Neighbour columns (lane+1; lane 63 takes the next chunk's lane 0) and
    // the chunks that are not uniformly in or out over all rows.
    uint32_t N[W];
    uint32_t live = 0;
#pragma unroll
    for (int c = 0; c < W; ++c) {
      const uint32_t own = B[c];
      uint32_t nb = __shfl_down(own, 1, kWave);
      const uint32_t nxt = (c + 1 < W) ? (uint32_t)__shfl(B[(c + 1 < W) ? c + 1 : c], 0, kWave) : E;
      if (lane == kWave - 1) nb = nxt;
      N[c] = nb;
      const uint32_t uo = own & fullmask, un = nb & fullmask;
      const bool all_out = (uo | un) == 0u;
      const bool all_in = (uo & un) == fullmask;
      if (__ballot(!all_out) != 0ull && __ballot(!all_in) != 0ull) live |= 1u << c;
    }

#pragma unroll 1
    for (int r = 0; r < R; ++r) {
      const uint32_t y = j0 + r;
      if (y >= g.cy) break;
      uint32_t cnt_a = 0, cnt_v = 0, cnt_q = 0;
      const uint32_t seg = xt + a.XT * (y + g.cy * k);
#pragma unroll
      for (int c = 0; c < W; ++c) {
        if (!((live >> c) & 1u)) continue;
        const uint32_t own = B[c], nb = N[c];
        const uint32_t x = x0 + 64u * c + lane;
        const uint32_t t = ((own >> r) & 0x30003u) | (((nb >> r) & 0x30003u) << 2);
        const uint32_t key = (t & 0xFu) | ((t >> 12) & 0xF0u);
        const uint32_t cfg = (key & 0x11u) | ((key >> 1) & 0x66u) | ((key << 2) & 0x88u);
        const bool active = (x < g.cx) && (cfg != 0u) && (cfg != 0xFFu);
        const uint64_t bal = __ballot(active);
        if (bal == 0ull) continue;
        if (lane == 0)
          a.gmap[(size_t)seg * W + c] = make_uint4((uint32_t)bal, (uint32_t)(bal >> 32), cnt_a, 0u);
        uint32_t nv = 0, nq = 0, lut2 = 0;
        if (active) {
          const uint32_t info = s_info[cfg];
          lut2 = use_variant1(g, s_info, info, x, y, k);
          const uint32_t b6 = ((cfg >> 2) ^ (cfg >> 6)) & 1u;
          const uint32_t b9 = ((cfg >> 5) ^ (cfg >> 6)) & 1u;
          const uint32_t b10 = ((cfg >> 7) ^ (cfg >> 6)) & 1u;
          nv = (lut2 ? ((info >> 3) & 7u) : (info & 7u)) + b6 + b9 + b10;
          const bool xin = x + 1 < g.cx, yin = y + 1 < g.cy, zin = k + 1 < g.cz;
          nq = (b6 & (uint32_t)(xin && yin)) + (b9 & (uint32_t)(xin && zin)) +
               (b10 & (uint32_t)(yin && zin));
        }
        const uint32_t packed = nv | (nq << 16);
        const uint32_t incl = wave_incl_scan(packed);
        const uint32_t excl = incl - packed;
        const uint32_t tot = __shfl(incl, kWave - 1, kWave);
        if (active) {
          const uint32_t rank = cnt_a + mbcnt64(bal);
          const uint32_t lvb = cnt_v + (excl & 0xFFFFu);
          const uint32_t lqb = cnt_q + (excl >> 16);
          a.slots[(size_t)seg * CAP + rank] = make_record(x - x0, cfg, lut2, lvb, lqb);
        }
        cnt_a += (uint32_t)__popcll(bal);
        cnt_v += tot & 0xFFFFu;
        cnt_q += tot >> 16;
      }
      if (lane == 0) a.seg_counts[seg] = make_uint2(cnt_a | (cnt_q << 16), cnt_v);
    }
  }
}

// Plane-tile classify: the main kernel for rows that are a multiple of 4
// points.  A 256-thread workgroup owns one tile = R cell rows x one x-tile
// (256*W4 cells) x KP cell planes.  Workgroups are dispatched round-robin over
// the 8 XCDs (block b runs on XCD b % 8), each with its own L2.  A column is
// one x-tile of one row block; columns form groups of B (the band option).
// In every plane block, XCD x takes groups x, x+8, x+16, ... and walks them
// in memory order, plane block after plane block, so the point plane shared
// by two plane blocks is read twice by the same XCD a few tiles apart and
// comes from its L2, and all XCDs stream through the same few planes at a
// time (a small working set for the address translation caches).  The groups
// left over when their count is not a multiple of 8 are dealt out over the
// XCDs eight plane blocks at a time, so every XCD gets the same number of
// tiles (within one group) and none of them finishes late.
//
// Phase 1: wave w converts point rows r = w, w+4, ... (R+1 of them, so every
// wave has the same share) of all KP+1 point planes.  A lane owns 4
// consecutive points of a 256-point chunk (one 16-byte load from a uniform
// row address plus a fixed lane offset); it packs their sign bits and the
// first bit of lane l+1 into a 5-bit value v5, pairs the values of planes
// p and p+1 into 16 bits and stores that pair into the LDS word of cell row
// r (low half) and of cell row r-1 (high half) of cell plane p:
//   byte 0: point row j,   plane k     byte 1: point row j,   plane k+1
//   byte 2: point row j+1, plane k     byte 3: point row j+1, plane k+1
// (the LDS has one padding row above and below, so every store is
// unconditional and every address is the wave's base plus a constant).
// Phase 2: each wave takes cell rows w, w+4, ...; one LDS read gives all 20
// corner bits of the lane's 4 cells.  For cell q, t = (w >> q) & 0x03030303
// holds the corner pairs (p0,p1), (p4,p5), (p3,p2), (p7,p6) in its bytes and
// (t * 0x41041) >> 18 gathers them into an 8-bit key, which indexes a table of
// configuration and counts.
constexpr int kPtThreads = 256;

// key bits p0 p1 p4 p5 p3 p2 p7 p6 -> configuration bits p0..p7
__host__ __device__ constexpr uint32_t key_to_cfg(uint32_t key) {
  return (key & 0x03u) | ((key >> 3) & 0x04u) | ((key >> 1) & 0x08u) | ((key << 2) & 0x30u) |
         ((key & 0x40u) << 1) | ((key & 0x80u) >> 1);
}

struct PtArgs {
  ClassifyArgs c;
  uint32_t B;    // columns per group
  uint32_t NKB;  // plane blocks
  uint32_t NRB;  // row blocks
};

// Cell rows per plane tile (R + 1 point rows, 2 per wave) and cell planes per
// tile (KP + 1 point planes), chosen so that all loads of a wave fit in
// registers: (R+1)/4 * (KP+1) * W4 float4 per lane.
constexpr uint32_t pt_rows(uint32_t) { return 7u; }
constexpr uint32_t pt_planes(uint32_t W4) { return W4 >= 4 ? 1u : (W4 == 2 ? 2u : 4u); }

template <int W4, int R, int KP>
struct PtShape {
  static_assert(W4 >= 1 && W4 <= 4, "x-tile is at most 1024 cells");
  static_assert((R + 1) % 4 == 0 && R <= 31, "point rows are split evenly over 4 waves");
  static_assert(KP >= 1 && KP <= 4, "one to four cell planes per tile");
  static constexpr uint32_t kRowWords = W4 * kWave;           // LDS words per cell row
  static constexpr uint32_t kPlaneWords = kRowWords * (R + 2);  // incl. two padding rows
  static constexpr uint32_t kLdsWords = KP * kPlaneWords;
};

// Two tables indexed by key, then the staging of a tile's row counts
// (kPtRowsMax uint2):
//   s_tab[key]        c_cfg_count of the key's configuration | cfg << 16
//   s_tab[256 + key]  an interior cell's contribution, variant 0, all three
//                     owned edges' rings in the grid: 1 | nv << 9 | nq << 20 (the
//                     packed counts the wave scan adds up) | (nv | nq << 3) << 24
//                     (its record byte); 0 for the inactive keys 0 and 255
constexpr uint32_t kPtRowsMax = 64;
constexpr uint32_t kPtTabWords = 512 + 2 * kPtRowsMax;
// This thread's entry of the first table (a global load) ...
__device__ __forceinline__ uint32_t pt_table_entry() {
  const uint32_t cfg = key_to_cfg(threadIdx.x);
  return (uint32_t)c_cfg_count[cfg] | (cfg << 16);
}
// ... its entry of the second, from it ...
__device__ __forceinline__ uint32_t pt_table2_entry(uint32_t tab) {
  const uint32_t cfg = tab >> 16, nv = tab & 0xFu, nq = (uint32_t)__popc((tab >> 8) & 7u);
  if (cfg == 0u || cfg == 0xFFu) return 0u;
  return 1u | (nv << 9) | (nq << 20) | ((nv | (nq << 3)) << 24);
}
// ... and both tables in LDS (the caller synchronises).
__device__ __forceinline__ void pt_store_tables(uint32_t *s_tab, uint32_t tab) {
  s_tab[threadIdx.x] = tab;
  s_tab[256 + threadIdx.x] = pt_table2_entry(tab);
}
__device__ __forceinline__ void pt_init_table(uint32_t *s_tab) { pt_store_tables(s_tab, pt_table_entry()); }

// Zeroes the state the scan expects to start from (first workgroup only).
__device__ __forceinline__ void zero_scan_state(const ClassifyArgs &a, uint32_t tid, uint32_t nthreads) {
  for (uint32_t i = tid; i < a.scan_state_words; i += nthreads) a.scan_state[i] = 0ull;
  for (uint32_t i = tid; i < a.seg_pad_n; i += nthreads) a.seg_pad[i] = make_uint2(0u, 0u);
  if (tid < 8) reinterpret_cast<uint32_t *>(a.totals)[tid] = 0u;
  zero_tickets(a.scan_ticket, tid);
}

// One plane tile (x-tile xt, row block rb, plane block kb) of the volume `a`.
// VEC: rows start on 16-byte boundaries and hold a multiple of 4 points, so a
// lane's 4 points are one 16-byte load; otherwise 4 separate loads.
template <int W4, int R, int KP, bool VEC>
__device__ __forceinline__ void classify_pt_tile(const ClassifyArgs &a, uint32_t xt, uint32_t rb,
                                                 uint32_t kb, uint32_t *s_w, uint32_t *s_tab,
                                                 uint32_t tab_entry = 0u, bool store_tab = false) {
  using S = PtShape<W4, R, KP>;
  static_assert(KP * R <= kPtRowsMax, "row counts are staged in kPtRowsMax LDS entries");
  uint2 *s_cnt = reinterpret_cast<uint2 *>(s_tab + 512);
  constexpr int kRowsPerWave = (R + 1) / 4;
  constexpr uint32_t CAP = 256u * W4;
  constexpr uint32_t kRowWords = S::kRowWords;
  constexpr uint32_t kPlaneWords = S::kPlaneWords;
  const uint32_t lane = lane_id(), wave = wave_id();
  const FieldDesc g = a.g;
  const float iso = g.iso;
  const uint32_t k = kb * KP;  // first cell plane of the tile
  const uint32_t j0 = rb * R;
  const bool tile_ok = (k < g.cz) && (j0 < g.cy);  // padding tiles do nothing
  const uint32_t x0 = xt * CAP;
  const size_t plane = (size_t)g.X * g.Y;
  const bool has_extra = x0 + CAP < g.X;

  // ---- phase 1.  All loads are issued first, from in-bounds addresses:
  // rows and planes past the grid are clamped to its last row / plane and
  // lanes past the row end to its last 4 points; the cells that would use
  // such values lie outside the grid and are never classified.
  if (tile_ok) {
    float4 v[kRowsPerWave][KP + 1][W4];
    float ve[kRowsPerWave][KP + 1];
    uint32_t xob[W4][VEC ? 1 : 4];  // byte offsets of the lane's points in a row
#pragma unroll
    for (int c = 0; c < W4; ++c) {
      if constexpr (VEC) {
        xob[c][0] = min(x0 + 256u * c + 4u * lane, g.X - 4u) * 4u;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) xob[c][e] = min(x0 + 256u * c + 4u * lane + (uint32_t)e, g.X - 1u) * 4u;
      }
    }
    const uint32_t xe = has_extra ? x0 + CAP : 0u;
#pragma unroll
    for (int i = 0; i < kRowsPerWave; ++i) {
      const size_t yoff = (size_t)min(j0 + wave + 4u * i, g.Y - 1u) * g.X;
#pragma unroll
      for (int p = 0; p <= KP; ++p) {
        const float *rowp = g.f + (size_t)min(k + p, g.Z - 1u) * plane + yoff;
        const char *rowb = reinterpret_cast<const char *>(rowp);
#pragma unroll
        for (int c = 0; c < W4; ++c) {
          if constexpr (VEC) {
            v[i][p][c] = *reinterpret_cast<const float4 *>(rowb + xob[c][0]);
          } else {
            v[i][p][c].x = *reinterpret_cast<const float *>(rowb + xob[c][0]);
            v[i][p][c].y = *reinterpret_cast<const float *>(rowb + xob[c][1]);
            v[i][p][c].z = *reinterpret_cast<const float *>(rowb + xob[c][2]);
            v[i][p][c].w = *reinterpret_cast<const float *>(rowb + xob[c][3]);
          }
        }
        ve[i][p] = rowp[xe];
      }
    }
    uint16_t *s_h = reinterpret_cast<uint16_t *>(s_w);
    const uint32_t base = wave * kRowWords + lane;
#pragma unroll
    for (int i = 0; i < kRowsPerWave; ++i) {
      uint32_t v5[KP + 1][W4];
#pragma unroll
      for (int p = 0; p <= KP; ++p) {
        uint32_t nib[W4];
#pragma unroll
        for (int c = 0; c < W4; ++c)
          nib[c] = (uint32_t)(v[i][p][c].x < iso) | ((uint32_t)(v[i][p][c].y < iso) << 1) |
                   ((uint32_t)(v[i][p][c].z < iso) << 2) | ((uint32_t)(v[i][p][c].w < iso) << 3);
        const uint32_t ext = (uint32_t)(ve[i][p] < iso);
#pragma unroll
        for (int c = 0; c < W4; ++c) {
          // first point of lane+1, or of the next chunk / next x-tile for lane 63
          uint32_t nb = lane_next(nib[c]);
          const uint32_t nxt =
              (c + 1 < W4) ? (uint32_t)__builtin_amdgcn_readfirstlane(nib[(c + 1 < W4) ? c + 1 : c]) : ext;
          if (lane == kWave - 1) nb = nxt;
          v5[p][c] = nib[c] | ((nb & 1u) << 4);
        }
      }
#pragma unroll
      for (int cp = 0; cp < KP; ++cp) {
#pragma unroll
        for (int c = 0; c < W4; ++c) {
          const uint16_t h = (uint16_t)(v5[cp][c] | (v5[cp + 1][c] << 8));
          // cell row r = wave + 4i sits at padded row r + 1; cell row r - 1 at r
          s_h[2u * (base + cp * kPlaneWords + (4u * i + 1u) * kRowWords + c * kWave)] = h;
          s_h[2u * (base + cp * kPlaneWords + (4u * i) * kRowWords + c * kWave) + 1u] = h;
        }
      }
    }
  }
  // the table entry was loaded before the field rows; storing it here keeps
  // its wait behind theirs
  if (store_tab) pt_store_tables(s_tab, tab_entry);
  __syncthreads();
  if (!tile_ok) return;  // uniform over the workgroup

  // ---- phase 2: each wave classifies cell rows wave, wave+4, ... of the
  // KP*R rows of the tile
  uint32_t xvalid[W4];  // bit q: cell x0 + 256c + 4 lane + q lies in the grid
#pragma unroll
  for (int c = 0; c < W4; ++c) {
    const uint32_t xb = x0 + 256u * c + 4u * lane;
    xvalid[c] = xb >= g.cx ? 0u : (g.cx - xb >= 4u ? 0xFu : (1u << (g.cx - xb)) - 1u);
  }
  for (uint32_t m = wave; m < (uint32_t)(KP * R); m += kPtThreads / kWave) {
    const uint32_t cp = m / R, r = m - cp * R;
    const uint32_t y = j0 + r, z = k + cp;
    if (y >= g.cy || z >= g.cz) continue;
    const bool yin = y + 1 < g.cy, zin = z + 1 < g.cz;
    uint32_t cnt_a = 0, cnt_v = 0, cnt_q = 0;
    const uint32_t seg = xt + a.XT * (y + g.cy * z);
#pragma unroll
    for (int c = 0; c < W4; ++c) {
      const uint32_t w = s_w[cp * kPlaneWords + (r + 1u) * kRowWords + c * kWave + lane];
      // all 20 corner bits equal -> none of the lane's 4 cells is active
      if (__ballot(w != 0u && w != 0x1F1F1F1Fu) == 0ull) continue;
      const uint32_t xb = x0 + 256u * c + 4u * lane;
      // per cell: key, and active (bit 0 of the second table entry)
      uint32_t keys = 0, lact = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t tq = (w >> q) & 0x03030303u;
        const uint32_t key = ((tq * 0x41041u) >> 18) & 0xFFu;
        keys |= key << (8 * q);
        lact |= (s_tab[256u + key] & 1u) << q;
      }
      lact &= xvalid[c];
      if (__ballot(lact != 0u) == 0ull) continue;
      // first table entry of each active cell (0 for inactive cells, so they
      // count nothing)
      uint32_t e4[4];
      uint32_t amb = 0;  // bit q: cell q is a 2B/3B configuration
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t ent = s_tab[(keys >> (8 * q)) & 0xFFu];
        e4[q] = ((lact >> q) & 1u) ? ent : 0u;
        amb |= ((e4[q] >> 11) & 1u) << q;
      }
      const bool any_amb = __ballot(amb != 0u) != 0ull;
      // 2B/3B cells (rare): the variant depends on the neighbour across the
      // listed face.  Inside the tile the neighbour's corner bits are in LDS
      // (the word one cell or lane over, or of the next row or plane); only a
      // neighbour outside the tile is read from the field.
      uint32_t l2m = 0;
      if (any_amb) {
        const uint32_t widx = cp * kPlaneWords + (r + 1u) * kRowWords + c * kWave + lane;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (!((amb >> q) & 1u)) continue;
          const uint32_t face = (e4[q] >> 12) & 7u;
          int nx = (int)(xb + q), ny = (int)y, nz = (int)z;
          bool in_tile;
          uint32_t nidx = widx, nq = (uint32_t)q;
          switch (face) {
            case 0:
              nx += 1;
              in_tile = q < 3 || lane + kWave * c + 1u < kWave * W4;
              nidx = widx + (q == 3 ? 1u : 0u);
              nq = (uint32_t)(q + 1) & 3u;
              break;
            case 1:
              nx -= 1;
              in_tile = q > 0 || lane + kWave * c > 0u;
              nidx = widx - (q == 0 ? 1u : 0u);
              nq = (uint32_t)(q + 3) & 3u;
              break;
            case 2: ny += 1; in_tile = r + 1u < (uint32_t)R; nidx = widx + kRowWords; break;
            case 3: ny -= 1; in_tile = r > 0u; nidx = widx - kRowWords; break;
            case 4: nz += 1; in_tile = cp + 1u < (uint32_t)KP; nidx = widx + kPlaneWords; break;
            default: nz -= 1; in_tile = cp > 0u; nidx = widx - kPlaneWords; break;
          }
          if (nx < 0 || ny < 0 || nz < 0 || nx >= (int)g.cx || ny >= (int)g.cy || nz >= (int)g.cz) continue;
          uint32_t ent;
          if (in_tile) {
            const uint32_t tq = (s_w[nidx] >> nq) & 0x03030303u;
            ent = s_tab[((tq * 0x41041u) >> 18) & 0xFFu];
          } else {
            ent = c_cfg_count[cell_cfg_global(g, (uint32_t)nx, (uint32_t)ny, (uint32_t)nz)];
          }
          l2m |= ((ent >> 11) & 1u) << q;
        }
      }
      // per cell: variant, vertex count, quad count; per-cell byte: nv (3
      // bits) | nq (2 bits) << 3 | variant << 5
      uint32_t pk = 0, info4 = 0, cfgs = 0;
      if (yin && zin && x0 + 256u * (c + 1u) < g.cx && !any_amb) {
        // interior chunk (every cell and its owned edges' rings in the grid)
        // without 2B/3B cells: the second table has it all
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t e2 = s_tab[256u + ((keys >> (8 * q)) & 0xFFu)];
          cfgs |= (e4[q] >> 16) << (8 * q);
          pk += e2 & 0x00FFFFFFu;
          info4 |= (e2 >> 24) << (8 * q);
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t e = e4[q] & 0xFFFFu, l2 = (l2m >> q) & 1u;
          cfgs |= (e4[q] >> 16) << (8 * q);
          const uint32_t nv = (e >> (4 * l2)) & 0xFu;
          const bool xin = xb + q + 1 < g.cx;
          const uint32_t qmask = (uint32_t)(xin && yin) | ((uint32_t)(xin && zin) << 1) |
                                 ((uint32_t)(yin && zin) << 2);
          const uint32_t nq = (uint32_t)__popc((e >> 8) & 7u & qmask);
          info4 |= (nv | (nq << 3) | (l2 << 5)) << (8 * q);
          pk += ((lact >> q) & 1u) | (nv << 9) | (nq << 20);
        }
      }
      const uint32_t incl = wave_incl_scan_dpp(pk);
      const uint32_t excl = incl - pk;
      const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)incl, kWave - 1);
      {
        // Group map: the 16 lanes of a DPP row hold the 64 cells of one group
        // (4 per lane); OR their nibbles into lane 15 of the row.
        uint32_t mlo = (lane & 8u) ? 0u : lact << (4u * (lane & 7u));
        uint32_t mhi = (lane & 8u) ? lact << (4u * (lane & 7u)) : 0u;
        mlo |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)mlo, kDppRowShr + 1, 0xF, 0xF, true);
        mhi |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)mhi, kDppRowShr + 1, 0xF, 0xF, true);
        mlo |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)mlo, kDppRowShr + 2, 0xF, 0xF, true);
        mhi |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)mhi, kDppRowShr + 2, 0xF, 0xF, true);
        mlo |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)mlo, kDppRowShr + 4, 0xF, 0xF, true);
        mhi |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)mhi, kDppRowShr + 4, 0xF, 0xF, true);
        mlo |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)mlo, kDppRowShr + 8, 0xF, 0xF, true);
        mhi |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)mhi, kDppRowShr + 8, 0xF, 0xF, true);
        if ((lane & 15u) == 15u && (mlo | mhi) != 0u) {
          const uint32_t first = cnt_a + (incl & 0x1FFu) - (uint32_t)(__popc(mlo) + __popc(mhi));
          a.gmap[(size_t)seg * (CAP / 64u) + 4u * c + (lane >> 4)] = make_uint4(mlo, mhi, first, 0u);
        }
      }
      if (lact) {
        uint32_t rank = cnt_a + (excl & 0x1FFu);
        uint32_t lvb = cnt_v + ((excl >> 9) & 0x7FFu);
        uint32_t lqb = cnt_q + (excl >> 20);
        uint64_t *dst = a.slots + (size_t)seg * CAP;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (!((lact >> q) & 1u)) continue;
          const uint32_t b = (info4 >> (8 * q)) & 0xFFu;
          dst[rank] = make_record(xb + q - x0, (cfgs >> (8 * q)) & 0xFFu, b >> 5, lvb, lqb);
          rank += 1;
          lvb += b & 7u;
          lqb += (b >> 3) & 3u;
        }
      }
      cnt_a += tot & 0x1FFu;
      cnt_v += (tot >> 9) & 0x7FFu;
      cnt_q += tot >> 20;
    }
    if (lane == 0) s_cnt[m] = make_uint2(cnt_a | (cnt_q << 16), cnt_v);
  }
  // the tile's row counts leave in one store (a store per row costs more)
  __syncthreads();
  if (threadIdx.x < (uint32_t)(KP * R)) {
    const uint32_t m = threadIdx.x, cp = m / R, r = m - cp * R;
    const uint32_t y = j0 + r, z = k + cp;
    if (y < g.cy && z < g.cz) a.seg_counts[xt + a.XT * (y + g.cy * z)] = s_cnt[m];
  }
}

// Single volume: block b -> tile through the XCD mapping (pt_tile_of).
template <int W4, int R, int KP, bool VEC>
__global__ __launch_bounds__(kPtThreads) __attribute__((amdgpu_waves_per_eu(W4 >= 4 ? 4 : 8, W4 >= 4 ? 4 : 8))) void k_classify_pt(PtArgs pa) {
  __shared__ uint32_t s_w[PtShape<W4, R, KP>::kLdsWords];
  __shared__ __align__(16) uint32_t s_tab[kPtTabWords];
  const ClassifyArgs &a = pa.c;
  const uint32_t tab = pt_table_entry();
  if (blockIdx.x == 0) {
    zero_scan_state(a, threadIdx.x, kPtThreads);
    if (a.counts_out != nullptr && threadIdx.x < 4) a.counts_out[threadIdx.x] = 0u;
  }
  const PtTile t = pt_tile_of(blockIdx.x, a.XT, pa.NRB, pa.NKB, pa.B);
  classify_pt_tile<W4, R, KP, VEC>(a, t.xt, t.rb, t.kb, s_w, s_tab, tab, true);
}

// ---------------------------------------------------------------------------
// Batches: many volumes in one launch sequence.  The segments of all volumes
// are concatenated (volume v owns segments [seg_off, seg_off + nsegs_v)), so
// the scan runs once over the whole batch and a volume's local vertex / quad
// numbers are the global ones minus the prefix at its first segment.

constexpr uint32_t kMaxBatch = 1024;  // volumes per batch call

struct BatchVol {
  ClassifyArgs c;         // per-volume field and tiling; record pointers start at seg_off
  uint32_t tile_off;      // first flattened classify tile
  uint32_t NRB, NKP;      // row blocks and plane blocks of the volume
  uint32_t vec;           // rows are 16-byte aligned multiples of 4 points
  uint32_t seg_off, nsegs;
  float mnx, mny, mnz, rgx, rgy, rgz;
  float *verts;
  uint64_t vcap;
  int32_t *quads;
  uint64_t qcap;
  uint32_t *counts;       // the volume's mdmc_counts
  uint32_t vbase, qbase;  // written by k_batch_counts: prefix at seg_off
  float *coords;          // X + Y + Z point coordinates (filled before the emit)
};

struct BatchArgs {
  BatchVol *vols;
  uint32_t nvol, total_tiles;
  ClassifyArgs zero;      // scan state of the batch (zeroed by block 0)
};

// Classify over the flattened tiles of all volumes (tiles of a volume in
// memory order): with a grid of at least the tile count a workgroup per tile,
// with a smaller grid each workgroup walks several (persistent).
template <int W4, int R, int KP>
__global__ __launch_bounds__(kPtThreads) __attribute__((amdgpu_waves_per_eu(W4 == 1 ? 8 : (W4 == 2 ? 6 : 4), W4 == 1 ? 8 : (W4 == 2 ? 6 : 4)))) void k_classify_batch(BatchArgs b) {
  __shared__ uint32_t s_w[PtShape<W4, R, KP>::kLdsWords];
  __shared__ __align__(16) uint32_t s_tab[kPtTabWords];
  __shared__ uint32_t s_toff[kMaxBatch + 1];
  pt_init_table(s_tab);
  for (uint32_t i = threadIdx.x; i <= b.nvol; i += kPtThreads)
    s_toff[i] = i < b.nvol ? b.vols[i].tile_off : b.total_tiles;
  if (blockIdx.x == 0) {
    zero_scan_state(b.zero, threadIdx.x, kPtThreads);
    if (threadIdx.x < 4) b.zero.counts_out[threadIdx.x] = 0u;  // the batch's scan counts
  }
  __syncthreads();
  // XCD x (workgroups b with b % 8 == x) takes a contiguous eighth of the
  // flattened tiles, so tiles that share halo rows and planes meet in its L2
  const uint32_t G = gridDim.x;
  uint32_t t = blockIdx.x, t_end = b.total_tiles, t_step = G;
  if (G >= 8u) {
    const uint32_t x = blockIdx.x & 7u;
    t = (uint32_t)((uint64_t)b.total_tiles * x / 8u) + (blockIdx.x >> 3);
    t_end = (uint32_t)((uint64_t)b.total_tiles * (x + 1u) / 8u);
    t_step = (G >> 3) + (x < (G & 7u) ? 1u : 0u);
  }
  for (; t < t_end; t += t_step) {
    // volume of tile t: the last v with tile_off[v] <= t
    uint32_t lo = 0, hi = b.nvol;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_toff[mid] <= t) lo = mid; else hi = mid;
    }
    const uint32_t v = __builtin_amdgcn_readfirstlane(lo);
    const BatchVol &bv = b.vols[v];
    const ClassifyArgs &a = bv.c;
    const uint32_t local = t - s_toff[v];
    const uint32_t xt = local % a.XT, u = local / a.XT;
    __syncthreads();  // the previous tile's LDS words are no longer read
    if (bv.vec)
      classify_pt_tile<W4, R, KP, true>(a, xt, u % bv.NRB, u / bv.NRB, s_w, s_tab);
    else
      classify_pt_tile<W4, R, KP, false>(a, xt, u % bv.NRB, u / bv.NRB, s_w, s_tab);
  }
}

// utils::ijk_to_xyz (include/utils.hpp:118-122): i / n * range + min, three
// correctly rounded float operations.
__device__ __forceinline__ float coordf(uint32_t i, uint32_t n, float range, float mn) {
  return (float)i / (float)n * range + mn;
}

// Coordinates of the point indices of a volume, per axis (X, then Y, then Z
// entries): the same three float operations per entry as the reference.
struct CoordArgs {
  float *out;
  uint32_t X, Y, Z, cx, cy, cz;
  float mnx, mny, mnz, rgx, rgy, rgz;
  uint32_t z_off;  // global index of the first z point (a slab of a larger grid; cz is then the grid's)
};

__device__ __forceinline__ void coords_fill(const CoordArgs &c, uint32_t i0, uint32_t stride) {
  const uint32_t n = c.X + c.Y + c.Z;
  for (uint32_t i = i0; i < n; i += stride) {
    float v;
    if (i < c.X) v = coordf(i, c.cx, c.rgx, c.mnx);
    else if (i < c.X + c.Y) v = coordf(i - c.X, c.cy, c.rgy, c.mny);
    else v = coordf(i - c.X - c.Y + c.z_off, c.cz, c.rgz, c.mnz);
    c.out[i] = v;
  }
}

__global__ __launch_bounds__(256) void k_coords(CoordArgs c) {
  coords_fill(c, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// ---------------------------------------------------------------------------
// Decoupled look-back scan over segment totals.  A tile is 4096 segments, 16
// per thread read as eight 16-byte loads (the totals array is padded to a
// whole number of tiles with zeros).  Four channels are scanned: active
// cells, vertices, quads and non-empty segments.  Outputs: the first vertex
// and quad of every segment; the list of non-empty segments with the index
// of their first active cell (ne[i] = {segment, first active}, ne[NE] =
// {nsegs, A}); for every chunk of 64 active cells the list position of the
// segment holding its first cell; the grand totals.

constexpr int kScanThreads = 256;
constexpr int kScanIpt = 8;  // segments per thread (16: 1 us slower per 512^3 extraction)
constexpr int kScanTile = kScanThreads * kScanIpt;
constexpr uint32_t kSpinLimit = 1u << 24;

struct ScanArgs {
  const uint2 *seg_counts;  // padded to ntiles * kScanTile entries
  uint32_t nsegs, ntiles;
  uint64_t max_chunks;      // capacity of chunk_ne
  uint2 *seg_vq;            // per segment: first vertex, first quad (same padding)
  uint2 *ne;                // non-empty segments: {segment, first active cell}; nsegs + 1
  uint32_t *chunk_ne;       // per 64 active cells: index into ne of its first cell's segment
  unsigned long long *state;  // 4 words per tile: (status << 32) | value
  uint32_t *ticket;
  Totals *totals;
  uint32_t *counts_out;  // optional user mdmc_counts
  CoordArgs co;          // coordinate tables, filled by the scan's workgroups
  uint32_t fill_coords;  // when the box is known at count time
};

__device__ __forceinline__ unsigned long long ld_state(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_state(unsigned long long *p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kScanThreads) void k_scan(ScanArgs a) {
  __shared__ uint32_t s_tile;
  __shared__ uint4 s_wave[kScanThreads / kWave];
  __shared__ uint4 s_prefix;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = wave_id();
  if (tid == 0) s_tile = atomicAdd(a.ticket, 1u);
  __syncthreads();
  const uint32_t t = s_tile;
  const uint32_t base = t * kScanTile + tid * kScanIpt;

  uint4 c[kScanIpt / 2];
  const uint4 *src = reinterpret_cast<const uint4 *>(a.seg_counts + base);
#pragma unroll
  for (int i = 0; i < kScanIpt / 2; ++i) c[i] = src[i];
  // per segment: .x = nact | nquad << 16, .y = nvert; two segments per uint4
  uint4 sum = make_uint4(0, 0, 0, 0);
#pragma unroll
  for (int i = 0; i < kScanIpt / 2; ++i) {
    sum.x += (c[i].x & 0xFFFFu) + (c[i].z & 0xFFFFu);
    sum.y += c[i].y + c[i].w;
    sum.z += (c[i].x >> 16) + (c[i].z >> 16);
    sum.w += ((c[i].x & 0xFFFFu) ? 1u : 0u) + ((c[i].z & 0xFFFFu) ? 1u : 0u);
  }
  uint4 inc;
  inc.x = wave_incl_scan_dpp(sum.x);
  inc.y = wave_incl_scan_dpp(sum.y);
  inc.z = wave_incl_scan_dpp(sum.z);
  inc.w = wave_incl_scan_dpp(sum.w);
  if (lane == kWave - 1) s_wave[wave] = inc;
  __syncthreads();
  uint4 woff = make_uint4(0, 0, 0, 0), agg = make_uint4(0, 0, 0, 0);
#pragma unroll
  for (int w = 0; w < kScanThreads / kWave; ++w) {
    const uint4 v = s_wave[w];
    if ((uint32_t)w < wave) { woff.x += v.x; woff.y += v.y; woff.z += v.z; woff.w += v.w; }
    agg.x += v.x; agg.y += v.y; agg.z += v.z; agg.w += v.w;
  }

  if (wave == 0) {
    unsigned long long *my = a.state + 4ull * t;
    uint4 excl = make_uint4(0, 0, 0, 0);
    const unsigned long long AGG = 1ull << 32, INC = 2ull << 32;
    if (t == 0) {
      if (lane == 0) {
        st_state(my + 0, INC | agg.x); st_state(my + 1, INC | agg.y);
        st_state(my + 2, INC | agg.z); st_state(my + 3, INC | agg.w);
      }
    } else {
      if (lane == 0) {
        st_state(my + 0, AGG | agg.x); st_state(my + 1, AGG | agg.y);
        st_state(my + 2, AGG | agg.z); st_state(my + 3, AGG | agg.w);
      }
      // Each channel word is self-describing (status, value) and a tile flips
      // its four words from aggregate to inclusive one at a time, so every
      // channel is resolved on its own: sum aggregates back to the nearest
      // tile whose word in *that* channel is inclusive.
      int64_t look = (int64_t)t - 1;
      bool timed_out = false;
      uint32_t done = 0;  // bit c: channel c resolved
      while (done != 0xFu) {
        const int64_t idx = look - (int64_t)lane;
        unsigned long long w[4] = {INC, INC, INC, INC};  // before tile 0: inclusive zeros
        if (idx >= 0) {
          const unsigned long long *p = a.state + 4ull * (uint64_t)idx;
          // the four channel words are read together; only words not yet
          // published are read again
#pragma unroll
          for (int ch = 0; ch < 4; ++ch)
            if (!((done >> ch) & 1u)) w[ch] = ld_state(p + ch);
          uint32_t spins = 0;
#pragma unroll
          for (int ch = 0; ch < 4; ++ch) {
            if ((done >> ch) & 1u) continue;
            while ((w[ch] >> 32) == 0 && ++spins < kSpinLimit) w[ch] = ld_state(p + ch);
          }
          if (spins >= kSpinLimit) timed_out = true;
        }
        uint32_t add[4];
#pragma unroll
        for (int ch = 0; ch < 4; ++ch) {
          add[ch] = 0;
          if ((done >> ch) & 1u) continue;
          const uint64_t incmask = __ballot((w[ch] >> 32) == 2);
          const int stop = incmask ? (__ffsll((long long)incmask) - 1) : kWave;
          const uint32_t val = ((int)lane <= stop && idx >= 0) ? (uint32_t)w[ch] : 0u;
          add[ch] = __shfl(wave_incl_scan(val), kWave - 1, kWave);
          if (incmask) done |= 1u << ch;
        }
        excl.x += add[0]; excl.y += add[1]; excl.z += add[2]; excl.w += add[3];
        look -= kWave;
      }
      if (__ballot(timed_out) && lane == 0 && a.counts_out)
        atomicOr(a.counts_out + 3, kStatusScanTimeout);
      if (lane == 0) {
        st_state(my + 0, INC | (uint32_t)(excl.x + agg.x));
        st_state(my + 1, INC | (uint32_t)(excl.y + agg.y));
        st_state(my + 2, INC | (uint32_t)(excl.z + agg.z));
        st_state(my + 3, INC | (uint32_t)(excl.w + agg.w));
        // the tile where a 32-bit total first passes 2^32 - 1 sees its
        // inclusive sum wrap (vertex and quad numbers are 32-bit, as the
        // reference's)
        const bool wraps = excl.x + agg.x < excl.x || excl.y + agg.y < excl.y || excl.z + agg.z < excl.z;
        if (wraps && a.counts_out) atomicOr(a.counts_out + 3, kStatusIndexRange);
      }
    }
    if (lane == 0) s_prefix = excl;
  }
  __syncthreads();
  const uint4 tp = s_prefix;
  // exclusive prefix of this thread's first segment
  uint32_t pa = tp.x + woff.x + inc.x - sum.x;
  uint32_t pv = tp.y + woff.y + inc.y - sum.y;
  uint32_t pq = tp.z + woff.z + inc.z - sum.z;
  uint32_t pe = tp.w + woff.w + inc.w - sum.w;
  bool bad = false;
  uint4 *dst = reinterpret_cast<uint4 *>(a.seg_vq + base);
#pragma unroll
  for (int i = 0; i < kScanIpt / 2; ++i) {
    uint4 o;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const uint32_t cx_ = h ? c[i].z : c[i].x, cy_ = h ? c[i].w : c[i].y;
      const uint32_t n = cx_ & 0xFFFFu;
      if (h) { o.z = pv; o.w = pq; } else { o.x = pv; o.y = pq; }
      if (n) {
        const uint32_t sg = base + 2 * i + h;
        if (pe < a.nsegs) a.ne[pe] = make_uint2(sg, pa); else bad = true;
        for (uint32_t gch = (pa + 63u) >> 6; (gch << 6) < pa + n; ++gch) {
          if (gch < a.max_chunks) a.chunk_ne[gch] = pe; else bad = true;
        }
        ++pe;
      }
      pa += n;
      pv += cy_;
      pq += cx_ >> 16;
    }
    dst[i] = o;
  }
  if (t == a.ntiles - 1 && tid == kScanThreads - 1) {
    // pa/pv/pq/pe now hold the grand totals
    Totals tot;
    tot.num_active = pa;
    tot.num_vertices = pv;
    tot.num_quads = pq;
    tot.num_nonempty = pe;
    tot.num_chunks = (pa + 63u) >> 6;
    tot.pad0 = tot.pad1 = tot.pad2 = 0;
    *a.totals = tot;
    if (pe <= a.nsegs) a.ne[pe] = make_uint2(a.nsegs, pa); else bad = true;
    if (a.counts_out) {
      a.counts_out[0] = pa;
      a.counts_out[1] = pv;
      a.counts_out[2] = pq;
    }
  }
  if (bad && a.counts_out) atomicOr(a.counts_out + 3, kStatusInternal);
  if (a.fill_coords) coords_fill(a.co, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// ---------------------------------------------------------------------------
// Emit: vertices and quads.  Each wave takes chunks of 64 consecutive active
// cells (in cell order), one lane per cell, and finds each cell's segment in
// the non-empty segment list (the chunk's first list entry comes from the
// scan; a window of 64 entries is searched with shuffles).  Per cell: edge
// crossing points, dual vertices (FP32, reference operation order) and
// quads.  A ring
// neighbour of an owned edge is active by construction; its record is found
// in O(1): the +x neighbour is the next record of the same segment (or the
// first of the next x-tile's segment), and in a neighbour row the rank of
// cell x is the group's first rank plus the population count of the group's
// active mask below x.

constexpr int kEmitThreads = 256;

// Per-volume part of the emit arguments.
struct EmitVol {
  FieldDesc g;
  const float *xt_c, *yt_c, *zt_c;     // coordinate of every point index per axis
  uint32_t XT, seg_off;                // x-tiles; first segment of the volume
  uint32_t vbase, qbase;               // global vertex / quad number of its first
  float *verts;
  uint64_t vcap;
  int32_t *quads;
  uint64_t qcap;
  uint32_t *status;
  // emit only cell planes [kmin, kmax); quad entries are vertex numbers plus
  // vshift (a slab of a larger grid, numbered globally)
  uint32_t kmin = 0, kmax = 0xFFFFFFFFu, vshift = 0;
};

struct EmitArgs {
  EmitVol v;  // the volume (single-volume launches)
  uint32_t cap, nsegs;
  uint64_t max_chunks;  // capacity of chunk_ne
  const uint2 *seg_counts;
  const uint2 *seg_vq;
  const uint64_t *slots;
  const uint4 *gmap;
  const uint2 *ne;
  const uint32_t *chunk_ne;
  const Totals *totals;
  uint32_t *emit_ctr;  // k_emit's chunk tickets: the ticket region (zero at launch)
};

// Arrays seen through the constant address space: loads at wave-uniform
// indices become scalar loads.  Only for arrays no kernel running at the
// same time writes.
__device__ __forceinline__ const __attribute__((address_space(4))) float *as_const(const float *p) {
  return (const __attribute__((address_space(4))) float *)p;
}
// Entry i of an array of uint2 through the constant address space.
__device__ __forceinline__ uint2 const_u2(const uint2 *p, uint32_t i) {
  const __attribute__((address_space(4))) uint32_t *w = (const __attribute__((address_space(4))) uint32_t *)p;
  return make_uint2(w[2 * i], w[2 * i + 1]);
}
// Its first word only.
__device__ __forceinline__ uint32_t const_u2x(const uint2 *p, uint32_t i) {
  return ((const __attribute__((address_space(4))) uint32_t *)p)[2 * i];
}

// One active cell (segment s, rank within it) of volume vol: its vertices and
// the quads of its owned edges.
__device__ __forceinline__ void emit_cell(const EmitArgs &a, const EmitVol &vol, uint32_t s, uint32_t rank,
                                          const uint64_t *s_lab, float *s_acc) {
  const FieldDesc &g = vol.g;
  const float iso = g.iso;
  const size_t sy = g.X, sz = (size_t)g.X * g.Y;
  const uint32_t cap = a.cap;
  const uint64_t rec = a.slots[(size_t)s * cap + rank];
  const uint32_t xl = rec_x(rec), cfg = rec_cfg(rec), lut2 = rec_lut2(rec);
  const uint32_t sl = s - vol.seg_off;
  const uint32_t xt = sl % vol.XT, row = sl / vol.XT;
  const uint32_t j = row % g.cy, k = row / g.cy;
  const uint32_t x = xt * cap + xl;
  if (x >= g.cx || xl >= cap) { atomicOr(vol.status, kStatusInternal); return; }
  if (k < vol.kmin || k >= vol.kmax) return;
  // The lanes of a chunk nearly always share one segment.  What depends only
  // on the segment (its bases and count, the row's y and z coordinates, the
  // ring segments' vertex bases below) comes through scalar loads for the
  // lanes in the wave's first segment and through vector loads for the rest
  // (these arrays are read-only in the emit).
  const uint32_t s0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)s);
  // s0u: s0 in a scalar register that the compiler cannot equate with s, so
  // that on the lanes where s == s0 it does not turn the loads below back
  // into per-lane vector loads at s
  uint32_t s0u = s0;
  asm volatile("" : "+s"(s0u));
  const uint32_t row0 = (s0u - vol.seg_off) / vol.XT;
  const uint32_t j0 = row0 % g.cy, k0 = row0 / g.cy;
  // (scalar loads first, outside any branch on the lane's segment)
  uint2 vq = const_u2(a.seg_vq, s0u);
  uint32_t nact = const_u2x(a.seg_counts, s0u) & 0xFFFFu;
  float Y0 = as_const(vol.yt_c)[j0], Y1 = as_const(vol.yt_c)[j0 + 1];
  float Z0 = as_const(vol.zt_c)[k0], Z1 = as_const(vol.zt_c)[k0 + 1];
  const bool lead = s == s0;
  if (!lead) {
    vq = a.seg_vq[s];
    nact = a.seg_counts[s].x & 0xFFFFu;
    Y0 = vol.yt_c[j];
    Y1 = vol.yt_c[j + 1];
    Z0 = vol.zt_c[k];
    Z1 = vol.zt_c[k + 1];
  }
  const uint32_t vb = vq.x + rec_lvb(rec) - vol.vbase;  // volume-local numbers
  const uint32_t qb = vq.y + rec_lqb(rec) - vol.qbase;

  const float *p = g.f + (size_t)k * sz + (size_t)j * sy + x;
  float v[8];
  v[0] = p[0];
  v[1] = p[1];
  v[2] = p[sy + 1];
  v[3] = p[sy];
  v[4] = p[sz];
  v[5] = p[sz + 1];
  v[6] = p[sz + sy + 1];
  v[7] = p[sz + sy];

  const float X0 = vol.xt_c[x], X1 = vol.xt_c[x + 1];
  // a rank past the segment's count means inconsistent count-pass data
  if (rank >= nact) { atomicOr(vol.status, kStatusInternal); return; }

  // Edge e runs from corner EF[e] to corner ET[e]: the orientation its
  // owning cell uses (owner edge 6: p2->p6, 9: p5->p6, 10: p6->p7), so the
  // point is bitwise the one the owner writes (dmc.cu:1468-1514).
  constexpr uint8_t EF[12] = {1, 1, 2, 0, 0, 1, 2, 3, 5, 5, 6, 4};
  constexpr uint8_t ET[12] = {0, 2, 3, 3, 4, 5, 6, 7, 4, 6, 7, 7};
  // Owner of each edge lies at a non-positive offset; an edge is skipped
  // when its owner would be outside the grid (dmc.cu:1539-1545).
  // bit 0: needs x>0, bit 1: needs y>0, bit 2: needs z>0
  constexpr uint8_t NEED[12] = {6, 4, 4, 5, 3, 2, 0, 1, 2, 0, 0, 1};
  const uint32_t have = (x > 0 ? 1u : 0u) | (j > 0 ? 2u : 0u) | (k > 0 ? 4u : 0u);

  const uint64_t labels = s_lab[lut2 * 256 + cfg];
  const uint16_t info = c_cfg_info[cfg];
  const uint32_t niso = lut2 ? ((info >> 3) & 7u) : (info & 7u);
  // Group sums live in LDS, component c of group q at s_acc[(3q + c) * T]
  // (T = threads per block; the caller offsets s_acc by the thread index).
  // They start at -0.0, so the first addition yields the point itself
  // bitwise, as the reference's "assign the first, add the rest" does.
  constexpr uint32_t T = kEmitThreads;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if ((uint32_t)q < niso) {
      s_acc[(3 * q + 0) * T] = -0.0f;
      s_acc[(3 * q + 1) * T] = -0.0f;
      s_acc[(3 * q + 2) * T] = -0.0f;
    }
  }
  uint32_t cnt = 0;  // 8 bits per group
  float ex[3] = {0, 0, 0}, ey[3] = {0, 0, 0}, ez[3] = {0, 0, 0};
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    const uint32_t lab = (uint32_t)(labels >> (4 * e)) & 0xFu;
    if (lab == 0xFu) continue;
    const int f = EF[e], to = ET[e];
    const float vf = v[f], vt = v[to];
    const float tt = (iso - vf) / (vt - vf);
    const float omt = 1.0f - tt;
    const float fx = (f == 1 || f == 2 || f == 5 || f == 6) ? X1 : X0;
    const float fy = (f == 2 || f == 3 || f == 6 || f == 7) ? Y1 : Y0;
    const float fz = (f >= 4) ? Z1 : Z0;
    const float tx = (to == 1 || to == 2 || to == 5 || to == 6) ? X1 : X0;
    const float ty = (to == 2 || to == 3 || to == 6 || to == 7) ? Y1 : Y0;
    const float tz = (to >= 4) ? Z1 : Z0;
    const float px = fx * omt + tx * tt;
    const float py = fy * omt + ty * tt;
    const float pz = fz * omt + tz * tt;
    if (e == 6) { ex[0] = px; ey[0] = py; ez[0] = pz; }
    if (e == 9) { ex[1] = px; ey[1] = py; ez[1] = pz; }
    if (e == 10) { ex[2] = px; ey[2] = py; ez[2] = pz; }
    if ((NEED[e] & have) != NEED[e]) continue;
    float *acc = s_acc + 3 * lab * T;
    acc[0] = acc[0] + px;
    acc[T] = acc[T] + py;
    acc[2 * T] = acc[2 * T] + pz;
    cnt += 1u << (8 * lab);
  }
  const uint32_t b6 = ((cfg >> 2) ^ (cfg >> 6)) & 1u;
  const uint32_t b9 = ((cfg >> 5) ^ (cfg >> 6)) & 1u;
  const uint32_t b10 = ((cfg >> 7) ^ (cfg >> 6)) & 1u;
  const uint32_t nv = niso + b6 + b9 + b10;
  if ((uint64_t)vb + nv <= vol.vcap) {
    float *out = vol.verts + 3ull * vb;
    const float qnan = __int_as_float(0x7FC00000);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if ((uint32_t)q < niso) {
        const uint32_t n = (cnt >> (8 * q)) & 0xFFu;
        const float d = (float)n;
        out[3 * q + 0] = n ? s_acc[(3 * q + 0) * T] / d : qnan;
        out[3 * q + 1] = n ? s_acc[(3 * q + 1) * T] / d : qnan;
        out[3 * q + 2] = n ? s_acc[(3 * q + 2) * T] / d : qnan;
      }
    }
    uint32_t wv = niso;
    if (b6) { out[3 * wv] = ex[0]; out[3 * wv + 1] = ey[0]; out[3 * wv + 2] = ez[0]; ++wv; }
    if (b9) { out[3 * wv] = ex[1]; out[3 * wv + 1] = ey[1]; out[3 * wv + 2] = ez[1]; ++wv; }
    if (b10) { out[3 * wv] = ex[2]; out[3 * wv + 1] = ey[2]; out[3 * wv + 2] = ez[2]; ++wv; }
  } else {
    atomicOr(vol.status, kStatusVertexOverflow);
  }

  // Quads (dmc.cu:1983-2047): rings in ccw order, reversed (0,3,2,1) when
  // the owned edge's first corner is not above iso.
  const bool xin = x + 1 < g.cx, yin = j + 1 < g.cy, zin = k + 1 < g.cz;
  const bool e6 = b6 && xin && yin, e9 = b9 && xin && zin, e10 = b10 && yin && zin;
  if (!(e6 || e9 || e10)) return;
  const uint32_t nq = (uint32_t)e6 + (uint32_t)e9 + (uint32_t)e10;
  if ((uint64_t)qb + nq > vol.qcap) { atomicOr(vol.status, kStatusQuadOverflow); return; }
  // Ring cells n = 0 (+1,0,0), 1 (0,+1,0), 2 (0,0,+1), 3 (+1,+1,0),
  // 4 (+1,0,+1), 5 (0,+1,+1).  A ring cell contains the crossing edge, so it
  // is active.  Cells 1, 2 and 5 are found through their segments' group
  // maps.  Cells 0, 3 and 4 follow from this cell, 1 and 2: a ring cell and
  // its -x neighbour in the ring are both active, so the +x one is the next
  // active cell of its row (the first of the next x-tile's segment when x is
  // the tile's last cell).  Each record's x field confirms the cell found.
  // A lookup the quads do not need reads this cell's segment instead, so that
  // every load is in bounds.
  const uint32_t need = (e6 ? 0x0Bu : 0u) | (e9 ? 0x15u : 0u) | (e10 ? 0x26u : 0u);
  const bool wrap = xl + 1u == cap;  // +x lies in the next x-tile's segment
  const uint32_t xl1 = wrap ? 0u : xl + 1u;
  const uint32_t dy = vol.XT, dz = vol.XT * g.cy;
  const uint32_t s1 = (need & 0x0Au) ? s + dy : s, s2 = (need & 0x14u) ? s + dz : s;
  const uint32_t s5 = (need & 0x20u) ? s + dy + dz : s;
  const uint32_t G = cap >> 6;
  const uint4 g1 = a.gmap[(size_t)s1 * G + (xl >> 6)], g2 = a.gmap[(size_t)s2 * G + (xl >> 6)];
  const uint4 g5 = a.gmap[(size_t)s5 * G + (xl >> 6)];
  // the first segment's ring segments (clamped: a lane that needs one has it
  // in the grid), scalar loads first as above
  const uint32_t last = a.nsegs - 1u;
  uint32_t vq0[6];
  vq0[1] = const_u2x(a.seg_vq, min(s0u + dy, last));
  vq0[2] = const_u2x(a.seg_vq, min(s0u + dz, last));
  vq0[5] = const_u2x(a.seg_vq, min(s0u + dy + dz, last));
  if (!lead) {
    vq0[1] = a.seg_vq[s1].x;
    vq0[2] = a.seg_vq[s2].x;
    vq0[5] = a.seg_vq[s5].x;
  }
  vq0[0] = vq.x;
  vq0[3] = vq0[1];
  vq0[4] = vq0[2];
  if (wrap && (need & 0x19u)) {  // the next x-tile's segments of rows y, y+1 and plane z+1
    vq0[0] = a.seg_vq[s + 1u].x;
    vq0[3] = a.seg_vq[s1 + 1u].x;
    vq0[4] = a.seg_vq[s2 + 1u].x;
  }
  const uint32_t bit = xl & 63u;
  const uint64_t below = (1ull << bit) - 1ull;
  const uint64_t m1 = (uint64_t)g1.x | ((uint64_t)g1.y << 32), m2 = (uint64_t)g2.x | ((uint64_t)g2.y << 32);
  const uint64_t m5 = (uint64_t)g5.x | ((uint64_t)g5.y << 32);
  const uint32_t r1 = g1.z + (uint32_t)__popcll(m1 & below), r2 = g2.z + (uint32_t)__popcll(m2 & below);
  const uint32_t r5 = g5.z + (uint32_t)__popcll(m5 & below);
  const uint32_t ns[6] = {wrap ? s + 1u : s, s1, s2, wrap ? s1 + 1u : s1, wrap ? s2 + 1u : s2, s5};
  const uint32_t nr[6] = {wrap ? 0u : rank + 1u, r1, r2, wrap ? 0u : r1 + 1u, wrap ? 0u : r2 + 1u, r5};
  const uint32_t nx[6] = {xl1, xl, xl, xl1, xl1, xl};
  const bool inmask[6] = {true, ((m1 >> bit) & 1ull) != 0, ((m2 >> bit) & 1ull) != 0, true, true,
                          ((m5 >> bit) & 1ull) != 0};
  const uint32_t *slot_lo = reinterpret_cast<const uint32_t *>(a.slots);  // record bits 0..31
  uint32_t nrec[6];
  bool act[6];
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    const bool ok_r = ((need >> n) & 1u) && inmask[n] && nr[n] < cap;
    nrec[n] = ok_r ? slot_lo[2 * ((size_t)ns[n] * cap + nr[n])] : 0u;
    act[n] = ok_r && rec_x(nrec[n]) == nx[n];
  }
  bool ok = true;
  // vertex of ring cell n for its local edge e
  auto vtx = [&](int n, int e, bool used) -> int32_t {
    const uint32_t m2 = nrec[n];
    const uint32_t lab = (uint32_t)(s_lab[rec_lut2(m2) * 256 + rec_cfg(m2)] >> (4 * e)) & 0xFu;
    if (used && (!act[n] || lab == 0xFu)) ok = false;
    return (int32_t)(vq0[n] + rec_lvb(m2) + lab - vol.vbase + vol.vshift);
  };
  int4 *qo = reinterpret_cast<int4 *>(vol.quads) + qb;
  uint32_t qn = 0;
  const int32_t self = (int32_t)(vb + vol.vshift);
  if (e6) {
    const int32_t m0 = self + (int32_t)((labels >> 24) & 0xFu);
    const int32_t m1 = vtx(0, 7, true), m2 = vtx(3, 4, true), m3 = vtx(1, 5, true);
    qo[qn++] = v[2] <= iso ? make_int4(m0, m1, m2, m3) : make_int4(m0, m3, m2, m1);
  }
  if (e9) {
    const int32_t m0 = self + (int32_t)((labels >> 36) & 0xFu);
    const int32_t m1 = vtx(2, 1, true), m2 = vtx(4, 3, true), m3 = vtx(0, 11, true);
    qo[qn++] = v[5] <= iso ? make_int4(m0, m1, m2, m3) : make_int4(m0, m3, m2, m1);
  }
  if (e10) {
    const int32_t m0 = self + (int32_t)((labels >> 40) & 0xFu);
    const int32_t m1 = vtx(1, 8, true), m2 = vtx(5, 0, true), m3 = vtx(2, 2, true);
    qo[qn++] = v[7] <= iso ? make_int4(m0, m1, m2, m3) : make_int4(m0, m3, m2, m1);
  }
  if (!ok) atomicOr(vol.status, kStatusLookup);
}

// Segment and rank of active cell ch*64 + lane: the chunk's first entry of
// the non-empty list comes from the scan, and a window of 64 entries is
// searched with shuffles (all lanes take part).  False for lanes past the
// last cell and for inconsistent inputs (reported in the status word).
__device__ __forceinline__ bool locate_cell(const EmitArgs &a, const Totals &tot, uint32_t num_ne,
                                            uint32_t ch, uint32_t lane, uint32_t &s, uint32_t &rank) {
  const uint32_t act = ch * 64u + lane;
  const uint32_t u0 = a.chunk_ne[ch];
  const uint2 wl = (u0 < num_ne && u0 + lane <= num_ne) ? a.ne[u0 + lane] : make_uint2(0u, 0xFFFFFFFFu);
  uint32_t m = 0;
#pragma unroll
  for (uint32_t step = 32; step >= 1; step >>= 1) {
    const uint32_t cand = m + step;
    const uint32_t wc = (uint32_t)__shfl((int)wl.y, (int)(cand & 63u), kWave);
    if (cand < 64u && wc <= act) m = cand;
  }
  const uint32_t ap = (uint32_t)__shfl((int)wl.y, (int)m, kWave);
  s = (uint32_t)__shfl((int)wl.x, (int)m, kWave);
  if (act >= tot.num_active) return false;
  rank = act - ap;
  // (the rank is checked against the segment's count by the caller, whose
  // loads that check rides along with)
  if (u0 >= num_ne || u0 + m >= num_ne || s >= a.nsegs || act < ap || rank >= a.cap) {
    atomicOr(a.v.status, kStatusInternal);
    return false;
  }
  return true;
}

// Chunks of a wave.  Workgroups are dispatched round-robin over the 8 XCDs
// (workgroup b runs on XCD b % 8); with at least 8 workgroups, XCD x takes
// the contiguous chunk range [x n / 8, (x + 1) n / 8) and its workgroups
// stride through it, so that neighbouring chunks, which share field rows and
// the records of neighbour segments, meet in one L2.
struct ChunkIter {
  uint32_t next, end, step;
};

__device__ __forceinline__ ChunkIter chunk_iter(uint32_t n, uint32_t wave) {
  constexpr uint32_t wpb = kEmitThreads / kWave;
  const uint32_t G = gridDim.x;
  if (G < 8u) return ChunkIter{blockIdx.x * wpb + wave, n, G * wpb};
  const uint32_t x = blockIdx.x & 7u, slot = blockIdx.x >> 3;
  const uint32_t gx = (G >> 3) + (x < (G & 7u) ? 1u : 0u);
  const uint32_t lo = (uint32_t)((uint64_t)n * x / 8u), hi = (uint32_t)((uint64_t)n * (x + 1u) / 8u);
  return ChunkIter{lo + slot * wpb + wave, hi, gx * wpb};
}

// Emit grid: at most this many workgroups per CU (MDMC_OPT_EMIT_BLOCKS_PER_CU
// overrides it).  The grid is sized for the chunk capacity because the count
// is still on the device when the emit is launched; on a surface-dense 1024^3
// field 128 per CU beat 8 by 8% (more, shorter-lived workgroups keep each
// XCD's chunks in dispatch order), and past 256 the workgroups that find no
// chunk cost a sparse extraction several microseconds.
constexpr int kEmitBlocksPerCu = 128;

// Whether any wave of this workgroup has a chunk (its wave 0 has the lowest
// one).  The grid is sized for the chunk capacity, so when the mesh is
// sparse most workgroups leave at once, before loading their tables.
__device__ __forceinline__ bool workgroup_has_chunks(uint32_t n) {
  const ChunkIter it = chunk_iter(n, 0);
  return it.next < it.end;
}

// Single-volume emit grid: a persistent grid of this many workgroups per CU
// (MDMC_OPT_EMIT_BLOCKS_PER_CU overrides it), the most that are resident at
// once at 8 waves per SIMD.
constexpr int kEmitPersistPerCu = 8;

// 8 waves per SIMD: the emit is latency-bound, the register cap costs a few
// scalar spills only.  Chunks come by ticket (emit_tickets, mdmc_tilemap.h):
// workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8), so the chunks in flight on an XCD stay one contiguous window
// that moves through its range in order, and neighbouring chunks, which share
// field rows and the records of neighbour segments, meet in its L2 while both
// are in flight; the next ticket is asked for before the current one's
// chunks run and read after them, so its latency hides under the work.  The
// placement only shapes locality: any mapping of workgroups to XCDs takes
// every chunk exactly once.  The grid is a multiple of 8 workgroups, and the
// ticket counters are zero at launch (the classify zeroes them; the host
// zeroes them again before an emit that follows another).
__global__ __launch_bounds__(kEmitThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_emit(EmitArgs a) {
  constexpr uint32_t wpb = kEmitThreads / kWave;
  __shared__ uint64_t s_lab[2 * 256];
  __shared__ float s_acc[12 * kEmitThreads];
  __shared__ uint32_t s_ticket[2];
  const Totals tot = *a.totals;
  // Counts are clamped to buffer capacities so that no access can leave the
  // scratch buffers even if the count pass produced inconsistent data.
  const uint32_t num_chunks = (uint32_t)min((uint64_t)tot.num_chunks, a.max_chunks);
  const EmitTickets et = emit_tickets(num_chunks, blockIdx.x & 7u, gridDim.x, wpb);
  uint64_t first = emit_ticket_chunk(et, blockIdx.x >> 3, wpb);
  if (first >= et.hi) return;  // (uniform: before the barrier)
  for (uint32_t i = threadIdx.x; i < 512; i += kEmitThreads) s_lab[i] = (&c_label[0][0])[i];
  uint32_t *ctr = a.emit_ctr + kEmitCtr * ((blockIdx.x & 7u) + 1u);
  __syncthreads();
  const uint32_t lane = lane_id(), wave = wave_id();
  const uint32_t num_ne = min(tot.num_nonempty, a.nsegs);
  for (uint32_t it = 0;; ++it) {
    // the next ticket is asked for before this one's chunks run and published
    // after them; two slots, because thread 0 writes the next iteration's
    // only after every thread has passed this iteration's barrier
    uint32_t next = 0;
    if (et.dynamic && threadIdx.x == 0) next = atomicAdd(ctr, 1u);
    const uint64_t ch = first + wave;
    if (ch < et.hi) {
      uint32_t s, rank;
      if (locate_cell(a, tot, num_ne, (uint32_t)ch, lane, s, rank)) emit_cell(a, a.v, s, rank, s_lab, s_acc + threadIdx.x);
    }
    if (!et.dynamic) break;
    if (threadIdx.x == 0) s_ticket[it & 1u] = next;
    __syncthreads();
    first = emit_ticket_chunk(et, (uint64_t)et.slots + s_ticket[it & 1u], wpb);
    if (first >= et.hi) break;
  }
}

// Zeroes the emit's ticket counters (one workgroup of 64 threads).
__global__ void k_zero_tickets(uint32_t *t) { zero_tickets(t, threadIdx.x); }


// ---------------------------------------------------------------------------
// Smoothing (num_smooth > 0): smooth_edge_vertices (dmc.cu:1868-1981) then
// calc_iso_vertices_kern again (dmc.cu:1517-1593, 2157-2169), per iteration.
// Both run over the emitted mesh, one lane per active cell in the emit's
// chunks.  k_smooth moves owned edge points (reading iso vertices only);
// k_reaverage recomputes iso vertices from the stored edge points (reading
// edge points only), so neither reads what it writes.
//
// The split of a convex ring is at the corner of the largest angle, taken as
// the smallest cosine (a.b) / (sqrt(a.a) sqrt(b.b)) with IEEE sqrt and
// division, first smallest wins (the reference takes acosf of rsqrtf-
// normalised vectors, dmc.cu:1818-1860); everything else is the reference's
// float operation sequence.

// Position of an active cell from its segment.
struct CellPos {
  uint32_t xl, cfg, lut2, x, j, k, vb, qb;
  uint32_t seg_vb;  // the segment's first vertex
  uint32_t nact;    // the segment's active cells
  float y1, z1;     // coordinates of point row j + 1 and plane k + 1
};

// As in the emit, what depends only on the segment (its bases and count, the
// row's coordinates) comes through scalar loads for the lanes in the wave's
// first segment and through vector loads for the others.
__device__ __forceinline__ CellPos cell_pos(const EmitArgs &a, const EmitVol &vol, uint32_t s, uint32_t rank,
                                            uint64_t &rec) {
  CellPos c;
  rec = a.slots[(size_t)s * a.cap + rank];
  c.xl = rec_x(rec);
  c.cfg = rec_cfg(rec);
  c.lut2 = rec_lut2(rec);
  const uint32_t sl = s - vol.seg_off;
  const uint32_t xt = sl % vol.XT, row = sl / vol.XT;
  c.j = row % vol.g.cy;
  c.k = row / vol.g.cy;
  c.x = xt * a.cap + c.xl;
  const uint32_t s0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)s);
  uint32_t s0u = s0;
  asm volatile("" : "+s"(s0u));
  const uint32_t row0 = (s0u - vol.seg_off) / vol.XT;
  uint2 vq = const_u2(a.seg_vq, s0u);
  uint32_t nact = const_u2x(a.seg_counts, s0u) & 0xFFFFu;
  float y1 = as_const(vol.yt_c)[row0 % vol.g.cy + 1u], z1 = as_const(vol.zt_c)[row0 / vol.g.cy + 1u];
  if (s != s0) {
    vq = a.seg_vq[s];
    nact = a.seg_counts[s].x & 0xFFFFu;
    y1 = vol.yt_c[c.j + 1u];
    z1 = vol.zt_c[c.k + 1u];
  }
  c.seg_vb = vq.x;
  c.nact = nact;
  c.y1 = y1;
  c.z1 = z1;
  c.vb = vq.x + rec_lvb(rec) - vol.vbase;
  c.qb = vq.y + rec_lqb(rec) - vol.qbase;
  return c;
}

// the same from a copy of c_cfg_info in LDS
__device__ __forceinline__ uint32_t iso_count_lds(const uint16_t *s_info, uint32_t cfg, uint32_t lut2) {
  const uint32_t info = s_info[cfg];
  return lut2 ? ((info >> 3) & 7u) : (info & 7u);
}

__device__ __forceinline__ uint32_t iso_count(uint32_t cfg, uint32_t lut2) {
  const uint32_t info = c_cfg_info[cfg];
  return lut2 ? ((info >> 3) & 7u) : (info & 7u);
}

// bipolar owned edges 6, 9, 10 of a configuration, as bits 0, 1, 2
__device__ __forceinline__ uint32_t owned_bipolar(uint32_t cfg) {
  return (((cfg >> 2) ^ (cfg >> 6)) & 1u) | ((((cfg >> 5) ^ (cfg >> 6)) & 1u) << 1) |
         ((((cfg >> 7) ^ (cfg >> 6)) & 1u) << 2);
}

struct P2 {
  float x, y;
};

// calc_cross_z_sign (dmc.cu:1741-1747)
__device__ __forceinline__ bool cross_nonneg(P2 l, P2 m, P2 r) {
  const float dx1 = r.x - m.x, dy1 = r.y - m.y, dx2 = l.x - m.x, dy2 = l.y - m.y;
  return dx1 * dy2 - dx2 * dy1 >= 0.0f;
}

__device__ __forceinline__ float corner_cos(P2 l, P2 m, P2 r) {
  const float ax = l.x - m.x, ay = l.y - m.y, bx = r.x - m.x, by = r.y - m.y;
  const float num = ax * bx + ay * by;
  return num / (sqrtf(ax * ax + ay * ay) * sqrtf(bx * bx + by * by));
}

// is_inside_triangle (dmc.cu:1894-1912)
__device__ __forceinline__ bool inside_tri(P2 p0, P2 p1, P2 p2, P2 t, float &al, float &be, float &ga) {
  const float v0x = p1.x - p0.x, v0y = p1.y - p0.y, v1x = p2.x - p0.x, v1y = p2.y - p0.y;
  const float v2x = t.x - p0.x, v2y = t.y - p0.y;
  const float d00 = v0x * v0x + v0y * v0y, d10 = v1x * v0x + v1y * v0y, d11 = v1x * v1x + v1y * v1y;
  const float d20 = v2x * v0x + v2y * v0y, d21 = v2x * v1x + v2y * v1y;
  const float inv = 1.0f / (d00 * d11 - d10 * d10);
  be = (d11 * d20 - d10 * d21) * inv;
  ga = (d00 * d21 - d10 * d20) * inv;
  al = 1.0f - be - ga;
  return (-1e-4 < (double)be) && (-1e-4 < (double)ga) && ((double)(be + ga) < 1.0 + 1e-4);
}

// v[t] for t in 0..3 by selects (no indexed register array)
__device__ __forceinline__ float sel4(uint32_t t, const float (&v)[4]) {
  const float lo = (t & 1u) ? v[1] : v[0], hi = (t & 1u) ? v[3] : v[2];
  return (t & 2u) ? hi : lo;
}

// One smoothing pass over the owned edge points of a chunk's cells (found:
// this lane has a cell).  The chunk's quads are spread over its lanes, one
// quad per lane per round: the lanes' quad counts are prefix-summed and lane
// L of round r takes quad r*64 + L from the lane that owns it (a shuffle
// search over the prefix).  Taking each cell's up-to-three quads on its own
// lane would run three divergent rounds in nearly every surface-dense chunk,
// where most cells own one quad.  Every lane takes part in the shuffles.
__device__ __forceinline__ void smooth_chunk(const EmitArgs &a, const EmitVol &vol, bool found, uint32_t s,
                                             uint32_t rank) {
  const FieldDesc &g = vol.g;
  uint32_t has = 0, bip = 0, qb = 0, ev0 = 0;
  float x1 = 0.0f, y1 = 0.0f, z1 = 0.0f;
  if (found) {
    uint64_t rec;
    const CellPos c = cell_pos(a, vol, s, rank, rec);
    if (c.x >= g.cx || rank >= c.nact) {
      atomicOr(vol.status, kStatusInternal);
    } else {
      bip = owned_bipolar(c.cfg);
      const bool xin = c.x + 1 < g.cx, yin = c.j + 1 < g.cy, zin = c.k + 1 < g.cz;
      has = ((bip & 1u) && xin && yin ? 1u : 0u) | ((bip & 2u) && xin && zin ? 2u : 0u) |
            ((bip & 4u) && yin && zin ? 4u : 0u);
      if (has) x1 = vol.xt_c[c.x + 1];
      y1 = c.y1;
      z1 = c.z1;
      qb = c.qb;
      ev0 = c.vb + iso_count(c.cfg, c.lut2);
    }
  }
  const uint32_t lane = lane_id();
  const uint32_t cnt = (uint32_t)__popc(has);
  const uint32_t incl = wave_incl_scan(cnt), off = incl - cnt;
  const uint32_t K = (uint32_t)__shfl((int)incl, kWave - 1, kWave);
  for (uint32_t base = 0; base < K; base += kWave) {
    const uint32_t jq = base + lane;
    // owner: the highest lane whose prefix is <= jq (for jq < K it owns jq)
    uint32_t m = 0;
#pragma unroll
    for (uint32_t step = 32; step >= 1; step >>= 1) {
      const uint32_t cand = m + step;
      const uint32_t oc = (uint32_t)__shfl((int)off, (int)(cand & 63u), kWave);
      if (cand < 64u && oc <= jq) m = cand;
    }
    const uint32_t o_off = (uint32_t)__shfl((int)off, (int)m, kWave);
    const uint32_t o_has = (uint32_t)__shfl((int)has, (int)m, kWave);
    const uint32_t o_bip = (uint32_t)__shfl((int)bip, (int)m, kWave);
    const uint32_t o_qb = (uint32_t)__shfl((int)qb, (int)m, kWave);
    const uint32_t o_ev0 = (uint32_t)__shfl((int)ev0, (int)m, kWave);
    const float o_x1 = __shfl(x1, (int)m, kWave), o_y1 = __shfl(y1, (int)m, kWave), o_z1 = __shfl(z1, (int)m, kWave);
    if (jq >= K) continue;
    // the owner's k-th quad is on its k-th set bit l of `has`
    const uint32_t kq = jq - o_off;
    uint32_t h = o_has;
    if (kq >= 1u) h &= h - 1u;
    if (kq >= 2u) h &= h - 1u;
    const uint32_t l = (uint32_t)__builtin_ctz(h);
    const uint32_t q = o_qb + kq;
    const uint32_t e_vtx = o_ev0 + (uint32_t)__popc(o_bip & ((1u << l) - 1u));
    if ((uint64_t)q >= vol.qcap || (uint64_t)e_vtx >= vol.vcap) { atomicOr(vol.status, kStatusInternal); continue; }
    const int4 rq = reinterpret_cast<const int4 *>(vol.quads)[q];
    const int32_t iv[4] = {rq.x, rq.y, rq.z, rq.w};
    // the ring's four iso vertices, read once (k_smooth writes edge points
    // only, so they do not change while it runs); the edge's plane: x-y for
    // e6 (l = 0), x-z for e9, y-z for e10
    float vx[4], vy[4], vz[4];
    P2 P[4];
    bool bad = false;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      bad |= iv[p] < 0 || (uint64_t)iv[p] >= vol.vcap;
      const float *vp = vol.verts + 3ull * (uint32_t)(bad ? 0 : iv[p]);
      vx[p] = vp[0];
      vy[p] = vp[1];
      vz[p] = vp[2];
      P[p].x = l == 2u ? vy[p] : vx[p];
      P[p].y = l == 0u ? vy[p] : vz[p];
    }
    if (bad) { atomicOr(vol.status, kStatusInternal); continue; }
    // calc_quadrilateral_signs (dmc.cu:1749-1780)
    uint32_t npos = 0, ipos = 0, ineg = 0;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      if (cross_nonneg(P[(p + 3) & 3], P[p], P[(p + 1) & 3])) { ++npos; ipos = p; }
      else ineg = p;
    }
    if (npos == 2) continue;  // self-crossing ring (dmc.cu:1783-1786)
    uint32_t s0;
    if (npos == 0 || npos == 4) {
      float best = corner_cos(P[3], P[0], P[1]);
      s0 = 0;
#pragma unroll
      for (int p = 1; p < 4; ++p) {
        const float cs = corner_cos(P[(p + 3) & 3], P[p], P[(p + 1) & 3]);
        if (cs < best) { best = cs; s0 = p; }
      }
    } else {
      s0 = npos < 2 ? ipos : ineg;  // the concave corner
    }
    const uint32_t s1 = (s0 + 2) & 3;
    const P2 o = l == 0u ? P2{o_x1, o_y1} : (l == 1u ? P2{o_x1, o_z1} : P2{o_y1, o_z1});
    float al, be, ga;
    uint32_t ta, tb, tc;
    if (inside_tri(P[s0], P[(s0 + 1) & 3], P[s1], o, al, be, ga)) {
      ta = s0; tb = (s0 + 1) & 3; tc = s1;
    } else if (inside_tri(P[s1], P[(s1 + 1) & 3], P[s0], o, al, be, ga)) {
      ta = s1; tb = (s1 + 1) & 3; tc = s0;
    } else {
      continue;
    }
    float *E = vol.verts + 3ull * e_vtx;
    const float r0 = al * sel4(ta, vx) + be * sel4(tb, vx) + ga * sel4(tc, vx);
    const float r1 = al * sel4(ta, vy) + be * sel4(tb, vy) + ga * sel4(tc, vy);
    const float r2 = al * sel4(ta, vz) + be * sel4(tb, vz) + ga * sel4(tc, vz);
    E[0] = r0;
    E[1] = r1;
    E[2] = r2;
  }
}

// Iso vertices from the stored edge points: for each edge of the cell in
// order 0..11 with a vertex and an in-grid owner, the owner's stored point
// (owner cells n = 0 (-1,0,0), 1 (0,-1,0), 2 (0,0,-1), 3 (-1,-1,0),
// 4 (-1,0,-1), 5 (0,-1,-1), found like the quad ring cells), summed per
// group from -0.0 and divided by the count; groups with no point keep their
// value.
__device__ __forceinline__ void reaverage_cell(const EmitArgs &a, const EmitVol &vol, uint32_t s, uint32_t rank,
                                               const uint64_t *s_lab, const uint16_t *s_info, float *s_acc) {
  uint64_t rec;
  const CellPos c = cell_pos(a, vol, s, rank, rec);
  const FieldDesc &g = vol.g;
  const uint32_t cap = a.cap;
  if (c.x >= g.cx || c.xl >= cap || rank >= c.nact) {
    atomicOr(vol.status, kStatusInternal);
    return;
  }
  const uint64_t labels = s_lab[c.lut2 * 256 + c.cfg];
  const uint32_t niso = iso_count_lds(s_info, c.cfg, c.lut2);
  // owner cell (7 = this cell) and the owner's edge shift (0: e6, 1: e9, 2: e10)
  constexpr uint8_t OWN_N[12] = {5, 2, 2, 4, 3, 1, 7, 0, 1, 7, 7, 0};
  constexpr uint8_t OWN_L[12] = {2, 1, 2, 1, 0, 0, 0, 0, 2, 1, 2, 1};
  constexpr uint8_t NEED[12] = {6, 4, 4, 5, 3, 2, 0, 1, 2, 0, 0, 1};  // bit 0: x>0, 1: y>0, 2: z>0
  const uint32_t have = (c.x > 0 ? 1u : 0u) | (c.j > 0 ? 2u : 0u) | (c.k > 0 ? 4u : 0u);
  uint32_t need = 0;
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    const uint32_t lab = (uint32_t)(labels >> (4 * e)) & 0xFu;
    if (lab != 0xFu && (NEED[e] & have) == NEED[e] && OWN_N[e] < 7) need |= 1u << OWN_N[e];
  }
  // Owners 1 (0,-1,0), 2 (0,0,-1) and 5 (0,-1,-1) are found through their
  // segments' group maps.  An owner needed here contains a crossing edge, so
  // it is active, and so are the other cells around that edge: owners 0
  // (-1,0,0), 3 (-1,-1,0) and 4 (-1,0,-1) are the active cells just before
  // this cell, owner 1 and owner 2 in their rows (when x is the first cell
  // of its x-tile they lie in the previous x-tile's segments and are found
  // through its group maps instead).  Each record's x field confirms the
  // cell found.  A lookup the cell does not need reads this cell's segment,
  // so that every load is in bounds.
  const bool wrap = c.xl == 0;  // -x lies in the previous x-tile's segment
  const uint32_t xlm = wrap ? cap - 1u : c.xl - 1u;
  const uint32_t dy = vol.XT, dz = vol.XT * g.cy;
  const uint32_t need1 = need & 0x0Au, need2 = need & 0x14u, need5 = need & 0x20u, needw = need & 0x19u;
  const uint32_t ns[6] = {wrap ? s - 1u : s, need1 ? s - dy : s, need2 ? s - dz : s,
                          wrap ? s - 1u - dy : s - dy, wrap ? s - 1u - dz : s - dz, need5 ? s - dy - dz : s};
  const uint32_t G = cap >> 6;
  const uint4 g1 = a.gmap[(size_t)ns[1] * G + (c.xl >> 6)], g2 = a.gmap[(size_t)ns[2] * G + (c.xl >> 6)];
  const uint4 g5 = a.gmap[(size_t)ns[5] * G + (c.xl >> 6)];
  uint32_t vq0[6];
  vq0[1] = a.seg_vq[ns[1]].x;
  vq0[2] = a.seg_vq[ns[2]].x;
  vq0[5] = a.seg_vq[ns[5]].x;
  const uint32_t bit = c.xl & 63u;
  const uint64_t below = (1ull << bit) - 1ull;
  const uint64_t m1 = (uint64_t)g1.x | ((uint64_t)g1.y << 32), m2 = (uint64_t)g2.x | ((uint64_t)g2.y << 32);
  const uint64_t m5 = (uint64_t)g5.x | ((uint64_t)g5.y << 32);
  const uint32_t r1 = g1.z + (uint32_t)__popcll(m1 & below), r2 = g2.z + (uint32_t)__popcll(m2 & below);
  const uint32_t r5 = g5.z + (uint32_t)__popcll(m5 & below);
  uint32_t nr[6] = {rank - 1u, r1, r2, r1 - 1u, r2 - 1u, r5};
  bool inmask[6] = {rank > 0u, ((m1 >> bit) & 1ull) != 0, ((m2 >> bit) & 1ull) != 0, r1 > 0u, r2 > 0u,
                    ((m5 >> bit) & 1ull) != 0};
  vq0[0] = c.seg_vb;
  vq0[3] = vq0[1];
  vq0[4] = vq0[2];
  if (wrap && needw) {  // owners 0, 3, 4: the last cells of the previous x-tile's segments
    const uint32_t gw = (cap - 1u) >> 6, bw = (cap - 1u) & 63u;
    const uint64_t bwl = (1ull << bw) - 1ull;
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      if (n != 0 && n != 3 && n != 4) continue;
      const uint32_t sn = ((need >> n) & 1u) ? ns[n] : s;  // in bounds when not needed
      const uint4 gx = a.gmap[(size_t)sn * G + gw];
      const uint64_t mx = (uint64_t)gx.x | ((uint64_t)gx.y << 32);
      nr[n] = gx.z + (uint32_t)__popcll(mx & bwl);
      inmask[n] = ((mx >> bw) & 1ull) != 0;
      vq0[n] = a.seg_vq[sn].x;
    }
  }
  const uint32_t nx[6] = {xlm, c.xl, c.xl, xlm, xlm, c.xl};
  const uint32_t *slot_lo = reinterpret_cast<const uint32_t *>(a.slots);
  uint32_t own_base[8], own_bip[8];  // first edge point of each owner, its bipolar bits
  bool ok = true;
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    const bool use = (need >> n) & 1u;
    const bool ok_r = use && inmask[n] && nr[n] < cap;
    const uint32_t m = slot_lo[2 * ((size_t)(ok_r ? ns[n] : s) * cap + (ok_r ? nr[n] : rank))];
    if (use && !(ok_r && rec_x(m) == nx[n])) ok = false;
    own_base[n] = vq0[n] + rec_lvb(m) - vol.vbase + iso_count_lds(s_info, rec_cfg(m), rec_lut2(m));
    own_bip[n] = owned_bipolar(rec_cfg(m));
  }
  own_base[6] = 0;
  own_bip[6] = 0;
  own_base[7] = c.vb + niso;
  own_bip[7] = owned_bipolar(c.cfg);
  constexpr uint32_t T = kEmitThreads;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if ((uint32_t)q < niso) {
      s_acc[(3 * q + 0) * T] = -0.0f;
      s_acc[(3 * q + 1) * T] = -0.0f;
      s_acc[(3 * q + 2) * T] = -0.0f;
    }
  }
  if ((uint64_t)c.vb + niso > vol.vcap) { atomicOr(vol.status, kStatusVertexOverflow); return; }
  // every edge point's index first, then all the loads (vertex 0, which
  // exists, for edges that do not contribute), then the sums in edge order:
  // the loads are in flight together instead of one edge at a time
  uint32_t use = 0, pidx[12];
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    pidx[e] = 0;
    const uint32_t lab = (uint32_t)(labels >> (4 * e)) & 0xFu;
    if (lab == 0xFu || (NEED[e] & have) != NEED[e]) continue;
    const uint32_t on = OWN_N[e], ol = OWN_L[e];
    if (!((own_bip[on] >> ol) & 1u)) { ok = false; continue; }
    const uint32_t idx = own_base[on] + (uint32_t)__popc(own_bip[on] & ((1u << ol) - 1u));
    if ((uint64_t)idx >= vol.vcap) { ok = false; continue; }
    pidx[e] = idx;
    use |= 1u << e;
  }
  float px[12], py[12], pz[12];
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    const float *pnt = vol.verts + 3ull * pidx[e];
    px[e] = pnt[0];
    py[e] = pnt[1];
    pz[e] = pnt[2];
  }
  uint32_t cnt = 0;
#pragma unroll
  for (int e = 0; e < 12; ++e) {
    if (!((use >> e) & 1u)) continue;
    const uint32_t lab = (uint32_t)(labels >> (4 * e)) & 0xFu;
    float *acc = s_acc + 3 * lab * T;
    acc[0] = acc[0] + px[e];
    acc[T] = acc[T] + py[e];
    acc[2 * T] = acc[2 * T] + pz[e];
    cnt += 1u << (8 * lab);
  }
  float *out = vol.verts + 3ull * c.vb;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if ((uint32_t)q < niso) {
      const uint32_t n = (cnt >> (8 * q)) & 0xFFu;
      if (n) {
        const float d = (float)n;
        out[3 * q + 0] = s_acc[(3 * q + 0) * T] / d;
        out[3 * q + 1] = s_acc[(3 * q + 1) * T] / d;
        out[3 * q + 2] = s_acc[(3 * q + 2) * T] / d;
      }
    }
  }
  if (!ok) atomicOr(vol.status, kStatusLookup);
}

__global__ __launch_bounds__(kEmitThreads) void k_smooth(EmitArgs a) {
  const uint32_t lane = lane_id(), wave = wave_id();
  const Totals tot = *a.totals;
  const uint32_t num_chunks = (uint32_t)min((uint64_t)tot.num_chunks, a.max_chunks);
  const uint32_t num_ne = min(tot.num_nonempty, a.nsegs);
  const ChunkIter it = chunk_iter(num_chunks, wave);
  for (uint32_t ch = it.next; ch < it.end; ch += it.step) {
    uint32_t s = 0, rank = 0;
    const bool found = locate_cell(a, tot, num_ne, ch, lane, s, rank);
    smooth_chunk(a, a.v, found, s, rank);
  }
}

__global__ __launch_bounds__(kEmitThreads) void k_reaverage(EmitArgs a) {
  __shared__ uint64_t s_lab[2 * 256];
  __shared__ uint16_t s_info[256];  // c_cfg_info: seven lookups per cell
  __shared__ float s_acc[12 * kEmitThreads];
  const Totals tot = *a.totals;
  const uint32_t num_chunks = (uint32_t)min((uint64_t)tot.num_chunks, a.max_chunks);
  if (!workgroup_has_chunks(num_chunks)) return;  // (uniform: before the barrier)
  for (uint32_t i = threadIdx.x; i < 512; i += kEmitThreads) s_lab[i] = (&c_label[0][0])[i];
  for (uint32_t i = threadIdx.x; i < 256; i += kEmitThreads) s_info[i] = c_cfg_info[i];
  __syncthreads();
  const uint32_t lane = lane_id(), wave = wave_id();
  const uint32_t num_ne = min(tot.num_nonempty, a.nsegs);
  const ChunkIter it = chunk_iter(num_chunks, wave);
  for (uint32_t ch = it.next; ch < it.end; ch += it.step) {
    uint32_t s, rank;
    if (!locate_cell(a, tot, num_ne, ch, lane, s, rank)) continue;
    reaverage_cell(a, a.v, s, rank, s_lab, s_info, s_acc + threadIdx.x);
  }
}

// ---------------------------------------------------------------------------
// Batch count and emit.

struct BatchCountArgs {
  const BatchVol *vols;
  uint32_t nvol, nsegs;
  const uint2 *seg_vq;
  const uint2 *ne;              // non-empty segments {segment, first active}, sentinel last
  const Totals *totals;
  uint2 *base;                  // per volume: global vertex / quad number of its first
  const uint32_t *batch_counts;  // the batch scan's counts (status bits in [3])
};

// Active cells before segment s: the first active of the first non-empty
// segment at or after s (the list ends with {nsegs, total}).
__device__ __forceinline__ uint32_t actives_before(const uint2 *ne, uint32_t num_ne, uint32_t s) {
  uint32_t lo = 0, hi = num_ne;  // ne[num_ne] is the sentinel
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (ne[mid].x < s) lo = mid + 1; else hi = mid;
  }
  return ne[lo].y;
}

// One slab of an out-of-core extraction, after its count: the count words,
// the vertex / quad bases of segments s_lo and s_hi (the slab's own cell
// planes) and its active cells, in one small record for one host copy:
// out[0..3] counts, out[4..7] base(s_lo).xy base(s_hi).xy, out[8] actives.
__global__ void k_slab_info(const uint2 *seg_vq, const uint2 *ne, const Totals *totals, uint32_t nsegs,
                            const uint32_t *counts, uint32_t s_lo, uint32_t s_hi, uint32_t *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t num_ne = min(totals->num_nonempty, nsegs);
  for (int i = 0; i < 4; ++i) out[i] = counts[i];
  const uint2 lo = seg_vq[s_lo], hi = seg_vq[s_hi];
  out[4] = lo.x;
  out[5] = lo.y;
  out[6] = hi.x;
  out[7] = hi.y;
  out[8] = actives_before(ne, num_ne, s_hi) - actives_before(ne, num_ne, s_lo);
}

// One thread per volume: its totals and the bases of its numbering.
__global__ __launch_bounds__(256) void k_batch_counts(BatchCountArgs b) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= b.nvol) return;
  const BatchVol &bv = b.vols[v];
  const uint32_t num_ne = min(b.totals->num_nonempty, b.nsegs);
  const uint2 v0 = b.seg_vq[bv.seg_off], v1 = b.seg_vq[bv.seg_off + bv.nsegs];
  b.base[v] = v0;
  if (bv.counts) {
    bv.counts[0] = bv.nsegs ? actives_before(b.ne, num_ne, bv.seg_off + bv.nsegs) - actives_before(b.ne, num_ne, bv.seg_off)
                            : 0u;
    bv.counts[1] = v1.x - v0.x;
    bv.counts[2] = v1.y - v0.y;
    bv.counts[3] = b.batch_counts[3];
  }
}

// After a batch emit: status bits the emit could only attribute to the batch
// (a chunk whose cell could not be located) go to every volume's counts, the
// words callers read.
__global__ __launch_bounds__(256) void k_batch_status(const BatchVol *vols, uint32_t nvol,
                                                      const uint32_t *batch_status) {
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvol) return;
  const uint32_t st = *batch_status;
  uint32_t *const cnt = vols[v].counts;
  if (st && cnt) atomicOr(cnt + 3, st);
}

// One workgroup per volume: its coordinate tables.
__global__ __launch_bounds__(256) void k_coords_batch(const BatchVol *vols, uint32_t nvol) {
  for (uint32_t v = blockIdx.x; v < nvol; v += gridDim.x) {
    const BatchVol &bv = vols[v];
    if (bv.nsegs == 0) continue;
    CoordArgs c{bv.coords, bv.c.g.X, bv.c.g.Y, bv.c.g.Z, bv.c.g.cx, bv.c.g.cy, bv.c.g.cz,
                bv.mnx, bv.mny, bv.mnz, bv.rgx, bv.rgy, bv.rgz};
    coords_fill(c, threadIdx.x, blockDim.x);
  }
}

struct BatchEmitArgs {
  EmitArgs a;  // arrays of the whole batch; a.v.status is the batch status word
  const BatchVol *vols;
  const uint2 *base;
  uint32_t nvol;
};

// Emit over the active cells of all volumes; a lane finds its volume by
// binary search over the volumes' first segments.
__global__ __launch_bounds__(kEmitThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_emit_batch(BatchEmitArgs b) {
  __shared__ uint64_t s_lab[2 * 256];
  __shared__ float s_acc[12 * kEmitThreads];
  // first segment of each volume (the search below reads entries 1 .. nvol-1;
  // kMaxBatch entries keep the workgroup's LDS at 20 KB, 8 workgroups per CU)
  __shared__ uint32_t s_soff[kMaxBatch];
  const EmitArgs &a = b.a;
  const Totals tot = *a.totals;
  const uint32_t num_chunks = (uint32_t)min((uint64_t)tot.num_chunks, a.max_chunks);
  if (!workgroup_has_chunks(num_chunks)) return;  // (uniform: before the barrier)
  for (uint32_t i = threadIdx.x; i < 512; i += kEmitThreads) s_lab[i] = (&c_label[0][0])[i];
  for (uint32_t i = threadIdx.x; i < b.nvol; i += kEmitThreads) s_soff[i] = b.vols[i].seg_off;
  __syncthreads();
  const uint32_t lane = lane_id(), wave = wave_id();
  const uint32_t num_ne = min(tot.num_nonempty, a.nsegs);
  const ChunkIter it = chunk_iter(num_chunks, wave);
  for (uint32_t ch = it.next; ch < it.end; ch += it.step) {
    uint32_t s, rank;
    if (!locate_cell(a, tot, num_ne, ch, lane, s, rank)) continue;
    // the last volume whose first segment is <= s (empty volumes share it)
    uint32_t lo = 0, hi = b.nvol;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_soff[mid] <= s) lo = mid; else hi = mid;
    }
    // the lanes of one volume at a time (64 consecutive active cells rarely
    // span two volumes), so that the volume's descriptor is wave-uniform and
    // comes in through scalar loads instead of occupying vector registers
    for (;;) {
      const uint32_t v = __builtin_amdgcn_readfirstlane(lo);
      if (lo != v) continue;
      const __attribute__((address_space(4))) BatchVol *bv = (const __attribute__((address_space(4))) BatchVol *)b.vols + v;
      const __attribute__((address_space(4))) uint2 *bb = (const __attribute__((address_space(4))) uint2 *)b.base + v;
      EmitVol e;
      e.g.f = bv->c.g.f;
      e.g.X = bv->c.g.X;
      e.g.Y = bv->c.g.Y;
      e.g.Z = bv->c.g.Z;
      e.g.cx = bv->c.g.cx;
      e.g.cy = bv->c.g.cy;
      e.g.cz = bv->c.g.cz;
      e.g.iso = bv->c.g.iso;
      float *const co = bv->coords;
      e.xt_c = co;
      e.yt_c = co + e.g.X;
      e.zt_c = co + e.g.X + e.g.Y;
      e.XT = bv->c.XT;
      e.seg_off = bv->seg_off;
      e.vbase = bb->x;
      e.qbase = bb->y;
      e.verts = bv->verts;
      e.vcap = e.verts ? bv->vcap : 0;
      e.quads = bv->quads;
      e.qcap = e.quads ? bv->qcap : 0;
      uint32_t *const cnt = bv->counts;
      e.status = cnt ? cnt + 3 : a.v.status;
      emit_cell(a, e, s, rank, s_lab, s_acc + threadIdx.x);
      break;
    }
  }
}

// ---------------------------------------------------------------------------
// Per-cell records in ascending cell order (parity and debugging): one
// workgroup per non-empty segment (grid-stride over the list).

struct CellsArgs {
  uint32_t XT, cap, cx, nsegs;
  const uint2 *seg_counts;
  const uint64_t *slots;
  const uint2 *seg_vq;
  const uint2 *ne;
  const Totals *totals;
  uint32_t *cell_index;
  uint8_t *cfg;
  uint8_t *flags;
  uint32_t *vbegin;
  uint64_t capacity;
};

__global__ __launch_bounds__(256) void k_cells(CellsArgs a) {
  const uint32_t num_ne = min(a.totals->num_nonempty, a.nsegs);
  for (uint32_t i = blockIdx.x; i < num_ne; i += gridDim.x) {
    const uint2 e = a.ne[i];
    const uint32_t sg = e.x;
    if (sg >= a.nsegs) continue;
    const uint32_t n = min(a.seg_counts[sg].x & 0xFFFFu, a.cap);
    const uint32_t vb = a.seg_vq[sg].x;
    const uint32_t xt = sg % a.XT, row = sg / a.XT;
    for (uint32_t r = threadIdx.x; r < n; r += blockDim.x) {
      const uint64_t pos = (uint64_t)e.y + r;
      if (pos >= a.capacity) break;
      const uint64_t rec = a.slots[(size_t)sg * a.cap + r];
      a.cell_index[pos] = xt * a.cap + rec_x(rec) + a.cx * row;
      a.cfg[pos] = (uint8_t)rec_cfg(rec);
      a.flags[pos] = (uint8_t)rec_lut2(rec);
      a.vbegin[pos] = vb + rec_lvb(rec);
    }
  }
}

// ---------------------------------------------------------------------------
// Synthetic fields (definitions in mdmc_fields.h).

// Blob tables, ldb = blob_ldb(nblobs) entries per blob dimension, entries
// past nblobs zero: ex is [ldb][X] (read per lane), ey and ez are transposed,
// [Y][ldb] and [Z][ldb], so that the factors of one row are contiguous and
// come in through a few wide scalar loads; amp has ldb entries.
struct BlobArgs {
  float *out;
  uint32_t X, Y, Z, nblobs, ldb;
  const double *ex, *ey, *ez, *amp;
  uint32_t per;  // planes per task
};

// Generators work in wave tasks of 64 consecutive columns x; a lane owns one
// column, and a task takes one row y in `per` consecutive planes z.  The loop
// over the task's planes is uniform over the wave, so per-plane values are
// scalar loads, the stores of a plane are one contiguous 256-byte run, and
// what depends on x and y only is computed once per task.  Consecutive tasks
// take consecutive column blocks, then rows, then plane blocks.  `per` is at
// most kGenPer, fewer for a field too small to give the chip enough waves (a
// task's planes run one after another).
constexpr uint32_t kGenPer = 16;
constexpr uint32_t kGenThreads = 256;
constexpr uint32_t kGenWaves = kGenThreads / kWave;
constexpr uint32_t kBlobRegs = 32;  // blobs kept in registers (more: per-point path)

// blob table row stride for a blob count
__host__ __device__ inline uint32_t blob_ldb(uint32_t nblobs) {
  return nblobs <= kBlobRegs ? kBlobRegs : nblobs;
}

__host__ __device__ inline uint64_t gen_tasks(uint32_t X, uint32_t Y, uint32_t Z, uint32_t per) {
  return (uint64_t)((X + 63u) / 64u) * Y * ((Z + per - 1u) / per);
}

// Wave-uniform reads through the constant address space compile to scalar
// loads (the tables are not written while a generator runs).
typedef const __attribute__((address_space(4))) double *uniform_dptr;
__device__ __forceinline__ uniform_dptr as_uniform(const double *p) { return (uniform_dptr)p; }

// A task: the lane's column x, its row y, and its planes z0 .. z0 + nz - 1
// (nz = 0 past the last task).
struct PlaneTask {
  uint32_t x, y, z0, nz;
};

__device__ __forceinline__ PlaneTask plane_task(uint32_t X, uint32_t Y, uint32_t Z, uint32_t per, uint64_t task) {
  const uint32_t ncc = (X + 63u) / 64u;
  PlaneTask t;
  t.x = (uint32_t)(task % ncc) * 64u + (threadIdx.x & 63u);
  const uint64_t rest = task / ncc;
  t.y = (uint32_t)(rest % Y);
  const uint64_t zb = rest / Y;
  t.z0 = (uint32_t)min<uint64_t>(zb * per, Z);
  t.nz = min(per, Z - t.z0);
  return t;
}

// Blob sums: a lane keeps (amp[b] * ex[b][x]) * ey[y][b] for every blob in
// registers and, plane by plane, adds it times ez[z][b] blob by blob with one
// FMA each (fma((amp*ex) * ey, ez, sum)), the same operations in the same
// order as the CPU generator.  Four planes go per pass (four independent FMA
// chains); blobs go in groups of 8 (one group of a plane's factors is one
// wide scalar load).  The zero entries that pad the last group add +0.0 to a
// sum that starts at +0.0 and so can never be -0.0, which leaves it
// unchanged.  Needs ldb == kBlobRegs.
__device__ __forceinline__ void blob_task(const BlobArgs &a, uint64_t task) {
  constexpr uint32_t P = 4;
  const PlaneTask t = plane_task(a.X, a.Y, a.Z, a.per, task);
  if (t.x >= a.X || t.nz == 0) return;
  const uniform_dptr amp = as_uniform(a.amp);
  const uniform_dptr ey = as_uniform(a.ey + (size_t)t.y * kBlobRegs);
  double txy[kBlobRegs];
#pragma unroll
  for (uint32_t g = 0; g < kBlobRegs / 8; ++g) {
#pragma unroll
    for (uint32_t i = 0; i < 8; ++i) txy[8 * g + i] = 0.0;
    if (8 * g < a.nblobs) {
#pragma unroll
      for (uint32_t i = 0; i < 8; ++i)
        txy[8 * g + i] = (amp[8 * g + i] * a.ex[(size_t)(8 * g + i) * a.X + t.x]) * ey[8 * g + i];
    }
  }
  const uint64_t plane = (uint64_t)a.X * a.Y;
  float *out = a.out + (uint64_t)t.z0 * plane + (uint64_t)t.y * a.X + t.x;
  for (uint32_t kz = 0; kz < t.nz; kz += P) {
    // planes past the task's last repeat it (summed, not stored)
    uniform_dptr ez[P];
#pragma unroll
    for (uint32_t j = 0; j < P; ++j) ez[j] = as_uniform(a.ez + (size_t)(t.z0 + min(kz + j, t.nz - 1u)) * kBlobRegs);
    double acc[P];
#pragma unroll
    for (uint32_t j = 0; j < P; ++j) acc[j] = 0.0;
#pragma unroll
    for (uint32_t g = 0; g < kBlobRegs / 8; ++g) {
      if (8 * g < a.nblobs) {
        double f[P][8];
#pragma unroll
        for (uint32_t j = 0; j < P; ++j)
#pragma unroll
          for (uint32_t i = 0; i < 8; ++i) f[j][i] = ez[j][8 * g + i];
#pragma unroll
        for (uint32_t i = 0; i < 8; ++i)
#pragma unroll
          for (uint32_t j = 0; j < P; ++j) acc[j] = __builtin_fma(txy[8 * g + i], f[j][i], acc[j]);
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < P; ++j)
      if (kz + j < t.nz) out[(size_t)(kz + j) * plane] = (float)acc[j];
  }
}

// Per-point path (any number of blobs).
__device__ __forceinline__ void blob_points(const BlobArgs &a, size_t i0, size_t stride) {
  const size_t n = (size_t)a.X * a.Y * a.Z;
  for (size_t id = i0; id < n; id += stride) {
    const uint32_t x = (uint32_t)(id % a.X);
    const size_t r = id / a.X;
    const uint32_t y = (uint32_t)(r % a.Y), z = (uint32_t)(r / a.Y);
    double acc = 0.0;
    for (uint32_t b = 0; b < a.nblobs; ++b)
      acc = __builtin_fma((a.amp[b] * a.ex[(size_t)b * a.X + x]) * a.ey[(size_t)y * a.ldb + b],
                          a.ez[(size_t)z * a.ldb + b], acc);
    a.out[id] = (float)acc;
  }
}

struct GyroidArgs {
  float *out;
  uint32_t X, Y, Z;
  const double *sx, *cx, *sy, *cy, *sz, *cz;
  uint64_t seed;
  double amp;
  uint32_t per;  // planes per task
};

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Gyroid plus hash noise of point id = x + X * (y + Y * z):
// sx*cy + sy*cz + sz*cx from the 1-D tables, in that order, then the noise of
// mix64(seed + (id + 1) * golden).  Along the task's planes id grows by X*Y,
// so the hash input grows by X*Y*golden (mod 2^64).
__device__ __forceinline__ void gyroid_task(const GyroidArgs &a, uint64_t task) {
  constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
  const PlaneTask t = plane_task(a.X, a.Y, a.Z, a.per, task);
  if (t.x >= a.X || t.nz == 0) return;
  const double sxx = a.sx[t.x], cxx = a.cx[t.x];
  const uniform_dptr sz = as_uniform(a.sz), cz = as_uniform(a.cz);
  const double syy = as_uniform(a.sy)[t.y];
  const double t1 = sxx * as_uniform(a.cy)[t.y];
  const uint64_t plane = (uint64_t)a.X * a.Y;
  const uint64_t id0 = (uint64_t)t.z0 * plane + (uint64_t)t.y * a.X + t.x;
  const uint64_t step = plane * kGolden;
  uint64_t hin = a.seed + (id0 + 1) * kGolden;
  float *out = a.out + id0;
  for (uint32_t kz = 0; kz < t.nz; ++kz, hin += step) {
    const uint32_t z = t.z0 + kz;
    const double gval = t1 + syy * cz[z] + sz[z] * cxx;
    const uint64_t h = mix64(hin);
    const double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);
    out[kz * plane] = (float)(gval + a.amp * (2.0 * u - 1.0));
  }
}

// The generator kernel: the fields of one or many volumes in one launch.
// Workgroup b belongs to the volume whose range [block_off, block_off +
// nblocks) holds it, so its waves share the volume and its kind.
struct GenVol {
  float *out;
  uint32_t X, Y, Z, kind;  // kind 1: blobs, 2: gyroid
  uint32_t nblobs, block_off, nblocks, per;  // per: planes per task
  const double *t[6];      // blobs: ex, ey, ez, amp; gyroid: sx, cx, sy, cy, sz, cz
  uint64_t seed;
  double amp;              // gyroid noise amplitude
};

__global__ __launch_bounds__(kGenThreads) void k_field_batch(const GenVol *vols, uint32_t nvol) {
  uint32_t lo = 0, hi = nvol;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (vols[mid].block_off <= blockIdx.x) lo = mid; else hi = mid;
  }
  const GenVol &v = vols[__builtin_amdgcn_readfirstlane(lo)];
  const uint32_t local = blockIdx.x - v.block_off;
  const uint64_t task = (uint64_t)local * kGenWaves + wave_id();
  if (v.kind == 1) {
    const BlobArgs ba{v.out, v.X, v.Y, v.Z, v.nblobs, blob_ldb(v.nblobs), v.t[0], v.t[1], v.t[2], v.t[3], v.per};
    if (v.nblobs <= kBlobRegs) {
      if (task < gen_tasks(v.X, v.Y, v.Z, v.per)) blob_task(ba, task);
    } else {
      blob_points(ba, (size_t)local * blockDim.x + threadIdx.x, (size_t)v.nblocks * blockDim.x);
    }
    return;
  }
  if (task < gen_tasks(v.X, v.Y, v.Z, v.per)) {
    const GyroidArgs ga{v.out, v.X, v.Y, v.Z, v.t[0], v.t[1], v.t[2], v.t[3], v.t[4], v.t[5], v.seed, v.amp, v.per};
    gyroid_task(ga, task);
  }
}

}  // namespace mdmc
