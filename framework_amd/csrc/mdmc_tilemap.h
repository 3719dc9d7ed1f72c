// Workgroup -> tile map of the plane-tile classify (k_classify_pt), and the
// chunk tickets of the persistent emit (k_emit).
//
// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8).  A column is one x-tile of one row block; columns form groups of B.
// In every plane block XCD x takes groups x, x+8, x+16, ... and walks them in
// memory order, plane block after plane block; the groups left over when
// their count is not a multiple of 8 are dealt out over the XCDs eight plane
// blocks at a time.  The launch has 8 * B * ceil(NKB * groups / 8) blocks.
//
// Plain C++ (no HIP types) so that a host test can check it.
#ifndef MDMC_TILEMAP_H
#define MDMC_TILEMAP_H

#include <cstdint>

#if defined(__HIPCC__)
#define MDMC_HD __host__ __device__
#else
#define MDMC_HD
#endif

namespace mdmc {

// Tile of block b (XCD b % 8, b / 8-th of that XCD's walk); blocks past the
// XCD's share get plane block NKB, a tile outside the grid.
struct PtTile {
  uint32_t xt, rb, kb;
};
MDMC_HD inline PtTile pt_tile_of(uint32_t b, uint32_t XT, uint32_t NRB, uint32_t NKB, uint32_t B) {
  const uint32_t x = b & 7u, s = b >> 3;
  const uint32_t cols = XT * NRB;
  const uint32_t ng = (cols + B - 1u) / B;  // column groups
  const uint32_t q = ng >> 3, r = ng & 7u;  // groups per XCD per plane block; left over
  const uint32_t gs = s / B, w = s - gs * B;  // group tile of the walk, column in the group
  const uint32_t per = 8u * q + r;          // group tiles per XCD per 8 plane blocks
  const uint32_t S = gs / per, o = gs - S * per;
  PtTile t{0u, 0u, NKB};
  if ((uint64_t)S * 8u >= NKB) return t;
  const uint32_t nk = NKB - 8u * S < 8u ? NKB - 8u * S : 8u;
  uint32_t kb, g;
  if (o < nk * q) {  // the XCD's own groups, plane block by plane block
    kb = 8u * S + o / q;
    g = (o % q) * 8u + x;
  } else {  // left-over group tiles e = x, x + 8, ... of these plane blocks
    const uint32_t e = x + 8u * (o - nk * q);
    if (e >= nk * r) return t;
    kb = 8u * S + e / r;
    g = 8u * q + e % r;
  }
  const uint32_t col = g * B + w;
  if (col >= cols) return t;
  t.xt = col % XT;
  t.rb = col / XT;
  t.kb = kb;
  return t;
}

// Chunk tickets of the persistent emit (k_emit).  The workgroups b with
// b % 8 == x (XCD x) take the contiguous chunk range [x n / 8, (x + 1) n / 8)
// in tickets of one chunk per wave: ticket t is chunks [lo + W t, lo + W t +
// W) for W waves per workgroup.  Workgroup b's first ticket is its slot b / 8
// (no atomic), its later ones are slots + c for the values c of counter x,
// taken only when the first tickets do not already cover the range
// (dynamic).  A workgroup stops at the first ticket that starts at or past
// hi.  G, the number of workgroups, is a multiple of 8.
struct EmitTickets {
  uint32_t lo, hi, slots;
  bool dynamic;
};
MDMC_HD inline EmitTickets emit_tickets(uint32_t n, uint32_t x, uint32_t G, uint32_t W) {
  EmitTickets e;
  e.lo = (uint32_t)((uint64_t)n * x / 8u);
  e.hi = (uint32_t)((uint64_t)n * (x + 1u) / 8u);
  e.slots = G >> 3;
  e.dynamic = e.lo + (uint64_t)e.slots * W < e.hi;
  return e;
}
// first chunk of ticket t
MDMC_HD inline uint64_t emit_ticket_chunk(const EmitTickets &e, uint64_t t, uint32_t W) { return e.lo + t * W; }

}  // namespace mdmc

#endif
