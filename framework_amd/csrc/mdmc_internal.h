// Internal types of the mdmc host code: tile plan, scratch layout and the
// dispatch of the classify kernel over its compile-time tile shapes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mdmc_kernels.hpp"

namespace mdmc {

// Tiling of one volume.  Cells are split into segments: one cell row of one
// x-tile (64*W cells) in one plane, numbered xt + XT*(y + cy*z) so that
// segment order is the reference's cell order (x fastest).
struct Plan {
  uint32_t X, Y, Z, cx, cy, cz;
  uint32_t W, R, Kc, cap, XT, NRB, NKC, ntiles;
  uint64_t nsegs, nslots, max_chunks, classify_blocks;
  uint64_t ncoord;  // point coordinates kept for the emit: X + Y + Z (summed over a batch)
  int empty;
  int vec4;    // rows are a multiple of 4 points
  int zmarch;  // use the z-marching classify kernel
  uint32_t ptR, ptKP;        // plane tiles: cell rows, cell planes
  uint32_t band, NKBp, NRBp; // plane tiles: columns per group, plane blocks, row blocks
  uint64_t pt_blocks;        // plane-tile grid
};

struct Scratch {
  uint2 *seg_counts;
  uint2 *seg_vq;
  uint2 *ne;
  uint32_t *chunk_ne;
  unsigned long long *scan_state;
  uint32_t *ticket;
  Totals *totals;
  uint64_t *slots;
  uint4 *gmap;
  float *coords;
};

// opt: the context's option values (MDMC_OPT_*), or null for defaults.
int plan_tiling(const uint32_t dims[3], Plan *p, const int *opt);
constexpr int kNumOptions = 6;

// Tile shapes with a compiled classify kernel.
inline bool tiling_supported(uint32_t W, uint32_t R) {
  const bool w_ok = W == 1 || W == 2 || W == 4 || W == 8 || W == 16;
  const bool r_ok = R == 2 || R == 4 || R == 8 || (R == 12 && W <= 8);
  return w_ok && r_ok;
}

template <int W, int R>
inline void launch_classify_t(const ClassifyArgs &a, uint32_t blocks, hipStream_t st) {
  hipLaunchKernelGGL((k_classify<W, R>), dim3(blocks), dim3(64), 0, st, a);
}

template <int W>
inline int launch_classify_w(uint32_t R, const ClassifyArgs &a, uint32_t blocks, hipStream_t st) {
  switch (R) {
    case 2: launch_classify_t<W, 2>(a, blocks, st); return 0;
    case 4: launch_classify_t<W, 4>(a, blocks, st); return 0;
    case 8: launch_classify_t<W, 8>(a, blocks, st); return 0;
    case 12:
      if constexpr (W <= 8) { launch_classify_t<W, 12>(a, blocks, st); return 0; }
      return -1;
    default: return -1;
  }
}

template <int W4, int R, int KP>
inline void launch_classify_pt_t(bool vec, const PtArgs &a, uint32_t blocks, hipStream_t st) {
  if (vec)
    hipLaunchKernelGGL((k_classify_pt<W4, R, KP, true>), dim3(blocks), dim3(kPtThreads), 0, st, a);
  else
    hipLaunchKernelGGL((k_classify_pt<W4, R, KP, false>), dim3(blocks), dim3(kPtThreads), 0, st, a);
}

// Compiled plane-tile shapes besides the defaults (pt_rows, pt_planes).
inline bool pt_shape_supported(uint32_t W4, int R, int KP) {
  if (W4 == 1) return (R == 7 && KP == 4) || (R == 15 && KP == 2);
  if (W4 == 2) return (R == 7 && (KP == 2 || KP == 3)) || (R == 15 && (KP == 1 || KP == 2));
  return (uint32_t)R == pt_rows(W4) && (uint32_t)KP == pt_planes(W4);
}
// cell planes of the 15-row shapes by default
inline uint32_t pt_planes15(uint32_t W4) { return W4 == 1 ? 2u : 1u; }

// Plane-tile kernel (W = 4, 8 or 16); vec: 16-byte row loads.
inline int launch_classify_pt(uint32_t W, uint32_t R, uint32_t KP, bool vec, const PtArgs &a,
                              uint32_t blocks, hipStream_t st) {
  switch (W) {
    case 4:
      if (R == 7 && KP == 4) { launch_classify_pt_t<1, 7, 4>(vec, a, blocks, st); return 0; }
      if (R == 15 && KP == 2) { launch_classify_pt_t<1, 15, 2>(vec, a, blocks, st); return 0; }
      return -1;
    case 8:
      if (R == 7 && KP == 2) { launch_classify_pt_t<2, 7, 2>(vec, a, blocks, st); return 0; }
      if (R == 7 && KP == 3) { launch_classify_pt_t<2, 7, 3>(vec, a, blocks, st); return 0; }
      if (R == 15 && KP == 1) { launch_classify_pt_t<2, 15, 1>(vec, a, blocks, st); return 0; }
      if (R == 15 && KP == 2) { launch_classify_pt_t<2, 15, 2>(vec, a, blocks, st); return 0; }
      return -1;
    case 16: launch_classify_pt_t<4, (int)pt_rows(4), (int)pt_planes(4)>(vec, a, blocks, st); return 0;
    default: return -1;
  }
}

// z-marching kernel for any row length.
inline int launch_classify(uint32_t W, uint32_t R, const ClassifyArgs &a, uint32_t blocks,
                           hipStream_t st) {
  switch (W) {
    case 1: return launch_classify_w<1>(R, a, blocks, st);
    case 2: return launch_classify_w<2>(R, a, blocks, st);
    case 4: return launch_classify_w<4>(R, a, blocks, st);
    case 8: return launch_classify_w<8>(R, a, blocks, st);
    case 16: return launch_classify_w<16>(R, a, blocks, st);
    default: return -1;
  }
}

}  // namespace mdmc
