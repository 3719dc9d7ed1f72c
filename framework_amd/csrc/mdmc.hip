// mdmc — host side of the C ABI declared in include/mdmc.h.
//
// Owns the case tables in constant memory, per-context scratch in HBM, the
// tile planner and the launches of the kernels in mdmc_kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "../../include/mdmc.h"
#include "dmc_luts.h"
#include "mdmc_fields.h"
#include "mdmc_kernels.hpp"
#include "mdmc_internal.h"
#include "mdmc_io.h"

namespace {

thread_local std::string g_last_error;
std::mutex g_setup_mutex;
std::mutex g_host_entry_mutex;  // the host entries share the default context
uint64_t g_setup_mask = 0;  // bit d: tables uploaded on device d
mdmc_ctx *g_default_ctx = nullptr;
mdmc_ctx *g_ooc_ctx = nullptr;  // the out-of-core entry's own buffers (kept to its budget)

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess)                                                               \
      return fail(MDMC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));     \
  } while (0)

int current_device(int *dev) {
  HIP_TRY(hipGetDevice(dev));
  return MDMC_OK;
}

bool is_setup_on(int dev) {
  std::lock_guard<std::mutex> lk(g_setup_mutex);
  return dev >= 0 && dev < 64 && ((g_setup_mask >> dev) & 1ull);
}

// Device copies of the tables in the reference's formats (one block per
// device, mdmc_setup_device_luts), guarded by g_setup_mutex.
void *g_ref_tables[64] = {};

// The no-allocation entry's scratch on each device (the address of the
// `scratch` array below, taken by mdmc_setup on that device) and the
// device's compute units, guarded by g_setup_mutex.
void *g_noalloc_ptr[64] = {};
int g_noalloc_cus[64] = {};

// Byte offsets of the tables in that block.
enum : size_t {
  kTabLut1 = 0, kTabLut2 = 3072, kTabNv1 = 6144, kTabNv2 = 6400, kTabAmb = 6656, kTabAmbFace = 6692,
  kTabOpposite = 6728, kTabFaceDir = 6734, kTabEdgeOwner = 6740, kTabRing = 6752, kTabLocal = 6764,
  kTabBytes = 6784
};

// The tables of the reference's 11-pointer setup (dmc.cu:45-661), built from
// the case-table builder and the cube's geometry (corner offsets, edge and
// face corner lists) rather than typed in.
void build_reference_tables(uint8_t *out) {
  std::memset(out, 0, kTabBytes);
  mdmc_luts L;
  mdmc_build_luts(&L);
  std::memcpy(out + kTabLut1, L.edge_label[0], 3072);
  std::memcpy(out + kTabLut2, L.edge_label[1], 3072);
  std::memcpy(out + kTabNv1, L.num_vertex[0], 256);
  std::memcpy(out + kTabNv2, L.num_vertex[1], 256);
  int n = 0;
  for (int c = 0; c < 256; ++c)
    if (L.amb_face[c] != MDMC_NOT_AMBIGUOUS) {
      out[kTabAmb + n] = (uint8_t)c;
      out[kTabAmbFace + n] = L.amb_face[c];
      ++n;
    }
  // the opposite face shares no corner
  for (int f = 0; f < 6; ++f)
    for (int g = 0; g < 6; ++g) {
      bool shared = false;
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) shared |= MDMC_FACE_C[f][a] == MDMC_FACE_C[g][b];
      if (!shared) out[kTabOpposite + f] = (uint8_t)g;
    }
  for (int f = 0; f < 6; ++f) out[kTabFaceDir + f] = MDMC_FACE_DIR[f];
  // An edge along axis A at offsets (b, c) in {0,1} on the other two axes is
  // owned by the cell at offset (b - 1, c - 1) on those axes, where it is the
  // edge at (1, 1): edge 10 for x, 9 for y, 6 for z (the edges of corner 6).
  const int owner_edge[3] = {10, 9, 6};
  for (int e = 0; e < 12; ++e) {
    const uint8_t *p = MDMC_CORNER_D[MDMC_EDGE_C[e][0]], *q = MDMC_CORNER_D[MDMC_EDGE_C[e][1]];
    int axis = 0;
    while (p[axis] == q[axis]) ++axis;
    uint8_t code = 0;
    for (int d = 0; d < 3; ++d)
      if (d != axis && p[d] == 0) code |= (uint8_t)(0x80u >> d);
    out[kTabEdgeOwner + e] = code ? (uint8_t)(code | owner_edge[axis]) : 0xFFu;
  }
  // Ring of owned edge l along axis A (z, y, x for 6, 9, 10): with B, C the
  // next axes in cyclic order, the cells at (B, C) offsets (0,0), (1,0),
  // (1,1), (0,1) -- counter-clockwise looking along +A -- each holding the
  // edge at (1 - dB, 1 - dC).
  const int ring_axis[3] = {2, 1, 0};
  const int ring_d[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
  for (int l = 0; l < 3; ++l) {
    const int A = ring_axis[l], B = (A + 1) % 3, C = (A + 2) % 3;
    for (int k = 0; k < 4; ++k) {
      for (int e = 0; e < 12; ++e) {
        const uint8_t *p = MDMC_CORNER_D[MDMC_EDGE_C[e][0]], *q = MDMC_CORNER_D[MDMC_EDGE_C[e][1]];
        if (p[A] != q[A] && p[B] == 1 - ring_d[k][0] && q[B] == p[B] && p[C] == 1 - ring_d[k][1] && q[C] == p[C])
          out[kTabRing + 4 * l + k] = (uint8_t)e;
      }
    }
    out[kTabLocal + l] = out[kTabRing + 4 * l];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Planning

namespace mdmc {

static int env_int(const char *name, int dflt) {
  const char *s = std::getenv(name);
  return s ? std::atoi(s) : dflt;
}

int plan_tiling(const uint32_t dims[3], Plan *p, const int *opt) {
  std::memset(p, 0, sizeof(*p));
  p->X = dims[0];
  p->Y = dims[1];
  p->Z = dims[2];
  if (p->X < 2 || p->Y < 2 || p->Z < 2) {
    p->empty = 1;
    return MDMC_OK;
  }
  const uint64_t npts = (uint64_t)p->X * p->Y * p->Z;
  if (npts > 0xFFFFFFFFull) return fail(MDMC_ERR_TOO_LARGE, "grid has more than 2^32-1 points");
  p->cx = p->X - 1;
  p->cy = p->Y - 1;
  p->cz = p->Z - 1;
  // Classify kernel: plane tiles (x-tiles of 256, 512 or 1024 cells), or the
  // z-marching kernel when MDMC_CLASSIFY=zmarch (x-tiles from 64 cells; the
  // two produce identical records, which the tests use as a cross-check).
  const char *ck = std::getenv("MDMC_CLASSIFY");
  p->zmarch = (ck != nullptr && std::strcmp(ck, "zmarch") == 0) ? 1 : 0;
  // x-tile: smallest 64*W covering the row, at most 512 cells for plane tiles
  // (a 1024-cell tile holds twice the registers and halves the occupancy:
  // two 512-cell x-tiles classify a 1024^3 gyroid in 1.56 ms instead of 1.91)
  // and 1024 for the z-marching kernel.
  uint32_t W = p->zmarch ? 1 : 4;
  while (W < (p->zmarch ? 16u : 8u) && 64u * W < p->cx) W <<= 1;
  uint32_t R = (W >= 8) ? 4 : 8;
  if (const char *t = std::getenv("MDMC_TILE")) {  // "W,R,Kc" override for tuning
    int w = 0, r = 0, kc = 0;
    if (std::sscanf(t, "%d,%d,%d", &w, &r, &kc) == 3) {
      W = (uint32_t)w;
      R = (uint32_t)r;
      p->Kc = (uint32_t)kc;
    }
  }
  if (!tiling_supported(W, R)) return fail(MDMC_ERR_INVALID, "unsupported tile shape");
  p->W = W;
  p->R = R;
  p->cap = 64u * W;
  p->XT = (p->cx + p->cap - 1) / p->cap;
  p->NRB = (p->cy + R - 1) / R;
  if (p->Kc == 0) {
    // Enough waves to cover the chip several times over: ~8 per CU.
    const uint64_t target = (uint64_t)env_int("MDMC_TARGET_WAVES", 4096);
    const uint64_t cols = (uint64_t)p->XT * p->NRB;
    uint64_t nkc = (target + cols - 1) / cols;
    if (nkc < 1) nkc = 1;
    uint32_t Kc = (uint32_t)((p->cz + nkc - 1) / nkc);
    if (Kc < 4) Kc = std::min<uint32_t>(4, p->cz);
    p->Kc = Kc;
  }
  p->NKC = (p->cz + p->Kc - 1) / p->Kc;
  p->nsegs = (uint64_t)p->XT * p->cy * p->cz;
  p->nslots = p->nsegs * p->cap;
  const uint64_t cells = (uint64_t)p->cx * p->cy * p->cz;
  p->max_chunks = (cells + 63) / 64 + 1;
  p->ncoord = (uint64_t)p->X + p->Y + p->Z;
  // whole scan tiles with at least one padding segment (its vertex / quad
  // prefix is the grand total)
  p->ntiles = (uint32_t)(p->nsegs / kScanTile + 1);
  p->classify_blocks = (uint64_t)p->XT * p->NRB * p->NKC;
  p->vec4 = (p->X % 4 == 0) && env_int("MDMC_VEC4", 1) != 0;
  p->ptR = pt_rows(W / 4);
  p->ptKP = pt_planes(W / 4);
  // context options (0: default): cell rows and planes per tile, columns per group
  const int r_opt = opt ? opt[MDMC_OPT_PT_ROWS] : 0;
  if (r_opt == 15 && pt_shape_supported(W / 4, 15, (int)pt_planes15(W / 4))) {
    p->ptR = 15;
    p->ptKP = pt_planes15(W / 4);
  }
  const int kp_opt = opt ? opt[MDMC_OPT_PT_PLANES] : 0;
  if (kp_opt > 0 && pt_shape_supported(W / 4, (int)p->ptR, kp_opt)) p->ptKP = (uint32_t)kp_opt;
  const uint32_t rows = p->ptR;
  const uint32_t nrb = (p->cy + rows - 1) / rows;
  const int band_opt = opt ? opt[MDMC_OPT_PT_BAND] : 0;
  p->band = band_opt > 0 && band_opt <= 64 ? (uint32_t)band_opt : 1u;
  p->NRBp = nrb;
  p->NKBp = (p->cz + p->ptKP - 1) / p->ptKP;
  // 8 XCDs, each at most ceil(NKB * groups / 8) group tiles of `band` columns
  const uint64_t ng = ((uint64_t)p->XT * nrb + p->band - 1) / p->band;
  p->pt_blocks = 8ull * p->band * (((uint64_t)p->NKBp * ng + 7) / 8);
  return MDMC_OK;
}

}  // namespace mdmc

// ---------------------------------------------------------------------------
// The no-allocation entry's scratch: a fixed array in device memory, one copy
// per GPU, held until the process ends.  Host code never passes the array's
// name to a kernel (on the host it is a stand-in, not a device address): it
// uses the address hipGetSymbolAddress gives for each device (g_noalloc_ptr).
static __device__ __align__(256) unsigned char scratch[MDMC_NOALLOC_SCRATCH_BYTES];

// ---------------------------------------------------------------------------
// Context and scratch

struct mdmc_ctx {
  int device = 0;
  int num_cus = 256;
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  bool scratch_external = false;  // memory given by mdmc_ctx_set_scratch
  mdmc::Scratch s{};
  mdmc::Plan plan{};
  bool counted = false;
  bool coords_valid = false;  // coordinate tables hold coords_box for the counted plan
  bool emitted = false;       // emit_args describe the last emit of the counted plan
  // device buffers of the host entry (mdmc_run)
  void *run_field = nullptr, *run_counts = nullptr, *run_v = nullptr, *run_q = nullptr;
  size_t run_field_bytes = 0, run_counts_bytes = 0, run_v_bytes = 0, run_q_bytes = 0;
  // device buffers of the out-of-core entry: two field slabs (sized exactly
  // for the last call's slabs), its counts and slab record, a slab's mesh
  void *ooc_field[2] = {nullptr, nullptr}, *ooc_counts = nullptr, *ooc_v = nullptr, *ooc_q = nullptr;
  size_t ooc_field_bytes[2] = {0, 0}, ooc_v_bytes = 0, ooc_q_bytes = 0;
  mdmc::EmitArgs emit_args{};
  float coords_box[6] = {};
  const float *counted_field = nullptr;
  float iso = 0.f;
  // batches: device table of BatchVol + bases + the batch's own counts, and
  // two pinned staging tables (count, emit) uploaded with async copies
  void *bdev = nullptr;
  size_t bdev_bytes = 0;
  mdmc::BatchVol *bhost[2] = {nullptr, nullptr};
  size_t bhost_cap = 0;  // entries per staging table
  hipEvent_t bhost_ev[2] = {nullptr, nullptr};
  bool bhost_ev_used[2] = {false, false};
  bool batch_counted = false;
  uint32_t batch_n = 0, batch_W4 = 0, batch_KP = 0, batch_tiles = 0;
  uint64_t batch_key = 0;
  mdmc::Plan batch_plan{};
  // batched field generation: pinned staging + device copy of GenVol table and
  // the per-axis tables
  void *gdev = nullptr;
  size_t gdev_bytes = 0;
  void *ghost = nullptr;
  size_t ghost_bytes = 0;
  hipEvent_t ghost_ev = nullptr;
  bool ghost_ev_used = false;
  // stream capture: a context used inside a capture belongs to that graph
  // (its kernels and copies hold pointers to the scratch and staging
  // buffers), so every later call outside that capture is refused
  bool captured = false;
  unsigned long long capture_id = 0;
  bool capturing_now = false;  // the current call's stream is capturing
  bool tickets_zero = false;   // the emit's ticket counters are zero (a count ran after the last emit)
  // size the scratch exactly (no headroom): set for the out-of-core
  // context, which keeps to a device budget
  bool exact_scratch = false;
  // tuning options (0 = default)
  int opt[mdmc::kNumOptions] = {};
  // per-kernel event timing
  bool profiling = false;
  std::vector<hipEvent_t> events;  // pool, used in (start, stop) pairs
  std::vector<int> event_kernel;   // kernel id of each used pair
  size_t events_used = 0;          // pairs in use
};

namespace mdmc {

// Brackets one launch with HIP events when the context is profiling.
struct ProfScope {
  mdmc_ctx *ctx;
  hipStream_t st;
  hipEvent_t stop = nullptr;
  ProfScope(mdmc_ctx *c, int kernel, hipStream_t s) : ctx(c), st(s) {
    if (!ctx || !ctx->profiling) return;
    const size_t pair = ctx->events_used;
    while (ctx->events.size() < 2 * (pair + 1)) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) { (void)hipGetLastError(); return; }
      ctx->events.push_back(e);
    }
    if (ctx->event_kernel.size() < pair + 1) ctx->event_kernel.resize(pair + 1);
    ctx->event_kernel[pair] = kernel;
    if (hipEventRecord(ctx->events[2 * pair], st) != hipSuccess) { (void)hipGetLastError(); return; }
    stop = ctx->events[2 * pair + 1];
    ctx->events_used = pair + 1;
  }
  ~ProfScope() {
    if (stop) (void)hipEventRecord(stop, st);
  }
};

static size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Called first by every entry point that launches work on a context.  While
// `st` is capturing, the context becomes the property of that capture (the
// graph's nodes hold pointers to its scratch, staging tables and volume
// table, which a later call would rewrite, regrow or free); calls outside
// that capture, eager or in another capture, are refused from then on.  The
// context stays usable by the graph until it is destroyed.
int capture_guard(mdmc_ctx *ctx, hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  unsigned long long id = 0;
  HIP_TRY(hipStreamGetCaptureInfo(st, &cs, &id));
  ctx->capturing_now = cs != hipStreamCaptureStatusNone;
  if (ctx->capturing_now) {
    if (ctx->captured && ctx->capture_id != id)
      return fail(MDMC_ERR_INVALID, "context already belongs to another captured graph (use one context per graph)");
    ctx->captured = true;
    ctx->capture_id = id;
    return MDMC_OK;
  }
  if (ctx->captured)
    return fail(MDMC_ERR_INVALID,
                "context belongs to a captured graph whose nodes use its buffers; use another context outside it");
  return MDMC_OK;
}

struct ScratchLayout {
  size_t counts, vq, ne, chunk, state, ticket, totals, slots, gmap, coords, total;
};

ScratchLayout scratch_layout(const Plan &p) {
  ScratchLayout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += align256(bytes);
    return o;
  };
  const size_t seg_padded = (size_t)p.ntiles * kScanTile;  // the scan reads whole tiles
  L.counts = take(seg_padded * sizeof(uint2));
  L.vq = take(seg_padded * sizeof(uint2));
  L.ne = take((p.nsegs + 1) * sizeof(uint2));
  L.chunk = take(p.max_chunks * sizeof(uint32_t));
  L.state = take((size_t)p.ntiles * 4 * sizeof(unsigned long long));
  L.ticket = take(kTicketBytes);
  L.totals = take(sizeof(Totals));
  L.slots = take(p.nslots * sizeof(uint64_t));
  L.gmap = take(p.nsegs * (p.cap / 64) * sizeof(uint4));
  L.coords = take(p.ncoord * sizeof(float));
  L.total = off;
  return L;
}

int ensure_scratch(mdmc_ctx *ctx, const Plan &p) {
  Scratch s{};
  const ScratchLayout lay = scratch_layout(p);
  const size_t off = lay.total;
  const size_t o_counts = lay.counts, o_vq = lay.vq, o_ne = lay.ne, o_chunk = lay.chunk, o_state = lay.state,
               o_ticket = lay.ticket, o_totals = lay.totals, o_slots = lay.slots, o_gmap = lay.gmap,
               o_coords = lay.coords;
  if (off > ctx->scratch_bytes && ctx->scratch_external)
    return fail(MDMC_ERR_CAPACITY, "the scratch memory given to the context is too small");
  const bool regrow = off > ctx->scratch_bytes;
  if (regrow && ctx->capturing_now)
    return fail(MDMC_ERR_INVALID, "scratch must be sized by a call on this context before stream capture");
  if (regrow) {
    if (ctx->scratch) {
      hipError_t e = hipFree(ctx->scratch);
      (void)e;
      ctx->scratch = nullptr;
      ctx->scratch_bytes = 0;
    }
    // headroom for slightly larger grids, unless the context keeps to a budget
    const size_t want = ctx->exact_scratch ? off : off + off / 8;
    if (hipMalloc(&ctx->scratch, want) != hipSuccess) {
      (void)hipGetLastError();
      if (hipMalloc(&ctx->scratch, off) != hipSuccess) {
        (void)hipGetLastError();
        ctx->scratch = nullptr;
        return fail(MDMC_ERR_NOMEM, "scratch allocation failed");
      }
      ctx->scratch_bytes = off;
    } else {
      ctx->scratch_bytes = want;
    }
  }
  char *b = static_cast<char *>(ctx->scratch);
  s.seg_counts = reinterpret_cast<uint2 *>(b + o_counts);
  s.seg_vq = reinterpret_cast<uint2 *>(b + o_vq);
  s.ne = reinterpret_cast<uint2 *>(b + o_ne);
  s.chunk_ne = reinterpret_cast<uint32_t *>(b + o_chunk);
  s.scan_state = reinterpret_cast<unsigned long long *>(b + o_state);
  s.ticket = reinterpret_cast<uint32_t *>(b + o_ticket);
  s.totals = reinterpret_cast<Totals *>(b + o_totals);
  s.slots = reinterpret_cast<uint64_t *>(b + o_slots);
  s.gmap = reinterpret_cast<uint4 *>(b + o_gmap);
  s.coords = reinterpret_cast<float *>(b + o_coords);
  ctx->s = s;
  return MDMC_OK;
}

__global__ void k_zero_counts(uint32_t *c, Totals *t) {
  if (threadIdx.x < 4 && c) c[threadIdx.x] = 0u;
  if (threadIdx.x < 8 && t) reinterpret_cast<uint32_t *>(t)[threadIdx.x] = 0u;
}

// Coordinate tables of a plan for a box: xyz_range = xyz_max - xyz_min in
// float (dmc.cu:2149).
CoordArgs coord_args(const Plan &p, const Scratch &s, const float mn[3], const float mx[3]) {
  return CoordArgs{s.coords, p.X, p.Y, p.Z, p.cx, p.cy, p.cz, mn[0], mn[1], mn[2],
                   mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
}

// Grid of the single-volume emit (k_emit): its persistent workgroups, a
// multiple of 8 (one chunk ticket counter per XCD), no more than one wave per
// chunk of the capacity needs.
uint32_t emit_blocks(const mdmc_ctx *ctx, uint64_t max_chunks) {
  const int bpc = ctx->opt[MDMC_OPT_EMIT_BLOCKS_PER_CU] > 0 ? ctx->opt[MDMC_OPT_EMIT_BLOCKS_PER_CU] : kEmitPersistPerCu;
  const uint64_t want = (max_chunks + kEmitThreads / 64 - 1) / (kEmitThreads / 64);
  const uint64_t b = std::min<uint64_t>(want, (uint64_t)ctx->num_cus * (uint64_t)bpc);
  return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(8, (b + 7) / 8 * 8), 0x7FFFFFF8ull);
}

// box (optional): xyz_min, xyz_max; with it the scan also fills the
// coordinate tables the emit reads.
int launch_count(mdmc_ctx *ctx, const Plan &p, const float *d_field, float iso,
                 mdmc_counts *d_counts, hipStream_t st, const float *box_min = nullptr,
                 const float *box_max = nullptr) {
  const Scratch &s = ctx->s;
  ctx->coords_valid = false;
  ctx->emitted = false;
  if (p.empty) {
    hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, st,
                       reinterpret_cast<uint32_t *>(d_counts), s.totals);
    HIP_TRY(hipGetLastError());
    return MDMC_OK;
  }
  ClassifyArgs ca{};
  ca.g = FieldDesc{d_field, p.X, p.Y, p.Z, p.cx, p.cy, p.cz, iso};
  ca.XT = p.XT;
  ca.NRB = p.NRB;
  ca.NKC = p.NKC;
  ca.Kc = p.Kc;
  ca.seg_counts = s.seg_counts;
  ca.slots = s.slots;
  ca.gmap = s.gmap;
  ca.scan_state = s.scan_state;
  ca.scan_state_words = p.ntiles * 4;
  ca.seg_pad = s.seg_counts + p.nsegs;
  ca.seg_pad_n = (uint32_t)((uint64_t)p.ntiles * kScanTile - p.nsegs);
  ca.scan_ticket = s.ticket;
  ca.totals = s.totals;
  ca.counts_out = reinterpret_cast<uint32_t *>(d_counts);
  int rc;
  {
    ProfScope ps(ctx, 0, st);
    if (!p.zmarch) {
      PtArgs pa{ca, p.band, p.NKBp, p.NRBp};
      const bool vec = p.vec4 && ((uintptr_t)d_field % 16 == 0);
      rc = launch_classify_pt(p.W, p.ptR, p.ptKP, vec, pa, (uint32_t)p.pt_blocks, st);
    } else {
      rc = launch_classify(p.W, p.R, ca, (uint32_t)p.classify_blocks, st);
    }
  }
  if (rc != MDMC_OK) return fail(MDMC_ERR_INVALID, "no classify kernel for this tile shape");
  HIP_TRY(hipGetLastError());
  ctx->tickets_zero = true;

  ScanArgs sa{};
  sa.seg_counts = s.seg_counts;
  sa.nsegs = (uint32_t)p.nsegs;
  sa.ntiles = p.ntiles;
  sa.max_chunks = p.max_chunks;
  sa.seg_vq = s.seg_vq;
  sa.ne = s.ne;
  sa.chunk_ne = s.chunk_ne;
  sa.state = s.scan_state;
  sa.ticket = s.ticket;
  sa.totals = s.totals;
  sa.counts_out = reinterpret_cast<uint32_t *>(d_counts);
  if (box_min && box_max) {
    sa.co = coord_args(p, s, box_min, box_max);
    sa.fill_coords = 1;
    std::memcpy(ctx->coords_box, box_min, 3 * sizeof(float));
    std::memcpy(ctx->coords_box + 3, box_max, 3 * sizeof(float));
    ctx->coords_valid = true;
  }
  {
    ProfScope ps(ctx, 1, st);
    hipLaunchKernelGGL(k_scan, dim3(p.ntiles), dim3(kScanThreads), 0, st, sa);
  }
  HIP_TRY(hipGetLastError());
  return MDMC_OK;
}


int launch_emit(mdmc_ctx *ctx, const Plan &p, const float *d_field, float iso,
                const float mn[3], const float mx[3], float *d_vertices, uint64_t vcap,
                int32_t *d_quads, uint64_t qcap, mdmc_counts *d_counts, hipStream_t st) {
  if (p.empty) return MDMC_OK;
  const Scratch &s = ctx->s;
  EmitArgs ea{};
  ea.v.g = FieldDesc{d_field, p.X, p.Y, p.Z, p.cx, p.cy, p.cz, iso};
  const CoordArgs co = coord_args(p, s, mn, mx);
  // the count already filled the tables for this box
  const bool have_coords = ctx->coords_valid && std::memcmp(ctx->coords_box, mn, 3 * sizeof(float)) == 0 &&
                           std::memcmp(ctx->coords_box + 3, mx, 3 * sizeof(float)) == 0;
  ea.v.xt_c = s.coords;
  ea.v.yt_c = s.coords + p.X;
  ea.v.zt_c = s.coords + p.X + p.Y;
  ea.v.XT = p.XT;
  ea.cap = p.cap;
  ea.nsegs = (uint32_t)p.nsegs;
  ea.max_chunks = p.max_chunks;
  ea.seg_counts = s.seg_counts;
  ea.seg_vq = s.seg_vq;
  ea.slots = s.slots;
  ea.gmap = s.gmap;
  ea.ne = s.ne;
  ea.chunk_ne = s.chunk_ne;
  ea.totals = s.totals;
  ea.v.verts = d_vertices;
  ea.v.vcap = d_vertices ? vcap : 0;
  ea.v.quads = d_quads;
  ea.v.qcap = d_quads ? qcap : 0;
  ea.v.status = reinterpret_cast<uint32_t *>(d_counts) + 3;
  {
    ProfScope ps(ctx, 2, st);
    if (!have_coords) {
      hipLaunchKernelGGL(k_coords, dim3((uint32_t)std::min<uint64_t>((p.ncoord + 255) / 256, 1024)), dim3(256), 0,
                         st, co);
      // the tables now hold this box: a later emit of the same count with
      // another box must rewrite them, one with this box may reuse them
      std::memcpy(ctx->coords_box, mn, 3 * sizeof(float));
      std::memcpy(ctx->coords_box + 3, mx, 3 * sizeof(float));
      ctx->coords_valid = true;
    }
    ea.emit_ctr = s.ticket;
    // the count zeroed the ticket counters; an emit after another (or one
    // being captured, whose replays may follow one) zeroes them first
    if (!ctx->tickets_zero || ctx->capturing_now) hipLaunchKernelGGL(k_zero_tickets, dim3(1), dim3(64), 0, st, s.ticket);
    hipLaunchKernelGGL(k_emit, dim3(emit_blocks(ctx, p.max_chunks)), dim3(kEmitThreads), 0, st, ea);
    ctx->tickets_zero = false;
  }
  HIP_TRY(hipGetLastError());
  ctx->emitted = true;
  ctx->emit_args = ea;
  return MDMC_OK;
}

// Smoothing iterations over the last emit's mesh (the given outputs must be
// the ones that emit wrote).
int launch_smooth(mdmc_ctx *ctx, float *d_vertices, uint64_t vcap, int32_t *d_quads, uint64_t qcap,
                  uint32_t num_smooth, mdmc_counts *d_counts, hipStream_t st) {
  if (num_smooth == 0 || ctx->plan.empty) return MDMC_OK;
  EmitArgs ea = ctx->emit_args;
  ea.v.verts = d_vertices;
  ea.v.vcap = d_vertices ? vcap : 0;
  ea.v.quads = d_quads;
  ea.v.qcap = d_quads ? qcap : 0;
  ea.v.status = reinterpret_cast<uint32_t *>(d_counts) + 3;
  const uint64_t want = (ctx->plan.max_chunks + kEmitThreads / 64 - 1) / (kEmitThreads / 64);
  const int bpc = ctx->opt[MDMC_OPT_EMIT_BLOCKS_PER_CU] > 0 ? ctx->opt[MDMC_OPT_EMIT_BLOCKS_PER_CU] : kEmitBlocksPerCu;
  const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)ctx->num_cus * bpc));
  ProfScope ps(ctx, 3, st);
  for (uint32_t it = 0; it < num_smooth; ++it) {
    hipLaunchKernelGGL(k_smooth, dim3(blocks), dim3(kEmitThreads), 0, st, ea);
    hipLaunchKernelGGL(k_reaverage, dim3(blocks), dim3(kEmitThreads), 0, st, ea);
  }
  HIP_TRY(hipGetLastError());
  return MDMC_OK;
}

// ---------------------------------------------------------------------------
// Batches

constexpr uint32_t kBatchRows = 7;  // pt_rows for every x-tile width

uint32_t batch_planes(const mdmc_ctx *ctx, uint32_t W4) {
  const int kp = ctx->opt[MDMC_OPT_BATCH_PLANES];
  if (W4 == 1 && (kp == 2 || kp == 4)) return (uint32_t)kp;
  if (W4 == 2 && (kp == 1 || kp == 2)) return (uint32_t)kp;
  return pt_planes(W4);
}

template <int W4, int KP>
void launch_classify_batch_t(const BatchArgs &b, uint32_t blocks, hipStream_t st) {
  hipLaunchKernelGGL((k_classify_batch<W4, (int)kBatchRows, KP>), dim3(blocks), dim3(kPtThreads), 0, st, b);
}

int launch_classify_batch(uint32_t W4, uint32_t KP, const BatchArgs &b, uint32_t blocks, hipStream_t st) {
  if (W4 == 1 && KP == 4) { launch_classify_batch_t<1, 4>(b, blocks, st); return 0; }
  if (W4 == 1 && KP == 2) { launch_classify_batch_t<1, 2>(b, blocks, st); return 0; }
  if (W4 == 2 && KP == 2) { launch_classify_batch_t<2, 2>(b, blocks, st); return 0; }
  if (W4 == 2 && KP == 1) { launch_classify_batch_t<2, 1>(b, blocks, st); return 0; }
  if (W4 == 4 && KP == 1) { launch_classify_batch_t<4, 1>(b, blocks, st); return 0; }
  return -1;
}

uint64_t batch_key_of(const mdmc_volume *vols, uint32_t n) {
  uint64_t h = 0x9E3779B97F4A7C15ull ^ n;
  auto mix = [&](uint64_t v) { h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2); };
  for (uint32_t v = 0; v < n; ++v) {
    mix((uint64_t)(uintptr_t)vols[v].field);
    mix(((uint64_t)vols[v].dims[0] << 32) | vols[v].dims[1]);
    mix(vols[v].dims[2]);
    uint32_t iso;
    std::memcpy(&iso, &vols[v].iso, 4);
    mix(iso);
  }
  return h;
}

// Device table (n BatchVol), bases (n uint2), batch counts (4 words); two
// pinned staging tables of n entries.
int ensure_batch_buffers(mdmc_ctx *ctx, uint32_t n) {
  const size_t o_base = align256(n * sizeof(BatchVol));
  const size_t o_cnt = o_base + align256(n * sizeof(uint2));
  const size_t need = o_cnt + 256;
  if ((need > ctx->bdev_bytes || n > ctx->bhost_cap) && ctx->capturing_now)
    return fail(MDMC_ERR_INVALID, "batch tables must be sized by a call on this context before stream capture");
  if (need > ctx->bdev_bytes) {
    if (ctx->bdev) (void)hipFree(ctx->bdev);
    ctx->bdev = nullptr;
    ctx->bdev_bytes = 0;
    if (hipMalloc(&ctx->bdev, need) != hipSuccess) {
      (void)hipGetLastError();
      return fail(MDMC_ERR_NOMEM, "batch table allocation failed");
    }
    ctx->bdev_bytes = need;
  }
  if (n > ctx->bhost_cap) {
    for (int i = 0; i < 2; ++i) {
      if (ctx->bhost_ev_used[i]) HIP_TRY(hipEventSynchronize(ctx->bhost_ev[i]));
      if (ctx->bhost[i]) (void)hipHostFree(ctx->bhost[i]);
      ctx->bhost[i] = nullptr;
    }
    ctx->bhost_cap = 0;
    for (int i = 0; i < 2; ++i) {
      if (hipHostMalloc(reinterpret_cast<void **>(&ctx->bhost[i]), kMaxBatch * sizeof(BatchVol)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(MDMC_ERR_NOMEM, "pinned staging allocation failed");
      }
    }
    ctx->bhost_cap = kMaxBatch;
  }
  for (int i = 0; i < 2; ++i)
    if (!ctx->bhost_ev[i]) HIP_TRY(hipEventCreateWithFlags(&ctx->bhost_ev[i], hipEventDisableTiming));
  return MDMC_OK;
}

BatchVol *batch_table_dev(mdmc_ctx *ctx) { return static_cast<BatchVol *>(ctx->bdev); }
uint2 *batch_base_dev(mdmc_ctx *ctx, uint32_t n) {
  return reinterpret_cast<uint2 *>(static_cast<char *>(ctx->bdev) + align256(n * sizeof(BatchVol)));
}
uint32_t *batch_counts_dev(mdmc_ctx *ctx, uint32_t n) {
  return reinterpret_cast<uint32_t *>(static_cast<char *>(ctx->bdev) + align256(n * sizeof(BatchVol)) +
                                      align256(n * sizeof(uint2)));
}

// Staging table `slot` may be rewritten once its previous upload has run (not
// while a graph is being captured: the captured copy reads it at replay).
int staging_acquire(mdmc_ctx *ctx, int slot, hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIP_TRY(hipStreamIsCapturing(st, &cs));
  if (cs == hipStreamCaptureStatusNone && ctx->bhost_ev_used[slot])
    HIP_TRY(hipEventSynchronize(ctx->bhost_ev[slot]));
  return MDMC_OK;
}

int staging_upload(mdmc_ctx *ctx, int slot, uint32_t n, hipStream_t st) {
  HIP_TRY(hipMemcpyAsync(ctx->bdev, ctx->bhost[slot], n * sizeof(BatchVol), hipMemcpyHostToDevice, st));
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIP_TRY(hipStreamIsCapturing(st, &cs));
  if (cs == hipStreamCaptureStatusNone) {
    HIP_TRY(hipEventRecord(ctx->bhost_ev[slot], st));
    ctx->bhost_ev_used[slot] = true;
  }
  return MDMC_OK;
}

__global__ void k_zero_words(uint32_t *w, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) w[i] = 0u;
}

// Fills staging table `slot` for the batch; with_outputs: copy the output
// pointers and capacities (the count pass leaves them out when absent).
int fill_batch_table(mdmc_ctx *ctx, int slot, const mdmc_volume *vols, uint32_t n, mdmc_counts *d_counts,
                     uint32_t W4, uint32_t KP, const Scratch &S, uint64_t *nsegs_total, uint64_t *cells_total,
                     uint32_t *tiles_total) {
  const uint32_t cap = 256u * W4;
  uint64_t seg_off = 0, tile_off = 0, cells = 0, coord_off = 0;
  for (uint32_t v = 0; v < n; ++v) {
    const mdmc_volume &mv = vols[v];
    BatchVol &b = ctx->bhost[slot][v];
    std::memset(&b, 0, sizeof(b));
    const uint32_t X = mv.dims[0], Y = mv.dims[1], Z = mv.dims[2];
    const bool empty = X < 2 || Y < 2 || Z < 2;
    const uint32_t cx = empty ? 0 : X - 1, cy = empty ? 0 : Y - 1, cz = empty ? 0 : Z - 1;
    const uint32_t XT = empty ? 1 : (cx + cap - 1) / cap;
    const uint32_t NRB = empty ? 0 : (cy + kBatchRows - 1) / kBatchRows;
    const uint32_t NKP = empty ? 0 : (cz + KP - 1) / KP;
    const uint64_t nsegs = (uint64_t)XT * cy * cz;
    b.c.g = FieldDesc{mv.field, X, Y, Z, cx, cy, cz, mv.iso};
    b.c.XT = XT;
    b.c.seg_counts = S.seg_counts + seg_off;
    b.c.slots = S.slots + seg_off * cap;
    b.c.gmap = S.gmap + seg_off * (cap / 64);
    b.tile_off = (uint32_t)tile_off;
    b.NRB = NRB;
    b.NKP = NKP;
    b.vec = (X % 4 == 0 && (reinterpret_cast<uintptr_t>(mv.field) & 15u) == 0) ? 1u : 0u;
    b.seg_off = (uint32_t)seg_off;
    b.nsegs = (uint32_t)nsegs;
    b.mnx = mv.xyz_min[0];
    b.mny = mv.xyz_min[1];
    b.mnz = mv.xyz_min[2];
    // xyz_range = xyz_max - xyz_min in float (dmc.cu:2149)
    b.rgx = mv.xyz_max[0] - mv.xyz_min[0];
    b.rgy = mv.xyz_max[1] - mv.xyz_min[1];
    b.rgz = mv.xyz_max[2] - mv.xyz_min[2];
    b.verts = mv.vertices;
    b.vcap = mv.vertices ? mv.vertex_capacity : 0;
    b.quads = mv.quads;
    b.qcap = mv.quads ? mv.quad_capacity : 0;
    b.counts = reinterpret_cast<uint32_t *>(d_counts + v);
    b.coords = empty ? nullptr : S.coords + coord_off;
    if (!empty) coord_off += (uint64_t)X + Y + Z;
    seg_off += nsegs;
    tile_off += (uint64_t)XT * NRB * NKP;
    cells += (uint64_t)cx * cy * cz;
  }
  *nsegs_total = seg_off;
  *cells_total = cells;
  *tiles_total = (uint32_t)tile_off;
  return MDMC_OK;
}

// x-tile width and the aggregate plan (what the scratch buffers and the scan
// see) of a batch.
int batch_layout(const mdmc_volume *vols, uint32_t n, bool check_fields, Plan *agg, uint32_t *W4_out,
                 uint32_t *KP_out) {
  for (uint32_t v = 0; v < n; ++v) {
    const mdmc_volume &mv = vols[v];
    if (mv.dims[0] < 2 || mv.dims[1] < 2 || mv.dims[2] < 2) continue;
    if (check_fields && !mv.field) return fail(MDMC_ERR_INVALID, "null field in batch");
    if ((uint64_t)mv.dims[0] * mv.dims[1] * mv.dims[2] > 0xFFFFFFFFull)
      return fail(MDMC_ERR_TOO_LARGE, "grid has more than 2^32-1 points");
  }
  // x-tile width: every volume may span several x-tiles, so take the width
  // with the fewest padded cells over the batch (lanes past a row's end do
  // no work), the wider on a tie (fewer tiles)
  uint32_t W4 = 1;
  uint64_t best = ~0ull;
  for (uint32_t w : {2u, 1u}) {  // at most 512-cell x-tiles, as for single volumes
    uint64_t padded = 0;
    for (uint32_t v = 0; v < n; ++v) {
      const mdmc_volume &mv = vols[v];
      if (mv.dims[0] < 2 || mv.dims[1] < 2 || mv.dims[2] < 2) continue;
      const uint64_t c = 256ull * w;
      padded += (mv.dims[0] - 1 + c - 1) / c * c * (mv.dims[1] - 1) * (uint64_t)(mv.dims[2] - 1);
    }
    if (padded < best) {
      best = padded;
      W4 = w;
    }
  }
  const uint32_t KP = pt_planes(W4);
  const uint32_t cap = 256u * W4;
  uint64_t nseg = 0, tiles = 0, cells = 0, ncoord = 0;
  for (uint32_t v = 0; v < n; ++v) {
    const mdmc_volume &mv = vols[v];
    if (mv.dims[0] < 2 || mv.dims[1] < 2 || mv.dims[2] < 2) continue;
    const uint64_t XT = (mv.dims[0] - 1 + cap - 1) / cap;
    nseg += XT * (mv.dims[1] - 1) * (uint64_t)(mv.dims[2] - 1);
    tiles += XT * ((mv.dims[1] - 1 + kBatchRows - 1) / kBatchRows) * ((mv.dims[2] - 1 + KP - 1) / KP);
    cells += (uint64_t)(mv.dims[0] - 1) * (mv.dims[1] - 1) * (mv.dims[2] - 1);
    ncoord += (uint64_t)mv.dims[0] + mv.dims[1] + mv.dims[2];
  }
  if (nseg >= 0x7FFFFFFFull || tiles >= 0x7FFFFFFFull) return fail(MDMC_ERR_TOO_LARGE, "batch has too many cell rows");
  *agg = Plan{};
  agg->cap = cap;
  agg->nsegs = nseg;
  agg->nslots = nseg * cap;
  agg->ntiles = (uint32_t)(nseg / kScanTile + 1);
  agg->max_chunks = (cells + 63) / 64 + 1;
  agg->classify_blocks = tiles;
  agg->ncoord = ncoord;
  agg->empty = nseg == 0;
  *W4_out = W4;
  *KP_out = KP;
  return MDMC_OK;
}

int batch_count(mdmc_ctx *ctx, const mdmc_volume *vols, uint32_t n, mdmc_counts *d_counts, hipStream_t st) {
  Plan agg{};
  uint32_t W4 = 1, KP = 1;
  int rc0 = batch_layout(vols, n, true, &agg, &W4, &KP);
  if (rc0 != MDMC_OK) return rc0;
  KP = batch_planes(ctx, W4);
  const uint64_t nseg_est = agg.nsegs;
  int rc = ensure_batch_buffers(ctx, n);
  if (rc != MDMC_OK) return rc;
  ctx->batch_counted = false;
  ctx->counted = false;  // the scratch now holds this batch
  ctx->emitted = false;
  if (nseg_est == 0) {  // nothing to extract: zero every volume's counts
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(256), 0, st, reinterpret_cast<uint32_t *>(d_counts), 4u * n);
    HIP_TRY(hipGetLastError());
    ctx->batch_n = n;
    ctx->batch_key = batch_key_of(vols, n);
    ctx->batch_plan = Plan{};
    ctx->batch_plan.empty = 1;
    ctx->batch_counted = true;
    return MDMC_OK;
  }
  rc = staging_acquire(ctx, 0, st);
  if (rc != MDMC_OK) return rc;
  rc = ensure_scratch(ctx, agg);
  if (rc != MDMC_OK) return rc;
  const Scratch &S = ctx->s;
  uint64_t nsegs = 0, cells2 = 0;
  uint32_t tiles = 0;
  rc = fill_batch_table(ctx, 0, vols, n, d_counts, W4, KP, S, &nsegs, &cells2, &tiles);
  if (rc != MDMC_OK) return rc;
  rc = staging_upload(ctx, 0, n, st);
  if (rc != MDMC_OK) return rc;

  uint32_t *bcounts = batch_counts_dev(ctx, n);
  BatchArgs ba{};
  ba.vols = batch_table_dev(ctx);
  ba.nvol = n;
  ba.total_tiles = tiles;
  ba.zero.scan_state = S.scan_state;
  ba.zero.scan_state_words = agg.ntiles * 4;
  ba.zero.seg_pad = S.seg_counts + nsegs;
  ba.zero.seg_pad_n = (uint32_t)((uint64_t)agg.ntiles * kScanTile - nsegs);
  ba.zero.scan_ticket = S.ticket;
  ba.zero.totals = S.totals;
  ba.zero.counts_out = bcounts;
  {
    ProfScope ps(ctx, 0, st);
    // default: a workgroup per tile (each XCD's workgroups take its share of
    // the tiles in order); the option makes the grid persistent
    const int bpc = ctx->opt[MDMC_OPT_CLASSIFY_BLOCKS_PER_CU];
    const uint64_t want = bpc > 0 ? (uint64_t)ctx->num_cus * bpc : ((uint64_t)tiles + 7) / 8 * 8;
    const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, ((uint64_t)tiles + 7) / 8 * 8));
    if (launch_classify_batch(W4, KP, ba, blocks, st) != 0) return fail(MDMC_ERR_INVALID, "no batch classify kernel");
  }
  HIP_TRY(hipGetLastError());
  ScanArgs sa{};
  sa.seg_counts = S.seg_counts;
  sa.nsegs = (uint32_t)nsegs;
  sa.ntiles = agg.ntiles;
  sa.max_chunks = agg.max_chunks;
  sa.seg_vq = S.seg_vq;
  sa.ne = S.ne;
  sa.chunk_ne = S.chunk_ne;
  sa.state = S.scan_state;
  sa.ticket = S.ticket;
  sa.totals = S.totals;
  sa.counts_out = bcounts;
  {
    ProfScope ps(ctx, 1, st);
    hipLaunchKernelGGL(k_scan, dim3(agg.ntiles), dim3(kScanThreads), 0, st, sa);
  }
  HIP_TRY(hipGetLastError());
  BatchCountArgs bc{};
  bc.vols = batch_table_dev(ctx);
  bc.nvol = n;
  bc.nsegs = (uint32_t)nsegs;
  bc.seg_vq = S.seg_vq;
  bc.ne = S.ne;
  bc.totals = S.totals;
  bc.base = batch_base_dev(ctx, n);
  bc.batch_counts = bcounts;
  {
    ProfScope ps(ctx, 3, st);
    hipLaunchKernelGGL(k_batch_counts, dim3((n + 255) / 256), dim3(256), 0, st, bc);
  }
  HIP_TRY(hipGetLastError());
  agg.nsegs = nsegs;
  ctx->batch_plan = agg;
  ctx->batch_n = n;
  ctx->batch_W4 = W4;
  ctx->batch_KP = KP;
  ctx->batch_tiles = tiles;
  ctx->batch_key = batch_key_of(vols, n);
  ctx->batch_counted = true;
  return MDMC_OK;
}

int batch_emit(mdmc_ctx *ctx, const mdmc_volume *vols, uint32_t n, mdmc_counts *d_counts, hipStream_t st,
               bool upload) {
  if (!ctx->batch_counted || ctx->batch_n != n || ctx->batch_key != batch_key_of(vols, n))
    return fail(MDMC_ERR_INVALID, "batch emit without a matching batch count on this context");
  const Plan &agg = ctx->batch_plan;
  if (agg.empty) return MDMC_OK;
  const Scratch &S = ctx->s;
  if (upload) {
    int rc = staging_acquire(ctx, 1, st);
    if (rc != MDMC_OK) return rc;
    uint64_t nsegs = 0, cells = 0;
    uint32_t tiles = 0;
    rc = fill_batch_table(ctx, 1, vols, n, d_counts, ctx->batch_W4, ctx->batch_KP, S, &nsegs, &cells, &tiles);
    if (rc != MDMC_OK) return rc;
    rc = staging_upload(ctx, 1, n, st);
    if (rc != MDMC_OK) return rc;
  }
  BatchEmitArgs be{};
  be.a.cap = agg.cap;
  be.a.nsegs = (uint32_t)agg.nsegs;
  be.a.max_chunks = agg.max_chunks;
  be.a.seg_counts = S.seg_counts;
  be.a.seg_vq = S.seg_vq;
  be.a.slots = S.slots;
  be.a.gmap = S.gmap;
  be.a.ne = S.ne;
  be.a.chunk_ne = S.chunk_ne;
  be.a.totals = S.totals;
  be.a.v.status = batch_counts_dev(ctx, n) + 3;
  be.vols = batch_table_dev(ctx);
  be.base = batch_base_dev(ctx, n);
  be.nvol = n;
  {
    ProfScope ps(ctx, 2, st);
    const int bpc = ctx->opt[MDMC_OPT_EMIT_BLOCKS_PER_CU] > 0 ? ctx->opt[MDMC_OPT_EMIT_BLOCKS_PER_CU] : kEmitBlocksPerCu;
    const uint64_t want = (agg.max_chunks + kEmitThreads / 64 - 1) / (kEmitThreads / 64);
    const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)ctx->num_cus * bpc));
    hipLaunchKernelGGL(k_coords_batch, dim3(std::min<uint32_t>(n, (uint32_t)ctx->num_cus * 4)), dim3(256), 0, st,
                       static_cast<const BatchVol *>(batch_table_dev(ctx)), n);
    hipLaunchKernelGGL(k_emit_batch, dim3(blocks), dim3(kEmitThreads), 0, st, be);
    hipLaunchKernelGGL(k_batch_status, dim3((n + 255) / 256), dim3(256), 0, st,
                       static_cast<const BatchVol *>(batch_table_dev(ctx)), n,
                       static_cast<const uint32_t *>(batch_counts_dev(ctx, n) + 3));
  }
  HIP_TRY(hipGetLastError());
  return MDMC_OK;
}

}  // namespace mdmc

using namespace mdmc;

// ---------------------------------------------------------------------------
// C ABI

int mdmc::set_error(int code, const char *msg) { return fail(code, msg); }

extern "C" {

const char *mdmc_status_string(int status) {
  switch (status) {
    case MDMC_OK: return "ok";
    case MDMC_ERR_INVALID: return "invalid argument";
    case MDMC_ERR_HIP: return "HIP runtime error";
    case MDMC_ERR_NOMEM: return "out of memory";
    case MDMC_ERR_CAPACITY: return "output capacity too small";
    case MDMC_ERR_TOO_LARGE: return "grid too large";
    case MDMC_ERR_NOT_SETUP: return "mdmc_setup() not called";
    case MDMC_ERR_INTERNAL: return "device invariant check failed";
    case MDMC_ERR_REJECTED: return "request rejected by admission";
    default: return "unknown status";
  }
}

const char *mdmc_last_error(void) { return g_last_error.c_str(); }


int mdmc_setup(void) {
  int dev = 0;
  int rc = current_device(&dev);
  if (rc != MDMC_OK) return rc;
  if (dev < 0 || dev >= 64) return fail(MDMC_ERR_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lk(g_setup_mutex);
  if ((g_setup_mask >> dev) & 1ull) return MDMC_OK;
  mdmc_luts L;
  mdmc_build_luts(&L);
  static uint64_t packed[2][256];
  static uint16_t info[256];
  mdmc_pack_luts(&L, packed, info);
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_label), packed, sizeof(packed)));
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_cfg_info), info, sizeof(info)));
  static uint16_t count[256];
  for (int c = 0; c < 256; ++c) {
    const int b6 = ((c >> 2) ^ (c >> 6)) & 1, b9 = ((c >> 5) ^ (c >> 6)) & 1, b10 = ((c >> 7) ^ (c >> 6)) & 1;
    const int nbip = b6 + b9 + b10;
    uint16_t w = (uint16_t)((L.num_vertex[0][c] + nbip) | ((L.num_vertex[1][c] + nbip) << 4) |
                            ((b6 | (b9 << 1) | (b10 << 2)) << 8));
    if (L.amb_face[c] != MDMC_NOT_AMBIGUOUS)
      w = (uint16_t)(w | (1u << 11) | ((unsigned)MDMC_FACE_DIR[L.amb_face[c]] << 12));
    count[c] = w;
  }
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_cfg_count), count, sizeof(count)));
  if (!g_noalloc_ptr[dev]) {
    void *ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(MDMC_ERR_HIP, "no device address for the no-allocation scratch");
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) {
      (void)hipGetLastError();
      cus = 256;
    }
    g_noalloc_cus[dev] = cus;
    g_noalloc_ptr[dev] = ptr;
  }
  g_setup_mask |= 1ull << dev;
  return MDMC_OK;
}

int mdmc_cleanup(void) {
  int dev = 0;
  int rc = current_device(&dev);
  if (rc != MDMC_OK) return rc;
  {
    std::lock_guard<std::mutex> entry(g_host_entry_mutex);
    if (g_default_ctx) {
      mdmc_ctx_destroy(g_default_ctx);
      g_default_ctx = nullptr;
    }
    if (g_ooc_ctx) {
      mdmc_ctx_destroy(g_ooc_ctx);
      g_ooc_ctx = nullptr;
    }
  }
  staged_release();
  std::lock_guard<std::mutex> lk(g_setup_mutex);
  if (dev >= 0 && dev < 64) {
    g_setup_mask &= ~(1ull << dev);
    if (g_ref_tables[dev]) {
      (void)hipFree(g_ref_tables[dev]);
      g_ref_tables[dev] = nullptr;
    }
  }
  return MDMC_OK;
}

static void fill_table_pointers(uint8_t *b, mdmc_device_luts *o) {
  o->config_edge_lut1 = b + kTabLut1;
  o->config_edge_lut2 = b + kTabLut2;
  o->num_vertex_lut1 = b + kTabNv1;
  o->num_vertex_lut2 = b + kTabNv2;
  o->config_2B_3B_lut = b + kTabAmb;
  o->config_2B_3B_ambiguous_face = b + kTabAmbFace;
  o->opposite_face_lut = b + kTabOpposite;
  o->face_to_check_dir_lut = b + kTabFaceDir;
  o->edge_belonged_voxel_lut = b + kTabEdgeOwner;
  o->circular_edge_lut = b + kTabRing;
  o->voxel_local_edges = b + kTabLocal;
}

int mdmc_setup_device_luts(mdmc_device_luts *out) {
  if (!out) return fail(MDMC_ERR_INVALID, "null argument");
  std::memset(out, 0, sizeof(*out));
  int rc = mdmc_setup();
  if (rc != MDMC_OK) return rc;
  int dev = 0;
  rc = current_device(&dev);
  if (rc != MDMC_OK) return rc;
  std::lock_guard<std::mutex> lk(g_setup_mutex);
  if (!g_ref_tables[dev]) {
    uint8_t host[kTabBytes];
    build_reference_tables(host);
    void *d = nullptr;
    if (hipMalloc(&d, kTabBytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(MDMC_ERR_NOMEM, "table allocation failed");
    }
    if (hipMemcpy(d, host, kTabBytes, hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipFree(d);
      return fail(MDMC_ERR_HIP, "table upload failed");
    }
    g_ref_tables[dev] = d;
  }
  fill_table_pointers(static_cast<uint8_t *>(g_ref_tables[dev]), out);
  return MDMC_OK;
}

int mdmc_reference_luts_host(uint8_t *out) {
  static_assert(kTabBytes == MDMC_REF_LUTS_BYTES, "table block size");
  if (!out) return fail(MDMC_ERR_INVALID, "null argument");
  build_reference_tables(out);
  return MDMC_OK;
}

int mdmc_cleanup_device_luts(const mdmc_device_luts *luts) {
  if (!luts) return fail(MDMC_ERR_INVALID, "null argument");
  int dev = 0;
  int rc = current_device(&dev);
  if (rc != MDMC_OK) return rc;
  {
    std::lock_guard<std::mutex> lk(g_setup_mutex);
    mdmc_device_luts mine;
    std::memset(&mine, 0, sizeof(mine));
    if (g_ref_tables[dev]) fill_table_pointers(static_cast<uint8_t *>(g_ref_tables[dev]), &mine);
    if (!g_ref_tables[dev] || std::memcmp(&mine, luts, sizeof(mine)) != 0)
      return fail(MDMC_ERR_INVALID, "table pointers were not handed out by mdmc_setup_device_luts on this device");
  }
  return mdmc_cleanup();  // frees the table block too
}

int mdmc_ctx_create(mdmc_ctx **out) {
  if (!out) return fail(MDMC_ERR_INVALID, "null ctx pointer");
  *out = nullptr;
  int rc = mdmc_setup();
  if (rc != MDMC_OK) return rc;
  mdmc_ctx *ctx = new (std::nothrow) mdmc_ctx();
  if (!ctx) return fail(MDMC_ERR_NOMEM, "ctx allocation failed");
  HIP_TRY(hipGetDevice(&ctx->device));
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) == hipSuccess &&
      cus > 0)
    ctx->num_cus = cus;
  *out = ctx;
  return MDMC_OK;
}

int mdmc_ctx_profile(mdmc_ctx *ctx, int enable) {
  if (!ctx) return fail(MDMC_ERR_INVALID, "null ctx");
  ctx->profiling = enable != 0;
  return MDMC_OK;
}

int mdmc_ctx_profile_read(mdmc_ctx *ctx, double ms_total[4], uint32_t launches[4]) {
  if (!ctx || !ms_total || !launches) return fail(MDMC_ERR_INVALID, "null argument");
  for (int i = 0; i < 4; ++i) { ms_total[i] = 0.0; launches[i] = 0; }
  for (size_t pr = 0; pr < ctx->events_used; ++pr) {
    HIP_TRY(hipEventSynchronize(ctx->events[2 * pr + 1]));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->events[2 * pr], ctx->events[2 * pr + 1]));
    const int k = std::min(3, std::max(0, ctx->event_kernel[pr]));
    ms_total[k] += ms;
    launches[k] += 1;
  }
  ctx->events_used = 0;
  return MDMC_OK;
}

void mdmc_ctx_destroy(mdmc_ctx *ctx) {
  if (!ctx) return;
  for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
  if (ctx->scratch && !ctx->scratch_external) (void)hipFree(ctx->scratch);
  for (void *b : {ctx->run_field, ctx->run_counts, ctx->run_v, ctx->run_q, ctx->ooc_field[0], ctx->ooc_field[1],
                  ctx->ooc_counts, ctx->ooc_v, ctx->ooc_q})
    if (b) (void)hipFree(b);
  if (ctx->bdev) (void)hipFree(ctx->bdev);
  if (ctx->gdev) (void)hipFree(ctx->gdev);
  if (ctx->ghost) (void)hipHostFree(ctx->ghost);
  if (ctx->ghost_ev) (void)hipEventDestroy(ctx->ghost_ev);
  for (int i = 0; i < 2; ++i) {
    if (ctx->bhost[i]) (void)hipHostFree(ctx->bhost[i]);
    if (ctx->bhost_ev[i]) (void)hipEventDestroy(ctx->bhost_ev[i]);
  }
  delete ctx;
}

static int count_impl(mdmc_ctx *ctx, const float *d_field, const uint32_t dims[3], float iso,
                      mdmc_counts *d_counts, void *stream, const float *box_min, const float *box_max) {
  if (!ctx || !dims || !d_counts) return fail(MDMC_ERR_INVALID, "null argument");
  if (!is_setup_on(ctx->device)) return fail(MDMC_ERR_NOT_SETUP, "mdmc_setup() not called");
  int rc = capture_guard(ctx, static_cast<hipStream_t>(stream));
  if (rc != MDMC_OK) return rc;
  Plan p;
  rc = plan_tiling(dims, &p, ctx->opt);
  if (rc != MDMC_OK) return rc;
  if (!p.empty && !d_field) return fail(MDMC_ERR_INVALID, "null field");
  rc = ensure_scratch(ctx, p);
  if (rc != MDMC_OK) return rc;
  rc = launch_count(ctx, p, d_field, iso, d_counts, static_cast<hipStream_t>(stream), box_min, box_max);
  if (rc != MDMC_OK) return rc;
  ctx->plan = p;
  ctx->counted = true;
  ctx->counted_field = d_field;
  ctx->iso = iso;
  return MDMC_OK;
}

int mdmc_count_device(mdmc_ctx *ctx, const float *d_field, const uint32_t dims[3], float iso,
                      mdmc_counts *d_counts, void *stream) {
  return count_impl(ctx, d_field, dims, iso, d_counts, stream, nullptr, nullptr);
}

int mdmc_emit_device(mdmc_ctx *ctx, const float *d_field, const float xyz_min[3],
                     const float xyz_max[3], float *d_vertices, uint64_t vertex_capacity,
                     int32_t *d_quads, uint64_t quad_capacity, mdmc_counts *d_counts,
                     void *stream) {
  if (!ctx || !xyz_min || !xyz_max || !d_counts) return fail(MDMC_ERR_INVALID, "null argument");
  if (!ctx->counted) return fail(MDMC_ERR_INVALID, "emit without a preceding count");
  if (d_field != ctx->counted_field) return fail(MDMC_ERR_INVALID, "emit field differs from counted field");
  const int grc = capture_guard(ctx, static_cast<hipStream_t>(stream));
  if (grc != MDMC_OK) return grc;
  return launch_emit(ctx, ctx->plan, d_field, ctx->iso, xyz_min, xyz_max, d_vertices, vertex_capacity,
                     d_quads, quad_capacity, d_counts, static_cast<hipStream_t>(stream));
}

int mdmc_extract_device(mdmc_ctx *ctx, const float *d_field, const uint32_t dims[3],
                        const float xyz_min[3], const float xyz_max[3], float iso,
                        float *d_vertices, uint64_t vertex_capacity, int32_t *d_quads,
                        uint64_t quad_capacity, mdmc_counts *d_counts, void *stream) {
  if (!xyz_min || !xyz_max) return fail(MDMC_ERR_INVALID, "null argument");
  int rc = count_impl(ctx, d_field, dims, iso, d_counts, stream, xyz_min, xyz_max);
  if (rc != MDMC_OK) return rc;
  return mdmc_emit_device(ctx, d_field, xyz_min, xyz_max, d_vertices, vertex_capacity, d_quads,
                          quad_capacity, d_counts, stream);
}

int mdmc_extract_device_noalloc(const float *d_field, const uint32_t dims[3], const float xyz_min[3],
                                const float xyz_max[3], float iso, float *d_vertices, uint64_t vertex_capacity,
                                int32_t *d_quads, uint64_t quad_capacity, mdmc_counts *d_counts, void *stream) {
  if (!dims || !xyz_min || !xyz_max || !d_counts) return fail(MDMC_ERR_INVALID, "null argument");
  int dev = 0;
  int rc = current_device(&dev);
  if (rc != MDMC_OK) return rc;
  if (dev < 0 || dev >= 64) return fail(MDMC_ERR_INVALID, "device index out of range");
  void *ptr = nullptr;
  int cus = 0;
  {
    std::lock_guard<std::mutex> lk(g_setup_mutex);
    if ((g_setup_mask >> dev) & 1ull) {
      ptr = g_noalloc_ptr[dev];
      cus = g_noalloc_cus[dev];
    }
  }
  if (!ptr) return fail(MDMC_ERR_NOT_SETUP, "mdmc_setup() not called on this device");
  Plan p;
  rc = plan_tiling(dims, &p, nullptr);
  if (rc != MDMC_OK) return rc;
  if (!p.empty && !d_field) return fail(MDMC_ERR_INVALID, "null field");
  const size_t need = p.empty ? 0 : scratch_layout(p).total;
  if (need > MDMC_NOALLOC_SCRATCH_BYTES)
    return fail(MDMC_ERR_CAPACITY, "grid needs " + std::to_string(need) +
                                       " bytes of scratch; the no-allocation entry has " +
                                       std::to_string(MDMC_NOALLOC_SCRATCH_BYTES));
  // a context on the stack over the fixed array: its scratch is external, so
  // nothing below allocates (the launch helpers only carve and launch)
  mdmc_ctx c;
  c.device = dev;
  c.num_cus = cus;
  c.scratch = ptr;
  c.scratch_bytes = MDMC_NOALLOC_SCRATCH_BYTES;
  c.scratch_external = true;
  rc = ensure_scratch(&c, p);
  if (rc != MDMC_OK) return rc;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  rc = launch_count(&c, p, d_field, iso, d_counts, st, xyz_min, xyz_max);
  if (rc != MDMC_OK) return rc;
  return launch_emit(&c, p, d_field, iso, xyz_min, xyz_max, d_vertices, vertex_capacity, d_quads, quad_capacity,
                     d_counts, st);
}

int mdmc_smooth_device(mdmc_ctx *ctx, float *d_vertices, uint64_t vertex_capacity, const int32_t *d_quads,
                       uint64_t quad_capacity, uint32_t num_smooth, mdmc_counts *d_counts, void *stream) {
  if (!ctx || !d_counts) return fail(MDMC_ERR_INVALID, "null argument");
  if (num_smooth == 0) return MDMC_OK;
  if (!ctx->emitted) return fail(MDMC_ERR_INVALID, "smoothing without a preceding emit on this context");
  const int grc = capture_guard(ctx, static_cast<hipStream_t>(stream));
  if (grc != MDMC_OK) return grc;
  return launch_smooth(ctx, d_vertices, vertex_capacity, const_cast<int32_t *>(d_quads), quad_capacity, num_smooth,
                       d_counts, static_cast<hipStream_t>(stream));
}

int mdmc_ctx_set_option(mdmc_ctx *ctx, int option, int value) {
  if (!ctx) return fail(MDMC_ERR_INVALID, "null ctx");
  if (option < 0 || option >= kNumOptions) return fail(MDMC_ERR_INVALID, "unknown option");
  ctx->opt[option] = value;
  return MDMC_OK;
}

static int check_batch_args(mdmc_ctx *ctx, const mdmc_volume *vols, uint32_t n, mdmc_counts *d_counts,
                            void *stream) {
  if (!ctx || !vols || !d_counts) return fail(MDMC_ERR_INVALID, "null argument");
  if (n == 0 || n > kMaxBatch) return fail(MDMC_ERR_INVALID, "batch size must be 1..1024");
  if (!is_setup_on(ctx->device)) return fail(MDMC_ERR_NOT_SETUP, "mdmc_setup() not called");
  return capture_guard(ctx, static_cast<hipStream_t>(stream));
}

int mdmc_scratch_bytes(const uint32_t dims[3], uint64_t *bytes) {
  if (!dims || !bytes) return fail(MDMC_ERR_INVALID, "null argument");
  Plan p;
  const int rc = plan_tiling(dims, &p, nullptr);  // the options do not change the layout
  if (rc != MDMC_OK) return rc;
  *bytes = p.empty ? 0 : scratch_layout(p).total;
  return MDMC_OK;
}

int mdmc_batch_scratch_bytes(const mdmc_volume *vols, uint32_t n, uint64_t *bytes) {
  if (!vols || !bytes) return fail(MDMC_ERR_INVALID, "null argument");
  if (n == 0 || n > kMaxBatch) return fail(MDMC_ERR_INVALID, "batch size must be 1..1024");
  Plan agg{};
  uint32_t W4 = 1, KP = 1;
  int rc = batch_layout(vols, n, false, &agg, &W4, &KP);
  if (rc != MDMC_OK) return rc;
  *bytes = agg.empty ? 0 : scratch_layout(agg).total;
  return MDMC_OK;
}

int mdmc_ctx_set_scratch(mdmc_ctx *ctx, void *d_scratch, uint64_t bytes) {
  if (!ctx) return fail(MDMC_ERR_INVALID, "null ctx");
  if (ctx->captured) return fail(MDMC_ERR_INVALID, "context belongs to a captured graph whose nodes use its scratch");
  if (ctx->scratch && !ctx->scratch_external) HIP_TRY(hipFree(ctx->scratch));
  ctx->scratch = d_scratch;
  ctx->scratch_bytes = d_scratch ? bytes : 0;
  ctx->scratch_external = d_scratch != nullptr;
  ctx->counted = false;
  ctx->batch_counted = false;
  return MDMC_OK;
}

int mdmc_batch_count_device(mdmc_ctx *ctx, const mdmc_volume *vols, uint32_t n, mdmc_counts *d_counts,
                            void *stream) {
  int rc = check_batch_args(ctx, vols, n, d_counts, stream);
  if (rc != MDMC_OK) return rc;
  return batch_count(ctx, vols, n, d_counts, static_cast<hipStream_t>(stream));
}

int mdmc_batch_emit_device(mdmc_ctx *ctx, const mdmc_volume *vols, uint32_t n, mdmc_counts *d_counts,
                           void *stream) {
  int rc = check_batch_args(ctx, vols, n, d_counts, stream);
  if (rc != MDMC_OK) return rc;
  return batch_emit(ctx, vols, n, d_counts, static_cast<hipStream_t>(stream), true);
}

int mdmc_batch_extract_device(mdmc_ctx *ctx, const mdmc_volume *vols, uint32_t n, mdmc_counts *d_counts,
                              void *stream) {
  int rc = check_batch_args(ctx, vols, n, d_counts, stream);
  if (rc != MDMC_OK) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  rc = batch_count(ctx, vols, n, d_counts, st);
  if (rc != MDMC_OK) return rc;
  return batch_emit(ctx, vols, n, d_counts, st, false);
}

int mdmc_cells_device(mdmc_ctx *ctx, uint32_t *d_cell_index, uint8_t *d_cfg, uint8_t *d_flags,
                      uint32_t *d_vertex_begin, uint64_t capacity, void *stream) {
  if (!ctx || !d_cell_index || !d_cfg || !d_flags || !d_vertex_begin)
    return fail(MDMC_ERR_INVALID, "null argument");
  if (!ctx->counted) return fail(MDMC_ERR_INVALID, "no counted extraction on this context");
  const int grc = capture_guard(ctx, static_cast<hipStream_t>(stream));
  if (grc != MDMC_OK) return grc;
  const Plan &p = ctx->plan;
  if (p.empty) return MDMC_OK;
  CellsArgs ca{};
  ca.XT = p.XT;
  ca.cap = p.cap;
  ca.cx = p.cx;
  ca.nsegs = (uint32_t)p.nsegs;
  ca.seg_counts = ctx->s.seg_counts;
  ca.slots = ctx->s.slots;
  ca.seg_vq = ctx->s.seg_vq;
  ca.ne = ctx->s.ne;
  ca.totals = ctx->s.totals;
  ca.cell_index = d_cell_index;
  ca.cfg = d_cfg;
  ca.flags = d_flags;
  ca.vbegin = d_vertex_begin;
  ca.capacity = capacity;
  hipLaunchKernelGGL(k_cells, dim3(std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)p.nsegs, (uint32_t)ctx->num_cus * 8))), dim3(256), 0, static_cast<hipStream_t>(stream), ca);
  HIP_TRY(hipGetLastError());
  return MDMC_OK;
}

int mdmc_run(const float *field, const uint32_t dims[3], const float xyz_min[3],
             const float xyz_max[3], float iso, uint32_t num_smooth, mdmc_mesh *out) {
  if (!out || !dims || !xyz_min || !xyz_max) return fail(MDMC_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> entry(g_host_entry_mutex);
  std::memset(out, 0, sizeof(*out));
  int dev = 0;
  int rc = current_device(&dev);
  if (rc != MDMC_OK) return rc;
  if (!is_setup_on(dev)) return fail(MDMC_ERR_NOT_SETUP, "mdmc_setup() not called");
  if (!g_default_ctx) {
    rc = mdmc_ctx_create(&g_default_ctx);
    if (rc != MDMC_OK) return rc;
  }
  mdmc_ctx *ctx = g_default_ctx;
  Plan p;
  rc = plan_tiling(dims, &p, ctx->opt);
  if (rc != MDMC_OK) return rc;
  if (p.empty) return MDMC_OK;
  if (!field) return fail(MDMC_ERR_INVALID, "null field");
  const size_t npts = (size_t)p.X * p.Y * p.Z;
  // device buffers of the host entry stay with the default context and grow
  // as needed; the field and the mesh cross PCIe through pinned chunks
  auto grow = [&](void **buf, size_t *have, size_t want) -> int {
    if (want <= *have) return MDMC_OK;
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *have = 0;
    if (hipMalloc(buf, want) != hipSuccess) {
      (void)hipGetLastError();
      return fail(MDMC_ERR_NOMEM, "device buffer allocation failed");
    }
    *have = want;
    return MDMC_OK;
  };
  rc = grow(&ctx->run_field, &ctx->run_field_bytes, npts * sizeof(float));
  if (rc == MDMC_OK) rc = grow(&ctx->run_counts, &ctx->run_counts_bytes, sizeof(mdmc_counts));
  if (rc != MDMC_OK) return rc;
  float *d_field = static_cast<float *>(ctx->run_field);
  mdmc_counts *d_counts = static_cast<mdmc_counts *>(ctx->run_counts);
  rc = staged_copy(const_cast<float *>(field), d_field, npts * sizeof(float), 0);
  if (rc != MDMC_OK) return fail(rc, "field upload failed");
  rc = mdmc_count_device(ctx, d_field, dims, iso, d_counts, nullptr);
  if (rc != MDMC_OK) return rc;
  mdmc_counts h{};
  HIP_TRY(hipMemcpy(&h, d_counts, sizeof(h), hipMemcpyDeviceToHost));
  if (h.status != 0) return fail(MDMC_ERR_INTERNAL, "device status after count");
  const size_t vbytes = (size_t)h.num_vertices * 3 * sizeof(float), qbytes = (size_t)h.num_quads * 4 * sizeof(int32_t);
  rc = grow(&ctx->run_v, &ctx->run_v_bytes, std::max<size_t>(vbytes, 256));
  if (rc == MDMC_OK) rc = grow(&ctx->run_q, &ctx->run_q_bytes, std::max<size_t>(qbytes, 256));
  if (rc != MDMC_OK) return rc;
  float *d_v = static_cast<float *>(ctx->run_v);
  int32_t *d_q = static_cast<int32_t *>(ctx->run_q);
  rc = mdmc_emit_device(ctx, d_field, xyz_min, xyz_max, d_v, h.num_vertices, d_q, h.num_quads, d_counts, nullptr);
  if (rc != MDMC_OK) return rc;
  rc = mdmc_smooth_device(ctx, d_v, h.num_vertices, d_q, h.num_quads, num_smooth, d_counts, nullptr);
  if (rc != MDMC_OK) return rc;
  mdmc_counts h2{};
  HIP_TRY(hipMemcpy(&h2, d_counts, sizeof(h2), hipMemcpyDeviceToHost));
  if (h2.status != 0) return fail(MDMC_ERR_INTERNAL, "device status after emit");
  out->num_active = h.num_active;
  out->num_vertices = h.num_vertices;
  out->num_quads = h.num_quads;
  out->vertices = static_cast<float *>(std::malloc(std::max<size_t>(1, vbytes)));
  out->quads = static_cast<int32_t *>(std::malloc(std::max<size_t>(1, qbytes)));
  if (!out->vertices || !out->quads) {
    mdmc_mesh_free(out);
    return fail(MDMC_ERR_NOMEM, "host mesh allocation failed");
  }
  rc = staged_copy(out->vertices, d_v, vbytes, 1);
  if (rc == MDMC_OK) rc = staged_copy(out->quads, d_q, qbytes, 1);
  if (rc != MDMC_OK) {
    mdmc_mesh_free(out);
    return fail(rc, "mesh download failed");
  }
  return MDMC_OK;
}

int mdmc_run_device(const float *d_field, const uint32_t dims[3], const float xyz_min[3],
                    const float xyz_max[3], float iso, uint32_t num_smooth, void *stream,
                    mdmc_device_mesh *out) {
  if (!out || !dims || !xyz_min || !xyz_max) return fail(MDMC_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> entry(g_host_entry_mutex);
  std::memset(out, 0, sizeof(*out));
  int dev = 0;
  int rc = current_device(&dev);
  if (rc != MDMC_OK) return rc;
  if (!is_setup_on(dev)) return fail(MDMC_ERR_NOT_SETUP, "mdmc_setup() not called");
  if (!g_default_ctx) {
    rc = mdmc_ctx_create(&g_default_ctx);
    if (rc != MDMC_OK) return rc;
  }
  mdmc_ctx *ctx = g_default_ctx;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  if (!ctx->run_counts) {
    if (hipMalloc(&ctx->run_counts, sizeof(mdmc_counts)) != hipSuccess) {
      (void)hipGetLastError();
      ctx->run_counts = nullptr;
      return fail(MDMC_ERR_NOMEM, "device buffer allocation failed");
    }
    ctx->run_counts_bytes = sizeof(mdmc_counts);
  }
  mdmc_counts *d_counts = static_cast<mdmc_counts *>(ctx->run_counts);
  rc = mdmc_count_device(ctx, d_field, dims, iso, d_counts, stream);
  if (rc != MDMC_OK) return rc;
  if (ctx->plan.empty) return MDMC_OK;
  mdmc_counts h{};
  HIP_TRY(hipMemcpyAsync(&h, d_counts, sizeof(h), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (h.status != 0) return fail(MDMC_ERR_INTERNAL, "device status after count");
  if (h.num_vertices == 0 && h.num_quads == 0) {
    out->num_active = h.num_active;
    return MDMC_OK;
  }
  const size_t vbytes = (size_t)h.num_vertices * 3 * sizeof(float), qbytes = (size_t)h.num_quads * 4 * sizeof(int32_t);
  if (vbytes && hipMalloc(reinterpret_cast<void **>(&out->vertices), vbytes) != hipSuccess) {
    (void)hipGetLastError();
    out->vertices = nullptr;
    return fail(MDMC_ERR_NOMEM, "device mesh allocation failed");
  }
  if (qbytes && hipMalloc(reinterpret_cast<void **>(&out->quads), qbytes) != hipSuccess) {
    (void)hipGetLastError();
    out->quads = nullptr;
    mdmc_device_mesh_free(out);
    return fail(MDMC_ERR_NOMEM, "device mesh allocation failed");
  }
  auto bail = [&](int code) {
    mdmc_device_mesh_free(out);
    return code;
  };
  rc = mdmc_emit_device(ctx, d_field, xyz_min, xyz_max, out->vertices, h.num_vertices, out->quads, h.num_quads,
                        d_counts, stream);
  if (rc != MDMC_OK) return bail(rc);
  rc = mdmc_smooth_device(ctx, out->vertices, h.num_vertices, out->quads, h.num_quads, num_smooth, d_counts, stream);
  if (rc != MDMC_OK) return bail(rc);
  mdmc_counts h2{};
  if (hipMemcpyAsync(&h2, d_counts, sizeof(h2), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    (void)hipGetLastError();
    return bail(fail(MDMC_ERR_HIP, "stream error after emit"));
  }
  if (h2.status != 0) return bail(fail(MDMC_ERR_INTERNAL, "device status after emit"));
  out->num_active = h.num_active;
  out->num_vertices = h.num_vertices;
  out->num_quads = h.num_quads;
  return MDMC_OK;
}

void mdmc_device_mesh_free(mdmc_device_mesh *mesh) {
  if (!mesh) return;
  if (mesh->vertices) (void)hipFree(mesh->vertices);
  if (mesh->quads) (void)hipFree(mesh->quads);
  std::memset(mesh, 0, sizeof(*mesh));
}

// Point planes [c0-1, c0+S+2] (clipped) of a slab of S cell planes, and the
// largest S whose field slab and scratch fit the budget.
static int ooc_slab_planes(const uint32_t dims[3], uint64_t budget, const int *opt, uint32_t *S_out) {
  const uint32_t X = dims[0], Y = dims[1], Z = dims[2];
  if (X < 2 || Y < 2 || Z < 2) {
    *S_out = 0;
    return MDMC_OK;
  }
  const uint32_t cz = Z - 1;
  auto need = [&](uint32_t S) -> uint64_t {
    const uint32_t npl = S < cz ? std::min(S + 4, Z) : Z;  // a middle slab is the largest
    const uint32_t sd[3] = {X, Y, npl};
    Plan p;
    if (plan_tiling(sd, &p, opt) != MDMC_OK) return ~0ull;
    // two field slabs when there is more than one (the next one is uploaded
    // while the current one is extracted)
    return (S < cz ? 2ull : 1ull) * 4ull * X * Y * npl + scratch_layout(p).total;
  };
  if (need(1) > budget) return fail(MDMC_ERR_NOMEM, "budget below one slab of one cell plane");
  uint32_t lo = 1, hi = cz;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo + 1) / 2;
    if (need(mid) <= budget) lo = mid; else hi = mid - 1;
  }
  *S_out = lo;
  return MDMC_OK;
}

int mdmc_out_of_core_slab_planes(const uint32_t dims[3], uint64_t device_budget, uint32_t *slab_planes) {
  if (!dims || !slab_planes) return fail(MDMC_ERR_INVALID, "null argument");
  return ooc_slab_planes(dims, device_budget, nullptr, slab_planes);
}

// Out-of-core: the grid goes through the device in z-slabs of S cell
// planes.  Slab [c0, c1) is classified with point planes [c0-1, c1+2]
// (clipped to the grid): one plane below and two above give every cell of
// the slab, and every ring cell in plane c1, its 2B/3B neighbours, so their
// configurations, variants and vertex counts are those of the whole grid.
// In the slab's own numbering, the vertices of planes >= c0 are the whole
// grid's numbering shifted by a constant (the slab starts where the previous
// slabs' vertices end), so the emit writes only planes [c0, c1) and adds that
// shift to every quad entry.  Coordinates use the whole grid's divisors.
int mdmc_run_out_of_core(const float *field, const uint32_t dims[3], const float xyz_min[3],
                         const float xyz_max[3], float iso, uint64_t device_budget, mdmc_mesh *out) {
  if (!out || !dims || !xyz_min || !xyz_max) return fail(MDMC_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> entry(g_host_entry_mutex);
  std::memset(out, 0, sizeof(*out));
  int dev = 0;
  int rc = current_device(&dev);
  if (rc != MDMC_OK) return rc;
  if (!is_setup_on(dev)) return fail(MDMC_ERR_NOT_SETUP, "mdmc_setup() not called");
  // a context of its own, so that no buffer of the in-core entries counts
  // against the budget, with its scratch sized exactly
  if (!g_ooc_ctx) {
    rc = mdmc_ctx_create(&g_ooc_ctx);
    if (rc != MDMC_OK) return rc;
    g_ooc_ctx->exact_scratch = true;
  }
  mdmc_ctx *ctx = g_ooc_ctx;
  const uint32_t X = dims[0], Y = dims[1], Z = dims[2];
  if (X < 2 || Y < 2 || Z < 2) return MDMC_OK;
  if (!field) return fail(MDMC_ERR_INVALID, "null field");
  const uint32_t cz = Z - 1;
  const size_t plane = (size_t)X * Y;
  uint32_t S = 0;
  rc = ooc_slab_planes(dims, device_budget, ctx->opt, &S);
  if (rc != MDMC_OK) return rc;
  {
    // the scratch of the largest slab, exactly: every slab's layout fits in
    // it, and a larger scratch left by an earlier call is given back first
    const uint32_t sd[3] = {X, Y, S < cz ? std::min(S + 4, Z) : Z};
    Plan pmax;
    rc = plan_tiling(sd, &pmax, ctx->opt);
    if (rc != MDMC_OK) return rc;
    const size_t want = scratch_layout(pmax).total;
    if (ctx->scratch && ctx->scratch_bytes != want) {
      (void)hipFree(ctx->scratch);
      ctx->scratch = nullptr;
      ctx->scratch_bytes = 0;
    }
    rc = ensure_scratch(ctx, pmax);
    if (rc != MDMC_OK) return rc;
  }
  // point planes of a slab of cell planes [c0, c0 + n)
  auto slab_planes = [&](uint32_t c0, uint32_t n, uint32_t *lo) {
    const uint32_t l = c0 > 0 ? c0 - 1 : 0, hi = std::min(c0 + n + 2, Z - 1);
    *lo = l;
    return hi - l + 1;
  };
  // the device buffers stay with the default context, so that repeated calls
  // on grids of one size allocate nothing
  void **const d_fbuf = ctx->ooc_field;
  void *&d_v = ctx->ooc_v, *&d_q = ctx->ooc_q;
  float *hv = nullptr;
  int32_t *hq = nullptr;
  uint64_t hv_cap = 0, hq_cap = 0, G = 0, GQ = 0, A = 0;
  std::thread uploader;
  int upload_rc = MDMC_OK;
  auto release = [&]() {
    if (uploader.joinable()) uploader.join();
  };
  auto bail = [&](int code, const char *msg) {
    release();
    std::free(hv);
    std::free(hq);
    return fail(code, msg);
  };
  // exact: reallocate unless the buffer has exactly the wanted size (a field
  // slab never holds more than the budget allowed for this call)
  auto grow_dev = [&](void **b, size_t *have, size_t want, bool exact) -> bool {
    if (exact ? want == *have : want <= *have) return true;
    if (*b) (void)hipFree(*b);
    *b = nullptr;
    *have = 0;
    if (hipMalloc(b, want) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    *have = want;
    return true;
  };
  auto grow_host = [&](void **b, uint64_t *cap, uint64_t want, size_t elem) -> bool {
    if (want <= *cap) return true;
    const uint64_t n = std::max<uint64_t>(want, *cap * 2);
    void *nb = std::realloc(*b, std::max<size_t>(1, (size_t)n * elem));
    if (!nb) return false;
    *b = nb;
    *cap = n;
    return true;
  };
  // counts (4 words) then the slab record of k_slab_info (9 words)
  if (!ctx->ooc_counts && hipMalloc(&ctx->ooc_counts, 16 * sizeof(uint32_t)) != hipSuccess) {
    (void)hipGetLastError();
    ctx->ooc_counts = nullptr;
    return bail(MDMC_ERR_NOMEM, "counts allocation failed");
  }
  mdmc_counts *d_counts = static_cast<mdmc_counts *>(ctx->ooc_counts);
  uint32_t *d_info = static_cast<uint32_t *>(ctx->ooc_counts) + 4;
  // slab i computes from field buffer i % 2 while a host thread uploads slab
  // i + 1 into the other buffer
  const uint32_t nslabs = (cz + S - 1) / S;
  const size_t max_npl = std::min<size_t>((size_t)S + 4, Z);
  for (int b2 = 0; b2 < 2; ++b2) {
    if (b2 < (int)nslabs) {
      if (!grow_dev(&d_fbuf[b2], &ctx->ooc_field_bytes[b2], 4 * plane * max_npl, true))
        return bail(MDMC_ERR_NOMEM, "field slab allocation failed");
    } else if (d_fbuf[b2]) {  // a single slab: no second buffer
      (void)hipFree(d_fbuf[b2]);
      d_fbuf[b2] = nullptr;
      ctx->ooc_field_bytes[b2] = 0;
    }
  }
  auto upload = [&](uint32_t i) {
    uint32_t lo;
    const uint32_t n = slab_planes(i * S, std::min(S, cz - i * S), &lo);
    return staged_copy(const_cast<float *>(field) + plane * lo, d_fbuf[i % 2], 4 * plane * n, 0);
  };
  rc = upload(0);
  if (rc != MDMC_OK) return bail(rc, "field slab upload failed");
  for (uint32_t i = 0; i < nslabs; ++i) {
    const uint32_t c0 = i * S, c1 = std::min(c0 + S, cz);
    uint32_t plo;
    const uint32_t npl = slab_planes(c0, c1 - c0, &plo);
    const uint32_t h = c0 - plo;
    if (uploader.joinable()) uploader.join();
    if (upload_rc != MDMC_OK) return bail(upload_rc, "field slab upload failed");
    if (i + 1 < nslabs)
      uploader = std::thread([&, i, dev] {
        (void)hipSetDevice(dev);
        upload_rc = upload(i + 1);
      });
    void *d_field = d_fbuf[i % 2];
    const uint32_t sd[3] = {X, Y, npl};
    Plan p;
    rc = plan_tiling(sd, &p, ctx->opt);
    if (rc != MDMC_OK) return bail(rc, "slab plan failed");
    rc = ensure_scratch(ctx, p);
    if (rc != MDMC_OK) return bail(rc, "slab scratch failed");
    const float *df = static_cast<const float *>(d_field);
    rc = launch_count(ctx, p, df, iso, d_counts, nullptr);
    if (rc != MDMC_OK) return bail(rc, "slab count failed");
    ctx->plan = p;
    ctx->counted = true;
    ctx->counted_field = df;
    ctx->iso = iso;
    // the count status, the vertex / quad bases of the first segment of
    // planes h and h + (c1 - c0), and the slab's active cells: one copy
    const uint64_t sp = (uint64_t)p.XT * p.cy;
    const uint64_t s_lo = h * sp, s_hi = (uint64_t)(h + (c1 - c0)) * sp;
    hipLaunchKernelGGL(k_slab_info, dim3(1), dim3(64), 0, nullptr, ctx->s.seg_vq, ctx->s.ne, ctx->s.totals,
                       (uint32_t)p.nsegs, reinterpret_cast<const uint32_t *>(d_counts), (uint32_t)s_lo,
                       (uint32_t)s_hi, d_info);
    uint32_t info[9] = {};
    if (hipMemcpy(info, d_info, sizeof(info), hipMemcpyDeviceToHost) != hipSuccess)
      return bail(MDMC_ERR_HIP, "slab record");
    if (info[3]) return bail(MDMC_ERR_INTERNAL, "slab count status");
    const uint2 b_lo = make_uint2(info[4], info[5]), b_hi = make_uint2(info[6], info[7]);
    A += info[8];
    mdmc_counts hc{};
    const uint32_t Vs = b_hi.x - b_lo.x, Qs = b_hi.y - b_lo.y;
    if (!grow_dev(&d_v, &ctx->ooc_v_bytes, std::max<size_t>(256, 12ull * Vs), false) ||
        !grow_dev(&d_q, &ctx->ooc_q_bytes, std::max<size_t>(256, 16ull * Qs), false))
      return bail(MDMC_ERR_NOMEM, "slab output allocation failed");
    // emit planes [h, h + c1 - c0) with global coordinates and numbering
    const Scratch &sc = ctx->s;
    EmitArgs ea{};
    ea.v.g = FieldDesc{df, p.X, p.Y, p.Z, p.cx, p.cy, p.cz, iso};
    CoordArgs co = coord_args(p, sc, xyz_min, xyz_max);
    co.cz = cz;
    co.z_off = plo;
    ea.v.xt_c = sc.coords;
    ea.v.yt_c = sc.coords + p.X;
    ea.v.zt_c = sc.coords + p.X + p.Y;
    ea.v.XT = p.XT;
    ea.v.vbase = b_lo.x;
    ea.v.qbase = b_lo.y;
    ea.v.kmin = h;
    ea.v.kmax = h + (c1 - c0);
    ea.v.vshift = (uint32_t)G;
    ea.cap = p.cap;
    ea.nsegs = (uint32_t)p.nsegs;
    ea.max_chunks = p.max_chunks;
    ea.seg_counts = sc.seg_counts;
    ea.seg_vq = sc.seg_vq;
    ea.slots = sc.slots;
    ea.gmap = sc.gmap;
    ea.ne = sc.ne;
    ea.chunk_ne = sc.chunk_ne;
    ea.totals = sc.totals;
    ea.v.verts = static_cast<float *>(d_v);
    ea.v.vcap = Vs;
    ea.v.quads = static_cast<int32_t *>(d_q);
    ea.v.qcap = Qs;
    ea.v.status = reinterpret_cast<uint32_t *>(d_counts) + 3;
    hipLaunchKernelGGL(k_coords, dim3((uint32_t)std::min<uint64_t>((p.ncoord + 255) / 256, 1024)), dim3(256), 0,
                       nullptr, co);
    ea.emit_ctr = sc.ticket;  // zeroed by this slab's count
    hipLaunchKernelGGL(k_emit, dim3(emit_blocks(ctx, p.max_chunks)), dim3(kEmitThreads), 0, nullptr, ea);
    ctx->tickets_zero = false;
    if (hipGetLastError() != hipSuccess) return bail(MDMC_ERR_HIP, "slab emit launch");
    ctx->emitted = false;  // the slab mesh is not the context's mesh
    if (hipMemcpy(&hc, d_counts, sizeof(hc), hipMemcpyDeviceToHost) != hipSuccess || hc.status)
      return bail(MDMC_ERR_INTERNAL, "slab emit status");
    if (G + Vs > 0xFFFFFFFFull) return bail(MDMC_ERR_TOO_LARGE, "more than 2^32-1 vertices");
    if (!grow_host(reinterpret_cast<void **>(&hv), &hv_cap, G + Vs, 12) ||
        !grow_host(reinterpret_cast<void **>(&hq), &hq_cap, GQ + Qs, 16))
      return bail(MDMC_ERR_NOMEM, "host mesh allocation failed");
    rc = staged_copy(hv + 3 * G, d_v, 12ull * Vs, 1);
    if (rc == MDMC_OK) rc = staged_copy(hq + 4 * GQ, d_q, 16ull * Qs, 1);
    if (rc != MDMC_OK) return bail(rc, "slab mesh download failed");
    G += Vs;
    GQ += Qs;
  }
  release();
  if (upload_rc != MDMC_OK) {
    std::free(hv);
    std::free(hq);
    return fail(upload_rc, "field slab upload failed");
  }
  out->vertices = hv ? hv : static_cast<float *>(std::malloc(1));
  out->quads = hq ? hq : static_cast<int32_t *>(std::malloc(1));
  out->num_vertices = (uint32_t)G;
  out->num_quads = (uint32_t)GQ;
  out->num_active = (uint32_t)A;
  return MDMC_OK;
}

void mdmc_mesh_free(mdmc_mesh *mesh) {
  if (!mesh) return;
  std::free(mesh->vertices);
  std::free(mesh->quads);
  std::memset(mesh, 0, sizeof(*mesh));
}

// ---- synthetic fields ------------------------------------------------------

int mdmc_field_blobs_device(mdmc_ctx *ctx, float *d_out, const uint32_t dims[3], uint64_t seed,
                            uint32_t nblobs, double sigma_lo, double sigma_hi, double amp_lo,
                            double amp_hi, void *stream) {
  if (!ctx || !d_out || !dims) return fail(MDMC_ERR_INVALID, "null argument");
  if (!dims[0] || !dims[1] || !dims[2]) return MDMC_OK;
  mdmc_gen_spec g{};
  g.out = d_out;
  std::memcpy(g.dims, dims, sizeof(g.dims));
  g.kind = MDMC_FIELD_BLOBS;
  g.seed = seed;
  g.nblobs = nblobs;
  g.sigma_lo = sigma_lo;
  g.sigma_hi = sigma_hi;
  g.amp_lo = amp_lo;
  g.amp_hi = amp_hi;
  return mdmc_batch_generate_device(ctx, &g, 1, stream);
}

int mdmc_field_gyroid_device(mdmc_ctx *ctx, float *d_out, const uint32_t dims[3], double period,
                             uint64_t seed, double noise_amp, void *stream) {
  if (!ctx || !d_out || !dims) return fail(MDMC_ERR_INVALID, "null argument");
  if (!dims[0] || !dims[1] || !dims[2]) return MDMC_OK;
  mdmc_gen_spec g{};
  g.out = d_out;
  std::memcpy(g.dims, dims, sizeof(g.dims));
  g.kind = MDMC_FIELD_GYROID;
  g.seed = seed;
  g.period = period;
  g.noise_amp = noise_amp;
  return mdmc_batch_generate_device(ctx, &g, 1, stream);
}

int mdmc_batch_generate_device(mdmc_ctx *ctx, const mdmc_gen_spec *specs, uint32_t n, void *stream) {
  if (!ctx || !specs) return fail(MDMC_ERR_INVALID, "null argument");
  if (n == 0 || n > MDMC_MAX_BATCH) return fail(MDMC_ERR_INVALID, "batch size must be 1..1024");
  hipStream_t st = static_cast<hipStream_t>(stream);
  // layout: GenVol[n] then the doubles of every volume's tables
  size_t ndoubles = 0;
  for (uint32_t v = 0; v < n; ++v) {
    const mdmc_gen_spec &g = specs[v];
    if (g.kind == MDMC_FIELD_BLOBS)
      ndoubles += (size_t)blob_ldb(g.nblobs) * ((size_t)g.dims[0] + g.dims[1] + g.dims[2] + 1);
    else if (g.kind == MDMC_FIELD_GYROID)
      ndoubles += 2 * ((size_t)g.dims[0] + g.dims[1] + g.dims[2]);
    else
      return fail(MDMC_ERR_INVALID, "unknown field kind");
    if (!g.out && (uint64_t)g.dims[0] * g.dims[1] * g.dims[2]) return fail(MDMC_ERR_INVALID, "null output");
  }
  const size_t head = align256(n * sizeof(GenVol));
  const size_t bytes = head + ndoubles * sizeof(double);
  int grc = capture_guard(ctx, st);
  if (grc != MDMC_OK) return grc;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  HIP_TRY(hipStreamIsCapturing(st, &cs));
  if (cs == hipStreamCaptureStatusNone && ctx->ghost_ev_used) HIP_TRY(hipEventSynchronize(ctx->ghost_ev));
  if ((bytes > ctx->ghost_bytes || bytes > ctx->gdev_bytes) && cs != hipStreamCaptureStatusNone)
    return fail(MDMC_ERR_INVALID, "generator tables must be sized by a call on this context before stream capture");
  if (bytes > ctx->ghost_bytes) {
    if (ctx->ghost) (void)hipHostFree(ctx->ghost);
    ctx->ghost = nullptr;
    ctx->ghost_bytes = 0;
    const size_t want = std::max<size_t>(bytes + bytes / 4, 1u << 20);
    if (hipHostMalloc(&ctx->ghost, want) != hipSuccess) {
      (void)hipGetLastError();
      return fail(MDMC_ERR_NOMEM, "pinned staging allocation failed");
    }
    ctx->ghost_bytes = want;
  }
  if (bytes > ctx->gdev_bytes) {
    if (ctx->gdev) (void)hipFree(ctx->gdev);
    ctx->gdev = nullptr;
    ctx->gdev_bytes = 0;
    if (hipMalloc(&ctx->gdev, bytes + bytes / 4) != hipSuccess) {
      (void)hipGetLastError();
      return fail(MDMC_ERR_NOMEM, "generator table allocation failed");
    }
    ctx->gdev_bytes = bytes + bytes / 4;
  }
  if (!ctx->ghost_ev) HIP_TRY(hipEventCreateWithFlags(&ctx->ghost_ev, hipEventDisableTiming));
  GenVol *hv = static_cast<GenVol *>(ctx->ghost);
  double *ht = reinterpret_cast<double *>(static_cast<char *>(ctx->ghost) + head);
  const double *dt = reinterpret_cast<const double *>(static_cast<char *>(ctx->gdev) + head);
  size_t off = 0;
  uint32_t blocks = 0;
  for (uint32_t v = 0; v < n; ++v) {
    const mdmc_gen_spec &g = specs[v];
    GenVol &G = hv[v];
    std::memset(&G, 0, sizeof(G));
    const uint32_t X = g.dims[0], Y = g.dims[1], Z = g.dims[2];
    G.out = g.out;
    G.X = X;
    G.Y = Y;
    G.Z = Z;
    G.kind = (uint32_t)g.kind;
    G.seed = g.seed;
    const uint64_t npts = (uint64_t)X * Y * Z;
    G.block_off = blocks;
    uint64_t nb64 = 0;
    // planes per task: the most that still gives the chip 8 waves per SIMD (a
    // small field in tasks of 16 planes is bound by the planes' latency)
    const bool blobs = g.kind == MDMC_FIELD_BLOBS;
    const uint64_t min_tasks = (uint64_t)ctx->num_cus * 32;
    uint32_t per = kGenPer;
    while (npts && per > 1 && gen_tasks(X, Y, Z, per) < min_tasks) per >>= 1;
    G.per = per;
    if (npts && blobs && g.nblobs > kBlobRegs)
      nb64 = std::min<uint64_t>((npts + 255) / 256, 4096);
    else if (npts)
      nb64 = (gen_tasks(X, Y, Z, per) + kGenWaves - 1) / kGenWaves;
    if (nb64 + blocks > 0x7FFFFFFFull) return fail(MDMC_ERR_TOO_LARGE, "fields too large for the generator grid");
    G.nblocks = (uint32_t)nb64;
    blocks += G.nblocks;
    double *h = ht + off;
    const double *d = dt + off;
    if (g.kind == MDMC_FIELD_BLOBS) {
      // ex [ldb][X]; ey [Y][ldb] and ez [Z][ldb] transposed; amp [ldb]; zero padded
      const uint32_t nb = g.nblobs, ldb = blob_ldb(nb);
      G.nblobs = nb;
      const size_t tab = (size_t)ldb * ((size_t)X + Y + Z + 1);
      double *ex = h, *ey = ex + (size_t)ldb * X, *ez = ey + (size_t)ldb * Y, *amp = ez + (size_t)ldb * Z;
      std::memset(h, 0, sizeof(double) * tab);
      std::vector<double> line(std::max(Y, Z));
      for (uint32_t b = 0; b < nb; ++b) {
        mdmc_blob bp = mdmc_blob_params(g.seed, b, g.dims, g.sigma_lo, g.sigma_hi, g.amp_lo, g.amp_hi);
        mdmc_gauss_1d(ex + (size_t)b * X, X, bp.cx, bp.sigma);
        mdmc_gauss_1d(line.data(), Y, bp.cy, bp.sigma);
        for (uint32_t y = 0; y < Y; ++y) ey[(size_t)y * ldb + b] = line[y];
        mdmc_gauss_1d(line.data(), Z, bp.cz, bp.sigma);
        for (uint32_t z = 0; z < Z; ++z) ez[(size_t)z * ldb + b] = line[z];
        amp[b] = bp.amp;
      }
      G.t[0] = d;
      G.t[1] = d + (size_t)ldb * X;
      G.t[2] = d + (size_t)ldb * ((size_t)X + Y);
      G.t[3] = d + (size_t)ldb * ((size_t)X + Y + Z);
      off += tab;
    } else {
      double *sx = h, *cx = sx + X, *sy = cx + X, *cy = sy + Y, *sz = cy + Y, *cz = sz + Z;
      mdmc_trig_1d(sx, cx, X, g.period);
      mdmc_trig_1d(sy, cy, Y, g.period);
      mdmc_trig_1d(sz, cz, Z, g.period);
      G.t[0] = d;
      G.t[1] = d + X;
      G.t[2] = d + 2 * (size_t)X;
      G.t[3] = d + 2 * (size_t)X + Y;
      G.t[4] = d + 2 * (size_t)X + 2 * (size_t)Y;
      G.t[5] = d + 2 * (size_t)X + 2 * (size_t)Y + Z;
      G.amp = g.noise_amp;
      off += 2 * ((size_t)X + Y + Z);
    }
  }
  if (blocks == 0) return MDMC_OK;
  HIP_TRY(hipMemcpyAsync(ctx->gdev, ctx->ghost, bytes, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_field_batch, dim3(blocks), dim3(kGenThreads), 0, st, static_cast<const GenVol *>(ctx->gdev),
                     n);
  HIP_TRY(hipGetLastError());
  // the next call rewrites both table copies once this kernel has read them
  if (cs == hipStreamCaptureStatusNone) {
    HIP_TRY(hipEventRecord(ctx->ghost_ev, st));
    ctx->ghost_ev_used = true;
  }
  return MDMC_OK;
}

}  // extern "C"
