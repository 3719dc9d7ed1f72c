// Request loop for serving extractions (mdmc_server_* in include/mdmc.h).
//
// Memory: one device allocation of the budget, carved by a first-fit pool.
// A request's footprint is its field (when the server generates it), its
// share of the batch scratch and the largest mesh its dims allow (4 iso
// vertices and 3 edge points per cell, one quad per owned edge whose ring
// lies in the grid).  A request is admitted when its field and worst-case
// output regions can be carved from the pool; its batch then runs generate
// -> count -> emit as one stream-ordered sequence (no host round trip between
// the count and the emit), and once it has finished each output region is
// cut down to the exact size the count gave.  So an admitted request always
// has room for its output: none is rejected after its count.  A request whose
// footprint exceeds the whole budget is rejected when it is submitted.
//
// Progress: a worker thread owned by the server schedules (every phase
// change is noticed within microseconds, whatever the caller is doing);
// submit, poll, release and the rest take the server's mutex and wake it.
//
// Scheduling: `num_streams` lanes, each with its own stream, extraction
// context and launcher thread.  For an idle lane the worker takes the queued
// requests in arrival order (up to max_batch, as long as their regions can be
// carved) and hands the batch to the lane's thread, which generates the
// fields, counts and emits them and queues the copy of the counts back (the
// generator tables and the launches cost host time per request, so the lanes
// do this in parallel and outside the server's mutex); the worker that sees
// that sequence finish trims the outputs, frees fields and scratch and hands
// out the results.  A request of at least `solo_cells` cells runs alone on a
// lane through the single-volume extraction instead; batches take the
// smaller requests queued behind it, and it starts on the next free lane
// once it heads the queue.  Lanes run independently, so a batch started for
// late arrivals overlaps with the batches already on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/mdmc.h"
#include "mdmc_pool.h"

namespace {

constexpr uint64_t kNoOffset = ~0ull;
constexpr uint64_t kPoolAlign = 256;
// Requests of at least this many cells run alone through the single-volume
// kernels: on one large volume they beat the batch kernels (their classify
// tiles and scan are shaped for one grid; a 384^3 gyroid takes 507 us alone
// and 587 us as a batch of one), while small requests share launches.
constexpr uint64_t kSoloCellsDefault = 200ull * 200 * 200;

struct Req {
  uint64_t id = 0;
  mdmc_request spec{};
  int status = MDMC_OK;
  int state = 0;  // 0 queued, 1 in a batch, 2 finished, 3 returned to the caller
  double t_submit = 0, t_start = 0, t_done = 0;
  uint64_t field_off = kNoOffset, vert_off = kNoOffset, quad_off = kNoOffset;
  uint64_t vmax = 0, qmax = 0;  // worst-case mesh of the dims
  uint64_t footprint = 0, scratch_est = 0;
  bool solo = false;  // runs alone through the single-volume kernels
  const float *field = nullptr;
  float *verts = nullptr;
  int32_t *quads = nullptr;
  mdmc_counts counts{};
};

struct Batch {
  std::vector<Req *> reqs;
  std::vector<mdmc_volume> vols;
  uint64_t scratch_off = kNoOffset, scratch_bytes = 0;
  void *base_ptr = nullptr;  // the scratch block's device address
  bool solo = false;         // one request, through the single-volume extraction
};

// A lane: a stream and an extraction context, and a thread of its own that
// issues the work of the batches handed to it (building generator tables and
// launching kernels cost host time per request; lanes do it in parallel).
struct Lane {
  hipStream_t stream = nullptr;
  mdmc_ctx *ctx = nullptr;
  mdmc_counts *d_counts = nullptr;  // device, MDMC_MAX_BATCH entries
  mdmc_counts *h_counts = nullptr;  // pinned
  hipEvent_t ev = nullptr;          // end of the batch's sequence
  Batch *batch = nullptr;
  // phase (guarded by the server's mutex): 0 idle, 1 handed to the lane's
  // thread, 2 launched (launch_rc says whether the launches went through)
  int phase = 0;
  int launch_rc = MDMC_OK;
  std::thread thread;
  std::condition_variable cv;  // with the server's mutex: a batch was handed over, or stop
  std::mutex ctx_mu;           // held by the lane's thread while it uses ctx and stream
};

// MDMC_SERVER_TRACE=1: one line per scheduling event on stderr.
bool trace_on() {
  static const bool on = [] {
    const char *e = std::getenv("MDMC_SERVER_TRACE");
    return e && e[0] && e[0] != '0';
  }();
  return on;
}
#define TRACE(...)                              \
  do {                                          \
    if (trace_on()) {                           \
      std::fprintf(stderr, "[mdmc_server] " __VA_ARGS__); \
      std::fputc('\n', stderr);                \
    }                                           \
  } while (0)

uint64_t field_bytes(const mdmc_request &r) {
  return 4ull * r.dims[0] * r.dims[1] * r.dims[2];
}

uint64_t round_up(uint64_t v) { return (v + kPoolAlign - 1) / kPoolAlign * kPoolAlign; }

// The largest mesh a grid of these dims can have: 4 iso vertices (the
// configurations 0x5A and 0xA5) and 3 edge points per cell, and a quad per
// owned edge (6, 9, 10) whose four-cell ring lies in the grid
// (dmc.cu:1382-1406, 1696-1708).  A checkerboard field reaches both.
void worst_case_mesh(const uint32_t d[3], uint64_t *V, uint64_t *Q) {
  if (d[0] < 2 || d[1] < 2 || d[2] < 2) {
    *V = *Q = 0;
    return;
  }
  const uint64_t cx = d[0] - 1, cy = d[1] - 1, cz = d[2] - 1;
  *V = 7 * cx * cy * cz;
  *Q = (cx - 1) * (cy - 1) * cz + (cx - 1) * cy * (cz - 1) + cx * (cy - 1) * (cz - 1);
}

}  // namespace

struct mdmc_server {
  mdmc_server_config cfg{};
  char *base = nullptr;
  mdmc::Pool pool;
  std::deque<Lane> lanes;  // (a deque: lanes hold a mutex and never move)
  std::deque<Req *> queue;
  std::unordered_map<uint64_t, Req *> reqs;
  std::deque<Req *> finished;  // not yet returned by poll
  uint64_t next_id = 1;
  uint64_t solo_cells = kSoloCellsDefault;  // 0: no request runs alone
  uint64_t submitted = 0, completed = 0, rejected = 0, batches = 0, peak = 0;
  double busy = 0, busy_since = -1;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  // worker thread and its synchronisation (everything above is guarded by mu)
  std::mutex mu;
  std::condition_variable cv_work;  // submit / release / stop
  std::condition_variable cv_done;  // a result finished or the lanes went idle
  std::thread worker;
  std::atomic<uint64_t> epoch{0};   // bumped (under mu) by calls that may unblock work
  std::atomic<bool> stop{false};
  bool idle = false;                // the worker waits for a call (nothing it can do)
  uint64_t idle_epoch = 0;          // the epoch it went idle at
  int device = 0;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  int active_lanes() const {
    int a = 0;
    for (const Lane &l : lanes) a += l.batch != nullptr;
    return a;
  }
  void note_busy() {
    const bool b = active_lanes() > 0;
    const double t = now();
    if (b && busy_since < 0) busy_since = t;
    if (!b && busy_since >= 0) {
      busy += t - busy_since;
      busy_since = -1;
    }
  }
};

namespace {

bool pool_alloc(mdmc_server *s, uint64_t bytes, uint64_t *off) {
  if (!s->pool.alloc(std::max<uint64_t>(bytes, kPoolAlign), off)) return false;
  s->peak = std::max(s->peak, s->pool.in_use());
  return true;
}

void pool_free(mdmc_server *s, uint64_t *off) {
  if (*off != kNoOffset) s->pool.release(*off);
  *off = kNoOffset;
}

// Cuts a region down to `bytes` (nothing: the region goes back to the pool).
void pool_trim(mdmc_server *s, uint64_t *off, uint64_t bytes) {
  if (*off == kNoOffset) return;
  if (bytes == 0) {
    pool_free(s, off);
    return;
  }
  s->pool.shrink(*off, bytes);
}

uint64_t cells_of(const uint32_t d[3]) {
  return d[0] < 2 || d[1] < 2 || d[2] < 2 ? 0 : (uint64_t)(d[0] - 1) * (d[1] - 1) * (d[2] - 1);
}

// Scratch one volume needs: exact for a request that runs alone; for one
// that may join a batch, the estimate in its footprint (a batch's exact
// scratch is carved when it starts).
uint64_t scratch_estimate(const mdmc_request &r, bool solo) {
  uint64_t b = 0;
  if (solo) return mdmc_scratch_bytes(r.dims, &b) == MDMC_OK ? b : 0;
  mdmc_volume v{};
  std::memcpy(v.dims, r.dims, sizeof(v.dims));
  if (mdmc_batch_scratch_bytes(&v, 1, &b) != MDMC_OK) return 0;
  return b;
}

void finish(mdmc_server *s, Req *r, int status) {
  if (r->status == MDMC_OK) r->status = status;
  r->state = 2;
  r->t_done = s->now();
  s->finished.push_back(r);
  s->completed += 1;
  if (r->status == MDMC_ERR_REJECTED) s->rejected += 1;
}

// Frees fields and scratch of a finished batch, and the output regions of
// its failed requests.  The lane's stream must have finished the batch's
// work (on an error path the caller synchronises it first).
void retire_batch(mdmc_server *s, Lane &lane) {
  Batch *b = lane.batch;
  for (Req *r : b->reqs) {
    pool_free(s, &r->field_off);
    if (r->status != MDMC_OK) {
      pool_free(s, &r->vert_off);
      pool_free(s, &r->quad_off);
      r->verts = nullptr;
      r->quads = nullptr;
    }
  }
  pool_free(s, &b->scratch_off);
  delete b;
  lane.batch = nullptr;
  lane.phase = 0;
}

// Ends every request of the lane's batch with `rc` after anything already
// queued on the lane's stream has run (its kernels may still use the
// batch's regions, which go back to the pool here).  The lane's thread must
// be done with the batch (phase 2).
void fail_batch(mdmc_server *s, Lane &lane, int rc) {
  if (hipStreamSynchronize(lane.stream) != hipSuccess) (void)hipGetLastError();
  for (Req *r : lane.batch->reqs) finish(s, r, rc);
  retire_batch(s, lane);
}

// Carves a queued request's regions: its field (generated requests) and its
// worst-case vertex and quad arrays.  False, holding nothing, if they do not
// fit now.
bool carve_request(mdmc_server *s, Req *r) {
  if (r->spec.field_kind != MDMC_FIELD_DEVICE && !pool_alloc(s, field_bytes(r->spec), &r->field_off)) return false;
  if ((r->vmax && !pool_alloc(s, 12 * r->vmax, &r->vert_off)) ||
      (r->qmax && !pool_alloc(s, 16 * r->qmax, &r->quad_off))) {
    pool_free(s, &r->field_off);
    pool_free(s, &r->vert_off);
    return false;
  }
  return true;
}

// Gives a request's regions back and returns it to the head of the queue.
void uncarve_request(mdmc_server *s, Req *r) {
  pool_free(s, &r->field_off);
  pool_free(s, &r->vert_off);
  pool_free(s, &r->quad_off);
  r->field = nullptr;
  r->verts = nullptr;
  r->quads = nullptr;
}

// Takes a carved request out of the queue into batch `b`.
void admit(mdmc_server *s, Batch *b, Req *r) {
  s->queue.erase(std::find(s->queue.begin(), s->queue.end(), r));
  r->state = 1;
  r->t_start = s->now();
  r->field = r->spec.field_kind != MDMC_FIELD_DEVICE ? reinterpret_cast<const float *>(s->base + r->field_off)
                                                     : r->spec.field;
  r->verts = r->vmax ? reinterpret_cast<float *>(s->base + r->vert_off) : nullptr;
  r->quads = r->qmax ? reinterpret_cast<int32_t *>(s->base + r->quad_off) : nullptr;
  b->reqs.push_back(r);
}

// Generator specs of the batch's requests whose field the server makes.
std::vector<mdmc_gen_spec> gen_specs(const Batch *b) {
  std::vector<mdmc_gen_spec> gens;
  for (const Req *r : b->reqs) {
    if (r->spec.field_kind == MDMC_FIELD_DEVICE) continue;
    mdmc_gen_spec g{};
    g.out = const_cast<float *>(r->field);
    std::memcpy(g.dims, r->spec.dims, sizeof(g.dims));
    g.kind = r->spec.field_kind;
    g.seed = r->spec.seed;
    g.nblobs = r->spec.nblobs;
    g.sigma_lo = r->spec.sigma_lo;
    g.sigma_hi = r->spec.sigma_hi;
    g.amp_lo = r->spec.amp_lo;
    g.amp_hi = r->spec.amp_hi;
    g.period = r->spec.period;
    g.noise_amp = r->spec.noise_amp;
    gens.push_back(g);
  }
  return gens;
}

// Gives a formed batch to an idle lane's thread.
void hand_over(mdmc_server *s, Lane &lane, Batch *b) {
  lane.batch = b;
  lane.phase = 1;
  lane.launch_rc = MDMC_OK;
  s->batches += 1;
  lane.cv.notify_one();
}

// The batch's work on the lane's stream, issued by the lane's thread without
// the server's mutex: generate the fields, count and emit (the outputs have
// their worst-case size, so no host round trip between them), copy the
// counts back, record the end event.
int launch(Lane &lane, const Batch *b) {
  const uint32_t n = (uint32_t)b->reqs.size();
  const std::vector<mdmc_gen_spec> gens = gen_specs(b);
  // the context works in the batch's scratch (none: a batch of empty volumes,
  // which needs none; the context then never touches a stale block)
  int rc = mdmc_ctx_set_scratch(lane.ctx, b->scratch_bytes ? b->base_ptr : nullptr, b->scratch_bytes);
  if (rc == MDMC_OK && !gens.empty()) {
    rc = mdmc_batch_generate_device(lane.ctx, gens.data(), (uint32_t)gens.size(), lane.stream);
    TRACE("  generate %zu rc %d", gens.size(), rc);
  }
  if (rc == MDMC_OK && b->solo) {
    const Req *r = b->reqs[0];
    rc = mdmc_extract_device(lane.ctx, r->field, r->spec.dims, r->spec.xyz_min, r->spec.xyz_max, r->spec.iso,
                             r->verts, r->vmax, r->quads, r->qmax, lane.d_counts, lane.stream);
    TRACE("  extract alone rc %d", rc);
  } else if (rc == MDMC_OK) {
    rc = mdmc_batch_extract_device(lane.ctx, b->vols.data(), n, lane.d_counts, lane.stream);
    TRACE("  extract rc %d", rc);
  }
  if (rc == MDMC_OK &&
      (hipMemcpyAsync(lane.h_counts, lane.d_counts, n * sizeof(mdmc_counts), hipMemcpyDeviceToHost,
                      lane.stream) != hipSuccess ||
       hipEventRecord(lane.ev, lane.stream) != hipSuccess)) {
    (void)hipGetLastError();
    rc = MDMC_ERR_HIP;
  }
  return rc;
}

// Starts the request at the head of the queue alone on an idle lane, through
// the single-volume extraction, if its regions and its scratch can be carved.
int start_solo(mdmc_server *s, Lane &lane) {
  Req *r = s->queue.front();
  Batch *b = new (std::nothrow) Batch();
  if (!b) return MDMC_ERR_NOMEM;
  if (!carve_request(s, r)) {
    delete b;
    return MDMC_OK;
  }
  const uint64_t sbytes = r->scratch_est;
  if (sbytes && !pool_alloc(s, sbytes, &b->scratch_off)) {
    uncarve_request(s, r);
    delete b;
    return MDMC_OK;
  }
  admit(s, b, r);
  TRACE("start %llu alone on stream %p (scratch %llu, in use %llu)", (unsigned long long)r->id,
        (void *)lane.stream, (unsigned long long)sbytes, (unsigned long long)s->pool.in_use());
  b->solo = true;
  b->scratch_bytes = sbytes;
  b->base_ptr = sbytes ? s->base + b->scratch_off : nullptr;
  hand_over(s, lane, b);
  return MDMC_OK;
}

// Starts a batch on an idle lane from the head of the queue.  Requests join
// in arrival order while their regions and the scratch of the batch with
// them can be carved (the scratch is re-carved for each candidate), so a
// started batch always holds everything it needs.  Requests that run alone
// are passed over (each starts on a free lane once it heads the queue, and
// while it heads the queue and does not fit, no batch starts either, so it
// is never starved of memory).
int start_batch(mdmc_server *s, Lane &lane) {
  if (!s->queue.empty() && s->queue.front()->solo) return start_solo(s, lane);
  Batch *b = new (std::nothrow) Batch();
  if (!b) return MDMC_ERR_NOMEM;
  uint64_t sbytes = 0;
  size_t qi = 0;
  while (qi < s->queue.size() && b->reqs.size() < s->cfg.max_batch) {
    Req *r = s->queue[qi];
    if (r->solo) {  // runs alone once it heads the queue; smaller ones may pass it
      ++qi;
      continue;
    }
    if (!carve_request(s, r)) break;  // wait for memory (arrival order)
    mdmc_volume v{};
    std::memcpy(v.dims, r->spec.dims, sizeof(v.dims));
    b->vols.push_back(v);
    uint64_t want = 0;
    int rc = mdmc_batch_scratch_bytes(b->vols.data(), (uint32_t)b->vols.size(), &want);
    bool ok = rc == MDMC_OK;
    if (ok && want != sbytes) {
      // give the old block back first, so that the new one may reuse it; if
      // the new one does not fit, the old size fits again where it was freed
      pool_free(s, &b->scratch_off);
      if (want && !pool_alloc(s, want, &b->scratch_off)) {
        ok = false;
        if (sbytes && !pool_alloc(s, sbytes, &b->scratch_off)) sbytes = 0;  // cannot happen: it was just freed
      } else {
        sbytes = want;
      }
    }
    if (!ok) {
      b->vols.pop_back();
      uncarve_request(s, r);
      break;
    }
    admit(s, b, r);
    mdmc_volume &bv = b->vols.back();
    bv.field = r->field;
    bv.iso = r->spec.iso;
    std::memcpy(bv.xyz_min, r->spec.xyz_min, sizeof(bv.xyz_min));
    std::memcpy(bv.xyz_max, r->spec.xyz_max, sizeof(bv.xyz_max));
    bv.vertices = r->verts;
    bv.vertex_capacity = r->vmax;
    bv.quads = r->quads;
    bv.quad_capacity = r->qmax;
  }
  const uint32_t n = (uint32_t)b->reqs.size();
  if (n == 0) {
    pool_free(s, &b->scratch_off);
    delete b;
    return MDMC_OK;
  }
  TRACE("start batch of %u on stream %p (scratch %llu, in use %llu)", n, (void *)lane.stream,
        (unsigned long long)sbytes, (unsigned long long)s->pool.in_use());
  b->scratch_bytes = sbytes;
  b->base_ptr = sbytes ? s->base + b->scratch_off : nullptr;
  hand_over(s, lane, b);
  return MDMC_OK;
}

// One pass over the lanes; returns true if anything changed.
bool pump(mdmc_server *s) {
  bool progress = false;
  for (Lane &lane : s->lanes) {
    Batch *b = lane.batch;
    if (!b || lane.phase != 2) continue;  // idle, or its thread is still launching
    if (lane.launch_rc != MDMC_OK) {
      progress = true;
      fail_batch(s, lane, lane.launch_rc);
      continue;
    }
    const hipError_t q = hipEventQuery(lane.ev);
    if (q == hipErrorNotReady) continue;
    progress = true;
    if (q != hipSuccess) {
      (void)hipGetLastError();
      fail_batch(s, lane, MDMC_ERR_HIP);
      continue;
    }
    for (size_t i = 0; i < b->reqs.size(); ++i) {
      Req *r = b->reqs[i];
      r->counts = lane.h_counts[i];
      if (r->counts.status) r->status = MDMC_ERR_INTERNAL;
      if (r->status == MDMC_OK) {
        // keep exactly the mesh the count gave
        pool_trim(s, &r->vert_off, 12ull * r->counts.num_vertices);
        pool_trim(s, &r->quad_off, 16ull * r->counts.num_quads);
        if (r->vert_off == kNoOffset) r->verts = nullptr;
        if (r->quad_off == kNoOffset) r->quads = nullptr;
      }
      finish(s, r, MDMC_OK);
    }
    TRACE("batch of %zu done", b->reqs.size());
    retire_batch(s, lane);
  }
  // an idle lane takes whatever is queued at once: holding a lane back for a
  // fuller batch lowered neither the GPU time per request nor the tail
  for (Lane &lane : s->lanes) {
    if (lane.batch || s->queue.empty()) continue;
    if (start_batch(s, lane) != MDMC_OK) break;
    if (lane.batch) progress = true;
  }
  // With no batch in flight and no region in use (results all released), a
  // request whose footprint fits the budget always starts; if the head still
  // did not, it cannot (the footprint check at submission makes this a
  // safeguard): reject it rather than wait for a wake-up that cannot come.
  if (!s->queue.empty() && s->active_lanes() == 0 && s->pool.in_use() == 0) {
    Req *r = s->queue.front();
    s->queue.pop_front();
    TRACE("reject %llu: does not fit an empty pool", (unsigned long long)r->id);
    finish(s, r, MDMC_ERR_REJECTED);
    progress = true;
  }
  s->note_busy();
  return progress;
}

// The worker: pumps the lanes under the lock; while a phase is on the device
// it watches the lanes' events WITHOUT the lock (only the worker records or
// retires them), so callers are never held up by its polling; with nothing in
// flight it sleeps until a call bumps the epoch.
void worker_loop(mdmc_server *s) {
  (void)hipSetDevice(s->device);
  std::vector<hipEvent_t> watch;
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->stop.load()) {
    const uint64_t seen = s->epoch.load();
    s->idle = false;
    bool moved = false;
    while (pump(s)) moved = true;
    watch.clear();
    bool launching = false;
    for (const Lane &l : s->lanes) {
      if (l.batch && l.phase == 2) watch.push_back(l.ev);
      launching = launching || (l.batch && l.phase == 1);
    }
    if (moved) s->cv_done.notify_all();
    if (watch.empty()) {
      // nothing on the device to watch: wait for a call or for a lane's
      // thread to finish launching (either bumps the epoch)
      if (!launching) {
        s->idle = true;
        s->idle_epoch = seen;
        s->cv_done.notify_all();
      }
      s->cv_work.wait(lk, [&] { return s->stop.load() || s->epoch.load() != seen; });
      continue;
    }
    lk.unlock();
    for (;;) {
      bool ready = false;
      for (hipEvent_t e : watch)
        if (hipEventQuery(e) != hipErrorNotReady) {
          ready = true;
          break;
        }
      if (ready || s->stop.load() || s->epoch.load() != seen) break;
      std::this_thread::yield();
    }
    lk.lock();
  }
  s->idle = true;
  s->cv_done.notify_all();
}

void wake(mdmc_server *s) {
  s->epoch.fetch_add(1);
  s->cv_work.notify_one();
}

// A lane's thread: launches each batch handed to it, then reports back to
// the worker (which watches the batch's end event).
void lane_loop(mdmc_server *s, Lane *lane) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    lane->cv.wait(lk, [&] { return s->stop.load() || lane->phase == 1; });
    if (lane->phase != 1) break;  // stopping
    const Batch *b = lane->batch;
    lk.unlock();
    int rc;
    {
      std::lock_guard<std::mutex> cl(lane->ctx_mu);
      rc = launch(*lane, b);
    }
    lk.lock();
    lane->launch_rc = rc;
    lane->phase = 2;
    wake(s);
  }
}

}  // namespace

extern "C" {

int mdmc_server_create(const mdmc_server_config *cfg, mdmc_server **out) {
  if (!cfg || !out) return MDMC_ERR_INVALID;
  *out = nullptr;
  if (cfg->max_batch == 0 || cfg->max_batch > MDMC_MAX_BATCH || cfg->num_streams == 0 || cfg->num_streams > 8 ||
      cfg->memory_budget < (1ull << 20))
    return MDMC_ERR_INVALID;
  int rc = mdmc_setup();
  if (rc != MDMC_OK) return rc;
  mdmc_server *s = new (std::nothrow) mdmc_server();
  if (!s) return MDMC_ERR_NOMEM;
  s->cfg = *cfg;
  if (hipMalloc(reinterpret_cast<void **>(&s->base), cfg->memory_budget) != hipSuccess) {
    (void)hipGetLastError();
    delete s;
    return MDMC_ERR_NOMEM;
  }
  s->pool.reset(cfg->memory_budget);
  for (uint32_t i = 0; i < cfg->num_streams; ++i) s->lanes.emplace_back();
  for (Lane &l : s->lanes) {
    bool ok = hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&l.ev, hipEventDisableTiming) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&l.d_counts), MDMC_MAX_BATCH * sizeof(mdmc_counts)) == hipSuccess &&
              hipHostMalloc(reinterpret_cast<void **>(&l.h_counts), MDMC_MAX_BATCH * sizeof(mdmc_counts)) ==
                  hipSuccess &&
              mdmc_ctx_create(&l.ctx) == MDMC_OK;
    if (!ok) {
      (void)hipGetLastError();
      mdmc_server_destroy(s);
      return MDMC_ERR_NOMEM;
    }
  }
  if (hipGetDevice(&s->device) != hipSuccess) {
    (void)hipGetLastError();
    mdmc_server_destroy(s);
    return MDMC_ERR_HIP;
  }
  try {
    for (Lane &l : s->lanes) l.thread = std::thread(lane_loop, s, &l);
    s->worker = std::thread(worker_loop, s);
  } catch (...) {
    mdmc_server_destroy(s);
    return MDMC_ERR_NOMEM;
  }
  *out = s;
  return MDMC_OK;
}

void mdmc_server_destroy(mdmc_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop.store(true);
    for (Lane &l : s->lanes) l.cv.notify_all();
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (Lane &l : s->lanes)
    if (l.thread.joinable()) l.thread.join();
  for (Lane &l : s->lanes) {
    if (l.stream) (void)hipStreamSynchronize(l.stream);
    if (l.batch) delete l.batch;
    if (l.ctx) mdmc_ctx_destroy(l.ctx);
    if (l.d_counts) (void)hipFree(l.d_counts);
    if (l.h_counts) (void)hipHostFree(l.h_counts);
    if (l.ev) (void)hipEventDestroy(l.ev);
    if (l.stream) (void)hipStreamDestroy(l.stream);
  }
  for (auto &kv : s->reqs) delete kv.second;
  if (s->base) (void)hipFree(s->base);
  delete s;
}

int mdmc_server_submit(mdmc_server *s, const mdmc_request *req, uint64_t *id) {
  if (!s || !req || !id) return MDMC_ERR_INVALID;
  if (req->field_kind != MDMC_FIELD_DEVICE && req->field_kind != MDMC_FIELD_BLOBS &&
      req->field_kind != MDMC_FIELD_GYROID)
    return MDMC_ERR_INVALID;
  if (req->field_kind == MDMC_FIELD_DEVICE && !req->field &&
      (uint64_t)req->dims[0] * req->dims[1] * req->dims[2] > 0)
    return MDMC_ERR_INVALID;
  Req *r = new (std::nothrow) Req();
  if (!r) return MDMC_ERR_NOMEM;
  std::lock_guard<std::mutex> lk(s->mu);
  r->id = s->next_id++;
  TRACE("submit %llu dims %u %u %u kind %d", (unsigned long long)r->id, req->dims[0], req->dims[1], req->dims[2],
        req->field_kind);
  r->spec = *req;
  r->t_submit = s->now();
  const uint64_t fb = field_bytes(*req);
  r->solo = s->solo_cells && cells_of(req->dims) >= s->solo_cells;
  r->scratch_est = scratch_estimate(*req, r->solo);
  worst_case_mesh(req->dims, &r->vmax, &r->qmax);
  r->footprint = (req->field_kind != MDMC_FIELD_DEVICE ? round_up(fb) : 0) + round_up(r->scratch_est) +
                 round_up(12 * r->vmax) + round_up(16 * r->qmax);
  s->reqs[r->id] = r;
  s->submitted += 1;
  *id = r->id;
  if (r->footprint > s->pool.size()) {
    finish(s, r, MDMC_ERR_REJECTED);
    s->cv_done.notify_all();
    return MDMC_OK;
  }
  s->queue.push_back(r);
  wake(s);
  return MDMC_OK;
}

int mdmc_server_poll(mdmc_server *s, int wait, mdmc_result *out, uint32_t max, uint32_t *n) {
  if (!s || !n || (max && !out)) return MDMC_ERR_INVALID;
  *n = 0;
  std::unique_lock<std::mutex> lk(s->mu);
  if (wait) {
    // until a result is there, or the worker has gone idle since the last
    // call that could start work (then no result comes without another call)
    s->cv_done.wait(lk, [&] {
      return !s->finished.empty() || (s->idle && s->idle_epoch == s->epoch.load());
    });
  }
  while (*n < max && !s->finished.empty()) {
    Req *r = s->finished.front();
    s->finished.pop_front();
    r->state = 3;
    mdmc_result &o = out[(*n)++];
    std::memset(&o, 0, sizeof(o));
    o.id = r->id;
    o.tag = r->spec.tag;
    o.status = r->status;
    o.counts = r->counts;
    o.vertices = r->verts;
    o.quads = r->quads;
    o.t_submit = r->t_submit;
    o.t_start = r->t_start;
    o.t_done = r->t_done;
  }
  return MDMC_OK;
}

int mdmc_server_copy_result(mdmc_server *s, uint64_t id, float *h_vertices, int32_t *h_quads) {
  if (!s) return MDMC_ERR_INVALID;
  const float *dv = nullptr;
  const int32_t *dq = nullptr;
  uint64_t V = 0, Q = 0;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end() || it->second->state < 2) return MDMC_ERR_INVALID;
    Req *r = it->second;
    if (r->status != MDMC_OK) return r->status;
    dv = r->verts;
    dq = r->quads;
    V = r->counts.num_vertices;
    Q = r->counts.num_quads;
  }
  // the result stays put until the caller releases it
  if (h_vertices && dv && hipMemcpy(h_vertices, dv, 12ull * V, hipMemcpyDeviceToHost) != hipSuccess)
    return MDMC_ERR_HIP;
  if (h_quads && dq && hipMemcpy(h_quads, dq, 16ull * Q, hipMemcpyDeviceToHost) != hipSuccess)
    return MDMC_ERR_HIP;
  return MDMC_OK;
}

int mdmc_server_release(mdmc_server *s, uint64_t id) {
  if (!s) return MDMC_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end() || it->second->state < 2) return MDMC_ERR_INVALID;
  Req *r = it->second;
  pool_free(s, &r->vert_off);
  pool_free(s, &r->quad_off);
  if (r->state == 2) {  // still in the finished list: drop it there too
    auto f = std::find(s->finished.begin(), s->finished.end(), r);
    if (f != s->finished.end()) s->finished.erase(f);
  }
  s->reqs.erase(it);
  delete r;
  wake(s);  // freed memory may let queued work start
  return MDMC_OK;
}

int mdmc_server_set_option(mdmc_server *s, int option, int value) {
  if (!s) return MDMC_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (option == MDMC_SERVER_OPT_SOLO_CELLS) {
    if (value < 0) return MDMC_ERR_INVALID;
    s->solo_cells = (uint64_t)value;
    return MDMC_OK;
  }
  for (Lane &l : s->lanes) {
    std::lock_guard<std::mutex> cl(l.ctx_mu);  // not while the lane's thread launches
    const int rc = mdmc_ctx_set_option(l.ctx, option, value);
    if (rc != MDMC_OK) return rc;
  }
  return MDMC_OK;
}

int mdmc_server_get_stats(mdmc_server *s, mdmc_server_stats *st) {
  if (!s || !st) return MDMC_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  s->note_busy();
  std::memset(st, 0, sizeof(*st));
  st->submitted = s->submitted;
  st->completed = s->completed;
  st->rejected = s->rejected;
  st->batches = s->batches;
  st->queued = s->queue.size();
  uint64_t inflight = 0;
  for (const Lane &l : s->lanes) inflight += l.batch ? l.batch->reqs.size() : 0;
  st->in_flight = inflight;
  st->memory_in_use = s->pool.in_use();
  st->memory_peak = s->peak;
  st->memory_budget = s->pool.size();
  st->busy_seconds = s->busy + (s->busy_since >= 0 ? s->now() - s->busy_since : 0.0);
  return MDMC_OK;
}

double mdmc_server_now(mdmc_server *s) { return s ? s->now() : 0.0; }

}  // extern "C"
