// Host I/O either side of the extraction: staged host<->device copies for
// the host entry (mdmc_run) and the OBJ writer.
//
// Staged copies: pageable host memory cannot be the source or target of a
// DMA, so the copy goes through pinned chunks.  A few host threads each own
// two pinned chunks and a stream; thread t moves chunks t, t + T, t + 2T, ...
// (memcpy into its free pinned chunk, then an async DMA on its stream; for
// downloads the other way round), so host copies of one thread overlap the
// DMAs of the others and of its own previous chunk.  All streams run at the
// default priority.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/mdmc.h"
#include "mdmc_fmt.h"
#include "mdmc_io.h"

namespace {

constexpr int kIoThreads = 4;
constexpr size_t kIoChunk = 4u << 20;  // bytes per pinned chunk

struct IoLane {
  hipStream_t stream = nullptr;
  void *pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
};

struct IoStaging {
  int device = -1;
  IoLane lanes[kIoThreads];
  bool ready = false;
};

// one staging set per direction: an upload and a download (the next slab in
// and the last slab's mesh out, say) run at the same time over full-duplex PCIe
std::mutex g_io_mutex[2];
IoStaging g_io[2];

void io_release(IoStaging &io) {
  for (IoLane &l : io.lanes) {
    if (l.stream) (void)hipStreamSynchronize(l.stream);
    for (int b = 0; b < 2; ++b) {
      if (l.pin[b]) (void)hipHostFree(l.pin[b]);
      if (l.ev[b]) (void)hipEventDestroy(l.ev[b]);
      l.pin[b] = nullptr;
      l.ev[b] = nullptr;
    }
    if (l.stream) (void)hipStreamDestroy(l.stream);
    l.stream = nullptr;
  }
  io.ready = false;
  io.device = -1;
}

int io_init(IoStaging &io, int device) {
  if (io.ready && io.device == device) return MDMC_OK;
  if (io.ready) io_release(io);
  for (IoLane &l : io.lanes) {
    if (hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking) != hipSuccess) goto fail;
    for (int b = 0; b < 2; ++b) {
      if (hipHostMalloc(&l.pin[b], kIoChunk) != hipSuccess) goto fail;
      if (hipEventCreateWithFlags(&l.ev[b], hipEventDisableTiming) != hipSuccess) goto fail;
    }
  }
  io.ready = true;
  io.device = device;
  return MDMC_OK;
fail:
  (void)hipGetLastError();
  io_release(io);
  return MDMC_ERR_NOMEM;
}

// dir 0: host -> device, 1: device -> host.  Returns the first HIP error.
hipError_t io_lane_copy(IoLane &l, int t, int nthreads, char *host, char *dev, size_t bytes, int dir) {
  const size_t nchunks = (bytes + kIoChunk - 1) / kIoChunk;
  bool pending[2] = {false, false};
  size_t pend_off[2] = {0, 0}, pend_len[2] = {0, 0};
  int b = 0;
  hipError_t err = hipSuccess;
  for (size_t c = (size_t)t; c < nchunks; c += (size_t)nthreads) {
    const size_t off = c * kIoChunk, len = bytes - off < kIoChunk ? bytes - off : kIoChunk;
    if (pending[b]) {
      err = hipEventSynchronize(l.ev[b]);
      if (err != hipSuccess) return err;
      if (dir == 1) std::memcpy(host + pend_off[b], l.pin[b], pend_len[b]);
      pending[b] = false;
    }
    if (dir == 0) {
      std::memcpy(l.pin[b], host + off, len);
      err = hipMemcpyAsync(dev + off, l.pin[b], len, hipMemcpyHostToDevice, l.stream);
    } else {
      err = hipMemcpyAsync(l.pin[b], dev + off, len, hipMemcpyDeviceToHost, l.stream);
    }
    if (err != hipSuccess) return err;
    err = hipEventRecord(l.ev[b], l.stream);
    if (err != hipSuccess) return err;
    pending[b] = true;
    pend_off[b] = off;
    pend_len[b] = len;
    b ^= 1;
  }
  for (int bb = 0; bb < 2; ++bb) {  // drain both chunks
    if (!pending[bb]) continue;
    err = hipEventSynchronize(l.ev[bb]);
    if (err != hipSuccess) return err;
    if (dir == 1) std::memcpy(host + pend_off[bb], l.pin[bb], pend_len[bb]);
    pending[bb] = false;
  }
  return hipSuccess;
}

}  // namespace

namespace mdmc {

// Synchronous staged copy between pageable host memory and device memory on
// the current device (dir 0: to the device, 1: to the host).
int staged_copy(void *host, void *dev, size_t bytes, int dir) {
  if (bytes == 0) return MDMC_OK;
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) {
    (void)hipGetLastError();
    return MDMC_ERR_HIP;
  }
  if (dir != 0 && dir != 1) return MDMC_ERR_INVALID;
  IoStaging &io = g_io[dir];
  std::lock_guard<std::mutex> lk(g_io_mutex[dir]);
  int rc = io_init(io, device);
  if (rc != MDMC_OK) return rc;
  const size_t nchunks = (bytes + kIoChunk - 1) / kIoChunk;
  const int nthreads = (int)(nchunks < (size_t)kIoThreads ? nchunks : (size_t)kIoThreads);
  hipError_t errs[kIoThreads];
  std::vector<std::thread> pool;
  for (int t = 1; t < nthreads; ++t)
    pool.emplace_back([&, t] {
      (void)hipSetDevice(device);
      errs[t] = io_lane_copy(io.lanes[t], t, nthreads, static_cast<char *>(host), static_cast<char *>(dev),
                             bytes, dir);
    });
  errs[0] = io_lane_copy(io.lanes[0], 0, nthreads, static_cast<char *>(host), static_cast<char *>(dev), bytes,
                         dir);
  for (std::thread &th : pool) th.join();
  for (int t = 0; t < nthreads; ++t)
    if (errs[t] != hipSuccess) {
      (void)hipGetLastError();
      return MDMC_ERR_HIP;
    }
  return MDMC_OK;
}

void staged_release() {
  for (int d = 0; d < 2; ++d) {
    std::lock_guard<std::mutex> lk(g_io_mutex[d]);
    if (g_io[d].ready) io_release(g_io[d]);
  }
}

}  // namespace mdmc

namespace {

// OBJ text of vertices [i0, i1) or of the two triangles of quads [i0, i1),
// one line (or two) per element, into dst; returns the bytes written.
constexpr size_t kObjLineMax = 128;  // a vertex line or a quad's two lines fit

size_t obj_vertices(char *dst, const float *v, uint64_t i0, uint64_t i1) {
  size_t used = 0;
  // "v x y z" with six significant digits (the reference writes floats
  // through std::ostream at its default precision, dmc.cu:2238-2246): the
  // bytes of printf's "v %g %g %g\n" (mdmc_fmt.h)
  for (uint64_t i = i0; i < i1; ++i) {
    char *o = dst + used;
    size_t n = 0;
    for (int c = 0; c < 3; ++c) {
      o[n++] = c ? ' ' : 'v';
      if (!c) o[n++] = ' ';
      n += mdmc::fmt_g(o + n, v[3 * i + c]);
    }
    o[n++] = '\n';
    used += n;
  }
  return used;
}

size_t obj_faces(char *dst, const int32_t *quads, uint64_t i0, uint64_t i1) {
  size_t used = 0;
  // two triangles (q0,q1,q2), (q2,q3,q0) per quad, 1-based indices
  for (uint64_t q = i0; q < i1; ++q) {
    const uint32_t a = (uint32_t)quads[4 * q] + 1u, b = (uint32_t)quads[4 * q + 1] + 1u,
                   c = (uint32_t)quads[4 * q + 2] + 1u, d = (uint32_t)quads[4 * q + 3] + 1u;
    const uint32_t idx[6] = {a, b, c, c, d, a};
    char *o = dst + used;
    size_t n = 0;
    for (int k = 0; k < 6; ++k) {
      if (k % 3 == 0) o[n++] = 'f';
      o[n++] = ' ';
      n += mdmc::fmt_u32(o + n, idx[k]);
      if (k % 3 == 2) o[n++] = '\n';
    }
    used += n;
  }
  return used;
}

}  // namespace

// Formatting the text is what costs (about 100 ns per vertex line with the
// integer formatter of mdmc_fmt.h, a few hundred through snprintf), so chunks
// of lines are formatted by several threads, a round of chunks at a time (memory
// stays at a round's buffers however large the mesh), and written in order:
// the bytes are those of one thread writing line after line.
extern "C" int mdmc_write_obj(const char *path, const float *vertices, uint64_t num_vertices, const int32_t *quads,
                              uint64_t num_quads) {
  if (!path || (num_vertices && !vertices) || (num_quads && !quads))
    return mdmc::set_error(MDMC_ERR_INVALID, "mdmc_write_obj: null path or mesh pointer");
  // every face must name an existing vertex; checked before the file is made
  for (uint64_t k = 0; k < 4 * num_quads; ++k)
    if (quads[k] < 0 || (uint64_t)quads[k] >= num_vertices)
      return mdmc::set_error(MDMC_ERR_INVALID, "mdmc_write_obj: quad index outside the vertex array");
  FILE *f = std::fopen(path, "wb");
  if (!f) return mdmc::set_error(MDMC_ERR_INVALID, "mdmc_write_obj: cannot open the output file");
  constexpr uint64_t kLines = 16384;  // elements per chunk
  const uint64_t nvc = (num_vertices + kLines - 1) / kLines, nqc = (num_quads + kLines - 1) / kLines;
  const uint64_t chunks = nvc + nqc;
  const unsigned hw = std::thread::hardware_concurrency();
  const uint64_t T = std::max<uint64_t>(1, std::min<uint64_t>({16, hw ? hw : 1, chunks}));
  std::vector<std::vector<char>> buf(T, std::vector<char>(kLines * kObjLineMax));
  std::vector<size_t> used(T, 0);
  auto format = [&](uint64_t c, uint64_t t) {
    if (c < nvc)
      used[t] = obj_vertices(buf[t].data(), vertices, c * kLines, std::min(num_vertices, (c + 1) * kLines));
    else
      used[t] = obj_faces(buf[t].data(), quads, (c - nvc) * kLines, std::min(num_quads, (c - nvc + 1) * kLines));
  };
  bool ok = true;
  for (uint64_t c0 = 0; c0 < chunks && ok; c0 += T) {
    const uint64_t n = std::min(T, chunks - c0);
    std::vector<std::thread> th;
    try {
      for (uint64_t t = 1; t < n; ++t) th.emplace_back(format, c0 + t, t);
    } catch (...) {
      for (uint64_t t = 1 + th.size(); t < n; ++t) format(c0 + t, t);  // no thread: format here
    }
    format(c0, 0);
    for (std::thread &x : th) x.join();
    for (uint64_t t = 0; t < n && ok; ++t)
      if (used[t] && std::fwrite(buf[t].data(), 1, used[t], f) != used[t]) ok = false;
  }
  if (std::fclose(f) != 0) ok = false;
  return ok ? MDMC_OK : mdmc::set_error(MDMC_ERR_INVALID, "mdmc_write_obj: write to the output file failed");
}
