/*
 * mdmc — MI355X-native Dual Marching Cubes, C ABI.
 *
 * Drop-in for the extraction path of the reference DMC_CUDA library
 * (DMC_CUDA/DMC_CUDA/include/dmc.hpp:40-57): grid dims, scalar-field pointer
 * and isovalue in; vertex array and quad index list out.  Plain C types only;
 * HIP streams are passed as `void*` (a hipStream_t, NULL = default stream).
 *
 * Output layout (identical to the reference, dmc.cu:901-921, 1478-1513,
 * 2011-2047): active cells in ascending cell index; each contributes its iso
 * (dual) vertices by table label, then the crossing points of its bipolar
 * owned edges 6, 9, 10 in that order.  Quads are int32 x4 vertex indices; the
 * reference's triangles are (q0,q1,q2),(q2,q3,q0).  Vertex xyz follows
 * utils::ijk_to_xyz (include/utils.hpp:118-122): i / cells * (max-min) + min.
 *
 * Every entry point returns an MDMC_* status; none calls exit().
 */
#ifndef MDMC_H
#define MDMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MDMC_OK = 0,
  MDMC_ERR_INVALID = -1,   /* bad argument (null pointer, dims) */
  MDMC_ERR_HIP = -2,       /* a HIP runtime call failed */
  MDMC_ERR_NOMEM = -3,     /* device or host allocation failed */
  MDMC_ERR_CAPACITY = -4,  /* caller-provided output buffer too small */
  MDMC_ERR_TOO_LARGE = -5, /* grid beyond the 32-bit index range */
  MDMC_ERR_NOT_SETUP = -6, /* mdmc_setup() not called */
  MDMC_ERR_INTERNAL = -7,  /* a device-side invariant check failed */
  MDMC_ERR_REJECTED = -8   /* request larger than the serving memory budget */
};

/* Device-side status bits reported in mdmc_counts.status. */
enum {
  MDMC_STATUS_VERTEX_OVERFLOW = 1, /* vertices beyond vertex_capacity dropped */
  MDMC_STATUS_QUAD_OVERFLOW = 2,   /* quads beyond quad_capacity dropped */
  MDMC_STATUS_LOOKUP_FAILED = 4,   /* a ring neighbour was not found */
  MDMC_STATUS_SCAN_TIMEOUT = 8,    /* look-back spin limit hit */
  MDMC_STATUS_INTERNAL = 16,       /* inconsistent count-pass data (bug) */
  MDMC_STATUS_INDEX_RANGE = 32     /* more than 2^32-1 active cells, vertices or quads */
};

/* Sizes of one extraction.  Written by the device; read after a stream sync. */
typedef struct mdmc_counts {
  uint32_t num_active;   /* cells with 0 < config < 255 */
  uint32_t num_vertices; /* iso vertices + owned edge points */
  uint32_t num_quads;
  uint32_t status;       /* MDMC_STATUS_* bits */
} mdmc_counts;

/* Host-side mesh returned by mdmc_run (free with mdmc_mesh_free). */
typedef struct mdmc_mesh {
  float *vertices;   /* 3 * num_vertices */
  int32_t *quads;    /* 4 * num_quads */
  uint32_t num_vertices;
  uint32_t num_quads;
  uint32_t num_active;
} mdmc_mesh;

/* ---- reference-compatible entry points ---------------------------------- */

/* Replaces dmc::setup_device_luts() (include/dmc.hpp:40, dmc.cu:713-779):
 * builds the case tables and uploads them to device constant memory on the
 * current device.  Idempotent. */
int mdmc_setup(void);

/* Replaces dmc::cleanup_device_luts() (include/dmc.hpp:48, dmc.cu:781-811):
 * releases the library's cached device state.  Idempotent. */
int mdmc_cleanup(void);

/* Device copies of the case tables in the reference's array formats, as the
 * reference's 11-pointer setup hands them to its caller (the tables of
 * dmc.cu:45-661).  All entries are bytes. */
typedef struct mdmc_device_luts {
  uint8_t *config_edge_lut1;            /* [256][12] edge -> iso-vertex label, 0xff none */
  uint8_t *config_edge_lut2;            /* [256][12] the 2B/3B variant */
  uint8_t *num_vertex_lut1;             /* [256] */
  uint8_t *num_vertex_lut2;             /* [256] */
  uint8_t *config_2B_3B_lut;            /* [36] the 2B/3B configurations, ascending */
  uint8_t *config_2B_3B_ambiguous_face; /* [36] their ambiguous face, same order */
  uint8_t *opposite_face_lut;           /* [6] */
  uint8_t *face_to_check_dir_lut;       /* [6] POS_X=0, NEG_X, POS_Y, NEG_Y, POS_Z, NEG_Z=5 */
  uint8_t *edge_belonged_voxel_lut;     /* [12] 0x80|0x40|0x20 = owner at -x|-y|-z, low bits
                                           = owner's edge; 0xff = the cell's own edge */
  uint8_t *circular_edge_lut;           /* [3][4] ring of edges 6, 9, 10 (ccw) */
  uint8_t *voxel_local_edges;           /* [3] 6, 9, 10 */
} mdmc_device_luts;

/* Replaces the 11-pointer dmc::setup_device_luts(iso_vertex_m_type**, ...)
 * (include/dmc.hpp:42-46, dmc.cu:721-779): mdmc_setup() plus device copies
 * of the tables above on the current device, owned by the library (the same
 * pointers on every call until cleanup).  The extraction itself does not read
 * these copies (it uses its own constant-memory forms); they are there for
 * callers that read the tables.  One difference: the 2B/3B list is in
 * ascending configuration order (the reference lists 2B then 3B cases). */
int mdmc_setup_device_luts(mdmc_device_luts *out);

/* Replaces the 11-pointer dmc::cleanup_device_luts(iso_vertex_m_type*, ...)
 * (include/dmc.hpp:50-54, dmc.cu:793-811): frees the table copies and runs
 * mdmc_cleanup().  MDMC_ERR_INVALID, freeing nothing, when a pointer is not
 * the one mdmc_setup_device_luts handed out on this device. */
int mdmc_cleanup_device_luts(const mdmc_device_luts *luts);

/* The same tables in host memory (no device call), one block of
 * MDMC_REF_LUTS_BYTES: the fields of mdmc_device_luts in declaration order at
 * byte offsets 0, 3072, 6144, 6400, 6656, 6692, 6728, 6734, 6740, 6752, 6764
 * (the device copies have the same layout). */
#define MDMC_REF_LUTS_BYTES 6784
int mdmc_reference_luts_host(uint8_t *out);

/* Replaces dmc::run_dmc(vertices, triangles, grid, xyz_min, xyz_max, iso,
 * num_smooth) (include/dmc.hpp:56-57, dmc.cu:2069-2208).  `field` is host
 * memory, dims[3] are point counts (x fastest).  num_smooth smoothing
 * iterations follow the extraction (dmc.cu:2157-2169).  The mesh is
 * allocated by the library. */
int mdmc_run(const float *field, const uint32_t dims[3], const float xyz_min[3],
             const float xyz_max[3], float iso, uint32_t num_smooth, mdmc_mesh *out);

/* mdmc_run for grids larger than the device memory one wants to give it (no
 * counterpart in the reference, which copies the whole grid to the device,
 * dmc.cu:2077-2079): the host field goes through the device in z-slabs
 * sized so that two slabs (the next one is uploaded while the current one is
 * extracted) and the scratch of one fit `device_budget` bytes (the mesh
 * buffers of one slab come on top).  The mesh is identical to mdmc_run's
 * with num_smooth = 0 (smoothing needs the whole mesh and is not offered
 * here).  Like mdmc_run's, these device buffers stay allocated between calls
 * (the field slabs at exactly this call's size) until mdmc_cleanup, so
 * repeated calls on one grid size allocate nothing. */
int mdmc_run_out_of_core(const float *field, const uint32_t dims[3], const float xyz_min[3],
                         const float xyz_max[3], float iso, uint64_t device_budget, mdmc_mesh *out);
/* Cell planes per slab mdmc_run_out_of_core uses for these dims and budget
 * (MDMC_ERR_NOMEM if not even one plane fits).  Host-only arithmetic. */
int mdmc_out_of_core_slab_planes(const uint32_t dims[3], uint64_t device_budget, uint32_t *slab_planes);

void mdmc_mesh_free(mdmc_mesh *mesh);

/* A mesh in device memory, allocated by the library (mdmc_run_device). */
typedef struct mdmc_device_mesh {
  float *vertices;   /* device, 3 * num_vertices */
  int32_t *quads;    /* device, 4 * num_quads */
  uint32_t num_vertices;
  uint32_t num_quads;
  uint32_t num_active;
} mdmc_device_mesh;

/* The device-resident form of run_dmc (no counterpart in the reference, which
 * takes a host grid, dmc.cu:2077-2079): `d_field` is device memory on the
 * current device; count, exact output allocation (hipMalloc), emit and
 * num_smooth smoothing iterations run on `stream` (a hipStream_t, NULL for
 * the default stream), and the call returns once the mesh is complete.
 * Shares the library's default context with mdmc_run (calls serialise).
 * An empty mesh has null pointers. */
int mdmc_run_device(const float *d_field, const uint32_t dims[3], const float xyz_min[3],
                    const float xyz_max[3], float iso, uint32_t num_smooth, void *stream,
                    mdmc_device_mesh *out);
void mdmc_device_mesh_free(mdmc_device_mesh *mesh);

/* Writes a mesh as Wavefront OBJ, like the reference's dump_obj
 * (dmc.cu:2238-2246): "v x y z" per vertex (six significant digits, as the
 * reference's std::ostream default) and two 1-based triangle faces
 * (q0,q1,q2), (q2,q3,q0) per quad.  Host memory only; no device call.
 * MDMC_ERR_INVALID, with no file written, if a quad index lies outside
 * [0, num_vertices). */
int mdmc_write_obj(const char *path, const float *vertices, uint64_t num_vertices, const int32_t *quads,
                   uint64_t num_quads);

const char *mdmc_status_string(int status);

/* Message of the last failing call on this thread. */
const char *mdmc_last_error(void);

/* ---- device-resident path (no counterpart in the reference, whose only
 * entry copies host memory every call; this is the same computation with the
 * field and outputs resident in HBM and stream-ordered launches) ---------- */

typedef struct mdmc_ctx mdmc_ctx;

/* A context owns scratch buffers on the current device; one context per
 * concurrent stream/thread. */
int mdmc_ctx_create(mdmc_ctx **ctx);
void mdmc_ctx_destroy(mdmc_ctx *ctx);

/* Classify + scan.  Counts are written to device memory `d_counts`; nothing
 * is synchronised. */
int mdmc_count_device(mdmc_ctx *ctx, const float *d_field, const uint32_t dims[3], float iso,
                      mdmc_counts *d_counts, void *stream);

/* Emit vertices and quads for the extraction last counted on `ctx` (same
 * field, dims, iso, stream).  Writes beyond a capacity are dropped and flagged
 * in d_counts->status. */
int mdmc_emit_device(mdmc_ctx *ctx, const float *d_field, const float xyz_min[3],
                     const float xyz_max[3], float *d_vertices, uint64_t vertex_capacity,
                     int32_t *d_quads, uint64_t quad_capacity, mdmc_counts *d_counts,
                     void *stream);

/* mdmc_count_device followed by mdmc_emit_device. */
int mdmc_extract_device(mdmc_ctx *ctx, const float *d_field, const uint32_t dims[3],
                        const float xyz_min[3], const float xyz_max[3], float iso,
                        float *d_vertices, uint64_t vertex_capacity, int32_t *d_quads,
                        uint64_t quad_capacity, mdmc_counts *d_counts, void *stream);

/* mdmc_extract_device without a context and without any allocation (no
 * counterpart in the reference; for callers that cannot allocate device
 * memory: code under stream capture, tight serving loops, harnesses that
 * forbid it).  The scratch is a fixed device array inside the library, one
 * per GPU, of MDMC_NOALLOC_SCRATCH_BYTES; its address is taken by
 * mdmc_setup() on the current device, which must have run there.  Grids whose
 * scratch does not fit (roughly: more than 256^3 points, or rows longer than
 * 256 cells in grids of more than about 128^2 rows) return MDMC_ERR_CAPACITY
 * before anything is launched.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (other streams, other host threads) race and
 * must be serialised by the caller.  Results are identical to
 * mdmc_extract_device's. */
#define MDMC_NOALLOC_SCRATCH_BYTES (256ull << 20)
int mdmc_extract_device_noalloc(const float *d_field, const uint32_t dims[3], const float xyz_min[3],
                                const float xyz_max[3], float iso, float *d_vertices, uint64_t vertex_capacity,
                                int32_t *d_quads, uint64_t quad_capacity, mdmc_counts *d_counts, void *stream);

/* Smoothing iterations over the mesh the last mdmc_emit_device /
 * mdmc_extract_device of this context wrote (pass the same outputs): each
 * iteration is smooth_edge_vertices (dmc.cu:1868-1981) followed by
 * calc_iso_vertices_kern (dmc.cu:1517-1593), as in run_dmc's num_smooth loop
 * (dmc.cu:2157-2169).  Quads are unchanged.  The split corner of a convex
 * ring is found by comparing cosines with IEEE sqrt and division (the
 * reference uses acosf and rsqrtf).  Stream-ordered; status bits in
 * d_counts->status. */
int mdmc_smooth_device(mdmc_ctx *ctx, float *d_vertices, uint64_t vertex_capacity, const int32_t *d_quads,
                       uint64_t quad_capacity, uint32_t num_smooth, mdmc_counts *d_counts, void *stream);

/* Per-active-cell records of the last count on `ctx`, in ascending cell
 * order: cell index1D, config, flags (bit 0 = uses the variant-1 table) and
 * vertex_begin.  `capacity` entries at most.  For parity checks against the
 * reference's _VoxelInfo records (dmc.cu:814-1006). */
int mdmc_cells_device(mdmc_ctx *ctx, uint32_t *d_cell_index, uint8_t *d_cfg, uint8_t *d_flags,
                      uint32_t *d_vertex_begin, uint64_t capacity, void *stream);

/* ---- per-kernel timing (no counterpart in the reference, which has no
 * timers).  While enabled, every kernel the context launches is bracketed by
 * HIP events on its stream.  Kernel ids: 0 classify, 1 scan, 2 emit, 3 other. */
int mdmc_ctx_profile(mdmc_ctx *ctx, int enable);
/* Synchronises the recorded events, returns the summed milliseconds and the
 * launch count per kernel id, and clears the record. */
int mdmc_ctx_profile_read(mdmc_ctx *ctx, double ms_total[4], uint32_t launches[4]);

/* ---- batches (no counterpart in the reference, which extracts one grid per
 * call; for callers with many volumes: one launch sequence for all of them,
 * capturable in a HIP graph after a first call has sized the buffers) ---- */

#define MDMC_MAX_BATCH 1024

/* One volume of a batch.  `field` is device memory; the outputs may be NULL
 * for mdmc_batch_count_device.  Vertex and quad numbers are local to the
 * volume (each volume's mesh is exactly what mdmc_extract_device gives). */
typedef struct mdmc_volume {
  const float *field;
  uint32_t dims[3];
  float iso;
  float xyz_min[3];
  float xyz_max[3];
  float *vertices;          /* 3 * vertex_capacity floats */
  uint64_t vertex_capacity;
  int32_t *quads;           /* 4 * quad_capacity int32 */
  uint64_t quad_capacity;
} mdmc_volume;

/* Classify + scan all volumes; d_counts[v] (device, n entries) receives
 * volume v's counts.  n <= MDMC_MAX_BATCH. */
int mdmc_batch_count_device(mdmc_ctx *ctx, const mdmc_volume *vols, uint32_t n, mdmc_counts *d_counts,
                            void *stream);
/* Emit the batch last counted on ctx (same fields, dims and isovalues, in the
 * same order) into the outputs named in vols. */
int mdmc_batch_emit_device(mdmc_ctx *ctx, const mdmc_volume *vols, uint32_t n, mdmc_counts *d_counts,
                           void *stream);
/* Count + emit with the capacities given in vols (overflow is flagged per
 * volume in d_counts[v].status, never written). */
int mdmc_batch_extract_device(mdmc_ctx *ctx, const mdmc_volume *vols, uint32_t n, mdmc_counts *d_counts,
                              void *stream);

/* Synthetic fields for many volumes in one launch (bitwise the fields of
 * mdmc_field_blobs_device / mdmc_field_gyroid_device with the same
 * parameters). */
enum { MDMC_FIELD_DEVICE = 0, MDMC_FIELD_BLOBS = 1, MDMC_FIELD_GYROID = 2 };
typedef struct mdmc_gen_spec {
  float *out;        /* device, dims[0]*dims[1]*dims[2] floats, x fastest */
  uint32_t dims[3];
  int kind;          /* MDMC_FIELD_BLOBS or MDMC_FIELD_GYROID */
  uint64_t seed;
  uint32_t nblobs;   /* blobs */
  double sigma_lo, sigma_hi, amp_lo, amp_hi;
  double period, noise_amp;  /* gyroid */
} mdmc_gen_spec;
int mdmc_batch_generate_device(mdmc_ctx *ctx, const mdmc_gen_spec *specs, uint32_t n, void *stream);

/* Device scratch bytes a context needs for mdmc_count_device /
 * mdmc_extract_device on a grid of these dims, for mdmc_ctx_set_scratch (no
 * counterpart in the reference, which allocates per call; host arithmetic). */
int mdmc_scratch_bytes(const uint32_t dims[3], uint64_t *bytes);
/* Device scratch bytes mdmc_batch_count_device needs for this batch (fields
 * may be NULL here; only dims are read). */
int mdmc_batch_scratch_bytes(const mdmc_volume *vols, uint32_t n, uint64_t *bytes);
/* Make the context use caller-owned device memory as scratch (NULL: back to
 * its own allocations).  A call that needs more than `bytes` returns
 * MDMC_ERR_CAPACITY instead of allocating. */
int mdmc_ctx_set_scratch(mdmc_ctx *ctx, void *d_scratch, uint64_t bytes);

/* Tuning options of a context (0 restores the default). */
enum {
  MDMC_OPT_BATCH_PLANES = 0,           /* cell planes per batch classify tile */
  MDMC_OPT_CLASSIFY_BLOCKS_PER_CU = 1, /* persistent batch classify grid (0: a workgroup per tile) */
  MDMC_OPT_EMIT_BLOCKS_PER_CU = 2,     /* emit grid, workgroups per CU: single-volume persistent grid
                                          (default 8), batch and smoothing at most (default 128) */
  MDMC_OPT_PT_PLANES = 3,              /* cell planes per single-volume classify tile */
  MDMC_OPT_PT_BAND = 4,                /* single-volume classify: columns per XCD group */
  MDMC_OPT_PT_ROWS = 5                 /* cell rows per single-volume classify tile: 7 or 15 */
};
int mdmc_ctx_set_option(mdmc_ctx *ctx, int option, int value);

/* ---- serving (no counterpart in the reference): a request loop that admits
 * extraction jobs by their device-memory footprint, batches the admitted ones
 * (continuous batching over a few streams, so that jobs arriving while
 * others run are started as memory and streams free up) and keeps fields and
 * results resident.  A worker thread of the server launches work and notices
 * completions; every function may be called from any host thread. -------- */

typedef struct mdmc_server mdmc_server;

typedef struct mdmc_server_config {
  uint64_t memory_budget;  /* bytes of device memory for fields, scratch and results */
  uint32_t max_batch;      /* volumes per launch sequence, 1..MDMC_MAX_BATCH */
  uint32_t num_streams;    /* launch sequences in flight, 1..8 */
} mdmc_server_config;

typedef struct mdmc_request {
  uint32_t dims[3];
  float iso;
  float xyz_min[3];
  float xyz_max[3];
  int field_kind;        /* MDMC_FIELD_DEVICE (caller's field, kept alive until the
                            result is returned) or a generator kind */
  const float *field;    /* MDMC_FIELD_DEVICE */
  uint64_t seed;         /* generators (see mdmc_gen_spec) */
  uint32_t nblobs;
  double sigma_lo, sigma_hi, amp_lo, amp_hi;
  double period, noise_amp;
  uint64_t tag;          /* returned with the result */
} mdmc_request;

typedef struct mdmc_result {
  uint64_t id, tag;
  int status;            /* MDMC_OK, MDMC_ERR_REJECTED (footprint over budget), ... */
  mdmc_counts counts;
  const float *vertices; /* device memory of the server, valid until release */
  const int32_t *quads;
  double t_submit, t_start, t_done;  /* seconds on the server clock */
} mdmc_result;

typedef struct mdmc_server_stats {
  uint64_t submitted, completed, rejected, batches;
  uint64_t queued, in_flight;
  uint64_t memory_in_use, memory_peak, memory_budget;
  double busy_seconds;   /* time with at least one launch sequence in flight */
} mdmc_server_stats;

int mdmc_server_create(const mdmc_server_config *cfg, mdmc_server **out);
void mdmc_server_destroy(mdmc_server *srv);
/* Queues a request; *id identifies its result. */
int mdmc_server_submit(mdmc_server *srv, const mdmc_request *req, uint64_t *id);
/* Returns up to `max` finished results.  wait != 0: block until at least
 * one result is ready, or until the server has nothing left that it could
 * finish without another call (nothing in flight, nothing it can admit). */
int mdmc_server_poll(mdmc_server *srv, int wait, mdmc_result *out, uint32_t max, uint32_t *n);
/* Copies a finished result's mesh to host memory (sizes from its counts). */
int mdmc_server_copy_result(mdmc_server *srv, uint64_t id, float *h_vertices, int32_t *h_quads);
/* Frees a finished result's device memory. */
int mdmc_server_release(mdmc_server *srv, uint64_t id);
int mdmc_server_get_stats(mdmc_server *srv, mdmc_server_stats *st);
/* Server options, set with mdmc_server_set_option; they apply to requests
 * submitted afterwards. */
enum {
  /* requests of at least this many cells run alone, through the
   * single-volume kernels, which are faster than the batch kernels on one
   * large volume (0: every request may join a batch) */
  MDMC_SERVER_OPT_SOLO_CELLS = 100
};
/* Sets a server option (MDMC_SERVER_OPT_*), or a launch option (MDMC_OPT_*)
 * on every lane's context. */
int mdmc_server_set_option(mdmc_server *srv, int option, int value);
/* Seconds on the server clock (the clock of t_submit / t_done). */
double mdmc_server_now(mdmc_server *srv);

/* ---- synthetic fields on the device ------------------------------------- */

/* Sum of `nblobs` Gaussian blobs (centres uniform in the grid, sigma uniform
 * in [sigma_lo, sigma_hi], amplitude uniform in [amp_lo, amp_hi]) from a
 * SplitMix64 stream `seed`; in double precision, each blob's term added with
 * one fused multiply-add, rounded to float at the end. */
int mdmc_field_blobs_device(mdmc_ctx *ctx, float *d_out, const uint32_t dims[3], uint64_t seed,
                            uint32_t nblobs, double sigma_lo, double sigma_hi, double amp_lo,
                            double amp_hi, void *stream);

/* sin(wx)cos(wy)+sin(wy)cos(wz)+sin(wz)cos(wx), w = 2 pi / period, plus
 * uniform hash noise in [-noise_amp, noise_amp). */
int mdmc_field_gyroid_device(mdmc_ctx *ctx, float *d_out, const uint32_t dims[3], double period,
                             uint64_t seed, double noise_amp, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* MDMC_H */
