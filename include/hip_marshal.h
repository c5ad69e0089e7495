/*
 * hip_marshal.h - C ABI of the MI355X (gfx950) in-place layout-conversion library.
 *
 * Every entry point of the reference's OpenCL header include/cl_marshal.h:26-49 has a
 * `hip_`-prefixed twin here with the same argument meaning and the same error convention:
 *
 *   - cl_command_queue  -> hipStream_t (0 = the null stream)
 *   - cl_mem            -> void* device pointer owned by the caller
 *   - cl_ulong *        -> uint64_t * (elapsed nanoseconds, ACCUMULATED with +=)
 *   - return value      -> false on success, true on error (diagnostic on stderr)
 *
 * The plain twins move 4-byte elements, as the reference does when compiled with SP=1
 * (src/cl_marshal.cc:37-43). The `_es` variants take the element size in bytes
 * (1, 2, 4, 8 or 16; elements are opaque bit patterns) and 64-bit dimensions.
 *
 * Timing: when profiling is on (the default, like LIBMARSHAL_OCL_PROFILE in the reference's
 * src/CMakeLists.txt:2,8-10) every call synchronizes its stream: a non-NULL `elapsed_ns`
 * receives the wall time of the whole call measured with HIP events, and a NULL pointer makes
 * the call print "[Profile] <name> <ns>ns; <GB/s> GB/s" to stderr (include/cl_profile.h:43-56).
 * With profiling off (hip_marshal_set_profiling(0)) calls are fully stream-ordered and never
 * synchronize, and `elapsed_ns` is left untouched.
 *
 * Scratch: cycle-following stages need visited-bit flags, the stripe route a buffer of one
 * or more stripes and the through-scratch forms one of a group of blocks (each at most
 * "st_scratch_mb", 8 GiB by default: 2.8% of a 288 GB MI355X). All come from
 * one per-device pool owned by the library that only grows (hipMalloc and hipFree happen on
 * growth only, never per call once the pool is large enough; hip_marshal_reserve()
 * pre-sizes it). Flags are cleared with a fill kernel on the caller's stream. A stage holds
 * the pool's mutex while it enqueues, so concurrent host threads take turns; when a stage
 * is enqueued on a different stream than the pool's previous user, that stream first waits
 * (hipStreamWaitEvent) for everything the previous user had enqueued, so two calls on two
 * streams never use the pool at the same time on the GPU - their pool-using stages run one
 * after the other. Under stream capture the cross-stream wait is not recorded, so a
 * captured graph must not run concurrently with other pool users on the same device; the
 * pool does not grow under capture (reserve it first), and a pool buffer handed out under
 * capture is kept, not freed, when the pool later grows. Steps in the cluster-staged form
 * (route tag "0100_CL-") keep their barrier words in a small buffer per stream (16 KiB or
 * more, allocated on the first such step outside capture; every launch leaves it zeroed);
 * under capture they use pool scratch instead. hip_marshal_finalize frees the pool, the
 * cycle-leader lists, the segment tables and those buffers: graphs that captured library
 * calls must not be replayed after it.
 */
#ifndef LIBMARSHAL_AMD_HIP_MARSHAL_H_
#define LIBMARSHAL_AMD_HIP_MARSHAL_H_

#include <stddef.h>
#include <stdint.h>
#include <stdbool.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------------
 * Twins of include/cl_marshal.h (4-byte elements, 32-bit dimensions).
 * ------------------------------------------------------------------------------------- */

/* AoS [height][width] -> ASTA [height/tile][width][tile] with the LDS-staged kernel.
 * Replaces cl_aos_asta_bs, include/cl_marshal.h:26-27 (src/cl_marshal.cc:793-800). */
bool hip_aos_asta_bs(hipStream_t stream, void *src, int height, int width, int tile_size);

/* Same conversion by cycle following with LDS bit flags; R spreads flags over R times more
 * words, P pads one word per 32. Replaces cl_aos_asta_pttwac, include/cl_marshal.h:28-29
 * (src/cl_marshal.cc:803-812). */
bool hip_aos_asta_pttwac(hipStream_t stream, void *src, int height, int width, int tile_size,
                         int R, int P);

/* LDS-staged form, falling back to cycle following when it fails.
 * Replaces cl_aos_asta, include/cl_marshal.h:30-31 (src/cl_marshal.cc:814-818). */
bool hip_aos_asta(hipStream_t stream, void *src, int height, int width, int tile_size, int R);

/* Full in-place transpose of an (A*a) x (B*b) row-major matrix into (B*b) x (A*a).
 * stages selects the elementary-stage sequence: 2, 22, 3, 32, 4 as in the reference
 * (src/cl_marshal.cc:569-767); an infeasible choice falls back to the element-granular
 * route (src/cl_marshal.cc:769-775). stages = 0 lets the planner choose.
 * Replaces cl_transpose, include/cl_marshal.h:32-33 (src/cl_marshal.cc:561-776). */
bool hip_transpose(hipStream_t stream, void *src, int A, int a, int B, int b, int R, int stages,
                   uint64_t *elapsed_ns);

/* 010: [A][a][B] -> [A][B][a], whole tile staged in LDS.
 * Replaces cl_transpose_010_bs, include/cl_marshal.h:35-36 (src/cl_marshal.cc:136-277). */
bool hip_transpose_010_bs(hipStream_t stream, void *src, int A, int a, int B,
                          uint64_t *elapsed_ns);

/* 010 by cycle following with the tile's visited bits in LDS.
 * Replaces cl_transpose_010_pttwac, include/cl_marshal.h:37-38 (src/cl_marshal.cc:280-336). */
bool hip_transpose_010_pttwac(hipStream_t stream, void *src, int A, int a, int B,
                              uint64_t *elapsed_ns, int R, int P);

/* 100: [A][B][b] -> [B][A][b], cycle following over super-elements of b elements with
 * global visited bits. Replaces cl_transpose_100, include/cl_marshal.h:40-41
 * (src/cl_marshal.cc:779-782 -> 340-558). */
bool hip_transpose_100(hipStream_t stream, void *src, int A, int B, int b, uint64_t *elapsed_ns);

/* 0100: [A][a][B][b] -> [A][B][a][b] (A independent 100 instances).
 * Replaces cl_transpose_0100, include/cl_marshal.h:43-44 (src/cl_marshal.cc:785-788). */
bool hip_transpose_0100(hipStream_t stream, void *src, int A, int a, int B, int b,
                        uint64_t *elapsed_ns);

/* In-place row padding [y_size][x_size] -> [y_size][pad_size] (new columns zeroed) and its
 * inverse [y_size][pad_size] -> [y_size][x_size] (tail zeroed). The buffer must hold
 * y_size*pad_size elements. Replace cl_padding / cl_unpadding, include/cl_marshal.h:46-49
 * (src/cl_marshal.cc:821-956, src/cl/cl_aos_asta.cl:813-938). */
bool hip_padding(hipStream_t stream, void *src, int x_size, int y_size, int pad_size,
                 uint64_t *elapsed_ns);
bool hip_unpadding(hipStream_t stream, void *src, int x_size, int y_size, int pad_size,
                   uint64_t *elapsed_ns);

/* Releases the scratch pools, cycle-leader lists and segment tables of every device and
 * forgets measured choices (autotuned shapes and classes). Replaces cl_marshal_finalize,
 * src/cl_marshal.cc:116-119 (exported there but missing from its header). */
void hip_marshal_finalize(void);

/* ---------------------------------------------------------------------------------------
 * Element-size-generic variants (elem_bytes in {1,2,4,8,16}, 64-bit dimensions).
 * Same semantics as the twins above; no counterpart in the reference, whose element type
 * is fixed at compile time (src/cl_marshal.cc:37-43).
 * ------------------------------------------------------------------------------------- */
bool hip_transpose_010_bs_es(hipStream_t stream, void *src, size_t elem_bytes, int64_t A,
                             int64_t a, int64_t B, uint64_t *elapsed_ns);
bool hip_transpose_010_pttwac_es(hipStream_t stream, void *src, size_t elem_bytes, int64_t A,
                                 int64_t a, int64_t B, uint64_t *elapsed_ns, int R, int P);
/* 100 / 0100 pick their form: a single-instance 100 over 1-, 2-, 4-, 8- or 16-byte
 * super-elements (buffer aligned to that size) is planned as a full transpose; otherwise,
 * within the scratch cap, groups of instances go through scratch (a copy and an out-of-place
 * gather); else the in-place forms (row/column decomposition, cycle following). The
 * AoS/SoA conversions below (method 0) choose the same way; hip_marshal_last_route names
 * the form taken. */
bool hip_transpose_100_es(hipStream_t stream, void *src, size_t elem_bytes, int64_t A,
                          int64_t B, int64_t b, uint64_t *elapsed_ns);
bool hip_transpose_0100_es(hipStream_t stream, void *src, size_t elem_bytes, int64_t A,
                           int64_t a, int64_t B, int64_t b, uint64_t *elapsed_ns);
bool hip_transpose_es(hipStream_t stream, void *src, size_t elem_bytes, int64_t A, int64_t a,
                      int64_t B, int64_t b, int R, int stages, uint64_t *elapsed_ns);
/* Full transpose of an M x N matrix; the planner picks the tile factors and the route. */
bool hip_transpose_full_es(hipStream_t stream, void *src, size_t elem_bytes, int64_t M,
                           int64_t N, uint64_t *elapsed_ns);
/* AoS [height][width] -> ASTA [height/tile][width][tile] (010) and back (ASTA -> AoS is
 * 010 with a and B exchanged). method: 0 = LDS-staged when the tile fits LDS, else through
 * scratch within the scratch cap, else split / row-column / cycle following in place;
 * 1 = LDS-staged only, 2 = cycle following only. */
bool hip_aos_asta_es(hipStream_t stream, void *src, size_t elem_bytes, int64_t height,
                     int64_t width, int64_t tile_size, int method, uint64_t *elapsed_ns);
bool hip_asta_aos_es(hipStream_t stream, void *src, size_t elem_bytes, int64_t height,
                     int64_t width, int64_t tile_size, int method, uint64_t *elapsed_ns);
/* SoA [width][height] -> ASTA [height/tile][width][tile] is 100(width, height/tile, tile)
 * (test/cl_unittest.cc:200); ASTA -> SoA is 100(height/tile, width, tile). */
bool hip_soa_asta_es(hipStream_t stream, void *src, size_t elem_bytes, int64_t height,
                     int64_t width, int64_t tile_size, uint64_t *elapsed_ns);
bool hip_asta_soa_es(hipStream_t stream, void *src, size_t elem_bytes, int64_t height,
                     int64_t width, int64_t tile_size, uint64_t *elapsed_ns);
bool hip_padding_es(hipStream_t stream, void *src, size_t elem_bytes, int64_t x_size,
                    int64_t y_size, int64_t pad_size, uint64_t *elapsed_ns);
bool hip_unpadding_es(hipStream_t stream, void *src, size_t elem_bytes, int64_t x_size,
                      int64_t y_size, int64_t pad_size, uint64_t *elapsed_ns);

/* ---------------------------------------------------------------------------------------
 * No-allocation variant. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same result as hip_transpose_0100_es (A instances of [a][B][b] ->
 * [B][a][b], elem_bytes in {1,2,4,8,16}) but restricted to the in-place forms (LDS-staged,
 * row/column shuffles, LDS bits, global bits) and with the visited bits of the global-bit form
 * in a fixed device array of HIP_MARSHAL_NOALLOC_SCRATCH_BYTES (32 MiB) that each device holds
 * from the first call on it until the process ends. A step whose bits exceed that is rejected
 * with an error (true) before anything runs. The call allocates no device memory, creates no
 * events and does not time itself (no elapsed argument); it is stream-ordered and never
 * synchronizes. Every call on one device shares that device's array and clears the bits it
 * uses before its kernels run: calls that overlap on one device (concurrent streams, or host
 * threads) race and must be serialized by the caller.
 * ------------------------------------------------------------------------------------- */
#define HIP_MARSHAL_NOALLOC_SCRATCH_BYTES (32u << 20)
bool hip_transpose_0100_noalloc_es(hipStream_t stream, void *src, size_t elem_bytes, int64_t A,
                                   int64_t a, int64_t B, int64_t b);

/* ---------------------------------------------------------------------------------------
 * Batched form: many conversions of uneven shapes and element sizes in one graph launch.
 * No counterpart in the reference (it launches one kernel per stage per matrix).
 * hip_marshal_batch_create plans every descriptor (the planner picks tile factors, routes and
 * kernel forms), groups the stages of all matrices by level and kernel form, and uploads the
 * work lists; it returns NULL on an invalid descriptor. hip_marshal_batch_run converts every
 * buffer in place: with use_graph the sequence (bit clearing + one persistent kernel per
 * group and level) is built into a hipGraph on the first run, node by node (the library
 * captures no stream, so other threads may use the device meanwhile), and launched
 * afterwards; without it the same kernels are launched one by one. Matrices whose best route
 * has no grouped form (the stripe route, and from "batch_solo_kb" the larger in-place and
 * through-scratch forms) run their plain stage lists on side streams ("batch_solo_lanes",
 * balanced by bytes) forked from the run's stream as the graph starts and joined back to it.
 * Each buffer ends bit-identical to what the plain entry point gives for it alone. Buffers
 * must not overlap; a batch's own visited-bit space and scratch are private to it, so
 * different batches may run concurrently on different streams, from any threads, but one
 * batch must not run twice at once.
 * ------------------------------------------------------------------------------------- */
enum {
  HIP_MARSHAL_OP_TRANSPOSE = 0, /* rows x cols -> cols x rows                     */
  HIP_MARSHAL_OP_AOS_ASTA = 1,  /* [rows][cols] -> [rows/tile][cols][tile]        */
  HIP_MARSHAL_OP_SOA_ASTA = 2,  /* [cols][rows] -> [rows/tile][cols][tile]        */
  HIP_MARSHAL_OP_ASTA_AOS = 3,  /* inverse of AOS_ASTA                            */
  HIP_MARSHAL_OP_ASTA_SOA = 4   /* inverse of SOA_ASTA                            */
};
typedef struct hip_marshal_desc {
  void *ptr;          /* device buffer, naturally aligned to min(elem_bytes, 16) */
  int32_t op;         /* HIP_MARSHAL_OP_* */
  int32_t elem_bytes; /* 1, 2, 4, 8 or 16 */
  int64_t rows;       /* transpose: M; AoS/SoA/ASTA: height (number of structures) */
  int64_t cols;       /* transpose: N; AoS/SoA/ASTA: width (fields per structure) */
  int64_t tile;       /* ASTA tile width (divides rows); ignored by TRANSPOSE */
} hip_marshal_desc;
typedef struct hip_marshal_batch hip_marshal_batch;
hip_marshal_batch *hip_marshal_batch_create(const hip_marshal_desc *descs, int64_t count,
                                            int use_graph);
bool hip_marshal_batch_run(hip_marshal_batch *batch, hipStream_t stream, uint64_t *elapsed_ns);
/* "levels", "stages", "kernels", "flag_bytes", "bytes" (algorithmic bytes of one run),
 * "solo" (matrices run with their plain stage lists on the side streams), "arena_matrices",
 * "arena_rounds", "arena_bytes" (the through-scratch rounds). */
int64_t hip_marshal_batch_info(const hip_marshal_batch *batch, const char *key);
const char *hip_marshal_batch_route(const hip_marshal_batch *batch, int64_t index);
void hip_marshal_batch_destroy(hip_marshal_batch *batch);

/* ---------------------------------------------------------------------------------------
 * Serving: a request loop with continuous batching. No counterpart in the reference.
 * hip_marshal_server_create starts a scheduler thread that owns `streams` HIP streams.
 * Submitted requests convert device-resident buffers in place. Each is planned when it is
 * submitted, its scratch capped by the knob "serve_scratch_mb" (8192 by default; 0 keeps every
 * request on the in-place routes, whose scratch is visited bits or one stripe). Requests are
 * admitted in arrival order while the device memory they would hold fits hbm_budget_bytes:
 * the buffers of admitted, unfinished requests plus a bound on their visited bits, plus every
 * scratch buffer the server holds (the one shared by the large request that runs at a time,
 * each stream's own), counting the growth an admission needs. They are packed into dynamic
 * batches of at most max_batch requests and launched
 * on an idle stream, so requests arriving while others run are admitted as budget and
 * streams free up and run concurrently with them. A request whose buffer and scratch exceed
 * the budget is rejected (-1). The caller keeps each buffer alive and untouched until its request is done
 * and must not submit overlapping buffers concurrently. Results are bit-identical to the
 * plain entry points run on each buffer alone.
 * ------------------------------------------------------------------------------------- */
typedef struct hip_marshal_server hip_marshal_server;
typedef struct hip_marshal_server_stats {
  int64_t completed, failed, batches, max_batch;
  double p50_ms, p99_ms, mean_ms, max_ms; /* arrival -> completion */
  double bytes;                           /* algorithmic bytes moved (2 x each buffer) */
  double span_s;                          /* first arrival -> last completion */
} hip_marshal_server_stats;
hip_marshal_server *hip_marshal_server_create(size_t hbm_budget_bytes, int streams,
                                              int max_batch);
/* Submission with the arrival time on hip_marshal_server_now()'s clock (0 = now). Returns
 * the request id, or -1 for an invalid request. */
int64_t hip_marshal_server_submit(hip_marshal_server *server, const hip_marshal_desc *request);
int64_t hip_marshal_server_submit_at(hip_marshal_server *server,
                                     const hip_marshal_desc *request, double arrival_s);
double hip_marshal_server_now(void);
/* 0 queued, 1 running, 2 done, -1 failed or unknown. */
int hip_marshal_server_status(hip_marshal_server *server, int64_t id);
/* false once the request is done; true on failure or timeout. */
bool hip_marshal_server_wait(hip_marshal_server *server, int64_t id, double timeout_s);
bool hip_marshal_server_drain(hip_marshal_server *server, double timeout_s);
bool hip_marshal_server_request_times(hip_marshal_server *server, int64_t id, double *arrival,
                                      double *launch, double *done);
void hip_marshal_server_get_stats(hip_marshal_server *server, hip_marshal_server_stats *out);
void hip_marshal_server_reset_stats(hip_marshal_server *server);
/* Waits for every submitted request, then releases the streams and buffers. */
void hip_marshal_server_destroy(hip_marshal_server *server);

/* ---------------------------------------------------------------------------------------
 * Library configuration and introspection (no counterpart in the reference, where these
 * are compile-time macros: MAX_MEM src/cl_marshal.cc:45-59, LIBMARSHAL_OCL_PROFILE).
 * ------------------------------------------------------------------------------------- */
/* LDS bytes a stage may use per workgroup (default: the device's per-workgroup maximum,
 * 160 KiB on gfx950). Lowering it forces cycle-following routes; 0 restores the default. */
void hip_marshal_set_lds_budget(size_t bytes);
size_t hip_marshal_get_lds_budget(void);
/* 1 = time calls with events and synchronize (default), 0 = asynchronous. */
void hip_marshal_set_profiling(int on);
/* Cluster-staged steps (route tags 0100_CL / 0100_CLF) hold an instance in a few
 * workgroups that wait for each other at a barrier. The wait is bounded ("cl_spin"): a
 * cluster whose workgroups cannot all become resident in time (other work holding the CUs)
 * gives up and leaves its instance unconverted. Calls with profiling on (they synchronize)
 * then return true with a message; asynchronous callers, batches and servers ask here.
 * Waits for the current device, returns the clusters that gave up on it since the last
 * check (0 normally). The count is per device, not per call: with several threads on one
 * device, a give-up is reported by whichever check (this one or a synchronizing call's)
 * comes next. No counterpart in the reference. */
uint64_t hip_marshal_cluster_aborts(void);
/* Launch-shape knobs: "bs_upt" (4, 8, 16 units per lane per round of the LDS-staged 010),
 * "bs_blocks_per_cu", "pt010_blocks_per_cu", "pt100_blocks_per_cu", "rc_row_blocks_per_cu",
 * "rc_col_blocks_per_cu" (grid caps), "pt100_vw_max" (lanes per cycle-following walker),
 * "lds0100_min_inst", "rc_max_super_bytes" (route thresholds), and for the stripe route
 * "st_super_bytes" (target super-element bytes), "st_max_height" (at most this many
 * elements per super-element), "st_scratch_mb" (scratch cap of the
 * stripe and through-scratch forms; 0 switches them off), "gather_tile_kb" (LDS tile of the
 * out-of-place super-element gather),
 * "st_tile_kb" (LDS tile of the out-of-place transpose), "oop_elem_tile_kb" (its tile when
 * it stores single elements into tiles spanning a whole short side; 0 = "st_tile_kb"),
 * "oop_pitched" (through scratch with misaligned destination rows: 0 copy first, 2 transpose
 * first into 16-byte-aligned scratch rows and copy back, 1 (default) the latter for 1- and
 * 2-byte elements and from 512 MiB), "st_pt_vw", "st_rotate" (0/1), "copy_blocks_per_cu"
 * (grid cap of the 16-byte copy), "oop_min_tiles" (calls that have the device to
 * themselves: through-scratch tiles shrink, down to 4 KiB, until a group makes at least
 * this many; 0 = no minimum),
 * "oop_min_rows" (at least this many tile rows in the
 * out-of-place transpose of 1- to 4-byte elements; 0 = no minimum),
 * "copy_nt" (0/1: non-temporal source loads in the copy into scratch), "copy_nt_store_mb"
 * (copies of at least this many MiB store non-temporally; default 256), "oop_nt" (bit 0:
 * non-temporal loads, bit 1: non-temporal stores in the out-of-place transpose; default 2),
 * "oop_order" (its tile order when tiles are at least as wide as tall: 0 row-major, k >= 1
 * tile columns in groups of k walked row by row; default 2),
 * "pad_nt" (0/1: non-temporal loads and stores in the vector padding kernel),
 * "pl_min_bytes", "pl_max_order", "pl_min_walkers" (when a 100 / 0100 step takes the
 * cycle-leader form: super-elements of at least this many bytes, an order of R modulo R*C - 1
 * at most this, at least this many walkers),
 * "st_wstore" (0/1: window stores for every element size), "st_skew" (0 off, 1 for 1- and
 * 2-byte elements (default), 2 up to 8-byte ones: skewed tiles for misaligned destination
 * rows), "gather_row_bytes" (destination rows shorter than this and misaligned take the
 * gather), "st_tile_rows" (0 = no cap on tile rows), "pad_shape" (0-4: workgroup shape of
 * the vector padding kernel), "autotune"
 * (0/1: measure each new full-transpose shape on its first call) and "batch_solo_kb"
 * (matrices of at least this many KiB whose route is the square or through-scratch one run
 * on the batch's side streams instead of in a grouped kernel; below 1 MiB only in batches
 * of at least 1 GiB in all), "batch_solo_lanes" (how many side streams, 1-8, default 4),
 * and for the segmented cycle walk of 100 / 0100 steps (route tag "0100_SEG-": every cycle cut
 * into segments walked in parallel, one pass, a side buffer of each segment's start)
 * "seg_len" (slots per segment, default 16), "seg_min_bytes" (super-elements of at least this
 * many bytes, default 64), "seg_max_slots" (grids of at most this many slots, default 2^24),
 * "seg_blocks_per_cu", "seg_nt" (bit 0 / 1: non-temporal loads / stores, default 0),
 * "seg_xcd" (0/1: keep each XCD on a contiguous run of segments, default 1) and "seg_force"
 * (0/1: take it wherever it can run, default 0); for the cluster-staged form of 100 / 0100
 * steps (route tag "0100_CL-": a few workgroups hold each instance in LDS and registers, meet
 * at a barrier and write it back transposed, one pass, for super-elements of at least 64
 * bytes) "cl_max_wgs" (workgroups per instance, at most 16; 0 switches the form off),
 * "cl_min_bytes" (smallest instance, default 256 KiB), "cl_min_inst" (fewest instances,
 * default 64) and "cl_spin" (bound of the barrier wait in polls; a cluster that reaches it
 * leaves its instance unchanged).
 * Replaces the reference's hard-coded work-group tables (src/cl_marshal.cc:165-231,
 * 388-465). The setter returns true on an unknown key or a value outside [min, 2^30] (min is
 * 0 for "st_scratch_mb", "st_rotate", "st_wstore", "st_skew", "copy_nt", "oop_nt", "pad_nt",
 * "oop_elem_tile_kb", "oop_pitched", "oop_min_rows", "oop_min_tiles", "copy_nt_store_mb",
 * "gather_row_bytes", "st_tile_rows", "pad_shape", "autotune", "seg_nt", "seg_xcd",
 * "seg_force", "oop_order", "cl_max_wgs", "cl_min_bytes" and "cl_min_inst", 1 otherwise); the
 * getter
 * returns 0 for an unknown key. */
bool hip_marshal_set_tuning(const char *key, int64_t value);
/* Measured autotuner for an M x N transpose of elem_bytes elements: the planner's cheapest
 * candidate routes and the best of each route family are run on `buf` (on `stream`), each
 * timed and followed by the inverse transpose, so `buf` ends bit-identical to how it started
 * (if a candidate fails part-way, a device snapshot taken first, when the buffer fits the
 * scratch cap, is copied back, and true is returned).
 * The fastest becomes the plan for this shape (hip_transpose_full_es, batches and
 * hip_marshal_plan_transpose use it from then on); *stages receives its stage code. With the
 * knob "autotune" set to 1, hip_transpose_full_es does this by itself on the first call for
 * each shape. Synchronizes the stream; must not be called during graph capture. No
 * counterpart in the reference (its test picks tiles with a fixed heuristic,
 * test/cl_unittest.cc:396-495). */
bool hip_marshal_autotune(hipStream_t stream, void *buf, size_t elem_bytes, int64_t M,
                          int64_t N, int *stages);
/* Measured autotuner per shape class. For an AoS/SoA <-> ASTA descriptor, every form its step
 * can take (the 010 step: LDS-staged with three grid sizes, through scratch, split, row/column,
 * cycle following; the 100 step: cycle leaders, the segmented walk with three segment
 * lengths, through scratch, row/column, LDS bits, global bits) is run on `desc->ptr` and timed,
 * each followed by the inverse step, so the buffer ends bit-identical; the fastest becomes the
 * choice for its class (step kind, element or super-element size, power-of-two buckets of the
 * dimensions and of the bytes), used by every later plan in the class: plain calls, batches,
 * the server. *form receives the chosen form code. A transpose descriptor (or a 100 over 1-16
 * byte super-elements) is tuned as by hip_marshal_autotune for its exact shape, and *form
 * receives the stage code. The buffer is first copied to a device snapshot (when it fits the
 * scratch cap); if a candidate fails, the snapshot is copied back and true is returned.
 * Synchronizes the stream; not under graph capture. With the knob "autotune" set,
 * hip_marshal_batch_create tunes each class of its AoS/SoA <-> ASTA descriptors (once per
 * class) this way. */
bool hip_marshal_autotune_desc(hipStream_t stream, const hip_marshal_desc *desc, int *form);
int64_t hip_marshal_get_tuning(const char *key);
/* The planner's choice for an M x N transpose of elem_bytes elements: tile factors a | M,
 * b | N and the stage sequence (2, 22, 3, 32; 5 = row/column shuffles, 6/7 = stripes of
 * rows/columns through bounded scratch, 8 = square tile-pair swaps, 9 = through scratch;
 * -1 = element-granular fallback). Host-only
 * (no device work). Replaces the tile heuristic of the reference's test
 * (test/cl_unittest.cc:396-495, "Alg. 5" and "Alg. 3"). */
bool hip_marshal_plan_transpose(int64_t M, int64_t N, size_t elem_bytes, int64_t *a,
                                int64_t *b, int *stages);
/* Route tag of the last transpose, 010 / 0100 / 100 or AoS / SoA call on this thread
 * ("3.1-010_BS", "010_OOP-", "0100_FULL:SQ-" ...). */
const char *hip_marshal_last_route(void);
/* Grows the current device's flag pool to at least `bytes` (false on success). */
bool hip_marshal_reserve(size_t bytes);
/* Cycles of length >= 2 of the in-place transpose of an R x C grid (the 100 map), counted
 * from gcds without walking them, when the order of R modulo R*C - 1 (which every cycle length
 * divides; stored in *order) is at most max_order and R*C < 2^32; else 0. These are the
 * walkers of the claim-free cycle-leader form of 100 / 0100 (route tag "0100_CYC-"). No
 * counterpart in the reference, which counts cycles by walking them
 * (Transposition::GetNumCycles, include/plan.hpp:57-73). Host only. */
uint64_t hip_marshal_cycle_count(int64_t R, int64_t C, uint32_t max_order, uint32_t *order);
/* Bytes of visited-bit flags the given stage needs (0 for LDS-staged stages). */
size_t hip_marshal_flag_bytes_0100(int64_t A, int64_t a, int64_t B, int64_t b,
                                   size_t elem_bytes);
/* Per-launch kernel timing. With timing on, every kernel launch of the library is bracketed by
 * a pair of HIP events on the stream it runs on (nothing is recorded under stream capture).
 * hip_marshal_kernel_times waits for the recorded launches, writes one entry per kernel family
 * (in order of first launch, at most max_entries) with the launch count, the summed device
 * time and the summed algorithmic bytes (2 x the bytes each launch moves into place), clears
 * the record and returns the number of families. Turning timing on or off also clears it.
 * No counterpart in the reference (its Profiling class times one kernel per call,
 * include/cl_profile.h:23-56). */
typedef struct hip_marshal_kernel_time {
  char name[48];
  int64_t launches;
  double ms;
  double bytes;
} hip_marshal_kernel_time;
void hip_marshal_set_kernel_timing(int on);
int hip_marshal_kernel_times(hip_marshal_kernel_time *out, int max_entries);

/* Library version string. */
const char *hip_marshal_version(void);

#ifdef __cplusplus
}
#endif

#endif /* LIBMARSHAL_AMD_HIP_MARSHAL_H_ */
