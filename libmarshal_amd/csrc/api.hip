// api.hip - C ABI, stage composer and planner of the gfx950 layout-conversion library.
//
// Maps the reference's host layer (src/cl_marshal.cc, src/marshal.cu, include/plan.hpp) onto
// HIP: stages are enqueued on the caller's stream; visited-bit flags come from a per-device
// pool and are cleared by a fill kernel on the same stream (the reference allocates them per
// call and uploads zeros from the host, src/cl_marshal.cc:353-361).
#include "kernels.h"
#include "../../include/hip_marshal.h"
#include "../../include/marshal.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <array>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

namespace lm {

// ---------------------------------------------------------------------------------------
// Device information and configuration
// ---------------------------------------------------------------------------------------
namespace {
std::mutex g_dev_mu;
DeviceInfo g_dev[64];
std::atomic<size_t> g_lds_budget{0};  // 0 = device maximum
std::atomic<int> g_profiling{1};
thread_local std::string t_route;
thread_local int t_restricted = 0;
thread_local uint64_t t_scratch_cap = UINT64_MAX;  // ScratchCapScope override
}  // namespace

ScratchCapScope::ScratchCapScope(uint64_t bytes) : prev_(t_scratch_cap) {
  t_scratch_cap = std::min(t_scratch_cap, bytes);
}
ScratchCapScope::~ScratchCapScope() { t_scratch_cap = prev_; }
// Scratch cap of the through-scratch and stripe forms for stages planned on this thread.
static uint64_t scratch_cap_bytes() {
  return std::min<uint64_t>(uint64_t(tuning("st_scratch_mb")) << 20, t_scratch_cap);
}

RestrictedPlanning::RestrictedPlanning() { t_restricted++; }
RestrictedPlanning::~RestrictedPlanning() { t_restricted--; }
bool planning_restricted() { return t_restricted > 0; }

const DeviceInfo &device_info() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  std::lock_guard<std::mutex> lk(g_dev_mu);
  DeviceInfo &d = g_dev[dev];
  if (d.id < 0) {
    d.id = dev;
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        v > 0)
      d.cus = v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) ==
            hipSuccess &&
        v > 0)
      d.lds_per_block = size_t(v);
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) == hipSuccess) d.hbm_bytes = total;
  }
  return d;
}

static size_t lds_budget() {
  size_t dev = device_info().lds_per_block;
  size_t b = g_lds_budget.load();
  return (b == 0 || b > dev) ? dev : b;
}

// ---------------------------------------------------------------------------------------
// Launch-shape knobs. Defaults come from measurements on MI355X; hip_marshal_set_tuning
// overrides them (the reference hard-codes its equivalents per vendor,
// src/cl_marshal.cc:165-231 and 388-465).
// ---------------------------------------------------------------------------------------
namespace {
struct Knob {
  const char *key;
  std::atomic<int64_t> value;
  int64_t min = 1;  // smallest accepted value (0 for on/off switches)
};
Knob g_knobs[] = {
    {"bs_upt", {16}},
    {"bs_blocks_per_cu", {192}},
    {"pt010_blocks_per_cu", {8}},
    {"pt010_wave", {1}, 0},      // 1: the 010 cycle-following form owns one tile per wave
    {"pt010_lead_order", {64}, 0},  // its leader-only form when every cycle is this short
    {"pt010_lead_blocks_per_cu", {256}},
    {"pt010_lead_cached", {2}, 0},  // its leaders: 0 searched per workgroup, 1 the cached list
                                    // of the a x B grid, 2 that list ordered for few sectors
    {"pt100_vw_max", {8}},
    {"pt100_blocks_per_cu", {32}},
    {"lds0100_min_inst", {32}},
    {"rc_row_blocks_per_cu", {2}},
    {"rc_max_super_bytes", {64}},
    {"rc_col_blocks_per_cu", {1}},
    {"st_scratch_mb", {8192}, 0},  // 0: no routes that need scratch beyond visited bits
    {"st_super_bytes", {8192}},
    {"st_max_height", {4096}},
    {"st_tile_kb", {32}},
    {"oop_elem_tile_kb", {8}, 0},  // 0: element-store tiles take st_tile_kb too
    {"oop_pitched", {1}, 0},
    {"oop_min_rows", {256}, 0},
    {"oop_min_tiles", {512}, 0},
    {"st_pt_vw", {16}},
    {"st_rotate", {1}, 0},
    {"autotune", {0}, 0},
    {"pad_shape", {0}, 0},
    {"st_wstore", {0}, 0},
    {"st_tile_rows", {0}, 0},
    {"gather_tile_kb", {24}},
    {"gather_row_bytes", {256}, 0},
    {"st_skew", {1}, 0},
    {"copy_nt", {1}, 0},
    {"copy_blocks_per_cu", {64}},
    {"copy_nt_store_mb", {256}, 0},  // 0: non-temporal stores for every copy
    {"oop_nt", {2}, 0},
    {"cl_max_wgs", {16}, 0},  // workgroups per instance of the cluster-staged 100 (<= 16; 0: off)
    {"cl_min_bytes", {256 << 10}, 0},  // its smallest instance
    {"cl_min_inst", {64}, 0},  // and fewest instances
    {"cl_fuse", {1}, 0},  // a 010 then a 0100 over its tiles as one cluster-staged pass
    {"cl_spin", {1 << 22}},  // bound of its barrier wait, in polls (~2 s)
    {"oop_order", {2}, 0},  // tile order of the out-of-place transpose: 0 row-major, k: groups of k tile columns
    {"pad_nt", {1}, 0},
    {"pl_min_bytes", {128}},
    {"pl_max_order", {64}},
    {"pl_min_walkers", {16384}},
    {"batch_solo_kb", {64}},
    {"batch_solo_lanes", {4}},
    {"batch_arena_mb", {1024}, 0},    // a batch's through-scratch round (0: no rounds)
    {"batch_arena_max_mb", {16}, 0},  // largest matrix taken into them
    {"batch_level_lanes", {4}},       // branch streams for the groups of one level (1: none)
    {"serve_scratch_mb", {8192}, 0},  // per-request scratch cap of server requests
    {"seg_len", {16}},           // slots per segment of the segmented cycle walk
    {"seg_min_bytes", {64}},     // super-elements of at least this many bytes take it
    {"seg_max_slots", {1 << 24}},  // grids of at most this many slots (host-built tables)
    {"seg_blocks_per_cu", {64}},
    {"seg_nt", {0}, 0},          // bit 0: non-temporal loads, bit 1: stores, in the walk
    {"seg_xcd", {1}, 0},         // 1: each XCD takes a contiguous run of walkers
    {"seg_force", {0}, 0},       // 1: take the segmented walk wherever it can run
    {"seg_sync_slots", {1 << 20}},  // larger grids get their tables built in the background
};
Knob *find_knob(const char *key) {
  if (!key) return nullptr;
  for (Knob &k : g_knobs)
    if (strcmp(k.key, key) == 0) return &k;
  return nullptr;
}
}  // namespace

int64_t tuning(const char *key) {
  Knob *k = find_knob(key);
  return k ? k->value.load() : 0;
}

// ---------------------------------------------------------------------------------------
// Flag pool: one growable device buffer per device. A call holds the pool's mutex while it
// enqueues; use by a different stream than the previous user is ordered after that user's
// last enqueued work with an event, so flags are never cleared under a running stage.
// ---------------------------------------------------------------------------------------
namespace {
struct Pool {
  std::mutex mu;
  void *ptr = nullptr;
  size_t cap = 0;
  hipStream_t last = nullptr;
  hipEvent_t ev = nullptr;
  bool has_ev = false;
  // Handed out under stream capture: a captured graph may hold the address, so growth keeps
  // the old buffer (in `retired`) until hip_marshal_finalize instead of freeing it.
  bool captured = false;
  std::vector<void *> retired;
};
Pool g_pool[64];

Pool &pool_here() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  return g_pool[dev];
}

bool pool_grow(Pool &p, size_t bytes) {
  if (p.cap >= bytes) return false;
  if (p.ptr) {
    if (p.has_ev) LM_CHECK_HIP(hipEventSynchronize(p.ev));
    if (p.captured) {
      p.retired.push_back(p.ptr);
      p.captured = false;
    } else {
      LM_CHECK_HIP(hipFree(p.ptr));
    }
    p.ptr = nullptr;
    p.cap = 0;
  }
  size_t want = std::max<size_t>(bytes + bytes / 4, size_t(4) << 20);
  want = (want + 4095) & ~size_t(4095);
  LM_CHECK_HIP(hipMalloc(&p.ptr, want));
  p.cap = want;
  return false;
}

// Cycle-leader lists by (device, R, C): a list depends only on the grid, so a shape that
// recurs skips the search. Entries live until hip_marshal_finalize, up to kLeaderCacheBytes
// in all. An entry is created only outside stream capture; later users wait for the event
// that follows its search (or, while capturing, use it only once that search has finished).
struct LeaderEntry {
  uint32_t *list = nullptr;
  hipEvent_t ready = nullptr;
  size_t bytes = 0;
  std::vector<uint32_t> host;  // an ordered list's source of its upload (kept until finalize)
  // set by a lookup outside capture that finds the search finished: under capture no event may
  // be queried (that would invalidate the capture), so only such an entry is used there
  bool done = false;
};
std::mutex g_lc_mu;
// key: (device, R, C, 0) for a searched list, (device, a, B, es) for an ordered 010 list
std::map<std::tuple<int, uint64_t, uint64_t, uint64_t>, LeaderEntry> g_lc;
size_t g_lc_bytes = 0;
constexpr size_t kLeaderCacheBytes = size_t(256) << 20;

}  // namespace

const uint32_t *cached_leaders(uint64_t R, uint64_t C, const PlPlan &pc, hipStream_t st) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  (void)hipStreamIsCapturing(st, &cs);
  const bool capturing = cs != hipStreamCaptureStatusNone;
  std::lock_guard<std::mutex> lk(g_lc_mu);
  const auto key = std::make_tuple(dev, R, C, uint64_t(0));
  auto it = g_lc.find(key);
  if (it != g_lc.end()) {
    if (capturing) return it->second.done ? it->second.list : nullptr;
    if (!it->second.done && hipEventQuery(it->second.ready) == hipSuccess) it->second.done = true;
    if (!it->second.done && hipStreamWaitEvent(st, it->second.ready, 0) != hipSuccess)
      return nullptr;
    return it->second.list;
  }
  const size_t bytes = size_t(pc.cycles + 1) * 4;
  if (capturing || g_lc_bytes + bytes > kLeaderCacheBytes) return nullptr;
  LeaderEntry e;
  e.bytes = bytes;
  if (hipMalloc(reinterpret_cast<void **>(&e.list), bytes) != hipSuccess) return nullptr;
  if (hipEventCreateWithFlags(&e.ready, hipEventDisableTiming) != hipSuccess ||
      launch_find_leaders(e.list, R, C, pc, st) != hipSuccess ||
      hipEventRecord(e.ready, st) != hipSuccess) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(e.list);
    if (e.ready) (void)hipEventDestroy(e.ready);
    return nullptr;
  }
  g_lc[key] = e;
  g_lc_bytes += bytes;
  return e.list;
}

const uint32_t *cached_ordered_leaders(uint64_t a, uint64_t B, size_t es, const PlPlan &pc,
                                       hipStream_t st, uint32_t *entries) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || a > 0xffffffffu || B > 0xffffffffu) return nullptr;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  (void)hipStreamIsCapturing(st, &cs);
  const bool capturing = cs != hipStreamCaptureStatusNone;
  const auto key = std::make_tuple(dev, a, B, uint64_t(es));
  // A known entry: ordered list (or a shape remembered as not ordered).
  auto lookup = [&](LeaderEntry &f) -> const uint32_t * {
    if (!f.list) return nullptr;
    *entries = uint32_t(f.bytes / 4);
    if (capturing) return f.done ? f.list : nullptr;
    if (!f.done && hipEventQuery(f.ready) == hipSuccess) f.done = true;
    if (!f.done && hipStreamWaitEvent(st, f.ready, 0) != hipSuccess) return nullptr;
    return f.list;
  };
  {
    std::lock_guard<std::mutex> lk(g_lc_mu);
    auto it = g_lc.find(key);
    if (it != g_lc.end()) return lookup(it->second);
  }
  if (capturing) return nullptr;
  // ordered outside the lock (milliseconds for large tiles); a thread that got there first wins
  std::vector<uint32_t> host = order_010_leaders(uint32_t(a), uint32_t(B), es, pc.cycles);
  std::lock_guard<std::mutex> lk(g_lc_mu);
  auto it = g_lc.find(key);
  if (it != g_lc.end()) return lookup(it->second);
  const size_t bytes = host.size() * 4;
  if (host.empty() || g_lc_bytes + bytes > kLeaderCacheBytes) {
    g_lc[key] = LeaderEntry{};  // remembered: not ordered again until hip_marshal_finalize
    return nullptr;
  }
  LeaderEntry e;
  e.host = std::move(host);
  e.bytes = bytes;
  if (hipMalloc(reinterpret_cast<void **>(&e.list), bytes) != hipSuccess) return nullptr;
  if (hipEventCreateWithFlags(&e.ready, hipEventDisableTiming) != hipSuccess ||
      hipMemcpyAsync(e.list, e.host.data(), bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipEventRecord(e.ready, st) != hipSuccess) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(e.list);
    if (e.ready) (void)hipEventDestroy(e.ready);
    return nullptr;
  }
  *entries = uint32_t(e.host.size());
  const uint32_t *list = e.list;
  g_lc_bytes += bytes;
  g_lc[key] = std::move(e);
  return list;
}

namespace {

struct Lease {
  Pool &p;
  std::unique_lock<std::mutex> lk;
  hipStream_t st;
  bool ok = false;
  Lease(Pool &pool, hipStream_t s) : p(pool), lk(pool.mu), st(s) {}
  // Returns a pointer to at least `bytes` of flag space, ordered after earlier users.
  uint32_t *get(size_t bytes) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(st, &cs);
    if (cs != hipStreamCaptureStatusNone && p.cap < bytes) {
      // Growing would allocate during capture; reserve first (hip_marshal_reserve).
      fprintf(stderr, "[libmarshal] scratch pool too small under stream capture (need %zu "
                      "bytes; call hip_marshal_reserve first)\n", bytes);
      return nullptr;
    }
    if (pool_grow(p, bytes)) return nullptr;
    if (cs != hipStreamCaptureStatusNone) p.captured = true;
    if (p.has_ev && p.last != st && cs == hipStreamCaptureStatusNone &&
        hipStreamWaitEvent(st, p.ev, 0) != hipSuccess) {
      // without the wait this stream could clear bits another stream's stage still uses
      fprintf(stderr, "[libmarshal] scratch pool: ordering after its previous user failed\n");
      return nullptr;
    }
    ok = true;
    return static_cast<uint32_t *>(p.ptr);
  }
  ~Lease() {
    if (!ok) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(st, &cs);
    if (cs != hipStreamCaptureStatusNone) return;
    if (!p.ev) (void)hipEventCreateWithFlags(&p.ev, hipEventDisableTiming);
    if (p.ev && hipEventRecord(p.ev, st) == hipSuccess) {
      p.has_ev = true;
      p.last = st;
    }
  }
};
}  // namespace

// ---------------------------------------------------------------------------------------
// Timing in the reference's convention (include/cl_profile.h:23-56): accumulate ns into a
// caller's counter, or print "[Profile] <id> <ns>ns; <GB/s> GB/s" when it is NULL.
// ---------------------------------------------------------------------------------------
namespace {
struct Timer {
  hipStream_t st;
  const char *id;
  double bytes;
  uint64_t *out;
  bool on;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  Timer(hipStream_t s, const char *name, double moved_bytes, uint64_t *elapsed)
      : st(s), id(name), bytes(moved_bytes), out(elapsed), on(g_profiling.load() != 0) {
    if (!on) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(st, &cs);
    if (cs != hipStreamCaptureStatusNone) { on = false; return; }
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      on = false;
      return;
    }
    (void)hipEventRecord(e0, st);
  }
  bool finish(bool err) {
    if (!on) return err;
    bool ok = !err && hipEventRecord(e1, st) == hipSuccess &&
              hipEventSynchronize(e1) == hipSuccess;
    // The call's work has finished: a cluster-staged step that gave up its bounded barrier
    // wait (since the last check) left its instances unconverted, which is an error.
    if (ok && cluster_gave_up_new() != 0) {
      fprintf(stderr, "[libmarshal] %s: a cluster-staged step gave up its barrier wait (its "
                      "workgroups could not all be resident); instances left unconverted\n", id);
      err = true;
    }
    float ms = 0.f;
    if (ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) {
      uint64_t ns = uint64_t(double(ms) * 1e6 + 0.5);
      if (out) {
        *out += ns;
      } else {
        fprintf(stderr, "[Profile] %s %llu ns", id, (unsigned long long)ns);
        if (bytes > 0 && ns > 0) fprintf(stderr, "; %g GB/s", bytes / double(ns));
        fprintf(stderr, "\n");
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    on = false;
    return err;
  }
  ~Timer() {
    if (on) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
    }
  }
};

bool valid_es(size_t es) { return es == 1 || es == 2 || es == 4 || es == 8 || es == 16; }

// Product of positive dimensions; false (with a message) when any is < 1 or it overflows.
bool dims_ok(const char *who, std::initializer_list<int64_t> dims, size_t es, uint64_t *elems) {
  unsigned __int128 prod = 1;
  for (int64_t d : dims) {
    if (d < 1) {
      fprintf(stderr, "[libmarshal] %s: dimensions must be positive\n", who);
      return false;
    }
    prod *= uint64_t(d);
    if (prod * es > ((unsigned __int128)1 << 62)) {
      fprintf(stderr, "[libmarshal] %s: matrix too large\n", who);
      return false;
    }
  }
  if (!valid_es(es)) {
    fprintf(stderr, "[libmarshal] %s: element size must be 1, 2, 4, 8 or 16 bytes\n", who);
    return false;
  }
  *elems = uint64_t(prod);
  return true;
}

bool aligned_ok(const char *who, const void *p, size_t es) {
  size_t al = es < 16 ? es : 16;
  if (reinterpret_cast<uintptr_t>(p) % al) {
    fprintf(stderr, "[libmarshal] %s: buffer not aligned to %zu bytes\n", who, al);
    return false;
  }
  if (!p) {
    fprintf(stderr, "[libmarshal] %s: null buffer\n", who);
    return false;
  }
  return true;
}

uint32_t grid_cap(int64_t per_cu) { return grid_blocks(per_cu); }

}  // namespace

// ---------------------------------------------------------------------------------------
// Stage records: a route is a list of elementary stages with their chosen form. A plain call
// builds the list and runs it on its stream; a batch collects the lists of many matrices.
// ---------------------------------------------------------------------------------------
size_t current_lds_budget() { return lds_budget(); }

void set_last_route(const std::string &route) { t_route = route; }

// Workgroups for a grid-stride launch: CUs x a per-CU knob, computed in 64 bits and clamped
// to [1, 2^24] (knobs accept up to 2^30, so the product could wrap in 32 bits).
uint32_t grid_blocks(int64_t per_cu) {
  const uint64_t v =
      uint64_t(std::max(device_info().cus, 1)) * uint64_t(std::max<int64_t>(per_cu, 1));
  return uint32_t(std::min<uint64_t>(v, uint64_t(1) << 24));
}

// Which form a 010 step can use under the current LDS budget.
M010 choose_010(size_t es, uint64_t A, uint64_t a, uint64_t B, int R, int P,
                uint64_t base_addr) {
  Bs010Plan pl{};
  if (plan_010_bs(es, A, a, B, lds_budget(), base_addr, int(tuning("bs_upt")), &pl))
    return M010::BS;
  if (a * B < (1ull << 31) && pt010_flag_words(a, B, R, P) * 4 <= lds_budget())
    return M010::PTTWAC;
  return M010::Global;
}

// Pitched form of a through-scratch stage: when the destination rows are not 16-byte aligned
// (and long enough for the tiled transpose), each group of blocks is transposed into scratch
// rows at a 16-byte multiple pitch, with vector stores, and then copied back with the rows
// back to back, in aligned 16-byte chunks. The plain form would copy first and then store
// the transposed rows through skewed tiles (1- and 2-byte elements) or element by element.
// Measured on MI355X, the pitched form is 6-33 % faster for 1- and 2-byte elements and for
// stages of 1 GB and more, and 1-5 % slower for 4- and 8-byte elements below that, so knob
// "oop_pitched" = 1 takes it for 1- and 2-byte elements and from 512 MiB (2: always, 0:
// never). Returns the scratch row pitch, or 0 for the plain form.
uint64_t oop_pitch_for(const unsigned char *buf, uint64_t inst, uint64_t R, uint64_t C,
                       uint64_t S) {
  const uint64_t row_bytes = R * S;
  const uintptr_t b = reinterpret_cast<uintptr_t>(buf);
  const int64_t mode = tuning("oop_pitched");
  if (mode == 0) return 0;
  if (mode == 1 && S > 2 && inst * R * C * S < (uint64_t(512) << 20)) return 0;
  if (!(S == 1 || S == 2 || S == 4 || S == 8 || S == 16) || b % S) return 0;
  if (((b | row_bytes) & 15u) == 0) return 0;
  if (row_bytes < std::max<uint64_t>(16, uint64_t(tuning("gather_row_bytes")))) return 0;
  return (row_bytes + 15) & ~uint64_t(15);
}

static uint64_t oop_pitch(const StageRec &r) {
  return oop_pitch_for(r.buf, r.d0, r.d1, r.d2, r.d3);
}

// Uses the pitch fixed when the stage was planned (StageRec::pitch), so scratch sized from a
// plan covers the run even if the knobs change in between.
static uint64_t oop_scratch_bytes(const StageRec &r) {
  const uint64_t P = r.pitch;
  return P ? r.chunk * r.d2 * P : r.chunk * r.d1 * r.d2 * r.d3;
}

// ---------------------------------------------------------------------------------------
// Measured choices per shape class (hip_marshal_autotune_desc). A class is the stage kind,
// the element or super-element size and the power-of-two buckets of the dimensions and of the
// total bytes; its entry names the form the planner takes for every stage of that class (when
// the form can run there) and, for LDS-staged and segmented stages, a launch parameter.
// While a candidate is measured, a thread-local choice overrides the table.
// ---------------------------------------------------------------------------------------
namespace {
std::mutex g_class_mu;
std::map<ClassKey, ClassChoice> g_class;
thread_local const ClassChoice *t_force = nullptr;

int64_t lg2(uint64_t v) {
  int64_t l = 0;
  while (v > 1) {
    v >>= 1;
    l++;
  }
  return l;
}
int64_t unit_of(uint64_t S) { return S % 16 == 0 ? 16 : S % 8 == 0 ? 8 : S % 4 == 0 ? 4 : S % 2 == 0 ? 2 : 1; }
}  // namespace

ClassKey class_key_010(size_t es, uint64_t A, uint64_t a, uint64_t B) {
  return {1, int64_t(es), lg2(a), lg2(B), lg2(A * a * B * es), 0};
}
ClassKey class_key_0100(uint64_t inst, uint64_t R, uint64_t C, uint64_t S) {
  return {2, lg2(S), unit_of(S), lg2(R), lg2(C), lg2(inst * R * C * S)};
}
bool class_tuned(const ClassKey &k) {
  std::lock_guard<std::mutex> lk(g_class_mu);
  return g_class.count(k) != 0;
}
ClassChoice class_choice(const ClassKey &k) {
  if (t_force) return *t_force;
  std::lock_guard<std::mutex> lk(g_class_mu);
  auto it = g_class.find(k);
  return it == g_class.end() ? ClassChoice{} : it->second;
}

// Through scratch: groups of blocks are copied to scratch and moved back out of place,
// [inst][R][C] super-elements of S bytes -> [inst][C][R]; every byte is read twice and
// written twice by streaming kernels. One block must fit the scratch cap ("st_scratch_mb").
// No counterpart in the reference, which has only in-place forms.
static bool push_oop(std::vector<StageRec> &out, unsigned char *buf, size_t es, uint64_t inst,
                     uint64_t R, uint64_t C, uint64_t S) {
  if (inst == 0 || R <= 1 || C <= 1) return false;
  const uint64_t cap = scratch_cap_bytes();
  const double block = double(R) * double(C) * double(S);
  if (block > double(cap) || !oop_gather_fits(S)) return false;
  StageRec r{};
  r.kind = StageKind::OOP;
  r.es = es;
  r.buf = buf;
  r.d0 = inst;
  r.d1 = R;
  r.d2 = C;
  r.d3 = S;
  r.chunk = std::max<uint64_t>(1, std::min<uint64_t>(inst, cap / uint64_t(block)));
  r.pitch = oop_pitch(r);
  out.push_back(r);
  return true;
}

void push_0100(std::vector<StageRec> &out, unsigned char *buf, size_t es, uint64_t inst,
               uint64_t R, uint64_t C, uint64_t b, bool allow_oop, const char **tag) {
  if (tag) *tag = "";
  if (inst == 0 || R <= 1 || C <= 1) return;  // 1 x C and R x 1 grids are unchanged
  StageRec r{};
  r.es = es;
  r.buf = buf;
  r.d0 = inst;
  r.d1 = R;
  r.d2 = C;
  r.d3 = b;
  plan_100(es, inst, R, C, b, &r.pt, 0, reinterpret_cast<uintptr_t>(buf));
  const uint64_t S = b * es;
  const ClassChoice cc = class_choice(class_key_0100(inst, R, C, S));
  // A 010 over [inst * R][b][C] just before this step (the reference's 3-stage route: its
  // stages 2 and 3) fuses with it into one cluster-staged pass: [R][b][C] -> [C][R][b] per
  // instance, each tile transposed on its way into LDS. Two passes become one.
  if (!out.empty() && !planning_restricted() && tuning("cl_fuse") &&
      tuning("cl_max_wgs") > 0 && (cc.form == kFormModel || cc.form == kForm0100Cl)) {
    const StageRec &p = out.back();
    ClPlan fp;
    if ((p.kind == StageKind::BS010 || p.kind == StageKind::PT010) && p.buf == buf &&
        p.es == es && p.d0 == inst * R && p.d1 == b && p.d2 == C &&
        plan_cluster_fused(inst, R, b, C, es, reinterpret_cast<uintptr_t>(buf),
                           uint64_t(std::min<int64_t>(16, tuning("cl_max_wgs"))), &fp) &&
        inst * fp.G >= uint64_t(tuning("cl_min_inst"))) {
      StageRec f = r;
      f.kind = StageKind::CL;
      f.cl = fp;
      out.pop_back();
      out.push_back(f);
      if (tag) *tag = "0100_CLF-";
      return;
    }
  }
  // Many small instances: visited bits of each instance in LDS, one workgroup per instance.
  const bool lds_form = inst >= uint64_t(tuning("lds0100_min_inst")) &&
                        R * C * r.pt.nsub < (1ull << 31) &&
                        lds0100_bytes(R, C, r.pt) <= lds_budget();
  // Short cycles (small order of R modulo R*C - 1) over super-elements of some size: one
  // in-place pass with no claims, no scratch beyond the leader list.
  StageRec cy = r;
  bool cyc_ok = false;
  if (allow_oop && (S >= uint64_t(tuning("pl_min_bytes")) || cc.form == kForm0100Cyc)) {
    PlPlan pc;
    if (plan_100_leaders(R, C, uint32_t(tuning("pl_max_order")), &pc) &&
        (pc.cycles * inst * r.pt.nsub >= uint64_t(tuning("pl_min_walkers")) ||
         cc.form == kForm0100Cyc) && pc.cycles >= 1) {
      cy.kind = StageKind::PT100C;
      cy.pc = pc;
      cyc_ok = true;
    }
  }
  // Long cycles over super-elements of some size: the segmented walk, one pass with a side
  // buffer of each segment's start piece (worth it while that buffer stays under a quarter of
  // the data; grids of many short cycles have better forms).
  // (buf == nullptr: pricing a candidate route only; the table is not built, its segment
  // count is estimated.)
  StageRec sg = r;
  sg.kind = StageKind::SEG;
  bool seg_ok = false;
  const uint32_t seg_len =
      uint32_t(cc.form == kForm0100Seg && cc.param > 0 ? cc.param : tuning("seg_len"));
  if (!planning_restricted() &&
      (S >= uint64_t(tuning("seg_min_bytes")) || cc.form == kForm0100Seg) &&
      plan_100_segments(inst, R, C, S, reinterpret_cast<uintptr_t>(buf), seg_len,
                        uint64_t(tuning("seg_max_slots")), uint64_t(tuning("seg_sync_slots")),
                        buf != nullptr, &sg.sg))
    seg_ok = 4.0 * double(seg_save_bytes(inst, S, sg.sg)) <=
             double(inst) * double(R) * double(C) * double(S);
  // Instances that fit a cluster of workgroups: staged on chip, one pass (not under
  // restricted planning: the batch's grouped kernels have no cluster form).
  StageRec cs = r;
  cs.kind = StageKind::CL;
  // Measured on MI355X it beats the other forms only with many instances of a few hundred KB
  // (100 instances of 100 x 72 super-elements of 72 bytes: 30 us against 41 us segmented);
  // single instances of 1-5 MB and many small ones ran 1.1-5x slower than the cycle-leader or
  // segmented forms (one workgroup per CU holding a large share is latency-bound).
  const bool cl_ok = !planning_restricted() && tuning("cl_max_wgs") > 0 &&
                     inst >= uint64_t(tuning("cl_min_inst")) &&
                     double(R) * double(C) * double(S) >= double(tuning("cl_min_bytes")) &&
                     plan_100_cluster(inst, R, C, S, reinterpret_cast<uintptr_t>(buf),
                                      uint64_t(std::min<int64_t>(16, tuning("cl_max_wgs"))),
                                      &cs.cl);
  std::vector<StageRec> o;
  const bool oop_ok = allow_oop && push_oop(o, buf, es, inst, R, C, S);
  // A single 100 over short super-elements is a full R x C transpose of S-byte elements: the
  // LDS-staged row / column decomposition beats cycle following there.
  std::vector<StageRec> rc;
  const bool rc_ok = inst == 1 && (S < uint64_t(tuning("rc_max_super_bytes")) ||
                                   cc.form == kForm0100Rc) &&
                     push_rc(rc, buf, es, R, C, S);
  StageRec l = r;
  l.kind = lds_form ? StageKind::PT0100L : StageKind::PT100G;
  auto emit = [&](const StageRec &x, const char *t) {
    out.push_back(x);
    if (tag) *tag = t;
  };
  // A measured choice for this class, where its form can run.
  switch (cc.form) {
    case kForm0100Cyc: if (cyc_ok) return emit(cy, "0100_CYC-"); break;
    case kForm0100Seg: if (seg_ok) return emit(sg, "0100_SEG-"); break;
    case kForm0100Oop: if (oop_ok) return emit(o[0], "0100_OOP-"); break;
    case kForm0100Rc:
      if (rc_ok) {
        out.insert(out.end(), rc.begin(), rc.end());
        if (tag) *tag = "0100_RC-";
        return;
      }
      break;
    case kForm0100Lds: if (lds_form) return emit(l, "0100_LDS-"); break;
    case kForm0100Bits: {
      StageRec g = r;
      g.kind = StageKind::PT100G;
      return emit(g, "0100_PTTWAC-");
    }
    case kForm0100Cl: if (cl_ok) return emit(cs, "0100_CL-"); break;
    default: break;
  }
  // The model's choice: the cluster form where it applies (one pass at streaming rates;
  // "seg_force" keeps the segmented walk first).
  if (seg_ok && tuning("seg_force")) return emit(sg, "0100_SEG-");
  if (cl_ok) return emit(cs, "0100_CL-");
  if (cyc_ok && S >= uint64_t(tuning("pl_min_bytes")) &&
      cy.pc.cycles * inst * r.pt.nsub >= uint64_t(tuning("pl_min_walkers")))
    return emit(cy, "0100_CYC-");
  if (oop_ok) {
    double inplace = lds_form ? stage_bw_tbs(l) : 0.0;
    if (seg_ok) inplace = std::max(inplace, tuning("seg_force") ? 1e9 : stage_bw_tbs(sg));
    if (inplace < stage_bw_tbs(o[0])) return emit(o[0], "0100_OOP-");
  }
  if (seg_ok && (stage_bw_tbs(sg) > stage_bw_tbs(l) || tuning("seg_force")))
    return emit(sg, "0100_SEG-");
  if (rc_ok && S < uint64_t(tuning("rc_max_super_bytes"))) {
    out.insert(out.end(), rc.begin(), rc.end());
    if (tag) *tag = "0100_RC-";
    return;
  }
  emit(l, lds_form ? "0100_LDS-" : "0100_PTTWAC-");
}

bool plan_step_0100(std::vector<StageRec> &out, unsigned char *buf, size_t es, uint64_t inst,
                    uint64_t R, uint64_t C, uint64_t b, bool allow_oop, std::string &tag) {
  const uint64_t S = b * es;
  if (inst == 1 && R > 1 && C > 1 && (S == 1 || S == 2 || S == 4 || S == 8 || S == 16) &&
      reinterpret_cast<uintptr_t>(buf) % S == 0) {
    std::string route;
    if (!plan_transpose_stages(out, buf, S, R, C, route, allow_oop)) return false;
    tag = "0100_FULL:" + route;
    return true;
  }
  const char *t = "";
  push_0100(out, buf, es, inst, R, C, b, allow_oop, &t);
  tag = t;
  return true;
}

// A 010 whose tile is too large for LDS, split along B = B1 * B2:
//   [A][a][B1][B2] --0100(A, a, B1, B2)--> [A][B1][a][B2] --010(A*B1, a, B2)--> [A][B1][B2][a]
// The first stage moves B2-element super-elements (coalesced cycle following), the second is
// LDS-staged. B2 is the largest divisor of B whose a x B2 tile fits LDS, with super-elements
// of at least 64 bytes. Not a reference route: the reference runs such tiles with the
// element-granular cycle-following 010 (src/cl_marshal.cc:280-336, 814-818).
static bool push_010_split(std::vector<StageRec> &out, unsigned char *buf, size_t es,
                           uint64_t A, uint64_t a, uint64_t B) {
  Bs010Plan pl{};
  uint64_t best = 0;
  for (uint64_t d = 1; d * d <= B; d++) {
    if (B % d) continue;
    for (uint64_t B2 : {d, B / d}) {
      if (B2 <= best || B2 >= B || B2 * es < 64) continue;
      if (plan_010_bs(es, A * (B / B2), a, B2, lds_budget(), reinterpret_cast<uintptr_t>(buf),
                      int(tuning("bs_upt")), &pl))
        best = B2;
    }
  }
  if (!best) return false;
  push_0100(out, buf, es, A, a, B / best, best);
  StageRec r{};
  r.es = es;
  r.buf = buf;
  r.d0 = A * (B / best);
  r.d1 = a;
  r.d2 = best;
  r.R = 1;
  r.P = 1;
  plan_010_bs(es, r.d0, a, best, lds_budget(), reinterpret_cast<uintptr_t>(buf),
              int(tuning("bs_upt")), &r.bs);
  r.kind = StageKind::BS010;
  out.push_back(r);
  return true;
}

bool push_010(std::vector<StageRec> &out, unsigned char *buf, size_t es, uint64_t A,
              uint64_t a, uint64_t B, M010 m, int R, int P, const char **tag, bool allow_oop) {
  if (tag) *tag = "";
  if (A == 0 || a <= 1 || B <= 1) return true;  // 1 x B and a x 1 tiles are unchanged
  int64_t bs_grid = 0;  // a measured grid for the LDS-staged form (0: the knob)
  if (m == M010::Auto) {
    // A measured choice for this class, where its form can run.
    const ClassChoice cc = class_choice(class_key_010(es, A, a, B));
    Bs010Plan pl{};
    switch (cc.form) {
      case kForm010Bs:
        if (plan_010_bs(es, A, a, B, lds_budget(), reinterpret_cast<uintptr_t>(buf),
                        int(tuning("bs_upt")), &pl)) {
          m = M010::BS;
          bs_grid = cc.param;
        }
        break;
      case kForm010Oop:
        if (allow_oop && push_oop(out, buf, es, A, a, B, es)) {
          if (tag) *tag = "010_OOP-";
          return true;
        }
        break;
      case kForm010Split:
        if (push_010_split(out, buf, es, A, a, B)) {
          if (tag) *tag = "010_SPLIT-";
          return true;
        }
        break;
      case kForm010Rc:
        if (push_rc(out, buf, es, a, B, es, A)) {
          if (tag) *tag = "010_RC-";
          return true;
        }
        break;
      case kForm010Cyc:
        if (R >= 1 && P >= 0 && a * B < (1ull << 31) &&
            pt010_flag_words(a, B, R, P) * 4 <= lds_budget())
          m = M010::PTTWAC;
        break;
      default: break;
    }
  }
  if (m == M010::Auto) {
    m = choose_010(es, A, a, B, R, P, reinterpret_cast<uintptr_t>(buf));
    // Through scratch (two streaming passes) when allowed, for tiles too large for LDS, for
    // LDS-staged runs that would move single bytes or halves (no dword grouping: a tile of
    // over half the LDS with a*B*es not a multiple of 4) and, from 1 MB on, for LDS-staged
    // runs that would fill fewer than half the CUs with tiles; else the in-place forms below.
    bool oop = m != M010::BS;
    if (!oop && allow_oop) {
      Bs010Plan pl{};
      if (plan_010_bs(es, A, a, B, lds_budget(), reinterpret_cast<uintptr_t>(buf),
                      int(tuning("bs_upt")), &pl))
        oop = (es < 4 && pl.G == 1) ||
              (double(A) * double(a) * double(B) * double(es) >= double(1 << 20) &&
               pl.nchunks * 2 < uint64_t(device_info().cus));
    }
    if (oop && allow_oop && push_oop(out, buf, es, A, a, B, es)) {
      if (tag) *tag = "010_OOP-";
      return true;
    }
    if (m != M010::BS && push_010_split(out, buf, es, A, a, B)) {
      if (tag) *tag = "010_SPLIT-";
      return true;
    }
    // Otherwise each a x B tile by the row / column decomposition (A instances), when its
    // rows fit LDS: in place, LDS-staged, no element-granular cycle following.
    if (m != M010::BS && push_rc(out, buf, es, a, B, es, A)) {
      if (tag) *tag = "010_RC-";
      return true;
    }
  }
  StageRec r{};
  r.es = es;
  r.buf = buf;
  r.d0 = A;
  r.d1 = a;
  r.d2 = B;
  r.R = R;
  r.P = P;
  if (m == M010::BS) {
    if (!plan_010_bs(es, A, a, B, lds_budget(), reinterpret_cast<uintptr_t>(buf),
                     int(tuning("bs_upt")), &r.bs)) {
      fprintf(stderr, "[libmarshal] 010 BS: tile %llux%llu does not fit the LDS budget\n",
              (unsigned long long)a, (unsigned long long)B);
      return false;
    }
    r.kind = StageKind::BS010;
    r.grid = bs_grid;
    if (tag) *tag = "010_BS-";
    out.push_back(r);
    return true;
  }
  if (m == M010::PTTWAC) {
    if (R < 1 || P < 0 || a * B >= (1ull << 31) ||
        pt010_flag_words(a, B, R, P) * 4 > lds_budget()) {
      fprintf(stderr, "[libmarshal] 010 PTTWAC: flags of a %llux%llu tile do not fit LDS\n",
              (unsigned long long)a, (unsigned long long)B);
      return false;
    }
    r.kind = StageKind::PT010;
    if (tag) *tag = "010_PTTWAC-";
    out.push_back(r);
    return true;
  }
  // 010 on [A][a][B] is 0100 with single-element super-elements: global visited bits.
  if (tag) *tag = "010_G-";
  push_0100(out, buf, es, A, a, B, 1);
  return true;
}

bool push_rc(std::vector<StageRec> &out, unsigned char *buf, size_t es, uint64_t M,
             uint64_t N, uint64_t S, uint64_t inst) {
  if (M <= 1 || N <= 1) return true;  // a vector is its own transpose
  RcPlan pl{};
  if (!plan_rc(es, M, N, S, lds_budget(), &pl, inst)) return false;
  StageRec r{};
  r.es = es;
  r.buf = buf;
  r.d0 = inst;
  r.d1 = M;
  r.d2 = N;
  r.d3 = S;
  r.rc = pl;
  if (pl.c > 1) {
    r.kind = StageKind::RC_COL1;
    out.push_back(r);
  }
  r.kind = StageKind::RC_ROW;
  out.push_back(r);
  r.kind = StageKind::RC_COL2;
  out.push_back(r);
  return true;
}

bool push_st(std::vector<StageRec> &out, unsigned char *buf, size_t es, uint64_t M,
             uint64_t N, int variant) {
  if (M <= 1 || N <= 1) return true;
  StageRec r{};
  if (!plan_st(es, M, N, variant, uint64_t(tuning("st_super_bytes")),
               scratch_cap_bytes(), tuning("st_pt_vw"),
               tuning("st_rotate") != 0, &r.stp, reinterpret_cast<uintptr_t>(buf)))
    return false;
  r.kind = StageKind::ST;
  r.es = es;
  r.buf = buf;
  r.d1 = M;
  r.d2 = N;
  out.push_back(r);
  return true;
}

// Composer: the stage sequences of cl_transpose (src/cl_marshal.cc:561-776).
bool route_stages(std::vector<StageRec> &out, unsigned char *buf, size_t es, uint64_t A,
                  uint64_t a, uint64_t B, uint64_t b, int R, int stages, std::string &route) {
  route.clear();
  const char *tag = "";
  auto s010 = [&](uint64_t A_, uint64_t a_, uint64_t B_) {
    bool ok = push_010(out, buf, es, A_, a_, B_, M010::Auto, R, 1, &tag);
    route += tag;
    return ok;
  };
  switch (stages) {
    case 2:  // AaBb -> ABba: 010(A, a, Bb), then 100(A, Bb, a)   (src/cl_marshal.cc:569-601)
      if (!s010(A, a, B * b)) return false;
      push_0100(out, buf, es, 1, A, B * b, a);
      route += "2.1-";
      return true;
    case 22:  // 100(Aa, B, b), then 010(B, Aa, b)               (src/cl_marshal.cc:602-634)
      push_0100(out, buf, es, 1, A * a, B, b);
      if (!s010(B, A * a, b)) return false;
      route += "2.2-";
      return true;
    case 3:  // AaBb -> BAab -> BAba -> BbAa                      (src/cl_marshal.cc:636-677)
      if (a > 1 && b > 1) {
        push_0100(out, buf, es, 1, A * a, B, b);
        if (!s010(B * A, a, b)) return false;
        push_0100(out, buf, es, B, A, b, a);
        route += "3.1-";
        return true;
      }
      break;
    case 32:  // AaBb -> ABab -> ABba -> BbAa                     (src/cl_marshal.cc:679-720)
      if (a > 1 && b > 1) {
        push_0100(out, buf, es, A, a, B, b);
        if (!s010(A * B, a, b)) return false;
        push_0100(out, buf, es, 1, A, B * b, a);
        route += "3.2-";
        return true;
      }
      break;
    case 5:  // row / column decomposition (no counterpart in the reference)
      if (push_rc(out, buf, es, A * a, B * b, es)) {
        route += "RC-";
        return true;
      }
      break;
    case 9:  // matrices within the scratch cap: copy to scratch, transpose back (no counterpart)
      if (push_oop(out, buf, es, 1, A * a, B * b, es)) {
        route += "OOP-";
        return true;
      }
      break;
    case 8:  // square: one pass of tile-pair swaps (no counterpart)
      if (A * a == B * b && sq_fits(es, A * a)) {
        StageRec r{};
        r.kind = StageKind::SQ;
        r.es = es;
        r.buf = buf;
        r.d1 = A * a;
        r.d2 = B * b;
        out.push_back(r);
        route += "SQ-";
        return true;
      }
      break;
    case 6:  // stripes of rows / of columns through bounded scratch (no counterpart)
    case 7:
      if (push_st(out, buf, es, A * a, B * b, stages - 6)) {
        route += stages == 6 ? "ST_R-" : "ST_C-";
        return true;
      }
      break;
    case 4:  // Karlsson: 0100, 010, 100(A, B, ba), 0100          (src/cl_marshal.cc:722-767)
      push_0100(out, buf, es, A, a, B, b);
      if (!s010(B * A, a, b)) return false;
      push_0100(out, buf, es, 1, A, B, b * a);
      push_0100(out, buf, es, B, A, b, a);
      route += "4-";
      return true;
    default:
      break;
  }
  // Element-granular fallback: 100 over the whole matrix (src/cl_marshal.cc:769-775). When
  // the rows and columns fit LDS, the row / column decomposition replaces it.
  if (push_rc(out, buf, es, A * a, B * b, es)) {
    route += "RC-";
    return true;
  }
  route += "Fallback-";
  push_0100(out, buf, es, 1, A * a, B * b, 1);
  return true;
}

size_t stages_scratch_bytes(const std::vector<StageRec> &stages) {
  size_t m = 0;
  for (const StageRec &r : stages) {
    if (r.kind == StageKind::PT100G) m = std::max<size_t>(m, r.pt.flag_words * 4);
    if (r.kind == StageKind::PT100C) m = std::max<size_t>(m, size_t(r.pc.cycles + 1) * 4);
    if (r.kind == StageKind::ST) m = std::max<size_t>(m, r.stp.scratch_bytes);
    if (r.kind == StageKind::OOP) m = std::max<size_t>(m, size_t(oop_scratch_bytes(r)));
    if (r.kind == StageKind::SEG) m = std::max<size_t>(m, seg_save_bytes(r.d0, r.d3 * r.es, r.sg));
    if (r.kind == StageKind::CL) m = std::max<size_t>(m, size_t(cluster_flag_words(r.d0)) * 4);
  }
  return m;
}

// Runs stage records in order on one stream. Scratch (visited bits, the stripe route's
// buffer) comes from `scr` when given (at least stages_scratch_bytes, private to the caller),
// else from the device pool.
// `alone`: nothing else of this library runs on the device meanwhile (a plain call, or a
// server request with every other stream idle), which only tunes launch shapes.
static bool run_stages_impl(const std::vector<StageRec> &stages, hipStream_t st, bool alone,
                            unsigned char *scr) {
  for (const StageRec &r : stages) {
    switch (r.kind) {
      case StageKind::BS010:
        LM_CHECK_HIP(launch_010_bs(r.buf, r.es, r.d0, r.d1, r.d2, r.bs,
                                   grid_cap(r.grid ? r.grid : tuning("bs_blocks_per_cu")), st));
        break;
      case StageKind::PT010:
        LM_CHECK_HIP(launch_010_pttwac(r.buf, r.es, r.d0, r.d1, r.d2, r.R, r.P,
                                       grid_cap(tuning("pt010_blocks_per_cu")), st));
        break;
      case StageKind::PT0100L:
        LM_CHECK_HIP(launch_0100_lds(r.buf, r.d0, r.d1, r.d2, r.d3 * r.es, r.pt, grid_cap(8),
                                     st));
        break;
      case StageKind::RC_ROW:
      case StageKind::RC_COL1:
      case StageKind::RC_COL2: {
        int pass = r.kind == StageKind::RC_ROW ? 0 : r.kind == StageKind::RC_COL1 ? 1 : 2;
        LM_CHECK_HIP(launch_rc(r.buf, r.d1, r.d2, r.rc, pass,
                               grid_cap(uint32_t(tuning(pass ? "rc_col_blocks_per_cu"
                                                             : "rc_row_blocks_per_cu"))),
                               st));
        break;
      }
      case StageKind::SQ:
        LM_CHECK_HIP(launch_sq_transpose(r.buf, r.es, r.d1, st));
        break;
      case StageKind::OOP: {
        // groups of `chunk` blocks: copy to scratch, then out of place back into their place
        // (or, pitched: out of place into scratch, then copied back)
        const uint64_t S = r.d3, block = r.d1 * r.d2 * S;
        const uint64_t P = r.pitch;  // decided when planned: the scratch was sized with it
        std::unique_ptr<Lease> lease;
        unsigned char *tmp = scr;
        if (!tmp) {
          lease.reset(new Lease(pool_here(), st));
          tmp = reinterpret_cast<unsigned char *>(lease->get(oop_scratch_bytes(r)));
        }
        if (!tmp) return true;
        // A call that runs alone: tiles small enough that a group makes at least
        // "oop_min_tiles" of them (one workgroup each, on 256 CUs), down to 4 KiB. 512
        // measured 2-7 % faster on 2-17 MB matrices (1031 x 997 f32 481 -> 511 GB/s,
        // 4000 x 1000 f32 1142 -> 1211), but where several matrices run at once (a batch's
        // side streams) the larger tiles did better (its 1-16 MB class 8 % slower).
        size_t tile = size_t(tuning("st_tile_kb")) << 10;
        const int64_t mt = alone ? tuning("oop_min_tiles") : 0;
        if (mt) {
          const uint64_t per = std::min(r.chunk, r.d0) * block / uint64_t(mt);
          tile = std::min<size_t>(tile, std::max<size_t>(4096, size_t(per)));
        }
        for (uint64_t g0 = 0; g0 < r.d0; g0 += r.chunk) {
          const uint64_t n = std::min(r.chunk, r.d0 - g0);
          unsigned char *dst = r.buf + g0 * block;
          if (P) {
            // When the source and the scratch together fit the last-level cache
            // ("copy_nt_store_mb", 256 MiB), both passes store cached and the stage works in
            // that cache: 63 MB u8 +13 %, 67 MB u16 +4 %; a 268 MB stage was 6 % slower
            // cached. Below 32 MiB it measured 6 % slower too (21 MB u8), so small stages keep
            // the non-temporal stores.
            const int64_t knt = tuning("oop_nt");
            const uint64_t bytes = n * block;
            const bool cached = 2 * bytes < (uint64_t(tuning("copy_nt_store_mb")) << 20) &&
                                bytes >= (uint64_t(32) << 20);
            const int nt = int(cached ? (knt & 1) : knt);
            LM_CHECK_HIP(launch_oop_transpose(dst, tmp, S, n, r.d1, r.d2, block, r.d2 * P, tile,
                                              st, 0, P, nt));
            LM_CHECK_HIP(launch_unpitch(dst, tmp, r.d1 * S, P, n * r.d2, st,
                                        cached ? 0 : -1));
            continue;
          }
          LM_CHECK_HIP(launch_copy(tmp, dst, n * block, st));
          // the tiled transpose for 1- to 16-byte elements, unless destination rows are
          // short and not 16-byte aligned: then the gather, whose tiles span whole columns
          // and so write one contiguous range each
          const uint64_t row_bytes = r.d1 * S;
          const bool aligned_rows =
              ((reinterpret_cast<uintptr_t>(dst) | row_bytes) & 15u) == 0;
          if ((S == 1 || S == 2 || S == 4 || S == 8 || S == 16) &&
              reinterpret_cast<uintptr_t>(dst) % S == 0 &&
              (aligned_rows || row_bytes >= uint64_t(tuning("gather_row_bytes"))))
            LM_CHECK_HIP(launch_oop_transpose(tmp, dst, S, n, r.d1, r.d2, block, block, tile, st));
          else
            LM_CHECK_HIP(launch_oop_gather(tmp, dst, S, n, r.d1, r.d2,
                                           size_t(tuning("gather_tile_kb")) << 10, st));
        }
        break;
      }
      case StageKind::ST: {
        std::unique_ptr<Lease> lease;
        unsigned char *buf = scr;
        if (!buf) {
          lease.reset(new Lease(pool_here(), st));
          buf = reinterpret_cast<unsigned char *>(lease->get(r.stp.scratch_bytes));
        }
        if (!buf) return true;
        LM_CHECK_HIP(run_st(r.buf, r.stp, buf, size_t(tuning("st_tile_kb")) << 10,
                            grid_cap(tuning("pt100_blocks_per_cu")), st));
        break;
      }
      case StageKind::PT100C: {
        const uint32_t cap = grid_cap(tuning("pt100_blocks_per_cu"));
        const uint32_t *cached = cached_leaders(r.d1, r.d2, r.pc, st);
        if (cached) {
          LM_CHECK_HIP(launch_leader_walk(r.buf, cached, r.d0, r.d1, r.d2, r.d3 * r.es, r.pt,
                                          r.pc, cap, st));
          break;
        }
        std::unique_ptr<Lease> lease;
        uint32_t *list = reinterpret_cast<uint32_t *>(scr);
        if (!list) {
          lease.reset(new Lease(pool_here(), st));
          list = lease->get((r.pc.cycles + 1) * 4);
        }
        if (!list) return true;
        LM_CHECK_HIP(launch_100_leaders(r.buf, list, r.d0, r.d1, r.d2, r.d3 * r.es, r.pt, r.pc,
                                        cap, st));
        break;
      }
      case StageKind::SEG: {
        std::unique_ptr<Lease> lease;
        unsigned char *save = scr;
        if (!save) {
          lease.reset(new Lease(pool_here(), st));
          save = reinterpret_cast<unsigned char *>(
              lease->get(seg_save_bytes(r.d0, r.d3 * r.es, r.sg)));
        }
        if (!save) return true;
        LM_CHECK_HIP(launch_100_segments(r.buf, save, r.d0, r.d1, r.d2, r.d3 * r.es, r.sg,
                                         grid_cap(tuning("seg_blocks_per_cu")),
                                         uint32_t(tuning("seg_nt")), uint32_t(tuning("seg_xcd")),
                                         st));
        break;
      }
      case StageKind::CL: {
        std::unique_ptr<Lease> lease;
        uint32_t *flags = reinterpret_cast<uint32_t *>(scr);
        if (!flags) {
          lease.reset(new Lease(pool_here(), st));
          flags = lease->get(cluster_flag_words(r.d0) * 4);
        }
        if (!flags) return true;
        LM_CHECK_HIP(launch_100_cluster(r.buf, flags, r.d0, r.d1, r.d2, r.d3 * r.es, r.cl, st));
        break;
      }
      case StageKind::PT100G: {
        std::unique_ptr<Lease> lease;
        uint32_t *flags = reinterpret_cast<uint32_t *>(scr);
        if (!flags) {
          lease.reset(new Lease(pool_here(), st));
          flags = lease->get(r.pt.flag_words * 4);
        }
        if (!flags) return true;
        LM_CHECK_HIP(launch_fill_u32(flags, r.pt.flag_words, 0, st));
        LM_CHECK_HIP(launch_100_pttwac(r.buf, flags, r.d0, r.d1, r.d2, r.d3 * r.es, r.pt,
                                       grid_cap(tuning("pt100_blocks_per_cu")), st));
        break;
      }
    }
  }
  return false;
}

bool run_stages(const std::vector<StageRec> &stages, hipStream_t st) {
  return run_stages_impl(stages, st, true, nullptr);
}

bool run_stages_on(const std::vector<StageRec> &stages, hipStream_t st, unsigned char *scratch,
                   bool alone) {
  return run_stages_impl(stages, st, alone, scratch);
}

// One 0100 / 010 step on its own (entry points of single stages and the AoS / SoA
// conversions); the route taken is what hip_marshal_last_route reports afterwards.
static bool stage_0100(unsigned char *buf, size_t es, uint64_t inst, uint64_t R, uint64_t C,
                       uint64_t b, hipStream_t st) {
  std::vector<StageRec> v;
  std::string tag;
  if (!plan_step_0100(v, buf, es, inst, R, C, b, true, tag)) return true;
  t_route = tag;
  return run_stages(v, st);
}

static bool stage_010(unsigned char *buf, size_t es, uint64_t A, uint64_t a, uint64_t B,
                      M010 m, int R, int P, hipStream_t st) {
  std::vector<StageRec> v;
  const char *tag = "";
  if (!push_010(v, buf, es, A, a, B, m, R, P, &tag, m == M010::Auto)) return true;
  t_route = tag;
  return run_stages(v, st);
}

// ---------------------------------------------------------------------------------------
// Planner: tile factors and route for an M x N transpose.
// Every candidate route is expanded into the stage list it would run (the same records a
// call executes), and each stage is priced as 2 x bytes / the bandwidth its kernel form
// reaches on MI355X (measured): LDS-staged 010 ~5.5 TB/s; cycle following over S-byte
// super-elements is bound by device-scope atomic claims (one per move, ~5*10^9/s), i.e. by
// ~10 GB/s per byte of S, up to ~3 TB/s; LDS-bit forms ~0.7-1.5 TB/s; the row pass of the
// row/column decomposition ~3 TB/s and each of its column passes ~0.55 TB/s.
// Candidates: routes 2, 22, 3, 32 over divisor pairs, the row/column decomposition and the
// element-granular fallback. Choices are cached per (M, N, element size).
// ---------------------------------------------------------------------------------------
namespace {
std::vector<uint64_t> divisors(uint64_t n) {
  std::vector<uint64_t> lo, hi;
  for (uint64_t d = 1; d * d <= n; d++) {
    if (n % d) continue;
    lo.push_back(d);
    if (d * d != n) hi.push_back(n / d);
  }
  lo.insert(lo.end(), hi.rbegin(), hi.rend());
  return lo;
}

// At most `cap` divisors, spread over the whole range (always keeping 1 and n).
std::vector<uint64_t> some_divisors(uint64_t n, size_t cap) {
  std::vector<uint64_t> d = divisors(n);
  if (d.size() <= cap) return d;
  std::vector<uint64_t> out;
  for (size_t i = 0; i < cap; i++) out.push_back(d[i * (d.size() - 1) / (cap - 1)]);
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

}  // namespace

double stage_bw_tbs(const StageRec &r) {
  size_t es = r.es;
  switch (r.kind) {
    case StageKind::BS010: {
      if (es >= 4) return 5.5;
      const uint64_t G = uint64_t(r.bs.G);
      if (G <= 1) return es == 2 ? 4.5 : 2.8;
      // Dword units: both sides are carry-free when either extent is a whole number of units
      // (the chunk is then staged in the other side's order), one side otherwise.
      const int whole = (r.d1 % G == 0 || r.d2 % G == 0) ? 2 : 1;
      if (es == 2) return whole == 2 ? 5.3 : whole == 1 ? 4.6 : 4.3;
      return whole == 2 ? 4.7 : whole == 1 ? 3.4 : 2.7;
    }
    case StageKind::PT010: return 1.5;
    case StageKind::PT0100L: return std::min(3.0, 0.01 * double(r.d3 * es));
    case StageKind::PT100G: {
      double S = double(r.d3 * es);
      return std::min(3.0, 0.12 + 0.011 * S);
    }
    case StageKind::PT100C: return std::min(5.0, 0.02 * double(r.d3 * es));
    case StageKind::SEG: {
      // One pass of S-byte pieces at scattered slots: a piece at an offset aligned to
      // a = gcd(S, 128) touches (S + 128 - a) / 128 cache lines on average, so the useful share
      // of the lines moved is S / (S + 128 - a). Measured on MI355X (GB/s): 64 B 3100-3900,
      // 72 B 2100-2500, 128 B 5050, 256 B 4870, 288 B 3300, 1 KiB 5440, 4 KiB 5700. Narrow
      // units (pieces aligned to < 8 bytes) cost more instructions per byte. The side buffer
      // adds a read and a write of every segment start.
      const uint64_t S = r.d3 * es;
      uint64_t a = 128;
      while (a > 1 && S % a) a >>= 1;
      const double useful = double(S) / double(S + 128 - a);
      const int ub = r.sg.ub;
      const double unit = ub >= 8 ? 1.0 : ub == 4 ? 0.6 : ub == 2 ? 0.35 : 0.2;
      const double extra = 2.0 * double(r.sg.nseg) / double(std::max<uint64_t>(r.d1 * r.d2, 1));
      return std::min(5.5, 6.5 * useful) * unit / (1.0 + extra);
    }
    case StageKind::CL: {
      // A contiguous read of every instance and one scattered write of S-byte runs: a run at
      // an offset aligned to a = gcd(S, 128) touches (S + 128 - a) / 128 cache lines.
      const uint64_t S = r.d3 * es;
      uint64_t a = 128;
      while (a > 1 && S % a) a >>= 1;
      const double useful = double(S) / double(S + 128 - a);
      // Measured on MI355X: 3.5 TB/s for 100 instances of 518 KB (72-byte super-elements);
      // 2.2-2.5 TB/s for the fused form (its tiles are transposed on the way into LDS), which
      // replaces two passes.
      const double bw = 2.0 / (1.0 / 5.5 + 1.0 / std::min(5.5, 6.5 * useful));
      return std::min(bw, r.cl.fes ? 2.3 : 3.5);
    }
    case StageKind::RC_ROW: return r.rc.ub >= 4 ? 3.0 : r.rc.ub == 2 ? 2.0 : 1.2;
    case StageKind::RC_COL1:
    case StageKind::RC_COL2: return r.rc.ub >= 4 ? 0.55 : r.rc.ub == 2 ? 0.4 : 0.3;
    case StageKind::SQ: return es >= 4 ? 5.0 : es == 2 ? 4.0 : 3.0;
    case StageKind::OOP: {  // a copy pass and an out-of-place transpose / gather pass
      const uint64_t S = r.d3;
      if (S == 1 || S == 2 || S == 4 || S == 8 || S == 16)
        return 1.0 / (1.0 / 5.5 + 1.0 / (S >= 4 ? 4.5 : S == 2 ? 3.5 : 2.5));
      const uint64_t u = (S & 15u) == 0 ? 16 : (S & 7u) == 0 ? 8 : (S & 3u) == 0 ? 4 : (S & 1u) == 0 ? 2 : 1;
      return 1.0 / (1.0 / 5.5 + 1.0 / (u >= 8 ? 4.0 : u == 4 ? 3.5 : u == 2 ? 3.0 : 2.0));
    }
    case StageKind::ST: {  // copy + out-of-place transpose + 100 (+ padding and a 2-D copy)
      const StPlan &p = r.stp;
      StageRec pt{};
      pt.kind = StageKind::PT100G;
      pt.es = es;
      pt.d3 = p.h;
      double t = 1.0 / 5.5 + 1.0 / (es >= 4 ? 4.5 : es == 2 ? 3.5 : 2.5);
      if (p.q > 1) t += 1.0 / stage_bw_tbs(pt);
      if (p.r) t += 1.0 / (es >= 4 ? 3.0 : es == 2 ? 2.0 : 1.0);
      return 1.0 / t;
    }
  }
  return 0.1;
}

namespace {

double stage_bytes(const StageRec &r) {
  switch (r.kind) {
    case StageKind::BS010:
    case StageKind::PT010: return double(r.d0) * double(r.d1) * double(r.d2) * double(r.es);
    case StageKind::PT100G:
    case StageKind::PT100C:
    case StageKind::PT0100L:
    case StageKind::SEG:
    case StageKind::CL:
      return double(r.d0) * double(r.d1) * double(r.d2) * double(r.d3) * double(r.es);
    case StageKind::RC_ROW:
    case StageKind::RC_COL1:
    case StageKind::RC_COL2:  // d0 instances of d1 x d2 elements of d3 bytes
    case StageKind::OOP:
      return double(std::max<uint64_t>(r.d0, 1)) * double(r.d1) * double(r.d2) * double(r.d3);
    default:  // ST: d1 x d2 elements of es bytes
      return double(r.d1) * double(r.d2) * double(r.es);
  }
}

// Modelled time (relative units) of a stage list.
double route_cost(const std::vector<StageRec> &v) {
  double t = 0;
  for (const StageRec &r : v) t += stage_bytes(r) / stage_bw_tbs(r);
  return t;
}

// LIBMARSHAL_PLAN_DEBUG=1 prints every candidate route and its modelled cost per byte.
bool plan_debug() {
  static const bool on = [] {
    const char *e = getenv("LIBMARSHAL_PLAN_DEBUG");
    return e && *e && *e != '0';
  }();
  return on;
}

std::mutex g_plan_mu;
std::map<std::tuple<uint64_t, uint64_t, size_t, size_t, bool, bool>, Choice> g_plan_cache;
std::map<std::tuple<uint64_t, uint64_t, size_t, size_t, bool, bool>, Choice> g_tuned;  // measured
}  // namespace

// Every candidate route the model considers, cheapest first (one entry per stage code and
// tile factors).
static std::vector<Choice> rank_routes(size_t es, uint64_t M, uint64_t N, bool allow_st) {
  std::vector<Choice> all;
  std::string route;
  auto consider = [&](uint64_t A, uint64_t a, uint64_t B, uint64_t b, int stages) {
    std::vector<StageRec> v;
    if (!route_stages(v, nullptr, es, A, a, B, b, 1, stages, route)) return;
    if (route.find("Fallback") != std::string::npos && stages != -1) return;
    if (route.find("RC-") == std::string::npos && stages == 5) return;
    if (route.find("ST_") == std::string::npos && (stages == 6 || stages == 7)) return;
    if (route.find("SQ-") == std::string::npos && stages == 8) return;
    if (route.find("OOP-") == std::string::npos && stages == 9) return;
    double c = route_cost(v);
    if (plan_debug())
      fprintf(stderr, "[libmarshal] plan %llux%llu es=%zu: stages %d a=%llu b=%llu %s cost %.4g\n",
              (unsigned long long)M, (unsigned long long)N, es, stages, (unsigned long long)a,
              (unsigned long long)b, route.c_str(), c / double(M * N * es));
    Choice ch;
    ch.cost = c;
    ch.stages = stages;
    ch.a = (stages == 22) ? M : a;
    ch.b = (stages == 2) ? N : b;
    all.push_back(ch);
  };
  consider(1, M, 1, N, -1);  // element-granular fallback
  consider(1, M, 1, N, 5);   // row / column decomposition
  if (allow_st && M == N) consider(1, M, 1, N, 8);  // square: tile-pair swaps
  if (allow_st) consider(1, M, 1, N, 9);            // through scratch (within its cap)
  if (allow_st) {            // stripes through bounded scratch
    consider(1, M, 1, N, 6);
    consider(1, M, 1, N, 7);
  }
  std::vector<uint64_t> dm = some_divisors(M, 24), dn = some_divisors(N, 24);
  for (uint64_t a : dm) {
    if (a > 1 && a < M) consider(M / a, a, 1, N, 2);
    for (uint64_t b : dn) {
      if (a == 1 && b > 1 && b < N) consider(1, M, N / b, b, 22);
      if (a > 1 && b > 1 && a < M && b < N) {
        consider(M / a, a, N / b, b, 3);
        consider(M / a, a, N / b, b, 32);
      }
    }
  }
  std::stable_sort(all.begin(), all.end(),
                   [](const Choice &x, const Choice &y) { return x.cost < y.cost; });
  return all;
}

Choice plan_full(size_t es, uint64_t M, uint64_t N, bool allow_st) {
  Choice best;
  best.a = M;
  best.b = N;
  best.stages = -1;
  if (M <= 1 || N <= 1) return best;
  auto key = std::make_tuple(M, N, es, lds_budget(), allow_st, planning_restricted());
  {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) return it->second;
    it = g_plan_cache.find(key);
    if (it != g_plan_cache.end()) return it->second;
  }
  std::vector<Choice> all = rank_routes(es, M, N, allow_st);
  if (!all.empty()) best = all.front();
  std::lock_guard<std::mutex> lk(g_plan_mu);
  if (g_plan_cache.size() > 65536) g_plan_cache.clear();
  g_plan_cache[key] = best;
  return best;
}

// The stage list of a planner choice.
static bool choice_stages(std::vector<StageRec> &out, unsigned char *buf, size_t es,
                          uint64_t M, uint64_t N, const Choice &c, std::string &route) {
  if (c.stages == 2) return route_stages(out, buf, es, M / c.a, c.a, 1, N, 1, 2, route);
  if (c.stages == 22) return route_stages(out, buf, es, 1, M, N / c.b, c.b, 1, 22, route);
  if (c.stages == 3 || c.stages == 32)
    return route_stages(out, buf, es, M / c.a, c.a, N / c.b, c.b, 1, c.stages, route);
  if (c.stages >= 5 && c.stages <= 9)
    return route_stages(out, buf, es, 1, M, 1, N, 1, c.stages, route);
  return route_stages(out, buf, es, 1, M, 1, N, 1, -1, route);
}

// ---------------------------------------------------------------------------------------
// Measured autotuner for full transposes: the model's cheapest candidates plus the best of
// each route family (stripes by rows and by columns, row/column decomposition, tiled) are
// run on the caller's buffer, each timed with events on the caller's stream and followed by
// the inverse transpose, which restores the buffer bit for bit. The fastest becomes the
// plan for (M, N, element size) from then on (hip_marshal_plan_transpose reports it).
// ---------------------------------------------------------------------------------------
static bool autotune_full(hipStream_t st, unsigned char *buf, size_t es, uint64_t M,
                          uint64_t N, bool allow_st, Choice *out) {
  std::vector<Choice> all = rank_routes(es, M, N, allow_st);
  if (all.empty()) return true;
  std::vector<Choice> cand;
  auto add = [&](const Choice &c) {
    for (const Choice &d : cand)
      if (d.stages == c.stages && d.a == c.a && d.b == c.b) return;
    cand.push_back(c);
  };
  for (size_t i = 0; i < all.size() && i < 3; i++) add(all[i]);
  auto family_best = [&](auto pred) {
    for (const Choice &c : all)
      if (pred(c)) {
        add(c);
        return;
      }
  };
  family_best([](const Choice &c) { return c.stages == 6; });
  family_best([](const Choice &c) { return c.stages == 7; });
  family_best([](const Choice &c) { return c.stages == 5; });
  family_best([](const Choice &c) { return c.stages == 8 || c.stages == 9; });
  family_best([](const Choice &c) {
    return c.stages == 2 || c.stages == 22 || c.stages == 3 || c.stages == 32;
  });
  // The restoring transpose and every candidate's scratch, reserved before timing.
  std::vector<StageRec> inv;
  std::string route;
  if (!choice_stages(inv, buf, es, N, M, plan_full(es, N, M, allow_st), route)) return true;
  size_t need = stages_scratch_bytes(inv);
  std::vector<std::vector<StageRec>> lists(cand.size());
  for (size_t i = 0; i < cand.size(); i++) {
    if (!choice_stages(lists[i], buf, es, M, N, cand[i], route)) lists[i].clear();
    need = std::max(need, stages_scratch_bytes(lists[i]));
  }
  {
    Pool &p = pool_here();
    std::lock_guard<std::mutex> lk(p.mu);
    if (pool_grow(p, need)) return true;
  }
  // A device snapshot of the buffer (when it fits the scratch cap), copied back if a candidate
  // fails part-way; without one a failure is reported as leaving the buffer unrestored.
  const uint64_t bytes = M * N * es;
  void *snap = nullptr;
  if (bytes <= (uint64_t(tuning("st_scratch_mb")) << 20) &&
      hipMalloc(&snap, bytes) == hipSuccess &&
      hipMemcpyAsync(snap, buf, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess) {
    (void)hipFree(snap);
    snap = nullptr;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float best_ms = 1e30f;
  size_t best = SIZE_MAX;
  bool err = hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess;
  const bool warm = double(M) * double(N) * double(es) < double(1u << 30);
  for (size_t i = 0; i < cand.size() && !err; i++) {
    if (lists[i].empty()) continue;
    float ms = 0;
    if (warm)  // first launches of a kernel form carry one-time costs: one untimed round trip
      err = run_stages(lists[i], st) || run_stages(inv, st);
    err = err || hipEventRecord(e0, st) != hipSuccess || run_stages(lists[i], st) ||
          hipEventRecord(e1, st) != hipSuccess || run_stages(inv, st) ||
          hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess || hipStreamSynchronize(st) != hipSuccess;
    if (plan_debug())
      fprintf(stderr, "[libmarshal] autotune %llux%llu es=%zu: stages %d a=%llu b=%llu %.3f ms\n",
              (unsigned long long)M, (unsigned long long)N, es, cand[i].stages,
              (unsigned long long)cand[i].a, (unsigned long long)cand[i].b, ms);
    if (!err && ms < best_ms) {
      best_ms = ms;
      best = i;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (err) {
    (void)hipStreamSynchronize(st);
    const bool restored =
        snap && hipMemcpy(buf, snap, bytes, hipMemcpyDeviceToDevice) == hipSuccess;
    fprintf(stderr, "[libmarshal] autotune %llux%llu: a candidate failed; the buffer %s\n",
            (unsigned long long)M, (unsigned long long)N,
            restored ? "was restored"
            : snap   ? "was not restored (the copy back failed)"
                     : "was not restored (larger than the scratch cap)");
  }
  if (snap) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(snap);
  }
  if (err || best == SIZE_MAX) return true;
  Choice c = cand[best];
  c.cost = double(best_ms);
  std::lock_guard<std::mutex> lk(g_plan_mu);
  g_tuned[std::make_tuple(M, N, es, lds_budget(), allow_st, planning_restricted())] = c;
  *out = c;
  return false;
}

// ---------------------------------------------------------------------------------------
// Measured autotuner per class of 010 / 0100 steps (AoS/SoA <-> ASTA conversions): every
// form the class could take (LDS-staged with three grid sizes, through scratch, the split,
// row/column shuffles, cycle following; for 0100 also cycle leaders, the segmented walk with
// three segment lengths, LDS bits, global bits) is planned for the caller's buffer, run and
// timed, and followed by the inverse step (planned by the model), which restores the buffer
// bit for bit. The fastest becomes the class's entry. Before the first candidate runs, the
// buffer is copied to a device snapshot (when one fits the scratch cap); if any launch fails,
// the snapshot is copied back and the call reports the failure. No counterpart in the
// reference (its kernels and work-group sizes are fixed per vendor,
// src/cl_marshal.cc:165-231, 388-465).
// ---------------------------------------------------------------------------------------
// The route tag a step planned in a given form reports.
static const char *form_tag(int form) {
  switch (form) {
    case kForm010Bs: return "010_BS-";
    case kForm010Oop: return "010_OOP-";
    case kForm010Split: return "010_SPLIT-";
    case kForm010Rc: return "010_RC-";
    case kForm010Cyc: return "010_PTTWAC-";
    case kForm0100Cyc: return "0100_CYC-";
    case kForm0100Seg: return "0100_SEG-";
    case kForm0100Oop: return "0100_OOP-";
    case kForm0100Rc: return "0100_RC-";
    case kForm0100Lds: return "0100_LDS-";
    case kForm0100Bits: return "0100_PTTWAC-";
    case kForm0100Cl: return "0100_CL-";
  }
  return "";
}

static bool autotune_step(hipStream_t st, const hip_marshal_desc &d, ClassChoice *out,
                          int *route_code) {
  const size_t es = size_t(d.elem_bytes);
  unsigned char *buf = static_cast<unsigned char *>(d.ptr);
  const uint64_t h = uint64_t(d.rows), w = uint64_t(d.cols), t = uint64_t(d.tile);
  const bool is010 = d.op == HIP_MARSHAL_OP_AOS_ASTA || d.op == HIP_MARSHAL_OP_ASTA_AOS;
  const bool fwd = d.op == HIP_MARSHAL_OP_AOS_ASTA || d.op == HIP_MARSHAL_OP_SOA_ASTA;
  // 010(A, a, B) and its inverse 010(A, B, a); 100(R, C, b) and its inverse 100(C, R, b)
  const uint64_t A = h / t, a = is010 ? (fwd ? t : w) : 0, B = is010 ? (fwd ? w : t) : 0;
  const uint64_t R = is010 ? 0 : (fwd ? w : h / t), C = is010 ? 0 : (fwd ? h / t : w), b = t;
  const uint64_t S = b * es;
  if (is010 ? (a <= 1 || B <= 1 || A == 0) : (R <= 1 || C <= 1)) {  // nothing moves
    *out = ClassChoice{};
    if (route_code) *route_code = 0;
    return false;
  }
  if (!is010 && (S == 1 || S == 2 || S == 4 || S == 8 || S == 16) &&
      reinterpret_cast<uintptr_t>(buf) % S == 0) {
    // a single 100 over 1-16-byte super-elements is a full transpose: tune that shape
    Choice c;
    if (autotune_full(st, buf, S, R, C, true, &c)) return true;
    out->form = kFormModel;
    if (route_code) *route_code = c.stages;
    return false;
  }
  ClassKey key;
  std::vector<ClassChoice> cands;
  if (is010) {
    key = class_key_010(es, A, a, B);
    cands = {{kForm010Bs, 64}, {kForm010Bs, 192}, {kForm010Bs, 512}, {kForm010Oop, 0},
             {kForm010Split, 0}, {kForm010Rc, 0}, {kForm010Cyc, 0}};
  } else {
    key = class_key_0100(1, R, C, S);
    cands = {{kForm0100Cyc, 0}, {kForm0100Seg, 16}, {kForm0100Seg, 32}, {kForm0100Seg, 64},
             {kForm0100Oop, 0},  {kForm0100Rc, 0},   {kForm0100Lds, 0},  {kForm0100Bits, 0},
             {kForm0100Cl, 0}};
  }
  auto plan = [&](std::vector<StageRec> &v, std::string &tag, bool inverse) {
    if (is010) {
      const char *tg = "";
      const bool ok = inverse ? push_010(v, buf, es, A, B, a, M010::Auto, 1, 1, &tg, true)
                              : push_010(v, buf, es, A, a, B, M010::Auto, 1, 1, &tg, true);
      tag = tg;
      return ok;
    }
    return inverse ? plan_step_0100(v, buf, es, 1, C, R, b, true, tag)
                   : plan_step_0100(v, buf, es, 1, R, C, b, true, tag);
  };
  std::vector<StageRec> inv;
  std::string itag;
  if (!plan(inv, itag, true)) {
    fprintf(stderr, "[libmarshal] autotune: the inverse step cannot be planned\n");
    return true;
  }
  const uint64_t bytes = h * w * es;
  void *snap = nullptr;
  if (bytes <= (uint64_t(tuning("st_scratch_mb")) << 20) &&
      hipMalloc(&snap, bytes) == hipSuccess &&
      hipMemcpyAsync(snap, buf, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess) {
    (void)hipFree(snap);
    snap = nullptr;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool err = hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess;
  float best_ms = 1e30f;
  int best = -1;
  std::vector<std::string> seen;
  for (size_t i = 0; i < cands.size() && !err; i++) {
    std::vector<StageRec> v;
    std::string tag;
    t_force = &cands[i];
    const bool ok = plan(v, tag, false);
    t_force = nullptr;
    if (!ok || v.empty() || tag != form_tag(cands[i].form)) continue;  // cannot run here
    const std::string id = tag + ":" + std::to_string(cands[i].param);
    if (std::find(seen.begin(), seen.end(), id) != seen.end()) continue;
    seen.push_back(id);
    {
      Pool &p = pool_here();
      std::lock_guard<std::mutex> lk(p.mu);
      if (pool_grow(p, std::max(stages_scratch_bytes(v), stages_scratch_bytes(inv)))) {
        err = true;
        break;
      }
    }
    float ms = 0;
    if (bytes < (uint64_t(1) << 30))  // one untimed round trip: first launches of a form
      err = run_stages(v, st) || run_stages(inv, st);
    err = err || hipEventRecord(e0, st) != hipSuccess || run_stages(v, st) ||
          hipEventRecord(e1, st) != hipSuccess || run_stages(inv, st) ||
          hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess || hipStreamSynchronize(st) != hipSuccess;
    if (plan_debug())
      fprintf(stderr, "[libmarshal] autotune class %s: form %d param %lld %s %.3f ms\n",
              is010 ? "010" : "0100", cands[i].form, (long long)cands[i].param, tag.c_str(), ms);
    if (!err && ms < best_ms) {
      best_ms = ms;
      best = int(i);
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (err) {
    (void)hipStreamSynchronize(st);
    if (snap && hipMemcpy(buf, snap, bytes, hipMemcpyDeviceToDevice) == hipSuccess) {
      fprintf(stderr, "[libmarshal] autotune: a candidate failed; the buffer was restored\n");
    } else {
      fprintf(stderr, "[libmarshal] autotune: a candidate failed; the buffer was not "
                      "restored (%s)\n", snap ? "the copy back failed" : "larger than the scratch cap");
    }
  }
  if (snap) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(snap);
  }
  if (err) return true;
  if (best < 0) {  // nothing to move (unit tiles) or no form can run: the model decides
    *out = ClassChoice{};
    if (route_code) *route_code = 0;
    return false;
  }
  {
    std::lock_guard<std::mutex> lk(g_class_mu);
    g_class[key] = cands[size_t(best)];
  }
  *out = cands[size_t(best)];
  if (route_code) *route_code = 0;
  return false;
}

// Stage list of an M x N transpose as chosen by the planner (stages = 0).
bool plan_transpose_stages(std::vector<StageRec> &out, unsigned char *buf, size_t es,
                           uint64_t M, uint64_t N, std::string &route, bool allow_st) {
  return choice_stages(out, buf, es, M, N, plan_full(es, M, N, allow_st), route);
}

}  // namespace lm

using namespace lm;

// =======================================================================================
// C ABI
// =======================================================================================
extern "C" {

const char *hip_marshal_version(void) { return "libmarshal-amd 0.1 (gfx950)"; }

void hip_marshal_set_lds_budget(size_t bytes) { g_lds_budget.store(bytes); }

bool hip_marshal_plan_transpose(int64_t M, int64_t N, size_t es, int64_t *a, int64_t *b,
                                int *stages) {
  if (M < 1 || N < 1 || !valid_es(es)) return true;
  Choice c = plan_full(es, uint64_t(M), uint64_t(N));
  if (a) *a = int64_t(c.a);
  if (b) *b = int64_t(c.b);
  if (stages) *stages = c.stages;
  return false;
}

bool hip_marshal_autotune(hipStream_t st, void *buf, size_t es, int64_t M, int64_t N,
                          int *stages) {
  uint64_t n;
  if (!dims_ok("hip_marshal_autotune", {M, N}, es, &n) ||
      !aligned_ok("hip_marshal_autotune", buf, es))
    return true;
  Choice c = plan_full(es, uint64_t(M), uint64_t(N), true);
  if (M > 1 && N > 1 &&
      autotune_full(st, static_cast<unsigned char *>(buf), es, uint64_t(M), uint64_t(N), true, &c))
    return true;
  if (stages) *stages = c.stages;
  return false;
}

bool hip_marshal_autotune_desc(hipStream_t st, const hip_marshal_desc *d, int *form) {
  if (!d) return true;
  const size_t es = size_t(d->elem_bytes);
  uint64_t n;
  if (!dims_ok("hip_marshal_autotune_desc", {d->rows, d->cols}, es, &n) ||
      !aligned_ok("hip_marshal_autotune_desc", d->ptr, es))
    return true;
  if (d->op < HIP_MARSHAL_OP_TRANSPOSE || d->op > HIP_MARSHAL_OP_ASTA_SOA ||
      (d->op != HIP_MARSHAL_OP_TRANSPOSE && (d->tile < 1 || d->rows % d->tile))) {
    fprintf(stderr, "[libmarshal] hip_marshal_autotune_desc: invalid descriptor\n");
    return true;
  }
  if (d->op == HIP_MARSHAL_OP_TRANSPOSE) {
    int stages = 0;
    if (hip_marshal_autotune(st, d->ptr, es, d->rows, d->cols, &stages)) return true;
    if (form) *form = stages;
    return false;
  }
  ClassChoice c;
  int code = 0;
  if (autotune_step(st, *d, &c, &code)) return true;
  if (form) *form = c.form == kFormModel ? code : c.form;
  return false;
}

bool hip_marshal_set_tuning(const char *key, int64_t value) {
  Knob *k = find_knob(key);
  if (!k || value < k->min || value > (int64_t(1) << 30)) {
    fprintf(stderr, "[libmarshal] hip_marshal_set_tuning: bad key or value\n");
    return true;
  }
  k->value.store(value);
  // Route thresholds and caps change what the model's candidates are: plan again.
  std::lock_guard<std::mutex> lk(g_plan_mu);
  g_plan_cache.clear();
  return false;
}

int64_t hip_marshal_get_tuning(const char *key) { return tuning(key); }
size_t hip_marshal_get_lds_budget(void) { return lds_budget(); }
void hip_marshal_set_profiling(int on) { g_profiling.store(on ? 1 : 0); }

uint64_t hip_marshal_cluster_aborts(void) {
  if (hipDeviceSynchronize() != hipSuccess) return 0;
  return cluster_gave_up_new();
}
const char *hip_marshal_last_route(void) { return t_route.c_str(); }

bool hip_marshal_reserve(size_t bytes) {
  Pool &p = pool_here();
  std::lock_guard<std::mutex> lk(p.mu);
  return pool_grow(p, bytes);
}

uint64_t hip_marshal_cycle_count(int64_t R, int64_t C, uint32_t max_order, uint32_t *order) {
  PlPlan pc;
  if (R < 2 || C < 2 || !plan_100_leaders(uint64_t(R), uint64_t(C), max_order, &pc)) return 0;
  if (order) *order = pc.order;
  return pc.cycles;
}

size_t hip_marshal_flag_bytes_0100(int64_t A, int64_t a, int64_t B, int64_t b,
                                   size_t elem_bytes) {
  if (A < 1 || a < 1 || B < 1 || b < 1 || !valid_es(elem_bytes)) return 0;
  Pt100Plan pl{};
  plan_100(elem_bytes, uint64_t(A), uint64_t(a), uint64_t(B), uint64_t(b), &pl);
  return size_t(pl.flag_words * 4);
}

void hip_marshal_finalize(void) {
  free_segment_tables();
  free_cluster_flags();
  {  // measured choices go with the rest of the library's state
    std::lock_guard<std::mutex> lk(g_class_mu);
    g_class.clear();
  }
  {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    g_tuned.clear();
    g_plan_cache.clear();
  }
  {
    std::lock_guard<std::mutex> lk(g_lc_mu);
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (auto &kv : g_lc) {
      (void)hipSetDevice(std::get<0>(kv.first));
      if (kv.second.ready) {
        (void)hipEventSynchronize(kv.second.ready);
        (void)hipEventDestroy(kv.second.ready);
      }
      if (kv.second.list) (void)hipFree(kv.second.list);
    }
    g_lc.clear();
    g_lc_bytes = 0;
    (void)hipSetDevice(cur);
  }
  for (int d = 0; d < 64; d++) {
    Pool &p = g_pool[d];
    std::lock_guard<std::mutex> lk(p.mu);
    if (!p.ptr && !p.ev && p.retired.empty()) continue;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(d);
    if (p.has_ev) (void)hipEventSynchronize(p.ev);
    if (p.ptr) (void)hipFree(p.ptr);
    for (void *q : p.retired) (void)hipFree(q);
    p.retired.clear();
    p.captured = false;
    if (p.ev) (void)hipEventDestroy(p.ev);
    p.ptr = nullptr;
    p.cap = 0;
    p.ev = nullptr;
    p.has_ev = false;
    p.last = nullptr;
    (void)hipSetDevice(cur);
  }
}

// ---- element-size-generic entry points ---------------------------------------------------

bool hip_transpose_010_bs_es(hipStream_t st, void *src, size_t es, int64_t A, int64_t a,
                             int64_t B, uint64_t *ns) {
  uint64_t n;
  if (!dims_ok("hip_transpose_010_bs", {A, a, B}, es, &n) ||
      !aligned_ok("hip_transpose_010_bs", src, es))
    return true;
  Timer tm(st, "AOS-ASTA BS", 2.0 * n * es, ns);
  return tm.finish(stage_010(static_cast<unsigned char *>(src), es, A, a, B, M010::BS, 1, 0,
                             st));
}

bool hip_transpose_010_pttwac_es(hipStream_t st, void *src, size_t es, int64_t A, int64_t a,
                                 int64_t B, uint64_t *ns, int R, int P) {
  uint64_t n;
  if (!dims_ok("hip_transpose_010_pttwac", {A, a, B}, es, &n) ||
      !aligned_ok("hip_transpose_010_pttwac", src, es))
    return true;
  if (R < 1 || R > 32 || P < 0) {
    fprintf(stderr, "[libmarshal] hip_transpose_010_pttwac: need 1 <= R <= 32, P >= 0\n");
    return true;
  }
  Timer tm(st, "AOS-ASTA PTTWAC", 2.0 * n * es, ns);
  return tm.finish(stage_010(static_cast<unsigned char *>(src), es, A, a, B, M010::PTTWAC, R,
                             P, st));
}

bool hip_transpose_0100_es(hipStream_t st, void *src, size_t es, int64_t A, int64_t a,
                           int64_t B, int64_t b, uint64_t *ns) {
  uint64_t n;
  if (!dims_ok("hip_transpose_0100", {A, a, B, b}, es, &n) ||
      !aligned_ok("hip_transpose_0100", src, es))
    return true;
  Timer tm(st, "Transpose 0100/100 PTTWAC", 2.0 * n * es, ns);
  return tm.finish(stage_0100(static_cast<unsigned char *>(src), es, A, a, B, b, st));
}

bool hip_transpose_100_es(hipStream_t st, void *src, size_t es, int64_t A, int64_t B,
                          int64_t b, uint64_t *ns) {
  return hip_transpose_0100_es(st, src, es, 1, A, B, b, ns);
}

bool hip_transpose_es(hipStream_t st, void *src, size_t es, int64_t A, int64_t a, int64_t B,
                      int64_t b, int R, int stages, uint64_t *ns) {
  uint64_t n;
  if (!dims_ok("hip_transpose", {A, a, B, b}, es, &n) ||
      !aligned_ok("hip_transpose", src, es))
    return true;
  if (R < 1 || R > 32) {
    fprintf(stderr, "[libmarshal] hip_transpose: need 1 <= R <= 32\n");
    return true;
  }
  unsigned char *buf = static_cast<unsigned char *>(src);
  std::vector<StageRec> list;
  std::string route;
  if (stages == 0 && tuning("autotune")) {  // measure once per shape, on this buffer
    const uint64_t M = uint64_t(A) * a, N = uint64_t(B) * b;
    bool tuned;
    {
      std::lock_guard<std::mutex> lk(g_plan_mu);
      tuned = g_tuned.count(std::make_tuple(M, N, es, lds_budget(), true, planning_restricted())) != 0;
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(st, &cs);
    Choice c;
    if (!tuned && cs == hipStreamCaptureStatusNone && M > 1 && N > 1)
      autotune_full(st, buf, es, M, N, true, &c);
  }
  bool ok = stages == 0
                ? plan_transpose_stages(list, buf, es, uint64_t(A) * a, uint64_t(B) * b, route)
                : route_stages(list, buf, es, A, a, B, b, R, stages, route);
  t_route = route;
  if (!ok) return true;
  Timer tm(st, "Transpose", 2.0 * n * es, ns);
  return tm.finish(run_stages(list, st));
}

bool hip_transpose_full_es(hipStream_t st, void *src, size_t es, int64_t M, int64_t N,
                           uint64_t *ns) {
  return hip_transpose_es(st, src, es, 1, M, 1, N, 1, 0, ns);
}

static bool aos_common(const char *who, hipStream_t st, void *src, size_t es, int64_t h,
                       int64_t w, int64_t t, int method, uint64_t *ns, bool inverse) {
  uint64_t n;
  if (!dims_ok(who, {h, w, t}, es, &n) || !aligned_ok(who, src, es)) return true;
  if (h % t) {
    fprintf(stderr, "[libmarshal] %s: height %lld is not a multiple of tile %lld\n", who,
            (long long)h, (long long)t);
    return true;
  }
  n = uint64_t(h) * uint64_t(w);
  uint64_t A = uint64_t(h / t), a = uint64_t(t), B = uint64_t(w);
  if (inverse) std::swap(a, B);  // ASTA [h/t][w][t] -> AoS [h/t][t][w] is 010(h/t, w, t)
  Timer tm(st, inverse ? "ASTA-AOS" : "AOS-ASTA", 2.0 * n * es, ns);
  unsigned char *buf = static_cast<unsigned char *>(src);
  bool err;
  if (method == 1) {
    err = stage_010(buf, es, A, a, B, M010::BS, 1, 1, st);
  } else if (method == 2) {
    err = stage_010(buf, es, A, a, B, M010::PTTWAC, 1, 1, st);
  } else {
    // LDS-staged first, cycle following when it cannot run (src/cl_marshal.cc:814-818).
    err = stage_010(buf, es, A, a, B, M010::Auto, 1, 1, st);
  }
  return tm.finish(err);
}

bool hip_aos_asta_es(hipStream_t st, void *src, size_t es, int64_t h, int64_t w, int64_t t,
                     int method, uint64_t *ns) {
  return aos_common("hip_aos_asta", st, src, es, h, w, t, method, ns, false);
}

bool hip_asta_aos_es(hipStream_t st, void *src, size_t es, int64_t h, int64_t w, int64_t t,
                     int method, uint64_t *ns) {
  return aos_common("hip_asta_aos", st, src, es, h, w, t, method, ns, true);
}

bool hip_soa_asta_es(hipStream_t st, void *src, size_t es, int64_t h, int64_t w, int64_t t,
                     uint64_t *ns) {
  if (t < 1 || h % t) {
    fprintf(stderr, "[libmarshal] hip_soa_asta: height must be a positive multiple of tile\n");
    return true;
  }
  return hip_transpose_0100_es(st, src, es, 1, w, h / t, t, ns);
}

bool hip_asta_soa_es(hipStream_t st, void *src, size_t es, int64_t h, int64_t w, int64_t t,
                     uint64_t *ns) {
  if (t < 1 || h % t) {
    fprintf(stderr, "[libmarshal] hip_asta_soa: height must be a positive multiple of tile\n");
    return true;
  }
  return hip_transpose_0100_es(st, src, es, 1, h / t, w, t, ns);
}

static bool pad_common(const char *who, hipStream_t st, void *src, size_t es, int64_t x,
                       int64_t y, int64_t pad, bool unpad, uint64_t *ns) {
  uint64_t n;
  if (!dims_ok(who, {y, pad}, es, &n) || !aligned_ok(who, src, es)) return true;
  if (x < 1 || x > pad) {
    fprintf(stderr, "[libmarshal] %s: need 1 <= x_size <= pad_size\n", who);
    return true;
  }
  Timer tm(st, unpad ? "Unpadding" : "Padding", 2.0 * double(x) * y * es, ns);
  Lease lease(pool_here(), st);
  uint32_t *flags = lease.get(pad_flag_words(n, es) * 4);
  if (!flags) return tm.finish(true);
  hipError_t e = launch_padding(static_cast<unsigned char *>(src), es, x, y, pad, unpad, flags,
                                st);
  if (e != hipSuccess) {
    fprintf(stderr, "[libmarshal] %s: %s\n", who, hipGetErrorString(e));
    return tm.finish(true);
  }
  return tm.finish(false);
}

bool hip_padding_es(hipStream_t st, void *src, size_t es, int64_t x, int64_t y, int64_t pad,
                    uint64_t *ns) {
  return pad_common("hip_padding", st, src, es, x, y, pad, false, ns);
}

bool hip_unpadding_es(hipStream_t st, void *src, size_t es, int64_t x, int64_t y,
                      int64_t pad, uint64_t *ns) {
  return pad_common("hip_unpadding", st, src, es, x, y, pad, true, ns);
}

// ---- twins of include/cl_marshal.h (4-byte elements) ------------------------------------

bool hip_transpose_010_bs(hipStream_t st, void *src, int A, int a, int B, uint64_t *ns) {
  return hip_transpose_010_bs_es(st, src, 4, A, a, B, ns);
}

bool hip_transpose_010_pttwac(hipStream_t st, void *src, int A, int a, int B, uint64_t *ns,
                              int R, int P) {
  return hip_transpose_010_pttwac_es(st, src, 4, A, a, B, ns, R, P);
}

bool hip_transpose_100(hipStream_t st, void *src, int A, int B, int b, uint64_t *ns) {
  return hip_transpose_0100_es(st, src, 4, 1, A, B, b, ns);
}

bool hip_transpose_0100(hipStream_t st, void *src, int A, int a, int B, int b, uint64_t *ns) {
  return hip_transpose_0100_es(st, src, 4, A, a, B, b, ns);
}

bool hip_transpose(hipStream_t st, void *src, int A, int a, int B, int b, int R, int stages,
                   uint64_t *ns) {
  return hip_transpose_es(st, src, 4, A, a, B, b, R, stages, ns);
}

bool hip_aos_asta_bs(hipStream_t st, void *src, int height, int width, int tile_size) {
  if (tile_size < 1) return true;
  return hip_transpose_010_bs_es(st, src, 4, height / tile_size, tile_size, width, nullptr);
}

bool hip_aos_asta_pttwac(hipStream_t st, void *src, int height, int width, int tile_size,
                         int R, int P) {
  if (tile_size < 1 || height % tile_size) {
    fprintf(stderr, "[libmarshal] hip_aos_asta_pttwac: height must be a multiple of tile\n");
    return true;
  }
  return hip_transpose_010_pttwac_es(st, src, 4, height / tile_size, tile_size, width, nullptr,
                                     R, P);
}

bool hip_aos_asta(hipStream_t st, void *src, int height, int width, int tile_size, int R) {
  // BS first; only on failure the cycle-following form (src/cl_marshal.cc:814-818).
  return hip_aos_asta_bs(st, src, height, width, tile_size) &&
         hip_aos_asta_pttwac(st, src, height, width, tile_size, R, 1);
}

bool hip_padding(hipStream_t st, void *src, int x, int y, int pad, uint64_t *ns) {
  return hip_padding_es(st, src, 4, x, y, pad, ns);
}

bool hip_unpadding(hipStream_t st, void *src, int x, int y, int pad, uint64_t *ns) {
  return hip_unpadding_es(st, src, 4, x, y, pad, ns);
}

// ---- twins of include/marshal.h (null stream, asynchronous) -----------------------------

static bool gpu_010(void *src, size_t es, int h, int w, int t, M010 m) {
  if (t < 1 || h < 1 || w < 1 || h % t) {
    fprintf(stderr, "[libmarshal] gpu_aos_asta: height must be a positive multiple of tile\n");
    return true;
  }
  if (!aligned_ok("gpu_aos_asta", src, es)) return true;
  return stage_010(static_cast<unsigned char *>(src), es, uint64_t(h / t), uint64_t(t),
                   uint64_t(w), m, 1, 0, nullptr);
}

bool gpu_aos_asta_bs_float(float *src, int h, int w, int t, clock_t *) {
  return gpu_010(src, 4, h, w, t, M010::BS);
}
bool gpu_aos_asta_bs_double(double *src, int h, int w, int t, clock_t *) {
  return gpu_010(src, 8, h, w, t, M010::BS);
}
bool gpu_aos_asta_pttwac(float *src, int h, int w, int t, clock_t *) {
  return gpu_010(src, 4, h, w, t, M010::PTTWAC);
}
bool gpu_soa_asta_pttwac(float *src, int h, int w, int t, clock_t *) {
  if (t < 1 || h < 1 || w < 1 || h % t) {
    fprintf(stderr, "[libmarshal] gpu_soa_asta_pttwac: height must be a positive multiple of tile\n");
    return true;
  }
  if (!aligned_ok("gpu_soa_asta_pttwac", src, 4)) return true;
  return stage_0100(reinterpret_cast<unsigned char *>(src), 4, 1, uint64_t(w),
                    uint64_t(h / t), uint64_t(t), nullptr);
}

}  // extern "C"
