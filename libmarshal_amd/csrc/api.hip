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
#include <cstring>
#include <mutex>
#include <string>
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
}  // namespace

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
};
Knob g_knobs[] = {
    {"bs_upt", {16}},
    {"bs_blocks_per_cu", {12}},
    {"pt010_blocks_per_cu", {8}},
    {"pt100_blocks_per_cu", {16}},
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
    LM_CHECK_HIP(hipFree(p.ptr));
    p.ptr = nullptr;
    p.cap = 0;
  }
  size_t want = std::max<size_t>(bytes + bytes / 4, size_t(4) << 20);
  want = (want + 4095) & ~size_t(4095);
  LM_CHECK_HIP(hipMalloc(&p.ptr, want));
  p.cap = want;
  return false;
}

struct Lease {
  Pool &p;
  std::unique_lock<std::mutex> lk;
  hipStream_t st;
  bool ok = false;
  Lease(Pool &pool, hipStream_t s) : p(pool), lk(pool.mu), st(s) {}
  // Returns a pointer to at least `bytes` of flag space, ordered after earlier users.
  uint32_t *get(size_t bytes) {
    if (pool_grow(p, bytes)) return nullptr;
    if (p.has_ev && p.last != st) {
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      hipStreamIsCapturing(st, &cs);
      if (cs == hipStreamCaptureStatusNone) hipStreamWaitEvent(st, p.ev, 0);
    }
    ok = true;
    return static_cast<uint32_t *>(p.ptr);
  }
  ~Lease() {
    if (!ok) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    hipStreamIsCapturing(st, &cs);
    if (cs != hipStreamCaptureStatusNone) return;
    if (!p.ev) hipEventCreateWithFlags(&p.ev, hipEventDisableTiming);
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
    hipStreamIsCapturing(st, &cs);
    if (cs != hipStreamCaptureStatusNone) { on = false; return; }
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      on = false;
      return;
    }
    hipEventRecord(e0, st);
  }
  bool finish(bool err) {
    if (!on) return err;
    bool ok = !err && hipEventRecord(e1, st) == hipSuccess &&
              hipEventSynchronize(e1) == hipSuccess;
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
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    on = false;
    return err;
  }
  ~Timer() {
    if (on) {
      hipEventDestroy(e0);
      hipEventDestroy(e1);
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

uint32_t grid_cap(uint32_t per_cu) { return uint32_t(device_info().cus) * per_cu; }

}  // namespace

// ---------------------------------------------------------------------------------------
// Stages (no timing, no argument checks beyond what the kernels need)
// ---------------------------------------------------------------------------------------
enum class M010 { Auto, BS, PTTWAC, Global };

static bool stage_0100(unsigned char *buf, size_t es, uint64_t inst, uint64_t R, uint64_t C,
                       uint64_t b, hipStream_t st) {
  if (inst == 0 || R * C <= 1) return false;
  if (R == 1 || C == 1) return false;  // 1 x C or R x 1 grids are their own transpose
  Pt100Plan pl{};
  plan_100(es, inst, R, C, b, &pl);
  Lease lease(pool_here(), st);
  uint32_t *flags = lease.get(pl.flag_words * 4);
  if (!flags) return true;
  LM_CHECK_HIP(launch_fill_u32(flags, pl.flag_words, 0, st));
  LM_CHECK_HIP(launch_100_pttwac(buf, flags, inst, R, C, b * es, pl,
                                 grid_cap(uint32_t(tuning("pt100_blocks_per_cu"))), st));
  return false;
}

// Which form a 010 step can use under the current LDS budget.
static M010 choose_010(size_t es, uint64_t A, uint64_t a, uint64_t B, int R, int P,
                       const void *buf) {
  Bs010Plan pl{};
  if (plan_010_bs(es, A, a, B, lds_budget(), reinterpret_cast<uintptr_t>(buf), int(tuning("bs_upt")),
                  &pl))
    return M010::BS;
  if (a * B < (1ull << 31) && pt010_flag_words(a, B, R, P) * 4 <= lds_budget())
    return M010::PTTWAC;
  return M010::Global;
}

static bool stage_010(unsigned char *buf, size_t es, uint64_t A, uint64_t a, uint64_t B,
                      M010 m, int R, int P, hipStream_t st, const char **tag) {
  if (A == 0 || a <= 1 || B <= 1) {  // 1 x B and a x 1 tiles are unchanged by 010
    if (tag) *tag = "";
    return false;
  }
  if (m == M010::Auto) m = choose_010(es, A, a, B, R, P, buf);
  if (m == M010::BS) {
    Bs010Plan pl{};
    if (!plan_010_bs(es, A, a, B, lds_budget(), reinterpret_cast<uintptr_t>(buf),
                     int(tuning("bs_upt")), &pl)) {
      fprintf(stderr, "[libmarshal] 010 BS: tile %llux%llu does not fit the LDS budget\n",
              (unsigned long long)a, (unsigned long long)B);
      return true;
    }
    if (tag) *tag = "010_BS-";
    LM_CHECK_HIP(launch_010_bs(buf, es, A, a, B, pl,
                               grid_cap(uint32_t(tuning("bs_blocks_per_cu"))), st));
    return false;
  }
  if (m == M010::PTTWAC) {
    if (R < 1 || P < 0 || a * B >= (1ull << 31) ||
        pt010_flag_words(a, B, R, P) * 4 > lds_budget()) {
      fprintf(stderr, "[libmarshal] 010 PTTWAC: flags of a %llux%llu tile do not fit LDS\n",
              (unsigned long long)a, (unsigned long long)B);
      return true;
    }
    if (tag) *tag = "010_PTTWAC-";
    LM_CHECK_HIP(launch_010_pttwac(buf, es, A, a, B, R, P,
                                   grid_cap(uint32_t(tuning("pt010_blocks_per_cu"))), st));
    return false;
  }
  // 010 on [A][a][B] is 0100 with single-element super-elements: global visited bits.
  if (tag) *tag = "010_G-";
  return stage_0100(buf, es, A, a, B, 1, st);
}

// ---------------------------------------------------------------------------------------
// Composer: stage sequences of cl_transpose (src/cl_marshal.cc:561-776).
// ---------------------------------------------------------------------------------------
static bool compose(unsigned char *buf, size_t es, uint64_t A, uint64_t a, uint64_t B,
                    uint64_t b, int R, int stages, hipStream_t st, std::string &route) {
  route.clear();
  const char *tag = "";
  auto s010 = [&](uint64_t A_, uint64_t a_, uint64_t B_) {
    bool e = stage_010(buf, es, A_, a_, B_, M010::Auto, R, 1, st, &tag);
    route += tag;
    return e;
  };
  switch (stages) {
    case 2:  // AaBb -> ABba: 010(A, a, Bb), then 100(A, Bb, a)   (src/cl_marshal.cc:569-601)
      if (s010(A, a, B * b)) return true;
      if (stage_0100(buf, es, 1, A, B * b, a, st)) return true;
      route += "2.1-";
      return false;
    case 22:  // 100(Aa, B, b), then 010(B, Aa, b)               (src/cl_marshal.cc:602-634)
      if (stage_0100(buf, es, 1, A * a, B, b, st)) return true;
      if (s010(B, A * a, b)) return true;
      route += "2.2-";
      return false;
    case 3:  // AaBb -> BAab -> BAba -> BbAa                      (src/cl_marshal.cc:636-677)
      if (a > 1 && b > 1) {
        if (stage_0100(buf, es, 1, A * a, B, b, st)) return true;
        if (s010(B * A, a, b)) return true;
        if (stage_0100(buf, es, B, A, b, a, st)) return true;
        route += "3.1-";
        return false;
      }
      break;
    case 32:  // AaBb -> ABab -> ABba -> BbAa                     (src/cl_marshal.cc:679-720)
      if (a > 1 && b > 1) {
        if (stage_0100(buf, es, A, a, B, b, st)) return true;
        if (s010(A * B, a, b)) return true;
        if (stage_0100(buf, es, 1, A, B * b, a, st)) return true;
        route += "3.2-";
        return false;
      }
      break;
    case 4:  // Karlsson: 0100, 010, 100(A, B, ba), 0100          (src/cl_marshal.cc:722-767)
      if (stage_0100(buf, es, A, a, B, b, st)) return true;
      if (s010(B * A, a, b)) return true;
      if (stage_0100(buf, es, 1, A, B, b * a, st)) return true;
      if (stage_0100(buf, es, B, A, b, a, st)) return true;
      route += "4-";
      return false;
    default:
      break;
  }
  // Element-granular fallback: 100 over the whole matrix (src/cl_marshal.cc:769-775).
  route += "Fallback-";
  return stage_0100(buf, es, 1, A * a, B * b, 1, st);
}

// ---------------------------------------------------------------------------------------
// Planner: tile factors and route for an M x N transpose.
// Cost model in "HBM pass equivalents": an LDS-staged pass costs 1; a cycle-following pass
// over super-elements of S bytes costs 1 + 96/S (short moves waste most of each line);
// LDS-flag and global-flag 010 passes move single elements and cost like S = es.
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

double cost_pt(uint64_t S) { return 1.0 + 96.0 / double(S); }

double cost_010(size_t es, uint64_t A, uint64_t a, uint64_t B) {
  if (a <= 1 || B <= 1) return 0.0;
  Bs010Plan pl{};
  if (plan_010_bs(es, A, a, B, lds_budget(), 0, 16, &pl)) return 1.0;
  return 3.0 * cost_pt(es);
}

struct Choice {
  double cost = 1e300;
  uint64_t a = 1, b = 1;
  int stages = 0;
};
}  // namespace

static Choice plan_full(size_t es, uint64_t M, uint64_t N) {
  Choice best;
  best.cost = cost_pt(es);  // fallback
  best.a = M;
  best.b = N;
  best.stages = -1;
  std::vector<uint64_t> dm = divisors(M), dn = divisors(N);
  for (uint64_t a : dm) {
    uint64_t A = M / a;
    // route 2: 010(A, a, N) + 100(A, N, a)
    if (a > 1) {
      double c = cost_010(es, A, a, N) + cost_pt(a * es);
      if (c < best.cost) best = {c, a, N, 2};
    }
    for (uint64_t b : dn) {
      uint64_t B = N / b;
      if (a == 1 && b > 1) {
        // route 22 with a = M: 100(M, B, b) + 010(B, M, b)
        double c = cost_pt(b * es) + cost_010(es, B, M, b);
        if (c < best.cost) best = {c, M, b, 22};
      }
      if (a > 1 && b > 1 && a < M && b < N) {
        double c3 = cost_pt(b * es) + cost_010(es, B * A, a, b) + cost_pt(a * es);
        if (c3 < best.cost) best = {c3, a, b, 3};
      }
    }
  }
  return best;
}

}  // namespace lm

using namespace lm;

// =======================================================================================
// C ABI
// =======================================================================================
extern "C" {

const char *hip_marshal_version(void) { return "libmarshal-amd 0.1 (gfx950)"; }

void hip_marshal_set_lds_budget(size_t bytes) { g_lds_budget.store(bytes); }

bool hip_marshal_set_tuning(const char *key, int64_t value) {
  Knob *k = find_knob(key);
  if (!k || value < 1 || value > 1024) {
    fprintf(stderr, "[libmarshal] hip_marshal_set_tuning: bad key or value\n");
    return true;
  }
  k->value.store(value);
  return false;
}

int64_t hip_marshal_get_tuning(const char *key) { return tuning(key); }
size_t hip_marshal_get_lds_budget(void) { return lds_budget(); }
void hip_marshal_set_profiling(int on) { g_profiling.store(on ? 1 : 0); }
const char *hip_marshal_last_route(void) { return t_route.c_str(); }

bool hip_marshal_reserve(size_t bytes) {
  Pool &p = pool_here();
  std::lock_guard<std::mutex> lk(p.mu);
  return pool_grow(p, bytes);
}

size_t hip_marshal_flag_bytes_0100(int64_t A, int64_t a, int64_t B, int64_t b,
                                   size_t elem_bytes) {
  if (A < 1 || a < 1 || B < 1 || b < 1 || !valid_es(elem_bytes)) return 0;
  Pt100Plan pl{};
  plan_100(elem_bytes, uint64_t(A), uint64_t(a), uint64_t(B), uint64_t(b), &pl);
  return size_t(pl.flag_words * 4);
}

void hip_marshal_finalize(void) {
  for (int d = 0; d < 64; d++) {
    Pool &p = g_pool[d];
    std::lock_guard<std::mutex> lk(p.mu);
    if (!p.ptr && !p.ev) continue;
    int cur = 0;
    hipGetDevice(&cur);
    hipSetDevice(d);
    if (p.has_ev) hipEventSynchronize(p.ev);
    if (p.ptr) hipFree(p.ptr);
    if (p.ev) hipEventDestroy(p.ev);
    p.ptr = nullptr;
    p.cap = 0;
    p.ev = nullptr;
    p.has_ev = false;
    p.last = nullptr;
    hipSetDevice(cur);
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
                             st, nullptr));
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
                             P, st, nullptr));
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
  uint64_t A_ = A, a_ = a, B_ = B, b_ = b;
  int stg = stages;
  if (stg == 0) {
    Choice c = plan_full(es, A_ * a_, B_ * b_);
    if (c.stages == -1) {
      stg = -1;
    } else {
      stg = c.stages;
      if (stg == 2) { A_ = A_ * a_ / c.a; a_ = c.a; b_ = B_ * b_; B_ = 1; }
      else if (stg == 22) { b_ = c.b; B_ = uint64_t(B) * b / c.b; A_ = 1; a_ = uint64_t(A) * a; }
      else { uint64_t M = A_ * a_, N = B_ * b_; a_ = c.a; A_ = M / c.a; b_ = c.b; B_ = N / c.b; }
    }
  }
  Timer tm(st, "Transpose", 2.0 * n * es, ns);
  std::string route;
  bool err = compose(static_cast<unsigned char *>(src), es, A_, a_, B_, b_, R, stg, st, route);
  t_route = route;
  return tm.finish(err);
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
    err = stage_010(buf, es, A, a, B, M010::BS, 1, 1, st, nullptr);
  } else if (method == 2) {
    err = stage_010(buf, es, A, a, B, M010::PTTWAC, 1, 1, st, nullptr);
  } else {
    // LDS-staged first, cycle following when it cannot run (src/cl_marshal.cc:814-818).
    err = stage_010(buf, es, A, a, B, M010::Auto, 1, 1, st, nullptr);
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
  uint32_t *flags = lease.get(pad_flag_words(n) * 4);
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
                   uint64_t(w), m, 1, 0, nullptr, nullptr);
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
