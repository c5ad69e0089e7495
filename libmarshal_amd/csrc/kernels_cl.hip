// kernels_cl.hip - 100 / 0100 steps whose instances fit on a few compute units: a cluster of
// G workgroups holds one instance ([R][C] super-elements of S bytes) in LDS and registers,
// meets at a barrier, and writes the instance back as [C][R]. Every byte is read once and
// written once, in place, with no cycle following, no visited bits and no side buffer.
//
// Not a reference algorithm: the reference moves such grids by cycle following with visited
// bits (_transpose_100, src/cl/cl_aos_asta.cl:348-569, behind transpose_100 and
// transpose_0100 at :721-761). It is used for super-elements of at least 64 bytes, where each
// super-element is one contiguous destination run.
//
// Workgroups take cluster slots by a dynamic id (an atomic counter, taken when the workgroup
// starts), so ids follow the order in which workgroups became resident: at any time only the
// cluster holding the highest id taken can be incomplete, and its missing members start as
// soon as any resident workgroup of a complete cluster exits. The barrier wait therefore
// ends as long as the device can hold G workgroups of this launch at once (G is capped at 16
// and the launch needs one CU per workgroup), whatever else runs beside it. As a safety net
// the wait is bounded: a cluster whose members have not all arrived after the bound gives up
// as a whole (no member of it stores anything, so its instance is left unchanged).
//
// Fused form (FES > 0): the instance is [R][a][b] elements of FES bytes, R tiles of a x b, and
// the step is the 010 of every tile followed by the 0100 of the [R][b] grid of a-element
// super-elements, i.e. [R][a][b] -> [b][R][a] (the reference's 3-stage route runs these as its
// stages 2 and 3). Members hold whole tiles in LDS only; a tile's elements are written to LDS
// already transposed ([b][a]), so what LDS holds is the 010's output, and the store is the
// 0100's. One pass instead of two.
//
// The launch leaves its barrier words zero again (the last member to leave a cluster's
// barrier clears that cluster's words, the last workgroup to finish clears the counters), so
// a buffer kept for one stream needs clearing only once, when it is allocated.
#include "kernels.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

namespace lm {

namespace {
constexpr uint32_t kClAbort = 0x80000000u;
constexpr int kClThreads = 1024;
// Pieces held in registers per lane (16-byte pieces: 12, so that the kernel stays within the
// 128 VGPRs a lane has when 16 waves share a CU, without spilling).
constexpr int cl_regs(uint32_t ub) { return ub == 16 ? 12 : 16; }

struct ClParams {
  unsigned char *base;
  uint32_t *flags;      // [0]: dynamic workgroup ids, [1]: workgroups done, [2 + c]: arrivals
                        // at cluster c's barrier, [2 + inst + c]: members past it
  uint32_t *gave_up;    // the device's count of clusters that gave up their wait (or null)
  uint64_t inst_bytes;  // R * C * S
  uint64_t wgs;         // instances * G (the grid)
  uint64_t inst;
  uint32_t R, C, S, G;
  uint32_t chunk;       // bytes per workgroup, a multiple of the piece size
  uint32_t lds_pieces;  // pieces staged in LDS per workgroup (the rest in registers)
  uint32_t spin;        // bound of the barrier wait, in polls
  uint32_t fa, fb;      // fused form: tiles of fa x fb elements (the 010 before the 0100)
  FastDiv32 div_S, div_C, div_G, div_ab, div_b;
};

// Instance byte offset o of a piece -> its destination offset: super-element u = o / S at row
// r = u / C, column c = u - r * C, byte w = o - u * S goes to (c * R + r) * S + w.
__device__ __forceinline__ uint64_t cl_dest(const ClParams &p, uint32_t o) {
  const uint32_t u = p.div_S.div(o), w = o - u * p.S;
  const uint32_t r = p.div_C.div(u), c = u - r * p.C;
  return (uint64_t(c) * p.R + r) * p.S + w;
}

// Store: every piece to its destination; consecutive lanes write consecutive pieces of a
// super-element, so each super-element is one contiguous run of S bytes.
template <int UB, int NT, int NR>
__device__ __forceinline__ void cl_store(const ClParams &p, unsigned char *ib,
                                         const typename Elem<UB>::T *L,
                                         const typename Elem<UB>::T (&reg)[NR], uint32_t lo,
                                         uint32_t nl, uint32_t np) {
  using V = typename Elem<UB>::T;
  for (uint32_t i = threadIdx.x; i < nl; i += NT)
    *reinterpret_cast<V *>(ib + cl_dest(p, lo + i * UB)) = L[i];
#pragma unroll
  for (int k = 0; k < NR; k++) {
    const uint32_t i = nl + uint32_t(k) * NT + threadIdx.x;
    if (i < np) *reinterpret_cast<V *>(ib + cl_dest(p, lo + i * UB)) = reg[k];
  }
}

// Fused form's load: the chunk (whole tiles of a x b elements) arrives as FLV-byte vectors
// (each inside one tile row, since b * FES is a multiple of FLV) and every element is written
// to LDS at its place in the transposed tile: tile t, element (i, j) -> t * a * b + j * a + i.
template <int FES, int FLV, int NT>
__device__ __forceinline__ void cl_load_fused(const ClParams &p, const unsigned char *src,
                                              unsigned char *lds, uint32_t bytes) {
  using E = typename Elem<FES>::T;
  using LV = typename Elem<FLV>::T;
  constexpr uint32_t EPV = FLV / FES;
  const uint32_t nv = bytes / FLV;
  const LV *sv = reinterpret_cast<const LV *>(src);
  E *L = reinterpret_cast<E *>(lds);
  const uint32_t a = p.fa, ab = p.fa * p.fb;
  constexpr int KB = 8;
  for (uint32_t i0 = threadIdx.x; i0 < nv; i0 += KB * NT) {
    union {
      LV v;
      E e[EPV];
    } t[KB];
#pragma unroll
    for (int k = 0; k < KB; k++) {
      const uint32_t i = i0 + uint32_t(k) * NT;
      if (i < nv) t[k].v = sv[i];
    }
#pragma unroll
    for (int k = 0; k < KB; k++) {
      const uint32_t i = i0 + uint32_t(k) * NT;
      if (i >= nv) continue;
      const uint32_t e = i * EPV;  // first element of the vector within the chunk
      const uint32_t tl = p.div_ab.div(e), rem = e - tl * ab;
      const uint32_t ri = p.div_b.div(rem), cj = rem - ri * p.fb;
      const uint32_t o = tl * ab + cj * a + ri;
#pragma unroll
      for (uint32_t x = 0; x < EPV; x++) L[o + x * a] = t[k].e[x];
    }
  }
}

template <int UB, int NT, int FES, int FLV>
__global__ __launch_bounds__(NT) void k_100_cluster(ClParams p) {
  using V = typename Elem<UB>::T;
  extern __shared__ __align__(16) unsigned char lds[];
  __shared__ uint32_t s_id, s_go;
  if (threadIdx.x == 0)
    s_id = __hip_atomic_fetch_add(p.flags, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const uint32_t id = s_id;
  if (id >= p.wgs) return;  // never: the grid has exactly `wgs` workgroups (and ids start at 0)
  const uint32_t cl = p.div_G.div(id), mem = id - cl * p.G;
  unsigned char *ib = p.base + uint64_t(cl) * p.inst_bytes;
  const uint32_t lo = mem * p.chunk;  // < inst_bytes for every member
  const uint32_t hi = uint32_t(std::min<uint64_t>(uint64_t(lo) + p.chunk, p.inst_bytes));
  const uint32_t np = (hi - lo) / UB;  // S is a multiple of UB, so is every chunk end
  const uint32_t nl = np < p.lds_pieces ? np : p.lds_pieces;
  const V *src = reinterpret_cast<const V *>(ib + lo);
  V *L = reinterpret_cast<V *>(lds);

  // Load: the registers' share first (no waits), then the LDS share in batches of KB pieces
  // per lane, every load of a batch issued before any of them is waited for. (Fused form:
  // everything in LDS, transposed per tile on the way in.)
  constexpr int NR = FES ? 1 : cl_regs(UB);
  V reg[NR];
  if (FES) {
    cl_load_fused<FES ? FES : 4, FLV ? FLV : 4, NT>(p, ib + lo, lds, hi - lo);
  } else {
#pragma unroll
    for (int k = 0; k < NR; k++) {
      const uint32_t i = nl + uint32_t(k) * NT + threadIdx.x;
      if (i < np) reg[k] = src[i];
    }
  }
  constexpr int KB = 8;
  for (uint32_t i0 = threadIdx.x; !FES && i0 < nl; i0 += KB * NT) {
    V t[KB];
#pragma unroll
    for (int k = 0; k < KB; k++) {
      const uint32_t i = i0 + uint32_t(k) * NT;
      if (i < nl) t[k] = src[i];
    }
#pragma unroll
    for (int k = 0; k < KB; k++) {
      const uint32_t i = i0 + uint32_t(k) * NT;
      if (i < nl) L[i] = t[k];
    }
  }
  wait_vm_loads();  // every piece of this workgroup has arrived
  __syncthreads();

  // Cluster barrier: all members have loaded before any member stores. All or none of the
  // members store: a member that gives up sets the abort bit, only while the count is short.
  if (threadIdx.x == 0) {
    uint32_t *w = p.flags + 2 + cl;
    const uint32_t old = __hip_atomic_fetch_add(w, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t go = (old & kClAbort) ? 0u : 1u;
    for (uint32_t n = 0; go;) {
      uint32_t v = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (v & kClAbort) {
        go = 0;
      } else if (v >= p.G) {
        break;
      } else if (n >= p.spin) {
        if (__hip_atomic_compare_exchange_strong(w, &v, v | kClAbort, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          go = 0;
          // counted once per cluster (by the member that set the bit), for the host to report
          if (p.gave_up)
            __hip_atomic_fetch_add(p.gave_up, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      } else {
        n++;
        __builtin_amdgcn_s_sleep(2);
      }
    }
    s_go = go;
    // past the barrier (or given up): the last member to get here clears the cluster's words
    uint32_t *lv = p.flags + 2 + p.inst + cl;
    if (__hip_atomic_fetch_add(lv, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 == p.G) {
      __hip_atomic_store(w, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(lv, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __syncthreads();
  if (s_go) cl_store<UB, NT>(p, ib, L, reg, lo, nl, FES ? nl : np);
  // the last workgroup of the launch clears the counters (every workgroup took its id first)
  if (threadIdx.x == 0 &&
      __hip_atomic_fetch_add(p.flags + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 ==
          p.wgs) {
    __hip_atomic_store(p.flags, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p.flags + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}


// Clusters that gave up their barrier wait, counted per device since the library was loaded
// (the instances they held were left unconverted). Its address is looked up once per device,
// with that device current; cluster_gave_up_new reports the growth since its last call.
static __device__ uint32_t cl_gave_up;
std::mutex g_gu_mu;
uint32_t *g_gu_addr[64] = {};
uint32_t g_gu_seen[64] = {};
bool g_gu_used[64] = {};

uint32_t *gave_up_here(int dev) {
  std::lock_guard<std::mutex> lk(g_gu_mu);
  if (dev < 0 || dev >= 64) return nullptr;
  if (!g_gu_addr[dev]) {
    void *a = nullptr;
    if (hipGetSymbolAddress(&a, HIP_SYMBOL(cl_gave_up)) != hipSuccess || !a) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_gu_addr[dev] = static_cast<uint32_t *>(a);
  }
  g_gu_used[dev] = true;
  return g_gu_addr[dev];
}

size_t cl_lds_bytes() {
  // LDS per workgroup: the device's limit, less room for the two static words, rounded down
  const size_t dev = device_info().lds_per_block;
  return dev > 1024 ? (dev - 256) / 16 * 16 : 0;
}
}  // namespace

bool plan_100_cluster(uint64_t inst, uint64_t R, uint64_t C, uint64_t S, uint64_t base_addr,
                      uint64_t max_wgs, ClPlan *out) {
  if (inst == 0 || R <= 1 || C <= 1 || S < 64) return false;
  const uint32_t ub = (S % 16 == 0 && base_addr % 16 == 0) ? 16
                      : (S % 8 == 0 && base_addr % 8 == 0) ? 8
                      : (S % 4 == 0 && base_addr % 4 == 0) ? 4 : 0;
  if (!ub) return false;
  const double ib = double(R) * double(C) * double(S);
  if (ib >= double(1u << 31)) return false;
  const uint64_t inst_bytes = R * C * S;
  const uint64_t lds = cl_lds_bytes();
  const uint64_t cap = lds + uint64_t(cl_regs(ub)) * ub * kClThreads;  // bytes per workgroup
  if (lds == 0) return false;
  // As few members as hold the instance: measured on 100 instances of 100 x 72 super-elements
  // of 72 bytes, splitting them over 3, 4, 8 or 16 workgroups instead of 2 took 11 %, 11 %,
  // 36 % and 84 % longer (each workgroup's phases are latency-bound, not bandwidth-bound).
  const uint64_t g0 = (inst_bytes + cap - 1) / cap;
  if (g0 > max_wgs || g0 == 0) return false;
  uint64_t chunk = (inst_bytes + g0 - 1) / g0;
  chunk = (chunk + ub - 1) / ub * ub;
  const uint64_t G = (inst_bytes + chunk - 1) / chunk;  // every member's chunk starts inside
  if (chunk > cap || G > max_wgs || inst * G >= (1ull << 31)) return false;
  out->G = uint32_t(G);
  out->chunk = uint32_t(chunk);
  out->ub = ub;
  out->lds_pieces = uint32_t(std::min<uint64_t>(lds, chunk) / ub);
  out->lds_bytes = size_t(out->lds_pieces) * ub;
  return true;
}

bool plan_cluster_fused(uint64_t inst, uint64_t R, uint64_t a, uint64_t b, size_t es,
                        uint64_t base_addr, uint64_t max_wgs, ClPlan *out) {
  if (inst == 0 || R <= 1 || a <= 1 || b <= 1 || !(es == 4 || es == 8)) return false;
  const uint64_t S = a * es, row = b * es, tile = a * b * es;
  if (S < 64 || base_addr % es) return false;
  const uint32_t ub = (S % 16 == 0 && base_addr % 16 == 0) ? 16
                      : (S % 8 == 0 && base_addr % 8 == 0) ? 8 : 4;
  const uint32_t lv = (row % 16 == 0 && base_addr % 16 == 0) ? 16
                      : (row % 8 == 0 && base_addr % 8 == 0) ? 8 : 0;
  if (!lv || ub < es) return false;
  const double ib = double(R) * double(tile);
  if (ib >= double(1u << 31)) return false;
  const uint64_t lds = cl_lds_bytes();
  if (lds == 0 || tile > lds) return false;
  const uint64_t per = lds / tile;               // tiles a member can hold
  const uint64_t G = (R + per - 1) / per;
  const uint64_t tpm = (R + G - 1) / G;          // tiles per member, every chunk non-empty
  if (G > max_wgs || inst * G >= (1ull << 31)) return false;
  out->G = uint32_t(G);
  out->chunk = uint32_t(tpm * tile);
  out->ub = ub;
  out->lds_pieces = uint32_t(tpm * tile / ub);
  out->lds_bytes = size_t(tpm * tile);
  out->fa = uint32_t(a);
  out->fb = uint32_t(b);
  out->fes = uint32_t(es);
  out->flv = lv;
  return true;
}

uint64_t cluster_flag_words(uint64_t inst) { return 2 + 2 * inst; }

uint64_t cluster_gave_up_new() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  uint32_t *a = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_gu_mu);
    if (!g_gu_used[dev]) return 0;  // no cluster launch on this device yet
    a = g_gu_addr[dev];
  }
  uint32_t v = 0;
  if (!a || hipMemcpy(&v, a, sizeof(v), hipMemcpyDeviceToHost) != hipSuccess) return 0;
  std::lock_guard<std::mutex> lk(g_gu_mu);
  const uint32_t d = v - g_gu_seen[dev];
  g_gu_seen[dev] = v;
  return d;
}

// Barrier words kept per (device, stream): zero whenever no launch of this file runs on the
// stream, since every launch leaves them zero. A larger need replaces the buffer (the old one
// is kept until free_cluster_flags, as launches on the stream may still use it).
namespace {
struct ClFlags {
  uint32_t *ptr = nullptr;
  uint64_t words = 0;
};
std::mutex g_clf_mu;
std::map<std::pair<int, hipStream_t>, ClFlags> g_clf;
std::vector<std::pair<int, uint32_t *>> g_clf_retired;
}  // namespace

void free_cluster_flags() {
  std::lock_guard<std::mutex> lk(g_clf_mu);
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto &kv : g_clf) {
    (void)hipSetDevice(kv.first.first);
    (void)hipDeviceSynchronize();
    (void)hipFree(kv.second.ptr);
  }
  for (auto &r : g_clf_retired) {
    (void)hipSetDevice(r.first);
    (void)hipFree(r.second);
  }
  g_clf.clear();
  g_clf_retired.clear();
  (void)hipSetDevice(cur);
}

// The stream's kept buffer of at least `words` zero words (cleared here when new), or null
// under stream capture (a graph must own its scratch) or when allocation fails.
static uint32_t *stream_cluster_flags(uint64_t words, hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
    return nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_clf_mu);
  ClFlags &f = g_clf[{dev, st}];
  if (f.words >= words) return f.ptr;
  const uint64_t want = std::max<uint64_t>(words, 4096);
  uint32_t *p = nullptr;
  if (hipMalloc(&p, want * 4) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  if (launch_fill_u32(p, want, 0, st) != hipSuccess) {
    (void)hipFree(p);
    return nullptr;
  }
  if (f.ptr) g_clf_retired.push_back({dev, f.ptr});
  f.ptr = p;
  f.words = want;
  return p;
}

hipError_t launch_100_cluster(unsigned char *base, uint32_t *flags, uint64_t inst, uint64_t R,
                              uint64_t C, uint64_t S, const ClPlan &pl, hipStream_t st) {
  if (inst == 0 || R <= 1 || C <= 1) return hipSuccess;
  if (pl.G == 0 || pl.chunk == 0 || (reinterpret_cast<uintptr_t>(base) % pl.ub) ||
      S % pl.ub || uint64_t(pl.chunk) * pl.G < R * C * S || pl.lds_bytes > cl_lds_bytes())
    return hipErrorInvalidValue;
  // `flags` (the caller's scratch) is used, after clearing, only when the stream has no kept
  // buffer (under stream capture)
  uint32_t *kept = stream_cluster_flags(cluster_flag_words(inst), st);
  if (!kept) {
    if (!flags) return hipErrorInvalidValue;
    hipError_t e = launch_fill_u32(flags, cluster_flag_words(inst), 0, st);
    if (e != hipSuccess) return e;
  }
  KernelSpan span(st, "k_100_cluster", 2.0 * double(inst) * double(R) * double(C) * double(S));
  ClParams p{};
  p.base = base;
  p.flags = kept ? kept : flags;
  int dev = 0;
  p.gave_up = hipGetDevice(&dev) == hipSuccess ? gave_up_here(dev) : nullptr;
  p.inst = inst;
  p.inst_bytes = R * C * S;
  p.wgs = inst * pl.G;
  p.R = uint32_t(R);
  p.C = uint32_t(C);
  p.S = uint32_t(S);
  p.G = pl.G;
  p.chunk = pl.chunk;
  p.lds_pieces = pl.lds_pieces;
  p.spin = uint32_t(std::max<int64_t>(1, tuning("cl_spin")));
  p.div_S = FastDiv32(uint32_t(S));
  p.div_C = FastDiv32(uint32_t(C));
  p.div_G = FastDiv32(pl.G);
  p.fa = pl.fa;
  p.fb = pl.fb;
  if (pl.fes) {  // fused: the instance is R tiles of fa x fb elements, units of fa elements
    if (uint64_t(pl.fa) * pl.fes != S || C != pl.fb || pl.chunk % (pl.fa * pl.fb * pl.fes))
      return hipErrorInvalidValue;
    p.div_ab = FastDiv32(pl.fa * pl.fb);
    p.div_b = FastDiv32(pl.fb);
  }
  auto go = [&](auto kern) {
    if (pl.lds_bytes > 65536) {
      hipError_t a = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         int(pl.lds_bytes));
      if (a != hipSuccess) return a;
    }
    hipLaunchKernelGGL(kern, dim3(uint32_t(p.wgs)), dim3(kClThreads), pl.lds_bytes, st, p);
    return hipGetLastError();
  };
  auto fused = [&](auto ubt) -> hipError_t {
    constexpr int U = decltype(ubt)::value;
    if (pl.fes == 4 && pl.flv == 16) return go(k_100_cluster<U, kClThreads, 4, 16>);
    if (pl.fes == 4 && pl.flv == 8) return go(k_100_cluster<U, kClThreads, 4, 8>);
    if (pl.fes == 8 && pl.flv == 16) return go(k_100_cluster<U, kClThreads, 8, 16>);
    if (pl.fes == 8 && pl.flv == 8) return go(k_100_cluster<U, kClThreads, 8, 8>);
    return hipErrorInvalidValue;
  };
  if (pl.fes) {
    switch (pl.ub) {
      case 16: return fused(std::integral_constant<int, 16>{});
      case 8: return fused(std::integral_constant<int, 8>{});
      case 4: return fused(std::integral_constant<int, 4>{});
    }
    return hipErrorInvalidValue;
  }
  switch (pl.ub) {
    case 16: return go(k_100_cluster<16, kClThreads, 0, 0>);
    case 8: return go(k_100_cluster<8, kClThreads, 0, 0>);
    case 4: return go(k_100_cluster<4, kClThreads, 0, 0>);
  }
  return hipErrorInvalidValue;
}

}  // namespace lm
