// kernels_seg.hip - 100 / 0100 by cycle segments: one pass, no visited bits, no claims.
//
// The 100 map of an R x C grid of S-byte super-elements sends slot s < m = R*C - 1 to
// s*R mod m. Its cycles can be few and long (1800 x 100: 14 cycles, the longest 12857 slots),
// which leaves the claim-based forms with little parallelism, and the cycle-leader form with
// one serial walk per cycle. Here every cycle is cut into segments of at most `seg_len`
// consecutive slots p_0, p_1 = next(p_0), ..., and the segment's walker moves the content of
// p_j to p_{j+1} for j = 0 .. len-1 (p_len is the next segment's start, or the cycle's first
// start for its last segment). Two launches:
//   1. k_seg_save copies every segment's start piece to a side buffer (1/seg_len of the data);
//   2. k_seg_walk: each walker takes its start content from the side buffer and walks its
//      segment a window at a time: it loads the next W slots of the segment (the addresses
//      follow from the index map alone, so all W loads are in flight together), waits for
//      them, then stores each carried piece one slot further on.
// Every slot is stored exactly once (by the walker of the segment holding its predecessor),
// and its original content is read before that store: segment starts by launch 1, every other
// slot by its own walker in the window that also stores it. Cycles are disjoint and segments
// of one cycle share only their boundary slots, so no two walkers touch the same bytes within
// a launch and no atomics are needed.
//
// The segment table (start slot and length of every segment of a grid) is built on the host by
// walking the cycles once, O(R*C), and kept per (R, C, seg_len) with a copy on each device.
// Not a reference algorithm: the reference claims every slot with an atomic bit
// (src/cl/cl_aos_asta.cl:475-566) or, in its leader-only variant, walks each cycle serially
// (src/cl/cl_aos_asta.cl:221-236).
#include "kernels.h"
#include "kernels_dev.h"

#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <thread>
#include <tuple>
#include <vector>

namespace lm {

struct SegParams {
  unsigned char *base;
  unsigned char *save;
  const uint32_t *table;  // 2 words per segment: start slot, length
  uint64_t RC, S, nseg, nsub, sub_bytes, walkers;
  uint32_t R, C, vw;  // vw: lanes per walker (any width up to 64; 64 / vw walkers per wave)
  uint32_t nt;   // bit 0: non-temporal loads, bit 1: non-temporal stores (walk)
  uint32_t xcd;  // 1: consecutive walkers on one XCD (workgroup b runs on XCD b mod 8)
  FastDiv64 div_nsub, div_nseg;
  FastDiv32 div_C32;
};

__device__ __forceinline__ uint32_t seg_next(const SegParams &p, uint32_t s) {
  const uint32_t q = p.div_C32.div(s);
  return (s - q * p.C) * p.R + q;
}

// Workgroups are dealt to the 8 XCDs round robin (b -> XCD b mod 8). With `xcd` set, the
// workgroups of each XCD take a contiguous run of walkers, so the instances (and the cache
// lines) an XCD works on stay in its own L2.
__device__ __forceinline__ uint64_t seg_block(const SegParams &p) {
  const uint32_t b = blockIdx.x;
  if (!p.xcd) return b;
  const uint32_t nb = gridDim.x, q = nb >> 3, r = nb & 7u, x = b & 7u, i = b >> 3;
  return uint64_t(x) * q + (x < r ? x : r) + i;
}

template <int UB, int UPL, bool NT>
__device__ __forceinline__ void seg_ld(typename Elem<UB>::T (&d)[UPL], const unsigned char *src,
                                       uint32_t sub, uint32_t vw, uint32_t pbytes) {
  using V = typename Elem<UB>::T;
#pragma unroll
  for (int k = 0; k < UPL; k++) {
    const uint32_t off = (sub + k * vw) * UB;
    if (off < pbytes) {
      if constexpr (NT)
        d[k] = __builtin_nontemporal_load(reinterpret_cast<const V *>(src + off));
      else
        d[k] = *reinterpret_cast<const V *>(src + off);
    }
  }
}

template <int UB, int UPL, bool NT>
__device__ __forceinline__ void seg_st(unsigned char *dst, const typename Elem<UB>::T (&d)[UPL],
                                       uint32_t sub, uint32_t vw, uint32_t pbytes) {
  using V = typename Elem<UB>::T;
#pragma unroll
  for (int k = 0; k < UPL; k++) {
    const uint32_t off = (sub + k * vw) * UB;
    if (off < pbytes) {
      if constexpr (NT)
        __builtin_nontemporal_store(d[k], reinterpret_cast<V *>(dst + off));
      else
        *reinterpret_cast<V *>(dst + off) = d[k];
    }
  }
}

template <int UB, int UPL>
__device__ __forceinline__ void seg_load(typename Elem<UB>::T (&d)[UPL], const unsigned char *src,
                                         uint32_t sub, uint32_t vw, uint32_t pbytes) {
  using V = typename Elem<UB>::T;
#pragma unroll
  for (int k = 0; k < UPL; k++) {
    const uint32_t off = (sub + k * vw) * UB;
    if (off < pbytes) d[k] = *reinterpret_cast<const V *>(src + off);
  }
}

template <int UB, int UPL>
__device__ __forceinline__ void seg_store(unsigned char *dst, const typename Elem<UB>::T (&d)[UPL],
                                          uint32_t sub, uint32_t vw, uint32_t pbytes) {
  using V = typename Elem<UB>::T;
#pragma unroll
  for (int k = 0; k < UPL; k++) {
    const uint32_t off = (sub + k * vw) * UB;
    if (off < pbytes) *reinterpret_cast<V *>(dst + off) = d[k];
  }
}

// Walker w = ((instance * nseg) + segment) * nsub + piece; vw lanes per walker, the walkers
// of a wave side by side (lanes past the last whole walker idle).
struct SegWalker {
  uint64_t x, g, c;
  uint32_t pbytes;
};

__device__ __forceinline__ SegWalker seg_walker(const SegParams &p, uint64_t w) {
  SegWalker r;
  const uint64_t rest = p.div_nsub.div(w);
  r.c = w - rest * p.nsub;
  r.x = p.div_nseg.div(rest);
  r.g = rest - r.x * p.nseg;
  const uint64_t left = p.S - r.c * p.sub_bytes;
  r.pbytes = uint32_t(left < p.sub_bytes ? left : p.sub_bytes);
  return r;
}

template <int UB, int UPL>
__global__ __launch_bounds__(256) void k_seg_save(SegParams p) {
  using V = typename Elem<UB>::T;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t vw = p.vw, gpw = 64u / vw, me = lane / vw, sub = lane - me * vw;
  if (me >= gpw) return;  // idle lanes past the wave's last whole walker
  const uint64_t wave = seg_block(p) * 4 + (threadIdx.x >> 6);
  const uint64_t nwaves = uint64_t(gridDim.x) * 4;
  for (uint64_t wb = wave * gpw; wb < p.walkers; wb += nwaves * gpw) {
    const uint64_t w = wb + me;
    if (w >= p.walkers) continue;
    const SegWalker k = seg_walker(p, w);
    const uint32_t s0 = p.table[2 * k.g];
    if (uint64_t(s0) >= p.RC) continue;  // never true for a valid table; keeps accesses in bounds
    V d[UPL];
    load_piece<UB, UPL>(d, p.base + k.x * p.RC * p.S + uint64_t(s0) * p.S + k.c * p.sub_bytes,
                        sub, vw, k.pbytes);
    seg_store<UB, UPL>(p.save + (k.x * p.nseg + k.g) * p.S + k.c * p.sub_bytes, d, sub, vw,
                       k.pbytes);
  }
}

template <int UB, int UPL, int W, bool NTL, bool NTS>
__global__ __launch_bounds__(256) void k_seg_walk(SegParams p) {
  using V = typename Elem<UB>::T;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t vw = p.vw, gpw = 64u / vw, me = lane / vw, sub = lane - me * vw;
  if (me >= gpw) return;  // idle lanes past the wave's last whole walker
  const uint64_t wave = seg_block(p) * 4 + (threadIdx.x >> 6);
  const uint64_t nwaves = uint64_t(gridDim.x) * 4;
  const uint32_t m = uint32_t(p.RC - 1);
  for (uint64_t wb = wave * gpw; wb < p.walkers; wb += nwaves * gpw) {
    const uint64_t w = wb + me;
    if (w >= p.walkers) continue;
    const SegWalker k = seg_walker(p, w);
    uint32_t t = p.table[2 * k.g];
    const uint32_t len = p.table[2 * k.g + 1];
    if (t == 0 || t >= m || len > 0x7fffffffu) continue;  // never true for a valid table
    unsigned char *ib = p.base + k.x * p.RC * p.S + k.c * p.sub_bytes;
    V carry[UPL];
    seg_load<UB, UPL>(carry, p.save + (k.x * p.nseg + k.g) * p.S + k.c * p.sub_bytes, sub, vw,
                      k.pbytes);
    for (uint32_t i = 0; i < len; i += W) {
      const uint32_t n = len - i < uint32_t(W) ? len - i : uint32_t(W);
      uint32_t q[W];
      V d[W][UPL];
#pragma unroll
      for (int j = 0; j < W; j++) {
        if (uint32_t(j) < n) {
          t = seg_next(p, t);
          q[j] = t;
          // p_len (the next segment's start) is stored here but read by launch 1 only
          if (i + uint32_t(j) + 1 < len) seg_ld<UB, UPL, NTL>(d[j], ib + uint64_t(t) * p.S, sub, vw, k.pbytes);
        }
      }
      wait_vm_loads();  // every slot of the window is read before any of them is stored
#pragma unroll
      for (int j = 0; j < W; j++) {
        if (uint32_t(j) < n) {
          if (j == 0)
            seg_st<UB, UPL, NTS>(ib + uint64_t(q[0]) * p.S, carry, sub, vw, k.pbytes);
          else
            seg_st<UB, UPL, NTS>(ib + uint64_t(q[j]) * p.S, d[j - 1], sub, vw, k.pbytes);
        }
      }
      if (n == uint32_t(W)) {
#pragma unroll
        for (int u = 0; u < UPL; u++) carry[u] = d[W - 1][u];
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// Host: segment tables
// ---------------------------------------------------------------------------------------
namespace {
struct SegTable {
  std::vector<uint32_t> host;  // start, length per segment
  uint64_t nseg = 0;
  uint32_t *dev[64] = {};
};
std::mutex g_seg_mu;
std::map<std::tuple<uint64_t, uint64_t, uint32_t>, std::shared_ptr<SegTable>> g_seg;
size_t g_seg_bytes = 0;
constexpr size_t kSegCacheBytes = size_t(256) << 20;
// Tables of grids above the synchronous limit are built on background threads: the call that
// first meets such a grid plans another form instead of waiting for the walk.
std::set<std::tuple<uint64_t, uint64_t, uint32_t>> g_building;
struct Builders {
  std::vector<std::thread> th;
  ~Builders() {
    for (std::thread &t : th)
      if (t.joinable()) t.join();
  }
};
Builders g_builders;
thread_local int t_sync_build = 0;  // SegSyncBuild scopes alive on this thread

std::shared_ptr<SegTable> build_table(uint64_t R, uint64_t C, uint32_t L) {
  auto tb = std::make_shared<SegTable>();
  const uint64_t RC = R * C, m = RC - 1;
  std::vector<uint64_t> seen((RC + 63) / 64, 0);
  auto nxt = [&](uint64_t s) { return (s % C) * R + s / C; };
  for (uint64_t s = 1; s < m; s++) {
    if ((seen[s >> 6] >> (s & 63)) & 1u) continue;
    const size_t first = tb->host.size();
    uint64_t t = s, len = 0;
    do {
      seen[t >> 6] |= uint64_t(1) << (t & 63);
      if (len % L == 0) {
        tb->host.push_back(uint32_t(t));
        tb->host.push_back(0);
      }
      len++;
      t = nxt(t);
    } while (t != s);
    if (len < 2) {  // a fixed point
      tb->host.resize(first);
      continue;
    }
    for (size_t k = first, j = 0; k < tb->host.size(); k += 2, j++)
      tb->host[k + 1] = uint32_t(std::min<uint64_t>(L, len - j * L));
  }
  tb->nseg = tb->host.size() / 2;
  return tb;
}
}  // namespace

static void seg_shape(uint64_t S, uint64_t base_addr, SegPlan &sp) {
  const uint64_t al = S | base_addr;
  const int ub = (al % 16 == 0) ? 16 : (al % 8 == 0) ? 8 : (al % 4 == 0) ? 4 : (al % 2 == 0) ? 2 : 1;
  const uint64_t nU = S / uint64_t(ub);
  sp.ub = ub;
  if (nU <= 64) {
    sp.vw = uint32_t(nU);  // one unit per lane, 64 / nU walkers per wave
    sp.upl = 1;
    sp.nsub = 1;
    sp.sub_bytes = S;
  } else {
    sp.vw = 64;
    const uint64_t nsub = (nU + 127) / 128, per = (nU + nsub - 1) / nsub;
    sp.upl = per > 64 ? 2 : 1;
    sp.sub_bytes = per * uint64_t(ub);
    sp.nsub = (S + sp.sub_bytes - 1) / sp.sub_bytes;
  }
}

static std::shared_ptr<SegTable> insert_table(std::tuple<uint64_t, uint64_t, uint32_t> key,
                                              std::shared_ptr<SegTable> tb) {
  const size_t bytes = tb->host.size() * 4;
  std::lock_guard<std::mutex> lk(g_seg_mu);
  g_building.erase(key);
  auto it = g_seg.find(key);
  if (it != g_seg.end()) return it->second;
  if (g_seg_bytes + 2 * bytes > kSegCacheBytes) return nullptr;
  g_seg[key] = tb;
  g_seg_bytes += 2 * bytes;
  return tb;
}

bool plan_100_segments(uint64_t inst, uint64_t R, uint64_t C, uint64_t S, uint64_t base_addr,
                       uint32_t seg_len, uint64_t max_slots, uint64_t sync_slots, bool build,
                       SegPlan *out) {
  if (inst == 0 || R < 2 || C < 2 || S == 0 || seg_len < 2 || seg_len > 4096) return false;
  const uint64_t RC = R * C;
  if (RC > max_slots || RC >= (uint64_t(1) << 31)) return false;
  if (!build) {
    SegPlan sp;
    sp.seg_len = seg_len;
    sp.nseg = (RC + seg_len - 1) / seg_len;
    seg_shape(S, base_addr, sp);
    *out = sp;
    return true;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return false;
  std::shared_ptr<SegTable> tb;
  {
    std::lock_guard<std::mutex> lk(g_seg_mu);
    auto it = g_seg.find(std::make_tuple(R, C, seg_len));
    if (it != g_seg.end()) tb = it->second;
  }
  if (!tb) {
    const auto key = std::make_tuple(R, C, seg_len);
    if (RC > sync_slots && !t_sync_build) {  // walk it in the background; plan another form
      std::lock_guard<std::mutex> lk(g_seg_mu);
      if (g_building.insert(key).second)
        g_builders.th.emplace_back([key] { insert_table(key, build_table(std::get<0>(key), std::get<1>(key), std::get<2>(key))); });
      return false;
    }
    tb = insert_table(key, build_table(R, C, seg_len));
    if (!tb) return false;
  }
  if (tb->nseg == 0) return false;
  {
    std::lock_guard<std::mutex> lk(g_seg_mu);
    if (!tb->dev[dev]) {
      // A synchronous upload, once per grid and device; it fails under stream capture, and
      // then this form is not planned.
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      if (hipStreamIsCapturing(nullptr, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
        return false;
      uint32_t *d = nullptr;
      if (hipMalloc(&d, tb->host.size() * 4) != hipSuccess) return false;
      if (hipMemcpy(d, tb->host.data(), tb->host.size() * 4, hipMemcpyHostToDevice) !=
          hipSuccess) {
        (void)hipFree(d);
        return false;
      }
      tb->dev[dev] = d;
    }
  }
  SegPlan sp;
  sp.table = tb->dev[dev];
  sp.nseg = tb->nseg;
  sp.seg_len = seg_len;
  seg_shape(S, base_addr, sp);
  *out = sp;
  return true;
}

SegSyncBuild::SegSyncBuild() { t_sync_build++; }
SegSyncBuild::~SegSyncBuild() { t_sync_build--; }

size_t seg_save_bytes(uint64_t inst, uint64_t S, const SegPlan &sp) {
  return size_t((inst * sp.nseg * S + 255) & ~uint64_t(255));
}

void free_segment_tables() {
  std::vector<std::thread> th;
  {
    std::lock_guard<std::mutex> lk(g_seg_mu);
    th.swap(g_builders.th);
  }
  for (std::thread &t : th)
    if (t.joinable()) t.join();
  std::lock_guard<std::mutex> lk(g_seg_mu);
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto &kv : g_seg)
    for (int d = 0; d < 64; d++)
      if (kv.second->dev[d]) {
        (void)hipSetDevice(d);
        (void)hipDeviceSynchronize();
        (void)hipFree(kv.second->dev[d]);
        kv.second->dev[d] = nullptr;
      }
  g_seg.clear();
  g_seg_bytes = 0;
  (void)hipSetDevice(cur);
}

namespace {
template <int UB, int UPL>
hipError_t launch_seg_t(const SegParams &p, uint32_t blocks, hipStream_t st) {
  constexpr int W = UPL == 1 ? 16 : 8;
  {
    KernelSpan span(st, "k_seg_save", 0.0);
    hipLaunchKernelGGL((k_seg_save<UB, UPL>), dim3(blocks), dim3(256), 0, st, p);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  KernelSpan span(st, "k_seg_walk", 2.0 * double(p.walkers / p.nsub / p.nseg) * double(p.RC) * double(p.S));
  switch (p.nt & 3u) {
    case 0: hipLaunchKernelGGL((k_seg_walk<UB, UPL, W, false, false>), dim3(blocks), dim3(256), 0, st, p); break;
    case 1: hipLaunchKernelGGL((k_seg_walk<UB, UPL, W, true, false>), dim3(blocks), dim3(256), 0, st, p); break;
    case 2: hipLaunchKernelGGL((k_seg_walk<UB, UPL, W, false, true>), dim3(blocks), dim3(256), 0, st, p); break;
    default: hipLaunchKernelGGL((k_seg_walk<UB, UPL, W, true, true>), dim3(blocks), dim3(256), 0, st, p); break;
  }
  return hipGetLastError();
}

template <int UB>
hipError_t launch_seg_ub(const SegParams &p, int upl, uint32_t blocks, hipStream_t st) {
  if (upl == 1) return launch_seg_t<UB, 1>(p, blocks, st);
  if (upl == 2) return launch_seg_t<UB, 2>(p, blocks, st);
  return hipErrorInvalidValue;
}
}  // namespace

hipError_t launch_100_segments(unsigned char *base, unsigned char *save, uint64_t inst,
                               uint64_t R, uint64_t C, uint64_t S, const SegPlan &sp,
                               uint32_t max_blocks, uint32_t nt, uint32_t xcd, hipStream_t st) {
  if (inst == 0 || sp.nseg == 0) return hipSuccess;
  if (!sp.table || !save || R * C >= (uint64_t(1) << 31)) return hipErrorInvalidValue;
  SegParams p{};
  p.base = base;
  p.save = save;
  p.table = sp.table;
  p.RC = R * C;
  p.S = S;
  p.nseg = sp.nseg;
  p.nsub = sp.nsub;
  p.sub_bytes = sp.sub_bytes;
  p.walkers = inst * sp.nseg * sp.nsub;
  p.R = uint32_t(R);
  p.C = uint32_t(C);
  p.vw = sp.vw;
  p.nt = nt;
  p.xcd = xcd;
  p.div_nsub = FastDiv64(sp.nsub);
  p.div_nseg = FastDiv64(sp.nseg);
  p.div_C32 = FastDiv32(uint32_t(C));
  const uint64_t per_block = 4ull * (64u / sp.vw);
  const uint64_t blocks64 = std::min<uint64_t>((p.walkers + per_block - 1) / per_block,
                                               std::max<uint32_t>(max_blocks, 1));
  const uint32_t blocks = uint32_t(blocks64);
  switch (sp.ub) {
    case 1: return launch_seg_ub<1>(p, sp.upl, blocks, st);
    case 2: return launch_seg_ub<2>(p, sp.upl, blocks, st);
    case 4: return launch_seg_ub<4>(p, sp.upl, blocks, st);
    case 8: return launch_seg_ub<8>(p, sp.upl, blocks, st);
    case 16: return launch_seg_ub<16>(p, sp.upl, blocks, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace lm
