// kernels_010.hip - elementary transposition 010: [A][a][B] -> [A][B][a] for gfx950.
//
// Two forms, as in the reference (src/cl/cl_aos_asta.cl:53-84 and 202-278):
//   * k_010_bs      the LDS-staged ("barrier-sync") form: a chunk of whole tiles is read
//                   with coalesced loads into LDS, one barrier, and written back coalesced in
//                   the transposed order. Single pass over HBM.
//   * k_010_pttwac  cycle following inside a tile whose data stays in HBM and whose visited
//                   bits live in LDS (the reference's REMAP layout), for tiles too large to
//                   stage.
#include "kernels.h"
#include "kernels_dev.h"

#include <algorithm>
#include <type_traits>
#include <utility>
#include <vector>

namespace lm {

// ---------------------------------------------------------------------------------------
// LDS-staged 010.
//
// Layout in LDS: tile t of the chunk occupies a*Bp elements, row i at (t*a + i)*Bp, where
// Bp = B rounded up to an odd number. The gather reads consecutive output positions
// (consecutive i of one column j) from addresses Bp apart, so with an odd pitch each 32-bit
// bank is hit once per 32 lanes. The stage-in writes consecutive j, consecutive banks.
//
// Each lane moves a "unit" of G elements: G = 4/ES for 1- and 2-byte elements (so global
// accesses are dwords), 1 otherwise. Tile/row/column coordinates are carried in mixed radix
// and advanced by constant deltas; the only divisions are one per lane per chunk.
// ---------------------------------------------------------------------------------------

template <int ES, int G, int NT, int UPT, int LAYOUT>
__global__ __launch_bounds__(NT) void k_010_bs(Bs010Params p) {
  using E = typename Elem<ES>::T;
  using U = typename UnitOf<ES, G>::T;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  E *lds = reinterpret_cast<E *>(lds_raw);
  uint64_t c = blockIdx.x;
  if (c >= p.nchunks) return;
  U v[UPT];
  bs_load_round<ES, G, NT, UPT>(p, c, 0, v);
  while (true) {
    bs_stage_in<ES, G, NT, UPT, LAYOUT>(p, c, v, lds);
    __syncthreads();
    // Prefetch the next chunk (a disjoint region) while this one is written back.
    const uint64_t cn = c + gridDim.x;
    if (cn < p.nchunks) bs_load_round<ES, G, NT, UPT>(p, cn, 0, v);
    bs_gather_out<ES, G, NT, LAYOUT>(p, c, lds);
    __syncthreads();
    if (cn >= p.nchunks) break;
    c = cn;
  }
}

// ---------------------------------------------------------------------------------------
// Cycle-following 010 with visited bits in LDS.
//
// Element s of the a x B tile goes to next(s) = (s mod B)*a + floor(s/B). Bits use the
// reference's REMAP placement (src/cl/cl_aos_asta.cl:242-244): word = s mod sh, bit = s/sh,
// where sh = R*ceil(aB/32) (+P per 32 words); R > 1 spreads the bits over more words.
//
// Claim protocol (src/cl/cl_aos_asta.cl:238-272): a lane reads its start element, then tests
// the start's bit; while it keeps winning, it reads the next element, sets that element's
// bit with an LDS atomic OR and, only if the bit was clear, stores the carried value there.
// A slot is stored only by the lane that set its bit, and every value a lane carries was read
// before any store to its slot could happen: the start element is read (and the load
// waited for) before its bit is tested, and later elements are read before their bits are
// set. A cheap test of the start bit before the start element is read skips finished starts
// without touching HBM.
// ---------------------------------------------------------------------------------------
template <int ES, int NT>
__global__ __launch_bounds__(NT) void k_010_pttwac(Pt010Params p) {
  extern __shared__ __align__(16) uint32_t lflags[];
  for (uint64_t t = blockIdx.x; t < p.A; t += gridDim.x) pt010_tile<ES, NT>(p, t, lflags);
}

// Same stage with one tile per wave: 4 waves per workgroup, each with its own bit region of
// `sh_stride` words, no workgroup barrier (pt010_wave_tile).
template <int ES>
__global__ __launch_bounds__(256) void k_010_pt_wave(Pt010Params p, uint32_t sh_stride) {
  extern __shared__ __align__(16) uint32_t lflags[];
  const uint32_t wv = threadIdx.x >> 6;
  uint32_t *wflags = lflags + wv * sh_stride;
  const uint64_t nw = uint64_t(gridDim.x) * 4;
  for (uint64_t t = uint64_t(blockIdx.x) * 4 + wv; t < p.A; t += nw)
    pt010_wave_tile<ES>(p, t, wflags);
}

// Same stage for tiles whose cycles are all short (the order k of a modulo a*B - 1 is small):
// the reference's leader-only variant (src/cl/cl_aos_asta.cl:221-236), where each cycle is
// moved by the one work-item that starts at its smallest slot, rebuilt for gfx950. The tile map
// is the same for every tile, so each workgroup finds the leaders once: every lane walks the
// index map from its slots (at most k steps, no data), a wave64 ballot and a popcount prefix
// compact the leaders into an LDS list. Each wave then owns whole tiles (no workgroup barrier
// per tile): lane i takes leader i of a round of 64, loads the W slots of its cycle that follow
// (all W loads in flight together), waits, and stores each value one slot on; the value that
// returns to the leader is stored last. No visited bits: cycles are disjoint, so no two lanes
// touch the same element. A cached list may instead hold one start of any slot per cycle,
// ordered so that the lanes of a round touch few sectors (order_010_leaders), with kNoLeader
// entries for lanes that have no cycle in their round.
template <int ES, int W>
__global__ __launch_bounds__(256) void k_010_pt_lead(Pt010Params p, uint32_t order,
                                                      uint32_t max_leaders, FastDiv32 div_nl,
                                                      const uint32_t *cached) {
  using E = typename Elem<ES>::T;
  extern __shared__ __align__(16) uint32_t lds[];
  uint32_t *count = lds;      // [0]: leaders found
  uint32_t *list = lds + 4;   // the leaders
  const uint32_t tid = threadIdx.x, lane = tid & 63u;
  const uint32_t a = p.a, B = p.B, last = p.n - 1;
  auto nextp = [&](uint32_t s) {
    const uint32_t q = p.div_B.div(s);
    return (s - q * B) * a + q;
  };
  if (tid == 0) count[0] = cached ? max_leaders : 0;
  __syncthreads();
  // The tile map is the 100 map of the a x B grid: when that grid's leader list is cached,
  // it is copied in instead of searched for.
  for (uint32_t k = tid; cached && k < max_leaders; k += 256) list[k] = cached[k];
  for (uint32_t s0 = 0; !cached && s0 < last; s0 += 256) {  // uniform over the workgroup
    const uint32_t s = s0 + tid;
    bool lead = false;
    if (s < last) {
      uint32_t t = nextp(s);
      lead = t != s;
      for (uint32_t j = 1; lead && t != s; j++) {
        if (t < s || j > order) {
          lead = false;
          break;
        }
        t = nextp(t);
      }
    }
    const uint64_t m = __ballot(lead);
    if (m) {
      const uint32_t first = uint32_t(__builtin_ctzll(m));
      uint32_t base = 0;
      if (lane == first) base = atomicAdd(count, uint32_t(__builtin_popcountll(m)));
      base = __shfl(base, int(first));
      const uint32_t k = base + uint32_t(__builtin_popcountll(m & ((1ull << lane) - 1)));
      if (lead && k < max_leaders) list[k] = s;
    }
  }
  __syncthreads();
  const uint32_t nl = max_leaders;  // every leader is found: the count is exact (host-side)
  if (nl == 0 || count[0] != nl) return;  // never true; keeps the list reads in bounds
  const uint32_t wv = tid >> 6;
  const uint64_t nw = uint64_t(gridDim.x) * 4;
  const uint64_t w0 = uint64_t(blockIdx.x) * 4 + wv;
  if (w0 >= p.A) return;
  // The wave's (tile, leader) pairs in one flat sequence, 64 at a time: no lane idles at the
  // end of a tile. Pair i is leader i mod nl of the wave's tile i / nl.
  const uint64_t tiles = (p.A - w0 + nw - 1) / nw;
  const uint32_t pairs = uint32_t(tiles * nl);  // < 2^32: checked by the launcher
  for (uint32_t i = lane; i < pairs; i += 64) {
    const uint32_t k = div_nl.div(i);  // the wave's k-th tile
    const uint64_t t = w0 + uint64_t(k) * nw;
    {
      E *in = reinterpret_cast<E *>(p.base) + t * uint64_t(p.n);
      const uint32_t s = list[i - k * nl];
      if (s == kNoLeader) continue;  // padding of an ordered list
      E carry = in[s];
      uint32_t c = s;
      bool done = false;
      while (!done) {
        uint32_t q[W];
        E d[W];
        uint32_t n = 0;
#pragma unroll
        for (int j = 0; j < W; j++) {
          if (!done) {
            c = nextp(c);
            q[j] = c;
            n = uint32_t(j) + 1;
            if (c == s) done = true;  // back at the leader: its value was read first
            else d[j] = in[c];
          }
        }
        wait_vm_loads();  // every slot of the window is read before any is stored
#pragma unroll
        for (int j = 0; j < W; j++) {
          if (uint32_t(j) < n) in[q[j]] = j == 0 ? carry : d[j - 1];
        }
        if (!done) carry = d[W - 1];
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------------------
namespace {

template <int ES, int G, int NT, int UPT, int LAYOUT>
hipError_t launch_bs_t(const Bs010Params &p, size_t lds_bytes, uint32_t blocks,
                       hipStream_t st) {
  auto k = k_010_bs<ES, G, NT, UPT, LAYOUT>;
  if (lds_bytes > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       int(lds_bytes));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3(blocks), dim3(NT), lds_bytes, st, p);
  return hipGetLastError();
}

template <int ES, int NT>
hipError_t launch_pt_t(const Pt010Params &p, size_t lds_bytes, uint32_t blocks,
                       hipStream_t st) {
  auto k = k_010_pttwac<ES, NT>;
  if (lds_bytes > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       int(lds_bytes));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3(blocks), dim3(NT), lds_bytes, st, p);
  return hipGetLastError();
}

template <int ES>
hipError_t launch_ptw_t(const Pt010Params &p, uint32_t stride, uint32_t blocks, hipStream_t st) {
  auto k = k_010_pt_wave<ES>;
  const size_t lds = size_t(stride) * 4 * 4;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3(blocks), dim3(256), lds, st, p, stride);
  return hipGetLastError();
}

template <int ES>
hipError_t launch_ptl_t(const Pt010Params &p, uint32_t order, uint32_t max_leaders,
                        uint32_t blocks, const uint32_t *cached, hipStream_t st) {
  constexpr int W = ES >= 16 ? 4 : 8;
  auto k = k_010_pt_lead<ES, W>;
  const size_t lds = (size_t(max_leaders) + 4) * 4;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k, dim3(blocks), dim3(256), lds, st, p, order, max_leaders,
                     FastDiv32(max_leaders), cached);
  return hipGetLastError();
}

}  // namespace

namespace {
// splitmix64: the descent's moves come from a fixed seed, so an order is reproducible.
struct Mix64 {
  uint64_t x;
  uint64_t next() {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  uint32_t below(uint32_t m) { return uint32_t(next() % m); }
};

// The distinct sectors of one (group, step) with their lane counts (at most 64 pairs).
struct SectorSet {
  std::vector<std::pair<uint32_t, uint32_t>> v;
  int add(uint32_t s) {  // 1 when the sector is new to the set
    for (auto &e : v)
      if (e.first == s) {
        e.second++;
        return 0;
      }
    v.push_back({s, 1u});
    return 1;
  }
  int remove(uint32_t s) {  // -1 when its last lane left
    for (size_t i = 0; i < v.size(); i++)
      if (v[i].first == s) {
        if (--v[i].second == 0) {
          v[i] = v.back();
          v.pop_back();
          return -1;
        }
        return 0;
      }
    return 0;
  }
};
}  // namespace

// Lanes of one wave instruction that touch the same 64-byte sector cost one L2 request
// between them (the stores, written through the L1 cache, are most of this kernel's L2
// requests). So the cycles are placed in groups of 64 lanes that share sectors at every step:
//   1. any slot of a cycle may start its walk; each cycle takes the start whose walk gives
//      the smallest key, where the key interleaves the bits of the walk's sector ids, high
//      bits first, over all steps (so cycles whose walks stay in the same sectors at every
//      step get close keys);
//   2. the cycles, sorted by key, fill groups of 64 in turn;
//   3. a bounded descent with a fixed seed swaps two cycles of different groups or moves a
//      cycle's start, keeping only moves that lower the sum over groups and steps of
//      distinct sectors.
// On the 64 x 16 tile of 4-byte elements this sum falls from 305 (smallest slots in
// increasing order) to about 250 per tile.
std::vector<uint32_t> order_010_leaders(uint32_t a, uint32_t B, size_t es, uint64_t cycles) {
  constexpr uint32_t G = 64, kSector = 64, kMaxLen = 64;
  const uint64_t n64 = uint64_t(a) * B;
  if (a < 2 || B < 2 || es == 0 || n64 > (uint64_t(1) << 22) || cycles == 0 ||
      cycles > (uint64_t(1) << 15))
    return {};
  const uint32_t n = uint32_t(n64), last = n - 1;
  auto next = [&](uint32_t s) { return (s % B) * a + s / B; };
  // every cycle of length >= 2 of [1, last), its slots in walk order
  std::vector<uint8_t> seen(n, 0);
  std::vector<uint32_t> slots, off{0};
  for (uint32_t s = 1; s < last; s++) {
    if (seen[s]) continue;
    uint32_t t = s, len = 0;
    do {
      seen[t] = 1;
      slots.push_back(t);
      len++;
      t = next(t);
    } while (t != s && len <= kMaxLen);
    if (t != s) return {};  // a cycle longer than kMaxLen: not ordered
    if (len == 1) slots.pop_back();
    else off.push_back(uint32_t(slots.size()));
  }
  const uint32_t nc = uint32_t(off.size() - 1);
  if (nc != cycles) return {};
  auto len_of = [&](uint32_t c) { return off[c + 1] - off[c]; };
  uint32_t L = 0;
  for (uint32_t c = 0; c < nc; c++) L = std::max(L, len_of(c));
  auto sector = [&](uint32_t c, uint32_t rot, uint32_t j) {
    return uint32_t(uint64_t(slots[off[c] + (rot + j) % len_of(c)]) * es / kSector);
  };
  const uint32_t max_sector = uint32_t((n64 * es - 1) / kSector);
  uint32_t sb = 1;
  while (sb < 32 && (max_sector >> sb) != 0) sb++;
  // 1. each cycle's start: the smallest key over all its slots. The key's bits, taken plane by
  // plane from the sector ids' top bit down, step by step within a plane, are kept as one
  // 64-bit word per plane (step j at bit 63 - j; L <= 64), so starting the walk r slots later
  // rotates every word's top `len` bits left by r.
  std::vector<uint64_t> keys(size_t(nc) * sb), base(sb), cand(sb);
  std::vector<uint32_t> rot_of(nc, 0);
  for (uint32_t c = 0; c < nc; c++) {
    const uint32_t len = len_of(c);
    std::fill(base.begin(), base.end(), 0ull);
    for (uint32_t j = 0; j < len; j++) {
      const uint32_t sec = sector(c, 0, j);
      for (uint32_t b = 0; b < sb; b++)
        if ((sec >> b) & 1u) base[sb - 1 - b] |= uint64_t(1) << (63 - j);
    }
    uint64_t *best = &keys[size_t(c) * sb];
    std::copy(base.begin(), base.end(), best);
    const uint64_t mask = len == 64 ? ~0ull : (uint64_t(1) << len) - 1;
    for (uint32_t r = 1; r < len; r++) {
      for (uint32_t w = 0; w < sb; w++) {
        const uint64_t x = base[w] >> (64 - len);
        cand[w] = (((x << r) | (x >> (len - r))) & mask) << (64 - len);
      }
      if (std::lexicographical_compare(cand.begin(), cand.end(), best, best + sb)) {
        std::copy(cand.begin(), cand.end(), best);
        rot_of[c] = r;
      }
    }
  }
  // 2. sorted by key into groups
  std::vector<uint32_t> order(nc);
  for (uint32_t c = 0; c < nc; c++) order[c] = c;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    const uint64_t *kx = &keys[size_t(x) * sb], *ky = &keys[size_t(y) * sb];
    return std::lexicographical_compare(kx, kx + sb, ky, ky + sb);
  });
  const uint32_t ng = (nc + G - 1) / G;
  std::vector<uint32_t> grp_of(nc);
  std::vector<SectorSet> sets(size_t(ng) * L);
  auto put = [&](uint32_t c, uint32_t g, uint32_t rot, bool add) {
    int d = 0;
    for (uint32_t j = 0; j < len_of(c); j++) {
      SectorSet &ss = sets[size_t(g) * L + j];
      d += add ? ss.add(sector(c, rot, j)) : ss.remove(sector(c, rot, j));
    }
    return d;
  };
  for (uint32_t i = 0; i < nc; i++) {
    grp_of[order[i]] = i / G;
    put(order[i], i / G, rot_of[order[i]], true);
  }
  // 3. descent (each move costs about 4 L set updates of up to 64 entries; at most ~10^7 such
  // steps, a few milliseconds once per tile shape and element size)
  if (ng > 1 || L > 1) {
    const uint64_t iters =
        std::min<uint64_t>(uint64_t(50) * nc, uint64_t(10000000) / (uint64_t(L) * 256 + 1));
    Mix64 rng{0x5eedull + n64 * 131 + es};
    for (uint64_t it = 0; it < iters; it++) {
      if ((rng.next() & 1) || ng == 1) {
        const uint32_t c = rng.below(nc), len = len_of(c);
        const uint32_t r2 = rng.below(len), g = grp_of[c], r = rot_of[c];
        if (r2 == r) continue;
        const int d = put(c, g, r, false) + put(c, g, r2, true);
        if (d < 0) {
          rot_of[c] = r2;
        } else {
          put(c, g, r2, false);
          put(c, g, r, true);
        }
      } else {
        const uint32_t x = rng.below(nc), y = rng.below(nc);
        const uint32_t gx = grp_of[x], gy = grp_of[y];
        if (gx == gy) continue;
        const int d = put(x, gx, rot_of[x], false) + put(y, gy, rot_of[y], false) +
                      put(x, gy, rot_of[x], true) + put(y, gx, rot_of[y], true);
        if (d < 0) {
          grp_of[x] = gy;
          grp_of[y] = gx;
        } else {
          put(x, gy, rot_of[x], false);
          put(y, gx, rot_of[y], false);
          put(x, gx, rot_of[x], true);
          put(y, gy, rot_of[y], true);
        }
      }
    }
  }
  // the list: group by group, each padded to 64 entries
  std::vector<uint32_t> list(size_t(ng) * G, kNoLeader), fill(ng, 0);
  for (uint32_t i = 0; i < nc; i++) {
    const uint32_t c = order[i], g = grp_of[c];
    if (fill[g] >= G) return {};  // never: swaps keep every group's size
    list[size_t(g) * G + fill[g]++] = slots[off[c] + rot_of[c]];
  }
  // self-check: the walks from the listed starts cover every slot of [1, last) that is not a
  // fixed point exactly once
  std::fill(seen.begin(), seen.end(), 0);
  uint64_t covered = 0;
  for (uint32_t s : list) {
    if (s == kNoLeader) continue;
    uint32_t t = s;
    do {
      if (t == 0 || t >= last || seen[t]) return {};
      seen[t] = 1;
      covered++;
      t = next(t);
    } while (t != s);
  }
  if (covered != slots.size()) return {};
  return list;
}

uint32_t bs010_pitch(uint32_t B) { return (B & 1u) ? B : B + 1; }

bool plan_010_bs(size_t es, uint64_t A, uint64_t a, uint64_t B, size_t lds_budget,
                 uint64_t base_addr, int upt, Bs010Plan *out) {
  if (A == 0 || a == 0 || B == 0) return false;
  if (upt != 4 && upt != 8 && upt != 16) upt = 16;
  uint64_t Bp = bs010_pitch(uint32_t(std::min<uint64_t>(B, 0xffffffffu)));
  uint64_t tile_lds = a * Bp * es;
  if (a > 0xffffu * 16 || B > 0xffffu * 16 || tile_lds > lds_budget) return false;
  uint64_t n = a * B;
  // Unit width: dwords for sub-dword elements when the buffer and the chunk stride allow.
  int G = 1;
  uint64_t g0 = 1;
  if (es < 4) {
    uint64_t nb = n * es;
    uint64_t gcd4 = (nb % 4 == 0) ? 4 : (nb % 2 == 0 ? 2 : 1);
    g0 = 4 / gcd4;
    if ((base_addr % 4) == 0 && g0 * tile_lds <= lds_budget) G = int(4 / es);
    else g0 = 1;
  }
  // Dword units on a side whose extent is not a whole number of units: the chunk sits in LDS
  // unpadded, in that side's order, so that side copies units as dwords, while the other
  // side, whose extent is whole, moves each unit's G elements at fixed offsets.
  // B not whole: input order; the gather's lanes, G rows apart, step B dwords (distinct banks
  // for odd B). a not whole (B whole): output order; the stage-in's lanes, G columns apart,
  // step a dwords.
  int layout = 0;
  if (G > 1 && B % uint64_t(G) != 0) layout = 1;
  else if (G > 1 && a % uint64_t(G) != 0) layout = 2;
  if (layout) {
    Bp = B;
    tile_lds = a * Bp * es;
  }
  // At most 16 KiB per round (256 lanes x upt units of G x es bytes): 8- and 16-byte elements
  // take 8 and 4 units per lane. With 16 units, 16-byte elements made 64 KiB chunks, two
  // workgroups per CU by LDS, and ran 5.8 TB/s against 6.2 with 4 units.
  while (upt > 4 && uint64_t(256) * upt * G * es > (16u << 10)) upt /= 2;
  // One round of 256 threads x upt units per chunk when tiles are small.
  uint64_t target = uint64_t(256) * upt * G;
  uint64_t g = std::max<uint64_t>(1, target / n);
  g = std::min<uint64_t>(g, lds_budget / tile_lds);
  g = std::min<uint64_t>(g, A);
  g = std::max<uint64_t>(g0, (g / g0) * g0);
  if (g * tile_lds > lds_budget) g = g0;
  out->G = G;
  out->g = g;
  out->upt = upt;
  out->Bp = uint32_t(Bp);
  out->layout = layout;
  out->lds_bytes = size_t(g * tile_lds);
  out->nchunks = (A + g - 1) / g;
  return true;
}

namespace {
template <int ES, int G, int LAYOUT>
hipError_t launch_bs_l(const Bs010Params &p, const Bs010Plan &pl, uint32_t blocks,
                       hipStream_t st) {
  switch (pl.upt) {
    case 4: return launch_bs_t<ES, G, 256, 4, LAYOUT>(p, pl.lds_bytes, blocks, st);
    case 8: return launch_bs_t<ES, G, 256, 8, LAYOUT>(p, pl.lds_bytes, blocks, st);
    default: return launch_bs_t<ES, G, 256, 16, LAYOUT>(p, pl.lds_bytes, blocks, st);
  }
}

template <int ES, int G>
hipError_t launch_bs_g(const Bs010Params &p, const Bs010Plan &pl, uint32_t blocks,
                       hipStream_t st) {
  if constexpr (G > 1) {
    if (pl.layout == 1) return launch_bs_l<ES, G, 1>(p, pl, blocks, st);
    if (pl.layout == 2) return launch_bs_l<ES, G, 2>(p, pl, blocks, st);
  }
  return launch_bs_l<ES, G, 0>(p, pl, blocks, st);
}
}  // namespace

Bs010Params make_bs010_params(unsigned char *base, uint64_t A, uint64_t a, uint64_t B,
                              const Bs010Plan &pl) {
  constexpr int NT = 256;
  Bs010Params p{};
  p.base = base;
  p.A = A;
  p.a = uint32_t(a);
  p.B = uint32_t(B);
  p.Bp = pl.Bp;
  p.n = uint32_t(a * B);
  p.g = uint32_t(pl.g);
  p.nchunks = pl.nchunks;
  p.div_n = FastDiv32(p.n);
  p.div_B = FastDiv32(p.B);
  p.div_a = FastDiv32(p.a);
  // Per-unit advance: NT units of G elements, minus the G-1 single steps taken inside.
  uint32_t D = uint32_t(NT * pl.G - (pl.G - 1));
  p.din_t = D / p.n;
  p.din_i = (D % p.n) / p.B;
  p.din_j = (D % p.n) % p.B;
  p.dout_t = D / p.n;
  p.dout_j = (D % p.n) / p.a;
  p.dout_i = (D % p.n) % p.a;
  // When B (a) is a multiple of G, a unit never crosses an LDS row in the stage-in (a column
  // in the gather), so its G elements take fixed LDS offsets with no carries.
  p.aligned = 0;
  if (pl.G > 1 && pl.layout != 2) {  // output order: both sides need no runtime choice
    const uint64_t G = uint64_t(pl.G);
    if (pl.layout == 0 && B % G == 0) p.aligned |= 1u;
    if (a % G == 0) p.aligned |= 2u;
  }
  return p;
}

hipError_t launch_010_bs(unsigned char *base, size_t es, uint64_t A, uint64_t a, uint64_t B,
                         const Bs010Plan &pl, uint32_t max_blocks, hipStream_t st) {
  const Bs010Params p = make_bs010_params(base, A, a, B, pl);
  uint32_t blocks = uint32_t(std::min<uint64_t>(pl.nchunks, max_blocks));
  KernelSpan span(st, "k_010_bs", 2.0 * double(A) * double(a) * double(B) * double(es));
  switch (es) {
    case 1: return pl.G == 4 ? launch_bs_g<1, 4>(p, pl, blocks, st)
                             : launch_bs_g<1, 1>(p, pl, blocks, st);
    case 2: return pl.G == 2 ? launch_bs_g<2, 2>(p, pl, blocks, st)
                             : launch_bs_g<2, 1>(p, pl, blocks, st);
    case 4: return launch_bs_g<4, 1>(p, pl, blocks, st);
    case 8: return launch_bs_g<8, 1>(p, pl, blocks, st);
    case 16: return launch_bs_g<16, 1>(p, pl, blocks, st);
  }
  return hipErrorInvalidValue;
}

uint64_t pt010_flag_words(uint64_t a, uint64_t B, int R, int P) {
  uint64_t sh = uint64_t(R) * ((a * B + 31) / 32);
  sh += (sh >> 5) * uint64_t(P);
  return sh ? sh : 1;
}

hipError_t launch_010_pttwac(unsigned char *base, size_t es, uint64_t A, uint64_t a,
                             uint64_t B, int R, int P, uint32_t max_blocks, hipStream_t st) {
  constexpr int NT = 256;
  KernelSpan span(st, "k_010_pttwac", 2.0 * double(A) * double(a) * double(B) * double(es));
  Pt010Params p{};
  p.base = base;
  p.A = A;
  p.a = uint32_t(a);
  p.B = uint32_t(B);
  p.n = uint32_t(a * B);
  p.sh = uint32_t(pt010_flag_words(a, B, R, P));
  p.div_B = FastDiv32(p.B);
  p.div_sh = FastDiv32(p.sh);
  size_t lds = size_t(p.sh) * 4;
  // Short cycles everywhere (order of a modulo a*B - 1 at most "pt010_lead_order"): the
  // leader-only form, one tile per wave, no bits.
  PlPlan pc;
  const int64_t lead_order = tuning("pt010_lead_order");
  if (lead_order > 0 && a >= 2 && B >= 2 &&
      plan_100_leaders(a, B, uint32_t(std::min<int64_t>(lead_order, 4096)), &pc) &&
      pc.cycles >= 1 && (pc.cycles + 4) * 4 <= current_lds_budget() &&
      pc.cycles <= (uint64_t(1) << 15)) {
    const uint32_t lblocks = uint32_t(std::max<uint64_t>(
        1, std::min<uint64_t>((A + 3) / 4, grid_blocks(tuning("pt010_lead_blocks_per_cu")))));
    const uint64_t nw = uint64_t(lblocks) * 4;
    // "pt010_lead_cached": 0 every workgroup searches the leaders; 1 they come from the cached
    // list of the a x B grid; 2 from the list ordered for few sectors per instruction
    // (cached per tile shape and element size), when it fits the LDS budget
    const int64_t lead_src = tuning("pt010_lead_cached");
    uint32_t ml = uint32_t(pc.cycles);
    const uint32_t *cached = nullptr;
    if (lead_src == 2) {
      uint32_t entries = 0;
      cached = cached_ordered_leaders(a, B, es, pc, st, &entries);
      if (cached && (uint64_t(entries) + 4) * 4 <= current_lds_budget()) ml = entries;
      else cached = nullptr;
    }
    if (!cached && lead_src >= 1) cached = cached_leaders(a, B, pc, st);
    if ((A + nw - 1) / nw * ml >= (uint64_t(1) << 32)) return hipErrorInvalidValue;
    switch (es) {
      case 1: return launch_ptl_t<1>(p, pc.order, ml, lblocks, cached, st);
      case 2: return launch_ptl_t<2>(p, pc.order, ml, lblocks, cached, st);
      case 4: return launch_ptl_t<4>(p, pc.order, ml, lblocks, cached, st);
      case 8: return launch_ptl_t<8>(p, pc.order, ml, lblocks, cached, st);
      case 16: return launch_ptl_t<16>(p, pc.order, ml, lblocks, cached, st);
    }
    return hipErrorInvalidValue;
  }
  // One tile per wave when four bit regions fit the LDS budget (knob "pt010_wave" = 1).
  const uint32_t stride = (p.sh + 3u) & ~3u;
  if (tuning("pt010_wave") && size_t(stride) * 16 <= current_lds_budget()) {
    const uint32_t wblocks = uint32_t(std::max<uint64_t>(1, std::min<uint64_t>((A + 3) / 4, max_blocks)));
    switch (es) {
      case 1: return launch_ptw_t<1>(p, stride, wblocks, st);
      case 2: return launch_ptw_t<2>(p, stride, wblocks, st);
      case 4: return launch_ptw_t<4>(p, stride, wblocks, st);
      case 8: return launch_ptw_t<8>(p, stride, wblocks, st);
      case 16: return launch_ptw_t<16>(p, stride, wblocks, st);
    }
    return hipErrorInvalidValue;
  }
  uint32_t blocks = uint32_t(std::min<uint64_t>(A, max_blocks));
  switch (es) {
    case 1: return launch_pt_t<1, NT>(p, lds, blocks, st);
    case 2: return launch_pt_t<2, NT>(p, lds, blocks, st);
    case 4: return launch_pt_t<4, NT>(p, lds, blocks, st);
    case 8: return launch_pt_t<8, NT>(p, lds, blocks, st);
    case 16: return launch_pt_t<16, NT>(p, lds, blocks, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace lm
