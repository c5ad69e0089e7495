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
// touch the same element.
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
    if ((A + nw - 1) / nw * pc.cycles >= (uint64_t(1) << 32)) return hipErrorInvalidValue;
    const uint32_t ml = uint32_t(pc.cycles);
    const uint32_t *cached = tuning("pt010_lead_cached") ? cached_leaders(a, B, pc, st) : nullptr;
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
