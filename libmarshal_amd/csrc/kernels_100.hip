// kernels_100.hip - elementary transpositions 100 ([A][B][b] -> [B][A][b]) and 0100
// ([I][A][B][b] -> [I][B][A][b], I independent instances) for gfx950, by cycle following over
// super-elements of b elements with visited bits in global memory.
//
// Reference: _transpose_100 / _transpose_100_b and their 0100 instance offsets,
// src/cl/cl_aos_asta.cl:348-801, launched by _cl_transpose_0100, src/cl_marshal.cc:340-558.
//
// Work decomposition (this file's own design):
//   * A "walker" is a group of VW lanes (VW a power of two <= 64) that carries one piece of a
//     super-element around its cycle. A super-element of S bytes is split into `nsub` pieces
//     when it exceeds what 64 lanes hold in registers; every (instance, slot, piece) has its
//     own visited bit, so pieces of one cycle are followed independently.
//   * Each lane of a walker holds UPL units of UB bytes (UB = the largest power of two <= 16
//     that divides S), so the VW lanes read and write one contiguous piece coalesced.
//   * Walkers are the start points (instance, slot, piece); waves sweep them in order with a
//     grid stride, so the set of active start points advances uniformly.
//   * The group's leader lane performs the bit test/claim with a device-scope atomic and the
//     result reaches the other lanes through one wave64 ballot.
//
// Claim protocol (src/cl/cl_aos_asta.cl:475-566): a walker reads its start piece, then tests
// the start slot's bit; while it keeps winning it reads the next slot's piece, sets that
// slot's bit (atomic OR) and stores the carried piece there only if the bit was clear. Only
// the winner of a slot's bit stores to it, so any piece a walker carries is the slot's
// original content: later slots are read before their bits are set, and the start piece is
// read, and the load completed (s_waitcnt), before the start bit is tested. Losing walkers
// stop; the winner of the same slot carries the identical value on. Bits are read and set
// only with device-scope atomics, which are coherent across the XCDs' separate L2 caches;
// a stale cached copy of a slot can only ever hold the slot's original content, which is
// exactly what a walker that wins that slot needs.
#include "kernels.h"
#include "kernels_dev.h"

#include <algorithm>
#include <vector>

namespace lm {

template <int UB, int UPL, bool I64>
__global__ __launch_bounds__(256) void k_100_pttwac(Pt100Params p) {
  const uint64_t gpw = 64u >> p.vw_log2;  // walkers per wave
  const uint64_t wave = uint64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint64_t nwaves = uint64_t(gridDim.x) * (blockDim.x >> 6);
  for (uint64_t wb = wave * gpw; wb < p.walkers; wb += nwaves * gpw)
    pt100_walkers<UB, UPL, I64>(p, wb);
}

// ---------------------------------------------------------------------------------------
// 0100 with one instance per workgroup and the instance's visited bits in LDS.
//
// When an instance's grid is small (R*C*nsub bits fit the LDS budget) and there are many
// instances, claims are workgroup-scope LDS atomics instead of device-scope atomics on a
// small, heavily shared region of global memory, and the instance's data stays in the L2 of
// the one XCD that works on it. Same claim protocol as k_100_pttwac; walkers are groups of
// VW lanes, start points (slot, piece) are strided over the workgroup's walkers.
// ---------------------------------------------------------------------------------------
template <int UB, int UPL, int NT>
__global__ __launch_bounds__(NT) void k_0100_lds(Pt100Params p, uint64_t inst) {
  extern __shared__ __align__(16) uint32_t lflags[];
  for (uint64_t x = blockIdx.x; x < inst; x += gridDim.x) pt0100_instance<UB, UPL, NT>(p, x, lflags);
}

__global__ void k_fill_u32(uint32_t *p, uint64_t words, uint32_t value) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (; i < words; i += stride) p[i] = value;
}

static uint32_t fill_blocks(uint64_t words) {
  return uint32_t(std::min<uint64_t>((words + 255) / 256, uint64_t(device_info().cus) * 16));
}

hipError_t launch_fill_u32(uint32_t *p, uint64_t words, uint32_t value, hipStream_t st) {
  if (words == 0) return hipSuccess;
  KernelSpan span(st, "k_fill_u32", 0.0);
  hipLaunchKernelGGL(k_fill_u32, dim3(fill_blocks(words)), dim3(256), 0, st, p, words, value);
  return hipGetLastError();
}

hipError_t add_fill_u32_node(hipGraph_t graph, const hipGraphNode_t *deps, size_t ndeps,
                             uint32_t *p, uint64_t words, uint32_t value, hipGraphNode_t *node) {
  if (words == 0) return hipErrorInvalidValue;
  void *args[] = {&p, &words, &value};
  hipKernelNodeParams kp{};
  kp.func = reinterpret_cast<void *>(k_fill_u32);
  kp.gridDim = dim3(fill_blocks(words));
  kp.blockDim = dim3(256);
  kp.sharedMemBytes = 0;
  kp.kernelParams = args;
  kp.extra = nullptr;
  return hipGraphAddKernelNode(node, graph, deps, ndeps, &kp);
}

static uint64_t next_pow2(uint64_t v) {
  uint64_t r = 1;
  while (r < v) r <<= 1;
  return r;
}

void plan_100(size_t es, uint64_t inst, uint64_t R, uint64_t C, uint64_t b, Pt100Plan *out,
              int64_t vw_max, uint64_t base_addr) {
  uint64_t S = b * es;
  const uint64_t al = S | base_addr;
  int ub = (al % 16 == 0) ? 16 : (al % 8 == 0) ? 8 : (al % 4 == 0) ? 4 : (al % 2 == 0) ? 2 : 1;
  uint64_t nU = S / ub;
  uint64_t nsub = 1, per = nU;
  int upl = 1;
  uint32_t vwl = 0;
  uint64_t vmax = uint64_t(vw_max > 0 ? vw_max : tuning("pt100_vw_max"));
  if (vmax < 1 || vmax > 64 || (vmax & (vmax - 1))) vmax = 64;
  uint32_t vmaxl = 0;
  while ((1ull << vmaxl) < vmax) vmaxl++;
  if (nU <= vmax) {
    uint64_t vw = next_pow2(nU);
    while ((1ull << vwl) < vw) vwl++;
  } else {
    vwl = vmaxl;
    const uint64_t cap = vmax * 8;
    if (nU > cap) {
      nsub = (nU + cap - 1) / cap;
      per = (nU + nsub - 1) / nsub;
    }
    upl = int(next_pow2((per + vmax - 1) / vmax));
  }
  out->ub = ub;
  out->upl = upl;
  out->vw_log2 = vwl;
  out->nsub = nsub;
  out->sub_bytes = per * ub;
  // nsub may leave the last piece empty when per*ub*(nsub-1) >= S; recount pieces.
  out->nsub = (S + out->sub_bytes - 1) / out->sub_bytes;
  out->flag_words = (inst * R * C * out->nsub + 31) / 32;
}

namespace {
template <int UB, int UPL>
hipError_t launch_100_t(const Pt100Params &p, uint32_t blocks, hipStream_t st) {
  if (p.RC < (1ull << 32)) {
    hipLaunchKernelGGL((k_100_pttwac<UB, UPL, false>), dim3(blocks), dim3(256), 0, st, p);
  } else {
    hipLaunchKernelGGL((k_100_pttwac<UB, UPL, true>), dim3(blocks), dim3(256), 0, st, p);
  }
  return hipGetLastError();
}

template <int UB>
hipError_t launch_100_ub(const Pt100Params &p, int upl, uint32_t blocks, hipStream_t st) {
  switch (upl) {
    case 1: return launch_100_t<UB, 1>(p, blocks, st);
    case 2: return launch_100_t<UB, 2>(p, blocks, st);
    case 4: return launch_100_t<UB, 4>(p, blocks, st);
    case 8: return launch_100_t<UB, 8>(p, blocks, st);
  }
  return hipErrorInvalidValue;
}
}  // namespace

namespace {
template <int UB, int UPL>
hipError_t launch_lds_t(const Pt100Params &p, uint64_t inst, size_t lds, uint32_t blocks,
                        int nt, hipStream_t st) {
  if (nt == 1024) {
    auto k = k_0100_lds<UB, UPL, 1024>;
    if (lds > 65536) {
      hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, dim3(blocks), dim3(1024), lds, st, p, inst);
  } else {
    auto k = k_0100_lds<UB, UPL, 256>;
    if (lds > 65536) {
      hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, dim3(blocks), dim3(256), lds, st, p, inst);
  }
  return hipGetLastError();
}

template <int UB>
hipError_t launch_lds_ub(const Pt100Params &p, uint64_t inst, int upl, size_t lds,
                         uint32_t blocks, int nt, hipStream_t st) {
  switch (upl) {
    case 1: return launch_lds_t<UB, 1>(p, inst, lds, blocks, nt, st);
    case 2: return launch_lds_t<UB, 2>(p, inst, lds, blocks, nt, st);
    case 4: return launch_lds_t<UB, 4>(p, inst, lds, blocks, nt, st);
    case 8: return launch_lds_t<UB, 8>(p, inst, lds, blocks, nt, st);
  }
  return hipErrorInvalidValue;
}
}  // namespace

size_t lds0100_bytes(uint64_t R, uint64_t C, const Pt100Plan &pl) {
  return size_t((R * C * pl.nsub + 31) / 32 * 4);
}

hipError_t launch_0100_lds(unsigned char *base, uint64_t inst, uint64_t R, uint64_t C,
                           uint64_t S, const Pt100Plan &pl, uint32_t max_blocks,
                           hipStream_t st) {
  Pt100Params p{};
  p.base = base;
  p.flags = nullptr;
  p.R = R;
  p.C = C;
  p.RC = R * C;
  p.S = S;
  p.nsub = pl.nsub;
  p.sub_bytes = pl.sub_bytes;
  p.walkers = p.RC * pl.nsub;
  p.vw_log2 = pl.vw_log2;
  p.div_C64 = FastDiv64(C);
  p.div_RC = FastDiv64(p.RC);
  p.div_nsub = FastDiv64(pl.nsub);
  p.div_C32 = FastDiv32(uint32_t(C));
  size_t lds = lds0100_bytes(R, C, pl);
  KernelSpan span(st, "k_0100_lds", 2.0 * double(inst) * double(R) * double(C) * double(S));
  // Few instances: make each workgroup wide so that the machine still has enough walkers.
  int nt = inst < uint64_t(device_info().cus) * 2 ? 1024 : 256;
  uint32_t blocks = uint32_t(std::min<uint64_t>(inst, max_blocks));
  switch (pl.ub) {
    case 1: return launch_lds_ub<1>(p, inst, pl.upl, lds, blocks, nt, st);
    case 2: return launch_lds_ub<2>(p, inst, pl.upl, lds, blocks, nt, st);
    case 4: return launch_lds_ub<4>(p, inst, pl.upl, lds, blocks, nt, st);
    case 8: return launch_lds_ub<8>(p, inst, pl.upl, lds, blocks, nt, st);
    case 16: return launch_lds_ub<16>(p, inst, pl.upl, lds, blocks, nt, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_100_pttwac(unsigned char *base, uint32_t *flags, uint64_t inst, uint64_t R,
                             uint64_t C, uint64_t S, const Pt100Plan &pl, uint32_t max_blocks,
                             hipStream_t st, uint32_t rot_r, uint32_t rot_c) {
  Pt100Params p{};
  p.rot_r = uint32_t(rot_r % (R ? R : 1));
  p.rot_c = uint32_t(rot_c % (C ? C : 1));
  p.base = base;
  p.flags = flags;
  p.R = R;
  p.C = C;
  p.RC = R * C;
  p.S = S;
  p.nsub = pl.nsub;
  p.sub_bytes = pl.sub_bytes;
  p.walkers = inst * p.RC * pl.nsub;
  p.vw_log2 = pl.vw_log2;
  p.div_C64 = FastDiv64(C);
  p.div_RC = FastDiv64(p.RC);
  p.div_nsub = FastDiv64(pl.nsub);
  p.div_C32 = FastDiv32(uint32_t(std::min<uint64_t>(C, 0xffffffffu)));
  if (p.walkers == 0) return hipSuccess;
  uint64_t per_block = 4ull * (64u >> pl.vw_log2);
  uint64_t blocks = std::min<uint64_t>((p.walkers + per_block - 1) / per_block, max_blocks);
  if (blocks == 0) return hipErrorInvalidValue;  // walkers but no workgroups: never skip work
  KernelSpan span(st, "k_100_pttwac", 2.0 * double(inst) * double(R) * double(C) * double(S));
  switch (pl.ub) {
    case 1: return launch_100_ub<1>(p, pl.upl, uint32_t(blocks), st);
    case 2: return launch_100_ub<2>(p, pl.upl, uint32_t(blocks), st);
    case 4: return launch_100_ub<4>(p, pl.upl, uint32_t(blocks), st);
    case 8: return launch_100_ub<8>(p, pl.upl, uint32_t(blocks), st);
    case 16: return launch_100_ub<16>(p, pl.upl, uint32_t(blocks), st);
  }
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------------------
// 100 / 0100 by cycle leaders (no visited bits, no claims).
//
// On an R x C grid with m = R*C - 1, slot s < m moves to s*R mod m (slot m stays). If
// R^k = 1 (mod m), every cycle length divides k. The slots with period dividing d are the
// s in [0, m) with m | s (R^d - 1): gcd(m, R^d - 1) of them, so the number of slots of exact
// period d follows by Moebius inversion over the divisors of d, and the number of cycles of
// length d is that count / d. A slot is the leader of its cycle when the walk from it (at
// most k steps) meets no smaller slot before returning to it. Not a reference algorithm: the
// reference claims every slot with an atomic bit (src/cl/cl_aos_asta.cl:475-566).
// ---------------------------------------------------------------------------------------
namespace {
uint64_t mulmod64(uint64_t a, uint64_t b, uint64_t m) {
  return uint64_t((unsigned __int128)a * b % m);
}
uint64_t gcd64(uint64_t a, uint64_t b) {
  while (b) {
    const uint64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}
int mobius(uint32_t d) {
  int r = 1;
  for (uint32_t q = 2; q * q <= d; q++) {
    if (d % q) continue;
    d /= q;
    if (d % q == 0) return 0;
    r = -r;
  }
  return d > 1 ? -r : r;
}
}  // namespace

bool plan_100_leaders(uint64_t R, uint64_t C, uint32_t max_order, PlPlan *out) {
  if (R < 2 || C < 2 || R * C >= (1ull << 32) || max_order < 1) return false;
  const uint64_t m = R * C - 1;
  const uint64_t r = R % m;
  uint64_t pw = r;
  uint32_t k = 0;
  for (uint32_t j = 1; j <= max_order; j++) {
    if (pw == 1) {
      k = j;
      break;
    }
    pw = mulmod64(pw, r, m);
  }
  if (!k) return false;
  // F[d]: slots in [1, m) whose period divides d (d | k)
  std::vector<int64_t> F(k + 1, 0);
  uint64_t rd = 1;
  for (uint32_t d = 1; d <= k; d++) {
    rd = mulmod64(rd, r, m);  // R^d mod m
    if (k % d) continue;
    F[d] = int64_t(gcd64(m, (rd + m - 1) % m)) - 1;
  }
  uint64_t cycles = 0;
  for (uint32_t d = 2; d <= k; d++) {
    if (k % d) continue;
    int64_t e = 0;
    for (uint32_t f = 1; f <= d; f++)
      if (d % f == 0) e += mobius(d / f) * F[f];
    cycles += uint64_t(e) / d;
  }
  out->cycles = cycles;
  out->order = k;
  return true;
}

struct PlParams {
  unsigned char *base;
  uint32_t *leaders;   // [0, cycles): the list; [cycles]: its counter
  uint64_t R, C, RC, S, nsub, sub_bytes, cycles, walkers;
  uint32_t m, order, vw_log2;
  FastDiv64 div_nsub, div_cycles;
  FastDiv32 div_C32;
};

__device__ __forceinline__ uint32_t pl_next(const PlParams &p, uint32_t s) {
  const uint32_t q = p.div_C32.div(s);
  return (s - q * uint32_t(p.C)) * uint32_t(p.R) + q;
}

// Appends every cycle leader of slots [1, m) to the list. A workgroup takes chunks of
// kLeadChunk slots, gathers their leaders in LDS (one LDS counter update per wave), then
// reserves room in the list with one global counter update per chunk.
constexpr uint32_t kLeadChunk = 4096;
__global__ __launch_bounds__(256) void k_find_leaders(PlParams p) {
  __shared__ uint32_t buf[kLeadChunk];
  __shared__ uint32_t cnt, base;
  const uint32_t lane = threadIdx.x & 63u;
  for (uint64_t c0 = uint64_t(blockIdx.x) * kLeadChunk; c0 < p.m;
       c0 += uint64_t(gridDim.x) * kLeadChunk) {  // uniform over the workgroup
    if (threadIdx.x == 0) cnt = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kLeadChunk; i += 256) {
      const uint64_t s64 = c0 + i;
      const uint32_t s = uint32_t(s64);
      bool lead = false;
      if (s64 >= 1 && s64 < p.m) {
        uint32_t t = pl_next(p, s);
        lead = t != s;
        for (uint32_t j = 1; lead && t != s; j++) {
          if (t < s || j > p.order) {
            lead = false;
            break;
          }
          t = pl_next(p, t);
        }
      }
      const uint64_t mask = __ballot(lead);
      if (mask) {
        const uint32_t first = uint32_t(__builtin_ctzll(mask));
        uint32_t wb = 0;
        if (lane == first) wb = atomicAdd(&cnt, uint32_t(__builtin_popcountll(mask)));
        wb = __shfl(wb, int(first));
        if (lead) buf[wb + uint32_t(__builtin_popcountll(mask & ((1ull << lane) - 1)))] = s;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0)
      base = cnt ? __hip_atomic_fetch_add(&p.leaders[p.cycles], cnt, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT)
                 : 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < cnt; i += 256)
      if (uint64_t(base) + i < p.cycles) p.leaders[base + i] = buf[i];
    __syncthreads();
  }
}

// One walker (VW lanes) per (instance, leader, piece): carries the piece around the cycle.
// The walk reads each slot before it writes it; cycles are disjoint, so no other walker
// touches these bytes. A cycle of L slots takes L - 1 steps of the loop.
template <int UB, int UPL>
__global__ __launch_bounds__(256) void k_100_leaders(PlParams p) {
  using V = typename Elem<UB>::T;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t vwl = p.vw_log2, vw = 1u << vwl, sub = lane & (vw - 1u);
  const uint64_t gpw = 64u >> vwl;  // walkers per wave
  const uint64_t wave = uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const uint64_t nwaves = uint64_t(gridDim.x) * 4;
  for (uint64_t wb = wave * gpw; wb < p.walkers; wb += nwaves * gpw) {
    const uint64_t w = wb + (lane >> vwl);
    if (w >= p.walkers) continue;
    const uint64_t rest = p.div_nsub.div(w), c = w - rest * p.nsub;
    const uint64_t x = p.div_cycles.div(rest), l = rest - x * p.cycles;
    const uint32_t s0 = p.leaders[l];
    if (s0 == 0 || s0 >= p.m) continue;  // never true for a complete list; keeps accesses in bounds
    unsigned char *ib = p.base + x * p.RC * p.S + c * p.sub_bytes;
    const uint32_t pbytes = uint32_t(std::min<uint64_t>(p.sub_bytes, p.S - c * p.sub_bytes));
    // `a` goes to slot t, whose own content `nx` is already loaded; the slot after t is
    // loaded before `a` is stored, so two loads of the chain are in flight.
    V a[UPL], nx[UPL], nn[UPL];
    load_piece<UB, UPL>(a, ib + uint64_t(s0) * p.S, sub, vw, pbytes);
    uint32_t t = pl_next(p, s0);  // != s0: leaders head cycles of length >= 2
    load_piece<UB, UPL>(nx, ib + uint64_t(t) * p.S, sub, vw, pbytes);
    for (uint32_t j = 0; j < p.order; j++) {  // a cycle has at most `order` slots
      const uint32_t tn = pl_next(p, t);
      const bool last = tn == s0;
      if (!last) load_piece<UB, UPL>(nn, ib + uint64_t(tn) * p.S, sub, vw, pbytes);
      store_piece<UB, UPL>(ib + uint64_t(t) * p.S, a, sub, vw, pbytes);
      if (last) {
        store_piece<UB, UPL>(ib + uint64_t(s0) * p.S, nx, sub, vw, pbytes);
        break;
      }
#pragma unroll
      for (int k = 0; k < UPL; k++) {
        a[k] = nx[k];
        nx[k] = nn[k];
      }
      t = tn;
    }
  }
}

namespace {
template <int UB>
hipError_t launch_pl_ub(const PlParams &p, int upl, uint32_t blocks, hipStream_t st) {
  switch (upl) {
    case 1: hipLaunchKernelGGL((k_100_leaders<UB, 1>), dim3(blocks), dim3(256), 0, st, p); break;
    case 2: hipLaunchKernelGGL((k_100_leaders<UB, 2>), dim3(blocks), dim3(256), 0, st, p); break;
    case 4: hipLaunchKernelGGL((k_100_leaders<UB, 4>), dim3(blocks), dim3(256), 0, st, p); break;
    case 8: hipLaunchKernelGGL((k_100_leaders<UB, 8>), dim3(blocks), dim3(256), 0, st, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
}  // namespace

namespace {
PlParams make_pl(unsigned char *base, uint32_t *list, uint64_t inst, uint64_t R, uint64_t C,
                 uint64_t S, const Pt100Plan &pl, const PlPlan &pc) {
  PlParams p{};
  p.base = base;
  p.leaders = list;
  p.R = R;
  p.C = C;
  p.RC = R * C;
  p.S = S;
  p.nsub = pl.nsub;
  p.sub_bytes = pl.sub_bytes;
  p.cycles = pc.cycles;
  p.walkers = inst * pc.cycles * pl.nsub;
  p.m = uint32_t(R * C - 1);
  p.order = pc.order;
  p.vw_log2 = pl.vw_log2;
  p.div_nsub = FastDiv64(pl.nsub);
  p.div_cycles = FastDiv64(pc.cycles);
  p.div_C32 = FastDiv32(uint32_t(C));
  return p;
}
}  // namespace

hipError_t launch_find_leaders(uint32_t *list, uint64_t R, uint64_t C, const PlPlan &pc,
                               hipStream_t st) {
  if (pc.cycles == 0) return hipSuccess;
  if (R * C >= (1ull << 32) || pc.cycles >= (1ull << 32)) return hipErrorInvalidValue;
  Pt100Plan none{};
  none.nsub = 1;
  const PlParams p = make_pl(nullptr, list, 1, R, C, 1, none, pc);
  hipError_t e = launch_fill_u32(list + pc.cycles, 1, 0, st);
  if (e != hipSuccess) return e;
  const uint64_t chunks = (uint64_t(p.m) + kLeadChunk - 1) / kLeadChunk;
  const uint32_t fblocks = uint32_t(std::min<uint64_t>(chunks, uint64_t(device_info().cus) * 8));
  KernelSpan span(st, "k_find_leaders", 0.0);
  hipLaunchKernelGGL(k_find_leaders, dim3(fblocks), dim3(256), 0, st, p);
  return hipGetLastError();
}

hipError_t launch_leader_walk(unsigned char *base, const uint32_t *list, uint64_t inst,
                              uint64_t R, uint64_t C, uint64_t S, const Pt100Plan &pl,
                              const PlPlan &pc, uint32_t max_blocks, hipStream_t st) {
  if (inst == 0 || pc.cycles == 0) return hipSuccess;
  if (R * C >= (1ull << 32) || pc.cycles >= (1ull << 32)) return hipErrorInvalidValue;
  const PlParams p = make_pl(base, const_cast<uint32_t *>(list), inst, R, C, S, pl, pc);
  const uint64_t per_block = 4ull * (64u >> pl.vw_log2);
  const uint32_t blocks =
      uint32_t(std::min<uint64_t>((p.walkers + per_block - 1) / per_block, max_blocks));
  if (blocks == 0) return hipErrorInvalidValue;
  KernelSpan span(st, "k_100_leaders", 2.0 * double(inst) * double(R) * double(C) * double(S));
  switch (pl.ub) {
    case 1: return launch_pl_ub<1>(p, pl.upl, blocks, st);
    case 2: return launch_pl_ub<2>(p, pl.upl, blocks, st);
    case 4: return launch_pl_ub<4>(p, pl.upl, blocks, st);
    case 8: return launch_pl_ub<8>(p, pl.upl, blocks, st);
    case 16: return launch_pl_ub<16>(p, pl.upl, blocks, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_100_leaders(unsigned char *base, uint32_t *scratch, uint64_t inst,
                              uint64_t R, uint64_t C, uint64_t S, const Pt100Plan &pl,
                              const PlPlan &pc, uint32_t max_blocks, hipStream_t st) {
  hipError_t e = launch_find_leaders(scratch, R, C, pc, st);
  if (e != hipSuccess) return e;
  return launch_leader_walk(base, scratch, inst, R, C, S, pl, pc, max_blocks, st);
}

}  // namespace lm
