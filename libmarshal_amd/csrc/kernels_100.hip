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

#include <algorithm>

namespace lm {

template <bool I64>
__device__ __forceinline__ uint64_t next_slot(uint64_t s, const Pt100Params &p) {
  if constexpr (I64) {
    uint64_t q = p.div_C64.div(s);
    return (s - q * p.C) * p.R + q;
  } else {
    uint32_t s32 = uint32_t(s);
    uint32_t q = p.div_C32.div(s32);
    return uint64_t((s32 - q * uint32_t(p.C)) * uint32_t(p.R) + q);
  }
}

template <int UB, int UPL>
__device__ __forceinline__ void load_piece(typename Elem<UB>::T (&d)[UPL],
                                           const unsigned char *src, uint32_t sub,
                                           uint32_t vw, uint32_t pbytes) {
  using V = typename Elem<UB>::T;
#pragma unroll
  for (int k = 0; k < UPL; k++) {
    uint32_t off = (sub + k * vw) * UB;
    if (off < pbytes) d[k] = *reinterpret_cast<const V *>(src + off);
  }
}

template <int UB, int UPL>
__device__ __forceinline__ void store_piece(unsigned char *dst,
                                            const typename Elem<UB>::T (&d)[UPL],
                                            uint32_t sub, uint32_t vw, uint32_t pbytes) {
  using V = typename Elem<UB>::T;
#pragma unroll
  for (int k = 0; k < UPL; k++) {
    uint32_t off = (sub + k * vw) * UB;
    if (off < pbytes) *reinterpret_cast<V *>(dst + off) = d[k];
  }
}

template <int UB, int UPL, bool I64>
__global__ __launch_bounds__(256) void k_100_pttwac(Pt100Params p) {
  using V = typename Elem<UB>::T;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t vwl = p.vw_log2;
  const uint32_t vw = 1u << vwl;
  const uint32_t sub = lane & (vw - 1u);
  const uint32_t leader = lane & ~(vw - 1u);
  const bool is_leader = (sub == 0);
  const uint64_t gpw = 64u >> vwl;  // walkers per wave
  const uint64_t wave = uint64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint64_t nwaves = uint64_t(gridDim.x) * (blockDim.x >> 6);
  const uint64_t RC = p.RC, S = p.S;

  for (uint64_t wb = wave * gpw; wb < p.walkers; wb += nwaves * gpw) {
    const uint64_t w = wb + (lane >> vwl);
    bool walking = w < p.walkers;
    uint64_t x = 0, s = 0, c = 0, nxt = 0;
    if (walking) {
      uint64_t sc = p.div_nsub.div(w);
      c = w - sc * p.nsub;
      x = p.div_RC.div(sc);
      s = sc - x * RC;
      nxt = next_slot<I64>(s, p);
      if (nxt == s) walking = false;
    }
    unsigned char *ibase = p.base + x * RC * S + c * p.sub_bytes;
    const uint64_t fbase = x * RC;
    const uint32_t pbytes = uint32_t(std::min<uint64_t>(p.sub_bytes, S - c * p.sub_bytes));

    // Skip start points whose slot is already final without touching their data.
    uint32_t f = 0;
    if (walking && is_leader) f = flag_load(&p.flags[w >> 5]);
    uint64_t bal = __ballot(walking && is_leader && ((f >> (w & 31)) & 1u));
    if ((bal >> leader) & 1u) walking = false;

    V d1[UPL], d2[UPL];
    if (walking) load_piece<UB, UPL>(d1, ibase + s * S, sub, vw, pbytes);
    wait_vm_loads();
    if (walking && is_leader) f = flag_load(&p.flags[w >> 5]);
    bal = __ballot(walking && is_leader && ((f >> (w & 31)) & 1u));
    if ((bal >> leader) & 1u) walking = false;

    while (walking) {
      load_piece<UB, UPL>(d2, ibase + nxt * S, sub, vw, pbytes);
      const uint64_t fi = (fbase + nxt) * p.nsub + c;
      const uint32_t mask = 1u << (fi & 31);
      uint32_t old = 0;
      if (is_leader) old = flag_or(&p.flags[fi >> 5], mask);
      uint64_t lost = __ballot(is_leader && (old & mask));
      if ((lost >> leader) & 1u) break;
      store_piece<UB, UPL>(ibase + nxt * S, d1, sub, vw, pbytes);
#pragma unroll
      for (int k = 0; k < UPL; k++) d1[k] = d2[k];
      nxt = next_slot<I64>(nxt, p);
    }
  }
}

__global__ void k_fill_u32(uint32_t *p, uint64_t words, uint32_t value) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (; i < words; i += stride) p[i] = value;
}

hipError_t launch_fill_u32(uint32_t *p, uint64_t words, uint32_t value, hipStream_t st) {
  if (words == 0) return hipSuccess;
  uint64_t blocks = std::min<uint64_t>((words + 255) / 256, uint64_t(device_info().cus) * 16);
  hipLaunchKernelGGL(k_fill_u32, dim3(uint32_t(blocks)), dim3(256), 0, st, p, words, value);
  return hipGetLastError();
}

static uint64_t next_pow2(uint64_t v) {
  uint64_t r = 1;
  while (r < v) r <<= 1;
  return r;
}

void plan_100(size_t es, uint64_t inst, uint64_t R, uint64_t C, uint64_t b, Pt100Plan *out) {
  uint64_t S = b * es;
  int ub = (S % 16 == 0) ? 16 : (S % 8 == 0) ? 8 : (S % 4 == 0) ? 4 : (S % 2 == 0) ? 2 : 1;
  uint64_t nU = S / ub;
  uint64_t nsub = 1, per = nU;
  int upl = 1;
  uint32_t vwl = 0;
  if (nU <= 64) {
    uint64_t vw = next_pow2(nU);
    while ((1ull << vwl) < vw) vwl++;
  } else {
    vwl = 6;
    const uint64_t cap = 64 * 8;
    if (nU > cap) {
      nsub = (nU + cap - 1) / cap;
      per = (nU + nsub - 1) / nsub;
    }
    upl = int(next_pow2((per + 63) / 64));
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

hipError_t launch_100_pttwac(unsigned char *base, uint32_t *flags, uint64_t inst, uint64_t R,
                             uint64_t C, uint64_t S, const Pt100Plan &pl, uint32_t max_blocks,
                             hipStream_t st) {
  Pt100Params p{};
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
  uint64_t per_block = 4ull * (64u >> pl.vw_log2);
  uint64_t blocks = std::min<uint64_t>((p.walkers + per_block - 1) / per_block, max_blocks);
  if (blocks == 0) return hipSuccess;
  switch (pl.ub) {
    case 1: return launch_100_ub<1>(p, pl.upl, uint32_t(blocks), st);
    case 2: return launch_100_ub<2>(p, pl.upl, uint32_t(blocks), st);
    case 4: return launch_100_ub<4>(p, pl.upl, uint32_t(blocks), st);
    case 8: return launch_100_ub<8>(p, pl.upl, uint32_t(blocks), st);
    case 16: return launch_100_ub<16>(p, pl.upl, uint32_t(blocks), st);
  }
  return hipErrorInvalidValue;
}

}  // namespace lm
