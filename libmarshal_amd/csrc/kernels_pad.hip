// kernels_pad.hip - in-place row padding [y][x] -> [y][pad] (new columns zeroed) and
// unpadding [y][pad] -> [y][x] (freed tail zeroed) for gfx950.
//
// Reference: padding / unpadding, src/cl/cl_aos_asta.cl:813-938, launched by
// src/cl_marshal.cc:821-956. The reference gives every workgroup a dynamic id (atom_add) and
// chains them: group k stores only after group k-1 has loaded, and signals k+1 only after
// its own wait, so the signal crosses every group in sequence.
//
// This version keeps the dynamic ids and the "store only after every earlier group has
// loaded" rule, but publishes the two facts separately, as in a decoupled look-back: a group
// marks itself LOADED as soon as its loads have returned, then walks back over earlier
// groups until it finds one marked COMPLETE (that group and all before it have loaded),
// marks itself COMPLETE and stores. No group waits on a later one, and a group that got its
// id earlier is already resident, so every wait ends.
//
// Why earlier groups are the only hazard: padding moves every element to an address >= its
// source and is processed from the tail; unpadding moves every element to an address <= its
// source and is processed from the front. In both cases a later group's sources lie entirely
// on its own side of its destination range, so only earlier groups can still need to read
// the region a group overwrites.
#include "kernels.h"

#include <algorithm>

namespace lm {

constexpr int kPadThreads = 256;
constexpr int kPadPer = 16;  // elements per thread
constexpr uint64_t kPadTile = uint64_t(kPadThreads) * kPadPer;
constexpr uint32_t kLoaded = 1, kComplete = 2;
constexpr int kPadVecPer = 8;  // 16-byte vectors per lane in the 16-byte form
constexpr uint64_t kPadVecTile = uint64_t(kPadThreads) * kPadVecPer;

// Enough state words for either form: the element form's groups of kPadTile elements and the
// 16-byte form's groups of kPadVecTile vectors.
uint64_t pad_flag_words(uint64_t total_elems, size_t es) {
  const uint64_t a = (total_elems + kPadTile - 1) / kPadTile;
  const uint64_t b = (total_elems * es + kPadVecTile * 16 - 1) / (kPadVecTile * 16);
  return 1 + std::max(a, b);
}

struct PadParams {
  unsigned char *base;
  uint64_t x, y, pad;
  uint64_t total;  // y * pad
  uint64_t ntiles;
  uint32_t *flags;  // [0] = id counter, [1 + k] = state of group k
  FastDiv64 div_pad, div_x;
  bool unpad;
};

// Wait until every group before k (in processing order) has loaded, then mark k COMPLETE.
// Group k marks itself LOADED first, so later groups can skip over it. The first wave reads
// the states of 64 predecessors per step (lane l reads group j - l): the walk ends at the
// nearest COMPLETE group when every group between it and k is at least LOADED, and moves a
// whole window back when all 64 are LOADED. States only ever grow (0 -> LOADED ->
// COMPLETE), so re-reading a window after a short sleep is always safe, and every wait ends
// because each earlier group already holds its id and is running or done.
__device__ __forceinline__ void pad_lookback(uint32_t *flags, uint64_t k) {
  if (threadIdx.x < 64) {
    uint32_t *st = flags + 1;
    const uint32_t lane = threadIdx.x;
    if (k == 0) {
      if (lane == 0)
        __hip_atomic_store(&st[0], kComplete, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (lane == 0)
        __hip_atomic_store(&st[k], kLoaded, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const int64_t top0 = int64_t(k) - 1;  // newest predecessor
      int64_t top = top0;
      uint64_t below = 0;
      while (true) {
        const int64_t j = top - int64_t(lane);
        const uint32_t s = j >= 0 ? __hip_atomic_load(&st[j], __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                                  : kComplete;  // nothing precedes group 0
        const uint64_t done = __ballot(s == kComplete);
        const uint64_t waiting = __ballot(s != kComplete && s != kLoaded);
        if (done) {
          below = (done & (~done + 1)) - 1;  // lanes before the nearest COMPLETE
          if ((waiting & below) == 0) break;
        } else if (waiting == 0) {
          top -= 64;
          continue;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      // Every group seen LOADED on the way is now known to have all its predecessors loaded
      // too: mark the nearest ones COMPLETE on their behalf (a true, idempotent statement) so
      // later walks stop early instead of crossing every group still in flight.
      const int64_t jb = top - int64_t(lane);
      if (((below >> lane) & 1u) && jb >= 0)
        __hip_atomic_store(&st[jb], kComplete, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const int64_t j0 = top0 - int64_t(lane);
      if (top != top0 && j0 >= 0)
        __hip_atomic_store(&st[j0], kComplete, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (lane == 0)
        __hip_atomic_store(&st[k], kComplete, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __syncthreads();
}

__device__ __forceinline__ uint32_t pad_group_id(uint32_t *flags) {
  __shared__ uint32_t sid;
  if (threadIdx.x == 0)
    sid = __hip_atomic_fetch_add(&flags[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  return sid;
}

template <int ES>
__global__ __launch_bounds__(kPadThreads) void k_pad(PadParams p) {
  using E = typename Elem<ES>::T;
  const uint64_t k = pad_group_id(p.flags);
  uint64_t lo, hi;
  if (!p.unpad) {  // tail first
    hi = p.total - k * kPadTile;
    lo = hi > kPadTile ? hi - kPadTile : 0;
  } else {         // front first
    lo = k * kPadTile;
    hi = std::min<uint64_t>(lo + kPadTile, p.total);
  }
  E *m = reinterpret_cast<E *>(p.base);
  E v[kPadPer];
#pragma unroll
  for (int j = 0; j < kPadPer; j++) {
    uint64_t d = lo + threadIdx.x + uint64_t(j) * kPadThreads;
    E val{};
    if (d < hi) {
      if (!p.unpad) {
        uint64_t r = p.div_pad.div(d), c = d - r * p.pad;
        if (c < p.x) val = m[r * p.x + c];
      } else if (d < p.y * p.x) {
        uint64_t r = p.div_x.div(d), c = d - r * p.x;
        val = m[r * p.pad + c];
      }
    }
    v[j] = val;
  }
  wait_vm_loads();
  __syncthreads();
  pad_lookback(p.flags, k);
#pragma unroll
  for (int j = 0; j < kPadPer; j++) {
    uint64_t d = lo + threadIdx.x + uint64_t(j) * kPadThreads;
    if (d < hi) m[d] = v[j];
  }
}

// ---- 16-byte form (buffer base 16-byte aligned) ------------------------------------------
// Each lane produces whole 16-byte destination vectors. A vector that lies inside one row's
// data is a contiguous 16-byte source window at any byte alignment: it is read as the one or
// two aligned 16-byte chunks that cover it and shifted into place (v_alignbyte). Vectors that
// touch a row end, the pad, or the end of the buffer are done element by element.

template <int ES>
__global__ __launch_bounds__(kPadThreads) void k_pad_vec(PadParams p) {
  using E = typename Elem<ES>::T;
  constexpr uint64_t EPV = 16 / ES;
  union Vec {
    U128 v;
    E e[EPV];
  };
  const uint64_t k = pad_group_id(p.flags);
  const uint64_t nvec = (p.total * ES + 15) / 16;
  const uint64_t bytes = p.total * ES;
  uint64_t lo, hi;
  if (!p.unpad) {  // tail first
    hi = nvec - k * kPadVecTile;
    lo = hi > kPadVecTile ? hi - kPadVecTile : 0;
  } else {         // front first
    lo = k * kPadVecTile;
    hi = std::min<uint64_t>(lo + kPadVecTile, nvec);
  }
  const uint64_t X = p.x, P = p.pad;
  const uint64_t live = p.unpad ? p.y * X : p.total;  // destination elements that carry data
  const E *m = reinterpret_cast<const E *>(p.base);
  const uint64_t rowlen = p.unpad ? X : P;
  const uint64_t src_pitch = p.unpad ? P : X;
  Vec v[kPadVecPer];
  U128 wlo[kPadVecPer], whi[kPadVecPer];
  uint64_t wsb[kPadVecPer];
  bool win[kPadVecPer];
  // Issue every load first (window chunks, or single elements near row ends and the pad),
  // then shift the windows into place: the loads of all vectors overlap.
#pragma unroll
  for (int j = 0; j < kPadVecPer; j++) {
    const uint64_t vi = lo + threadIdx.x + uint64_t(j) * kPadThreads;
    v[j].v = U128{0u, 0u, 0u, 0u};
    win[j] = false;
    wsb[j] = 0;
    if (vi >= hi) continue;
    const uint64_t d = vi * EPV;
    // Row and column of the first destination element.
    const uint64_t r = p.unpad ? p.div_x.div(d) : p.div_pad.div(d);
    const uint64_t c = d - r * rowlen;
    const uint64_t sb = (r * src_pitch + c) * ES;
    if (c + EPV <= X && d + EPV <= live && (sb & ~uint64_t(15)) + 32 <= bytes) {
      win[j] = true;
      wsb[j] = sb;
      window_raw(p.base, sb, wlo[j], whi[j]);
    } else {
#pragma unroll
      for (uint32_t e = 0; e < EPV; e++) {
        const uint64_t dd = d + e;
        if (dd >= live) break;
        const uint64_t r2 = p.unpad ? p.div_x.div(dd) : p.div_pad.div(dd);
        const uint64_t c2 = dd - r2 * rowlen;
        if (c2 < X) v[j].e[e] = m[r2 * src_pitch + c2];
      }
    }
  }
#pragma unroll
  for (int j = 0; j < kPadVecPer; j++)
    if (win[j]) v[j].v = window_join(wsb[j], wlo[j], whi[j]);
  wait_vm_loads();
  __syncthreads();
  pad_lookback(p.flags, k);
#pragma unroll
  for (int j = 0; j < kPadVecPer; j++) {
    const uint64_t vi = lo + threadIdx.x + uint64_t(j) * kPadThreads;
    if (vi >= hi) continue;
    if ((vi + 1) * 16 <= bytes) {
      *reinterpret_cast<U128 *>(p.base + vi * 16) = v[j].v;
    } else {
      E *out = reinterpret_cast<E *>(p.base);
      for (uint32_t e = 0; e < EPV; e++)
        if (vi * EPV + e < p.total) out[vi * EPV + e] = v[j].e[e];
    }
  }
}

hipError_t launch_padding(unsigned char *base, size_t es, uint64_t x, uint64_t y, uint64_t pad,
                          bool unpad, uint32_t *flags, hipStream_t st) {
  PadParams p{};
  p.base = base;
  p.x = x;
  p.y = y;
  p.pad = pad;
  p.total = y * pad;
  p.ntiles = (p.total + kPadTile - 1) / kPadTile;
  p.flags = flags;
  p.unpad = unpad;
  p.div_pad = FastDiv64(pad ? pad : 1);
  p.div_x = FastDiv64(x ? x : 1);
  if (p.ntiles == 0) return hipSuccess;
  hipError_t e = launch_fill_u32(flags, pad_flag_words(p.total, es), 0, st);
  if (e != hipSuccess) return e;
  if ((reinterpret_cast<uintptr_t>(base) & 15u) == 0) {
    const uint64_t nvec = (p.total * es + 15) / 16;
    dim3 grid(uint32_t((nvec + kPadVecTile - 1) / kPadVecTile)), blk(kPadThreads);
    switch (es) {
      case 1: hipLaunchKernelGGL(k_pad_vec<1>, grid, blk, 0, st, p); break;
      case 2: hipLaunchKernelGGL(k_pad_vec<2>, grid, blk, 0, st, p); break;
      case 4: hipLaunchKernelGGL(k_pad_vec<4>, grid, blk, 0, st, p); break;
      case 8: hipLaunchKernelGGL(k_pad_vec<8>, grid, blk, 0, st, p); break;
      case 16: hipLaunchKernelGGL(k_pad_vec<16>, grid, blk, 0, st, p); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  dim3 grid(uint32_t(p.ntiles)), blk(kPadThreads);
  switch (es) {
    case 1: hipLaunchKernelGGL(k_pad<1>, grid, blk, 0, st, p); break;
    case 2: hipLaunchKernelGGL(k_pad<2>, grid, blk, 0, st, p); break;
    case 4: hipLaunchKernelGGL(k_pad<4>, grid, blk, 0, st, p); break;
    case 8: hipLaunchKernelGGL(k_pad<8>, grid, blk, 0, st, p); break;
    case 16: hipLaunchKernelGGL(k_pad<16>, grid, blk, 0, st, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace lm
