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

uint64_t pad_flag_words(uint64_t total_elems) {
  return 1 + (total_elems + kPadTile - 1) / kPadTile;
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

template <int ES>
__global__ __launch_bounds__(kPadThreads) void k_pad(PadParams p) {
  using E = typename Elem<ES>::T;
  __shared__ uint32_t sid;
  if (threadIdx.x == 0) sid = __hip_atomic_fetch_add(&p.flags[0], 1u, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const uint64_t k = sid;
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
  if (threadIdx.x == 0) {
    uint32_t *st = p.flags + 1;
    if (k == 0) {
      __hip_atomic_store(&st[0], kComplete, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      __hip_atomic_store(&st[k], kLoaded, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint64_t j = k - 1;
      while (true) {
        uint32_t s = __hip_atomic_load(&st[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (s == kComplete) break;
        if (s == kLoaded) {
          --j;  // j >= 1 here: group 0 only ever publishes kComplete
          continue;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      __hip_atomic_store(&st[k], kComplete, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kPadPer; j++) {
    uint64_t d = lo + threadIdx.x + uint64_t(j) * kPadThreads;
    if (d < hi) m[d] = v[j];
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
  hipError_t e = launch_fill_u32(flags, pad_flag_words(p.total), 0, st);
  if (e != hipSuccess) return e;
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
