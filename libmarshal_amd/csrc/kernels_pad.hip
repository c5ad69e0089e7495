// kernels_pad.hip - in-place row padding [y][x] -> [y][pad] (new columns zeroed) and
// unpadding [y][pad] -> [y][x] (freed tail zeroed) for gfx950.
//
// Reference: padding / unpadding, src/cl/cl_aos_asta.cl:813-938, launched by
// src/cl_marshal.cc:821-956. The reference gives every workgroup a dynamic id (atom_add) and
// chains them: group k stores only after group k-1 has loaded, and signals k+1 only after
// its own wait, so the signal crosses every group in sequence.
//
// This version keeps the dynamic ids but drops the chain: a group marks itself LOADED as
// soon as its loads have returned and then waits only for the groups that read from its own
// destination range (one short id range, computed from the row maps; see pad_reader_ids).
// No group waits on a later one, and a group that got its id earlier is already resident,
// so every wait ends.
//
// Why readers are always earlier groups: padding moves every element to an address >= its
// source and is processed from the tail; unpadding moves every element to an address <= its
// source and is processed from the front. In both cases the groups that read a range write
// on the far side of it, and so come first.
#include "kernels.h"

#include <algorithm>

namespace lm {

constexpr int kPadThreads = 256;
constexpr int kPadPer = 16;  // elements per thread
constexpr uint64_t kPadTile = uint64_t(kPadThreads) * kPadPer;
constexpr uint32_t kLoaded = 1;
// 16-byte form: NT lanes x VPL vectors per group (NT*VPL a power of two, >= 1024); the
// state words are sized for the smallest group, 1024 vectors (16 KiB).
constexpr uint64_t kPadVecMinTile = 1024;
static_assert(kPadTile == 4096, "element-form group size is 2^12");

// Enough state words for either form: the element form's groups of kPadTile elements and the
// 16-byte form's groups of at least kPadVecMinTile vectors.
uint64_t pad_flag_words(uint64_t total_elems, size_t es) {
  const uint64_t a = (total_elems + kPadTile - 1) / kPadTile;
  const uint64_t b = (total_elems * es + kPadVecMinTile * 16 - 1) / (kPadVecMinTile * 16);
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
  uint32_t nt;  // vector form: non-temporal loads and stores
};

// Groups are numbered by a dynamic id in processing order (tail first for padding, front
// first for unpadding). A group may only store once every group that reads from its
// destination range has loaded. Those readers form one contiguous id range, computed here
// from the row maps, and always precede the group: padding moves every element to an address
// >= its source and processes from the tail; unpadding moves every element to an address <=
// its source and processes from the front, so a reader of a range is always processed
// before the range's writer. Waiting only for these few groups, instead of for every earlier
// group, keeps one slow load from holding up every store behind it.
//
// [Dlo, Dhi): destination elements of the group; ids count groups of VT units of U elements
// (U, VT powers of two), nunits units in all. Returns glo > ghi when nobody reads the range.
__device__ __forceinline__ void pad_reader_ids(const PadParams &p, uint64_t Dlo, uint64_t Dhi,
                                               uint32_t u_log2, uint32_t vt_log2,
                                               uint64_t nunits, int64_t &glo, int64_t &ghi) {
  glo = 1;
  ghi = 0;
  const uint64_t X = p.x, P = p.pad, Y = p.y;
  if (!p.unpad) {  // sources are compact positions s < Y*X; reader of s writes (s/X)*P + s%X
    if (Dlo >= Y * X) return;
    const uint64_t plo = Dlo, phi = std::min<uint64_t>(Dhi, Y * X) - 1;
    const uint64_t rlo = p.div_x.div(plo), rhi = p.div_x.div(phi);
    const uint64_t dlo = rlo * P + (plo - rlo * X), dhi = rhi * P + (phi - rhi * X);
    glo = int64_t((nunits - 1 - (dhi >> u_log2)) >> vt_log2);
    ghi = int64_t((nunits - 1 - (dlo >> u_log2)) >> vt_log2);
  } else {  // sources are padded positions with column < X; reader writes row*X + column
    uint64_t row = p.div_pad.div(Dlo), c = Dlo - row * P;
    if (c >= X) {
      row++;
      c = 0;
    }
    const uint64_t row2 = p.div_pad.div(Dhi - 1);
    uint64_t c2 = Dhi - 1 - row2 * P;
    if (c2 >= X) c2 = X - 1;
    if (row > row2 || (row == row2 && c > c2)) return;
    glo = int64_t(((row * X + c) >> u_log2) >> vt_log2);
    ghi = int64_t(((row2 * X + c2) >> u_log2) >> vt_log2);
  }
}

// Marks group k LOADED (its loads have returned), then waits until groups glo..ghi (other
// than k) are LOADED. They hold smaller ids, so they are running or done and every wait ends.
__device__ __forceinline__ void pad_wait_readers(uint32_t *flags, uint64_t k, int64_t glo,
                                                 int64_t ghi) {
  if (threadIdx.x < 64) {
    uint32_t *st = flags + 1;
    const uint32_t lane = threadIdx.x;
    if (lane == 0)
      __hip_atomic_store(&st[k], kLoaded, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (ghi > int64_t(k) - 1) ghi = int64_t(k) - 1;
    if (glo < 0) glo = 0;
    for (int64_t b = glo; b <= ghi; b += 64) {
      const int64_t j = b + int64_t(lane);
      while (true) {
        const uint32_t s = j <= ghi ? __hip_atomic_load(&st[j], __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT)
                                    : kLoaded;
        if (__ballot(s != kLoaded) == 0) break;
        __builtin_amdgcn_s_sleep(1);
      }
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
  int64_t glo, ghi;
  pad_reader_ids(p, lo, hi, 0, 12, p.total, glo, ghi);  // kPadTile = 2^12 elements
  pad_wait_readers(p.flags, k, glo, ghi);
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

// NTM: non-temporal loads and stores (a template argument, so that the two load forms never
// share result registers: a select between them would make every load wait for the last).
template <int ES, int NT, int VPL, bool NTM>
__global__ __launch_bounds__(NT) void k_pad_vec(PadParams p) {
  constexpr uint64_t kPadVecTile = uint64_t(NT) * VPL;
  constexpr uint32_t kPadVecTileLog2 = __builtin_ctzll(kPadVecTile);
  static_assert((kPadVecTile & (kPadVecTile - 1)) == 0 && kPadVecTile >= kPadVecMinTile, "");
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
  Vec v[VPL];
  U128 wlo[VPL], whi[VPL];
  uint64_t wsb[VPL];
  bool win[VPL];
  // Issue every window load first, then the single-element loads of the vectors near row
  // ends and the pad, then shift the windows into place: the loads of all vectors overlap.
  // Registers that receive window chunks get no other value before the join (an assignment
  // there would have to wait for the load in flight).
#pragma unroll
  for (int j = 0; j < VPL; j++) {
    const uint64_t vi = lo + threadIdx.x + uint64_t(j) * NT;
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
      window_raw<NTM>(p.base, sb, wlo[j], whi[j]);
    }
  }
#pragma unroll
  for (int j = 0; j < VPL; j++) {
    const uint64_t vi = lo + threadIdx.x + uint64_t(j) * NT;
    if (win[j] || vi >= hi) continue;
    v[j].v = U128{0u, 0u, 0u, 0u};
    const uint64_t d = vi * EPV;
    // row and column of the first element, then stepped along (one division per vector)
    uint64_t r2 = p.unpad ? p.div_x.div(d) : p.div_pad.div(d);
    uint64_t c2 = d - r2 * rowlen;
    // A loop, not unrolled: unrolled, the u8 form needs 124 VGPRs instead of 66 (this path
    // runs only at row ends and the pad).
#pragma unroll 1
    for (uint32_t e = 0; e < EPV; e++) {
      if (d + e >= live) break;
      if (c2 < X) v[j].e[e] = m[r2 * src_pitch + c2];
      if (++c2 == rowlen) {
        c2 = 0;
        r2++;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < VPL; j++)
    if (win[j]) v[j].v = window_join(wsb[j], wlo[j], whi[j]);
  wait_vm_loads();
  __syncthreads();
  int64_t glo, ghi;
  constexpr uint32_t kEpvLog2 = ES == 1 ? 4 : ES == 2 ? 3 : ES == 4 ? 2 : ES == 8 ? 1 : 0;
  pad_reader_ids(p, lo * EPV, std::min<uint64_t>(hi * EPV, p.total), kEpvLog2, kPadVecTileLog2,
                 nvec, glo, ghi);
  pad_wait_readers(p.flags, k, glo, ghi);
#pragma unroll
  for (int j = 0; j < VPL; j++) {
    const uint64_t vi = lo + threadIdx.x + uint64_t(j) * NT;
    if (vi >= hi) continue;
    if ((vi + 1) * 16 <= bytes) {
      if (NTM)
        __builtin_nontemporal_store(v[j].v, reinterpret_cast<U128 *>(p.base + vi * 16));
      else
        *reinterpret_cast<U128 *>(p.base + vi * 16) = v[j].v;
    } else {
      E *out = reinterpret_cast<E *>(p.base);
      for (uint32_t e = 0; e < EPV; e++)
        if (vi * EPV + e < p.total) out[vi * EPV + e] = v[j].e[e];
    }
  }
}

namespace {
template <int ES, int NT, int VPL>
hipError_t launch_pad_vec_t(const PadParams &p, uint64_t nvec, hipStream_t st) {
  const uint64_t tile = uint64_t(NT) * VPL;
  const dim3 grid(uint32_t((nvec + tile - 1) / tile));
  if (p.nt)
    hipLaunchKernelGGL((k_pad_vec<ES, NT, VPL, true>), grid, dim3(NT), 0, st, p);
  else
    hipLaunchKernelGGL((k_pad_vec<ES, NT, VPL, false>), grid, dim3(NT), 0, st, p);
  return hipGetLastError();
}
template <int ES>
hipError_t launch_pad_vec(const PadParams &p, uint64_t nvec, int64_t shape, hipStream_t st) {
  switch (shape) {
    case 1: return launch_pad_vec_t<ES, 256, 4>(p, nvec, st);
    case 2: return launch_pad_vec_t<ES, 256, 8>(p, nvec, st);
    case 3: return launch_pad_vec_t<ES, 1024, 2>(p, nvec, st);
    case 4: return launch_pad_vec_t<ES, 1024, 4>(p, nvec, st);
    default: return launch_pad_vec_t<ES, 512, 4>(p, nvec, st);
  }
}
}  // namespace

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
  p.nt = uint32_t(tuning("pad_nt"));
  if (p.ntiles == 0) return hipSuccess;
  hipError_t e = launch_fill_u32(flags, pad_flag_words(p.total, es), 0, st);
  if (e != hipSuccess) return e;
  KernelSpan span(st, unpad ? "k_unpad" : "k_pad", 2.0 * double(x) * double(y) * double(es));
  if ((reinterpret_cast<uintptr_t>(base) & 15u) == 0) {
    const uint64_t nvec = (p.total * es + 15) / 16;
    // lanes x vectors per lane: 0: 512 x 4 (default), 1: 256 x 4, 2: 256 x 8,
    // 3: 1024 x 2, 4: 1024 x 4
    const int64_t shape = tuning("pad_shape");
    switch (es) {
      case 1: return launch_pad_vec<1>(p, nvec, shape, st);
      case 2: return launch_pad_vec<2>(p, nvec, shape, st);
      case 4: return launch_pad_vec<4>(p, nvec, shape, st);
      case 8: return launch_pad_vec<8>(p, nvec, shape, st);
      case 16: return launch_pad_vec<16>(p, nvec, shape, st);
      default: return hipErrorInvalidValue;
    }
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
