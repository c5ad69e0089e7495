// common.h - shared device/host helpers of the gfx950 layout-conversion library.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <cstdio>

namespace lm {

// ---------------------------------------------------------------------------------------
// Division by a run-time invariant divisor with one high multiply.
// Granlund & Montgomery, "Division by Invariant Integers using Multiplication" (PLDI 1994),
// Figure 4.1: with l = ceil(log2 d), m' = floor(2^N (2^l - d) / d) + 1,
//   t1 = mulhi(m', n);  q = (t1 + ((n - t1) >> sh1)) >> sh2,  sh1 = min(l,1), sh2 = max(l-1,0)
// exact for every N-bit n and 1 <= d < 2^N. The reference divides with the hardware divide
// in its index map (include/plan.hpp:54-56); here the divisor is fixed per launch.
// ---------------------------------------------------------------------------------------
struct FastDiv32 {
  uint32_t d = 1, m = 0, sh1 = 0, sh2 = 0;
  FastDiv32() = default;
  __host__ explicit FastDiv32(uint32_t dd) : d(dd) {
    uint32_t l = 0;
    while (l < 32 && (uint64_t(1) << l) < dd) l++;
    uint64_t num = ((uint64_t(1) << l) - dd) << 32;  // < 2^64 since 2^l - d < 2^32
    m = uint32_t(num / dd + 1);
    sh1 = l < 1 ? l : 1;
    sh2 = l > 1 ? l - 1 : 0;
  }
  __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const {
    uint32_t t = uint32_t((uint64_t(m) * n) >> 32);
    return (t + ((n - t) >> sh1)) >> sh2;
  }
};

struct FastDiv64 {
  uint64_t d = 1, m = 0;
  uint32_t sh1 = 0, sh2 = 0;
  FastDiv64() = default;
  __host__ explicit FastDiv64(uint64_t dd) : d(dd) {
    uint32_t l = 0;
    while (l < 64 && (uint64_t(1) << l) < dd) l++;
    unsigned __int128 two_l = (unsigned __int128)1 << l;
    unsigned __int128 num = (two_l - dd) << 64;
    m = uint64_t(num / dd + 1);
    sh1 = l < 1 ? l : 1;
    sh2 = l > 1 ? l - 1 : 0;
  }
  __host__ __device__ __forceinline__ uint64_t div(uint64_t n) const {
    uint64_t t = uint64_t(((unsigned __int128)m * n) >> 64);
    return (t + ((n - t) >> sh1)) >> sh2;
  }
};

// ---------------------------------------------------------------------------------------
// Element storage by size. Elements are opaque: they are only loaded and stored.
// ---------------------------------------------------------------------------------------
template <int ES> struct Elem;
template <> struct Elem<1> { using T = uint8_t; };
template <> struct Elem<2> { using T = uint16_t; };
template <> struct Elem<4> { using T = uint32_t; };
template <> struct Elem<8> { using T = uint64_t; };
typedef uint32_t U128 __attribute__((ext_vector_type(4)));
template <> struct Elem<16> { using T = U128; };

// Sizes of the device this library targets (gfx950 / MI355X).
constexpr int kWave = 64;

// Host-side error reporting in the reference's style: a message on stderr, `true` returned.
#define LM_CHECK_HIP(expr)                                                              \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess) {                                                             \
      fprintf(stderr, "[libmarshal] %s:%d %s: %s\n", __FILE__, __LINE__, #expr,         \
              hipGetErrorString(_e));                                                   \
      return true;                                                                      \
    }                                                                                   \
  } while (0)

// Wait until every outstanding vector-memory load of this wave has returned its data.
__device__ __forceinline__ void wait_vm_loads() {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// Visited-bit flags in global memory are only touched with device-scope (agent) atomics so
// that every XCD sees one coherent copy; the data moves themselves need no ordering beyond
// "a start element is read before its bit is tested" (see the cycle-following kernels).
__device__ __forceinline__ uint32_t flag_load(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t flag_or(uint32_t *p, uint32_t mask) {
  return __hip_atomic_fetch_or(p, mask, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// 16 bytes starting at byte offset s of `base` (any alignment): the one or two aligned 16-byte
// chunks covering them are loaded and shifted into place with v_alignbyte. When s is not a
// multiple of 16 the load covers bytes up to (s & ~15) + 32, which the caller keeps in bounds.
// Split form: window_raw issues the loads, window_join shifts (call it after other loads have
// been issued so their latencies overlap).
template <bool NT = false>
__device__ __forceinline__ void window_raw(const unsigned char *base, uint64_t s, U128 &lo,
                                           U128 &hi) {
  const uint64_t a = s & ~uint64_t(15);
  const U128 *pa = reinterpret_cast<const U128 *>(base + a);
  lo = NT ? __builtin_nontemporal_load(pa) : *pa;
  if (s & 15u) hi = NT ? __builtin_nontemporal_load(pa + 1) : pa[1];
}
__device__ __forceinline__ U128 window_join(uint64_t s, const U128 &lo, const U128 &hi) {
  const uint32_t o = uint32_t(s & 15u);
  if (o == 0) return lo;
  const uint32_t w[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
  const uint32_t q = o >> 2, b = o & 3u;
  uint32_t v[5];
#pragma unroll
  for (int i = 0; i < 5; i++)
    v[i] = q == 0 ? w[i] : q == 1 ? w[i + 1] : q == 2 ? w[i + 2] : w[i + 3];
  U128 r;
  r.x = __builtin_amdgcn_alignbyte(v[1], v[0], b);
  r.y = __builtin_amdgcn_alignbyte(v[2], v[1], b);
  r.z = __builtin_amdgcn_alignbyte(v[3], v[2], b);
  r.w = __builtin_amdgcn_alignbyte(v[4], v[3], b);
  return r;
}

// NT: the aligned chunks are read with non-temporal loads (data read once).
template <bool NT = false>
__device__ __forceinline__ U128 load_window16(const unsigned char *base, uint64_t s) {
  const uint64_t a = s & ~uint64_t(15);
  const uint32_t o = uint32_t(s & 15u);
  const U128 *pa = reinterpret_cast<const U128 *>(base + a);
  const U128 lo = NT ? __builtin_nontemporal_load(pa) : *pa;
  if (o == 0) return lo;
  const U128 hi = NT ? __builtin_nontemporal_load(pa + 1) : pa[1];
  const uint32_t w[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
  const uint32_t q = o >> 2, b = o & 3u;
  uint32_t v[5];
#pragma unroll
  for (int i = 0; i < 5; i++)
    v[i] = q == 0 ? w[i] : q == 1 ? w[i + 1] : q == 2 ? w[i + 2] : w[i + 3];
  U128 r;
  r.x = __builtin_amdgcn_alignbyte(v[1], v[0], b);
  r.y = __builtin_amdgcn_alignbyte(v[2], v[1], b);
  r.z = __builtin_amdgcn_alignbyte(v[3], v[2], b);
  r.w = __builtin_amdgcn_alignbyte(v[4], v[3], b);
  return r;
}

// Work lists of grouped kernels: item i owns units [prefix[i], prefix[i + 1]) (every item at
// least one), and right after prefix[n] a lookup table holds, for every kFindStep units, the
// item owning unit j * kFindStep (the last item past the end). find_item returns the item
// that owns unit q (q < prefix[n]): two table reads bound a binary search that is empty
// whenever the items are longer than kFindStep units.
constexpr uint32_t kFindLog2 = 6, kFindStep = 1u << kFindLog2;
__device__ __forceinline__ uint32_t find_item(const uint64_t *prefix, uint32_t n, uint64_t q) {
  const uint64_t *lut = prefix + n + 1;
  const uint64_t j = q >> kFindLog2;
  uint32_t lo = uint32_t(lut[j]), hi = uint32_t(lut[j + 1]) + 1;  // prefix[lo] <= q < prefix[hi]
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (prefix[mid] <= q) lo = mid;
    else hi = mid;
  }
  return lo;
}

}  // namespace lm
