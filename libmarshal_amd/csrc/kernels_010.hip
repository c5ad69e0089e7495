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

template <int ES, int G> struct UnitOf { using T = typename Elem<ES>::T; };
template <> struct UnitOf<1, 4> { using T = uint32_t; };
template <> struct UnitOf<2, 2> { using T = uint32_t; };

template <int ES, int G>
__device__ __forceinline__ typename Elem<ES>::T unit_get(const typename UnitOf<ES, G>::T &v,
                                                         int k) {
  if constexpr (G == 1) {
    return v;
  } else {
    return typename Elem<ES>::T(v >> (8 * ES * k));
  }
}

template <int ES, int G>
__device__ __forceinline__ void unit_put(typename UnitOf<ES, G>::T &v, int k,
                                         typename Elem<ES>::T e) {
  if constexpr (G == 1) {
    v = e;
  } else {
    v |= uint32_t(e) << (8 * ES * k);
  }
}

// Units of chunk `c` owned by this thread in round `r` are u = r*NT*UPT + tid + k*NT.
template <int ES, int G, int NT, int UPT>
__device__ __forceinline__ void bs_load_round(const Bs010Params &p, uint64_t c, uint32_t r,
                                              typename UnitOf<ES, G>::T (&v)[UPT]) {
  using U = typename UnitOf<ES, G>::T;
  const uint64_t t0 = c * p.g;
  const uint32_t tiles = uint32_t(std::min<uint64_t>(p.g, p.A - t0));
  const uint32_t nU = tiles * p.n / G;
  const U *gu = reinterpret_cast<const U *>(p.base + t0 * uint64_t(p.n) * ES);
#pragma unroll
  for (int k = 0; k < UPT; k++) {
    uint32_t u = r * (NT * UPT) + threadIdx.x + k * NT;
    if (u < nU) v[k] = __builtin_nontemporal_load(gu + u);
  }
}

template <int ES, int G, int NT, int UPT>
__global__ __launch_bounds__(NT) void k_010_bs(Bs010Params p) {
  using E = typename Elem<ES>::T;
  using U = typename UnitOf<ES, G>::T;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  E *lds = reinterpret_cast<E *>(lds_raw);
  const uint32_t tid = threadIdx.x;
  const uint32_t a = p.a, B = p.B, Bp = p.Bp, n = p.n;
  constexpr uint32_t RU = NT * UPT;  // units per round

  uint64_t c = blockIdx.x;
  if (c >= p.nchunks) return;
  U v[UPT];
  bs_load_round<ES, G, NT, UPT>(p, c, 0, v);

  while (true) {
    const uint64_t t0 = c * p.g;
    const uint32_t tiles = uint32_t(std::min<uint64_t>(p.g, p.A - t0));
    const uint32_t C = tiles * n;
    const uint32_t nU = C / G;
    unsigned char *gbase = p.base + t0 * uint64_t(n) * ES;
    U *gu = reinterpret_cast<U *>(gbase);
    const uint32_t rounds = (nU + RU - 1) / RU;

    // ---- stage in: registers -> padded LDS rows (round 0 was loaded ahead of time)
    for (uint32_t r = 0; r < rounds; r++) {
      if (r > 0) bs_load_round<ES, G, NT, UPT>(p, c, r, v);
      uint32_t f = (r * RU + tid) * G;
      uint32_t t = p.div_n.div(f);
      uint32_t rr = f - t * n;
      uint32_t i = p.div_B.div(rr);
      uint32_t j = rr - i * B;
#pragma unroll
      for (int k = 0; k < UPT; k++) {
        uint32_t u = r * RU + tid + k * NT;
        if (u < nU) {
#pragma unroll
          for (int e = 0; e < G; e++) {
            lds[(t * a + i) * Bp + j] = unit_get<ES, G>(v[k], e);
            if (e < G - 1) {
              if (++j == B) { j = 0; if (++i == a) { i = 0; ++t; } }
            }
          }
        }
        j += p.din_j; if (j >= B) { j -= B; ++i; }
        i += p.din_i; if (i >= a) { i -= a; ++t; }
        t += p.din_t;
      }
    }
    if constexpr (G > 1) {
      for (uint32_t f2 = nU * G + tid; f2 < C; f2 += NT) {
        uint32_t tt = p.div_n.div(f2), rr = f2 - tt * n;
        uint32_t ii = p.div_B.div(rr), jj = rr - ii * B;
        lds[(tt * a + ii) * Bp + jj] = reinterpret_cast<E *>(gbase)[f2];
      }
    }
    __syncthreads();

    // ---- prefetch the next chunk (a disjoint region) while this one is written back
    const uint64_t cn = c + gridDim.x;
    if (cn < p.nchunks) bs_load_round<ES, G, NT, UPT>(p, cn, 0, v);

    // ---- gather out: output position o = t*n + j*a + i reads LDS (t, i, j)
    {
      uint32_t o = tid * G;
      uint32_t t = p.div_n.div(o);
      uint32_t rr = o - t * n;
      uint32_t j = p.div_a.div(rr);
      uint32_t i = rr - j * a;
      for (uint32_t u = tid; u < nU; u += NT) {
        U w{};
#pragma unroll
        for (int e = 0; e < G; e++) {
          unit_put<ES, G>(w, e, lds[(t * a + i) * Bp + j]);
          if (e < G - 1) {
            if (++i == a) { i = 0; if (++j == B) { j = 0; ++t; } }
          }
        }
        __builtin_nontemporal_store(w, gu + u);
        i += p.dout_i; if (i >= a) { i -= a; ++j; }
        j += p.dout_j; if (j >= B) { j -= B; ++t; }
        t += p.dout_t;
      }
      if constexpr (G > 1) {
        for (uint32_t o2 = nU * G + tid; o2 < C; o2 += NT) {
          uint32_t tt = p.div_n.div(o2), rr2 = o2 - tt * n;
          uint32_t jj = p.div_a.div(rr2), ii = rr2 - jj * a;
          reinterpret_cast<E *>(gbase)[o2] = lds[(tt * a + ii) * Bp + jj];
        }
      }
    }
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
  using E = typename Elem<ES>::T;
  extern __shared__ __align__(16) uint32_t lflags[];
  const uint32_t tid = threadIdx.x;
  const uint32_t a = p.a, B = p.B, n = p.n, sh = p.sh;
  for (uint64_t t = blockIdx.x; t < p.A; t += gridDim.x) {
    E *in = reinterpret_cast<E *>(p.base) + t * uint64_t(n);
    for (uint32_t w = tid; w < sh; w += NT) lflags[w] = 0;
    __syncthreads();
    const uint32_t last = n - 1;
    for (uint32_t s = tid; s < last; s += NT) {
      uint32_t q = p.div_B.div(s);
      uint32_t nxt = (s - q * B) * a + q;
      if (nxt == s) continue;
      uint32_t bi = p.div_sh.div(s);
      uint32_t word = s - bi * sh, mask = 1u << bi;
      if (__hip_atomic_load(&lflags[word], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) &
          mask)
        continue;
      E d1 = in[s];
      wait_vm_loads();
      if (__hip_atomic_load(&lflags[word], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) &
          mask)
        continue;
      while (true) {
        E d2 = in[nxt];
        uint32_t bj = p.div_sh.div(nxt);
        uint32_t wj = nxt - bj * sh, mj = 1u << bj;
        uint32_t old = __hip_atomic_fetch_or(&lflags[wj], mj, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP);
        if (old & mj) break;
        in[nxt] = d1;
        d1 = d2;
        q = p.div_B.div(nxt);
        nxt = (nxt - q * B) * a + q;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------------------
namespace {

template <int ES, int G, int NT, int UPT>
hipError_t launch_bs_t(const Bs010Params &p, size_t lds_bytes, uint32_t blocks,
                       hipStream_t st) {
  auto k = k_010_bs<ES, G, NT, UPT>;
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
  out->lds_bytes = size_t(g * tile_lds);
  out->nchunks = (A + g - 1) / g;
  return true;
}

namespace {
template <int ES, int G>
hipError_t launch_bs_g(const Bs010Params &p, const Bs010Plan &pl, uint32_t blocks,
                       hipStream_t st) {
  switch (pl.upt) {
    case 4: return launch_bs_t<ES, G, 256, 4>(p, pl.lds_bytes, blocks, st);
    case 8: return launch_bs_t<ES, G, 256, 8>(p, pl.lds_bytes, blocks, st);
    default: return launch_bs_t<ES, G, 256, 16>(p, pl.lds_bytes, blocks, st);
  }
}
}  // namespace

hipError_t launch_010_bs(unsigned char *base, size_t es, uint64_t A, uint64_t a, uint64_t B,
                         const Bs010Plan &pl, uint32_t max_blocks, hipStream_t st) {
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
  uint32_t blocks = uint32_t(std::min<uint64_t>(pl.nchunks, max_blocks));
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
