// kernels_dev.h - device-side helpers shared by the single-matrix and batched kernels.
#pragma once

#include "kernels.h"

#include <algorithm>

namespace lm {

// A lane's access unit in the LDS-staged 010: a dword of 4/ES elements for 1- and 2-byte
// elements, one element otherwise.
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


// Destination of super-element slot s in an R x C grid: (s mod C)*R + floor(s/C), with the
// optional rotations: ((s mod C + rot_c) mod C)*R + (floor(s/C) + rot_r) mod R. Any such map
// is a permutation of the slots, so cycle following applies unchanged.
template <bool I64>
__device__ __forceinline__ uint64_t next_slot(uint64_t s, const Pt100Params &p) {
  if constexpr (I64) {
    uint64_t q = p.div_C64.div(s);
    uint64_t col = s - q * p.C;
    if (p.rot_r) {
      q += p.rot_r;
      if (q >= p.R) q -= p.R;
    }
    if (p.rot_c) {
      col += p.rot_c;
      if (col >= p.C) col -= p.C;
    }
    return col * p.R + q;
  } else {
    uint32_t s32 = uint32_t(s);
    uint32_t q = p.div_C32.div(s32);
    uint32_t col = s32 - q * uint32_t(p.C);
    if (p.rot_r) {
      q += p.rot_r;
      if (q >= uint32_t(p.R)) q -= uint32_t(p.R);
    }
    if (p.rot_c) {
      col += p.rot_c;
      if (col >= uint32_t(p.C)) col -= uint32_t(p.C);
    }
    return uint64_t(col * uint32_t(p.R) + q);
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
    if (off < pbytes) d[k] = __builtin_nontemporal_load(reinterpret_cast<const V *>(src + off));
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
    if (off < pbytes) __builtin_nontemporal_store(d[k], reinterpret_cast<V *>(dst + off));
  }
}


// ---- LDS-staged 010 ----
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

// LAYOUT (dword units only): 0 tile rows at pitch Bp, 1 the chunk in input order, 2 in
// output order (both unpadded; see plan_010_bs).
template <int ES, int G, int NT, int UPT, int LAYOUT>
__device__ __forceinline__ void bs_stage_in(const Bs010Params &p, uint64_t c,
                                            typename UnitOf<ES, G>::T (&v)[UPT],
                                            typename Elem<ES>::T *lds) {
  using E = typename Elem<ES>::T;
  const uint32_t tid = threadIdx.x;
  const uint32_t a = p.a, B = p.B, Bp = p.Bp, n = p.n;
  constexpr uint32_t RU = NT * UPT;  // units per round
  const uint64_t t0 = c * p.g;
  const uint32_t tiles = uint32_t(std::min<uint64_t>(p.g, p.A - t0));
  const uint32_t C = tiles * n;
  const uint32_t nU = C / G;
  const unsigned char *gbase = p.base + t0 * uint64_t(n) * ES;
  const uint32_t rounds = (nU + RU - 1) / RU;
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
        if constexpr (G > 1 && LAYOUT == 2) {
          // Output order, unpadded: element (t, i, j) at (t*B + j)*a + i; B is a multiple of
          // G, so the unit's G elements are columns j..j+G-1 of row i, a elements apart.
          E *col = lds + (t * B + j) * a + i;
#pragma unroll
          for (int e = 0; e < G; e++) col[e * a] = unit_get<ES, G>(v[k], e);
          j += G - 1;
        } else if constexpr (G > 1 && LAYOUT == 1) {
          // Unpadded rows: unit u is elements u*G .. u*G+G-1 of the chunk in LDS too, so
          // (t, i, j), which this form does not use, need not be kept in step.
          reinterpret_cast<typename UnitOf<ES, G>::T *>(lds)[u] = v[k];
        } else if (G > 1 && (p.aligned & 1u)) {
          // B is a multiple of G, so the unit's G elements share one LDS row. The rows sit
          // at an odd pitch, so the unit is not dword-aligned in LDS: the empty asm keeps the
          // compiler from fusing the element stores into one misaligned dword store.
          E *row = lds + (t * a + i) * Bp + j;
#pragma unroll
          for (int e = 0; e < G; e++) {
            row[e] = unit_get<ES, G>(v[k], e);
            asm volatile("" ::: "memory");
          }
          j += G - 1;
        } else {
#pragma unroll
          for (int e = 0; e < G; e++) {
            lds[(t * a + i) * Bp + j] = unit_get<ES, G>(v[k], e);
            if (e < G - 1) {
              if (++j == B) { j = 0; if (++i == a) { i = 0; ++t; } }
            }
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
      lds[(tt * a + ii) * Bp + jj] = reinterpret_cast<const E *>(gbase)[f2];
    }
  }
}

// Output position o = t*n + j*a + i of chunk c reads LDS (t, i, j); coalesced stores.
template <int ES, int G, int NT, int LAYOUT>
__device__ __forceinline__ void bs_gather_out(const Bs010Params &p, uint64_t c,
                                              const typename Elem<ES>::T *lds) {
  using E = typename Elem<ES>::T;
  using U = typename UnitOf<ES, G>::T;
  const uint32_t tid = threadIdx.x;
  const uint32_t a = p.a, B = p.B, Bp = p.Bp, n = p.n;
  const uint64_t t0 = c * p.g;
  const uint32_t tiles = uint32_t(std::min<uint64_t>(p.g, p.A - t0));
  const uint32_t C = tiles * n;
  const uint32_t nU = C / G;
  unsigned char *gbase = p.base + t0 * uint64_t(n) * ES;
  U *gu = reinterpret_cast<U *>(gbase);
  uint32_t o = tid * G;
  uint32_t t = p.div_n.div(o);
  uint32_t rr = o - t * n;
  uint32_t j = p.div_a.div(rr);
  uint32_t i = rr - j * a;
  for (uint32_t u = tid; u < nU; u += NT) {
    U w{};
    if constexpr (G > 1 && LAYOUT == 2) {
      w = reinterpret_cast<const U *>(lds)[u];  // output order: unit u is LDS unit u
    } else if (G > 1 && (p.aligned & 2u)) {
      // a is a multiple of G, so the unit's G elements are rows i..i+G-1 of column j.
      const E *col = lds + (t * a + i) * Bp + j;
#pragma unroll
      for (int e = 0; e < G; e++) unit_put<ES, G>(w, e, col[e * Bp]);
      i += G - 1;
    } else {
#pragma unroll
      for (int e = 0; e < G; e++) {
        unit_put<ES, G>(w, e, lds[(t * a + i) * Bp + j]);
        if (e < G - 1) {
          if (++i == a) { i = 0; if (++j == B) { j = 0; ++t; } }
        }
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

// ---- 010 cycle following, LDS bits ----
// One a x B tile by cycle following with the tile's bits in LDS (flags: sh words).
template <int ES, int NT>
__device__ __forceinline__ void pt010_tile(const Pt010Params &p, uint64_t t, uint32_t *lflags) {
  using E = typename Elem<ES>::T;
  const uint32_t tid = threadIdx.x;
  const uint32_t a = p.a, B = p.B, n = p.n, sh = p.sh;
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

// One a x B tile by cycle following, owned by one wave (no workgroup barrier), with the tile's
// bits in the wave's own LDS region (sh words). The wave sweeps the tile's slots 64 at a time;
// a wave64 ballot picks the lanes whose start slot is still unclaimed, and each of them walks
// its cycle with the claim protocol of pt010_tile, keeping the load of the slot after the one
// it is claiming in flight (a value is stored only after its own slot's claim is won, and only
// the winner of a slot stores to it, so a prefetched value is either the slot's original
// content or never used).
template <int ES>
__device__ __forceinline__ void pt010_wave_tile(const Pt010Params &p, uint64_t t,
                                                uint32_t *wflags) {
  using E = typename Elem<ES>::T;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t a = p.a, B = p.B, n = p.n, sh = p.sh;
  E *in = reinterpret_cast<E *>(p.base) + t * uint64_t(n);
  for (uint32_t w = lane; w < sh; w += 64)
    __hip_atomic_store(&wflags[w], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  auto nextp = [&](uint32_t s) {
    const uint32_t q = p.div_B.div(s);
    return (s - q * B) * a + q;
  };
  auto bit = [&](uint32_t s, uint32_t &word) {
    const uint32_t bi = p.div_sh.div(s);
    word = s - bi * sh;
    return 1u << bi;
  };
  const uint32_t last = n - 1;
  for (uint32_t s0 = 0; s0 < last; s0 += 64) {  // uniform over the wave
    const uint32_t s = s0 + lane;
    uint32_t word = 0, mask = 0, nxt = 0;
    bool walk = s < last;
    if (walk) {
      nxt = nextp(s);
      walk = nxt != s;
    }
    if (walk) {
      mask = bit(s, word);
      walk = !(__hip_atomic_load(&wflags[word], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) &
               mask);
    }
    if (!__ballot(walk)) continue;  // every start of this block is claimed: nothing to load
    E d1{}, d2{};
    if (walk) {
      d1 = in[s];
      d2 = in[nxt];
    }
    wait_vm_loads();  // the start value is read before its bit is tested
    if (walk)
      walk = !(__hip_atomic_load(&wflags[word], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT) &
               mask);
    while (walk) {
      const uint32_t nn = nextp(nxt);
      const E d3 = in[nn];  // in flight while nxt is claimed and stored
      uint32_t wj = 0;
      const uint32_t mj = bit(nxt, wj);
      const uint32_t old =
          __hip_atomic_fetch_or(&wflags[wj], mj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      if (old & mj) break;
      in[nxt] = d1;
      d1 = d2;
      d2 = d3;
      nxt = nn;
    }
  }
}

// ---- 100 / 0100 cycle following ----
// Walkers wb .. wb + 64/VW - 1 of a 100/0100 stage with global visited bits (one wave).
template <int UB, int UPL, bool I64>
__device__ __forceinline__ void pt100_walkers(const Pt100Params &p, uint64_t wb) {
  using V = typename Elem<UB>::T;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t vwl = p.vw_log2;
  const uint32_t vw = 1u << vwl;
  const uint32_t sub = lane & (vw - 1u);
  const uint32_t leader = lane & ~(vw - 1u);
  const bool is_leader = (sub == 0);
  const uint64_t RC = p.RC, S = p.S;
  {
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

// One instance x of a 0100 stage with its visited bits in LDS (whole workgroup).
template <int UB, int UPL, int NT>
__device__ __forceinline__ void pt0100_instance(const Pt100Params &p, uint64_t x,
                                                uint32_t *lflags) {
  using V = typename Elem<UB>::T;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t vwl = p.vw_log2;
  const uint32_t vw = 1u << vwl;
  const uint32_t sub = lane & (vw - 1u);
  const uint32_t leader = lane & ~(vw - 1u);
  const bool is_leader = (sub == 0);
  const uint32_t walkers = NT >> vwl;                 // walkers in the workgroup
  const uint32_t me = threadIdx.x >> vwl;              // this lane's walker
  const uint64_t RC = p.RC, S = p.S;
  const uint32_t starts = uint32_t(RC * p.nsub);      // start points per instance
  const uint32_t words = (starts + 31) / 32;

  {
    for (uint32_t i = threadIdx.x; i < words; i += NT) lflags[i] = 0;
    __syncthreads();
    unsigned char *xbase = p.base + x * RC * S;
    for (uint32_t wb = 0; wb < starts; wb += walkers) {
      const uint32_t w = wb + me;
      bool walking = w < starts;
      uint32_t s = 0, c = 0;
      uint64_t nxt = 0;
      if (walking) {
        s = uint32_t(p.div_nsub.div(w));
        c = w - s * uint32_t(p.nsub);
        nxt = next_slot<false>(s, p);
        if (nxt == s) walking = false;
      }
      unsigned char *ibase = xbase + uint64_t(c) * p.sub_bytes;
      const uint32_t pbytes = uint32_t(std::min<uint64_t>(p.sub_bytes, S - c * p.sub_bytes));
      uint32_t f = 0;
      if (walking && is_leader)
        f = __hip_atomic_load(&lflags[w >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      uint64_t bal = __ballot(walking && is_leader && ((f >> (w & 31)) & 1u));
      if ((bal >> leader) & 1u) walking = false;
      V d1[UPL], d2[UPL];
      if (walking) load_piece<UB, UPL>(d1, ibase + uint64_t(s) * S, sub, vw, pbytes);
      wait_vm_loads();
      if (walking && is_leader)
        f = __hip_atomic_load(&lflags[w >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      bal = __ballot(walking && is_leader && ((f >> (w & 31)) & 1u));
      if ((bal >> leader) & 1u) walking = false;
      while (walking) {
        load_piece<UB, UPL>(d2, ibase + nxt * S, sub, vw, pbytes);
        const uint32_t fi = uint32_t(nxt) * uint32_t(p.nsub) + c;
        const uint32_t mask = 1u << (fi & 31);
        uint32_t old = 0;
        if (is_leader)
          old = __hip_atomic_fetch_or(&lflags[fi >> 5], mask, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_WORKGROUP);
        uint64_t lost = __ballot(is_leader && (old & mask));
        if ((lost >> leader) & 1u) break;
        store_piece<UB, UPL>(ibase + nxt * S, d1, sub, vw, pbytes);
#pragma unroll
        for (int k = 0; k < UPL; k++) d1[k] = d2[k];
        nxt = next_slot<false>(nxt, p);
      }
    }
    __syncthreads();
  }
}

// ---- row / column decomposition passes ----
// Source column of output column jp in row r of pass R.
__device__ __forceinline__ uint32_t rc_row_src(const RcParams &p, uint32_t r, uint32_t rc,
                                               uint32_t rm, uint32_t jp) {
  uint32_t jpc = p.div_c.div(jp);
  uint32_t jm = jp - jpc * p.c;
  uint32_t q = jm <= rm ? rm - jm : rm + p.c - jm;
  uint32_t u = jm <= rm ? rc : (rc ? rc - 1 : p.a - 1);
  uint32_t umod = u - p.div_b.div(u) * p.b;
  uint32_t x = jpc >= umod ? jpc - umod : jpc + p.b - umod;
  uint64_t prod = uint64_t(x) * p.ainv;
  uint32_t s = uint32_t(prod - p.div_b64.div(prod) * p.b);
  return q * p.b + s;
}

// Source row of output (ip, jp) in pass C2.
__device__ __forceinline__ uint32_t rc_col2_src(const RcParams &p, uint32_t ip, uint32_t jp) {
  uint64_t D = uint64_t(ip) * p.N + jp;
  uint64_t jj = p.div_M64.div(D);
  uint32_t i = uint32_t(D - jj * p.M);
  uint32_t r = i + p.div_b.div(uint32_t(jj));
  return r >= p.M ? r - p.M : r;
}

// Source row of output (r, j) in pass C1.
__device__ __forceinline__ uint32_t rc_col1_src(const RcParams &p, uint32_t r, uint32_t j) {
  uint32_t k = p.div_b.div(j);  // < c <= M
  return r >= k ? r - k : r + p.M - k;
}

// ---- row pass: `rows` consecutive rows per workgroup iteration ----------------------------
template <int UB, int NT>
__device__ __forceinline__ void rc_row_block(const RcParams &p, uint64_t blk,
                                             typename Elem<UB>::T *lds) {
  using V = typename Elem<UB>::T;
  const uint32_t tid = threadIdx.x;
  const uint32_t r0 = uint32_t(blk) * p.rows;  // global row (instances back to back)
  const uint32_t rows = std::min<uint32_t>(p.rows, uint32_t(p.inst * p.M) - r0);
  const uint32_t upr = p.N * p.m;  // units per row
  const uint32_t total = rows * upr;
  V *g = reinterpret_cast<V *>(p.base) + uint64_t(r0) * upr;
  constexpr int K = 16;
  uint32_t v = tid;
  for (; v + (K - 1) * NT < total; v += K * NT) {
    V t[K];
#pragma unroll
    for (int k = 0; k < K; k++) t[k] = __builtin_nontemporal_load(g + v + k * NT);
#pragma unroll
    for (int k = 0; k < K; k++) lds[v + k * NT] = t[k];
  }
  for (; v < total; v += NT) lds[v] = __builtin_nontemporal_load(g + v);
  __syncthreads();
  for (uint32_t o = tid; o < total; o += NT) {
    uint32_t rr = p.div_upr.div(o);
    uint32_t rem = o - rr * upr;
    uint32_t jp = p.div_m.div(rem);
    uint32_t k = rem - jp * p.m;
    uint32_t gr = r0 + rr;
    uint32_t r = gr - p.div_M32.div(gr) * p.M;  // row within its instance
    uint32_t rc = p.div_c.div(r);
    uint32_t j = rc_row_src(p, r, rc, r - rc * p.c, jp);
    __builtin_nontemporal_store(lds[rr * upr + j * p.m + k], g + o);
  }
  __syncthreads();
}

// ---- column pass: a block of w columns over all M rows ------------------------------------
// cbt counts column blocks over all instances: instance cbt / ncolblk, block cbt % ncolblk.
template <int UB, int NT, bool SHUFFLE>
__device__ __forceinline__ void rc_col_block(const RcParams &p, uint64_t cbt,
                                             typename Elem<UB>::T *lds) {
  using V = typename Elem<UB>::T;
  const uint32_t tid = threadIdx.x;
  const uint64_t x = p.div_ncolblk.div(cbt);
  const uint32_t cb = uint32_t(cbt - x * p.ncolblk);
  const uint32_t J = cb * p.w;
  const uint32_t w = std::min<uint32_t>(p.w, p.N - J);
  const uint32_t seg = w * p.m;        // units per row segment
  const uint32_t pitch = p.w * p.m;    // LDS units per row
  const uint32_t upr = p.N * p.m;
  const uint32_t total = p.M * seg;
  V *g = reinterpret_cast<V *>(p.base) + x * uint64_t(p.M) * upr + uint64_t(J) * p.m;
  // stage in: unit v of the block is (row v / seg, unit v % seg)
  {
    FastDiv32 dseg = seg == p.w * p.m ? p.div_pitch : p.div_seg_last;
    constexpr int K = 16;
    uint32_t v = tid;
    for (; v + (K - 1) * NT < total; v += K * NT) {
      V t[K];
#pragma unroll
      for (int k = 0; k < K; k++) {
        uint32_t vv = v + k * NT;
        uint32_t row = dseg.div(vv), x = vv - row * seg;
        t[k] = __builtin_nontemporal_load(g + uint64_t(row) * upr + x);
      }
#pragma unroll
      for (int k = 0; k < K; k++) {
        uint32_t vv = v + k * NT;
        uint32_t row = dseg.div(vv), x = vv - row * seg;
        lds[row * pitch + x] = t[k];
      }
    }
    for (; v < total; v += NT) {
      uint32_t row = dseg.div(v), x = v - row * seg;
      lds[row * pitch + x] = __builtin_nontemporal_load(g + uint64_t(row) * upr + x);
    }
    __syncthreads();
    for (uint32_t o = tid; o < total; o += NT) {
      uint32_t ip = dseg.div(o), x = o - ip * seg;
      uint32_t jl = p.div_m.div(x);
      uint32_t jp = J + jl;
      uint32_t r = SHUFFLE ? rc_col2_src(p, ip, jp) : rc_col1_src(p, ip, jp);
      __builtin_nontemporal_store(lds[r * pitch + x], g + uint64_t(ip) * upr + x);
    }
    __syncthreads();
  }
}

}  // namespace lm
