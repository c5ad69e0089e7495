// kernels_st.hip - stripe route: full M x N transposes whose rows and columns are too long
// for LDS, with a bounded scratch buffer.
//
// Row-stripe form (M = q*h + r): the top q*h rows are q stripes of h x N. Each stripe is
// copied to scratch and transposed out of place back into its own place ([q][h][N] ->
// [q][N][h]); one cycle-following 100 over the q x N grid of h-element super-elements then
// gives [N][q][h] = the transpose of the top part. When r > 0 the bottom r x N rows are
// first transposed into scratch, the top's N rows of q*h elements are padded in place to M,
// and the saved N x r block fills the pad columns.
//
// Column-stripe form (N = q*h + r): the mirror image. When r > 0 the last r columns are
// saved to scratch and the rows unpadded from N to q*h; 100 over the M x q grid of h-element
// super-elements gives [q][M][h]; each M x h block is transposed out of place through
// scratch into h x M; the saved M x r block, transposed, fills the last r rows.
//
// Every element is moved a bounded number of times by streaming kernels (copy, tiled
// out-of-place transpose, padding) and once by cycle following over large super-elements,
// so prime or otherwise awkward dimensions need no element-granular cycle following. The
// scratch buffer holds one stripe or a group of stripes (up to a cap of a few GB), not a
// second copy of the matrix. Not a reference algorithm: the reference transposes such
// matrices only by element-granular cycle following (src/cl_marshal.cc:769-775); its padding
// kernels (src/cl/cl_aos_asta.cl:813-938) are the padding/unpadding used here.
#include "kernels.h"

#include <algorithm>
#include <type_traits>

namespace lm {

// ---- tiled out-of-place transpose of `blocks` R x C matrices ----------------------------
// A tile of TR x TC source elements is staged transposed in LDS: row c of the LDS tile holds
// source column c (TR elements), so a destination row segment is one contiguous LDS row and
// is read back with 16-byte LDS reads. LDS rows have a pitch that is a multiple of 128 bytes
// and byte offsets inside a row are XOR-swizzled with a per-row mask taken from c >> SW: the
// 16-byte chunk index with bits 0-2, the 4-byte word inside the chunk with bits 3-4 (the
// 8-byte half for 8-byte elements). SW is 0 when lanes load single elements (consecutive c)
// and log2(16 / ES) when each lane loads 16 bytes (c steps by 16 / ES): either way the lanes
// of one LDS write hit distinct banks, and 16-byte reads along a row stay conflict-free.
// Source rows are read as 16-byte windows at any byte alignment (two aligned loads and a
// byte shift), so rows of an odd number of bytes still load one vector per lane.

// Tile row and column of the rem-th tile of a block. order 0: row-major over tiles (tiles
// along a source row come one after another). order k >= 1: the tile columns form groups of
// k; a group is walked row by row, k tiles along each row, so the tiles in flight together
// cover whole destination rows (all tile rows of a few tile columns) while each source row
// is still read in runs of k tiles.
__device__ __forceinline__ void oop_tile_rc(const OopParams &p, uint64_t rem, uint64_t &tr,
                                            uint64_t &tc) {
  if (p.order == 0) {
    tr = rem / p.ntc;
    tc = rem - tr * p.ntc;
    return;
  }
  const uint64_t k = p.order;
  const uint64_t grp = rem / (k * p.ntr);
  const uint64_t w = std::min<uint64_t>(k, p.ntc - grp * k);  // width of this group
  const uint64_t in = rem - grp * k * p.ntr;
  tr = in / w;
  tc = grp * k + (in - tr * w);
}

template <int ES, int SW>
__device__ __forceinline__ uint32_t oop_mask(uint32_t c) {
  const uint32_t cs = c >> SW;
  const uint32_t hi = (cs & 7u) << 4;
  if (ES <= 4) return hi | (((cs >> 3) & 3u) << 2);
  if (ES == 8) return hi | (((cs >> 3) & 1u) << 3);
  return hi;
}

// Stores bytes [s, s + n) of a 16-byte value to p + s (n = 1, 2, 4 or 8; s a multiple of n).
__device__ __forceinline__ void store_piece(unsigned char *p, const U128 &v, uint32_t s,
                                            uint32_t n) {
  const uint32_t dw = s >> 2;
  const uint32_t lo = dw == 0 ? v.x : dw == 1 ? v.y : dw == 2 ? v.z : v.w;
  if (n == 8) {
    const uint32_t hi = dw == 0 ? v.y : v.w;
    *reinterpret_cast<uint64_t *>(p + s) = (uint64_t(hi) << 32) | lo;
  } else if (n == 4) {
    *reinterpret_cast<uint32_t *>(p + s) = lo;
  } else if (n == 2) {
    *reinterpret_cast<uint16_t *>(p + s) = uint16_t(lo >> ((s & 2u) * 8));
  } else {
    p[s] = uint8_t(lo >> ((s & 3u) * 8));
  }
}

// Destination chunks that hold a row boundary, for the skewed form: with rows of R * ES
// bytes back to back over all blocks, the aligned chunk holding the start of row k (when
// misaligned; its first bytes end row k - 1) and, for k = fix_rows, the one holding the end
// of the last row. Each is built from the source (out of place, so still intact) and
// written once: whole when it lies inside the destination range, else only its bytes inside.
template <int ES>
__device__ __forceinline__ void oop_fix(const OopParams &p, uint64_t k) {
  if (k > p.fix_rows) return;
  const uint64_t RB = p.R * ES, total = p.fix_rows * RB, at = k * RB;
  const uint32_t mis = uint32_t((reinterpret_cast<uintptr_t>(p.dst) + at) & 15u);
  if (mis == 0) return;
  const int64_t a = int64_t(at) - mis;  // chunk start, relative to dst
  const unsigned char *src = p.src + p.src_off;
  // source of byte x of destination row kk (block kk / C, source column kk mod C)
  auto row_src = [&](uint64_t kk) {
    const uint64_t g = p.div_C.div(kk), cc = kk - g * p.C;
    return src + g * p.src_stride + cc * ES;
  };
  const unsigned char *prev = k ? row_src(k - 1) : nullptr;         // row k - 1
  const unsigned char *cur = k < p.fix_rows ? row_src(k) : nullptr;  // row k
  union {
    U128 v;
    uint8_t b[16];
  } w;
  w.v = U128{0, 0, 0, 0};
#pragma unroll
  for (uint32_t i = 0; i < 16; i++) {
    const bool first = i < mis;
    const unsigned char *rs = first ? prev : cur;
    const uint64_t x = first ? RB - mis + i : i - mis;  // byte within that row
    if (rs) w.b[i] = rs[(x / ES) * p.src_pitch + (x % ES)];
  }
  if (a >= 0 && uint64_t(a) + 16 <= total) {
    *reinterpret_cast<U128 *>(p.dst + a) = w.v;
  } else {
#pragma unroll
    for (uint32_t i = 0; i < 16; i++) {
      const int64_t q = a + int64_t(i);
      if (q >= 0 && uint64_t(q) < total) p.dst[q] = w.b[i];
    }
  }
}

// OUT: 1 = destination rows 16-byte aligned (vector stores); 0 = not aligned (window stores
// for 1-/2-byte elements or when wstore is set, else element stores); 2 = not aligned, skewed
// tiles: the tile covers, in every destination row, the aligned 16-byte chunks from its first
// row on, so it loads 16 / ES - 1 source rows above its own and writes whole chunks only; the
// chunks holding a row boundary are written by the launch's extra workgroups (oop_fix).
// Window loads of the 16-byte-per-lane load path, split so that a lane issues the loads of
// KB windows before it waits for any of them: slot k covers tile window i = i0 + k * NT
// (tile row i >> vl, 16-byte column i & (2^vl - 1)).
template <int KB>
struct OopWin {
  U128 lo[KB], hi[KB];
  uint64_t sb[KB];
  uint32_t how[KB];  // 0: nothing, 1: window loaded, 2: element loads (end of source)
};

// The loaded registers get no value on the paths that skip a load (a zero written there would
// have to wait for the loads in flight); they are read only where the load ran.
template <int ES, uint32_t KOFS, int NT, int KB>
__device__ __forceinline__ void oop_issue(const OopParams &p, OopWin<KB> &w, uint32_t i0,
                                          uint32_t vl, uint32_t total, uint64_t r0,
                                          uint64_t c0, uint32_t nr, uint32_t nc,
                                          uint64_t sblk) {
  constexpr uint32_t EPV = 16 / ES;
  const uint32_t vpr = 1u << vl;
  const bool ntl = p.nt & 1u;
#pragma unroll
  for (uint32_t k = 0; k < KB; k++) {
    const uint32_t i = i0 + k * NT;
    const uint32_t r = i >> vl, v = i & (vpr - 1u);
    const int64_t rg = int64_t(r0) + r - KOFS;
    const bool use =
        i < total && rg >= 0 && rg < int64_t(p.R) && (KOFS || r < nr) && v * EPV < nc;
    const uint64_t sb = sblk + uint64_t(rg) * p.src_pitch + (c0 + v * EPV) * ES;
    const uint64_t need = (sb & 15u) ? (sb & ~uint64_t(15)) + 32 : sb + 16;
    w.sb[k] = sb;
    w.how[k] = !use ? 0u : need <= p.src_end ? 1u : 2u;
    if (w.how[k] == 1) {
      const U128 *pa = reinterpret_cast<const U128 *>(p.src + (sb & ~uint64_t(15)));
      w.lo[k] = ntl ? __builtin_nontemporal_load(pa) : *pa;
      if (sb & 15u) w.hi[k] = ntl ? __builtin_nontemporal_load(pa + 1) : pa[1];
    }
  }
}

// Source row r, columns v * EPV .. (of the tile) -> LDS rows c, element r (swizzled).
template <int ES, int SW, int NT, int KB>
__device__ __forceinline__ void oop_fill(const OopParams &p, unsigned char *lds,
                                         const OopWin<KB> &w, uint32_t i0, uint32_t vl,
                                         uint32_t nc) {
  using E = typename Elem<ES>::T;
  constexpr uint32_t EPV = 16 / ES;
  const uint32_t vpr = 1u << vl;
  const uint32_t PB = p.PB;
#pragma unroll
  for (uint32_t k = 0; k < KB; k++) {
    if (w.how[k] == 0) continue;
    const uint32_t i = i0 + k * NT;
    const uint32_t r = i >> vl, v = i & (vpr - 1u);
    const uint64_t sb = w.sb[k];
    union {
      U128 x;
      E e[EPV];
    } u;
    if (w.how[k] == 1) {
      u.x = window_join(sb, w.lo[k], w.hi[k]);
    } else {  // window would run past the readable end: whole elements only
      u.x = U128{0, 0, 0, 0};
#pragma unroll
      for (uint32_t e = 0; e < EPV; e++)
        if (v * EPV + e < nc) u.e[e] = *reinterpret_cast<const E *>(p.src + sb + e * ES);
    }
#pragma unroll
    for (uint32_t e = 0; e < EPV; e++) {
      const uint32_t c = v * EPV + e;  // c < TC; columns >= nc are never stored
      *reinterpret_cast<E *>(lds + c * PB + ((r * ES) ^ oop_mask<ES, SW>(c))) = u.e[e];
    }
  }
}

// Writes a staged tile (source rows r0 .. r0 + nr, columns c0 .. c0 + nc of block `dst`'s
// source) transposed: destination row c0 + c, elements r0 .. r0 + nr.
template <int ES, int SW, int OUT, int NT>
__device__ __forceinline__ void oop_store(const OopParams &p, const unsigned char *lds,
                                          unsigned char *dst, uint64_t r0, uint64_t c0,
                                          uint32_t nr, uint32_t nc) {
  using E = typename Elem<ES>::T;
  constexpr uint32_t EPV = 16 / ES;
  constexpr uint32_t EPV_LOG2 = ES == 1 ? 4 : ES == 2 ? 3 : ES == 4 ? 2 : ES == 8 ? 1 : 0;
  constexpr bool VOUT = OUT == 1;
  (void)EPV;
  const uint32_t TR = 1u << p.tr_log2, TC = 1u << p.tc_log2;
  const uint32_t PB = p.PB;
  (void)TR;
  // Store: destination row c0 + c holds source column c0 + c; its segment starts at r0.
  if (VOUT) {  // destination rows and tile rows are 16-byte aligned
    const uint32_t vl = p.tr_log2 - EPV_LOG2;
    const uint32_t vpr = 1u << vl;
    for (uint32_t i = threadIdx.x; i < (TC << vl); i += NT) {
      const uint32_t c = i >> vl, j = i & (vpr - 1u);
      if (c < nc && j * EPV < nr) {
        const uint32_t m = oop_mask<ES, SW>(c);
        const U128 y = *reinterpret_cast<const U128 *>(lds + c * PB + ((j << 4) ^ (m & ~15u)));
        // Undo the word swizzle inside the chunk: logical word w sits at w ^ s.
        const uint32_t s = (m >> 2) & 3u;
        U128 x;
        x.x = (s == 0) ? y.x : (s == 1) ? y.y : (s == 2) ? y.z : y.w;
        x.y = (s == 0) ? y.y : (s == 1) ? y.x : (s == 2) ? y.w : y.z;
        x.z = (s == 0) ? y.z : (s == 1) ? y.w : (s == 2) ? y.x : y.y;
        x.w = (s == 0) ? y.w : (s == 1) ? y.z : (s == 2) ? y.y : y.x;
        U128 *q = reinterpret_cast<U128 *>(dst + (c0 + c) * p.dst_pitch + r0 * ES + (j << 4));
        if (p.nt & 2u)
          __builtin_nontemporal_store(x, q);
        else
          *q = x;
      }
    }
  } else if (OUT == 2) {
    // Chunk j of destination row c starts at row byte x0 = r0 * ES - m + 16 j (m: the row's
    // misalignment), which is LDS row byte 16 - m + 16 j.
    const uint32_t vl = p.tr_log2 - EPV_LOG2;  // TR >= EPV
    const uint32_t vpr = 1u << vl;
    const int64_t rowbytes = int64_t(p.R) * ES;
    for (uint32_t i = threadIdx.x; i < (TC << vl); i += NT) {
      const uint32_t c = i >> vl, j = i & (vpr - 1u);
      if (c >= nc) continue;
      unsigned char *row = dst + (c0 + c) * p.dst_pitch;
      const uint32_t m = uint32_t(reinterpret_cast<uintptr_t>(row) & 15u);
      const int64_t x0 = int64_t(r0) * ES - m + 16 * int64_t(j);
      if (x0 < 0 || x0 + 16 > rowbytes) continue;  // holds a row boundary
      const uint32_t mk = oop_mask<ES, SW>(c);
      const uint32_t s = (mk >> 2) & 3u;
      const unsigned char *lrow = lds + c * PB;
      auto chunk = [&](uint32_t q) {  // logical 16-byte chunk q of this LDS row
        const U128 y = *reinterpret_cast<const U128 *>(lrow + ((q << 4) ^ (mk & ~15u)));
        U128 x;
        x.x = (s == 0) ? y.x : (s == 1) ? y.y : (s == 2) ? y.z : y.w;
        x.y = (s == 0) ? y.y : (s == 1) ? y.x : (s == 2) ? y.w : y.z;
        x.z = (s == 0) ? y.z : (s == 1) ? y.w : (s == 2) ? y.x : y.y;
        x.w = (s == 0) ? y.w : (s == 1) ? y.z : (s == 2) ? y.y : y.x;
        return x;
      };
      const uint32_t o = 16u - m + 16u * j;
      const U128 lo = chunk(o >> 4);
      const U128 hi = (o & 15u) ? chunk((o >> 4) + 1) : lo;
      if (p.nt & 2u)
        __builtin_nontemporal_store(window_join(o, lo, hi), reinterpret_cast<U128 *>(row + x0));
      else
        *reinterpret_cast<U128 *>(row + x0) = window_join(o, lo, hi);
    }
  } else if (ES <= 2 || p.wstore) {
    // Destination rows not 16-byte aligned: per row, the elements before the first aligned
    // address and after the last whole vector are stored one by one (lane k == nb of the
    // row); the vectors in between are aligned 16-byte stores, each shifted together from
    // the two LDS chunks it straddles (lanes k < nb).
    // A row of TR elements spans TR*ES/16 vectors: a misaligned one has at most one fewer
    // whole vector, so its head/tail lane still fits.
    const uint32_t vl = p.tr_log2 > EPV_LOG2 ? p.tr_log2 - EPV_LOG2 : 0;
    const uint32_t vpr = 1u << vl;
    for (uint32_t i = threadIdx.x; i < (TC << vl); i += NT) {
      const uint32_t c = i >> vl, k = i & (vpr - 1u);
      if (c >= nc) continue;
      unsigned char *row = dst + (c0 + c) * p.dst_pitch + r0 * ES;
      const uint32_t L = nr * ES;                                         // bytes in the row
      const uint32_t head = uint32_t((16u - (reinterpret_cast<uintptr_t>(row) & 15u)) & 15u);
      const uint32_t nb = L > head ? (L - head) / 16 : 0;                 // whole vectors
      const uint32_t m = oop_mask<ES, SW>(c);
      const uint32_t s = (m >> 2) & 3u;
      const unsigned char *lrow = lds + c * PB;
      auto chunk = [&](uint32_t j) {  // logical 16-byte chunk j of this LDS row
        const U128 y = *reinterpret_cast<const U128 *>(lrow + ((j << 4) ^ (m & ~15u)));
        U128 x;
        x.x = (s == 0) ? y.x : (s == 1) ? y.y : (s == 2) ? y.z : y.w;
        x.y = (s == 0) ? y.y : (s == 1) ? y.x : (s == 2) ? y.w : y.z;
        x.z = (s == 0) ? y.z : (s == 1) ? y.w : (s == 2) ? y.x : y.y;
        x.w = (s == 0) ? y.w : (s == 1) ? y.z : (s == 2) ? y.y : y.x;
        return x;
      };
      if (k < nb) {
        const uint32_t o = head + 16 * k;  // row byte offset of this vector
        const U128 lo = chunk(o >> 4);
        U128 hi = lo;
        if (o & 15u) hi = chunk((o >> 4) + 1);  // o + 15 < L: still inside the row
        *reinterpret_cast<U128 *>(row + o) = window_join(o, lo, hi);
      } else if (k == nb) {  // head and tail elements one by one
        const uint32_t tail0 = head + 16 * nb;
        for (uint32_t b = 0; b < L; b += ES) {
          if (b >= head && b < tail0) b = tail0;
          if (b >= L) break;
          const uint32_t r = b / ES;
          *reinterpret_cast<E *>(row + b) =
              *reinterpret_cast<const E *>(lrow + ((r * ES) ^ m));
        }
      }
    }
  } else {  // 4- to 8-byte elements: element stores already fill whole lines per row
    for (uint32_t i = threadIdx.x; i < (TC << p.tr_log2); i += NT) {
      const uint32_t c = i >> p.tr_log2, r = i & (TR - 1u);
      if (c < nc && r < nr)
        *reinterpret_cast<E *>(dst + (c0 + c) * p.dst_pitch + (r0 + r) * ES) =
            *reinterpret_cast<const E *>(lds + c * PB + ((r * ES) ^ oop_mask<ES, SW>(c)));
    }
  }
}

// One workgroup's share of a launch: block bx of the launch's grid (tile or fix-up unit). The
// single-matrix kernel passes blockIdx.x; the grouped kernel of a batch passes the block's
// index within the matrix it serves.
template <int ES, bool VIN, int OUT, int NT>
__device__ __forceinline__ void oop_block(const OopParams &p, uint64_t bx, unsigned char *lds) {
  using E = typename Elem<ES>::T;
  constexpr uint32_t EPV = 16 / ES;  // elements per 16-byte vector
  constexpr uint32_t EPV_LOG2 = ES == 1 ? 4 : ES == 2 ? 3 : ES == 4 ? 2 : ES == 8 ? 1 : 0;
  constexpr int SW = VIN ? int(EPV_LOG2) : 0;
  constexpr uint32_t KOFS = OUT == 2 ? EPV : 0;  // LDS element index of source row r0
  if (OUT == 2 && bx >= p.per_xcd * 8) {
    oop_fix<ES>(p, (bx - p.per_xcd * 8) * uint64_t(NT) + threadIdx.x);
    return;
  }
  // Consecutive tiles go to the same XCD (workgroups are dealt round-robin over 8 XCDs), so
  // partial lines shared by neighbouring tiles meet in one L2.
  const uint64_t t = (bx & 7u) * p.per_xcd + (bx >> 3);
  if (t >= p.tiles) return;
  const uint64_t g = t / p.per_block;
  const uint64_t rem = t - g * p.per_block;
  uint64_t tr, tc;
  oop_tile_rc(p, rem, tr, tc);
  const uint32_t TR = 1u << p.tr_log2, TC = 1u << p.tc_log2;
  const uint64_t r0 = tr << p.tr_log2, c0 = tc << p.tc_log2;
  const uint32_t nr = uint32_t(std::min<uint64_t>(TR, p.R - r0));
  const uint32_t nc = uint32_t(std::min<uint64_t>(TC, p.C - c0));
  const uint64_t sblk = p.src_off + g * p.src_stride;  // offset of this block's row 0 in src
  unsigned char *dst = p.dst + g * p.dst_stride;
  const uint32_t PB = p.PB;

  // Load: source row r0 + r - KOFS, column c -> LDS row c, element r.
  const uint32_t TRL = TR + KOFS;  // LDS elements per row that are loaded
  if (ES == 1 && VIN && (TRL & 3u) == 0) {
    // 1-byte elements, four rows at a time: a lane loads the 16-byte windows of four
    // consecutive rows over the same 16 columns, transposes each 4 x 4 byte block in
    // registers and stores one dword per column (its four rows) to LDS: 16 dword stores per
    // 64 bytes instead of 64 byte stores. Rows outside the tile load a clamped row instead;
    // their LDS bytes are never stored to the destination as elements.
    const uint32_t vl = p.tc_log2 - 4;
    const uint32_t vpr = 1u << vl;
    for (uint32_t i = threadIdx.x; i < ((TRL >> 2) << vl); i += NT) {
      const uint32_t rb = (i >> vl) << 2, v = i & (vpr - 1u);
      if (v * 16 >= nc) continue;
      U128 lo[4], hi[4];
      uint32_t off[4];
#pragma unroll
      for (uint32_t k = 0; k < 4; k++) {
        int64_t rg = int64_t(r0) + int64_t(rb + k) - KOFS;
        rg = rg < 0 ? 0 : rg >= int64_t(p.R) ? int64_t(p.R) - 1 : rg;
        const uint64_t sb = sblk + uint64_t(rg) * p.src_pitch + c0 + v * 16;
        const uint64_t a = sb & ~uint64_t(15);
        // the upper chunk only when the window needs it and it starts before src_end (bytes
        // at or past src_end belong to no element of the tile)
        const uint64_t a2 = ((sb & 15u) && a + 16 < p.src_end) ? a + 16 : a;
        if (p.nt & 1u) {
          lo[k] = __builtin_nontemporal_load(reinterpret_cast<const U128 *>(p.src + a));
          hi[k] = __builtin_nontemporal_load(reinterpret_cast<const U128 *>(p.src + a2));
        } else {
          lo[k] = *reinterpret_cast<const U128 *>(p.src + a);
          hi[k] = *reinterpret_cast<const U128 *>(p.src + a2);
        }
        off[k] = uint32_t(sb & 15u);
      }
      uint32_t w[4][4];  // w[k][g]: bytes of columns 4g..4g+3 of row rb + k
#pragma unroll
      for (uint32_t k = 0; k < 4; k++) {
        const U128 x = window_join(off[k], lo[k], hi[k]);
        w[k][0] = x.x;
        w[k][1] = x.y;
        w[k][2] = x.z;
        w[k][3] = x.w;
      }
#pragma unroll
      for (uint32_t gq = 0; gq < 4; gq++) {
        // 4 x 4 byte transpose: t[j] holds column 4 gq + j of rows rb..rb+3, row rb lowest
        const uint32_t l01 = __builtin_amdgcn_perm(w[1][gq], w[0][gq], 0x05010400u);
        const uint32_t h01 = __builtin_amdgcn_perm(w[1][gq], w[0][gq], 0x07030602u);
        const uint32_t l23 = __builtin_amdgcn_perm(w[3][gq], w[2][gq], 0x05010400u);
        const uint32_t h23 = __builtin_amdgcn_perm(w[3][gq], w[2][gq], 0x07030602u);
        const uint32_t t[4] = {__builtin_amdgcn_perm(l23, l01, 0x05040100u),
                               __builtin_amdgcn_perm(l23, l01, 0x07060302u),
                               __builtin_amdgcn_perm(h23, h01, 0x05040100u),
                               __builtin_amdgcn_perm(h23, h01, 0x07060302u)};
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) {
          const uint32_t c = v * 16 + gq * 4 + j;  // c < TC; columns >= nc are never stored
          *reinterpret_cast<uint32_t *>(lds + c * PB + (rb ^ oop_mask<ES, SW>(c))) = t[j];
        }
      }
    }
  } else if (VIN) {  // one 16-byte window per lane (TC >= EPV)
    // Two windows per batch: with four the kernel needed 73-77 VGPRs (three 512-lane
    // workgroups per CU), with two 39-48 (four, the LDS limit for 32 KiB tiles). Measured:
    // f32 through-scratch transposes +3-7 % (2500 x 5100: 4.83 -> 5.19 TB/s), 3a's f64 stripes
    // unchanged, u8 within 1 %.
    constexpr uint32_t KB = 2;
    const uint32_t vl = p.tc_log2 - EPV_LOG2;
    const uint32_t total = TRL << vl;
    for (uint32_t i0 = threadIdx.x; i0 < total; i0 += KB * NT) {
      OopWin<KB> w;
      oop_issue<ES, KOFS, NT>(p, w, i0, vl, total, r0, c0, nr, nc, sblk);
      oop_fill<ES, SW, NT>(p, lds, w, i0, vl, nc);
    }
  } else {
    for (uint32_t i = threadIdx.x; i < (TRL << p.tc_log2); i += NT) {
      const uint32_t r = i >> p.tc_log2, c = i & (TC - 1u);
      const int64_t rg = int64_t(r0) + r - KOFS;
      if (rg >= 0 && rg < int64_t(p.R) && (KOFS || r < nr) && c < nc)
        *reinterpret_cast<E *>(lds + c * PB + ((r * ES) ^ oop_mask<ES, SW>(c))) =
            *reinterpret_cast<const E *>(p.src + sblk + uint64_t(rg) * p.src_pitch +
                                         (c0 + c) * ES);
    }
  }
  __syncthreads();
  oop_store<ES, SW, OUT, NT>(p, lds, dst, r0, c0, nr, nc);
}

template <int ES, bool VIN, int OUT, int NT>
__global__ __launch_bounds__(NT) void k_oop_transpose(OopParams p) {
  extern __shared__ __align__(16) unsigned char lds[];
  oop_block<ES, VIN, OUT, NT>(p, blockIdx.x, lds);
}

// ---- square in-place transpose: swap tile pairs ------------------------------------------
// For M == N the transpose is a set of independent swaps: tile (I, J) with its mirror (J, I).
// One workgroup loads both tiles of a pair (one for a diagonal tile) into LDS, transposed
// and swizzled as above, and writes each to the other's place: every element is read once
// and written once, in one pass. Workgroups of the lower triangle (I > J) exit at once.
template <int ES>
struct SqTile {
  static constexpr uint32_t T = ES == 1 ? 128 : ES <= 4 ? 64 : 32;  // tile side (elements)
  static constexpr uint32_t PB = T * ES < 128 ? 128 : T * ES;       // LDS row pitch (bytes)
};

template <int ES, bool VEC, int NT>
__device__ __forceinline__ void sq_load(unsigned char *lds, const unsigned char *base,
                                        uint64_t n, uint32_t r0, uint32_t c0, uint32_t nr,
                                        uint32_t nc) {
  using E = typename Elem<ES>::T;
  constexpr uint32_t T = SqTile<ES>::T, PB = SqTile<ES>::PB, EPV = 16 / ES;
  constexpr uint32_t EPV_LOG2 = ES == 1 ? 4 : ES == 2 ? 3 : ES == 4 ? 2 : ES == 8 ? 1 : 0;
  constexpr int SW = VEC ? int(EPV_LOG2) : 0;
  if (VEC) {  // rows and tile columns 16-byte aligned
    constexpr uint32_t VPR = T / EPV;
    for (uint32_t i = threadIdx.x; i < T * VPR; i += NT) {
      const uint32_t r = i / VPR, v = i % VPR;
      if (r < nr && v * EPV < nc) {
        union {
          U128 x;
          E e[EPV];
        } u;
        u.x = *reinterpret_cast<const U128 *>(base + ((uint64_t(r0) + r) * n + c0 + v * EPV) * ES);
#pragma unroll
        for (uint32_t k = 0; k < EPV; k++) {
          const uint32_t c = v * EPV + k;
          *reinterpret_cast<E *>(lds + c * PB + ((r * ES) ^ oop_mask<ES, SW>(c))) = u.e[k];
        }
      }
    }
  } else {
    for (uint32_t i = threadIdx.x; i < T * T; i += NT) {
      const uint32_t r = i / T, c = i % T;
      if (r < nr && c < nc)
        *reinterpret_cast<E *>(lds + c * PB + ((r * ES) ^ oop_mask<ES, SW>(c))) =
            *reinterpret_cast<const E *>(base + ((uint64_t(r0) + r) * n + c0 + c) * ES);
    }
  }
}

// Writes the staged tile (source rows r0.., columns c0..; nr x nc) transposed: destination
// row c0 + c, columns r0 .. r0 + nr.
template <int ES, bool VEC, int NT>
__device__ __forceinline__ void sq_store(const unsigned char *lds, unsigned char *base,
                                         uint64_t n, uint32_t r0, uint32_t c0, uint32_t nr,
                                         uint32_t nc) {
  using E = typename Elem<ES>::T;
  constexpr uint32_t T = SqTile<ES>::T, PB = SqTile<ES>::PB, EPV = 16 / ES;
  constexpr uint32_t EPV_LOG2 = ES == 1 ? 4 : ES == 2 ? 3 : ES == 4 ? 2 : ES == 8 ? 1 : 0;
  constexpr int SW = VEC ? int(EPV_LOG2) : 0;
  if (VEC) {
    constexpr uint32_t VPR = T / EPV;
    for (uint32_t i = threadIdx.x; i < T * VPR; i += NT) {
      const uint32_t c = i / VPR, j = i % VPR;
      if (c < nc && j * EPV < nr) {
        const uint32_t m = oop_mask<ES, SW>(c);
        const U128 y = *reinterpret_cast<const U128 *>(lds + c * PB + ((j << 4) ^ (m & ~15u)));
        const uint32_t s = (m >> 2) & 3u;
        U128 x;
        x.x = (s == 0) ? y.x : (s == 1) ? y.y : (s == 2) ? y.z : y.w;
        x.y = (s == 0) ? y.y : (s == 1) ? y.x : (s == 2) ? y.w : y.z;
        x.z = (s == 0) ? y.z : (s == 1) ? y.w : (s == 2) ? y.x : y.y;
        x.w = (s == 0) ? y.w : (s == 1) ? y.z : (s == 2) ? y.y : y.x;
        *reinterpret_cast<U128 *>(base + ((uint64_t(c0) + c) * n + r0 + j * EPV) * ES) = x;
      }
    }
  } else {
    for (uint32_t i = threadIdx.x; i < T * T; i += NT) {
      const uint32_t c = i / T, r = i % T;
      if (c < nc && r < nr)
        *reinterpret_cast<E *>(base + ((uint64_t(c0) + c) * n + r0 + r) * ES) =
            *reinterpret_cast<const E *>(lds + c * PB + ((r * ES) ^ oop_mask<ES, SW>(c)));
    }
  }
}

template <int ES, bool VEC, int NT>
__global__ __launch_bounds__(NT) void k_sq_swap(unsigned char *base, uint64_t n) {
  constexpr uint32_t T = SqTile<ES>::T, PB = SqTile<ES>::PB;
  __shared__ __align__(16) unsigned char lds[2 * T * PB];
  const uint32_t I = blockIdx.y, J = blockIdx.x;
  if (I > J) return;
  const uint32_t ri = I * T, rj = J * T;
  const uint32_t ni = uint32_t(std::min<uint64_t>(T, n - ri));
  const uint32_t nj = uint32_t(std::min<uint64_t>(T, n - rj));
  sq_load<ES, VEC, NT>(lds, base, n, ri, rj, ni, nj);                    // tile (I, J)
  if (I != J) sq_load<ES, VEC, NT>(lds + T * PB, base, n, rj, ri, nj, ni);  // tile (J, I)
  __syncthreads();
  sq_store<ES, VEC, NT>(lds, base, n, ri, rj, ni, nj);  // (I, J)^T -> place of (J, I)
  if (I != J) sq_store<ES, VEC, NT>(lds + T * PB, base, n, rj, ri, nj, ni);
}

hipError_t launch_sq_transpose(unsigned char *base, size_t es, uint64_t n, hipStream_t st) {
  if (n < 2) return hipSuccess;
  KernelSpan span(st, "k_sq_swap", 2.0 * double(n) * double(n) * double(es));
  auto go = [&](auto tag, auto vec) -> hipError_t {
    constexpr int E = decltype(tag)::value;
    constexpr bool V = decltype(vec)::value;
    constexpr uint32_t T = SqTile<E>::T;
    const uint64_t nt = (n + T - 1) / T;
    if (nt > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_sq_swap<E, V, 256>), dim3(uint32_t(nt), uint32_t(nt)), dim3(256), 0,
                       st, base, n);
    return hipGetLastError();
  };
  const bool vec = ((reinterpret_cast<uintptr_t>(base) | (n * es)) & 15u) == 0;
  using std::integral_constant;
  switch (es) {
    case 1: return vec ? go(integral_constant<int, 1>{}, std::true_type{})
                       : go(integral_constant<int, 1>{}, std::false_type{});
    case 2: return vec ? go(integral_constant<int, 2>{}, std::true_type{})
                       : go(integral_constant<int, 2>{}, std::false_type{});
    case 4: return vec ? go(integral_constant<int, 4>{}, std::true_type{})
                       : go(integral_constant<int, 4>{}, std::false_type{});
    case 8: return vec ? go(integral_constant<int, 8>{}, std::true_type{})
                       : go(integral_constant<int, 8>{}, std::false_type{});
    case 16: return go(integral_constant<int, 16>{}, std::true_type{});
  }
  return hipErrorInvalidValue;
}

bool sq_fits(size_t es, uint64_t n) {
  const uint64_t T = es == 1 ? 128 : es <= 4 ? 64 : 32;
  return n >= 2 && (n + T - 1) / T <= 65535;
}

// ---- byte copies ------------------------------------------------------------------------
// Streaming 16-byte copy: four vectors in flight per lane; NTL reads the source with
// non-temporal loads (measured 6.2 TB/s against 4.6 for one plain vector per lane on a cold
// 1 GiB buffer).
constexpr int kCopyUnroll = 4;
// NTL / NTS: non-temporal loads / stores.
template <bool NTL, bool NTS>
__global__ __launch_bounds__(256) void k_copy16(U128 *dst, const U128 *src, uint64_t n) {
  const uint64_t stride = uint64_t(gridDim.x) * 256 * kCopyUnroll;
  for (uint64_t b = uint64_t(blockIdx.x) * 256 * kCopyUnroll + threadIdx.x; b < n; b += stride) {
    U128 v[kCopyUnroll];
#pragma unroll
    for (int u = 0; u < kCopyUnroll; u++) {
      const uint64_t i = b + uint64_t(u) * 256;
      if (i < n) v[u] = NTL ? __builtin_nontemporal_load(src + i) : src[i];
    }
#pragma unroll
    for (int u = 0; u < kCopyUnroll; u++) {
      const uint64_t i = b + uint64_t(u) * 256;
      if (i < n) {
        if (NTS)
          __builtin_nontemporal_store(v[u], dst + i);
        else
          dst[i] = v[u];
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_copy1(unsigned char *dst, const unsigned char *src,
                                               uint64_t n) {
  uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
  const uint64_t stride = uint64_t(gridDim.x) * 256;
  for (; i < n; i += stride) dst[i] = src[i];
}

// rows x width bytes between pitched layouts.
__global__ __launch_bounds__(256) void k_copy2d(unsigned char *dst, uint64_t dpitch,
                                                const unsigned char *src, uint64_t spitch,
                                                uint64_t width, uint64_t rows,
                                                FastDiv64 div_w) {
  const uint64_t n = width * rows;
  uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
  const uint64_t stride = uint64_t(gridDim.x) * 256;
  for (; i < n; i += stride) {
    const uint64_t r = div_w.div(i), c = i - r * width;
    dst[r * dpitch + c] = src[r * spitch + c];
  }
}

// Copy between any two addresses: the destination in aligned 16-byte chunks (the source read
// as aligned vectors or, at another alignment, as 16-byte windows), the bytes before the
// destination's first aligned address and after its last whole chunk by thread 0.
template <bool SRC_ALIGNED, bool NTL, bool NTS>
__global__ __launch_bounds__(256) void k_copy_any(unsigned char *dst, const unsigned char *src,
                                                  uint64_t bytes, uint32_t head, uint64_t nb) {
  const uint64_t stride = uint64_t(gridDim.x) * 256;
  U128 *d = reinterpret_cast<U128 *>(dst + head);  // 16-byte aligned
  // source of chunk i: bytes from src + head + 16 i, as an offset from an aligned base
  const unsigned char *sbase =
      reinterpret_cast<const unsigned char *>(reinterpret_cast<uintptr_t>(src + head) &
                                              ~uintptr_t(15));
  const uint64_t soff = reinterpret_cast<uintptr_t>(src + head) & 15u;
  for (uint64_t b = uint64_t(blockIdx.x) * 256 * kCopyUnroll + threadIdx.x; b < nb;
       b += stride * kCopyUnroll) {
    U128 v[kCopyUnroll];
#pragma unroll
    for (int u = 0; u < kCopyUnroll; u++) {
      const uint64_t i = b + uint64_t(u) * 256;
      if (i < nb)
        v[u] = !SRC_ALIGNED ? load_window16(sbase, soff + 16 * i)
               : NTL ? __builtin_nontemporal_load(reinterpret_cast<const U128 *>(sbase) + i)
                     : reinterpret_cast<const U128 *>(sbase)[i];
    }
#pragma unroll
    for (int u = 0; u < kCopyUnroll; u++) {
      const uint64_t i = b + uint64_t(u) * 256;
      if (i < nb) {
        if (NTS)
          __builtin_nontemporal_store(v[u], d + i);
        else
          d[i] = v[u];
      }
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (uint32_t b = 0; b < head; b++) dst[b] = src[b];
    for (uint64_t b = head + 16 * nb; b < bytes; b++) dst[b] = src[b];
  }
}

hipError_t launch_copy(unsigned char *dst, const unsigned char *src, uint64_t bytes,
                       hipStream_t st) {
  if (bytes == 0) return hipSuccess;
  KernelSpan span(st, "k_copy", 2.0 * double(bytes));
  const uint32_t cap = uint32_t(device_info().cus) * 16;
  const uint32_t cap16 = uint32_t(device_info().cus) * uint32_t(tuning("copy_blocks_per_cu"));
  const uint64_t per_block = 256 * kCopyUnroll;
  // Stores: non-temporal from "copy_nt_store_mb" on. Smaller copies land in the 256 MB
  // last-level cache, where the pass that reads them back finds them (measured 4-5 % faster
  // for 52-168 MB through-scratch transposes; 10 % slower for a 403 MB one).
  const int64_t nts_mb = tuning("copy_nt_store_mb");
  const bool nts = bytes >= (uint64_t(nts_mb) << 20);
  const bool ntl = tuning("copy_nt") != 0;
  if (bytes >= 64 && ((uintptr_t(dst) | uintptr_t(src) | bytes) & 15u) != 0) {
    const uint32_t head = uint32_t((16u - (uintptr_t(dst) & 15u)) & 15u);
    const uint64_t nb = (bytes - head) / 16;
    const uint32_t blocks = uint32_t(
        std::min<uint64_t>(std::max<uint64_t>((nb + per_block - 1) / per_block, 1), cap16));
    auto go = [&](auto k) {
      hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, st, dst, src, bytes, head, nb);
    };
    if (((uintptr_t(src) + head) & 15u) != 0)
      nts ? go(k_copy_any<false, false, true>) : go(k_copy_any<false, false, false>);
    else if (ntl)
      nts ? go(k_copy_any<true, true, true>) : go(k_copy_any<true, true, false>);
    else
      nts ? go(k_copy_any<true, false, true>) : go(k_copy_any<true, false, false>);
    return hipGetLastError();
  }
  if (((uintptr_t(dst) | uintptr_t(src) | bytes) & 15u) == 0) {
    const uint64_t n = bytes / 16;
    const uint32_t blocks = uint32_t(std::min<uint64_t>((n + per_block - 1) / per_block, cap16));
    auto go = [&](auto k) {
      hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, st, reinterpret_cast<U128 *>(dst),
                         reinterpret_cast<const U128 *>(src), n);
    };
    if (ntl)
      nts ? go(k_copy16<true, true>) : go(k_copy16<true, false>);
    else
      nts ? go(k_copy16<false, true>) : go(k_copy16<false, false>);
  } else {
    const uint32_t blocks = uint32_t(std::min<uint64_t>((bytes + 255) / 256, cap));
    hipLaunchKernelGGL(k_copy1, dim3(blocks), dim3(256), 0, st, dst, src, bytes);
  }
  return hipGetLastError();
}

// Contiguous copy out of a pitched buffer: destination bytes [0, total), rows of RB bytes
// back to back, from source rows of RB bytes at pitch P (a multiple of 16; src 16-byte
// aligned). The destination is written in aligned 16-byte chunks. A chunk inside one row is a
// 16-byte window of that source row; a chunk that holds a row boundary (RB >= 16: at most
// one) takes its bytes past the boundary from a second window that ends the row's padding and
// starts the next row. The bytes before the destination's first aligned address and after
// its last whole chunk are copied one by one by thread 0.
__device__ __forceinline__ uint64_t unpitch_src(uint64_t d, uint64_t RB, uint64_t P,
                                                const FastDiv64 &div_rb) {
  const uint64_t row = div_rb.div(d);
  return row * P + (d - row * RB);
}

// Destination chunks [rd * 256 * kCopyUnroll, +256 * kCopyUnroll) (round rd).
template <bool NTS>
__device__ __forceinline__ void unpitch_round(unsigned char *dst, const unsigned char *src,
                                              uint64_t RB, uint64_t P, uint32_t head,
                                              uint64_t nb, const FastDiv64 &div_rb,
                                              uint64_t rd) {
  U128 *d16 = reinterpret_cast<U128 *>(dst + head);  // 16-byte aligned
  {
    const uint64_t b = rd * 256 * kCopyUnroll + threadIdx.x;
    U128 lo[kCopyUnroll], hi[kCopyUnroll];
    uint64_t s[kCopyUnroll];
    uint32_t split[kCopyUnroll];  // bytes of the chunk before a row boundary (16: none)
    // Phase 1: the window of each chunk's first row (branch-free; a lane past the end loads
    // the source's first chunk).
#pragma unroll
    for (int u = 0; u < kCopyUnroll; u++) {
      const uint64_t i = b + uint64_t(u) * 256;
      const uint64_t dd = head + 16 * (i < nb ? i : 0);
      const uint64_t row = div_rb.div(dd), off = dd - row * RB;
      s[u] = row * P + off;
      split[u] = uint32_t(off + 16 <= RB ? 16 : RB - off);
      const uint64_t a = s[u] & ~uint64_t(15);
      const U128 *pa = reinterpret_cast<const U128 *>(src + a);
      lo[u] = pa[0];
      hi[u] = (s[u] & 15u) ? pa[1] : pa[0];
    }
#pragma unroll
    for (int u = 0; u < kCopyUnroll; u++) {
      const uint64_t i = b + uint64_t(u) * 256;
      if (i >= nb) continue;
      U128 v = window_join(s[u], lo[u], hi[u]);
      const uint32_t sp = split[u];
      if (sp < 16) {  // row boundary inside the chunk: the rest comes from the next row
        const uint64_t s2 = s[u] - (s[u] % P) + P - sp;  // so that byte sp is next row's 0
        const U128 w = load_window16(src, s2);
        const uint32_t vw[4] = {v.x, v.y, v.z, v.w}, ww[4] = {w.x, w.y, w.z, w.w};
        uint32_t o[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
          const uint32_t lo_b = 4 * k;
          if (sp >= lo_b + 4) {
            o[k] = vw[k];
          } else if (sp <= lo_b) {
            o[k] = ww[k];
          } else {
            const uint32_t m = (1u << (8 * (sp - lo_b))) - 1u;
            o[k] = (vw[k] & m) | (ww[k] & ~m);
          }
        }
        v = U128{o[0], o[1], o[2], o[3]};
      }
      if (NTS)
        __builtin_nontemporal_store(v, d16 + i);
      else
        d16[i] = v;
    }
  }
}

// Bytes before the destination's first aligned address and after its last whole chunk.
__device__ __forceinline__ void unpitch_ends(unsigned char *dst, const unsigned char *src,
                                             uint64_t RB, uint64_t P, uint64_t total,
                                             uint32_t head, uint64_t nb,
                                             const FastDiv64 &div_rb) {
  for (uint64_t x = 0; x < head && x < total; x++) dst[x] = src[unpitch_src(x, RB, P, div_rb)];
  for (uint64_t x = head + 16 * nb; x < total; x++) dst[x] = src[unpitch_src(x, RB, P, div_rb)];
}

template <bool NTS>
__global__ __launch_bounds__(256) void k_unpitch(unsigned char *dst,
                                                 const unsigned char *src, uint64_t RB,
                                                 uint64_t P, uint64_t total, uint32_t head,
                                                 uint64_t nb, FastDiv64 div_rb) {
  const uint64_t rounds = (nb + 256 * kCopyUnroll - 1) / (256 * kCopyUnroll);
  for (uint64_t rd = blockIdx.x; rd < rounds; rd += gridDim.x)
    unpitch_round<NTS>(dst, src, RB, P, head, nb, div_rb, rd);
  if (blockIdx.x == 0 && threadIdx.x == 0) unpitch_ends(dst, src, RB, P, total, head, nb, div_rb);
}

hipError_t launch_unpitch(unsigned char *dst, const unsigned char *src, uint64_t RB,
                          uint64_t P, uint64_t rows, hipStream_t st, int nt_store) {
  const uint64_t total = RB * rows;
  if (total == 0) return hipSuccess;
  if (RB < 16 || P < RB || (P & 15u) || (reinterpret_cast<uintptr_t>(src) & 15u))
    return hipErrorInvalidValue;
  KernelSpan span(st, "k_unpitch", 2.0 * double(total));
  const uint32_t head =
      uint32_t(std::min<uint64_t>((16u - (reinterpret_cast<uintptr_t>(dst) & 15u)) & 15u,
                                  total));
  const uint64_t nb = (total - head) / 16;
  const uint64_t per_block = 256 * kCopyUnroll;
  const uint32_t blocks = uint32_t(std::min<uint64_t>(
      std::max<uint64_t>((nb + per_block - 1) / per_block, 1),
      uint64_t(device_info().cus) * uint64_t(tuning("copy_blocks_per_cu"))));
  if (nt_store < 0) nt_store = (tuning("oop_nt") & 2) ? 1 : 0;
  if (nt_store)
    hipLaunchKernelGGL(k_unpitch<true>, dim3(blocks), dim3(256), 0, st, dst, src, RB, P,
                       total, head, nb, FastDiv64(RB));
  else
    hipLaunchKernelGGL(k_unpitch<false>, dim3(blocks), dim3(256), 0, st, dst, src, RB, P,
                       total, head, nb, FastDiv64(RB));
  return hipGetLastError();
}

hipError_t launch_copy2d(unsigned char *dst, uint64_t dpitch, const unsigned char *src,
                         uint64_t spitch, uint64_t width, uint64_t rows, hipStream_t st) {
  if (width == 0 || rows == 0) return hipSuccess;
  const uint64_t n = width * rows;
  KernelSpan span(st, "k_copy2d", 2.0 * double(n));
  const uint32_t blocks =
      uint32_t(std::min<uint64_t>((n + 255) / 256, uint64_t(device_info().cus) * 16));
  hipLaunchKernelGGL(k_copy2d, dim3(blocks), dim3(256), 0, st, dst, dpitch, src, spitch, width,
                     rows, FastDiv64(width));
  return hipGetLastError();
}

// ---- host: out-of-place transpose launcher ----------------------------------------------
namespace {
uint32_t log2_floor(uint64_t v) {
  uint32_t l = 0;
  while ((uint64_t(2) << l) <= v) l++;
  return l;
}
uint32_t log2_ceil(uint64_t v) {
  uint32_t l = 0;
  while ((uint64_t(1) << l) < v) l++;
  return l;
}

constexpr int kOopThreads = 512;

template <int ES, bool VIN, int OUT>
hipError_t launch_oop_t(const OopParams &p, size_t lds, hipStream_t st) {
  auto k = k_oop_transpose<ES, VIN, OUT, kOopThreads>;
  const uint64_t fix_blocks = OUT == 2 ? (p.fix_rows + 1 + kOopThreads - 1) / kOopThreads : 0;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
  }
  if (p.per_xcd * 8 + fix_blocks >= (1ull << 32)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k, dim3(uint32_t(p.per_xcd * 8 + fix_blocks)), dim3(kOopThreads), lds, st,
                     p);
  return hipGetLastError();
}

template <int ES>
hipError_t launch_oop_es(const OopParams &p, bool vin, int out, size_t lds, hipStream_t st) {
  if (out == 2) {
    if (ES == 16) return hipErrorInvalidValue;  // 16-byte rows are always aligned
    return vin ? launch_oop_t<ES, true, 2>(p, lds, st) : launch_oop_t<ES, false, 2>(p, lds, st);
  }
  if (vin && out) return launch_oop_t<ES, true, 1>(p, lds, st);
  if (vin) return launch_oop_t<ES, true, 0>(p, lds, st);
  if (out) return launch_oop_t<ES, false, 1>(p, lds, st);
  return launch_oop_t<ES, false, 0>(p, lds, st);
}
}  // namespace

hipError_t prepare_oop_transpose(const unsigned char *src, unsigned char *dst, size_t es,
                                 uint64_t blocks, uint64_t R, uint64_t C, uint64_t src_stride,
                                 uint64_t dst_stride, size_t tile_bytes, uint64_t src_end,
                                 uint64_t dst_pitch, int nt, OopLaunch *o) {
  *o = OopLaunch{};
  if (blocks == 0 || R == 0 || C == 0) return hipSuccess;
  if (!(es == 1 || es == 2 || es == 4 || es == 8 || es == 16)) return hipErrorInvalidValue;
  if (dst_pitch == 0) dst_pitch = R * es;
  if (dst_pitch < R * es) return hipErrorInvalidValue;
  auto al = [](uint64_t v) { return (v & 15u) == 0; };
  const bool dst_al = al(reinterpret_cast<uintptr_t>(dst)) && al(dst_pitch) && al(dst_stride);
  const int64_t skew_knob = tuning("st_skew");
  const bool skew_es = es < 16 && skew_knob > 0 && (es <= 2 || skew_knob >= 2);
  // Tile: cover the shorter side whole when a tile of it stays within `tile_bytes` with at
  // least 64 bytes along the other side; otherwise a square-ish tile.
  const uint64_t min_other = std::max<uint64_t>(1, 64 / es);
  uint32_t trl, tcl;
  auto pick = [&](size_t bytes) {  // true when the tile covers the shorter side whole
    const uint64_t cap_el = std::max<uint64_t>(bytes / es, 256);
    if (R <= C && (uint64_t(1) << log2_ceil(R)) * min_other <= cap_el) {
      trl = log2_ceil(R);
      tcl = log2_floor(cap_el >> trl);
      return true;
    }
    if (C < R && (uint64_t(1) << log2_ceil(C)) * min_other <= cap_el) {
      tcl = log2_ceil(C);
      trl = log2_floor(cap_el >> tcl);
      return true;
    }
    const uint32_t l = log2_floor(cap_el);
    tcl = (l + 1) / 2;
    trl = l - tcl;
    return false;
  };
  // Element stores (misaligned destination rows, no skewed form for this size) into tiles
  // that span a whole short side (the stripes of a column-stripe route) run faster with
  // smaller tiles: on 131071 x 256 f64 stripes, 16 x 256 tiles (32 KiB) reach about
  // 4.2 TB/s and 32 x 32 tiles (8 KiB) about 4.9. Square-ish tiles over both long sides
  // lose from the same change, so they keep `tile_bytes`.
  if (pick(tile_bytes) && !dst_al && !skew_es) {
    const int64_t ekb = tuning("oop_elem_tile_kb");
    if (ekb > 0 && (size_t(ekb) << 10) < tile_bytes) {
      tile_bytes = size_t(ekb) << 10;
      pick(tile_bytes);
    }
  }
  // Never larger than the matrix side rounded up to a power of two; optionally fewer tile
  // rows (longer contiguous row segments per tile).
  trl = std::min(trl, log2_ceil(R));
  tcl = std::min(tcl, log2_ceil(C));
  // Elements of up to 4 bytes: at least "oop_min_rows" tile rows (longer destination row
  // segments per tile), keeping the area and one 16-byte vector per source row segment.
  // 256 measured 1-10 % faster on 1-, 2- and 4-byte through-scratch transposes (no change on
  // the stripe route's u8 stripes); 8- and 16-byte ones lost about 2 %, so they keep theirs.
  const int64_t min_rows = tuning("oop_min_rows");
  if (min_rows > 0 && es <= 4) {
    const uint32_t want = log2_floor(uint64_t(min_rows));
    const uint32_t tc_min = es >= 16 ? 0 : log2_floor(16 / es);
    while (trl < want && trl < log2_ceil(R) && tcl > tc_min) {
      trl++;
      tcl--;
    }
  }
  const int64_t max_rows = tuning("st_tile_rows");
  if (max_rows > 0) {
    const uint32_t cap = log2_floor(uint64_t(max_rows));
    if (trl > cap) {
      tcl = std::min(tcl + (trl - cap), log2_ceil(C));
      trl = cap;
    }
  }
  const uint32_t EPV = uint32_t(16 / es);
  const bool vout = dst_al && (1u << trl) >= EPV;
  // Skewed tiles for misaligned destination rows ("st_skew": 0 off, 1 for 1- and 2-byte
  // elements, 2 up to 8-byte ones); rows back to back and long enough that a chunk holds at
  // most one row boundary.
  const bool skew = !vout && skew_es && (1u << trl) >= EPV && R * es >= 32 &&
                    dst_pitch == R * es && (blocks == 1 || dst_stride == R * C * es);
  // LDS rows: TR elements (plus 16 / es above them when skewed) padded to a multiple of 128
  // bytes; TC rows within the tile budget.
  const uint32_t PB =
      uint32_t(((((uint64_t(1) << trl) + (skew ? EPV : 0)) * es) + 127) / 128 * 128);
  while (tcl > 0 && (uint64_t(1) << tcl) * PB > std::max<size_t>(tile_bytes, PB)) tcl--;
  OopParams p{};
  const uintptr_t sa = reinterpret_cast<uintptr_t>(src);
  p.src = reinterpret_cast<const unsigned char *>(sa & ~uintptr_t(15));
  p.src_off = sa & 15u;
  p.dst = dst;
  p.src_stride = src_stride;
  p.dst_stride = dst_stride;
  p.src_pitch = C * es;
  p.dst_pitch = dst_pitch;
  p.src_end = p.src_off + (src_end ? src_end : (blocks - 1) * src_stride + R * C * es);
  p.R = R;
  p.C = C;
  p.tr_log2 = trl;
  p.tc_log2 = tcl;
  p.PB = PB;
  p.wstore = uint32_t(tuning("st_wstore"));
  p.nt = uint32_t(nt >= 0 ? nt : tuning("oop_nt"));
  // Tiles at least as wide as they are tall walk the tile columns in groups ("oop_order"
  // tiles wide, tile rows inside a group), so the destination rows of the tiles in flight are
  // written whole: on the 1024 x 131071 f64 stripes of a 98304 x 131071 transpose, 5.3-5.4
  // TB/s against 5.0 row-major. Taller tiles (1- to 4-byte elements) have source segments
  // shorter than their destination segments and keep the row-major order (4.55 against 4.4
  // TB/s on 4096 x 199999 u8 stripes).
  p.order = tcl >= trl ? uint32_t(tuning("oop_order")) : 0u;
  const uint32_t TC = 1u << tcl;
  p.ntr = (R + (1ull << trl) - 1) >> trl;
  p.ntc = (C + (1ull << tcl) - 1) >> tcl;
  p.per_block = p.ntr * p.ntc;
  p.tiles = p.per_block * blocks;
  p.per_xcd = (p.tiles + 7) / 8;
  if (p.per_xcd * 8 >= (1ull << 32)) return hipErrorInvalidValue;
  const size_t lds = size_t(TC) * PB;
  if (lds > device_info().lds_per_block) return hipErrorInvalidValue;
  const bool vin = TC >= EPV;
  const int out = skew ? 2 : vout ? 1 : 0;
  if (skew) {
    p.fix_rows = blocks * C;
    p.div_C = FastDiv64(C);
  }
  if (out == 2 && es == 16) return hipErrorInvalidValue;  // 16-byte rows are always aligned
  const uint64_t fix_blocks = out == 2 ? (p.fix_rows + 1 + kOopThreads - 1) / kOopThreads : 0;
  if (p.per_xcd * 8 + fix_blocks >= (1ull << 32)) return hipErrorInvalidValue;
  o->p = p;
  o->lds = lds;
  o->grid = p.per_xcd * 8 + fix_blocks;
  o->es = int(es);
  o->vin = vin;
  o->out = out;
  return hipSuccess;
}

hipError_t launch_oop_transpose(const unsigned char *src, unsigned char *dst, size_t es,
                                uint64_t blocks, uint64_t R, uint64_t C, uint64_t src_stride,
                                uint64_t dst_stride, size_t tile_bytes, hipStream_t st,
                                uint64_t src_end, uint64_t dst_pitch, int nt) {
  if (blocks == 0 || R == 0 || C == 0) return hipSuccess;
  OopLaunch o;
  hipError_t e = prepare_oop_transpose(src, dst, es, blocks, R, C, src_stride, dst_stride,
                                       tile_bytes, src_end, dst_pitch, nt, &o);
  if (e != hipSuccess) return e;
  KernelSpan span(st, "k_oop_transpose", 2.0 * double(blocks) * double(R) * double(C) * double(es));
  e = hipErrorInvalidValue;
  switch (es) {
    case 1: e = launch_oop_es<1>(o.p, o.vin, o.out, o.lds, st); break;
    case 2: e = launch_oop_es<2>(o.p, o.vin, o.out, o.lds, st); break;
    case 4: e = launch_oop_es<4>(o.p, o.vin, o.out, o.lds, st); break;
    case 8: e = launch_oop_es<8>(o.p, o.vin, o.out, o.lds, st); break;
    case 16: e = launch_oop_es<16>(o.p, o.vin, o.out, o.lds, st); break;
  }
  return e;
}

// ---- out-of-place 100 over super-elements of any byte size ------------------------------
// A tile of TR source rows x TC super-elements is staged in LDS as it lies in the source:
// each source row segment (TC * S bytes at any byte alignment) arrives as the aligned 16-byte
// chunks that cover it, stored from the start of its LDS row, so LDS row rr holds its segment
// shifted by that segment's source address mod 16 (which steps by C * S mod 16 per row).
// The destination is written as aligned 16-byte chunks, each gathered from LDS in units of
// U bytes. Destination row c of the tile (TR * S bytes, source column c) is one segment;
// when a tile spans whole columns (TR == R) the tile's destination rows are adjacent, and
// the whole tile (TC * R * S bytes) is one segment, so short rows cost no partial chunks.
// Unit x of a segment is byte (x mod S) of super-element (row, column) = (x / S, c), or in
// whole-column tiles ((x mod RS) / S, x / RS). A chunk only partly inside its segment (the
// rest belongs to a neighbouring tile) is stored in aligned pieces of 8, 4, 2 and 1 bytes.

// Block bx of the launch's grid (see oop_block).
template <int U, bool FC, int NT>
__device__ __forceinline__ void gather_block(const GatherParams &p, uint64_t bx,
                                             unsigned char *lds) {
  using T = typename Elem<U>::T;
  const uint64_t t = (bx & 7u) * p.per_xcd + (bx >> 3);
  if (t >= p.tiles) return;
  const uint64_t g = t / p.per_inst;
  const uint64_t rem = t - g * p.per_inst;
  const uint64_t tr = rem / p.ntc, tc = rem - tr * p.ntc;
  const uint64_t r0 = tr * p.TR, c0 = tc * p.TC;
  const uint32_t nr = uint32_t(std::min<uint64_t>(p.TR, p.R - r0));
  const uint32_t nc = uint32_t(std::min<uint64_t>(p.TC, p.C - c0));
  const uint32_t S = p.S, LP = p.LP;
  const uint64_t CS = p.C * S;
  const uint64_t sbase = g * p.inst_bytes + r0 * CS + uint64_t(c0) * S;  // tile row 0 segment
  const uint32_t seg = nc * S;
  for (uint32_t i = threadIdx.x; i < p.TR * p.nch; i += NT) {
    const uint32_t r = p.div_nch.div(i), j = i - r * p.nch;
    if (r >= nr) break;  // r grows with i
    const uint64_t sb = sbase + r * CS;
    const uint64_t a = (sb & ~uint64_t(15)) + 16u * j;
    if (a < sb + seg)
      *reinterpret_cast<U128 *>(lds + r * LP + 16u * j) =
          *reinterpret_cast<const U128 *>(p.src + a);
  }
  __syncthreads();
  const uint64_t RS64 = p.R * S;
  unsigned char *dbase = p.dst + g * p.inst_bytes + c0 * RS64 + r0 * S;
  const uint32_t RS = FC ? nr * S : 0;       // bytes of one destination row (whole columns)
  const uint32_t nseg = FC ? 1u : nc;
  const uint32_t L = FC ? nc * RS : nr * S;  // bytes per segment
  const uint32_t o0 = uint32_t(sbase & 15u), cs16 = uint32_t(CS & 15u);
  for (uint32_t i = threadIdx.x; i < nseg * p.nchd; i += NT) {
    const uint32_t sg = FC ? 0u : p.div_nchd.div(i), k = i - sg * p.nchd;
    if (sg >= nseg) break;  // sg grows with i
    unsigned char *row = dbase + uint64_t(sg) * RS64;
    const int32_t lead = int32_t(reinterpret_cast<uintptr_t>(row) & 15u);
    const int32_t x0 = int32_t(16u * k) - lead;  // segment byte of this chunk's first byte
    if (x0 >= int32_t(L)) continue;
    const uint32_t xs = uint32_t(std::max(x0, 0));
    const uint32_t xe = std::min<uint32_t>(uint32_t(x0 + 16), L);
    // Position of unit xs: LDS row rr (base rb, shift ro), column c, byte beta; la = its
    // LDS address. Kept incrementally from one unit to the next.
    uint32_t c = sg, rr, beta;
    if (FC) {
      c = p.div_RS.div(xs);
      const uint32_t w = xs - c * RS;
      rr = p.div_S.div(w);
      beta = w - rr * S;
    } else {
      rr = p.div_S.div(xs);
      beta = xs - rr * S;
    }
    uint32_t rb = rr * LP, ro = (o0 + rr * cs16) & 15u, cS = c * S;
    uint32_t la = rb + ro + cS + beta;
    auto next = [&]() {
      beta += U;
      la += U;
      if (beta == S) {
        beta = 0;
        rb += LP;
        ro = (ro + cs16) & 15u;
        if (FC && ++rr == nr) {
          rr = 0;
          rb = 0;
          ro = o0;
          cS += S;
        }
        la = rb + ro + cS;
      }
    };
    union {
      U128 v;
      T u[16 / U];
    } w;
    if (x0 >= 0 && xe == uint32_t(x0 + 16)) {  // whole aligned chunk
#pragma unroll
      for (int q = 0; q < 16 / U; q++) {
        w.u[q] = *reinterpret_cast<const T *>(lds + la);
        next();
      }
      *reinterpret_cast<U128 *>(row + x0) = w.v;
    } else {  // chunk bytes [xs - x0, xe - x0) only
      const uint32_t s0 = xs - uint32_t(x0), e0 = xe - uint32_t(x0);
      w.v = U128{0, 0, 0, 0};
#pragma unroll
      for (int q = 0; q < 16 / U; q++) {
        if (uint32_t(q * U) >= s0 && uint32_t(q * U) < e0) {
          w.u[q] = *reinterpret_cast<const T *>(lds + la);
          next();
        }
      }
      unsigned char *cb = row + x0;  // 16-byte aligned
      for (uint32_t s = s0; s < e0;) {
        const uint32_t n = ((s & 7u) == 0 && s + 8 <= e0)   ? 8u
                           : ((s & 3u) == 0 && s + 4 <= e0) ? 4u
                           : ((s & 1u) == 0 && s + 2 <= e0) ? 2u
                                                            : 1u;
        store_piece(cb, w.v, s, n);
        s += n;
      }
    }
  }
}

template <int U, bool FC, int NT>
__global__ __launch_bounds__(NT) void k_oop_gather(GatherParams p) {
  extern __shared__ __align__(16) unsigned char lds[];
  gather_block<U, FC, NT>(p, blockIdx.x, lds);
}

namespace {
constexpr int kGatherThreads = 512;

// Tile sides (super-elements) for S-byte super-elements within about `budget` bytes of LDS:
// square when both sides are long, else the short side whole and the other grown to fill.
void gather_tile(uint64_t S, uint64_t R, uint64_t C, size_t budget, uint32_t *TR, uint32_t *TC) {
  const uint64_t cells = std::max<uint64_t>(1, budget / S);  // super-elements per tile
  uint64_t t = 1;
  while ((t + 1) * (t + 1) <= cells && t < 1024) t++;
  uint64_t tr = std::min(R, t), tc = std::min(C, t);
  if (tr < t) tc = std::min<uint64_t>(C, std::min<uint64_t>(8192, std::max<uint64_t>(1, cells / tr)));
  if (tc < t) tr = std::min<uint64_t>(R, std::min<uint64_t>(8192, std::max<uint64_t>(1, cells / tc)));
  *TR = uint32_t(tr);
  *TC = uint32_t(tc);
}
}  // namespace

bool oop_gather_fits(uint64_t S) {
  return S >= 1 && S + 48 <= device_info().lds_per_block && S < (1ull << 24);
}

hipError_t prepare_oop_gather(const unsigned char *src, unsigned char *dst, uint64_t S,
                              uint64_t inst, uint64_t R, uint64_t C, size_t tile_bytes,
                              GatherLaunch *o) {
  *o = GatherLaunch{};
  if (inst == 0 || R == 0 || C == 0) return hipSuccess;
  if ((reinterpret_cast<uintptr_t>(src) & 15u) || !oop_gather_fits(S)) return hipErrorInvalidValue;
  GatherParams p{};
  p.src = src;
  p.dst = dst;
  p.R = R;
  p.C = C;
  p.S = uint32_t(S);
  p.inst_bytes = R * C * S;
  gather_tile(S, R, C, tile_bytes, &p.TR, &p.TC);
  // a row segment of TC * S bytes at any alignment spans at most (TC * S + 30) / 16 chunks
  p.fullcol = p.TR == R ? 1u : 0u;
  p.nch = uint32_t((uint64_t(p.TC) * S + 30) / 16);
  // chunks of one destination segment at any alignment: a row of TR * S bytes, or the whole
  // tile when it spans whole columns
  p.nchd = uint32_t(((p.fullcol ? uint64_t(p.TC) * R : uint64_t(p.TR)) * S + 30) / 16);
  p.LP = 16u * p.nch;
  if ((p.LP / 16) % 2 == 0) p.LP += 16;  // odd number of chunks: rows start on different banks
  const size_t lds = size_t(p.TR) * p.LP;
  if (lds > device_info().lds_per_block) return hipErrorInvalidValue;
  p.div_nch = FastDiv32(p.nch);
  p.div_nchd = FastDiv32(p.nchd);
  p.div_S = FastDiv32(p.S);
  p.div_RS = FastDiv32(uint32_t(std::min<uint64_t>(R * S, 0xffffffffu)));
  p.ntc = (C + p.TC - 1) / p.TC;
  p.per_inst = ((R + p.TR - 1) / p.TR) * p.ntc;
  p.tiles = p.per_inst * inst;
  p.per_xcd = (p.tiles + 7) / 8;
  if (p.per_xcd * 8 >= (1ull << 32)) return hipErrorInvalidValue;
  uint64_t u = 16;
  while (u > 1 && ((S | reinterpret_cast<uintptr_t>(dst)) & (u - 1))) u >>= 1;
  o->p = p;
  o->lds = lds;
  o->grid = p.per_xcd * 8;
  o->u = int(u);
  o->fc = p.fullcol != 0;
  return hipSuccess;
}

hipError_t launch_oop_gather(const unsigned char *src, unsigned char *dst, uint64_t S,
                             uint64_t inst, uint64_t R, uint64_t C, size_t tile_bytes,
                             hipStream_t st) {
  if (inst == 0 || R == 0 || C == 0) return hipSuccess;
  GatherLaunch o;
  hipError_t e0 = prepare_oop_gather(src, dst, S, inst, R, C, tile_bytes, &o);
  if (e0 != hipSuccess) return e0;
  KernelSpan span(st, "k_oop_gather", 2.0 * double(inst) * double(R) * double(C) * double(S));
  const GatherParams &p = o.p;
  const size_t lds = o.lds;
  auto go = [&](auto kern) -> hipError_t {
    if (lds > 65536) {
      hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(uint32_t(o.grid)), dim3(kGatherThreads), lds, st, p);
    return hipGetLastError();
  };
  auto by_u = [&](auto fc) -> hipError_t {
    constexpr bool F = decltype(fc)::value;
    switch (o.u) {
      case 16: return go(k_oop_gather<16, F, kGatherThreads>);
      case 8: return go(k_oop_gather<8, F, kGatherThreads>);
      case 4: return go(k_oop_gather<4, F, kGatherThreads>);
      case 2: return go(k_oop_gather<2, F, kGatherThreads>);
      default: return go(k_oop_gather<1, F, kGatherThreads>);
    }
  };
  return o.fc ? by_u(std::true_type{}) : by_u(std::false_type{});
}

// ---- host: planning ---------------------------------------------------------------------
namespace {
size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Best stripe height for dimension D (the other one is L): a divisor of D near the target
// super-element size when one exists, else the target itself with a ragged remainder.
// Returns false when no height keeps a stripe within the scratch limit.
bool pick_height(size_t es, uint64_t D, uint64_t L, uint64_t target_bytes, uint64_t max_stripe,
                 uint64_t *h_out) {
  uint64_t target = std::max<uint64_t>(1, target_bytes / es);
  const uint64_t hmax = uint64_t(std::max<int64_t>(1, tuning("st_max_height")));
  target = std::min<uint64_t>(target, hmax);
  while (target > 1 && target * L * es > max_stripe) target >>= 1;
  if (target * L * es > max_stripe) return false;
  if (D < 2 * target) return false;  // fewer than two stripes: not worth a scratch copy
  // A divisor in [target/2, 2*target] (within the scratch limit) avoids the ragged step.
  uint64_t best = 0;
  for (uint64_t h = target / 2 == 0 ? 1 : target / 2; h <= 2 * target && h <= hmax; h++) {
    if (D % h || h * L * es > max_stripe || (h * es) % 16) continue;
    if (!best || (h > best && h <= target) || (best > target && h < best)) best = h;
  }
  *h_out = best ? best : target;
  return true;
}
}  // namespace

bool plan_st(size_t es, uint64_t M, uint64_t N, int variant, uint64_t target_bytes,
             uint64_t scratch_cap, int64_t pt_vw, bool rotate, StPlan *out,
             uint64_t base_addr) {
  if (M < 2 || N < 2) return false;
  const uint64_t D = variant == 0 ? M : N;   // striped dimension
  const uint64_t L = variant == 0 ? N : M;   // the other one
  uint64_t h = 0;
  const uint64_t max_stripe = std::max<uint64_t>(scratch_cap, 1);
  if (!pick_height(es, D, L, target_bytes, max_stripe, &h)) return false;
  StPlan p{};
  p.variant = variant;
  p.es = es;
  p.M = M;
  p.N = N;
  p.h = h;
  p.q = D / h;
  p.r = D - p.q * h;
  const uint64_t stripe = h * L * es;
  // Rotation saves the copy pass but costs one launch per stripe: worth it once the stripes
  // need more than one scratch fill; smaller matrices take one copy and one transpose launch.
  p.rotate = rotate && p.q * stripe > scratch_cap;
  p.group = p.rotate ? 1 : std::max<uint64_t>(1, std::min<uint64_t>(p.q, scratch_cap / stripe));
  if (p.q > 1) {
    if (variant == 0) plan_100(es, 1, p.q, N, h, &p.pt, pt_vw, base_addr);
    else plan_100(es, 1, M, p.q, h, &p.pt, pt_vw, base_addr);
  }
  const uint64_t pt_words = p.q > 1 ? p.pt.flag_words : 0;
  const uint64_t pad_words = p.r ? pad_flag_words(M * N, es) : 0;
  p.stripe_off = 0;
  p.b_off = align256(p.group * stripe);
  p.f_off = p.b_off + align256(p.r * L * es);
  p.flag_words = std::max(pt_words, pad_words);
  p.scratch_bytes = p.f_off + align256(p.flag_words * 4);
  *out = p;
  return true;
}

hipError_t run_st(unsigned char *buf, const StPlan &p, unsigned char *scratch, size_t tile_bytes,
                  uint32_t pt_blocks, hipStream_t st) {
  const size_t es = p.es;
  const uint64_t M = p.M, N = p.N, h = p.h, q = p.q, r = p.r;
  const uint64_t total = M * N * es;
  unsigned char *S = scratch + p.stripe_off;
  unsigned char *B = scratch + p.b_off;
  uint32_t *F = reinterpret_cast<uint32_t *>(scratch + p.f_off);
  const bool rot = p.rotate && q > 1;
  hipError_t e = hipSuccess;
#define ST_TRY(x) do { e = (x); if (e != hipSuccess) return e; } while (0)
  if (p.variant == 0) {
    const uint64_t stripe = h * N * es;
    if (r)
      ST_TRY(launch_oop_transpose(buf + q * stripe, B, es, 1, r, N, 0, 0, tile_bytes, st,
                                  total - q * stripe));
    if (rot) {
      // Stripe k is transposed into slot k - 1 (stripe 0 into scratch, copied to the last
      // slot at the end): each launch reads one slot and writes another, so no copy pass.
      ST_TRY(launch_oop_transpose(buf, S, es, 1, h, N, 0, 0, tile_bytes, st, total));
      for (uint64_t k = 1; k < q; k++)
        ST_TRY(launch_oop_transpose(buf + k * stripe, buf + (k - 1) * stripe, es, 1, h, N, 0, 0,
                                    tile_bytes, st, total - k * stripe));
      ST_TRY(launch_copy(buf + (q - 1) * stripe, S, stripe, st));
    } else {
      for (uint64_t g0 = 0; g0 < q; g0 += p.group) {
        const uint64_t n = std::min(p.group, q - g0);
        ST_TRY(launch_copy(S, buf + g0 * stripe, n * stripe, st));
        ST_TRY(launch_oop_transpose(S, buf + g0 * stripe, es, n, h, N, stripe, stripe,
                                    tile_bytes, st));
      }
    }
    if (q > 1) {  // slot s holds stripe s + 1 (mod q) when rotated
      ST_TRY(launch_fill_u32(F, p.pt.flag_words, 0, st));
      ST_TRY(launch_100_pttwac(buf, F, 1, q, N, h * es, p.pt, pt_blocks, st, rot ? 1 : 0, 0));
    }
    if (r) {
      ST_TRY(launch_padding(buf, es, q * h, N, M, false, F, st));
      ST_TRY(launch_copy2d(buf + q * h * es, M * es, B, r * es, r * es, N, st));
    }
  } else {
    const uint64_t stripe = h * M * es;
    if (r) {
      ST_TRY(launch_copy2d(B, r * es, buf + q * h * es, N * es, r * es, M, st));
      ST_TRY(launch_padding(buf, es, q * h, M, N, true, F, st));
    }
    if (q > 1) {  // rotated: block k lands in slot k + 1 (mod q)
      ST_TRY(launch_fill_u32(F, p.pt.flag_words, 0, st));
      ST_TRY(launch_100_pttwac(buf, F, 1, M, q, h * es, p.pt, pt_blocks, st, 0, rot ? 1 : 0));
    }
    if (rot) {
      ST_TRY(launch_copy(S, buf, stripe, st));
      for (uint64_t k = 0; k + 1 < q; k++)
        ST_TRY(launch_oop_transpose(buf + (k + 1) * stripe, buf + k * stripe, es, 1, M, h, 0, 0,
                                    tile_bytes, st, total - (k + 1) * stripe));
      ST_TRY(launch_oop_transpose(S, buf + (q - 1) * stripe, es, 1, M, h, 0, 0, tile_bytes, st));
    } else {
      for (uint64_t g0 = 0; g0 < q; g0 += p.group) {
        const uint64_t n = std::min(p.group, q - g0);
        ST_TRY(launch_copy(S, buf + g0 * stripe, n * stripe, st));
        ST_TRY(launch_oop_transpose(S, buf + g0 * stripe, es, n, M, h, stripe, stripe,
                                    tile_bytes, st));
      }
    }
    if (r) ST_TRY(launch_oop_transpose(B, buf + q * stripe, es, 1, M, r, 0, 0, tile_bytes, st));
  }
#undef ST_TRY
  return hipSuccess;
}

// ---- grouped forms (a batch's through-scratch rounds) -------------------------------------
// Many launches of the kernels above as one: every workgroup walks the units of the work
// list (each item's launch grid, padded to a multiple of 8), finds its item by binary search
// and runs that launch's block. The grid is a multiple of 8, so a unit keeps the XCD it
// would have had in its own launch.
template <int ES, bool VIN, int OUT>
__global__ __launch_bounds__(kOopGroupThreads) void kg_oop_transpose(const OopParams *__restrict__ items,
                                                                     const uint64_t *__restrict__ prefix,
                                                                     uint32_t n, uint64_t total) {
  static_assert(kOopGroupThreads == kOopThreads, "fix-up units are sized for kOopThreads");
  extern __shared__ __align__(16) unsigned char lds[];
  for (uint64_t q = blockIdx.x; q < total; q += gridDim.x) {
    const uint32_t it = find_item(prefix, n, q);
    const OopParams p = items[it];
    oop_block<ES, VIN, OUT, kOopGroupThreads>(p, q - prefix[it], lds);
    __syncthreads();  // the next unit refills LDS
  }
}

template <int U, bool FC>
__global__ __launch_bounds__(kOopGroupThreads) void kg_oop_gather(const GatherParams *__restrict__ items,
                                                                  const uint64_t *__restrict__ prefix,
                                                                  uint32_t n, uint64_t total) {
  static_assert(kOopGroupThreads == kGatherThreads, "one block shape for both");
  extern __shared__ __align__(16) unsigned char lds[];
  for (uint64_t q = blockIdx.x; q < total; q += gridDim.x) {
    const uint32_t it = find_item(prefix, n, q);
    const GatherParams p = items[it];
    gather_block<U, FC, kOopGroupThreads>(p, q - prefix[it], lds);
    __syncthreads();
  }
}

// Unit u of a copy item: bytes [u * kCopyItemUnit, ...) of it as 16-byte vectors, all of a
// lane's in flight together (non-temporal loads, cached stores: the arena is read back next),
// the last unit's bytes past its last whole vector one per lane. NT lanes.
template <int NT>
__device__ __forceinline__ void copy_unit(const CopyItem &c, uint64_t u) {
  constexpr int UN = int(kCopyItemUnit / 16) / NT;
  static_assert(uint64_t(UN) * NT * 16 == kCopyItemUnit, "whole vectors per lane");
  const uint64_t b0 = u * kCopyItemUnit;
  if (b0 >= c.bytes) return;
  const uint64_t len = std::min<uint64_t>(kCopyItemUnit, c.bytes - b0);
  const uint32_t nv = uint32_t(len / 16);
  const U128 *s = reinterpret_cast<const U128 *>(c.src + b0);
  U128 *d = reinterpret_cast<U128 *>(c.dst + b0);
  U128 v[UN];
#pragma unroll
  for (int k = 0; k < UN; k++) {
    const uint32_t i = threadIdx.x + uint32_t(NT) * k;
    if (i < nv) v[k] = __builtin_nontemporal_load(s + i);  // the matrix is read once
  }
#pragma unroll
  for (int k = 0; k < UN; k++) {
    const uint32_t i = threadIdx.x + uint32_t(NT) * k;
    if (i < nv) d[i] = v[k];
  }
  const uint32_t tail = uint32_t(len) - 16u * nv;
  if (threadIdx.x < tail) c.dst[b0 + 16u * nv + threadIdx.x] = c.src[b0 + 16u * nv + threadIdx.x];
}

__global__ __launch_bounds__(256) void kg_copy(const CopyItem *__restrict__ items, const uint64_t *__restrict__ prefix,
                                               uint32_t n, uint64_t total) {
  for (uint64_t q = blockIdx.x; q < total; q += gridDim.x) {
    const uint32_t it = find_item(prefix, n, q);
    const CopyItem c = items[it];
    copy_unit<256>(c, q - prefix[it]);
  }
}

// Unit u of an item: round u of its unpitch copy; unit 0 also copies the end bytes.
template <bool NTS>
__global__ __launch_bounds__(256) void kg_unpitch(const UnpitchItem *__restrict__ items,
                                                  const uint64_t *__restrict__ prefix, uint32_t n,
                                                  uint64_t total) {
  static_assert(kUnpitchItemUnit == 256 * kCopyUnroll, "one round per unit");
  for (uint64_t q = blockIdx.x; q < total; q += gridDim.x) {
    const uint32_t it = find_item(prefix, n, q);
    const UnpitchItem c = items[it];
    const uint64_t u = q - prefix[it];
    if (u * kUnpitchItemUnit < c.nb)
      unpitch_round<NTS>(c.dst, c.src, c.RB, c.P, c.head, c.nb, c.div_rb, u);
    if (u == 0 && threadIdx.x == 0)
      unpitch_ends(c.dst, c.src, c.RB, c.P, c.total, c.head, c.nb, c.div_rb);
  }
}

UnpitchItem make_unpitch_item(unsigned char *dst, const unsigned char *src, uint64_t RB,
                              uint64_t P, uint64_t rows) {
  UnpitchItem c{};
  c.dst = dst;
  c.src = src;
  c.RB = RB;
  c.P = P;
  c.total = RB * rows;
  c.head = uint32_t(std::min<uint64_t>((16u - (reinterpret_cast<uintptr_t>(dst) & 15u)) & 15u,
                                       c.total));
  c.nb = (c.total - c.head) / 16;
  c.div_rb = FastDiv64(RB);
  return c;
}

const void *kg_unpitch_func(bool nts) {
  return nts ? reinterpret_cast<const void *>(kg_unpitch<true>)
             : reinterpret_cast<const void *>(kg_unpitch<false>);
}

namespace {
template <int ES>
const void *kg_oop_es(bool vin, int out) {
  if (out == 2)
    return ES == 16 ? nullptr
           : vin    ? reinterpret_cast<const void *>(kg_oop_transpose<ES, true, 2>)
                    : reinterpret_cast<const void *>(kg_oop_transpose<ES, false, 2>);
  if (vin && out) return reinterpret_cast<const void *>(kg_oop_transpose<ES, true, 1>);
  if (vin) return reinterpret_cast<const void *>(kg_oop_transpose<ES, true, 0>);
  if (out) return reinterpret_cast<const void *>(kg_oop_transpose<ES, false, 1>);
  return reinterpret_cast<const void *>(kg_oop_transpose<ES, false, 0>);
}
template <bool FC>
const void *kg_gather_fc(int u) {
  switch (u) {
    case 16: return reinterpret_cast<const void *>(kg_oop_gather<16, FC>);
    case 8: return reinterpret_cast<const void *>(kg_oop_gather<8, FC>);
    case 4: return reinterpret_cast<const void *>(kg_oop_gather<4, FC>);
    case 2: return reinterpret_cast<const void *>(kg_oop_gather<2, FC>);
    default: return reinterpret_cast<const void *>(kg_oop_gather<1, FC>);
  }
}
}  // namespace

const void *kg_oop_transpose_func(int es, bool vin, int out) {
  switch (es) {
    case 1: return kg_oop_es<1>(vin, out);
    case 2: return kg_oop_es<2>(vin, out);
    case 4: return kg_oop_es<4>(vin, out);
    case 8: return kg_oop_es<8>(vin, out);
    case 16: return kg_oop_es<16>(vin, out);
    default: return nullptr;
  }
}
const void *kg_oop_gather_func(int u, bool fc) {
  return fc ? kg_gather_fc<true>(u) : kg_gather_fc<false>(u);
}
const void *kg_copy_func() { return reinterpret_cast<const void *>(kg_copy); }

}  // namespace lm
