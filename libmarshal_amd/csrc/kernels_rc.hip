// kernels_rc.hip - full M x N transpose by row and column shuffles, every pass LDS-staged.
//
// With c = gcd(M, N), b = N / c and a = M / c, the in-place transpose factors into passes that
// each permute elements only within a row or only within a column (Catanzaro, Keller and
// Garland, "A Decomposition for In-place Matrix Transposition", PPoPP 2014; the gather forms
// below are this library's own derivation, checked by brute force over all M, N < 30):
//
//   C1 (only when c > 1)  new[r][j]   = old[(r - floor(j / b)) mod M][j]
//   R                     new[r][j']  = old[r][q*b + s]
//                           q = (r - j') mod c,  u = floor(((r - q) mod M) / c),
//                           s = ((floor(j' / c) - u) * a^-1) mod b      (a^-1 taken mod b)
//   C2                    new[i'][j'] = old[(D mod M + floor(floor(D / M) / b)) mod M][j'],
//                           D = i'*N + j'
//
// No factorisation of M or N is needed, so prime dimensions take 2 or 3 streaming passes
// instead of element-granular cycle following. A row pass stages whole rows in LDS; a column
// pass stages a block of w adjacent columns over all M rows. Elements are S bytes, moved as
// S/UB units of UB bytes. Not a reference algorithm: the reference transposes only through
// its tiled stages or the element-granular fallback (src/cl_marshal.cc:561-776).
#include "kernels_dev.h"

#include <algorithm>
#include <numeric>

namespace lm {

template <int UB, int NT>
__global__ __launch_bounds__(NT) void k_rc_row(RcParams p) {
  using V = typename Elem<UB>::T;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  V *lds = reinterpret_cast<V *>(lds_raw);
  const uint64_t nblk = (p.inst * p.M + p.rows - 1) / p.rows;
  for (uint64_t b = blockIdx.x; b < nblk; b += gridDim.x) rc_row_block<UB, NT>(p, b, lds);
}

// Column blocks handled by one XCD (workgroups are dealt round-robin over the 8 XCDs) are
// adjacent, so the 128-byte lines their row segments share stay in that XCD's L2.
template <int UB, int NT, bool SHUFFLE>
__global__ __launch_bounds__(NT) void k_rc_col(RcParams p) {
  using V = typename Elem<UB>::T;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  V *lds = reinterpret_cast<V *>(lds_raw);
  const uint64_t per = (p.ncb + 7) / 8;
  for (uint64_t q = blockIdx.x; q < per * 8; q += gridDim.x) {
    uint64_t cb = (q & 7) * per + (q >> 3);
    if (cb < p.ncb) rc_col_block<UB, NT, SHUFFLE>(p, cb, lds);
  }
}

// ---- host ---------------------------------------------------------------------------------

static uint64_t inv_mod(uint64_t a, uint64_t m) {
  if (m <= 1) return 0;
  int64_t t = 0, nt = 1, r = int64_t(m), nr = int64_t(a % m);
  while (nr) {
    int64_t q = r / nr;
    int64_t tmp = t - q * nt; t = nt; nt = tmp;
    tmp = r - q * nr; r = nr; nr = tmp;
  }
  return uint64_t(t < 0 ? t + int64_t(m) : t);
}

bool plan_rc(size_t es, uint64_t M, uint64_t N, uint64_t S, size_t lds_budget, RcPlan *out,
             uint64_t inst) {
  // S-byte elements made of UB-byte units; M, N < 2^31; rows and column blocks fit LDS.
  if (M < 2 || N < 2 || M >= (1ull << 31) || N >= (1ull << 31)) return false;
  if (inst < 1 || inst * M >= (1ull << 31)) return false;
  int ub = (S % 16 == 0) ? 16 : (S % 8 == 0) ? 8 : (S % 4 == 0) ? 4 : (S % 2 == 0) ? 2 : 1;
  (void)es;
  uint64_t row_bytes = N * S;
  if (row_bytes > lds_budget || M * S > lds_budget) return false;
  // Row pass: as many whole rows as fit ~64 KiB (at least one).
  uint64_t rows = std::max<uint64_t>(1, std::min<uint64_t>(M, (64 * 1024) / row_bytes));
  // Column pass: segments of >= 16 bytes (whole elements) when the block fits; a whole
  // 128-byte line per row when the columns are short.
  const uint64_t seg_target = M * 128 <= 40 * 1024 ? 128 : 16;
  uint64_t w = std::max<uint64_t>(1, (seg_target + S - 1) / S);
  while (w > 1 && M * w * S > lds_budget) w--;
  if (w > N) w = N;
  out->ub = ub;
  out->m = uint32_t(S / ub);
  out->rows = uint32_t(rows);
  out->w = uint32_t(w);
  out->c = uint32_t(std::gcd(M, N));
  out->inst = inst;
  out->row_lds = size_t(rows * row_bytes);
  out->col_lds = size_t(M * w * S);
  return true;
}

RcParams make_rc(unsigned char *base, uint64_t M, uint64_t N, const RcPlan &pl) {
  RcParams p{};
  p.base = base;
  p.M = uint32_t(M);
  p.N = uint32_t(N);
  p.c = pl.c;
  p.a = uint32_t(M / pl.c);
  p.b = uint32_t(N / pl.c);
  p.ainv = uint32_t(inv_mod(p.a, p.b));
  p.m = pl.m;
  p.rows = pl.rows;
  p.w = pl.w;
  p.ncolblk = uint32_t((N + pl.w - 1) / pl.w);
  p.inst = pl.inst;
  p.ncb = pl.inst * p.ncolblk;
  p.div_M32 = FastDiv32(uint32_t(M));
  p.div_ncolblk = FastDiv64(p.ncolblk);
  p.div_c = FastDiv32(p.c);
  p.div_b = FastDiv32(p.b);
  p.div_b64 = FastDiv64(p.b);
  p.div_m = FastDiv32(p.m);
  p.div_upr = FastDiv32(p.N * p.m);
  p.div_M64 = FastDiv64(M);
  p.div_pitch = FastDiv32(p.w * p.m);
  uint32_t lastw = uint32_t(N - uint64_t(p.ncolblk - 1) * pl.w);
  p.div_seg_last = FastDiv32(lastw * p.m);
  return p;
}

namespace {
template <int NT>
hipError_t set_lds(const void *k, size_t lds) {
  if (lds > 65536)
    return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  return hipSuccess;
}

template <int UB>
hipError_t launch_rc_ub(const RcParams &p, const RcPlan &pl, int pass, uint32_t max_blocks,
                        hipStream_t st) {
  constexpr int NT = 1024;
  if (pass == 0) {
    auto k = k_rc_row<UB, NT>;
    hipError_t e = set_lds<NT>(reinterpret_cast<const void *>(k), pl.row_lds);
    if (e != hipSuccess) return e;
    uint64_t nblk = (p.inst * p.M + p.rows - 1) / p.rows;
    hipLaunchKernelGGL(k, dim3(uint32_t(std::min<uint64_t>(nblk, max_blocks))), dim3(NT),
                       pl.row_lds, st, p);
  } else {
    auto k = pass == 1 ? k_rc_col<UB, NT, false> : k_rc_col<UB, NT, true>;
    hipError_t e = set_lds<NT>(reinterpret_cast<const void *>(k), pl.col_lds);
    if (e != hipSuccess) return e;
    uint32_t blocks = uint32_t(std::min<uint64_t>((p.ncb + 7) / 8 * 8,
                                                  (uint64_t(max_blocks) + 7) / 8 * 8));
    hipLaunchKernelGGL(k, dim3(blocks), dim3(NT), pl.col_lds, st, p);
  }
  return hipGetLastError();
}
}  // namespace

hipError_t launch_rc(unsigned char *base, uint64_t M, uint64_t N, const RcPlan &pl, int pass,
                     uint32_t max_blocks, hipStream_t st) {
  RcParams p = make_rc(base, M, N, pl);
  KernelSpan span(st, pass == 0 ? "k_rc_row" : "k_rc_col",
                  2.0 * double(pl.inst) * double(M) * double(N) * double(pl.m) * double(pl.ub));
  switch (pl.ub) {
    case 1: return launch_rc_ub<1>(p, pl, pass, max_blocks, st);
    case 2: return launch_rc_ub<2>(p, pl, pass, max_blocks, st);
    case 4: return launch_rc_ub<4>(p, pl, pass, max_blocks, st);
    case 8: return launch_rc_ub<8>(p, pl, pass, max_blocks, st);
    case 16: return launch_rc_ub<16>(p, pl, pass, max_blocks, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace lm
