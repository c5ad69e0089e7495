/*
 * oracle.c - CPU restatement of libmarshal's in-place layout conversions.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the oracle the HIP library is tested against: only
 * the test suite, the smoke check and the CPU-baseline leg of the benchmark load it. Nothing in the
 * library links, calls or falls back to it.
 *
 * Elements are opaque bit patterns of `es` bytes (1, 2, 4, 8 or 16). All indices are 64-bit.
 * Every function names the reference code it restates (paths relative to the reference tree).
 *
 * Pins (what checks each function independently of this file; see tests/test_oracle.py):
 *   orc_next / orc_num_cycles   closed form i*m mod (mn-1), brute-force cycle walk, the
 *                               reference KAT (test/cl_unittest.cc:705-709) and the
 *                               reference's own plan.hpp compiled into oracle/_ref.
 *   orc_010 / orc_100 / orc_0100 / orc_transpose
 *                               numpy reshape+transpose, committed golden vectors.
 *   orc_inplace_*               equality with the out-of-place ops on unique tags, over every
 *                               factorisation of small shapes (brute force).
 *   orc_padding / orc_unpadding numpy pad / slice.
 *   orc_tile_sizes              divisors > 1 by brute force and the reference's Factorize
 *                               compiled into oracle/_ref.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <unistd.h>

typedef int64_t i64;
typedef uint64_t u64;

/* ------------------------------------------------------------------------------------ */
/* Index map and cycle statistics                                                        */
/* ------------------------------------------------------------------------------------ */

/* Destination of element i when a row-major m x n matrix is transposed in place:
 * next(i) = i*m - (mn-1)*floor(i/n) = i*m mod (mn-1), with mn-1 fixed.
 * Restates Transposition::Next, include/plan.hpp:54-56, written as (i mod n)*m + i/n so
 * that it cannot overflow for m*n < 2^63. */
u64 orc_next(u64 i, u64 m, u64 n) {
  return (i % n) * m + i / n;
}

/* Number of non-trivial cycles of the m x n in-place transpose permutation.
 * Restates Transposition::GetNumCycles, include/plan.hpp:57-73: i is counted when it is the
 * smallest element of its cycle and not a fixed point. */
u64 orc_num_cycles(i64 m, i64 n) {
  u64 mn1 = (u64)m * (u64)n - 1, count = 0;
  for (u64 i = 1; i <= mn1; i++) {
    u64 j = orc_next(i, m, n);
    if (j == i) continue;
    while (j > i) j = orc_next(j, m, n);
    if (j == i) count++;
  }
  return count;
}

/* Length of the longest cycle (statistics for tests and planning). */
u64 orc_longest_cycle(i64 m, i64 n) {
  u64 mn1 = (u64)m * (u64)n - 1, best = 1;
  for (u64 i = 1; i < mn1; i++) {
    u64 j = orc_next(i, m, n), len = 1;
    while (j > i) { j = orc_next(j, m, n); len++; }
    if (j == i && len > best) best = len;
  }
  return best;
}

/* ------------------------------------------------------------------------------------ */
/* Out-of-place references                                                                */
/* ------------------------------------------------------------------------------------ */

/* 010: [A][a][B] -> [A][B][a]. Restates cpu_aos_asta, test/cl_unittest.cc:138-151
 * (dst[j + k*tile] = src[j*width + k] per tile), with tile = a and width = B. */
void orc_010(const void *src, void *dst, size_t es, i64 A, i64 a, i64 B) {
  const char *s = (const char *)src;
  char *d = (char *)dst;
  u64 n = (u64)a * (u64)B;
  for (i64 t = 0; t < A; t++) {
    const char *st = s + (u64)t * n * es;
    char *dt = d + (u64)t * n * es;
    for (i64 i = 0; i < a; i++)
      for (i64 j = 0; j < B; j++)
        memcpy(dt + ((u64)j * a + i) * es, st + ((u64)i * B + j) * es, es);
  }
}

/* 100: [A][B][b] -> [B][A][b]. Restates cpu_soa_asta, test/cl_unittest.cc:153-166
 * (dst[i][k][j] = src[k][i][j] with width = A, height/tile = B, tile = b). */
void orc_100(const void *src, void *dst, size_t es, i64 A, i64 B, i64 b) {
  const char *s = (const char *)src;
  char *d = (char *)dst;
  u64 sb = (u64)b * es;
  for (i64 x = 0; x < A; x++)
    for (i64 y = 0; y < B; y++)
      memcpy(d + ((u64)y * A + x) * sb, s + ((u64)x * B + y) * sb, sb);
}

/* 0100: [A][a][B][b] -> [A][B][a][b] = A instances of 100(a, B, b).
 * Restates the golden loop of test/cl_unittest.cc:693-696. */
void orc_0100(const void *src, void *dst, size_t es, i64 A, i64 a, i64 B, i64 b) {
  u64 inst = (u64)a * B * b * es;
  for (i64 x = 0; x < A; x++)
    orc_100((const char *)src + x * inst, (char *)dst + x * inst, es, a, B, b);
}

/* Out-of-place transpose of an M x N row-major matrix. The reference computes its golden as
 * cpu_aos_asta(src,dst,h,w,1) followed by cpu_soa_asta(dst,src,w,h,1)
 * (test/cl_unittest.cc:549-551): the first call is the identity and the second is
 * 100(M, N, 1). */
void orc_transpose(const void *src, void *dst, size_t es, i64 M, i64 N) {
  orc_100(src, dst, es, M, N, 1);
}

/* The reference's CPU golden for a full transpose exactly as its test runs it
 * (test/cl_unittest.cc:547-551), single-threaded: cpu_aos_asta(src, dst, h, w, 1), a 010 with
 * unit tiles (a copy into `tmp`), then cpu_soa_asta(dst, src, w, h, 1), the 100 that writes
 * the transpose back into `buf`. `tmp` holds M*N elements. Timed as the CPU baseline. */
void orc_golden_transpose(void *buf, void *tmp, size_t es, i64 M, i64 N) {
  orc_010(buf, tmp, es, M, 1, N);
  orc_100(tmp, buf, es, M, N, 1);
}

/* ------------------------------------------------------------------------------------ */
/* In-place restatements of the reference's device algorithms (sequential schedules)    */
/* ------------------------------------------------------------------------------------ */

/* 010 barrier-sync: every tile is scattered into a tile-sized store with the index map and
 * copied back. Restates BS_marshal, src/cl/cl_aos_asta.cl:53-84 (store[next] = in[i];
 * barrier; in[i] = store[i]). Returns 0, or -1 on allocation failure. */
int orc_inplace_010_bs(void *buf, size_t es, i64 A, i64 a, i64 B) {
  u64 n = (u64)a * B;
  char *store = (char *)malloc(n * es != 0 ? n * es : 1);
  if (!store) return -1;
  for (i64 t = 0; t < A; t++) {
    char *in = (char *)buf + (u64)t * n * es;
    for (u64 i = 0; i < n; i++)
      memcpy(store + orc_next(i, a, B) * es, in + i * es, es);
    memcpy(in, store, n * es);
  }
  free(store);
  return 0;
}

/* One element-granular cycle-following pass over a single m x n tile with visited bits.
 * `flag_word(idx)` / `flag_bit(idx)` place the bits (the reference's REMAP or narrowing
 * layouts). The sequential schedule visits start points in increasing order, exactly the
 * claim protocol of src/cl/cl_aos_asta.cl:238-272: read the start element, test its bit,
 * then repeatedly read the next element, set its bit, and write the carried value only if
 * the bit was clear. */
typedef struct {
  u64 sh_sz;  /* words (REMAP layout) */
  int remap;  /* 1: word = idx % sh_sz, bit = idx / sh_sz; 0: word = idx/32, bit = idx%32 */
} flag_layout;

static inline void fl_pos(const flag_layout *L, u64 idx, u64 *w, unsigned *bit) {
  if (L->remap) { *w = idx % L->sh_sz; *bit = (unsigned)(idx / L->sh_sz); }
  else { *w = idx >> 5; *bit = (unsigned)(idx & 31); }
}

static void cycle_follow(char *base, size_t es, u64 m, u64 n, uint32_t *flags,
                         const flag_layout *L) {
  u64 last = m * n - 1;
  char d1[16], d2[16];
  for (u64 s = 0; s < last; s++) {
    u64 nxt = orc_next(s, m, n);
    if (nxt == s) continue;
    memcpy(d1, base + s * es, es);
    u64 w; unsigned bit;
    fl_pos(L, s, &w, &bit);
    int done = (flags[w] >> bit) & 1u;
    while (!done) {
      memcpy(d2, base + nxt * es, es);
      fl_pos(L, nxt, &w, &bit);
      done = (flags[w] >> bit) & 1u;
      flags[w] |= 1u << bit;
      if (!done) memcpy(base + nxt * es, d1, es);
      memcpy(d1, d2, es);
      nxt = orc_next(nxt, m, n);
    }
  }
}

/* 010 cycle following with visited bits per tile in the REMAP layout (word = idx % sh_sz,
 * bit = idx / sh_sz, sh_sz = R*ceil(aB/32) + P*(that>>5)). Restates transpose_010_PTTWAC,
 * src/cl/cl_aos_asta.cl:202-278, and the sizing of src/cl_marshal.cc:304-305.
 * Returns 0, -1 on allocation failure, -2 on invalid R. */
int orc_inplace_010_pttwac(void *buf, size_t es, i64 A, i64 a, i64 B, int R, int P) {
  if (R < 1 || R > 32 || (R & (R - 1))) return -2;
  u64 n = (u64)a * B;
  u64 sh = (u64)R * ((n + 31) / 32);
  sh += (sh >> 5) * (u64)P;
  if (sh == 0) sh = 1;
  uint32_t *flags = (uint32_t *)malloc(sh * sizeof(uint32_t));
  if (!flags) return -1;
  flag_layout L = {sh, 1};
  for (i64 t = 0; t < A; t++) {
    memset(flags, 0, sh * sizeof(uint32_t));
    cycle_follow((char *)buf + (u64)t * n * es, es, a, B, flags, &L);
  }
  free(flags);
  return 0;
}

/* 0100 [A][a][B][b] -> [A][B][a][b] by cycle following over super-elements of b elements,
 * with 32 visited bits per word ("narrowing", N = 32). Restates _transpose_100 and its
 * 0100 instance offsets, src/cl/cl_aos_asta.cl:348-569 and 740-761 (flags zeroed per call,
 * src/cl_marshal.cc:353-361). A super-element is an opaque element of b*es bytes. */
int orc_inplace_0100_pttwac(void *buf, size_t es, i64 A, i64 a, i64 B, i64 b) {
  u64 g = (u64)a * B, sb = (u64)b * es;
  u64 words = (g + 31) / 32;
  uint32_t *flags = (uint32_t *)malloc((words ? words : 1) * sizeof(uint32_t));
  char *d1 = (char *)malloc(sb ? sb : 1), *d2 = (char *)malloc(sb ? sb : 1);
  if (!flags || !d1 || !d2) { free(flags); free(d1); free(d2); return -1; }
  for (i64 x = 0; x < A; x++) {
    char *base = (char *)buf + (u64)x * g * sb;
    memset(flags, 0, words * sizeof(uint32_t));
    u64 last = g - 1;
    for (u64 s = 0; s < last; s++) {
      u64 nxt = orc_next(s, a, B);
      if (nxt == s) continue;
      memcpy(d1, base + s * sb, sb);
      int done = (flags[s >> 5] >> (s & 31)) & 1u;
      while (!done) {
        memcpy(d2, base + nxt * sb, sb);
        done = (flags[nxt >> 5] >> (nxt & 31)) & 1u;
        flags[nxt >> 5] |= 1u << (nxt & 31);
        if (!done) memcpy(base + nxt * sb, d1, sb);
        memcpy(d1, d2, sb);
        nxt = orc_next(nxt, a, B);
      }
    }
  }
  free(flags); free(d1); free(d2);
  return 0;
}

/* 100 is the single-instance 0100 (src/cl_marshal.cc:779-782). */
int orc_inplace_100_pttwac(void *buf, size_t es, i64 A, i64 B, i64 b) {
  return orc_inplace_0100_pttwac(buf, es, 1, A, B, b);
}

/* Route the composer took, written as the reference's stderr tags. */
static void route_put(char *route, size_t cap, const char *tag) {
  if (!route || !cap) return;
  size_t l = strlen(route);
  snprintf(route + l, cap > l ? cap - l : 0, "%s", tag);
}

/* 010 step of a composed route: LDS-staged when the tile fits `max_mem` elements, else
 * cycle following (src/cl_marshal.cc:578-585, 615-622, 653-660, 696-703). */
static int step_010(void *buf, size_t es, i64 A, i64 a, i64 B, i64 max_mem, int R,
                    char *route, size_t cap) {
  if ((i64)a * B <= max_mem) {
    route_put(route, cap, "010_BS-");
    return orc_inplace_010_bs(buf, es, A, a, B);
  }
  route_put(route, cap, "010_PTTWAC-");
  return orc_inplace_010_pttwac(buf, es, A, a, B, R, 1);
}

/* Full in-place transpose of (A*a) x (B*b) into (B*b) x (A*a) by the stage sequence
 * `stages`. Restates cl_transpose, src/cl_marshal.cc:561-776, including its feasibility
 * tests written in elements against `max_mem` (the reference's MAX_MEM,
 * src/cl_marshal.cc:45-59) and the element-granular fallback (769-775). `sp` selects the
 * single-precision constants of those tests (64 vs 32). The 4-stage feasibility uses the
 * plan.hpp classes (include/plan.hpp:95-157) with `lds_bytes` as the local-memory size.
 * Returns 0 on success. */
int orc_inplace_transpose(void *buf, size_t es, i64 A, i64 a, i64 B, i64 b, int R,
                          int stages, i64 max_mem, int sp, i64 lds_bytes, char *route,
                          size_t cap) {
  if (route && cap) route[0] = 0;
  i64 slack = sp ? 64 : 32;
  int r = 0;
  if (stages == 2) {
    i64 f = ((a * B * b + 31) / 32);
    if (f + ((f >> 5) * 1) <= max_mem && a < (max_mem - slack) / 2) {
      if (a > 1) r = step_010(buf, es, A, a, B * b, max_mem, R, route, cap);
      if (r) return r;
      r = orc_inplace_100_pttwac(buf, es, A, B * b, a);
      route_put(route, cap, "2.1-");
      return r;
    }
  }
  if (stages == 22) {
    i64 f = ((b * A * a + 31) / 32);
    if (f + ((f >> 5) * 1) <= max_mem && b < (max_mem - slack) / 2) {
      r = orc_inplace_100_pttwac(buf, es, A * a, B, b);
      if (r) return r;
      if (b > 1) r = step_010(buf, es, B, A * a, b, max_mem, R, route, cap);
      route_put(route, cap, "2.2-");
      return r;
    }
  }
  if (stages == 3 || stages == 32) {
    i64 f = ((a * b + 31) / 32);
    if (f + ((f >> 5) * 1) <= max_mem && b < (max_mem - slack) / 2 &&
        a < (max_mem - slack) / 2 && a > 1 && b > 1) {
      if (stages == 3) {
        /* AaBb -> BAab -> BAba -> BbAa (src/cl_marshal.cc:636-677) */
        r = orc_inplace_100_pttwac(buf, es, A * a, B, b);
        if (!r) r = step_010(buf, es, B * A, a, b, max_mem, R, route, cap);
        if (!r) r = orc_inplace_0100_pttwac(buf, es, B, A, b, a);
        route_put(route, cap, "3.1-");
      } else {
        /* AaBb -> ABab -> ABba -> BbAa (src/cl_marshal.cc:679-720) */
        r = orc_inplace_0100_pttwac(buf, es, A, a, B, b);
        if (!r) r = step_010(buf, es, A * B, a, b, max_mem, R, route, cap);
        if (!r) r = orc_inplace_100_pttwac(buf, es, A, B * b, a);
        route_put(route, cap, "3.2-");
      }
      return r;
    }
  }
  if (stages == 4) {
    /* Karlsson's method (src/cl_marshal.cc:722-767) with plan.hpp feasibility. */
    i64 lds_words = lds_bytes / 4;
    int f1 = (b * 6 * 4 * 2 + 24) < lds_bytes - 512;                  /* T0100(A,a,B,b) */
    int f2 = a * b < lds_words - 512;                                  /* T010_BS(a,b) */
    int f2p = (a * b / 16) < lds_words;                                /* T010_PTTWAC */
    int f3 = (b * a * 6 * 4 * 2 + 24) < lds_bytes - 512;              /* T0100(1,A,B,ba) */
    int f4 = (a * 6 * 4 * 2 + 24) < lds_bytes - 512;                  /* T0100(B,A,b,a) */
    if (f1 && (f2 || f2p) && f3 && f4) {
      r = orc_inplace_0100_pttwac(buf, es, A, a, B, b);
      if (!r) {
        if (f2) { route_put(route, cap, "010_BS-"); r = orc_inplace_010_bs(buf, es, B * A, a, b); }
        else { route_put(route, cap, "010_PTTWAC-"); r = orc_inplace_010_pttwac(buf, es, B * A, a, b, R, 1); }
      }
      if (!r) r = orc_inplace_100_pttwac(buf, es, A, B, b * a);
      if (!r) r = orc_inplace_0100_pttwac(buf, es, B, A, b, a);
      route_put(route, cap, "4-");
      return r;
    }
  }
  route_put(route, cap, "Fallback-");
  return orc_inplace_0100_pttwac(buf, es, 1, A * a, B * b, 1);
}

/* In-place row padding [y][x] -> [y][pad] with zeroed new columns. The reference moves rows
 * from the tail towards the front (src/cl/cl_aos_asta.cl:813-870); sequentially, walking
 * the destination from the end is the same schedule. Buffer holds y*pad elements. */
void orc_padding(void *buf, size_t es, i64 x, i64 y, i64 pad) {
  char *m = (char *)buf;
  for (i64 p = y * pad - 1; p >= 0; p--) {
    i64 r = p / pad, c = p % pad;
    if (c < x) memmove(m + (u64)p * es, m + ((u64)r * x + c) * es, es);
    else memset(m + (u64)p * es, 0, es);
  }
}

/* In-place row unpadding [y][pad] -> [y][x] with the freed tail [y*x, y*pad) zeroed.
 * Restates unpadding, src/cl/cl_aos_asta.cl:881-938 (front-to-back schedule). */
void orc_unpadding(void *buf, size_t es, i64 x, i64 y, i64 pad) {
  char *m = (char *)buf;
  for (i64 p = 0; p < y * x; p++) {
    i64 r = p / x, c = p % x;
    memmove(m + (u64)p * es, m + ((u64)r * pad + c) * es, es);
  }
  memset(m + (u64)y * x * es, 0, (u64)(y * pad - y * x) * es);
}

/* ------------------------------------------------------------------------------------ */
/* Tile-factor planning (host logic)                                                      */
/* ------------------------------------------------------------------------------------ */

/* All divisors t > 1 of n (n itself included), ascending. Restates the set produced by
 * Factorize::tiling_options, include/plan.hpp:163-219 (the reference enumerates them in
 * recursion order; its only caller sorts them, test/cl_unittest.cc:427-430).
 * Returns the count; writes at most cap values. */
int orc_tile_sizes(i64 n, i64 *out, int cap) {
  int cnt = 0;
  if (n < 2) return 0;
  /* ascending: small divisors, then the paired large ones in reverse */
  i64 small[4096]; int ns = 0;
  i64 large[4096]; int nl = 0;
  for (i64 d = 1; d * d <= n; d++) {
    if (n % d) continue;
    if (d > 1 && ns < 4096) small[ns++] = d;
    if (d * d != n && nl < 4096) large[nl++] = n / d;
  }
  for (int i = 0; i < ns; i++) { if (cnt < cap) out[cnt] = small[i]; cnt++; }
  for (int i = nl - 1; i >= 0; i--) { if (cnt < cap) out[cnt] = large[i]; cnt++; }
  return cnt;
}

/* Super-element sizes for a full transpose, "Algorithm 5" of the reference's test
 * (test/cl_unittest.cc:396-467): the largest divisor of h (resp. w) in [min_limit,
 * max_limit]; while none exists the dimension is grown by one row (column) of padding.
 * Outputs aa, bb and the paddings; returns 0. */
int orc_choose_tiles(i64 h, i64 w, i64 min_limit, i64 max_limit, i64 *aa, i64 *bb,
                     i64 *pad_h, i64 *pad_w) {
  i64 ph = 0, pw = 0, A_ = -1, B_ = -1;
  static i64 opts[1 << 15];
  while (A_ < 0 || B_ < 0) {
    if (A_ < 0) {
      int c = orc_tile_sizes(h + ph, opts, 1 << 15);
      if (c > (1 << 15)) c = 1 << 15;
      for (int j = c - 1; j >= 0; j--)
        if (opts[j] >= min_limit && opts[j] <= max_limit) { A_ = opts[j]; break; }
    }
    if (B_ < 0) {
      int c = orc_tile_sizes(w + pw, opts, 1 << 15);
      if (c > (1 << 15)) c = 1 << 15;
      for (int j = c - 1; j >= 0; j--)
        if (opts[j] >= min_limit && opts[j] <= max_limit) { B_ = opts[j]; break; }
    }
    if (A_ < 0) ph++;
    if (B_ < 0) pw++;
  }
  *aa = A_; *bb = B_; *pad_h = ph; *pad_w = pw;
  return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Multi-threaded restatements (CPU baseline)                                             */
/* ------------------------------------------------------------------------------------ */

static int nthreads_or_all(int nt) {
  if (nt > 0) return nt;
  long c = sysconf(_SC_NPROCESSORS_ONLN);
  return c > 0 ? (int)c : 1;
}

typedef struct {
  char *buf; size_t es; i64 A, a, B; int tid, nt; int err;
} bs_job;

static void *bs_worker(void *p) {
  bs_job *j = (bs_job *)p;
  u64 n = (u64)j->a * j->B;
  char *store = (char *)malloc(n * j->es != 0 ? n * j->es : 1);
  if (!store) { j->err = -1; return NULL; }
  for (i64 t = j->tid; t < j->A; t += j->nt) {
    char *in = j->buf + (u64)t * n * j->es;
    for (u64 i = 0; i < n; i++)
      memcpy(store + orc_next(i, j->a, j->B) * j->es, in + i * j->es, j->es);
    memcpy(in, store, n * j->es);
  }
  free(store);
  return NULL;
}

/* BS_marshal (src/cl/cl_aos_asta.cl:53-84) with one host thread per work-group slot: tiles
 * are distributed round-robin over `nthreads` threads (0 = all online CPUs). */
int orc_par_010_bs(void *buf, size_t es, i64 A, i64 a, i64 B, int nthreads) {
  int nt = nthreads_or_all(nthreads);
  pthread_t th[512];
  bs_job jobs[512];
  if (nt > 512) nt = 512;
  for (int t = 0; t < nt; t++) {
    jobs[t] = (bs_job){(char *)buf, es, A, a, B, t, nt, 0};
    pthread_create(&th[t], NULL, bs_worker, &jobs[t]);
  }
  int err = 0;
  for (int t = 0; t < nt; t++) { pthread_join(th[t], NULL); if (jobs[t].err) err = jobs[t].err; }
  return err;
}

/* Relaxed atomic element copies: slots are read while other threads may write them, and a
 * value read that way is only ever used when this thread's claim on the slot succeeded. */
static inline void ld_rlx(void *dst, const char *src, size_t es) {
  switch (es) {
    case 1: { uint8_t v = __atomic_load_n((const uint8_t *)src, __ATOMIC_RELAXED); memcpy(dst, &v, 1); break; }
    case 2: { uint16_t v = __atomic_load_n((const uint16_t *)src, __ATOMIC_RELAXED); memcpy(dst, &v, 2); break; }
    case 4: { uint32_t v = __atomic_load_n((const uint32_t *)src, __ATOMIC_RELAXED); memcpy(dst, &v, 4); break; }
    default:
      for (size_t o = 0; o < es; o += 8) {
        uint64_t v = __atomic_load_n((const uint64_t *)(src + o), __ATOMIC_RELAXED);
        memcpy((char *)dst + o, &v, 8);
      }
  }
}
static inline void st_rlx(char *dst, const void *src, size_t es) {
  switch (es) {
    case 1: { uint8_t v; memcpy(&v, src, 1); __atomic_store_n((uint8_t *)dst, v, __ATOMIC_RELAXED); break; }
    case 2: { uint16_t v; memcpy(&v, src, 2); __atomic_store_n((uint16_t *)dst, v, __ATOMIC_RELAXED); break; }
    case 4: { uint32_t v; memcpy(&v, src, 4); __atomic_store_n((uint32_t *)dst, v, __ATOMIC_RELAXED); break; }
    default:
      for (size_t o = 0; o < es; o += 8) {
        uint64_t v; memcpy(&v, (const char *)src + o, 8);
        __atomic_store_n((uint64_t *)(dst + o), v, __ATOMIC_RELAXED);
      }
  }
}

typedef struct {
  char *buf; size_t es; u64 m, n, sb; uint32_t *flags; u64 start, stride; int err;
} pt_job;

/* The reference's racy claim protocol (src/cl/cl_aos_asta.cl:475-566): read the start
 * super-element, test its bit, then follow the cycle claiming destinations with an atomic
 * fetch-or; a value is written only into a slot whose bit this thread set. Super-elements
 * are moved in 8-byte (or element-size) relaxed atomic pieces. */
static void *pt_worker(void *p) {
  pt_job *j = (pt_job *)p;
  u64 last = j->m * j->n - 1;
  size_t sb = j->sb;
  char *d1 = (char *)malloc(sb), *d2 = (char *)malloc(sb);
  if (!d1 || !d2) { free(d1); free(d2); j->err = -1; return NULL; }
  size_t piece = (sb % 8 == 0) ? 8 : (j->es < 8 ? j->es : 8);
  for (u64 s = j->start; s < last; s += j->stride) {
    u64 nxt = orc_next(s, j->m, j->n);
    if (nxt == s) continue;
    for (size_t o = 0; o < sb; o += piece) ld_rlx(d1 + o, j->buf + s * sb + o, piece);
    uint32_t f = __atomic_load_n(&j->flags[s >> 5], __ATOMIC_ACQUIRE);
    int done = (f >> (s & 31)) & 1u;
    while (!done) {
      for (size_t o = 0; o < sb; o += piece) ld_rlx(d2 + o, j->buf + nxt * sb + o, piece);
      uint32_t mask = 1u << (nxt & 31);
      uint32_t old = __atomic_fetch_or(&j->flags[nxt >> 5], mask, __ATOMIC_ACQ_REL);
      done = (old & mask) != 0;
      if (!done)
        for (size_t o = 0; o < sb; o += piece) st_rlx(j->buf + nxt * sb + o, d1 + o, piece);
      char *t = d1; d1 = d2; d2 = t;
      nxt = orc_next(nxt, j->m, j->n);
    }
  }
  free(d1); free(d2);
  return NULL;
}

/* 0100 by parallel cycle following on `nthreads` host threads (0 = all online CPUs). */
int orc_par_0100_pttwac(void *buf, size_t es, i64 A, i64 a, i64 B, i64 b, int nthreads) {
  int nt = nthreads_or_all(nthreads);
  if (nt > 512) nt = 512;
  u64 g = (u64)a * B, sb = (u64)b * es, words = (g + 31) / 32;
  uint32_t *flags = (uint32_t *)malloc((words ? words : 1) * 4);
  if (!flags) return -1;
  pthread_t th[512];
  pt_job jobs[512];
  int err = 0;
  for (i64 x = 0; x < A; x++) {
    memset(flags, 0, words * 4);
    for (int t = 0; t < nt; t++) {
      jobs[t] = (pt_job){(char *)buf + (u64)x * g * sb, es, (u64)a, (u64)B, sb, flags,
                         (u64)t, (u64)nt, 0};
      pthread_create(&th[t], NULL, pt_worker, &jobs[t]);
    }
    for (int t = 0; t < nt; t++) { pthread_join(th[t], NULL); if (jobs[t].err) err = jobs[t].err; }
  }
  free(flags);
  return err;
}

/* The 3-stage route of src/cl_marshal.cc:636-677 on `nthreads` host threads. */
int orc_par_transpose_3(void *buf, size_t es, i64 A, i64 a, i64 B, i64 b, int nthreads) {
  int r = orc_par_0100_pttwac(buf, es, 1, A * a, B, b, nthreads);
  if (!r) r = orc_par_010_bs(buf, es, B * A, a, b, nthreads);
  if (!r) r = orc_par_0100_pttwac(buf, es, B, A, b, a, nthreads);
  return r;
}

/* Number of online host CPUs. */
int orc_num_cpus(void) { return nthreads_or_all(0); }
