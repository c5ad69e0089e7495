// kernels.h - kernel parameter blocks and host launchers shared by the library's sources.
#pragma once

#include "common.h"

#include <array>
#include <string>
#include <vector>

namespace lm {

// ---- 010, LDS-staged ------------------------------------------------------------------
struct Bs010Params {
  unsigned char *base;
  uint64_t A;                       // tiles
  uint32_t a, B, Bp, n, g;          // tile rows/cols, LDS row pitch, a*B, tiles per chunk
  uint64_t nchunks;
  FastDiv32 div_n, div_B, div_a;
  uint32_t din_t, din_i, din_j;     // per-iteration advance in (tile, row, col) radix
  uint32_t dout_t, dout_j, dout_i;  // per-iteration advance in (tile, col, row) radix
  // G > 1, rows at pitch Bp: bit 0 stage-in by rows (B % G == 0), bit 1 gather by columns
  // (a % G == 0). The unpadded layouts are template parameters (Bs010Plan::layout).
  uint32_t aligned;
};

struct Bs010Plan {
  int G;              // elements per lane access
  int upt;            // units per thread per round (4, 8 or 16)
  uint64_t g;         // tiles per chunk
  uint32_t Bp;        // LDS row pitch (elements)
  int layout = 0;     // dword units: 0 rows at pitch Bp, 1 the chunk unpadded in input order,
                      // 2 unpadded in output order
  size_t lds_bytes;   // dynamic LDS per workgroup
  uint64_t nchunks;
};

uint32_t bs010_pitch(uint32_t B);
// Chooses the chunking of an LDS-staged 010; false when one tile does not fit `lds_budget`.
bool plan_010_bs(size_t es, uint64_t A, uint64_t a, uint64_t B, size_t lds_budget,
                 uint64_t base_addr, int upt, Bs010Plan *out);
// Kernel parameters of one LDS-staged 010 over `base` with plan `pl`.
Bs010Params make_bs010_params(unsigned char *base, uint64_t A, uint64_t a, uint64_t B,
                              const Bs010Plan &pl);
hipError_t launch_010_bs(unsigned char *base, size_t es, uint64_t A, uint64_t a, uint64_t B,
                         const Bs010Plan &pl, uint32_t max_blocks, hipStream_t st);

// ---- 010, cycle following with LDS bits -------------------------------------------------
struct Pt010Params {
  unsigned char *base;
  uint64_t A;
  uint32_t a, B, n, sh;
  FastDiv32 div_B, div_sh;
};
uint64_t pt010_flag_words(uint64_t a, uint64_t B, int R, int P);
// Picks its form: leader-only (every cycle short, "pt010_lead_order"), else visited bits with
// one tile per wave ("pt010_wave") or per workgroup.
hipError_t launch_010_pttwac(unsigned char *base, size_t es, uint64_t A, uint64_t a,
                             uint64_t B, int R, int P, uint32_t max_blocks, hipStream_t st);

// ---- 100 / 0100, cycle following with global bits ---------------------------------------
struct Pt100Params {
  unsigned char *base;
  uint32_t *flags;
  uint64_t R, C, RC;      // super-element grid of one instance: R rows x C columns
  uint64_t S;             // super-element bytes
  uint64_t nsub;          // pieces per super-element (each piece has its own bit)
  uint64_t sub_bytes;     // bytes per piece (multiple of the unit; the last may be shorter)
  uint64_t walkers;       // instances * RC * nsub
  uint32_t vw_log2;       // lanes per walker = 1 << vw_log2 (<= 64)
  uint32_t rot_r, rot_c;  // rotations folded into the map (0 for a plain 100; see next_slot)
  FastDiv64 div_C64, div_RC, div_nsub;
  FastDiv32 div_C32;
};

struct Pt100Plan {
  int ub;                 // unit bytes (1,2,4,8,16)
  int upl;                // units per lane (1,2,4,8)
  uint32_t vw_log2;
  uint64_t nsub, sub_bytes;
  uint64_t flag_words;    // 32-bit words of visited bits
};
// vw_max > 0 overrides the "pt100_vw_max" knob (lanes per walker; pieces of up to 8 units of
// 16 bytes per lane).
// base_addr (0: unknown, taken as aligned): units never exceed the alignment of the buffer.
void plan_100(size_t es, uint64_t inst, uint64_t R, uint64_t C, uint64_t b, Pt100Plan *out,
              int64_t vw_max = 0, uint64_t base_addr = 0);
// rot_r / rot_c (< R / < C): the slot in row i, column j moves to row (j + rot_c) mod C,
// column (i + rot_r) mod R of the transposed grid (a plain 100 has both 0).
hipError_t launch_100_pttwac(unsigned char *base, uint32_t *flags, uint64_t inst, uint64_t R,
                             uint64_t C, uint64_t S, const Pt100Plan &pl, uint32_t max_blocks,
                             hipStream_t st, uint32_t rot_r = 0, uint32_t rot_c = 0);

// ---- 100 / 0100 by cycle leaders: no visited bits ----------------------------------------
// When the multiplicative order k of R modulo R*C - 1 is small, every cycle of the 100 map
// has a length dividing k. Each cycle's smallest slot (its leader) is found on the device by
// walking at most k steps from every slot; then one walker per (instance, leader, piece)
// moves the whole cycle with no claims. `cycles` (of length >= 2, per instance) comes from
// gcds on the host, which also sizes the leader list.
struct PlPlan {
  uint64_t cycles = 0;
  uint32_t order = 0;
};
// False unless R, C >= 2, R*C < 2^32 and the order is at most max_order.
bool plan_100_leaders(uint64_t R, uint64_t C, uint32_t max_order, PlPlan *out);
// The device list of cycle leaders of the R x C grid's map (pc from plan_100_leaders), kept
// per (device, R, C) and ordered after its search on `st`; null when it cannot be had now
// (allocation failed, cache full, or first use under stream capture).
const uint32_t *cached_leaders(uint64_t R, uint64_t C, const PlPlan &pc, hipStream_t st);
// Padding entry of an ordered leader list: the lane that draws it has no cycle in that round.
constexpr uint32_t kNoLeader = 0xffffffffu;
// One start slot per cycle (of length >= 2) of the 010 tile map s -> (s mod B) * a + s / B of
// an a x B tile, ordered for the leader-only 010 with es-byte elements: groups of 64 (one wave
// instruction each) whose lanes touch few distinct 64-byte sectors at every step of their
// walks, padded with kNoLeader to a whole number of groups. `cycles` is the host's count (a
// list that disagrees is not returned); empty when the tile is too large to order.
std::vector<uint32_t> order_010_leaders(uint32_t a, uint32_t B, size_t es, uint64_t cycles);
// That list on the device, kept per (device, a, B, es) like cached_leaders; *entries gets its
// length (a multiple of 64). Null when it cannot be had now.
const uint32_t *cached_ordered_leaders(uint64_t a, uint64_t B, size_t es, const PlPlan &pc,
                                       hipStream_t st, uint32_t *entries);
// scratch: cycles + 1 32-bit words (the leader list and its counter).
hipError_t launch_100_leaders(unsigned char *base, uint32_t *scratch, uint64_t inst,
                              uint64_t R, uint64_t C, uint64_t S, const Pt100Plan &pl,
                              const PlPlan &pc, uint32_t max_blocks, hipStream_t st);
// The two halves: the search (into `list`, cycles + 1 words) and the walk over a list.
hipError_t launch_find_leaders(uint32_t *list, uint64_t R, uint64_t C, const PlPlan &pc,
                               hipStream_t st);
hipError_t launch_leader_walk(unsigned char *base, const uint32_t *list, uint64_t inst,
                              uint64_t R, uint64_t C, uint64_t S, const Pt100Plan &pl,
                              const PlPlan &pc, uint32_t max_blocks, hipStream_t st);

// 0100 with the visited bits of one instance in LDS (one workgroup per instance).
size_t lds0100_bytes(uint64_t R, uint64_t C, const Pt100Plan &pl);
hipError_t launch_0100_lds(unsigned char *base, uint64_t inst, uint64_t R, uint64_t C,
                           uint64_t S, const Pt100Plan &pl, uint32_t max_blocks,
                           hipStream_t st);

// ---- full transpose by row / column shuffles (LDS-staged passes) -------------------------
struct RcPlan {
  int ub;             // unit bytes
  uint32_t m;         // units per element
  uint32_t rows;      // rows per workgroup iteration of the row pass
  uint32_t w;         // columns per block of a column pass
  uint32_t c;         // gcd(M, N)
  uint64_t inst;      // independent M x N matrices stored back to back
  size_t row_lds, col_lds;
};
struct RcParams {
  unsigned char *base;
  uint64_t inst, ncb;  // instances; column blocks over all instances (inst * ncolblk)
  FastDiv32 div_M32;   // local row of a global row index
  FastDiv64 div_ncolblk;
  uint32_t M, N, c, a, b, ainv, m, rows, w, ncolblk;
  FastDiv32 div_c, div_b, div_m, div_upr, div_pitch, div_seg_last;
  FastDiv64 div_b64, div_M64;
};
// False when M or N is < 2 or >= 2^31, inst * M >= 2^31, or a row or a one-element-wide
// column of S-byte elements does not fit `lds_budget`. `inst` matrices of M x N are
// transposed independently (each in its own place).
bool plan_rc(size_t es, uint64_t M, uint64_t N, uint64_t S, size_t lds_budget, RcPlan *out,
             uint64_t inst = 1);
RcParams make_rc(unsigned char *base, uint64_t M, uint64_t N, const RcPlan &pl);
// pass: 0 = row shuffle, 1 = column rotation, 2 = column shuffle.
hipError_t launch_rc(unsigned char *base, uint64_t M, uint64_t N, const RcPlan &pl, int pass,
                     uint32_t max_blocks, hipStream_t st);

// ---- utilities --------------------------------------------------------------------------
hipError_t launch_fill_u32(uint32_t *p, uint64_t words, uint32_t value, hipStream_t st);
// The same fill as a kernel node of `graph` after `deps` (for graphs built node by node).
hipError_t add_fill_u32_node(hipGraph_t graph, const hipGraphNode_t *deps, size_t ndeps,
                             uint32_t *p, uint64_t words, uint32_t value, hipGraphNode_t *node);
// Byte copy (16-byte vectors when both ends and the size allow) and a pitched 2-D copy.
hipError_t launch_copy(unsigned char *dst, const unsigned char *src, uint64_t bytes,
                       hipStream_t st);
hipError_t launch_copy2d(unsigned char *dst, uint64_t dpitch, const unsigned char *src,
                         uint64_t spitch, uint64_t width, uint64_t rows, hipStream_t st);
// Copy of `rows` rows of RB bytes (RB >= 16) from src at pitch P (a multiple of 16, src
// 16-byte aligned) to dst with the rows back to back; dst is written in aligned 16-byte
// chunks whatever its alignment.
hipError_t launch_unpitch(unsigned char *dst, const unsigned char *src, uint64_t RB,
                          uint64_t P, uint64_t rows, hipStream_t st, int nt_store = -1);
// Parameters of one out-of-place transpose launch (k_oop_transpose) and of one out-of-place
// super-element gather (k_oop_gather); both kernels take a block index, so a grouped kernel
// can run many such launches as one (every matrix's blocks in one work list).
struct OopParams {
  const unsigned char *src;         // 16-byte aligned base; the first block starts at src_off
  unsigned char *dst;
  uint64_t src_off;
  uint64_t src_stride, dst_stride;  // bytes between consecutive blocks
  uint64_t src_pitch, dst_pitch;    // bytes between rows (source rows of C, destination of R)
  uint64_t src_end;                 // bytes readable from src (bounds the window over-read)
  uint64_t R, C;                    // each block R x C (source) -> C x R (destination)
  uint32_t tr_log2, tc_log2;        // tile TR x TC elements, both powers of two
  uint32_t PB;                      // LDS row pitch in bytes (multiple of 128)
  uint32_t wstore;                  // 4-/8-byte elements: aligned vector stores on misaligned rows
  uint64_t ntr, ntc, per_block, tiles, per_xcd;
  // skewed form: workgroups from per_xcd * 8 on write the chunks holding row boundaries
  uint64_t fix_rows;                // rows of all blocks (one more chunk: the end)
  FastDiv64 div_C;
  uint32_t nt;                      // bit 0: non-temporal source loads; bit 1: stores
  uint32_t order;                   // tile order in a block (see oop_tile_rc)
};
struct GatherParams {
  const unsigned char *src;  // 16-byte aligned
  unsigned char *dst;
  uint64_t inst_bytes;       // R * C * S
  uint64_t R, C;
  uint32_t S, TR, TC, LP;    // LP: LDS row pitch (multiple of 16)
  uint32_t nch, nchd;        // 16-byte chunks per LDS row / per destination segment
  uint32_t fullcol;          // TR == R: one destination segment per tile (kernel form FC)
  FastDiv32 div_nch, div_nchd, div_S, div_RS;
  uint64_t ntc, per_inst, tiles, per_xcd;
};
struct OopLaunch {
  OopParams p{};
  size_t lds = 0;
  uint64_t grid = 0;  // workgroups of the single launch (tiles, plus fix-up units when out = 2)
  int es = 0;
  bool vin = false;
  int out = 0;
};
struct GatherLaunch {
  GatherParams p{};
  size_t lds = 0;
  uint64_t grid = 0;
  int u = 1;
  bool fc = false;
};
// One contiguous byte copy of a grouped copy kernel (dst and src 16-byte aligned).
struct CopyItem {
  unsigned char *dst;
  const unsigned char *src;
  uint64_t bytes;
};
constexpr uint64_t kCopyItemUnit = 16384;  // bytes per work unit of the grouped copy
// Parameters, template and grid of the launch launch_oop_transpose / launch_oop_gather would
// make (grid 0: nothing to do).
hipError_t prepare_oop_transpose(const unsigned char *src, unsigned char *dst, size_t es,
                                 uint64_t blocks, uint64_t R, uint64_t C, uint64_t src_stride,
                                 uint64_t dst_stride, size_t tile_bytes, uint64_t src_end,
                                 uint64_t dst_pitch, int nt, OopLaunch *o);
hipError_t prepare_oop_gather(const unsigned char *src, unsigned char *dst, uint64_t S,
                              uint64_t inst, uint64_t R, uint64_t C, size_t tile_bytes,
                              GatherLaunch *o);
// Grouped kernels (items, prefix of units per item, item count, total units): units of an
// item are its launch's blocks, padded to a multiple of 8 so that with a grid that is a
// multiple of 8 a unit keeps its XCD.
const void *kg_oop_transpose_func(int es, bool vin, int out);
const void *kg_oop_gather_func(int u, bool fc);
const void *kg_copy_func();
constexpr int kOopGroupThreads = 512;
// One pitched-to-contiguous copy of a grouped unpitch kernel (see launch_unpitch).
struct UnpitchItem {
  unsigned char *dst;
  const unsigned char *src;  // 16-byte aligned, rows of RB bytes at pitch P
  uint64_t RB, P, total, nb;
  uint32_t head;
  FastDiv64 div_rb;
};
constexpr uint64_t kUnpitchItemUnit = 1024;  // 16-byte destination chunks per work unit
UnpitchItem make_unpitch_item(unsigned char *dst, const unsigned char *src, uint64_t RB,
                              uint64_t P, uint64_t rows);
const void *kg_unpitch_func(bool nts);
// Scratch row pitch of a through-scratch step over [inst][R][C] S-byte units of `buf` when
// it takes the pitched order (transpose into 16-byte-aligned scratch rows, then one
// contiguous copy back); 0 for the plain order (copy, then transpose back).
uint64_t oop_pitch_for(const unsigned char *buf, uint64_t inst, uint64_t R, uint64_t C,
                       uint64_t S);

// Out-of-place transpose of `blocks` R x C matrices (src and dst must not overlap), tiles of
// at most `tile_bytes` staged in LDS. src_end (0: just the blocks) is how many bytes from
// src may be read: 16-byte window loads of misaligned rows read up to 31 bytes past a row.
// dst_pitch (0: R * es) is the byte distance between destination rows; nt (-1: the "oop_nt"
// knob) selects non-temporal loads (bit 0) and stores (bit 1).
hipError_t launch_oop_transpose(const unsigned char *src, unsigned char *dst, size_t es,
                                uint64_t blocks, uint64_t R, uint64_t C, uint64_t src_stride,
                                uint64_t dst_stride, size_t tile_bytes, hipStream_t st,
                                uint64_t src_end = 0, uint64_t dst_pitch = 0, int nt = -1);

// Out-of-place 100 over super-elements of any size: `inst` blocks of R x C super-elements of
// S bytes, [inst][R][C] in src -> [inst][C][R] in dst (no overlap). src must be 16-byte
// aligned; dst any address. Tiles of about `tile_bytes` are staged in LDS; the destination is
// assembled in units of the largest power of two (up to 16) dividing S and dst.
bool oop_gather_fits(uint64_t S);
hipError_t launch_oop_gather(const unsigned char *src, unsigned char *dst, uint64_t S,
                             uint64_t inst, uint64_t R, uint64_t C, size_t tile_bytes,
                             hipStream_t st);

// ---- square transpose: one pass of tile-pair swaps ----------------------------------------
bool sq_fits(size_t es, uint64_t n);
hipError_t launch_sq_transpose(unsigned char *base, size_t es, uint64_t n, hipStream_t st);

// ---- stripe route (full transpose through bounded scratch) -------------------------------
struct StPlan {
  int variant;             // 0: stripes of h rows (M = q*h + r); 1: of h columns (N = q*h + r)
  size_t es;
  uint64_t M, N, h, q, r;
  uint64_t group;          // stripes per scratch fill
  bool rotate;             // stripes written one slot down, rotation folded into the 100
  Pt100Plan pt;            // the 100 over h-element super-elements (when q > 1)
  size_t stripe_off, b_off, f_off, scratch_bytes;
  uint64_t flag_words;
};
// False when no stripe height keeps one stripe within `scratch_cap` bytes or the striped
// dimension holds fewer than two stripes.
// rotate: when the stripes need more than one scratch fill, each stripe is transposed into
// its neighbour's slot (no copy pass) and the 100 folds the one-slot rotation into its map;
// otherwise (or with rotate false) stripes go through scratch in groups.
bool plan_st(size_t es, uint64_t M, uint64_t N, int variant, uint64_t target_bytes,
             uint64_t scratch_cap, int64_t pt_vw, bool rotate, StPlan *out,
             uint64_t base_addr = 0);
// `scratch` must hold plan.scratch_bytes and not overlap `buf`.
hipError_t run_st(unsigned char *buf, const StPlan &p, unsigned char *scratch, size_t tile_bytes,
                  uint32_t pt_blocks, hipStream_t st);

// ---- padding / unpadding ----------------------------------------------------------------
// flags must hold pad_flag_words(...) words; they are initialised by the launcher.
uint64_t pad_flag_words(uint64_t total_elems, size_t es);
hipError_t launch_padding(unsigned char *base, size_t es, uint64_t x, uint64_t y, uint64_t pad,
                          bool unpad, uint32_t *flags, hipStream_t st);

// Run-time tuning knobs (launch shapes). Keys: "bs_upt", "bs_blocks_per_cu",
// "pt010_blocks_per_cu", "pt100_blocks_per_cu".
int64_t tuning(const char *key);

// ---- 100 / 0100 by cycle segments: one pass, no bits, no claims -------------------------
// Every cycle of the grid's 100 map is cut into segments of at most seg_len slots; a first
// launch saves each segment's start piece to a side buffer, a second walks the segments, each
// carrying its pieces one slot on. The table of segments (start, length) is built on the host
// once per (R, C, seg_len) and uploaded once per device (false under stream capture, or when
// R*C exceeds max_slots).
struct SegPlan {
  const uint32_t *table = nullptr;  // device copy: start slot and length of each segment
  uint64_t nseg = 0;                // segments per instance
  uint32_t seg_len = 0;
  int ub = 16, upl = 1;             // unit bytes, units per lane
  uint32_t vw = 1;                  // lanes per walker (1-64)
  uint64_t nsub = 1, sub_bytes = 0; // pieces per super-element and their size
};
// build = false: no table (for pricing a candidate); the segment count is estimated. A grid
// of more than sync_slots slots whose table does not exist yet gets it built on a background
// thread, and this call returns false (the caller plans another form).
bool plan_100_segments(uint64_t inst, uint64_t R, uint64_t C, uint64_t S, uint64_t base_addr,
                       uint32_t seg_len, uint64_t max_slots, uint64_t sync_slots, bool build,
                       SegPlan *out);
// While one of these is alive on a thread, segment tables of any size are built on that thread
// when first needed (planning that is not latency-critical, such as a batch's creation).
struct SegSyncBuild {
  SegSyncBuild();
  ~SegSyncBuild();
};
// Bytes of the side buffer (every segment's start piece, per instance).
size_t seg_save_bytes(uint64_t inst, uint64_t S, const SegPlan &sp);
hipError_t launch_100_segments(unsigned char *base, unsigned char *save, uint64_t inst,
                               uint64_t R, uint64_t C, uint64_t S, const SegPlan &sp,
                               uint32_t max_blocks, uint32_t nt, uint32_t xcd, hipStream_t st);
void free_segment_tables();

// ---- 100 / 0100 by workgroup clusters: one instance staged on chip ---------------------
// G workgroups (one CU each) hold an instance of R x C super-elements of S >= 64 bytes in LDS
// and registers, meet at a barrier and write it back transposed: one pass, in place, no bits.
// Barrier words: a buffer kept per stream (every launch leaves it zero); under stream capture
// the caller's `flags` (cluster_flag_words(inst) 32-bit words), cleared by the launch.
struct ClPlan {
  uint32_t G = 0;           // workgroups per instance (at most max_wgs)
  uint32_t chunk = 0;       // instance bytes per workgroup
  uint32_t ub = 16;         // piece bytes (16, 8 or 4)
  uint32_t lds_pieces = 0;  // pieces per workgroup staged in LDS
  size_t lds_bytes = 0;
  // fused form (fes > 0): a 010 over tiles of fa x fb elements of fes bytes, then the 0100
  // over the [R][fb] grid of fa-element super-elements, in one pass (LDS only)
  uint32_t fa = 0, fb = 0, fes = 0, flv = 0;
};
bool plan_100_cluster(uint64_t inst, uint64_t R, uint64_t C, uint64_t S, uint64_t base_addr,
                      uint64_t max_wgs, ClPlan *out);
// The fused form for `inst` instances of R tiles of a x b elements of es (4 or 8) bytes.
bool plan_cluster_fused(uint64_t inst, uint64_t R, uint64_t a, uint64_t b, size_t es,
                        uint64_t base_addr, uint64_t max_wgs, ClPlan *out);
uint64_t cluster_flag_words(uint64_t inst);
// Clusters of the cluster-staged form that gave up their bounded barrier wait on the current
// device since the last call (their instances were left unconverted); reads a device word,
// so the work in question must have finished. 0 when no cluster launch ran on the device.
uint64_t cluster_gave_up_new();
hipError_t launch_100_cluster(unsigned char *base, uint32_t *flags, uint64_t inst, uint64_t R,
                              uint64_t C, uint64_t S, const ClPlan &pl, hipStream_t st);
void free_cluster_flags();

// ---- stage records (a route as a list of elementary stages) -----------------------------
enum class M010 { Auto, BS, PTTWAC, Global };
enum class StageKind : int {
  BS010, PT010, PT100G, PT0100L, PT100C, RC_ROW, RC_COL1, RC_COL2, ST, SQ, OOP, SEG, CL
};
struct StageRec {
  StageKind kind;
  size_t es;
  unsigned char *buf;
  uint64_t d0, d1, d2, d3;  // 010: (A, a, B); 0100: (instances, R, C, b)
  int R, P;                 // 010 cycle following: bit spreading / padding
  Bs010Plan bs;
  Pt100Plan pt;
  RcPlan rc;                // RC_*: instances in d0, (M, N) in d1, d2, element bytes in d3
  StPlan stp;               // ST: the whole stripe route (M, N in d1, d2)
  uint64_t chunk;           // OOP: blocks per scratch fill (d0 blocks of d1 x d2, d3 bytes each)
  uint64_t pitch;           // OOP: scratch row pitch of the pitched form (0: plain form)
  PlPlan pc;                // PT100C: cycle count and order (pt: the walker shape)
  SegPlan sg;               // SEG: segment table and walker shape
  ClPlan cl;                // CL: cluster shape
  int64_t grid;             // BS010: workgroups per CU chosen for its class (0: the knob)
};
// Forms a measured class choice can name (ClassChoice::form; 0 = the cost model decides).
enum : int {
  kFormModel = 0,
  kForm010Bs = 1, kForm010Oop = 2, kForm010Split = 3, kForm010Rc = 4, kForm010Cyc = 5,
  kForm0100Cyc = 11, kForm0100Seg = 12, kForm0100Oop = 13, kForm0100Rc = 14,
  kForm0100Lds = 15, kForm0100Bits = 16, kForm0100Cl = 17,
};
struct ClassChoice {
  int form = kFormModel;
  int64_t param = 0;  // BS: workgroups per CU; SEG: slots per segment (0: the knob)
};
// While one of these is alive on a thread, planning on that thread uses only the forms whose
// scratch is visited bits or nothing (no side buffers, no device tables): the batch's grouped
// kernels and the no-alloc entry point plan this way.
struct RestrictedPlanning {
  RestrictedPlanning();
  ~RestrictedPlanning();
};
bool planning_restricted();
// While one of these is alive on a thread, the through-scratch and stripe forms planned on that
// thread use at most `bytes` of scratch (below the "st_scratch_mb" knob when that is lower).
struct ScratchCapScope {
  explicit ScratchCapScope(uint64_t bytes);
  ~ScratchCapScope();

 private:
  uint64_t prev_;
};
struct Choice {
  double cost = 1e300;
  uint64_t a = 1, b = 1;
  // 2, 22, 3, 32, 5 (rows/columns), 6/7 (stripes), 8 (square), 9 (through scratch),
  // -1 (element-granular)
  int stages = 0;
};
size_t current_lds_budget();
M010 choose_010(size_t es, uint64_t A, uint64_t a, uint64_t B, int R, int P,
                uint64_t base_addr);
// allow_oop admits the through-scratch form (plain execution only: it needs scratch of up to
// a scratch cap and has no grouped form for batches).
void push_0100(std::vector<StageRec> &out, unsigned char *buf, size_t es, uint64_t inst,
               uint64_t R, uint64_t C, uint64_t b, bool allow_oop = false,
               const char **tag = nullptr);
bool push_010(std::vector<StageRec> &out, unsigned char *buf, size_t es, uint64_t A,
              uint64_t a, uint64_t B, M010 m, int R, int P, const char **tag,
              bool allow_oop = false);
// One 0100 step called on its own: push_0100, except that a single-instance 100 whose
// super-elements are 1, 2, 4, 8 or 16 bytes (with the buffer aligned to that size) is
// planned as a full transpose of such elements. `tag` receives the route.
bool plan_step_0100(std::vector<StageRec> &out, unsigned char *buf, size_t es, uint64_t inst,
                    uint64_t R, uint64_t C, uint64_t b, bool allow_oop, std::string &tag);
double stage_bw_tbs(const StageRec &r);
bool route_stages(std::vector<StageRec> &out, unsigned char *buf, size_t es, uint64_t A,
                  uint64_t a, uint64_t B, uint64_t b, int R, int stages, std::string &route);
// allow_st = false leaves out the stripe route (it needs scratch beyond the visited bits and
// has no group-kernel form for batches).
bool plan_transpose_stages(std::vector<StageRec> &out, unsigned char *buf, size_t es,
                           uint64_t M, uint64_t N, std::string &route, bool allow_st = true);
Choice plan_full(size_t es, uint64_t M, uint64_t N, bool allow_st = true);
bool run_stages(const std::vector<StageRec> &stages, hipStream_t st);
// Same, with caller-owned scratch of at least stages_scratch_bytes(stages) bytes.
// alone: nothing else of this library runs on the device meanwhile (tunes launch shapes only).
bool run_stages_on(const std::vector<StageRec> &stages, hipStream_t st, unsigned char *scratch,
                   bool alone = false);
size_t stages_scratch_bytes(const std::vector<StageRec> &stages);
// Per-launch timing (hip_marshal_set_kernel_timing): a launcher opens one of these around its
// launch; `bytes` is the launch's algorithmic traffic (2 x the bytes it moves into place).
class KernelSpan {
 public:
  KernelSpan(hipStream_t st, const char *name, double bytes);
  ~KernelSpan();
  KernelSpan(const KernelSpan &) = delete;
  KernelSpan &operator=(const KernelSpan &) = delete;

 private:
  hipStream_t st_;
  int64_t idx_ = -1;
};

// Workgroups of a grid-stride launch: CUs x per_cu, clamped to [1, 2^24].
uint32_t grid_blocks(int64_t per_cu);

// Measured class choices (hip_marshal_autotune_desc) as the planner consults them.
using ClassKey = std::array<int64_t, 6>;
ClassKey class_key_010(size_t es, uint64_t A, uint64_t a, uint64_t B);
ClassKey class_key_0100(uint64_t inst, uint64_t R, uint64_t C, uint64_t S);
ClassChoice class_choice(const ClassKey &k);
bool class_tuned(const ClassKey &k);

// Sets what hip_marshal_last_route reports on this thread.
void set_last_route(const std::string &route);
// Appends the row / column decomposition of an M x N transpose of S-byte elements; false
// when it does not fit the LDS budget.
bool push_rc(std::vector<StageRec> &out, unsigned char *buf, size_t es, uint64_t M,
             uint64_t N, uint64_t S, uint64_t inst = 1);

// Device properties used for launch sizing.
struct DeviceInfo {
  int id = -1;
  int cus = 256;
  size_t lds_per_block = 160 * 1024;
  size_t hbm_bytes = 0;
};
const DeviceInfo &device_info();

}  // namespace lm
