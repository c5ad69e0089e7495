// kernels.h - kernel parameter blocks and host launchers shared by the library's sources.
#pragma once

#include "common.h"

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
};

struct Bs010Plan {
  int G;              // elements per lane access
  int upt;            // units per thread per round (4, 8 or 16)
  uint64_t g;         // tiles per chunk
  uint32_t Bp;        // LDS row pitch (elements)
  size_t lds_bytes;   // dynamic LDS per workgroup
  uint64_t nchunks;
};

uint32_t bs010_pitch(uint32_t B);
// Chooses the chunking of an LDS-staged 010; false when one tile does not fit `lds_budget`.
bool plan_010_bs(size_t es, uint64_t A, uint64_t a, uint64_t B, size_t lds_budget,
                 uint64_t base_addr, int upt, Bs010Plan *out);
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
void plan_100(size_t es, uint64_t inst, uint64_t R, uint64_t C, uint64_t b, Pt100Plan *out);
hipError_t launch_100_pttwac(unsigned char *base, uint32_t *flags, uint64_t inst, uint64_t R,
                             uint64_t C, uint64_t S, const Pt100Plan &pl, uint32_t max_blocks,
                             hipStream_t st);

// ---- utilities --------------------------------------------------------------------------
hipError_t launch_fill_u32(uint32_t *p, uint64_t words, uint32_t value, hipStream_t st);

// ---- padding / unpadding ----------------------------------------------------------------
// flags must hold pad_flag_words(...) words; they are initialised by the launcher.
uint64_t pad_flag_words(uint64_t total_elems);
hipError_t launch_padding(unsigned char *base, size_t es, uint64_t x, uint64_t y, uint64_t pad,
                          bool unpad, uint32_t *flags, hipStream_t st);

// Run-time tuning knobs (launch shapes). Keys: "bs_upt", "bs_blocks_per_cu",
// "pt010_blocks_per_cu", "pt100_blocks_per_cu".
int64_t tuning(const char *key);

// Device properties used for launch sizing.
struct DeviceInfo {
  int id = -1;
  int cus = 256;
  size_t lds_per_block = 160 * 1024;
  size_t hbm_bytes = 0;
};
const DeviceInfo &device_info();

}  // namespace lm
