// batch.h - host-side planning of grouped work lists (shared by the batch and the server).
#pragma once

#include "kernels.h"
#include "../../include/hip_marshal.h"

#include <string>
#include <tuple>
#include <vector>

namespace lm {

// Kernel form + template instantiation of a group. kind: 0 LDS-staged 010, 1 010 cycle
// following with LDS bits, 2 100 with global bits, 3 0100 with LDS bits, 4-6 row/column
// passes; through-scratch rounds: 7 copy into the arena, 8 out-of-place transpose from it,
// 9 gather from it, 10 pitched transpose into it, 11 unpitch copy from it.
struct GroupKey {
  int kind, a, b, c;
  bool operator<(const GroupKey &o) const {
    return std::tie(kind, a, b, c) < std::tie(o.kind, o.a, o.b, o.c);
  }
};

// One persistent kernel launch: the work list of every stage of one form in one level.
struct Group {
  GroupKey key;
  std::vector<unsigned char> items;  // packed parameter blocks
  size_t item_size = 0;
  std::vector<uint64_t> prefix{0};   // prefix sum of work units per item
  std::vector<uint64_t> lut;         // item of every kFindStep-th unit (stored after prefix)
  size_t lds = 0;
  size_t items_off = 0, prefix_off = 0;  // offsets in the parameter buffer
  const void *func = nullptr;
  uint32_t grid = 0;
  uint32_t block = 256;
};

struct BatchPlan {
  std::vector<std::vector<Group>> levels;
  uint64_t flag_words = 0;  // visited bits of every global-bit stage
  size_t param_bytes = 0;   // size of the parameter buffer
  int64_t stages = 0;
  double bytes = 0;         // algorithmic bytes of one run (2 x every matrix)
  std::vector<std::string> routes;
  // Matrices whose best route is the stripe route (no grouped form): their plain stage lists,
  // dealt over kSoloLanes side streams (balanced by bytes) and run one after another on each,
  // overlapping the grouped kernels and each other.
  std::vector<std::vector<StageRec>> solo;
  std::vector<int> solo_lane;
  std::vector<double> solo_bytes;
  size_t solo_scratch[8] = {};  // per lane: the most any of its lists needs
  int solo_lanes = 0;           // lanes in use
  // Through-scratch rounds: matrices copied into a batch-owned arena by one grouped copy, then
  // moved back out of place by grouped transpose / gather kernels (their levels are the last
  // ones; item pointers into the arena hold arena offsets until build_image).
  uint64_t arena_bytes = 0;
  int arena_rounds = 0;
  int64_t arena_matrices = 0;
};
constexpr int kSoloLanes = 8;  // at most; the "batch_solo_lanes" knob picks how many

constexpr int kLevelLanes = 8;  // at most; the "batch_level_lanes" knob picks how many

// Side streams, events and scratch for the solo lists of a batch, and the branch streams on
// which the independent groups of one level run side by side (joined before the next level).
struct SoloExec {
  hipStream_t st[kSoloLanes] = {};
  hipEvent_t join[kSoloLanes] = {};
  unsigned char *scr[kSoloLanes] = {};
  hipEvent_t fork = nullptr;
  int level_lanes = 0;  // branch streams created (0: every group on the run's stream)
  hipStream_t lst[kLevelLanes] = {};
  hipEvent_t lfork = nullptr, ljoin[kLevelLanes] = {};
};

// allow_solo: matrices best served by the stripe route go to `solo` (the batched form);
// otherwise every matrix is grouped (the server runs large requests on its own).
bool plan_batch(const hip_marshal_desc *descs, int64_t count, BatchPlan &b, const char *who,
                bool allow_solo = false);
// Stage list of one descriptor; allow_st admits the stripe route (plain execution only).
bool plan_desc(const hip_marshal_desc &d, std::vector<StageRec> &out, std::string &route,
               bool allow_st);
// Host image of the parameter buffer, with the flag offsets turned into pointers into `flags`
// and the arena offsets into pointers into `arena`.
void build_image(const BatchPlan &b, const uint32_t *flags, std::vector<unsigned char> &host,
                 unsigned char *arena = nullptr);
// Clears the flags and launches every group, level by level, on `st`. Solo lists (if any)
// run on the side streams of `solo` (ordered after the work already on `st`, which waits for
// all of them at the end).
bool enqueue_batch(const BatchPlan &b, void *params, uint32_t *flags, hipStream_t st,
                   const SoloExec *solo = nullptr);

}  // namespace lm
