// batch.hip - many layout conversions of uneven shapes and element sizes in one graph launch.
//
// A batch is planned once: every matrix's conversion becomes its route (a list of stage
// records, the same ones a plain call runs). Stage k of every matrix forms level k. Within a
// level, stages that use the same kernel form (and template instantiation) form a group,
// executed by ONE persistent kernel over a work list: each workgroup (or wave) claims work
// units (chunks, tiles, instances, walker sets) of any matrix in the group, looks up its
// matrix from a prefix sum of units (find_item), and runs the same device body as the
// single-matrix kernel. Levels are ordered; groups inside a level are independent and run
// side by side (on branch streams, or as parallel chains of graph nodes). The whole sequence -
// clearing every matrix's visited bits, then the group kernels level by level - is built once
// into a hipGraph, node by node (no stream capture), so a run is one graph launch.
//
// Three kinds of matrix share a run:
// - grouped in-place stages (the levels above);
// - through-scratch rounds: matrices of up to "batch_arena_max_mb" whose best route is not a
//   single streaming pass are copied (or, pitched, transposed) into a batch-owned arena by
//   one grouped launch per round, then moved back out of place by grouped launches; rounds
//   of at most "batch_arena_mb" follow one another on the run's stream;
// - side-stream matrices (stripe route, and from "batch_solo_kb" the square, cycle-leader,
//   cluster, segmented and larger through-scratch forms) run their plain stage lists on
//   side streams forked from the run's stream when the graph is launched and joined to it
//   after (the run's stream waits for both).
//
// No counterpart in the reference, which launches one kernel per stage per matrix and waits
// after each (src/cl_marshal.cc:561-776, include/cl_profile.h:23-24).
#include "batch.h"
#include "kernels_dev.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <tuple>
#include <vector>

namespace lm {

// ---- group kernels -------------------------------------------------------------------------

template <int ES, int G, int UPT, int LAYOUT>
__global__ __launch_bounds__(256) void kg_010_bs(const Bs010Params *__restrict__ items,
                                                 const uint64_t *__restrict__ prefix, uint32_t n,
                                                 uint64_t total) {
  using E = typename Elem<ES>::T;
  using U = typename UnitOf<ES, G>::T;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  E *lds = reinterpret_cast<E *>(lds_raw);
  uint64_t q = blockIdx.x;
  if (q >= total) return;
  uint32_t it = find_item(prefix, n, q);
  Bs010Params p = items[it];
  uint64_t c = q - prefix[it];
  U v[UPT];
  bs_load_round<ES, G, 256, UPT>(p, c, 0, v);
  while (true) {
    bs_stage_in<ES, G, 256, UPT, LAYOUT>(p, c, v, lds);
    __syncthreads();
    // Prefetch the next claimed chunk (a disjoint region, possibly of another matrix).
    const uint64_t qn = q + gridDim.x;
    Bs010Params pn = p;
    uint64_t cn = 0;
    if (qn < total) {
      uint32_t in = find_item(prefix, n, qn);
      pn = items[in];
      cn = qn - prefix[in];
      bs_load_round<ES, G, 256, UPT>(pn, cn, 0, v);
    }
    bs_gather_out<ES, G, 256, LAYOUT>(p, c, lds);
    __syncthreads();
    if (qn >= total) break;
    q = qn;
    p = pn;
    c = cn;
  }
}

template <int ES>
__global__ __launch_bounds__(256) void kg_010_pttwac(const Pt010Params *__restrict__ items,
                                                     const uint64_t *__restrict__ prefix, uint32_t n,
                                                     uint64_t total) {
  extern __shared__ __align__(16) uint32_t lflags[];
  for (uint64_t q = blockIdx.x; q < total; q += gridDim.x) {
    uint32_t it = find_item(prefix, n, q);
    const Pt010Params p = items[it];
    pt010_tile<ES, 256>(p, q - prefix[it], lflags);
  }
}

template <int UB, int UPL, bool I64>
__global__ __launch_bounds__(256) void kg_100(const Pt100Params *__restrict__ items, const uint64_t *__restrict__ prefix,
                                              uint32_t n, uint64_t tasks) {
  const uint64_t wave = uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const uint64_t nwaves = uint64_t(gridDim.x) * 4;
  for (uint64_t task = wave; task < tasks; task += nwaves) {
    uint32_t it = find_item(prefix, n, task);
    const Pt100Params p = items[it];
    const uint64_t gpw = 64u >> p.vw_log2;
    pt100_walkers<UB, UPL, I64>(p, (task - prefix[it]) * gpw);
  }
}

template <int UB, int UPL>
__global__ __launch_bounds__(256) void kg_0100_lds(const Pt100Params *__restrict__ items,
                                                   const uint64_t *__restrict__ prefix, uint32_t n,
                                                   uint64_t total) {
  extern __shared__ __align__(16) uint32_t lflags[];
  for (uint64_t q = blockIdx.x; q < total; q += gridDim.x) {
    uint32_t it = find_item(prefix, n, q);
    const Pt100Params p = items[it];
    pt0100_instance<UB, UPL, 256>(p, q - prefix[it], lflags);
  }
}

template <int UB>
__global__ __launch_bounds__(1024) void kg_rc_row(const RcParams *__restrict__ items, const uint64_t *__restrict__ prefix,
                                                 uint32_t n, uint64_t total) {
  using V = typename Elem<UB>::T;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  V *lds = reinterpret_cast<V *>(lds_raw);
  for (uint64_t q = blockIdx.x; q < total; q += gridDim.x) {
    uint32_t it = find_item(prefix, n, q);
    const RcParams p = items[it];
    rc_row_block<UB, 1024>(p, q - prefix[it], lds);
  }
}

// Column blocks: every item's unit count is padded to a multiple of 8 and the grid is a
// multiple of 8, so (unit mod 8) is this workgroup's XCD and each XCD sweeps a contiguous
// range of an item's column blocks.
template <int UB, bool SHUFFLE>
__global__ __launch_bounds__(1024) void kg_rc_col(const RcParams *__restrict__ items, const uint64_t *__restrict__ prefix,
                                                 uint32_t n, uint64_t total) {
  using V = typename Elem<UB>::T;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  V *lds = reinterpret_cast<V *>(lds_raw);
  for (uint64_t q = blockIdx.x; q < total; q += gridDim.x) {
    uint32_t it = find_item(prefix, n, q);
    const RcParams p = items[it];
    const uint64_t local = q - prefix[it];
    const uint64_t per = (p.ncb + 7) / 8;
    const uint64_t cb = (local & 7) * per + (local >> 3);
    if (cb < p.ncb) rc_col_block<UB, 1024, SHUFFLE>(p, cb, lds);
  }
}

// ---- host side -----------------------------------------------------------------------------

namespace {

template <typename P>
void add_item(Group &g, const P &p, uint64_t units) {
  if (units == 0) return;
  g.item_size = sizeof(P);
  size_t off = g.items.size();
  g.items.resize(off + sizeof(P));
  std::memcpy(g.items.data() + off, &p, sizeof(P));
  g.prefix.push_back(g.prefix.back() + units);
}

Bs010Params make_bs(const StageRec &r) {
  return make_bs010_params(r.buf, r.d0, r.d1, r.d2, r.bs);
}

Pt010Params make_pt010(const StageRec &r) {
  Pt010Params p{};
  p.base = r.buf;
  p.A = r.d0;
  p.a = uint32_t(r.d1);
  p.B = uint32_t(r.d2);
  p.n = uint32_t(r.d1 * r.d2);
  p.sh = uint32_t(pt010_flag_words(r.d1, r.d2, r.R, r.P));
  p.div_B = FastDiv32(p.B);
  p.div_sh = FastDiv32(p.sh);
  return p;
}

Pt100Params make_pt100(const StageRec &r, uint32_t *flags) {
  Pt100Params p{};
  p.base = r.buf;
  p.flags = flags;
  p.R = r.d1;
  p.C = r.d2;
  p.RC = r.d1 * r.d2;
  p.S = r.d3 * r.es;
  p.nsub = r.pt.nsub;
  p.sub_bytes = r.pt.sub_bytes;
  p.walkers = r.d0 * p.RC * r.pt.nsub;
  p.vw_log2 = r.pt.vw_log2;
  p.div_C64 = FastDiv64(p.C);
  p.div_RC = FastDiv64(p.RC);
  p.div_nsub = FastDiv64(p.nsub);
  p.div_C32 = FastDiv32(uint32_t(std::min<uint64_t>(p.C, 0xffffffffu)));
  return p;
}

// The out-of-place form of a descriptor: [inst][R][C] units of S bytes -> [inst][C][R]
// (false when the conversion moves nothing).
bool oop_shape(const hip_marshal_desc &d, uint64_t &inst, uint64_t &R, uint64_t &C,
               uint64_t &S) {
  const uint64_t h = uint64_t(d.rows), w = uint64_t(d.cols);
  const uint64_t t = uint64_t(std::max<int64_t>(d.tile, 1)), es = uint64_t(d.elem_bytes);
  switch (d.op) {
    case HIP_MARSHAL_OP_TRANSPOSE: inst = 1; R = h; C = w; S = es; break;
    case HIP_MARSHAL_OP_AOS_ASTA: inst = h / t; R = t; C = w; S = es; break;     // 010(h/t, t, w)
    case HIP_MARSHAL_OP_ASTA_AOS: inst = h / t; R = w; C = t; S = es; break;     // 010(h/t, w, t)
    case HIP_MARSHAL_OP_SOA_ASTA: inst = 1; R = w; C = h / t; S = t * es; break;  // 100(w, h/t, t)
    case HIP_MARSHAL_OP_ASTA_SOA: inst = 1; R = h / t; C = w; S = t * es; break;  // 100(h/t, w, t)
    default: return false;
  }
  return inst >= 1 && R > 1 && C > 1;
}

struct ArenaItem {
  int64_t desc;
  unsigned char *buf;
  uint64_t inst, R, C, S, bytes;
  uint64_t P;       // pitched order: arena row pitch (0: copy, then transpose back)
  uint64_t abytes;  // arena bytes it takes
  uint64_t off;     // its offset in the arena
};

uint64_t round8(uint64_t v) { return (v + 7) / 8 * 8; }

template <int ES, int G, int LAYOUT>
const void *bs_func_l(int upt) {
  switch (upt) {
    case 4: return reinterpret_cast<const void *>(kg_010_bs<ES, G, 4, LAYOUT>);
    case 8: return reinterpret_cast<const void *>(kg_010_bs<ES, G, 8, LAYOUT>);
    default: return reinterpret_cast<const void *>(kg_010_bs<ES, G, 16, LAYOUT>);
  }
}

template <int ES, int G>
const void *bs_func(int upt, int layout) {
  if constexpr (G > 1) {
    if (layout == 1) return bs_func_l<ES, G, 1>(upt);
    if (layout == 2) return bs_func_l<ES, G, 2>(upt);
  }
  return bs_func_l<ES, G, 0>(upt);
}

// upt_layout: units per lane * 4 + LDS layout (the group key's last field)
const void *bs_func_es(int es, int G, int upt_layout) {
  const int upt = upt_layout / 4, layout = upt_layout % 4;
  switch (es) {
    case 1: return G == 4 ? bs_func<1, 4>(upt, layout) : bs_func<1, 1>(upt, layout);
    case 2: return G == 2 ? bs_func<2, 2>(upt, layout) : bs_func<2, 1>(upt, layout);
    case 4: return bs_func<4, 1>(upt, layout);
    case 8: return bs_func<8, 1>(upt, layout);
    default: return bs_func<16, 1>(upt, layout);
  }
}

const void *pt010_func(int es) {
  switch (es) {
    case 1: return reinterpret_cast<const void *>(kg_010_pttwac<1>);
    case 2: return reinterpret_cast<const void *>(kg_010_pttwac<2>);
    case 4: return reinterpret_cast<const void *>(kg_010_pttwac<4>);
    case 8: return reinterpret_cast<const void *>(kg_010_pttwac<8>);
    default: return reinterpret_cast<const void *>(kg_010_pttwac<16>);
  }
}

template <int UB>
const void *pt100_func_ub(int upl, bool i64) {
  switch (upl) {
    case 1: return i64 ? reinterpret_cast<const void *>(kg_100<UB, 1, true>)
                       : reinterpret_cast<const void *>(kg_100<UB, 1, false>);
    case 2: return i64 ? reinterpret_cast<const void *>(kg_100<UB, 2, true>)
                       : reinterpret_cast<const void *>(kg_100<UB, 2, false>);
    case 4: return i64 ? reinterpret_cast<const void *>(kg_100<UB, 4, true>)
                       : reinterpret_cast<const void *>(kg_100<UB, 4, false>);
    default: return i64 ? reinterpret_cast<const void *>(kg_100<UB, 8, true>)
                        : reinterpret_cast<const void *>(kg_100<UB, 8, false>);
  }
}

const void *pt100_func(int ub, int upl, bool i64) {
  switch (ub) {
    case 1: return pt100_func_ub<1>(upl, i64);
    case 2: return pt100_func_ub<2>(upl, i64);
    case 4: return pt100_func_ub<4>(upl, i64);
    case 8: return pt100_func_ub<8>(upl, i64);
    default: return pt100_func_ub<16>(upl, i64);
  }
}

template <int UB>
const void *lds_func_ub(int upl) {
  switch (upl) {
    case 1: return reinterpret_cast<const void *>(kg_0100_lds<UB, 1>);
    case 2: return reinterpret_cast<const void *>(kg_0100_lds<UB, 2>);
    case 4: return reinterpret_cast<const void *>(kg_0100_lds<UB, 4>);
    default: return reinterpret_cast<const void *>(kg_0100_lds<UB, 8>);
  }
}

template <int UB>
const void *rc_func_ub(int pass) {
  switch (pass) {
    case 0: return reinterpret_cast<const void *>(kg_rc_row<UB>);
    case 1: return reinterpret_cast<const void *>(kg_rc_col<UB, false>);
    default: return reinterpret_cast<const void *>(kg_rc_col<UB, true>);
  }
}

const void *rc_func(int ub, int pass) {
  switch (ub) {
    case 1: return rc_func_ub<1>(pass);
    case 2: return rc_func_ub<2>(pass);
    case 4: return rc_func_ub<4>(pass);
    case 8: return rc_func_ub<8>(pass);
    default: return rc_func_ub<16>(pass);
  }
}

const void *lds_func(int ub, int upl) {
  switch (ub) {
    case 1: return lds_func_ub<1>(upl);
    case 2: return lds_func_ub<2>(upl);
    case 4: return lds_func_ub<4>(upl);
    case 8: return lds_func_ub<8>(upl);
    default: return lds_func_ub<16>(upl);
  }
}

}  // namespace

// The solo lists on the side streams, ordered after the work already on `st`; each lane's
// join event is recorded after its last list (wait_solo makes `st` wait for them).
static bool fork_solo(const BatchPlan &b, hipStream_t st, const SoloExec *solo) {
  if (b.solo.empty()) return false;
  if (!solo || !solo->fork) return true;
  LM_CHECK_HIP(hipEventRecord(solo->fork, st));
  for (int l = 0; l < b.solo_lanes; l++) {
    if (!solo->st[l] || !solo->join[l]) return true;
    LM_CHECK_HIP(hipStreamWaitEvent(solo->st[l], solo->fork, 0));
  }
  return false;
}

static bool run_solo(const BatchPlan &b, const SoloExec *solo) {
  if (b.solo.empty()) return false;
  for (size_t i = 0; i < b.solo.size(); i++) {
    const int l = b.solo_lane[i];
    if (run_stages_on(b.solo[i], solo->st[l], solo->scr[l])) return true;
  }
  for (int l = 0; l < b.solo_lanes; l++) LM_CHECK_HIP(hipEventRecord(solo->join[l], solo->st[l]));
  return false;
}

static bool wait_solo(const BatchPlan &b, hipStream_t st, const SoloExec *solo) {
  if (b.solo.empty()) return false;
  for (int l = 0; l < b.solo_lanes; l++) LM_CHECK_HIP(hipStreamWaitEvent(st, solo->join[l], 0));
  return false;
}

bool enqueue_batch(const BatchPlan &b, void *params, uint32_t *flags, hipStream_t st,
                   const SoloExec *solo) {
  if (fork_solo(b, st, solo) || run_solo(b, solo)) return true;
  if (b.flag_words) LM_CHECK_HIP(launch_fill_u32(flags, b.flag_words, 0, st));
  unsigned char *base = static_cast<unsigned char *>(params);
  const int lanes = solo ? solo->level_lanes : 0;
  for (auto &level : b.levels) {
    // The groups of one level touch disjoint matrices: with branch streams they run side by
    // side (a small group's few workgroups no longer wait for the groups before it), and the
    // level ends when every branch has joined the run's stream again.
    const int used = std::min<int>(lanes, int(level.size()));
    if (used > 1) {
      LM_CHECK_HIP(hipEventRecord(solo->lfork, st));
      for (int l = 0; l < used; l++) LM_CHECK_HIP(hipStreamWaitEvent(solo->lst[l], solo->lfork, 0));
    }
    for (size_t gi = 0; gi < level.size(); gi++) {
      const Group &g = level[gi];
      const void *items = base + g.items_off;
      const uint64_t *prefix = reinterpret_cast<const uint64_t *>(base + g.prefix_off);
      uint32_t n = uint32_t(g.prefix.size() - 1);
      uint64_t total = g.prefix.back();
      void *args[] = {&items, &prefix, &n, &total};
      hipStream_t gs = used > 1 ? solo->lst[gi % size_t(used)] : st;
      LM_CHECK_HIP(hipLaunchKernel(g.func, dim3(g.grid), dim3(g.block), args, g.lds, gs));
    }
    if (used > 1) {
      for (int l = 0; l < used; l++) {
        LM_CHECK_HIP(hipEventRecord(solo->ljoin[l], solo->lst[l]));
        LM_CHECK_HIP(hipStreamWaitEvent(st, solo->ljoin[l], 0));
      }
    }
  }
  return wait_solo(b, st, solo);
}

// The grouped part of a run (the flag clearing and every group, level by level) as a graph
// built node by node, with the dependencies the branch streams of enqueue_batch give: with
// `lanes` > 1 the groups of a level are dealt over that many chains, each chain's first group
// waits for the whole previous level, and a chain runs its groups one after another. No stream
// is captured, so building the graph neither restricts nor is disturbed by what other threads
// do with the device meanwhile (allocations, synchronous copies, their own captures).
static bool build_group_graph(const BatchPlan &b, void *params, uint32_t *flags, int lanes,
                              hipGraph_t *out) {
  hipGraph_t gr = nullptr;
  LM_CHECK_HIP(hipGraphCreate(&gr, 0));
  auto fail = [&](hipError_t e, const char *what) {
    fprintf(stderr, "[libmarshal] batch graph: %s: %s\n", what, hipGetErrorString(e));
    (void)hipGraphDestroy(gr);
    return true;
  };
  std::vector<hipGraphNode_t> prev;  // nodes the next level waits for
  if (b.flag_words) {
    hipGraphNode_t n = nullptr;
    hipError_t e = add_fill_u32_node(gr, nullptr, 0, flags, b.flag_words, 0, &n);
    if (e != hipSuccess) return fail(e, "flag clearing node");
    prev.push_back(n);
  }
  unsigned char *base = static_cast<unsigned char *>(params);
  for (auto &level : b.levels) {
    if (level.empty()) continue;
    const size_t used = size_t(std::max(1, std::min<int>(lanes, int(level.size()))));
    std::vector<hipGraphNode_t> cur(level.size(), nullptr);
    for (size_t gi = 0; gi < level.size(); gi++) {
      const Group &g = level[gi];
      const void *items = base + g.items_off;
      const uint64_t *prefix = reinterpret_cast<const uint64_t *>(base + g.prefix_off);
      uint32_t n = uint32_t(g.prefix.size() - 1);
      uint64_t total = g.prefix.back();
      void *args[] = {&items, &prefix, &n, &total};  // copied into the node when it is added
      hipKernelNodeParams kp{};
      kp.func = const_cast<void *>(g.func);
      kp.gridDim = dim3(g.grid);
      kp.blockDim = dim3(g.block);
      kp.sharedMemBytes = uint32_t(g.lds);
      kp.kernelParams = args;
      kp.extra = nullptr;
      const hipGraphNode_t *deps = gi >= used ? &cur[gi - used] : prev.data();
      const size_t nd = gi >= used ? 1 : prev.size();
      hipError_t e = hipGraphAddKernelNode(&cur[gi], gr, nd ? deps : nullptr, nd, &kp);
      if (e != hipSuccess) return fail(e, "kernel node");
    }
    // the last group of every chain (used <= cur.size())
    prev.assign(cur.end() - used, cur.end());
  }
  *out = gr;
  return false;
}

void build_image(const BatchPlan &b, const uint32_t *flags, std::vector<unsigned char> &host,
                 unsigned char *arena) {
  host.assign(b.param_bytes, 0);
  for (auto &lvl : b.levels)
    for (const Group &g : lvl) {
      std::memcpy(host.data() + g.items_off, g.items.data(), g.items.size());
      std::memcpy(host.data() + g.prefix_off, g.prefix.data(), g.prefix.size() * 8);
      std::memcpy(host.data() + g.prefix_off + g.prefix.size() * 8, g.lut.data(),
                  g.lut.size() * 8);
      unsigned char *img = host.data() + g.items_off;
      // arena offsets become pointers
      if (g.key.kind == 7)
        for (size_t off = 0; off < g.items.size(); off += sizeof(CopyItem)) {
          CopyItem c;
          std::memcpy(&c, img + off, sizeof(c));
          c.dst = arena + reinterpret_cast<uintptr_t>(c.dst);
          std::memcpy(img + off, &c, sizeof(c));
        }
      if (g.key.kind == 8 || g.key.kind == 10)  // from the arena / into it (pitched)
        for (size_t off = 0; off < g.items.size(); off += sizeof(OopParams)) {
          OopParams p;
          std::memcpy(&p, img + off, sizeof(p));
          if (g.key.kind == 8)
            p.src = arena + reinterpret_cast<uintptr_t>(p.src);
          else
            p.dst = arena + reinterpret_cast<uintptr_t>(p.dst);
          std::memcpy(img + off, &p, sizeof(p));
        }
      if (g.key.kind == 11)
        for (size_t off = 0; off < g.items.size(); off += sizeof(UnpitchItem)) {
          UnpitchItem u;
          std::memcpy(&u, img + off, sizeof(u));
          u.src = arena + reinterpret_cast<uintptr_t>(u.src);
          std::memcpy(img + off, &u, sizeof(u));
        }
      if (g.key.kind == 9)
        for (size_t off = 0; off < g.items.size(); off += sizeof(GatherParams)) {
          GatherParams p;
          std::memcpy(&p, img + off, sizeof(p));
          p.src = arena + reinterpret_cast<uintptr_t>(p.src);
          std::memcpy(img + off, &p, sizeof(p));
        }
      if (g.key.kind != 2) continue;  // global visited bits: offsets become pointers
      for (size_t off = 0; off < g.items.size(); off += sizeof(Pt100Params)) {
        Pt100Params p;
        std::memcpy(&p, host.data() + g.items_off + off, sizeof(p));
        p.flags = const_cast<uint32_t *>(flags) + reinterpret_cast<uintptr_t>(p.flags);
        std::memcpy(host.data() + g.items_off + off, &p, sizeof(p));
      }
    }
}

// The stage list of one descriptor (the route its plain entry point would take, except that
// the stripe route is only allowed when allow_st is set). False for an invalid descriptor.
bool plan_desc(const hip_marshal_desc &d, std::vector<StageRec> &out, std::string &route,
               bool allow_st) {
  size_t es = size_t(d.elem_bytes);
  unsigned char *buf = static_cast<unsigned char *>(d.ptr);
  bool ok = (es == 1 || es == 2 || es == 4 || es == 8 || es == 16) && d.ptr &&
            reinterpret_cast<uintptr_t>(d.ptr) % std::min<size_t>(es, 16) == 0 &&
            d.rows >= 1 && d.cols >= 1;
  if (ok && d.op != HIP_MARSHAL_OP_TRANSPOSE) ok = d.tile >= 1 && d.rows % d.tile == 0;
  if (!ok) return false;
  uint64_t h = uint64_t(d.rows), w = uint64_t(d.cols), t = uint64_t(d.tile);
  switch (d.op) {
    case HIP_MARSHAL_OP_TRANSPOSE:
      return plan_transpose_stages(out, buf, es, h, w, route, allow_st);
    case HIP_MARSHAL_OP_AOS_ASTA:    // 010(h/t, t, w)
    case HIP_MARSHAL_OP_ASTA_AOS: {  // 010(h/t, w, t)
      const bool fwd = d.op == HIP_MARSHAL_OP_AOS_ASTA;
      const char *tag = "";
      const bool ok = push_010(out, buf, es, h / t, fwd ? t : w, fwd ? w : t, M010::Auto, 1, 1,
                               &tag, allow_st);
      route = std::string(fwd ? "aos_asta:" : "asta_aos:") + tag;
      return ok;
    }
    case HIP_MARSHAL_OP_SOA_ASTA:    // 100(w, h/t, t)
    case HIP_MARSHAL_OP_ASTA_SOA: {  // 100(h/t, w, t)
      const bool fwd = d.op == HIP_MARSHAL_OP_SOA_ASTA;
      std::string tag;
      const bool ok = plan_step_0100(out, buf, es, 1, fwd ? w : h / t, fwd ? h / t : w, t,
                                     allow_st, tag);
      route = std::string(fwd ? "soa_asta:" : "asta_soa:") + tag;
      return ok;
    }
    default:
      return false;
  }
}

// Plans every descriptor and groups the stages; flags of global-bit stages are recorded as
// word offsets into the batch's flag space (see build_image).
bool plan_batch(const hip_marshal_desc *descs, int64_t count, BatchPlan &b, const char *who,
                bool allow_solo) {
  SegSyncBuild sync_tables;  // creation is not latency-critical: no background table builds
  // 1. routes
  std::vector<std::vector<StageRec>> per(count);
  b.routes.resize(count);
  // Size from which a matrix whose plain route is the square or a through-scratch form runs
  // on the side streams: "batch_solo_kb", except that a batch of less than 1 GiB in all keeps
  // its sub-MiB matrices grouped. There, with no large work to hide them behind, their
  // per-matrix launches cost more than the faster route saves (4096-matrix mix: 64 KiB
  // measured best; its sub-MB matrices alone ran at 206 GB/s grouped, 150 on side streams).
  double batch_bytes = 0;
  for (int64_t i = 0; i < count; i++)
    batch_bytes += double(descs[i].rows) * double(descs[i].cols) * double(descs[i].elem_bytes);
  double solo_min = double(uint64_t(tuning("batch_solo_kb")) << 10);
  if (batch_bytes < double(1ull << 30)) solo_min = std::max(solo_min, double(1u << 20));
  // Through-scratch rounds ("batch_arena_mb" per round, 0: off) for matrices of at most
  // "batch_arena_max_mb" whose best route is not a single streaming pass.
  const uint64_t arena_cap = uint64_t(tuning("batch_arena_mb")) << 20;
  const double arena_max =
      std::min(double(uint64_t(tuning("batch_arena_max_mb")) << 20), double(arena_cap));
  std::vector<ArenaItem> arena;
  for (int64_t i = 0; i < count; i++) {
    const hip_marshal_desc &d = descs[i];
    std::string route;
    if (allow_solo) {
      std::vector<StageRec> list;
      std::string r2;
      // The stripe route has no grouped form; the square, through-scratch, cycle-leader and
      // segmented forms beat the grouped ones from `solo_min` on (smaller matrices gain more
      // from sharing launches).
      const double mbytes = double(d.rows) * double(d.cols) * double(d.elem_bytes);
      auto has = [&](std::initializer_list<StageKind> ks) {
        for (const StageRec &r : list)
          for (StageKind k : ks)
            if (r.kind == k) return true;
        return false;
      };
      // (stage kinds, not route tags: a composer route's tag does not name its stage forms)
      if (plan_desc(d, list, r2, true)) {
        const bool stripe = has({StageKind::ST});
        const bool solo_form = mbytes >= solo_min && has({StageKind::SQ, StageKind::OOP,
                                                          StageKind::PT100C, StageKind::SEG,
                                                          StageKind::CL});
        // A single streaming pass as it would run here: the LDS-staged 010 (grouped below),
        // the square, cycle-leader and cluster forms on the side streams.
        const StageKind k0 = list.size() == 1 ? list[0].kind : StageKind::OOP;
        const bool one_pass =
            list.size() == 1 &&
            (k0 == StageKind::BS010 ||
             (solo_form && (k0 == StageKind::SQ || k0 == StageKind::PT100C ||
                            k0 == StageKind::CL)));
        ArenaItem ai{};
        if (!stripe && !one_pass && arena_cap && mbytes <= arena_max &&
            (reinterpret_cast<uintptr_t>(d.ptr) & 15u) == 0 &&
            oop_shape(d, ai.inst, ai.R, ai.C, ai.S) &&
            ((ai.S <= 16 && (ai.S & (ai.S - 1)) == 0) || oop_gather_fits(ai.S))) {
          ai.desc = i;
          ai.buf = static_cast<unsigned char *>(d.ptr);
          ai.bytes = ai.inst * ai.R * ai.C * ai.S;
          ai.P = oop_pitch_for(ai.buf, ai.inst, ai.R, ai.C, ai.S);
          ai.abytes = ai.P ? ai.inst * ai.C * ai.P : ai.bytes;
          arena.push_back(ai);
          b.routes[i] = "ARENA-";
          b.stages += 1;
          b.bytes += 2.0 * mbytes;
          continue;
        }
        if (stripe || solo_form) {
          b.solo.push_back(std::move(list));
          b.solo_bytes.push_back(mbytes);
          b.routes[i] = r2;
          b.stages += 1;
          b.bytes += 2.0 * mbytes;
          continue;
        }
      }
    }
    bool planned;
    {
      RestrictedPlanning restricted;  // grouped kernels: no side buffers or device tables
      planned = plan_desc(d, per[i], route, false);
    }
    if (!planned) {
      fprintf(stderr, "[libmarshal] %s: invalid descriptor %lld\n", who, (long long)i);
      return false;
    }
    uint64_t h = uint64_t(d.rows), w = uint64_t(d.cols);
    size_t es = size_t(d.elem_bytes);
    b.routes[i] = route;
    b.stages += int64_t(per[i].size());
    b.bytes += 2.0 * double(h) * double(w) * double(es);
  }
  // Solo lists: largest first onto the least loaded lane.
  if (!b.solo.empty()) {
    std::vector<size_t> order(b.solo.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return b.solo_bytes[x] > b.solo_bytes[y]; });
    double load[kSoloLanes] = {};
    b.solo_lane.assign(b.solo.size(), 0);
    const size_t lanes = size_t(std::min<int64_t>(kSoloLanes, tuning("batch_solo_lanes")));
    b.solo_lanes = int(std::min<size_t>(lanes, b.solo.size()));
    for (size_t i : order) {
      int l = 0;
      for (int k = 1; k < b.solo_lanes; k++)
        if (load[k] < load[l]) l = k;
      load[l] += b.solo_bytes[i];
      b.solo_lane[i] = l;
      b.solo_scratch[l] = std::max(b.solo_scratch[l], stages_scratch_bytes(b.solo[i]));
    }
  }
  // 2. flag space for every global-bit stage
  size_t maxlev = 0;
  for (auto &v : per) maxlev = std::max(maxlev, v.size());
  std::vector<std::vector<uint64_t>> flag_off(count);
  for (int64_t i = 0; i < count; i++) {
    for (const StageRec &r : per[i]) {
      flag_off[i].push_back(b.flag_words);
      if (r.kind == StageKind::PT100G) b.flag_words += (r.pt.flag_words + 63) / 64 * 64;
    }
  }
  // 3. groups per level
  const DeviceInfo &di = device_info();
  for (size_t lv = 0; lv < maxlev; lv++) {
    std::map<GroupKey, Group> groups;
    for (int64_t i = 0; i < count; i++) {
      if (lv >= per[i].size()) continue;
      const StageRec &r = per[i][lv];
      switch (r.kind) {
        case StageKind::BS010: {
          GroupKey k{0, int(r.es), r.bs.G, r.bs.upt * 4 + r.bs.layout};
          Group &g = groups[k];
          g.key = k;
          g.lds = std::max(g.lds, r.bs.lds_bytes);
          add_item(g, make_bs(r), r.bs.nchunks);
          break;
        }
        case StageKind::PT010: {
          GroupKey k{1, int(r.es), 0, 0};
          Group &g = groups[k];
          g.key = k;
          g.lds = std::max<size_t>(g.lds, pt010_flag_words(r.d1, r.d2, r.R, r.P) * 4);
          add_item(g, make_pt010(r), r.d0);
          break;
        }
        case StageKind::PT100G: {
          bool i64 = r.d1 * r.d2 >= (1ull << 32);
          GroupKey k{2, r.pt.ub, r.pt.upl, i64 ? 1 : 0};
          Group &g = groups[k];
          g.key = k;
          Pt100Params p = make_pt100(r, reinterpret_cast<uint32_t *>(uintptr_t(flag_off[i][lv])));
          uint64_t gpw = 64u >> p.vw_log2;
          add_item(g, p, (p.walkers + gpw - 1) / gpw);
          break;
        }
        case StageKind::RC_ROW:
        case StageKind::RC_COL1:
        case StageKind::RC_COL2: {
          int pass = r.kind == StageKind::RC_ROW ? 0 : r.kind == StageKind::RC_COL1 ? 1 : 2;
          GroupKey k{4 + pass, r.rc.ub, 0, 0};
          Group &g = groups[k];
          g.key = k;
          RcParams p = make_rc(r.buf, r.d1, r.d2, r.rc);
          if (pass == 0) {
            g.lds = std::max(g.lds, r.rc.row_lds);
            add_item(g, p, (p.inst * r.d1 + p.rows - 1) / p.rows);
          } else {
            g.lds = std::max(g.lds, r.rc.col_lds);
            add_item(g, p, (p.ncb + 7) / 8 * 8);
          }
          break;
        }
        case StageKind::PT0100L: {
          GroupKey k{3, r.pt.ub, r.pt.upl, 0};
          Group &g = groups[k];
          g.key = k;
          g.lds = std::max(g.lds, lds0100_bytes(r.d1, r.d2, r.pt));
          add_item(g, make_pt100(r, nullptr), r.d0);
          break;
        }
        case StageKind::ST:  // never planned for a batch (allow_st = false)
        case StageKind::SQ:
        case StageKind::OOP:
        case StageKind::PT100C:
        case StageKind::SEG:
        case StageKind::CL:
          fprintf(stderr, "[libmarshal] %s: stripe route in a batch\n", who);
          return false;
      }
    }
    std::vector<Group> lvl;
    for (auto &kv : groups) {
      Group g = kv.second;
      if (g.prefix.size() < 2) continue;
      uint64_t total = g.prefix.back();
      switch (g.key.kind) {
        case 0:
          g.func = bs_func_es(g.key.a, g.key.b, g.key.c);
          g.grid = uint32_t(std::min<uint64_t>(total, uint64_t(di.cus) * tuning("bs_blocks_per_cu")));
          break;
        case 1:
          g.func = pt010_func(g.key.a);
          g.grid = uint32_t(std::min<uint64_t>(total, uint64_t(di.cus) * tuning("pt010_blocks_per_cu")));
          break;
        case 2:
          g.func = pt100_func(g.key.a, g.key.b, g.key.c != 0);
          g.grid = uint32_t(std::min<uint64_t>((total + 3) / 4,
                                               uint64_t(di.cus) * tuning("pt100_blocks_per_cu")));
          break;
        case 3:
          g.func = lds_func(g.key.a, g.key.b);
          g.grid = uint32_t(std::min<uint64_t>(total, uint64_t(di.cus) * 8));
          break;
        case 4:
          g.block = 1024;
          g.func = rc_func(g.key.a, 0);
          g.grid = uint32_t(std::min<uint64_t>(total, uint64_t(di.cus) * tuning("rc_row_blocks_per_cu")));
          break;
        default:
          g.block = 1024;
          g.func = rc_func(g.key.a, g.key.kind - 4);
          g.grid = uint32_t(std::min<uint64_t>(total, uint64_t(di.cus) * tuning("rc_col_blocks_per_cu")));
          g.grid = (g.grid + 7) / 8 * 8;
          break;
      }
      if (g.lds > 65536 &&
          hipFuncSetAttribute(g.func, hipFuncAttributeMaxDynamicSharedMemorySize, int(g.lds)) !=
              hipSuccess) {
        fprintf(stderr, "[libmarshal] %s: LDS attribute failed\n", who);
        return false;
      }
      lvl.push_back(std::move(g));
    }
    b.levels.push_back(std::move(lvl));
  }
  // 3b. through-scratch rounds: first fit, largest first, into rounds of `arena_cap` bytes.
  // Round k: one grouped copy of its matrices into the arena, then the grouped out-of-place
  // kernels back into their buffers; the next round's copy waits for them (same stream).
  if (!arena.empty()) {
    std::stable_sort(arena.begin(), arena.end(),
                     [](const ArenaItem &x, const ArenaItem &y) { return x.abytes > y.abytes; });
    std::vector<uint64_t> fill;
    std::vector<std::vector<size_t>> rounds;
    for (size_t k = 0; k < arena.size(); k++) {
      const uint64_t need = (arena[k].abytes + 255) / 256 * 256;
      size_t r = 0;
      while (r < fill.size() && fill[r] + need > arena_cap) r++;
      if (r == fill.size()) {
        fill.push_back(0);
        rounds.emplace_back();
      }
      arena[k].off = fill[r];
      fill[r] += need;
      rounds[r].push_back(k);
      b.arena_bytes = std::max(b.arena_bytes, fill[r]);
    }
    b.arena_rounds = int(rounds.size());
    b.arena_matrices = int64_t(arena.size());
    const size_t tile = size_t(tuning("st_tile_kb")) << 10;
    const size_t gtile = size_t(tuning("gather_tile_kb")) << 10;
    const uint64_t grow = uint64_t(tuning("gather_row_bytes"));
    for (auto &rk : rounds) {
      // level A: copies into the arena and pitched transposes into it; level B: transposes,
      // gathers and unpitch copies back into the buffers
      Group cp;
      cp.key = GroupKey{7, 0, 0, 0};
      std::map<GroupKey, Group> fwd, back;
      for (size_t k : rk) {
        const ArenaItem &a = arena[k];
        // arena-relative addresses; build_image adds the arena's base
        unsigned char *scr = reinterpret_cast<unsigned char *>(uintptr_t(a.off));
        const uint64_t block = a.R * a.C * a.S, row_bytes = a.R * a.S;
        if (a.P) {
          OopLaunch o;
          if (prepare_oop_transpose(a.buf, scr, a.S, a.inst, a.R, a.C, block, a.C * a.P, tile,
                                    0, a.P, -1, &o) != hipSuccess ||
              !kg_oop_transpose_func(o.es, o.vin, o.out)) {
            fprintf(stderr, "[libmarshal] %s: no grouped transpose for descriptor %lld\n", who,
                    (long long)a.desc);
            return false;
          }
          const UnpitchItem u = make_unpitch_item(a.buf, scr, row_bytes, a.P, a.inst * a.C);
          const int nts = (tuning("oop_nt") & 2) ? 1 : 0;
          GroupKey gk{10, o.es, o.vin ? 1 : 0, o.out};
          Group &g = fwd[gk];
          g.key = gk;
          g.lds = std::max(g.lds, o.lds);
          add_item(g, o.p, round8(o.grid));
          GroupKey uk{11, nts, 0, 0};
          Group &ug = back[uk];
          ug.key = uk;
          add_item(ug, u, std::max<uint64_t>(1, (u.nb + kUnpitchItemUnit - 1) / kUnpitchItemUnit));
          continue;
        }
        add_item(cp, CopyItem{scr, a.buf, a.bytes}, (a.bytes + kCopyItemUnit - 1) / kCopyItemUnit);
        const bool aligned_rows = ((reinterpret_cast<uintptr_t>(a.buf) | row_bytes) & 15u) == 0;
        const bool unit = a.S == 1 || a.S == 2 || a.S == 4 || a.S == 8 || a.S == 16;
        if (unit && (aligned_rows || row_bytes >= grow)) {
          OopLaunch o;
          if (prepare_oop_transpose(scr, a.buf, a.S, a.inst, a.R, a.C, block, block, tile, 0, 0,
                                    -1, &o) != hipSuccess ||
              !kg_oop_transpose_func(o.es, o.vin, o.out)) {
            fprintf(stderr, "[libmarshal] %s: no grouped transpose for descriptor %lld\n", who,
                    (long long)a.desc);
            return false;
          }
          GroupKey gk{8, o.es, o.vin ? 1 : 0, o.out};
          Group &g = back[gk];
          g.key = gk;
          g.lds = std::max(g.lds, o.lds);
          add_item(g, o.p, round8(o.grid));
        } else {
          GatherLaunch o;
          if (prepare_oop_gather(scr, a.buf, a.S, a.inst, a.R, a.C, gtile, &o) != hipSuccess) {
            fprintf(stderr, "[libmarshal] %s: no grouped gather for descriptor %lld\n", who,
                    (long long)a.desc);
            return false;
          }
          GroupKey gk{9, o.u, o.fc ? 1 : 0, 0};
          Group &g = back[gk];
          g.key = gk;
          g.lds = std::max(g.lds, o.lds);
          add_item(g, o.p, round8(o.grid));
        }
      }
      auto finish = [&](std::map<GroupKey, Group> &gs, std::vector<Group> &lvl) {
        for (auto &kv : gs) {
          Group g = kv.second;
          switch (g.key.kind) {
            case 8:
            case 10:
              g.func = kg_oop_transpose_func(g.key.a, g.key.b != 0, g.key.c);
              g.block = kOopGroupThreads;
              g.grid = uint32_t(std::min<uint64_t>(g.prefix.back(), uint64_t(di.cus) * 4));
              break;
            case 9:
              g.func = kg_oop_gather_func(g.key.a, g.key.b != 0);
              g.block = kOopGroupThreads;
              g.grid = uint32_t(std::min<uint64_t>(g.prefix.back(), uint64_t(di.cus) * 4));
              break;
            default:  // 11: unpitch copies
              g.func = kg_unpitch_func(g.key.a != 0);
              g.block = 256;
              g.grid = uint32_t(std::min<uint64_t>(g.prefix.back(), uint64_t(di.cus) * 8));
              break;
          }
          if (g.lds > 65536 &&
              hipFuncSetAttribute(g.func, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  int(g.lds)) != hipSuccess) {
            fprintf(stderr, "[libmarshal] %s: LDS attribute failed\n", who);
            return false;
          }
          lvl.push_back(std::move(g));
        }
        return true;
      };
      std::vector<Group> la, lb;
      if (cp.prefix.size() > 1) {
        cp.func = kg_copy_func();
        cp.block = 256;
        cp.grid = uint32_t(std::min<uint64_t>(cp.prefix.back(), uint64_t(di.cus) * 8));
        la.push_back(std::move(cp));
      }
      if (!finish(fwd, la) || !finish(back, lb)) return false;
      b.levels.push_back(std::move(la));
      b.levels.push_back(std::move(lb));
    }
  }
  // 4. layout of every work list in one parameter buffer
  b.param_bytes = 0;
  for (auto &lvl : b.levels)
    for (Group &g : lvl) {
      g.items_off = b.param_bytes;
      b.param_bytes += (g.items.size() + 255) / 256 * 256;
      // find_item's table: the owner of unit j * kFindStep, for j up to one past the end
      const uint64_t total = g.prefix.back();
      const size_t n = g.prefix.size() - 1;
      g.lut.assign(size_t(total / kFindStep) + 2, uint64_t(n - 1));
      size_t it = 0;
      for (size_t j = 0; j < g.lut.size() && j * kFindStep < total; j++) {
        while (g.prefix[it + 1] <= j * kFindStep) it++;
        g.lut[j] = it;
      }
      g.prefix_off = b.param_bytes;
      b.param_bytes += ((g.prefix.size() + g.lut.size()) * 8 + 255) / 256 * 256;
    }
  return true;
}

}  // namespace lm

using namespace lm;

struct hip_marshal_batch {
  BatchPlan b;
  void *params = nullptr;     // device copy of all item arrays and prefix sums
  uint32_t *flags = nullptr;  // visited bits of every global-bit stage
  unsigned char *arena = nullptr;  // through-scratch rounds
  SoloExec solo;  // side streams, events and scratch of the stripe-route matrices
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool use_graph = true;
};

namespace lm {
// Measures each shape class of the AoS/SoA <-> ASTA descriptors once (on its first
// descriptor's buffer, which ends bit-identical), on a private stream, before the batch is
// planned (hip_marshal_autotune_desc). Full transposes keep the cost model's routes: their
// measured choices are per exact shape, which a batch of many shapes would spend most of its
// creation on.
bool autotune_classes(const hip_marshal_desc *descs, int64_t count) {
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return true;
  std::set<ClassKey> done;
  bool err = false;
  for (int64_t i = 0; i < count && !err; i++) {
    const hip_marshal_desc &d = descs[i];
    const uint64_t h = uint64_t(std::max<int64_t>(d.rows, 1)),
                   w = uint64_t(std::max<int64_t>(d.cols, 1));
    const uint64_t t = uint64_t(std::max<int64_t>(d.tile, 1));
    const size_t es = size_t(d.elem_bytes);
    ClassKey k{};
    switch (d.op) {
      case HIP_MARSHAL_OP_TRANSPOSE: continue;
      case HIP_MARSHAL_OP_AOS_ASTA: k = class_key_010(es, h / t, t, w); break;
      case HIP_MARSHAL_OP_ASTA_AOS: k = class_key_010(es, h / t, w, t); break;
      case HIP_MARSHAL_OP_SOA_ASTA: k = class_key_0100(1, w, h / t, t * es); break;
      case HIP_MARSHAL_OP_ASTA_SOA: k = class_key_0100(1, h / t, w, t * es); break;
      default: continue;
    }
    if (!done.insert(k).second || class_tuned(k)) continue;
    int form = 0;
    err = hip_marshal_autotune_desc(st, &d, &form);
  }
  (void)hipStreamSynchronize(st);
  (void)hipStreamDestroy(st);
  return err;
}
}  // namespace lm

extern "C" {

hip_marshal_batch *hip_marshal_batch_create(const hip_marshal_desc *descs, int64_t count,
                                            int use_graph) {
  const char *who = "hip_marshal_batch_create";
  if (!descs || count < 1) {
    fprintf(stderr, "[libmarshal] %s: empty batch\n", who);
    return nullptr;
  }
  if (tuning("autotune") && autotune_classes(descs, count)) {
    fprintf(stderr, "[libmarshal] %s: autotuning failed\n", who);
    return nullptr;
  }
  auto *hb = new hip_marshal_batch();
  hb->use_graph = use_graph != 0;
  BatchPlan &b = hb->b;
  if (!plan_batch(descs, count, b, who, true)) {
    hip_marshal_batch_destroy(hb);
    return nullptr;
  }
  // every device call checked; the first failure is named in the message
  const char *failed = nullptr;
  hipError_t ferr = hipSuccess;
  auto ok_ = [&](hipError_t e, const char *what) {
    if (e != hipSuccess && !failed) {
      failed = what;
      ferr = e;
    }
    return e == hipSuccess;
  };
  bool ok = true;
  if (!b.solo.empty()) {
    ok = ok_(hipEventCreateWithFlags(&hb->solo.fork, hipEventDisableTiming), "hipEventCreate");
    for (int l = 0; ok && l < b.solo_lanes; l++) {
      ok = ok_(hipStreamCreateWithFlags(&hb->solo.st[l], hipStreamNonBlocking),
               "hipStreamCreate") &&
           ok_(hipEventCreateWithFlags(&hb->solo.join[l], hipEventDisableTiming),
               "hipEventCreate");
      if (ok && b.solo_scratch[l])
        ok = ok_(hipMalloc(reinterpret_cast<void **>(&hb->solo.scr[l]), b.solo_scratch[l]),
                 "hipMalloc (side-stream scratch)");
    }
  }
  // branch streams for levels of more than one group
  size_t widest = 0;
  for (auto &l : b.levels) widest = std::max(widest, l.size());
  const int want = int(std::min<int64_t>(kLevelLanes, tuning("batch_level_lanes")));
  if (ok && widest > 1 && want > 1) {
    const int n = int(std::min<size_t>(size_t(want), widest));
    ok = ok_(hipEventCreateWithFlags(&hb->solo.lfork, hipEventDisableTiming), "hipEventCreate");
    for (int l = 0; ok && l < n; l++)
      ok = ok_(hipStreamCreateWithFlags(&hb->solo.lst[l], hipStreamNonBlocking),
               "hipStreamCreate") &&
           ok_(hipEventCreateWithFlags(&hb->solo.ljoin[l], hipEventDisableTiming),
               "hipEventCreate");
    if (ok) hb->solo.level_lanes = n;
  }
  if (ok && b.flag_words) ok = ok_(hipMalloc(&hb->flags, b.flag_words * 4), "hipMalloc (flags)");
  if (ok && b.arena_bytes)
    ok = ok_(hipMalloc(reinterpret_cast<void **>(&hb->arena), b.arena_bytes),
             "hipMalloc (arena)");
  if (ok && b.param_bytes) ok = ok_(hipMalloc(&hb->params, b.param_bytes), "hipMalloc (params)");
  if (ok && b.param_bytes) {
    std::vector<unsigned char> host;
    build_image(b, hb->flags, host, hb->arena);
    ok = ok_(hipMemcpy(hb->params, host.data(), b.param_bytes, hipMemcpyHostToDevice),
             "hipMemcpy (work lists)");
  }
  if (!ok) {
    fprintf(stderr, "[libmarshal] %s: %s failed: %s\n", who, failed ? failed : "setup",
            hipGetErrorString(ferr));
    hip_marshal_batch_destroy(hb);
    return nullptr;
  }
  return hb;
}

bool hip_marshal_batch_run(hip_marshal_batch *hb, hipStream_t st, uint64_t *elapsed_ns) {
  if (!hb) return true;
  BatchPlan &b = hb->b;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (elapsed_ns) {
    LM_CHECK_HIP(hipEventCreate(&e0));
    LM_CHECK_HIP(hipEventCreate(&e1));
    LM_CHECK_HIP(hipEventRecord(e0, st));
  }
  if (hb->use_graph) {
    if (!hb->exec) {
      hipGraph_t g = nullptr;
      if (build_group_graph(b, hb->params, hb->flags, hb->solo.level_lanes, &g)) return true;
      hb->graph = g;
      LM_CHECK_HIP(hipGraphInstantiate(&hb->exec, hb->graph, nullptr, nullptr, 0));
    }
    // the solo lists start with the graph, on the side streams, and the run's stream waits
    // for both
    if (fork_solo(b, st, &hb->solo)) return true;
    LM_CHECK_HIP(hipGraphLaunch(hb->exec, st));
    if (run_solo(b, &hb->solo) || wait_solo(b, st, &hb->solo)) return true;
  } else if (enqueue_batch(b, hb->params, hb->flags, st, &hb->solo)) {
    return true;
  }
  if (elapsed_ns) {
    float ms = 0;
    LM_CHECK_HIP(hipEventRecord(e1, st));
    LM_CHECK_HIP(hipEventSynchronize(e1));
    LM_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    *elapsed_ns += uint64_t(double(ms) * 1e6 + 0.5);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  return false;
}

int64_t hip_marshal_batch_info(const hip_marshal_batch *hb, const char *key) {
  if (!hb || !key) return -1;
  const BatchPlan &b = hb->b;
  if (!strcmp(key, "levels")) return int64_t(b.levels.size());
  if (!strcmp(key, "stages")) return b.stages;
  if (!strcmp(key, "flag_bytes")) return int64_t(b.flag_words * 4);
  if (!strcmp(key, "bytes")) return int64_t(b.bytes);
  if (!strcmp(key, "solo")) return int64_t(b.solo.size());
  if (!strcmp(key, "arena_matrices")) return b.arena_matrices;
  if (!strcmp(key, "arena_rounds")) return int64_t(b.arena_rounds);
  if (!strcmp(key, "arena_bytes")) return int64_t(b.arena_bytes);
  if (!strcmp(key, "kernels")) {
    int64_t k = b.flag_words ? 1 : 0;
    for (auto &l : b.levels) k += int64_t(l.size());
    return k;
  }
  return -1;
}

const char *hip_marshal_batch_route(const hip_marshal_batch *hb, int64_t i) {
  if (!hb || i < 0 || i >= int64_t(hb->b.routes.size())) return "";
  return hb->b.routes[size_t(i)].c_str();
}

void hip_marshal_batch_destroy(hip_marshal_batch *hb) {
  if (!hb) return;
  if (hb->exec) (void)hipGraphExecDestroy(hb->exec);
  if (hb->graph) (void)hipGraphDestroy(hb->graph);
  if (hb->params) (void)hipFree(hb->params);
  if (hb->flags) (void)hipFree(hb->flags);
  if (hb->arena) (void)hipFree(hb->arena);
  for (int l = 0; l < kSoloLanes; l++) {
    if (hb->solo.st[l]) (void)hipStreamSynchronize(hb->solo.st[l]);
    if (hb->solo.scr[l]) (void)hipFree(hb->solo.scr[l]);
    if (hb->solo.join[l]) (void)hipEventDestroy(hb->solo.join[l]);
    if (hb->solo.st[l]) (void)hipStreamDestroy(hb->solo.st[l]);
  }
  if (hb->solo.fork) (void)hipEventDestroy(hb->solo.fork);
  for (int l = 0; l < kLevelLanes; l++) {
    if (hb->solo.lst[l]) (void)hipStreamSynchronize(hb->solo.lst[l]);
    if (hb->solo.ljoin[l]) (void)hipEventDestroy(hb->solo.ljoin[l]);
    if (hb->solo.lst[l]) (void)hipStreamDestroy(hb->solo.lst[l]);
  }
  if (hb->solo.lfork) (void)hipEventDestroy(hb->solo.lfork);
  delete hb;
}

}  // extern "C"
