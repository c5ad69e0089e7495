// noalloc.hip - the one entry point that allocates no device memory.
//
// hip_transpose_0100_noalloc_es runs the 100 / 0100 step of hip_transpose_0100_es, limited to
// the in-place forms (LDS-staged 010, row/column shuffles, LDS bits, global bits), with the
// visited bits of the global-bit form in a fixed device array instead of the library's pool.
// It calls no hipMalloc / hipFree (and no event or stream creation): for callers that cannot
// allocate (code under stream capture, tight serving loops). The array lives for the whole
// process on every device the entry point runs on; every call on a device shares it, so calls
// that overlap on one device race (each call clears the bits it uses before its kernels run).
#include "kernels.h"
#include "../../include/hip_marshal.h"

#include <mutex>
#include <string>
#include <vector>

namespace lm {
namespace {

constexpr size_t kScratchBytes = size_t(32) << 20;  // HIP_MARSHAL_NOALLOC_SCRATCH_BYTES

// The visited-bit space of the no-alloc entry point: 32 MiB per device.
static __device__ __align__(256) uint32_t scratch[kScratchBytes / 4];

// Device address of `scratch` on each device, looked up once per device with that device
// current. (The array's name in host code is a host-side stand-in, not a device address.)
std::mutex g_mu;
void *g_addr[64] = {};

void *scratch_here(int dev) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_addr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "[libmarshal] hip_transpose_0100_noalloc: no address for the scratch "
                      "array on device %d: %s\n", dev, hipGetErrorString(e));
      return nullptr;
    }
    g_addr[dev] = p;
  }
  return g_addr[dev];
}

}  // namespace
}  // namespace lm

using namespace lm;

extern "C" {

bool hip_transpose_0100_noalloc_es(hipStream_t st, void *src, size_t es, int64_t A, int64_t a,
                                   int64_t B, int64_t b) {
  const char *who = "hip_transpose_0100_noalloc";
  if (A < 1 || a < 1 || B < 1 || b < 1) {
    fprintf(stderr, "[libmarshal] %s: dimensions must be positive\n", who);
    return true;
  }
  if (!(es == 1 || es == 2 || es == 4 || es == 8 || es == 16)) {
    fprintf(stderr, "[libmarshal] %s: element size must be 1, 2, 4, 8 or 16 bytes\n", who);
    return true;
  }
  const size_t al = es < 16 ? es : 16;
  if (!src || reinterpret_cast<uintptr_t>(src) % al) {
    fprintf(stderr, "[libmarshal] %s: null or misaligned buffer\n", who);
    return true;
  }
  const double bytes = double(A) * double(a) * double(B) * double(b) * double(es);
  if (bytes > double(int64_t(1) << 62)) {
    fprintf(stderr, "[libmarshal] %s: matrix too large\n", who);
    return true;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    fprintf(stderr, "[libmarshal] %s: no current device\n", who);
    return true;
  }
  unsigned char *scr = static_cast<unsigned char *>(scratch_here(dev));
  if (!scr) return true;
  // In-place forms only (allow_oop = false): no through-scratch copies, no cycle-leader lists
  // (whose cache allocates); the only scratch is the global-bit form's visited bits.
  std::vector<StageRec> list;
  std::string tag;
  RestrictedPlanning restricted;  // no side buffers or device tables: visited bits only
  if (!plan_step_0100(list, static_cast<unsigned char *>(src), es, uint64_t(A), uint64_t(a),
                      uint64_t(B), uint64_t(b), false, tag))
    return true;
  for (const StageRec &r : list) {
    if (r.kind == StageKind::OOP || r.kind == StageKind::ST || r.kind == StageKind::PT100C ||
        r.kind == StageKind::SEG) {
      fprintf(stderr, "[libmarshal] %s: internal error: planned a form that needs a pool\n", who);
      return true;
    }
  }
  const size_t need = stages_scratch_bytes(list);
  if (need > kScratchBytes) {
    fprintf(stderr,
            "[libmarshal] %s: this step needs %zu bytes of visited bits, more than the %zu of "
            "the no-alloc scratch array (use hip_transpose_0100_es)\n",
            who, need, kScratchBytes);
    return true;
  }
  set_last_route(tag);
  return run_stages_on(list, st, scr, true);
}

}  // extern "C"
