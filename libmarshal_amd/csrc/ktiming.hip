// ktiming.hip - optional per-launch timing of the library's kernels.
//
// With hip_marshal_set_kernel_timing(1) every launcher brackets its launch with a pair of HIP
// events on the launch's own stream and records the kernel family, the stream and the launch's
// algorithmic bytes (2 x the bytes it moves into place). hip_marshal_kernel_times waits for
// the recorded events and returns, per family, the launch count, the summed device time and
// the summed algorithmic bytes. Nothing is recorded while the stream is being captured, and
// nothing at all while timing is off (one relaxed atomic load per launch).
#include "kernels.h"
#include "../../include/hip_marshal.h"

#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace lm {
namespace {
std::atomic<int> g_on{0};
struct Rec {
  const char *name;
  double bytes;
  hipEvent_t e0, e1;
};
std::mutex g_mu;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_free;

hipEvent_t take_event() {
  if (!g_free.empty()) {
    hipEvent_t e = g_free.back();
    g_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

KernelSpan::KernelSpan(hipStream_t st, const char *name, double bytes) : st_(st) {
  if (!g_on.load(std::memory_order_relaxed)) return;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
  std::lock_guard<std::mutex> lk(g_mu);
  hipEvent_t e0 = take_event(), e1 = take_event();
  if (!e0 || !e1 || hipEventRecord(e0, st) != hipSuccess) {
    if (e0) g_free.push_back(e0);
    if (e1) g_free.push_back(e1);
    return;
  }
  idx_ = int64_t(g_recs.size());
  g_recs.push_back(Rec{name, bytes, e0, e1});
}

KernelSpan::~KernelSpan() {
  if (idx_ < 0) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (size_t(idx_) < g_recs.size()) (void)hipEventRecord(g_recs[size_t(idx_)].e1, st_);
}

}  // namespace lm

using namespace lm;

extern "C" {

void hip_marshal_set_kernel_timing(int on) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (Rec &r : g_recs) {
    (void)hipEventSynchronize(r.e1);
    g_free.push_back(r.e0);
    g_free.push_back(r.e1);
  }
  g_recs.clear();
  g_on.store(on ? 1 : 0);
}

int hip_marshal_kernel_times(hip_marshal_kernel_time *out, int max_entries) {
  std::lock_guard<std::mutex> lk(g_mu);
  std::map<std::string, hip_marshal_kernel_time> agg;
  std::vector<std::string> order;
  for (Rec &r : g_recs) {
    float ms = 0.f;
    if (hipEventSynchronize(r.e1) != hipSuccess || hipEventElapsedTime(&ms, r.e0, r.e1) != hipSuccess)
      ms = 0.f;
    auto it = agg.find(r.name);
    if (it == agg.end()) {
      hip_marshal_kernel_time t{};
      strncpy(t.name, r.name, sizeof(t.name) - 1);
      it = agg.emplace(r.name, t).first;
      order.push_back(r.name);
    }
    it->second.launches += 1;
    it->second.ms += double(ms);
    it->second.bytes += r.bytes;
    g_free.push_back(r.e0);
    g_free.push_back(r.e1);
  }
  g_recs.clear();
  int n = 0;
  for (const std::string &k : order) {
    if (n >= max_entries) break;
    if (out) out[n] = agg[k];
    n++;
  }
  return n;
}

}  // extern "C"
