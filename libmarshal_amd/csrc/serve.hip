// serve.hip - a request loop with continuous batching for in-place layout conversions.
//
// Clients submit conversions of device-resident buffers. Requests are admitted in arrival
// order while the device memory they hold stays within the HBM budget: the buffers of admitted,
// unfinished requests (plus a bound on their visited bits) and every scratch allocation the
// server holds (the one shared scratch of the large request that runs at a time, each stream's
// scratch, visited bits and work lists), counting what an admission would have to grow. Each
// request is planned once, at submission, with its scratch capped by the "serve_scratch_mb"
// knob, so its exact scratch need is known before it is admitted. Small requests admitted
// together form one dynamic batch of grouped persistent kernels (the batched form's work
// lists) on an idle stream; a small request admitted alone, and every request of 64 MiB or
// more, runs with its plain stage list and the stream's own scratch. Large requests run one
// at a time (sharing HBM bandwidth between them only delays each); small requests may go
// ahead of a large one that waits for its turn. A submission that finds a stream idle is
// started on the submitting thread; otherwise the scheduler thread starts it as streams and
// budget free up. Completion is observed with an event per stream, polled closely while work
// is in flight; latency is arrival -> completion.
//
// No counterpart in the reference (a single-call library, src/cl_marshal.cc).
#include "batch.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <string>
#include <vector>

namespace lm {

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
  hip_marshal_desc d;
  size_t bytes = 0;    // buffer + a bound on its visited bits
  size_t scratch = 0;  // scratch of its plain stage list
  std::vector<StageRec> plain;  // planned at submission
  double t_arrive = 0, t_launch = 0, t_done = 0;
  int state = 0;  // 0 queued, 1 running, 2 done, -1 failed
};

struct Slot {
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  void *params = nullptr;
  size_t params_cap = 0;
  unsigned char *host = nullptr;  // pinned staging for the parameter upload
  size_t host_cap = 0;
  uint32_t *flags = nullptr;
  size_t flags_cap = 0;
  unsigned char *scr = nullptr;  // scratch of a request run alone (visited bits, stripes)
  size_t scr_cap = 0;
  bool busy = false;
  bool launching = false;  // work is being enqueued (lock released): not yet retirable
  bool solo = false;       // runs one large request
  std::vector<int64_t> reqs;
  size_t bytes = 0;
};

}  // namespace

struct Server {
  size_t budget = 0;
  int max_batch = 256;
  size_t solo_bytes = size_t(64) << 20;  // requests this large run alone, on the plain path
  uint64_t scratch_cap = 0;              // per-request scratch cap of the plain forms
  int device = 0;
  // scratch of the large request that runs at a time (one at most), shared by all streams
  unsigned char *solo_scr = nullptr;
  size_t solo_cap = 0;
  std::atomic<size_t> held{0};  // device bytes of every server allocation (budget accounting)
  std::mutex mu;
  std::condition_variable cv;
  std::deque<int64_t> queue;
  std::unordered_map<int64_t, Req> reqs;
  int64_t next_id = 0;
  size_t inflight = 0;
  int solo_running = 0;  // large requests on the GPU (at most one at a time)
  std::vector<Slot> slots;
  std::thread th;
  bool stop = false;
  // statistics
  std::vector<double> lat;
  double bytes_done = 0;
  int64_t batches = 0, failed = 0, max_in_batch = 0;
  double t_first = -1, t_last = 0;
};

namespace {

size_t footprint(const hip_marshal_desc &d) {
  size_t data = size_t(d.rows) * size_t(d.cols) * size_t(d.elem_bytes);
  return data + data / 32 + 4096;  // buffer + an upper bound on its visited bits
}

// Grows a server-owned buffer to at least `want` bytes (a quarter more, at least 1 MiB, for
// small ones; device buffers of 64 MiB and more exactly, so large scratch holds no slack);
// `held` tracks the device bytes the server holds.
bool grow(void **p, size_t *cap, size_t want, bool pinned, std::atomic<size_t> *held) {
  if (*cap >= want) return true;
  size_t n = want >= (size_t(64) << 20) && !pinned ? want
                                                   : std::max<size_t>(want + want / 4, size_t(1) << 20);
  if (*p) {
    if (pinned) (void)hipHostFree(*p);
    else (void)hipFree(*p);
    if (!pinned && held) *held -= *cap;
    *p = nullptr;
    *cap = 0;
  }
  hipError_t e = pinned ? hipHostMalloc(p, n, hipHostMallocDefault) : hipMalloc(p, n);
  if (e != hipSuccess) return false;
  *cap = n;
  if (!pinned && held) *held += n;
  return true;
}

// Growth of a server buffer of capacity `cap` to hold `want` bytes (0 when it fits).
size_t growth(size_t cap, size_t want) { return want > cap ? want : 0; }

// Plans and launches one dynamic batch on an idle slot. Called without the lock held.
bool launch_batch(Server &s, Slot &sl, const std::vector<hip_marshal_desc> &descs) {
  BatchPlan plan;
  if (!plan_batch(descs.data(), int64_t(descs.size()), plan, "hip_marshal_server")) return false;
  if (!grow(reinterpret_cast<void **>(&sl.flags), &sl.flags_cap, plan.flag_words * 4 + 4, false,
            &s.held) ||
      !grow(&sl.params, &sl.params_cap, plan.param_bytes + 256, false, &s.held) ||
      !grow(reinterpret_cast<void **>(&sl.host), &sl.host_cap, plan.param_bytes + 256, true,
            nullptr))
    return false;
  std::vector<unsigned char> img;
  build_image(plan, sl.flags, img);
  if (!img.empty()) {
    std::memcpy(sl.host, img.data(), img.size());
    if (hipMemcpyAsync(sl.params, sl.host, img.size(), hipMemcpyHostToDevice, sl.st) !=
        hipSuccess)
      return false;
  }
  if (enqueue_batch(plan, sl.params, sl.flags, sl.st)) return false;
  return hipEventRecord(sl.ev, sl.st) == hipSuccess;
}

// A large (or lone) request runs with the stage list its plain entry point would use (stripe
// route included), on the slot's stream with the slot's own scratch: nothing to amortise,
// and per-request kernels beat the grouped forms there.
bool launch_solo(Server &s, Slot &sl, const std::vector<StageRec> &plain, size_t need,
                 bool large, bool alone) {
  unsigned char *scr = nullptr;
  if (large) {  // one large request at a time: the shared scratch
    if (need && !grow(reinterpret_cast<void **>(&s.solo_scr), &s.solo_cap, need, false, &s.held))
      return false;
    scr = s.solo_scr;
  } else {
    if (need && !grow(reinterpret_cast<void **>(&sl.scr), &sl.scr_cap, need, false, &s.held))
      return false;
    scr = sl.scr;
  }
  if (run_stages_on(plain, sl.st, scr, alone)) return false;
  return hipEventRecord(sl.ev, sl.st) == hipSuccess;
}

// Starts `ids` (already taken off the queue) on the idle slot `sl`. Called with the lock held;
// releases it while the work is enqueued and returns with it held. `plain`: one request on
// its plain stage list, else a dynamic batch of grouped kernels.
void start_on_slot(Server &s, Slot &sl, const std::vector<int64_t> &ids,
                   const std::vector<hip_marshal_desc> &descs, size_t add, bool plain,
                   bool solo, std::unique_lock<std::mutex> &lk) {
  const double t = now_s();
  for (int64_t id : ids) {
    s.reqs[id].state = 1;
    s.reqs[id].t_launch = t;
  }
  s.inflight += add;
  sl.busy = true;
  sl.launching = true;
  sl.solo = solo;
  sl.bytes = add;
  sl.reqs = ids;
  if (solo) s.solo_running++;
  bool alone = true;  // no other stream busy: the request has the device to itself
  for (const Slot &o : s.slots)
    if (&o != &sl && o.busy) alone = false;
  // copied while the lock is held: the request map may change once it is released
  std::vector<StageRec> plain_list;
  size_t plain_scratch = 0;
  if (plain) {
    plain_list = s.reqs[ids[0]].plain;
    plain_scratch = s.reqs[ids[0]].scratch;
  }
  s.batches++;
  s.max_in_batch = std::max<int64_t>(s.max_in_batch, int64_t(ids.size()));
  lk.unlock();
  int prev = -1;
  (void)hipGetDevice(&prev);
  // the server's device must be current for its launches; if it cannot be made so, the
  // requests fail rather than run elsewhere
  const bool dev_ok = prev == s.device || hipSetDevice(s.device) == hipSuccess;
  const bool ok = dev_ok && (plain ? launch_solo(s, sl, plain_list, plain_scratch, solo, alone)
                                   : launch_batch(s, sl, descs));
  if (prev != s.device && prev >= 0) (void)hipSetDevice(prev);
  lk.lock();
  sl.launching = false;
  if (!ok) {  // mark failed once the stream has drained
    (void)hipStreamSynchronize(sl.st);
    for (int64_t id : ids) {
      s.reqs[id].state = -1;
      s.failed++;
    }
    s.inflight -= add;
    if (sl.solo) s.solo_running--;
    sl.busy = false;
    sl.solo = false;
    sl.bytes = 0;
    sl.reqs.clear();
    s.cv.notify_all();
  }
}

// Retires finished slots. Lock held.
bool retire(Server &s) {
  bool progressed = false;
  for (Slot &sl : s.slots) {
    if (!sl.busy || sl.launching) continue;
    hipError_t q = hipEventQuery(sl.ev);
    if (q == hipErrorNotReady) continue;
    double t = now_s();
    for (int64_t id : sl.reqs) {
      Req &r = s.reqs[id];
      r.t_done = t;
      r.state = q == hipSuccess ? 2 : -1;
      if (q == hipSuccess) {
        s.lat.push_back(t - r.t_arrive);
        s.bytes_done += 2.0 * double(r.d.rows) * double(r.d.cols) * double(r.d.elem_bytes);
      } else {
        s.failed++;
      }
      s.t_last = std::max(s.t_last, t);
    }
    s.inflight -= sl.bytes;
    if (sl.solo) s.solo_running--;
    sl.bytes = 0;
    sl.reqs.clear();
    sl.busy = false;
    sl.solo = false;
    progressed = true;
  }
  if (progressed) s.cv.notify_all();
  return progressed;
}

// Frees the buffers of idle streams (and the shared large-request scratch when no large
// request runs), returning their bytes to the budget. Lock held; called only when nothing runs.
void release_idle(Server &s) {
  for (Slot &sl : s.slots) {
    if (sl.busy) continue;
    if (sl.flags) (void)hipFree(sl.flags);
    if (sl.params) (void)hipFree(sl.params);
    if (sl.scr) (void)hipFree(sl.scr);
    s.held -= sl.flags_cap + sl.params_cap + sl.scr_cap;
    sl.flags = nullptr;
    sl.params = nullptr;
    sl.scr = nullptr;
    sl.flags_cap = sl.params_cap = sl.scr_cap = 0;
  }
  if (s.solo_running == 0 && s.solo_scr) {
    (void)hipFree(s.solo_scr);
    s.held -= s.solo_cap;
    s.solo_scr = nullptr;
    s.solo_cap = 0;
  }
}

// Admits queued requests onto idle slots while the budget allows. Lock held (released while
// launching). Order: the queue head if it can start; a large request (>= solo_bytes) waits
// while another large one runs, and small requests behind it may then go ahead of it, so
// large requests never share the GPU with each other and small ones never wait behind them.
// Small requests taken together form one dynamic batch; one taken alone runs on its plain
// stage list.
void admit(Server &s, std::unique_lock<std::mutex> &lk) {
  while (!s.queue.empty()) {
    Slot *free_slot = nullptr;
    for (Slot &sl : s.slots)
      if (!sl.busy) {
        free_slot = &sl;
        break;
      }
    if (!free_slot) return;
    bool idle = s.inflight == 0;
    for (const Slot &sl : s.slots) idle = idle && !sl.busy;
    std::vector<int64_t> ids;
    std::vector<hip_marshal_desc> descs;
    size_t add = 0;
    bool solo = false;
    const Req &head = s.reqs[s.queue.front()];
    size_t held = s.held.load();
    if (idle) {
      // Nothing runs: the head request always fits alone (checked at submission), so make
      // room for it by freeing idle buffers if the server's own allocations are in the way.
      if (head.bytes + head.scratch + held > s.budget) {
        release_idle(s);
        held = s.held.load();
      }
    }
    if (head.bytes >= s.solo_bytes && s.solo_running == 0) {
      // its buffer, and the shared scratch grown to what it needs
      if (!idle && s.inflight + head.bytes + held + growth(s.solo_cap, head.scratch) > s.budget)
        return;  // wait for budget to free up (an idle server always starts the head)
      solo = true;
      ids.push_back(s.queue.front());
      descs.push_back(head.d);
      add = head.bytes;
      s.queue.pop_front();
    } else {
      for (auto it = s.queue.begin(); it != s.queue.end() && int(ids.size()) < s.max_batch;) {
        const Req &r = s.reqs[*it];
        if (r.bytes >= s.solo_bytes) {  // a large one waiting for its turn stays queued
          ++it;
          continue;
        }
        // (a request taken alone runs on its plain list with the stream's own scratch)
        const size_t extra = ids.empty() ? growth(free_slot->scr_cap, r.scratch) : 0;
        if (s.inflight + add + r.bytes + held + extra > s.budget && !(idle && ids.empty()))
          break;
        add += r.bytes;
        ids.push_back(*it);
        descs.push_back(r.d);
        it = s.queue.erase(it);
      }
      if (ids.empty()) return;
    }
    start_on_slot(s, *free_slot, ids, descs, add, solo || ids.size() == 1, solo, lk);
  }
}

void scheduler(Server *sp) {
  Server &s = *sp;
  (void)hipSetDevice(s.device);
  std::unique_lock<std::mutex> lk(s.mu);
  while (true) {
    retire(s);
    admit(s, lk);
    bool any_busy = false;
    for (Slot &sl : s.slots) any_busy |= sl.busy;
    if (s.stop && s.queue.empty() && !any_busy) break;
    if (any_busy) {
      // Work on the GPU: poll its completion closely (a timed wait oversleeps by tens of
      // microseconds), at the cost of one host core while requests are in flight.
      lk.unlock();
      for (int i = 0; i < 256; i++) __builtin_ia32_pause();
      lk.lock();
    } else {
      s.cv.wait_for(lk, std::chrono::milliseconds(1));  // woken by a submission
    }
  }
}

}  // namespace
}  // namespace lm

using namespace lm;

struct hip_marshal_server {
  Server s;
};

extern "C" {

hip_marshal_server *hip_marshal_server_create(size_t hbm_budget_bytes, int streams,
                                              int max_batch) {
  if (hbm_budget_bytes == 0 || streams < 1 || streams > 16 || max_batch < 1) {
    fprintf(stderr, "[libmarshal] hip_marshal_server_create: bad arguments\n");
    return nullptr;
  }
  auto *h = new hip_marshal_server();
  Server &s = h->s;
  s.budget = hbm_budget_bytes;
  s.max_batch = max_batch;
  s.scratch_cap = uint64_t(tuning("serve_scratch_mb")) << 20;
  (void)hipGetDevice(&s.device);
  s.slots.resize(size_t(streams));
  for (Slot &sl : s.slots) {
    hipError_t e = hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming);
    if (e != hipSuccess) {
      fprintf(stderr, "[libmarshal] hip_marshal_server_create: stream creation failed: %s\n",
              hipGetErrorString(e));
      delete h;
      return nullptr;
    }
  }
  s.th = std::thread(scheduler, &s);
  return h;
}

double hip_marshal_server_now(void) { return now_s(); }

int64_t hip_marshal_server_submit_at(hip_marshal_server *h, const hip_marshal_desc *d,
                                     double arrival) {
  if (!h || !d) return -1;
  Server &s = h->s;
  size_t es = size_t(d->elem_bytes);
  bool ok = (es == 1 || es == 2 || es == 4 || es == 8 || es == 16) && d->ptr && d->rows >= 1 &&
            d->cols >= 1 && d->op >= 0 && d->op <= 4 &&
            reinterpret_cast<uintptr_t>(d->ptr) % std::min<size_t>(es, 16) == 0;
  if (ok && d->op != HIP_MARSHAL_OP_TRANSPOSE) ok = d->tile >= 1 && d->rows % d->tile == 0;
  size_t fp = ok ? footprint(*d) : 0;
  // The plain stage list, planned here once: in-place and through-scratch forms with scratch
  // capped per request, no segment tables (their first upload would stall the other streams).
  std::vector<StageRec> plain;
  if (ok) {
    RestrictedPlanning no_tables;
    ScratchCapScope cap(s.scratch_cap);
    std::string route;
    ok = plan_desc(*d, plain, route, true);
  }
  const size_t scr = ok ? stages_scratch_bytes(plain) : 0;
  if (!ok || fp + scr > s.budget) {
    fprintf(stderr, "[libmarshal] hip_marshal_server_submit: invalid request or larger than "
                    "the HBM budget\n");
    return -1;
  }
  std::unique_lock<std::mutex> lk(s.mu);
  int64_t id = s.next_id++;
  Req &r = s.reqs[id];
  r.d = *d;
  r.bytes = fp;
  r.scratch = scr;
  r.plain = std::move(plain);
  r.t_arrive = arrival > 0 ? arrival : now_s();
  if (s.t_first < 0 || r.t_arrive < s.t_first) s.t_first = r.t_arrive;
  s.queue.push_back(id);
  // Start it right here when a stream is idle: no hand-over to the scheduler thread.
  admit(s, lk);
  s.cv.notify_all();
  return id;
}

int64_t hip_marshal_server_submit(hip_marshal_server *h, const hip_marshal_desc *d) {
  return hip_marshal_server_submit_at(h, d, 0.0);
}

int hip_marshal_server_status(hip_marshal_server *h, int64_t id) {
  if (!h) return -1;
  std::lock_guard<std::mutex> lk(h->s.mu);
  auto it = h->s.reqs.find(id);
  return it == h->s.reqs.end() ? -1 : it->second.state;
}

bool hip_marshal_server_wait(hip_marshal_server *h, int64_t id, double timeout_s) {
  if (!h) return true;
  Server &s = h->s;
  std::unique_lock<std::mutex> lk(s.mu);
  auto done = [&] {
    auto it = s.reqs.find(id);
    return it == s.reqs.end() || it->second.state == 2 || it->second.state == -1;
  };
  if (!s.cv.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) return true;
  auto it = s.reqs.find(id);
  return it == s.reqs.end() || it->second.state != 2;
}

bool hip_marshal_server_drain(hip_marshal_server *h, double timeout_s) {
  if (!h) return true;
  Server &s = h->s;
  std::unique_lock<std::mutex> lk(s.mu);
  auto idle = [&] {
    if (!s.queue.empty()) return false;
    for (Slot &sl : s.slots)
      if (sl.busy) return false;
    return true;
  };
  return !s.cv.wait_for(lk, std::chrono::duration<double>(timeout_s), idle);
}

bool hip_marshal_server_request_times(hip_marshal_server *h, int64_t id, double *arrival,
                                      double *launch, double *done) {
  if (!h) return true;
  std::lock_guard<std::mutex> lk(h->s.mu);
  auto it = h->s.reqs.find(id);
  if (it == h->s.reqs.end()) return true;
  if (arrival) *arrival = it->second.t_arrive;
  if (launch) *launch = it->second.t_launch;
  if (done) *done = it->second.t_done;
  return false;
}

void hip_marshal_server_get_stats(hip_marshal_server *h, hip_marshal_server_stats *out) {
  if (!h || !out) return;
  Server &s = h->s;
  std::lock_guard<std::mutex> lk(s.mu);
  std::memset(out, 0, sizeof(*out));
  out->completed = int64_t(s.lat.size());
  out->failed = s.failed;
  out->batches = s.batches;
  out->max_batch = s.max_in_batch;
  out->bytes = s.bytes_done;
  out->span_s = s.t_first >= 0 && s.t_last > s.t_first ? s.t_last - s.t_first : 0;
  if (!s.lat.empty()) {
    std::vector<double> v = s.lat;
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) { return v[std::min(v.size() - 1, size_t(p * (v.size() - 1) + 0.5))]; };
    double sum = 0;
    for (double x : v) sum += x;
    out->p50_ms = pct(0.50) * 1e3;
    out->p99_ms = pct(0.99) * 1e3;
    out->mean_ms = sum / double(v.size()) * 1e3;
    out->max_ms = v.back() * 1e3;
  }
}

void hip_marshal_server_reset_stats(hip_marshal_server *h) {
  if (!h) return;
  Server &s = h->s;
  std::lock_guard<std::mutex> lk(s.mu);
  s.lat.clear();
  s.bytes_done = 0;
  s.batches = s.failed = s.max_in_batch = 0;
  s.t_first = -1;
  s.t_last = 0;
  for (auto it = s.reqs.begin(); it != s.reqs.end();) {
    if (it->second.state == 2 || it->second.state == -1) it = s.reqs.erase(it);
    else ++it;
  }
}

void hip_marshal_server_destroy(hip_marshal_server *h) {
  if (!h) return;
  Server &s = h->s;
  {
    std::lock_guard<std::mutex> lk(s.mu);
    s.stop = true;
    s.cv.notify_all();
  }
  if (s.th.joinable()) s.th.join();
  if (s.solo_scr) (void)hipFree(s.solo_scr);
  for (Slot &sl : s.slots) {
    if (sl.st) (void)hipStreamSynchronize(sl.st);
    if (sl.flags) (void)hipFree(sl.flags);
    if (sl.scr) (void)hipFree(sl.scr);
    if (sl.params) (void)hipFree(sl.params);
    if (sl.host) (void)hipHostFree(sl.host);
    sl.flags = nullptr;
    sl.scr = nullptr;
    sl.params = nullptr;
    if (sl.ev) (void)hipEventDestroy(sl.ev);
    if (sl.st) (void)hipStreamDestroy(sl.st);
  }
  delete h;
}

}  // extern "C"
