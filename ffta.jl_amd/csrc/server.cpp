// server.cpp -- continuous-batching request loop for fft/bfft requests.
//
// Requests (any precision, direction and shape the plan interface accepts)
// are submitted from any thread and served by one dispatcher thread:
//   * admission by a device memory pool: each in-flight request holds its
//     plan scratch (and, for host payloads, its device staging) in an arena
//     carved from one allocation made at creation, so the loop never calls
//     hipMalloc; a request larger than the whole pool is rejected;
//   * continuous batching: requests that arrive while others run are admitted
//     as soon as pool space and a stream slot free up, in arrival order;
//     requests of one 1-D plan pending together run as one coalesced group
//     (the lines of one batch, one set of work-list launches);
//   * overlap: micro-batches go round-robin to a set of non-blocking HIP
//     streams, so transfers and kernels of different requests overlap;
//   * completion: an event per micro-batch, polled by the dispatcher; the
//     latency of a request runs from its (scheduled) arrival to the moment
//     the dispatcher sees its event complete.
// Plans are cached per (precision, direction, shape) and shared by all
// streams (executions use caller-provided scratch, so they are reentrant).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/fftahip.h"
#include "plan_internal.h"

namespace {

using fftahip::set_error;
using Clock = std::chrono::steady_clock;

double now_s() { return std::chrono::duration<double>(Clock::now().time_since_epoch()).count(); }

// First-fit arena over [0, size) with coalescing, 256-byte granules.
class Arena {
 public:
  explicit Arena(size_t size) : size_(size) { free_[0] = size; }
  bool alloc(size_t bytes, size_t* off) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes;
        const size_t start = it->first + bytes;
        free_.erase(it);
        if (rest) free_[start] = rest;
        used_ += bytes;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }
  size_t size() const { return size_; }

 private:
  size_t size_;
  size_t used_ = 0;
  std::map<size_t, size_t> free_;
};

struct Request {
  uint64_t id = 0;
  int prec = 32, dir = -1, ndims = 1;
  int64_t dims[8] = {0};
  const void* in = nullptr;
  void* out = nullptr;
  bool host = false;
  double t_arrive = 0;
  size_t bytes = 0;  // payload bytes (one array)
  bool done = false;
  fftahip_status status = FFTAHIP_SUCCESS;
  double latency_ms = 0;
};

struct Flight {
  std::vector<Request*> reqs;
  int stream = 0;
  hipEvent_t ev = nullptr;
  size_t off = 0, bytes = 0;  // arena allocation
};

using PlanKey = std::vector<int64_t>;

}  // namespace

struct fftahip_server_s {
  int device = 0;
  void* pool = nullptr;
  std::unique_ptr<Arena> arena;
  std::vector<hipStream_t> streams;
  std::vector<int> stream_inflight;
  int max_per_stream = 8;
  std::map<PlanKey, fftahip_plan> plans;
  std::map<PlanKey, size_t> scratch_bytes;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> pending;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  std::vector<Flight> inflight;
  std::atomic<uint64_t> next_id{1};
  bool stop = false;
  std::thread worker;

  // stats
  std::vector<double> lat_ms;
  uint64_t completed = 0, rejected = 0, failed = 0;
  // requests that ran inside coalesced groups, and the number of groups
  uint64_t coalesced = 0, groups = 0;
  bool coalesce = true;
  double t_first = 0, t_last = 0;
  double payload_bytes = 0;
  size_t peak_pool = 0;
};

namespace {

PlanKey key_of(const Request& r) {
  PlanKey k = {r.prec, r.dir, r.ndims};
  for (int i = 0; i < r.ndims; ++i) k.push_back(r.dims[i]);
  return k;
}

// Requests of one plan that may run as one coalesced launch set: 1-D with
// device payloads.  A group is the largest power of two (<= kMaxGroup) of
// such requests among the first kScan pending ones, so that each plan keeps
// at most five many-vector schedules.
constexpr size_t kScan = 256;
constexpr size_t kMaxGroup = 32;

bool coalescable(const Request& r) { return r.ndims == 1 && !r.host; }
bool same_plan(const Request& a, const Request& b) {
  return a.prec == b.prec && a.dir == b.dir && a.ndims == b.ndims && a.dims[0] == b.dims[0];
}

// Plan and scratch size for a request's shape (created on first use).  Only
// the dispatcher thread inserts plans, so a new plan is built with the lock
// released (`lk`, held on entry and on return): plan creation runs table
// kernels and synchronises (milliseconds for a new Bluestein length), and
// submitters, waiters and stats readers must not stall behind it.
fftahip_status get_plan(fftahip_server_s* S, const Request& r, fftahip_plan* p, size_t* scratch,
                        std::unique_lock<std::mutex>& lk) {
  PlanKey k = key_of(r);
  auto it = S->plans.find(k);
  if (it == S->plans.end()) {
    std::vector<int64_t> region(r.ndims);
    for (int i = 0; i < r.ndims; ++i) region[i] = i + 1;
    fftahip_plan plan = nullptr;
    size_t sb = 0;
    lk.unlock();
    fftahip_status st = fftahip_plan_create(&plan, r.prec, r.dir, r.ndims, r.dims, r.ndims, region.data(),
                                            FFTAHIP_DEFAULT_BLUESTEIN_CUTOFF);
    if (!st) st = fftahip::plan_scratch_bytes(plan, r.ndims, r.dims, &sb);
    // the schedules of coalesced groups of this plan, built now rather than
    // on first use inside admission
    if (!st && coalescable(r))
      for (size_t K = 2; K <= kMaxGroup && !st; K *= 2) {
        size_t a = 0, b = 0;
        st = fftahip::plan_many_bytes(plan, (int)K, &a, &b, nullptr);
      }
    if (st && plan) fftahip_plan_destroy(plan);
    lk.lock();
    if (st) return st;
    it = S->plans.emplace(k, plan).first;
    S->scratch_bytes[k] = sb;
  }
  *p = it->second;
  *scratch = S->scratch_bytes[k];
  return FFTAHIP_SUCCESS;
}

size_t need_bytes(const Request& r, size_t scratch) {
  size_t b = scratch;
  if (r.host) b += 2 * ((r.bytes + 255) & ~(size_t)255);
  return b;
}

// Issue one admitted request on a stream.  Returns a failure status if the
// launch could not be issued (the request is then completed with it).
fftahip_status issue(fftahip_server_s* S, Request* r, fftahip_plan plan, size_t scratch, char* base,
                     hipStream_t s) {
  char* sc = base;
  const void* din = r->in;
  void* dout = r->out;
  if (r->host) {
    char* dx = base + scratch;
    char* dy = dx + ((r->bytes + 255) & ~(size_t)255);
    if (hipMemcpyAsync(dx, r->in, r->bytes, hipMemcpyHostToDevice, s) != hipSuccess)
      return set_error(FFTAHIP_ERR_HIP, "hipMemcpyAsync H2D failed");
    din = dx;
    dout = dy;
  }
  fftahip_status st = fftahip::plan_execute_scratch(plan, r->ndims, r->dims, din, dout, scratch ? sc : nullptr, s);
  if (st) return st;
  if (r->host && hipMemcpyAsync(r->out, dout, r->bytes, hipMemcpyDeviceToHost, s) != hipSuccess)
    return set_error(FFTAHIP_ERR_HIP, "hipMemcpyAsync D2H failed");
  (void)S;
  return FFTAHIP_SUCCESS;
}

void finish(fftahip_server_s* S, Request* r, fftahip_status st, double t) {
  r->status = st;
  r->latency_ms = (t - r->t_arrive) * 1e3;
  r->done = true;
  if (st == FFTAHIP_SUCCESS) {
    S->completed++;
    S->lat_ms.push_back(r->latency_ms);
    S->payload_bytes += 2.0 * r->bytes;
  } else if (st == FFTAHIP_ERR_CAPACITY) {
    S->rejected++;
  } else {
    S->failed++;
  }
  S->t_last = std::max(S->t_last, t);
}

void worker_loop(fftahip_server_s* S) {
  (void)hipSetDevice(S->device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    // 1. poll completions
    const double t = now_s();
    for (size_t i = 0; i < S->inflight.size();) {
      Flight& f = S->inflight[i];
      if (hipEventQuery(f.ev) == hipSuccess) {
        for (Request* r : f.reqs) finish(S, r, FFTAHIP_SUCCESS, t);
        S->arena->release(f.off, f.bytes);
        S->stream_inflight[f.stream]--;
        (void)hipEventDestroy(f.ev);
        S->inflight[i] = S->inflight.back();
        S->inflight.pop_back();
        S->cv_done.notify_all();
      } else {
        ++i;
      }
    }
    // 2. admit pending requests in arrival order while pool and streams allow
    while (!S->pending.empty()) {
      Request* r = S->pending.front();
      fftahip_plan plan = nullptr;
      size_t scratch = 0;
      fftahip_status st = get_plan(S, *r, &plan, &scratch, lk);
      if (st) {
        S->pending.pop_front();
        finish(S, r, st, now_s());
        S->cv_done.notify_all();
        continue;
      }
      const size_t need1 = need_bytes(*r, scratch);
      if (need1 > S->arena->size()) {
        S->pending.pop_front();
        set_error(FFTAHIP_ERR_CAPACITY, "request exceeds the device pool");
        finish(S, r, FFTAHIP_ERR_CAPACITY, now_s());
        S->cv_done.notify_all();
        continue;
      }
      // stream with the fewest micro-batches in flight
      int sidx = 0;
      for (int i = 1; i < (int)S->streams.size(); ++i)
        if (S->stream_inflight[i] < S->stream_inflight[sidx]) sidx = i;
      if (S->stream_inflight[sidx] >= S->max_per_stream) break;
      hipStream_t s = S->streams[sidx];
      // coalesce the requests of r's plan pending behind it
      std::vector<Request*> group{r};
      if (S->coalesce && coalescable(*r)) {
        size_t scanned = 0;
        for (auto it = S->pending.begin() + 1;
             it != S->pending.end() && scanned < kScan && group.size() < kMaxGroup; ++it, ++scanned)
          if (coalescable(**it) && same_plan(**it, *r)) group.push_back(*it);
      }
      size_t K = 1;
      while (K * 2 <= group.size()) K *= 2;
      group.resize(K);
      size_t need = need1, scr_many = 0, desc_many = 0;
      if (K > 1) {
        // the first use of a group size builds its schedule (and may wait
        // for table kernels): without the lock, as plan creation
        lk.unlock();
        const fftahip_status ms = fftahip::plan_many_bytes(plan, (int)K, &scr_many, &desc_many, s);
        lk.lock();
        if (ms) {
          K = 1;
          group.resize(1);
        } else {
          need = ((scr_many + 255) & ~(size_t)255) + desc_many;
        }
      }
      size_t off = 0;
      if (!S->arena->alloc(need, &off)) break;  // wait for memory to free up
      S->peak_pool = std::max(S->peak_pool, S->arena->used());
      Flight f;
      f.stream = sidx;
      f.off = off;
      f.bytes = need;
      f.reqs = group;
      S->pending.pop_front();
      for (size_t g = 1; g < K; ++g) S->pending.erase(std::find(S->pending.begin(), S->pending.end(), group[g]));
      if (K > 1) {
        std::vector<const void*> ins(K);
        std::vector<void*> outs(K);
        for (size_t g = 0; g < K; ++g) {
          ins[g] = group[g]->in;
          outs[g] = group[g]->out;
        }
        char* base = (char*)S->pool + off;
        st = fftahip::plan_execute_many(plan, (int)K, ins.data(), outs.data(), scr_many ? base : nullptr,
                                        base + ((scr_many + 255) & ~(size_t)255), s);
        S->coalesced += K;
        S->groups++;
      } else {
        st = issue(S, r, plan, scratch, (char*)S->pool + off, s);
      }
      if (st) {
        (void)hipStreamSynchronize(s);
        S->arena->release(off, need);
        for (Request* q : group) finish(S, q, st, now_s());
        S->cv_done.notify_all();
        continue;
      }
      (void)hipEventCreateWithFlags(&f.ev, hipEventDisableTiming);
      (void)hipEventRecord(f.ev, s);
      S->stream_inflight[sidx]++;
      S->inflight.push_back(f);
    }
    if (S->stop && S->pending.empty() && S->inflight.empty()) break;
    if (S->inflight.empty()) {
      S->cv_work.wait(lk, [S] { return S->stop || !S->pending.empty(); });
    } else {
      // poll events; release the lock briefly so submitters can enqueue
      lk.unlock();
      std::this_thread::yield();
      lk.lock();
    }
  }
}

}  // namespace

extern "C" {

fftahip_status fftahip_server_create(fftahip_server* out, int64_t pool_bytes, int nstreams) {
  if (!out) return set_error(FFTAHIP_ERR_ARGUMENT, "server pointer is NULL");
  *out = nullptr;
  if (pool_bytes <= 0) return set_error(FFTAHIP_ERR_ARGUMENT, "pool_bytes must be > 0");
  if (nstreams <= 0) nstreams = 4;
  auto S = std::make_unique<fftahip_server_s>();
  (void)hipGetDevice(&S->device);
  {
    const char* v = std::getenv("FFTAHIP_COALESCE");
    S->coalesce = !(v && v[0] == '0');
  }
  hipError_t e = hipMalloc(&S->pool, (size_t)pool_bytes);
  if (e != hipSuccess) return set_error(FFTAHIP_ERR_ALLOC, std::string("server pool: ") + hipGetErrorString(e));
  S->arena = std::make_unique<Arena>((size_t)pool_bytes);
  for (int i = 0; i < nstreams; ++i) {
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
      for (hipStream_t q : S->streams) (void)hipStreamDestroy(q);
      (void)hipFree(S->pool);
      return set_error(FFTAHIP_ERR_HIP, "hipStreamCreate failed");
    }
    S->streams.push_back(s);
    S->stream_inflight.push_back(0);
  }
  fftahip_server_s* raw = S.get();
  S->worker = std::thread(worker_loop, raw);
  *out = S.release();
  return FFTAHIP_SUCCESS;
}

static fftahip_status submit_at(fftahip_server s, int precision, int direction, int ndims, const int64_t* dims,
                                const void* in, void* out, int host, double t_arrive, uint64_t* ticket) {
  if (!s || !ticket) return set_error(FFTAHIP_ERR_ARGUMENT, "NULL argument");
  if (precision != 32 && precision != 64) return set_error(FFTAHIP_ERR_ARGUMENT, "precision must be 32 or 64");
  if (direction != -1 && direction != 1) return set_error(FFTAHIP_ERR_ARGUMENT, "direction must be -1 or +1");
  if (ndims < 1 || ndims > 8) return set_error(FFTAHIP_ERR_ARGUMENT, "1 to 8 dimensions");
  auto r = std::make_unique<Request>();
  r->prec = precision;
  r->dir = direction;
  r->ndims = ndims;
  int64_t n = 1;
  for (int i = 0; i < ndims; ++i) {
    if (dims[i] <= 0) return set_error(FFTAHIP_ERR_DIMENSION_MISMATCH, "Array length must be strictly positive");
    r->dims[i] = dims[i];
    n *= dims[i];
  }
  r->bytes = (size_t)n * (precision == 32 ? 8 : 16);
  r->in = in;
  r->out = out;
  r->host = host != 0;
  r->t_arrive = t_arrive;
  r->id = s->next_id++;
  *ticket = r->id;
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->t_first == 0 || t_arrive < s->t_first) s->t_first = t_arrive;
    s->pending.push_back(r.get());
    s->reqs[r->id] = std::move(r);
  }
  s->cv_work.notify_one();
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_server_submit(fftahip_server s, int precision, int direction, int ndims, const int64_t* dims,
                                     const void* in, void* out, int host_buffers, uint64_t* ticket) {
  return submit_at(s, precision, direction, ndims, dims, in, out, host_buffers, now_s(), ticket);
}

fftahip_status fftahip_server_replay(fftahip_server s, int64_t count, const double* t_offset, const int* precision,
                                     const int* direction, const int* ndims, const int64_t* dims, void* const* in,
                                     void* const* out, int host_buffers, uint64_t* tickets) {
  if (!s) return set_error(FFTAHIP_ERR_ARGUMENT, "NULL server");
  if (count < 0 || (count > 0 && (!t_offset || !precision || !direction || !ndims || !dims || !in || !out ||
                                   !tickets)))
    return set_error(FFTAHIP_ERR_ARGUMENT, "NULL argument");
  // dims holds 3 entries per request
  for (int64_t i = 0; i < count; ++i)
    if (ndims[i] < 1 || ndims[i] > 3) return set_error(FFTAHIP_ERR_ARGUMENT, "replay requests have 1 to 3 dimensions");
  const double t0 = now_s();
  for (int64_t i = 0; i < count; ++i) {
    const double due = t0 + t_offset[i];
    while (now_s() < due) {
    }
    fftahip_status st = submit_at(s, precision[i], direction[i], ndims[i], dims + 3 * i, in[i], out[i], host_buffers,
                                  due, &tickets[i]);
    if (st) return st;
  }
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_server_test(fftahip_server s, uint64_t ticket, int* done) {
  if (!s || !done) return set_error(FFTAHIP_ERR_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return set_error(FFTAHIP_ERR_ARGUMENT, "unknown ticket");
  *done = it->second->done ? 1 : 0;
  return it->second->done ? it->second->status : FFTAHIP_SUCCESS;
}

fftahip_status fftahip_server_wait(fftahip_server s, uint64_t ticket, double* latency_ms) {
  if (!s) return set_error(FFTAHIP_ERR_ARGUMENT, "NULL server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return set_error(FFTAHIP_ERR_ARGUMENT, "unknown ticket");
  Request* r = it->second.get();
  s->cv_done.wait(lk, [r] { return r->done; });
  if (latency_ms) *latency_ms = r->latency_ms;
  if (r->status == FFTAHIP_ERR_CAPACITY) return set_error(FFTAHIP_ERR_CAPACITY, "request exceeds the device pool");
  return r->status;
}

fftahip_status fftahip_server_drain(fftahip_server s) {
  if (!s) return set_error(FFTAHIP_ERR_ARGUMENT, "NULL server");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [s] { return s->pending.empty() && s->inflight.empty(); });
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_server_stats_get(fftahip_server s, fftahip_server_stats* out) {
  if (!s || !out) return set_error(FFTAHIP_ERR_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> g(s->mu);
  std::vector<double> v = s->lat_ms;
  std::sort(v.begin(), v.end());
  auto pct = [&](double q) {
    if (v.empty()) return 0.0;
    size_t i = (size_t)std::min<double>(v.size() - 1, std::max(0.0, q * v.size() - 1e-9));
    return v[i];
  };
  std::memset(out, 0, sizeof(*out));
  out->completed = s->completed;
  out->rejected = s->rejected;
  out->failed = s->failed;
  out->pending = s->pending.size();
  out->p50_ms = pct(0.50);
  out->p99_ms = pct(0.99);
  out->max_ms = v.empty() ? 0 : v.back();
  double sum = 0;
  for (double x : v) sum += x;
  out->mean_ms = v.empty() ? 0 : sum / v.size();
  const double span = s->t_last - s->t_first;
  out->seconds = span;
  out->requests_per_s = span > 0 ? s->completed / span : 0;
  out->payload_gbps = span > 0 ? s->payload_bytes / span / 1e9 : 0;
  out->pool_peak_bytes = (int64_t)s->peak_pool;
  out->plans = (int64_t)s->plans.size();
  out->coalesced = (int64_t)s->coalesced;
  out->groups = (int64_t)s->groups;
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_server_reset_stats(fftahip_server s) {
  if (!s) return set_error(FFTAHIP_ERR_ARGUMENT, "NULL server");
  std::lock_guard<std::mutex> g(s->mu);
  s->lat_ms.clear();
  s->completed = s->rejected = s->failed = 0;
  s->coalesced = s->groups = 0;
  s->t_first = s->t_last = 0;
  s->payload_bytes = 0;
  s->peak_pool = s->arena->used();
  for (auto it = s->reqs.begin(); it != s->reqs.end();) {
    if (it->second->done) it = s->reqs.erase(it);
    else ++it;
  }
  return FFTAHIP_SUCCESS;
}

fftahip_status fftahip_server_destroy(fftahip_server s) {
  if (!s) return FFTAHIP_SUCCESS;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (hipStream_t q : s->streams) (void)hipStreamDestroy(q);
  for (auto& kv : s->plans) fftahip_plan_destroy(kv.second);
  (void)hipFree(s->pool);
  delete s;
  return FFTAHIP_SUCCESS;
}

}  // extern "C"
