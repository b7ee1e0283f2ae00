// serve.cpp -- the continuous-batching loop of the served Poisson solves,
// native (no counterpart in the reference).
//
// Requests arrive over time with their right-hand sides resident in HBM.  A
// request is admitted once it has arrived, a stream slot is free and its work
// area fits a first-fit arena over one caller-given HBM region; its solve
// (p3dfft_poisson_solve) is queued on the slot's stream, so admitted solves
// of different shapes overlap on the device.  When the oldest waiting request
// does not fit, up to `lookahead` later arrived requests that do may go ahead
// of it, at most `max_bypass` times per waiting request (no head-of-line
// blocking, no starvation).  Completion is polled with HIP events, whose GPU
// timestamps give each request's latency (arrival to completion).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "p3dfft_amd_internal.hpp"

using namespace p3amd;

extern "C" int p3dfft_poisson_solve(int nx, int ny, int nz, p3amd::real_t* field, void* work, size_t work_bytes,
                                    void* stream);

namespace {

// first-fit allocator over [0, size), 256-byte granules, adjacent free blocks merged
class Arena {
 public:
  explicit Arena(size_t size) { if (size) free_[0] = size; }
  bool alloc(size_t n, size_t* off) {
    n = (n + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < n) continue;
      *off = it->first;
      const size_t rest = it->second - n, at = it->first + n;
      free_.erase(it);
      if (rest) free_[at] = rest;
      used_ += n;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = (n + 255) & ~(size_t)255;
    used_ -= n;
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t peak() const { return peak_; }

 private:
  std::map<size_t, size_t> free_;
  size_t used_ = 0, peak_ = 0;
};

}  // namespace

extern "C" {

// Serve n requests: request k solves field `order[k]` (shape shapes[3i..3i+2],
// field fields[i], work area work_bytes[i]) and arrives arrivals[k] seconds
// after the start (non-decreasing).  Work areas come from the HBM region
// [arena, arena + arena_bytes).  slots streams (1 and sequential != 0: one
// request at a time, synchronised after each launch).  Outputs: latency[k]
// in seconds, the makespan (last completion - first arrival) and the peak
// work bytes in use.  Returns 0, or -1 with the message in
// p3dfft_amd_last_error.
int p3dfft_serve_run(int n, const int* shapes, p3amd::real_t* const* fields, const size_t* work_bytes,
                     const int* order, const double* arrivals, int slots, void* arena, size_t arena_bytes,
                     int sequential, int lookahead, int max_bypass, double* latency, double* makespan,
                     size_t* peak) {
  if (n <= 0 || !shapes || !fields || !work_bytes || !order || !arrivals || !latency || slots <= 0 ||
      (!arena && arena_bytes)) {
    set_error("p3dfft_serve_run: invalid arguments");
    return -1;
  }
  for (int k = 0; k < n; k++) {
    const int i = order[k];
    if (i < 0 || ((work_bytes[i] + 255) & ~(size_t)255) > arena_bytes) {
      set_error("p3dfft_serve_run: a request's work area exceeds the arena (or a bad request index)");
      return -1;
    }
  }
  const int nslots = sequential ? 1 : slots;
  std::vector<hipStream_t> st(nslots, nullptr);
  std::vector<hipEvent_t> ev(nslots, nullptr);
  hipEvent_t t0 = nullptr;
  int rc = -1;
  try {
    for (int s = 0; s < nslots; s++) {
      check_hip(hipStreamCreateWithFlags(&st[s], hipStreamNonBlocking), "hipStreamCreate");
      check_hip(hipEventCreate(&ev[s]), "hipEventCreate");
    }
    check_hip(hipEventCreate(&t0), "hipEventCreate");
    Arena ar(arena_bytes);
    struct Busy {
      int k = -1;
      size_t off = 0, bytes = 0;
    };
    std::vector<Busy> busy(nslots);
    std::vector<int> free_slots;
    for (int s = nslots - 1; s >= 0; s--) free_slots.push_back(s);
    std::deque<int> pending;
    for (int k = 0; k < n; k++) pending.push_back(k);
    int done = 0, bypassed = 0;
    double last_ms = 0.0;
    check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    check_hip(hipEventRecord(t0, nullptr), "hipEventRecord");
    check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const auto start = std::chrono::steady_clock::now();
    auto launch = [&](int k, size_t off, int slot) {
      const int i = order[k];
      if (p3dfft_poisson_solve(shapes[3 * i], shapes[3 * i + 1], shapes[3 * i + 2], fields[i],
                               (char*)arena + off, work_bytes[i], (void*)st[slot]) != 0)
        throw Error{std::string(), -1};  // the solve set the message
      check_hip(hipEventRecord(ev[slot], st[slot]), "hipEventRecord");
      busy[slot].k = k;
      busy[slot].off = off;
      busy[slot].bytes = work_bytes[i];
      if (sequential) check_hip(hipStreamSynchronize(st[slot]), "hipStreamSynchronize");
    };
    while (done < n) {
      bool progress = false;
      for (int s = 0; s < nslots; s++) {
        if (busy[s].k < 0) continue;
        const hipError_t q = hipEventQuery(ev[s]);
        if (q == hipErrorNotReady) continue;
        check_hip(q, "hipEventQuery");
        float ms = 0.0f;
        check_hip(hipEventElapsedTime(&ms, t0, ev[s]), "hipEventElapsedTime");
        const int k = busy[s].k;
        latency[k] = std::max(ms * 1e-3 - arrivals[k], 0.0);
        last_ms = std::max(last_ms, (double)ms);
        ar.release(busy[s].off, busy[s].bytes);
        busy[s].k = -1;
        free_slots.push_back(s);
        done++;
        progress = true;
      }
      const double now = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
      while (!pending.empty() && !free_slots.empty() && arrivals[pending.front()] <= now) {
        size_t off = 0;
        int k = pending.front();
        if (ar.alloc(work_bytes[order[k]], &off)) {
          pending.pop_front();
          bypassed = 0;
        } else {
          // the oldest request waits for memory; admit a later arrived one that fits
          if (sequential || bypassed >= max_bypass) break;
          k = -1;
          const int lim = std::min<int>(lookahead, (int)pending.size());
          for (int j = 1; j < lim; j++) {
            const int kj = pending[j];
            if (arrivals[kj] > now) break;
            if (ar.alloc(work_bytes[order[kj]], &off)) {
              k = kj;
              pending.erase(pending.begin() + j);
              bypassed++;
              break;
            }
          }
          if (k < 0) break;
        }
        const int slot = free_slots.back();
        free_slots.pop_back();
        launch(k, off, slot);
        progress = true;
      }
      if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    if (makespan) *makespan = last_ms * 1e-3 - arrivals[0];
    if (peak) *peak = ar.peak();
    rc = 0;
  } catch (const Error& e) {
    if (!e.msg.empty()) set_error(e.msg);
    (void)hipDeviceSynchronize();
  }
  for (int s = 0; s < nslots; s++) {
    if (st[s]) (void)hipStreamSynchronize(st[s]);
    if (ev[s]) (void)hipEventDestroy(ev[s]);
    if (st[s]) (void)hipStreamDestroy(st[s]);
  }
  if (t0) (void)hipEventDestroy(t0);
  return rc;
}

}  // extern "C"
