// serve.hip -- continuous-batching request loop.
//
// One host thread owns the device work.  Requests wait in a FIFO queue and
// are admitted when a resident slot is free and their device footprint fits
// the HBM budget (an arena carved out of one allocation, best-fit with
// coalescing, so admission never calls hipMalloc).  A request that does not
// fit lets smaller ones behind it be admitted (backfill) until it has waited
// `max_head_wait_s`; after that nothing overtakes it.  Every tick advances
// all resident lattices together by k steps, k being the smallest remaining
// step count (capped by steps_per_tick), so lattices leave exactly at their
// step count; finished lattices are exported (FP64 diagnostics, optionally
// rho/u) and their memory returns to the arena before the next admission.
// A lattice that goes unstable (a density that is not a positive finite
// number, flagged by the step kernel) leaves at the end of the tick that saw
// it, with status NFLOW_EDIVERGED.
// A lattice's arithmetic never depends on what else is resident, so a
// request gives the same result alone and in any mix.  With a tau floor set,
// requests whose relaxation time lies below it are kept in FP32 storage: they
// form a second batch, stepped in its own launch on its own stream each tick.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "errors.hpp"
#include "nflow.h"

namespace nflow {

class Arena {
 public:
  int open(size_t bytes) {
    size_ = bytes / 256 * 256;
    if (size_ == 0) return fail(NFLOW_EINVAL, "server budget too small");
    if (hipMalloc(&base_, size_) != hipSuccess) {
      (void)hipGetLastError();
      base_ = nullptr;
      return fail(NFLOW_ENOMEM, "hipMalloc of the %zu-byte serving arena failed", size_);
    }
    free_[0] = size_;
    return 0;
  }
  ~Arena() {
    if (base_) (void)hipFree(base_);
  }
  size_t size() const { return size_; }
  // best fit; returns nullptr when no free range is large enough
  void* alloc(size_t bytes) {
    bytes = (bytes + 255) / 256 * 256;
    auto best = free_.end();
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= bytes && (best == free_.end() || it->second < best->second)) best = it;
    if (best == free_.end()) return nullptr;
    const size_t off = best->first, len = best->second;
    free_.erase(best);
    if (len > bytes) free_[off + bytes] = len - bytes;
    used_ += bytes;
    return static_cast<char*>(base_) + off;
  }
  void release(void* p, size_t bytes) {
    bytes = (bytes + 255) / 256 * 256;
    size_t off = static_cast<char*>(p) - static_cast<char*>(base_);
    used_ -= bytes;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + bytes == next->first) {
      bytes += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += bytes;
        return;
      }
    }
    free_[off] = bytes;
  }
  size_t used() const { return used_; }

 private:
  void* base_ = nullptr;
  size_t size_ = 0, used_ = 0;
  std::map<size_t, size_t> free_;  // offset -> length
};

struct Pending {
  nflow_request req;
  double submit_s;
  int which;  // batch: 0 the server's storage, 1 FP32
};

struct Resident {
  nflow_request req;
  double submit_s, admit_s;
  long long target;
};

class Server {
 public:
  Batch batches[2];  // [0] the server's storage, [1] FP32 (below the tau floor)
  Arena arena;
  double tau_floor = 0.0;
  int device = 0;
  int max_resident = 64;
  int steps_per_tick = 100;
  double max_head_wait_s = 2.0;
  std::chrono::steady_clock::time_point t0;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Pending> queue;    // submitted, not yet seen by the worker (under mu)
  std::deque<Pending> waiting;  // seen by the worker, not admitted (worker only)
  std::deque<nflow_result> results;
  std::map<unsigned long long, std::vector<float>> fields;
  std::vector<Resident> resident[2];  // parallel to batches[w].lats
  nflow_server_stats st{};
  bool stop = false;
  int fatal = 0;
  std::string fatal_msg;
  std::thread worker;

  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

  void loop();
  void admit(double t);
  void reject(const Pending& p, int status, double t);
  int tick();
  void finish(int w, size_t i, bool diverged);
  size_t resident_count() const { return resident[0].size() + resident[1].size(); }
};

void Server::reject(const Pending& p, int status, double t) {
  nflow_result r{};
  r.tag = p.req.tag;
  r.status = status;
  r.nx = p.req.input.nx;
  r.ny = p.req.input.ny;
  r.nz = p.req.input.nz;
  r.submit_s = p.submit_s;
  r.admit_s = r.done_s = t;
  std::lock_guard<std::mutex> lk(mu);
  results.push_back(r);
  st.rejected++;
  cv_done.notify_all();
}

// Moves new submissions behind the ones already waiting, then admits in
// FIFO order with backfill.  Only the lock-protected hand-over touches the
// submitters' queue; lattice set-up (device work) runs without the lock, so
// submit and poll never wait for the device.
void Server::admit(double t) {
  {
    std::lock_guard<std::mutex> lk(mu);
    while (!queue.empty()) {
      waiting.push_back(queue.front());
      queue.pop_front();
    }
  }
  bool head_blocked = false;
  for (auto it = waiting.begin(); it != waiting.end();) {
    if ((int)resident_count() >= max_resident) break;
    const int w = it->which;
    Batch& batch = batches[w];
    const size_t bytes = lattice_bytes(it->req.input, batch.storage);
    if (bytes > arena.size()) {  // can never fit
      reject(*it, NFLOW_EBUDGET, t);
      it = waiting.erase(it);
      continue;
    }
    const bool is_head = it == waiting.begin();
    if (head_blocked && t - waiting.front().submit_s > max_head_wait_s) break;  // no overtaking an old head
    void* mem = arena.alloc(bytes);
    if (!mem) {
      if (is_head) head_blocked = true;
      ++it;
      continue;
    }
    const int idx = (int)batch.lats.size();
    int rc = batch.add_lattice(it->req.input, {}, mem);
    if (!rc) rc = batch.init_lattice(idx, nullptr);
    if (rc) {
      if ((int)batch.lats.size() > idx) batch.remove_lattice(idx);
      arena.release(mem, bytes);
      reject(*it, rc, t);
      it = waiting.erase(it);
      continue;
    }
    resident[w].push_back(Resident{it->req, it->submit_s, t, (long long)it->req.nsteps});
    it = waiting.erase(it);
    std::lock_guard<std::mutex> lk(mu);
    st.admitted++;
    st.peak_resident = std::max<int>(st.peak_resident, (int)resident_count());
    st.peak_bytes = std::max(st.peak_bytes, arena.used());
  }
}

void Server::finish(int w, size_t i, bool diverged) {
  Batch& batch = batches[w];
  HostLattice& h = batch.lats[i];
  Resident& r = resident[w][i];
  nflow_result res{};
  res.tag = r.req.tag;
  res.nx = h.in.nx;
  res.ny = h.in.ny;
  res.nz = h.in.nz;
  res.steps = h.steps_done;
  res.submit_s = r.submit_s;
  res.admit_s = r.admit_s;
  double d[5] = {0, 0, 0, 0, 0};
  std::vector<float> f;
  int rc;
  // the lattice leaves: the export works inside its own (idle) population
  // buffer, so finishing allocates no device memory outside the arena
  if (r.req.want_fields && !diverged) {
    const size_t n = h.dev.n;
    f.resize(4 * n);
    rc = batch.export_one((int)i, nullptr, f.data(), f.data() + n, f.data() + 2 * n, f.data() + 3 * n, d, nullptr,
                          true);
  } else {
    rc = batch.export_one((int)i, nullptr, nullptr, nullptr, nullptr, nullptr, d, nullptr, true);
  }
  bool finite = true;
  for (double v : d) finite = finite && std::isfinite(v);
  res.status = rc ? rc : (diverged || !finite) ? NFLOW_EDIVERGED : 0;
  res.mass = d[0];
  res.mom[0] = d[1];
  res.mom[1] = d[2];
  res.mom[2] = d[3];
  res.drho_l1 = d[4];
  res.done_s = now();
  const long long upd = (long long)h.dev.n * h.steps_done;
  const bool keep_fields = r.req.want_fields && res.status == 0;
  void* mem = h.mem;
  const size_t bytes = h.bytes;
  batch.remove_lattice((int)i);
  arena.release(mem, bytes);
  resident[w].erase(resident[w].begin() + i);  // r and h are invalid from here on
  std::lock_guard<std::mutex> lk(mu);
  if (keep_fields) fields[res.tag] = std::move(f);
  results.push_back(res);
  st.completed++;
  st.node_updates += upd;
  cv_done.notify_all();
}

// Advances every resident lattice by k steps: one launch per step for each
// non-empty batch, the two batches on their own streams; step_s accumulates
// the host time from the first launch to the end of the last.
int Server::tick() {
  long long k = steps_per_tick;
  for (int w = 0; w < 2; ++w)
    for (size_t i = 0; i < resident[w].size(); ++i)
      k = std::min(k, resident[w][i].target - batches[w].lats[i].steps_done);
  const double t0 = now();
  if (k > 0)
    for (int w = 0; w < 2; ++w)
      if (!resident[w].empty()) {
        int rc = batches[w].run((int)k);
        if (rc) return rc;
      }
  for (int w = 0; w < 2; ++w)
    if (batches[w].stream) NFLOW_HIP(hipStreamSynchronize(batches[w].stream));
  if (k > 0) {
    const double dt = now() - t0;
    std::lock_guard<std::mutex> lk(mu);  // stats are read by nflow_server_stats_get
    st.launches += k;
    st.step_s += dt;
    st.ticks++;
  }
  return 0;
}

void Server::loop() {
  (void)hipSetDevice(device);
  double busy_from = -1.0;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu);
      if (stop) break;  // destroy abandons queued and resident work
      if (resident_count() == 0 && queue.empty() && waiting.empty()) {
        if (busy_from >= 0.0) {
          st.busy_s += now() - busy_from;
          busy_from = -1.0;
        }
        cv_work.wait_for(lk, std::chrono::milliseconds(50));
        continue;
      }
    }
    const double t = now();
    admit(t);
    if (resident_count() == 0) {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait_for(lk, std::chrono::milliseconds(5));
      continue;
    }
    if (busy_from < 0.0) busy_from = t;
    int rc = tick();
    if (rc) {
      std::lock_guard<std::mutex> lk(mu);
      fatal = rc;
      fatal_msg = last_error_slot();
      // fail everything still held
      for (int w = 0; w < 2; ++w)
        for (auto& r : resident[w]) {
          nflow_result res{};
          res.tag = r.req.tag;
          res.status = rc;
          res.submit_s = r.submit_s;
          res.admit_s = r.admit_s;
          res.done_s = now();
          results.push_back(res);
        }
      for (const auto* q : {&waiting, &queue})
        for (const auto& p : *q) {
          nflow_result res{};
          res.tag = p.req.tag;
          res.status = rc;
          res.submit_s = p.submit_s;
          results.push_back(res);
        }
      waiting.clear();
      queue.clear();
      resident[0].clear();
      resident[1].clear();
      cv_done.notify_all();
      break;
    }
    for (int w = 0; w < 2; ++w) {
      if (resident[w].empty()) continue;
      std::vector<int> diverged;
      rc = batches[w].diverged_flags(diverged);
      if (rc) diverged.assign(resident[w].size(), 0);  // the export that follows reports the device error
      for (size_t i = resident[w].size(); i-- > 0;)
        if (diverged[i] || batches[w].lats[i].steps_done >= resident[w][i].target) finish(w, i, diverged[i] != 0);
    }
  }
  if (busy_from >= 0.0) {
    std::lock_guard<std::mutex> lk(mu);
    st.busy_s += now() - busy_from;
  }
}

}  // namespace nflow

using namespace nflow;

struct nflow_server {
  Server s;
};

extern "C" int nflow_server_create(size_t hbm_budget_bytes, int storage, int max_resident, int steps_per_tick,
                                   nflow_server** out) {
  if (!out || max_resident < 1 || steps_per_tick < 1) return fail(NFLOW_EINVAL, "nflow_server_create: bad argument");
  if (storage < NFLOW_STORAGE_F32 || storage > NFLOW_STORAGE_BF16) return fail(NFLOW_EINVAL, "bad storage");
  *out = nullptr;
  nflow_server* sv = new nflow_server();
  Server& s = sv->s;
  s.batches[0].storage = storage;
  s.batches[1].storage = NFLOW_STORAGE_F32;
  s.max_resident = max_resident;
  s.steps_per_tick = steps_per_tick;
  nflow_tuning t = s.batches[0].tuning;
  t.graph_steps = 0;  // the resident set changes between ticks
  // measured on a 48-request FP16 stream (cavities 48-256, channels
  // 256-1024 x 128 x 128): 512-node items with XCD-ordered dealing 67.8-68.4
  // GLUPS, 512 alone 67.0, 256 with XCD order 66.6, 256 alone 63.8-64.3
  t.block = 512;
  t.xcd_order = 1;
  s.batches[0].tuning = t;
  s.batches[1].tuning = t;
  int rc = hipGetDevice(&s.device) == hipSuccess ? 0 : fail(NFLOW_EDEVICE, "hipGetDevice");
  if (!rc) rc = s.batches[0].open();
  if (!rc && storage != NFLOW_STORAGE_F32) rc = s.batches[1].open();
  if (!rc) rc = s.arena.open(hbm_budget_bytes);
  if (rc) {
    delete sv;
    return rc;
  }
  s.t0 = std::chrono::steady_clock::now();
  s.worker = std::thread([&s] { s.loop(); });
  *out = sv;
  return 0;
}

extern "C" int nflow_server_set_tau_floor(nflow_server* sv, double tau_min) {
  if (!sv || !(tau_min >= 0.0)) return fail(NFLOW_EINVAL, "nflow_server_set_tau_floor: bad argument");
  std::lock_guard<std::mutex> lk(sv->s.mu);
  sv->s.tau_floor = tau_min;
  return 0;
}

extern "C" int nflow_server_submit(nflow_server* sv, const nflow_request* r) {
  if (!sv || !r) return fail(NFLOW_EINVAL, "nflow_server_submit: null argument");
  if (r->nsteps < 0) return fail(NFLOW_EINVAL, "nsteps must be >= 0");
  Server& s = sv->s;
  int rc = validate_input(r->input, s.batches[0].storage);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(s.mu);
  if (s.fatal) return fail(s.fatal, "server stopped: %s", s.fatal_msg.c_str());
  if (s.stop) return fail(NFLOW_ESTATE, "server is shutting down");
  const int w = (s.batches[0].storage != NFLOW_STORAGE_F32 && r->input.tau < s.tau_floor) ? 1 : 0;
  Pending p{*r, s.now(), w};
  p.req.input.storage = s.batches[w].storage;
  s.queue.push_back(p);
  s.st.submitted++;
  s.cv_work.notify_all();
  return 0;
}

extern "C" int nflow_server_poll(nflow_server* sv, nflow_result* out, int max_results, int wait_ms) {
  if (!sv || (!out && max_results > 0)) return fail(NFLOW_EINVAL, "nflow_server_poll: null argument");
  Server& s = sv->s;
  std::unique_lock<std::mutex> lk(s.mu);
  if (s.results.empty() && wait_ms > 0)
    s.cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s.results.empty(); });
  int n = 0;
  while (n < max_results && !s.results.empty()) {
    out[n++] = s.results.front();
    s.results.pop_front();
  }
  return n;
}

extern "C" int nflow_server_fetch_macro(nflow_server* sv, unsigned long long tag, float* rho, float* ux, float* uy,
                                        float* uz) {
  if (!sv || !rho || !ux || !uy || !uz) return fail(NFLOW_EINVAL, "nflow_server_fetch_macro: null argument");
  Server& s = sv->s;
  std::lock_guard<std::mutex> lk(s.mu);
  auto it = s.fields.find(tag);
  if (it == s.fields.end()) return fail(NFLOW_ENOTFOUND, "no fields kept for tag %llu", tag);
  const size_t n = it->second.size() / 4;
  std::memcpy(rho, it->second.data(), n * sizeof(float));
  std::memcpy(ux, it->second.data() + n, n * sizeof(float));
  std::memcpy(uy, it->second.data() + 2 * n, n * sizeof(float));
  std::memcpy(uz, it->second.data() + 3 * n, n * sizeof(float));
  s.fields.erase(it);
  return 0;
}

extern "C" int nflow_server_drain(nflow_server* sv) {
  if (!sv) return fail(NFLOW_EINVAL, "null server");
  Server& s = sv->s;
  std::unique_lock<std::mutex> lk(s.mu);
  for (;;) {
    if (s.fatal) return fail(s.fatal, "server stopped: %s", s.fatal_msg.c_str());
    if (s.queue.empty() && s.st.completed + s.st.rejected >= s.st.submitted) return 0;
    s.cv_done.wait_for(lk, std::chrono::milliseconds(20));
  }
}

extern "C" int nflow_server_stats_get(nflow_server* sv, nflow_server_stats* out) {
  if (!sv || !out) return fail(NFLOW_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(sv->s.mu);
  *out = sv->s.st;
  return 0;
}

extern "C" void nflow_server_destroy(nflow_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lk(sv->s.mu);
    sv->s.stop = true;
    sv->s.cv_work.notify_all();
  }
  if (sv->s.worker.joinable()) sv->s.worker.join();
  delete sv;
}
