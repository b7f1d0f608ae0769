// Serve loop: continuous batching of independent SPH simulations ("jobs")
// over one device-resident particle pool.
//
// The pool holds `capacity` particle slots.  A job occupies a contiguous
// range [off, off + n) and a simulation slot (its SimDev); unused ranges are
// holes.  Every tick runs ONE force pass over the whole pool: the radix sort
// keys a job's particles by coff + Hilbert(cell) with coff = off * 16 + 1 and a
// hole slot t by t * 16, so the sort keeps every job inside its own range and
// every hole in place, and each job's cells never collide with another's.
// Each job carries its own pass modes (first pass: initial kNN + forces
// without a kick; then regular passes = one reference Step each), so jobs at
// different points of their lives share every launch.  A job admitted with
// nsteps runs nsteps + 1 passes (the reference's Step 0 computes forces
// twice, sim/sph.go:70-76) and is read back by ID when its last pass ends;
// its range and slot are reused from the next tick on.
//
// Admission is FIFO: the head of the queue waits until a slot and a
// first-fit range of n slots are free.  One native worker thread owns the
// device; clients submit and wait from any thread.  The cell grid of a job
// is the one a standalone handle uses (about 4 cells per particle), so a job
// produces exactly the results of sph_create + sph_step with the same
// parameters, whatever else shares the pool.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "sph_sim.hpp"

using namespace sph;

namespace {

constexpr uint32_t kHoleCpp = 16;  // cell keys reserved per pool slot
constexpr int kNumRes = 7;         // fields kept for a finished job
constexpr int kResFields[kNumRes] = {SPH_X, SPH_Y, SPH_VX, SPH_VY, SPH_RHO, SPH_E, SPH_H};

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

int res_index(int field) {
  for (int i = 0; i < kNumRes; ++i)
    if (kResFields[i] == field) return i;
  return -1;
}

struct Job {
  int64_t id = 0;
  sph_params p = {};
  int64_t n = 0;
  int32_t nsteps = 0;
  std::vector<double> in;  // x, y, vx, vy, e: 5 rows of n
  int32_t state = SPH_JOB_QUEUED;
  int slot = -1;
  int64_t off = -1;
  int32_t passes = 0;  // passes completed; the job ends after nsteps + 1
  Clock::time_point t_submit, t_start, t_done;
  std::vector<double> out;  // kNumRes rows of n
  double red[7] = {};       // reductions (sum e, sum rho, E, Px, Py, dt_last), |v| last
  std::string error;
};

}  // namespace

struct sph_server {
  sph_sim core;  // pool state: np = capacity, nsims = max_jobs
  int64_t cap = 0;
  int max_jobs = 0;
  uint32_t flags = 0;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job*> queue;
  std::map<int64_t, Job*> jobs;
  int64_t next_id = 1;
  bool stop = false;
  int fatal = 0;
  std::string fatal_msg;
  std::thread worker;

  // worker state
  std::vector<Job*> slots;                          // running job per slot
  std::vector<std::pair<int64_t, int64_t>> holes;   // free ranges (off, len), sorted
  std::vector<SimDev> hsim;                         // host copy of each slot's SimDev
  std::vector<SimDev> hsim_rb;                      // readback of finishing slots
  std::vector<Item> h_items, h_items_all;
  std::vector<int> h_rf, h_rc;
  std::vector<uint32_t> iota;
  Item* d_items = nullptr;
  Item* d_items_all = nullptr;
  int* d_rf = nullptr;
  int* d_rc = nullptr;
  double* d_stage = nullptr;  // kNumRes rows of capacity
  std::vector<int> h_err;  // error flags: [0] any, [1 + slot] per job slot
  KnnStats h_stats[kStatSlots] = {};
  unsigned long long seen_failures = 0;
  // slots freed and item lists changed at the end of a tick, uploaded at the
  // start of the next one (host buffers of a pending copy are not reused)
  std::vector<int> to_clear;
  bool items_dirty = false;
  SimDev zero_sim = {};

  // statistics
  int64_t n_done = 0, n_failed = 0, particle_steps = 0, ticks = 0;
  double occupancy_sum = 0;
  double tick_ms_sum = 0;
  double gpu_ms_sum = 0;
  hipEvent_t ev_tick[2] = {nullptr, nullptr};
  bool have_t0 = false;
  Clock::time_point t0, t_last;
  std::vector<double> lat_ms;    // submit to done, per finished job
  std::vector<double> queue_ms;  // submit to admission
  std::vector<double> run_ms;    // admission to done
};

namespace {

int64_t first_fit(sph_server* s, int64_t n) {
  for (auto& h : s->holes)
    if (h.second >= n) return h.first;
  return -1;
}

void take_range(sph_server* s, int64_t off, int64_t n) {
  for (size_t i = 0; i < s->holes.size(); ++i) {
    auto& h = s->holes[i];
    if (h.first == off) {
      h.first += n;
      h.second -= n;
      if (h.second == 0) s->holes.erase(s->holes.begin() + i);
      return;
    }
  }
}

void give_range(sph_server* s, int64_t off, int64_t n) {
  auto& H = s->holes;
  auto it = std::lower_bound(H.begin(), H.end(), std::make_pair(off, (int64_t)0));
  it = H.insert(it, {off, n});
  size_t i = it - H.begin();
  if (i + 1 < H.size() && H[i].first + H[i].second == H[i + 1].first) {
    H[i].second += H[i + 1].second;
    H.erase(H.begin() + i + 1);
  }
  if (i > 0 && H[i - 1].first + H[i - 1].second == H[i].first) {
    H[i - 1].second += H[i].second;
    H.erase(H.begin() + i);
  }
}

void push_items(std::vector<Item>& v, int sim, int64_t off, int64_t n) {
  for (int64_t b = 0; b < n; b += kItem) {
    Item it;
    it.sim = sim;
    it.first = off + b;
    it.count = (int)((n - b) < kItem ? (n - b) : kItem);
    v.push_back(it);
  }
}

int64_t max_items(int64_t cap, int max_jobs) { return cap / kItem + 3 * (int64_t)max_jobs + 4; }

// Work items of the running jobs (grouped by slot), plus the holes.
int upload_items(sph_server* s) {
  s->h_items.clear();
  s->h_items_all.clear();
  std::fill(s->h_rf.begin(), s->h_rf.end(), 0);
  std::fill(s->h_rc.begin(), s->h_rc.end(), 0);
  for (int q = 0; q < s->max_jobs; ++q) {
    Job* j = s->slots[q];
    if (!j) continue;
    s->h_rf[q] = (int)s->h_items.size();
    push_items(s->h_items, q, j->off, j->n);
    s->h_rc[q] = (int)s->h_items.size() - s->h_rf[q];
  }
  // keys, sort and permute cover [0, hi), hi = end of the highest job
  // (first fit keeps jobs low in the pool)
  int64_t hi = 0;
  for (Job* j : s->slots)
    if (j && j->off + j->n > hi) hi = j->off + j->n;
  s->h_items_all = s->h_items;
  for (auto& h : s->holes)
    if (h.first < hi) push_items(s->h_items_all, -1, h.first, std::min(h.second, hi - h.first));
  PassBuffers& pb = s->core.pb;
  pb.any_open = 0;
  pb.red_parts_max = 1;
  for (int q = 0; q < s->max_jobs; ++q) {
    if (s->slots[q] && s->hsim[q].open) pb.any_open = 1;
    pb.red_parts_max = std::max(pb.red_parts_max, red_parts(s->h_rc[q]));
  }
  pb.nitems = (int)s->h_items.size();
  pb.nitems_all = (int)s->h_items_all.size();
  pb.np = hi;
  int bits = 1;
  while ((1LL << bits) < hi * (int64_t)kHoleCpp) ++bits;  // keys lie in [0, 16 hi)
  pb.key_bits = bits;
  const hipStream_t st = s->core.st;
  HIPCHK(hipMemcpyAsync(s->d_items, s->h_items.data(), s->h_items.size() * sizeof(Item),
                        hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(s->d_items_all, s->h_items_all.data(),
                        s->h_items_all.size() * sizeof(Item), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(s->d_rf, s->h_rf.data(), s->max_jobs * sizeof(int), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(s->d_rc, s->h_rc.data(), s->max_jobs * sizeof(int), hipMemcpyHostToDevice, st));
  return 0;
}

// Place an admitted job's particles in its range (IDs 0..n-1 in input
// order) and its parameters in its slot.
int upload_job(sph_server* s, Job* j) {
  sph_sim& c = s->core;
  const hipStream_t st = c.st;
  const int64_t n = j->n, o = j->off;
  SimDev& d = s->hsim[j->slot];
  fill_simdev(d, &j->p, n, grid_log2(n, 4.0));
  init_open_box(d, j->in.data(), j->in.data() + n, n);
  d.off = o;
  d.coff = o * kHoleCpp + 1;
  d.pmode = kPermInit;
  d.fmode = kForceNoKick;
  const double inf = INFINITY;
  std::memcpy(&d.cfl_min_bits, &inf, sizeof inf);
  HIPCHK(hipMemcpyAsync(c.pb.sims + j->slot, &d, sizeof d, hipMemcpyHostToDevice, st));
  const size_t b = n * sizeof(double);
  const Fields& f = c.F[c.cur];
  HIPCHK(hipMemcpyAsync(c.P[0].x + o, j->in.data(), b, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(c.P[0].y + o, j->in.data() + n, b, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(f.vx + o, j->in.data() + 2 * n, b, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(f.vy + o, j->in.data() + 3 * n, b, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(f.e + o, j->in.data() + 4 * n, b, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(f.id + o, s->iota.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  double* zero[] = {f.ax, f.ay, f.edot, f.vpx, f.vpy, f.ep, f.h, f.rho, f.c};
  for (double* a : zero) HIPCHK(hipMemsetAsync(a + o, 0, b, st));
  return 0;
}

// After its first pass a job switches to regular steps.
int upload_modes(sph_server* s, int slot) {
  SimDev& d = s->hsim[slot];
  d.pmode = kPermPredict;
  d.fmode = kForceKick;
  char* dst = (char*)(s->core.pb.sims + slot) + offsetof(SimDev, pmode);
  HIPCHK(hipMemcpyAsync(dst, &d.pmode, 2 * sizeof(int32_t), hipMemcpyHostToDevice, s->core.st));
  return 0;
}

int clear_slot(sph_server* s, int slot) {
  HIPCHK(hipMemcpyAsync(s->core.pb.sims + slot, &s->zero_sim, sizeof(SimDev),
                        hipMemcpyHostToDevice, s->core.st));
  return 0;
}

// Fields and reductions of a job whose last pass was just enqueued.
int enqueue_readback(sph_server* s, Job* j) {
  sph_sim& c = s->core;
  j->out.resize((size_t)kNumRes * j->n);
  for (int r = 0; r < kNumRes; ++r) {
    double* dst = s->d_stage + (int64_t)r * s->cap + j->off;
    int rc = enqueue_read_field(c.pb, kResFields[r], j->slot, dst, c.st);
    if (rc) return fail(SPH_EDEVICE, "readback launch: %s", hipGetErrorString((hipError_t)rc));
    HIPCHK(hipMemcpyAsync(j->out.data() + (size_t)r * j->n, dst, j->n * sizeof(double),
                          hipMemcpyDeviceToHost, c.st));
  }
  HIPCHK(hipMemcpyAsync(&s->hsim_rb[j->slot], c.pb.sims + j->slot, sizeof(SimDev),
                        hipMemcpyDeviceToHost, c.st));
  return 0;
}

void fail_job(sph_server* s, Job* j, const std::string& msg) {
  j->state = SPH_JOB_FAILED;
  j->error = msg;
  j->t_done = Clock::now();
  s->n_failed += 1;
}

void fail_all_running(sph_server* s, const std::string& msg) {
  for (int q = 0; q < s->max_jobs; ++q)
    if (s->slots[q]) fail_job(s, s->slots[q], msg);
}

// One tick: admission, one pass over the pool, completions.
// Returns false when the worker should stop.
bool tick(sph_server* s) {
  sph_sim& c = s->core;
  std::vector<Job*> admitted;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    bool any = false;
    for (Job* j : s->slots) any = any || j;
    s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty() || any; });
    if (s->stop) return false;
    while (!s->queue.empty()) {
      Job* j = s->queue.front();
      int slot = -1;
      for (int q = 0; q < s->max_jobs && slot < 0; ++q)
        if (!s->slots[q]) slot = q;
      const int64_t off = slot < 0 ? -1 : first_fit(s, j->n);
      if (off < 0) break;
      s->queue.pop_front();
      take_range(s, off, j->n);
      j->slot = slot;
      j->off = off;
      j->state = SPH_JOB_RUNNING;
      j->t_start = Clock::now();
      s->slots[slot] = j;
      admitted.push_back(j);
    }
  }
  const Clock::time_point t_tick = Clock::now();
  int rc = 0;
  if (s->ev_tick[0]) (void)hipEventRecord(s->ev_tick[0], c.st);  // timing only
  for (int q : s->to_clear)
    if (!rc && !s->slots[q]) rc = clear_slot(s, q);
  s->to_clear.clear();
  for (Job* j : admitted)
    if (!rc) rc = upload_job(s, j);
  if (!rc && (s->items_dirty || !admitted.empty())) rc = upload_items(s);
  s->items_dirty = false;
  std::vector<Job*> finishing;
  int64_t occupied = 0;
  for (Job* j : s->slots) {
    if (!j || rc) continue;
    occupied += j->n;
    if (j->passes == 1) rc = upload_modes(s, j->slot);
    if (j->passes + 1 == j->nsteps + 1) finishing.push_back(j);
  }
  if (!rc) {
    c.bind();
    rc = enqueue_pass(c.pb, kPerSim, kPerSim, c.st, nullptr);
    if (rc) rc = fail(SPH_EDEVICE, "pass launch: %s", hipGetErrorString((hipError_t)rc));
    c.cur ^= 1;
    c.bind();
  }
  for (Job* j : finishing)
    if (!rc) rc = enqueue_readback(s, j);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(s->h_err.data(), c.pb.err, s->h_err.size() * sizeof(int),
                                  hipMemcpyDeviceToHost, c.st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(s->h_stats, c.pb.stats, sizeof s->h_stats, hipMemcpyDeviceToHost, c.st);
    if (e == hipSuccess && s->ev_tick[1]) e = hipEventRecord(s->ev_tick[1], c.st);
    if (e == hipSuccess) e = hipStreamSynchronize(c.st);
    if (e != hipSuccess) rc = fail(SPH_EDEVICE, "tick: %s", hipGetErrorString(e));
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (rc) {
    s->fatal = rc;
    s->fatal_msg = sph_last_error();
    fail_all_running(s, s->fatal_msg);
    for (Job* j : s->queue) {
      j->state = SPH_JOB_FAILED;
      j->error = s->fatal_msg;
      s->n_failed += 1;
    }
    s->queue.clear();
    s->cv_done.notify_all();
    return false;
  }
  s->ticks += 1;
  s->occupancy_sum += (double)occupied / (double)s->cap;
  s->tick_ms_sum += ms_between(t_tick, Clock::now());
  float gms = 0.f;
  if (s->ev_tick[0] && hipEventElapsedTime(&gms, s->ev_tick[0], s->ev_tick[1]) == hipSuccess)
    s->gpu_ms_sum += gms;
  const unsigned long long failures = sum_stats(s->h_stats).failures;
  bool flagged = s->h_err[0] != 0;
  for (int q = 0; q < s->max_jobs && !flagged; ++q) flagged = s->h_err[1 + q] != 0;
  if (flagged || failures != s->seen_failures) {
    // each flag names the job slot whose particle or search failed: only
    // those jobs fail, the others of the pass go on
    bool owned = false;
    for (int q = 0; q < s->max_jobs; ++q) {
      const int f = s->h_err[1 + q];
      if (!f || !s->slots[q]) continue;
      fail_job(s, s->slots[q], (f & 1) ? "a particle left the periodic images [-L, 2L) of the box"
                                       : "kNN found fewer than k neighbours");
      owned = true;
    }
    if (!owned)
      fail_all_running(s, s->h_err[0] ? "a particle left the periodic images [-L, 2L) of the box"
                                      : "kNN found fewer than k neighbours");
    s->seen_failures = failures;
    if (flagged && hipMemsetAsync(c.pb.err, 0, s->h_err.size() * sizeof(int), c.st) != hipSuccess) {
      s->fatal = SPH_EDEVICE;
      s->fatal_msg = "could not clear the error flags";
    }
  }
  const Clock::time_point now = Clock::now();
  for (int q = 0; q < s->max_jobs; ++q) {
    Job* j = s->slots[q];
    if (!j) continue;
    if (j->state == SPH_JOB_RUNNING) {
      j->passes += 1;
      if (j->passes < j->nsteps + 1) continue;
      const SimDev& d = s->hsim_rb[q];
      for (int r = 0; r < kNumRed; ++r) j->red[r] = d.red[r];
      j->red[5] = d.dt_last;
      j->red[6] = d.last_speed;
      j->state = SPH_JOB_DONE;
      j->t_done = now;
      s->n_done += 1;
      s->particle_steps += j->n * (int64_t)j->nsteps;
      s->lat_ms.push_back(ms_between(j->t_submit, now));
      s->queue_ms.push_back(ms_between(j->t_submit, j->t_start));
      s->run_ms.push_back(ms_between(j->t_start, now));
    }
    s->t_last = now;
    give_range(s, j->off, j->n);
    s->slots[q] = nullptr;
    s->to_clear.push_back(q);
    s->items_dirty = true;
  }
  s->cv_done.notify_all();
  return s->fatal == 0;
}

void worker_main(sph_server* s) {
  // the device was selected and checked when the server was created
  (void)hipSetDevice(s->core.device);
  while (tick(s)) {
  }
  std::lock_guard<std::mutex> lk(s->mu);
  s->stop = true;
  s->cv_done.notify_all();
}

Job* find_job(sph_server* s, int64_t id) {
  auto it = s->jobs.find(id);
  return it == s->jobs.end() ? nullptr : it->second;
}

}  // namespace

extern "C" int sph_server_create(int64_t capacity, int32_t max_jobs, int32_t kmax,
                                 int32_t precision, uint32_t flags, sph_server** out) {
  if (!out) return fail(SPH_EINVAL, "null output handle");
  *out = nullptr;
  if (capacity < 2 || capacity >= (1LL << 27))
    return fail(SPH_EINVAL, "capacity must be in [2, 2^27), got %lld", (long long)capacity);
  if (max_jobs < 1 || max_jobs > 65536) return fail(SPH_EINVAL, "max_jobs must be in [1, 65536]");
  if (kmax < 2 || kmax > kMaxK) return fail(SPH_EINVAL, "kmax must be in [2, %d]", kMaxK);
  if (precision != SPH_PREC_FP64 && precision != SPH_PREC_FP32C)
    return fail(SPH_EINVAL, "unknown precision %d", precision);
  if ((flags & ~SPH_FLAG_BITWISE) != 0) return fail(SPH_EINVAL, "unsupported server flags");
  if ((flags & SPH_FLAG_BITWISE) && precision != SPH_PREC_FP64)
    return fail(SPH_EINVAL, "bitwise reference order needs FP64 precision");
  auto* s = new sph_server();
  auto bail = [&](int code) {
    delete s;
    return code;
  };
  sph_sim& c = s->core;
  s->cap = capacity;
  s->max_jobs = max_jobs;
  s->flags = flags;
  HIPCHK(hipGetDevice(&c.device));
  c.np = capacity;
  c.kmax = kmax;
  c.prec = precision;
  c.bitwise = (flags & SPH_FLAG_BITWISE) ? 1 : 0;
  if (hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(SPH_EDEVICE, "hipStreamCreate failed"));
  PassBuffers& pb = c.pb;
  const int64_t nmax = max_items(capacity, max_jobs);
  pb.nsims = max_jobs;
  pb.kmax = (kmax + 3) & ~3;
  pb.kout = kmax;
  pb.prec = precision;
  pb.bitwise = c.bitwise;
  pb.np = capacity;
  pb.ncells = capacity * (int64_t)kHoleCpp;
  int bits = 1;
  while ((1LL << bits) < pb.ncells) ++bits;
  pb.key_bits = bits;
  pb.hole_cpp = kHoleCpp;
  pb.walk = kWalkDefault;
  pb.have_h = c.bitwise ? 0 : 1;  // jobs without h take the exact search
  c.cell_capacity = pb.ncells + 1;
  SimDev* d_sims = nullptr;
  int rc = 0;
  if ((rc = c.alloc(&s->d_items, nmax)) || (rc = c.alloc(&s->d_items_all, nmax)) ||
      (rc = c.alloc(&d_sims, max_jobs)) || (rc = c.alloc(&s->d_rf, max_jobs)) ||
      (rc = c.alloc(&s->d_rc, max_jobs)) || (rc = alloc_state(&c, capacity, (int)nmax)) ||
      (rc = c.alloc(&s->d_stage, (int64_t)kNumRes * capacity)))
    return bail(rc);
  pb.items = s->d_items;
  pb.items_all = s->d_items_all;
  pb.sims = d_sims;
  pb.red_first = s->d_rf;
  pb.red_count = s->d_rc;
  pb.nitems = 0;
  pb.nitems_all = 0;
  if (hipMemsetAsync(d_sims, 0, max_jobs * sizeof(SimDev), c.st) != hipSuccess ||
      hipMemsetAsync(pb.stats, 0, kStatSlots * sizeof(KnnStats), c.st) != hipSuccess ||
      hipMemsetAsync(pb.walk_miss, 0, (int64_t)max_jobs * kMissSlots * kMissStride * 4, c.st) !=
          hipSuccess ||
      hipMemsetAsync(pb.run_count, 0, sizeof(uint32_t), c.st) != hipSuccess ||
      hipMemsetAsync(pb.err, 0, (1 + max_jobs) * sizeof(int), c.st) != hipSuccess ||
      hipMemsetAsync(pb.nbr, 0, nbr_entries(capacity, pb.kmax) * sizeof(uint32_t), c.st) != hipSuccess ||
      hipStreamSynchronize(c.st) != hipSuccess)
    return bail(fail(SPH_EDEVICE, "server init failed"));
  s->slots.assign(max_jobs, nullptr);
  s->h_err.assign(1 + max_jobs, 0);
  s->hsim.assign(max_jobs, SimDev{});
  s->hsim_rb.assign(max_jobs, SimDev{});
  s->h_rf.assign(max_jobs, 0);
  s->h_rc.assign(max_jobs, 0);
  s->iota.resize(capacity);
  for (int64_t i = 0; i < capacity; ++i) s->iota[i] = (uint32_t)i;
  s->holes.push_back({0, capacity});
  for (auto& e : s->ev_tick)
    if (hipEventCreate(&e) != hipSuccess) return bail(fail(SPH_EDEVICE, "hipEventCreate failed"));
  c.bind();
  s->worker = std::thread(worker_main, s);
  *out = s;
  return 0;
}

extern "C" void sph_server_destroy(sph_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)hipStreamSynchronize(s->core.st);  // teardown
  for (auto& e : s->ev_tick)
    if (e) (void)hipEventDestroy(e);
  for (auto& kv : s->jobs) delete kv.second;
  delete s;
}

extern "C" int sph_server_submit(sph_server* s, const sph_params* p, int64_t n, const double* x,
                                 const double* y, const double* vx, const double* vy,
                                 const double* e, int32_t nsteps, int64_t* job_id) {
  if (!s || !job_id) return fail(SPH_EINVAL, "null argument");
  int rc = check_params(p, n);
  if (rc) return rc;
  if (n > s->cap) return fail(SPH_ENOMEM, "job of %lld particles exceeds the pool capacity %lld",
                              (long long)n, (long long)s->cap);
  if (p->k > s->core.kmax) return fail(SPH_EINVAL, "k = %d exceeds the server's kmax", p->k);
  if (p->precision != s->core.prec || (p->flags & SPH_FLAG_BITWISE) != (s->flags & SPH_FLAG_BITWISE))
    return fail(SPH_EINVAL, "precision and order flags must match the server's");
  if (nsteps < 1) return fail(SPH_EINVAL, "a job runs at least one step");
  if (!x || !y) return fail(SPH_EINVAL, "positions must not be null");
  auto* j = new Job();
  j->p = *p;
  j->n = n;
  j->nsteps = nsteps;
  j->in.assign((size_t)5 * n, 0.0);
  const double* src[5] = {x, y, vx, vy, e};
  for (int r = 0; r < 5; ++r)
    if (src[r]) std::memcpy(j->in.data() + (size_t)r * n, src[r], n * sizeof(double));
  for (int64_t i = 0; i < 2 * n; ++i)
    if (!std::isfinite(j->in[i])) {
      delete j;
      return fail(SPH_ERANGE, "non-finite position of particle %lld", (long long)(i % n));
    }
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop || s->fatal) {
    delete j;
    return fail(SPH_ESTATE, "server stopped: %s", s->fatal_msg.c_str());
  }
  j->id = s->next_id++;
  j->t_submit = Clock::now();
  if (!s->have_t0) {
    s->have_t0 = true;
    s->t0 = j->t_submit;
  }
  s->jobs[j->id] = j;
  s->queue.push_back(j);
  *job_id = j->id;
  s->cv_work.notify_one();
  return 0;
}

extern "C" int sph_server_wait(sph_server* s, int64_t job_id, double timeout_s) {
  if (!s) return fail(SPH_EINVAL, "null handle");
  std::unique_lock<std::mutex> lk(s->mu);
  Job* j = find_job(s, job_id);
  if (!j) return fail(SPH_EINVAL, "unknown job %lld", (long long)job_id);
  auto ready = [&] { return j->state == SPH_JOB_DONE || j->state == SPH_JOB_FAILED || s->stop; };
  if (timeout_s < 0) s->cv_done.wait(lk, ready);
  else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
    return fail(SPH_ESTATE, "job %lld not finished within %.3f s", (long long)job_id, timeout_s);
  if (j->state == SPH_JOB_FAILED) return fail(SPH_ERANGE, "job %lld failed: %s", (long long)job_id, j->error.c_str());
  if (j->state != SPH_JOB_DONE) return fail(SPH_ESTATE, "server stopped before job %lld ended", (long long)job_id);
  return 0;
}

extern "C" int sph_server_drain(sph_server* s) {
  if (!s) return fail(SPH_EINVAL, "null handle");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] {
    if (s->stop) return true;
    if (!s->queue.empty()) return false;
    for (Job* j : s->slots)
      if (j) return false;
    return true;
  });
  if (s->fatal) return fail(s->fatal, "server stopped: %s", s->fatal_msg.c_str());
  return 0;
}

extern "C" int sph_server_job_info(sph_server* s, int64_t job_id, int32_t* state, double info[4]) {
  if (!s) return fail(SPH_EINVAL, "null handle");
  std::lock_guard<std::mutex> lk(s->mu);
  Job* j = find_job(s, job_id);
  if (!j) return fail(SPH_EINVAL, "unknown job %lld", (long long)job_id);
  if (state) *state = j->state;
  if (info) {
    const bool started = j->state != SPH_JOB_QUEUED;
    const bool ended = j->state == SPH_JOB_DONE || j->state == SPH_JOB_FAILED;
    info[0] = started ? ms_between(j->t_submit, j->t_start) : -1.0;
    info[1] = ended && started ? ms_between(j->t_start, j->t_done) : -1.0;
    info[2] = ended ? ms_between(j->t_submit, j->t_done) : -1.0;
    info[3] = (double)(j->passes > 0 ? j->passes - 1 : 0);
  }
  return 0;
}

extern "C" int sph_server_read_result(sph_server* s, int64_t job_id, int32_t field, double* out) {
  if (!s || !out) return fail(SPH_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  Job* j = find_job(s, job_id);
  if (!j) return fail(SPH_EINVAL, "unknown job %lld", (long long)job_id);
  if (j->state != SPH_JOB_DONE) return fail(SPH_ESTATE, "job %lld has not finished", (long long)job_id);
  const int r = res_index(field);
  if (r < 0) return fail(SPH_EINVAL, "field %d is not kept for finished jobs", field);
  std::memcpy(out, j->out.data() + (size_t)r * j->n, j->n * sizeof(double));
  return 0;
}

extern "C" int sph_server_job_reductions(sph_server* s, int64_t job_id, double out[7]) {
  if (!s || !out) return fail(SPH_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  Job* j = find_job(s, job_id);
  if (!j) return fail(SPH_EINVAL, "unknown job %lld", (long long)job_id);
  if (j->state != SPH_JOB_DONE) return fail(SPH_ESTATE, "job %lld has not finished", (long long)job_id);
  for (int r = 0; r < 7; ++r) out[r] = j->red[r];
  return 0;
}

extern "C" int sph_server_release(sph_server* s, int64_t job_id) {
  if (!s) return fail(SPH_EINVAL, "null handle");
  std::lock_guard<std::mutex> lk(s->mu);
  Job* j = find_job(s, job_id);
  if (!j) return fail(SPH_EINVAL, "unknown job %lld", (long long)job_id);
  if (j->state != SPH_JOB_DONE && j->state != SPH_JOB_FAILED)
    return fail(SPH_ESTATE, "job %lld is still queued or running", (long long)job_id);
  s->jobs.erase(job_id);
  delete j;
  return 0;
}

extern "C" int sph_server_stats(sph_server* s, double out[SPH_SERVER_NSTATS]) {
  if (!s || !out) return fail(SPH_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  // nearest-rank percentile of a copy of v
  auto pct = [](std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    size_t i = (size_t)std::ceil(q * v.size());
    i = i > 0 ? i - 1 : 0;
    return v[i < v.size() ? i : v.size() - 1];
  };
  const double wall = s->have_t0 && s->n_done ? ms_between(s->t0, s->t_last) / 1e3 : 0.0;
  int64_t queued = (int64_t)s->queue.size(), running = 0;
  for (Job* j : s->slots) running += j ? 1 : 0;
  out[0] = (double)s->n_done;
  out[1] = (double)s->n_failed;
  out[2] = (double)s->particle_steps;
  out[3] = wall;
  out[4] = wall > 0 ? s->particle_steps / wall : 0.0;
  out[5] = pct(s->lat_ms, 0.50);
  out[6] = pct(s->lat_ms, 0.99);
  out[7] = s->ticks ? s->occupancy_sum / s->ticks : 0.0;
  out[8] = (double)s->ticks;
  out[9] = (double)(queued + running);
  out[10] = s->ticks ? s->tick_ms_sum / s->ticks : 0.0;
  out[11] = s->ticks ? s->gpu_ms_sum / s->ticks : 0.0;
  out[12] = pct(s->queue_ms, 0.50);
  out[13] = pct(s->queue_ms, 0.99);
  out[14] = pct(s->run_ms, 0.50);
  out[15] = pct(s->run_ms, 0.99);
  return 0;
}
