// Internal handle of the C ABI (include/sphugo.h), shared by the batch
// path and the serve loop.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "sph_host.hpp"
#include "sphugo.h"

namespace sph {

int fail(int code, const char* fmt, ...);

#define HIPCHK(call)                                                                \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      return sph::fail(SPH_EDEVICE, "%s: %s", #call, hipGetErrorString(e_));        \
  } while (0)

}  // namespace sph

using sph::Fields;
using sph::Item;
using sph::KnnStats;
using sph::PassBuffers;
using sph::Positions;
using sph::SimDev;

struct sph_sim {
  int device = 0;
  std::vector<SimDev> hsims;
  std::vector<sph_params> hparams;  // creation parameters, per simulation
  std::vector<int64_t> sim_n;
  std::vector<Item> hitems;
  int64_t np = 0;
  int kmax = 0;
  int prec = 0;
  int bitwise = 0;
  bool timing = false;
  hipStream_t st = nullptr;
  hipEvent_t ev[2 * SPH_NTIMERS] = {};
  double t_ms[SPH_NTIMERS] = {};
  int64_t t_cnt[SPH_NTIMERS] = {};
  int cur = 0;
  Fields F[2] = {};
  Positions P[2] = {};
  PassBuffers pb = {};
  int64_t step = 0;
  std::vector<void*> allocs;
  double* d_scratch = nullptr;     // np doubles for readback
  uint32_t* d_nids = nullptr;      // neighbour readback (lazy)
  double* d_nd2 = nullptr;
  double* d_npx = nullptr;         // neighbour image positions (lazy)
  double* d_npy = nullptr;
  KnnStats last_stats = {};
  // graph replay of a regular step, one per field-buffer parity
  bool graph_on = false;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  double grid_cpp = 4.0;        // cells per particle of the cell grid
  int64_t cell_capacity = 0;    // cstart entries allocated

  template <class T>
  int alloc(T** p, int64_t count) {
    void* q = nullptr;
    size_t bytes = (size_t)(count > 0 ? count : 1) * sizeof(T);
    if (hipMalloc(&q, bytes) != hipSuccess)
      return sph::fail(SPH_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
    allocs.push_back(q);
    *p = (T*)q;
    return 0;
  }
  void drop_graphs() {
    for (auto& g : gexec)
      if (g) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
  }
  ~sph_sim() {
    drop_graphs();
    // teardown: nothing is left to report a failure to
    for (void* q : allocs) (void)hipFree(q);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (st) (void)hipStreamDestroy(st);
  }
  void bind() {
    pb.fcur = F[cur];
    pb.fnew = F[cur ^ 1];
    pb.pos_cur = P[0];
  }
  // Exchange everything with another handle (a handle rebuilt in place:
  // the caller's pointer stays valid).  Lists every member above.
  void swap_state(sph_sim& o) {
    std::swap(device, o.device);
    hsims.swap(o.hsims);
    hparams.swap(o.hparams);
    sim_n.swap(o.sim_n);
    hitems.swap(o.hitems);
    std::swap(np, o.np);
    std::swap(kmax, o.kmax);
    std::swap(prec, o.prec);
    std::swap(bitwise, o.bitwise);
    std::swap(timing, o.timing);
    std::swap(st, o.st);
    std::swap(ev, o.ev);
    std::swap(t_ms, o.t_ms);
    std::swap(t_cnt, o.t_cnt);
    std::swap(cur, o.cur);
    std::swap(F, o.F);
    std::swap(P, o.P);
    std::swap(pb, o.pb);
    std::swap(step, o.step);
    allocs.swap(o.allocs);
    std::swap(d_scratch, o.d_scratch);
    std::swap(d_nids, o.d_nids);
    std::swap(d_nd2, o.d_nd2);
    std::swap(d_npx, o.d_npx);
    std::swap(d_npy, o.d_npy);
    std::swap(last_stats, o.last_stats);
    std::swap(graph_on, o.graph_on);
    std::swap(gexec, o.gexec);
    std::swap(grid_cpp, o.grid_cpp);
    std::swap(cell_capacity, o.cell_capacity);
  }
};

namespace sph {
// shared helpers (sph_api.hip)
void fill_simdev(SimDev& d, const sph_params* p, int64_t n, int lg);
void set_grid(SimDev& d, int lg);
// box of the open axes around host positions (creation guess; the device
// sets it again before every pass) and the first-pass radius guess
void init_open_box(SimDev& d, const double* x, const double* y, int64_t n);
int check_params(const sph_params* p, int64_t n);
int fetch_sims(sph_sim* s);
int check_device_flags(sph_sim* s);
int alloc_state(sph_sim* h, int64_t np, int nitems);
int grid_log2(int64_t n, double cpp);
}  // namespace sph
