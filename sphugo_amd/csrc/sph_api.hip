// C ABI of the SPH step (declared in include/sphugo.h).
//
// One handle = one batch of independent simulations sharing device buffers
// (a single simulation is a batch of one).  State lives on the device as
// structure-of-arrays in cell order; host arrays are only touched by
// sph_create, sph_read_* and sph_write_field.
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sph_sim.hpp"

using namespace sph;

thread_local std::string g_err;

namespace sph {
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
}  // namespace sph

static bool is_open(const double a[2]) { return a[0] == -DBL_MAX; }

extern "C" int sph_default_params(sph_params* p) {
  if (!p) return fail(SPH_EINVAL, "null params");
  std::memset(p, 0, sizeof *p);
  p->k = 32;
  p->kernel = SPH_KERNEL_MONAGHAN;
  p->precision = SPH_PREC_FP64;
  p->gamma = 1.66666;
  p->mass = 1;
  p->dt_half = 0.001;
  p->cfl = 0.3;
  p->hor[0] = -DBL_MAX; p->hor[1] = DBL_MAX;
  p->vert[0] = -DBL_MAX; p->vert[1] = DBL_MAX;
  p->refl[0] = -DBL_MAX; p->refl[1] = DBL_MAX;
  p->refl[2] = -DBL_MAX; p->refl[3] = DBL_MAX;
  p->visc_alpha = 0.75;
  p->visc_beta = 1.5;
  p->visc_eta2 = 0.01;
  return 0;
}

int sph::check_params(const sph_params* p, int64_t n) {
  if (!p) return fail(SPH_EINVAL, "null params");
  if (n < 2) return fail(SPH_EINVAL, "need at least 2 particles, got %lld", (long long)n);
  if (n >= (1LL << 28)) return fail(SPH_EINVAL, "too many particles in one batch: %lld", (long long)n);
  if (p->k < 2 || p->k > kMaxK) return fail(SPH_EINVAL, "k must be in [2, %d], got %d", kMaxK, p->k);
  if (p->kernel != SPH_KERNEL_MONAGHAN && p->kernel != SPH_KERNEL_WENDLAND)
    return fail(SPH_EINVAL, "unknown kernel %d", p->kernel);
  if (p->precision != SPH_PREC_FP64 && p->precision != SPH_PREC_FP32C)
    return fail(SPH_EINVAL, "unknown precision %d", p->precision);
  if ((p->flags & SPH_FLAG_BITWISE) && p->precision != SPH_PREC_FP64)
    return fail(SPH_EINVAL, "bitwise reference order needs FP64 precision");
  // an axis is open when its lower bound is -MaxFloat64, and then its upper
  // bound must be +MaxFloat64 (the reference panics otherwise,
  // sim/nearest-neighbour.go:39-55); an upper bound of +MaxFloat64 under a
  // finite lower one is refused the same way
  const bool ox = is_open(p->hor), oy = is_open(p->vert);
  if (ox != (p->hor[1] == DBL_MAX) || oy != (p->vert[1] == DBL_MAX))
    return fail(SPH_EINVAL, "cannot have open and periodic boundary on one axis at the same time");
  if ((!ox && (!(p->hor[1] > p->hor[0]) || !std::isfinite(p->hor[1] - p->hor[0]))) ||
      (!oy && (!(p->vert[1] > p->vert[0]) || !std::isfinite(p->vert[1] - p->vert[0]))))
    return fail(SPH_EINVAL, "periodic box must have finite positive size");
  const int64_t images = (ox ? 1 : 3) * (oy ? 1 : 3);
  if (images * (n - 1) < p->k)
    return fail(SPH_EINVAL, "fewer than k candidates over the %lld images", (long long)images);
  if (!(p->gamma > 1) && !(p->gamma < 1)) return fail(SPH_EINVAL, "gamma must not be 1");
  if (p->dt_half <= 0 && !(p->cfl > 0)) return fail(SPH_EINVAL, "CFL number must be positive");
  // the viscosity's softening (the reference's constant 0.01, sim/sph.go:379-383)
  // keeps mu finite as two approaching particles meet; the FP32C force
  // evaluates it for every pair, so it must be a positive finite number
  if (!(p->visc_eta2 > 0) || !std::isfinite(p->visc_eta2))
    return fail(SPH_EINVAL, "visc_eta2 must be positive and finite, got %g", p->visc_eta2);
  return 0;
}

// log2 of cells per axis: a power of two with about `cpp` cells per
// particle at the mean density (Hilbert-ordered cells; coarser levels of the
// same grid serve the denser regions' neighbours and the sparser regions'
// search boxes).
int sph::grid_log2(int64_t n, double cpp) {
  int lg = 0;
  while (lg < 13 && (double)(1LL << (2 * lg)) < cpp * (double)n) ++lg;
  return lg;
}

void sph::fill_simdev(SimDev& d, const sph_params* p, int64_t n, int lg) {
  const int g = 1 << lg;
  std::memset(&d, 0, sizeof d);
  d.n = (int32_t)n;
  d.k = p->k;
  d.kernel = p->kernel;
  d.prec = p->precision;
  d.flags = p->flags;
  d.use_cfl = p->dt_half <= 0 ? 1 : 0;
  d.gx = g;
  d.gy = g;
  d.lg = lg;
  d.open = (is_open(p->hor) ? 1 : 0) | (is_open(p->vert) ? 2 : 0);
  // an open axis gets a unit placeholder here: its box is set around the
  // positions (sph::open_box), on the host for a creation guess and on the
  // device before every pass
  d.x0 = (d.open & 1) ? 0.0 : p->hor[0];
  d.y0 = (d.open & 2) ? 0.0 : p->vert[0];
  d.lx = (d.open & 1) ? 1.0 : p->hor[1] - p->hor[0];
  d.ly = (d.open & 2) ? 1.0 : p->vert[1] - p->vert[0];
  d.ivx = g / d.lx;
  d.ivy = g / d.ly;
  d.sx = d.lx / g;
  d.sy = d.ly / g;
  d.gamma = p->gamma;
  d.mass = p->mass;
  d.cfl = p->cfl;
  d.gxa = p->accel[0];
  d.gya = p->accel[1];
  d.hor0 = p->hor[0];
  d.hor1 = p->hor[1];
  d.vert0 = p->vert[0];
  d.vert1 = p->vert[1];
  for (int i = 0; i < 4; ++i) d.refl[i] = p->refl[i];
  d.alpha = p->visc_alpha;
  d.beta = p->visc_beta;
  d.eta2 = p->visc_eta2;
  d.fpref = p->kernel == SPH_KERNEL_WENDLAND ? kWendlandFPref : kMonaghanPref;
  d.dfpref = p->kernel == SPH_KERNEL_WENDLAND ? kWendlandDFPref : kMonaghanPref;
  d.cfac = p->gamma * (p->gamma - 1);
  d.dt_fixed = p->dt_half;
  d.dt_half = p->dt_half > 0 ? p->dt_half : 0.0;
  d.walk_delta = kWalkDefault.delta;
  d.walk_delta0 = kWalkDefault.delta;
  d.r_init = 1.2 * std::sqrt(p->k * d.lx * d.ly / (M_PI * (double)n));
}

void sph::init_open_box(SimDev& d, const double* x, const double* y, int64_t n) {
  if (!d.open) return;
  double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
  for (int64_t i = 0; i < n; ++i) {
    x0 = std::fmin(x0, x[i]);
    x1 = std::fmax(x1, x[i]);
    y0 = std::fmin(y0, y[i]);
    y1 = std::fmax(y1, y[i]);
  }
  open_box(d, x0, x1, y0, y1);
  // first-pass search radius guess over the occupied area
  d.r_init = 1.2 * std::sqrt(d.k * d.lx * d.ly / (M_PI * (double)n));
}

void sph::set_grid(SimDev& d, int lg) {
  const int g = 1 << lg;
  d.lg = lg;
  d.gx = g;
  d.gy = g;
  d.ivx = g / d.lx;
  d.ivy = g / d.ly;
  d.sx = d.lx / g;
  d.sy = d.ly / g;
}

// Cell offsets and key width of all simulations for `cpp` cells per particle.
static int64_t layout_cells(sph_sim* h, double cpp) {
  int64_t coff = 0;
  for (size_t s = 0; s < h->hsims.size(); ++s) {
    set_grid(h->hsims[s], grid_log2(h->sim_n[s], cpp));
    h->hsims[s].coff = coff;
    coff += (int64_t)h->hsims[s].gx * h->hsims[s].gy;
  }
  int bits = 1;
  while ((1LL << bits) < coff) ++bits;  // keys lie in [0, coff)
  h->pb.ncells = coff;
  h->pb.key_bits = bits;
  h->grid_cpp = cpp;
  return coff;
}

// Per-particle device state of a handle holding np particle slots in
// `nitems` work items (fields, positions, cell ranges, neighbour lists,
// force records, kNN and sort scratch).  cell_capacity must be set.
int sph::alloc_state(sph_sim* h, int64_t np, int nitems) {
  PassBuffers& pb = h->pb;
  int rc = 0;
  for (int q = 0; q < 2; ++q) {
    Fields& f = h->F[q];
    double** fl[] = {&f.vx, &f.vy, &f.e, &f.ax, &f.ay, &f.edot, &f.vpx, &f.vpy, &f.ep, &f.h, &f.rho, &f.c};
    for (double** a : fl)
      if ((rc = h->alloc(a, np))) return rc;
    if ((rc = h->alloc(&f.id, np))) return rc;
  }
  if ((rc = h->alloc(&h->P[0].x, np)) || (rc = h->alloc(&h->P[0].y, np)) ||
      (rc = h->alloc(&pb.xyp, np)) || (rc = h->alloc(&pb.uxy, np)))
    return rc;
  if ((rc = h->alloc(&pb.wrap, np)) || (rc = h->alloc(&pb.cstart, h->cell_capacity)) ||
      (rc = h->alloc(&pb.nbr, nbr_entries(np, pb.kmax))) ||
      (rc = h->alloc((char**)&pb.rec,
                     np * (int64_t)(h->prec == SPH_PREC_FP32C ? sizeof(RecF) : sizeof(RecD)))) ||
      (rc = h->alloc(&pb.partials, (int64_t)nitems * kNumRed)) ||
      (rc = h->alloc(&pb.stats, kStatSlots)) || (rc = h->alloc(&pb.err, 1 + pb.nsims)) ||
      (rc = h->alloc(&pb.walk_miss, (int64_t)pb.nsims * kMissSlots * kMissStride)) ||
      (rc = h->alloc(&pb.red_part, (int64_t)pb.nsims * kRedSplit * kNumRed)) ||
      (rc = h->alloc(&pb.retry, np)) || (rc = h->alloc(&pb.retry_count, 1)) ||
      (rc = h->alloc(&pb.runs, 4 * (int64_t)kRunQueue)) || (rc = h->alloc(&pb.run_count, 1)) ||
      (rc = h->alloc(&pb.knn_cnt, np)) || (rc = h->alloc(&pb.knn_ann, (np + 63) / 64 * 64 * kAnn)) ||
      (rc = h->alloc(&pb.sort.keys, np)) || (rc = h->alloc(&pb.sort.vals, np)) ||
      (rc = h->alloc(&pb.sort.keys_alt, np)) || (rc = h->alloc(&pb.sort.vals_alt, np)) ||
      (rc = h->alloc(&pb.sort.counts, sort_counts_size(np))) || (rc = h->alloc(&pb.sort.tot, 256)) ||
      (rc = h->alloc(&h->d_scratch, np)))
    return rc;
  return 0;
}

static const double kTuneCpp[] = {0.5, 1.0, 2.0, 4.0, 8.0, 16.0};

// Build a handle holding `count` simulations.
static int create_batch(int count, const sph_params* const* ps, const int64_t* ns,
                        const double* const* xs, const double* const* ys,
                        const double* const* vxs, const double* const* vys,
                        const double* const* es, const uint32_t* const* ids, sph_sim** out) {
  if (!out) return fail(SPH_EINVAL, "null output handle");
  *out = nullptr;
  if (count < 1) return fail(SPH_EINVAL, "empty batch");
  int64_t np = 0;
  int kmax = 0;
  for (int s = 0; s < count; ++s) {
    int rc = check_params(ps[s], ns[s]);
    if (rc) return rc;
    if (ps[s]->precision != ps[0]->precision || (ps[s]->flags & SPH_FLAG_BITWISE) !=
                                                    (ps[0]->flags & SPH_FLAG_BITWISE))
      return fail(SPH_EINVAL, "precision and order flags must agree across a batch");
    np += ns[s];
    kmax = ps[s]->k > kmax ? ps[s]->k : kmax;
  }
  if (np >= (1LL << 28)) return fail(SPH_EINVAL, "too many particles in one batch");
  auto* h = new sph_sim();
  h->np = np;
  h->kmax = kmax;
  h->prec = ps[0]->precision;
  h->bitwise = (ps[0]->flags & SPH_FLAG_BITWISE) ? 1 : 0;
  h->timing = (ps[0]->flags & SPH_FLAG_TIMING) != 0;
  int rc = 0;
  auto bail = [&](int code) {
    delete h;
    return code;
  };
  if (hipGetDevice(&h->device) != hipSuccess)
    return bail(fail(SPH_EDEVICE, "hipGetDevice failed"));
  if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(SPH_EDEVICE, "hipStreamCreate failed"));
  if (h->timing)
    for (auto& e : h->ev)
      if (hipEventCreate(&e) != hipSuccess) return bail(fail(SPH_EDEVICE, "hipEventCreate failed"));
  // simulations, cells, items
  int64_t off = 0, coff = 0;
  std::vector<int> red_first(count), red_count(count);
  for (int s = 0; s < count; ++s) {
    SimDev d;
    const int lg = grid_log2(ns[s], 4.0);
    fill_simdev(d, ps[s], ns[s], lg);
    if (d.open) {
      if (!xs[s] || !ys[s]) return bail(fail(SPH_EINVAL, "positions must not be null"));
      init_open_box(d, xs[s], ys[s], ns[s]);
      h->pb.any_open = 1;
    }
    d.off = off;
    h->hsims.push_back(d);
    h->hparams.push_back(*ps[s]);
    h->sim_n.push_back(ns[s]);
    red_first[s] = (int)h->hitems.size();
    for (int64_t b = 0; b < ns[s]; b += kItem) {
      Item it;
      it.sim = s;
      it.first = off + b;
      it.count = (int)((ns[s] - b) < kItem ? (ns[s] - b) : kItem);
      h->hitems.push_back(it);
    }
    red_count[s] = (int)h->hitems.size() - red_first[s];
    h->pb.red_parts_max = std::max(h->pb.red_parts_max, red_parts(red_count[s]));
    off += ns[s];
  }
  // cell storage for the finest grid the autotuner may choose
  for (double cpp : kTuneCpp) {
    const int64_t c = layout_cells(h, cpp);
    coff = c > coff ? c : coff;
  }
  if (coff >= (1LL << 32) - 1) return bail(fail(SPH_EINVAL, "too many cells"));
  h->cell_capacity = coff + 1;
  layout_cells(h, 4.0);
  PassBuffers& pb = h->pb;
  pb.nsims = count;
  pb.nitems = (int)h->hitems.size();
  pb.kmax = (kmax + 3) & ~3;
  pb.kout = kmax;
  pb.prec = h->prec;
  pb.bitwise = h->bitwise;
  pb.np = np;
  pb.walk = kWalkDefault;
  // device memory
  Item* d_items = nullptr;
  SimDev* d_sims = nullptr;
  int *d_rf = nullptr, *d_rc = nullptr;
  if ((rc = h->alloc(&d_items, pb.nitems)) || (rc = h->alloc(&d_sims, count)) ||
      (rc = h->alloc(&d_rf, count)) || (rc = h->alloc(&d_rc, count)))
    return bail(rc);
  if ((rc = alloc_state(h, np, pb.nitems))) return bail(rc);
  pb.items = d_items;
  pb.items_all = d_items;
  pb.nitems_all = pb.nitems;
  pb.hole_cpp = 0;
  pb.sims = d_sims;
  pb.red_first = d_rf;
  pb.red_count = d_rc;
  // upload
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->st) == hipSuccess;
  };
  bool ok = up(d_items, h->hitems.data(), h->hitems.size() * sizeof(Item)) &&
            up(d_sims, h->hsims.data(), h->hsims.size() * sizeof(SimDev)) &&
            up(d_rf, red_first.data(), count * sizeof(int)) &&
            up(d_rc, red_count.data(), count * sizeof(int));
  if (!ok) return bail(fail(SPH_EDEVICE, "upload failed"));
  for (int q = 0; q < 2; ++q) {
    Fields& f = h->F[q];
    double* zero[] = {f.vx, f.vy, f.e, f.ax, f.ay, f.edot, f.vpx, f.vpy, f.ep, f.h, f.rho, f.c};
    for (double* a : zero)
      if (hipMemsetAsync(a, 0, np * sizeof(double), h->st) != hipSuccess)
        return bail(fail(SPH_EDEVICE, "memset failed"));
  }
  if (hipMemsetAsync(pb.stats, 0, kStatSlots * sizeof(KnnStats), h->st) != hipSuccess ||
      hipMemsetAsync(pb.walk_miss, 0, (int64_t)pb.nsims * kMissSlots * kMissStride * 4, h->st) !=
          hipSuccess ||
      hipMemsetAsync(pb.run_count, 0, sizeof(uint32_t), h->st) != hipSuccess ||
      hipMemsetAsync(pb.err, 0, (1 + pb.nsims) * sizeof(int), h->st) != hipSuccess ||
      hipMemsetAsync(pb.nbr, 0, nbr_entries(np, pb.kmax) * sizeof(uint32_t), h->st) != hipSuccess)
    return bail(fail(SPH_EDEVICE, "memset failed"));
  std::vector<uint32_t> idbuf;
  for (int s = 0; s < count; ++s) {
    const int64_t o = h->hsims[s].off, n = ns[s];
    if (!xs[s] || !ys[s]) return bail(fail(SPH_EINVAL, "positions must not be null"));
    for (int64_t i = 0; i < n; ++i)
      if (!std::isfinite(xs[s][i]) || !std::isfinite(ys[s][i]))
        return bail(fail(SPH_ERANGE, "non-finite position of particle %lld", (long long)i));
    ok = up(h->P[0].x + o, xs[s], n * sizeof(double)) && up(h->P[0].y + o, ys[s], n * sizeof(double));
    if (vxs && vxs[s]) ok = ok && up(h->F[0].vx + o, vxs[s], n * sizeof(double));
    if (vys && vys[s]) ok = ok && up(h->F[0].vy + o, vys[s], n * sizeof(double));
    if (es && es[s]) ok = ok && up(h->F[0].e + o, es[s], n * sizeof(double));
    if (!ok) return bail(fail(SPH_EDEVICE, "upload failed"));
    idbuf.resize(n);
    if (ids && ids[s]) {
      std::vector<char> seen(n, 0);
      for (int64_t i = 0; i < n; ++i) {
        if (ids[s][i] >= (uint64_t)n || seen[ids[s][i]])
          return bail(fail(SPH_EINVAL, "ids must be a permutation of 0..n-1"));
        seen[ids[s][i]] = 1;
        idbuf[i] = ids[s][i];
      }
    } else {
      for (int64_t i = 0; i < n; ++i) idbuf[i] = (uint32_t)i;
    }
    if (hipMemcpy(h->F[0].id + o, idbuf.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice) !=
        hipSuccess)
      return bail(fail(SPH_EDEVICE, "upload failed"));
  }
  h->bind();
  if (enqueue_init_sims(pb, h->st)) return bail(fail(SPH_EDEVICE, "init kernel failed"));
  if (hipStreamSynchronize(h->st) != hipSuccess)
    return bail(fail(SPH_EDEVICE, "create: %s", hipGetErrorString(hipGetLastError())));
  *out = h;
  return 0;
}

extern "C" int sph_create(const sph_params* p, int64_t n, const double* x, const double* y,
                          const double* vx, const double* vy, const double* e,
                          const uint32_t* id, sph_sim** out) {
  return create_batch(1, &p, &n, &x, &y, &vx, &vy, &e, &id, out);
}

extern "C" void sph_destroy(sph_sim* s) {
  if (!s) return;
  if (s->st) (void)hipStreamSynchronize(s->st);  // teardown
  delete s;
}

static int run_pass(sph_sim* s, int perm_mode, int force_mode) {
  s->bind();
  int rc = enqueue_pass(s->pb, perm_mode, force_mode, s->st, s->timing ? s->ev : nullptr);
  if (rc) return fail(SPH_EDEVICE, "kernel launch failed: %s", hipGetErrorString((hipError_t)rc));
  s->cur ^= 1;
  s->pb.have_h = 1;
  s->bind();
  if (s->timing) {
    HIPCHK(hipStreamSynchronize(s->st));
    for (int t = 0; t < SPH_NTIMERS; ++t) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, s->ev[2 * t], s->ev[2 * t + 1]) == hipSuccess) {
        s->t_ms[t] += ms;
        s->t_cnt[t] += 1;
      }
    }
  }
  return 0;
}

int sph::check_device_flags(sph_sim* s) {
  int err = 0;
  KnnStats slots[kStatSlots];
  HIPCHK(hipMemcpyAsync(&err, s->pb.err, sizeof(int), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipMemcpyAsync(slots, s->pb.stats, sizeof slots, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  const KnnStats st = sum_stats(slots);
  s->last_stats = st;
  if (err) return fail(SPH_ERANGE, "a particle left the periodic images [-L, 2L) of the box");
  if (st.failures) return fail(SPH_ERANGE, "kNN: %llu searches found fewer than k neighbours",
                               (unsigned long long)st.failures);
  return 0;
}

// One regular step from a captured graph (captured on first use for the
// current field-buffer parity).  The kernels read every per-step value
// (half step, counters) from device memory, so replays need no host input.
static int graph_step(sph_sim* s) {
  const int g = s->cur;
  if (!s->gexec[g]) {
    s->bind();
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_pass(s->pb, kPermPredict, kForceKick, s->st, nullptr);
    hipError_t e = hipStreamEndCapture(s->st, &graph);
    if (rc || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return fail(SPH_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&s->gexec[g], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(SPH_EDEVICE, "graph instantiate: %s", hipGetErrorString(e));
  }
  HIPCHK(hipGraphLaunch(s->gexec[g], s->st));
  s->cur ^= 1;
  s->bind();
  return 0;
}

extern "C" int sph_step_async(sph_sim* s, int32_t nsteps) {
  if (!s) return fail(SPH_EINVAL, "null handle");
  if (nsteps < 0) return fail(SPH_EINVAL, "negative step count");
  for (int32_t i = 0; i < nsteps; ++i) {
    int rc;
    if (s->step == 0 && (rc = run_pass(s, kPermInit, kForceNoKick))) return rc;
    if (s->graph_on && !s->timing && s->pb.have_h) rc = graph_step(s);
    else rc = run_pass(s, kPermPredict, kForceKick);
    if (rc) return rc;
    s->step += 1;
  }
  return 0;
}

extern "C" int sph_set_graph(sph_sim* s, int32_t on) {
  if (!s) return fail(SPH_EINVAL, "null handle");
  HIPCHK(hipStreamSynchronize(s->st));
  s->drop_graphs();
  s->graph_on = on != 0;
  return 0;
}

// Autotuner: times force passes over the candidate grid resolutions on the
// current state (only before the first step) and keeps the fastest.  The
// step-0 pass that follows searches every query exactly, as an untuned
// handle's does.  The grid changes the order in which neighbours are found
// and so the summation order: results stay bit-identical in the bitwise
// form and within the 1e-12 (FP64) and 1e-5 (FP32C density) contracts
// otherwise; neighbour sets and h are the same.
extern "C" int sph_autotune(sph_sim* s, double* best_ms, double* best_cpp) {
  if (!s) return fail(SPH_EINVAL, "null handle");
  if (s->step != 0) return fail(SPH_ESTATE, "autotune runs before the first step");
  HIPCHK(hipStreamSynchronize(s->st));
  s->drop_graphs();
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  double best = 1e30, bcpp = s->grid_cpp;
  const bool timing = s->timing;
  s->timing = false;
  // a pass to set h on the current grid, then the faster of two timed passes
  // (one pass's time varies by a few per cent, as much as the choices differ)
  auto timed_pass = [&](float* ms) -> int {
    int r = run_pass(s, kPermInit, kForceNoKick);
    if (r) return r;
    *ms = 1e30f;
    for (int rep = 0; rep < 2; ++rep) {
      (void)hipEventRecord(e0, s->st);  // a failed record shows in the sync
      if ((r = run_pass(s, kPermInit, kForceNoKick))) return r;
      (void)hipEventRecord(e1, s->st);
      if (hipEventSynchronize(e1) != hipSuccess) return fail(SPH_EDEVICE, "autotune pass failed");
      float t = 0.f;
      if (hipEventElapsedTime(&t, e0, e1) != hipSuccess)
        return fail(SPH_EDEVICE, "autotune timing failed");
      *ms = t < *ms ? t : *ms;
    }
    return 0;
  };
  int rc = 0;
  for (double cpp : kTuneCpp) {
    // keep the device-updated fields (half step, CFL minimum, totals)
    if ((rc = fetch_sims(s))) break;
    layout_cells(s, cpp);
    if (hipMemcpyAsync(s->pb.sims, s->hsims.data(), s->hsims.size() * sizeof(SimDev),
                       hipMemcpyHostToDevice, s->st) != hipSuccess) {
      rc = fail(SPH_EDEVICE, "upload failed");
      break;
    }
    float ms = 0.f;
    if ((rc = timed_pass(&ms))) break;
    if (ms < best) {
      best = ms;
      bcpp = cpp;
    }
  }
  // then the walk's tile shape on that grid, one knob at a time (sub-tiles,
  // first lane group, split threshold); the bracket width stays: it answers
  // to the motion between passes, which passes at rest cannot show
  if (!rc && !s->bitwise) {
    if ((rc = fetch_sims(s)) == 0) {
      layout_cells(s, bcpp);
      if (hipMemcpyAsync(s->pb.sims, s->hsims.data(), s->hsims.size() * sizeof(SimDev),
                         hipMemcpyHostToDevice, s->st) != hipSuccess)
        rc = fail(SPH_EDEVICE, "upload failed");
    }
    WalkKnobs bk = s->pb.walk;
    float bms = 0.f;
    if (!rc) rc = timed_pass(&bms);
    const int subs[] = {8, 4}, groups[] = {64, 32}, splits[] = {24, 32, 16};
    for (int knob = 0; knob < 3 && !rc; ++knob) {
      const int* vals = knob == 0 ? subs : (knob == 1 ? groups : splits);
      const int nv = knob == 2 ? 3 : 2;
      for (int i = 0; i < nv && !rc; ++i) {
        WalkKnobs kk = bk;
        int& slot = knob == 0 ? kk.nsub : (knob == 1 ? kk.group : kk.split);
        if (slot == vals[i]) continue;
        slot = vals[i];
        s->pb.walk = kk;
        float ms = 0.f;
        if ((rc = timed_pass(&ms))) break;
        if (ms < bms) {
          bms = ms;
          bk = kk;
        }
      }
    }
    s->pb.walk = bk;
    if (!rc) best = bms;
  }
  s->timing = timing;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if ((rc = fetch_sims(s))) return rc;
  layout_cells(s, bcpp);
  HIPCHK(hipMemcpyAsync(s->pb.sims, s->hsims.data(), s->hsims.size() * sizeof(SimDev),
                        hipMemcpyHostToDevice, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  s->pb.have_h = 0;  // step 0 searches exactly, as without tuning
  if (best_ms) *best_ms = best;
  if (best_cpp) *best_cpp = bcpp;
  return check_device_flags(s);
}

static bool valid_knobs(const WalkKnobs& k) {
  return k.delta >= 0.005 && k.delta <= 0.1 && k.split >= 8 && k.split <= 64 &&
         (k.group == 16 || k.group == 32 || k.group == 64) && (k.nsub == 4 || k.nsub == 8);
}

extern "C" int sph_get_tuning(sph_sim* s, double out[5]) {
  if (!s || !out) return fail(SPH_EINVAL, "null argument");
  int rc = fetch_sims(s);  // the bracket adapted by the passes so far
  if (rc) return rc;
  out[0] = s->grid_cpp;
  out[1] = s->hsims[0].walk_delta;
  out[2] = s->pb.walk.split;
  out[3] = s->pb.walk.group;
  out[4] = s->pb.walk.nsub;
  return 0;
}

extern "C" int sph_set_tuning(sph_sim* s, const double in[5]) {
  if (!s || !in) return fail(SPH_EINVAL, "null argument");
  bool known_grid = false;
  for (double cpp : kTuneCpp) known_grid = known_grid || cpp == in[0];
  WalkKnobs k = kWalkDefault;
  k.delta = in[1];
  k.split = (int32_t)in[2];
  k.group = (int32_t)in[3];
  k.nsub = (int32_t)in[4];
  if (!known_grid || !valid_knobs(k) || k.split != in[2] || k.group != in[3] || k.nsub != in[4])
    return fail(SPH_EINVAL, "tuning out of range (grid 0.5..16 cells per particle, delta "
                            "0.005..0.1, split 8..64, group 16/32/64, sub-tiles 4/8)");
  int rc = fetch_sims(s);  // waits for the handle's work; keeps device scalars
  if (rc) return rc;
  s->drop_graphs();
  layout_cells(s, in[0]);
  s->pb.walk = k;
  for (SimDev& d : s->hsims) d.walk_delta = d.walk_delta0 = k.delta;  // where every bracket restarts
  HIPCHK(hipMemcpyAsync(s->pb.sims, s->hsims.data(), s->hsims.size() * sizeof(SimDev),
                        hipMemcpyHostToDevice, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return 0;
}

extern "C" int sph_synchronize(sph_sim* s) {
  if (!s) return fail(SPH_EINVAL, "null handle");
  HIPCHK(hipStreamSynchronize(s->st));
  return check_device_flags(s);
}

extern "C" int sph_step(sph_sim* s, int32_t nsteps) {
  int rc = sph_step_async(s, nsteps);
  if (rc) return rc;
  return sph_synchronize(s);
}

extern "C" int sph_calculate_forces(sph_sim* s) {
  if (!s) return fail(SPH_EINVAL, "null handle");
  int rc = run_pass(s, kPermKeep, kForceNoKick);
  if (rc) return rc;
  return sph_synchronize(s);
}

extern "C" int sph_read_field(sph_sim* s, int32_t field, double* out) {
  if (!s || !out) return fail(SPH_EINVAL, "null argument");
  if (field < 0 || field >= SPH_NFIELDS) return fail(SPH_EINVAL, "unknown field %d", field);
  s->bind();
  int rc = enqueue_read_field(s->pb, field, -1, s->d_scratch, s->st);
  if (rc) return fail(SPH_EDEVICE, "readback launch failed");
  HIPCHK(hipMemcpyAsync(out, s->d_scratch, s->np * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return 0;
}

extern "C" int sph_write_field(sph_sim* s, int32_t field, const double* in) {
  if (!s || !in) return fail(SPH_EINVAL, "null argument");
  if (field < 0 || field > SPH_VPY || field == SPH_RHO || field == SPH_C)
    return fail(SPH_EINVAL, "field %d cannot be written", field);
  if (field == SPH_X || field == SPH_Y)
    for (int64_t i = 0; i < s->np; ++i)
      if (!std::isfinite(in[i])) return fail(SPH_ERANGE, "non-finite position");
  s->bind();
  HIPCHK(hipMemcpyAsync(s->d_scratch, in, s->np * sizeof(double), hipMemcpyHostToDevice, s->st));
  int rc = enqueue_write_field(s->pb, field, -1, s->d_scratch, s->st);
  if (rc) return fail(SPH_EDEVICE, "upload launch failed");
  HIPCHK(hipStreamSynchronize(s->st));
  return 0;
}

extern "C" int sph_read_neighbours(sph_sim* s, uint32_t* ids, double* d2) {
  if (!s || !ids) return fail(SPH_EINVAL, "null argument");
  const int64_t total = s->np * s->pb.kout;
  int rc = 0;
  if (!s->d_nids && ((rc = s->alloc(&s->d_nids, total)) || (rc = s->alloc(&s->d_nd2, total))))
    return rc;
  s->bind();
  rc = enqueue_read_neighbours(s->pb, s->d_nids, s->d_nd2, nullptr, nullptr, s->st);
  if (rc) return fail(SPH_EDEVICE, "neighbour readback launch failed");
  HIPCHK(hipMemcpyAsync(ids, s->d_nids, total * sizeof(uint32_t), hipMemcpyDeviceToHost, s->st));
  if (d2) HIPCHK(hipMemcpyAsync(d2, s->d_nd2, total * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return 0;
}

extern "C" int sph_read_neighbour_pos(sph_sim* s, double* px, double* py) {
  if (!s || !px || !py) return fail(SPH_EINVAL, "null argument");
  const int64_t total = s->np * s->pb.kout;
  int rc = 0;
  if (!s->d_nids && ((rc = s->alloc(&s->d_nids, total)) || (rc = s->alloc(&s->d_nd2, total))))
    return rc;
  if (!s->d_npx && ((rc = s->alloc(&s->d_npx, total)) || (rc = s->alloc(&s->d_npy, total))))
    return rc;
  s->bind();
  rc = enqueue_read_neighbours(s->pb, s->d_nids, s->d_nd2, s->d_npx, s->d_npy, s->st);
  if (rc) return fail(SPH_EDEVICE, "neighbour readback launch failed");
  HIPCHK(hipMemcpyAsync(px, s->d_npx, total * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipMemcpyAsync(py, s->d_npy, total * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return 0;
}

int sph::fetch_sims(sph_sim* s) {
  HIPCHK(hipMemcpyAsync(s->hsims.data(), s->pb.sims, s->hsims.size() * sizeof(SimDev),
                        hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return 0;
}

extern "C" int sph_reductions(sph_sim* s, double out[6]) {
  if (!s || !out) return fail(SPH_EINVAL, "null argument");
  int rc = fetch_sims(s);
  if (rc) return rc;
  const SimDev& d = s->hsims[0];
  for (int r = 0; r < kNumRed; ++r) out[r] = d.red[r];
  out[5] = d.dt_last;
  return 0;
}

extern "C" int sph_total_energy(sph_sim* s, double* out) {
  double r[6];
  int rc = sph_reductions(s, r);
  if (!rc) *out = r[0];
  return rc;
}

extern "C" int sph_total_density(sph_sim* s, double* out) {
  double r[6];
  int rc = sph_reductions(s, r);
  if (!rc) *out = r[1];
  return rc;
}

extern "C" int sph_total_momentum(sph_sim* s, double* out) {
  if (!s || !out) return fail(SPH_EINVAL, "null argument");
  int rc = fetch_sims(s);
  if (!rc) *out = s->hsims[0].last_speed;
  return rc;
}

extern "C" int64_t sph_current_step(const sph_sim* s) { return s ? s->step : -1; }
extern "C" int64_t sph_num_particles(const sph_sim* s) { return s ? s->np : -1; }

extern "C" int sph_next_dt_half(sph_sim* s, double* out) {
  if (!s || !out) return fail(SPH_EINVAL, "null argument");
  int rc = fetch_sims(s);
  if (!rc) *out = s->hsims[0].dt_half;
  return rc;
}

extern "C" int sph_kernel_times(sph_sim* s, double ms[SPH_NTIMERS], int64_t counts[SPH_NTIMERS]) {
  if (!s || !ms) return fail(SPH_EINVAL, "null argument");
  for (int t = 0; t < SPH_NTIMERS; ++t) {
    ms[t] = s->t_cnt[t] ? s->t_ms[t] / s->t_cnt[t] : 0.0;
    if (counts) counts[t] = s->t_cnt[t];
  }
  return 0;
}

extern "C" int sph_reset_kernel_times(sph_sim* s) {
  if (!s) return fail(SPH_EINVAL, "null handle");
  for (int t = 0; t < SPH_NTIMERS; ++t) {
    s->t_ms[t] = 0;
    s->t_cnt[t] = 0;
  }
  return 0;
}

extern "C" int sph_set_timing(sph_sim* s, int32_t on) {
  if (!s) return fail(SPH_EINVAL, "null handle");
  if (on && !s->ev[0])
    for (auto& e : s->ev)
      if (hipEventCreate(&e) != hipSuccess) return fail(SPH_EDEVICE, "hipEventCreate failed");
  s->timing = on != 0;
  return 0;
}

extern "C" int sph_knn_stats(sph_sim* s, int64_t* retries, int64_t* ties) {
  if (!s) return fail(SPH_EINVAL, "null handle");
  KnnStats slots[kStatSlots];
  HIPCHK(hipMemcpyAsync(slots, s->pb.stats, sizeof slots, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  const KnnStats st = sum_stats(slots);
  if (retries) *retries = (int64_t)st.retries;
  if (ties) *ties = (int64_t)st.ties;
  return 0;
}

extern "C" int sph_knn_counters(sph_sim* s, int64_t out[7]) {
  if (!s || !out) return fail(SPH_EINVAL, "null argument");
  KnnStats slots[kStatSlots];
  HIPCHK(hipMemcpyAsync(slots, s->pb.stats, sizeof slots, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  const KnnStats st = sum_stats(slots);
  const unsigned long long v[7] = {st.retries, st.ties, st.failures, st.miss_lo,
                                   st.miss_hi, st.miss_ann, st.split_fallback};
  for (int i = 0; i < 7; ++i) out[i] = (int64_t)v[i];
  return 0;
}

extern "C" int sph_batch_create(int32_t count, const sph_params* params, const int64_t* n,
                                const double* const* x, const double* const* y,
                                const double* const* vx, const double* const* vy,
                                const double* const* e, sph_sim** out) {
  if (!params || !n || !x || !y || count < 1) return fail(SPH_EINVAL, "bad batch arguments");
  std::vector<const sph_params*> ps(count);
  for (int i = 0; i < count; ++i) ps[i] = params + i;
  return create_batch(count, ps.data(), n, x, y, vx, vy, e, nullptr, out);
}

extern "C" int32_t sph_batch_count(const sph_sim* s) { return s ? (int32_t)s->hsims.size() : -1; }

extern "C" int64_t sph_batch_num_particles(const sph_sim* s, int32_t sim) {
  if (!s || sim < 0 || sim >= (int32_t)s->hsims.size()) return -1;
  return s->sim_n[sim];
}

extern "C" int sph_batch_read_field(sph_sim* s, int32_t sim, int32_t field, double* out) {
  if (!s || !out) return fail(SPH_EINVAL, "null argument");
  if (sim < 0 || sim >= (int32_t)s->hsims.size()) return fail(SPH_EINVAL, "no simulation %d", sim);
  if (field < 0 || field >= SPH_NFIELDS) return fail(SPH_EINVAL, "unknown field %d", field);
  s->bind();
  if (enqueue_read_field(s->pb, field, sim, s->d_scratch, s->st))
    return fail(SPH_EDEVICE, "readback launch failed");
  HIPCHK(hipMemcpyAsync(out, s->d_scratch, s->sim_n[sim] * sizeof(double), hipMemcpyDeviceToHost,
                        s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return 0;
}

extern "C" int sph_batch_write_field(sph_sim* s, int32_t sim, int32_t field, const double* in) {
  if (!s || !in) return fail(SPH_EINVAL, "null argument");
  if (sim < 0 || sim >= (int32_t)s->hsims.size()) return fail(SPH_EINVAL, "no simulation %d", sim);
  if (field < 0 || field > SPH_VPY || field == SPH_RHO || field == SPH_C)
    return fail(SPH_EINVAL, "field %d cannot be written", field);
  s->bind();
  HIPCHK(hipMemcpyAsync(s->d_scratch, in, s->sim_n[sim] * sizeof(double), hipMemcpyHostToDevice,
                        s->st));
  if (enqueue_write_field(s->pb, field, sim, s->d_scratch, s->st))
    return fail(SPH_EDEVICE, "upload launch failed");
  HIPCHK(hipStreamSynchronize(s->st));
  return 0;
}

extern "C" int sph_batch_reductions(sph_sim* s, int32_t sim, double out[6]) {
  if (!s || !out) return fail(SPH_EINVAL, "null argument");
  if (sim < 0 || sim >= (int32_t)s->hsims.size()) return fail(SPH_EINVAL, "no simulation %d", sim);
  int rc = fetch_sims(s);
  if (rc) return rc;
  const SimDev& d = s->hsims[sim];
  for (int r = 0; r < kNumRed; ++r) out[r] = d.red[r];
  out[5] = d.dt_last;
  return 0;
}

// ------------------------------------------------------------------------
// Checkpoint / resume: header, creation parameters, the simulation's device
// scalars, then every per-particle array in slot order and the IDs.  Slot
// order (not ID order) is what the stable sort of the next pass starts from,
// so a loaded handle continues bit for bit.

namespace {
constexpr char kCkptMagic[8] = {'S', 'P', 'H', 'C', 'K', 'P', 'T', '1'};
constexpr uint32_t kCkptVersion = 1;
constexpr int kCkptFields = 14;  // x y vx vy e ax ay edot vpx vpy ep h rho c

struct CkptHeader {
  char magic[8];
  uint32_t version;
  uint32_t header_bytes;
  int64_t n;
  int64_t step;
  uint8_t have_h;
  uint8_t walk_split;  // the walk's tile shape (zero in images written
  uint8_t walk_group;  // before it was kept: the defaults)
  uint8_t walk_nsub;
  uint32_t params_bytes;
  uint32_t simdev_bytes;
  uint32_t nfields;
  double grid_cpp;
  double walk_delta;
};
static_assert(sizeof(CkptHeader) == 64, "checkpoint header is 64 bytes");

int64_t ckpt_bytes(int64_t n) {
  return (int64_t)(sizeof(CkptHeader) + sizeof(sph_params) + sizeof(SimDev)) +
         n * (int64_t)(kCkptFields * sizeof(double) + sizeof(uint32_t));
}

// device arrays of the payload, in payload order (the state between steps:
// positions in P[0], everything else in the current field set)
void ckpt_arrays(sph_sim* s, double* a[kCkptFields]) {
  const Fields& f = s->F[s->cur];
  double* v[kCkptFields] = {s->P[0].x, s->P[0].y, f.vx,  f.vy, f.e, f.ax,  f.ay,
                            f.edot,    f.vpx,     f.vpy, f.ep, f.h, f.rho, f.c};
  for (int i = 0; i < kCkptFields; ++i) a[i] = v[i];
}

// The fields of a simulation record that follow from its parameters, its
// size and its grid (everything but what the device updates between steps,
// and the box and first-pass radius of open axes, which are set around the
// positions).
bool same_record(const SimDev& a, const SimDev& b) {
  bool same = a.off == b.off && a.coff == b.coff && a.n == b.n && a.k == b.k &&
              a.kernel == b.kernel && a.prec == b.prec && a.flags == b.flags &&
              a.use_cfl == b.use_cfl && a.gx == b.gx && a.gy == b.gy && a.lg == b.lg &&
              a.pmode == b.pmode && a.fmode == b.fmode && a.open == b.open;
  auto eq = [](double u, double v) { return std::memcmp(&u, &v, sizeof u) == 0; };
  const double* pa[] = {&a.gamma, &a.mass, &a.cfl, &a.gxa, &a.gya, &a.hor0, &a.hor1, &a.vert0,
                        &a.vert1, &a.refl[0], &a.refl[1], &a.refl[2], &a.refl[3], &a.alpha,
                        &a.beta, &a.eta2, &a.fpref, &a.dfpref, &a.cfac, &a.dt_fixed};
  const double* pb[] = {&b.gamma, &b.mass, &b.cfl, &b.gxa, &b.gya, &b.hor0, &b.hor1, &b.vert0,
                        &b.vert1, &b.refl[0], &b.refl[1], &b.refl[2], &b.refl[3], &b.alpha,
                        &b.beta, &b.eta2, &b.fpref, &b.dfpref, &b.cfac, &b.dt_fixed};
  for (size_t i = 0; i < sizeof pa / sizeof pa[0]; ++i) same = same && eq(*pa[i], *pb[i]);
  if (!(a.open & 1))
    same = same && eq(a.x0, b.x0) && eq(a.lx, b.lx) && eq(a.ivx, b.ivx) && eq(a.sx, b.sx);
  if (!(a.open & 2))
    same = same && eq(a.y0, b.y0) && eq(a.ly, b.ly) && eq(a.ivy, b.ivy) && eq(a.sy, b.sy);
  if (!a.open) same = same && eq(a.r_init, b.r_init);
  return same;
}

// The scalars a device pass updates (half steps, CFL minimum, totals).
void take_device_scalars(SimDev& d, const SimDev& from) {
  d.dt_half = from.dt_half;
  d.dt_last = from.dt_last;
  d.cfl_min_bits = from.cfl_min_bits;
  for (int r = 0; r < kNumRed; ++r) d.red[r] = from.red[r];
  d.last_speed = from.last_speed;
  d.walk_delta = from.walk_delta;
  d.walk_delta0 = from.walk_delta0;
}
}  // namespace

extern "C" int sph_checkpoint_size(sph_sim* s, int64_t* bytes) {
  if (!s || !bytes) return fail(SPH_EINVAL, "null argument");
  if (s->hsims.size() != 1) return fail(SPH_EINVAL, "a checkpoint holds one simulation, not a batch");
  *bytes = ckpt_bytes(s->np);
  return 0;
}

extern "C" int sph_checkpoint_save(sph_sim* s, void* buf, int64_t bytes) {
  if (!s || !buf) return fail(SPH_EINVAL, "null argument");
  if (s->hsims.size() != 1) return fail(SPH_EINVAL, "a checkpoint holds one simulation, not a batch");
  const int64_t n = s->np;
  if (bytes < ckpt_bytes(n))
    return fail(SPH_EINVAL, "checkpoint buffer of %lld bytes, %lld needed", (long long)bytes,
                (long long)ckpt_bytes(n));
  int rc = fetch_sims(s);  // also waits for the handle's work
  if (rc) return rc;
  CkptHeader hd;
  std::memset(&hd, 0, sizeof hd);
  std::memcpy(hd.magic, kCkptMagic, sizeof hd.magic);
  hd.version = kCkptVersion;
  hd.header_bytes = sizeof hd;
  hd.n = n;
  hd.step = s->step;
  hd.have_h = (uint8_t)s->pb.have_h;
  hd.params_bytes = sizeof(sph_params);
  hd.simdev_bytes = sizeof(SimDev);
  hd.nfields = kCkptFields;
  hd.grid_cpp = s->grid_cpp;
  hd.walk_delta = s->pb.walk.delta;
  hd.walk_split = (uint8_t)s->pb.walk.split;
  hd.walk_group = (uint8_t)s->pb.walk.group;
  hd.walk_nsub = (uint8_t)s->pb.walk.nsub;
  char* p = (char*)buf;
  std::memcpy(p, &hd, sizeof hd);
  p += sizeof hd;
  std::memcpy(p, &s->hparams[0], sizeof(sph_params));
  p += sizeof(sph_params);
  std::memcpy(p, &s->hsims[0], sizeof(SimDev));
  p += sizeof(SimDev);
  double* a[kCkptFields];
  ckpt_arrays(s, a);
  for (int i = 0; i < kCkptFields; ++i) {
    HIPCHK(hipMemcpyAsync(p, a[i], n * sizeof(double), hipMemcpyDeviceToHost, s->st));
    p += n * sizeof(double);
  }
  HIPCHK(hipMemcpyAsync(p, s->F[s->cur].id, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return 0;
}

extern "C" int sph_checkpoint_load(const void* buf, int64_t bytes, sph_sim** out) {
  if (!buf || !out) return fail(SPH_EINVAL, "null argument");
  *out = nullptr;
  CkptHeader hd;
  if (bytes < (int64_t)sizeof hd) return fail(SPH_EINVAL, "not a checkpoint: %lld bytes", (long long)bytes);
  std::memcpy(&hd, buf, sizeof hd);
  if (std::memcmp(hd.magic, kCkptMagic, sizeof hd.magic) != 0)
    return fail(SPH_EINVAL, "not a checkpoint (bad magic)");
  if (hd.version != kCkptVersion || hd.header_bytes != sizeof hd ||
      hd.params_bytes != sizeof(sph_params) || hd.simdev_bytes != sizeof(SimDev) ||
      hd.nfields != kCkptFields)
    return fail(SPH_EINVAL, "checkpoint version %u or layout is not this build's", hd.version);
  if (hd.n < 1 || hd.n >= (1LL << 28) || hd.step < 0 || (hd.have_h != 0 && hd.have_h != 1))
    return fail(SPH_EINVAL, "corrupt checkpoint header");
  const int64_t n = hd.n;
  if (bytes < ckpt_bytes(n))
    return fail(SPH_EINVAL, "truncated checkpoint: %lld bytes, %lld needed", (long long)bytes,
                (long long)ckpt_bytes(n));
  bool known_grid = false;
  for (double cpp : kTuneCpp) known_grid = known_grid || cpp == hd.grid_cpp;
  if (!known_grid) return fail(SPH_EINVAL, "corrupt checkpoint header (grid)");
  WalkKnobs knobs = kWalkDefault;
  if (hd.walk_nsub != 0) {
    knobs.delta = hd.walk_delta;
    knobs.split = hd.walk_split;
    knobs.group = hd.walk_group;
    knobs.nsub = hd.walk_nsub;
    if (!valid_knobs(knobs)) return fail(SPH_EINVAL, "corrupt checkpoint header (walk shape)");
  }
  const char* p = (const char*)buf + sizeof hd;
  sph_params prm;
  std::memcpy(&prm, p, sizeof prm);
  p += sizeof prm;
  SimDev sd;
  std::memcpy(&sd, p, sizeof sd);
  p += sizeof sd;
  // The parameters must be valid, and the simulation record must be the one
  // they give (grid, box, physics, offsets): only the scalars the device
  // updates between steps are taken from the image.  Checked before any
  // device call.
  int rc = check_params(&prm, n);
  if (rc) return fail(SPH_EINVAL, "corrupt checkpoint (parameters): %s", sph_last_error());
  {
    SimDev ref;
    fill_simdev(ref, &prm, n, grid_log2(n, hd.grid_cpp));
    if (!same_record(sd, ref))
      return fail(SPH_EINVAL, "corrupt checkpoint (simulation record does not match its parameters)");
    if (!std::isfinite(sd.dt_half) || sd.dt_half < 0 || !std::isfinite(sd.dt_last))
      return fail(SPH_EINVAL, "corrupt checkpoint (half step)");
    if (!(sd.walk_delta >= 0.005 && sd.walk_delta <= kDeltaMax && sd.walk_delta0 >= 0.005 &&
          sd.walk_delta0 <= kDeltaMax))
      return fail(SPH_EINVAL, "corrupt checkpoint (kNN bracket)");
  }
  // aligned host copies of what sph_create takes; the rest goes straight up
  std::vector<double> init(5 * n);
  for (int i = 0; i < 5; ++i) std::memcpy(init.data() + i * n, p + i * n * sizeof(double), n * sizeof(double));
  std::vector<uint32_t> ids(n);
  std::memcpy(ids.data(), p + kCkptFields * n * sizeof(double), n * sizeof(uint32_t));
  const double* xs[5];
  for (int i = 0; i < 5; ++i) xs[i] = init.data() + i * n;
  const uint32_t* idp = ids.data();
  const sph_params* pp = &prm;
  sph_sim* h = nullptr;
  rc = create_batch(1, &pp, &n, &xs[0], &xs[1], &xs[2], &xs[3], &xs[4], &idp, &h);
  if (rc) return rc;
  auto bail = [&](int code) {
    sph_destroy(h);
    return code;
  };
  layout_cells(h, hd.grid_cpp);
  h->pb.walk = knobs;
  take_device_scalars(h->hsims[0], sd);
  double* a[kCkptFields];
  ckpt_arrays(h, a);
  for (int i = 5; i < kCkptFields; ++i)
    if (hipMemcpyAsync(a[i], p + i * n * sizeof(double), n * sizeof(double), hipMemcpyHostToDevice,
                       h->st) != hipSuccess)
      return bail(fail(SPH_EDEVICE, "upload failed"));
  if (hipMemcpyAsync(h->pb.sims, h->hsims.data(), sizeof(SimDev), hipMemcpyHostToDevice, h->st) !=
          hipSuccess ||
      hipStreamSynchronize(h->st) != hipSuccess)
    return bail(fail(SPH_EDEVICE, "upload failed"));
  h->step = hd.step;
  h->pb.have_h = hd.have_h;
  h->bind();
  *out = h;
  return 0;
}

// Sources (sim/sph.go:71-86): m particles appended with the next IDs.  The
// handle is rebuilt around the grown state (slot order kept, new particles
// after it) and swapped into the caller's handle.
extern "C" int sph_append(sph_sim* s, int64_t m, const double* x, const double* y,
                          const double* vx, const double* vy, const double* e) {
  if (!s) return fail(SPH_EINVAL, "null handle");
  if (m < 0 || (m > 0 && (!x || !y))) return fail(SPH_EINVAL, "append: bad arguments");
  if (m == 0) return 0;
  if (s->hsims.size() != 1)
    return fail(SPH_EINVAL, "particles are appended to one simulation, not a batch");
  const int64_t n = s->np, n2 = n + m;
  if (n2 >= (1LL << 28)) return fail(SPH_EINVAL, "too many particles in one simulation");
  for (int64_t i = 0; i < m; ++i)
    if (!std::isfinite(x[i]) || !std::isfinite(y[i]))
      return fail(SPH_ERANGE, "non-finite position of appended particle %lld", (long long)i);
  int rc = fetch_sims(s);
  if (rc) return rc;
  std::vector<double> a((size_t)kCkptFields * n2, 0.0);
  std::vector<uint32_t> ids(n2);
  double* src[kCkptFields];
  ckpt_arrays(s, src);
  for (int f = 0; f < kCkptFields; ++f)
    HIPCHK(hipMemcpyAsync(a.data() + f * n2, src[f], n * sizeof(double), hipMemcpyDeviceToHost,
                          s->st));
  HIPCHK(hipMemcpyAsync(ids.data(), s->F[s->cur].id, n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                        s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  for (int64_t i = 0; i < m; ++i) {
    a[0 * n2 + n + i] = x[i];
    a[1 * n2 + n + i] = y[i];
    a[2 * n2 + n + i] = vx ? vx[i] : 0.0;
    a[3 * n2 + n + i] = vy ? vy[i] : 0.0;
    a[4 * n2 + n + i] = e ? e[i] : 0.0;
    ids[n + i] = (uint32_t)(n + i);
  }
  const sph_params prm = s->hparams[0];
  const sph_params* pp = &prm;
  const double* xs[5];
  for (int i = 0; i < 5; ++i) xs[i] = a.data() + i * n2;
  const uint32_t* idp = ids.data();
  sph_sim* h = nullptr;
  if ((rc = create_batch(1, &pp, &n2, &xs[0], &xs[1], &xs[2], &xs[3], &xs[4], &idp, &h)))
    return rc;
  auto bail = [&](int code) {
    sph_destroy(h);
    return code;
  };
  layout_cells(h, s->grid_cpp);  // same cells per particle and walk shape as before
  h->pb.walk = s->pb.walk;
  // device-updated scalars carry over (the next half step was set by the
  // last pass, over the particles that had neighbours)
  take_device_scalars(h->hsims[0], s->hsims[0]);
  double* dst[kCkptFields];
  ckpt_arrays(h, dst);
  for (int f = 5; f < kCkptFields; ++f)
    if (hipMemcpyAsync(dst[f], a.data() + f * n2, n2 * sizeof(double), hipMemcpyHostToDevice,
                       h->st) != hipSuccess)
      return bail(fail(SPH_EDEVICE, "upload failed"));
  if (hipMemcpyAsync(h->pb.sims, h->hsims.data(), sizeof(SimDev), hipMemcpyHostToDevice, h->st) !=
          hipSuccess ||
      hipStreamSynchronize(h->st) != hipSuccess)
    return bail(fail(SPH_EDEVICE, "upload failed"));
  if (s->timing && (rc = sph_set_timing(h, 1))) return bail(rc);
  h->step = s->step;
  h->pb.have_h = s->pb.have_h;
  h->graph_on = s->graph_on;
  for (int t = 0; t < SPH_NTIMERS; ++t) {
    h->t_ms[t] = s->t_ms[t];
    h->t_cnt[t] = s->t_cnt[t];
  }
  h->bind();
  s->swap_state(*h);
  sph_destroy(h);  // the old buffers
  return 0;
}

extern "C" const char* sph_last_error(void) { return g_err.c_str(); }
extern "C" const char* sph_version(void) { return "sphugo_amd 0.1 (gfx950)"; }
