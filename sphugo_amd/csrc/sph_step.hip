// SPH force pass and integrator kernels for gfx950 (wave64).
//
//   k_keys          cell key of each particle's pass position: Hilbert index
//                   of a power-of-two grid plus the simulation's cell offset
//                   (replaces Treebuild/Partition, sim/core.go:126-224)
//   sort_pairs      stable LSD radix sort of (key, index) (sph_sort.hip)
//   k_permute       gather the state into cell order, apply the drift-1 /
//                   predictor step (sim/sph.go:108-117) or the step-0
//                   initialisation (:96-101), write the cell ranges and
//                   (FP32C) the fixed-point positions
//   k_fill_runs     the cell ranges of sparse grid stretches k_permute queued
//   k_knn           tile walk: a wavefront takes 64 consecutive queries,
//                   stages the candidates of their cells through LDS and
//                   sorts them against each query's bracket around its last
//                   h: certain neighbours go to the list, the annulus to a
//                   short side list; misses go to a retry list
//   k_knn_finish    thread per query: exact selection of the annulus,
//                   density and sound speed, the force record
//                   (sim/nearest-neighbour.go:28-165, sim/sph.go:306-323,
//                   421-429)
//   k_knn_fallback  exact wavefront-per-query search for the retry list (or
//                   for every query: first pass, bitwise order)
//   k_force         thread per particle: pressure + artificial viscosity +
//                   EDot over the packed records of its neighbours, fused
//                   with kick, drift 2, periodic wrap and reflection
//                   (sim/sph.go:327-401,122-193), partial sums and the CFL
//                   minimum
//   k_reduce_part,  per-simulation FP64 totals in a fixed order, the next
//   k_reduce        CFL half step and the next kNN walk bracket
//
// All FP64 arithmetic on the parity path is evaluated in the reference's
// association order; the library is compiled with -ffp-contract=off so no
// product is fused into an add there.  Only the FP32C pair terms enable
// contraction, block by block.
#include <type_traits>

#include "sph_device.hpp"
#include "sph_host.hpp"

namespace sph {

// ------------------------------------------------------------------------
// binning
struct CellIdx {
  int64_t ux, uy;  // unwrapped cell
  bool ok;
};

__device__ inline CellIdx cell_of(const SimDev* S, double x, double y) {
  double fx = floor((x - S->x0) * S->ivx);
  double fy = floor((y - S->y0) * S->ivy);
  const double lox = -2.0 * S->gx, hix = 3.0 * S->gx - 1;
  const double loy = -2.0 * S->gy, hiy = 3.0 * S->gy - 1;
  // an open axis has no images: its box (k_box) holds every pass position
  // inside cells [0, G) with a wide margin; the clamp only pins the index
  if (S->open & 1) fx = fx < 0 ? 0.0 : (fx > S->gx - 1 ? (double)(S->gx - 1) : fx);
  if (S->open & 2) fy = fy < 0 ? 0.0 : (fy > S->gy - 1 ? (double)(S->gy - 1) : fy);
  CellIdx c;
  c.ok = (fx >= lox) && (fx <= hix) && (fy >= loy) && (fy <= hiy);
  c.ux = c.ok ? (int64_t)fx : 0;
  c.uy = c.ok ? (int64_t)fy : 0;
  return c;
}

// Cell span [a, b] of a search at a level with nc cells per axis, limited to
// the images the reference visits: [-2 nc, 3 nc) on a periodic axis (the
// -L, 0, +L images of positions that may sit up to a box outside it), the
// box itself on an open axis (sim/nearest-neighbour.go:39-55).
__device__ inline void clamp_span(int64_t& a, int64_t& b, int64_t nc, bool open) {
  const int64_t lo = open ? 0 : -2 * nc, hi = open ? nc - 1 : 3 * nc - 1;
  if (a < lo) a = lo;
  if (b > hi) b = hi;
  if (b < a) b = a - 1;  // empty span
}

// Box of each open simulation's pass positions, in two launches: per work
// item the range of the positions k_keys bins (same predictor arithmetic),
// then per simulation the range over its items and the grid box (open_box).
// Periodic simulations and unused pool slots are skipped.
__global__ __launch_bounds__(kItem) void k_bbox(const Item* __restrict__ items, int nitems,
                                                const SimDev* __restrict__ sims,
                                                const double* __restrict__ X,
                                                const double* __restrict__ Y,
                                                const double* __restrict__ VX,
                                                const double* __restrict__ VY, int mode,
                                                double* __restrict__ part) {
  __shared__ double red[4][4];
  const int item_idx = blockIdx.x;
  const Item it = items[item_idx];
  if (it.sim < 0) return;
  const SimDev* S = sims + it.sim;
  if (!S->open) return;
  if (mode < 0) mode = S->pmode;
  double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
  if ((int)threadIdx.x < it.count) {
    const int64_t t = it.first + threadIdx.x;
    double x = X[t], y = Y[t];
    if (mode == kPermPredict) {
      const double dt = S->dt_half;
      x = x + VX[t] * dt;
      y = y + VY[t] * dt;
    }
    x0 = x1 = x;
    y0 = y1 = y;
  }
  for (int o = 32; o > 0; o >>= 1) {
    x0 = fmin(x0, __shfl_xor(x0, o, 64));
    x1 = fmax(x1, __shfl_xor(x1, o, 64));
    y0 = fmin(y0, __shfl_xor(y0, o, 64));
    y1 = fmax(y1, __shfl_xor(y1, o, 64));
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[wv][0] = x0;
    red[wv][1] = x1;
    red[wv][2] = y0;
    red[wv][3] = y1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      x0 = fmin(x0, red[w][0]);
      x1 = fmax(x1, red[w][1]);
      y0 = fmin(y0, red[w][2]);
      y1 = fmax(y1, red[w][3]);
    }
    double* p = part + (int64_t)item_idx * kNumRed;
    p[0] = x0;
    p[1] = x1;
    p[2] = y0;
    p[3] = y1;
  }
}

__global__ __launch_bounds__(256) void k_box(SimDev* __restrict__ sims,
                                             const double* __restrict__ part,
                                             const int* __restrict__ item_first,
                                             const int* __restrict__ item_count) {
  __shared__ double red[4][4];
  SimDev* S = sims + blockIdx.x;
  if (S->n == 0 || !S->open) return;
  const int i0 = item_first[blockIdx.x], ni = item_count[blockIdx.x];
  double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
  for (int i = threadIdx.x; i < ni; i += 256) {
    const double* p = part + (int64_t)(i0 + i) * kNumRed;
    x0 = fmin(x0, p[0]);
    x1 = fmax(x1, p[1]);
    y0 = fmin(y0, p[2]);
    y1 = fmax(y1, p[3]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    x0 = fmin(x0, __shfl_xor(x0, o, 64));
    x1 = fmax(x1, __shfl_xor(x1, o, 64));
    y0 = fmin(y0, __shfl_xor(y0, o, 64));
    y1 = fmax(y1, __shfl_xor(y1, o, 64));
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[wv][0] = x0;
    red[wv][1] = x1;
    red[wv][2] = y0;
    red[wv][3] = y1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      x0 = fmin(x0, red[w][0]);
      x1 = fmax(x1, red[w][1]);
      y0 = fmin(y0, red[w][2]);
      y1 = fmax(y1, red[w][3]);
    }
    SimDev d = *S;
    open_box(d, x0, x1, y0, y1);
    S->x0 = d.x0;
    S->lx = d.lx;
    S->ivx = d.ivx;
    S->sx = d.sx;
    S->y0 = d.y0;
    S->ly = d.ly;
    S->ivy = d.ivy;
    S->sy = d.sy;
  }
}

__global__ __launch_bounds__(kItem) void k_keys(const Item* __restrict__ items, int nitems,
                                                const SimDev* __restrict__ sims,
                                                const double* __restrict__ X,
                                                const double* __restrict__ Y,
                                                const double* __restrict__ VX,
                                                const double* __restrict__ VY, int mode,
                                                uint32_t* __restrict__ keys, int* err,
                                                uint32_t* __restrict__ retry_count,
                                                uint32_t* __restrict__ run_count,
                                                uint32_t hole_cpp) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *retry_count = 0u;
    *run_count = 0u;
  }
  const Item it = items[xcd_item(nitems)];
  if ((int)threadIdx.x >= it.count) return;
  const int64_t t = it.first + threadIdx.x;
  if (it.sim < 0) {
    // unused pool slot: a key of its own position keeps it in place
    keys[t] = (uint32_t)(t * hole_cpp);
    return;
  }
  const SimDev* S = sims + it.sim;
  if (mode < 0) mode = S->pmode;
  double x = X[t], y = Y[t];
  if (mode == kPermPredict) {
    const double dt = S->dt_half;
    x = x + VX[t] * dt;
    y = y + VY[t] * dt;
  }
  CellIdx c = cell_of(S, x, y);
  if (!c.ok) {
    atomicOr(err, 1);
    atomicOr(err + 1 + it.sim, 1);
  }
  const int64_t cx = c.ux - (floordiv_pow2(c.ux, S->lg) << S->lg);
  const int64_t cy = c.uy - (floordiv_pow2(c.uy, S->lg) << S->lg);
  keys[t] = (uint32_t)(S->coff + hilbert2((uint32_t)cx, (uint32_t)cy, S->lg));
}

// Queues the cells [lo, hi] of a permute wave for k_fill_runs, in chunks of
// at most kRunChunk cells (wave-uniform call): cell q of a chunk starts at
// t0 plus the number of the wave's nv keys (skeys[t0 .. t0 + nv)) below q.
// Returns false, leaving the cells to the caller, when the queue is full
// (its reserved slots below the queue's end then hold empty chunks).
__device__ inline bool queue_cells(uint32_t* __restrict__ runs, uint32_t* __restrict__ run_count,
                                   int64_t lo, int64_t hi, int64_t t0, int nv, int lane) {
  const uint32_t nch = (uint32_t)((hi - lo + kRunChunk) / kRunChunk);
  uint32_t at = 0;
  if (lane == 0) at = atomicAdd(run_count, nch);
  at = __shfl(at, 0, 64);
  const bool fits = at + nch <= (uint32_t)kRunQueue;
  for (uint32_t c = (uint32_t)lane; c < nch && at + c < (uint32_t)kRunQueue; c += 64) {
    const int64_t f = lo + (int64_t)c * kRunChunk;
    const int64_t l = f + kRunChunk - 1 < hi ? f + kRunChunk - 1 : hi;
    uint32_t* r = runs + 4 * (int64_t)(at + c);
    r[0] = fits ? (uint32_t)f : 1u;
    r[1] = fits ? (uint32_t)l : 0u;
    r[2] = (uint32_t)t0;
    r[3] = (uint32_t)nv;
  }
  return fits;
}

// The queued chunks of cells of this pass, a block per chunk: the chunk's
// wave keys in LDS, each cell a binary search over them.
__global__ __launch_bounds__(256) void k_fill_runs(const uint32_t* __restrict__ runs,
                                                   const uint32_t* __restrict__ run_count,
                                                   const uint32_t* __restrict__ skeys,
                                                   uint32_t* __restrict__ cstart) {
  __shared__ uint32_t keys[64];
  const uint32_t c = *run_count;
  const uint32_t n = c < (uint32_t)kRunQueue ? c : (uint32_t)kRunQueue;
  for (uint32_t r = blockIdx.x; r < n; r += gridDim.x) {  // uniform per block
    const int64_t f = runs[4 * (int64_t)r], l = runs[4 * (int64_t)r + 1];
    const uint32_t t0 = runs[4 * (int64_t)r + 2];
    const int nv = (int)runs[4 * (int64_t)r + 3];
    __syncthreads();
    if ((int)threadIdx.x < nv) keys[threadIdx.x] = skeys[(int64_t)t0 + threadIdx.x];
    __syncthreads();
    for (int64_t q = f + threadIdx.x; q <= l; q += 256) {
      int a = 0, b = nv;  // the number of keys below q
      while (a < b) {
        const int mid = (a + b) >> 1;
        if ((int64_t)keys[mid] >= q) b = mid; else a = mid + 1;
      }
      cstart[q] = t0 + (uint32_t)a;
    }
  }
}

__global__ __launch_bounds__(kItem) void k_permute(
    const Item* __restrict__ items, int nitems, const SimDev* __restrict__ sims,
    const uint32_t* __restrict__ perm, const uint32_t* __restrict__ skeys, int mode,
    Positions pin, Fields fin, double2* __restrict__ xyp, uint2* __restrict__ uxy, Fields fout,
    uint8_t* __restrict__ wrap, uint32_t* __restrict__ cstart, uint32_t* __restrict__ runs,
    uint32_t* __restrict__ run_count) {
  const Item it = items[xcd_item(nitems)];
  if (it.sim < 0) return;  // unused pool slot: stays in place, owns no cells
  const SimDev* S = sims + it.sim;
  const bool valid = (int)threadIdx.x < it.count;
  const int64_t t = it.first + (valid ? threadIdx.x : 0);
  // Cell ranges of this simulation's cells [coff, coff + G^2]: cell q starts
  // at the first particle whose key is >= q; the simulation's last particle
  // closes the range.  The wave writes the contiguous block of cells
  // (key[t0 - 1], key[t_last]] of its particles together, coalesced, each
  // cell finding its particle by a binary search over the lanes' keys; a
  // block wider than kWaveCells (a sparse stretch of the grid) goes to
  // k_fill_runs instead, which spreads it over many blocks.  Cells outside
  // every simulation's range are never read.
  {
    const int lane = threadIdx.x & 63;
    const int w0 = threadIdx.x & ~63;
    if (w0 < it.count) {
      const int64_t t0 = it.first + w0;
      const int nv = (it.count - w0) < 64 ? (it.count - w0) : 64;
      const uint32_t kt = skeys[t0 + (lane < nv ? lane : nv - 1)];
      const int64_t lo = (t0 == S->off ? S->coff : (int64_t)skeys[t0 - 1] + 1);
      const int64_t hi = (int64_t)__shfl(kt, nv - 1, 64);
      if (hi - lo + 1 <= kWaveCells || !queue_cells(runs, run_count, lo, hi, t0, nv, lane)) {
        for (int64_t qb = lo; qb <= hi; qb += 64) {  // uniform: every lane shuffles
          const int64_t q = qb + lane;
          int a = 0, b = nv - 1;  // smallest lane with key >= q (lane nv-1 qualifies)
          for (int st = 0; st < 6; ++st) {
            const int mid = (a + b) >> 1;
            const int64_t km = (int64_t)__shfl(kt, mid, 64);
            if (km >= q) b = mid; else a = mid + 1;
          }
          if (q <= hi) cstart[q] = (uint32_t)(t0 + a);
        }
      }
      const int64_t tl = t0 + nv - 1;
      if (tl == S->off + S->n - 1) {
        const int64_t end = S->coff + (int64_t)S->gx * S->gy;
        if (end - hi <= kWaveCells || !queue_cells(runs, run_count, hi + 1, end, tl + 1, 0, lane))
          for (int64_t q = hi + 1 + lane; q <= end; q += 64) cstart[q] = (uint32_t)(tl + 1);
      }
    }
  }
  if (valid) {
    const uint32_t i = perm[t];
    if (mode < 0) mode = S->pmode;
    double x = pin.x[i], y = pin.y[i];
    const double vx = fin.vx[i], vy = fin.vy[i], e = fin.e[i];
    double vpx, vpy, ep;
    if (mode == kPermPredict) {
      const double dt = S->dt_half;
      x = x + vx * dt;
      y = y + vy * dt;
      const double adx = fin.ax[i] * dt, ady = fin.ay[i] * dt;
      vpx = vx + adx;
      vpy = vy + ady;
      ep = e + fin.edot[i] * dt;
    } else if (mode == kPermInit) {
      vpx = vx; vpy = vy; ep = e;
    } else {
      vpx = fin.vpx[i]; vpy = fin.vpy[i]; ep = fin.ep[i];
    }
    xyp[t] = make_double2(x, y);
    // FP32C: the fixed-point form of the force record (units of L / 2^32,
    // wrapped), whose 32-bit differences are minimum-image displacements
    if (S->prec == kPrecFP32C)
      uxy[t] = make_uint2((uint32_t)(int64_t)floor((x - S->x0) * (4294967296.0 / S->lx)),
                          (uint32_t)(int64_t)floor((y - S->y0) * (4294967296.0 / S->ly)));
    fout.vx[t] = vx;
    fout.vy[t] = vy;
    fout.e[t] = e;
    fout.vpx[t] = vpx;
    fout.vpy[t] = vpy;
    fout.ep[t] = ep;
    fout.h[t] = fin.h[i];
    fout.id[t] = fin.id[i];
    CellIdx c = cell_of(S, x, y);
    const int64_t wx = floordiv_pow2(c.ux, S->lg), wy = floordiv_pow2(c.uy, S->lg);
    wrap[t] = (uint8_t)((wx + 2) | ((wy + 2) << 3));
  }
}

// ------------------------------------------------------------------------
// kNN.
//
// Candidates are the particles of the cells at a coarser level of the
// grid: a coarse cell (X, Y) at level c is one contiguous range of the
// sorted arrays, [cstart[m << 2(lg-c)], cstart[(m+1) << 2(lg-c)]) with m
// its Hilbert index at that level.  Periodic images follow the reference:
// a candidate j is seen through the query offsets o in {-L, 0, L}^2 and its
// squared distance is ((p_i + o) - p_j)^2 (sim/nearest-neighbour.go:57-61,
// 72,76), self excluded at every image (:79).  The k nearest are ordered by
// (d2, ID, image) so ties resolve the same way everywhere.

__device__ inline int level_for(const SimDev* S, double w, double hgt, double ncells) {
  int c = S->lg;
  while (c > 0 && ((double)(1 << c) * w > ncells * S->lx || (double)(1 << c) * hgt > ncells * S->ly))
    --c;
  return c;
}

__device__ inline void cell_range(const SimDev* S, const uint32_t* __restrict__ cstart, int c,
                                  uint32_t xw, uint32_t yw, uint32_t& a, uint32_t& b) {
  const int sh = 2 * (S->lg - c);
  const uint64_t m = (uint64_t)hilbert2(xw, yw, c) << sh;
  a = cstart[S->coff + m];
  b = cstart[S->coff + m + (1ull << sh)];
}

constexpr int kV1Cells = 128;  // coarse cells of one search (<= 9 x 9 used)
constexpr int kTileCells = 256;  // coarse cells of one tile box (<= 15 x 15 used)

// Per-wave scratch of the exact wavefront search.
struct SearchShared {
  double d2[kCandCap];
  uint32_t pk[kCandCap];
  uint32_t pst[kV1Cells];
  int ppref[kV1Cells + 1];
  int8_t ptx[kV1Cells];
  int8_t pty[kV1Cells];
  double sd2[kMaxK];
  uint32_t spk[kMaxK];
  double fterm[kMaxK];
};

// Per-wave scratch of the tile search (lane = query), KSUB sub-tiles.
template <int KSUB>
struct TileShared {
  uint32_t pst[kTileCells];
  int ppref[kTileCells + 1];
  int8_t ptx[kTileCells];
  int8_t pty[kTileCells];
  double2 sxy[65];           // staged candidates; entry 64 is a NaN sentinel
  uint32_t spk[65];
  double sbb[KSUB][4];       // box of each sub-tile (x0, x1, y0, y1)
  uint8_t slist[KSUB][64];   // staged candidates within reach of each sub-tile
  uint32_t slpk[KSUB][64];   // their packed entries, in the same order
  unsigned long long smask[KSUB];  // staged slots in each sub-tile's list
  unsigned long long selfm[64];    // per query: staged slots that are itself
};

struct SearchOut {
  double tau;   // k-th smallest squared distance
  double facc;  // sum of kernel values F(d/h) over the k neighbours
};

// Candidate of a piece: returns false if the reference does not visit
// this (particle, image) pair.
__device__ inline bool cand_image(int ptx, int pty, uint8_t wc, int& ix, int& iy) {
  const int nx = ptx - ((int)(wc & 7) - 2);
  const int ny = pty - ((int)(wc >> 3) - 2);
  ix = -nx;
  iy = -ny;
  return nx >= -1 && nx <= 1 && ny >= -1 && ny <= 1;
}

// Neighbour lists are stored in tiles of 64 consecutive queries: entry e of
// query t lives at tile(t) * kstride * 64 + (e / 4) * 256 + (t % 64) * 4 +
// e % 4 (u32 units), so 64 lanes reading the same 4-entry chunk of their
// rows touch one contiguous 1 KiB segment.  kstride is a multiple of 4.
struct NbrRow {
  uint32_t* base;
  __device__ uint32_t& operator[](int e) const { return base[(e >> 2) * 256 + (e & 3)]; }
  __device__ u32x4* chunk(int c) const { return (u32x4*)(base + c * 256); }
};
__device__ inline NbrRow nbr_row(uint32_t* nbr, int64_t t, int kstride) {
  return NbrRow{nbr + (t >> 6) * (int64_t)kstride * 64 + (t & 63) * 4};
}
__device__ inline NbrRow nbr_row(const uint32_t* nbr, int64_t t, int kstride) {
  return nbr_row(const_cast<uint32_t*>(nbr), t, kstride);
}

// Exact kNN of one query by the whole wavefront, starting from radius R.
// Writes the k entries of `row` (BITWISE: sorted descending by (d2, ID,
// image), the reference's queue order) and returns tau and the density
// kernel sum (BITWISE: summed farthest first, sim/sph.go:313-320).
template <bool BITWISE, bool FP32C>
__device__ SearchOut wave_search(const SimDev* S, int64_t t, double xq, double yq, double R,
                                 const double2* __restrict__ XY,
                                 const uint8_t* __restrict__ WR,
                                 const uint32_t* __restrict__ cstart,
                                 const uint32_t* __restrict__ ID, NbrRow row,
                                 SearchShared& sh, unsigned long long& nretry,
                                 unsigned long long& nties, unsigned long long& nfail) {
  const int lane = lane_id();
  const int k = S->k;
  const double lx = S->lx, ly = S->ly;
  // every image of every particle lies within 2 sqrt(2) max(lx, ly) of the query
  const double rcap = 3.0 * (lx > ly ? lx : ly);
  if (R > rcap) R = rcap;
  double r_lo = 0.0, r_hi = 0.0;
  int cnt = 0;
  double R2 = 0;
  int attempt = 0;
  for (;; ++attempt) {
    R2 = R * R;
    const double Rm = R * (1 + 1e-12) + 1e-15 * (lx + ly);
    const int c = level_for(S, 2 * Rm, 2 * Rm, 8.0);
    const int64_t nc = 1LL << c;
    const double icx = (double)nc / lx, icy = (double)nc / ly;
    const double scx = lx / (double)nc, scy = ly / (double)nc;
    int64_t X0 = (int64_t)floor((xq - S->x0 - Rm) * icx), X1 = (int64_t)floor((xq - S->x0 + Rm) * icx);
    int64_t Y0 = (int64_t)floor((yq - S->y0 - Rm) * icy), Y1 = (int64_t)floor((yq - S->y0 + Rm) * icy);
    clamp_span(X0, X1, nc, S->open & 1);
    clamp_span(Y0, Y1, nc, S->open & 2);
    const int nX = (int)(X1 - X0 + 1), nY = (int)(Y1 - Y0 + 1);
    const int ncell = nX * nY;
    int npieces = 0, ncand = 0;
    for (int cb = 0; cb < ncell; cb += 64) {
      const int ci = cb + lane;
      bool have = false;
      uint32_t a = 0, b = 0;
      int tx = 0, ty = 0;
      if (ci < ncell) {
        const int64_t Xc = X0 + ci % nX, Yc = Y0 + ci / nX;
        const double bx0 = S->x0 + (double)Xc * scx, by0 = S->y0 + (double)Yc * scy;
        const double dx = xq < bx0 ? bx0 - xq : (xq > bx0 + scx ? xq - (bx0 + scx) : 0.0);
        const double dy = yq < by0 ? by0 - yq : (yq > by0 + scy ? yq - (by0 + scy) : 0.0);
        if (dx * dx + dy * dy <= Rm * Rm) {
          tx = (int)floordiv_pow2(Xc, c);
          ty = (int)floordiv_pow2(Yc, c);
          cell_range(S, cstart, c, (uint32_t)(Xc - tx * nc), (uint32_t)(Yc - ty * nc), a, b);
          have = b > a;
        }
      }
      const unsigned long long m = ballot(have);
      const int len = have ? (int)(b - a) : 0;
      const int incl = wave_incl_scan(len);
      if (have) {
        const int pos = npieces + lane_prefix(m);
        sh.pst[pos] = a;
        sh.ppref[pos] = ncand + incl - len;
        sh.ptx[pos] = (int8_t)tx;
        sh.pty[pos] = (int8_t)ty;
      }
      npieces += __popcll(m);
      ncand += __shfl(incl, 63, 64);
    }
    if (lane == 0) sh.ppref[npieces] = ncand;
    __builtin_amdgcn_wave_barrier();
    int nsteps = 0;
    while ((1 << nsteps) < npieces) ++nsteps;
    cnt = 0;
    // Stage s takes candidate lane * nst + s: each lane walks a run of
    // consecutive candidates, so its piece is found by a binary search once
    // and then moves at most one step per stage (every piece is non-empty).
    const int nst = (ncand + 63) >> 6;
    int lo = 0;
    {
      const int u0 = lane * nst;
      if (u0 < ncand) {
        int hi = npieces - 1;
        for (int s = 0; s < nsteps; ++s) {
          const int mid = (lo + hi + 1) >> 1;
          if (sh.ppref[mid] <= u0) lo = mid; else hi = mid - 1;
        }
      }
    }
    for (int stage = 0; stage < nst; ++stage) {
      const int u = lane * nst + stage;
      bool acc = false;
      double d2 = 0;
      uint32_t pk = 0;
      if (u < ncand) {
        if (sh.ppref[lo + 1] <= u) ++lo;
        const uint32_t j = sh.pst[lo] + (uint32_t)(u - sh.ppref[lo]);
        // the position is loaded beside the wrap code, not after it
        int ix, iy;
        const bool vis = cand_image(sh.ptx[lo], sh.pty[lo], WR[j], ix, iy);
        const bool seen = vis && (int64_t)j != t;
        const double2 pj = XY[j];
        const double ex = (xq + (double)ix * lx) - pj.x;
        const double ey = (yq + (double)iy * ly) - pj.y;
        d2 = ex * ex + ey * ey;
        acc = seen && d2 <= R2;
        pk = (j << 4) | img_code(seen ? ix : 0, seen ? iy : 0);
      }
      const unsigned long long m = ballot(acc);
      if (acc) {
        const int pos = cnt + lane_prefix(m);
        if (pos < kCandCap) {
          sh.d2[pos] = d2;
          sh.pk[pos] = pk;
        }
      }
      cnt += __popcll(m);
    }
    __builtin_amdgcn_wave_barrier();
    if (cnt >= k && cnt <= kCandCap) break;
    if (cnt < k) {
      r_lo = R;
      double f = cnt > 0 ? sqrt((1.5 * k) / cnt) : 2.0;
      if (f < 1.25) f = 1.25;
      if (f > 4.0) f = 4.0;
      R = R * f;
      if (r_hi > 0 && R >= r_hi) R = sqrt(r_lo * r_hi);
      if (R > rcap) R = rcap;
    } else {
      r_hi = R;
      double f = sqrt((0.5 * (k + kCandCap)) / cnt);
      if (f > 0.9) f = 0.9;
      R = R * f;
      if (R <= r_lo) R = sqrt(r_lo * r_hi);
    }
    if (attempt > 60 || (cnt < k && r_lo >= rcap)) {
      if (lane == 0) ++nfail;
      break;
    }
  }
  if (attempt > 0 && lane == 0) ++nretry;
  SearchOut out;
  out.tau = 0;
  out.facc = 0;
  if (cnt < k || cnt > kCandCap) return out;
  // k-th smallest squared distance: ballot quickselect
  constexpr int NS = kCandCap / 64;
  double v[NS];
  uint32_t pkr[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int idx = s * 64 + lane;
    v[s] = idx < cnt ? sh.d2[idx] : INFINITY;
    pkr[s] = idx < cnt ? sh.pk[idx] : 0u;
  }
  double lo = -1.0, hi = INFINITY, tau = 0;
  int c_lt = 0, c_le = 0;
  for (;;) {
    int slot = -1, src = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const unsigned long long b = ballot(v[s] > lo && v[s] < hi);
      if (slot < 0 && b) {
        slot = s;
        src = __builtin_ctzll(b);
      }
    }
    double pv = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (s == slot) pv = readlane_d(v[s], src);
    c_lt = 0;
    c_le = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      c_lt += __popcll(ballot(v[s] < pv));
      c_le += __popcll(ballot(v[s] <= pv));
    }
    if (c_lt < k && k <= c_le) {
      tau = pv;
      break;
    }
    if (k <= c_lt) hi = pv; else lo = pv;
  }
  // ties at tau keep the smallest (ID, image)
  bool tsel[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) tsel[s] = false;
  if (c_le > k) {
    if (lane == 0) ++nties;
    uint64_t key2[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int idx = s * 64 + lane;
      key2[s] = ~0ull;
      if (idx < cnt && v[s] == tau)
        key2[s] = ((uint64_t)ID[pkr[s] >> 4] << 4) | img_order(pkr[s] & 15u);
    }
    for (int r = 0; r < k - c_lt; ++r) {
      uint64_t mn = ~0ull;
#pragma unroll
      for (int s = 0; s < NS; ++s) mn = key2[s] < mn ? key2[s] : mn;
      for (int o = 32; o > 0; o >>= 1) {
        const uint64_t other = __shfl_xor(mn, o, 64);
        mn = other < mn ? other : mn;
      }
#pragma unroll
      for (int s = 0; s < NS; ++s)
        if (key2[s] == mn) {
          tsel[s] = true;
          key2[s] = ~0ull;
        }
    }
  }
  const double h = sqrt(tau);
  const int kern = S->kernel;
  double facc = 0.0;
  float fs = 0.f, fc = 0.f;
  // FP32C: q^2 = d^2 / tau formed in FP64 before the FP32 square root.  A
  // rounded 1 / h shared by every term would move all q together, and
  // F(q) = (1 - q)^3 / 3 amplifies that by 3q / (1 - q): a particle whose
  // neighbours all sit near q = 1 (the edge of a dense clump) would lose
  // 1e-5 of its density.
  const double itau = 1.0 / tau;
  int emitted = 0;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int idx = s * 64 + lane;
    const bool sel = idx < cnt && (v[s] < tau || (v[s] == tau && (c_le == k || tsel[s])));
    const unsigned long long m = ballot(sel);
    if (sel) {
      const int pos = emitted + lane_prefix(m);
      if constexpr (!BITWISE) {
        row[pos] = pkr[s];
        if constexpr (FP32C) {
          // hardware square root (within 1 ulp), as in the walk's finish
          const float q = __builtin_amdgcn_sqrtf((float)(v[s] * itau));
          const float f = kern_Ff(kern, q > 1.f ? 1.f : q);
          const float y = f - fc;
          const float tt = fs + y;
          fc = (tt - fs) - y;
          fs = tt;
        } else {
          facc += kern_F(kern, sqrt(v[s]) / h);
        }
      } else {
        sh.sd2[pos] = v[s];
        sh.spk[pos] = pkr[s];
      }
    }
    emitted += __popcll(m);
  }
  if constexpr (!BITWISE) {
    if constexpr (FP32C) {
      // compensated tree sum: each level carries its rounding error
      float s_ = fs, c_ = -fc;
      for (int o = 32; o > 0; o >>= 1) {
        const float so = __shfl_xor(s_, o, 64);
        const float co = __shfl_xor(c_, o, 64);
        const float sum = s_ + so;
        const float bp = sum - s_;
        const float err = (s_ - (sum - bp)) + (so - bp);
        s_ = sum;
        c_ = c_ + co + err;
      }
      out.facc = (double)s_ + (double)c_;
    } else {
      out.facc = group_sum(facc, 64);
    }
  } else {
    __builtin_amdgcn_wave_barrier();
    double dv = lane < k ? sh.sd2[lane] : -1.0;
    uint32_t pv = lane < k ? sh.spk[lane] : 0u;
    uint32_t iv = lane < k ? ID[pv >> 4] : 0u;
    // bitonic sort, descending by (d2, ID, image)
    for (int size = 2; size <= 64; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        const double od = __shfl_xor(dv, stride, 64);
        const uint32_t op = __shfl_xor(pv, stride, 64);
        const uint32_t oi = __shfl_xor(iv, stride, 64);
        const bool other_greater =
            od > dv || (od == dv && (oi > iv || (oi == iv && img_order(op & 15u) > img_order(pv & 15u))));
        const bool lower = (lane & stride) == 0;
        const bool desc = (lane & size) == 0;
        const bool take = (lower == desc) ? other_greater : !other_greater;
        if (take && (od != dv || oi != iv || op != pv)) {
          dv = od;
          pv = op;
          iv = oi;
        }
      }
    }
    if (lane < k) {
      row[lane] = pv;
      sh.fterm[lane] = kern_F(kern, sqrt(dv) / h);
    }
    __builtin_amdgcn_wave_barrier();
    double acc = 0.0;
    for (int n = 0; n < k; ++n) acc += sh.fterm[n];
    out.facc = acc;
  }
  out.tau = tau;
  __builtin_amdgcn_wave_barrier();
  return out;
}

// Density, sound speed, h and the force record of query t.
template <bool FP32C>
__device__ inline void write_result(const SimDev* S, int64_t t, double xq, double yq, double tau,
                                    double facc, const double* __restrict__ EP,
                                    const double* __restrict__ VPX,
                                    const double* __restrict__ VPY, double* __restrict__ H,
                                    double* __restrict__ RHO, double* __restrict__ C,
                                    void* __restrict__ recp) {
  const double h = sqrt(tau);
  const double rho = S->fpref * S->mass * facc / (h * h);
  const double c = sqrt(S->cfac * EP[t]);
  H[t] = h;
  RHO[t] = rho;
  C[t] = c;
  const double B = c * c / (S->gamma * rho);
  if constexpr (FP32C) {
    RecF r;
    r.ux = (uint32_t)(int64_t)floor((xq - S->x0) * (4294967296.0 / S->lx));
    r.uy = (uint32_t)(int64_t)floor((yq - S->y0) * (4294967296.0 / S->ly));
    r.vpx = (float)VPX[t];
    r.vpy = (float)VPY[t];
    r.rho = (float)(0.5 * rho);
    r.c = (float)(0.5 * c);
    r.h = (float)(0.5 * h);
    r.B = (float)B;
    ((RecF*)recp)[t] = r;
  } else {
    RecD r;
    r.x = xq;
    r.y = yq;
    r.vpx = VPX[t];
    r.vpy = VPY[t];
    r.rho = 0.5 * rho;
    r.c = 0.5 * c;
    r.h = 0.5 * h;
    r.B = B;
    ((RecD*)recp)[t] = r;
  }
}

// FP32C: may the displacements of a query with support h come from the 32-bit
// fixed-point positions (units of L / 2^32)?  Their 32-bit difference is the
// minimum image while every neighbour lies within half a box (h < 0.45 L).
// The displacement is centred on the query's exact position (centred_step):
// what is left is the neighbour's flooring, uniform within half a step and
// independent between neighbours, below 2^-20 h per component while
// h > 2^-13 L.  Emulated on clusters reaching that threshold the density
// stays within 2e-6 of FP64.  Otherwise the FP64 positions give the
// displacement.
__device__ inline bool fixed_point_ok(double h, double lx, double ly) {
  const double lmin = lx < ly ? lx : ly;
  return h < 0.45 * lmin && h > 0x1p-13 * lmin;
}

// Offset, in fixed-point steps, from a particle's exact position to the
// centre of its fixed-point step (1/2 minus the fraction the floor dropped):
// the displacement to a neighbour j is (u_j - u_i) + this, whose error is
// j's flooring alone, centred on zero.
__device__ inline float centred_step(double x, double x0, double l) {
  const double f = (x - x0) * (4294967296.0 / l);
  return (float)(0.5 - (f - floor(f)));
}

// First-pass search radius: local density from the occupancy of the
// query's cell at a level holding a few particles.
__device__ inline double local_radius(const SimDev* S, const uint32_t* __restrict__ cstart,
                                      double xq, double yq) {
  const CellIdx ci = cell_of(S, xq, yq);
  const int64_t fx = ci.ux - (floordiv_pow2(ci.ux, S->lg) << S->lg);
  const int64_t fy = ci.uy - (floordiv_pow2(ci.uy, S->lg) << S->lg);
  int c = S->lg - 1 > 0 ? S->lg - 1 : 0;
  for (;; --c) {
    const int sh = S->lg - c;
    uint32_t a, b;
    cell_range(S, cstart, c, (uint32_t)(fx >> sh), (uint32_t)(fy >> sh), a, b);
    const double area = (S->lx / (double)(1LL << c)) * (S->ly / (double)(1LL << c));
    if (b - a >= 8 || c == 0) {
      const double dens = (double)(b - a > 0 ? b - a : 1) / area;
      const double r = sqrt(1.3 * S->k / (M_PI * dens));
      return r < S->r_init * 4 ? r : S->r_init * 4;
    }
  }
}

// kNN + density + sound speed + force record.  A wave takes 64 consecutive
// (Hilbert-ordered) queries.  Fast path (lane = query): the wave stages the
// candidates of the tile's bounding box (grown by each query's bracket
// radius) through LDS 64 at a time and every lane tests all of them against
// its bracket [h_prev (1 - delta), h_prev (1 + delta)]: nearer candidates
// are certain neighbours and go straight to the list; the annulus keeps at
// most kAnn entries (written straight to global memory, which keeps the
// wave's LDS small) from which the remaining ones are selected exactly.
// Queries without a previous h, or whose bracket misses, fall back to the
// exact wavefront search.  BITWISE uses the wavefront search for all.
// Five waves per SIMD (a few spilled registers): the walk is issue-bound
// with long LDS and gather latencies; its LDS (about 7 KB per wave) allows
// five blocks per CU.
template <bool BITWISE, bool FP32C, int KSUB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_knn(
    const Item* __restrict__ items, int nitems, const SimDev* __restrict__ sims,
    const double2* __restrict__ XY,
    const uint8_t* __restrict__ WR, const uint32_t* __restrict__ cstart,
    double* __restrict__ H, const double* __restrict__ EP, const double* __restrict__ VPX,
    const double* __restrict__ VPY, const uint32_t* __restrict__ ID,
    uint32_t* __restrict__ nbr, int kstride, double* __restrict__ RHO,
    double* __restrict__ C, void* __restrict__ recp, KnnStats* __restrict__ stats,
    uint2* __restrict__ retry, uint32_t* __restrict__ retry_count,
    uint32_t* __restrict__ cnt, uint32_t* __restrict__ ann, WalkKnobs kn,
    uint32_t* __restrict__ walk_miss) {
  constexpr int kSub = KSUB;       // sub-tiles of a wave
  constexpr int kSubW = 64 / KSUB; // consecutive queries per sub-tile
  __shared__ TileShared<KSUB> wsh[4];
  const int item_idx = xcd_item(nitems);
  const Item it = items[item_idx];
  const SimDev* S = sims + it.sim;
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  TileShared<KSUB>& ts = wsh[wv];
  const int k = S->k;
  const double lx = S->lx, ly = S->ly;
  // per-lane outcomes, counted with ballots at the end (each lane once)
  bool miss_lo = false, miss_hi = false, miss_ann = false, split_fb = false;

  const int qi = wv * 64 + lane;
  const bool valid = qi < it.count;
  const int64_t t = it.first + (valid ? qi : 0);
  const double2 pq = XY[t];
  const double xq = pq.x, yq = pq.y;
  const double hp = H[t];
  const NbrRow row = nbr_row(nbr, t, kstride);
  bool fast = !BITWISE && valid && hp > 0;
  bool resolved = false;
  double tau = 0.0;
  if constexpr (!BITWISE) {
    const double delta = S->walk_delta;
    const double Rlo = hp * (1 - delta), Rhi = hp * (1 + delta);
    const double Rlo2 = Rlo * Rlo, Rhi2 = Rhi * Rhi;
    const double Rm = Rhi * (1 + 1e-12) + 1e-15 * (lx + ly);
    int c_lo = 0, c_an = 0;
    // certain neighbours collect in cbuf and leave in 16-byte stores
    u32x4 cbuf = {0u, 0u, 0u, 0u};
    u32x4* cptr = row.chunk(0);  // next chunk of the row to fill
    const double xqm = xq - lx, xqp = xq + lx, yqm = yq - ly, yqp = yq + ly;
    // annulus candidates go straight out, entry e of query t at
    // tile(t) * 64 kAnn + 64 e + t % 64, so the lanes' writes of one entry
    // index fill the same lines
    uint32_t* annrow = ann + (t >> 6) * (int64_t)(64 * kAnn) + (t & 63);
    // Lane groups: aligned blocks of consecutive queries.  A block whose
    // box holds more than kn.split candidates per query is halved (cell
    // order can put two distant regions in one wave); blocks below 8 lanes
    // go to the wavefront search, which handles them at lower cost.
    unsigned long long pending = ballot(fast);
    const int gsz0 = kn.group;
    int gsz = gsz0;
    while (pending) {
      const int l0 = __builtin_ctzll(pending);
      const int ga = l0 & ~(gsz - 1);
      const bool grp = fast && ((pending >> lane) & 1ull) && lane >= ga && lane < ga + gsz;
      const unsigned long long gm = ballot(grp);
      double bx0 = grp ? xq - Rm : INFINITY, bx1 = grp ? xq + Rm : -INFINITY;
      double by0 = grp ? yq - Rm : INFINITY, by1 = grp ? yq + Rm : -INFINITY;
      for (int o = 32; o > 0; o >>= 1) {
        bx0 = fmin(bx0, __shfl_xor(bx0, o, 64));
        by0 = fmin(by0, __shfl_xor(by0, o, 64));
        bx1 = fmax(bx1, __shfl_xor(bx1, o, 64));
        by1 = fmax(by1, __shfl_xor(by1, o, 64));
      }
      const int c = level_for(S, bx1 - bx0, by1 - by0, 14.0);
      const int64_t nc = 1LL << c;
      const double icx = (double)nc / lx, icy = (double)nc / ly;
      int64_t X0 = (int64_t)floor((bx0 - S->x0) * icx), X1 = (int64_t)floor((bx1 - S->x0) * icx);
      int64_t Y0 = (int64_t)floor((by0 - S->y0) * icy), Y1 = (int64_t)floor((by1 - S->y0) * icy);
      clamp_span(X0, X1, nc, S->open & 1);
      clamp_span(Y0, Y1, nc, S->open & 2);
      const int nX = (int)(X1 - X0 + 1), nY = (int)(Y1 - Y0 + 1);
      const int ncell = nX * nY;
      int npieces = 0, ncand = 0;
      for (int cb = 0; cb < ncell; cb += 64) {
        const int ci = cb + lane;
        bool have = false;
        uint32_t pa = 0, pb = 0;
        int tx = 0, ty = 0;
        if (ci < ncell) {
          const int64_t Xc = X0 + ci % nX, Yc = Y0 + ci / nX;
          tx = (int)floordiv_pow2(Xc, c);
          ty = (int)floordiv_pow2(Yc, c);
          cell_range(S, cstart, c, (uint32_t)(Xc - tx * nc), (uint32_t)(Yc - ty * nc), pa, pb);
          have = pb > pa;
        }
        const unsigned long long m = ballot(have);
        const int len = have ? (int)(pb - pa) : 0;
        const int incl = wave_incl_scan(len);
        if (have) {
          const int pos = npieces + lane_prefix(m);
          ts.pst[pos] = pa;
          ts.ppref[pos] = ncand + incl - len;
          ts.ptx[pos] = (int8_t)tx;
          ts.pty[pos] = (int8_t)ty;
        }
        npieces += __popcll(m);
        ncand += __shfl(incl, 63, 64);
      }
      npieces = __builtin_amdgcn_readfirstlane(npieces);
      ncand = __builtin_amdgcn_readfirstlane(ncand);
      const int gs = __popcll(gm);
      if (ncand > kn.split * gs) {
        if (gsz > 8) {
          gsz >>= 1;
        } else {
          if (grp) {
            fast = false;
            split_fb = true;
          }
          pending &= ~gm;
          gsz = gsz0;
        }
        __builtin_amdgcn_wave_barrier();
        continue;
      }
      if (lane == 0) ts.ppref[npieces] = ncand;
      {
        // boxes of the sub-tiles (kSubW consecutive queries each)
        double sx0 = grp ? xq - Rm : INFINITY, sx1 = grp ? xq + Rm : -INFINITY;
        double sy0 = grp ? yq - Rm : INFINITY, sy1 = grp ? yq + Rm : -INFINITY;
        for (int o = kSubW >> 1; o > 0; o >>= 1) {
          sx0 = fmin(sx0, __shfl_xor(sx0, o, 64));
          sy0 = fmin(sy0, __shfl_xor(sy0, o, 64));
          sx1 = fmax(sx1, __shfl_xor(sx1, o, 64));
          sy1 = fmax(sy1, __shfl_xor(sy1, o, 64));
        }
        if ((lane & (kSubW - 1)) == 0) {
          ts.sbb[lane / kSubW][0] = sx0;
          ts.sbb[lane / kSubW][1] = sx1;
          ts.sbb[lane / kSubW][2] = sy0;
          ts.sbb[lane / kSubW][3] = sy1;
        }
        if (lane == 0) {
          ts.sxy[64] = make_double2(NAN, NAN);
          ts.spk[64] = 0u;
        }
      }
      __builtin_amdgcn_wave_barrier();
      const int my_sub = lane / kSubW;
      double sbx0[kSub], sbx1[kSub], sby0[kSub], sby1[kSub];
#pragma unroll
      for (int g = 0; g < kSub; ++g) {
        sbx0[g] = readlane_d(ts.sbb[g][0], 0);
        sbx1[g] = readlane_d(ts.sbb[g][1], 0);
        sby0[g] = readlane_d(ts.sbb[g][2], 0);
        sby1[g] = readlane_d(ts.sbb[g][3], 0);
      }
      int nsteps = 0;
      while ((1 << nsteps) < npieces) ++nsteps;
      ts.selfm[lane] = 0ull;
      const int64_t wave_t0 = it.first + (int64_t)wv * 64;
      // Stage s takes every nst-th candidate (lane l: candidate l * nst + s),
      // so each stage samples the whole box evenly and the eight sub-tile
      // lists come out about equally long: every lane runs the longest one.
      const int nst = (ncand + 63) >> 6;
      // piece of this lane's candidate: a binary search for the first
      // stage, then (candidate u + 1 after u, every piece non-empty) at most
      // one step forward per stage
      int lo = 0;
      {
        const int u0 = lane * nst;
        if (u0 < ncand) {
          int hi = npieces - 1;
          for (int s = 0; s < nsteps; ++s) {
            const int mid = (lo + hi + 1) >> 1;
            if (ts.ppref[mid] <= u0) lo = mid; else hi = mid - 1;
          }
        }
      }
      for (int base = 0; base < ncand; base += 64) {
        // stage 64 candidates
        const int u = lane * nst + (base >> 6);
        double cxv = NAN, cyv = 0.0;
        uint32_t cpk = 0;
        bool interior = true;
        if (u < ncand) {
          if (ts.ppref[lo + 1] <= u) ++lo;
          const uint32_t j = ts.pst[lo] + (uint32_t)(u - ts.ppref[lo]);
          int ix, iy;
          if (cand_image(ts.ptx[lo], ts.pty[lo], WR[j], ix, iy)) {
            const double2 pj = XY[j];
            cxv = pj.x;
            cyv = pj.y;
            interior = (ix == 0) && (iy == 0);
          }
          cpk = (j << 4) | img_code(ix, iy);
          // an image of one of this wave's own queries: mark it in that
          // query's self mask (ordered after the query's own clear below)
          const int64_t so = (int64_t)j - wave_t0;
          if (so >= 0 && so < 64) atomicOr(&ts.selfm[so], 1ull << lane);
        }
        ts.sxy[lane] = make_double2(cxv, cyv);
        ts.spk[lane] = cpk;
        const bool all_interior = ballot(interior) == ~0ull;
        // per sub-tile, the staged candidates inside its box (in stage order)
        int nm = 0;
        {
          const double px = cxv - (double)img_ix(cpk & 15u) * lx;
          const double py = cyv - (double)img_iy(cpk & 15u) * ly;
#pragma unroll
          for (int g = 0; g < kSub; ++g) ts.slist[g][lane] = 64;
#pragma unroll
          for (int g = 0; g < kSub; ++g) {
            const bool in = px >= sbx0[g] && px <= sbx1[g] && py >= sby0[g] && py <= sby1[g];
            const unsigned long long m = ballot(in);
            if (in) {
              const int pos = lane_prefix(m);
              ts.slist[g][pos] = (uint8_t)lane;
              ts.slpk[g][pos] = cpk;
            }
            if (lane == 0) ts.smask[g] = m;
            const int c = __popcll(m);
            nm = c > nm ? c : nm;
          }
        }
        __builtin_amdgcn_wave_barrier();
        nm = (__builtin_amdgcn_readfirstlane(nm) + 3) & ~3;
        const uint8_t* mylist = ts.slist[my_sub];
        const uint32_t* mypk = ts.slpk[my_sub];
        // list positions of this query's own images (every image of the
        // query itself is excluded): a staged slot u in the sub-tile's list
        // sits at the count of list slots below u
        unsigned long long spos = 0ull;
        {
          const unsigned long long msub = ts.smask[my_sub];
          unsigned long long sm = ts.selfm[lane] & msub;
          ts.selfm[lane] = 0ull;
          while (sm) {
            const int us = __builtin_ctzll(sm);
            sm &= sm - 1ull;
            spos |= 1ull << __popcll(msub & ((1ull << us) - 1ull));
          }
        }
        // Blocks of 32 list positions: first every lane classifies its
        // candidates into bit masks, shifting one bit per position into
        // each (certain: inside Rlo, reach: up to Rhi), then emits the set
        // bits in list order.  Certain neighbours shift through a 4-entry
        // register buffer (oldest in x) that leaves as one 16-byte chunk
        // store every fourth entry.
        for (int mb = 0; mb < nm; mb += 32) {
          const int mend = nm < mb + 32 ? nm : mb + 32;
          uint32_t cmask = 0u, hmask = 0u;
// One position: compare into VCC, then add-with-carry mask + mask + VCC
// shifts the bit in (one instruction per mask where the compiler emits a
// select and a shift-or).  s_nop 1 gives the two wait states an FP64
// compare's VCC write needs before a VALU reads it.
#define SPH_CLASSIFY(D2)                                                                \
  {                                                                                     \
    const double d2v_ = (D2);                                                           \
    asm("v_cmp_lt_f64 vcc, %1, %2\n\ts_nop 1\n\tv_addc_co_u32 %0, vcc, %0, %0, vcc"         \
        : "+v"(cmask) : "v"(d2v_), "v"(Rlo2) : "vcc");                                  \
    asm("v_cmp_le_f64 vcc, %1, %2\n\ts_nop 1\n\tv_addc_co_u32 %0, vcc, %0, %0, vcc"         \
        : "+v"(hmask) : "v"(d2v_), "v"(Rhi2) : "vcc");                                  \
  }
          if (all_interior) {
            for (int m = mb; m < mend; m += 4) {
              const uint32_t i4 = *(const uint32_t*)(mylist + m);
              const uint32_t i0 = i4 & 255u, i1 = (i4 >> 8) & 255u, i2 = (i4 >> 16) & 255u,
                             i3 = i4 >> 24;
              const double2 e0 = ts.sxy[i0], e1 = ts.sxy[i1], e2 = ts.sxy[i2], e3 = ts.sxy[i3];
              const double a0 = xq - e0.x, b0 = yq - e0.y;
              const double a1 = xq - e1.x, b1 = yq - e1.y;
              const double a2 = xq - e2.x, b2 = yq - e2.y;
              const double a3 = xq - e3.x, b3 = yq - e3.y;
              SPH_CLASSIFY(a0 * a0 + b0 * b0);
              SPH_CLASSIFY(a1 * a1 + b1 * b1);
              SPH_CLASSIFY(a2 * a2 + b2 * b2);
              SPH_CLASSIFY(a3 * a3 + b3 * b3);
            }
          } else {
            for (int m = mb; m < mend; ++m) {
              const uint32_t im = mylist[m];
              const uint32_t pk = ts.spk[im];
              const int ix = img_ix(pk & 15u), iy = img_iy(pk & 15u);
              const double ax = ix == 0 ? xq : (ix < 0 ? xqm : xqp);
              const double ay = iy == 0 ? yq : (iy < 0 ? yqm : yqp);
              const double2 e = ts.sxy[im];
              const double ex = ax - e.x;
              const double ey = ay - e.y;
              SPH_CLASSIFY(ex * ex + ey * ey);
            }
          }
#undef SPH_CLASSIFY
          // position m of the block was shifted in at bit mend - 1 - m;
          // reversed and shifted down, it sits at bit m - mb
          const int sh = 32 - (mend - mb);
          const uint32_t keep = grp ? ~(uint32_t)(spos >> mb) : 0u;
          cmask = (__builtin_bitreverse32(cmask) >> sh) & keep;
          uint32_t amask = (__builtin_bitreverse32(hmask) >> sh) & keep & ~cmask;
          // the entries of the block's positions, read once per emitted bit
          const uint32_t* blpk = mypk + mb;
          while (cmask) {
            const uint32_t pk = blpk[__builtin_ctz(cmask)];
            cmask &= cmask - 1u;
            cbuf.x = cbuf.y;
            cbuf.y = cbuf.z;
            cbuf.z = cbuf.w;
            cbuf.w = pk;
            if ((c_lo & 3) == 3 && c_lo < k) {
              *cptr = cbuf;
              cptr += 64;
            }
            ++c_lo;
          }
          while (amask) {
            const uint32_t pk = blpk[__builtin_ctz(amask)];
            amask &= amask - 1u;
            if (c_an < kAnn) annrow[64 * c_an] = pk;
            ++c_an;
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
      pending &= ~gm;
      gsz = gsz0;
    }
    // hand the resolvable queries to k_knn_finish: certain neighbours are
    // in the row, annulus candidates go to `ann`
    if (fast) {
      miss_lo = c_lo >= k;
      miss_ann = !miss_lo && c_an > kAnn;
      miss_hi = !miss_lo && !miss_ann && c_lo + c_an < k;
    }
    if (fast && c_lo < k && c_an <= kAnn && c_lo + c_an >= k) {
      // the last c_lo % 4 certain entries sit at the top of the buffer
      const int rem = c_lo & 3;
      if (rem) {
        uint32_t* r0 = &row[c_lo & ~3];
        if (rem == 1) {
          r0[0] = cbuf.w;
        } else if (rem == 2) {
          r0[0] = cbuf.z;
          r0[1] = cbuf.w;
        } else {
          r0[0] = cbuf.y;
          r0[1] = cbuf.z;
          r0[2] = cbuf.w;
        }
      }
      resolved = true;
    }
    if (valid) cnt[t] = resolved ? ((uint32_t)c_lo | ((uint32_t)c_an << 8)) : 0xFFFFFFFFu;
  }
  // misses go to the retry list, searched by k_knn_fallback
  const bool need = valid && !resolved;
  const unsigned long long m = ballot(need);
  if (m) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(retry_count, (uint32_t)__popcll(m));
    base = readlane_u(base, 0);
    if (need) retry[base + lane_prefix(m)] = make_uint2((uint32_t)t, (uint32_t)item_idx);
  }
  // counters: queries with an h that needed the exact search, and why
  const unsigned long long nfall = __popcll(m & ballot(hp > 0));
  const unsigned long long nlo = __popcll(ballot(miss_lo)), nhi = __popcll(ballot(miss_hi));
  const unsigned long long nann = __popcll(ballot(miss_ann)), nsplit = __popcll(ballot(split_fb));
  if (lane == 0) {
    // the simulation's bracket misses (k_reduce adapts its delta)
    uint32_t* ms = walk_miss + ((int64_t)it.sim * kMissSlots + (blockIdx.x & (kMissSlots - 1))) *
                                   kMissStride;
    if (nlo + nhi) atomicAdd(ms, (uint32_t)(nlo + nhi));
    if (nann) atomicAdd(ms + 1, (uint32_t)nann);
    KnnStats* sp = stat_slot(stats);
    stat_add(&sp->retries, nfall);
    stat_add(&sp->miss_lo, nlo);
    stat_add(&sp->miss_hi, nhi);
    stat_add(&sp->miss_ann, nann);
    stat_add(&sp->split_fallback, nsplit);
  }
}

// Byte offset of record j of an array of T, from a packed entry
// pk = j << 4 | image (pk & ~15 is 16 j): one shift at most, so that the
// address is one shift-add on the array base.
template <typename T>
__device__ inline uint64_t rec_off(uint32_t pk) {
  constexpr int sz = (int)sizeof(T);
  static_assert(sz == 8 || sz == 16 || sz == 32 || sz == 64, "record size");
  const uint32_t m = pk & ~15u;
  if constexpr (sz == 8) return m >> 1;
  else if constexpr (sz == 16) return m;
  else return (uint64_t)m << (sz == 32 ? 1 : 2);
}

// Sources of the neighbour gathers, indexed by packed entry.  A gather runs
// in three steps over a round's entries: near() issues the first load and
// says whether it is the value, pin() waits for those loads, far() loads the
// rest.  An array: one global load.
template <typename T>
struct ArrSrc {
  const T* __restrict__ g;
  __device__ bool near(uint32_t pk, T& v) const {
    v = *(const T*)((const char*)g + rec_off<T>(pk));
    return true;
  }
  __device__ void pin(T&) const {}
  __device__ T far(uint32_t pk) const { return *(const T*)((const char*)g + rec_off<T>(pk)); }
};

// An array read through a block's LDS window: entries [wb, wb + W)
// (indices mod 2^32) come from the staged copy, the rest from the array.
// Particles are in Hilbert order, so most neighbours of a block's 256
// consecutive queries lie within a few hundred slots of them.  Every lane
// reads the window (slot 0 when outside it) and pin() passes the values
// through an empty asm statement, so that the compiler cannot fold the two
// loads into one flat load: a flat load completes out of order with the
// global loads, and every wait for a gather would also wait for the row
// chunk loaded behind it.
template <typename T, int W>
struct WinSrc {
  const T* __restrict__ g;
  const T* lds;
  uint32_t wb;
  __device__ bool near(uint32_t pk, T& v) const {
    const uint32_t wj = (pk >> 4) - wb;
    const bool in = wj < (uint32_t)W;
    v = lds[in ? wj : 0u];
    return in;
  }
  __device__ void pin(T& v) const {
    uint32_t* w = (uint32_t*)&v;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) asm("" : "+v"(w[i]));
  }
  __device__ T far(uint32_t pk) const { return *(const T*)((const char*)g + rec_off<T>(pk)); }
};

// Gathers the sources of n packed entries.
template <int N, typename S, typename T>
__device__ inline void gather(const S& src, const uint32_t* pks, T* pj) {
  bool in[N];
#pragma unroll
  for (int u = 0; u < N; ++u) in[u] = src.near(pks[u], pj[u]);
#pragma unroll
  for (int u = 0; u < N; ++u) src.pin(pj[u]);
#pragma unroll
  for (int u = 0; u < N; ++u)
    if (!in[u]) pj[u] = src.far(pks[u]);
}

// Stages entries [wb, wb + W) of `g` that lie in [lo, hi) into `lds` (the
// block's threads together; the caller synchronises).
template <typename T, int W>
__device__ inline void stage_window(T* lds, const T* __restrict__ g, int64_t wb, int64_t lo,
                                    int64_t hi) {
  for (int i = threadIdx.x; i < W; i += blockDim.x) {
    const int64_t j = wb + i;
    if (j >= lo && j < hi) lds[i] = g[j];
  }
}

struct NoFlush {
  __device__ void operator()() const {}
};

// Visits the k entries of a neighbour row, N (4 or 8) at a time: the N
// sources (src's records of the entries) are gathered together and the next N entries' row
// chunks are loaded before the current N are processed, so that the row and
// gather latencies overlap.  Full rounds gather their entries as they are; a
// last partial round gathers the query's own slot `self` for the entries past
// k.  NT: the row is read for the last time (non-temporal loads).  `flush`
// runs after each round of N entries (block-compensated sums).
template <int N, bool NT = false, typename S, typename Body, typename Flush = NoFlush>
__device__ inline void row_visit(const NbrRow row, int k, uint32_t self, const S& src,
                                 Body body, Flush flush = Flush()) {
  static_assert(N == 4 || N == 8, "one or two row chunks per round");
  using T = std::remove_cv_t<std::remove_reference_t<decltype(src.far(0u))>>;
  auto ld = [&](int c) {
    if constexpr (NT) return __builtin_nontemporal_load(row.chunk(c));
    else return *row.chunk(c);
  };
  // entries past k are never read from the chunks (they gather `self`), so a
  // chunk past the row is zeros rather than a copy that waits for a load
  const u32x4 zero = {0u, 0u, 0u, 0u};
  u32x4 qa = ld(0);
  u32x4 qb = (N == 8 && k > 4) ? ld(1) : zero;
  int n0 = 0;
  for (; n0 + N <= k; n0 += N) {
    const uint32_t pks[8] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w};
    T pj[N];
    gather<N>(src, pks, pj);
    if (n0 + N < k) {
      qa = ld((n0 + N) >> 2);
      if constexpr (N == 8) qb = n0 + 12 < k ? ld(((n0 + N) >> 2) + 1) : zero;
    }
#pragma unroll
    for (int u = 0; u < N; ++u) body(pks[u], pj[u]);
    flush();
  }
  if (n0 < k) {
    const uint32_t spk = self << 4;
    const uint32_t pks[8] = {qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w};
    uint32_t gpk[N];
#pragma unroll
    for (int u = 0; u < N; ++u) gpk[u] = n0 + u < k ? pks[u] : spk;
    T pj[N];
    gather<N>(src, gpk, pj);
#pragma unroll
    for (int u = 0; u < N; ++u) {
      if (n0 + u >= k) break;
      body(pks[u], pj[u]);
    }
    flush();
  }
}

// Block-compensated FP32 sum: the terms of one round (4 or 8 neighbours)
// add up plainly, and each round's partial joins the running sum through a
// Kahan step.  Besides the compensated sum's own error this adds at most
// (B - 1) 2^-24 of the sum of |terms| of a block of B, against four
// operations per term for a Kahan step per term.
struct BlockKahan {
  float s = 0.f, c = 0.f, p = 0.f;
  __device__ void add(float t) { p += t; }
  __device__ void add_mul(float a, float b) { p = __builtin_fmaf(a, b, p); }
  __device__ void flush() {
    const float y = p - c;
    const float tt = s + y;
    c = (tt - s) - y;
    s = tt;
    p = 0.f;
  }
  __device__ double value() const { return (double)s - (double)c; }
};

// Completes the fast-path queries: the r = k - c_lo smallest annulus
// candidates by (d2, ID, image) fill the row, then density, sound speed and
// the force record.  One thread per query.
template <bool FP32C>
__global__ __launch_bounds__(256) void k_knn_finish(
    const Item* __restrict__ items, int nitems, const SimDev* __restrict__ sims,
    const double2* __restrict__ XY, const uint2* __restrict__ UXY, double* __restrict__ H,
    const double* __restrict__ EP, const double* __restrict__ VPX,
    const double* __restrict__ VPY, const uint32_t* __restrict__ ID, uint32_t* __restrict__ nbr,
    int kstride, double* __restrict__ RHO, double* __restrict__ C, void* __restrict__ recp,
    const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ ann,
    KnnStats* __restrict__ stats) {
  const Item it = items[xcd_item(nitems)];
  const int lane = threadIdx.x & 63;
  const bool valid = (int)threadIdx.x < it.count;
  const int64_t t = it.first + (valid ? threadIdx.x : 0);
  const uint32_t cp = cnt[t];
  // the block's window of positions in LDS for the density sums, slots
  // [first - 128, first + 384) (most neighbours of Hilbert-ordered queries)
  constexpr int kWin = 512;
  constexpr int kBack = 128;
  __shared__ double2 wxy[kWin];
  __shared__ uint2 wuxy[FP32C ? kWin : 1];
  const int64_t wb = it.first - kBack;
  {
    const SimDev* S = sims + it.sim;
    if constexpr (FP32C) stage_window<uint2, kWin>(wuxy, UXY, wb, S->off, S->off + S->n);
    else stage_window<double2, kWin>(wxy, XY, wb, S->off, S->off + S->n);
    __syncthreads();
  }
  const WinSrc<double2, kWin> xys{XY, wxy, (uint32_t)wb};
  const WinSrc<uint2, FP32C ? kWin : 1> uxys{UXY, wuxy, (uint32_t)wb};
  unsigned long long nties = 0;
  if (valid && cp != 0xFFFFFFFFu) {
    const SimDev* S = sims + it.sim;
    const int k = S->k;
    const double lx = S->lx, ly = S->ly;
    const int c_lo = (int)(cp & 255u), c_an = (int)((cp >> 8) & 255u);
    const int r = k - c_lo;
    const double2 pq = XY[t];
    const double xq = pq.x, yq = pq.y;
    const NbrRow row = nbr_row(nbr, t, kstride);
    // annulus squared distances, recomputed exactly as in the walk; slots
    // from c_an on stand for the query itself at infinite distance, after
    // every entry.  Every query owns kAnn annulus slots, so FP32C reads all
    // of them unconditionally (no load waits behind a branch; the unused
    // ones hold stale values); FP64, whose density loop holds more registers,
    // keeps the occupancy of reading only the c_an it has.
    const uint32_t* ai = ann + (t >> 6) * (int64_t)(64 * kAnn) + (t & 63);  // stride 64
    const uint32_t self_pk = ((uint32_t)t << 4) | 5u;
    uint32_t ap[kAnn];
    double ad[kAnn];
#pragma unroll
    for (int e = 0; e < kAnn; ++e) {
      if constexpr (FP32C) {
        const uint32_t raw = ai[64 * e];
        ap[e] = e < c_an ? raw : self_pk;
      } else {
        ap[e] = e < c_an ? ai[64 * e] : self_pk;
      }
    }
#pragma unroll
    for (int e = 0; e < kAnn; ++e) {
      const uint32_t pk = ap[e];
      const double2 pj = XY[pk >> 4];
      // p_i + (-L) rounds as the walk's p_i - L, and p_i + 0 is p_i
      const double ax = xq + (double)img_ix(pk & 15u) * lx;
      const double ay = yq + (double)img_iy(pk & 15u) * ly;
      const double ex = ax - pj.x, ey = ay - pj.y;
      if constexpr (FP32C) {
        const double d2 = ex * ex + ey * ey;
        ad[e] = e < c_an ? d2 : INFINITY;
      } else {
        ad[e] = e < c_an ? ex * ex + ey * ey : INFINITY;
      }
    }
    // the r smallest by (d2, slot): entry e is taken when fewer than r
    // entries precede it in that order, and goes to row position c_lo plus
    // that count (a fixed 16 x 15 comparisons, where a selection loop runs
    // as often as the largest r of the wave)
    double tau = -1.0;
    uint32_t taken = 0u;
#pragma unroll
    for (int e = 0; e < kAnn; ++e) {
      int before = 0;
#pragma unroll
      for (int f = 0; f < kAnn; ++f)
        if (f != e) before += (f < e ? ad[f] <= ad[e] : ad[f] < ad[e]) ? 1 : 0;
      if (e < c_an && before < r) {
        row[c_lo + before] = ap[e];
        tau = ad[e] > tau ? ad[e] : tau;
        taken |= 1u << e;
      }
    }
    // an annulus entry left out with d2 == tau: resolve the tie exactly
    bool tie = false;
#pragma unroll
    for (int e = 0; e < kAnn; ++e)
      tie = tie || (e < c_an && !((taken >> e) & 1u) && ad[e] == tau);
    if (tie) {
      ++nties;
      int below = 0;
#pragma unroll
      for (int e = 0; e < kAnn; ++e) below += ad[e] < tau ? 1 : 0;
      // the k - c_lo - below slots after the entries below tau take the
      // tied entries with the smallest (ID, image order)
      uint64_t prev = 0;
      bool first = true;
      for (int s_ = c_lo + below; s_ < k; ++s_) {
        uint64_t best = ~0ull;
        uint32_t bpk = 0;
        for (int e = 0; e < c_an; ++e) {
          if (ad[e] != tau) continue;
          const uint64_t key = ((uint64_t)ID[ap[e] >> 4] << 4) | img_order(ap[e] & 15u);
          if ((first || key > prev) && key < best) {
            best = key;
            bpk = ap[e];
          }
        }
        row[s_] = bpk;
        prev = best;
        first = false;
      }
    }
    // density over the row's k neighbours
    const double h = sqrt(tau);
    const int kern = S->kernel;
    double facc;
    if constexpr (FP32C) {
      BlockKahan acc;
      kern_visit(kern, [&](auto kt) {
        using KT = decltype(kt);
        // q of the k-th neighbour may round to a few ulp past 1, where F is
        // of order ulp^3 (no clamp needed)
        auto kahan = [&](float q) { acc.add(KT::F(q)); };
        auto flush = [&]() { acc.flush(); };
        // displacements from the fixed-point positions (the minimum image,
        // exact to L / 2^32) while every neighbour lies within half a box,
        // scaled straight to units of h; otherwise from the FP64 positions.
        // Square roots in hardware (within 1 ulp): the FP32 form's error
        // budget holds them.
        if (fixed_point_ok(h, lx, ly)) {
          const uint2 me = UXY[t];
          const float gx = (float)(lx * 0x1p-32 / h), gy = (float)(ly * 0x1p-32 / h);
          const float cx = centred_step(xq, S->x0, lx) * gx, cy = centred_step(yq, S->y0, ly) * gy;
          row_visit<8>(row, k, (uint32_t)t, uxys, [&](uint32_t, uint2 pj) {
#pragma clang fp contract(fast)
            const float ex = __builtin_fmaf((float)(int32_t)(pj.x - me.x), gx, cx);
            const float ey = __builtin_fmaf((float)(int32_t)(pj.y - me.y), gy, cy);
            kahan(__builtin_amdgcn_sqrtf(__builtin_fmaf(ey, ey, ex * ex)));
          }, flush);
        } else {
          // q^2 in FP64 before the FP32 square root (see wave_search)
          const double itau = 1.0 / tau;
          row_visit<8>(row, k, (uint32_t)t, ArrSrc<double2>{XY}, [&](uint32_t pk, double2 pj) {
            const double ex = (xq + (double)img_ix(pk & 15u) * lx) - pj.x;
            const double ey = (yq + (double)img_iy(pk & 15u) * ly) - pj.y;
            kahan(__builtin_amdgcn_sqrtf((float)((ex * ex + ey * ey) * itau)));
          }, flush);
        }
      });
      facc = acc.value();
    } else {
      double acc = 0.0;
      const double hinv64 = 1.0 / h;
      row_visit<8>(row, k, (uint32_t)t, xys, [&](uint32_t pk, double2 pj) {
#pragma clang fp contract(fast)
        const double ex = (xq + (double)img_ix(pk & 15u) * lx) - pj.x;
        const double ey = (yq + (double)img_iy(pk & 15u) * ly) - pj.y;
        const double d2 = ex * ex + ey * ey;
        acc += kern_F_fma(kern, (d2 * rsq64(d2)) * hinv64);
      });
      facc = acc;
    }
    write_result<FP32C>(S, t, xq, yq, tau, facc, EP, VPX, VPY, H, RHO, C, recp);
  }
  for (int o = 32; o > 0; o >>= 1) nties += __shfl_xor(nties, o, 64);
  if (lane == 0) stat_add(&stat_slot(stats)->ties, nties);
}

// Exact wavefront search for the queries of the retry list (or, in
// all-queries mode, for every particle: first pass and bitwise order).
// Waves stride over the list, so the slow searches spread over the GPU.
// Latency-bound (a chain of dependent loads per query): held to four waves
// per SIMD (a few bytes spill) it hides that latency better than at the
// three its registers would otherwise allow.
template <bool BITWISE, bool FP32C>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_knn_fallback(
    const Item* __restrict__ items, int nitems, const SimDev* __restrict__ sims,
    const double2* __restrict__ XY,
    const uint8_t* __restrict__ WR, const uint32_t* __restrict__ cstart,
    double* __restrict__ H, const double* __restrict__ EP, const double* __restrict__ VPX,
    const double* __restrict__ VPY, const uint32_t* __restrict__ ID,
    uint32_t* __restrict__ nbr, int kstride, double* __restrict__ RHO,
    double* __restrict__ C, void* __restrict__ recp, KnnStats* __restrict__ stats,
    const uint2* __restrict__ retry, const uint32_t* __restrict__ retry_count, int all_mode,
    int* __restrict__ err) {
  __shared__ SearchShared wsh[4];
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  SearchShared& sh = wsh[wv];
  const int64_t nw = (int64_t)gridDim.x * 4;
  const int64_t total = all_mode ? (int64_t)nitems * kItem : (int64_t)(*retry_count);
  unsigned long long nretry = 0, nties = 0, nfail = 0;
  for (int64_t q = (int64_t)blockIdx.x * 4 + wv; q < total; q += nw) {
    int64_t t;
    int item;
    if (all_mode) {
      item = (int)(q / kItem);
      const int qi = (int)(q % kItem);
      if (qi >= items[item].count) continue;
      t = items[item].first + qi;
    } else {
      const uint2 e = retry[q];
      t = e.x;
      item = (int)e.y;
    }
    const SimDev* S = sims + items[item].sim;
    const double2 pq = XY[t];
    const double xq = pq.x, yq = pq.y, hp = H[t];
    const double R0 = hp > 0 ? hp * 1.1 : local_radius(S, cstart, xq, yq);
    const unsigned long long nfail0 = nfail;
    const SearchOut so = wave_search<BITWISE, FP32C>(S, t, xq, yq, R0, XY, WR, cstart, ID,
                                                     nbr_row(nbr, t, kstride), sh, nretry, nties, nfail);
    if (lane == 0 && nfail != nfail0) atomicOr(err + 1 + items[item].sim, 2);
    if (lane == 0 && so.tau > 0)
      write_result<FP32C>(S, t, xq, yq, so.tau, so.facc, EP, VPX, VPY, H, RHO, C, recp);
  }
  if (lane == 0) {
    KnnStats* sp = stat_slot(stats);
    stat_add(&sp->retries, nretry);
    stat_add(&sp->ties, nties);
    stat_add(&sp->failures, nfail);
  }
}

// ------------------------------------------------------------------------
// Force: one thread per particle, neighbours summed in list order (the
// reference's order when the list is sorted, BITWISE), then kick, drift 2,
// periodic wrap and reflection (sim/sph.go:327-401,122-193).
template <bool FP32C, bool EXACT>
__global__ __launch_bounds__(256) void k_force(
    const Item* __restrict__ items, int nitems, SimDev* __restrict__ sims,
    const void* __restrict__ recp, Fields f, const uint32_t* __restrict__ nbr, int kstride,
    int mode, Positions pout, double* __restrict__ partials, const double2* __restrict__ XY) {
  __shared__ double red[4][kNumRed + 1];
  const int item_idx = xcd_item(nitems);
  const Item it = items[item_idx];
  SimDev* S = sims + it.sim;
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const bool valid = (int)threadIdx.x < it.count;
  const int64_t t = it.first + (valid ? threadIdx.x : 0);
  const int k = S->k;
  const double mass = S->mass;
  const double lx = S->lx, ly = S->ly;
  const int kern = S->kernel;
  const double dt = S->dt_half;
  if (mode < 0) mode = S->fmode;
  // the block's window of neighbour records in LDS: slots [first - 128,
  // first + 384) (16 KB FP32C, 32 KB FP64), which hold about 85 % of the
  // neighbours of Hilbert-ordered particles; the rest are gathered from
  // memory.  Force time at 4M clustered, k = 64: FP64 1.17 -> 0.87 ms,
  // FP32C 0.57 -> 0.49 ms (with 4 records in flight instead of 8).
  constexpr int kWin = 512;
  constexpr int kBack = 128;
  using Rec = std::conditional_t<FP32C, RecF, RecD>;
  __shared__ Rec wrec[kWin];
  const int64_t wb = it.first - kBack;
  stage_window<Rec, kWin>(wrec, (const Rec*)recp, wb, S->off, S->off + S->n);
  __syncthreads();
  const WinSrc<Rec, kWin> wsrc{(const Rec*)recp, wrec, (uint32_t)wb};
  double xi, yi, hi, ci, rhoi, A;
  double sx = 0.0, sy = 0.0, se = 0.0;
  const NbrRow row = nbr_row(nbr, t, kstride);
  if constexpr (!FP32C) {
    const RecD* rec = (const RecD*)recp;
    const RecD me = rec[t];
    xi = me.x; yi = me.y; hi = 2.0 * me.h; ci = 2.0 * me.c; rhoi = 2.0 * me.rho; A = me.B;
    const double hih = me.h, cih = me.c, rhoih = me.rho;  // halves, for the pair means
    const double vpxi = me.vpx, vpyi = me.vpy;
    const double alpha = S->alpha, beta = S->beta, eta2 = S->eta2;
    const double hinv = EXACT ? 0.0 : 1.0 / hi;
    // 4 records (64 B each) in flight per lane, the next row chunk behind them
    row_visit<4, true>(row, k, (uint32_t)t, wsrc, [&](uint32_t pk, const RecD& nb) {
      const double ox = (double)img_ix(pk & 15u) * lx, oy = (double)img_iy(pk & 15u) * ly;
      const double rABx = (nb.x - ox) - xi, rABy = (nb.y - oy) - yi;
      const double vABx = nb.vpx - vpxi, vABy = nb.vpy - vpyi;
      if constexpr (EXACT) {
        // the reference's arithmetic, operation for operation
        const double ex = (xi + ox) - nb.x, ey = (yi + oy) - nb.y;
        const double d2 = ex * ex + ey * ey;
        const double dot = vABx * rABx + vABy * rABy;
        const double d = sqrt(d2);
        const double q = d / hi;
        const double dR = kern_DF(kern, q);
        double piAB = 0.0;
        if (dot < 0) {
          const double cAB = cih + nb.c;
          const double rhoAB = rhoih + nb.rho;
          const double hAB = hih + nb.h;
          const double muAB = dot * hAB / ((rABx * rABx + rABy * rABy) + eta2);
          piAB = (-alpha * cAB * muAB + beta * muAB * muAB) / rhoAB;
        }
        sx += rABx * (piAB + A + nb.B) * dR / d;
        sy += rABy * (piAB + A + nb.B) * dR / d;
        se += dot * dR;
      } else {
#pragma clang fp contract(fast)
        // within the 1e-12 contract of this form: the distance from r_AB
        // (one rounding away from the query-side difference) and fused
        // multiply-adds throughout
        const double r2 = rABx * rABx + rABy * rABy;
        const double dot = vABx * rABx + vABy * rABy;
        const double rinv = rsq64(r2);
        const double q = (r2 * rinv) * hinv;
        const double dR = kern_DF_fma(kern, q);
        double piAB = 0.0;
        if (dot < 0) {
          // pi = mu (beta mu - alpha cAB) / rhoAB with one reciprocal:
          // t = mu / rhoAB = dot hAB / ((r^2 + eta2) rhoAB)
          const double rhoAB = rhoih + nb.rho;
          const double t = dot * (hih + nb.h) * rcp64((r2 + eta2) * rhoAB);
          const double muAB = t * rhoAB;
          piAB = (beta * muAB - alpha * (cih + nb.c)) * t;
        }
        const double w = (piAB + A + nb.B) * dR * rinv;
        sx += rABx * w;
        sy += rABy * w;
        se += dot * dR;
      }
    });
  } else {
    const RecF* rec = (const RecF*)recp;
    const RecF me = rec[t];
    const double2 pme = XY[t];
    xi = pme.x; yi = pme.y;
    hi = 2.0 * (double)me.h;
    // displacements from the fixed-point difference (the minimum image) hold
    // while every neighbour lies within half a box and h spans enough steps
    // of L / 2^32; otherwise from FP64
    const bool fixed = fixed_point_ok(hi, lx, ly);
    const float fsx = (float)(lx * 0x1p-32), fsy = (float)(ly * 0x1p-32);
    const float csx = centred_step(xi, S->x0, lx) * fsx, csy = centred_step(yi, S->y0, ly) * fsy;
    ci = 2.0 * (double)me.c;
    rhoi = 2.0 * (double)me.rho;
    A = (double)me.B;
    // own rho, c, h halved as stored (the pair means are sums of halves)
    const float vpxi = me.vpx, vpyi = me.vpy, cif = me.c, rhoif = me.rho, hif = me.h, Af = me.B;
    const float hinv = __builtin_amdgcn_rcpf(2.0f * hif);
    const float alpha = (float)S->alpha, beta = (float)S->beta, eta2 = (float)S->eta2;
    // block-compensated sums (a Kahan step per round of 4 neighbours)
    BlockKahan ax_, ay_, ae_;
    auto flush = [&]() {
      ax_.flush();
      ay_.flush();
      ae_.flush();
    };
    kern_visit(kern, [&](auto kt) {
      using KT = decltype(kt);
      auto pair = [&](float rx, float ry, const RecF& nb) {
#pragma clang fp contract(fast)
        const float d2 = rx * rx + ry * ry;
        // Two particles within one fixed-point step (L / 2^32) can get a zero
        // displacement here (or one from the FP64 positions of coincident
        // particles); their pair term vanishes with the distance (DF(q) ~ q
        // and dot ~ r as r -> 0): the reciprocal square root of at least
        // FLT_MIN keeps it 0 (q = 0, DF(0) = 0, r = 0) rather than the
        // 0 * inf of rsq(0).
        const float dinv = __builtin_amdgcn_rsqf(__builtin_fmaxf(d2, 0x1p-126f));
        // q of the k-th neighbour may round to a few ulp past 1, where DF is
        // of order ulp^2 (no clamp needed)
        const float q = (d2 * dinv) * hinv;
        const float dR = KT::DF(q);
        const float vx_ = nb.vpx - vpxi, vy_ = nb.vpy - vpyi;
        const float dot = vx_ * rx + vy_ * ry;
        // viscosity where the pair approaches (dot < 0), for every pair from
        // min(dot, 0) (mu and pi vanish otherwise; the lanes of a wave take
        // both sides anyway): mu = dot hAB / (r^2 + eta2),
        // pi = mu (beta mu - alpha cAB) / rhoAB, with one reciprocal of
        // (r^2 + eta2) rhoAB
        const float cAB = cif + nb.c;
        const float rhoAB = rhoif + nb.rho;
        const float hAB = hif + nb.h;
        const float a_ = __builtin_fminf(dot, 0.f) * hAB;
        const float inv = __builtin_amdgcn_rcpf((d2 + eta2) * rhoAB);
        const float muAB = a_ * inv * rhoAB;
        const float piAB = __builtin_fmaf(beta, muAB, -alpha * cAB) * (a_ * inv);
        const float w = (piAB + Af + nb.B) * dR * dinv;
        ax_.add_mul(rx, w);
        ay_.add_mul(ry, w);
        ae_.add_mul(dot, dR);
      };
      // 4 records (32 B each) in flight per lane: with most of them in the LDS
      // window, the registers of more buy less than a fifth wave per SIMD.  A
      // wave whose queries all take fixed-point displacements (the usual case)
      // runs a loop without the FP64 form, so every record loads whole.
      if (__all(fixed)) {
        row_visit<4, true>(row, k, (uint32_t)t, wsrc, [&](uint32_t, const RecF& nb) {
          pair(__builtin_fmaf((float)(int32_t)(nb.ux - me.ux), fsx, csx),
               __builtin_fmaf((float)(int32_t)(nb.uy - me.uy), fsy, csy), nb);
        }, flush);
      } else {
        row_visit<4, true>(row, k, (uint32_t)t, wsrc, [&](uint32_t pk, const RecF& nb) {
          float rx, ry;
          if (fixed) {
            rx = __builtin_fmaf((float)(int32_t)(nb.ux - me.ux), fsx, csx);
            ry = __builtin_fmaf((float)(int32_t)(nb.uy - me.uy), fsy, csy);
          } else {
            const double ox = (double)img_ix(pk & 15u) * lx, oy = (double)img_iy(pk & 15u) * ly;
            const double2 pj = XY[pk >> 4];
            rx = (float)((pj.x - ox) - xi);
            ry = (float)((pj.y - oy) - yi);
          }
          pair(rx, ry, nb);
        }, flush);
      }
    });
    sx = ax_.value();
    sy = ay_.value();
    se = ae_.value();
  }
  double r_e = 0, r_rho = 0, r_E = 0, r_px = 0, r_py = 0, r_cfl = INFINITY;
  if (valid) {
    const double fac = mass * S->dfpref / (hi * hi * hi);
    const double ax = sx * fac + S->gxa;
    const double ay = sy * fac + S->gya;
    const double edot = A * se * mass;
    f.ax[t] = ax;
    f.ay[t] = ay;
    f.edot[t] = edot;
    double vx = f.vx[t], vy = f.vy[t], e = f.e[t];
    double x = xi, y = yi;
    if (mode == kForceKick) {
      // kick, sim/sph.go:122-127
      const double adx = ax * (2 * dt), ady = ay * (2 * dt);
      vx = vx + adx;
      vy = vy + ady;
      e = e + edot * 2 * dt;
      // drift 2, :130-135
      const double vdx = vx * dt, vdy = vy * dt;
      x = x + vdx;
      y = y + vdy;
      // periodic wrap, :147-167 (an X wrap skips the Y wrap)
      if (x < S->hor0) {
        x += (S->hor1 - S->hor0);
      } else if (x > S->hor1) {
        x -= (S->hor1 - S->hor0);
      } else if (y < S->vert0) {
        y += (S->vert1 - S->vert0);
      } else if (y > S->vert1) {
        y -= (S->vert1 - S->vert0);
      }
      // reflection, :170-193
      if (x < S->refl[0]) { x -= x - S->refl[0]; vx = -vx; }
      if (x > S->refl[1]) { x -= x - S->refl[1]; vx = -vx; }
      if (y < S->refl[2]) { y -= y - S->refl[2]; vy = -vy; }
      if (y > S->refl[3]) { y -= y - S->refl[3]; vy = -vy; }
      f.vx[t] = vx;
      f.vy[t] = vy;
      f.e[t] = e;
    }
    pout.x[t] = x;
    pout.y[t] = y;
    const double sp2 = vx * vx + vy * vy;
    r_cfl = (0.5 * hi) / (ci + sqrt(sp2));
    r_e = e;
    r_rho = rhoi;
    r_E = mass * (e + 0.5 * sp2);
    r_px = mass * vx;
    r_py = mass * vy;
  }
  // block reductions in a fixed order
  r_e = group_sum(r_e, 64);
  r_rho = group_sum(r_rho, 64);
  r_E = group_sum(r_E, 64);
  r_px = group_sum(r_px, 64);
  r_py = group_sum(r_py, 64);
  for (int o = 32; o > 0; o >>= 1) {
    const double other = __shfl_xor(r_cfl, o, 64);
    r_cfl = other < r_cfl ? other : r_cfl;
  }
  if (lane == 0) {
    red[wv][0] = r_e;
    red[wv][1] = r_rho;
    red[wv][2] = r_E;
    red[wv][3] = r_px;
    red[wv][4] = r_py;
    red[wv][5] = r_cfl;
  }
  __syncthreads();
  if (threadIdx.x < kNumRed) {
    const int r = threadIdx.x;
    partials[(int64_t)item_idx * kNumRed + r] = (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
  }
  if (threadIdx.x == 0) {
    double m = red[0][5];
    for (int w = 1; w < 4; ++w) m = red[w][5] < m ? red[w][5] : m;
    if (m < INFINITY) atomicMin(&S->cfl_min_bits, (unsigned long long)__double_as_longlong(m));
  }
}

// Totals over a simulation's work items in a fixed order.  A simulation of
// more than 1024 items is summed in two stages: red_parts blocks each sum a
// contiguous share of its items (k_reduce_part), then k_reduce adds those
// sums in block order; a smaller one is summed by k_reduce alone.  The order
// of every addition depends only on the simulation's own item count, so
// totals are reproducible and the same alone, batched and served.
__device__ inline void block_sum(double (*sh)[256], double (&acc)[kNumRed]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int r = 0; r < kNumRed; ++r) sh[r][tid] = acc[r];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o)
#pragma unroll
      for (int r = 0; r < kNumRed; ++r) sh[r][tid] += sh[r][tid + o];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_reduce_part(const SimDev* __restrict__ sims,
                                                     const double* __restrict__ partials,
                                                     const int* __restrict__ item_first,
                                                     const int* __restrict__ item_count,
                                                     double* __restrict__ red_part,
                                                     int parts_max) {
  __shared__ double sh[kNumRed][256];
  const int sim = blockIdx.x / parts_max, part = blockIdx.x % parts_max;
  if (sims[sim].n == 0) return;  // unused slot
  const int i0 = item_first[sim], ni = item_count[sim];
  const int nparts = red_parts(ni);
  if (nparts == 1 || part >= nparts) return;
  const int chunk = (ni + nparts - 1) / nparts;
  const int ia = part * chunk, ib = ia + chunk < ni ? ia + chunk : ni;
  double acc[kNumRed];
#pragma unroll
  for (int r = 0; r < kNumRed; ++r) acc[r] = 0.0;
#pragma unroll 4
  for (int i = ia + (int)threadIdx.x; i < ib; i += 256)
#pragma unroll
    for (int r = 0; r < kNumRed; ++r) acc[r] += partials[(int64_t)(i0 + i) * kNumRed + r];
  block_sum(sh, acc);
  if (threadIdx.x < kNumRed)
    red_part[((int64_t)sim * kRedSplit + part) * kNumRed + threadIdx.x] = sh[threadIdx.x][0];
}

// One block per simulation: the totals from the first stage's sums, the
// next CFL half step, the TotalMomentum quirk value and the next bracket.
__global__ __launch_bounds__(256) void k_reduce(SimDev* __restrict__ sims,
                                                const double* __restrict__ partials,
                                                const double* __restrict__ red_part,
                                                const int* __restrict__ item_first,
                                                const int* __restrict__ item_count, Fields f,
                                                int mode, uint32_t* __restrict__ walk_miss,
                                                int walked) {
  __shared__ double sh[kNumRed][256];
  __shared__ uint32_t miss[2][kMissSlots];
  SimDev* S = sims + blockIdx.x;
  if (S->n == 0) return;  // unused slot
  if (mode < 0) mode = S->fmode;
  const int tid = threadIdx.x;
  // the walk's miss counts of this pass (copy q by thread q), read and
  // cleared for the next pass
  if (tid < kMissSlots) {
    uint32_t* wm = walk_miss + ((int64_t)blockIdx.x * kMissSlots + tid) * kMissStride;
    miss[0][tid] = wm[0];
    miss[1][tid] = wm[1];
    wm[0] = 0u;
    wm[1] = 0u;
  }
  const int64_t last = S->off + S->n - 1;
  const double vx_last = tid == 0 ? f.vx[last] : 0.0, vy_last = tid == 0 ? f.vy[last] : 0.0;
  const int i0 = item_first[blockIdx.x], ni = item_count[blockIdx.x];
  const int nparts = red_parts(ni);
  double acc[kNumRed];
#pragma unroll
  for (int r = 0; r < kNumRed; ++r) acc[r] = 0.0;
  if (nparts == 1) {
    // unrolled so that many items' loads are in flight; the additions keep
    // their order
#pragma unroll 8
    for (int i = tid; i < ni; i += 256)
#pragma unroll
      for (int r = 0; r < kNumRed; ++r) acc[r] += partials[(int64_t)(i0 + i) * kNumRed + r];
  } else if (tid < nparts) {
#pragma unroll
    for (int r = 0; r < kNumRed; ++r)
      acc[r] = red_part[((int64_t)blockIdx.x * kRedSplit + tid) * kNumRed + r];
  }
  block_sum(sh, acc);
  if (tid == 0) {
#pragma unroll
    for (int r = 0; r < kNumRed; ++r) S->red[r] = sh[r][0];
    S->last_speed = sqrt(vx_last * vx_last + vy_last * vy_last);
    if (mode == kForceKick) S->dt_last = S->dt_half;
    uint32_t move = 0u, wide = 0u;
    for (int q = 0; q < kMissSlots; ++q) {
      move += miss[0][q];
      wide += miss[1][q];
    }
    if (walked) S->walk_delta = adapt_delta(S->walk_delta, S->walk_delta0, move, wide, S->n);
    const double m = __longlong_as_double((long long)S->cfl_min_bits);
    S->dt_half = S->use_cfl ? 0.5 * (S->cfl * m) : S->dt_fixed;
    S->cfl_min_bits = (unsigned long long)__double_as_longlong(INFINITY);
  }
}

__global__ void k_init_sims(SimDev* sims, int nsims) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsims) return;
  sims[s].cfl_min_bits = (unsigned long long)__double_as_longlong(INFINITY);
}

// ------------------------------------------------------------------------
int enqueue_pass(PassBuffers& b, int perm_mode, int force_mode, hipStream_t st,
                 hipEvent_t* ev) {
  auto mark = [&](int slot, int end) {
    if (ev) (void)hipEventRecord(ev[2 * slot + end], st);  // timing only
  };
  const dim3 blk(kItem);
  const dim3 grd(b.nitems);
  const dim3 grd_all(b.nitems_all);
  mark(0, 0);
  if (b.any_open) {
    hipLaunchKernelGGL(k_bbox, grd, blk, 0, st, b.items, b.nitems, b.sims, b.pos_cur.x,
                       b.pos_cur.y, b.fcur.vx, b.fcur.vy, perm_mode, b.partials);
    hipLaunchKernelGGL(k_box, dim3(b.nsims), dim3(256), 0, st, b.sims, b.partials, b.red_first,
                       b.red_count);
  }
  hipLaunchKernelGGL(k_keys, grd_all, blk, 0, st, b.items_all, b.nitems_all, b.sims, b.pos_cur.x,
                     b.pos_cur.y, b.fcur.vx, b.fcur.vy, perm_mode, b.sort.keys, b.err,
                     b.retry_count, b.run_count, b.hole_cpp);
  mark(0, 1);
  mark(1, 0);
  uint32_t *skeys = nullptr, *perm = nullptr;
  sort_pairs(b.sort, b.np, b.key_bits, st, true, &skeys, &perm);
  mark(1, 1);
  mark(2, 0);
  hipLaunchKernelGGL(k_permute, grd_all, blk, 0, st, b.items_all, b.nitems_all, b.sims, perm, skeys,
                     perm_mode, b.pos_cur, b.fcur, b.xyp, b.uxy, b.fnew, b.wrap, b.cstart, b.runs,
                     b.run_count);
  hipLaunchKernelGGL(k_fill_runs, dim3(1024), dim3(256), 0, st, b.runs, b.run_count, skeys,
                     b.cstart);
  mark(2, 1);
  mark(3, 0);
  {
    const bool fastpath = !b.bitwise && b.have_h;
    // retry lists: four times as many blocks as fit at once, so that blocks
    // that drew quick searches make room for others (a few % faster than
    // one resident round)
    const int fb_grid = fastpath ? 4096 : (b.nitems < 4096 ? b.nitems : 4096);
#define SPH_KNN(F32, NSUB)                                                                    \
  hipLaunchKernelGGL((k_knn<false, F32, NSUB>), grd, dim3(256), 0, st, b.items, b.nitems,      \
                     b.sims, b.xyp, b.wrap, b.cstart, b.fnew.h, b.fnew.ep,                      \
                     b.fnew.vpx, b.fnew.vpy, b.fnew.id, b.nbr, b.kmax, b.fnew.rho, b.fnew.c,    \
                     b.rec, b.stats, b.retry, b.retry_count, b.knn_cnt, b.knn_ann, b.walk,      \
                     b.walk_miss)
#define SPH_FB(BW, F32)                                                                       \
  hipLaunchKernelGGL((k_knn_fallback<BW, F32>), dim3(fb_grid), dim3(256), 0, st, b.items,     \
                     b.nitems, b.sims, b.xyp, b.wrap, b.cstart, b.fnew.h,                       \
                     b.fnew.ep, b.fnew.vpx, b.fnew.vpy, b.fnew.id, b.nbr, b.kmax, b.fnew.rho,   \
                     b.fnew.c, b.rec, b.stats, b.retry, b.retry_count, fastpath ? 0 : 1, b.err)
    mark(6, 0);
    if (fastpath) {
      const bool f32 = b.prec == kPrecFP32C;
      if (b.walk.nsub == 4) {
        if (f32) SPH_KNN(true, 4); else SPH_KNN(false, 4);
      } else {
        if (f32) SPH_KNN(true, 8); else SPH_KNN(false, 8);
      }
    }
    mark(6, 1);
    mark(7, 0);
    if (fastpath) {
      if (b.prec == kPrecFP32C)
        hipLaunchKernelGGL((k_knn_finish<true>), grd, dim3(256), 0, st, b.items, b.nitems, b.sims,
                           b.xyp, b.uxy, b.fnew.h, b.fnew.ep, b.fnew.vpx, b.fnew.vpy, b.fnew.id, b.nbr,
                           b.kmax, b.fnew.rho, b.fnew.c, b.rec, b.knn_cnt, b.knn_ann, b.stats);
      else
        hipLaunchKernelGGL((k_knn_finish<false>), grd, dim3(256), 0, st, b.items, b.nitems, b.sims,
                           b.xyp, b.uxy, b.fnew.h, b.fnew.ep, b.fnew.vpx, b.fnew.vpy, b.fnew.id, b.nbr,
                           b.kmax, b.fnew.rho, b.fnew.c, b.rec, b.knn_cnt, b.knn_ann, b.stats);
    }
    mark(7, 1);
    mark(8, 0);
    if (b.bitwise) SPH_FB(true, false);
    else if (b.prec == kPrecFP32C) SPH_FB(false, true);
    else SPH_FB(false, false);
    mark(8, 1);
#undef SPH_KNN
#undef SPH_FB
  }
  mark(3, 1);
  mark(4, 0);
#define SPH_FORCE(F32, EX)                                                                     \
  hipLaunchKernelGGL((k_force<F32, EX>), grd, dim3(256), 0, st, b.items, b.nitems, b.sims,     \
                     (const void*)b.rec, b.fnew, b.nbr, b.kmax, force_mode, b.pos_cur, b.partials, \
                     b.xyp)
  if (b.prec == kPrecFP32C) SPH_FORCE(true, false);
  else if (b.bitwise) SPH_FORCE(false, true);
  else SPH_FORCE(false, false);
#undef SPH_FORCE
  mark(4, 1);
  mark(5, 0);
  if (b.red_parts_max > 1)
    hipLaunchKernelGGL(k_reduce_part, dim3(b.nsims * b.red_parts_max), dim3(256), 0, st, b.sims,
                       b.partials, b.red_first, b.red_count, b.red_part, b.red_parts_max);
  hipLaunchKernelGGL(k_reduce, dim3(b.nsims), dim3(256), 0, st, b.sims, b.partials, b.red_part,
                     b.red_first, b.red_count, b.fnew, force_mode, b.walk_miss,
                     (!b.bitwise && b.have_h) ? 1 : 0);
  mark(5, 1);
  return (int)hipGetLastError();
}

// ------------------------------------------------------------------------
// readback / upload by particle ID
__device__ inline double* field_ptr(const Positions& p, const Fields& f, int field) {
  switch (field) {
    case 0: return p.x;
    case 1: return p.y;
    case 2: return f.vx;
    case 3: return f.vy;
    case 4: return f.rho;
    case 5: return f.c;
    case 6: return f.e;
    case 7: return f.edot;
    case 8: return f.ax;
    case 9: return f.ay;
    case 10: return f.ep;
    case 11: return f.vpx;
    case 12: return f.vpy;
    case 13: return f.h;
    default: return nullptr;
  }
}

// Field readback / upload by particle ID.  Row of particle t is
// off(sim) + ID(t); `sim` < 0 selects every simulation, otherwise rows are
// relative to that simulation's first particle.
__global__ void k_scatter_by_id(const Item* __restrict__ items, int nitems,
                                const SimDev* __restrict__ sims, Positions p, Fields f,
                                int field, int sim, double* __restrict__ out) {
  const Item it = items[blockIdx.x];
  if ((int)threadIdx.x >= it.count || (sim >= 0 && it.sim != sim)) return;
  const int64_t t = it.first + threadIdx.x;
  const int64_t base = sim >= 0 ? 0 : sims[it.sim].off;
  out[base + f.id[t]] = field_ptr(p, f, field)[t];
}

__global__ void k_gather_by_id(const Item* __restrict__ items, int nitems,
                               const SimDev* __restrict__ sims, Positions p, Fields f,
                               int field, int sim, const double* __restrict__ in) {
  const Item it = items[blockIdx.x];
  if ((int)threadIdx.x >= it.count || (sim >= 0 && it.sim != sim)) return;
  const int64_t t = it.first + threadIdx.x;
  const int64_t base = sim >= 0 ? 0 : sims[it.sim].off;
  field_ptr(p, f, field)[t] = in[base + f.id[t]];
}

int enqueue_read_field(const PassBuffers& b, int field, int sim, double* dev_out, hipStream_t st) {
  hipLaunchKernelGGL(k_scatter_by_id, dim3(b.nitems), dim3(kItem), 0, st, b.items, b.nitems, b.sims,
                     b.pos_cur, b.fcur, field, sim, dev_out);
  return (int)hipGetLastError();
}

int enqueue_write_field(PassBuffers& b, int field, int sim, const double* dev_in, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_by_id, dim3(b.nitems), dim3(kItem), 0, st, b.items, b.nitems, b.sims,
                     b.pos_cur, b.fcur, field, sim, dev_in);
  return (int)hipGetLastError();
}

// Neighbour rows by ID: IDs of the neighbours, recomputed squared distances
// (same formula as the search) and, if px is not null, the neighbours'
// positions seen through the query's image offset o (NNPos = p_j - o,
// sim/nearest-neighbour.go:81,152), sorted ascending by (d2, ID).  Entries of a
// row past the simulation's k (batches of mixed k) read ID 0xFFFFFFFF and
// NaN, so every output entry is written.
__global__ void k_read_nbrs(const Item* __restrict__ items, int nitems,
                            const SimDev* __restrict__ sims, const double2* __restrict__ XY, Fields f,
                            const uint32_t* __restrict__ nbr, int kstride, int kout,
                            uint32_t* __restrict__ ids, double* __restrict__ d2o,
                            double* __restrict__ px, double* __restrict__ py) {
  const Item it = items[blockIdx.x];
  if ((int)threadIdx.x >= it.count) return;
  const SimDev* S = sims + it.sim;
  const int64_t t = it.first + threadIdx.x;
  const int k = S->k;
  const int64_t row = (int64_t)f.id[t] + S->off;
  uint32_t* ro = ids + row * kout;
  double* rd = d2o + row * kout;
  double* rx = px ? px + row * kout : nullptr;
  double* ry = py ? py + row * kout : nullptr;
  const double xi = XY[t].x, yi = XY[t].y;
  for (int n = 0; n < k; ++n) {
    const uint32_t pk = nbr_row(nbr, t, kstride)[n];
    const uint32_t j = pk >> 4;
    const double ox = (double)img_ix(pk & 15u) * S->lx;
    const double oy = (double)img_iy(pk & 15u) * S->ly;
    const double2 pj = XY[j];
    const double ex = (xi + ox) - pj.x, ey = (yi + oy) - pj.y;
    const double d2 = ex * ex + ey * ey;
    const uint32_t id = f.id[j];
    int m = n;
    while (m > 0 && (rd[m - 1] > d2 || (rd[m - 1] == d2 && ro[m - 1] > id))) {
      rd[m] = rd[m - 1];
      ro[m] = ro[m - 1];
      if (rx) {
        rx[m] = rx[m - 1];
        ry[m] = ry[m - 1];
      }
      --m;
    }
    rd[m] = d2;
    ro[m] = id;
    if (rx) {
      rx[m] = pj.x - ox;
      ry[m] = pj.y - oy;
    }
  }
  for (int n = k; n < kout; ++n) {
    ro[n] = 0xFFFFFFFFu;
    rd[n] = NAN;
    if (rx) {
      rx[n] = NAN;
      ry[n] = NAN;
    }
  }
}

int enqueue_read_neighbours(const PassBuffers& b, uint32_t* dev_ids, double* dev_d2,
                            double* dev_px, double* dev_py, hipStream_t st) {
  hipLaunchKernelGGL(k_read_nbrs, dim3(b.nitems), dim3(kItem), 0, st, b.items, b.nitems, b.sims,
                     b.xyp, b.fcur, b.nbr, b.kmax, b.kout, dev_ids, dev_d2, dev_px, dev_py);
  return (int)hipGetLastError();
}

int enqueue_init_sims(PassBuffers& b, hipStream_t st) {
  hipLaunchKernelGGL(k_init_sims, dim3((b.nsims + 255) / 256), dim3(256), 0, st, b.sims,
                     b.nsims);
  return (int)hipGetLastError();
}

}  // namespace sph
