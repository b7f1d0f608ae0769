// SPH force pass and integrator kernels for gfx950 (wave64).
//
//   k_keys     row-major cell key of each particle's pass position
//              (replaces Treebuild/Partition, sim/core.go:126-224)
//   k_permute  gather the state into cell order, apply the drift-1 /
//              predictor step (sim/sph.go:108-117) or the step-0
//              initialisation (:96-101), build the cell ranges
//   k_knn      exact kNN, one wavefront per query, fused with density and
//              sound speed (sim/nearest-neighbour.go:28-165,
//              sim/sph.go:306-323,421-429)
//   k_force    pressure + artificial viscosity + EDot, one lane group per
//              particle, fused with kick, drift 2, periodic wrap and
//              reflection (sim/sph.go:327-401,122-193)
//   k_reduce   per-simulation FP64 totals and the CFL half step
//
// All FP64 arithmetic on the parity path is evaluated in the reference's
// association order; the library is compiled with -ffp-contract=off so no
// product is fused into an add.
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
  const double fx = floor((x - S->x0) * S->ivx);
  const double fy = floor((y - S->y0) * S->ivy);
  const double lox = -2.0 * S->gx, hix = 3.0 * S->gx - 1;
  const double loy = -2.0 * S->gy, hiy = 3.0 * S->gy - 1;
  CellIdx c;
  c.ok = (fx >= lox) && (fx <= hix) && (fy >= loy) && (fy <= hiy);
  c.ux = c.ok ? (int64_t)fx : 0;
  c.uy = c.ok ? (int64_t)fy : 0;
  return c;
}

__global__ __launch_bounds__(kItem) void k_keys(const Item* __restrict__ items, int nitems,
                                                const SimDev* __restrict__ sims,
                                                const double* __restrict__ X,
                                                const double* __restrict__ Y,
                                                const double* __restrict__ VX,
                                                const double* __restrict__ VY, int mode,
                                                uint32_t* __restrict__ keys, int* err) {
  const Item it = items[xcd_item(nitems)];
  if ((int)threadIdx.x >= it.count) return;
  const SimDev* S = sims + it.sim;
  const int64_t t = it.first + threadIdx.x;
  double x = X[t], y = Y[t];
  if (mode == kPermPredict) {
    const double dt = S->dt_half;
    x = x + VX[t] * dt;
    y = y + VY[t] * dt;
  }
  CellIdx c = cell_of(S, x, y);
  if (!c.ok) atomicOr(err, 1);
  const int64_t cx = c.ux - floordiv(c.ux, S->gx) * S->gx;
  const int64_t cy = c.uy - floordiv(c.uy, S->gy) * S->gy;
  keys[t] = (uint32_t)(S->coff + cy * S->gx + cx);
}

__global__ __launch_bounds__(kItem) void k_permute(
    const Item* __restrict__ items, int nitems, const SimDev* __restrict__ sims,
    const uint32_t* __restrict__ perm, const uint32_t* __restrict__ skeys, int mode,
    Positions pin, Fields fin, Positions pout, Fields fout, uint8_t* __restrict__ wrap,
    uint32_t* __restrict__ cstart, int64_t np, int64_t ncells) {
  const Item it = items[xcd_item(nitems)];
  if ((int)threadIdx.x >= it.count) return;
  const SimDev* S = sims + it.sim;
  const int64_t t = it.first + threadIdx.x;
  const uint32_t i = perm[t];
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
  pout.x[t] = x;
  pout.y[t] = y;
  fout.vx[t] = vx;
  fout.vy[t] = vy;
  fout.e[t] = e;
  fout.vpx[t] = vpx;
  fout.vpy[t] = vpy;
  fout.ep[t] = ep;
  fout.h[t] = fin.h[i];
  fout.id[t] = fin.id[i];
  fout.ax[t] = fin.ax[i];
  fout.ay[t] = fin.ay[i];
  fout.edot[t] = fin.edot[i];
  CellIdx c = cell_of(S, x, y);
  const int64_t wx = floordiv(c.ux, S->gx), wy = floordiv(c.uy, S->gy);
  wrap[t] = (uint8_t)((wx + 2) | ((wy + 2) << 3));
  // cell ranges: cells (key[t-1], key[t]] start at t
  const int64_t kt = skeys[t];
  const int64_t kp = t > 0 ? (int64_t)skeys[t - 1] : -1;
  for (int64_t q = kp + 1; q <= kt; ++q) cstart[q] = (uint32_t)t;
  if (t == np - 1)
    for (int64_t q = kt + 1; q <= ncells; ++q) cstart[q] = (uint32_t)np;
}

// ------------------------------------------------------------------------
// kNN: one wavefront per query.
//
// Search: all cells meeting the disk of radius R around the query, over
// the periodic images the reference visits (query offsets {-L,0,L}^2,
// sim/nearest-neighbour.go:57-61).  Rows of cells are contiguous particle
// ranges; each lane owns one row, splits it at the period boundaries into
// pieces, and the wave walks the concatenated candidates 64 at a time.
// Candidates with d2 <= R^2 are appended to a per-wave LDS list.  When at
// least k arrive the k smallest (d2, ID, image) are exact: every candidate
// not seen lies farther than R.  Otherwise R grows (or shrinks on overflow)
// and the search repeats.  The k-th key is found by a ballot quickselect.
constexpr int kMaxPieces = 5 * 64;

__device__ inline double pow2d(double x) { return x * x; }

struct KnnShared {
  double d2[kCandCap];
  uint32_t pk[kCandCap];
  uint32_t pst[kMaxPieces];
  int ppref[kMaxPieces + 1];
  int8_t ptx[kMaxPieces];
  int8_t pty[kMaxPieces];
  double fterm[kMaxK];
  double sd2[kMaxK];   // reference-order staging (bitwise mode)
  uint32_t spk[kMaxK];
};

template <bool BITWISE, bool FP32C>
__global__ __launch_bounds__(256) void k_knn(
    const Item* __restrict__ items, int nitems, const SimDev* __restrict__ sims,
    const double* __restrict__ X, const double* __restrict__ Y,
    const uint8_t* __restrict__ WR, const uint32_t* __restrict__ cstart,
    double* __restrict__ H, const double* __restrict__ EP, const uint32_t* __restrict__ ID,
    uint32_t* __restrict__ nbr, int kstride, double* __restrict__ RHO,
    double* __restrict__ C, KnnStats* __restrict__ stats) {
  __shared__ KnnShared sh_all[4];
  const Item it = items[xcd_item(nitems)];
  const SimDev* S = sims + it.sim;
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  KnnShared& sh = sh_all[wv];
  const int k = S->k;
  const int32_t gx = S->gx, gy = S->gy;
  const double lx = S->lx, ly = S->ly;
  unsigned long long nretry = 0, nties = 0, nfail = 0;

  for (int qq = 0; qq < 64; ++qq) {
    const int qi = wv * 64 + qq;
    if (qi >= it.count) break;
    const int64_t t = it.first + qi;
    const double xq = X[t], yq = Y[t];
    const double hp = H[t];
    double R = hp > 0 ? hp * 1.1 : S->r_init;
    const double rcap = 1.5 * (lx > ly ? lx : ly);
    if (R > rcap) R = rcap;
    double r_lo = 0.0, r_hi = 0.0;  // bracket: count < k at r_lo, overflow at r_hi
    int cnt = 0;
    double R2 = 0;
    int attempt = 0;
    for (;; ++attempt) {
      R2 = R * R;
      // culling radius: a hair larger than R so that binning rounding can
      // never hide a particle with d2 <= R^2
      const double Rm = R * (1 + 1e-12) + 1e-15 * (lx + ly);
      const double Rm2 = Rm * Rm;
      int64_t ry0 = (int64_t)floor((yq - S->y0 - Rm) * S->ivy);
      int64_t ry1 = (int64_t)floor((yq - S->y0 + Rm) * S->ivy);
      if (ry0 < -2LL * gy) ry0 = -2LL * gy;
      if (ry1 > 3LL * gy - 1) ry1 = 3LL * gy - 1;
      cnt = 0;
      for (int64_t rb = ry0; rb <= ry1; rb += 64) {
        // ---- pieces of this lane's row ----
        const int64_t ry = rb + lane;
        uint32_t pst[5];
        int plen[5];
        int ptx[5];
        int np = 0;
        int pty = 0;
        if (ry <= ry1) {
          const double by0 = S->y0 + (double)ry * S->sy;
          const double by1 = by0 + S->sy;
          const double dy = yq < by0 ? by0 - yq : (yq > by1 ? yq - by1 : 0.0);
          const double w2 = Rm2 - dy * dy;
          if (w2 >= 0) {
            const double w = sqrt(w2);
            int64_t cx0 = (int64_t)floor((xq - S->x0 - w) * S->ivx);
            int64_t cx1 = (int64_t)floor((xq - S->x0 + w) * S->ivx);
            if (cx0 < -2LL * gx) cx0 = -2LL * gx;
            if (cx1 > 3LL * gx - 1) cx1 = 3LL * gx - 1;
            const int64_t ty = floordiv(ry, gy);
            pty = (int)ty;
            const int64_t rowbase = S->coff + (ry - ty * gy) * gx;
            int64_t c = cx0;
#pragma unroll
            for (int p = 0; p < 5; ++p) {
              if (c <= cx1) {
                const int64_t tx = floordiv(c, gx);
                int64_t cend = (tx + 1) * gx - 1;
                if (cend > cx1) cend = cx1;
                const uint32_t a = cstart[rowbase + (c - tx * gx)];
                const uint32_t b = cstart[rowbase + (cend - tx * gx) + 1];
                if (b > a) {
                  pst[np] = a;
                  plen[np] = (int)(b - a);
                  ptx[np] = (int)tx;
                  ++np;
                }
                c = cend + 1;
              }
            }
          }
        }
        int tot = 0;
#pragma unroll
        for (int p = 0; p < 5; ++p) tot += p < np ? plen[p] : 0;
        const int pbase_incl = wave_incl_scan(np);
        const int cbase_incl = wave_incl_scan(tot);
        const int npieces = __shfl(pbase_incl, 63, 64);
        const int ncand = __shfl(cbase_incl, 63, 64);
        {
          int pi = pbase_incl - np;
          int ci = cbase_incl - tot;
#pragma unroll
          for (int p = 0; p < 5; ++p) {
            if (p < np) {
              sh.pst[pi] = pst[p];
              sh.ppref[pi] = ci;
              sh.ptx[pi] = (int8_t)ptx[p];
              sh.pty[pi] = (int8_t)pty;
              ++pi;
              ci += plen[p];
            }
          }
          if (lane == 63) sh.ppref[npieces] = ncand;
        }
        __builtin_amdgcn_wave_barrier();
        int nsteps = 0;
        while ((1 << nsteps) < npieces) ++nsteps;
        // ---- walk the candidates ----
        for (int base = 0; base < ncand; base += 64) {
          const int u = base + lane;
          bool acc = false;
          double d2 = 0;
          uint32_t pk = 0;
          if (u < ncand) {
            int lo = 0, hi = npieces - 1;
            for (int s = 0; s < nsteps; ++s) {
              const int mid = (lo + hi + 1) >> 1;
              if (sh.ppref[mid] <= u) lo = mid; else hi = mid - 1;
            }
            const uint32_t j = sh.pst[lo] + (uint32_t)(u - sh.ppref[lo]);
            const uint8_t wc = WR[j];
            const int nx = (int)sh.ptx[lo] - ((int)(wc & 7) - 2);
            const int ny = (int)sh.pty[lo] - ((int)(wc >> 3) - 2);
            if ((int64_t)j != t && nx >= -1 && nx <= 1 && ny >= -1 && ny <= 1) {
              // query offset o = -n * L, the reference's float64(i)*deltaX
              const double ox = (double)(-nx) * lx, oy = (double)(-ny) * ly;
              const double ex = (xq + ox) - X[j];
              const double ey = (yq + oy) - Y[j];
              d2 = ex * ex + ey * ey;
              acc = d2 <= R2;
              pk = (j << 4) | (uint32_t)((-nx + 1) * 3 + (-ny + 1));
            }
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
      }
      if (cnt >= k && cnt <= kCandCap) break;
      // new radius from the area law, kept inside the bracket
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
        ++nfail;
        break;
      }
    }
    if (attempt > 0) ++nretry;
    if (cnt < k || cnt > kCandCap) {
      // unreachable for valid input: record, emit an empty pass result
      if (lane == 0) {
        H[t] = 0.0;
        RHO[t] = 0.0;
        C[t] = 0.0;
      }
      continue;
    }
    // ---- k-th smallest squared distance: ballot quickselect ----
    double v[kCandCap / 64];
    uint32_t pkr[kCandCap / 64];
#pragma unroll
    for (int s = 0; s < kCandCap / 64; ++s) {
      const int idx = s * 64 + lane;
      v[s] = idx < cnt ? sh.d2[idx] : INFINITY;
      pkr[s] = idx < cnt ? sh.pk[idx] : 0u;
    }
    double lo = -1.0, hi = INFINITY, tau = 0;
    int c_lt = 0, c_le = 0;
    for (;;) {
      int slot = -1, src = 0;
#pragma unroll
      for (int s = 0; s < kCandCap / 64; ++s) {
        const unsigned long long b = ballot(v[s] > lo && v[s] < hi);
        if (slot < 0 && b) {
          slot = s;
          src = __builtin_ctzll(b);
        }
      }
      double pv = 0;
#pragma unroll
      for (int s = 0; s < kCandCap / 64; ++s)
        if (s == slot) pv = readlane_d(v[s], src);
      c_lt = 0;
      c_le = 0;
#pragma unroll
      for (int s = 0; s < kCandCap / 64; ++s) {
        c_lt += __popcll(ballot(v[s] < pv));
        c_le += __popcll(ballot(v[s] <= pv));
      }
      if (c_lt < k && k <= c_le) {
        tau = pv;
        break;
      }
      if (k <= c_lt) hi = pv; else lo = pv;
    }
    // ---- ties at tau: keep the smallest (ID, image) ----
    bool tsel[kCandCap / 64];
#pragma unroll
    for (int s = 0; s < kCandCap / 64; ++s) tsel[s] = false;
    if (c_le > k) {
      ++nties;
      uint64_t key2[kCandCap / 64];
#pragma unroll
      for (int s = 0; s < kCandCap / 64; ++s) {
        const int idx = s * 64 + lane;
        key2[s] = ~0ull;
        if (idx < cnt && v[s] == tau)
          key2[s] = ((uint64_t)ID[pkr[s] >> 4] << 4) | (pkr[s] & 15u);
      }
      for (int r = 0; r < k - c_lt; ++r) {
        uint64_t mn = ~0ull;
#pragma unroll
        for (int s = 0; s < kCandCap / 64; ++s) mn = key2[s] < mn ? key2[s] : mn;
        for (int o = 32; o > 0; o >>= 1) {
          const uint64_t other = __shfl_xor(mn, o, 64);
          mn = other < mn ? other : mn;
        }
#pragma unroll
        for (int s = 0; s < kCandCap / 64; ++s)
          if (key2[s] == mn) {
            tsel[s] = true;
            key2[s] = ~0ull;
          }
      }
    }
    // ---- emit the neighbour list, density, sound speed ----
    const double h = sqrt(tau);
    const int kern = S->kernel;
    double facc = 0.0;
    float fs = 0.f, fc = 0.f;
    const float hinv_f = 1.0f / (float)h;
    int emitted = 0;
#pragma unroll
    for (int s = 0; s < kCandCap / 64; ++s) {
      const int idx = s * 64 + lane;
      const bool sel = idx < cnt && (v[s] < tau || (v[s] == tau && (c_le == k || tsel[s])));
      const unsigned long long m = ballot(sel);
      if (sel) {
        const int pos = emitted + lane_prefix(m);
        const uint32_t pkv = pkr[s];
        if constexpr (!BITWISE) {
          nbr[t * kstride + pos] = pkv;
          if constexpr (FP32C) {
            const float q = sqrtf((float)v[s]) * hinv_f;
            const float f = kern_Ff(kern, q > 1.f ? 1.f : q);
            // Kahan-compensated accumulation of this lane's terms
            const float y = f - fc;
            const float tt = fs + y;
            fc = (tt - fs) - y;
            fs = tt;
          } else {
            facc += kern_F(kern, sqrt(v[s]) / h);
          }
        } else {
          sh.sd2[pos] = v[s];  // staged for the ordered sum
          sh.spk[pos] = pkv;
        }
      }
      emitted += __popcll(m);
    }
    double rho;
    if constexpr (!BITWISE) {
      double tot;
      if constexpr (FP32C) {
        // compensated tree sum: carry each level's rounding error
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
        tot = (double)s_ + (double)c_;
      } else {
        tot = group_sum(facc, 64);
      }
      rho = S->fpref * S->mass * tot / (h * h);
    } else {
      // reference order: queue index 0 = k-th (farthest) first,
      // sim/sph.go:313-320; sort the k entries descending by (d2, ID)
      __builtin_amdgcn_wave_barrier();
      double dv = lane < k ? sh.sd2[lane] : -1.0;
      uint32_t pv = lane < k ? sh.spk[lane] : 0u;
      uint32_t iv = lane < k ? ID[pv >> 4] : 0u;
      // bitonic sort, descending by (d2, id, image)
      for (int size = 2; size <= 64; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
          const double od = __shfl_xor(dv, stride, 64);
          const uint32_t op = __shfl_xor(pv, stride, 64);
          const uint32_t oi = __shfl_xor(iv, stride, 64);
          const bool other_greater =
              od > dv || (od == dv && (oi > iv || (oi == iv && (op & 15u) > (pv & 15u))));
          const bool lower = (lane & stride) == 0;
          const bool desc = (lane & size) == 0;
          const bool take = (lower == desc) ? other_greater : !other_greater;
          if (take && (od != dv || oi != iv || op != pv)) {
            dv = od; pv = op; iv = oi;
          }
        }
      }
      if (lane < k) {
        nbr[t * kstride + lane] = pv;
        sh.fterm[lane] = kern_F(kern, sqrt(dv) / h);
      }
      __builtin_amdgcn_wave_barrier();
      double acc = 0.0;
      for (int n = 0; n < k; ++n) acc += sh.fterm[n];
      rho = S->fpref * S->mass * acc / (h * h);
    }
    if (lane == 0) {
      H[t] = h;
      RHO[t] = rho;
      C[t] = sqrt(S->cfac * EP[t]);
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (lane == 0 && (nretry | nties | nfail)) {
    atomicAdd(&stats->retries, nretry);
    atomicAdd(&stats->ties, nties);
    atomicAdd(&stats->failures, nfail);
  }
}

// ------------------------------------------------------------------------
// Force: a group of L lanes (L = next power of two >= k) per particle, one
// neighbour per lane, then kick / drift 2 / wrap / reflection by the
// group's first lane.
template <bool BITWISE, bool FP32C>
__global__ __launch_bounds__(256) void k_force(
    const Item* __restrict__ items, int nitems, SimDev* __restrict__ sims, Positions p,
    Fields f, const uint32_t* __restrict__ nbr, int kstride, int mode, Positions pout,
    double* __restrict__ partials) {
  __shared__ double red[4][kNumRed + 1];
  __shared__ double terms[BITWISE ? 4 : 1][BITWISE ? 64 : 1][3];
  const int item_idx = xcd_item(nitems);
  const Item it = items[item_idx];
  SimDev* S = sims + it.sim;
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int k = S->k;
  int L = 1;
  while (L < k) L <<= 1;
  const int gpw = 64 / L;
  const int g = lane / L;
  const int n = lane % L;
  const double gamma = S->gamma, mass = S->mass;
  const double alpha = S->alpha, beta = S->beta, eta2 = S->eta2;
  const double lx = S->lx, ly = S->ly;
  const int kern = S->kernel;
  const double dt = S->dt_half;
  // per-lane partial reductions (leaders only contribute)
  double r_e = 0, r_rho = 0, r_E = 0, r_px = 0, r_py = 0;
  double r_cfl = INFINITY;

  const int wbeg = wv * 64;
  int wend = wbeg + 64;
  if (wend > it.count) wend = it.count;
  for (int base = wbeg; base < wend; base += gpw) {
    const int pi = base + g;
    const bool vp = pi < wend;
    const int64_t t = it.first + (vp ? pi : wbeg);
    const bool active = vp && n < k;
    const double xi = p.x[t], yi = p.y[t];
    const double vpxi = f.vpx[t], vpyi = f.vpy[t];
    const double rhoi = f.rho[t], ci = f.c[t], hi = f.h[t];
    const double A = ci * ci / (gamma * rhoi);
    double tx = 0.0, ty = 0.0, te = 0.0;
    if (active) {
      const uint32_t pk = nbr[t * kstride + n];
      const uint32_t j = pk >> 4;
      const int img = (int)(pk & 15u);
      const double ox = (double)(img / 3 - 1) * lx;
      const double oy = (double)(img % 3 - 1) * ly;
      const double xj = p.x[j], yj = p.y[j];
      const double ex = (xi + ox) - xj, ey = (yi + oy) - yj;
      const double d2 = ex * ex + ey * ey;
      const double rABx = (xj - ox) - xi, rABy = (yj - oy) - yi;
      const double vABx = f.vpx[j] - vpxi, vABy = f.vpy[j] - vpyi;
      const double cj = f.c[j], rhoj = f.rho[j], hj = f.h[j];
      if constexpr (!FP32C) {
        const double d = sqrt(d2);
        const double q = d / hi;
        const double dR = kern_DF(kern, q);
        const double B = cj * cj / (gamma * rhoj);
        const double dot = vABx * rABx + vABy * rABy;
        double piAB = 0.0;
        if (dot < 0) {
          const double cAB = 0.5 * (ci + cj);
          const double rhoAB = 0.5 * (rhoi + rhoj);
          const double hAB = 0.5 * (hi + hj);
          const double muAB = dot * hAB / ((rABx * rABx + rABy * rABy) + eta2);
          piAB = (-alpha * cAB * muAB + beta * muAB * muAB) / rhoAB;
        }
        tx = rABx * (piAB + A + B) * dR / d;
        ty = rABy * (piAB + A + B) * dR / d;
        te = dot * dR;
      } else {
        const float d = sqrtf((float)d2);
        const float q = d * (1.0f / (float)hi);
        const float dR = kern_DFf(kern, q > 1.f ? 1.f : q);
        const float cjf = (float)cj, rjf = (float)rhoj;
        const float B = cjf * cjf / ((float)gamma * rjf);
        const float rx = (float)rABx, ry = (float)rABy;
        const float vx_ = (float)vABx, vy_ = (float)vABy;
        const float dot = vx_ * rx + vy_ * ry;
        float piAB = 0.f;
        if (dot < 0.f) {
          const float cAB = 0.5f * ((float)ci + cjf);
          const float rhoAB = 0.5f * ((float)rhoi + rjf);
          const float hAB = 0.5f * (float)(hi + hj);
          const float muAB = dot * hAB / ((rx * rx + ry * ry) + (float)eta2);
          piAB = (-(float)alpha * cAB * muAB + (float)beta * muAB * muAB) / rhoAB;
        }
        const float sdr = (piAB + (float)A + B) * dR / d;
        tx = (double)(rx * sdr);
        ty = (double)(ry * sdr);
        te = (double)(dot * dR);
      }
    }
    double sx, sy, se;
    if constexpr (!BITWISE) {
      sx = group_sum(tx, L);
      sy = group_sum(ty, L);
      se = group_sum(te, L);
    } else {
      terms[wv][lane][0] = tx;
      terms[wv][lane][1] = ty;
      terms[wv][lane][2] = te;
      __builtin_amdgcn_wave_barrier();
      sx = 0.0; sy = 0.0; se = 0.0;
      if (n == 0) {
        for (int m = 0; m < k; ++m) {
          sx += terms[wv][g * L + m][0];
          sy += terms[wv][g * L + m][1];
          se += terms[wv][g * L + m][2];
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
    if (vp && n == 0) {
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
      const double cfl_t = (0.5 * hi) / (ci + sqrt(sp2));
      r_cfl = cfl_t < r_cfl ? cfl_t : r_cfl;
      r_e += e;
      r_rho += rhoi;
      r_E += mass * (e + 0.5 * sp2);
      r_px += mass * vx;
      r_py += mass * vy;
    }
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

// One block per simulation: totals over its items in a fixed order, the
// next CFL half step, the TotalMomentum quirk value.
__global__ __launch_bounds__(256) void k_reduce(SimDev* __restrict__ sims,
                                                const double* __restrict__ partials,
                                                const int* __restrict__ item_first,
                                                const int* __restrict__ item_count, Fields f,
                                                int mode) {
  __shared__ double sh[kNumRed][256];
  SimDev* S = sims + blockIdx.x;
  const int tid = threadIdx.x;
  const int i0 = item_first[blockIdx.x], ni = item_count[blockIdx.x];
  double acc[kNumRed];
#pragma unroll
  for (int r = 0; r < kNumRed; ++r) acc[r] = 0.0;
  for (int i = tid; i < ni; i += 256)
#pragma unroll
    for (int r = 0; r < kNumRed; ++r) acc[r] += partials[(int64_t)(i0 + i) * kNumRed + r];
#pragma unroll
  for (int r = 0; r < kNumRed; ++r) sh[r][tid] = acc[r];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o)
#pragma unroll
      for (int r = 0; r < kNumRed; ++r) sh[r][tid] += sh[r][tid + o];
    __syncthreads();
  }
  if (tid == 0) {
#pragma unroll
    for (int r = 0; r < kNumRed; ++r) S->red[r] = sh[r][0];
    const int64_t last = S->off + S->n - 1;
    const double vx = f.vx[last], vy = f.vy[last];
    S->last_speed = sqrt(vx * vx + vy * vy);
    if (mode == kForceKick) S->dt_last = S->dt_half;
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
    if (ev) hipEventRecord(ev[2 * slot + end], st);
  };
  const dim3 blk(kItem);
  const dim3 grd(b.nitems);
  mark(0, 0);
  hipLaunchKernelGGL(k_keys, grd, blk, 0, st, b.items, b.nitems, b.sims, b.pos_cur.x,
                     b.pos_cur.y, b.fcur.vx, b.fcur.vy, perm_mode, b.sort.keys, b.err);
  mark(0, 1);
  mark(1, 0);
  uint32_t *skeys = nullptr, *perm = nullptr;
  sort_pairs(b.sort, b.np, b.key_bits, st, true, &skeys, &perm);
  mark(1, 1);
  mark(2, 0);
  hipLaunchKernelGGL(k_permute, grd, blk, 0, st, b.items, b.nitems, b.sims, perm, skeys,
                     perm_mode, b.pos_cur, b.fcur, b.pos_pass, b.fnew, b.wrap, b.cstart, b.np,
                     b.ncells);
  mark(2, 1);
  mark(3, 0);
  const int knn_blk = 256;
#define SPH_KNN(BW, F32)                                                                    \
  hipLaunchKernelGGL((k_knn<BW, F32>), grd, dim3(knn_blk), 0, st, b.items, b.nitems, b.sims, \
                     b.pos_pass.x, b.pos_pass.y, b.wrap, b.cstart, b.fnew.h, b.fnew.ep,       \
                     b.fnew.id, b.nbr, b.kmax, b.fnew.rho, b.fnew.c, b.stats)
  if (b.bitwise) SPH_KNN(true, false);
  else if (b.prec == kPrecFP32C) SPH_KNN(false, true);
  else SPH_KNN(false, false);
#undef SPH_KNN
  mark(3, 1);
  mark(4, 0);
#define SPH_FORCE(BW, F32)                                                                  \
  hipLaunchKernelGGL((k_force<BW, F32>), grd, dim3(256), 0, st, b.items, b.nitems, b.sims,   \
                     b.pos_pass, b.fnew, b.nbr, b.kmax, force_mode, b.pos_cur, b.partials)
  if (b.bitwise) SPH_FORCE(true, false);
  else if (b.prec == kPrecFP32C) SPH_FORCE(false, true);
  else SPH_FORCE(false, false);
#undef SPH_FORCE
  mark(4, 1);
  mark(5, 0);
  hipLaunchKernelGGL(k_reduce, dim3(b.nsims), dim3(256), 0, st, b.sims, b.partials,
                     b.red_first, b.red_count, b.fnew, force_mode);
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

__global__ void k_scatter_by_id(Positions p, Fields f, int field, int64_t np,
                                double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np) return;
  out[f.id[t]] = field_ptr(p, f, field)[t];
}

__global__ void k_gather_by_id(Positions p, Fields f, int field, int64_t np,
                               const double* __restrict__ in) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np) return;
  field_ptr(p, f, field)[t] = in[f.id[t]];
}

int enqueue_read_field(const PassBuffers& b, int field, double* dev_out, hipStream_t st) {
  const int64_t nb = (b.np + 255) / 256;
  hipLaunchKernelGGL(k_scatter_by_id, dim3(nb), dim3(256), 0, st, b.pos_cur, b.fcur, field,
                     b.np, dev_out);
  return (int)hipGetLastError();
}

int enqueue_write_field(PassBuffers& b, int field, const double* dev_in, hipStream_t st) {
  const int64_t nb = (b.np + 255) / 256;
  hipLaunchKernelGGL(k_gather_by_id, dim3(nb), dim3(256), 0, st, b.pos_cur, b.fcur, field,
                     b.np, dev_in);
  return (int)hipGetLastError();
}

// neighbour rows by ID, IDs of the neighbours, recomputed squared distances
// (same formula as the search), sorted ascending by (d2, ID)
__global__ void k_read_nbrs(const Item* __restrict__ items, int nitems,
                            const SimDev* __restrict__ sims, Positions p, Fields f,
                            const uint32_t* __restrict__ nbr, int kstride,
                            uint32_t* __restrict__ ids, double* __restrict__ d2o) {
  const Item it = items[blockIdx.x];
  if ((int)threadIdx.x >= it.count) return;
  const SimDev* S = sims + it.sim;
  const int64_t t = it.first + threadIdx.x;
  const int k = S->k;
  const int64_t row = (int64_t)f.id[t] + S->off;
  uint32_t* ro = ids + row * kstride;
  double* rd = d2o + row * kstride;
  const double xi = p.x[t], yi = p.y[t];
  for (int n = 0; n < k; ++n) {
    const uint32_t pk = nbr[t * kstride + n];
    const uint32_t j = pk >> 4;
    const int img = (int)(pk & 15u);
    const double ox = (double)(img / 3 - 1) * S->lx;
    const double oy = (double)(img % 3 - 1) * S->ly;
    const double ex = (xi + ox) - p.x[j], ey = (yi + oy) - p.y[j];
    const double d2 = ex * ex + ey * ey;
    const uint32_t id = f.id[j];
    int m = n;
    while (m > 0 && (rd[m - 1] > d2 || (rd[m - 1] == d2 && ro[m - 1] > id))) {
      rd[m] = rd[m - 1];
      ro[m] = ro[m - 1];
      --m;
    }
    rd[m] = d2;
    ro[m] = id;
  }
}

int enqueue_read_neighbours(const PassBuffers& b, uint32_t* dev_ids, double* dev_d2,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_read_nbrs, dim3(b.nitems), dim3(kItem), 0, st, b.items, b.nitems, b.sims,
                     b.pos_pass, b.fcur, b.nbr, b.kmax, dev_ids, dev_d2);
  return (int)hipGetLastError();
}

int enqueue_init_sims(PassBuffers& b, hipStream_t st) {
  hipLaunchKernelGGL(k_init_sims, dim3((b.nsims + 255) / 256), dim3(256), 0, st, b.sims,
                     b.nsims);
  return (int)hipGetLastError();
}

}  // namespace sph
