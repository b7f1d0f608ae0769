// Device-side definitions shared by the SPH kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sph {

constexpr int kWave = 64;
constexpr int kItem = 256;        // particles per work item (one block)
constexpr int kMaxK = 64;         // neighbours per particle
constexpr int kCandCap = 256;     // kNN candidates held per wave
constexpr int kNumRed = 5;        // sum e, sum rho, E, Px, Py
constexpr int kRedSplit = 16;     // first-stage reduction blocks per simulation
// first-stage reduction blocks of a simulation with nitems work items (one
// per 1024 items, at most kRedSplit; 1: its totals are summed in one stage)
__host__ __device__ inline int red_parts(int nitems) {
  const int p = (nitems + 1023) / 1024;
  return p < 1 ? 1 : (p > kRedSplit ? kRedSplit : p);
}
// the walk's miss counts of a pass in progress, per simulation: kMissSlots
// copies, one 128-byte line each ([0]: h moved past the bracket, [1]:
// annulus overflow), blocks adding to copy blockIdx % kMissSlots
constexpr int kMissSlots = 16;
constexpr int kMissStride = 32;   // u32 per copy (one cache line)
constexpr int kAnn = 16;          // kNN annulus entries kept per query
// cell ranges: a permute wave whose particles span more than kWaveCells
// cells queues them, in chunks of at most kRunChunk cells, for a parallel
// fill (kRunQueue chunks per pass; a wave that finds the queue full writes
// its cells itself)
constexpr int kWaveCells = 4096;
constexpr int kRunChunk = 4096;
constexpr int kRunQueue = 65536;

// Go folds 6*40/(Pi*7), 4*7/(Pi*4), 8*7/(Pi*4) exactly and rounds once
// (sim/sph.go:266,275,293,303); these are the correctly rounded values.
constexpr double kMonaghanPref = 0x1.5d3b3e3583243p+3;
constexpr double kWendlandFPref = 0x1.1d34a60108f72p+1;
constexpr double kWendlandDFPref = 0x1.1d34a60108f72p+2;

enum : int32_t { kMonaghan = 0, kWendland = 1 };
enum : int32_t { kPrecFP64 = 0, kPrecFP32C = 1 };
enum : uint32_t { kFlagBitwise = 1u, kFlagTiming = 2u };

// Per-simulation parameters and device-updated scalars.
struct SimDev {
  int64_t off;      // first particle (global index)
  int64_t coff;     // first cell (global index)
  int32_t n;        // particles
  int32_t k;        // neighbours
  int32_t kernel;   // kMonaghan / kWendland
  int32_t prec;     // kPrecFP64 / kPrecFP32C
  uint32_t flags;
  int32_t use_cfl;
  int32_t gx, gy;   // cells per axis (gx = gy = 2^lg)
  int32_t lg;       // log2 of cells per axis (cells in Hilbert order)
  int32_t pmode;    // pass modes when a launch runs each simulation in its own
  int32_t fmode;    // mode (continuous batching): PermuteMode / ForceMode
  int32_t open;     // open axes (bit 0: x, bit 1: y; bounds +-MaxFloat64)
  double x0, y0;    // grid box origin: Hor[0], Vert[0] of a periodic axis; on
  double lx, ly;    // an open axis the box around the pass positions, set on
                    // the device before each pass (k_box)
  double ivx, ivy;  // cells per unit length
  double sx, sy;    // cell size
  double gamma, mass, cfl;
  double gxa, gya;  // Acceleration
  double hor0, hor1, vert0, vert1;
  double refl[4];   // L R U D
  double alpha, beta, eta2;
  double fpref, dfpref;
  double cfac;      // gamma*(gamma-1)
  double dt_fixed;
  double r_init;    // first-pass kNN radius guess
  // device-updated
  double dt_half;       // half step of the current step
  double dt_last;       // half step of the last completed step
  unsigned long long cfl_min_bits;  // atomicMin over non-negative doubles
  double red[kNumRed];  // reductions of the last pass
  double last_speed;    // |v| of the particle stored last
  // kNN walk bracket half-width of this simulation, adapted after each pass
  // from the queries its walk could not resolve (k_reduce): misses because
  // h moved past the bracket widen it, annulus overflows narrow it, and it
  // returns towards walk_delta0 (the configured value) when misses are rare
  double walk_delta;
  double walk_delta0;
};

// Bracket adaptation after a pass with n queries, from its walk's misses
// because h moved past the bracket (fewer than k within it, or k within its
// inner radius) and its annulus overflows: a miss rate above 1/128
// of either kind moves delta by a quarter (towards the kind that dominates),
// not past [kDeltaMin, kDeltaMax]; once both stay below a quarter of that
// rate it moves a tenth back towards the configured d0, and in between it
// holds (so a burst of motion widens it for as long as it lasts, and it
// does not settle at the edge of the widening threshold).  Neighbour sets and h do not depend on delta (every miss is
// searched exactly); the share of queries the walk resolves does, and
// outside the bitwise form so does the order in which a row's neighbours
// are summed.  The counts are integers, so the adaptation is deterministic.
constexpr double kDeltaMin = 0.0125, kDeltaMax = 0.1;
__host__ __device__ inline double adapt_delta(double d, double d0, uint32_t move, uint32_t wide,
                                              int32_t n) {
  const double thr = n / 128.0;
  if (move > thr && move >= wide) return d * 1.25 < kDeltaMax ? d * 1.25 : (d > kDeltaMax ? d : kDeltaMax);
  if (wide > thr) return d * 0.8 > kDeltaMin ? d * 0.8 : (d < kDeltaMin ? d : kDeltaMin);
  if (d == d0 || move > thr / 4 || wide > thr / 4) return d;
  const double t = d > d0 ? d * 0.9 : d * 1.1;
  return (d > d0) == (t > d0) ? t : d0;
}

// Grid box of a simulation's open axes around the range [lo, hi] of its pass
// positions per axis (the reference searches no periodic images on such an
// axis, sim/nearest-neighbour.go:39-55).  The box is widened by 2^-8 of the
// range plus 2^-40 of the coordinates' size, far beyond the rounding of the
// box and of floor((x - x0) * ivx), so every position lands strictly inside
// cells [0, 2^lg).  With both axes open the box is a square (square cells).
// An empty range takes the other axis's size (or 1).  Same arithmetic on the
// host (the creation guess) and the device (each pass).
__host__ __device__ inline void open_box(SimDev& d, double xlo, double xhi, double ylo,
                                         double yhi) {
  const double G = (double)(1 << d.lg);
  auto pad = [](double lo, double hi) {
    return (hi - lo) * 0x1p-8 + (fabs(lo) + fabs(hi)) * 0x1p-40 + 0x1p-1000;
  };
  const bool ox = (d.open & 1) != 0, oy = (d.open & 2) != 0;
  double wx = ox ? (xhi > xlo ? (xhi - xlo) + 2 * pad(xlo, xhi) : 0.0) : d.lx;
  double wy = oy ? (yhi > ylo ? (yhi - ylo) + 2 * pad(ylo, yhi) : 0.0) : d.ly;
  if (ox && oy) {
    double w = wx > wy ? wx : wy;
    if (!(w > 0)) w = 1.0;
    wx = w;
    wy = w;
  } else if (ox && !(wx > 0)) {
    wx = d.ly;
  } else if (oy && !(wy > 0)) {
    wy = d.lx;
  }
  if (ox) {
    d.x0 = 0.5 * (xlo + xhi) - 0.5 * wx;
    d.lx = wx;
    d.ivx = G / d.lx;
    d.sx = d.lx / G;
  }
  if (oy) {
    d.y0 = 0.5 * (ylo + yhi) - 0.5 * wy;
    d.ly = wy;
    d.ivy = G / d.ly;
    d.sy = d.ly / G;
  }
}

// Hilbert-curve index of cell (x, y) on a 2^lg x 2^lg grid.  The curve visits
// the four quadrants of every square one after the other (each quadrant
// turned so that consecutive cells share a side), so, as with a Morton code,
// the cells of a coarser level c are contiguous index ranges: cell (X, Y)
// of level c is [hilbert2(X, Y, c) << 2(lg - c), + 4^(lg - c)).  Unlike a
// Morton code it never jumps across the box, so consecutive particles stay
// spatially compact: on the clustered 4M layout (k = 64), the sub-tile boxes
// of the kNN walk hold 42 % fewer candidates per query and its 64-query
// tiles 34 % fewer, with none too wide to search.  The state is the
// orientation of the current square (a swap of x and y, a complement of
// both): each level reads one bit of x and y, emits a base-4 digit and
// updates the orientation.
__host__ __device__ inline uint32_t hilbert2(uint32_t x, uint32_t y, int lg) {
  uint32_t d = 0, swap = 0, comp = 0;
  for (int s = lg - 1; s >= 0; --s) {
    const uint32_t bx = (x >> s) & 1u, by = (y >> s) & 1u;
    const uint32_t rx = (swap ? by : bx) ^ comp;
    const uint32_t ry = (swap ? bx : by) ^ comp;
    d = (d << 2) | ((3u * rx) ^ ry);
    const uint32_t m = ry ^ 1u;
    swap ^= m;
    comp ^= m & rx;
  }
  return d;
}

// Image code of a neighbour entry: the query offset o = (ix*Lx, iy*Ly) of
// the reference's d2 = ((p_i + o) - p_j)^2, ix, iy in {-1, 0, 1}.
__device__ inline uint32_t img_code(int ix, int iy) {
  return (uint32_t)((ix + 1) | ((iy + 1) << 2));
}
__device__ inline int img_ix(uint32_t code) { return (int)(code & 3u) - 1; }
__device__ inline int img_iy(uint32_t code) { return (int)((code >> 2) & 3u) - 1; }
// order of images used by the oracle for (d2, ID, image) ties
__device__ inline uint32_t img_order(uint32_t code) {
  return (uint32_t)((img_ix(code) + 1) * 3 + (img_iy(code) + 1));
}

// Per-particle record gathered by the force kernel (written by the kNN
// kernel once density, sound speed and h are known).  rho, c and h are
// stored halved: the pair means 0.5 (a_i + a_j) are then one add, and equal
// (a_i / 2 + a_j / 2 rounds as (a_i + a_j) / 2: halving is exact).
struct alignas(16) RecD {
  double x, y, vpx, vpy, rho, c, h, B;  // rho, c, h halved; B = c^2 / (gamma rho)
};
// FP32C record: the position as 32-bit fixed point of the box (units of
// L / 2^32, wrapped modulo 2^32), so that the 32-bit difference of two
// records is the minimum-image displacement; the rest in FP32.
struct alignas(16) RecF {
  uint32_t ux, uy;
  float vpx, vpy, rho, c, h, B;  // rho, c, h halved, as in RecD
};

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

struct Item {
  int32_t sim;
  int32_t count;
  int64_t first;
};

// Tile-search shape of the kNN walk (the autotuner's knobs): bracket
// half-width around each query's last h, the largest candidate count per
// query of a lane group before it is halved, the lane group a wave starts
// with (queries searched together: 64, 32 or 16), and the sub-tiles a wave
// filters its staged candidates for (4 or 8).
struct WalkKnobs {
  double delta;
  int32_t split;
  int32_t group;
  int32_t nsub;
  int32_t pad;
};
constexpr WalkKnobs kWalkDefault = {0.025, 24, 64, 8, 0};

struct KnnStats {
  unsigned long long retries;         // searches needing a second radius
  unsigned long long ties;            // exact d2 ties at the k-th boundary
  unsigned long long failures;        // searches that found fewer than k
  unsigned long long miss_lo;         // bracket: k or more inside h(1-delta)
  unsigned long long miss_hi;         // bracket: fewer than k inside h(1+delta)
  unsigned long long miss_ann;        // bracket: annulus over capacity
  unsigned long long split_fallback;  // tile too wide even for 8 lanes
};

// kNN counters are spread over kStatSlots copies, chosen by block, so that
// thousands of waves do not serialise their atomics on one address; readers
// sum the copies.
constexpr int kStatSlots = 64;
__device__ inline KnnStats* stat_slot(KnnStats* base) {
  return base + (blockIdx.x & (kStatSlots - 1));
}
__device__ inline void stat_add(unsigned long long* p, unsigned long long v) {
  if (v) atomicAdd(p, v);
}

// Work-item index for this block, arranged so that each of the 8 XCDs
// (which receive blocks round-robin) walks a contiguous run of items and
// keeps its own L2 warm on one band of the particle array.
__device__ inline int xcd_item(int nitems) {
  const int b = blockIdx.x;
  const int q = nitems >> 3, r = nitems & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

__device__ inline int64_t floordiv(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}
// floor(a / 2^s): the arithmetic shift (every grid and level has 2^s cells
// per axis, so no 64-bit division is needed on the device)
__device__ inline int64_t floordiv_pow2(int64_t a, int s) { return a >> s; }

// Reference kernel functions, evaluated in Go's order with no contraction.
#pragma clang fp contract(off)
__device__ inline double kern_F(int32_t kernel, double q) {
  if (kernel == kWendland) return (1 - q) * (1 - q) * (1 - q) * (1 - q) * (1 + 4 * q);
  if (q < 0.5) return q * q * q - q * q + 1.0 / 6;
  return (1 - q) * (1 - q) * (1 - q) / 3;
}
__device__ inline double kern_DF(int32_t kernel, double q) {
  if (kernel == kWendland) return -10 * q * (1 - q) * (1 - q) * (1 - q);
  if (q < 0.5) return (3 * q * q - 2 * q);
  return -(1 - q) * (1 - q);
}
// FP64 kernel value with fused multiply-adds (the plain FP64 form's density,
// whose contract is 1e-12 relative; the bitwise form keeps kern_F)
__device__ inline double kern_F_fma(int32_t kernel, double q) {
  if (kernel == kWendland) {
    const double a = 1 - q, a2 = a * a;
    return a2 * a2 * fma(4.0, q, 1.0);
  }
  if (q < 0.5) return fma(q * q, q - 1, 1.0 / 6);
  const double a = 1 - q;
  return a * a * a * (1.0 / 3);
}
// FP64 derivative with fused multiply-adds (the plain FP64 form's force,
// whose contract is 1e-12 relative; the bitwise form keeps kern_DF)
__device__ inline double kern_DF_fma(int32_t kernel, double q) {
  if (kernel == kWendland) {
    const double a = 1 - q;
    return -10 * q * (a * a * a);
  }
  if (q < 0.5) return q * fma(3.0, q, -2.0);
  const double a = 1 - q;
  return -(a * a);
}
// FP32 forms (FP32C), fused multiply-adds where they fit.  The two pieces
// of the cubic spline are both evaluated and one is selected: the lanes of
// a wave take both sides anyway, and a select needs no exec-mask branch.
// FP32 forms of each kernel type.  The type is per simulation, so loops
// over neighbours are instantiated per type (kern_visit) rather than
// branching on it per pair.
struct CubicF {
  __device__ static float F(float q) {
    const float a = 1.f - q;
    const float lo = __builtin_fmaf(q * q, q - 1.f, 1.f / 6.f);
    const float hi = a * a * a * (1.f / 3.f);
    return q < 0.5f ? lo : hi;
  }
  __device__ static float DF(float q) {
    const float a = 1.f - q;
    const float lo = q * __builtin_fmaf(3.f, q, -2.f);
    const float hi = -a * a;
    return q < 0.5f ? lo : hi;
  }
};
struct WendlandF {
  __device__ static float F(float q) {
    const float a = 1.f - q;
    const float a2 = a * a;
    return a2 * a2 * __builtin_fmaf(4.f, q, 1.f);
  }
  __device__ static float DF(float q) {
    const float a = 1.f - q;
    return -10.f * q * a * a * a;
  }
};
__device__ inline float kern_Ff(int32_t kernel, float q) {
  return kernel == kWendland ? WendlandF::F(q) : CubicF::F(q);
}
__device__ inline float kern_DFf(int32_t kernel, float q) {
  return kernel == kWendland ? WendlandF::DF(q) : CubicF::DF(q);
}
// Calls f(CubicF{}) or f(WendlandF{}) for the simulation's kernel type.
template <typename Fn>
__device__ inline void kern_visit(int32_t kernel, Fn&& f) {
  if (kernel == kWendland) f(WendlandF{});
  else f(CubicF{});
}

// FP64 reciprocal and reciprocal square root from the hardware estimates
// refined by two Newton steps (explicit fma; results within about one ulp).
// Used by the non-bitwise FP64 form, whose contract is 1e-12 relative.
__device__ inline double rcp64(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}
__device__ inline double rsq64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  double r = fma(-hx * y, y, 0.5);
  y = fma(y, r, y);
  r = fma(-hx * y, y, 0.5);
  return fma(y, r, y);
}

// ---- wave helpers (wave64) ----
__device__ inline int lane_id() { return __lane_id(); }

__device__ inline unsigned long long ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// number of set bits of m in lanes below this one
__device__ inline int lane_prefix(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ inline double readlane_d(double v, int lane) {
  union { double d; int32_t i[2]; } u, r;
  u.d = v;
  r.i[0] = __builtin_amdgcn_readlane(u.i[0], lane);
  r.i[1] = __builtin_amdgcn_readlane(u.i[1], lane);
  return r.d;
}
__device__ inline uint32_t readlane_u(uint32_t v, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

// inclusive scan over the wave
__device__ inline int wave_incl_scan(int v) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (l >= o) v += t;
  }
  return v;
}

// butterfly sum over groups of `width` lanes (power of two); every lane of
// a group ends with the same value (a+b == b+a exactly)
__device__ inline double group_sum(double v, int width) {
  for (int o = width >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline float group_sumf(float v, int width) {
  for (int o = width >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace sph
