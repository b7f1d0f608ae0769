// Stable LSD radix sort of (cell key, particle index) pairs, 8-bit digits.
//
// Replaces the reference's per-pass Treebuild/Partition permutation
// (sim/core.go:126-224): particles are reordered by the Hilbert-curve
// index of their cell in a uniform 2^lg x 2^lg grid, so that every aligned
// square block of cells, at every level, is one contiguous range of the
// state arrays.
//
// Per pass: (1) block histograms of the digit, (2) one exclusive scan over
// the digit-major [256][blocks] table, (3) a stable scatter in which each
// wave ranks a contiguous quarter of the block's tile, 64 keys at a time, by
// matching digits with 9 ballots against per-wave digit counts in LDS; the
// waves' counts are then scanned once, the tile is gathered in digit order
// in LDS and written out so that each digit's run is contiguous.  Stability
// across passes makes the result independent of launch timing, so runs are
// bitwise reproducible.
#include "sph_device.hpp"
#include "sph_host.hpp"

namespace sph {

constexpr int kSortThreads = 256;
constexpr int kSortSub = 16;  // sub-tiles of 256 keys per block
constexpr int kSortTile = kSortThreads * kSortSub;

__global__ __launch_bounds__(kSortThreads) void k_sort_hist(
    const uint32_t* __restrict__ keys, int64_t n, int shift,
    uint32_t* __restrict__ counts, int nblocks) {
  __shared__ uint32_t hist[256];
  const int tid = threadIdx.x;
  hist[tid] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll 4
  for (int r = 0; r < kSortSub; ++r) {
    int64_t i = base + (int64_t)r * kSortThreads + tid;
    if (i < n) atomicAdd(&hist[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  counts[(int64_t)tid * nblocks + blockIdx.x] = hist[tid];
}

// Row d of the digit-major table counts[d][b]: exclusive scan over blocks b
// in place, row total to tot[d].  One block per digit.
__global__ __launch_bounds__(1024) void k_sort_rowscan(uint32_t* __restrict__ counts,
                                                       uint32_t* __restrict__ tot,
                                                       int nblocks) {
  __shared__ uint32_t part[1024];
  const int tid = threadIdx.x;
  uint32_t* row = counts + (int64_t)blockIdx.x * nblocks;
  const int per = (nblocks + 1023) / 1024;
  const int b = tid * per;
  uint32_t s = 0;
  for (int i = 0; i < per; ++i) s += (b + i < nblocks) ? row[b + i] : 0u;
  part[tid] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    uint32_t t = tid >= o ? part[tid - o] : 0u;
    __syncthreads();
    part[tid] += t;
    __syncthreads();
  }
  uint32_t run = part[tid] - s;
  for (int i = 0; i < per && b + i < nblocks; ++i) {
    const uint32_t c = row[b + i];
    row[b + i] = run;
    run += c;
  }
  if (tid == 1023) tot[blockIdx.x] = part[1023];
}

__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, int64_t n,
    int shift, const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ tot,
    int nblocks, int identity_vals) {
  constexpr int kWaves = kSortThreads / 64;
  constexpr int kPerWave = kSortTile / kWaves;  // contiguous keys ranked by one wave
  __shared__ uint32_t wcnt[kWaves][256];  // per wave: keys of digit d so far
  __shared__ uint32_t scan[2][256];
  __shared__ uint32_t gofs[256];    // global position of this block's digit-d keys
  __shared__ uint32_t lstart[256];  // their first position in the block's sorted tile
  __shared__ uint32_t skey[kSortTile];
  __shared__ uint32_t sval[kSortTile];
  const int tid = threadIdx.x;
  const int w = tid >> 6;
  const int lane = tid & 63;
  const int b = blockIdx.x;
#pragma unroll
  for (int q = 0; q < kWaves; ++q) wcnt[q][tid] = 0;
  // exclusive prefix of the digit totals, this block's row offset, and the
  // exclusive prefix of this block's own digit counts
  {
    const uint32_t t = tot[tid];
    const uint32_t o = offsets[(int64_t)tid * nblocks + b];
    const uint32_t cnt = (b + 1 < nblocks ? offsets[(int64_t)tid * nblocks + b + 1] : t) - o;
    scan[0][tid] = t;
    scan[1][tid] = cnt;
    __syncthreads();
    for (int s = 1; s < 256; s <<= 1) {
      const uint32_t a = tid >= s ? scan[0][tid - s] : 0u;
      const uint32_t c = tid >= s ? scan[1][tid - s] : 0u;
      __syncthreads();
      scan[0][tid] += a;
      scan[1][tid] += c;
      __syncthreads();
    }
    gofs[tid] = scan[0][tid] - t + o;
    lstart[tid] = scan[1][tid] - cnt;
  }
  const int64_t base = (int64_t)b * kSortTile;
  const int count = (int)((n - base) < kSortTile ? (n - base) : kSortTile);
  const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // the wave's keys and values, all in flight at once
  uint32_t kreg[kSortSub], vreg[kSortSub], loc[kSortSub];
#pragma unroll
  for (int r = 0; r < kSortSub; ++r) {
    const int li = w * kPerWave + r * 64 + lane;
    const int64_t i = base + li;
    kreg[r] = li < count ? kin[i] : 0u;
    vreg[r] = li < count ? (identity_vals ? (uint32_t)i : vin[i]) : 0u;
  }
  // stable ranking within the wave, 64 keys per round in key order: a key's
  // place among its wave's keys of the same digit is the digit's count so
  // far plus the equal-digit lanes below it (matched with 9 ballots); the
  // lowest such lane then advances the count.  One wave's LDS operations
  // complete in order, so every round reads the count the last one left.
#pragma unroll
  for (int r = 0; r < kSortSub; ++r) {
    const int li = w * kPerWave + r * 64 + lane;
    const bool valid = li < count;
    const uint32_t d = valid ? ((kreg[r] >> shift) & 255u) : 256u;
    unsigned long long peers = ~0ull;
#pragma unroll
    for (int bit = 0; bit < 9; ++bit) {
      const bool on = (d >> bit) & 1u;
      const unsigned long long bal = ballot(on);
      peers &= on ? bal : ~bal;
    }
    const unsigned long long below = peers & lt_mask;
    const uint32_t before = valid ? wcnt[w][d] : 0u;
    loc[r] = before + (uint32_t)__popcll(below);
    __builtin_amdgcn_wave_barrier();
    if (valid && below == 0ull) wcnt[w][d] = before + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // digit d of wave q starts after the block's earlier digits and the
  // digit-d keys of waves 0 .. q-1
  {
    uint32_t run = lstart[tid];
#pragma unroll
    for (int q = 0; q < kWaves; ++q) {
      const uint32_t c = wcnt[q][tid];
      wcnt[q][tid] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortSub; ++r) {
    const int li = w * kPerWave + r * 64 + lane;
    if (li < count) {
      const uint32_t lp = wcnt[w][(kreg[r] >> shift) & 255u] + loc[r];
      skey[lp] = kreg[r];
      sval[lp] = vreg[r];
    }
  }
  __syncthreads();
  // write out in tile order: runs of one digit go to consecutive addresses
  for (int li = tid; li < count; li += kSortThreads) {
    const uint32_t key = skey[li];
    const uint32_t d = (key >> shift) & 255u;
    const uint32_t dst = gofs[d] + ((uint32_t)li - lstart[d]);
    kout[dst] = key;
    vout[dst] = sval[li];
  }
}

int sort_pairs(SortBuffers& sb, int64_t n, int bits, hipStream_t st,
               bool first_identity, uint32_t** keys_out, uint32_t** vals_out) {
  const int nblocks = (int)((n + kSortTile - 1) / kSortTile);
  uint32_t* kin = sb.keys;
  uint32_t* vin = sb.vals;
  uint32_t* kout = sb.keys_alt;
  uint32_t* vout = sb.vals_alt;
  const int passes = bits <= 0 ? 1 : (bits + 7) / 8;
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    hipLaunchKernelGGL(k_sort_hist, dim3(nblocks), dim3(kSortThreads), 0, st, kin, n,
                       shift, sb.counts, nblocks);
    hipLaunchKernelGGL(k_sort_rowscan, dim3(256), dim3(1024), 0, st, sb.counts, sb.tot,
                       nblocks);
    hipLaunchKernelGGL(k_sort_scatter, dim3(nblocks), dim3(kSortThreads), 0, st, kin, vin,
                       kout, vout, n, shift, sb.counts, sb.tot, nblocks,
                       (p == 0 && first_identity) ? 1 : 0);
    uint32_t* t = kin; kin = kout; kout = t;
    t = vin; vin = vout; vout = t;
  }
  *keys_out = kin;
  *vals_out = vin;
  return (int)hipGetLastError();
}

int64_t sort_counts_size(int64_t n) {
  return (int64_t)256 * ((n + kSortTile - 1) / kSortTile);
}

}  // namespace sph
