// Host-side declarations shared by the kernel launchers and the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_device.hpp"

namespace sph {

struct SortBuffers {
  uint32_t* keys;
  uint32_t* vals;
  uint32_t* keys_alt;
  uint32_t* vals_alt;
  uint32_t* counts;  // [256][blocks]
  uint32_t* tot;     // [256]
};

int sort_pairs(SortBuffers& sb, int64_t n, int bits, hipStream_t st, bool first_identity,
               uint32_t** keys_out, uint32_t** vals_out);
int64_t sort_counts_size(int64_t n);

// One structure-of-arrays set of per-particle fields (positions separate).
struct Fields {
  double *vx, *vy, *e, *ax, *ay, *edot, *vpx, *vpy, *ep, *h, *rho, *c;
  uint32_t* id;
};

struct Positions {
  double *x, *y;
};

// -1 (kPerSim) selects each simulation's own mode (SimDev::pmode / fmode)
enum PermuteMode : int { kPerSim = -1, kPermPredict = 0, kPermInit = 1, kPermKeep = 2 };
enum ForceMode : int { kForceKick = 0, kForceNoKick = 1 };

struct PassBuffers {
  const Item* items;     // work items of the simulations
  int nitems;
  const Item* items_all; // plus unused pool ranges (sim = -1), for keys/permute
  int nitems_all;
  uint32_t hole_cpp;     // key of an unused slot t is t * hole_cpp
  SimDev* sims;
  int nsims;
  int kmax;             // row stride of the neighbour list (multiple of 4)
  int kout;             // row stride of neighbour readback (largest k)
  int prec;             // kPrecFP64 / kPrecFP32C (uniform over a batch)
  int bitwise;          // reference-order sums (FP64 only)
  const int* red_first; // per simulation: first item
  const int* red_count; // per simulation: item count
  int64_t np;           // particles
  int64_t ncells;       // cells (all simulations)
  int key_bits;
  Positions pos_cur;    // positions at rest (P0)
  double2* xyp;         // positions of the pass, interleaved (x, y)
  uint2* uxy;           // FP32C: the same as 32-bit fixed point of the box
  Fields fcur;          // fields, current order
  Fields fnew;          // fields, order of this pass
  uint8_t* wrap;        // wrap code per particle of the pass
  uint32_t* cstart;     // ncells + 1
  uint32_t* nbr;        // np * kmax packed (index << 4 | image code)
  void* rec;            // np force records (RecD or RecF)
  double* partials;     // nitems * kNumRed
  KnnStats* stats;      // kStatSlots copies
  uint32_t* walk_miss;  // per simulation: the walk's miss counts (kMissSlots lines)
  double* red_part;     // per simulation: kRedSplit first-stage sums (kNumRed each)
  int red_parts_max;    // the largest red_parts over the simulations (1: one stage)
  uint2* retry;           // kNN retry list (np entries)
  uint32_t* retry_count;
  uint32_t* runs;          // queued cell chunks of the pass: (first, last, t0, nv) x kRunQueue
  uint32_t* run_count;
  uint32_t* knn_cnt;       // per query: certain | annulus << 8, or ~0 (retry)
  uint32_t* knn_ann;       // per query: annulus candidates (kAnn entries)
  int have_h;             // every particle has an h from an earlier pass
  int any_open;           // some simulation has an open axis (k_bbox, k_box run)
  WalkKnobs walk;         // tile-search shape of the kNN walk
  int* err;             // device error flags: [0] any, [1 + s] simulation s (1: a
                        // particle left the images, 2: kNN found fewer than k)
  SortBuffers sort;
};

// Enqueue one force pass (keys, sort, permute, kNN + density, force,
// reductions) on `st`.  Timing events (may be null) are recorded around
// each kernel class.
// Sum of the per-block copies of the kNN counters.
inline KnnStats sum_stats(const KnnStats* v) {
  KnnStats r = {};
  for (int i = 0; i < kStatSlots; ++i) {
    r.retries += v[i].retries;
    r.ties += v[i].ties;
    r.failures += v[i].failures;
    r.miss_lo += v[i].miss_lo;
    r.miss_hi += v[i].miss_hi;
    r.miss_ann += v[i].miss_ann;
    r.split_fallback += v[i].split_fallback;
  }
  return r;
}

// u32 entries of the neighbour lists of np queries (64-query tiles)
inline int64_t nbr_entries(int64_t np, int kstride) { return (np + 63) / 64 * 64 * kstride; }

int enqueue_pass(PassBuffers& b, int perm_mode, int force_mode, hipStream_t st,
                 hipEvent_t* ev /* 2*SPH_NTIMERS or null */);

int enqueue_read_field(const PassBuffers& b, int field, int sim, double* dev_out, hipStream_t st);
int enqueue_write_field(PassBuffers& b, int field, int sim, const double* dev_in, hipStream_t st);
// neighbour rows by ID (ids, d2 and, if dev_px is not null, the image
// positions NNPos), ascending by (d2, ID); np * kout entries each
int enqueue_read_neighbours(const PassBuffers& b, uint32_t* dev_ids, double* dev_d2,
                            double* dev_px, double* dev_py, hipStream_t st);
int enqueue_init_sims(PassBuffers& b, hipStream_t st);

}  // namespace sph
