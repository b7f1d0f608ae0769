/*
 * sphugo_amd: MI355X-native SPH step (k-nearest-neighbour search, density,
 * pressure + artificial-viscosity force, leapfrog DKD integration) behind a
 * plain C ABI.  Device-resident structure-of-arrays state; hand-written HIP
 * kernels for gfx950.
 *
 * Every entry point returns 0 on success or a negative SPH_E* code; the
 * message of the last failure on the calling thread is sph_last_error().
 * The library never aborts.  A handle is not re-entrant: one caller thread
 * at a time (the reference holds Simulation.IsBusy across Step,
 * sim/sph.go:66,197); distinct handles are independent.
 *
 * Reference interfaces replaced (bbeni/sphugo, Go package `sim`):
 *   sph_create            MakeSimulationFromConf        sim/sph.go:41-54
 *   sph_step              (*Simulation).Step            sim/sph.go:64-198
 *   sph_calculate_forces  (*Simulation).CalculateForces sim/sph.go:403-435
 *   sph_read_field        Root.Particles[i].{Pos,Vel,Rho,C,E,EDot,VDot,
 *                         EPred,VPred,NNDists[0]}        sim/core.go:17-42
 *                         (the Animator's readback, sim/animator.go:55-176)
 *   sph_read_neighbours   Particle.NearestNeighbours/NNDists after
 *                         FindNearestNeighboursPeriodic sim/nearest-neighbour.go:28-67
 *   sph_total_*           TotalEnergy/TotalDensity/TotalMomentum
 *                                                        sim/sph.go:441-463
 *   sph_params            SphConfig fields read by the step
 *                                                        sim/config-parser.go:111-149
 */
#ifndef SPHUGO_H
#define SPHUGO_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define SPH_OK 0
#define SPH_EINVAL -1   /* bad argument or parameter */
#define SPH_EDEVICE -2  /* HIP runtime error */
#define SPH_ENOMEM -3   /* device allocation failed or admission refused */
#define SPH_ESTATE -4   /* call not valid in the current state */
#define SPH_ERANGE -5   /* particle left the supported range (e.g. NaN) */

/* kernels (sim/sph.go:258,284; config words "Monahan", "Wendtland") */
#define SPH_KERNEL_MONAGHAN 0
#define SPH_KERNEL_WENDLAND 1

/* arithmetic of the density and force sums */
#define SPH_PREC_FP64 0  /* plain FP64 throughout */
#define SPH_PREC_FP32C 1 /* FP32 pair arithmetic, compensated sums; FP64 kNN
                            distances, h, positions, CFL and reductions */

/* flags */
#define SPH_FLAG_BITWISE 1u /* FP64: neighbour sums in the reference's order
                               (farthest first, sim/sph.go:313,343) */
#define SPH_FLAG_TIMING 2u  /* record per-kernel HIP event timings */

/* fields (same numbering in the oracle) */
#define SPH_X 0
#define SPH_Y 1
#define SPH_VX 2
#define SPH_VY 3
#define SPH_RHO 4
#define SPH_C 5
#define SPH_E 6
#define SPH_EDOT 7
#define SPH_AX 8
#define SPH_AY 9
#define SPH_EPRED 10
#define SPH_VPX 11
#define SPH_VPY 12
#define SPH_H 13 /* k-th neighbour distance, NNDists[0] */
#define SPH_NFIELDS 14

/* per-kernel timer slots (SPH_FLAG_TIMING) */
#define SPH_T_KEYS 0
#define SPH_T_SORT 1
#define SPH_T_PERMUTE 2
#define SPH_T_KNN_DENSITY 3
#define SPH_T_FORCE 4
#define SPH_T_REDUCE 5
/* the three launches inside SPH_T_KNN_DENSITY */
#define SPH_T_KNN_WALK 6     /* tile walk of the bracketed fast path */
#define SPH_T_KNN_FINISH 7   /* annulus selection, density, force records */
#define SPH_T_KNN_FALLBACK 8 /* exact wavefront search of the remaining queries */
#define SPH_NTIMERS 9

/* Values the step reads from SphConfig (sim/config-parser.go:111-149) plus
 * the parameters the reference hard-codes (sim/core.go:14 NN_SIZE,
 * sim/sph.go:379-383 viscosity).  hor/vert equal to {-DBL_MAX, DBL_MAX}
 * mean an open axis, as in the reference. */
typedef struct sph_params {
  int32_t k;          /* neighbours per particle, 2..64 (reference: 32) */
  int32_t kernel;     /* SPH_KERNEL_* */
  int32_t precision;  /* SPH_PREC_* */
  uint32_t flags;     /* SPH_FLAG_* */
  double gamma;       /* Gamma (reference default 1.66666) */
  double mass;        /* ParticleMass */
  double dt_half;     /* DeltaTHalf; <= 0 selects the CFL half step */
  double cfl;         /* CFL number for dt_half <= 0 */
  double accel[2];    /* Acceleration (gravity) */
  double hor[2];      /* HorPeriodicity */
  double vert[2];     /* VertPeriodicity */
  double refl[4];     /* Reflections L, R, U, D */
  double visc_alpha;  /* 0.75 in the reference */
  double visc_beta;   /* 1.5 */
  double visc_eta2;   /* 0.01; must be > 0 */
} sph_params;

typedef struct sph_sim sph_sim;

/* Reference defaults (MakeConfig, sim/config-parser.go:131-149, with k=32
 * and the viscosity constants of sim/sph.go:379-383). */
int sph_default_params(sph_params* p);

/* Upload n particles (positions, velocities, specific internal energy).
 * vx, vy, e may be NULL (zero).  id may be NULL (IDs 0..n-1); otherwise IDs
 * must be a permutation of 0..n-1.  The caller keeps ownership of the host
 * arrays.  MakeSimulationFromConf, sim/sph.go:41-54. */
int sph_create(const sph_params* p, int64_t n, const double* x, const double* y,
               const double* vx, const double* vy, const double* e,
               const uint32_t* id, sph_sim** out);
void sph_destroy(sph_sim* s);

/* Sources (sim/sph.go:71-86, the spawners of sim/config-parser.go:58-102
 * running on the host): append m particles with the next IDs n..n+m-1.
 * vx, vy, e may be NULL (zero); every other field of a new particle starts
 * at zero, as Go's zero value, and its first neighbour search is the exact
 * one.  Call between steps (the reference spawns at the start of Step); the
 * half step already set for the next step is kept.  One simulation per
 * handle (SPH_EINVAL for a batch). */
int sph_append(sph_sim* s, int64_t m, const double* x, const double* y,
               const double* vx, const double* vy, const double* e);

/* nsteps calls of Simulation.Step (sim/sph.go:64-198), including the extra
 * force pass of step 0; returns when the device has finished. */
int sph_step(sph_sim* s, int32_t nsteps);
/* Same, without waiting for the device (pair with sph_synchronize). */
int sph_step_async(sph_sim* s, int32_t nsteps);
int sph_synchronize(sph_sim* s);

/* Simulation.CalculateForces (sim/sph.go:403-435) on the current state:
 * kNN, density, sound speed from EPred, force from VPred. */
int sph_calculate_forces(sph_sim* s);

/* Copy one field of every particle to host memory, indexed by particle ID. */
int sph_read_field(sph_sim* s, int32_t field, double* out);
/* Overwrite one field (X..VPY) of every particle, indexed by ID. */
int sph_write_field(sph_sim* s, int32_t field, const double* in);
/* Neighbours of the last pass: ids[ID*k + i] and d2[ID*k + i] (d2 may be
 * NULL), rows sorted ascending by (squared distance, ID).  In a batch whose
 * simulations have different k, rows are k_max long and entries past a
 * simulation's k read ID 0xFFFFFFFF and d2 NaN. */
int sph_read_neighbours(sph_sim* s, uint32_t* ids, double* d2);
/* NNPos of the last pass (Particle.NNPos, sim/core.go:38, filled at
 * sim/nearest-neighbour.go:81,152): the neighbours' positions seen through the
 * query's periodic image, p_j - o, in the row order of sph_read_neighbours
 * (so the reference's NNPos[0], the farthest, is row entry k-1 and NNPos[1]
 * entry k-2; AnimatorGL.AddFrame reads those two, sim/animator.go:171-172). */
int sph_read_neighbour_pos(sph_sim* s, double* px, double* py);
/* The same rows as sph_read_neighbours, staged through a fixed device
 * buffer instead of buffers allocated on the first call: no device memory
 * is allocated or freed.  No counterpart in the reference (whose neighbour
 * arrays live in the Go particle slice); for callers that cannot allocate
 * device memory (stream capture, serving loops, harnesses that forbid it).
 * Limit: particles x k_max <= 2^20 entries (e.g. 16,384 particles at
 * k = 64); beyond it SPH_EINVAL.  The buffer (12 MiB per GPU, held until the
 * process ends) is shared by every call on a GPU: calls that overlap on one
 * GPU race. */
int sph_read_neighbours_noalloc(sph_sim* s, uint32_t* ids, double* d2);

/* sum e, sum rho, E = sum m(e + v^2/2), Px, Py, last dt_half (FP64, fixed
 * reduction order). */
int sph_reductions(sph_sim* s, double out[6]);
/* reference totals: TotalEnergy = sum e, TotalDensity = sum rho,
 * TotalMomentum = |v| of the particle stored last (sim/sph.go:457-463) */
int sph_total_energy(sph_sim* s, double* out);
int sph_total_density(sph_sim* s, double* out);
int sph_total_momentum(sph_sim* s, double* out);

int64_t sph_current_step(const sph_sim* s);
int64_t sph_num_particles(const sph_sim* s);
/* CFL half step the next step will use (when dt_half <= 0). */
int sph_next_dt_half(sph_sim* s, double* out);

/* Average device time per launch of each kernel class since the last
 * reset (SPH_FLAG_TIMING), in milliseconds, and launch counts. */
int sph_kernel_times(sph_sim* s, double ms[SPH_NTIMERS], int64_t counts[SPH_NTIMERS]);
int sph_reset_kernel_times(sph_sim* s);
/* Switch per-kernel timing on (1) or off (0) for later steps.  Timing
 * waits for the device after each pass, so whole-step throughput is
 * measured with it off. */
int sph_set_timing(sph_sim* s, int32_t on);

/* Number of kNN searches that needed more than one radius, and the
 * number of exact squared-distance ties met at the k-th boundary, summed
 * over all passes since sph_create. */
int sph_knn_stats(sph_sim* s, int64_t* retries, int64_t* ties);
/* All kNN counters since sph_create: retries, ties, failures, bracket
 * misses (k inside the inner radius, fewer than k inside the outer radius,
 * annulus over capacity), tiles sent to the wavefront search. */
int sph_knn_counters(sph_sim* s, int64_t out[7]);

/* ---- checkpoint and resume (no reference counterpart: the reference keeps
 * no state on disk; its Animator holds rendered frames only,
 * sim/animator.go:111-130) ----
 * A checkpoint is one byte image: a 64-byte header ("SPHCKPT1", version, N,
 * step, ...), the creation parameters, the simulation's device scalars
 * (half step, CFL minimum, totals) and the 14 per-particle FP64 arrays
 * (x y vx vy e ax ay edot vpx vpy epred h rho c) in the device's slot order
 * followed by the particle IDs (u32).  A handle loaded from it continues
 * bit for bit as the saved one would have.  One simulation per handle
 * (SPH_EINVAL for a batch); the caller owns the buffer.  Load validates the
 * header and sizes and returns SPH_EINVAL for anything else. */
int sph_checkpoint_size(sph_sim* s, int64_t* bytes);
int sph_checkpoint_save(sph_sim* s, void* buf, int64_t bytes);
int sph_checkpoint_load(const void* buf, int64_t bytes, sph_sim** out);

/* ---- batches of independent simulations (no reference counterpart) ----
 * One handle holding `count` simulations; every kernel of a step covers all
 * of them in one launch over (simulation, tile) work items.  Precision and
 * the bitwise flag are shared; k, kernel, physics and box are per
 * simulation.  x[i], y[i], vx[i], vy[i], e[i] hold simulation i's n[i]
 * particles (vx, vy, e or their entries may be NULL).  sph_step and the
 * whole-handle readbacks work on the concatenation of all simulations. */
int sph_batch_create(int32_t count, const sph_params* params, const int64_t* n,
                     const double* const* x, const double* const* y,
                     const double* const* vx, const double* const* vy,
                     const double* const* e, sph_sim** out);
int32_t sph_batch_count(const sph_sim* s);
int64_t sph_batch_num_particles(const sph_sim* s, int32_t sim);
int sph_batch_read_field(sph_sim* s, int32_t sim, int32_t field, double* out);
int sph_batch_write_field(sph_sim* s, int32_t sim, int32_t field, const double* in);
int sph_batch_reductions(sph_sim* s, int32_t sim, double out[6]);

/* Replay regular steps from a captured hipGraph (1) or launch them one by
 * one (0).  Results are identical either way. */
int sph_set_graph(sph_sim* s, int32_t on);
/* Time force passes over the candidate cell grid resolutions (cells per
 * particle 0.5 ... 16) on the current state and keep the fastest, then tune
 * the kNN walk's tile shape on that grid one knob at a time: sub-tiles per
 * wave (8, 4), the lane group a wave starts with (64, 32 queries), and the
 * candidates per query above which a group is halved (24, 32, 16).  Only
 * before the first step (SPH_ESTATE otherwise).  Either output may be NULL.
 * Outside the bitwise form the grid and shape change the order in which
 * neighbours are summed (results within the precision form's contract). */
int sph_autotune(sph_sim* s, double* best_ms, double* best_cells_per_particle);
/* The grid and walk shape in use: cells per particle, bracket half-width
 * around each query's last h, split threshold, first lane group, sub-tiles.
 * The bracket adapts per simulation after every pass (widened by a quarter
 * when more than 1/128 of the queries fell outside it because h moved,
 * narrowed by a fifth when more than 1/128 overflow its annulus, within
 * 0.0125..0.1, moved a tenth back towards the value set here once both
 * kinds stay below 1/512, and held in between);
 * the value returned is simulation 0's, where its next pass starts (waits
 * for the handle's work). */
int sph_get_tuning(sph_sim* s, double out[5]);
/* Set them (same order and ranges: a grid the autotuner may choose, bracket
 * 0.005..0.1, split 8..64, group 16/32/64, sub-tiles 4/8); SPH_EINVAL
 * otherwise.  The bracket restarts from the given value in every
 * simulation.  Any time between steps. */
int sph_set_tuning(sph_sim* s, const double in[5]);

/* ---- serve loop: continuous batching (no reference counterpart) ----
 * A native worker thread steps every admitted job in one shared device
 * pool: each tick is one force pass over all resident jobs, which may be at
 * different steps of their lives; a job that ends frees its particle range
 * for the next one in the queue (FIFO, first fit).  A job's results equal
 * those of sph_create + sph_step(nsteps) with the same inputs.  All calls
 * are thread-safe. */
typedef struct sph_server sph_server;

#define SPH_JOB_QUEUED 0
#define SPH_JOB_RUNNING 1
#define SPH_JOB_DONE 2
#define SPH_JOB_FAILED 3

/* capacity: particle slots of the pool; max_jobs: jobs resident at once;
 * kmax: largest k of a job; precision and flags (SPH_FLAG_BITWISE only) are
 * shared by all jobs.  Uses the device current on the calling thread. */
int sph_server_create(int64_t capacity, int32_t max_jobs, int32_t kmax, int32_t precision,
                      uint32_t flags, sph_server** out);
/* Stops the worker; jobs still queued or running are abandoned. */
void sph_server_destroy(sph_server* s);
/* Queue a job of n particles for nsteps >= 1 steps (inputs are copied; vx,
 * vy, e may be NULL).  IDs are 0..n-1 in input order. */
int sph_server_submit(sph_server* s, const sph_params* p, int64_t n, const double* x,
                      const double* y, const double* vx, const double* vy, const double* e,
                      int32_t nsteps, int64_t* job_id);
/* Wait for a job (timeout_s < 0: no limit; SPH_ESTATE on timeout). */
int sph_server_wait(sph_server* s, int64_t job_id, double timeout_s);
/* Wait until no job is queued or running. */
int sph_server_drain(sph_server* s);
/* state (SPH_JOB_*); info = queue wait ms, run ms, submit-to-done latency ms
 * (-1 where not reached), steps completed. */
int sph_server_job_info(sph_server* s, int64_t job_id, int32_t* state, double info[4]);
/* One field of a finished job, indexed by ID: SPH_X, SPH_Y, SPH_VX, SPH_VY,
 * SPH_RHO, SPH_E or SPH_H. */
int sph_server_read_result(sph_server* s, int64_t job_id, int32_t field, double* out);
/* sph_reductions of a finished job, then TotalMomentum (|v| of the particle
 * stored last, as sph_total_momentum). */
int sph_server_job_reductions(sph_server* s, int64_t job_id, double out[7]);
/* Forget a finished or failed job and its results. */
int sph_server_release(sph_server* s, int64_t job_id);
/* jobs done, jobs failed, particle-steps done, seconds from the first submit
 * to the last completion, particle-steps/s over that span, p50 and p99
 * submit-to-done latency (ms), mean pool occupancy, ticks, jobs queued or
 * running now, mean tick duration (ms: uploads, one pass, readbacks) and
 * its device part (ms between events at the tick's start and end), then the
 * two parts of the latency: p50 and p99 of the queue wait (submit to
 * admission, ms) and of the run time (admission to done, ms). */
#define SPH_SERVER_NSTATS 16
int sph_server_stats(sph_server* s, double out[SPH_SERVER_NSTATS]);

const char* sph_last_error(void);
const char* sph_version(void);

#ifdef __cplusplus
}
#endif
#endif
