/*
 * CPU oracle for the sphugo SPH step.  TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's algorithm (Go package `sim`):
 *   sim/core.go:10-15,93-105,126-312       tree (Partition, Treebuild, BoundingSpheres)
 *   sim/nearest-neighbour.go:15-165        kNN (tree search, sentinel queue)
 *   sim/sph.go:64-198,237-435,441-463      kernels, density, force, DKD step, totals
 *   sim/linear-algebra.go:36-71            Vec2 helpers, DistSq, Dist (Hypot)
 *   sim/config-parser.go:111-149           defaults
 * compiled with -ffp-contract=off so every product and sum rounds once, as
 * Go's gc compiler does on amd64 (no FMA contraction, left-to-right order).
 *
 * Only the tests, the smoke check and the CPU baseline leg of the benchmark
 * load this library.  The product (libsphugo_amd.so) never does.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* kNN modes */
#define ORC_KNN_EXACT 0      /* exact kNN, (d2, ID, image) order, cell grid */
#define ORC_KNN_TREE 1       /* faithful tree search incl. merge bug + sentinel */
#define ORC_KNN_TREE_FIXED 2 /* faithful tree search, corrected sphere merge */
#define ORC_KNN_BRUTE 3      /* exact kNN, brute force over all 9 images */

#define ORC_KERNEL_MONAGHAN 0
#define ORC_KERNEL_WENDLAND 1

/* field numbering shared with the product's C ABI */
enum {
  ORC_X = 0, ORC_Y, ORC_VX, ORC_VY, ORC_RHO, ORC_C, ORC_E, ORC_EDOT,
  ORC_AX, ORC_AY, ORC_EPRED, ORC_VPX, ORC_VPY, ORC_H, ORC_NFIELDS
};

typedef struct {
  int32_t k;          /* NN_SIZE (sim/core.go:14), runtime here */
  int32_t kernel;     /* ORC_KERNEL_* (sim/sph.go:258,284) */
  int32_t knn_mode;   /* ORC_KNN_* */
  int32_t use_cfl;    /* 1: dt_half from the CFL condition each step */
  double gamma;       /* SphConfig.Gamma */
  double mass;        /* SphConfig.ParticleMass */
  double dt_half;     /* SphConfig.DeltaTHalf (used when use_cfl == 0) */
  double cfl;         /* CFL number (no reference counterpart) */
  double accel[2];    /* SphConfig.Acceleration */
  double hor[2];      /* SphConfig.HorPeriodicity, +-DBL_MAX = open */
  double vert[2];     /* SphConfig.VertPeriodicity */
  double refl[4];     /* SphConfig.Reflections L, R, U, D */
  double visc_alpha;  /* sim/sph.go:380 (0.75) */
  double visc_beta;   /* sim/sph.go:381 (1.5) */
  double visc_eta2;   /* sim/sph.go:382 (0.01) */
} orc_params;

typedef struct orc_sim orc_sim;
typedef struct orc_tree orc_tree;

/* reference defaults (sim/config-parser.go:131-149) + sim/sph.go:379-383 */
void orc_default_params(orc_params* p);

/* ---- unit-level entry points (pinned by the reference's own tests) ---- */
/* Partition (sim/core.go:126-164): permutes x,y in place, returns len(a). */
int64_t orc_partition(double* x, double* y, int64_t n, int32_t orientation,
                      double middle);
/* Kernel function values and prefactors (sim/sph.go:258-304). */
double orc_kernel_F(int32_t kernel, double q);
double orc_kernel_DF(int32_t kernel, double q);
double orc_kernel_Fpref(int32_t kernel);
double orc_kernel_DFpref(int32_t kernel);
double orc_hypot(double p, double q);

/* MakeCells (sim/core.go:93-105): tree over n particles with root [0,1]^2. */
orc_tree* orc_tree_make(const double* x, const double* y, int64_t n,
                        int32_t fixed_merge);
int64_t orc_tree_ncells(const orc_tree* t);
/* per cell: begin,end (particle slice), lower,upper (child ids or -1),
 * ll(x,y), ur(x,y), bcenter(x,y), bradius.  Arrays sized ncells. */
void orc_tree_cells(const orc_tree* t, int64_t* begin, int64_t* end,
                    int64_t* lower, int64_t* upper, double* box4,
                    double* sphere3);
/* particle positions in tree order */
void orc_tree_positions(const orc_tree* t, double* x, double* y);
void orc_tree_free(orc_tree* t);

/* ---- simulation ---- */
orc_sim* orc_create(const orc_params* p, int64_t n, const double* x,
                    const double* y, const double* vx, const double* vy,
                    const double* e);
void orc_destroy(orc_sim* s);
/* Sources (sim/sph.go:71-86): append m particles with the next IDs (vx, vy,
 * e may be NULL).  0 on success. */
int32_t orc_append(orc_sim* s, int64_t m, const double* x, const double* y,
                   const double* vx, const double* vy, const double* e);
void orc_set_threads(orc_sim* s, int32_t nthreads);
/* Simulation.Step() nsteps times (sim/sph.go:64-198). 0 on success. */
int32_t orc_step(orc_sim* s, int32_t nsteps);
/* Simulation.CalculateForces() on the current state (sim/sph.go:403-435). */
int32_t orc_calculate_forces(orc_sim* s);
/* read / write one field for all particles, indexed by particle ID */
void orc_get_field(const orc_sim* s, int32_t field, double* out);
void orc_set_field(orc_sim* s, int32_t field, const double* in);
/* neighbour IDs and squared distances of the last pass, indexed by ID,
 * each row sorted ascending by (d2, ID).  Returns number of nil entries. */
int64_t orc_get_neighbours(const orc_sim* s, int64_t* ids, double* d2);
/* per ID: sum over the neighbours of |term| of ax (which 0), ay (1) or
 * edot (2), with the same prefactor as the sum (last force pass) */
void orc_get_term_scale(const orc_sim* s, int32_t which, double* out);
/* NNPos of the last pass (p_j minus the query's image offset), rows
 * ascending by (d2, ID) as orc_get_neighbours */
void orc_get_neighbour_pos(const orc_sim* s, double* px, double* py);
/* sum_e, sum_rho, E_tot, Px, Py, last dt_half */
void orc_reductions(const orc_sim* s, double out[6]);
/* reference totals (sim/sph.go:441-463) */
double orc_total_energy(const orc_sim* s);
double orc_total_density(const orc_sim* s);
double orc_total_momentum(const orc_sim* s);
int64_t orc_current_step(const orc_sim* s);
/* CFL half step the next Step() will use (use_cfl == 1) */
double orc_cfl_dt_half(const orc_sim* s);
/* number of exact (d2) ties at the k-th boundary in the last exact pass */
int64_t orc_boundary_ties(const orc_sim* s);
const char* orc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
