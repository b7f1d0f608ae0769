/*
 * CPU oracle for the sphugo SPH step.  TEST INFRASTRUCTURE ONLY: loaded by
 * tests/, by the smoke check and by the benchmark's CPU baseline leg, never
 * by the product library.
 *
 * Restates the reference (Go package `sim`) line by line; every function
 * cites the file:line it follows.  Build with -ffp-contract=off (see
 * oracle/Makefile): Go's gc compiler on amd64 rounds every * and + on its
 * own and evaluates left to right, and so does this file.
 *
 * Parity status of each part (how it is pinned) is recorded in the test
 * modules tests/test_oracle_*.py and tests/test_golden.py.  orc_hypot is
 * held to what the mathematics fixes (tests/test_oracle_hypot.py: exact
 * triples, symmetry, exact power-of-two scaling, no overflow, Go's special
 * cases, 2 ulp against extended precision), but its last-bit rounding
 * against Go's own math.Hypot stays unpinned (only Go could produce those
 * vectors); it only steers the faithful tree's pruning, never the exact kNN
 * the device path is checked against.
 */
#include "sph_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static __thread char g_err[256];
static void set_err(const char* m) {
  snprintf(g_err, sizeof g_err, "%s", m);
}
const char* orc_last_error(void) { return g_err; }

/* sim/core.go:10-15 */
#define MAX_PARTICLES_PER_CELL 8
#define SPLIT_FRACTION 0.5

/* Go folds these constant expressions exactly and rounds once
 * (sim/sph.go:266,275,293,303); the hex values are the correctly rounded
 * results, one ulp away from a naive C evaluation for Wendland. */
#define MONAGHAN_PREF 0x1.5d3b3e3583243p+3  /* 6*40/(Pi*7) */
#define WENDLAND_FPREF 0x1.1d34a60108f72p+1 /* 4*7/(Pi*4) */
#define WENDLAND_DFPREF 0x1.1d34a60108f72p+2 /* 8*7/(Pi*4) */

/* ------------------------------------------------------------------ */
/* sim/linear-algebra.go:64-71 and Go's math.Hypot (pure-Go form:
 * p*Sqrt(1+(q/p)^2) with p >= q >= 0); only used for tree pruning. */
double orc_hypot(double p, double q) {
  p = fabs(p);
  q = fabs(q);
  if (isinf(p) || isinf(q)) return INFINITY;
  if (isnan(p) || isnan(q)) return NAN;
  if (p < q) { double t = p; p = q; q = t; }
  if (p == 0) return 0;
  q = q / p;
  return p * sqrt(1 + q * q);
}

/* ------------------------------------------------------------------ */
/* Kernels, sim/sph.go:258-304 */
double orc_kernel_F(int32_t kernel, double q) {
  if (kernel == ORC_KERNEL_WENDLAND) {
    /* sim/sph.go:285-291 */
    return (1 - q) * (1 - q) * (1 - q) * (1 - q) * (1 + 4 * q);
  }
  /* sim/sph.go:259-264 */
  if (q < 0.5) return q * q * q - q * q + 1.0 / 6;
  return (1 - q) * (1 - q) * (1 - q) / 3;
}

double orc_kernel_DF(int32_t kernel, double q) {
  if (kernel == ORC_KERNEL_WENDLAND) {
    /* sim/sph.go:295-301 */
    return -10 * q * (1 - q) * (1 - q) * (1 - q);
  }
  /* sim/sph.go:268-273 */
  if (q < 0.5) return (3 * q * q - 2 * q);
  return -(1 - q) * (1 - q);
}

double orc_kernel_Fpref(int32_t kernel) {
  return kernel == ORC_KERNEL_WENDLAND ? WENDLAND_FPREF : MONAGHAN_PREF;
}
double orc_kernel_DFpref(int32_t kernel) {
  return kernel == ORC_KERNEL_WENDLAND ? WENDLAND_DFPREF : MONAGHAN_PREF;
}

/* ------------------------------------------------------------------ */
/* Particle, sim/core.go:17-42.  The per-particle NN arrays live in side
 * tables indexed by the particle's stable ID, so that they travel with the
 * particle when the tree permutes the array (as the Go struct fields do). */
typedef struct {
  double x, y, vx, vy, rho, c, e, edot, ax, ay, ep, vpx, vpy;
  /* sum over the neighbour terms of |term| for ax, ay and edot, scaled as
   * the sums are (test measure: error per particle relative to the size of
   * the terms it cancels; not read by the step) */
  double sax, say, sedot;
  int64_t id;
} OP;

typedef struct {
  int64_t begin, end;      /* Particles slice */
  double llx, lly, urx, ury; /* LowerLeft, UpperRight */
  double bcx, bcy, brad;   /* BCenter, BRadius */
  int64_t lower, upper;    /* children, -1 = nil */
} OCell;

typedef struct {
  OCell* c;
  int64_t n, cap;
} CellVec;

static int64_t cell_new(CellVec* v) {
  if (v->n == v->cap) {
    v->cap = v->cap ? 2 * v->cap : 1024;
    v->c = (OCell*)realloc(v->c, (size_t)v->cap * sizeof(OCell));
  }
  memset(&v->c[v->n], 0, sizeof(OCell));
  v->c[v->n].lower = v->c[v->n].upper = -1;
  return v->n++;
}

static void swap_op(OP* a, OP* b) { OP t = *a; *a = *b; *b = t; }

/* Partition, sim/core.go:126-164.  Vertical (orientation 0) compares Y with
 * the i==j fix-up inside the loop; Horizontal compares X with it after. */
static int64_t partition_ops(OP* ps, int64_t n, int32_t orientation,
                             double middle) {
  int64_t i = 0, j = n - 1;
  if (orientation == 0) {
    while (i < j) {
      while (i < j && ps[i].y <= middle) i++;
      while (i < j && ps[j].y > middle) j--;
      if (ps[i].y > ps[j].y) swap_op(&ps[i], &ps[j]);
      if (i == j && middle > ps[i].y) i++;
    }
  } else {
    while (i < j) {
      while (i < j && ps[i].x <= middle) i++;
      while (i < j && ps[j].x > middle) j--;
      if (ps[i].x > ps[j].x) swap_op(&ps[i], &ps[j]);
    }
    if (i == j && middle > ps[i].x) i++;
  }
  return i;
}

int64_t orc_partition(double* x, double* y, int64_t n, int32_t orientation,
                      double middle) {
  OP* ps = (OP*)calloc((size_t)(n > 0 ? n : 1), sizeof(OP));
  for (int64_t i = 0; i < n; i++) { ps[i].x = x[i]; ps[i].y = y[i]; }
  int64_t a = partition_ops(ps, n, orientation, middle);
  for (int64_t i = 0; i < n; i++) { x[i] = ps[i].x; y[i] = ps[i].y; }
  free(ps);
  return a;
}

/* Treebuild, sim/core.go:172-224.  `root` persists across passes (the Go
 * Simulation keeps sim.Root); its children are reassigned only when the
 * corresponding side is non-empty (stale-child quirk, reported as an
 * error when it would bite). */
static int treebuild(CellVec* cv, OP* ps, int64_t ci, int32_t orientation) {
  OCell* r = &cv->c[ci];
  /* dirty fix, sim/core.go:176-178 */
  if (r->lly == r->ury && r->llx == r->urx) return 0;
  double mid;
  if (orientation == 0)
    mid = SPLIT_FRACTION * r->lly + (1 - SPLIT_FRACTION) * r->ury;
  else
    mid = SPLIT_FRACTION * r->llx + (1 - SPLIT_FRACTION) * r->urx;
  int64_t b0 = r->begin, n = r->end - r->begin;
  int64_t na = partition_ops(ps + b0, n, orientation, mid);
  int64_t nb = n - na;
  if (na > 0) {
    int64_t lo = cell_new(cv);
    r = &cv->c[ci];
    OCell* L = &cv->c[lo];
    L->begin = b0; L->end = b0 + na;
    L->llx = r->llx; L->lly = r->lly; L->urx = r->urx; L->ury = r->ury;
    if (orientation == 0) L->ury = mid; else L->urx = mid;
    r->lower = lo;
    if (na > MAX_PARTICLES_PER_CELL)
      if (treebuild(cv, ps, lo, 1 - orientation)) return -1;
  }
  if (nb > 0) {
    int64_t up = cell_new(cv);
    r = &cv->c[ci];
    OCell* U = &cv->c[up];
    U->begin = b0 + na; U->end = b0 + n;
    U->llx = r->llx; U->lly = r->lly; U->urx = r->urx; U->ury = r->ury;
    if (orientation == 0) U->lly = mid; else U->llx = mid;
    r->upper = up;
    if (nb > MAX_PARTICLES_PER_CELL)
      if (treebuild(cv, ps, up, 1 - orientation)) return -1;
  }
  return 0;
}

/* BoundingSpheres, sim/core.go:229-312.  fixed != 0 replaces the two-circle
 * merge (wrong when one child circle contains the other, :300-311) with the
 * correct enclosing circle; used only to attribute kNN disagreements. */
static void bounding_spheres(CellVec* cv, const OP* ps, int64_t ci,
                             int fixed) {
  OCell* r = &cv->c[ci];
  if (r->upper < 0 && r->lower < 0) {
    int64_t n = r->end - r->begin;
    const OP* P = ps + r->begin;
    if (n == 1) {
      r->bcx = P[0].x; r->bcy = P[0].y; r->brad = 0;
    } else {
      double dmax = 0.0, ax = 0, ay = 0, bx = 0, by = 0;
      for (int64_t a = 0; a < n; a++)
        for (int64_t b = 0; b < n; b++) {
          double xx = P[b].x - P[a].x, xy = P[b].y - P[a].y;
          double d = xx * xx + xy * xy;
          if (d > dmax) { dmax = d; ax = P[a].x; ay = P[a].y; bx = P[b].x; by = P[b].y; }
        }
      double rx = (bx - ax) * 0.5, ry = (by - ay) * 0.5;
      r->bcx = rx + ax; r->bcy = ry + ay;
      double r2 = rx * rx + ry * ry;
      for (int64_t a = 0; a < n; a++) {
        double xx = r->bcx - P[a].x, xy = r->bcy - P[a].y;
        double d = xx * xx + xy * xy;
        if (d > r2) r2 = d;
      }
      r->brad = sqrt(r2);
    }
    return;
  }
  if (r->upper >= 0) bounding_spheres(cv, ps, r->upper, fixed);
  if (r->lower >= 0) bounding_spheres(cv, ps, r->lower, fixed);
  r = &cv->c[ci];
  if (r->upper < 0) {
    r->brad = cv->c[r->lower].brad; r->bcx = cv->c[r->lower].bcx; r->bcy = cv->c[r->lower].bcy;
    return;
  }
  if (r->lower < 0) {
    r->brad = cv->c[r->upper].brad; r->bcx = cv->c[r->upper].bcx; r->bcy = cv->c[r->upper].bcy;
    return;
  }
  const OCell* U = &cv->c[r->upper];
  const OCell* L = &cv->c[r->lower];
  double abx = U->bcx - L->bcx, aby = U->bcy - L->bcy;
  double abn = sqrt(abx * abx + aby * aby);
  double rA = L->brad, rB = U->brad;
  if (fixed && abn + rA <= rB) { r->bcx = U->bcx; r->bcy = U->bcy; r->brad = rB; return; }
  if (fixed && abn + rB <= rA) { r->bcx = L->bcx; r->bcy = L->bcy; r->brad = rA; return; }
  double rC = (rA + rB + abn) * 0.5;
  double f = (rB - rC) / abn;
  double mx = abx * f, my = aby * f;
  r->bcx = mx + U->bcx; r->bcy = my + U->bcy;
  r->brad = rC;
}

/* ------------------------------------------------------------------ */
struct orc_tree {
  CellVec cv;
  OP* ps;
  int64_t n;
};

orc_tree* orc_tree_make(const double* x, const double* y, int64_t n,
                        int32_t fixed_merge) {
  orc_tree* t = (orc_tree*)calloc(1, sizeof(orc_tree));
  t->n = n;
  t->ps = (OP*)calloc((size_t)(n > 0 ? n : 1), sizeof(OP));
  for (int64_t i = 0; i < n; i++) { t->ps[i].x = x[i]; t->ps[i].y = y[i]; t->ps[i].id = i; }
  int64_t root = cell_new(&t->cv);
  OCell* r = &t->cv.c[root];
  r->begin = 0; r->end = n; r->llx = 0; r->lly = 0; r->urx = 1; r->ury = 1;
  treebuild(&t->cv, t->ps, root, 0); /* MakeCells, sim/core.go:93-105 */
  bounding_spheres(&t->cv, t->ps, root, fixed_merge);
  return t;
}
int64_t orc_tree_ncells(const orc_tree* t) { return t->cv.n; }
void orc_tree_cells(const orc_tree* t, int64_t* begin, int64_t* end,
                    int64_t* lower, int64_t* upper, double* box4,
                    double* sphere3) {
  for (int64_t i = 0; i < t->cv.n; i++) {
    const OCell* c = &t->cv.c[i];
    begin[i] = c->begin; end[i] = c->end; lower[i] = c->lower; upper[i] = c->upper;
    box4[4 * i] = c->llx; box4[4 * i + 1] = c->lly; box4[4 * i + 2] = c->urx; box4[4 * i + 3] = c->ury;
    sphere3[3 * i] = c->bcx; sphere3[3 * i + 1] = c->bcy; sphere3[3 * i + 2] = c->brad;
  }
}
void orc_tree_positions(const orc_tree* t, double* x, double* y) {
  for (int64_t i = 0; i < t->n; i++) { x[i] = t->ps[i].x; y[i] = t->ps[i].y; }
}
void orc_tree_free(orc_tree* t) {
  if (!t) return;
  free(t->cv.c); free(t->ps); free(t);
}

/* ------------------------------------------------------------------ */
struct orc_sim {
  orc_params p;
  int64_t n;
  int32_t k;
  OP* ps;          /* Root.Particles, permuted by the tree each pass */
  double* nnd;     /* [id*k + i] NNDists */
  int64_t* nni;    /* [id*k + i] NearestNeighbours as index into ps, -1 nil */
  double* nnpx;    /* [id*k + i] NNPos */
  double* nnpy;
  int32_t* nnimg;  /* [id*k + i] image code (ix+1)*3+(iy+1), exact modes */
  double* nnd2;    /* [id*k + i] squared distance key before the sqrt */
  CellVec cv;
  int64_t root;
  int64_t step;    /* CurrentStep */
  double dt_used;
  int32_t nthreads;
  int64_t ties;
  int64_t* pos_of_id; /* scratch: ps index of each id */
};

void orc_default_params(orc_params* p) {
  memset(p, 0, sizeof *p);
  p->k = 32;
  p->kernel = ORC_KERNEL_MONAGHAN;
  p->knn_mode = ORC_KNN_EXACT;
  p->use_cfl = 0;
  p->gamma = 1.66666;
  p->mass = 1;
  p->dt_half = 0.001;
  p->cfl = 0.3;
  p->hor[0] = -DBL_MAX; p->hor[1] = DBL_MAX;
  p->vert[0] = -DBL_MAX; p->vert[1] = DBL_MAX;
  p->refl[0] = -DBL_MAX; p->refl[1] = DBL_MAX;
  p->refl[2] = -DBL_MAX; p->refl[3] = DBL_MAX;
  p->visc_alpha = 0.75; p->visc_beta = 1.5; p->visc_eta2 = 0.01;
}

orc_sim* orc_create(const orc_params* p, int64_t n, const double* x,
                    const double* y, const double* vx, const double* vy,
                    const double* e) {
  if (n <= 0 || p->k <= 0) { set_err("orc_create: n and k must be positive"); return NULL; }
  orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
  s->p = *p;
  s->n = n;
  s->k = p->k;
  s->nthreads = 1;
  s->ps = (OP*)calloc((size_t)n, sizeof(OP));
  size_t nk = (size_t)n * (size_t)p->k;
  s->nnd = (double*)calloc(nk, sizeof(double));
  s->nni = (int64_t*)malloc(nk * sizeof(int64_t));
  for (size_t i = 0; i < nk; i++) s->nni[i] = -1;
  s->nnpx = (double*)calloc(nk, sizeof(double));
  s->nnpy = (double*)calloc(nk, sizeof(double));
  s->nnimg = (int32_t*)calloc(nk, sizeof(int32_t));
  s->nnd2 = (double*)calloc(nk, sizeof(double));
  s->pos_of_id = (int64_t*)malloc((size_t)n * sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) {
    OP* q = &s->ps[i];
    q->x = x[i]; q->y = y[i];
    q->vx = vx ? vx[i] : 0; q->vy = vy ? vy[i] : 0;
    q->e = e ? e[i] : 0;
    q->id = i;
  }
  /* MakeCells(ps, Vertical): root [0,1]^2 (sim/core.go:93-105) */
  s->root = cell_new(&s->cv);
  OCell* r = &s->cv.c[s->root];
  r->begin = 0; r->end = n; r->llx = 0; r->lly = 0; r->urx = 1; r->ury = 1;
  if (p->knn_mode == ORC_KNN_TREE || p->knn_mode == ORC_KNN_TREE_FIXED) {
    if (treebuild(&s->cv, s->ps, s->root, 0)) { orc_destroy(s); return NULL; }
    bounding_spheres(&s->cv, s->ps, s->root, p->knn_mode == ORC_KNN_TREE_FIXED);
  }
  return s;
}

/* Sources, sim/sph.go:71-86: the particles a spawner produced are appended
 * to Root.Particles (next IDs; velocity and energy as given, every other
 * field zero, as the Go zero value), and the root is rebuilt by MakeCells
 * (sim/core.go:93-105): a fresh cell over [0,1]^2 without children. */
int32_t orc_append(orc_sim* s, int64_t m, const double* x, const double* y,
                   const double* vx, const double* vy, const double* e) {
  if (m < 0 || (m > 0 && (!x || !y))) { set_err("orc_append: bad arguments"); return -1; }
  if (m == 0) return 0;
  const int64_t n = s->n, n2 = n + m;
  const size_t k = (size_t)s->k, nk = (size_t)n * k, nk2 = (size_t)n2 * k;
  OP* ps = (OP*)realloc(s->ps, (size_t)n2 * sizeof(OP));
  double* nnd = (double*)realloc(s->nnd, nk2 * sizeof(double));
  int64_t* nni = (int64_t*)realloc(s->nni, nk2 * sizeof(int64_t));
  double* nnpx = (double*)realloc(s->nnpx, nk2 * sizeof(double));
  double* nnpy = (double*)realloc(s->nnpy, nk2 * sizeof(double));
  int32_t* nnimg = (int32_t*)realloc(s->nnimg, nk2 * sizeof(int32_t));
  double* nnd2 = (double*)realloc(s->nnd2, nk2 * sizeof(double));
  int64_t* poi = (int64_t*)realloc(s->pos_of_id, (size_t)n2 * sizeof(int64_t));
  if (ps) s->ps = ps;
  if (nnd) s->nnd = nnd;
  if (nni) s->nni = nni;
  if (nnpx) s->nnpx = nnpx;
  if (nnpy) s->nnpy = nnpy;
  if (nnimg) s->nnimg = nnimg;
  if (nnd2) s->nnd2 = nnd2;
  if (poi) s->pos_of_id = poi;
  if (!ps || !nnd || !nni || !nnpx || !nnpy || !nnimg || !nnd2 || !poi) {
    set_err("orc_append: out of memory");
    return -1;
  }
  memset(&s->ps[n], 0, (size_t)m * sizeof(OP));
  for (int64_t i = 0; i < m; i++) {
    OP* q = &s->ps[n + i];
    q->x = x[i]; q->y = y[i];
    q->vx = vx ? vx[i] : 0; q->vy = vy ? vy[i] : 0;
    q->e = e ? e[i] : 0;
    q->id = n + i;
  }
  memset(s->nnd + nk, 0, (nk2 - nk) * sizeof(double));
  for (size_t i = nk; i < nk2; i++) s->nni[i] = -1;
  memset(s->nnpx + nk, 0, (nk2 - nk) * sizeof(double));
  memset(s->nnpy + nk, 0, (nk2 - nk) * sizeof(double));
  memset(s->nnimg + nk, 0, (nk2 - nk) * sizeof(int32_t));
  memset(s->nnd2 + nk, 0, (nk2 - nk) * sizeof(double));
  s->n = n2;
  s->cv.n = 1;
  OCell* r = &s->cv.c[s->root];
  r->begin = 0; r->end = n2; r->llx = 0; r->lly = 0; r->urx = 1; r->ury = 1;
  r->lower = r->upper = -1;
  return 0;
}

void orc_destroy(orc_sim* s) {
  if (!s) return;
  free(s->ps); free(s->nnd); free(s->nni); free(s->nnpx); free(s->nnpy);
  free(s->nnimg); free(s->nnd2); free(s->cv.c); free(s->pos_of_id); free(s);
}

void orc_set_threads(orc_sim* s, int32_t nthreads) {
  s->nthreads = nthreads < 1 ? 1 : nthreads;
}

/* ------------------------------------------------------------------ */
/* Tree kNN, sim/nearest-neighbour.go:28-165 */

/* NNQueueInsert, sim/nearest-neighbour.go:139-153 */
static void nn_insert(orc_sim* s, int64_t id, double dist, int64_t nb,
                      double rpx, double rpy, int32_t img, double d2) {
  int k = s->k;
  double* D = s->nnd + id * k;
  int64_t* I = s->nni + id * k;
  double* PX = s->nnpx + id * k;
  double* PY = s->nnpy + id * k;
  int32_t* IM = s->nnimg + id * k;
  double* D2 = s->nnd2 + id * k;
  int i = 1;
  for (; i < k && D[i] > dist; i++) {
    D[i - 1] = D[i]; I[i - 1] = I[i]; PX[i - 1] = PX[i]; PY[i - 1] = PY[i];
    IM[i - 1] = IM[i]; D2[i - 1] = D2[i];
  }
  D[i - 1] = dist; I[i - 1] = nb; PX[i - 1] = rpx; PY[i - 1] = rpy;
  IM[i - 1] = img; D2[i - 1] = d2;
}

/* findNNRec, sim/nearest-neighbour.go:70-119 */
static void find_nn_rec(orc_sim* s, int64_t pi, int64_t ci, double ox,
                        double oy, int32_t img) {
  const OCell* r = &s->cv.c[ci];
  OP* me = &s->ps[pi];
  int64_t id = me->id;
  double px = me->x + ox, py = me->y + oy;
  if (r->upper < 0 && r->lower < 0) {
    for (int64_t j = r->begin; j < r->end; j++) {
      double dx = px - s->ps[j].x, dy = py - s->ps[j].y;
      double d2 = dx * dx + dy * dy;
      if (d2 < s->nnd[id * s->k] && pi != j)
        nn_insert(s, id, d2, j, s->ps[j].x - ox, s->ps[j].y - oy, img, d2);
    }
    return;
  }
  if (r->upper >= 0 && r->lower >= 0) {
    const OCell* U = &s->cv.c[r->upper];
    const OCell* L = &s->cv.c[r->lower];
    double du = orc_hypot(U->bcx - px, U->bcy - py);
    double dl = orc_hypot(L->bcx - px, L->bcy - py);
    double maxd = sqrt(s->nnd[id * s->k]);
    int64_t up = r->upper, lo = r->lower;
    if (dl < du) {
      if (dl - L->brad < maxd) find_nn_rec(s, pi, lo, ox, oy, img);
      if (du - U->brad < maxd) find_nn_rec(s, pi, up, ox, oy, img);
    } else {
      if (du - U->brad < maxd) find_nn_rec(s, pi, up, ox, oy, img);
      if (dl - L->brad < maxd) find_nn_rec(s, pi, lo, ox, oy, img);
    }
    return;
  }
  if (r->upper >= 0) find_nn_rec(s, pi, r->upper, ox, oy, img);
  if (r->lower >= 0) find_nn_rec(s, pi, r->lower, ox, oy, img);
}

/* image range of one axis, sim/nearest-neighbour.go:32-55 */
static int image_range(const double per[2], int* lo, int* hi, double* delta) {
  *lo = -1; *hi = 1; *delta = per[1] - per[0];
  if (per[0] == -DBL_MAX) {
    *lo = 0; *hi = 0; *delta = 0;
    if (per[1] != DBL_MAX) return -1;
  }
  return 0;
}

/* FindNearestNeighboursPeriodic, sim/nearest-neighbour.go:28-67 */
static int knn_tree_one(orc_sim* s, int64_t pi) {
  int64_t id = s->ps[pi].id;
  int k = s->k;
  /* NNQueueInitSentinel, :155-165 (keys = previous sqrt distances + 0.4) */
  for (int i = 0; i < k; i++) s->nnd[id * k + i] += 0.4;
  int is, ie, js, je;
  double dx, dy;
  if (image_range(s->p.hor, &is, &ie, &dx) || image_range(s->p.vert, &js, &je, &dy)) {
    set_err("cannot have open and periodic boundary at same time");
    return -1;
  }
  for (int i = is; i <= ie; i++)
    for (int j = js; j <= je; j++)
      find_nn_rec(s, pi, s->root, (double)i * dx, (double)j * dy, (i + 1) * 3 + (j + 1));
  for (int i = 0; i < k; i++) s->nnd[id * k + i] = sqrt(s->nnd[id * k + i]);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Exact kNN: the k smallest (d2, ID, image) over every image of every other
 * particle, with d2 computed by the reference formula
 * ((p_i + o) - p_j)^2 (sim/nearest-neighbour.go:59,72,76).  Stored in the
 * reference's queue order: index 0 = k-th (largest), descending. */
typedef struct { double d2; int64_t id; int32_t img; int64_t j; } Cand;

static int cand_less(const Cand* a, const Cand* b) {
  if (a->d2 != b->d2) return a->d2 < b->d2;
  if (a->id != b->id) return a->id < b->id;
  return a->img < b->img;
}

/* bounded list of the k best, kept sorted ascending */
static void best_insert(Cand* best, int* nb, int k, const Cand* c) {
  if (*nb == k && !cand_less(c, &best[k - 1])) return;
  int i = (*nb < k) ? (*nb)++ : k - 1;
  while (i > 0 && cand_less(c, &best[i - 1])) { best[i] = best[i - 1]; i--; }
  best[i] = *c;
}

typedef struct {
  int gx, gy;          /* cells per axis */
  double x0, y0, lx, ly; /* grid box */
  double ivx, ivy;     /* cells per unit length */
  int perx, pery;      /* periodic axes */
  int64_t* start;      /* gx*gy + 1 */
  int64_t* idx;        /* ps indices by cell */
  int32_t* wx;         /* per ps index: wrap count of its unwrapped cell */
  int32_t* wy;
} OGrid;

static int64_t floordiv(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) q--;
  return q;
}

static void grid_build(orc_sim* s, OGrid* g) {
  int64_t n = s->n;
  g->perx = s->p.hor[0] != -DBL_MAX;
  g->pery = s->p.vert[0] != -DBL_MAX;
  double bx0 = 0, bx1 = 0, by0 = 0, by1 = 0;
  if (!g->perx || !g->pery) {
    bx0 = bx1 = s->ps[0].x; by0 = by1 = s->ps[0].y;
    for (int64_t i = 1; i < n; i++) {
      if (s->ps[i].x < bx0) bx0 = s->ps[i].x;
      if (s->ps[i].x > bx1) bx1 = s->ps[i].x;
      if (s->ps[i].y < by0) by0 = s->ps[i].y;
      if (s->ps[i].y > by1) by1 = s->ps[i].y;
    }
  }
  if (g->perx) { g->x0 = s->p.hor[0]; g->lx = s->p.hor[1] - s->p.hor[0]; }
  else { g->x0 = bx0; g->lx = (bx1 - bx0) > 0 ? (bx1 - bx0) * (1 + 1e-9) : 1; }
  if (g->pery) { g->y0 = s->p.vert[0]; g->ly = s->p.vert[1] - s->p.vert[0]; }
  else { g->y0 = by0; g->ly = (by1 - by0) > 0 ? (by1 - by0) * (1 + 1e-9) : 1; }
  int side = (int)floor(sqrt((double)n / 2.0));
  if (side < 1) side = 1;
  if (side > 4096) side = 4096;
  g->gx = g->gy = side;
  g->ivx = g->gx / g->lx;
  g->ivy = g->gy / g->ly;
  int64_t nc = (int64_t)g->gx * g->gy;
  g->start = (int64_t*)calloc((size_t)nc + 1, sizeof(int64_t));
  g->idx = (int64_t*)malloc((size_t)n * sizeof(int64_t));
  g->wx = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  g->wy = (int32_t*)malloc((size_t)n * sizeof(int32_t));
  int64_t* cell = (int64_t*)malloc((size_t)n * sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) {
    int64_t ux = (int64_t)floor((s->ps[i].x - g->x0) * g->ivx);
    int64_t uy = (int64_t)floor((s->ps[i].y - g->y0) * g->ivy);
    int64_t cx, cy;
    if (g->perx) { g->wx[i] = (int32_t)floordiv(ux, g->gx); cx = ux - (int64_t)g->wx[i] * g->gx; }
    else { g->wx[i] = 0; cx = ux < 0 ? 0 : (ux >= g->gx ? g->gx - 1 : ux); }
    if (g->pery) { g->wy[i] = (int32_t)floordiv(uy, g->gy); cy = uy - (int64_t)g->wy[i] * g->gy; }
    else { g->wy[i] = 0; cy = uy < 0 ? 0 : (uy >= g->gy ? g->gy - 1 : uy); }
    cell[i] = cy * g->gx + cx;
    g->start[cell[i] + 1]++;
  }
  for (int64_t c = 0; c < nc; c++) g->start[c + 1] += g->start[c];
  int64_t* fill = (int64_t*)malloc((size_t)nc * sizeof(int64_t));
  memcpy(fill, g->start, (size_t)nc * sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) g->idx[fill[cell[i]]++] = i;
  free(fill);
  free(cell);
}

static void grid_free(OGrid* g) { free(g->start); free(g->idx); free(g->wx); free(g->wy); }

static void visit_cell(const orc_sim* s, const OGrid* g, int64_t pi, int64_t ucx,
                       int64_t ucy, Cand* best, int* nb, double dx, double dy) {
  int64_t cx, cy, tx = 0, ty = 0;
  if (g->perx) { tx = floordiv(ucx, g->gx); cx = ucx - tx * g->gx; }
  else { if (ucx < 0 || ucx >= g->gx) return; cx = ucx; }
  if (g->pery) { ty = floordiv(ucy, g->gy); cy = ucy - ty * g->gy; }
  else { if (ucy < 0 || ucy >= g->gy) return; cy = ucy; }
  int64_t c = cy * g->gx + cx;
  const OP* me = &s->ps[pi];
  for (int64_t t = g->start[c]; t < g->start[c + 1]; t++) {
    int64_t j = g->idx[t];
    if (j == pi) continue; /* self excluded at every image */
    int64_t nx = g->perx ? tx - g->wx[j] : 0;
    int64_t ny = g->pery ? ty - g->wy[j] : 0;
    if (nx < -1 || nx > 1 || ny < -1 || ny > 1) continue;
    double ox = (double)(-nx) * dx, oy = (double)(-ny) * dy;
    double ex = (me->x + ox) - s->ps[j].x, ey = (me->y + oy) - s->ps[j].y;
    Cand cd;
    cd.d2 = ex * ex + ey * ey;
    cd.id = s->ps[j].id;
    cd.img = (int32_t)((-nx + 1) * 3 + (-ny + 1));
    cd.j = j;
    best_insert(best, nb, s->k, &cd);
  }
}

static void store_exact(orc_sim* s, int64_t pi, const Cand* best, int nb) {
  int k = s->k;
  int64_t id = s->ps[pi].id;
  double dxp, dyp;
  int a, b;
  image_range(s->p.hor, &a, &b, &dxp);
  image_range(s->p.vert, &a, &b, &dyp);
  for (int i = 0; i < k; i++) {
    /* descending: queue slot i holds the (k-1-i)-th smallest */
    int src = k - 1 - i;
    if (src >= nb) { s->nni[id * k + i] = -1; s->nnd[id * k + i] = 0; continue; }
    const Cand* c = &best[src];
    int ix = c->img / 3 - 1, iy = c->img % 3 - 1;
    double ox = (double)ix * dxp, oy = (double)iy * dyp;
    s->nnd2[id * k + i] = c->d2;
    s->nnd[id * k + i] = sqrt(c->d2);
    s->nni[id * k + i] = c->j;
    s->nnimg[id * k + i] = c->img;
    s->nnpx[id * k + i] = s->ps[c->j].x - ox;
    s->nnpy[id * k + i] = s->ps[c->j].y - oy;
  }
}

static int knn_exact_all(orc_sim* s, int brute) {
  int k = s->k;
  int64_t n = s->n;
  double dx, dy;
  int is, ie, js, je;
  if (image_range(s->p.hor, &is, &ie, &dx) || image_range(s->p.vert, &js, &je, &dy)) {
    set_err("cannot have open and periodic boundary at same time");
    return -1;
  }
  int64_t navail = (int64_t)(ie - is + 1) * (je - js + 1) * (n - 1);
  if (navail < k) { set_err("exact kNN: fewer than k candidates"); return -1; }
  OGrid g;
  memset(&g, 0, sizeof g);
  if (!brute) grid_build(s, &g);
  int64_t ties = 0;
  int fail = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(s->nthreads) reduction(+ : ties)
  for (int64_t pi = 0; pi < n; pi++) {
    Cand* best = (Cand*)malloc((size_t)(k + 1) * sizeof(Cand));
    int nb = 0;
    const OP* me = &s->ps[pi];
    if (brute) {
      for (int64_t j = 0; j < n; j++) {
        if (j == pi) continue;
        for (int i = is; i <= ie; i++)
          for (int jj = js; jj <= je; jj++) {
            double ox = (double)i * dx, oy = (double)jj * dy;
            double ex = (me->x + ox) - s->ps[j].x, ey = (me->y + oy) - s->ps[j].y;
            Cand cd;
            cd.d2 = ex * ex + ey * ey;
            cd.id = s->ps[j].id;
            cd.img = (i + 1) * 3 + (jj + 1);
            cd.j = j;
            best_insert(best, &nb, k, &cd);
          }
      }
    } else {
      int64_t qx = (int64_t)floor((me->x - g.x0) * g.ivx);
      int64_t qy = (int64_t)floor((me->y - g.y0) * g.ivy);
      double sx = g.lx / g.gx, sy = g.ly / g.gy;
      double smin = sx < sy ? sx : sy;
      int64_t rmax = 2 * (g.gx > g.gy ? g.gx : g.gy) + 2;
      for (int64_t r = 0; r <= rmax; r++) {
        if (r == 0) visit_cell(s, &g, pi, qx, qy, best, &nb, dx, dy);
        else {
          for (int64_t t = -r; t <= r; t++) {
            visit_cell(s, &g, pi, qx + t, qy - r, best, &nb, dx, dy);
            visit_cell(s, &g, pi, qx + t, qy + r, best, &nb, dx, dy);
          }
          for (int64_t t = -r + 1; t <= r - 1; t++) {
            visit_cell(s, &g, pi, qx - r, qy + t, best, &nb, dx, dy);
            visit_cell(s, &g, pi, qx + r, qy + t, best, &nb, dx, dy);
          }
        }
        /* every unvisited cell lies farther than r*smin from the query */
        double lim = (double)r * smin * (1 - 1e-9);
        if (nb == k && best[k - 1].d2 < lim * lim) break;
      }
    }
    if (nb < k) {
#pragma omp atomic write
      fail = 1;
    }
    if (nb == k && k >= 2 && best[k - 1].d2 == best[k - 2].d2) ties++;
    store_exact(s, pi, best, nb);
    free(best);
  }
  if (!brute) grid_free(&g);
  s->ties = ties;
  if (fail) { set_err("exact kNN: fewer than k neighbours found"); return -1; }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Density2D, sim/sph.go:306-323 */
static double density_one(const orc_sim* s, const OP* p) {
  int k = s->k;
  const double* D = s->nnd + p->id * k;
  double maxR = D[0];
  double acc = 0.0;
  for (int i = 0; i < k; i++) {
    double x = D[i] / maxR;
    acc += orc_kernel_F(s->p.kernel, x);
  }
  return orc_kernel_Fpref(s->p.kernel) * s->p.mass * acc / (maxR * maxR);
}

/* AccelerationAndEDot2D, sim/sph.go:327-401 (alpha, beta, eta^2 are
 * parameters here; the reference hard-codes 0.75, 1.5, 0.01) */
static void force_one(orc_sim* s, OP* p) {
  int k = s->k;
  double gamma = s->p.gamma;
  const double* D = s->nnd + p->id * k;
  const int64_t* I = s->nni + p->id * k;
  const double* PX = s->nnpx + p->id * k;
  const double* PY = s->nnpy + p->id * k;
  double maxR = D[0];
  double contributionA = p->c * p->c / (gamma * p->rho);
  double contributionB = 0.0;
  double dRKernel = 0.0;
  double acc_ax = 0.0, acc_ay = 0.0, acc_edot = 0.0;
  double abs_ax = 0.0, abs_ay = 0.0, abs_edot = 0.0;
  double alpha = s->p.visc_alpha, beta = s->p.visc_beta, etaSq = s->p.visc_eta2;
  for (int i = 0; i < k; i++) {
    if (I[i] < 0) break; /* nil neighbour, :346-349 */
    const OP* nn = &s->ps[I[i]];
    double q = D[i] / maxR;
    dRKernel = orc_kernel_DF(s->p.kernel, q);
    contributionB = nn->c * nn->c / (gamma * nn->rho);
    double vABx = nn->vpx - p->vpx, vABy = nn->vpy - p->vpy;
    double rABx = PX[i] - p->x, rABy = PY[i] - p->y;
    double dot = vABx * rABx + vABy * rABy;
    double piAB = 0.0;
    if (dot < 0) {
      double cAB = 0.5 * (p->c + nn->c);
      double rhoAB = 0.5 * (p->rho + nn->rho);
      double hAB = 0.5 * (D[0] + s->nnd[nn->id * k]);
      double muAB = dot * hAB / ((rABx * rABx + rABy * rABy) + etaSq);
      piAB = (-alpha * cAB * muAB + beta * muAB * muAB) / rhoAB;
    }
    const double tx = rABx * (piAB + contributionA + contributionB) * dRKernel / D[i];
    const double ty = rABy * (piAB + contributionA + contributionB) * dRKernel / D[i];
    acc_ax += tx;
    acc_ay += ty;
    acc_edot += dot * dRKernel;
    abs_ax += fabs(tx);
    abs_ay += fabs(ty);
    abs_edot += fabs(dot * dRKernel);
  }
  double f = s->p.mass * orc_kernel_DFpref(s->p.kernel) / (maxR * maxR * maxR);
  p->ax = acc_ax * f + s->p.accel[0];
  p->ay = acc_ay * f + s->p.accel[1];
  p->edot = contributionA * acc_edot * s->p.mass;
  p->sax = abs_ax * fabs(f);
  p->say = abs_ay * fabs(f);
  p->sedot = fabs(contributionA * s->p.mass) * abs_edot;
}

/* CalculateForces, sim/sph.go:403-435 */
int32_t orc_calculate_forces(orc_sim* s) {
  int mode = s->p.knn_mode;
  int64_t n = s->n;
  if (mode == ORC_KNN_TREE || mode == ORC_KNN_TREE_FIXED) {
    /* Treebuild + BoundingSpheres on the persistent root, :406-408 */
    s->cv.n = 1;
    OCell* r = &s->cv.c[s->root];
    int had_lo = r->lower >= 0, had_up = r->upper >= 0;
    r->lower = r->upper = -1;
    if (treebuild(&s->cv, s->ps, s->root, 0)) return -1;
    r = &s->cv.c[s->root];
    if ((had_lo && r->lower < 0) || (had_up && r->upper < 0)) {
      /* Go would keep the previous pass's child (sim/core.go:189-223) */
      set_err("tree: stale root child quirk is not reproduced");
      return -1;
    }
    bounding_spheres(&s->cv, s->ps, s->root, mode == ORC_KNN_TREE_FIXED);
    int fail = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(s->nthreads)
    for (int64_t i = 0; i < n; i++)
      if (knn_tree_one(s, i)) {
#pragma omp atomic write
        fail = 1;
      }
    if (fail) return -1;
  } else {
    if (knn_exact_all(s, mode == ORC_KNN_BRUTE)) return -1;
  }
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
  for (int64_t i = 0; i < n; i++) s->ps[i].rho = density_one(s, &s->ps[i]);
  double factor = s->p.gamma * (s->p.gamma - 1);
  for (int64_t i = 0; i < n; i++) s->ps[i].c = sqrt(factor * s->ps[i].ep);
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
  for (int64_t i = 0; i < n; i++) force_one(s, &s->ps[i]);
  return 0;
}

/* CFL half step (no reference counterpart):
 *   dt_half = 0.5 * (cfl * min_i (0.5*h_i) / (c_i + |v_i|)),
 * h_i the k-th neighbour distance and c_i of the last force pass, v_i the
 * current velocity.  min is order independent. */
double orc_cfl_dt_half(const orc_sim* s) {
  double m = INFINITY;
  for (int64_t i = 0; i < s->n; i++) {
    const OP* p = &s->ps[i];
    double h = s->nnd[p->id * s->k];
    if (!(h > 0)) continue; /* appended since the last pass: no neighbours yet */
    double v = sqrt(p->vx * p->vx + p->vy * p->vy);
    double t = (0.5 * h) / (p->c + v);
    if (t < m) m = t;
  }
  return 0.5 * (s->p.cfl * m);
}

/* Step, sim/sph.go:64-198 (a source's particles come in by orc_append
 * before the step, sim/sph.go:71-86) */
int32_t orc_step(orc_sim* s, int32_t nsteps) {
  int64_t n = s->n;
  for (int32_t st = 0; st < nsteps; st++) {
    if (s->step == 0) {
      for (int64_t i = 0; i < n; i++) {
        s->ps[i].vpx = s->ps[i].vx; s->ps[i].vpy = s->ps[i].vy;
        s->ps[i].ep = s->ps[i].e;
      }
      if (orc_calculate_forces(s)) return -1;
    }
    double dtHalf = s->p.use_cfl ? orc_cfl_dt_half(s) : s->p.dt_half;
    s->dt_used = dtHalf;
    for (int64_t i = 0; i < n; i++) { /* drift 1, :108-117 */
      OP* p = &s->ps[i];
      double vdx = p->vx * dtHalf, vdy = p->vy * dtHalf;
      p->x = p->x + vdx; p->y = p->y + vdy;
      double adx = p->ax * dtHalf, ady = p->ay * dtHalf;
      p->vpx = p->vx + adx; p->vpy = p->vy + ady;
      p->ep = p->e + p->edot * dtHalf;
    }
    if (orc_calculate_forces(s)) return -1;
    for (int64_t i = 0; i < n; i++) { /* kick, :122-127 */
      OP* p = &s->ps[i];
      double adx = p->ax * (2 * dtHalf), ady = p->ay * (2 * dtHalf);
      p->vx = p->vx + adx; p->vy = p->vy + ady;
      p->e = p->e + p->edot * 2 * dtHalf;
    }
    for (int64_t i = 0; i < n; i++) { /* drift 2, :130-135 */
      OP* p = &s->ps[i];
      double vdx = p->vx * dtHalf, vdy = p->vy * dtHalf;
      p->x = p->x + vdx; p->y = p->y + vdy;
    }
    const double* H = s->p.hor;
    const double* V = s->p.vert;
    for (int64_t i = 0; i < n; i++) { /* periodic wrap, :147-167 */
      OP* p = &s->ps[i];
      if (p->x < H[0]) { p->x += (H[1] - H[0]); continue; }
      if (p->x > H[1]) { p->x -= (H[1] - H[0]); continue; }
      if (p->y < V[0]) { p->y += (V[1] - V[0]); continue; }
      if (p->y > V[1]) { p->y -= (V[1] - V[0]); }
    }
    const double* R = s->p.refl;
    for (int64_t i = 0; i < n; i++) { /* reflection, :170-193 */
      OP* p = &s->ps[i];
      if (p->x < R[0]) { p->x -= p->x - R[0]; p->vx = -p->vx; }
      if (p->x > R[1]) { p->x -= p->x - R[1]; p->vx = -p->vx; }
      if (p->y < R[2]) { p->y -= p->y - R[2]; p->vy = -p->vy; }
      if (p->y > R[3]) { p->y -= p->y - R[3]; p->vy = -p->vy; }
    }
    s->step += 1;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
static double* field_ptr(OP* p, int32_t f) {
  switch (f) {
    case ORC_X: return &p->x;
    case ORC_Y: return &p->y;
    case ORC_VX: return &p->vx;
    case ORC_VY: return &p->vy;
    case ORC_RHO: return &p->rho;
    case ORC_C: return &p->c;
    case ORC_E: return &p->e;
    case ORC_EDOT: return &p->edot;
    case ORC_AX: return &p->ax;
    case ORC_AY: return &p->ay;
    case ORC_EPRED: return &p->ep;
    case ORC_VPX: return &p->vpx;
    case ORC_VPY: return &p->vpy;
    default: return NULL;
  }
}

void orc_get_field(const orc_sim* s, int32_t field, double* out) {
  for (int64_t i = 0; i < s->n; i++) {
    OP* p = &s->ps[i];
    if (field == ORC_H) out[p->id] = s->nnd[p->id * s->k];
    else out[p->id] = *field_ptr(p, field);
  }
}

void orc_set_field(orc_sim* s, int32_t field, const double* in) {
  for (int64_t i = 0; i < s->n; i++) {
    OP* p = &s->ps[i];
    if (field == ORC_H) s->nnd[p->id * s->k] = in[p->id];
    else *field_ptr(p, field) = in[p->id];
  }
}

static int cmp_pair(const void* a, const void* b) {
  const double* x = (const double*)a;
  const double* y = (const double*)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
  return 0;
}

int64_t orc_get_neighbours(const orc_sim* s, int64_t* ids, double* d2) {
  int k = s->k;
  int64_t nil = 0;
  double* tmp = (double*)malloc((size_t)k * 2 * sizeof(double));
  for (int64_t i = 0; i < s->n; i++) {
    int64_t id = s->ps[i].id;
    int m = 0;
    for (int t = 0; t < k; t++) {
      int64_t j = s->nni[id * k + t];
      if (j < 0) { nil++; continue; }
      double dd = s->nnd2[id * k + t]; /* squared key at insertion */
      tmp[2 * m] = dd;
      tmp[2 * m + 1] = (double)s->ps[j].id;
      m++;
    }
    qsort(tmp, (size_t)m, 2 * sizeof(double), cmp_pair);
    for (int t = 0; t < k; t++) {
      ids[id * k + t] = t < m ? (int64_t)tmp[2 * t + 1] : -1;
      if (d2) d2[id * k + t] = t < m ? tmp[2 * t] : NAN;
    }
  }
  free(tmp);
  return nil;
}

void orc_get_term_scale(const orc_sim* s, int32_t which, double* out) {
  for (int64_t i = 0; i < s->n; i++) {
    const OP* p = &s->ps[i];
    out[p->id] = which == 0 ? p->sax : (which == 1 ? p->say : p->sedot);
  }
}

/* NNPos (sim/nearest-neighbour.go:81,152, the neighbour's position seen through
 * the query's image offset, p_j - o), rows in the order of
 * orc_get_neighbours: ascending by (d2, ID). */
typedef struct {
  double d2, id, px, py;
} NbrPos;
static int cmp_nbrpos(const void* a, const void* b) {
  const NbrPos* x = (const NbrPos*)a;
  const NbrPos* y = (const NbrPos*)b;
  if (x->d2 != y->d2) return x->d2 < y->d2 ? -1 : 1;
  if (x->id != y->id) return x->id < y->id ? -1 : 1;
  return 0;
}

void orc_get_neighbour_pos(const orc_sim* s, double* px, double* py) {
  int k = s->k;
  NbrPos* tmp = (NbrPos*)malloc((size_t)k * sizeof(NbrPos));
  for (int64_t i = 0; i < s->n; i++) {
    int64_t id = s->ps[i].id;
    int m = 0;
    for (int t = 0; t < k; t++) {
      int64_t j = s->nni[id * k + t];
      if (j < 0) continue;
      tmp[m].d2 = s->nnd2[id * k + t];
      tmp[m].id = (double)s->ps[j].id;
      tmp[m].px = s->nnpx[id * k + t];
      tmp[m].py = s->nnpy[id * k + t];
      m++;
    }
    qsort(tmp, (size_t)m, sizeof(NbrPos), cmp_nbrpos);
    for (int t = 0; t < k; t++) {
      px[id * k + t] = t < m ? tmp[t].px : NAN;
      py[id * k + t] = t < m ? tmp[t].py : NAN;
    }
  }
  free(tmp);
}

double orc_total_energy(const orc_sim* s) { /* sim/sph.go:441-447 */
  double tot = 0.0;
  for (int64_t i = 0; i < s->n; i++) tot += s->ps[i].e;
  return tot;
}
double orc_total_density(const orc_sim* s) { /* sim/sph.go:449-455 */
  double tot = 0.0;
  for (int64_t i = 0; i < s->n; i++) tot += s->ps[i].rho;
  return tot;
}
double orc_total_momentum(const orc_sim* s) { /* sim/sph.go:457-463 (quirk) */
  double tot = 0.0;
  for (int64_t i = 0; i < s->n; i++)
    tot = sqrt(s->ps[i].vx * s->ps[i].vx + s->ps[i].vy * s->ps[i].vy);
  return tot;
}

void orc_reductions(const orc_sim* s, double out[6]) {
  /* sums in ID order so the result does not depend on the tree order */
  double se = 0, srho = 0, et = 0, px = 0, py = 0;
  int64_t* pos = s->pos_of_id;
  for (int64_t i = 0; i < s->n; i++) pos[s->ps[i].id] = i;
  double m = s->p.mass;
  for (int64_t id = 0; id < s->n; id++) {
    const OP* p = &s->ps[pos[id]];
    se += p->e;
    srho += p->rho;
    et += m * (p->e + 0.5 * (p->vx * p->vx + p->vy * p->vy));
    px += m * p->vx;
    py += m * p->vy;
  }
  out[0] = se; out[1] = srho; out[2] = et; out[3] = px; out[4] = py;
  out[5] = s->dt_used;
}

int64_t orc_current_step(const orc_sim* s) { return s->step; }
int64_t orc_boundary_ties(const orc_sim* s) { return s->ties; }
