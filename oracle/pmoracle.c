/*
 * pmoracle.c -- FP64 CPU restatement of ProtoMol's force-evaluation hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and CPU baseline for the HIP
 * implementation.  Nothing in the product path links or loads this file.
 *
 * Each function names the reference file and lines (ProtoMol 2.1,
 * framework/) whose loop order and arithmetic it restates.  The restatement
 * is written so that, compiled with gcc -O2 and no -ffast-math/-ffp-contract,
 * it reproduces the reference's rounding for the cell binning and the pair
 * loop.  See pmoracle.h for which test pins each function.
 */
#include "pmoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* base/pmconstants.cpp:41-75 */
#define ORC_EPSILON 1.0e-14
#define ORC_MAX_INT_2 (INT_MAX / 2)
#define ORC_PI 3.14159265358979323846
#define ORC_INV_TIMEFACTOR 0.02045482828087295
#define ORC_BOLTZMANN 0.001987191

/* ------------------------------------------------------------------ PBC */
/* topology/PeriodicBoundaryConditions.cpp:34-77 for an orthogonal box
 * e1=(Lx,0,0), e2=(0,Ly,0), e3=(0,0,Lz). */
typedef struct {
  double e[3];    /* myE1.x, myE2.y, myE3.z */
  double er[3];   /* myE1r.x, myE2r.y, myE3r.z */
  double origin[3];
  double min[3], max[3];
  double h[3], h2[3];
  double d;       /* myD */
} orc_pbc;

static void pbc_set(orc_pbc* p, const double* box, const double* origin) {
  double lx = box[0], ly = box[1], lz = box[2];
  /* a1 = e2 x e3, a2 = e3 x e1, a3 = e1 x e2 (base/Vector3D.h:103-106) with
   * zero off-diagonal entries; e1r = a1 / e1.dot(a1) etc. */
  double a1x = ly * lz - 0.0 * 0.0;
  double a2y = lz * lx - 0.0 * 0.0;
  double a3z = lx * ly - 0.0 * 0.0;
  double dot1 = lx * a1x + 0.0 * 0.0 + 0.0 * 0.0;
  double dot2 = 0.0 * 0.0 + ly * a2y + 0.0 * 0.0;
  double dot3 = 0.0 * 0.0 + 0.0 * 0.0 + lz * a3z;
  int k;
  p->e[0] = lx; p->e[1] = ly; p->e[2] = lz;
  p->er[0] = a1x / dot1;
  p->er[1] = a2y / dot2;
  p->er[2] = a3z / dot3;
  for (k = 0; k < 3; k++) {
    double s = p->e[k];  /* (e1+e2+e3) component k: L_k + 0 + 0 */
    double a = origin[k] - s * 0.5;
    double b = origin[k] + s * 0.5;
    p->origin[k] = origin[k];
    p->min[k] = a < b ? a : b;
    p->max[k] = a > b ? a : b;
    p->h[k] = p->max[k] - p->min[k];
    p->h2[k] = p->h[k] * 0.5;
  }
  {
    double dx = (lx * 0.5) * (lx * 0.5);
    double dy = (ly * 0.5) * (ly * 0.5);
    double dz = (lz * 0.5) * (lz * 0.5);
    double m = dy < dz ? dy : dz;
    p->d = dx < m ? dx : m;
  }
}

/* PeriodicBoundaryConditions::minimalDifference(c1,c2,distSquared),
 * topology/PeriodicBoundaryConditions.h:76-117 (orthogonal branch). */
static inline void pbc_min_diff(const orc_pbc* p, const double* c1,
                                const double* c2, double* d, double* r2) {
  double x = c2[0] - c1[0], y = c2[1] - c1[1], z = c2[2] - c1[2];
  double s = x * x + y * y + z * z;
  if (s > p->d) {
    if (x < -p->h2[0]) x += p->h[0]; else if (x > p->h2[0]) x -= p->h[0];
    if (y < -p->h2[1]) y += p->h[1]; else if (y > p->h2[1]) y -= p->h[1];
    if (z < -p->h2[2]) z += p->h[2]; else if (z > p->h2[2]) z -= p->h[2];
    s = x * x + y * y + z * z;
    if (s > p->d) {
      x -= p->e[0] * rint(p->er[0] * x);
      y -= p->e[1] * rint(p->er[1] * y);
      z -= p->e[2] * rint(p->er[2] * z);
      s = x * x + y * y + z * z;
    }
  }
  d[0] = x; d[1] = y; d[2] = z;
  *r2 = s;
}

/* PeriodicBoundaryConditions::minimalPosition, .h:120-160 (orthogonal). */
static inline void pbc_min_pos(const orc_pbc* p, const double* c, double* out) {
  double x = c[0] - p->origin[0], y = c[1] - p->origin[1], z = c[2] - p->origin[2];
  if (x * x + y * y + z * z <= p->d) {
    out[0] = x; out[1] = y; out[2] = z;
    return;
  }
  if (x < -p->h2[0]) x += p->h[0]; else if (x > p->h2[0]) x -= p->h[0];
  if (y < -p->h2[1]) y += p->h[1]; else if (y > p->h2[1]) y -= p->h[1];
  if (z < -p->h2[2]) z += p->h[2]; else if (z > p->h2[2]) z -= p->h[2];
  if (x * x + y * y + z * z <= p->d) {
    out[0] = x; out[1] = y; out[2] = z;
    return;
  }
  {
    double tx = p->e[0] * rint(p->er[0] * x);
    double ty = p->e[1] * rint(p->er[1] * y);
    double tz = p->e[2] * rint(p->er[2] * z);
    out[0] = x - tx; out[1] = y - ty; out[2] = z - tz;
  }
}

/* ------------------------------------------------------------- cell grid */
typedef struct {
  int n[3];
  double real[3], rreal[3];
  double delta[3];   /* origin - min */
  int maxn[3], n1[3], maxn1[3];
} orc_grid;

/* CubicCellManager::initialize (topology/CubicCellManager.cpp:25-42) and
 * ArrayCellListStructure::initialize (ArrayCellListStructure.cpp:24-112). */
static void grid_set(orc_grid* g, const orc_pbc* p, double cellsize) {
  int k;
  for (k = 0; k < 3; k++) {
    double d = p->max[k] - p->min[k];
    double f = floor(d / cellsize);
    g->real[k] = d / (1.0 > f ? 1.0 : f);
    g->rreal[k] = 1.0 / g->real[k];
    {
      int nn = (int)floor(d / g->real[k] + ORC_EPSILON);
      g->n[k] = nn > 1 ? nn : 1;
    }
    g->delta[k] = p->origin[k] - p->min[k];
    g->maxn[k] = ORC_MAX_INT_2 - ORC_MAX_INT_2 % g->n[k];
    g->n1[k] = -g->n[k] / 2;
    g->maxn1[k] = g->maxn[k] - g->n1[k];
  }
}

/* buildCellLists (topology/buildCellLists.cpp:11-38) with
 * CubicCellManager::findCell (CubicCellManager.h:61-64).  Out-of-range cell
 * coordinates (the reference's undefined behaviour at x == +L/2) are clamped
 * into the grid. */
static int cell_of(const orc_pbc* p, const orc_grid* g, const double* x, int* c3) {
  double mp[3];
  int k;
  pbc_min_pos(p, x, mp);
  for (k = 0; k < 3; k++) {
    double v = g->delta[k] + mp[k];
    double f = floor(v * g->rreal[k]);
    int c;
    if (!(f >= 0.0)) c = 0;                     /* negative or NaN */
    else if (f >= (double)g->n[k]) c = g->n[k] - 1;
    else c = (int)f;
    c3[k] = c;
  }
  return (c3[0] * g->n[1] + c3[1]) * g->n[2] + c3[2];
}

int orc_cells(int n, const double* pos, const double* box, const double* origin,
              double cellsize, int* dims, double* realcell, int* cell,
              int* head, int* next, int ncap) {
  orc_pbc p;
  orc_grid g;
  int i, nc, c3[3];
  pbc_set(&p, box, origin);
  grid_set(&g, &p, cellsize);
  nc = g.n[0] * g.n[1] * g.n[2];
  if (dims) { dims[0] = g.n[0]; dims[1] = g.n[1]; dims[2] = g.n[2]; }
  if (realcell) { realcell[0] = g.real[0]; realcell[1] = g.real[1]; realcell[2] = g.real[2]; }
  if (!head || nc > ncap) return nc > ncap ? -1 : nc;
  for (i = 0; i < nc; i++) head[i] = -1;
  for (i = n - 1; i >= 0; i--) {
    int c = cell_of(&p, &g, pos + 3 * i, c3);
    if (cell) cell[i] = c;
    if (next) next[i] = head[c];
    head[c] = i;
  }
  return nc;
}

/* ---------------------------------------------------------- delta list */
static int cmp_cell3(const void* a, const void* b) {
  const int* x = (const int*)a;
  const int* y = (const int*)b;
  /* CubicCellLocation::operator< (topology/CubicCellLocation.h:56-64) */
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
  if (x[2] != y[2]) return x[2] < y[2] ? -1 : 1;
  return 0;
}

/* CellListEnumerator<PBC,Cubic>::initialize
 * (topology/CellListEnumerator_periodicBoundaries.h:37-108) with
 * ArrayCellListStructure::basisCell (ArrayCellListStructure.h:117-122). */
static int build_deltas(const orc_grid* g, double cutoff, int** out) {
  double cutoff2 = cutoff * cutoff;
  int nx = (int)(cutoff / g->real[0] + 1.0 + ORC_EPSILON);
  int ny = (int)(cutoff / g->real[1] + 1.0 + ORC_EPSILON);
  int nz = (int)(cutoff / g->real[2] + 1.0 + ORC_EPSILON);
  int n0 = nx < g->n[0] ? nx : g->n[0];
  int n1 = ny < g->n[1] ? ny : g->n[1];
  int n2 = nz < g->n[2] ? nz : g->n[2];
  double xx = g->real[0] * g->real[0];
  double yy = g->real[1] * g->real[1];
  double zz = g->real[2] * g->real[2];
  int cap = (2 * n0 + 1) * (2 * n1 + 1) * (2 * n2 + 1);
  int* list = (int*)malloc(sizeof(int) * 3 * (size_t)cap);
  int cnt = 0, k, l, m, t;
  for (k = -n0; k <= n0; k++) {
    int x = abs(k) - 1;
    double d0;
    if (x < 0) x = 0;
    d0 = x * x * xx;
    for (l = -n1; l <= n1; l++) {
      int y = abs(l) - 1;
      double d1;
      if (y < 0) y = 0;
      d1 = d0 + y * y * yy;
      for (m = -n2; m <= n2; m++) {
        int z = abs(m) - 1;
        if (z < 0) z = 0;
        if (d1 + z * z * zz < cutoff2) {
          int b[3], dup = 0;
          b[0] = (k + g->maxn1[0]) % g->n[0] + g->n1[0];
          b[1] = (l + g->maxn1[1]) % g->n[1] + g->n1[1];
          b[2] = (m + g->maxn1[2]) % g->n[2] + g->n1[2];
          if (b[0] == 0 && b[1] == 0 && b[2] == 0) continue;  /* zero seeded in the set */
          for (t = 0; t < cnt; t++)
            if (list[3 * t] == b[0] && list[3 * t + 1] == b[1] && list[3 * t + 2] == b[2]) { dup = 1; break; }
          if (!dup) {
            list[3 * cnt] = b[0]; list[3 * cnt + 1] = b[1]; list[3 * cnt + 2] = b[2];
            cnt++;
          }
        }
      }
    }
  }
  qsort(list, (size_t)cnt, sizeof(int) * 3, cmp_cell3);
  *out = list;
  return cnt;
}

int orc_delta_list(const double* box, const double* origin, double cellsize,
                   double cutoff, int* out, int cap) {
  orc_pbc p;
  orc_grid g;
  int* list;
  int cnt, i;
  pbc_set(&p, box, origin);
  grid_set(&g, &p, cellsize);
  cnt = build_deltas(&g, cutoff, &list);
  if (cnt > cap) { free(list); return -cnt; }
  for (i = 0; i < 3 * cnt; i++) out[i] = list[i];
  free(list);
  return cnt;
}

/* ----------------------------------------------------------- exclusions */
typedef struct {
  int n;
  int* start;   /* per low atom, CSR over (partner, class) */
  int* partner;
  int* cls;
  int* tab_a;   /* table order (ExclusionTable::myTable) */
  int* tab_b;
  int* tab_c;
  int ntab;
} orc_excl;

/* ExclusionTable::add/check (topology/ExclusionTable.cpp:112-134,
 * ExclusionTable.h:130-146): pairs are stored with a1 < a2, the last add of a
 * pair wins for check(); the table keeps, in the order of the adds, the first
 * add with class != NONE and a1 != a2 of every pair (mySet orders
 * ExclusionPair by (a1, a2) only, ExclusionTable.cpp:20-28). */
static void excl_build(orc_excl* e, int n, int nexcl, const int* ea,
                       const int* eb, const int* ec) {
  int i, t;
  int* fill = (int*)calloc((size_t)n + 1, sizeof(int));
  size_t cap = (size_t)(nexcl > 0 ? nexcl : 1);
  e->n = n;
  e->start = (int*)calloc((size_t)n + 1, sizeof(int));
  e->partner = (int*)malloc(sizeof(int) * cap);
  e->cls = (int*)malloc(sizeof(int) * cap);
  e->tab_a = (int*)malloc(sizeof(int) * cap);
  e->tab_b = (int*)malloc(sizeof(int) * cap);
  e->tab_c = (int*)malloc(sizeof(int) * cap);
  e->ntab = 0;
  for (t = 0; t < nexcl; t++) {
    int a = ea[t] < eb[t] ? ea[t] : eb[t];
    e->start[a + 1]++;
  }
  for (i = 0; i < n; i++) e->start[i + 1] += e->start[i];
  for (i = 0; i < n; i++) fill[i] = e->start[i];
  for (t = 0; t < nexcl; t++) {
    int a = ea[t], b = eb[t], c = ec[t], u, seen = 0;
    if (a > b) { int s = a; a = b; b = s; }
    /* the row keeps every add in order: check() takes the last match */
    for (u = e->start[a]; u < fill[a]; u++)
      if (e->partner[u] == b && e->cls[u] != ORC_EXCL_NONE) { seen = 1; break; }
    e->partner[fill[a]] = b;
    e->cls[fill[a]] = c;
    fill[a]++;
    if (c != ORC_EXCL_NONE && a != b && !seen) {
      e->tab_a[e->ntab] = a; e->tab_b[e->ntab] = b; e->tab_c[e->ntab] = c; e->ntab++;
    }
  }
  free(fill);
}

static inline int excl_check(const orc_excl* e, int a, int b) {
  int u, r = ORC_EXCL_NONE;
  if (a > b) { int s = a; a = b; b = s; }
  for (u = e->start[a]; u < e->start[a + 1]; u++)
    if (e->partner[u] == b) r = e->cls[u];
  return r;
}

static void excl_free(orc_excl* e) {
  free(e->start); free(e->partner); free(e->cls);
  free(e->tab_a); free(e->tab_b); free(e->tab_c);
}

/* ------------------------------------------------------ pair functors */
/* mathutilities.cpp:29-40, Abramowitz & Stegun 7.1.26 */
static inline double poly5(double ar) {
  const double E1 = 0.254829592, E2 = -0.284496736, E3 = 1.421413741;
  const double E4 = -1.453152027, E5 = 1.061405429, PP = 0.3275911;
  double t = 1.0 / (1.0 + PP * ar);
  return ((t) * (E1 + (t) * (E2 + (t) * (E3 + (t) * (E4 + (t) * E5)))));
}

typedef struct {
  orc_term t;
  double cut2;            /* cutoffSquared() */
  /* C2 constants, C2SwitchingFunction.cpp:15-23 */
  double s_on2, sw1, sw2, sw3;
  /* C1 constants, C1SwitchingFunction.cpp:19-26 */
  double c15_1, c05_3, c15_3;
  /* EwaldReal, CoulombEwaldRealForce.cpp:9-18 */
  double two_alpha_pi;
} orc_term_state;

static void term_init(orc_term_state* s, const orc_term* t) {
  double c = t->cutoff;
  s->t = *t;
  s->cut2 = c * c;
  s->s_on2 = t->switchon * t->switchon;
  {
    double d = c * c - t->switchon * t->switchon;
    double d3 = d * d * d;   /* power<3> */
    s->sw1 = 1.0 / d3;
    s->sw2 = c * c - 3.0 * t->switchon * t->switchon;
    s->sw3 = 4.0 / d3;
  }
  s->c15_1 = 1.5 / c;
  s->c05_3 = 0.5 / (c * c * c);
  s->c15_3 = 1.5 / (c * c * c);
  s->two_alpha_pi = 2.0 * t->alpha / sqrt(ORC_PI);
}

/* Switching functions: CutoffSwitchingFunction.h:64-68,
 * C1SwitchingFunction.h:64-75, C2SwitchingFunction.h:64-82. */
static inline void swf_eval(const orc_term_state* s, double r2, double* v, double* d) {
  switch (s->t.swf) {
    case ORC_SWF_C2:
      *d = 0.0;
      if (r2 > s->cut2) {
        *v = 0.0;
      } else if (r2 >= s->s_on2) {
        double c2 = s->cut2 - r2;
        double c4 = c2 * (s->sw2 + 2.0 * r2);
        *v = s->sw1 * (c2 * c4);
        *d = s->sw3 * (c2 * c2 - c4);
      } else {
        *v = 1.0;
      }
      break;
    case ORC_SWF_C1:
      if (r2 > s->cut2) { *v = 0.0; *d = 0.0; return; }
      {
        double dist = sqrt(r2);
        *v = 1.0 - dist * (s->c15_1 - r2 * s->c05_3);
        *d = -s->c15_1 / dist + dist * s->c15_3;
      }
      break;
    default:
      *d = 0;
      *v = (r2 > s->cut2 ? 0.0 : 1.0);
      break;
  }
}

static inline int term_dist_r2(const orc_term_state* s) {
  return s->t.pot != ORC_POT_EWALDREAL;   /* DIST_R2 enums */
}

static inline int term_modify(const orc_term_state* s) {
  return s->t.swf != ORC_SWF_CUTOFF;      /* MODIFY enums */
}

/* LennardJonesForce.h:37-74, CoulombForce.h:41-57, CoulombEwaldRealForce.h:45-73 */
static inline void pot_eval(const orc_term_state* s, double* energy, double* force,
                            double r2, double rr2, int i, int j, int excl,
                            const double* q, const int* type, int ntypes,
                            const double* ljA, const double* ljB,
                            const double* ljA14, const double* ljB14, double csf) {
  if (s->t.pot == ORC_POT_LJ) {
    int ti = type[i], tj = type[j];
    double A, B, r6, r12, r6B, r12A;
    if (excl != ORC_EXCL_MODIFIED) { A = ljA[ti * ntypes + tj]; B = ljB[ti * ntypes + tj]; }
    else { A = ljA14[ti * ntypes + tj]; B = ljB14[ti * ntypes + tj]; }
    r6 = rr2 * rr2 * rr2;
    r12 = r6 * r6;
    r6B = B * r6;
    r12A = A * r12;
    *energy = r12A - r6B;
    *force = 12.0 * r12A * rr2 - 6.0 * r6B * rr2;
  } else if (s->t.pot == ORC_POT_COULOMB) {
    double e = q[i] * q[j] * sqrt(rr2);
    if (excl == ORC_EXCL_MODIFIED) e *= csf;
    *energy = e;
    *force = e * rr2;
  } else {
    double qq = q[i] * q[j];
    double r, rr, ar, e, en;
    if (csf != 1.0 && excl == ORC_EXCL_MODIFIED) qq *= csf;
    r = sqrt(r2);
    rr = 1.0 / r;
    ar = s->t.alpha * r;
    e = qq * exp(-ar * ar);
    en = poly5(ar) * e * rr;
    *energy = en;
    *force = ((en + s->two_alpha_pi * e) * rr * rr);
  }
}

static inline void add_virial(double* en, const double* f, const double* d) {
  /* ScalarStructure::addVirial (base/ScalarStructure.cpp:41-63) */
  double xy = f[0] * d[1], xz = f[0] * d[2], yz = f[1] * d[2];
  en[ORC_E_VXX] += f[0] * d[0];
  en[ORC_E_VXY] += xy;
  en[ORC_E_VXZ] += xz;
  en[ORC_E_VYX] += xy;
  en[ORC_E_VYY] += f[1] * d[1];
  en[ORC_E_VYZ] += yz;
  en[ORC_E_VZX] += xz;
  en[ORC_E_VZY] += yz;
  en[ORC_E_VZZ] += f[2] * d[2];
}

typedef struct {
  int n;
  const double* pos;
  const double* q;
  const int* type;
  int ntypes;
  const double *ljA, *ljB, *ljA14, *ljB14;
  const int* molecule;
  double csf;
  int nterms;
  orc_term_state ts[2];
  double cut2;
  int want_virial;
  orc_excl ex;
  orc_pbc pbc;
} orc_pairctx;

/* OneAtomPair::doOneAtomPair (OneAtomPair.h:127-188) and
 * OneAtomPairTwo::doOneAtomPair (OneAtomPairTwo.h:131-199).
 * Returns 1 if the pair passed the cutoff and exclusion tests. */
static inline int do_pair(const orc_pairctx* c, int i, int j, double* f, double* en) {
  double d[3], r2, rr2, e1, f1, e2 = 0.0, f2 = 0.0, fij[3], ftot;
  int mi, mj, excl;
  pbc_min_diff(&c->pbc, c->pos + 3 * i, c->pos + 3 * j, d, &r2);
  if (r2 > c->cut2) return 0;
  mi = c->molecule[i];
  mj = c->molecule[j];
  excl = (mi == mj) ? excl_check(&c->ex, i, j) : ORC_EXCL_NONE;
  if (excl == ORC_EXCL_FULL) return 0;
  {
    int any_r2 = term_dist_r2(&c->ts[0]) || (c->nterms > 1 && term_dist_r2(&c->ts[1]));
    rr2 = any_r2 ? 1.0 / r2 : 1.0;
  }
  pot_eval(&c->ts[0], &e1, &f1, r2, rr2, i, j, excl, c->q, c->type, c->ntypes,
           c->ljA, c->ljB, c->ljA14, c->ljB14, c->csf);
  if (c->nterms > 1)
    pot_eval(&c->ts[1], &e2, &f2, r2, rr2, i, j, excl, c->q, c->type, c->ntypes,
             c->ljA, c->ljB, c->ljA14, c->ljB14, c->csf);
  if (c->nterms == 1) {
    if (term_modify(&c->ts[0])) {
      double sv, sd;
      swf_eval(&c->ts[0], r2, &sv, &sd);
      f1 = f1 * sv - e1 * sd;
      e1 = e1 * sv;
    }
    en[c->ts[0].t.pot == ORC_POT_LJ ? ORC_E_LJ : ORC_E_COULOMB] += e1;
    ftot = f1;
  } else {
    if (term_modify(&c->ts[0]) || term_modify(&c->ts[1])) {
      double sv, sd;
      swf_eval(&c->ts[0], r2, &sv, &sd);
      f1 = f1 * sv - e1 * sd;
      e1 = e1 * sv;
      swf_eval(&c->ts[1], r2, &sv, &sd);
      f2 = f2 * sv - e2 * sd;
      e2 = e2 * sv;
    }
    en[c->ts[0].t.pot == ORC_POT_LJ ? ORC_E_LJ : ORC_E_COULOMB] += e1;
    en[c->ts[1].t.pot == ORC_POT_LJ ? ORC_E_LJ : ORC_E_COULOMB] += e2;
    ftot = f1 + f2;
  }
  fij[0] = d[0] * ftot; fij[1] = d[1] * ftot; fij[2] = d[2] * ftot;
  f[3 * i] -= fij[0]; f[3 * i + 1] -= fij[1]; f[3 * i + 2] -= fij[2];
  f[3 * j] += fij[0]; f[3 * j + 1] += fij[1]; f[3 * j + 2] += fij[2];
  if (c->want_virial) add_virial(en, fij, d);
  return 1;
}

static int pairctx_init(orc_pairctx* c, int n, const double* pos, const double* q,
                        const int* type, int ntypes, const double* ljA,
                        const double* ljB, const double* ljA14, const double* ljB14,
                        const int* molecule, int nexcl, const int* ex_a,
                        const int* ex_b, const int* ex_cls, double csf,
                        const double* box, const double* origin, int nterms,
                        const orc_term* terms, int want_virial) {
  int t;
  if (nterms < 1 || nterms > 2) return -1;
  c->n = n; c->pos = pos; c->q = q; c->type = type; c->ntypes = ntypes;
  c->ljA = ljA; c->ljB = ljB; c->ljA14 = ljA14; c->ljB14 = ljB14;
  c->molecule = molecule; c->csf = csf; c->nterms = nterms;
  c->want_virial = want_virial;
  for (t = 0; t < nterms; t++) term_init(&c->ts[t], &terms[t]);
  c->cut2 = c->ts[0].cut2;
  if (nterms > 1 && c->ts[1].cut2 > c->cut2) c->cut2 = c->ts[1].cut2;  /* OneAtomPairTwo.h:48 */
  excl_build(&c->ex, n, nexcl, ex_a, ex_b, ex_cls);
  pbc_set(&c->pbc, box, origin);
  return 0;
}

/* Visit the cell pairs in the reference enumerator's order
 * (CellListEnumerator_periodicBoundaries.h:143-179, NonbondedCutoffForce.h:86-109). */
typedef void (*cellpair_fn)(void* user, int ci, int cj, int same);

static void enumerate_cell_pairs(const orc_grid* g, const int* head, const int* deltas,
                                 int nd, cellpair_fn fn, void* user) {
  int nc = g->n[0] * g->n[1] * g->n[2];
  int ci;
  for (ci = 0; ci < nc; ci++) {
    int cx, cy, cz, t;
    if (head[ci] < 0) continue;
    cx = ci / (g->n[1] * g->n[2]);
    cy = (ci / g->n[2]) % g->n[1];
    cz = ci % g->n[2];
    fn(user, ci, ci, 1);
    for (t = 0; t < nd; t++) {
      /* findPeriodic (ArrayCellListStructure.h:108-114) */
      int x = (cx + deltas[3 * t] + g->maxn[0]) % g->n[0];
      int y = (cy + deltas[3 * t + 1] + g->maxn[1]) % g->n[1];
      int z = (cz + deltas[3 * t + 2] + g->maxn[2]) % g->n[2];
      int cj = (x * g->n[1] + y) * g->n[2] + z;
      if (head[cj] < 0) continue;
      if (cj < ci) continue;     /* i > j: skipped by the do-while */
      fn(user, ci, cj, 0);
    }
  }
}

typedef struct {
  const orc_pairctx* c;
  const int* head;
  const int* next;
  double* f;
  double* en;
  long long* cand;
  long long* inrc;
} cutoff_user;

static void cutoff_cellpair(void* u, int ci, int cj, int same) {
  cutoff_user* cu = (cutoff_user*)u;
  int i, j;
  long long nc = 0, ni = 0;
  for (i = cu->head[ci]; i != -1; i = cu->next[i]) {
    for (j = (same ? cu->next[i] : cu->head[cj]); j != -1; j = cu->next[j]) {
      nc++;
      ni += do_pair(cu->c, i, j, cu->f, cu->en);
    }
  }
  if (cu->cand) cu->cand[ci] += nc;
  if (cu->inrc) cu->inrc[ci] += ni;
}

int orc_cutoff(int n, const double* pos, const double* q, const int* type,
               int ntypes, const double* ljA, const double* ljB,
               const double* ljA14, const double* ljB14, const int* molecule,
               int nexcl, const int* ex_a, const int* ex_b, const int* ex_cls,
               double csf, const double* box, const double* origin,
               double cellsize, int nterms, const orc_term* terms,
               double algo_cutoff, int want_virial, double* forces,
               double* energies, long long* cand, long long* inrc) {
  orc_pairctx c;
  orc_grid g;
  int *head, *next, *deltas, nd, nc;
  cutoff_user cu;
  if (pairctx_init(&c, n, pos, q, type, ntypes, ljA, ljB, ljA14, ljB14, molecule,
                   nexcl, ex_a, ex_b, ex_cls, csf, box, origin, nterms, terms,
                   want_virial) != 0)
    return -1;
  grid_set(&g, &c.pbc, cellsize);
  nc = g.n[0] * g.n[1] * g.n[2];
  head = (int*)malloc(sizeof(int) * (size_t)nc);
  next = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  orc_cells(n, pos, box, origin, cellsize, NULL, NULL, NULL, head, next, nc);
  nd = build_deltas(&g, algo_cutoff, &deltas);
  cu.c = &c; cu.head = head; cu.next = next; cu.f = forces; cu.en = energies;
  cu.cand = cand; cu.inrc = inrc;
  if (cand) memset(cand, 0, sizeof(long long) * (size_t)nc);
  if (inrc) memset(inrc, 0, sizeof(long long) * (size_t)nc);
  enumerate_cell_pairs(&g, head, deltas, nd, cutoff_cellpair, &cu);
  free(deltas); free(head); free(next);
  excl_free(&c.ex);
  return 0;
}

/* OpenMP variant: cell packages with per-thread force/energy buffers,
 * reduced at the end (the structure of Parallel::getNumberOfPackages,
 * parallel/Parallel.cpp:723-734, without MPI). */
int orc_cutoff_omp(int nthreads, int n, const double* pos, const double* q,
                   const int* type, int ntypes, const double* ljA,
                   const double* ljB, const double* ljA14, const double* ljB14,
                   const int* molecule, int nexcl, const int* ex_a,
                   const int* ex_b, const int* ex_cls, double csf,
                   const double* box, const double* origin, double cellsize,
                   int nterms, const orc_term* terms, double algo_cutoff,
                   double* forces, double* energies) {
  orc_pairctx c;
  orc_grid g;
  int *head, *next, *deltas, nd, nc, used = 1;
  if (pairctx_init(&c, n, pos, q, type, ntypes, ljA, ljB, ljA14, ljB14, molecule,
                   nexcl, ex_a, ex_b, ex_cls, csf, box, origin, nterms, terms, 0) != 0)
    return -1;
  grid_set(&g, &c.pbc, cellsize);
  nc = g.n[0] * g.n[1] * g.n[2];
  head = (int*)malloc(sizeof(int) * (size_t)nc);
  next = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  orc_cells(n, pos, box, origin, cellsize, NULL, NULL, NULL, head, next, nc);
  nd = build_deltas(&g, algo_cutoff, &deltas);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
  {
#pragma omp single
    used = omp_get_num_threads();
  }
#endif
  {
    double* fbuf = (double*)calloc((size_t)used * 3 * (size_t)n, sizeof(double));
    double* ebuf = (double*)calloc((size_t)used * ORC_E_COUNT, sizeof(double));
    int ci;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (ci = 0; ci < nc; ci++) {
      int tid = 0;
      cutoff_user cu;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      if (head[ci] < 0) continue;
      cu.c = &c; cu.head = head; cu.next = next;
      cu.f = fbuf + (size_t)tid * 3 * n; cu.en = ebuf + (size_t)tid * ORC_E_COUNT;
      cu.cand = NULL; cu.inrc = NULL;
      {
        int cx = ci / (g.n[1] * g.n[2]);
        int cy = (ci / g.n[2]) % g.n[1];
        int cz = ci % g.n[2];
        int t;
        cutoff_cellpair(&cu, ci, ci, 1);
        for (t = 0; t < nd; t++) {
          int x = (cx + deltas[3 * t] + g.maxn[0]) % g.n[0];
          int y = (cy + deltas[3 * t + 1] + g.maxn[1]) % g.n[1];
          int z = (cz + deltas[3 * t + 2] + g.maxn[2]) % g.n[2];
          int cj = (x * g.n[1] + y) * g.n[2] + z;
          if (head[cj] < 0 || cj < ci) continue;
          cutoff_cellpair(&cu, ci, cj, 0);
        }
      }
    }
    {
      int t, k;
      for (t = 0; t < used; t++) {
        for (k = 0; k < 3 * n; k++) forces[k] += fbuf[(size_t)t * 3 * n + k];
        for (k = 0; k < ORC_E_COUNT; k++) energies[k] += ebuf[(size_t)t * ORC_E_COUNT + k];
      }
    }
    free(fbuf); free(ebuf);
  }
  free(deltas); free(head); free(next);
  excl_free(&c.ex);
  return used;
}

/* ------------------------------------------------------------- B-spline */
/* BSpline::set (base/BSpline.cpp:64-117) */
void orc_bspline(int order, double w, double* theta, double* dtheta) {
  int k, j;
  if (order == 1) { theta[0] = 1; dtheta[0] = 0; return; }
  if (order == 2) { theta[1] = w; theta[0] = 1.0 - w; dtheta[1] = 1; dtheta[0] = -1; return; }
  theta[order - 1] = 0.0;
  theta[1] = w;
  theta[0] = 1.0 - w;
  for (k = 2; k < order - 1; k++) {
    double div = 1.0 / k;
    theta[k] = div * w * theta[k - 1];
    for (j = 1; j < k; j++)
      theta[k - j] = div * ((w + j) * theta[k - j - 1] + (k + 1 - j - w) * theta[k - j]);
    theta[0] = div * (1 - w) * theta[0];
  }
  dtheta[0] = -theta[0];
  for (j = 1; j <= order - 1; j++) dtheta[j] = theta[j - 1] - theta[j];
  {
    double div = 1.0 / (order - 1.0);
    theta[order - 1] = div * w * theta[order - 2];
    for (j = 1; j < order - 1; j++)
      theta[order - 1 - j] = div * ((w + j) * theta[order - j - 2] + (order - j - w) * theta[order - 1 - j]);
    theta[0] = div * (1 - w) * theta[0];
  }
}

/* ------------------------------------------------------------------ FFT */
/* Mixed-radix decimation-in-time recursion (Cooley-Tukey), out of place on a
 * strided 1D sequence.  Unnormalised; w[k] = exp(sign*2*pi*i*k/n) for the
 * full length n; a sub-transform of length m uses w[k*(n/m)]. */
static void fft1d_rec(const double* in, double* out, int n, int stride,
                      const double* w, int wstep, double* scratch) {
  int p, m, k, r, s, nfull = n * wstep;
  if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
  p = 0;
  if (n % 4 == 0 && n != 8) p = 4;
  else if (n % 2 == 0) p = 2;
  else if (n % 3 == 0) p = 3;
  else if (n % 5 == 0) p = 5;
  else {
    /* direct DFT for a prime length */
    for (k = 0; k < n; k++) {
      double re = 0, im = 0;
      for (r = 0; r < n; r++) {
        int idx = (int)(((long long)r * k % n) * wstep);
        double c = w[2 * idx], sn = w[2 * idx + 1];
        double xr = in[2 * r * stride], xi = in[2 * r * stride + 1];
        re += xr * c - xi * sn;
        im += xr * sn + xi * c;
      }
      out[2 * k] = re; out[2 * k + 1] = im;
    }
    return;
  }
  m = n / p;
  for (r = 0; r < p; r++)
    fft1d_rec(in + 2 * r * stride, out + 2 * r * m, m, stride * p, w, wstep * p, scratch);
  for (k = 0; k < m; k++) {
    double tr[5], ti[5];
    for (r = 0; r < p; r++) {
      int idx = (r * k * wstep) % nfull;
      double c = w[2 * idx], sn = w[2 * idx + 1];
      double xr = out[2 * (r * m + k)], xi = out[2 * (r * m + k) + 1];
      tr[r] = xr * c - xi * sn;
      ti[r] = xr * sn + xi * c;
    }
    for (s = 0; s < p; s++) {
      double re = 0, im = 0;
      for (r = 0; r < p; r++) {
        int idx = (((r * s) % p) * m * wstep) % nfull;
        double c = w[2 * idx], sn = w[2 * idx + 1];
        re += tr[r] * c - ti[r] * sn;
        im += tr[r] * sn + ti[r] * c;
      }
      scratch[2 * (s * m + k)] = re;
      scratch[2 * (s * m + k) + 1] = im;
    }
  }
  memcpy(out, scratch, sizeof(double) * 2 * (size_t)n);
}

static void fft_axis(double* data, int n, int stride, int sign, double* buf,
                     double* out, double* scratch, const double* w) {
  int k;
  for (k = 0; k < n; k++) { buf[2 * k] = data[2 * (size_t)k * stride]; buf[2 * k + 1] = data[2 * (size_t)k * stride + 1]; }
  fft1d_rec(buf, out, n, 1, w, 1, scratch);
  for (k = 0; k < n; k++) { data[2 * (size_t)k * stride] = out[2 * k]; data[2 * (size_t)k * stride + 1] = out[2 * k + 1]; }
  (void)sign;
}

static double* twiddles(int n, int sign) {
  double* w = (double*)malloc(sizeof(double) * 2 * (size_t)n);
  int k;
  for (k = 0; k < n; k++) {
    double a = sign * 2.0 * ORC_PI * (double)k / n;
    w[2 * k] = cos(a);
    w[2 * k + 1] = sin(a);
  }
  return w;
}

void orc_fft3d(double* data, int nx, int ny, int nz, int sign) {
  int nmax = nx > ny ? nx : ny, i, j;
  double *buf, *out, *scratch, *wx, *wy, *wz;
  if (nz > nmax) nmax = nz;
  buf = (double*)malloc(sizeof(double) * 2 * (size_t)nmax);
  out = (double*)malloc(sizeof(double) * 2 * (size_t)nmax);
  scratch = (double*)malloc(sizeof(double) * 2 * (size_t)nmax);
  wx = twiddles(nx, sign); wy = twiddles(ny, sign); wz = twiddles(nz, sign);
  for (i = 0; i < nx; i++)
    for (j = 0; j < ny; j++)
      fft_axis(data + 2 * ((size_t)(i * ny + j) * nz), nz, 1, sign, buf, out, scratch, wz);
  for (i = 0; i < nx; i++)
    for (j = 0; j < nz; j++)
      fft_axis(data + 2 * ((size_t)i * ny * nz + j), ny, nz, sign, buf, out, scratch, wy);
  for (i = 0; i < ny; i++)
    for (j = 0; j < nz; j++)
      fft_axis(data + 2 * ((size_t)i * nz + j), nx, ny * nz, sign, buf, out, scratch, wx);
  free(buf); free(out); free(scratch); free(wx); free(wy); free(wz);
}

/* ------------------------------------------------------------------ PME */
double orc_pme_alpha(double rc, double accuracy) {
  /* NonbondedPMEwaldSystemForce.h:564-586 */
  double alpha = 1.0, low = 0.0, high;
  int i;
  while (erfc(alpha * rc) / rc >= accuracy) alpha *= 2.0;
  high = alpha;
  for (i = 0; i < 100; ++i) {
    alpha = 0.5 * (low + high);
    if (erfc(alpha * rc) / rc >= accuracy) low = alpha; else high = alpha;
  }
  return alpha;
}

/* Grid::dftmod (base/Grid.h:650-665) */
static void dftmod(int order, int n, const double* theta, double* mod) {
  int i, j;
  for (i = 0; i < n; i++) {
    double sc = 0.0, ss = 0.0;
    for (j = 0; j < order; j++) {
      double x = ORC_PI * 2.0 * i * j / (double)n;
      sc += theta[j] * cos(x);
      ss += theta[j] * sin(x);
    }
    mod[i] = 1.0 / (sc * sc + ss * ss);
  }
}

int orc_pme(int n, const double* pos, const double* q, const int* molecule,
            int nexcl, const int* ex_a, const int* ex_b, const int* ex_cls,
            double csf, const double* box, const double* origin,
            double cellsize, int nx, int ny, int nz, int order, double rc,
            double accuracy, double alpha, int flags, int want_virial,
            double* forces, double* energies, double* parts, double* qgrid) {
  orc_pbc pbc;
  double lx = box[0], ly = box[1], lz = box[2];
  double v, a2, two_alpha_pi;
  double e_real = 0, e_recip = 0, e_intra = 0, e_self = 0, e_charged = 0;
  int do_real = flags & 1, do_recip = flags & 2, do_corr = flags & 4;
  int i;
  (void)molecule;
  if (order < 2 || (order % 2) != 0 || nx < order || ny < order || nz < order || rc <= 0 || accuracy <= 0)
    return -1;  /* NonbondedPMEwaldSystemForce.h:1117-1130 */
  pbc_set(&pbc, box, origin);
  /* initialize (:499-648) */
  if (alpha <= 0.0) alpha = orc_pme_alpha(rc, accuracy);
  v = fabs(lx * ly * lz);
  a2 = alpha * alpha;
  two_alpha_pi = 2.0 * alpha / sqrt(ORC_PI);
  if (do_corr) {
    double qs = 0.0, qn = 0.0;
    for (i = 0; i < n; i++) qs += q[i] * q[i];
    e_self = -qs * alpha / sqrt(ORC_PI);
    for (i = 0; i < n; i++) qn += q[i];
    if (fabs(qn * 0.00268283) > 1.0e-5) e_charged = -ORC_PI / (2.0 * v * a2) * qn * qn;
  }
  if (do_corr) {
    /* correctionTerm (:942-981) over the exclusion table */
    orc_excl ex;
    int t;
    excl_build(&ex, n, nexcl, ex_a, ex_b, ex_cls);
    if (want_virial) {
      energies[ORC_E_VXX] += e_charged;
      energies[ORC_E_VYY] += e_charged;
      energies[ORC_E_VZZ] += e_charged;
    }
    for (t = 0; t < ex.ntab; t++) {
      int a = ex.tab_a[t], b = ex.tab_b[t];
      double d[3], r2, qq, r, rr, e, fij[3], s;
      pbc_min_diff(&pbc, pos + 3 * a, pos + 3 * b, d, &r2);
      qq = q[a] * q[b];
      if (ex.tab_c[t] == ORC_EXCL_MODIFIED) qq *= 1 - csf;
      r = sqrt(r2);
      rr = 1 / r;
      e = erf(alpha * r) * rr;
      e_intra -= qq * e;
      s = (qq * (two_alpha_pi * exp(-a2 * r2) - e) * rr * rr);
      fij[0] = d[0] * s; fij[1] = d[1] * s; fij[2] = d[2] * s;
      forces[3 * a] -= fij[0]; forces[3 * a + 1] -= fij[1]; forces[3 * a + 2] -= fij[2];
      forces[3 * b] += fij[0]; forces[3 * b + 1] += fij[1]; forces[3 * b + 2] += fij[2];
      if (want_virial) add_virial(energies, fij, d);
    }
    excl_free(&ex);
  }
  if (do_real) {
    /* realTerm (:658-768): cell pairs x lattice images, CutoffSwitchingFunction */
    orc_grid g;
    orc_excl ex;
    int *head, *next, *deltas, nd, nc, nl = 0, k, l, m, ci;
    double lat[3 * 1024];
    double rc2 = rc * rc;
    excl_build(&ex, n, nexcl, ex_a, ex_b, ex_cls);
    grid_set(&g, &pbc, cellsize);
    nc = g.n[0] * g.n[1] * g.n[2];
    head = (int*)malloc(sizeof(int) * (size_t)nc);
    next = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    orc_cells(n, pos, box, origin, cellsize, NULL, NULL, NULL, head, next, nc);
    nd = build_deltas(&g, rc, &deltas);
    /* lattice vectors |n| <= 2rc sorted by norm, (0,0,0) first
     * (PeriodicBoundaryConditions.cpp:80-111) */
    {
      int bk = (int)floor(rc / lx + 1.0), bl = (int)floor(rc / ly + 1.0), bm = (int)floor(rc / lz + 1.0);
      lat[0] = lat[1] = lat[2] = 0.0;
      nl = 1;
      for (k = -bk; k <= bk; k++)
        for (l = -bl; l <= bl; l++)
          for (m = -bm; m <= bm; m++) {
            double cx = lx * k, cy = ly * l, cz = lz * m;
            if ((k || l || m) && sqrt(cx * cx + cy * cy + cz * cz) <= 2 * rc && nl < 1024) {
              lat[3 * nl] = cx; lat[3 * nl + 1] = cy; lat[3 * nl + 2] = cz; nl++;
            }
          }
      /* stable insertion sort by norm^2 of entries 1..nl-1 */
      for (k = 2; k < nl; k++) {
        double t0 = lat[3 * k], t1 = lat[3 * k + 1], t2 = lat[3 * k + 2];
        double nk = t0 * t0 + t1 * t1 + t2 * t2;
        l = k - 1;
        while (l >= 1 && lat[3 * l] * lat[3 * l] + lat[3 * l + 1] * lat[3 * l + 1] + lat[3 * l + 2] * lat[3 * l + 2] > nk) {
          lat[3 * (l + 1)] = lat[3 * l]; lat[3 * (l + 1) + 1] = lat[3 * l + 1]; lat[3 * (l + 1) + 2] = lat[3 * l + 2];
          l--;
        }
        lat[3 * (l + 1)] = t0; lat[3 * (l + 1) + 1] = t1; lat[3 * (l + 1) + 2] = t2;
      }
    }
    for (ci = 0; ci < nc; ci++) {
      int cx, cy, cz, t, pass;
      if (head[ci] < 0) continue;
      cx = ci / (g.n[1] * g.n[2]);
      cy = (ci / g.n[2]) % g.n[1];
      cz = ci % g.n[2];
      for (pass = -1; pass < nd; pass++) {
        int cj, same, ii, jj;
        if (pass < 0) { cj = ci; same = 1; }
        else {
          int x = (cx + deltas[3 * pass] + g.maxn[0]) % g.n[0];
          int y = (cy + deltas[3 * pass + 1] + g.maxn[1]) % g.n[1];
          int z = (cz + deltas[3 * pass + 2] + g.maxn[2]) % g.n[2];
          cj = (x * g.n[1] + y) * g.n[2] + z;
          if (head[cj] < 0 || cj < ci) continue;
          same = 0;
        }
        for (ii = head[ci]; ii != -1; ii = next[ii]) {
          double qi = q[ii], fi[3] = {0, 0, 0};
          int mi = molecule[ii];
          for (jj = (same ? ii : head[cj]); jj != -1; jj = next[jj]) {
            double rmin[3], r2d;
            int mj = molecule[jj], excl;
            pbc_min_diff(&pbc, pos + 3 * ii, pos + 3 * jj, rmin, &r2d);
            excl = (mi == mj) ? excl_check(&ex, ii, jj) : ORC_EXCL_NONE;
            if (ii == jj) excl = ORC_EXCL_FULL;
            for (t = 0; t < nl; t++) {
              double rij[3], rs, r, qq, rr, ar, e, energy, force, sv, sd, fij[3];
              if (excl == ORC_EXCL_FULL) { excl = ORC_EXCL_NONE; continue; }
              rij[0] = rmin[0] + lat[3 * t]; rij[1] = rmin[1] + lat[3 * t + 1]; rij[2] = rmin[2] + lat[3 * t + 2];
              rs = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
              if (rs > rc2) continue;
              r = sqrt(rs);
              qq = qi * q[jj];
              if (excl == ORC_EXCL_MODIFIED) qq *= csf;
              rr = 1.0 / r;
              ar = alpha * r;
              e = qq * exp(-ar * ar);
              energy = poly5(ar) * e * rr;
              force = ((energy + two_alpha_pi * e) * rr * rr);
              sv = (rs > rc2 ? 0.0 : 1.0); sd = 0;   /* CutoffSwitchingFunction */
              force = force * sv - energy * sd;
              energy = energy * sv;
              e_real += energy;
              fij[0] = rij[0] * force; fij[1] = rij[1] * force; fij[2] = rij[2] * force;
              fi[0] -= fij[0]; fi[1] -= fij[1]; fi[2] -= fij[2];
              forces[3 * jj] += fij[0]; forces[3 * jj + 1] += fij[1]; forces[3 * jj + 2] += fij[2];
              if (want_virial) add_virial(energies, fij, rij);
              excl = ORC_EXCL_NONE;
            }
          }
          forces[3 * ii] += fi[0]; forces[3 * ii + 1] += fi[1]; forces[3 * ii + 2] += fi[2];
        }
      }
    }
    free(deltas); free(head); free(next);
    excl_free(&ex);
  }
  if (do_recip) {
    /* reciprocalTerm (:879-932) with Grid<BSpline> (base/Grid.h) */
    size_t ng = (size_t)nx * ny * nz;
    double* Q = (double*)calloc(2 * ng, sizeof(double));
    double *th = (double*)malloc(sizeof(double) * 6 * (size_t)order * (size_t)(n > 0 ? n : 1));
    int* i0s = (int*)malloc(sizeof(int) * 3 * (size_t)(n > 0 ? n : 1));
    double hxr = nx / lx, hyr = ny / ly, hzr = nz / lz;
    int offx = -(order - 1) / 2 + nx, offy = -(order - 1) / 2 + ny, offz = -(order - 1) / 2 + nz;
    double *modx = (double*)malloc(sizeof(double) * nx), *mody = (double*)malloc(sizeof(double) * ny), *modz = (double*)malloc(sizeof(double) * nz);
    double *expx = (double*)malloc(sizeof(double) * nx), *expy = (double*)malloc(sizeof(double) * ny), *expz = (double*)malloc(sizeof(double) * nz);
    double fac = ORC_PI * ORC_PI / (alpha * alpha);
    double wr = 1.0 / lx, lr = 1.0 / ly, hr = 1.0 / lz;
    double vg = lx * ly * lz;
    double energy = 0, vxx = 0, vxy = 0, vxz = 0, vyy = 0, vyz = 0, vzz = 0;
    /* Grid::initialize (:557-648) */
    for (i = 0; i < nx; i++) { int i0 = i <= nx / 2 ? i : i - nx; expx[i] = exp(-fac * ((i0 * wr) * (i0 * wr))); }
    for (i = 0; i < ny; i++) { int i0 = i <= ny / 2 ? i : i - ny; expy[i] = exp(-fac * ((i0 * lr) * (i0 * lr))); }
    for (i = 0; i <= nz / 2; i++) { expz[i] = exp(-fac * ((i * hr) * (i * hr))); }
    {
      double t0[64], d0[64];
      orc_bspline(order, 0.0, t0, d0);
      dftmod(order, nx, t0, modx);
      dftmod(order, ny, t0, mody);
      dftmod(order, nz, t0, modz);
    }
    /* anterpolateCharge (:138-190) */
    for (i = 0; i < n; i++) {
      double x = pos[3 * i] * hxr, y = pos[3 * i + 1] * hyr, z = pos[3 * i + 2] * hzr;
      int ix, iy, iz, a, b, c;
      double *tx = th + (size_t)6 * order * i, *ty = tx + order, *tz = ty + order;
      double *dx = tz + order, *dy = dx + order, *dz = dy + order;
      while (x < 0.0) x += nx;
      while (x >= nx) x -= nx;
      while (y < 0.0) y += ny;
      while (y >= ny) y -= ny;
      while (z < 0.0) z += nz;
      while (z >= nz) z -= nz;
      ix = (int)x; iy = (int)y; iz = (int)z;
      i0s[3 * i] = ix + offx; i0s[3 * i + 1] = iy + offy; i0s[3 * i + 2] = iz + offz;
      orc_bspline(order, x - ix, tx, dx);
      orc_bspline(order, y - iy, ty, dy);
      orc_bspline(order, z - iz, tz, dz);
      for (a = 0; a < order; a++) {
        double qa = q[i] * tx[a];
        int i1 = (a + i0s[3 * i]) % nx;
        for (b = 0; b < order; b++) {
          double ab = qa * ty[b];
          int j1 = (b + i0s[3 * i + 1]) % ny;
          for (c = 0; c < order; c++) {
            int k1 = (c + i0s[3 * i + 2]) % nz;
            Q[2 * (((size_t)i1 * ny + j1) * nz + k1)] += ab * tz[c];
          }
        }
      }
    }
    if (qgrid) for (size_t t = 0; t < ng; t++) qgrid[t] = Q[2 * t];
    orc_fft3d(Q, nx, ny, nz, -1);   /* fftBack: exp(-i) */
    /* scalarSum (:276-389) */
    {
      double piVr = 1.0 / (ORC_PI * vg);
      int ii, jj, kk;
      for (ii = 0; ii < nx; ii++) {
        int i0 = ii <= nx / 2 ? ii : ii - nx;
        double mi = i0 * wr;
        double ex = expx[ii] * piVr;
        for (jj = 0; jj < ny; jj++) {
          int j0 = jj <= ny / 2 ? jj : jj - ny;
          double mxy = modx[ii] * mody[jj];
          double mj = j0 * lr;
          double mij = mi * mi + mj * mj;
          double exy = ex * expy[jj];
          for (kk = (ii != 0 || jj != 0 ? 0 : 1); kk <= nz / 2; kk++) {
            int k0 = kk <= nz / 2 ? kk : kk - nz;
            double mk = k0 * hr;
            double m2 = mij + mk * mk;
            double theta = mxy * modz[kk] * exy * expz[kk] / m2;
            size_t idx = ((size_t)ii * ny + jj) * nz + kk;
            double qq = Q[2 * idx] * Q[2 * idx] + Q[2 * idx + 1] * Q[2 * idx + 1];
            double e = qq * theta;
            double vv = 2.0 * (1.0 / m2 + fac);
            if (kk > 0 && ((kk != nz / 2) || (nz & 1))) {
              size_t w = ((size_t)((nx - ii) % nx) * ny + (ny - jj) % ny) * nz + (nz - kk) % nz;
              e *= 2.0;
              Q[2 * w] *= theta;
              Q[2 * w + 1] *= theta;
            }
            energy += e;
            if (want_virial) {
              vxx += e * (1.0 - vv * mi * mi);
              vxy -= e * (vv * mi * mj);
              vxz -= e * (vv * mi * mk);
              vyy += e * (1.0 - vv * mj * mj);
              vyz -= e * (vv * mj * mk);
              vzz += e * (1.0 - vv * mk * mk);
            }
            Q[2 * idx] *= theta;
            Q[2 * idx + 1] *= theta;
          }
        }
      }
      Q[0] = 0.0; Q[1] = 0.0;
      if (want_virial) {
        energies[ORC_E_VXX] += vxx * 0.5; energies[ORC_E_VXY] += vxy * 0.5; energies[ORC_E_VXZ] += vxz * 0.5;
        energies[ORC_E_VYX] += vxy * 0.5; energies[ORC_E_VYY] += vyy * 0.5; energies[ORC_E_VYZ] += vyz * 0.5;
        energies[ORC_E_VZX] += vxz * 0.5; energies[ORC_E_VZY] += vyz * 0.5; energies[ORC_E_VZZ] += vzz * 0.5;
      }
      e_recip = energy * 0.5;
    }
    orc_fft3d(Q, nx, ny, nz, +1);   /* fftForward: exp(+i) */
    /* interpolateForce (:193-228) */
    for (i = 0; i < n; i++) {
      double *tx = th + (size_t)6 * order * i, *ty = tx + order, *tz = ty + order;
      double *dx = tz + order, *dy = dx + order, *dz = dy + order;
      double fx = 0, fy = 0, fz = 0;
      int a, b, c;
      for (a = 0; a < order; a++) {
        int i1 = (a + i0s[3 * i]) % nx;
        for (b = 0; b < order; b++) {
          int j1 = (b + i0s[3 * i + 1]) % ny;
          double xij = dx[a] * ty[b];
          double yij = tx[a] * dy[b];
          double zij = tx[a] * ty[b];
          for (c = 0; c < order; c++) {
            int k1 = (c + i0s[3 * i + 2]) % nz;
            double term = Q[2 * (((size_t)i1 * ny + j1) * nz + k1)];
            fx -= term * xij * tz[c];
            fy -= term * yij * tz[c];
            fz -= term * zij * dz[c];
          }
        }
      }
      forces[3 * i] += fx * hxr * q[i];
      forces[3 * i + 1] += fy * hyr * q[i];
      forces[3 * i + 2] += fz * hzr * q[i];
    }
    free(Q); free(th); free(i0s);
    free(modx); free(mody); free(modz); free(expx); free(expy); free(expz);
  }
  energies[ORC_E_COULOMB] += e_real + e_recip + e_intra + e_self + e_charged;
  if (parts) {
    parts[0] = e_real; parts[1] = e_recip; parts[2] = e_intra;
    parts[3] = e_self; parts[4] = e_charged; parts[5] = alpha;
  }
  return 0;
}

/* ------------------------------------------------------------ integrator */
void orc_leapfrog_kickdrift(int n, double dt, double fac, const double* mass,
                            const double* f, double* v, double* x) {
  /* LeapfrogIntegrator::doHalfKickdoDrift / doKickdoDrift (:46-104) */
  double h = dt * ORC_INV_TIMEFACTOR;
  int i, k;
  for (i = 0; i < n; i++)
    for (k = 0; k < 3; k++) {
      if (fac == 0.5) v[3 * i + k] += f[3 * i + k] * h * 0.5 / mass[i];
      else v[3 * i + k] += f[3 * i + k] * h / mass[i];
      x[3 * i + k] += v[3 * i + k] * h;
    }
}

void orc_half_kick(int n, double dt, const double* mass, const double* f, double* v) {
  /* STSIntegrator::doHalfKick (integrators/STSIntegrator.cpp) */
  double h = 0.5 * dt * ORC_INV_TIMEFACTOR;
  int i, k;
  for (i = 0; i < n; i++)
    for (k = 0; k < 3; k++) v[3 * i + k] += f[3 * i + k] * h / mass[i];
}

void orc_langevin_half_kick(int n, double dt, double temperature, double gamma,
                            const double* mass, const double* f,
                            const double* xi, double* v) {
  /* LangevinLeapfrogIntegrator.cpp:36 (gamma conversion) and :58-79 */
  double g = gamma / (1000 * ORC_INV_TIMEFACTOR);
  double h = dt * ORC_INV_TIMEFACTOR;
  double fdt = (1.0 - exp(-0.5 * g * h)) / g;
  double vdt = exp(-0.5 * g * h);
  double ndt = sqrt((1.0 - exp(-g * h)) / (2.0 * g));
  double fc = 2 * ORC_BOLTZMANN * temperature * g;
  int i, k;
  for (i = 0; i < n; i++) {
    double m = mass[i];
    double s = sqrt(fc / m);
    for (k = 0; k < 3; k++)
      v[3 * i + k] = v[3 * i + k] * vdt + f[3 * i + k] * fdt / m + xi[3 * i + k] * s * ndt;
  }
}

double orc_kinetic_energy(int n, const double* mass, const double* v) {
  /* topologyutilities.cpp:182-191: 0.5 * sum m v^2 */
  double ke = 0.0;
  int i;
  for (i = 0; i < n; i++)
    ke += mass[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
  return 0.5 * ke;
}

/* ------------------------------------------------------------- Hessian */
/* 3x3 pair block of the switched potential, Hessian.cpp:456-624 with
 * ReducedHessLennardJones.cpp:7-50 / ReducedHessCoulomb.cpp:8-42 and the
 * switching hessians C2SwitchingFunction.cpp:45-53, C1SwitchingFunction.cpp:39-49. */
static void pair_hessian(const orc_term_state* s, const double* rij, double a,
                         double rawE, double rawF, double sv, double sd,
                         int i, int j, int excl, const double* q, const int* type,
                         int ntypes, const double* ljA, const double* ljB,
                         const double* ljA14, const double* ljB14, double csf,
                         double H[9]) {
  double rr[9], mz[9];
  int u, w;
  for (u = 0; u < 3; u++) for (w = 0; w < 3; w++) rr[3 * u + w] = rij[u] * rij[w];
  for (u = 0; u < 9; u++) mz[u] = 0.0;
  if (sv > 0.0 && sv < 1.0) {
    if (s->t.swf == ORC_SWF_C2) {
      double c2 = s->cut2, s2 = s->s_on2;
      double tm1 = c2 - s2, tm2 = tm1 * tm1 * tm1;
      double tm3 = 12.0 * (a - c2) * (a - s2) / tm2;
      double tm4 = 24.0 * (2.0 * a - c2 - s2) / tm2;
      for (u = 0; u < 9; u++) mz[u] = rr[u] * tm4;
      mz[0] += tm3; mz[4] += tm3; mz[8] += tm3;
    } else if (s->t.swf == ORC_SWF_C1) {
      double c = s->t.cutoff, rn = sqrt(a);
      double tm1 = 1.5 / (c * rn), tm2 = 1.5 * (1.0 / (c * c * c)) * rn;
      double tm3 = -tm1 + tm2, tm4 = tm1 + tm2;
      for (u = 0; u < 9; u++) mz[u] = (rr[u] / a) * tm4;
      mz[0] += tm3; mz[4] += tm3; mz[8] += tm3;
    }
  }
  if (s->t.pot == ORC_POT_LJ) {
    int ti = type[i], tj = type[j];
    double A = ljA[ti * ntypes + tj], B = ljB[ti * ntypes + tj];
    double a4 = a * a * a * a, a7 = a4 * a * a * a, a5 = a4 * a, a8 = a4 * a4;
    double tm1, tm2;
    if (excl == ORC_EXCL_MODIFIED) { A = ljA14[ti * ntypes + tj]; B = ljB14[ti * ntypes + tj]; }
    tm1 = 6 * B / a4 - 12 * A / a7;
    tm2 = -48 * B / a5 + 168 * A / a8;
    for (u = 0; u < 9; u++) H[u] = rr[u] * tm2;
    H[0] += tm1; H[4] += tm1; H[8] += tm1;
    for (u = 0; u < 9; u++) H[u] *= sv;
  } else {
    double cs = (excl == ORC_EXCL_MODIFIED) ? csf : 1.0;
    double na = sqrt(a);
    double tm1 = -cs * q[i] * q[j] / a / na;
    for (u = 0; u < 9; u++) {
      double id = (u == 0 || u == 4 || u == 8) ? 1.0 : 0.0;
      H[u] = (id - rr[u] * (3 / a)) * (tm1 * sv);
    }
  }
  for (u = 0; u < 9; u++) H[u] += mz[u] * rawE - rr[u] * (2 * rawF * sd);
}

int orc_hessian_vec(int n, const double* pos, const double* q, const int* type,
                    int ntypes, const double* ljA, const double* ljB,
                    const double* ljA14, const double* ljB14,
                    const int* molecule, int nexcl, const int* ex_a,
                    const int* ex_b, const int* ex_cls, double csf,
                    const double* box, const double* origin, int nterms,
                    const orc_term* terms, const double* v, double* hv) {
  orc_pbc pbc;
  orc_excl ex;
  orc_term_state ts[2];
  int i, j, t;
  (void)molecule;
  if (nterms < 1 || nterms > 2) return -1;
  pbc_set(&pbc, box, origin);
  excl_build(&ex, n, nexcl, ex_a, ex_b, ex_cls);
  for (t = 0; t < nterms; t++) {
    term_init(&ts[t], &terms[t]);
    if (terms[t].pot == ORC_POT_EWALDREAL) { excl_free(&ex); return -2; }  /* no reference Hessian */
  }
  for (i = 0; i < 3 * n; i++) hv[i] = 0.0;
  for (t = 0; t < nterms; t++) {
    for (i = 0; i < n; i++) {
      for (j = i + 1; j < n; j++) {
        int ec = excl_check(&ex, i, j);
        double rij[3], a, rawE = 0.0, rawF = 0.0, sv = 1.0, sd = 0.0, H[9], dv[3];
        int u;
        if (ec == ORC_EXCL_FULL) continue;
        pbc_min_diff(&pbc, pos + 3 * i, pos + 3 * j, rij, &a);
        /* Hessian.cpp uses the switching function when one is configured;
         * Cutoff switching takes the value 1/0 branch without derivatives */
        swf_eval(&ts[t], a, &sv, &sd);
        if (ts[t].t.swf == ORC_SWF_CUTOFF) sd = 0.0;
        if (sv > 0.0 && sv < 1.0)
          pot_eval(&ts[t], &rawE, &rawF, a, 1.0 / a, i, j, ec, q, type, ntypes,
                   ljA, ljB, ljA14, ljB14, csf);
        pair_hessian(&ts[t], rij, a, rawE, rawF, sv, sd, i, j, ec, q, type, ntypes,
                     ljA, ljB, ljA14, ljB14, csf, H);
        dv[0] = v[3 * i] - v[3 * j]; dv[1] = v[3 * i + 1] - v[3 * j + 1]; dv[2] = v[3 * i + 2] - v[3 * j + 2];
        for (u = 0; u < 3; u++) {
          double s = H[3 * u] * dv[0] + H[3 * u + 1] * dv[1] + H[3 * u + 2] * dv[2];
          hv[3 * i + u] += s;
          hv[3 * j + u] -= s;
        }
      }
    }
  }
  excl_free(&ex);
  return 0;
}

/* ------------------------------------------------------------- bonded */
static double nrm(const double* v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

void orc_bonded(int n, const double* pos, const double* box, const double* origin, int nb,
                const int* ba, const int* bb, const double* bk, const double* br0, int na,
                const int* aa, const int* ab, const int* ac, const double* ak, const double* at0,
                const double* akub, const double* arub, int which, int want_virial,
                double* forces, double* energies) {
  orc_pbc p;
  int k, c;
  (void)n;
  pbc_set(&p, box, origin);
  if (which & 1) {
    /* BondSystemForce::calcBond (forces/BondSystemForce.h:111-145) */
    for (k = 0; k < nb; k++) {
      int a1 = ba[k], a2 = bb[k];
      double r12[3], r2, r, dpotdr, f1[3];
      pbc_min_diff(&p, pos + 3 * a2, pos + 3 * a1, r12, &r2);
      r = nrm(r12);
      dpotdr = 2.0 * bk[k] * (r - br0[k]);
      for (c = 0; c < 3; c++) f1[c] = r12[c] * (-dpotdr / r);
      for (c = 0; c < 3; c++) { forces[3 * a1 + c] += f1[c]; forces[3 * a2 + c] -= f1[c]; }
      energies[ORC_E_BOND] += bk[k] * (r - br0[k]) * (r - br0[k]);
      if (want_virial) add_virial(energies, f1, r12);
    }
  }
  if (which & 2) {
    /* AngleSystemForce::calcAngle (forces/AngleSystemForce.h:113-200) */
    for (k = 0; k < na; k++) {
      int a1 = aa[k], a2 = ab[k], a3 = ac[k];
      double r12[3], r32[3], r13[3], t2, d12, d32, d13, cr[3], theta, st, ct, dpot;
      double f1[3], f2[3], f3[3], eh, eu;
      pbc_min_diff(&p, pos + 3 * a2, pos + 3 * a1, r12, &t2);
      pbc_min_diff(&p, pos + 3 * a2, pos + 3 * a3, r32, &t2);
      pbc_min_diff(&p, pos + 3 * a3, pos + 3 * a1, r13, &t2);
      d12 = nrm(r12); d32 = nrm(r32); d13 = nrm(r13);
      cr[0] = r12[1] * r32[2] - r12[2] * r32[1];
      cr[1] = r12[2] * r32[0] - r12[0] * r32[2];
      cr[2] = r12[0] * r32[1] - r12[1] * r32[0];
      theta = atan2(nrm(cr), r12[0] * r32[0] + r12[1] * r32[1] + r12[2] * r32[2]);
      st = sin(theta); ct = cos(theta);
      dpot = 2.0 * ak[k] * (theta - at0[k]);
      for (c = 0; c < 3; c++) {
        double dr12 = r12[c] / d12, dr32 = r32[c] / d32, dr13 = r13[c] / d13;
        double dth1 = (dr12 * ct - dr32) / (st * d12);
        double dth3 = (dr32 * ct - dr12) / (st * d32);
        double ub1 = dr13 * (2.0 * akub[k] * (d13 - arub[k]));
        f1[c] = -dth1 * dpot - ub1;
        f3[c] = -dth3 * dpot - (-ub1);
        f2[c] = -f1[c] - f3[c];
      }
      for (c = 0; c < 3; c++) {
        forces[3 * a1 + c] += f1[c];
        forces[3 * a2 + c] += f2[c];
        forces[3 * a3 + c] += f3[c];
      }
      eh = ak[k] * (theta - at0[k]) * (theta - at0[k]);
      eu = akub[k] * (d13 - arub[k]) * (d13 - arub[k]);
      energies[ORC_E_ANGLE] += eh + eu;
      if (want_virial) {
        double xy = f1[0] * r12[1] + f3[0] * r32[1];
        double xz = f1[0] * r12[2] + f3[0] * r32[2];
        double yz = f1[1] * r12[2] + f3[1] * r32[2];
        energies[ORC_E_VXX] += f1[0] * r12[0] + f3[0] * r32[0];
        energies[ORC_E_VXY] += xy;
        energies[ORC_E_VXZ] += xz;
        energies[ORC_E_VYX] += xy;
        energies[ORC_E_VYY] += f1[1] * r12[1] + f3[1] * r32[1];
        energies[ORC_E_VYZ] += yz;
        energies[ORC_E_VZX] += xz;
        energies[ORC_E_VZY] += yz;
        energies[ORC_E_VZZ] += f1[2] * r12[2] + f3[2] * r32[2];
      }
    }
  }
}

static void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* MTorsionSystemForce::calcTorsion (forces/MTorsionSystemForce.h:34-153):
 * phi from the unit normals a = r12 x r23, b = r23 x r34, c = r23 x a; the
 * force through d(cos phi) when |sin phi| > 0.1, else through d(sin phi). */
void orc_torsions(int n, const double* pos, const double* box, const double* origin, int nt, const int* atoms,
                  const int* term_row, const double* tk, const int* period, const double* phase, int improper,
                  int want_virial, double* forces, double* energies) {
  orc_pbc p;
  int t, c, u;
  const double pi = 3.14159265358979323846;
  (void)n;
  pbc_set(&p, box, origin);
  for (t = 0; t < nt; t++) {
    const int a1 = atoms[4 * t], a2 = atoms[4 * t + 1], a3 = atoms[4 * t + 2], a4 = atoms[4 * t + 3];
    double r12[3], r23[3], r34[3], d2, a[3], b[3], cc[3], ra, rb, rc, cosp, sinp, phi, dpot = 0.0, e = 0.0;
    double f1[3], f2[3], f3[3];
    pbc_min_diff(&p, pos + 3 * a2, pos + 3 * a1, r12, &d2);
    pbc_min_diff(&p, pos + 3 * a3, pos + 3 * a2, r23, &d2);
    pbc_min_diff(&p, pos + 3 * a4, pos + 3 * a3, r34, &d2);
    cross3(r12, r23, a);
    cross3(r23, r34, b);
    cross3(r23, a, cc);
    ra = 1.0 / nrm(a); rb = 1.0 / nrm(b); rc = 1.0 / nrm(cc);
    for (c = 0; c < 3; c++) { a[c] *= ra; b[c] *= rb; cc[c] *= rc; }
    cosp = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    sinp = cc[0] * b[0] + cc[1] * b[1] + cc[2] * b[2];
    phi = -atan2(sinp, cosp);
    for (u = term_row[t]; u < term_row[t + 1]; u++) {
      if (period[u] > 0) {
        dpot -= period[u] * tk[u] * sin(period[u] * phi + phase[u]);
        e += tk[u] * (1.0 + cos(period[u] * phi + phase[u]));
      } else {
        double diff = phi - phase[u];
        if (diff < -pi) diff += 2 * pi;
        else if (diff > pi) diff -= 2 * pi;
        dpot += 2.0 * tk[u] * diff;
        e += tk[u] * diff * diff;
      }
    }
    if (fabs(sinp) > 0.1) {
      double dA[3], dB[3], k1 = dpot / sinp;
      for (c = 0; c < 3; c++) { dA[c] = (a[c] * cosp - b[c]) * ra; dB[c] = (b[c] * cosp - a[c]) * rb; }
      f1[0] = k1 * (r23[1] * dA[2] - r23[2] * dA[1]);
      f1[1] = k1 * (r23[2] * dA[0] - r23[0] * dA[2]);
      f1[2] = k1 * (r23[0] * dA[1] - r23[1] * dA[0]);
      f3[0] = k1 * (r23[2] * dB[1] - r23[1] * dB[2]);
      f3[1] = k1 * (r23[0] * dB[2] - r23[2] * dB[0]);
      f3[2] = k1 * (r23[1] * dB[0] - r23[0] * dB[1]);
      f2[0] = k1 * (r12[2] * dA[1] - r12[1] * dA[2] + r34[1] * dB[2] - r34[2] * dB[1]);
      f2[1] = k1 * (r12[0] * dA[2] - r12[2] * dA[0] + r34[2] * dB[0] - r34[0] * dB[2]);
      f2[2] = k1 * (r12[1] * dA[0] - r12[0] * dA[1] + r34[0] * dB[1] - r34[1] * dB[0]);
    } else {
      double dC[3], dB[3], k1 = -dpot / cosp, x[3];
      for (c = 0; c < 3; c++) { dC[c] = (cc[c] * sinp - b[c]) * rc; dB[c] = (b[c] * sinp - cc[c]) * rb; }
      f1[0] = k1 * ((r23[1] * r23[1] + r23[2] * r23[2]) * dC[0] - r23[0] * r23[1] * dC[1] - r23[0] * r23[2] * dC[2]);
      f1[1] = k1 * ((r23[2] * r23[2] + r23[0] * r23[0]) * dC[1] - r23[1] * r23[2] * dC[2] - r23[1] * r23[0] * dC[0]);
      f1[2] = k1 * ((r23[0] * r23[0] + r23[1] * r23[1]) * dC[2] - r23[2] * r23[0] * dC[0] - r23[2] * r23[1] * dC[1]);
      cross3(dB, r23, x);
      for (c = 0; c < 3; c++) f3[c] = x[c] * k1;
      f2[0] = k1 * (-(r23[1] * r12[1] + r23[2] * r12[2]) * dC[0] + (2.0 * r23[0] * r12[1] - r12[0] * r23[1]) * dC[1] +
                    (2.0 * r23[0] * r12[2] - r12[0] * r23[2]) * dC[2] + dB[2] * r34[1] - dB[1] * r34[2]);
      f2[1] = k1 * (-(r23[2] * r12[2] + r23[0] * r12[0]) * dC[1] + (2.0 * r23[1] * r12[2] - r12[1] * r23[2]) * dC[2] +
                    (2.0 * r23[1] * r12[0] - r12[1] * r23[0]) * dC[0] + dB[0] * r34[2] - dB[2] * r34[0]);
      f2[2] = k1 * (-(r23[0] * r12[0] + r23[1] * r12[1]) * dC[2] + (2.0 * r23[2] * r12[0] - r12[2] * r23[0]) * dC[0] +
                    (2.0 * r23[2] * r12[1] - r12[2] * r23[1]) * dC[1] + dB[1] * r34[0] - dB[0] * r34[1]);
    }
    for (c = 0; c < 3; c++) {
      forces[3 * a1 + c] += f1[c];
      forces[3 * a2 + c] += f2[c] - f1[c];
      forces[3 * a3 + c] += f3[c] - f2[c];
      forces[3 * a4 + c] -= f3[c];
    }
    energies[improper ? ORC_E_IMPROPER : ORC_E_DIHEDRAL] += e;
    if (want_virial) {
      double xy = f1[0] * r12[1] + f2[0] * r23[1] + f3[0] * r34[1];
      double xz = f1[0] * r12[2] + f2[0] * r23[2] + f3[0] * r34[2];
      double yz = f1[1] * r12[2] + f2[1] * r23[2] + f3[1] * r34[2];
      energies[ORC_E_VXX] += f1[0] * r12[0] + f2[0] * r23[0] + f3[0] * r34[0];
      energies[ORC_E_VXY] += xy;
      energies[ORC_E_VXZ] += xz;
      energies[ORC_E_VYX] += xy;
      energies[ORC_E_VYY] += f1[1] * r12[1] + f2[1] * r23[1] + f3[1] * r34[1];
      energies[ORC_E_VYZ] += yz;
      energies[ORC_E_VZX] += xz;
      energies[ORC_E_VZY] += yz;
      energies[ORC_E_VZZ] += f1[2] * r12[2] + f2[2] * r23[2] + f3[2] * r34[2];
    }
  }
}

double orc_rem_delta(double ti, double tj, double ei, double ej) {
  /* rem.cpp:941-946 */
  return (1 / ORC_BOLTZMANN / ti - 1 / ORC_BOLTZMANN / tj) * (ej - ei);
}

/* ------------------------------------------------------ SHAKE / RATTLE */
/* topologyutilities.cpp:782-865 */
int orc_constraint_list(int n, const double* mass, int nb, const int* ba, const int* bb, const double* br0,
                        int na, const int* aa, const int* ab, const int* ac, const double* at0, int* c1, int* c2,
                        double* cr0) {
  int m = 0;
  for (int i = 0; i < nb; i++) {
    c1[m] = ba[i];
    c2[m] = bb[i];
    cr0[m] = br0[i];
    m++;
  }
  (void)n;
  for (int i = 0; i < na; i++) {
    const int a1 = aa[i], a2 = ab[i], a3 = ac[i];
    if (!(mass[a1] < 5 && mass[a3] < 5)) continue;
    const int p1a = a1 < a2 ? a1 : a2, p1b = a1 < a2 ? a2 : a1;
    const int p2a = a2 < a3 ? a2 : a3, p2b = a2 < a3 ? a3 : a2;
    int b1 = -1, b2 = -1;
    /* atoms[a].mybonds: the bonds of atom a in bond order */
    for (int j = 0; j < nb; j++) {
      if (ba[j] != a1 && bb[j] != a1) continue;
      const int s1a = ba[j] < bb[j] ? ba[j] : bb[j], s1b = ba[j] < bb[j] ? bb[j] : ba[j];
      if (s1a != p1a || s1b != p1b) continue;
      for (int k = 0; k < nb; k++) {
        if (ba[k] != a3 && bb[k] != a3) continue;
        const int s2a = ba[k] < bb[k] ? ba[k] : bb[k], s2b = ba[k] < bb[k] ? bb[k] : ba[k];
        if (s2a != p2a || s2b != p2b) continue;
        if (!(b1 > -1 || b2 > -1)) { b1 = j; b2 = k; }
      }
    }
    if (b1 > -1 && b2 > -1) {
      const double b = br0[b1], c = br0[b2], alpha = at0[i];
      c1[m] = a1;
      c2[m] = a3;
      cr0[m] = sqrt(b * b + c * c - 2 * b * c * cos(alpha));
      m++;
    }
  }
  return m;
}

/* ModifierMetaShake::calcError (ModifierMetaShake.cpp:12-31) */
static double shake_error(int m, const int* c1, const int* c2, const double* cr0, const double* x) {
  double error = 0;
  for (int i = 0; i < m; i++) {
    const int a1 = c1[i], a2 = c2[i];
    const double r0sq = cr0[i] * cr0[i];
    const double dx = x[3 * a2] - x[3 * a1], dy = x[3 * a2 + 1] - x[3 * a1 + 1], dz = x[3 * a2 + 2] - x[3 * a1 + 2];
    const double err = fabs((dx * dx + dy * dy + dz * dz) - r0sq) / (2.0 * r0sq);
    error += err;
  }
  return error /= m;
}

/* ModifierShake::doExecute (ModifierShake.cpp:21-103); the virial is not kept */
int orc_shake(int m, const int* c1, const int* c2, const double* cr0, const double* mass, double dt_fs,
              double eps, int max_iter, double* x, double* v, double* x_last) {
  double error = shake_error(m, c1, c2, cr0, x);
  const double dt = dt_fs / 48.88821290839616;   /* Constant::TIMEFACTOR */
  int iter = 0;
  while (error > eps) {
    for (int i = 0; i < m; i++) {
      const int a1 = c1[i], a2 = c2[i];
      const double restLength = cr0[i];
      double pab[3], rab[3], dp[3];
      for (int d = 0; d < 3; d++) pab[d] = x[3 * a1 + d] - x[3 * a2 + d];
      const double pabsq = pab[0] * pab[0] + pab[1] * pab[1] + pab[2] * pab[2];
      const double rabsq = restLength * restLength;
      const double diffsq = rabsq - pabsq;
      for (int d = 0; d < 3; d++) rab[d] = x_last[3 * a1 + d] - x_last[3 * a2 + d];
      const double rpab = rab[0] * pab[0] + rab[1] * pab[1] + rab[2] * pab[2];
      const double rM1 = 1 / mass[a1], rM2 = 1 / mass[a2];
      const double gab = diffsq / (2 * (rM1 + rM2) * rpab);
      for (int d = 0; d < 3; d++) dp[d] = rab[d] * gab;
      for (int d = 0; d < 3; d++) {
        x[3 * a1 + d] += dp[d] * rM1;
        x[3 * a2 + d] -= dp[d] * rM2;
      }
      for (int d = 0; d < 3; d++) dp[d] /= dt;
      for (int d = 0; d < 3; d++) {
        v[3 * a1 + d] += dp[d] * rM1;
        v[3 * a2 + d] -= dp[d] * rM2;
      }
    }
    error = shake_error(m, c1, c2, cr0, x);
    iter++;
    if (iter > max_iter) break;
  }
  /* myLastPositions = positions (every atom; only constrained ones are read) */
  for (int i = 0; i < m; i++) {
    for (int d = 0; d < 3; d++) {
      x_last[3 * c1[i] + d] = x[3 * c1[i] + d];
      x_last[3 * c2[i] + d] = x[3 * c2[i] + d];
    }
  }
  return iter;
}

/* ModifierMetaRattle::calcError (ModifierMetaRattle.cpp:12-31) */
static double rattle_error(int m, const int* c1, const int* c2, const double* x, const double* v) {
  double error = 0;
  for (int i = 0; i < m; i++) {
    const int a1 = c1[i], a2 = c2[i];
    double vab[3], rab[3];
    for (int d = 0; d < 3; d++) {
      vab[d] = v[3 * a1 + d] - v[3 * a2 + d];
      rab[d] = x[3 * a1 + d] - x[3 * a2 + d];
    }
    const double err = fabs(rab[0] * vab[0] + rab[1] * vab[1] + rab[2] * vab[2]);
    error += err;
  }
  return error /= m;
}

/* ModifierRattle::doExecute (ModifierRattle.cpp:20-85); the virial is not kept */
int orc_rattle(int m, const int* c1, const int* c2, const double* mass, double dt_fs, double eps, int max_iter,
               const double* x, double* v) {
  double error = rattle_error(m, c1, c2, x, v);
  const double dt = dt_fs / 48.88821290839616;
  int iter = 0;
  while (error > eps) {
    for (int i = 0; i < m; i++) {
      const int a1 = c1[i], a2 = c2[i];
      const double rM1 = 1 / mass[a1], rM2 = 1 / mass[a2];
      double rab[3], vab[3], dp[3];
      for (int d = 0; d < 3; d++) rab[d] = x[3 * a1 + d] - x[3 * a2 + d];
      const double rabsq = rab[0] * rab[0] + rab[1] * rab[1] + rab[2] * rab[2];
      for (int d = 0; d < 3; d++) vab[d] = v[3 * a1 + d] - v[3 * a2 + d];
      const double rvab = rab[0] * vab[0] + rab[1] * vab[1] + rab[2] * vab[2];
      const double gab = -rvab / (dt * (rM1 + rM2) * rabsq);
      for (int d = 0; d < 3; d++) dp[d] = rab[d] * gab;
      for (int d = 0; d < 3; d++) {
        v[3 * a1 + d] += dp[d] * dt * rM1;
        v[3 * a2 + d] -= dp[d] * dt * rM2;
      }
    }
    error = rattle_error(m, c1, c2, x, v);
    iter++;
    if (iter > max_iter) break;
  }
  return iter;
}
