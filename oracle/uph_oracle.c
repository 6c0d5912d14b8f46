/*
 * uph_oracle.c -- FP64 CPU restatement of the reference's per-step SPH hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is
 * compared against, and the CPU baseline that bench.py times.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the product
 * library (uph_amd/libuph_gpu.so) never links or calls it.
 *
 * It restates, in the reference's own pair-list + scatter form and FP64
 * throughout, the "config-mode" scheme:
 *   - neighbour contract of SearchNeighbors           Search.cc:5-44
 *   - W / grad W of the cubic spline                   derivatives.c:44-66
 *   - pair kernel evaluation                           derivatives.c:1063-1106
 *   - continuity                                       derivatives.c:1520-1567
 *   - predicting continuity (order dependent)          derivatives.c:1217-1252
 *   - Tait equation of state (feeds the momentum eq.)  derivatives.c:1674-1682
 *   - pressure gradient + predicting variant           derivatives.c:1109-1163, 1333-1377
 *   - Monaghan (1992) artificial viscosity added to both pressure terms
 *     (absent from the reference: alpha is parsed at System.c:75 but unused)
 *   - gravity                                          derivatives.c:1167-1178, 1379-1390
 *   - spring boundary force + predicting variant       derivatives.c:2396-2449, 1392-1446
 *   - boundary reset and moving-set kinematics         inputbp.c:117-150, 192-283
 *   - mirror ("virtual") particles by virtualType     derivatives.c:3193-4192
 *     (ranked by the neighbour search, never paired)   derivatives.c:1069-1085
 *   - leapfrog predictor / corrector                   sph.cc:79-168
 * with the fluctuation fields and epsilon terms zero (they multiply fields the
 * scheme does not evolve).  The reference's own build cannot be compiled in
 * this image (it needs <mkl.h> and libnabo), so nothing here is linked against
 * it; each function is pinned by the closed forms, brute-force checks and
 * invariants in tests/test_oracle.py.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
  double dx, h, rho0, gamma, c0, B, g;
  int gravity;
  double alpha_av, k_spring, c_damp, dt;
  int K, motion;
  double m_amp, m_omega;
  int kernel;                           /* 0: cubic spline (derivatives.c:44-66), 1: PolyOrderSeven (:767-823) */
  int mode;                             /* 0: config mode, 1: the reference's own scheme (ref mode) */
  double kappa, mu, eps;                /* ref mode: thermalDiffusivity0, viscosity0, epsilon */
  double lid_u, depth, flat_bottom;     /* lidVelocity, LocalDepth, FlatBottomLength (motions 3, 4, 5) */
  int vtype;                            /* virtualType: mirror particles of derivatives.c:3193-4192 */
  int stress_type;                      /* ref mode: stressType, 1 single-step viscous forces, 2 two-step
                                           (ClassicShearStress + ViscousForcesTwoStep) */
} orc_params;

typedef struct {
  int nf, nfixed, nmoving, n;
  orc_params p;
  double *x, *v, *vmin, *vloc, *rho, *rhomin, *rholoc, *P, *mass, *hsml;
  int *itype;
  double *R, *Rt, *a, *at;              /* ddensitydt, ddensitydt_approx, dvdt, dvdt_approx */
  /* ref mode: DynamicPressure_min, localPressure, dpdt, dpdt_approx, the fluctuation
   * fields, the pair kernel values w, spongeFactor */
  double *Pmin, *Ploc, *RP, *RPt, *drho, *dP, *dv, *w, *sponge;
  double *stress;                       /* ViscousStress{xx,xy,yx,yy} per particle (stressType 2) */
  double *gate_y;                       /* tempPositionGate (deck column 9) */
  /* mirror particles of the last derivatives() call: rows n .. n+ng-1 of x, and their
   * Velocity, localVelocity, density, localDensity, DynamicPressure, localPressure, mass */
  int ng, gcap;
  double *gv, *glv, *grho, *grl, *gP, *gPl, *gmass;
  double lbp, rbp;                      /* LeftBoundaryPosition, RightBoundaryPosition */
  double *deck_x, *deck_v, *deck_rho, *deck_P, *deck_lvel;
  int *ip, *jp;
  double *rij2, *dwdx;
  long cap;
  int niac;
  int itime;
  int nthreads;
  double *tmp;                          /* nthreads x 2n scatter buffers */
} orc_sim;

int orc_ghost_sections(int vtype, int *sec);

static int nthreads_of(int want) {
#ifdef _OPENMP
  return want > 0 ? want : omp_get_max_threads();
#else
  (void)want;
  return 1;
#endif
}

/* ------------------------------------------------------------------ kernel */

/* Cubic spline W and grad W (derivatives.c:44-66): alpha = 10/(7 pi h^2). */
void orc_cubic(const double *dx, double r, double h, double *w, double *dw) {
  const double alpha = 10.0 / (7. * M_PI * h * h);
  const double q = r / h;
  if (q >= 0.0 && q < 1.0) {
    double q2 = q * q;
    double q3 = q * q * q;
    *w = alpha * (1 - 1.5 * q2 + 0.75 * q3);
    for (int d = 0; d < 2; d++) dw[d] = alpha * (-3 + 2.25 * q) * dx[d] / (h * h);
  } else if (q >= 1.0 && q < 2.0) {
    double qp2 = (2 - q) * (2 - q);
    double qp3 = qp2 * (2 - q);
    *w = alpha * 0.25 * qp3;
    for (int d = 0; d < 2; d++) dw[d] = -alpha * 0.75 * qp2 * dx[d] / (h * r);
  } else {
    *w = 0.0;
    dw[0] = dw[1] = 0.0;
  }
}

/* PolyOrderSeven W (PolyOrderSevenConvolutionKernel, derivatives.c:767-784):
 * W = a (1 - q^2/4)^8 (2048 - 3072 q^2 + 2816 q^4 - 1984 q^6 + 1184 q^8 - 628 q^10 + 305 q^12),
 * a = 2145/(1347584 pi h^2), for q <= 2. */
double orc_poly7_w(double r, double h) {
  const double a = 2145.0 / (1347584.0 * M_PI * h * h);
  const double q = r / h;
  if (!(q >= 0.0 && q <= 2.0)) return 0.0;
  const double u = q * q, s = 1 - 0.25 * u;
  const double s2 = s * s, s4 = s2 * s2, s8 = s4 * s4;
  const double poly = 2048 - 3072 * u + 2816 * u * u - 1984 * u * u * u + 1184 * u * u * u * u -
                      628 * u * u * u * u * u + 305 * u * u * u * u * u * u;
  return a * s8 * poly;
}

/* PolyOrderSeven grad W (PolyOrderSevenDeconvolutionKernel, derivatives.c:786-823).
 * Its correction coefficients -131/329 and 150882/1840097 (:792-793) are integer
 * divisions and evaluate to 0, so only the first term is left:
 * grad W = -a0 (1 - q^2/4)^7 (2048 - 3584 q^2 + 3712 q^4 - 2912 q^6 + 1912 q^8 - 1106 q^10 + 305 q^12) dx / h^2,
 * a0 = 2145/(192512 pi h^2).  CalculateKernel (:1099-1101) pairs it with the W above. */
void orc_poly7(const double *dx, double r, double h, double *w, double *dw) {
  const double a0 = 2145.0 / (192512.0 * M_PI * h * h);
  const double q = r / h;
  *w = orc_poly7_w(r, h);
  if (!(q >= 0.0 && q <= 2.0)) {
    dw[0] = dw[1] = 0.0;
    return;
  }
  const double u = q * q, s = 1 - 0.25 * u;
  const double s2 = s * s, s4 = s2 * s2, s7 = s4 * s2 * s;
  const double poly = 2048 - 3584 * u + 3712 * u * u - 2912 * u * u * u + 1912 * u * u * u * u -
                      1106 * u * u * u * u * u + 305 * u * u * u * u * u * u;
  for (int d = 0; d < 2; d++) dw[d] = -a0 * s7 * poly * dx[d] / (h * h);
}

/* ------------------------------------------------------- neighbour search */

typedef struct { double d2; int j; } cand_t;

static int cand_less(const cand_t *a, const cand_t *b) {
  return a->d2 < b->d2 || (a->d2 == b->d2 && a->j < b->j);
}

static int cand_cmp(const void *pa, const void *pb) {
  const cand_t *a = pa, *b = pb;
  if (cand_less(a, b)) return -1;
  if (cand_less(b, a)) return 1;
  return 0;
}

static inline double dist2(const double *pos, int s, int j) {
  const double ex = pos[2 * s] - pos[2 * j];
  const double ey = pos[2 * s + 1] - pos[2 * j + 1];
  return ex * ex + ey * ey;
}

/*
 * Brute-force statement of the SearchNeighbors contract (Search.cc:5-44):
 * for each fluid query s the K nearest particles other than s itself, ranked
 * by (d2, index); keep j if j > s, d2 < dist and j < ParticleBoundary; emit in
 * rank order.  O(N^2 log N): for small inputs only.
 */
int orc_search_brute(const double *pos, int *ip, int *jp, double *rij2, int nf, int n, int pb,
                     int K, double dist, long cap) {
  cand_t *c = malloc(sizeof(cand_t) * (size_t)(n > 1 ? n : 1));
  long niac = 0;
  for (int s = 0; s < nf; s++) {
    int m = 0;
    for (int j = 0; j < n; j++) {
      if (j == s) continue;
      c[m].d2 = dist2(pos, s, j);
      c[m].j = j;
      m++;
    }
    qsort(c, m, sizeof(cand_t), cand_cmp);
    int kk = m < K ? m : K;
    for (int k = 0; k < kk; k++) {
      if (c[k].j > s && c[k].d2 < dist && c[k].j < pb) {
        if (niac >= cap) { free(c); return -1; }
        ip[niac] = s; jp[niac] = c[k].j; rij2[niac] = c[k].d2;
        niac++;
      }
    }
  }
  free(c);
  return (int)niac;
}

typedef struct {
  int ncx, ncy;
  double x0, y0, inv;
  int *start;   /* ncx*ncy + 1 */
  int *idx;     /* particle ids sorted by cell, ascending id within a cell */
} grid_t;

static void grid_build(grid_t *g, const double *pos, int n, double cell) {
  double lx = pos[0], ly = pos[1], hx = pos[0], hy = pos[1];
  for (int i = 1; i < n; i++) {
    if (pos[2 * i] < lx) lx = pos[2 * i];
    if (pos[2 * i] > hx) hx = pos[2 * i];
    if (pos[2 * i + 1] < ly) ly = pos[2 * i + 1];
    if (pos[2 * i + 1] > hy) hy = pos[2 * i + 1];
  }
  /* widely spread inputs: coarser cells keep the 3x3 stencil exact (cell >= radius) */
  while ((floor((hx - lx) / cell) + 1) * (floor((hy - ly) / cell) + 1) > 4.0 * n + 1024) cell *= 2.0;
  g->x0 = lx; g->y0 = ly; g->inv = 1.0 / cell;
  g->ncx = (int)floor((hx - lx) * g->inv) + 1;
  g->ncy = (int)floor((hy - ly) * g->inv) + 1;
  long nc = (long)g->ncx * g->ncy;
  g->start = calloc(nc + 1, sizeof(int));
  g->idx = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int *cell_of = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++) {
    int cx = (int)floor((pos[2 * i] - lx) * g->inv);
    int cy = (int)floor((pos[2 * i + 1] - ly) * g->inv);
    if (cx >= g->ncx) cx = g->ncx - 1;
    if (cy >= g->ncy) cy = g->ncy - 1;
    cell_of[i] = cy * g->ncx + cx;
    g->start[cell_of[i] + 1]++;
  }
  for (long c = 0; c < nc; c++) g->start[c + 1] += g->start[c];
  int *fill = malloc(sizeof(int) * (size_t)(nc > 0 ? nc : 1));
  for (long c = 0; c < nc; c++) fill[c] = g->start[c];
  for (int i = 0; i < n; i++) g->idx[fill[cell_of[i]]++] = i;
  free(fill);
  free(cell_of);
}

static void grid_free(grid_t *g) { free(g->start); free(g->idx); }

/* In-range candidates of query s (d2 < dist, j != s), sorted by (d2, j). */
static int grid_candidates(const grid_t *g, const double *pos, int s, double dist, cand_t **buf,
                           int *bufcap) {
  int cx = (int)floor((pos[2 * s] - g->x0) * g->inv);
  int cy = (int)floor((pos[2 * s + 1] - g->y0) * g->inv);
  if (cx >= g->ncx) cx = g->ncx - 1;
  if (cy >= g->ncy) cy = g->ncy - 1;
  int m = 0;
  for (int yy = cy - 1; yy <= cy + 1; yy++) {
    if (yy < 0 || yy >= g->ncy) continue;
    for (int xx = cx - 1; xx <= cx + 1; xx++) {
      if (xx < 0 || xx >= g->ncx) continue;
      int c = yy * g->ncx + xx;
      for (int t = g->start[c]; t < g->start[c + 1]; t++) {
        int j = g->idx[t];
        if (j == s) continue;
        double d2 = dist2(pos, s, j);
        if (!(d2 < dist)) continue;
        if (m == *bufcap) {
          *bufcap = *bufcap * 2 + 64;
          *buf = realloc(*buf, sizeof(cand_t) * (size_t)*bufcap);
        }
        (*buf)[m].d2 = d2;
        (*buf)[m].j = j;
        m++;
      }
    }
  }
  /* insertion sort: candidate lists are short */
  cand_t *c = *buf;
  for (int a = 1; a < m; a++) {
    cand_t key = c[a];
    int b = a - 1;
    while (b >= 0 && cand_less(&key, &c[b])) { c[b + 1] = c[b]; b--; }
    c[b + 1] = key;
  }
  return m;
}

/*
 * Cell-grid statement of the same contract.  When at most K particles lie
 * inside the search radius they are all among the K nearest, so only in-range
 * candidates need ranking; otherwise the first K of them by (d2, j) are kept.
 */
int orc_search(const double *pos, int *ip, int *jp, double *rij2, int nf, int n, int pb, int K,
               double dist, long cap, int nthreads) {
  if (nf <= 0) return 0;
  grid_t g;
  grid_build(&g, pos, n, sqrt(dist) * (1.0 + 1e-9));
  int *cnt = calloc((size_t)nf + 1, sizeof(int));
  int nt = nthreads_of(nthreads);
#pragma omp parallel num_threads(nt)
  {
    int bc = 128;
    cand_t *buf = malloc(sizeof(cand_t) * bc);
#pragma omp for schedule(dynamic, 256)
    for (int s = 0; s < nf; s++) {
      int m = grid_candidates(&g, pos, s, dist, &buf, &bc);
      int kk = m < K ? m : K;
      int q = 0;
      for (int k = 0; k < kk; k++)
        if (buf[k].j > s && buf[k].j < pb) q++;
      cnt[s + 1] = q;
    }
    free(buf);
  }
  long *off = malloc(sizeof(long) * ((size_t)nf + 1));
  off[0] = 0;
  for (int s = 0; s < nf; s++) off[s + 1] = off[s] + cnt[s + 1];
  long total = off[nf];
  if (total > cap) { free(off); free(cnt); grid_free(&g); return -1; }
#pragma omp parallel num_threads(nt)
  {
    int bc = 128;
    cand_t *buf = malloc(sizeof(cand_t) * bc);
#pragma omp for schedule(dynamic, 256)
    for (int s = 0; s < nf; s++) {
      int m = grid_candidates(&g, pos, s, dist, &buf, &bc);
      int kk = m < K ? m : K;
      long o = off[s];
      for (int k = 0; k < kk; k++)
        if (buf[k].j > s && buf[k].j < pb) {
          ip[o] = s; jp[o] = buf[k].j; rij2[o] = buf[k].d2;
          o++;
        }
    }
    free(buf);
  }
  free(off);
  free(cnt);
  grid_free(&g);
  return (int)total;
}

/* ---------------------------------------------------------------- system */

orc_sim *orc_create(int nf, int nfixed, int nmoving, const orc_params *p, const double *pos,
                    const double *vel, const double *rho, const double *pres, const double *mass,
                    const int *itype, const double *lvel, int nthreads) {
  orc_sim *s = calloc(1, sizeof(orc_sim));
  int n = nf + nfixed + nmoving;
  s->nf = nf; s->nfixed = nfixed; s->nmoving = nmoving; s->n = n;
  s->p = *p;
  s->nthreads = nthreads_of(nthreads);
  size_t n1 = (size_t)(n > 0 ? n : 1);
  /* at most one mirror particle per fluid particle and section */
  s->gcap = nf * orc_ghost_sections(p->vtype, NULL);
  const size_t g1 = (size_t)(s->gcap > 0 ? s->gcap : 1);
  s->x = calloc(2 * (n1 + g1), sizeof(double));
#define G(f, c) s->f = calloc(g1 * (c), sizeof(double))
  G(gv, 2); G(glv, 2); G(grho, 1); G(grl, 1); G(gP, 1); G(gPl, 1); G(gmass, 1);
#undef G
#define A(f, c) s->f = calloc(n1 * (c), sizeof(double))
  A(v, 2); A(vmin, 2); A(vloc, 2); A(rho, 1); A(rhomin, 1); A(rholoc, 1); A(P, 1);
  A(mass, 1); A(hsml, 1); A(R, 1); A(Rt, 1); A(a, 2); A(at, 2);
  A(Pmin, 1); A(Ploc, 1); A(RP, 1); A(RPt, 1); A(drho, 1); A(dP, 1); A(dv, 2); A(sponge, 1);
  A(gate_y, 1); A(stress, 4);
  A(deck_x, 2); A(deck_v, 2); A(deck_rho, 1); A(deck_P, 1); A(deck_lvel, 2);
#undef A
  s->itype = calloc(n1, sizeof(int));
  memcpy(s->x, pos, sizeof(double) * 2 * n);
  memcpy(s->v, vel, sizeof(double) * 2 * n);
  memcpy(s->vloc, lvel, sizeof(double) * 2 * n);
  memcpy(s->rho, rho, sizeof(double) * n);
  memcpy(s->P, pres, sizeof(double) * n);
  memcpy(s->mass, mass, sizeof(double) * n);
  memcpy(s->itype, itype, sizeof(int) * n);
  memcpy(s->deck_x, pos, sizeof(double) * 2 * n);
  memcpy(s->deck_v, vel, sizeof(double) * 2 * n);
  memcpy(s->deck_rho, rho, sizeof(double) * n);
  memcpy(s->deck_P, pres, sizeof(double) * n);
  memcpy(s->deck_lvel, lvel, sizeof(double) * 2 * n);
  for (int i = 0; i < n; i++) {
    s->hsml[i] = p->h;
    if (p->mode == 1) {
      /* input.c:109-110: localDensity = rho0, localPressure = 0 */
      s->rholoc[i] = p->rho0;
      s->Ploc[i] = 0.0;
      /* input.c:137-143 with System.c:333-335 (spongeFactor0 = 0.2, spongePosition0 = 7,
       * waveAbsorption = 1); SpongeLayerWaveReflectionAbsorption acts for movingBoundary == 1 */
      s->sponge[i] = (i < nf && p->motion == 1 && pos[2 * i] >= 7.0) ? 0.2 : 0.0;
    } else {
      /* config-mode choice: localDensity starts equal to density (input.c:109 sets rho0) */
      s->rholoc[i] = i < nf ? rho[i] : p->rho0;
    }
  }
  memcpy(s->rhomin, rho, sizeof(double) * n);
  memcpy(s->vmin, vel, sizeof(double) * 2 * n);
  s->cap = (long)nf * (p->K > 0 ? p->K : 1) + 16;
  s->ip = malloc(sizeof(int) * s->cap);
  s->jp = malloc(sizeof(int) * s->cap);
  s->rij2 = malloc(sizeof(double) * s->cap);
  s->dwdx = malloc(sizeof(double) * 2 * s->cap);
  s->w = malloc(sizeof(double) * s->cap);
  s->tmp = malloc(sizeof(double) * 2 * n1 * s->nthreads);
  s->itime = 0;
  return s;
}

void orc_destroy(orc_sim *s) {
  if (!s) return;
  free(s->x); free(s->v); free(s->vmin); free(s->vloc); free(s->rho); free(s->rhomin);
  free(s->rholoc); free(s->P); free(s->mass); free(s->hsml); free(s->R); free(s->Rt);
  free(s->a); free(s->at); free(s->deck_x); free(s->deck_v); free(s->deck_rho); free(s->deck_P);
  free(s->deck_lvel); free(s->itype); free(s->ip); free(s->jp); free(s->rij2); free(s->dwdx);
  free(s->tmp); free(s->Pmin); free(s->Ploc); free(s->RP); free(s->RPt); free(s->drho);
  free(s->dP); free(s->dv); free(s->w); free(s->sponge); free(s->gate_y); free(s->stress);
  free(s->gv); free(s->glv); free(s->grho); free(s->grl); free(s->gP); free(s->gPl); free(s->gmass);
  free(s);
}

/* Boundary reset and wall kinematics (inputbp.c:117-150, 194-207, 285-290). */
static void boundary_reset(orc_sim *s, double t) {
  const orc_params *p = &s->p;
  for (int i = s->nf; i < s->n; i++) {
    s->x[2 * i] = s->deck_x[2 * i]; s->x[2 * i + 1] = s->deck_x[2 * i + 1];
    s->v[2 * i] = s->deck_v[2 * i]; s->v[2 * i + 1] = s->deck_v[2 * i + 1];
    s->vloc[2 * i] = s->deck_lvel[2 * i]; s->vloc[2 * i + 1] = s->deck_lvel[2 * i + 1];
    s->rho[i] = s->deck_rho[i];
    s->P[i] = s->deck_P[i];
    s->hsml[i] = p->h;
    s->rholoc[i] = p->rho0;
    s->Ploc[i] = 0.0;
  }
  /* LeftBoundaryPosition: the piston's or the sloshing tank's, 0 in inputbp's no-motion
   * branch (inputbp.c:279-281); the lid and gate branches leave it (and
   * RightBoundaryPosition) unset -- 0 here */
  s->lbp = 0.0;
  if (p->motion == 1 || p->motion == 2) {
    double X, V;
    if (p->motion == 1) {          /* reference piston, absolute position */
      X = 0.5 * p->m_amp * (1.0 - cos(p->m_omega * t));
      V = 0.5 * p->m_amp * p->m_omega * sin(p->m_omega * t);
    } else {                       /* sinusoidal offset of the deck position */
      X = p->m_amp * sin(p->m_omega * t);
      V = p->m_amp * p->m_omega * cos(p->m_omega * t);
    }
    for (int i = s->nf + s->nfixed; i < s->n; i++) {
      s->x[2 * i] = (p->motion == 1) ? X : s->deck_x[2 * i] + X;
      s->v[2 * i] = V;
      s->vloc[2 * i] = V;
    }
    if (p->motion == 1) s->lbp = X;
  } else if (p->motion == 3) {     /* lid (inputbp.c:209-223): tempPosition is the deck x */
    for (int i = s->nf + s->nfixed; i < s->n; i++) {
      s->x[2 * i] = p->lid_u * t + s->deck_x[2 * i];
      s->v[2 * i] = p->lid_u;
      s->vloc[2 * i] = p->lid_u;
    }
  } else if (p->motion == 4) {     /* gate (inputbp.c:225-243) */
    const double stop = 1.5 * (p->depth - 15.0 * p->dx) / p->lid_u;
    for (int i = s->nf + s->nfixed; i < s->n; i++) {
      if (t > 0.0 && t < stop) {
        s->x[2 * i + 1] = p->lid_u * t + s->gate_y[i];
        s->v[2 * i + 1] = p->lid_u;
      } else {
        s->x[2 * i + 1] = p->lid_u * stop + s->gate_y[i];
        s->v[2 * i + 1] = 0.0;
      }
    }
  } else if (p->motion == 5) {     /* sloshing tank (inputbp.c:244-276), by deck itype */
    double L = 0.5 * p->m_amp * (1.0 - cos(p->m_omega * t));
    double V = 0.5 * p->m_amp * p->m_omega * sin(p->m_omega * t);
    const double R = p->flat_bottom + L;
    const int first = s->nf + s->nfixed;
    for (int i = first; i < s->n; i++) {
      if (s->itype[i] == 2) {
        s->x[2 * i] = L;
        s->v[2 * i] = V;
      } else if (s->itype[i] == 3) {
        s->x[2 * i] = R;
        s->v[2 * i] = V;
      } else if (s->itype[i] == 1) {
        s->x[2 * i] = (i - first) * p->dx * 0.5 + L;
        s->v[2 * i] = V;
      } else {                     /* the loop zeroes the left wall's motion for later rows */
        L = 0.0;
        V = 0.0;
      }
    }
    s->lbp = L;
    s->rbp = R;
  }
}

/* tempPositionGate of every particle (deck column 9), for the gate motion. */
void orc_set_gate_y(orc_sim *s, const double *gate_y) {
  memcpy(s->gate_y, gate_y, sizeof(double) * s->n);
}

/* Monaghan (1992) viscous term Pi_ij, beta = 0, mean sound speed c0. */
static inline double av_pi(const orc_sim *s, int i, int j, const double *dx, double r2, double h) {
  const double dvx = s->v[2 * i] - s->v[2 * j];
  const double dvy = s->v[2 * i + 1] - s->v[2 * j + 1];
  const double vr = dvx * dx[0] + dvy * dx[1];
  if (!(vr < 0.0)) return 0.0;
  const double mu = h * vr / (r2 + 0.01 * h * h);
  const double rhobar = 0.5 * (s->rho[i] + s->rho[j]);
  return -s->p.alpha_av * s->p.c0 * mu / rhobar;
}

/* Per-thread scatter accumulation then thread-ordered reduction, the pattern of
 * derivatives.c:1114-1157. `width` is 1 (density) or 2 (vector). */
#define SCATTER_BEGIN(width)                                               \
  {                                                                        \
    const int W_ = (width);                                                \
    _Pragma("omp parallel num_threads(s->nthreads)")                       \
    {                                                                      \
      int tid_ = 0;                                                        \
      SCATTER_TID;                                                         \
      double *tb = s->tmp + (size_t)tid_ * 2 * s->n;                       \
      memset(tb, 0, sizeof(double) * W_ * s->n);                           \
      _Pragma("omp for schedule(static)")                                  \
      for (int k = 0; k < s->niac; k++) {
#define SCATTER_END(dst)                                                   \
      }                                                                    \
    }                                                                      \
    for (int th = 0; th < s->nthreads; th++) {                             \
      const double *tb = s->tmp + (size_t)th * 2 * s->n;                   \
      for (int q = 0; q < W_ * s->nf; q++) (dst)[q] += tb[q];              \
    }                                                                      \
  }
#ifdef _OPENMP
#define SCATTER_TID tid_ = omp_get_thread_num()
#else
#define SCATTER_TID (void)0
#endif

/* ------------------------------------------------------------ ref mode */

/* Bulk modulus of the pressure transport, rho0 c^2 + gamma P (derivatives.c:1290-1291, 1607-1608). */
static inline double ref_pp(const orc_sim *s, int i) {
  return s->p.rho0 * s->p.c0 * s->p.c0 + s->p.gamma * s->P[i];
}

/* PredictingDynamicPressureEquation (derivatives.c:1254-1331) -> RPt.  Note the
 * reference's volumes: volumei = mass[j]/density[i], volumej = mass[j]/density[j]. */
static void ref_predicting_pressure(orc_sim *s) {
  const orc_params *p = &s->p;
  SCATTER_BEGIN(1)
    const int i = s->ip[k], j = s->jp[k];
    const double *dw = &s->dwdx[2 * k];
    double dx[2], dv[2];
    for (int d = 0; d < 2; d++) {
      dv[d] = s->v[2 * i + d] - s->v[2 * j + d];
      dx[d] = s->x[2 * i + d] - s->x[2 * j + d];
    }
    double vdw = 0.0, rdwij = 0.0;
    for (int d = 0; d < 2; d++) {
      vdw += dv[d] * dw[d];
      rdwij += dx[d] * dw[d];
    }
    const double pi = ref_pp(s, i), pj = ref_pp(s, j);
    const double volumei = s->mass[j] / s->rho[i];
    const double volumej = s->mass[j] / s->rho[j];
    const double vj = s->mass[j] / s->rho[i];
    const double vi = s->mass[i] / s->rho[j];
    tb[i] += vj * pi * vdw;
    tb[j] += vi * pj * vdw;
    const double h = 0.5 * (s->hsml[i] + s->hsml[j]);
    const double eta2 = 0.01 * h * h;
    const double nu = rdwij / (s->rij2[k] + eta2);
    const double epj = 4.0 * pi / (pi + pj);
    const double epi = 4.0 * pj / (pi + pj);
    tb[i] += p->gamma * p->kappa * volumej * (s->P[i] - s->P[j]) * nu * epj;
    tb[j] += p->gamma * p->kappa * volumei * (s->P[j] - s->P[i]) * nu * epi;
    const double rho = (s->rho[i] + s->rho[j]) / (s->rho[i] * s->rho[j]);
    const double rhoi = (s->rho[i] - s->rho[j]) * rho;
    const double rhoj = (s->rho[j] - s->rho[i]) * rho;
    tb[i] -= p->kappa * volumej * pi * rhoi * nu;
    tb[j] -= p->kappa * pj * volumei * rhoj * nu;
  SCATTER_END(s->RPt)
}

/* LocalDensityFluctuation (derivatives.c:2100-2135) and LocalPressureFluctuation
 * (:2137-2171): f_i -= V_j (f_i - f_j) w, f_j += V_i (f_i - f_j) w. */
static void ref_scalar_fluctuation(orc_sim *s, const double *f, double *dst) {
  for (int i = 0; i < s->n; i++) dst[i] = 0.0;
  SCATTER_BEGIN(1)
    const int i = s->ip[k], j = s->jp[k];
    const double volumei = s->mass[i] / s->rho[i];
    const double volumej = s->mass[j] / s->rho[j];
    const double avgpart = (f[i] - f[j]) * s->w[k];
    tb[i] -= volumej * avgpart;
    tb[j] += volumei * avgpart;
  SCATTER_END(dst)
}

/* LocalVelocityFluctuation (derivatives.c:2060-2098).  Both of the reference's
 * updates go to particle i (the first of the pair); j receives nothing. */
static void ref_velocity_fluctuation(orc_sim *s) {
  for (int i = 0; i < 2 * s->n; i++) s->dv[i] = 0.0;
  SCATTER_BEGIN(2)
    const int i = s->ip[k], j = s->jp[k];
    for (int d = 0; d < 2; d++) {
      const double avgpart = (s->v[2 * i + d] - s->v[2 * j + d]) * s->w[k];
      tb[2 * i + d] -= s->mass[j] * avgpart / s->rholoc[i];
      tb[2 * i + d] += s->mass[i] * avgpart / s->rholoc[j];
    }
  SCATTER_END(s->dv)
}

/* DynamicPressureEquation (derivatives.c:1569-1672) -> RP. */
static void ref_pressure(orc_sim *s) {
  const orc_params *p = &s->p;
  SCATTER_BEGIN(1)
    const int i = s->ip[k], j = s->jp[k];
    const double *dw = &s->dwdx[2 * k];
    double dx[2], dvij[2], dvji[2];
    for (int d = 0; d < 2; d++) {
      dvij[d] = s->v[2 * i + d] - s->vloc[2 * j + d];
      dvji[d] = s->v[2 * j + d] - s->vloc[2 * i + d];
      dx[d] = s->x[2 * i + d] - s->x[2 * j + d];
    }
    double vdwij = 0.0, vdwji = 0.0, rdwij = 0.0;
    for (int d = 0; d < 2; d++) {
      vdwij += dvij[d] * dw[d];
      vdwji += dvji[d] * dw[d];
      rdwij += dx[d] * dw[d];
    }
    const double pi = ref_pp(s, i), pj = ref_pp(s, j);
    const double volumei = s->mass[i] / s->rho[i];
    const double volumej = s->mass[j] / s->rho[j];
    const double reducedVolumei = volumei * s->rholoc[i] / s->rho[j];
    const double reducedVolumej = volumej * s->rholoc[j] / s->rho[i];
    tb[i] += reducedVolumej * pi * vdwij;
    tb[j] -= reducedVolumei * pj * vdwji;
    const double h = 0.5 * (s->hsml[i] + s->hsml[j]);
    const double eta2 = 0.01 * h * h;
    const double nu = rdwij / (s->rij2[k] + eta2);
    const double epj = 4.0 * pi / (pi + pj);
    const double epi = 4.0 * pj / (pi + pj);
    tb[i] += p->gamma * p->kappa * volumej * (s->P[i] - s->Ploc[j]) * nu * epj;
    tb[j] += p->gamma * p->kappa * volumei * (s->P[j] - s->Ploc[i]) * nu * epi;
    const double rho = (s->rho[i] + s->rho[j]) / (s->rho[i] * s->rho[j]);
    const double rhoi = (s->rho[i] - s->rholoc[j]) * rho;
    const double rhoj = (s->rho[j] - s->rholoc[i]) * rho;
    tb[i] -= p->kappa * volumej * pi * rhoi * nu;
    tb[j] -= p->kappa * pj * volumei * rhoj * nu;
  SCATTER_END(s->RP)
}

/* ViscousForcesSingleStep (derivatives.c:1770-1839) -> a, and
 * PredictingViscousForcesSingleStep (:1448-1515) -> at.  dx = x_j - x_i,
 * dv = v_i - v_j; viscosity[] = viscosity0 (input.c:108, inputbp.c:288). */
static void ref_viscous(orc_sim *s, int predicting) {
  const orc_params *p = &s->p;
  SCATTER_BEGIN(2)
    const int i = s->ip[k], j = s->jp[k];
    const double *dw = &s->dwdx[2 * k];
    double dx[2], dvij[2], dvji[2];
    for (int d = 0; d < 2; d++) {
      dvij[d] = s->v[2 * i + d] - s->v[2 * j + d];
      dvji[d] = predicting ? dvij[d] : s->v[2 * j + d] - s->v[2 * i + d];
      dx[d] = s->x[2 * j + d] - s->x[2 * i + d];
    }
    const double h = 0.5 * (s->hsml[i] + s->hsml[j]);
    const double eta2 = 0.01 * h * h;
    const double invR2 = 1.0 / (s->rij2[k] + eta2);
    double rhoi, rhoj;
    if (predicting) {
      rhoi = 1.0 / (s->rho[i] * s->rho[i]);
      rhoj = 1.0 / (s->rho[j] * s->rho[j]);
    } else {
      rhoi = s->rholoc[j] / (s->rho[i] * s->rho[i] * s->rho[j]);
      rhoj = s->rholoc[i] / (s->rho[j] * s->rho[j] * s->rho[i]);
    }
    const double ei = s->rho[i] * p->mu, ej = s->rho[j] * p->mu;
    const double eff = 2.0 * ei * ej / (ei + ej);
    const double iS00 = eff * (dx[0] * dvij[0] - dx[1] * dvij[1]) * invR2;
    const double iS01 = eff * (dx[0] * dvij[1] + dx[1] * dvij[0]) * invR2;
    const double iS10 = eff * (dx[1] * dvij[0] + dx[0] * dvij[1]) * invR2;
    const double iS11 = eff * (dx[1] * dvij[1] - dx[0] * dvij[0]) * invR2;
    const double jS00 = eff * (dx[0] * dvji[0] - dx[1] * dvji[1]) * invR2;
    const double jS01 = eff * (dx[0] * dvji[1] + dx[1] * dvji[0]) * invR2;
    const double jS10 = eff * (dx[1] * dvji[0] + dx[0] * dvji[1]) * invR2;
    const double jS11 = eff * (dx[1] * dvji[1] - dx[0] * dvji[0]) * invR2;
    double ax = (iS00 * rhoi + jS00 * rhoj) * dw[0];
    ax += (iS01 * rhoi + jS01 * rhoj) * dw[1];
    double ay = (iS10 * rhoi + jS10 * rhoj) * dw[0];
    ay += (iS11 * rhoi + jS11 * rhoj) * dw[1];
    tb[2 * i] += s->mass[j] * ax;
    tb[2 * i + 1] += s->mass[j] * ay;
    tb[2 * j] -= s->mass[i] * ax;
    tb[2 * j + 1] -= s->mass[i] * ay;
  SCATTER_END(predicting ? s->at : s->a)
}

/* ClassicShearStress (derivatives.c:1686-1721), serial in pair order as there:
 * stress[4 i + {0,1,2,3}] = ViscousStress{xx,xy,yx,yy}[i], zeroed each call
 * (:3150-3153).  dvij = v_i - vloc_j, dvji = v_j - vloc_i; the reduced volume of j is
 * (m_j / rho_j) rholoc_j / rho_i and of i (m_i / rho_i) rholoc_i / rho_j;
 * viscosity[] = viscosity0 (input.c:108, inputbp.c:288). */
static void ref_shear_stress(orc_sim *s) {
  const double mu = s->p.mu;
  double *S = s->stress;
  memset(S, 0, sizeof(double) * 4 * s->n);
  for (int k = 0; k < s->niac; k++) {
    const int i = s->ip[k], j = s->jp[k];
    const double *dw = &s->dwdx[2 * k];
    double dvij[2], dvji[2];
    for (int d = 0; d < 2; d++) {
      dvij[d] = s->v[2 * i + d] - s->vloc[2 * j + d];
      dvji[d] = s->v[2 * j + d] - s->vloc[2 * i + d];
    }
    const double vi = s->mass[i] / s->rho[i], vj = s->mass[j] / s->rho[j];
    const double rvi = vi * s->rholoc[i] / s->rho[j];
    const double rvj = vj * s->rholoc[j] / s->rho[i];
    S[4 * i + 0] -= mu * rvj * (dvij[0] * dw[0] + dvij[0] * dw[0]);
    S[4 * j + 0] += mu * rvi * (dvji[0] * dw[0] + dvji[0] * dw[0]);
    S[4 * i + 1] -= mu * rvj * (dvij[0] * dw[1] + dvij[1] * dw[0]);
    S[4 * j + 1] += mu * rvi * (dvji[0] * dw[1] + dvji[1] * dw[0]);
    S[4 * i + 2] -= mu * rvj * (dvij[1] * dw[0] + dvij[0] * dw[1]);
    S[4 * j + 2] += mu * rvi * (dvji[1] * dw[0] + dvji[0] * dw[1]);
    S[4 * i + 3] -= mu * rvj * (dvij[1] * dw[1] + dvij[1] * dw[1]);
    S[4 * j + 3] += mu * rvi * (dvji[1] * dw[1] + dvji[1] * dw[1]);
  }
}

/* ViscousForcesTwoStep (derivatives.c:1724-1768) -> a: the divergence of the stresses
 * of both particles of a pair, weighted rholoc_j/(rho_i^2 rho_j) and
 * rholoc_i/(rho_j^2 rho_i); the second particle takes its share only when it is a
 * fluid particle. */
static void ref_viscous_two_step(orc_sim *s) {
  const double *S = s->stress;
  SCATTER_BEGIN(2)
    const int i = s->ip[k], j = s->jp[k];
    const double *dw = &s->dwdx[2 * k];
    const double rhoi = s->rholoc[j] / (s->rho[i] * s->rho[i] * s->rho[j]);
    const double rhoj = s->rholoc[i] / (s->rho[j] * s->rho[j] * s->rho[i]);
    double ax = (S[4 * i + 0] * rhoi + S[4 * j + 0] * rhoj) * dw[0];
    ax += (S[4 * i + 1] * rhoi + S[4 * j + 1] * rhoj) * dw[1];
    double ay = (S[4 * i + 2] * rhoi + S[4 * j + 2] * rhoj) * dw[0];
    ay += (S[4 * i + 3] * rhoi + S[4 * j + 3] * rhoj) * dw[1];
    tb[2 * i] += s->mass[j] * ax;
    tb[2 * i + 1] += s->mass[j] * ay;
    if (j < s->nf) {
      tb[2 * j] -= s->mass[i] * ax;
      tb[2 * j + 1] -= s->mass[i] * ay;
    }
  SCATTER_END(s->a)
}

/* Ref-mode sums that do not touch dvdt: the pressure rates and the fluctuation
 * fields (derivatives.c:4204-4237). */
static void ref_pressure_and_fluctuations(orc_sim *s) {
  memset(s->RP, 0, sizeof(double) * s->n);
  memset(s->RPt, 0, sizeof(double) * s->n);
  ref_predicting_pressure(s);
  ref_scalar_fluctuation(s, s->rho, s->drho);
  ref_scalar_fluctuation(s, s->P, s->dP);
  ref_velocity_fluctuation(s);
  ref_pressure(s);
}

/* ------------------------------------------------------- mirror particles */

/* Sections of derivatives.c:3193-4192: each is one loop over the fluid rows in
 * index order that appends the mirror image of every row meeting its test. */
enum {
  GS_LEFT0 = 1,    /* x about 0 (0.0 * FlatBottomLength), |x| < 2h                 :3474-3502 */
  GS_LEFT,         /* x about LeftBoundaryPosition, |x - LBP| < 2h                 :3198-3226 */
  GS_BOTTOM,       /* y about 0, y < 2h and x <= FlatBottomLength                  :3228-3257 */
  GS_CORNER_L,     /* both, at (LBP, 0): |x - LBP| < 2h and y < 2h                 :3259-3288 */
  GS_CORNER_R,     /* both, at (RightBoundaryPosition, 0)                          :3722-3751 */
  GS_RIGHT_FB,     /* x about FlatBottomLength, |x - FBL| < 2h                     :3596-3624 */
  GS_RIGHT,        /* x about RightBoundaryPosition                                :3753-3781 */
  GS_TOP_DX,       /* y -> LocalDepth + 2 |y - LocalDepth - dx|                    :3534-3563 */
  GS_TOP,          /* y -> LocalDepth + 2 |y - LocalDepth|                         :4095-4125 */
  GS_SLOPE         /* reflection in y = (x - FBL)/15, distance < 3h, x > FBL       :3290-3318 */
};

/* The section list of a virtualType (sec may be NULL); returns its length.
 * virtualType 2's sections are all commented out in the reference. */
int orc_ghost_sections(int vtype, int *sec) {
  static const int t1[] = {GS_LEFT, GS_BOTTOM, GS_CORNER_L, GS_SLOPE};                  /* :3195 */
  static const int t3[] = {GS_LEFT0, GS_BOTTOM, GS_TOP_DX, GS_CORNER_L, GS_RIGHT_FB};   /* :3471 */
  static const int t4[] = {GS_LEFT0, GS_BOTTOM, GS_CORNER_L, GS_RIGHT_FB};              /* :3783, :3907 */
  static const int t5[] = {GS_LEFT, GS_CORNER_L, GS_CORNER_R, GS_RIGHT};                /* :3626 */
  static const int t7[] = {GS_LEFT0, GS_BOTTOM, GS_TOP, GS_CORNER_L, GS_RIGHT_FB};      /* :4032 */
  const int *t = NULL;
  int m = 0;
  switch (vtype) {
    case 1: t = t1; m = 4; break;
    case 3: t = t3; m = 5; break;
    case 4: case 8: t = t4; m = 4; break;
    case 5: t = t5; m = 4; break;
    case 7: t = t7; m = 5; break;
    default: break;
  }
  if (sec) for (int k = 0; k < m; k++) sec[k] = t[k];
  return m;
}

/* Appends the mirror rows of the current fluid state after the n particle rows of
 * x, in the reference's order, with the fields derivatives.c copies into them;
 * sets s->ng (NumberOfVirtualParticles, :4192). */
static void ghosts_generate(orc_sim *s) {
  const orc_params *p = &s->p;
  int sec[8];
  const int m = orc_ghost_sections(p->vtype, sec);
  const double slope = 1.0 / 15.0;   /* System.c:309 */
  const double lbp = s->lbp, rbp = s->rbp, fbl = p->flat_bottom, H = p->depth;
  int ng = 0;
  for (int k = 0; k < m; k++) {
    for (int i = 0; i < s->nf; i++) {
      const double x = s->x[2 * i], y = s->x[2 * i + 1], h2 = 2.0 * s->hsml[i];
      double gx, gy, sx = 1.0, sy = 1.0;    /* ghost position, velocity signs */
      double vx = s->v[2 * i], vy = s->v[2 * i + 1], lx = s->vloc[2 * i], ly = s->vloc[2 * i + 1];
      int on = 0;
      switch (sec[k]) {
        case GS_LEFT0: {
          double d = fabs(x - 0.0 * fbl);
          on = d < h2; gx = 0.0 - d; gy = y; sx = -1.0;
          break;
        }
        case GS_LEFT: {
          double d = fabs(x - lbp);
          on = d < h2; gx = lbp - d; gy = y; sx = -1.0;
          break;
        }
        case GS_BOTTOM:
          on = (y < h2) && (x <= fbl); gx = x; gy = -y; sy = -1.0;
          break;
        case GS_CORNER_L: {
          double d = fabs(x - lbp), dy = y;
          on = (d < h2) && (dy < h2); gx = lbp - d; gy = -y; sx = sy = -1.0;
          break;
        }
        case GS_CORNER_R: {
          double d = fabs(x - rbp), dy = y;
          on = (d < h2) && (dy < h2); gx = rbp + d; gy = -y; sx = sy = -1.0;
          break;
        }
        case GS_RIGHT_FB: {
          double d = fabs(x - fbl);
          on = d < h2; gx = fbl + d; gy = y; sx = -1.0;
          break;
        }
        case GS_RIGHT: {
          double d = fabs(x - rbp);
          on = d < h2; gx = rbp + d; gy = y; sx = -1.0;
          break;
        }
        case GS_TOP_DX: {
          double d = fabs(y - H - p->dx);
          on = d < h2; gx = x; gy = H + 2 * d; sy = -1.0;
          break;
        }
        case GS_TOP: {
          double d = fabs(y - H);
          on = d < h2; gx = x; gy = H + 2 * d; sy = -1.0;
          break;
        }
        case GS_SLOPE: {
          double denoms = 1 + slope * slope;
          double nums = 1 - slope * slope;
          double yIntercept = slope * fbl;
          double d = (1.0 / sqrt(denoms)) * fabs(slope * x - y - yIntercept);
          on = (d < 3.0 * s->hsml[i]) && (x > fbl);
          gx = (nums * x + 2.0 * slope * y + 2.0 * slope * yIntercept) / denoms;
          gy = (2.0 * slope * x - nums * y - 2.0 * yIntercept) / denoms;
          if (on) {
            const double rvx = (nums * vx + 2.0 * slope * vy) / denoms;
            const double rvy = (2.0 * slope * vx - nums * vy) / denoms;
            const double rlx = (nums * lx + 2.0 * slope * ly) / denoms;
            const double rly = (2.0 * slope * lx - nums * ly) / denoms;
            vx = rvx; vy = rvy; lx = rlx; ly = rly;
          }
          break;
        }
        default:
          gx = gy = 0.0;
          break;
      }
      if (!on) continue;
      const int r = s->n + ng;
      s->x[2 * r] = gx;
      s->x[2 * r + 1] = gy;
      s->gv[2 * ng] = sx * vx; s->gv[2 * ng + 1] = sy * vy;
      s->glv[2 * ng] = sx * lx; s->glv[2 * ng + 1] = sy * ly;
      s->grho[ng] = s->rho[i];
      s->grl[ng] = s->rholoc[i];
      s->gP[ng] = s->P[i];
      s->gPl[ng] = s->Ploc[i];
      s->gmass[ng] = s->mass[i];
      ng++;
    }
  }
  s->ng = ng;
}

/* The mirror rows of the last derivatives() call (any pointer may be NULL);
 * returns their number. */
int orc_get_ghosts(const orc_sim *s, double *pos, double *vel, double *lvel, double *rho,
                   double *rholoc, double *P, double *Ploc, double *mass) {
  const int ng = s->ng;
  if (pos) memcpy(pos, s->x + 2 * (size_t)s->n, sizeof(double) * 2 * ng);
  if (vel) memcpy(vel, s->gv, sizeof(double) * 2 * ng);
  if (lvel) memcpy(lvel, s->glv, sizeof(double) * 2 * ng);
  if (rho) memcpy(rho, s->grho, sizeof(double) * ng);
  if (rholoc) memcpy(rholoc, s->grl, sizeof(double) * ng);
  if (P) memcpy(P, s->gP, sizeof(double) * ng);
  if (Ploc) memcpy(Ploc, s->gPl, sizeof(double) * ng);
  if (mass) memcpy(mass, s->gmass, sizeof(double) * ng);
  return ng;
}

int orc_ghost_capacity(const orc_sim *s) { return s->gcap; }

/*
 * One call of derivatives(sys, t) in config mode (derivatives.c:3137-4315 with
 * the scheme's substitutions): rates R, Rt, a, at are recomputed from scratch.
 */
int orc_derivatives(orc_sim *s, double t) {
  const orc_params *p = &s->p;
  const int n = s->n;
  memset(s->R, 0, sizeof(double) * n);
  memset(s->Rt, 0, sizeof(double) * n);
  memset(s->a, 0, sizeof(double) * 2 * n);
  memset(s->at, 0, sizeof(double) * 2 * n);
  boundary_reset(s, t);
  /* EquationOfState (derivatives.c:1674-1682); config mode feeds it to the momentum
   * terms, ref mode integrates DynamicPressure instead (its Tait Pressure is output
   * only).  Evaluated before the mirror rows copy the pressure of their sources. */
  if (p->mode == 0)
    for (int i = 0; i < n; i++) {
      const double density = s->rho[i] / p->rho0;
      s->P[i] = p->B * (pow(density, p->gamma) - 1.0);
    }
  ghosts_generate(s);

  /* CalculateKernel (derivatives.c:1063-1106) with the cubic spline: the search ranks
   * the mirror rows too (:1069-1074), the pair filter drops them */
  const int pb = s->nf + s->nfixed + s->nmoving - 1;
  int niac = orc_search(s->x, s->ip, s->jp, s->rij2, s->nf, n + s->ng, pb, p->K,
                        s->hsml[0] * s->hsml[0] * 4, s->cap, s->nthreads);
  if (niac < 0) return -1;
  s->niac = niac;
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int k = 0; k < niac; k++) {
    const int i = s->ip[k], j = s->jp[k];
    double dx[2], w;
    const double h = 0.5 * (s->hsml[i] + s->hsml[j]);
    for (int d = 0; d < 2; d++) dx[d] = s->x[2 * i + d] - s->x[2 * j + d];
    const double r = sqrt(s->rij2[k]);
    if (p->kernel == 1)
      orc_poly7(dx, r, h, &w, &s->dwdx[2 * k]);
    else
      orc_cubic(dx, r, h, &w, &s->dwdx[2 * k]);
    s->w[k] = w;
  }

  if (p->mode == 1) ref_pressure_and_fluctuations(s);
  /* ClassicShearStress runs in every ref-mode call (derivatives.c:4225); only stressType
   * 2 reads it */
  if (p->mode == 1) ref_shear_stress(s);

  /* PredictingDensityEquation (derivatives.c:1217-1252) */
  SCATTER_BEGIN(1)
    const int i = s->ip[k], j = s->jp[k];
    const double *dw = &s->dwdx[2 * k];
    double dv[2];
    for (int d = 0; d < 2; d++) dv[d] = s->v[2 * i + d] - s->vloc[2 * j + d];
    for (int d = 0; d < 2; d++) {
      double vdw = dv[d] * dw[d];
      tb[i] += s->mass[j] * vdw;
      tb[j] += s->mass[j] * vdw;
    }
  SCATTER_END(s->Rt)

  /* PredictingPressureGradientForce (derivatives.c:1333-1377) + artificial viscosity */
  SCATTER_BEGIN(2)
    const int i = s->ip[k], j = s->jp[k];
    const double *dw = &s->dwdx[2 * k];
    const double rhoi = 1.0 / (s->rho[i] * s->rho[i]);
    const double rhoj = 1.0 / (s->rho[j] * s->rho[j]);
    double pij = s->P[i] * rhoi + s->P[j] * rhoj;
    double dx[2] = {s->x[2 * i] - s->x[2 * j], s->x[2 * i + 1] - s->x[2 * j + 1]};
    pij += av_pi(s, i, j, dx, s->rij2[k], 0.5 * (s->hsml[i] + s->hsml[j]));
    for (int d = 0; d < 2; d++) {
      tb[2 * i + d] -= s->mass[j] * pij * dw[d];
      tb[2 * j + d] += s->mass[i] * pij * dw[d];
    }
  SCATTER_END(s->at)

  /* DensityEquation (derivatives.c:1520-1567) */
  SCATTER_BEGIN(1)
    const int i = s->ip[k], j = s->jp[k];
    const double *dw = &s->dwdx[2 * k];
    double dvij[2], dvji[2];
    for (int d = 0; d < 2; d++) {
      dvij[d] = s->v[2 * i + d] - s->vloc[2 * j + d];
      dvji[d] = s->v[2 * j + d] - s->vloc[2 * i + d];
    }
    for (int d = 0; d < 2; d++) {
      double vdwij = dvij[d] * dw[d];
      double vdwji = dvji[d] * dw[d];
      double volumei = s->mass[i] / s->rho[i];
      double volumej = s->mass[j] / s->rho[j];
      double mj = volumej * s->rholoc[j];
      double mi = volumei * s->rholoc[i];
      tb[i] += mj * vdwij;
      tb[j] -= mi * vdwji;
    }
  SCATTER_END(s->R)

  /* PressureGradientForce (derivatives.c:1109-1163) + artificial viscosity */
  SCATTER_BEGIN(2)
    const int i = s->ip[k], j = s->jp[k];
    const double *dw = &s->dwdx[2 * k];
    const double rhoi = s->rholoc[j] / (s->rho[i] * s->rho[i] * s->rho[j]);
    const double rhoj = s->rholoc[i] / (s->rho[j] * s->rho[j] * s->rho[i]);
    double pij = s->P[i] * rhoi + s->P[j] * rhoj;
    double dx[2] = {s->x[2 * i] - s->x[2 * j], s->x[2 * i + 1] - s->x[2 * j + 1]};
    pij += av_pi(s, i, j, dx, s->rij2[k], 0.5 * (s->hsml[i] + s->hsml[j]));
    for (int d = 0; d < 2; d++) {
      tb[2 * i + d] -= s->mass[j] * pij * dw[d];
      tb[2 * j + d] += s->mass[i] * pij * dw[d];
    }
  SCATTER_END(s->a)

  /* ref mode: stressType 2 takes the two-step viscous forces, any other the single-step
   * ones; the predicting form is single-step either way (derivatives.c:4231-4244) */
  if (p->mode == 1) {
    if (p->stress_type == 2)
      ref_viscous_two_step(s);
    else
      ref_viscous(s, 0);
    ref_viscous(s, 1);
  }

  /* BodyForce / PredictingBodyForce (derivatives.c:1167-1178, 1379-1390) */
  if (p->gravity == 1)
    for (int i = 0; i < s->nf; i++) {
      s->a[2 * i] -= 0.0; s->a[2 * i + 1] -= p->g;
      s->at[2 * i] -= 0.0; s->at[2 * i + 1] -= p->g;
    }

  /* SpongeLayerWaveReflectionAbsorption (derivatives.c:2827-2838) */
  if (p->mode == 1)
    for (int i = 0; i < s->nf; i++) s->a[2 * i] -= s->sponge[i] * s->v[2 * i];

  /* SpringBoundaryForces + PredictingSpringBoundaryForces (derivatives.c:2396-2449,
   * 1392-1446): identical formulas accumulated into a and at. */
  for (int pass = 0; pass < 2; pass++) {
    double *dst = pass == 0 ? s->a : s->at;
    const double r0 = 3.0 * p->dx;
    SCATTER_BEGIN(2)
      const int i = s->ip[k], j = s->jp[k];
      if ((s->itype[i] == 1) && ((s->itype[j] == 2) || (s->itype[j] == 3) || (s->itype[j] == 5))) {
        double dx[2], dv[2];
        double rr = 0.0, rvij = 0.0;
        for (int d = 0; d < 2; d++) {
          dx[d] = s->x[2 * i + d] - s->x[2 * j + d];
          dv[d] = s->v[2 * i + d] - s->v[2 * j + d];
          rr += dx[d] * dx[d];
          rvij += dx[d] * dv[d];
        }
        rr = sqrt(rr);
        if (rr < r0) {
          for (int d = 0; d < 2; d++) {
            tb[2 * i + d] -= p->k_spring * (r0 - rr) * dx[d] / (rr * s->mass[i]);
            tb[2 * i + d] -= p->c_damp * rvij * dx[d] / (rr * rr * s->mass[i]);
          }
        }
      }
    SCATTER_END(dst)
  }
  return 0;
}

/* Leapfrog time step (sph.cc:90-168).  Config mode: fluctuation fields zero and
 * DynamicPressure from the EOS.  Ref mode: DynamicPressure integrated with dpdt,
 * local fields carry the fluctuations (epsilon weighs the velocity's). */
int orc_step(orc_sim *s) {
  const double dt = s->p.dt;
  const int nf = s->nf, n = s->n;
  const int ref = s->p.mode == 1;
  const double eps = ref ? s->p.eps : 0.0;
  int itime = ++s->itime;
  double t = itime * dt;
  if (itime != 1) {
    memcpy(s->rhomin, s->rho, sizeof(double) * nf);
    memcpy(s->rholoc, s->rho, sizeof(double) * nf);
    memcpy(s->vloc, s->v, sizeof(double) * 2 * n);
    memcpy(s->vmin, s->v, sizeof(double) * 2 * n);
    if (ref) {
      memcpy(s->Pmin, s->P, sizeof(double) * nf);
      memcpy(s->Ploc, s->P, sizeof(double) * nf);
    }
    for (int i = 0; i < nf; i++) {
      s->rho[i] += 0.5 * dt * s->Rt[i];
      if (ref) {
        s->rholoc[i] = s->rholoc[i] + 0.5 * dt * s->Rt[i] + s->drho[i];
        s->P[i] += 0.5 * dt * s->RPt[i];
        s->Ploc[i] = s->Ploc[i] + 0.5 * dt * s->RPt[i] + s->dP[i];
      } else {
        s->rholoc[i] = s->rholoc[i] + 0.5 * dt * s->Rt[i];
      }
    }
    for (int i = 0; i < nf; i++)
      for (int d = 0; d < 2; d++) {
        s->vmin[2 * i + d] = s->v[2 * i + d];
        s->v[2 * i + d] += 0.5 * dt * s->at[2 * i + d];
        s->vloc[2 * i + d] = s->v[2 * i + d] + 0.5 * dt * s->at[2 * i + d];
        if (ref) s->vloc[2 * i + d] += eps * s->dv[2 * i + d];
      }
  }
  if (orc_derivatives(s, t) != 0) return -1;
  if (itime == 1) {
    for (int i = 0; i < nf; i++) {
      s->rholoc[i] = s->rho[i] + 0.5 * dt * s->R[i];
      if (ref) s->rholoc[i] += s->drho[i];
      s->rho[i] += 0.5 * dt * s->R[i];
      if (ref) {
        s->Ploc[i] = s->P[i] + 0.5 * dt * s->RP[i] + s->dP[i];
        s->P[i] += 0.5 * dt * s->RP[i];
      }
    }
    for (int i = 0; i < nf; i++)
      for (int d = 0; d < 2; d++) {
        s->vloc[2 * i + d] = s->v[2 * i + d] + 0.5 * dt * s->a[2 * i + d];
        if (ref) s->vloc[2 * i + d] += eps * s->dv[2 * i + d];
        s->v[2 * i + d] += 0.5 * dt * s->a[2 * i + d];
        s->x[2 * i + d] += dt * s->v[2 * i + d];
      }
  } else {
    for (int i = 0; i < nf; i++) {
      s->rholoc[i] = s->rhomin[i] + dt * s->R[i];
      if (ref) s->rholoc[i] += s->drho[i];
      s->rho[i] = s->rhomin[i] + dt * s->R[i];
      if (ref) {
        s->Ploc[i] = s->Pmin[i] + dt * s->RP[i] + s->dP[i];
        s->P[i] = s->Pmin[i] + dt * s->RP[i];
      }
    }
    for (int i = 0; i < nf; i++)
      for (int d = 0; d < 2; d++) {
        s->vloc[2 * i + d] = s->vmin[2 * i + d] + dt * s->a[2 * i + d];
        if (ref) s->vloc[2 * i + d] += eps * s->dv[2 * i + d];
        s->v[2 * i + d] = s->vmin[2 * i + d] + dt * s->a[2 * i + d];
        s->x[2 * i + d] += dt * s->v[2 * i + d];
      }
  }
  return 0;
}

int orc_run(orc_sim *s, int nsteps) {
  for (int k = 0; k < nsteps; k++)
    if (orc_step(s) != 0) return -1;
  return 0;
}

/* Accessors (all arrays in reference order, length n or 2n). */
void orc_get_state(const orc_sim *s, double *x, double *v, double *rho, double *vloc,
                   double *rholoc) {
  if (x) memcpy(x, s->x, sizeof(double) * 2 * s->n);
  if (v) memcpy(v, s->v, sizeof(double) * 2 * s->n);
  if (rho) memcpy(rho, s->rho, sizeof(double) * s->n);
  if (vloc) memcpy(vloc, s->vloc, sizeof(double) * 2 * s->n);
  if (rholoc) memcpy(rholoc, s->rholoc, sizeof(double) * s->n);
}

void orc_get_rates(const orc_sim *s, double *R, double *Rt, double *a, double *at, double *P) {
  if (R) memcpy(R, s->R, sizeof(double) * s->n);
  if (Rt) memcpy(Rt, s->Rt, sizeof(double) * s->n);
  if (a) memcpy(a, s->a, sizeof(double) * 2 * s->n);
  if (at) memcpy(at, s->at, sizeof(double) * 2 * s->n);
  if (P) memcpy(P, s->P, sizeof(double) * s->n);
}

/* Ref-mode fields: DynamicPressure, localPressure, dpdt, dpdt_approx, density and
 * pressure fluctuations (n each), velocity fluctuation (2n). */
/* ViscousStress{xx,xy,yx,yy} of the last derivatives() call, 4 per particle. */
void orc_get_stress(const orc_sim *s, double *stress) {
  memcpy(stress, s->stress, sizeof(double) * 4 * s->n);
}

void orc_get_ref(const orc_sim *s, double *P, double *Ploc, double *RP, double *RPt, double *drho,
                 double *dP, double *dv) {
  if (P) memcpy(P, s->P, sizeof(double) * s->n);
  if (Ploc) memcpy(Ploc, s->Ploc, sizeof(double) * s->n);
  if (RP) memcpy(RP, s->RP, sizeof(double) * s->n);
  if (RPt) memcpy(RPt, s->RPt, sizeof(double) * s->n);
  if (drho) memcpy(drho, s->drho, sizeof(double) * s->n);
  if (dP) memcpy(dP, s->dP, sizeof(double) * s->n);
  if (dv) memcpy(dv, s->dv, sizeof(double) * 2 * s->n);
}

/* Sets localDensity, localPressure and localVelocity (n, n, 2n; NULL keeps a field),
 * for tests that evaluate derivatives() at a chosen state. */
void orc_set_local(orc_sim *s, const double *rholoc, const double *Ploc, const double *vloc) {
  if (rholoc) memcpy(s->rholoc, rholoc, sizeof(double) * s->n);
  if (Ploc) memcpy(s->Ploc, Ploc, sizeof(double) * s->n);
  if (vloc) memcpy(s->vloc, vloc, sizeof(double) * 2 * s->n);
}

int orc_niac(const orc_sim *s) { return s->niac; }
int orc_itime(const orc_sim *s) { return s->itime; }
