/*
 * c3d_oracle.c -- FP64 CPU restatement of the compact3d per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
 * product library.
 *
 * Each function follows one subroutine of compact3d.f (file:line at its
 * definition) operation by operation: same loop structure per element, same
 * accumulation order into RUH/HH, same Thomas / bordered-LU sequences.  The
 * loops over independent lines run under OpenMP; reductions are summed in a
 * fixed order so that results do not depend on the thread count.
 *
 * Deliberate, value-preserving differences (each is exact):
 *  - CTDMA11/CTDMA33 factorisations depend only on (b, n); they are computed
 *    once with the reference's own sequence (compact3d.f:8959-8977) and reused.
 *  - The 36 SGS line-set solves of RHS1-3 only run when FILES != 0.  With
 *    FILES = 0 the reference feeds them SGSVIS == 0 (GETUP zeroes it,
 *    compact3d.f:908-912) and they add exact zeros.
 *  - Where the reference reads one element past a local work array and
 *    multiplies it by an exact zero flag (e.g. SGSVIS1(I,JP) at J=N2M,
 *    compact3d.f:6581), the element is taken as 0.
 *  - FFTW's 2-D DFT (GETDP, compact3d.f:9315-9333) is replaced by the
 *    textbook DFT definition evaluated with a mixed-radix FFT (c3o_dft2d).
 */
#include "c3d_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  int n;
  double b;
  double *beta, *gama, *p, *q; /* 1-based, length n+1 */
} cycfac;

struct c3o {
  int n1, n2, n3, n1m, n2m, n3m, n3mh;
  double re, alx, aly, alz, vol, dt, cks, files, wmodel;
  int ifwall;
  double dx1, dx1q, dx3, dx3q;
  double *y, *dy, *h, *hm, *hc, *hp, *dym, *dyc, *dyp, *yc; /* 0..n2 */
  int *ipa, *ima, *kpa, *kma;
  double *fjpa, *fjmu, *fjmv;
  double *ak1, *ak3, *pmj, *pcj, *ppj;
  double *ubc;        /* UBC(0:M1,0:M3,3,2) */
  double *tauw, *taub; /* (0:M1,0:M3,2) */
  double wsm[3];
  double taow;
  /* WMODEL = 2 (IWMLES): the planes of COMMON /IWMLES/ (iwm_* below), its
   * time step wm_dt, CFLMAX of /TSTEP/, and the number of points of the last
   * call that took the excessive-linear-correction branch without an
   * equilibrium stress of their own (see c3o_wallmodel_iwmles) */
  double *iwm;
  double wm_dt, cflmax;
  int iwm_fixed;
  cycfac fx[3], fz[3]; /* b = 6, 10, 22 along x (n1m) and z (n3m) */
};

/* ---------------------------------------------------------------- layout */
#define NI (o->n1 + 1)
#define NJ (o->n2 + 1)
#define NK (o->n3 + 1)
#define NU ((size_t)NI * NJ * NK)
#define IX(i, j, k) ((((size_t)(k)) * NJ + (size_t)(j)) * NI + (size_t)(i))
#define PX(i, j, k) \
  ((((size_t)(k) - 1) * o->n2 + ((size_t)(j) - 1)) * o->n1 + ((size_t)(i) - 1))
#define UBCX(i, k, n, s) \
  ((size_t)(i) + (size_t)NI * ((size_t)(k) + (size_t)NK * ((size_t)(n) - 1 + 3 * ((size_t)(s) - 1))))
#define TAUX(i, k, s) ((size_t)(i) + (size_t)NI * ((size_t)(k) + (size_t)NK * ((size_t)(s) - 1)))

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* ------------------------------------------------- periodic bordered LU */
/* CTDMA11 / CTDMA33 (compact3d.f:8952-9046) with A=C=1, B=b: factorisation */
static void cyc_factor(cycfac *f, int n, double b) {
  const double A = 1.0, B = b, C = 1.0;
  f->n = n;
  f->b = b;
  f->beta = dalloc(n + 1);
  f->gama = dalloc(n + 1);
  f->p = dalloc(n + 1);
  f->q = dalloc(n + 1);
  double *BETA = f->beta, *GAMA = f->gama, *P = f->p, *Q = f->q;
  BETA[1] = B;
  GAMA[1] = C / BETA[1];
  for (int i = 2; i <= n - 1; i++) {
    BETA[i] = B - A * GAMA[i - 1];
    GAMA[i] = C / BETA[i];
  }
  P[1] = C;
  Q[1] = A / BETA[1];
  for (int i = 2; i <= n - 2; i++) {
    P[i] = -P[i - 1] * GAMA[i - 1];
    Q[i] = -(A * Q[i - 1] / BETA[i]);
  }
  P[n - 1] = A - P[n - 2] * GAMA[n - 2];
  Q[n - 1] = (C - A * Q[n - 2]) / BETA[n - 1];
  P[n] = B;
  for (int i = 1; i <= n - 1; i++) P[n] = P[n] - P[i] * Q[i];
}

static void cyc_free(cycfac *f) {
  free(f->beta);
  free(f->gama);
  free(f->p);
  free(f->q);
}

/* CTDMA11 forward/backward substitution (compact3d.f:8979-8995); d,x 1-based */
static void cyc_solve(const cycfac *f, const double *d, double *x) {
  const int n = f->n;
  const double A = 1.0;
  const double *BETA = f->beta, *GAMA = f->gama, *P = f->p, *Q = f->q;
  double z[n + 1];
  z[1] = d[1] / BETA[1];
  for (int i = 2; i <= n - 1; i++) z[i] = (d[i] - A * z[i - 1]) / BETA[i];
  z[n] = d[n];
  for (int i = 1; i <= n - 1; i++) z[n] = z[n] - P[i] * z[i];
  z[n] = z[n] / P[n];
  x[n] = z[n];
  x[n - 1] = z[n - 1] - Q[n - 1] * x[n];
  for (int i = n - 2; i >= 1; i--) x[i] = z[i] - GAMA[i] * x[i + 1] - Q[i] * x[n];
}

void c3o_ctdma(int n, double b, const double *d0, double *x0) {
  cycfac f;
  cyc_factor(&f, n, b);
  double d[n + 1], x[n + 1];
  for (int i = 1; i <= n; i++) d[i] = d0[i - 1];
  cyc_solve(&f, d, x);
  for (int i = 1; i <= n; i++) x0[i - 1] = x[i];
  cyc_free(&f);
}

/* CTDMA1J / CTDMA3I single-line sequence (compact3d.f:8721-8750, 8766-8807);
 * a,b,c,d,x 1-based; x may alias d */
static void cyc_solve_general(int n, const double *a, const double *b,
                              const double *c, const double *d, double *x) {
  double gam[n + 1], p[n + 1], q[n + 1];
  double bet = 1.0 / b[1];
  gam[1] = c[1] * bet;
  x[1] = d[1] * bet;
  p[1] = c[n];
  q[1] = a[1] * bet;
  for (int k = 2; k <= n - 1; k++) {
    bet = 1.0 / (b[k] - a[k] * gam[k - 1]);
    gam[k] = c[k] * bet;
    p[k] = -p[k - 1] * gam[k - 1];
    q[k] = -(a[k] * bet * q[k - 1]);
    x[k] = (d[k] - a[k] * x[k - 1]) * bet;
  }
  bet = 1.0 / (b[n - 1] - a[n - 1] * gam[n - 2]);
  p[n - 1] = a[n] - p[n - 2] * gam[n - 2];
  q[n - 1] = (c[n - 1] - a[n - 1] * q[n - 2]) * bet;
  x[n] = d[n];
  p[n] = b[n];
  for (int k = 1; k <= n - 1; k++) {
    x[n] = x[n] - p[k] * x[k];
    p[n] = p[n] - p[k] * q[k];
  }
  x[n] = x[n] / p[n];
  gam[n - 1] = 0.0;
  for (int k = n - 1; k >= 1; k--) x[k] = x[k] - gam[k] * x[k + 1] - q[k] * x[n];
}

void c3o_ctdma_ref_general(int n, const double *a0, const double *b0,
                           const double *c0, const double *d0, double *x0) {
  double a[n + 1], b[n + 1], c[n + 1], x[n + 1];
  for (int i = 1; i <= n; i++) {
    a[i] = a0[i - 1];
    b[i] = b0[i - 1];
    c[i] = c0[i - 1];
    x[i] = d0[i - 1];
  }
  cyc_solve_general(n, a, b, c, x, x);
  for (int i = 1; i <= n; i++) x0[i - 1] = x[i];
}

/* index of the cached factor for b in {6,10,22} */
static const cycfac *fxb(const c3o *o, double b) {
  return b == 6.0 ? &o->fx[0] : (b == 10.0 ? &o->fx[1] : &o->fx[2]);
}
static const cycfac *fzb(const c3o *o, double b) {
  return b == 6.0 ? &o->fz[0] : (b == 10.0 ? &o->fz[1] : &o->fz[2]);
}

/* ------------------------------------------------------------------ setup */
void c3o_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* MESH (compact3d.f:302-443), INDICES (:446-494), INIWAVE (:9182-9228),
 * METRICPOISSON (:9239-9263).  cks replaces the hard-coded CKS=2.5 (:338). */
c3o *c3o_new(int n1, int n2, int n3, double re, double alx, double aly,
             double alz, double dt, double files, double wmodel, double cks) {
  c3o *o = (c3o *)calloc(1, sizeof(c3o));
  o->n1 = n1;
  o->n2 = n2;
  o->n3 = n3;
  o->n1m = n1 - 1;
  o->n2m = n2 - 1;
  o->n3m = n3 - 1;
  o->n3mh = o->n3m / 2 + 1;
  o->re = re;
  o->alx = alx;
  o->aly = aly;
  o->alz = alz;
  o->dt = dt;
  o->files = files;
  o->wmodel = wmodel;
  o->cks = cks;
  o->ifwall = 0;
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m, N2 = n2;
  o->y = dalloc(n2 + 2);
  o->dy = dalloc(n2 + 2);
  o->h = dalloc(n2 + 2);
  o->hm = dalloc(n2 + 2);
  o->hc = dalloc(n2 + 2);
  o->hp = dalloc(n2 + 2);
  o->dym = dalloc(n2 + 2);
  o->dyc = dalloc(n2 + 2);
  o->dyp = dalloc(n2 + 2);
  o->yc = dalloc(n2 + 2);
  o->fjpa = dalloc(n2 + 2);
  o->fjmu = dalloc(n2 + 2);
  o->fjmv = dalloc(n2 + 2);
  o->pmj = dalloc(n2 + 2);
  o->pcj = dalloc(n2 + 2);
  o->ppj = dalloc(n2 + 2);
  o->ipa = (int *)calloc(n1 + 2, sizeof(int));
  o->ima = (int *)calloc(n1 + 2, sizeof(int));
  o->kpa = (int *)calloc(n3 + 2, sizeof(int));
  o->kma = (int *)calloc(n3 + 2, sizeof(int));
  o->ak1 = dalloc(n1 + 2);
  o->ak3 = dalloc(n3 + 2);
  o->ubc = dalloc((size_t)(n1 + 1) * (n3 + 1) * 6);
  o->tauw = dalloc((size_t)(n1 + 1) * (n3 + 1) * 2);
  o->taub = dalloc((size_t)(n1 + 1) * (n3 + 1) * 2);
  o->iwm = wmodel == 2.0 ? dalloc((size_t)IW_NPL * (n1 - 1) * (n3 - 1)) : NULL;

  /* MESH: y grid (:338-361) */
  double *Y = o->y, *DY = o->dy, *H = o->h;
  const double CKS = cks, Y0 = 0.0;
  Y[0] = (wmodel == 0.0) ? 0.0 : Y0;
  for (int j = 1; j <= N2; j++) {
    if (wmodel == 0.0) {
      if (aly == 2.0)
        Y[j] = 1 - tanh(CKS * (1 - aly * (j - 1) / (N2 - 1))) / tanh(CKS);
      else
        Y[j] = 1 - tanh(CKS * (1 - 2. * aly * (j - 1) / (2 * N2 - 1))) / tanh(CKS);
    } else {
      Y[j] = (aly - Y0) * (j - 1) / (N2 - 1) + Y0;
    }
  }
  o->aly = Y[N2];
  for (int j = 1; j <= N2M; j++) o->yc[j] = 0.5 * (Y[j + 1] + Y[j]);
  o->yc[N2] = Y[N2];
  o->vol = o->alx * o->aly * o->alz;
  o->dx1 = (double)N1M / alx;
  o->dx3 = (double)N3M / alz;
  o->dx1q = o->dx1 * o->dx1;
  o->dx3q = o->dx3 * o->dx3;
  DY[1] = Y[2];
  for (int j = 2; j <= N2M; j++) {
    DY[j] = Y[j + 1] - Y[j];
    H[j] = 0.5 * (DY[j] + DY[j - 1]);
  }
  H[1] = 0.5 * DY[1];
  H[N2] = 0.5 * DY[N2M];
  double *HP = o->hp, *HC = o->hc, *HM = o->hm;
  for (int j = 2; j <= N2M - 1; j++) {
    HP[j] = 1.0 / H[j + 1] / DY[j];
    HC[j] = (H[j + 1] + H[j]) / H[j + 1] / H[j] / DY[j];
    HM[j] = 1.0 / H[j] / DY[j];
  }
  HP[1] = 4.0 * H[1] / H[2] / (H[1] + H[2]) / DY[1];
  HC[1] = 4.0 / H[2] / DY[1];
  HM[1] = 4.0 / (H[1] + H[2]) / DY[1];
  HP[N2M] = 4.0 / (H[N2M] + H[N2]) / DY[N2M];
  HC[N2M] = 4.0 / H[N2M] / DY[N2M];
  HM[N2M] = 4.0 * H[N2] / H[N2M] / (H[N2M] + H[N2]) / DY[N2M];
  if (wmodel != 0.0) {
    HP[1] = 1.0 / H[2] / DY[1];
    HC[1] = 1.0 / H[2] / DY[1];
    HM[1] = 0.0;
    HP[N2M] = 0.0;
    HC[N2M] = 1.0 / H[N2M] / DY[N2M];
    HM[N2M] = 1.0 / H[N2M] / DY[N2M];
  }
  for (int j = 2; j <= N2M; j++) {
    o->dyp[j] = 1.0 / H[j] / DY[j];
    o->dyc[j] = 1.0 / H[j] * (1.0 / DY[j] + 1.0 / DY[j - 1]);
    o->dym[j] = 1.0 / H[j] / DY[j - 1];
  }

  /* INDICES (:454-481) */
  for (int i = 1; i <= N1M; i++) {
    o->ipa[i] = i + 1;
    o->ima[i] = i - 1;
  }
  o->ipa[N1M] = 1;
  o->ima[1] = N1M;
  for (int k = 1; k <= N3M; k++) {
    o->kpa[k] = k + 1;
    o->kma[k] = k - 1;
  }
  o->kpa[N3M] = 1;
  o->kma[1] = N3M;
  {
    int jpa[n2 + 2], jmu[n2 + 2], jmv[n2 + 2];
    for (int j = 1; j <= N2M; j++) {
      jpa[j] = j + 1;
      jmu[j] = j - 1;
      jmv[j] = j - 1;
    }
    jpa[N2M] = N2M;
    jmu[1] = 1;
    jmv[2] = 2;
    for (int j = 1; j <= N2M; j++) {
      o->fjpa[j] = jpa[j] - j;
      o->fjmu[j] = j - jmu[j];
      o->fjmv[j] = j - jmv[j];
    }
  }

  /* INIWAVE (:9197-9222) */
  {
    const double PI = acos(-1.0);
    for (int i = 1; i <= N1M; i++) {
      int ii = i;
      if (i > N1M / 2 + 1) ii = i - N1M;
      double ani = (ii - 1) * 2.0 * PI;
      double c = cos(ani / N1M);
      double den = 11. + c;
      o->ak1[i] = 288. * o->dx1q * (1. - c) / (den * den);
    }
    for (int k = 1; k <= o->n3mh; k++) {
      double ank = (k - 1) * 2. * PI;
      double c = cos(ank / N3M);
      double den = 11. + c;
      o->ak3[k] = 288. * o->dx3q * (1. - c) / (den * den);
    }
  }
  /* METRICPOISSON (:9249-9260) */
  for (int jc = 1; jc <= N2M; jc++) {
    int jp = jc + 1;
    double fjum = o->fjmu[jc], fjup = o->fjpa[jc];
    o->pmj[jc] = fjum * fjup * (1.0 / DY[jc] / H[jc]) +
                 (1. - fjup) * (1.0 / DY[N2M] / H[N2M]);
    o->pcj[jc] = fjum * fjup * (-1.0 / DY[jc] / H[jc] - 1.0 / DY[jc] / H[jp]) +
                 (1. - fjup) * (-1.0 / DY[N2M] / H[N2M]) +
                 (1. - fjum) * (-1.0 / DY[1] / H[2]);
    o->ppj[jc] = fjum * fjup * (1.0 / DY[jc] / H[jp]) +
                 (1. - fjum) * (1.0 / DY[1] / H[2]);
  }
  const double bs[3] = {6.0, 10.0, 22.0};
  for (int t = 0; t < 3; t++) {
    cyc_factor(&o->fx[t], N1M, bs[t]);
    cyc_factor(&o->fz[t], N3M, bs[t]);
  }
  return o;
}

void c3o_free(c3o *o) {
  if (!o) return;
  double *ps[] = {o->y,   o->dy,  o->h,   o->hm,  o->hc,   o->hp,  o->dym,
                  o->dyc, o->dyp, o->yc,  o->fjpa, o->fjmu, o->fjmv, o->pmj,
                  o->pcj, o->ppj, o->ak1, o->ak3, o->ubc,  o->tauw, o->taub, o->iwm};
  for (size_t t = 0; t < sizeof(ps) / sizeof(ps[0]); t++) free(ps[t]);
  free(o->ipa);
  free(o->ima);
  free(o->kpa);
  free(o->kma);
  for (int t = 0; t < 3; t++) {
    cyc_free(&o->fx[t]);
    cyc_free(&o->fz[t]);
  }
  free(o);
}

void c3o_set_dt(c3o *o, double dt) { o->dt = dt; }
void c3o_set_ifwall(c3o *o, int ifwall) { o->ifwall = ifwall; }

int c3o_get_vec(const c3o *o, int which, double *out) {
  const double *src = NULL;
  int n = o->n2 + 1;
  switch (which) {
    case 0: src = o->y; break;
    case 1: src = o->dy; break;
    case 2: src = o->h; break;
    case 3: src = o->hm; break;
    case 4: src = o->hc; break;
    case 5: src = o->hp; break;
    case 6: src = o->dym; break;
    case 7: src = o->dyc; break;
    case 8: src = o->dyp; break;
    case 9: src = o->ak1; n = o->n1 + 1; break;
    case 10: src = o->ak3; n = o->n3mh + 1; break;
    case 11: src = o->pmj; break;
    case 12: src = o->pcj; break;
    case 13: src = o->ppj; break;
    case 14: src = o->fjpa; break;
    case 15: src = o->fjmu; break;
    case 16: src = o->fjmv; break;
    default: return -1;
  }
  memcpy(out, src, sizeof(double) * n);
  return n;
}

double c3o_get_scalar(const c3o *o, int which) {
  switch (which) {
    case 100: return o->dx1;
    case 101: return o->dx1q;
    case 102: return o->dx3;
    case 103: return o->dx3q;
    case 104: return o->aly;
    case 105: return o->vol;
    case 106: return o->taow;
    default: return NAN;
  }
}

void c3o_get_tau(const c3o *o, double *tauw, double *taub) {
  size_t n = (size_t)(o->n1 + 1) * (o->n3 + 1) * 2;
  if (tauw) memcpy(tauw, o->tauw, n * sizeof(double));
  if (taub) memcpy(taub, o->taub, n * sizeof(double));
}
void c3o_set_tau(c3o *o, const double *tauw, const double *taub) {
  size_t n = (size_t)(o->n1 + 1) * (o->n3 + 1) * 2;
  if (tauw) memcpy(o->tauw, tauw, n * sizeof(double));
  if (taub) memcpy(o->taub, taub, n * sizeof(double));
}
void c3o_get_wsm(const c3o *o, double *w) { memcpy(w, o->wsm, 3 * sizeof(double)); }
void c3o_set_wsm(c3o *o, const double *w) { memcpy(o->wsm, w, 3 * sizeof(double)); }

/* ------------------------------------------------------------------ BCOND */
/* compact3d.f:963-1001 */
void c3o_bcond(c3o *o, double *U, double *SGSVIS) {
  const int N1M = o->n1m, N3M = o->n3m, N2 = o->n2, N2M = o->n2m;
  const double ALY = o->aly;
  for (int k = 1; k <= N3M; k++)
    for (int i = 1; i <= N1M; i++) {
      o->ubc[UBCX(i, k, 1, 1)] = 0.0;
      o->ubc[UBCX(i, k, 2, 1)] = 0.0;
      o->ubc[UBCX(i, k, 3, 1)] = 0.0;
      U[IX(i, 0, k)] = o->ubc[UBCX(i, k, 1, 1)];
      U[NU + IX(i, 0, k)] = o->ubc[UBCX(i, k, 2, 1)];
      U[2 * NU + IX(i, 0, k)] = o->ubc[UBCX(i, k, 3, 1)];
      o->ubc[UBCX(i, k, 1, 2)] = U[IX(i, N2M, k)] * (2.0 - ALY);
      o->ubc[UBCX(i, k, 2, 2)] = 0.0;
      o->ubc[UBCX(i, k, 3, 2)] = U[2 * NU + IX(i, N2M, k)] * (2.0 - ALY);
      U[IX(i, N2, k)] = o->ubc[UBCX(i, k, 1, 2)];
      U[NU + IX(i, N2, k)] = o->ubc[UBCX(i, k, 2, 2)];
      U[2 * NU + IX(i, N2, k)] = o->ubc[UBCX(i, k, 3, 2)];
      if (o->files != 0.0 && SGSVIS) {
        SGSVIS[IX(i, 0, k)] = 0.0;
        SGSVIS[IX(i, N2, k)] = SGSVIS[IX(i, N2M, k)] * (2.0 - ALY);
      }
    }
}

/* helpers for periodic line RHS builders; all arrays 1-based */
#define WRAPM(i, n) ((i) == 1 ? (n) : (i) - 1)
#define WRAPP(i, n) ((i) == (n) ? 1 : (i) + 1)

/* ------------------------------------------------------------------- RHS1 */
/* compact3d.f:6063-6782 */
void c3o_rhs1(c3o *o, const double *U, const double *P, double *HH,
              const double *HH2, double *RUH1, double presg,
              const double *SGSVIS) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m, N2 = o->n2;
  const double RE = o->re, DX1 = o->dx1, DX1Q = o->dx1q, DX3 = o->dx3,
               DX3Q = o->dx3q, ALY = o->aly;
  const double *DY = o->dy, *H = o->h;
  const int IFWALL = o->ifwall;
  const double *V = U + NU, *W = U + 2 * NU;
  double *H1 = HH;
  const double *H21 = HH2;

  /* loop 10: viscous y (:6104-6129) */
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++) {
      int jp = j + 1, jm = j - 1;
      for (int i = 1; i <= N1M; i++) {
        double viscos = (1. / RE) * (o->hp[j] * U[IX(i, jp, k)] -
                                     o->hc[j] * U[IX(i, j, k)] +
                                     o->hm[j] * U[IX(i, jm, k)]);
        RUH1[IX(i, j, k)] = viscos;
      }
    }
  /* loop 20: D2/Dx2 (:6142-6157) */
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++) {
      double d[N1M + 1], x[N1M + 1];
      for (int i = 1; i <= N1M; i++) {
        int im = WRAPM(i, N1M), ip = WRAPP(i, N1M);
        d[i] = 12. * DX1Q * (U[IX(im, j, k)] - 2. * U[IX(i, j, k)] + U[IX(ip, j, k)]);
      }
      cyc_solve(fxb(o, 10.0), d, x);
      for (int i = 1; i <= N1M; i++) RUH1[IX(i, j, k)] = RUH1[IX(i, j, k)] + (1. / RE) * x[i];
    }
  /* loop 30: D2/Dz2 (:6170-6185) */
#pragma omp parallel for
  for (int i = 1; i <= N1M; i++)
    for (int j = 1; j <= N2M; j++) {
      double d[N3M + 1], x[N3M + 1];
      for (int k = 1; k <= N3M; k++) {
        int km = WRAPM(k, N3M), kp = WRAPP(k, N3M);
        d[k] = 12. * DX3Q * (U[IX(i, j, km)] - 2. * U[IX(i, j, k)] + U[IX(i, j, kp)]);
      }
      cyc_solve(fzb(o, 10.0), d, x);
      for (int k = 1; k <= N3M; k++) RUH1[IX(i, j, k)] = RUH1[IX(i, j, k)] + (1. / RE) * x[k];
    }
  /* loop 40: pressure gradient (:6196-6210) */
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++) {
      double d[N1M + 1], x[N1M + 1];
      for (int i = 1; i <= N1M; i++)
        d[i] = 24. * DX1 * (-P[PX(WRAPM(i, N1M), j, k)] + P[PX(i, j, k)]);
      cyc_solve(fxb(o, 22.0), d, x);
      for (int i = 1; i <= N1M; i++)
        RUH1[IX(i, j, k)] = RUH1[IX(i, j, k)] - x[i] - presg;
    }
  /* D(UV)/DY (:6224-6276) */
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++) {
    double *VV1 = dalloc((size_t)(N1M + 1) * (N2M + 2)); /* VV1(I,J), J up to N2M+1 */
    for (int j = 1; j <= N2M; j++) {
      double d[N1M + 1], x[N1M + 1];
      for (int i = 1; i <= N1M; i++)
        d[i] = 4.0 * (V[IX(WRAPM(i, N1M), j, k)] + V[IX(i, j, k)]);
      cyc_solve(fxb(o, 6.0), d, x);
      for (int i = 1; i <= N1M; i++) VV1[(size_t)j * (N1M + 1) + i] = x[i];
    }
    for (int j = 1; j <= N2M; j++) {
      int jp = j + 1, jm = j - 1;
      double fjum = o->fjmu[j], fjup = o->fjpa[j];
      for (int i = 1; i <= N1M; i++) {
        double u2 = fjup * 1.0 / H[jp] * (DY[j] / 2. * U[IX(i, jp, k)] + DY[jp] / 2. * U[IX(i, j, k)]) +
                    (1. - fjup) * U[IX(i, N2, k)];
        double u1 = fjum * 1.0 / H[j] * (DY[jm] / 2. * U[IX(i, j, k)] + DY[j] / 2. * U[IX(i, jm, k)]) +
                    (1. - fjum) * U[IX(i, 0, k)];
        double v2 = VV1[(size_t)jp * (N1M + 1) + i];
        double v1 = VV1[(size_t)j * (N1M + 1) + i];
        if (j == N2M) v2 = 0.0;
        if (j == 1) v1 = 0.0;
        H1[IX(i, j, k)] = H1[IX(i, j, k)] + 1. / DY[j] * (v2 * u2 - v1 * u1);
      }
    }
    free(VV1);
  }
  /* D(UU)/DX (:6295-6334) */
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++) {
      double d[N1M + 1], uu1[N1M + 1], x[N1M + 1];
      for (int i = 1; i <= N1M; i++)
        d[i] = 4.0 * (U[IX(i, j, k)] + U[IX(WRAPP(i, N1M), j, k)]);
      cyc_solve(fxb(o, 6.0), d, uu1);
      for (int i = 1; i <= N1M; i++) {
        int im = WRAPM(i, N1M);
        d[i] = 24. * DX1 * (uu1[i] * uu1[i] - uu1[im] * uu1[im]);
      }
      cyc_solve(fxb(o, 22.0), d, x);
      for (int i = 1; i <= N1M; i++) H1[IX(i, j, k)] = H1[IX(i, j, k)] + x[i];
    }
  /* D(UW)/DZ (:6360-6433) */
#pragma omp parallel for
  for (int j = 1; j <= N2M; j++) {
    double *UU3 = dalloc((size_t)(N1M + 1) * (N3M + 1));
    double *WW1 = dalloc((size_t)(N1M + 1) * (N3M + 1));
#define A2(a, i, k) a[(size_t)(k) * (N1M + 1) + (i)]
    for (int i = 1; i <= N1M; i++) {
      double d[N3M + 1], x[N3M + 1];
      for (int k = 1; k <= N3M; k++) d[k] = 4. * (U[IX(i, j, WRAPM(k, N3M))] + U[IX(i, j, k)]);
      cyc_solve(fzb(o, 6.0), d, x);
      for (int k = 1; k <= N3M; k++) A2(UU3, i, k) = x[k];
    }
    for (int k = 1; k <= N3M; k++) {
      double d[N1M + 1], x[N1M + 1];
      for (int i = 1; i <= N1M; i++) d[i] = 4. * (W[IX(WRAPM(i, N1M), j, k)] + W[IX(i, j, k)]);
      cyc_solve(fxb(o, 6.0), d, x);
      for (int i = 1; i <= N1M; i++) A2(WW1, i, k) = x[i];
    }
    for (int i = 1; i <= N1M; i++) {
      double d[N3M + 1], x[N3M + 1];
      for (int k = 1; k <= N3M; k++) {
        int kp = WRAPP(k, N3M);
        d[k] = 24. * DX3 * (-A2(UU3, i, k) * A2(WW1, i, k) + A2(UU3, i, kp) * A2(WW1, i, kp));
      }
      cyc_solve(fzb(o, 22.0), d, x);
      for (int k = 1; k <= N3M; k++) H1[IX(i, j, k)] = H1[IX(i, j, k)] + x[k];
    }
    free(UU3);
    free(WW1);
  }

  if (o->files != 0.0) {
    /* D(SGS*(DU/DX))/DX (:6445-6466) */
#pragma omp parallel for
    for (int k = 1; k <= N3M; k++)
      for (int j = 1; j <= N2M; j++) {
        double d[N1M + 1], x[N1M + 1];
        for (int i = 1; i <= N1M; i++)
          d[i] = 24. * DX1 * (U[IX(WRAPP(i, N1M), j, k)] - U[IX(i, j, k)]);
        cyc_solve(fxb(o, 22.0), d, x);
        for (int i = 1; i <= N1M; i++) {
          int im = WRAPM(i, N1M);
          d[i] = 24. * DX1 * (SGSVIS[IX(i, j, k)] * x[i] - SGSVIS[IX(im, j, k)] * x[im]);
        }
        double x2[N1M + 1];
        cyc_solve(fxb(o, 22.0), d, x2);
        for (int i = 1; i <= N1M; i++) H1[IX(i, j, k)] = H1[IX(i, j, k)] - 2. * x2[i];
      }
    /* D(SGS*(DU/DZ))/DZ (:6491-6540) */
#pragma omp parallel for
    for (int j = 1; j <= N2M; j++) {
      double *S2 = dalloc((size_t)(N1M + 1) * (N3M + 1));
      for (int k = 1; k <= N3M; k++) {
        double d[N1M + 1], x[N1M + 1];
        for (int i = 1; i <= N1M; i++)
          d[i] = 4. * (SGSVIS[IX(i, j, k)] + SGSVIS[IX(WRAPM(i, N1M), j, k)]);
        cyc_solve(fxb(o, 6.0), d, x);
        for (int i = 1; i <= N1M; i++) A2(S2, i, k) = x[i];
      }
      for (int i = 1; i <= N1M; i++) {
        double d[N3M + 1], x3[N3M + 1], x33[N3M + 1], x4[N3M + 1];
        for (int k = 1; k <= N3M; k++) d[k] = 4. * (A2(S2, i, k) + A2(S2, i, WRAPM(k, N3M)));
        cyc_solve(fzb(o, 6.0), d, x3);
        for (int k = 1; k <= N3M; k++)
          d[k] = 24. * DX3 * (U[IX(i, j, k)] - U[IX(i, j, WRAPM(k, N3M))]);
        cyc_solve(fzb(o, 22.0), d, x33);
        for (int k = 1; k <= N3M; k++) {
          int kp = WRAPP(k, N3M);
          d[k] = 24. * DX3 * (x33[kp] * x3[kp] - x33[k] * x3[k]);
        }
        cyc_solve(fzb(o, 22.0), d, x4);
        for (int k = 1; k <= N3M; k++) H1[IX(i, j, k)] = H1[IX(i, j, k)] - x4[k];
      }
      free(S2);
    }
    /* D(SGS*(DU/DY))/DY (:6551-6604) */
#pragma omp parallel for
    for (int k = 1; k <= N3M; k++) {
      double *S1 = dalloc((size_t)(N1M + 1) * (N2M + 2)); /* rows 0..N2M+1, 0 and N2M+1 stay 0 */
#define B2(a, i, j) a[(size_t)(j) * (N1M + 1) + (i)]
      for (int j = 1; j <= N2M; j++) {
        double d[N1M + 1], x[N1M + 1];
        for (int i = 1; i <= N1M; i++)
          d[i] = 4. * (SGSVIS[IX(i, j, k)] + SGSVIS[IX(WRAPM(i, N1M), j, k)]);
        cyc_solve(fxb(o, 6.0), d, x);
        for (int i = 1; i <= N1M; i++) B2(S1, i, j) = x[i];
      }
      for (int i = 1; i <= N1M; i++) {
        int im = o->ima[i];
        for (int j = 1; j <= N2M; j++) {
          int jp = j + 1, jm = j - 1;
          double fjum = o->fjmu[j], fjup = o->fjpa[j];
          double sgsj2 = fjup * 1. / H[jp] * (DY[j] / 2. * B2(S1, i, jp) + DY[jp] / 2. * B2(S1, i, j)) +
                         (1. - fjup) * 0.5 * (SGSVIS[IX(i, N2, k)] + SGSVIS[IX(im, N2, k)]);
          double sgsj1 = fjum * 1. / H[j] * (DY[jm] / 2. * B2(S1, i, j) + DY[j] / 2. * B2(S1, i, jm)) +
                         (1. - fjum) * 0.5 * (SGSVIS[IX(i, 0, k)] + SGSVIS[IX(im, 0, k)]);
          double uu22 = (U[IX(i, jp, k)] - U[IX(i, j, k)]) / H[jp];
          double uu11 = (U[IX(i, j, k)] - U[IX(i, jm, k)]) / H[j];
          double sgsduy = 1. / DY[j] * ((1. - IFWALL * (1. - fjup) * (ALY - 1.)) * sgsj2 * uu22 -
                                        (1. - IFWALL * (1. - fjum)) * sgsj1 * uu11);
          H1[IX(i, j, k)] = H1[IX(i, j, k)] - sgsduy;
        }
      }
      free(S1);
    }
    /* D(SGS*(DV/DX))/DY (:6622-6680) */
#pragma omp parallel for
    for (int k = 1; k <= N3M; k++) {
      double *UU12 = dalloc((size_t)(N1M + 1) * (N2M + 2));
      double *S12 = dalloc((size_t)(N1M + 1) * (N2M + 2));
      for (int j = 1; j <= N2M; j++) {
        double d[N1M + 1], x[N1M + 1];
        for (int i = 1; i <= N1M; i++)
          d[i] = 24. * DX1 * (V[IX(i, j, k)] - V[IX(WRAPM(i, N1M), j, k)]);
        cyc_solve(fxb(o, 22.0), d, x);
        for (int i = 1; i <= N1M; i++) B2(UU12, i, j) = x[i];
        for (int i = 1; i <= N1M; i++)
          d[i] = 4. * (SGSVIS[IX(i, j, k)] + SGSVIS[IX(WRAPM(i, N1M), j, k)]);
        cyc_solve(fxb(o, 6.0), d, x);
        for (int i = 1; i <= N1M; i++) B2(S12, i, j) = x[i];
      }
      for (int i = 1; i <= N1M; i++) {
        int im = o->ima[i];
        for (int j = 1; j <= N2M; j++) {
          int jp = j + 1, jm = j - 1;
          double fjum = o->fjmu[j], fjup = o->fjpa[j];
          double sgsj2 = fjup * 1.0 / H[jp] * (DY[j] / 2. * B2(S12, i, jp) + DY[jp] / 2. * B2(S12, i, j)) +
                         (1. - fjup) * 0.5 * (SGSVIS[IX(i, N2, k)] + SGSVIS[IX(im, N2, k)]);
          double sgsj1 = fjum * 1.0 / H[j] * (DY[jm] / 2. * B2(S12, i, j) + DY[j] / 2. * B2(S12, i, jm)) +
                         (1. - fjum) * 0.5 * (SGSVIS[IX(i, 0, k)] + SGSVIS[IX(im, 0, k)]);
          double sgsduxy = 1.0 / DY[j] * ((1. - IFWALL * (1. - fjup) * (ALY - 1.)) * sgsj2 * B2(UU12, i, jp) -
                                          (1. - IFWALL * (1. - fjum)) * sgsj1 * B2(UU12, i, j));
          H1[IX(i, j, k)] = H1[IX(i, j, k)] - sgsduxy;
        }
      }
      free(UU12);
      free(S12);
    }
    /* D(SGS*(DW/DX))/DZ (:6713-6764) */
#pragma omp parallel for
    for (int j = 1; j <= N2M; j++) {
      double *S13 = dalloc((size_t)(N1M + 1) * (N3M + 1));
      double *WW13 = dalloc((size_t)(N1M + 1) * (N3M + 1));
      for (int k = 1; k <= N3M; k++) {
        double d[N1M + 1], x[N1M + 1];
        for (int i = 1; i <= N1M; i++)
          d[i] = 4. * (SGSVIS[IX(i, j, k)] + SGSVIS[IX(WRAPM(i, N1M), j, k)]);
        cyc_solve(fxb(o, 6.0), d, x);
        for (int i = 1; i <= N1M; i++) A2(S13, i, k) = x[i];
        for (int i = 1; i <= N1M; i++)
          d[i] = 24. * DX1 * (W[IX(i, j, k)] - W[IX(WRAPM(i, N1M), j, k)]);
        cyc_solve(fxb(o, 22.0), d, x);
        for (int i = 1; i <= N1M; i++) A2(WW13, i, k) = x[i];
      }
      for (int i = 1; i <= N1M; i++) {
        double d[N3M + 1], x3[N3M + 1], x33[N3M + 1];
        for (int k = 1; k <= N3M; k++) d[k] = 4. * (A2(S13, i, k) + A2(S13, i, WRAPM(k, N3M)));
        cyc_solve(fzb(o, 6.0), d, x3);
        for (int k = 1; k <= N3M; k++) {
          int kp = WRAPP(k, N3M);
          d[k] = 24. * DX3 * (A2(WW13, i, kp) * x3[kp] - A2(WW13, i, k) * x3[k]);
        }
        cyc_solve(fzb(o, 22.0), d, x33);
        for (int k = 1; k <= N3M; k++) H1[IX(i, j, k)] = H1[IX(i, j, k)] - x33[k];
      }
      free(S13);
      free(WW13);
    }
  }
  /* AB2 + wall-stress BC (:6765-6775) */
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++) {
      double fjum = o->fjmu[j], fjup = o->fjpa[j];
      for (int i = 1; i <= N1M; i++) {
        double bc_down = -o->tauw[TAUX(i, k, 1)] / DY[1] * IFWALL;
        double bc_up = -o->tauw[TAUX(i, k, 2)] / DY[N2M] * IFWALL * (ALY - 1.);
        double bc = (1. - fjum) * bc_down + (1. - fjup) * bc_up;
        RUH1[IX(i, j, k)] = RUH1[IX(i, j, k)] - 3. / 2. * H1[IX(i, j, k)] +
                            1. / 2. * H21[IX(i, j, k)] + bc;
      }
    }
}

/* ------------------------------------------------------------------- RHS2 */
/* compact3d.f:6784-7396 */
void c3o_rhs2(c3o *o, const double *U, const double *P, double *HH,
              const double *HH2, double *RUH2, const double *SGSVIS) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m;
  const double RE = o->re, DX1 = o->dx1, DX1Q = o->dx1q, DX3 = o->dx3, DX3Q = o->dx3q;
  const double *DY = o->dy, *H = o->h;
  const double *Uu = U, *V = U + NU, *W = U + 2 * NU;
  double *H2c = HH + NU;
  const double *H22 = HH2 + NU;

  /* loop 10 (:6825-6865) */
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 2; j <= N2M; j++) {
      int jp = j + 1, jm = j - 1;
      for (int i = 1; i <= N1M; i++) {
        double viscos = (1. / RE) * (o->dyp[j] * V[IX(i, jp, k)] - o->dyc[j] * V[IX(i, j, k)] +
                                     o->dym[j] * V[IX(i, jm, k)]);
        double pressg2 = (P[PX(i, j, k)] - P[PX(i, jm, k)]) / H[j];
        RUH2[IX(i, j, k)] = -pressg2 + viscos;
        double v2 = 0.5 * (V[IX(i, jp, k)] + V[IX(i, j, k)]);
        double v1 = 0.5 * (V[IX(i, j, k)] + V[IX(i, jm, k)]);
        H2c[IX(i, j, k)] = H2c[IX(i, j, k)] + 1. / H[j] * (v2 * v2 - v1 * v1);
        double sgsj2 = SGSVIS[IX(i, j, k)];
        double sgsj1 = SGSVIS[IX(i, jm, k)];
        v2 = 1. / DY[j] * (V[IX(i, jp, k)] - V[IX(i, j, k)]);
        v1 = 1. / DY[jm] * (V[IX(i, j, k)] - V[IX(i, jm, k)]);
        double vis22 = 1.0 / H[j] * (sgsj2 * v2 - sgsj1 * v1);
        H2c[IX(i, j, k)] = H2c[IX(i, j, k)] - 2. * vis22;
      }
    }
  /* loop 20: D2/Dx2 (:6879-6895) */
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 2; j <= N2M; j++) {
      double d[N1M + 1], x[N1M + 1];
      for (int i = 1; i <= N1M; i++) {
        int im = WRAPM(i, N1M), ip = WRAPP(i, N1M);
        d[i] = 12. * DX1Q * (V[IX(im, j, k)] - 2. * V[IX(i, j, k)] + V[IX(ip, j, k)]);
      }
      cyc_solve(fxb(o, 10.0), d, x);
      for (int i = 1; i <= N1M; i++) RUH2[IX(i, j, k)] = RUH2[IX(i, j, k)] + (1. / RE) * x[i];
    }
  /* loop 30: D2/Dz2 (:6907-6923) */
#pragma omp parallel for
  for (int i = 1; i <= N1M; i++)
    for (int j = 2; j <= N2M; j++) {
      double d[N3M + 1], x[N3M + 1];
      for (int k = 1; k <= N3M; k++) {
        int km = WRAPM(k, N3M), kp = WRAPP(k, N3M);
        d[k] = 12. * DX3Q * (V[IX(i, j, km)] - 2. * V[IX(i, j, k)] + V[IX(i, j, kp)]);
      }
      cyc_solve(fzb(o, 10.0), d, x);
      for (int k = 1; k <= N3M; k++) RUH2[IX(i, j, k)] = RUH2[IX(i, j, k)] + (1. / RE) * x[k];
    }
  /* D(UV)/DX (:6942-7006) */
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 2; j <= N2M; j++) {
      double d[N1M + 1], vv1[N1M + 1], uu2[N1M + 1], x[N1M + 1];
      for (int i = 1; i <= N1M; i++) d[i] = 4. * (V[IX(WRAPM(i, N1M), j, k)] + V[IX(i, j, k)]);
      cyc_solve(fxb(o, 6.0), d, vv1);
      for (int i = 1; i <= N1M; i++)
        uu2[i] = 1. / H[j] * (DY[j] / 2. * Uu[IX(i, j - 1, k)] + DY[j - 1] / 2. * Uu[IX(i, j, k)]);
      for (int i = 1; i <= N1M; i++) {
        int ip = WRAPP(i, N1M);
        d[i] = 24. * DX1 * (-uu2[i] * vv1[i] + uu2[ip] * vv1[ip]);
      }
      cyc_solve(fxb(o, 22.0), d, x);
      for (int i = 1; i <= N1M; i++) H2c[IX(i, j, k)] = H2c[IX(i, j, k)] + x[i];
    }
  /* D(VW)/DZ (:7023-7084) */
#pragma omp parallel for
  for (int i = 1; i <= N1M; i++)
    for (int j = 2; j <= N2M; j++) {
      double d[N3M + 1], vv3[N3M + 1], ww2[N3M + 1], x[N3M + 1];
      for (int k = 1; k <= N3M; k++) d[k] = 4. * (V[IX(i, j, WRAPM(k, N3M))] + V[IX(i, j, k)]);
      cyc_solve(fzb(o, 6.0), d, vv3);
      for (int k = 1; k <= N3M; k++)
        ww2[k] = 1. / H[j] * (DY[j] / 2. * W[IX(i, j - 1, k)] + DY[j - 1] / 2. * W[IX(i, j, k)]);
      for (int k = 1; k <= N3M; k++) {
        int kp = WRAPP(k, N3M);
        d[k] = 24. * DX3 * (-vv3[k] * ww2[k] + vv3[kp] * ww2[kp]);
      }
      cyc_solve(fzb(o, 22.0), d, x);
      for (int k = 1; k <= N3M; k++) H2c[IX(i, j, k)] = H2c[IX(i, j, k)] + x[k];
    }

  if (o->files != 0.0) {
    /* D(SGS*(DU/DY))/DX (:7104-7158) */
#pragma omp parallel for
    for (int k = 1; k <= N3M; k++) {
      double *SV1 = dalloc((size_t)(N1M + 1) * (N2M + 2));
      for (int j = 1; j <= N2M; j++) {
        double d[N1M + 1], x[N1M + 1];
        for (int i = 1; i <= N1M; i++)
          d[i] = 4. * (SGSVIS[IX(i, j, k)] + SGSVIS[IX(WRAPM(i, N1M), j, k)]);
        cyc_solve(fxb(o, 6.0), d, x);
        for (int i = 1; i <= N1M; i++) B2(SV1, i, j) = x[i];
      }
      for (int j = 2; j <= N2M; j++) {
        double sv21[N1M + 1], vv21[N1M + 1], d[N1M + 1], x[N1M + 1];
        for (int i = 1; i <= N1M; i++) {
          sv21[i] = 1. / H[j] * (DY[j - 1] / 2. * B2(SV1, i, j) + DY[j] / 2. * B2(SV1, i, j - 1));
          vv21[i] = 1. / H[j] * (Uu[IX(i, j, k)] - Uu[IX(i, j - 1, k)]);
        }
        for (int i = 1; i <= N1M; i++) {
          int ip = WRAPP(i, N1M);
          d[i] = 24. * DX1 * (sv21[ip] * vv21[ip] - sv21[i] * vv21[i]);
        }
        cyc_solve(fxb(o, 22.0), d, x);
        for (int i = 1; i <= N1M; i++) H2c[IX(i, j, k)] = H2c[IX(i, j, k)] - x[i];
      }
      free(SV1);
    }
    /* D(SGS*(DW/DY))/DZ (:7177-7226) */
#pragma omp parallel for
    for (int i = 1; i <= N1M; i++) {
      double *S3 = dalloc((size_t)(N2M + 2) * (N3M + 1));
#define C2(a, j, k) a[(size_t)(k) * (N2M + 2) + (j)]
      for (int j = 1; j <= N2M; j++) {
        double d[N3M + 1], x[N3M + 1];
        for (int k = 1; k <= N3M; k++)
          d[k] = 4. * (SGSVIS[IX(i, j, WRAPM(k, N3M))] + SGSVIS[IX(i, j, k)]);
        cyc_solve(fzb(o, 6.0), d, x);
        for (int k = 1; k <= N3M; k++) C2(S3, j, k) = x[k];
      }
      for (int j = 2; j <= N2M; j++) {
        double s23[N3M + 1], vv32[N3M + 1], d[N3M + 1], x[N3M + 1];
        for (int k = 1; k <= N3M; k++) {
          s23[k] = 1. / H[j] * (DY[j] / 2. * C2(S3, j - 1, k) + DY[j - 1] / 2. * C2(S3, j, k));
          vv32[k] = 1. / H[j] * (W[IX(i, j, k)] - W[IX(i, j - 1, k)]);
        }
        for (int k = 1; k <= N3M; k++) {
          int kp = WRAPP(k, N3M);
          d[k] = 24. * DX3 * (s23[kp] * vv32[kp] - s23[k] * vv32[k]);
        }
        cyc_solve(fzb(o, 22.0), d, x);
        for (int k = 1; k <= N3M; k++) H2c[IX(i, j, k)] = H2c[IX(i, j, k)] - x[k];
      }
      free(S3);
    }
    /* D(SGS*(DV/DZ))/DZ (:7245-7304) */
#pragma omp parallel for
    for (int i = 1; i <= N1M; i++) {
      double *SV3 = dalloc((size_t)(N2M + 2) * (N3M + 1));
      double *VV23 = dalloc((size_t)(N2M + 2) * (N3M + 1));
      for (int j = 1; j <= N2M; j++) {
        double d[N3M + 1], x[N3M + 1];
        for (int k = 1; k <= N3M; k++)
          d[k] = 4. * (SGSVIS[IX(i, j, WRAPM(k, N3M))] + SGSVIS[IX(i, j, k)]);
        cyc_solve(fzb(o, 6.0), d, x);
        for (int k = 1; k <= N3M; k++) C2(SV3, j, k) = x[k];
        for (int k = 1; k <= N3M; k++)
          d[k] = 24. * DX3 * (V[IX(i, j, k)] - V[IX(i, j, WRAPM(k, N3M))]);
        cyc_solve(fzb(o, 22.0), d, x);
        for (int k = 1; k <= N3M; k++) C2(VV23, j, k) = x[k];
      }
      for (int j = 2; j <= N2M; j++) {
        double s32[N3M + 1], d[N3M + 1], x[N3M + 1];
        for (int k = 1; k <= N3M; k++)
          s32[k] = 1. / H[j] * (DY[j] / 2. * C2(SV3, j - 1, k) + DY[j - 1] / 2. * C2(SV3, j, k));
        for (int k = 1; k <= N3M; k++) {
          int kp = WRAPP(k, N3M);
          d[k] = 24. * DX3 * (s32[kp] * C2(VV23, j, kp) - s32[k] * C2(VV23, j, k));
        }
        cyc_solve(fzb(o, 22.0), d, x);
        for (int k = 1; k <= N3M; k++) H2c[IX(i, j, k)] = H2c[IX(i, j, k)] - x[k];
      }
      free(SV3);
      free(VV23);
    }
    /* D(SGS*(DV/DX))/DX (:7323-7384) */
#pragma omp parallel for
    for (int k = 1; k <= N3M; k++) {
      double *S12 = dalloc((size_t)(N1M + 1) * (N2M + 2));
      for (int j = 1; j <= N2M; j++) {
        double d[N1M + 1], x[N1M + 1];
        for (int i = 1; i <= N1M; i++)
          d[i] = 4. * (SGSVIS[IX(i, j, k)] + SGSVIS[IX(WRAPM(i, N1M), j, k)]);
        cyc_solve(fxb(o, 6.0), d, x);
        for (int i = 1; i <= N1M; i++) B2(S12, i, j) = x[i];
      }
      for (int j = 2; j <= N2M; j++) {
        double vv12[N1M + 1], s22[N1M + 1], d[N1M + 1], x[N1M + 1];
        for (int i = 1; i <= N1M; i++)
          d[i] = 24. * DX1 * (V[IX(i, j, k)] - V[IX(WRAPM(i, N1M), j, k)]);
        cyc_solve(fxb(o, 22.0), d, vv12);
        for (int i = 1; i <= N1M; i++)
          s22[i] = 1.0 / H[j] * (DY[j - 1] / 2. * B2(S12, i, j) + DY[j] / 2. * B2(S12, i, j - 1));
        for (int i = 1; i <= N1M; i++) {
          int ip = WRAPP(i, N1M);
          d[i] = 24. * DX1 * (s22[ip] * vv12[ip] - s22[i] * vv12[i]);
        }
        cyc_solve(fxb(o, 22.0), d, x);
        for (int i = 1; i <= N1M; i++) H2c[IX(i, j, k)] = H2c[IX(i, j, k)] - x[i];
      }
      free(S12);
    }
  }
  /* AB2 (:7386-7389) */
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 2; j <= N2M; j++)
      for (int i = 1; i <= N1M; i++)
        RUH2[IX(i, j, k)] = RUH2[IX(i, j, k)] - 3. / 2. * H2c[IX(i, j, k)] + 1. / 2. * H22[IX(i, j, k)];
}

/* ------------------------------------------------------------------- RHS3 */
/* compact3d.f:7399-8114 */
void c3o_rhs3(c3o *o, const double *U, const double *P, double *HH,
              const double *HH2, double *RUH3, double presg3,
              const double *SGSVIS) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m, N2 = o->n2;
  const double RE = o->re, DX1 = o->dx1, DX1Q = o->dx1q, DX3 = o->dx3,
               DX3Q = o->dx3q, ALY = o->aly;
  const double *DY = o->dy, *H = o->h;
  const int IFWALL = o->ifwall;
  const double *Uu = U, *V = U + NU, *W = U + 2 * NU;
  double *H3 = HH + 2 * NU;
  const double *H23 = HH2 + 2 * NU;

  /* loop 10 (:7442-7469) */
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++) {
      int jp = j + 1, jm = j - 1;
      for (int i = 1; i <= N1M; i++)
        RUH3[IX(i, j, k)] = (1. / RE) * (o->hp[j] * W[IX(i, jp, k)] - o->hc[j] * W[IX(i, j, k)] +
                                         o->hm[j] * W[IX(i, jm, k)]);
    }
  /* loop 20 (:7483-7498) */
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++) {
      double d[N1M + 1], x[N1M + 1];
      for (int i = 1; i <= N1M; i++) {
        int im = WRAPM(i, N1M), ip = WRAPP(i, N1M);
        d[i] = 12. * DX1Q * (W[IX(im, j, k)] - 2. * W[IX(i, j, k)] + W[IX(ip, j, k)]);
      }
      cyc_solve(fxb(o, 10.0), d, x);
      for (int i = 1; i <= N1M; i++) RUH3[IX(i, j, k)] = RUH3[IX(i, j, k)] + (1. / RE) * x[i];
    }
  /* loop 30 (:7510-7525) */
#pragma omp parallel for
  for (int i = 1; i <= N1M; i++)
    for (int j = 1; j <= N2M; j++) {
      double d[N3M + 1], x[N3M + 1];
      for (int k = 1; k <= N3M; k++) {
        int km = WRAPM(k, N3M), kp = WRAPP(k, N3M);
        d[k] = 12. * DX3Q * (W[IX(i, j, km)] - 2. * W[IX(i, j, k)] + W[IX(i, j, kp)]);
      }
      cyc_solve(fzb(o, 10.0), d, x);
      for (int k = 1; k <= N3M; k++) RUH3[IX(i, j, k)] = RUH3[IX(i, j, k)] + (1. / RE) * x[k];
    }
  /* loop 40: pressure gradient (:7536-7550) */
#pragma omp parallel for
  for (int i = 1; i <= N1M; i++)
    for (int j = 1; j <= N2M; j++) {
      double d[N3M + 1], x[N3M + 1];
      for (int k = 1; k <= N3M; k++)
        d[k] = 24. * DX3 * (-P[PX(i, j, WRAPM(k, N3M))] + P[PX(i, j, k)]);
      cyc_solve(fzb(o, 22.0), d, x);
      for (int k = 1; k <= N3M; k++) RUH3[IX(i, j, k)] = RUH3[IX(i, j, k)] - x[k] - presg3;
    }
  /* D(WV)/DY (:7563-7613) */
#pragma omp parallel for
  for (int i = 1; i <= N1M; i++) {
    double *VV3 = dalloc((size_t)(N2M + 2) * (N3M + 1));
    for (int j = 1; j <= N2M; j++) {
      double d[N3M + 1], x[N3M + 1];
      for (int k = 1; k <= N3M; k++) d[k] = 4.0 * (V[IX(i, j, WRAPM(k, N3M))] + V[IX(i, j, k)]);
      cyc_solve(fzb(o, 6.0), d, x);
      for (int k = 1; k <= N3M; k++) C2(VV3, j, k) = x[k];
    }
    for (int k = 1; k <= N3M; k++)
      for (int j = 1; j <= N2M; j++) {
        int jp = j + 1, jm = j - 1;
        double fjum = o->fjmu[j], fjup = o->fjpa[j];
        double w2 = fjup * 1.0 / H[jp] * (DY[jp] / 2. * W[IX(i, j, k)] + DY[j] / 2. * W[IX(i, jp, k)]) +
                    (1. - fjup) * W[IX(i, N2, k)];
        double w1 = fjum * 1.0 / H[j] * (DY[j] / 2. * W[IX(i, jm, k)] + DY[jm] / 2. * W[IX(i, j, k)]) +
                    (1. - fjum) * W[IX(i, 0, k)];
        double v2 = C2(VV3, jp, k), v1 = C2(VV3, j, k);
        if (j == N2M) v2 = 0.0;
        if (j == 1) v1 = 0.0;
        H3[IX(i, j, k)] = H3[IX(i, j, k)] + 1. / DY[j] * (v2 * w2 - v1 * w1);
      }
    free(VV3);
  }
  /* D(WW)/DZ (:7633-7672) */
#pragma omp parallel for
  for (int i = 1; i <= N1M; i++)
    for (int j = 1; j <= N2M; j++) {
      double d[N3M + 1], ww3[N3M + 1], x[N3M + 1];
      for (int k = 1; k <= N3M; k++) d[k] = 4. * (W[IX(i, j, k)] + W[IX(i, j, WRAPP(k, N3M))]);
      cyc_solve(fzb(o, 6.0), d, ww3);
      for (int k = 1; k <= N3M; k++) {
        int km = WRAPM(k, N3M);
        d[k] = 24. * DX3 * (ww3[k] * ww3[k] - ww3[km] * ww3[km]);
      }
      cyc_solve(fzb(o, 22.0), d, x);
      for (int k = 1; k <= N3M; k++) H3[IX(i, j, k)] = H3[IX(i, j, k)] + x[k];
    }
  /* D(UW)/DX (:7698-7771) */
#pragma omp parallel for
  for (int j = 1; j <= N2M; j++) {
    double *UU3 = dalloc((size_t)(N1M + 1) * (N3M + 1));
    double *WW1 = dalloc((size_t)(N1M + 1) * (N3M + 1));
    for (int i = 1; i <= N1M; i++) {
      double d[N3M + 1], x[N3M + 1];
      for (int k = 1; k <= N3M; k++) d[k] = 4. * (Uu[IX(i, j, WRAPM(k, N3M))] + Uu[IX(i, j, k)]);
      cyc_solve(fzb(o, 6.0), d, x);
      for (int k = 1; k <= N3M; k++) A2(UU3, i, k) = x[k];
    }
    for (int k = 1; k <= N3M; k++) {
      double d[N1M + 1], x[N1M + 1];
      for (int i = 1; i <= N1M; i++) d[i] = 4. * (W[IX(WRAPM(i, N1M), j, k)] + W[IX(i, j, k)]);
      cyc_solve(fxb(o, 6.0), d, x);
      for (int i = 1; i <= N1M; i++) A2(WW1, i, k) = x[i];
    }
    for (int k = 1; k <= N3M; k++) {
      double d[N1M + 1], x[N1M + 1];
      for (int i = 1; i <= N1M; i++) {
        int ip = WRAPP(i, N1M);
        d[i] = 24. * DX1 * (-A2(UU3, i, k) * A2(WW1, i, k) + A2(UU3, ip, k) * A2(WW1, ip, k));
      }
      cyc_solve(fxb(o, 22.0), d, x);
      for (int i = 1; i <= N1M; i++) H3[IX(i, j, k)] = H3[IX(i, j, k)] + x[i];
    }
    free(UU3);
    free(WW1);
  }

  if (o->files != 0.0) {
    /* D(SGS*(DW/DX))/DX (:7797-7847) */
#pragma omp parallel for
    for (int j = 1; j <= N2M; j++) {
      double *SW3 = dalloc((size_t)(N1M + 1) * (N3M + 1));
      for (int i = 1; i <= N1M; i++) {
        double d[N3M + 1], x[N3M + 1];
        for (int k = 1; k <= N3M; k++)
          d[k] = 4. * (SGSVIS[IX(i, j, WRAPM(k, N3M))] + SGSVIS[IX(i, j, k)]);
        cyc_solve(fzb(o, 6.0), d, x);
        for (int k = 1; k <= N3M; k++) A2(SW3, i, k) = x[k];
      }
      for (int k = 1; k <= N3M; k++) {
        double d[N1M + 1], x1[N1M + 1], x11[N1M + 1], x2[N1M + 1];
        for (int i = 1; i <= N1M; i++) d[i] = 4. * (A2(SW3, WRAPM(i, N1M), k) + A2(SW3, i, k));
        cyc_solve(fxb(o, 6.0), d, x1);
        for (int i = 1; i <= N1M; i++) d[i] = 24. * DX1 * (W[IX(i, j, k)] - W[IX(WRAPM(i, N1M), j, k)]);
        cyc_solve(fxb(o, 22.0), d, x11);
        for (int i = 1; i <= N1M; i++) {
          int ip = WRAPP(i, N1M);
          d[i] = 24. * DX1 * (x1[ip] * x11[ip] - x1[i] * x11[i]);
        }
        cyc_solve(fxb(o, 22.0), d, x2);
        for (int i = 1; i <= N1M; i++) H3[IX(i, j, k)] = H3[IX(i, j, k)] - x2[i];
      }
      free(SW3);
    }
    /* D(SGS*(DW/DY))/DY (:7858-7908) */
#pragma omp parallel for
    for (int i = 1; i <= N1M; i++) {
      double *SW2 = dalloc((size_t)(N2M + 2) * (N3M + 1));
      for (int j = 1; j <= N2M; j++) {
        double d[N3M + 1], x[N3M + 1];
        for (int k = 1; k <= N3M; k++)
          d[k] = 4. * (SGSVIS[IX(i, j, k)] + SGSVIS[IX(i, j, WRAPM(k, N3M))]);
        cyc_solve(fzb(o, 6.0), d, x);
        for (int k = 1; k <= N3M; k++) C2(SW2, j, k) = x[k];
      }
      for (int k = 1; k <= N3M; k++) {
        int km = o->kma[k];
        for (int j = 1; j <= N2M; j++) {
          int jp = j + 1, jm = j - 1;
          double fjum = o->fjmu[j], fjup = o->fjpa[j];
          double ww22 = 1. / H[jp] * (W[IX(i, jp, k)] - W[IX(i, j, k)]);
          double ww11 = 1. / H[j] * (W[IX(i, j, k)] - W[IX(i, jm, k)]);
          double sgsj1 = fjum * 1. / H[j] * (DY[jm] / 2. * C2(SW2, j, k) + DY[j] / 2. * C2(SW2, jm, k)) +
                         (1. - fjum) * 0.5 * (SGSVIS[IX(i, 0, k)] + SGSVIS[IX(i, 0, km)]);
          double sgsj2 = fjup * 1.0 / H[jp] * (DY[j] / 2. * C2(SW2, jp, k) + DY[jp] / 2. * C2(SW2, j, k)) +
                         (1. - fjup) * 0.5 * (SGSVIS[IX(i, N2, k)] + SGSVIS[IX(i, N2, km)]);
          double sgsdwy = 1. / DY[j] * ((1. - IFWALL * (1. - fjup) * (ALY - 1.)) * sgsj2 * ww22 -
                                        (1. - IFWALL * (1. - fjum)) * sgsj1 * ww11);
          H3[IX(i, j, k)] = H3[IX(i, j, k)] - sgsdwy;
        }
      }
      free(SW2);
    }
    /* D(SGS*(DW/DZ))/DZ (:7919-7940) */
#pragma omp parallel for
    for (int i = 1; i <= N1M; i++)
      for (int j = 1; j <= N2M; j++) {
        double d[N3M + 1], x33[N3M + 1], x2[N3M + 1];
        for (int k = 1; k <= N3M; k++)
          d[k] = 24. * DX3 * (W[IX(i, j, WRAPP(k, N3M))] - W[IX(i, j, k)]);
        cyc_solve(fzb(o, 22.0), d, x33);
        for (int k = 1; k <= N3M; k++) {
          int km = WRAPM(k, N3M);
          d[k] = 24. * DX3 * (x33[k] * SGSVIS[IX(i, j, k)] - x33[km] * SGSVIS[IX(i, j, km)]);
        }
        cyc_solve(fzb(o, 22.0), d, x2);
        for (int k = 1; k <= N3M; k++) H3[IX(i, j, k)] = H3[IX(i, j, k)] - 2. * x2[k];
      }
    /* D(SGS*(DU/DZ))/DX (:7972-8024) */
#pragma omp parallel for
    for (int j = 1; j <= N2M; j++) {
      double *UU31 = dalloc((size_t)(N1M + 1) * (N3M + 1));
      double *S31 = dalloc((size_t)(N1M + 1) * (N3M + 1));
      for (int i = 1; i <= N1M; i++) {
        double d[N3M + 1], x[N3M + 1];
        for (int k = 1; k <= N3M; k++)
          d[k] = 24. * DX3 * (Uu[IX(i, j, k)] - Uu[IX(i, j, WRAPM(k, N3M))]);
        cyc_solve(fzb(o, 22.0), d, x);
        for (int k = 1; k <= N3M; k++) A2(UU31, i, k) = x[k];
        for (int k = 1; k <= N3M; k++)
          d[k] = 4. * (SGSVIS[IX(i, j, k)] + SGSVIS[IX(i, j, WRAPM(k, N3M))]);
        cyc_solve(fzb(o, 6.0), d, x);
        for (int k = 1; k <= N3M; k++) A2(S31, i, k) = x[k];
      }
      for (int k = 1; k <= N3M; k++) {
        double d[N1M + 1], x1[N1M + 1], x[N1M + 1];
        for (int i = 1; i <= N1M; i++) d[i] = 4. * (A2(S31, i, k) + A2(S31, WRAPM(i, N1M), k));
        cyc_solve(fxb(o, 6.0), d, x1);
        for (int i = 1; i <= N1M; i++) {
          int ip = WRAPP(i, N1M);
          d[i] = 24. * DX1 * (x1[ip] * A2(UU31, ip, k) - x1[i] * A2(UU31, i, k));
        }
        cyc_solve(fxb(o, 22.0), d, x);
        for (int i = 1; i <= N1M; i++) H3[IX(i, j, k)] = H3[IX(i, j, k)] - x[i];
      }
      free(UU31);
      free(S31);
    }
    /* D(SGS*(DV/DZ))/DY, part 1 (:8042-8094) */
#pragma omp parallel for
    for (int i = 1; i <= N1M; i++) {
      double *VV32 = dalloc((size_t)(N2M + 2) * (N3M + 1));
      double *S32 = dalloc((size_t)(N2M + 2) * (N3M + 1));
      for (int j = 1; j <= N2M; j++) {
        double d[N3M + 1], x[N3M + 1];
        for (int k = 1; k <= N3M; k++)
          d[k] = 24. * DX3 * (V[IX(i, j, k)] - V[IX(i, j, WRAPM(k, N3M))]);
        cyc_solve(fzb(o, 22.0), d, x);
        for (int k = 1; k <= N3M; k++) C2(VV32, j, k) = x[k];
        for (int k = 1; k <= N3M; k++)
          d[k] = 4. * (SGSVIS[IX(i, j, k)] + SGSVIS[IX(i, j, WRAPM(k, N3M))]);
        cyc_solve(fzb(o, 6.0), d, x);
        for (int k = 1; k <= N3M; k++) C2(S32, j, k) = x[k];
      }
      for (int k = 1; k <= N3M; k++) {
        int km = o->kma[k];
        for (int j = 1; j <= N2M; j++) {
          int jp = j + 1, jm = j - 1;
          double fjum = o->fjmu[j], fjup = o->fjpa[j];
          double sgsj2 = fjup * 1.0 / H[jp] * (DY[j] / 2. * C2(S32, jp, k) + DY[jp] / 2. * C2(S32, j, k)) +
                         (1. - fjup) * 0.5 * (SGSVIS[IX(i, N2, k)] + SGSVIS[IX(i, N2, km)]);
          double sgsj1 = fjum * 1.0 / H[j] * (DY[jm] / 2. * C2(S32, j, k) + DY[j] / 2. * C2(S32, jm, k)) +
                         (1. - fjum) * 0.5 * (SGSVIS[IX(i, 0, k)] + SGSVIS[IX(i, 0, km)]);
          double sgsdvy = 1.0 / DY[j] * ((1. - IFWALL * (1. - fjup) * (ALY - 1.)) * sgsj2 * C2(VV32, jp, k) -
                                         (1. - IFWALL * (1. - fjum)) * sgsj1 * C2(VV32, j, k));
          H3[IX(i, j, k)] = H3[IX(i, j, k)] - sgsdvy;
        }
      }
      free(VV32);
      free(S32);
    }
  }
  /* wall-stress BC + AB2 (:8096-8103) */
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++) {
      double fjum = o->fjmu[j], fjup = o->fjpa[j];
      for (int i = 1; i <= N1M; i++) {
        double bc_down = -o->taub[TAUX(i, k, 1)] / DY[1] * IFWALL;
        double bc_up = -o->taub[TAUX(i, k, 2)] / DY[N2M] * IFWALL * (ALY - 1.);
        double bc = (1. - fjum) * bc_down + (1. - fjup) * bc_up;
        RUH3[IX(i, j, k)] = RUH3[IX(i, j, k)] - 3. / 2. * H3[IX(i, j, k)] +
                            1. / 2. * H23[IX(i, j, k)] + bc;
      }
    }
}

/* ------------------------------------------------------------ GETUH1/2/3 */
/* y-direction Thomas TDMAI (compact3d.f:8630-8659) on one line j=ns..nf;
 * a,b,c,d indexed by j; x may alias d */
static void tdmai_line(int ns, int nf, const double *a, const double *b,
                       const double *c, const double *d, double *x) {
  double gam[nf + 2];
  double bet = 1.0 / b[ns];
  gam[ns] = c[ns] * bet;
  x[ns] = d[ns] * bet;
  for (int j = ns + 1; j <= nf; j++) {
    bet = 1.0 / (b[j] - a[j] * gam[j - 1]);
    gam[j] = c[j] * bet;
    x[j] = (d[j] - a[j] * x[j - 1]) * bet;
  }
  for (int j = nf - 1; j >= ns; j--) x[j] = x[j] - x[j + 1] * gam[j];
}

/* y-sweep of GETUH1/3 (:8249-8282, :8508-8543) and GETUH2 (:8374-8411) */
static void getuh_y(c3o *o, double *UHc, const double *RUHc, int comp) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m;
  const double RE = o->re, DT = o->dt;
  const int ns = (comp == 2) ? 2 : 1;
  double am[N2M + 2], ac[N2M + 2], ap[N2M + 2];
  for (int j = ns; j <= N2M; j++) {
    double fjup = o->fjpa[j];
    if (comp == 2) {
      double fjum = o->fjmv[j];
      ap[j] = fjup * (-(0.5 * o->dyp[j] * (1. / RE))) * DT;
      ac[j] = 1.0 + (+0.5 * o->dyc[j] * (1. / RE)) * DT;
      am[j] = fjum * (-(0.5 * o->dym[j] * (1. / RE))) * DT;
    } else {
      double fjum = o->fjmu[j];
      ap[j] = fjup * (-(0.5 * o->hp[j] * (1. / RE))) * DT;
      ac[j] = 1.0 + (0.5 * o->hc[j] * (1. / RE)) * DT;
      am[j] = fjum * (-(0.5 * o->hm[j] * (1. / RE))) * DT;
    }
  }
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int i = 1; i <= N1M; i++) {
      double r[N2M + 2];
      for (int j = ns; j <= N2M; j++)
        r[j] = (comp == 2) ? DT * RUHc[IX(i, j, k)] : RUHc[IX(i, j, k)] * DT;
      tdmai_line(ns, N2M, am, ac, ap, r, r);
      for (int j = ns; j <= N2M; j++) UHc[IX(i, j, k)] = r[j];
    }
}

/* x-sweep (CTDMA1J, :8285-8311) and z-sweep (CTDMA3I, :8314-8338) */
static void getuh_x(c3o *o, double *UHc, int js) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m;
  const double RE = o->re, DT = o->dt, DX1Q = o->dx1q;
  double ai[N1M + 1], bi[N1M + 1];
  for (int i = 1; i <= N1M; i++) {
    ai[i] = 1.0 - DT / 2.0 * (1. / RE) * 12.0 * DX1Q * (1.0);
    bi[i] = 10.0 - DT / 2.0 * (1. / RE) * 12.0 * DX1Q * (-2.0);
  }
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = js; j <= N2M; j++) {
      double r[N1M + 1];
      for (int i = 1; i <= N1M; i++) {
        int im = WRAPM(i, N1M), ip = WRAPP(i, N1M);
        r[i] = UHc[IX(im, j, k)] + 10.0 * UHc[IX(i, j, k)] + UHc[IX(ip, j, k)];
      }
      cyc_solve_general(N1M, ai, bi, ai, r, r);
      for (int i = 1; i <= N1M; i++) UHc[IX(i, j, k)] = r[i];
    }
}

static void getuh_z(c3o *o, double *UHc, int js) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m;
  const double RE = o->re, DT = o->dt, DX3Q = o->dx3q;
  double ak[N3M + 1], bk[N3M + 1];
  for (int k = 1; k <= N3M; k++) {
    ak[k] = 1.0 - DT / 2.0 * (1. / RE) * 12.0 * DX3Q * (1.0);
    bk[k] = 10.0 - DT / 2.0 * (1. / RE) * 12.0 * DX3Q * (-2.0);
  }
#pragma omp parallel for
  for (int j = js; j <= N2M; j++)
    for (int i = 1; i <= N1M; i++) {
      double r[N3M + 1];
      for (int k = 1; k <= N3M; k++) {
        int km = WRAPM(k, N3M), kp = WRAPP(k, N3M);
        r[k] = UHc[IX(i, j, km)] + 10.0 * UHc[IX(i, j, k)] + UHc[IX(i, j, kp)];
      }
      cyc_solve_general(N3M, ak, bk, ak, r, r);
      for (int k = 1; k <= N3M; k++) UHc[IX(i, j, k)] = r[k];
    }
}

/* GETUH1 (compact3d.f:8218-8341): y, x, z; leaves delta in UH(.,1) */
void c3o_getuh1(c3o *o, const double *U, double *UH, const double *RUH1) {
  (void)U;
  getuh_y(o, UH, RUH1, 1);
  getuh_x(o, UH, 1);
  getuh_z(o, UH, 1);
}
/* GETUH2 (compact3d.f:8344-8475): y, x, z on J=2..N2M */
void c3o_getuh2(c3o *o, const double *U, double *UH, const double *RUH2) {
  (void)U;
  getuh_y(o, UH + NU, RUH2, 2);
  getuh_x(o, UH + NU, 2);
  getuh_z(o, UH + NU, 2);
}
/* GETUH3 (compact3d.f:8478-8627): y, z, x, then UH = U + delta for all three */
void c3o_getuh3(c3o *o, const double *U, double *UH, const double *RUH3) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m;
  getuh_y(o, UH + 2 * NU, RUH3, 3);
  getuh_z(o, UH + 2 * NU, 1);
  getuh_x(o, UH + 2 * NU, 1);
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++)
      for (int i = 1; i <= N1M; i++) {
        UH[IX(i, j, k)] = U[IX(i, j, k)] + UH[IX(i, j, k)];
        UH[2 * NU + IX(i, j, k)] = U[2 * NU + IX(i, j, k)] + UH[2 * NU + IX(i, j, k)];
      }
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 2; j <= N2M; j++)
      for (int i = 1; i <= N1M; i++)
        UH[NU + IX(i, j, k)] = U[NU + IX(i, j, k)] + UH[NU + IX(i, j, k)];
}

/* ------------------------------------------------------------------ RHSDP */
/* compact3d.f:8117-8214 */
void c3o_rhsdp(c3o *o, double *RDP, const double *UH) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m;
  const double DX1 = o->dx1, DX3 = o->dx3, DT = o->dt;
  const double *DY = o->dy;
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++) {
      int jp = j + 1;
      double fjum = o->fjmu[j], fjup = o->fjpa[j];
      for (int i = 1; i <= N1M; i++) {
        double divuh = (fjup * UH[NU + IX(i, jp, k)] - fjum * UH[NU + IX(i, j, k)]) / DY[j];
        double cbc = (1. - fjum) * o->ubc[UBCX(i, k, 2, 1)] / DY[j] -
                     (1. - fjup) * o->ubc[UBCX(i, k, 2, 2)] / DY[j];
        RDP[PX(i, j, k)] = (divuh - cbc) / DT;
      }
    }
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++) {
      double d[N1M + 1], x[N1M + 1];
      for (int i = 1; i <= N1M; i++)
        d[i] = 24. * DX1 * (UH[IX(WRAPP(i, N1M), j, k)] - UH[IX(i, j, k)]);
      cyc_solve(fxb(o, 22.0), d, x);
      for (int i = 1; i <= N1M; i++) RDP[PX(i, j, k)] = RDP[PX(i, j, k)] + x[i] / DT;
    }
#pragma omp parallel for
  for (int i = 1; i <= N1M; i++)
    for (int j = 1; j <= N2M; j++) {
      double d[N3M + 1], x[N3M + 1];
      for (int k = 1; k <= N3M; k++)
        d[k] = 24. * DX3 * (UH[2 * NU + IX(i, j, WRAPP(k, N3M))] - UH[2 * NU + IX(i, j, k)]);
      cyc_solve(fzb(o, 22.0), d, x);
      for (int k = 1; k <= N3M; k++) RDP[PX(i, j, k)] = RDP[PX(i, j, k)] + x[k] / DT;
    }
}

/* ------------------------------------------------------------ DFT (FFTW) */
typedef double complex cplx;

static int small_factor(int n) {
  if (n % 4 == 0) return 4;
  for (int r = 2; r * r <= n; r++)
    if (n % r == 0) return r;
  return n;
}

/* y[k] = sum_j x[j*xs] * exp(sign*2 pi i j k / n); tw[m] = exp(sign*2 pi i m / N),
 * n = N / s */
static void fft_rec(int n, int s, const cplx *x, int xs, cplx *y,
                    const cplx *tw, int N) {
  if (n == 1) {
    y[0] = x[0];
    return;
  }
  int r = small_factor(n);
  int m = n / r;
  for (int q = 0; q < r; q++) fft_rec(m, s * r, x + (size_t)q * xs, xs * r, y + (size_t)q * m, tw, N);
  cplx t[r];
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < r; q++) t[q] = y[(size_t)q * m + k] * tw[((long)q * k * s) % N];
    for (int p = 0; p < r; p++) {
      cplx acc = 0;
      for (int q = 0; q < r; q++) acc += t[q] * tw[((long)q * p * m * s) % N];
      y[(size_t)k + (size_t)p * m] = acc;
    }
  }
}

static void fft1d(int n, cplx *data, int stride, int sign, const cplx *tw) {
  cplx in[n], out[n];
  for (int j = 0; j < n; j++) in[j] = data[(size_t)j * stride];
  fft_rec(n, 1, in, 1, out, tw, n);
  for (int j = 0; j < n; j++) data[(size_t)j * stride] = out[j];
  (void)sign;
}

static cplx *make_tw(int n, int sign) {
  cplx *tw = (cplx *)malloc(sizeof(cplx) * n);
  const long double twopi = 6.283185307179586476925286766559L;
  for (int m = 0; m < n; m++) {
    long double a = sign * twopi * (long double)m / (long double)n;
    tw[m] = (double)cosl(a) + I * (double)sinl(a);
  }
  return tw;
}

void c3o_dft2d(int n1, int n3, double *ri, int sign) {
  cplx *c = (cplx *)ri;
  cplx *tw1 = make_tw(n1, sign), *tw3 = make_tw(n3, sign);
  for (int k = 0; k < n3; k++) fft1d(n1, c + (size_t)k * n1, 1, sign, tw1);
  for (int i = 0; i < n1; i++) fft1d(n3, c + i, n1, sign, tw3);
  free(tw1);
  free(tw3);
}

/* TDMAP (compact3d.f:9444-9459) */
static void tdmap(const double *a, const double *b, const double *c,
                  const double *r, int n, double *x) {
  double gam[n + 2];
  double bet = b[1];
  x[1] = r[1] / bet;
  for (int j = 2; j <= n; j++) {
    gam[j] = c[j - 1] / bet;
    bet = b[j] - a[j] * gam[j];
    x[j] = (r[j] - a[j] * x[j - 1]) / bet;
  }
  for (int j = n - 1; j >= 1; j--) x[j] = x[j] - gam[j + 1] * x[j + 1];
}

/* ------------------------------------------------------------------ GETDP */
/* compact3d.f:9271-9441 */
void c3o_getdp(c3o *o, double *DP, const double *RDP) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m, N3MH = o->n3mh;
  const int N1MH = N1M / 2 + 1;
  cplx *coef = (cplx *)malloc(sizeof(cplx) * (size_t)N1M * N3M * N2M);
  cplx *F = (cplx *)malloc(sizeof(cplx) * (size_t)N1M * N3MH * N2M); /* F(i,k,j) */
  cplx *tw1f = make_tw(N1M, -1), *tw3f = make_tw(N3M, -1);
  cplx *tw1b = make_tw(N1M, +1), *tw3b = make_tw(N3M, +1);
#define CO(i, k, j) coef[((size_t)(j - 1) * N3M + (k - 1)) * N1M + (i - 1)]
#define FF(i, k, j) F[((size_t)(j - 1) * N3MH + (k - 1)) * N1M + (i - 1)]
#pragma omp parallel for
  for (int j = 1; j <= N2M; j++) {
    for (int k = 1; k <= N3M; k++)
      for (int i = 1; i <= N1M; i++) CO(i, k, j) = RDP[PX(i, j, k)];
    cplx *pl = &CO(1, 1, j);
    for (int k = 0; k < N3M; k++) fft1d(N1M, pl + (size_t)k * N1M, 1, -1, tw1f);
    for (int i = 0; i < N1M; i++) fft1d(N3M, pl + i, N1M, -1, tw3f);
    for (int k = 1; k <= N3MH; k++)
      for (int i = 1; i <= N1M; i++) FF(i, k, j) = CO(i, k, j);
  }
  /* real parts (:9353-9377) and imaginary parts (:9381-9406) */
#pragma omp parallel for
  for (int k = 1; k <= N3MH; k++)
    for (int i = 1; i <= N1M; i++) {
      double cm[N2M + 2], cc[N2M + 2], cp[N2M + 2], cf[N2M + 2];
      for (int j = 1; j <= N2M; j++) {
        cm[j] = o->pmj[j];
        cc[j] = o->pcj[j] - o->ak3[k] - o->ak1[i];
        cp[j] = o->ppj[j];
        cf[j] = creal(FF(i, k, j));
        if (i == 1 && k == 1) {
          cm[1] = 0.0;
          cc[1] = 1.0;
          cp[1] = 0.0;
          cf[1] = 0.0;
        }
      }
      tdmap(cm, cc, cp, cf, N2M, cf);
      double fr[N2M + 2];
      for (int j = 1; j <= N2M; j++) fr[j] = cf[j];
      for (int j = 1; j <= N2M; j++) {
        cm[j] = o->pmj[j];
        cc[j] = o->pcj[j] - o->ak3[k] - o->ak1[i];
        cp[j] = o->ppj[j];
        cf[j] = cimag(FF(i, k, j));
        if ((i == 1 && k == 1) || (i == 1 && k == N3MH) || (i == N1MH && k == 1) ||
            (i == N1MH && k == N3MH)) {
          cm[j] = 0.0;
          cc[j] = 1.0;
          cp[j] = 0.0;
          cf[j] = 0.0;
        }
      }
      tdmap(cm, cc, cp, cf, N2M, cf);
      for (int j = 1; j <= N2M; j++) FF(i, k, j) = fr[j] + I * cf[j];
    }
  /* inverse (:9410-9435) */
#pragma omp parallel for
  for (int j = 1; j <= N2M; j++) {
    for (int k = 1; k <= N3MH; k++)
      for (int i = 1; i <= N1M; i++) CO(i, k, j) = FF(i, k, j);
    for (int k = N3MH + 1; k <= N3M; k++) {
      int kk = N3M - k + 2;
      for (int i = 1; i <= N1M; i++) {
        int ii = N1M - i + 2;
        if (i == 1) ii = 1;
        CO(i, k, j) = conj(CO(ii, kk, j));
      }
    }
    cplx *pl = &CO(1, 1, j);
    for (int k = 0; k < N3M; k++) fft1d(N1M, pl + (size_t)k * N1M, 1, +1, tw1b);
    for (int i = 0; i < N1M; i++) fft1d(N3M, pl + i, N1M, +1, tw3b);
    for (int k = 1; k <= N3M; k++)
      for (int i = 1; i <= N1M; i++) DP[PX(i, j, k)] = creal(CO(i, k, j)) / (double)(N1M * N3M);
  }
  free(coef);
  free(F);
  free(tw1f);
  free(tw3f);
  free(tw1b);
  free(tw3b);
#undef CO
#undef FF
}

/* ----------------------------------------------------------------- UPCALC */
/* compact3d.f:1052-1238 */
void c3o_upcalc(c3o *o, double *U, double *P, const double *UH,
                const double *DP, double *presg, double *presg3) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m, N2 = o->n2;
  const double DX1 = o->dx1, DX3 = o->dx3, DT = o->dt, VOL = o->vol;
  const double *DY = o->dy, *H = o->h;
  /* the sums run in the reference's serial loop order (k, j, i) */
  double flow1 = 0.0, dmpresg = 0.0;
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++)
      for (int i = 1; i <= N1M; i++) {
        int im = o->ima[i];
        flow1 = flow1 + U[IX(i, j, k)] * DY[j] / DX3 / DX1;
        dmpresg = dmpresg + UH[IX(i, j, k)] * DY[j] / DX3 / DX1 -
                  DT * (DP[PX(i, j, k)] - DP[PX(im, j, k)]) * DY[j] / DX3;
      }
  dmpresg = (dmpresg - flow1) / VOL / DT;
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++) {
      double d[N1M + 1], x[N1M + 1];
      for (int i = 1; i <= N1M; i++)
        d[i] = 24. * DX1 * (-DP[PX(WRAPM(i, N1M), j, k)] + DP[PX(i, j, k)]);
      cyc_solve(fxb(o, 22.0), d, x);
      for (int i = 1; i <= N1M; i++) U[IX(i, j, k)] = UH[IX(i, j, k)] - DT * x[i] - DT * dmpresg;
    }
  for (int k = 1; k <= N3M; k++)
    for (int i = 1; i <= N1M; i++) {
      U[IX(i, 0, k)] = o->ubc[UBCX(i, k, 1, 1)];
      U[IX(i, N2, k)] = o->ubc[UBCX(i, k, 1, 2)];
    }
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 2; j <= N2M; j++)
      for (int i = 1; i <= N1M; i++)
        U[NU + IX(i, j, k)] = UH[NU + IX(i, j, k)] - DT * (DP[PX(i, j, k)] - DP[PX(i, j - 1, k)]) / H[j];
  for (int k = 1; k <= N3M; k++)
    for (int i = 1; i <= N1M; i++) {
      U[NU + IX(i, 1, k)] = o->ubc[UBCX(i, k, 2, 1)];
      U[NU + IX(i, N2, k)] = o->ubc[UBCX(i, k, 2, 2)];
    }
  double flow3 = 0.0, dmpresg3 = 0.0;
  for (int k = 1; k <= N3M; k++) {
    int km = o->kma[k];
    for (int j = 1; j <= N2M; j++)
      for (int i = 1; i <= N1M; i++) {
        flow3 = flow3 + U[2 * NU + IX(i, j, k)] * DY[j] / DX1 / DX3;
        dmpresg3 = dmpresg3 + UH[2 * NU + IX(i, j, k)] * DY[j] / DX3 / DX1 -
                   DT * (DP[PX(i, j, k)] - DP[PX(i, j, km)]) * DY[j] / DX1;
      }
  }
  dmpresg3 = (dmpresg3 - flow3) / VOL / DT;
#pragma omp parallel for
  for (int i = 1; i <= N1M; i++)
    for (int j = 1; j <= N2M; j++) {
      double d[N3M + 1], x[N3M + 1];
      for (int k = 1; k <= N3M; k++)
        d[k] = 24. * DX3 * (-DP[PX(i, j, WRAPM(k, N3M))] + DP[PX(i, j, k)]);
      cyc_solve(fzb(o, 22.0), d, x);
      for (int k = 1; k <= N3M; k++)
        U[2 * NU + IX(i, j, k)] = UH[2 * NU + IX(i, j, k)] - DT * x[k] - DT * dmpresg3;
    }
  for (int k = 1; k <= N3M; k++)
    for (int i = 1; i <= N1M; i++) {
      U[2 * NU + IX(i, 0, k)] = o->ubc[UBCX(i, k, 3, 1)];
      U[2 * NU + IX(i, N2, k)] = o->ubc[UBCX(i, k, 3, 2)];
    }
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++)
      for (int i = 1; i <= N1M; i++) P[PX(i, j, k)] = P[PX(i, j, k)] + DP[PX(i, j, k)];
  *presg = *presg + dmpresg;
  *presg3 = *presg3 + dmpresg3;
}

/* --------------------------------------------------------------- DIVCHECK */
/* compact3d.f:9049-9131; DIV (1..N1M,1..N2M,1..N3M) optional, P-like layout
 * with extents (N1M,N2M,N3M) */
double c3o_divcheck(c3o *o, const double *U, double *DIVo) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m;
  const double DX1 = o->dx1, DX3 = o->dx3;
  const double *DY = o->dy;
  double *DIV = dalloc((size_t)N1M * N2M * N3M);
#define DV(i, j, k) DIV[(((size_t)(k) - 1) * N2M + ((j) - 1)) * N1M + ((i) - 1)]
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++)
      for (int i = 1; i <= N1M; i++)
        DV(i, j, k) = (U[NU + IX(i, j + 1, k)] - U[NU + IX(i, j, k)]) / DY[j];
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++) {
      double d[N1M + 1], x[N1M + 1];
      for (int i = 1; i <= N1M; i++) d[i] = 24. * DX1 * (U[IX(WRAPP(i, N1M), j, k)] - U[IX(i, j, k)]);
      cyc_solve(fxb(o, 22.0), d, x);
      for (int i = 1; i <= N1M; i++) DV(i, j, k) = DV(i, j, k) + x[i];
    }
  double mx[N1M + 1];
#pragma omp parallel for
  for (int i = 1; i <= N1M; i++) {
    double m = 0.0;
    for (int j = 1; j <= N2M; j++) {
      double d[N3M + 1], x[N3M + 1];
      for (int k = 1; k <= N3M; k++)
        d[k] = 24. * DX3 * (U[2 * NU + IX(i, j, WRAPP(k, N3M))] - U[2 * NU + IX(i, j, k)]);
      cyc_solve(fzb(o, 22.0), d, x);
      for (int k = 1; k <= N3M; k++) {
        DV(i, j, k) = DV(i, j, k) + x[k];
        m = fmax(fabs(DV(i, j, k)), m);
      }
    }
    mx[i] = m;
  }
  double divmax = 0.0;
  for (int i = 1; i <= N1M; i++) divmax = fmax(mx[i], divmax);
  if (DIVo) memcpy(DIVo, DIV, sizeof(double) * (size_t)N1M * N2M * N3M);
  free(DIV);
#undef DV
  return divmax;
}

/* -------------------------------------------------------------------- CFL */
/* compact3d.f:9138-9169 */
double c3o_cfl(c3o *o, const double *U) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m;
  const double DX1 = o->dx1, DX3 = o->dx3;
  const double *DY = o->dy;
  double mk[N3M + 1];
#pragma omp parallel for
  for (int k = 1; k <= N3M; k++) {
    int kp = o->kpa[k];
    double m = 0.0;
    for (int j = 1; j <= N2M; j++)
      for (int i = 1; i <= N1M; i++) {
        int ip = o->ipa[i];
        double cfll = fabs(U[IX(i, j, k)] + U[IX(ip, j, k)]) * 0.5 * DX1 +
                      fabs(U[NU + IX(i, j, k)] + U[NU + IX(i, j + 1, k)]) * 0.5 / DY[j] +
                      fabs(U[2 * NU + IX(i, j, k)] + U[2 * NU + IX(i, j, kp)]) * 0.5 * DX3;
        m = fmax(m, cfll);
      }
    mk[k] = m;
  }
  double cflm = 0.0;
  for (int k = 1; k <= N3M; k++) cflm = fmax(cflm, mk[k]);
  return cflm;
}

/* ------------------------------------------------------------------ CHKMF */
/* compact3d.f:9463-9503, including its use of U(1,J,K,1) (:9484) and
 * U(I,J,1,3) (:9495) */
void c3o_chkmf(c3o *o, const double *U, double *flow1, double *flow3) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m;
  const double DX1 = o->dx1, DX3 = o->dx3;
  const double *DY = o->dy;
  double f1 = 0.0, f3 = 0.0;
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++)
      for (int i = 1; i <= N1M; i++) f1 = f1 + U[IX(1, j, k)] * DY[j] / DX3 / DX1;
  f1 = f1 / o->alx / o->alz;
  for (int k = 1; k <= N3M; k++)
    for (int j = 1; j <= N2M; j++)
      for (int i = 1; i <= N1M; i++) f3 = f3 + U[2 * NU + IX(i, j, 1)] * DY[j] / DX1 / DX3;
  f3 = f3 / o->alx / o->alz;
  *flow1 = f1;
  *flow3 = f3;
}

/* ----------------------------------------------------------------- WALLSS */
/* compact3d.f:11077-11143 */
void c3o_wallss(c3o *o, const double *U) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m, N2 = o->n2;
  const double *DY = o->dy;
  const double ALY = o->aly;
  double nxz = 1.0 / (double)(N1M * N3M);
  double w1 = 0.0, w2 = 0.0, w3 = 0.0;
  for (int k = 1; k <= N3M; k++) {
    int kp = o->kpa[k];
    for (int i = 1; i <= N1M; i++) {
      int ip = o->ipa[i];
      const double *Vv = U + NU, *Ww = U + 2 * NU;
      double uhc = 0.5 * (U[IX(i, 1, k)] + U[IX(ip, 1, k)]);
      double uhw = 0.5 * (U[IX(i, 0, k)] + U[IX(ip, 0, k)]);
      double vhc = Vv[IX(i, 2, k)], vhw = Vv[IX(i, 1, k)];
      double whc = 0.5 * (Ww[IX(i, 1, k)] + Ww[IX(i, 1, kp)]);
      double whw = 0.5 * (Ww[IX(i, 0, k)] + Ww[IX(i, 0, kp)]);
      double wsb1 = (uhc - uhw) / (DY[1] * 0.5);
      double wsb2 = (vhc - vhw) / DY[1];
      double wsb3 = (whc - whw) / (DY[1] * 0.5);
      uhc = 0.5 * (U[IX(i, N2M, k)] + U[IX(ip, N2M, k)]);
      uhw = 0.5 * (U[IX(i, N2, k)] + U[IX(ip, N2, k)]);
      double wst1 = (uhc - uhw) / (DY[N2M] * 0.5);
      w1 = w1 + wsb1 + wst1 * (ALY - 1.0);
      w2 = w2 + wsb2 + wst1 * (ALY - 1.0);
      w3 = w3 + wsb3 + wst1 * (ALY - 1.0);
    }
  }
  o->wsm[0] = w1 * nxz / o->re / ALY;
  o->wsm[1] = w2 * nxz / o->re / ALY;
  o->wsm[2] = w3 * nxz / o->re / ALY;
}

/* ----------------------------------------------------------------- STRAIN */
/* compact3d.f:5558-5689; SR(1:N1,1:N3,6) column-major */
void c3o_strain(c3o *o, const double *U, double *SR, int j) {
  const int N1M = o->n1m, N3M = o->n3m, N2 = o->n2, N2M = o->n2m;
  const double DX1 = o->dx1, DX3 = o->dx3, WM = o->wmodel;
  const double *DY = o->dy, *H = o->h;
  const double *V = U + NU, *W = U + 2 * NU;
  const int jp = j + 1, jm = j - 1;
  const int fjum = (int)o->fjmu[j], fjup = (int)o->fjpa[j];
#define SRX(i, k, n) SR[((size_t)(n) - 1) * o->n1 * o->n3 + ((size_t)(k) - 1) * o->n1 + ((i) - 1)]
  for (int k = 1; k <= N3M; k++) {
    int kp = o->kpa[k], km = o->kma[k];
    for (int i = 1; i <= N1M; i++) {
      int ip = o->ipa[i], im = o->ima[i];
      double u2 = U[IX(ip, j, k)], u1 = U[IX(i, j, k)];
      SRX(i, k, 1) = DX1 * (u2 - u1);
      double ujp = 0.5 * (U[IX(i, jp, k)] + U[IX(ip, jp, k)]);
      double ujc = 0.5 * (U[IX(i, j, k)] + U[IX(ip, j, k)]);
      double ujm = 0.5 * (U[IX(i, jm, k)] + U[IX(ip, jm, k)]);
      if (WM == 0.) {
        u2 = fjup * 1.0 / H[jp] * (DY[j] / 2.0 * ujp + DY[jp] / 2.0 * ujc) +
             (1. - fjup) * 0.5 * (U[IX(i, N2, k)] + U[IX(ip, N2, k)]);
        u1 = fjum * 1.0 / H[j] * (DY[j] / 2.0 * ujm + DY[jm] / 2.0 * ujc) +
             (1. - fjum) * 0.5 * (U[IX(i, 0, k)] + U[IX(ip, 0, k)]);
      } else {
        u2 = fjup * 1.0 / H[jp] * (DY[j] / 2.0 * ujp + DY[jp] / 2.0 * ujc) + (1. - fjup) * ujc;
        u1 = fjum * 1.0 / H[j] * (DY[j] / 2.0 * ujm + DY[jm] / 2.0 * ujc) + (1. - fjum) * ujc;
      }
      double vip = 0.5 * (V[IX(ip, j, k)] + V[IX(ip, jp, k)]);
      double vic = 0.5 * (V[IX(i, j, k)] + V[IX(i, jp, k)]);
      double vim = 0.5 * (V[IX(im, j, k)] + V[IX(im, jp, k)]);
      double v2 = 0.5 * (vip + vic), v1 = 0.5 * (vim + vic);
      if (WM == 0.) {
        SRX(i, k, 2) = 0.5 * (DX1 * (v2 - v1) + 1.0 / DY[j] * (u2 - u1));
      } else {
        if (j == 1 || j == N2M)
          SRX(i, k, 2) = 0.5 * (DX1 * (v2 - v1) + 2.0 / DY[j] * (u2 - u1));
        else
          SRX(i, k, 2) = 0.5 * (DX1 * (v2 - v1) + 1.0 / DY[j] * (u2 - u1));
      }
      double ukp = 0.5 * (U[IX(i, j, kp)] + U[IX(ip, j, kp)]);
      double ukc = 0.5 * (U[IX(i, j, k)] + U[IX(ip, j, k)]);
      double ukm = 0.5 * (U[IX(i, j, km)] + U[IX(ip, j, km)]);
      u2 = 0.5 * (ukp + ukc);
      u1 = 0.5 * (ukm + ukc);
      double wip = 0.5 * (W[IX(ip, j, k)] + W[IX(ip, j, kp)]);
      double wic = 0.5 * (W[IX(i, j, k)] + W[IX(i, j, kp)]);
      double wim = 0.5 * (W[IX(im, j, k)] + W[IX(im, j, kp)]);
      double w2 = 0.5 * (wip + wic), w1 = 0.5 * (wim + wic);
      SRX(i, k, 3) = 0.5 * (DX1 * (w2 - w1) + DX3 * (u2 - u1));
      v2 = V[IX(i, jp, k)];
      v1 = V[IX(i, j, k)];
      SRX(i, k, 4) = 1.0 / DY[j] * (v2 - v1);
      double wjp = 0.5 * (W[IX(i, jp, k)] + W[IX(i, jp, kp)]);
      double wjc = 0.5 * (W[IX(i, j, k)] + W[IX(i, j, kp)]);
      double wjm = 0.5 * (W[IX(i, jm, k)] + W[IX(i, jm, kp)]);
      if (WM == 0.) {
        w2 = fjup * 1.0 / H[jp] * (DY[j] / 2.0 * wjp + DY[jp] / 2.0 * wjc) +
             (1. - fjup) * 0.5 * (W[IX(i, N2, k)] + W[IX(i, N2, kp)]);
        w1 = fjum * 1.0 / H[j] * (DY[j] / 2.0 * wjm + DY[jm] / 2.0 * wjc) +
             (1. - fjum) * 0.5 * (W[IX(i, 0, k)] + W[IX(i, 0, kp)]);
      } else {
        w2 = fjup * 1.0 / H[jp] * (DY[j] / 2.0 * wjp + DY[jp] / 2.0 * wjc) + (1. - fjup) * wjc;
        w1 = fjum * 1.0 / H[j] * (DY[j] / 2.0 * wjm + DY[jm] / 2.0 * wjc) + (1. - fjum) * wjc;
      }
      double vkp = 0.5 * (V[IX(i, j, kp)] + V[IX(i, jp, kp)]);
      double vkc = 0.5 * (V[IX(i, j, k)] + V[IX(i, jp, k)]);
      double vkm = 0.5 * (V[IX(i, j, km)] + V[IX(i, jp, km)]);
      v2 = 0.5 * (vkp + vkc);
      v1 = 0.5 * (vkm + vkc);
      if (WM == 0.) {
        SRX(i, k, 5) = 0.5 * (1.0 / DY[j] * (w2 - w1) + DX3 * (v2 - v1));
      } else {
        if (j == 1 || j == N2M)
          SRX(i, k, 5) = 0.5 * (2.0 / DY[j] * (w2 - w1) + DX3 * (v2 - v1));
        else
          SRX(i, k, 5) = 0.5 * (1.0 / DY[j] * (w2 - w1) + DX3 * (v2 - v1));
      }
      w2 = W[IX(i, j, kp)];
      w1 = W[IX(i, j, k)];
      SRX(i, k, 6) = DX3 * (w2 - w1);
    }
  }
#undef SRX
}

/* ------------------------------------------------------------------ SGSCS */
/* compact3d.f:1413-1488 */
void c3o_sgscs(c3o *o, const double *U, double *SGSVIS) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m;
  const double RE = o->re, DX1 = o->dx1, DX3 = o->dx3;
  const double *DY = o->dy, *Y = o->y;
  const double reninv = 1. / (double)RE;
  const double utau = sqrt(sqrt(o->wsm[0] * o->wsm[0] + o->wsm[2] * o->wsm[2]));
#pragma omp parallel for
  for (int j = 1; j <= N2M; j++) {
    double *SR = dalloc((size_t)o->n1 * o->n3 * 6);
#define SRX(i, k, n) SR[((size_t)(n) - 1) * o->n1 * o->n3 + ((size_t)(k) - 1) * o->n1 + ((i) - 1)]
    double dxyz = 1.0 / DX1 * DY[j] / DX3;
    c3o_strain(o, U, SR, j);
    for (int k = 1; k <= N3M; k++)
      for (int i = 1; i <= N1M; i++) {
        double s1 = SRX(i, k, 1), s2 = SRX(i, k, 2), s3 = SRX(i, k, 3);
        double s4 = SRX(i, k, 4), s5 = SRX(i, k, 5), s6 = SRX(i, k, 6);
        double sumsr = 2. * (s2 * s2) + s1 * s1 + 2. * (s3 * s3) + s4 * s4 + 2. * (s5 * s5) + s6 * s6;
        SGSVIS[IX(i, j, k)] = sqrt(2. * sumsr);
      }
    double cs = 0.1, a_plus = 26.0;
    double y_plus = (Y[j] + DY[j] * 0.5) * utau * RE;
    cs = cs * (1 - exp(-y_plus / a_plus));
    double smagc = cs * cs;
    double p23 = pow(dxyz, 2. / 3.);
    for (int k = 1; k <= N3M; k++)
      for (int i = 1; i <= N1M; i++) {
        double turbv = smagc * p23 * SGSVIS[IX(i, j, k)];
        SGSVIS[IX(i, j, k)] = fmax(turbv, -reninv);
      }
    free(SR);
#undef SRX
  }
}

/* ----------------------------------------------------------------- SGSDY */
/* compact3d.f:1250-1400: dynamic Smagorinsky model.  Per J plane: STRAIN,
 * |S| (kept in SGSVIS), COM = DELTA^2 |S| S_ij; test filter (FILTER) of COM
 * and S_ij; M_ij = 4 DELTA^2 |S~| S~_ij - (DELTA^2 |S| S_ij)~; cell-centre
 * velocities UC and their products, both filtered; L_ij = (UC_i UC_j)~ -
 * UC~_i UC~_j; C(J) = -0.5 <L:M>/<M:M> over the plane (serial sums in the
 * reference's K, I order); SGSVIS = max(C DELTA^2 |S|, -1/RE). */
void c3o_sgsdy(c3o *o, const double *U, double *SGSVIS, double *smagc_out) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m;
  const double RE = o->re, DX1 = o->dx1, DX3 = o->dx3;
  const double *DY = o->dy;
  const double nxz = 1.0 / (double)(N1M * N3M);
  const double reninv = 1. / (double)RE;
  const double *V = U + NU, *W = U + 2 * NU;
#pragma omp parallel for
  for (int j = 1; j <= N2M; j++) {
    const size_t pl = (size_t)o->n1 * o->n3;
    double *SR = dalloc(pl * 6), *COM = dalloc(pl * 6), *UC = dalloc(pl * 3);
#define DYX(A, i, k, n) A[((size_t)(n) - 1) * pl + ((size_t)(k) - 1) * o->n1 + ((i) - 1)]
    const double dxyz = 1.0 / DX1 * DY[j] / DX3;
    c3o_strain(o, U, SR, j);
    for (int k = 1; k <= N3M; k++)
      for (int i = 1; i <= N1M; i++) {
        const double delta2 = pow(dxyz, 2. / 3.);
        const double sumsr = 2. * (DYX(SR, i, k, 2) * DYX(SR, i, k, 2)) + DYX(SR, i, k, 1) * DYX(SR, i, k, 1) +
                             2. * (DYX(SR, i, k, 3) * DYX(SR, i, k, 3)) + DYX(SR, i, k, 4) * DYX(SR, i, k, 4) +
                             2. * (DYX(SR, i, k, 5) * DYX(SR, i, k, 5)) + DYX(SR, i, k, 6) * DYX(SR, i, k, 6);
        const double abssr = sqrt(2. * sumsr);
        SGSVIS[IX(i, j, k)] = abssr;
        const double delabs = delta2 * abssr;
        for (int m = 1; m <= 6; m++) DYX(COM, i, k, m) = delabs * DYX(SR, i, k, m);
      }
    c3o_filter(o, COM, 6);
    c3o_filter(o, SR, 6);
    for (int k = 1; k <= N3M; k++)
      for (int i = 1; i <= N1M; i++) {
        const double sumfsr = 2. * (DYX(SR, i, k, 2) * DYX(SR, i, k, 2)) + DYX(SR, i, k, 1) * DYX(SR, i, k, 1) +
                              2. * (DYX(SR, i, k, 3) * DYX(SR, i, k, 3)) + DYX(SR, i, k, 4) * DYX(SR, i, k, 4) +
                              2. * (DYX(SR, i, k, 5) * DYX(SR, i, k, 5)) + DYX(SR, i, k, 6) * DYX(SR, i, k, 6);
        const double absfsr = sqrt(2. * sumfsr);
        const double fdxyz = 4.0 * pow(dxyz, 2. / 3.);
        const double fdabs = fdxyz * absfsr;
        for (int m = 1; m <= 6; m++) DYX(SR, i, k, m) = fdabs * DYX(SR, i, k, m) - DYX(COM, i, k, m);
      }
    const int jp = j + 1;
    for (int k = 1; k <= N3M; k++) {
      const int kp = o->kpa[k];
      for (int i = 1; i <= N1M; i++) {
        const int ip = o->ipa[i];
        const double u1 = (U[IX(i, j, k)] + U[IX(ip, j, k)]) * 0.5;
        const double u2 = (V[IX(i, j, k)] + V[IX(i, jp, k)]) * 0.5;
        const double u3 = (W[IX(i, j, k)] + W[IX(i, j, kp)]) * 0.5;
        DYX(UC, i, k, 1) = u1;
        DYX(UC, i, k, 2) = u2;
        DYX(UC, i, k, 3) = u3;
        DYX(COM, i, k, 1) = u1 * u1;
        DYX(COM, i, k, 2) = u1 * u2;
        DYX(COM, i, k, 3) = u1 * u3;
        DYX(COM, i, k, 4) = u2 * u2;
        DYX(COM, i, k, 5) = u2 * u3;
        DYX(COM, i, k, 6) = u3 * u3;
      }
    }
    c3o_filter(o, UC, 3);
    c3o_filter(o, COM, 6);
    double sumnum = 0., sumden = 0.;
    for (int k = 1; k <= N3M; k++)
      for (int i = 1; i <= N1M; i++) {
        const double c1 = DYX(UC, i, k, 1), c2 = DYX(UC, i, k, 2), c3 = DYX(UC, i, k, 3);
        const double xl11 = DYX(COM, i, k, 1) - c1 * c1, xl12 = DYX(COM, i, k, 2) - c1 * c2;
        const double xl13 = DYX(COM, i, k, 3) - c1 * c3, xl22 = DYX(COM, i, k, 4) - c2 * c2;
        const double xl23 = DYX(COM, i, k, 5) - c2 * c3, xl33 = DYX(COM, i, k, 6) - c3 * c3;
        const double s1 = DYX(SR, i, k, 1), s2 = DYX(SR, i, k, 2), s3 = DYX(SR, i, k, 3);
        const double s4 = DYX(SR, i, k, 4), s5 = DYX(SR, i, k, 5), s6 = DYX(SR, i, k, 6);
        const double sqmm = 2. * (s2 * s2) + s1 * s1 + 2. * (s3 * s3) + s4 * s4 + 2. * (s5 * s5) + s6 * s6;
        const double sqlm = 2. * xl12 * s2 + xl11 * s1 + 2. * xl13 * s3 + xl22 * s4 + 2. * xl23 * s5 + xl33 * s6;
        sumden = sumden + sqmm;
        sumnum = sumnum + sqlm;
      }
    sumnum = sumnum * nxz;
    sumden = sumden * nxz;
    const double smagc = -0.5 * sumnum / sumden;
    if (smagc_out) smagc_out[j] = smagc;
    for (int k = 1; k <= N3M; k++)
      for (int i = 1; i <= N1M; i++) {
        const double turbv = smagc * pow(dxyz, 2. / 3.) * SGSVIS[IX(i, j, k)];
        SGSVIS[IX(i, j, k)] = fmax(turbv, -reninv);
      }
    free(SR);
    free(COM);
    free(UC);
#undef DYX
  }
}

/* ----------------------------------------------------------------- SGSSD */
/* The scale-dependent dynamic model, compact3d.f:1502-1948, with
 * COEFFICIENT (:2079-2097), DQRRT (:11548-11570) and CHHQR (:11571-11677).
 *
 * CHHQR finds the eigenvalues of an upper Hessenberg matrix by the
 * double-shift QR iteration (Francis; the "hqr" procedure of Wilkinson and
 * Reinsch), deflating where a subdiagonal entry is at most EPS times its two
 * diagonal neighbours.  DQRRT calls it with EPS = 1e-5 on the companion
 * matrix of x^5 + PA(1) x^4 + ... + PA(5), so the roots carry the error of
 * that loose deflation and depend on the iteration path: the restatement
 * keeps the reference's steps (no exceptional shifts, its reflector
 * formulas, at most 60 iterations per root) so that it reproduces them.
 * A (n x n) column-major, wr/wi the real and imaginary parts; returns 0 when
 * an eigenvalue took 60 iterations ('FAIL'), leaving the rest unset. */
static int chhqr(double *A, int n, double *wr, double *wi, double eps) {
#define HA(i, j) A[((j) - 1) * n + ((i) - 1)]
  int m = n, it = 0;
  for (;;) {
    if (m == 0) return 1;
    int l = m;
    while (l > 1 && fabs(HA(l, l - 1)) > eps * (fabs(HA(l - 1, l - 1)) + fabs(HA(l, l)))) l--;
    if (l == m) {
      wr[m - 1] = HA(m, m);
      wi[m - 1] = 0.0;
      m -= 1;
      it = 0;
      continue;
    }
    if (l == m - 1) {
      const double b = -(HA(m, m) + HA(m - 1, m - 1));
      const double c = HA(m, m) * HA(m - 1, m - 1) - HA(m, m - 1) * HA(m - 1, m);
      const double w = b * b - 4 * c;
      const double y = sqrt(fabs(w));
      if (w > 0.0) {
        const double xy = (b < 0.0) ? -1.0 : 1.0;
        wr[m - 1] = (-b - xy * y) / 2.0;
        wr[m - 2] = c / wr[m - 1];
        wi[m - 1] = 0.0;
        wi[m - 2] = 0.0;
      } else {
        wr[m - 1] = -b / 2.0;
        wr[m - 2] = wr[m - 1];
        wi[m - 1] = y / 2.0;
        wi[m - 2] = -wi[m - 1];
      }
      m -= 2;
      it = 0;
      continue;
    }
    if (it >= 60) return 0;
    it += 1;
    for (int j = l + 2; j <= m; j++) HA(j, j - 2) = 0.0;
    for (int j = l + 3; j <= m; j++) HA(j, j - 3) = 0.0;
    for (int k = l; k <= m - 1; k++) {
      double p, q, r, x, y;
      if (k != l) {
        p = HA(k, k - 1);
        q = HA(k + 1, k - 1);
        r = 0.0;
        if (k != m - 1) r = HA(k + 2, k - 1);
      } else {
        x = HA(m, m) + HA(m - 1, m - 1);
        y = HA(m - 1, m - 1) * HA(m, m) - HA(m - 1, m) * HA(m, m - 1);
        p = HA(l, l) * (HA(l, l) - x) + HA(l, l + 1) * HA(l + 1, l) + y;
        q = HA(l + 1, l) * (HA(l, l) + HA(l + 1, l + 1) - x);
        r = HA(l + 1, l) * HA(l + 2, l + 1);
      }
      if (fabs(p) + fabs(q) + fabs(r) != 0.0) {
        const double xy = (p < 0.0) ? -1.0 : 1.0;
        const double s = xy * sqrt(p * p + q * q + r * r);
        if (k != l) HA(k, k - 1) = -s;
        const double e = -q / s, f = -r / s;
        x = -p / s;
        y = -x - f * r / (p + s);
        const double g = e * r / (p + s);
        const double z = -x - e * q / (p + s);
        for (int j = k; j <= m; j++) {
          p = x * HA(k, j) + e * HA(k + 1, j);
          q = e * HA(k, j) + y * HA(k + 1, j);
          r = f * HA(k, j) + g * HA(k + 1, j);
          if (k != m - 1) {
            p = p + f * HA(k + 2, j);
            q = q + g * HA(k + 2, j);
            r = r + z * HA(k + 2, j);
            HA(k + 2, j) = r;
          }
          HA(k + 1, j) = q;
          HA(k, j) = p;
        }
        int jj = k + 3;
        if (jj >= m) jj = m;
        for (int i = l; i <= jj; i++) {
          p = x * HA(i, k) + e * HA(i, k + 1);
          q = e * HA(i, k) + y * HA(i, k + 1);
          r = f * HA(i, k) + g * HA(i, k + 1);
          if (k != m - 1) {
            p = p + f * HA(i, k + 2);
            q = q + g * HA(i, k + 2);
            r = r + z * HA(i, k + 2);
            HA(i, k + 2) = r;
          }
          HA(i, k + 1) = q;
          HA(i, k) = p;
        }
      }
    }
  }
#undef HA
}

/* DQRRT (:11548-11570): BETA = the largest real part of a root of
 * x^5 + PA(1) x^4 + ... + PA(5) in (0, 2), else 0.  Real parts of complex
 * roots count too.  Roots CHHQR leaves unset after a failure are taken as
 * 0 (not selected); in the reference they are uninitialised. */
double c3o_dqrrt(const double *pa, int n) {
  double B[25], xr[5] = {0, 0, 0, 0, 0}, xi[5];
  if (n > 5) return 0.0;
  for (int j = 1; j <= n; j++) B[(j - 1) * n] = -pa[j - 1];
  for (int i = 2; i <= n; i++)
    for (int j = 1; j <= n; j++) B[(j - 1) * n + (i - 1)] = 0.0;
  for (int i = 1; i <= n - 1; i++) B[(i - 1) * n + i] = 1.0;
  chhqr(B, n, xr, xi, 1.0E-5);
  double beta = 0.;
  for (int j = 0; j < n; j++)
    if (xr[j] > beta && xr[j] < 2.0) beta = xr[j];
  return beta;
}

/* one test-filter level of SGSSD's coefficient sums (:1544-1667 with FILTER,
 * :1672-1800 with FILTER4): out = A, B, C, D, E scaled as the reference */
static void sgssd_sums(c3o *o, const double *U, int j, int four, double *SR, double *COM, double *COM1,
                       double *SR1, double *UC, double out[5]) {
  const int N1M = o->n1m, N3M = o->n3m;
  const double DX1 = o->dx1, DX3 = o->dx3;
  const double nxz = 1.0 / (double)(N1M * N3M);
  const double *V = U + NU, *W = U + 2 * NU;
  const size_t pl = (size_t)o->n1 * o->n3;
#define DYX(A, i, k, n) A[((size_t)(n) - 1) * pl + ((size_t)(k) - 1) * o->n1 + ((i) - 1)]
  const double dxyz = 1.0 / DX1 * o->dy[j] / DX3;
  c3o_strain(o, U, SR, j);
  for (int k = 1; k <= N3M; k++)
    for (int i = 1; i <= N1M; i++) {
      const double delta2 = pow(dxyz, 2. / 3.);
      const double sumsr = 2. * (DYX(SR, i, k, 2) * DYX(SR, i, k, 2)) + DYX(SR, i, k, 1) * DYX(SR, i, k, 1) +
                           2. * (DYX(SR, i, k, 3) * DYX(SR, i, k, 3)) + DYX(SR, i, k, 4) * DYX(SR, i, k, 4) +
                           2. * (DYX(SR, i, k, 5) * DYX(SR, i, k, 5)) + DYX(SR, i, k, 6) * DYX(SR, i, k, 6);
      const double abssr = sqrt(2. * sumsr);
      const double delabs = delta2 * abssr;
      for (int m = 1; m <= 6; m++) DYX(COM1, i, k, m) = delabs * DYX(SR, i, k, m);
    }
  if (four) {
    c3o_filter4(o, COM1, 6);
    c3o_filter4(o, SR, 6);
  } else {
    c3o_filter(o, COM1, 6);
    c3o_filter(o, SR, 6);
  }
  for (int k = 1; k <= N3M; k++)
    for (int i = 1; i <= N1M; i++) {
      const double sumfsr = 2. * (DYX(SR, i, k, 2) * DYX(SR, i, k, 2)) + DYX(SR, i, k, 1) * DYX(SR, i, k, 1) +
                            2. * (DYX(SR, i, k, 3) * DYX(SR, i, k, 3)) + DYX(SR, i, k, 4) * DYX(SR, i, k, 4) +
                            2. * (DYX(SR, i, k, 5) * DYX(SR, i, k, 5)) + DYX(SR, i, k, 6) * DYX(SR, i, k, 6);
      const double absfsr = sqrt(2. * sumfsr);
      const double fdxyz = pow(dxyz, 2. / 3.);
      const double fdabs = fdxyz * absfsr;
      for (int m = 1; m <= 6; m++) DYX(SR1, i, k, m) = fdabs * DYX(SR, i, k, m);
    }
  const int jp = j + 1;
  for (int k = 1; k <= N3M; k++) {
    const int kp = o->kpa[k];
    for (int i = 1; i <= N1M; i++) {
      const int ip = o->ipa[i];
      const double u1 = (U[IX(i, j, k)] + U[IX(ip, j, k)]) * 0.5;
      const double u2 = (V[IX(i, j, k)] + V[IX(i, jp, k)]) * 0.5;
      const double u3 = (W[IX(i, j, k)] + W[IX(i, j, kp)]) * 0.5;
      DYX(UC, i, k, 1) = u1;
      DYX(UC, i, k, 2) = u2;
      DYX(UC, i, k, 3) = u3;
      DYX(COM, i, k, 1) = u1 * u1;
      DYX(COM, i, k, 2) = u1 * u2;
      DYX(COM, i, k, 3) = u1 * u3;
      DYX(COM, i, k, 4) = u2 * u2;
      DYX(COM, i, k, 5) = u2 * u3;
      DYX(COM, i, k, 6) = u3 * u3;
    }
  }
  if (four) {
    c3o_filter4(o, UC, 3);
    c3o_filter4(o, COM, 6);
  } else {
    c3o_filter(o, UC, 3);
    c3o_filter(o, COM, 6);
  }
  double a = 0.0, b = 0.0;
  for (int k = 1; k <= N3M; k++)
    for (int i = 1; i <= N1M; i++) {
      const double c1 = DYX(UC, i, k, 1), c2 = DYX(UC, i, k, 2), c3 = DYX(UC, i, k, 3);
      const double xl11 = DYX(COM, i, k, 1) - c1 * c1, xl12 = DYX(COM, i, k, 2) - c1 * c2;
      const double xl13 = DYX(COM, i, k, 3) - c1 * c3, xl22 = DYX(COM, i, k, 4) - c2 * c2;
      const double xl23 = DYX(COM, i, k, 5) - c2 * c3, xl33 = DYX(COM, i, k, 6) - c3 * c3;
      /* the accumulator is the first operand of the reference's sum */
      a = a + 2. * xl12 * DYX(SR1, i, k, 2) + xl11 * DYX(SR1, i, k, 1) + 2. * xl13 * DYX(SR1, i, k, 3) +
          xl22 * DYX(SR1, i, k, 4) + 2. * xl23 * DYX(SR1, i, k, 5) + xl33 * DYX(SR1, i, k, 6);
      b = b + 2. * xl12 * DYX(COM1, i, k, 2) + xl11 * DYX(COM1, i, k, 1) + 2. * xl13 * DYX(COM1, i, k, 3) +
          xl22 * DYX(COM1, i, k, 4) + 2. * xl23 * DYX(COM1, i, k, 5) + xl33 * DYX(COM1, i, k, 6);
    }
  double c = 0.0, d = 0.0, e = 0.0;
  for (int k = 1; k <= N3M; k++)
    for (int i = 1; i <= N1M; i++) {
#define C1(m) DYX(COM1, i, k, m)
#define S1(m) DYX(SR1, i, k, m)
      c = c + 2. * C1(2) * C1(2) + C1(1) * C1(1) + 2. * C1(3) * C1(3) + C1(4) * C1(4) + 2. * C1(5) * C1(5) +
          C1(6) * C1(6);
      d = d + 2. * S1(2) * S1(2) + S1(1) * S1(1) + 2. * S1(3) * S1(3) + S1(4) * S1(4) + 2. * S1(5) * S1(5) +
          S1(6) * S1(6);
      e = e + 2. * S1(2) * C1(2) + S1(1) * C1(1) + 2. * S1(3) * C1(3) + S1(4) * C1(4) + 2. * S1(5) * C1(5) +
          S1(6) * C1(6);
#undef C1
#undef S1
    }
  if (four) {
    out[0] = -32.0 * a * nxz;
    out[1] = -2.0 * b * nxz;
    out[2] = 4.0 * c * nxz;
    out[3] = 1024.0 * d * nxz;
    out[4] = 128.0 * e * nxz;
  } else {
    out[0] = -8.0 * a * nxz;
    out[1] = -2.0 * b * nxz;
    out[2] = 4.0 * c * nxz;
    out[3] = 64.0 * d * nxz;
    out[4] = 32.0 * e * nxz;
  }
#undef DYX
}

/* COEFFICIENT (:2089-2093): PA(1..5) from the two filter levels' sums */
void c3o_sd_coefficient(const double s1[5], const double s2[5], double pa[5]) {
  const double A1 = s1[0], B1 = s1[1], C1 = s1[2], D1 = s1[3], E1 = s1[4];
  const double A2 = s2[0], B2 = s2[1], C2 = s2[2], D2 = s2[3], E2 = s2[4];
  pa[4] = (B2 * C1 - B1 * C2) / (A1 * D2);
  pa[3] = (A1 * C2 - B2 * E1) / (A1 * D2);
  pa[2] = (B2 * D1 + B1 * E2 - A2 * C1) / (A1 * D2);
  pa[1] = (A2 * E1 - A1 * E2) / (A1 * D2);
  pa[0] = -(A2 * D1 + B1 * D2) / (A1 * D2);
}

/* one plane's sums of both filter levels (s1, s2: A..E), PA and BETA */
double c3o_sgssd_plane(c3o *o, const double *U, int j, double *s1, double *s2, double *pa) {
  const size_t pl = (size_t)o->n1 * o->n3;
  double *SR = dalloc(pl * 6), *COM = dalloc(pl * 6), *COM1 = dalloc(pl * 6), *SR1 = dalloc(pl * 6);
  double *UC = dalloc(pl * 3);
  sgssd_sums(o, U, j, 0, SR, COM, COM1, SR1, UC, s1);
  sgssd_sums(o, U, j, 1, SR, COM, COM1, SR1, UC, s2);
  c3o_sd_coefficient(s1, s2, pa);
  free(SR);
  free(COM);
  free(COM1);
  free(SR1);
  free(UC);
  return c3o_dqrrt(pa, 5);
}

/* SGSSD: per plane the sums of both filter levels, BETA from the quintic,
 * then SGSDY's Germano step with the test-filter factor 4*BETA (:1804-1930).
 * smagc_out[J], beta_out[J] (J = 1..N2M) when not NULL. */
void c3o_sgssd(c3o *o, const double *U, double *SGSVIS, double *smagc_out, double *beta_out) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m;
  const double RE = o->re, DX1 = o->dx1, DX3 = o->dx3;
  const double nxz = 1.0 / (double)(N1M * N3M);
  const double reninv = 1. / (double)RE;
  const double *V = U + NU, *W = U + 2 * NU;
#pragma omp parallel for
  for (int j = 1; j <= N2M; j++) {
    const size_t pl = (size_t)o->n1 * o->n3;
    double *SR = dalloc(pl * 6), *COM = dalloc(pl * 6), *COM1 = dalloc(pl * 6), *SR1 = dalloc(pl * 6);
    double *UC = dalloc(pl * 3);
#define DYX(A, i, k, n) A[((size_t)(n) - 1) * pl + ((size_t)(k) - 1) * o->n1 + ((i) - 1)]
    double s1[5], s2[5], pa[5];
    sgssd_sums(o, U, j, 0, SR, COM, COM1, SR1, UC, s1);
    sgssd_sums(o, U, j, 1, SR, COM, COM1, SR1, UC, s2);
    c3o_sd_coefficient(s1, s2, pa);
    const double beta = c3o_dqrrt(pa, 5);
    if (beta_out) beta_out[j] = beta;
    const double dxyz = 1.0 / DX1 * o->dy[j] / DX3;
    c3o_strain(o, U, SR, j);
    for (int k = 1; k <= N3M; k++)
      for (int i = 1; i <= N1M; i++) {
        const double delta2 = pow(dxyz, 2. / 3.);
        const double sumsr = 2. * (DYX(SR, i, k, 2) * DYX(SR, i, k, 2)) + DYX(SR, i, k, 1) * DYX(SR, i, k, 1) +
                             2. * (DYX(SR, i, k, 3) * DYX(SR, i, k, 3)) + DYX(SR, i, k, 4) * DYX(SR, i, k, 4) +
                             2. * (DYX(SR, i, k, 5) * DYX(SR, i, k, 5)) + DYX(SR, i, k, 6) * DYX(SR, i, k, 6);
        const double abssr = sqrt(2. * sumsr);
        SGSVIS[IX(i, j, k)] = abssr;
        const double delabs = delta2 * abssr;
        for (int m = 1; m <= 6; m++) DYX(COM, i, k, m) = delabs * DYX(SR, i, k, m);
      }
    c3o_filter(o, COM, 6);
    c3o_filter(o, SR, 6);
    for (int k = 1; k <= N3M; k++)
      for (int i = 1; i <= N1M; i++) {
        const double sumfsr = 2. * (DYX(SR, i, k, 2) * DYX(SR, i, k, 2)) + DYX(SR, i, k, 1) * DYX(SR, i, k, 1) +
                              2. * (DYX(SR, i, k, 3) * DYX(SR, i, k, 3)) + DYX(SR, i, k, 4) * DYX(SR, i, k, 4) +
                              2. * (DYX(SR, i, k, 5) * DYX(SR, i, k, 5)) + DYX(SR, i, k, 6) * DYX(SR, i, k, 6);
        const double absfsr = sqrt(2. * sumfsr);
        const double fdxyz = 4.0 * pow(dxyz, 2. / 3.) * beta;
        const double fdabs = fdxyz * absfsr;
        for (int m = 1; m <= 6; m++) DYX(SR, i, k, m) = fdabs * DYX(SR, i, k, m) - DYX(COM, i, k, m);
      }
    const int jp = j + 1;
    for (int k = 1; k <= N3M; k++) {
      const int kp = o->kpa[k];
      for (int i = 1; i <= N1M; i++) {
        const int ip = o->ipa[i];
        const double u1 = (U[IX(i, j, k)] + U[IX(ip, j, k)]) * 0.5;
        const double u2 = (V[IX(i, j, k)] + V[IX(i, jp, k)]) * 0.5;
        const double u3 = (W[IX(i, j, k)] + W[IX(i, j, kp)]) * 0.5;
        DYX(UC, i, k, 1) = u1;
        DYX(UC, i, k, 2) = u2;
        DYX(UC, i, k, 3) = u3;
        DYX(COM, i, k, 1) = u1 * u1;
        DYX(COM, i, k, 2) = u1 * u2;
        DYX(COM, i, k, 3) = u1 * u3;
        DYX(COM, i, k, 4) = u2 * u2;
        DYX(COM, i, k, 5) = u2 * u3;
        DYX(COM, i, k, 6) = u3 * u3;
      }
    }
    c3o_filter(o, UC, 3);
    c3o_filter(o, COM, 6);
    double sumnum = 0., sumden = 0.;
    for (int k = 1; k <= N3M; k++)
      for (int i = 1; i <= N1M; i++) {
        const double c1 = DYX(UC, i, k, 1), c2 = DYX(UC, i, k, 2), c3 = DYX(UC, i, k, 3);
        const double xl11 = DYX(COM, i, k, 1) - c1 * c1, xl12 = DYX(COM, i, k, 2) - c1 * c2;
        const double xl13 = DYX(COM, i, k, 3) - c1 * c3, xl22 = DYX(COM, i, k, 4) - c2 * c2;
        const double xl23 = DYX(COM, i, k, 5) - c2 * c3, xl33 = DYX(COM, i, k, 6) - c3 * c3;
        const double m1 = DYX(SR, i, k, 1), m2 = DYX(SR, i, k, 2), m3 = DYX(SR, i, k, 3);
        const double m4 = DYX(SR, i, k, 4), m5 = DYX(SR, i, k, 5), m6 = DYX(SR, i, k, 6);
        const double sqmm = 2. * (m2 * m2) + m1 * m1 + 2. * (m3 * m3) + m4 * m4 + 2. * (m5 * m5) + m6 * m6;
        const double sqlm = 2. * xl12 * m2 + xl11 * m1 + 2. * xl13 * m3 + xl22 * m4 + 2. * xl23 * m5 + xl33 * m6;
        sumden = sumden + sqmm;
        sumnum = sumnum + sqlm;
      }
    sumnum = sumnum * nxz;
    sumden = sumden * nxz;
    const double smagc = -0.5 * sumnum / sumden;
    if (smagc_out) smagc_out[j] = smagc;
    for (int k = 1; k <= N3M; k++)
      for (int i = 1; i <= N1M; i++) {
        const double turbv = smagc * pow(dxyz, 2. / 3.) * SGSVIS[IX(i, j, k)];
        SGSVIS[IX(i, j, k)] = fmax(turbv, -reninv) * 1.0;
      }
    free(SR);
    free(COM);
    free(COM1);
    free(SR1);
    free(UC);
#undef DYX
  }
}

/* ------------------------------------------------------ WALLMODEL_INSTANT */
/* compact3d.f:2319-2493 */
/* the log-law Newton iteration of WALLMODEL_INSTANT on row j at the wall
 * distance deltap, into layer `layer` of TAUW/TAUB; returns the sum of
 * tau_w over the plane */
static double instant_wall(c3o *o, const double *U, int j, double deltap, int layer) {
  const int N1M = o->n1m, N3M = o->n3m;
  const double RE = o->re;
  const double *W = U + 2 * NU;
  const double KAPA = 0.41, B = 5.2;
  double wsm1 = 0.0;
  for (int k = 1; k <= N3M; k++)
    for (int i = 1; i <= N1M; i++) {
      double u3 = U[IX(i, j, k)], w3 = W[IX(i, j, k)];
      double u_ave = sqrt(u3 * u3 + w3 * w3);
      double ut1 = 1.0, ut2 = 0.000001, f1, f2, f3, lam = 1.0;
      while (fabs(ut2 - ut1) > 1.0E-6) {
        ut1 = ut2;
        f1 = u_ave / ut1 - log(ut1 * deltap * RE) / KAPA - B;
        f2 = -u_ave / (ut1 * ut1) - 1. / KAPA / ut1;
        ut2 = ut1 - lam * f1 / f2;
        f3 = u_ave / ut2 - log(ut2 * deltap * RE) / KAPA - B;
        if (fabs(f3) > fabs(f1)) lam = 0.5 * lam;
      }
      double tauw1 = ut1 * ut1;
      wsm1 = wsm1 + tauw1;
      o->tauw[TAUX(i, k, layer)] = tauw1 * u3 / u_ave;
      o->taub[TAUX(i, k, layer)] = tauw1 * w3 / u_ave;
    }
  return wsm1;
}

void c3o_wallmodel_instant(c3o *o, const double *U) {
  const double *DY = o->dy;
  const double wsm1 = instant_wall(o, U, 3, DY[1] + DY[2] + DY[3] / 2.0, 1);
  o->wsm[0] = wsm1 / (double)(o->n1m * o->n3m);
  o->taow = o->wsm[0];
}

/* WMODEL = 6, an extension with no reference counterpart (the reference's WALLMODEL_INSTANT
 * models the bottom wall only, so a full channel's top wall has no wall
 * stress): WALLMODEL_INSTANT, then the same iteration at the top wall on
 * the third cell from it, J = N2M-2, at the distance DY(N2M) + DY(N2M-1) +
 * DY(N2M-2)/2, into TAUW(I,K,2) and TAUB(I,K,2).  WSM(1) = TAOW stay the
 * bottom wall's mean. */
void c3o_wallmodel_instant2(c3o *o, const double *U) {
  const double *DY = o->dy;
  const int N2M = o->n2m;
  c3o_wallmodel_instant(o, U);
  (void)instant_wall(o, U, N2M - 2, DY[N2M] + DY[N2M - 1] + DY[N2M - 2] / 2.0, 2);
}

/* ------------------------------------------------ WALLMODEL_IWMLES_SMOOTH */
/* compact3d.f:3386-4329 (WMODEL = 2 with Y0 = 0): the integral wall model of
 * Yang et al. (2015) for a smooth wall, at the bottom wall.  Every
 * IWM_SKIP-th step (NTIME multiple of 5) it time-filters the third cell's
 * velocity and the first cell's pressure fluctuation, forms the integrated
 * momentum balance of the wall layer (iwm_calc_lhs_smooth) and solves it per
 * point for the friction velocities (utx, uty) by Newton's method with a
 * finite-difference Jacobian (iwm_calc_wallstress_smooth); every step it
 * imposes the stored stresses.  State: the 43 planes of COMMON /IWMLES/ in
 * their COMMON order (IW_* in c3d_oracle.h) and wm_dt.
 *
 * Powers follow what the reference's compiler makes of them, measured on
 * the reference build: x**2.0 and x**2 are x*x, x**3.0 is x*x*x, x**4.0
 * and x**0.25 call pow.
 *
 * Three points differ from the reference's text, where it is undefined:
 *  - iwm_wallstress_smooth stores the spanwise stress in TAUW(I,K,3) of a
 *    (0:M1,0:M3,2) array (:3549); in COMMON /TAUXYZ/ that element is
 *    TAUB(I,K,1), which is where it is stored here (and read back for
 *    WSM(1), :3470).
 *  - the excessive-linear-correction branch (:4139-4147) takes the
 *    equilibrium stress equilutx/equiluty of the point, which the reference
 *    computes only when the Newton iteration failed (:4065-4099); otherwise
 *    it reads whatever value the local variable kept from an earlier point.
 *    Here the point's equilibrium stress is computed when it was not
 *    (o->iwm_fixed counts those points in the last call).
 *  - the equilibrium iterations (:3616-3627, :4077-4088) have no iteration
 *    limit in the reference; here they stop after IWM_EQ_MAXIT iterations. */
#define IWM_SKIP 5
#define IWM_EQ_MAXIT 10000
#define IWP(p, i, k) o->iwm[((size_t)(p) * N3M + (size_t)(k) - 1) * N1M + (size_t)(i) - 1]

/* the log-law friction velocity for the speed u_ave at the distance deltap
 * (KAPA = 0.4, B = 5; :3616-3627, :4077-4088) */
static double iwm_equil(double u_ave, double deltap, double re) {
  const double KAPA = 0.4, B = 5.0;
  double ut1 = 1.0, ut2 = 0.000001, f1, f2, f3, lam = 1.0;
  int it = 0;
  while (fabs(ut2 - ut1) > 1.0E-6 && it < IWM_EQ_MAXIT) {
    ut1 = ut2;
    f1 = u_ave / ut1 - log(ut1 * deltap * re) / KAPA - B;
    f2 = -u_ave / (ut1 * ut1) - 1. / KAPA / ut1;
    ut2 = ut1 - lam * f1 / f2;
    f3 = u_ave / ut2 - log(ut2 * deltap * re) / KAPA - B;
    if (fabs(f3) > fabs(f1)) lam = 0.5 * lam;
    it++;
  }
  return ut1;
}

/* iwm_init_smooth (:3561-3714) */
static void iwm_init(c3o *o, const double *U) {
  const int N1M = o->n1m, N3M = o->n3m;
  const double *DY = o->dy;
  const double vonk = 0.4, REINV = 1.0 / o->re;
  const double Dzp = DY[1] + DY[2] + DY[3] / 2.0;
  double uinit = 0.0;
  for (int k = 1; k <= N3M; k++)
    for (int i = 1; i <= N1M; i++) uinit = uinit + U[IX(i, 3, k)];
  uinit = uinit / (double)(N1M * N3M);
  const double usinit = iwm_equil(uinit, Dzp, o->re);
  for (size_t t = 0; t < (size_t)IW_NPL * N1M * N3M; t++) o->iwm[t] = 0.0;
  for (int k = 1; k <= N3M; k++)
    for (int i = 1; i <= N1M; i++) {
      IWP(IW_UTX, i, k) = usinit;
      IWP(IW_TAUWX, i, k) = usinit * usinit;
      IWP(IW_FLT, i, k) = uinit;
      IWP(IW_FLTM, i, k) = uinit;
      IWP(IW_INTE + 0, i, k) = uinit * Dzp;
      IWP(IW_INTE + 1, i, k) = uinit * uinit * Dzp;
      IWP(IW_INTEM + 0, i, k) = uinit * Dzp;
      IWP(IW_INTEM + 1, i, k) = uinit * uinit * Dzp;
      IWP(IW_LHS, i, k) = -uinit * Dzp;
      IWP(IW_LHS + 1, i, k) = -uinit * Dzp;
      IWP(IW_DUDZT, i, k) = usinit / vonk / Dzp;
      IWP(IW_FLTUS, i, k) = usinit;
      IWP(IW_TR, i, k) = (o->cflmax * o->alx / N1M / uinit) / (Dzp / vonk / usinit);
      IWP(IW_DZ, i, k) = Dzp;
      IWP(IW_DELI, i, k) = fmin(11.0 * REINV / usinit, Dzp);
      IWP(IW_DELV, i, k) = REINV / usinit;
    }
  o->wm_dt = IWM_SKIP * o->dt;
}

/* iwm_calc_lhs_smooth (:3718-3877) */
static void iwm_lhs(c3o *o, const double *U, const double *P, double presg, double presg3) {
  const int N1M = o->n1m, N3M = o->n3m;
  const double *W = U + 2 * NU;
  const double DX1 = o->dx1, DX3 = o->dx3, wm_dt = o->wm_dt;
  double p_bar = 0.0;
  for (int i = 1; i <= N1M; i++)
    for (int k = 1; k <= N3M; k++) p_bar = p_bar + P[PX(i, 1, k)];
  p_bar = p_bar / (double)N1M / (double)N3M;
  for (int i = 1; i <= N1M; i++)
    for (int k = 1; k <= N3M; k++) {
      const double tr = IWP(IW_TR, i, k);
      IWP(IW_FLTM, i, k) = IWP(IW_FLT, i, k);
      IWP(IW_FLTM + 1, i, k) = IWP(IW_FLT + 1, i, k);
      IWP(IW_FLT, i, k) = IWP(IW_FLT, i, k) * (1.0 - tr) + U[IX(i, 3, k)] * tr;
      IWP(IW_FLT + 1, i, k) = IWP(IW_FLT + 1, i, k) * (1.0 - tr) + W[IX(i, 3, k)] * tr;
      IWP(IW_FLTP, i, k) = IWP(IW_FLTP, i, k) * (1.0 - tr) + (P[PX(i, 1, k)] - p_bar) * tr;
    }
  for (int k = 1; k <= N3M; k++) {
    const int kp = o->kpa[k], km = o->kma[k];
    for (int i = 1; i <= N1M; i++) {
      const int ip = o->ipa[i], im = o->ima[i];
      const double Dz = IWP(IW_DZ, i, k);
      IWP(IW_UNSDY, i, k) = (IWP(IW_INTE + 0, i, k) - IWP(IW_INTEM + 0, i, k)) / wm_dt;
      IWP(IW_UNSDY + 1, i, k) = (IWP(IW_INTE + 2, i, k) - IWP(IW_INTEM + 2, i, k)) / wm_dt;
      const double uux = (IWP(IW_INTE + 1, ip, k) - IWP(IW_INTE + 1, im, k)) * DX1 / 2.0;
      const double uvy = (IWP(IW_INTE + 4, i, kp) - IWP(IW_INTE + 4, i, km)) * DX3 / 2.0;
      const double uvx = (IWP(IW_INTE + 4, ip, k) - IWP(IW_INTE + 4, im, k)) * DX1 / 2.0;
      const double vvy = (IWP(IW_INTE + 3, i, kp) - IWP(IW_INTE + 3, i, km)) * DX3 / 2.0;
      const double ux = (IWP(IW_INTE + 0, ip, k) - IWP(IW_INTE + 0, im, k)) * DX1 / 2.0;
      const double vy = (IWP(IW_INTE + 2, i, kp) - IWP(IW_INTE + 2, i, km)) * DX3 / 2.0;
      const double cx = uux + uvy - IWP(IW_FLTM, i, k) * (ux + vy);
      const double cy = uvx + vvy - IWP(IW_FLTM + 1, i, k) * (ux + vy);
      const double px = (IWP(IW_FLTP, ip, k) - IWP(IW_FLTP, im, k)) * DX1 / 2.0 * Dz + presg * Dz;
      const double py = (IWP(IW_FLTP, i, kp) - IWP(IW_FLTP, i, km)) * DX3 / 2.0 * Dz + presg3 * Dz;
      IWP(IW_CONV, i, k) = cx;
      IWP(IW_CONV + 1, i, k) = cy;
      IWP(IW_PRSG, i, k) = px;
      IWP(IW_PRSG + 1, i, k) = py;
      IWP(IW_LHS, i, k) = -IWP(IW_INTE + 0, i, k) + wm_dt * (cx + px - IWP(IW_DIFF, i, k));
      IWP(IW_LHS + 1, i, k) = -IWP(IW_INTE + 2, i, k) + wm_dt * (cy + py - IWP(IW_DIFF + 1, i, k));
    }
  }
}

/* iwm_slv_smooth (:3881-3926): the residuals of the integrated balance for
 * the friction velocities (utx, uty) */
static void iwm_slv(double lhsx, double lhsy, double Ux, double Uy, double Dz, double utx, double uty,
                    double re, double *fx, double *fy) {
  const double vonk = 0.4, reinv = 1.0 / re;
  const double us = pow(pow(utx, 4.0) + pow(uty, 4.0), 0.25);
  const double delv = reinv / us;
  const double deli = fmin(11.0 * reinv / us, Dz);
  const double Vel = sqrt(Ux * Ux + Uy * Uy);
  const double lg = log(deli / Dz);
  const double Ax = (Ux / us + Ux / Vel / vonk * lg - deli / delv * utx / us) / (1.0 - deli / Dz);
  const double Ay = (Uy / us + Uy / Vel / vonk * lg - deli / delv * uty / us) / (1.0 - deli / Dz);
  const double Cx = Ux / us - Ax;
  const double Cy = Uy / us - Ay;
  const double inteLu = 0.5 * utx * (deli * deli) / delv +
                        us * Dz * (0.5 * Ax * (1.0 - (deli * deli) / (Dz * Dz)) + Cx * (1.0 - deli / Dz) -
                                   1.0 / vonk * Ux / Vel * (1.0 - deli / Dz + deli / Dz * lg));
  const double inteLv = 0.5 * uty * (deli * deli) / delv +
                        us * Dz * (0.5 * Ay * (1.0 - (deli * deli) / (Dz * Dz)) + Cy * (1.0 - deli / Dz) -
                                   1.0 / vonk * Uy / Vel * (1.0 - deli / Dz + deli / Dz * lg));
  *fx = inteLu + lhsx;
  *fy = inteLv + lhsy;
}

/* iwm_calc_wallstress_smooth (:3930-4279) at one point; returns 1 when the
 * excessive-linear-correction branch found no equilibrium stress of the
 * point's own */
static int iwm_point(c3o *o, int i, int k) {
  const int N1M = o->n1m, N3M = o->n3m;
  const double *DY = o->dy;
  const double vonk = 0.4, re = o->re, reinv = 1.0 / re;
  const double deltap = DY[1] + DY[2] + DY[3] / 2.0;
  const int MaxIter = 1500;
  const double wm_tol = 0.000001, wm_eps = 0.000000001;
  const double lx = IWP(IW_LHS, i, k), ly = IWP(IW_LHS + 1, i, k);
  const double Ux = IWP(IW_FLT, i, k), Uy = IWP(IW_FLT + 1, i, k), Dz = IWP(IW_DZ, i, k);
  double utx = IWP(IW_UTX, i, k), uty = IWP(IW_UTY, i, k);
  double fx, fy, fxp, fyp;
  iwm_slv(lx, ly, Ux, Uy, Dz, utx, uty, re, &fx, &fy);
  int iter = 0, equil = 0, have_eq = 0, fixed = 0;
  while (fmax(fabs(fx), fabs(fy)) > wm_tol) {
    iwm_slv(lx, ly, Ux, Uy, Dz, utx + wm_eps, uty, re, &fxp, &fyp);
    const double a11 = (fxp - fx) / wm_eps, a21 = (fyp - fy) / wm_eps;
    iwm_slv(lx, ly, Ux, Uy, Dz, utx, uty + wm_eps, re, &fxp, &fyp);
    const double a12 = (fxp - fx) / wm_eps, a22 = (fyp - fy) / wm_eps;
    const double nx = utx - (a22 * fx - a12 * fy) / (a11 * a22 - a12 * a21);
    const double ny = uty - (-a21 * fx + a11 * fy) / (a11 * a22 - a12 * a21);
    utx = nx;
    uty = ny;
    iwm_slv(lx, ly, Ux, Uy, Dz, utx, uty, re, &fx, &fy);
    iter = iter + 1;
    if (iter > MaxIter) {
      equil = 1;
      break;
    }
  }
  if (utx - 1.0 == utx || uty - 1.0 == uty) equil = 1;
  double equilutx = 0.0, equiluty = 0.0;
  if (equil) {
    const double U_ave = sqrt(Ux * Ux + Uy * Uy);
    const double ut = iwm_equil(U_ave, deltap, re);
    const double TAUW1 = ut * ut;
    equilutx = TAUW1 * Ux / U_ave;
    equiluty = TAUW1 * Uy / U_ave;
    utx = sqrt(fabs(equilutx));
    uty = sqrt(fabs(equiluty));
    have_eq = 1;
  }
  /* the Newton (or equilibrium) friction velocities: everything below uses
   * these (wmputx, wmputy), also where the branch after the linear
   * correction stores other values in the state (:4110-4147) */
  const double putx = utx, puty = uty;
  const double utaup = pow(pow(putx, 4.0) + pow(puty, 4.0), 0.25);
  const double delv = reinv / utaup;
  const double deli = fmin(11.0 * reinv / utaup, Dz);
  const double Vel = sqrt(Ux * Ux + Uy * Uy);
  const double lg = log(deli / Dz);
  double Ax, Ay, Cx, Cy;
  if (equil) {
    Ax = Ay = Cx = Cy = 0.0;
  } else {
    Ax = (Ux / utaup + Ux / Vel / vonk * lg - deli / delv * putx / utaup) / (1.0 - deli / Dz);
    Ay = (Uy / utaup + Uy / Vel / vonk * lg - deli / delv * puty / utaup) / (1.0 - deli / Dz);
    Cx = Ux / utaup - Ax;
    Cy = Uy / utaup - Ay;
  }
  if (fabs(Ax) > 1.0 || fabs(Ay) > 1.0) {
    if (!have_eq) {
      const double U_ave = sqrt(Ux * Ux + Uy * Uy);
      const double ut = iwm_equil(U_ave, deltap, re);
      const double TAUW1 = ut * ut;
      equilutx = TAUW1 * Ux / U_ave;
      equiluty = TAUW1 * Uy / U_ave;
      fixed = 1;
    }
    equil = 1;
    utx = sqrt(fabs(equilutx));
    uty = sqrt(fabs(equiluty));
    Ax = Ay = Cx = Cy = 0.0;
  }
  IWP(IW_UTX, i, k) = utx;  /* the state: after the branch */
  IWP(IW_UTY, i, k) = uty;
  IWP(IW_AX, i, k) = Ax;
  IWP(IW_AY, i, k) = Ay;
  IWP(IW_CX, i, k) = Cx;
  IWP(IW_CY, i, k) = Cy;
  for (int q = 0; q < 5; q++) IWP(IW_INTEM + q, i, k) = IWP(IW_INTE + q, i, k);
  const double d2 = deli * deli, d3 = deli * deli * deli, z2 = Dz * Dz, z3 = Dz * Dz * Dz;
  const double dv2 = delv * delv, us2 = utaup * utaup, V2 = Vel * Vel, k2 = vonk * vonk;
  IWP(IW_INTE + 0, i, k) = 0.5 * putx * d2 / delv +
                           utaup * Dz * (0.5 * Ax * (1.0 - d2 / z2) + Cx * (1.0 - deli / Dz) -
                                         1.0 / vonk * Ux / Vel * (1.0 - deli / Dz + deli / Dz * lg));
  IWP(IW_INTE + 2, i, k) = 0.5 * puty * d2 / delv +
                           utaup * Dz * (0.5 * Ay * (1.0 - d2 / z2) + Cy * (1.0 - deli / Dz) -
                                         1.0 / vonk * Uy / Vel * (1.0 - deli / Dz + deli / Dz * lg));
  IWP(IW_INTE + 4, i, k) =
      1.0 / 3.0 * putx * puty * d3 / dv2 +
      us2 * Dz *
          (-1.0 / vonk * (Ax * Uy / Vel + Ay * Ux / Vel) * (0.25 - 0.25 * d2 / z2 + 0.5 * d2 / Dz * lg) -
           1.0 / vonk * (Cx * Uy / Vel + Cy * Ux / Vel) * (1.0 - deli / Dz + deli / Dz * lg) -
           Ux * Uy / V2 / k2 * (deli / Dz - 2.0 + deli / Dz * ((lg - 1.0) * (lg - 1.0))) +
           1.0 / 3.0 * Ax * Ay * (1.0 - d3 / z3) + 0.5 * (Ax * Cy + Ay * Cx) * (1.0 - d2 / z2) +
           Cx * Cy * (1.0 - deli / Dz));
  {
    const double tx = Cx - 1.0 / vonk * Ux / Vel, sx = Cx - Ux / Vel / vonk + Ux / Vel / vonk * lg;
    const double ty = Cy - 1.0 / vonk * Uy / Vel, sy = Cy - Uy / Vel / vonk + Uy / Vel / vonk * lg;
    IWP(IW_INTE + 1, i, k) =
        1.0 / 3.0 * (putx * putx) * d3 / dv2 +
        us2 * Dz *
            (-Ax * Ux / Vel / vonk * d2 / z2 * lg + Ax * (Cx - 1.0 / 2.0 / vonk * Ux / Vel) * (1.0 - d2 / z2) +
             Ax * Ax / 3.0 * (1.0 - d3 / z3) + tx * tx - deli / Dz * (sx * sx) +
             1.0 / k2 * (Ux * Ux) / V2 * (1.0 - deli / Dz));
    IWP(IW_INTE + 3, i, k) =
        1.0 / 3.0 * (puty * puty) * d3 / dv2 +
        us2 * Dz *
            (-Ay * Uy / Vel / vonk * d2 / z2 * lg + Ay * (Cy - 1.0 / 2.0 / vonk * Uy / Vel) * (1.0 - d2 / z2) +
             Ay * Ay / 3.0 * (1.0 - d3 / z3) + ty * ty - deli / Dz * (sy * sy) +
             1.0 / k2 * (Uy * Uy) / V2 * (1.0 - deli / Dz));
  }
  double tx, ty;
  if (deli < Dz) {
    tx = utaup / Dz * (Ax + 1.0 / vonk * Ux / Vel);
    ty = utaup / Dz * (Ay + 1.0 / vonk * Uy / Vel);
  } else {
    tx = putx / delv;
    ty = puty / delv;
  }
  const double bx = putx / delv, by = puty / delv;
  IWP(IW_DUDZT, i, k) = tx;
  IWP(IW_DUDZT + 1, i, k) = ty;
  IWP(IW_DUDZB, i, k) = bx;
  IWP(IW_DUDZB + 1, i, k) = by;
  const double dVelzT = fabs(Ux / Vel * tx + Uy / Vel * ty);
  if (deli < Dz) {
    IWP(IW_DIFF, i, k) = (reinv + (vonk * Dz) * (vonk * Dz) * dVelzT) * tx - reinv * bx;
    IWP(IW_DIFF + 1, i, k) = (reinv + (vonk * Dz) * (vonk * Dz) * dVelzT) * ty - reinv * by;
  } else {
    IWP(IW_DIFF, i, k) = 0.0;
    IWP(IW_DIFF + 1, i, k) = 0.0;
  }
  if (equil) {
    IWP(IW_TAUWX, i, k) = equilutx;
    IWP(IW_TAUWY, i, k) = equiluty;
  } else if (deli < Dz) {
    IWP(IW_TAUWX, i, k) = putx * putx;
    IWP(IW_TAUWY, i, k) = puty * puty;
  } else {
    IWP(IW_TAUWX, i, k) = reinv * Ux / Dz;
    IWP(IW_TAUWY, i, k) = reinv * Uy / Dz;
  }
  const double tR = IWP(IW_TR, i, k);
  IWP(IW_FLTUS, i, k) = IWP(IW_FLTUS, i, k) * (1.0 - tR) + utaup * tR;
  double ntr = o->wm_dt / (Dz / IWP(IW_FLTUS, i, k) / vonk);
  if (ntr > 1.0) ntr = 1.0;
  IWP(IW_TR, i, k) = ntr;
  return fixed;
}

void c3o_wallmodel_iwmles(c3o *o, const double *U, const double *P, double presg, double presg3, int ntime) {
  const int N1M = o->n1m, N3M = o->n3m;
  if (ntime == 1) iwm_init(o, U);
  /* iwm_wallstress_smooth (:3489-3557) */
  if (ntime % IWM_SKIP == 1)
    o->wm_dt = o->dt;
  else
    o->wm_dt = o->wm_dt + o->dt;
  o->iwm_fixed = 0;
  if (ntime % IWM_SKIP == 0) {
    iwm_lhs(o, U, P, presg, presg3);
    for (int i = 1; i <= N1M; i++)
      for (int k = 1; k <= N3M; k++) o->iwm_fixed += iwm_point(o, i, k);
  }
  for (int i = 1; i <= N1M; i++)
    for (int k = 1; k <= N3M; k++) {
      o->tauw[TAUX(i, k, 1)] = IWP(IW_TAUWX, i, k);
      o->taub[TAUX(i, k, 1)] = IWP(IW_TAUWY, i, k);
    }
  /* WSM(1) = TAOW, the mean of |tau| (:3463-3476) */
  double wsm1 = 0.0;
  for (int k = 1; k <= N3M; k++)
    for (int i = 1; i <= N1M; i++) {
      const double a = o->tauw[TAUX(i, k, 1)], b = o->taub[TAUX(i, k, 1)];
      wsm1 = wsm1 + sqrt(a * a + b * b);
    }
  o->wsm[0] = wsm1 / (double)(N1M * N3M);
  o->taow = o->wsm[0];
}

void c3o_iwm_state(const c3o *o, double *planes, double *wm_dt, int *fixed) {
  if (planes && o->iwm) memcpy(planes, o->iwm, sizeof(double) * IW_NPL * o->n1m * o->n3m);
  if (wm_dt) *wm_dt = o->wm_dt;
  if (fixed) *fixed = o->iwm_fixed;
}

void c3o_iwm_set(c3o *o, const double *planes, double wm_dt) {
  if (o->iwm) memcpy(o->iwm, planes, sizeof(double) * IW_NPL * o->n1m * o->n3m);
  o->wm_dt = wm_dt;
}

void c3o_set_cflmax(c3o *o, double cflmax) { o->cflmax = cflmax; }
#undef IWP

/* ------------------------------------------------------- WALLMODEL_WERNER */
/* compact3d.f:4815-4958 (WMODEL = 4): the Werner-Wengle power-law wall
 * stress at both walls from the first cell's centre velocity, a linear
 * (viscous) branch below VCASE = AB3*sqrt|WSM(1)|/11.81 (WSM(1) of the
 * previous call), then averaged onto the faces.  The reference takes the
 * I neighbour as IPA(K) and IMA(K) (:4849, :4925: the K index in IPA/IMA),
 * which is reproduced; it is defined for N3-1 <= N1-1. */
void c3o_wallmodel_werner(c3o *o, const double *U) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m;
  const double RE = o->re;
  const double *DY = o->dy;
  const double *W = U + 2 * NU;
  const double B = 0.1428571429;
  const double AB1 = 7.202125274, AB3 = 69.74057973, AB4 = 0.1376936317, AB5 = 1.75;
  double *TWE = dalloc((size_t)N1M * N3M * 4);
#define TE(i, k, l) TWE[((size_t)(l) - 1) * N1M * N3M + ((size_t)(k) - 1) * N1M + ((i) - 1)]
  for (int wall = 0; wall < 2; wall++) {
    const int j = wall == 0 ? 1 : N2M;
    const double deltap = DY[j];
    for (int k = 1; k <= N3M; k++) {
      const int kp = o->kpa[k];
      for (int i = 1; i <= N1M; i++) {
        const int ip = o->ipa[k];
        const double u1 = 0.5 * (U[IX(i, j, k)] + U[IX(ip, j, k)]);
        const double u2 = 0.5 * (W[IX(i, j, k)] + W[IX(i, j, kp)]);
        const double velocp = sqrt(u1 * u1 + u2 * u2);
        const double temp = sqrt(fabs(o->wsm[0])) / 11.81;
        const double vcase = AB3 * temp;
        double tauw1 = 0.0;
        if (velocp <= vcase) {
          tauw1 = 2. * velocp / deltap / RE;
        } else if (velocp > vcase) {
          tauw1 = AB1 * pow(1.0 / deltap / RE, 1. + B);
          tauw1 = tauw1 + AB4 * pow(1.0 / deltap / RE, B) * velocp;
          tauw1 = pow(tauw1, AB5);
        }
        TE(i, k, 2 * wall + 1) = tauw1 * (u1 / velocp);
        TE(i, k, 2 * wall + 2) = tauw1 * (u2 / velocp);
      }
    }
  }
  for (int k = 1; k <= N3M; k++) {
    const int kp = o->kma[k];
    for (int i = 1; i <= N1M; i++) {
      const int ip = o->ima[k];
      o->tauw[TAUX(i, k, 1)] = 0.5 * (TE(i, k, 1) + TE(ip, k, 1));
      o->taub[TAUX(i, k, 1)] = 0.5 * (TE(i, k, 2) + TE(i, kp, 2));
      o->tauw[TAUX(i, k, 2)] = 0.5 * (TE(i, k, 3) + TE(ip, k, 3));
      o->taub[TAUX(i, k, 2)] = 0.5 * (TE(i, k, 4) + TE(i, kp, 4));
    }
  }
  double w1 = 0.0;
  for (int k = 1; k <= N3M; k++)
    for (int i = 1; i <= N1M; i++) w1 = w1 + TE(i, k, 1);
  o->wsm[0] = w1 / (double)(N1M * N3M);
  o->wsm[1] = 0.0;
  o->wsm[2] = 0.0;
  o->taow = o->wsm[0];
#undef TE
  free(TWE);
}

/* ------------------------------------------------------------ FILTER(4) */
/* FILTER (compact3d.f:5190-5228): FILV(1:N1,1:N3,n) */
void c3o_filter(c3o *o, double *FILV, int n) {
  const int N1M = o->n1m, N3M = o->n3m, M1 = o->n1, M3 = o->n3;
  double *T = dalloc((size_t)M1 * M3);
#define FV(i, k, l) FILV[((size_t)(l) - 1) * M1 * M3 + ((size_t)(k) - 1) * M1 + ((i) - 1)]
#define TT(i, k) T[((size_t)(k) - 1) * M1 + ((i) - 1)]
  for (int l = 1; l <= n; l++) {
    for (int k = 1; k <= N3M; k++)
      for (int i = 1; i <= N1M; i++)
        TT(i, k) = (FV(o->ima[i], k, l) + 4.0 * FV(i, k, l) + FV(o->ipa[i], k, l)) / 6.0;
    for (int k = 1; k <= N3M; k++)
      for (int i = 1; i <= N1M; i++)
        FV(i, k, l) = (TT(i, o->kma[k]) + 4.0 * TT(i, k) + TT(i, o->kpa[k])) / 6.0;
  }
  free(T);
}

/* FILTER4 (compact3d.f:5481-5546) */
void c3o_filter4(c3o *o, double *FILV, int n) {
  const int N1M = o->n1m, N3M = o->n3m, M1 = o->n1, M3 = o->n3;
  double *T = dalloc((size_t)M1 * M3);
  int ip2[N1M + 1], im2[N1M + 1], kp2[N3M + 1], km2[N3M + 1];
  for (int i = 1; i <= N1M; i++) {
    ip2[i] = i + 2;
    im2[i] = i - 2;
  }
  ip2[N1M] = 2;
  ip2[N1M - 1] = 1;
  im2[1] = N1M - 1;
  im2[2] = N1M;
  for (int k = 1; k <= N3M; k++) {
    kp2[k] = k + 2;
    km2[k] = k - 2;
  }
  kp2[N3M] = 2;
  kp2[N3M - 1] = 1;
  km2[1] = N3M - 1;
  km2[2] = N3M;
  for (int l = 1; l <= n; l++) {
    for (int k = 1; k <= N3M; k++)
      for (int i = 1; i <= N1M; i++)
        TT(i, k) = (FV(im2[i], k, l) + 4.0 * FV(o->ima[i], k, l) + 2.0 * FV(i, k, l) +
                    4.0 * FV(o->ipa[i], k, l) + FV(ip2[i], k, l)) / 12.0;
    for (int k = 1; k <= N3M; k++)
      for (int i = 1; i <= N1M; i++)
        FV(i, k, l) = (TT(i, km2[k]) + 4.0 * TT(i, o->kma[k]) + 2.0 * TT(i, k) +
                       4.0 * TT(i, o->kpa[k]) + TT(i, kp2[k])) / 12.0;
  }
  free(T);
#undef FV
#undef TT
}

/* ------------------------------------------------------------- ENERGY */
/* The plane statistics of ENERGY (compact3d.f:9902): cell-centre vorticity
 * (:10005-10069) and, per J plane, the means and second moments of loop 20
 * (:10073-10158), in the reference's K, I summation order.  out[s*N2M +
 * J-1], s = 0 VM1 1 VM2 2 VM3 3 PM 4 WM1 5 WM2 6 WM3 7-12 VRMS1..6 (uu vv ww
 * uv uw vw) 13 PRMS 14 PV 15-17 WRMS1..3 18 U2V 19 VW2. */
void c3o_plane_stats(c3o *o, const double *U, const double *P, double *out) {
  const int N1M = o->n1m, N2M = o->n2m, N3M = o->n3m, N2 = o->n2;
  const double DX1 = o->dx1, DX3 = o->dx3;
  const double *DY = o->dy, *H = o->h;
  const double *V = U + NU, *W = U + 2 * NU;
  const double nxz = 1.0 / (double)(N1M * N3M);
#define PP(i, j, k) P[((size_t)(k) - 1) * o->n1 * o->n2 + ((size_t)(j) - 1) * o->n1 + ((i) - 1)]
#pragma omp parallel for
  for (int j = 1; j <= N2M; j++) {
    const int jp = j + 1, jm = j - 1;
    const double fjum = o->fjmu[j], fjup = o->fjpa[j];
    double acc[20];
    for (int q = 0; q < 20; q++) acc[q] = 0.0;
    for (int k = 1; k <= N3M; k++) {
      const int kp = o->kpa[k], km = o->kma[k];
      for (int i = 1; i <= N1M; i++) {
        const int ip = o->ipa[i], im = o->ima[i];
        /* vorticity at the cell centre (:10024-10069) */
        const double u1kp = 0.5 * (U[IX(i, j, kp)] + U[IX(ip, j, kp)]);
        const double u1km = 0.5 * (U[IX(i, j, km)] + U[IX(ip, j, km)]);
        const double u2kp = 0.5 * (V[IX(i, j, kp)] + V[IX(i, jp, kp)]);
        const double u2km = 0.5 * (V[IX(i, j, km)] + V[IX(i, jp, km)]);
        const double u2ip = 0.5 * (V[IX(ip, j, k)] + V[IX(ip, jp, k)]);
        const double u2im = 0.5 * (V[IX(im, j, k)] + V[IX(im, jp, k)]);
        const double u3ip = 0.5 * (W[IX(ip, j, k)] + W[IX(ip, j, kp)]);
        const double u3im = 0.5 * (W[IX(im, j, k)] + W[IX(im, j, kp)]);
        const double u1jp = 0.5 * (U[IX(i, jp, k)] + U[IX(ip, jp, k)]);
        const double u1jc = 0.5 * (U[IX(i, j, k)] + U[IX(ip, j, k)]);
        const double u1jm = 0.5 * (U[IX(i, jm, k)] + U[IX(ip, jm, k)]);
        double u12 = 0.5 / H[jp] * (DY[jp] * u1jc + DY[j] * u1jp);
        double u11 = 0.5 / H[j] * (DY[j] * u1jm + DY[jm] * u1jc);
        u12 = u12 * fjup + (1. - fjup) * (U[IX(i, N2, k)] + U[IX(ip, N2, k)]) * 0.5;
        u11 = u11 * fjum + (1. - fjum) * (U[IX(i, 0, k)] + U[IX(ip, 0, k)]) * 0.5;
        const double u3jp = 0.5 * (W[IX(i, jp, k)] + W[IX(i, jp, kp)]);
        const double u3jc = 0.5 * (W[IX(i, j, k)] + W[IX(i, j, kp)]);
        const double u3jm = 0.5 * (W[IX(i, jm, k)] + W[IX(i, jm, kp)]);
        double u32 = 0.5 / H[jp] * (DY[jp] * u3jc + DY[j] * u3jp);
        double u31 = 0.5 / H[j] * (DY[j] * u3jm + DY[jm] * u3jc);
        u32 = u32 * fjup + (1. - fjup) * (W[IX(i, N2, k)] + W[IX(i, N2, kp)]) * 0.5;
        u31 = u31 * fjum + (1. - fjum) * (W[IX(i, 0, k)] + W[IX(i, 0, kp)]) * 0.5;
        const double dv3dx2 = (u32 - u31) / DY[j];
        const double dv2dx3 = (u2kp - u2km) * 0.5 * DX3;
        const double dv3dx1 = (u3ip - u3im) * 0.5 * DX1;
        const double dv1dx3 = (u1kp - u1km) * 0.5 * DX3;
        const double dv2dx1 = (u2ip - u2im) * 0.5 * DX1;
        const double dv1dx2 = (u12 - u11) / DY[j];
        const double w1 = dv3dx2 - dv2dx3, w2 = dv1dx3 - dv3dx1, w3 = dv2dx1 - dv1dx2;
        /* loop 20 (:10114-10139) */
        const double v1 = (U[IX(i, j, k)] + U[IX(ip, j, k)]) * 0.5;
        const double v2 = (V[IX(i, j, k)] + V[IX(i, jp, k)]) * 0.5;
        const double v3 = (W[IX(i, j, k)] + W[IX(i, j, kp)]) * 0.5;
        const double p = PP(i, j, k);
        acc[0] += v1;
        acc[1] += v2;
        acc[2] += v3;
        acc[3] += p;
        acc[4] += w1;
        acc[5] += w2;
        acc[6] += w3;
        acc[7] += v1 * v1;
        acc[8] += v2 * v2;
        acc[9] += v3 * v3;
        acc[10] += v1 * v2;
        acc[11] += v1 * v3;
        acc[12] += v2 * v3;
        acc[13] += p * p;
        acc[14] += p * v2;
        acc[15] += w1 * w1;
        acc[16] += w2 * w2;
        acc[17] += w3 * w3;
        acc[18] += v1 * v1 * v2;
        acc[19] += v2 * (v3 * v3);
      }
    }
    for (int q = 0; q < 20; q++) out[(size_t)q * N2M + (j - 1)] = acc[q] * nxz;
  }
#undef PP
}

/* ------------------------------------------------------------------ GETUP */
/* compact3d.f:879-960 */
void c3o_getup(c3o *o, double *U, double *P, double *HH, double *SGSVIS,
               double *presg, double *presg3, int ntime) {
  const size_t nu = NU, np = (size_t)o->n1 * o->n2 * o->n3;
  double *UH = dalloc(3 * nu), *HH2 = dalloc(3 * nu), *RUH = dalloc(3 * nu);
  double *DP = dalloc(np), *RDP = dalloc(np);
  for (size_t t = 0; t < nu; t++) SGSVIS[t] = 0.0;
  memcpy(HH2, HH, sizeof(double) * 3 * nu);
  memset(HH, 0, sizeof(double) * 3 * nu);
  if (o->files == 1.0) c3o_sgsdy(o, U, SGSVIS, NULL);
  if (o->files == 2.0) c3o_sgscs(o, U, SGSVIS);
  if (o->files == 3.0) c3o_sgssd(o, U, SGSVIS, NULL, NULL);
  if (o->files != 0.0) {
    if (o->wmodel == 0.0) {
      o->ifwall = 0;
      c3o_wallss(o, U);
    } else if (o->wmodel == 1.0) {
      c3o_wallmodel_instant(o, U);
      o->ifwall = 1;
    } else if (o->wmodel == 2.0) {
      c3o_wallmodel_iwmles(o, U, P, *presg, *presg3, ntime);
      o->ifwall = 1;
    } else if (o->wmodel == 4.0) {
      c3o_wallmodel_werner(o, U);
      o->ifwall = 1;
    } else if (o->wmodel == 6.0) {
      c3o_wallmodel_instant2(o, U);
      o->ifwall = 1;
    }
  }
  c3o_bcond(o, U, SGSVIS);
  c3o_rhs1(o, U, P, HH, HH2, RUH, *presg, SGSVIS);
  c3o_rhs2(o, U, P, HH, HH2, RUH + nu, SGSVIS);
  c3o_rhs3(o, U, P, HH, HH2, RUH + 2 * nu, *presg3, SGSVIS);
  c3o_getuh1(o, U, UH, RUH);
  c3o_getuh2(o, U, UH, RUH + nu);
  c3o_getuh3(o, U, UH, RUH + 2 * nu);
  c3o_rhsdp(o, RDP, UH);
  c3o_getdp(o, DP, RDP);
  c3o_upcalc(o, U, P, UH, DP, presg, presg3);
  free(UH);
  free(HH2);
  free(RUH);
  free(DP);
  free(RDP);
}
