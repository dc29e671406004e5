// c3d_kernels.hip -- HIP kernels of the compact3d hot path for gfx950 (FP64 path).
//
// Pass structure of one GETUP (compact3d.f:879-960):
//   k_bcond                      BCOND                           :963-1001
//   k_rhs_x, k_rhs_z, k_rhs_x2,  RHS1/2/3 (x-line, z-line and    :6063-8114
//   k_rhs_y                        pointwise y parts, AB2)
//   k_adi_y, k_adi_line(x/z)     GETUH1/2/3 (TDMAI, CTDMA1J/3I)  :8218-8627
//   k_rhsdp_x, k_rhsdp_z         RHSDP                           :8117-8214
//   k_fft_x, k_fft_z, k_pois_y   GETDP (2-D DFT per plane, one   :9271-9441
//     k_fft_z, k_fft_xinv          Thomas solve per mode)
//   k_flow_red/k_flow_fin,       UPCALC (+ DIVCHECK and CFL      :1052-1238
//   k_upd_x, k_ghosts, k_upd_z     fused into the update passes)  :9049-9169
//
// Periodic compact operators are solved by tile_solve() (c3d_dev.h) on LDS
// tiles of L lines: x-lines are contiguous rows, z-lines are L consecutive i
// at one j, so every global access is a coalesced row access.
#include <math.h>

#include "c3d_impl.h"

// ------------------------------------------------------------------ BCOND
__global__ void k_bcond(Ctx c) {
  const Geo g = c.g;
  const long long n = (long long)g.nx * g.nz;
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = t % g.nx, k = t / g.nx;
  const Fields F = c.d;
  const double aly = c.s.aly;
  long long b0 = gidx(g, i, 0, k), bn = gidx(g, i, g.n2, k), bm = gidx(g, i, g.n2m, k);
  F.u[b0] = 0.0;
  F.v[b0] = 0.0;
  F.w[b0] = 0.0;
  double ut = F.u[bm] * (2.0 - aly);
  double wt = F.w[bm] * (2.0 - aly);
  F.ubc_u[t] = ut;
  F.ubc_w[t] = wt;
  F.u[bn] = ut;
  F.v[bn] = 0.0;
  F.w[bn] = wt;
  if (c.s.files) {
    F.sgs[b0] = 0.0;
    F.sgs[bn] = F.sgs[bm] * (2.0 - aly);
  }
}

// ------------------------------------------------------------ RHS x-pass
// per x-line (j,k), j = 1..N2M:
//   RUH1  = (1/RE) D2x u - D1x(P) - PRESG                 (:6142-6157, :6196-6210)
//   HH1   = D1x((Ix u)^2)                                  (:6295-6334)
//   RUH2  = (1/RE) D2x v                         (j>=2)   (:6879-6895)
//   VV1   = Ix v                 (stored)                  (:6228-6242, :6946-6959)
//   HH2   = D1x(UU2 * VV1)                       (j>=2)   (:6961-7004)
//   RUH3  = (1/RE) D2x w                                   (:7483-7498)
//   WW1   = Ix w                 (stored)                  (:6380-6394, :7718-7732)
__global__ __launch_bounds__(C3D_NT) void k_rhs_x(Ctx c) {
  extern __shared__ double sm[];
  const Geo g = c.g;
  const TileShape ts = c.o.tx;
  const int n = ts.n, L = ts.nl, LS = ts.ls, nL = n * L;
  double* Wt = sm;
  double* Kt = sm + n * LS;
  const int nlines = g.n2m * g.nz;
  const int ln0 = blockIdx.x * L;
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  const double presg = F.sc[SC_PRESG];
  double acc[MAXE];

#define XDEC                         \
  const int l = e / n, p = e % n;    \
  const int ln = ln0 + l;            \
  const bool ok = ln < nlines;       \
  const int j = 1 + (ok ? ln % g.n2m : 0); \
  const int k = ok ? ln / g.n2m : 0; \
  const long long b = gidx(g, 0, j, k); \
  const int pm = wrapm(p, n), pp = wrapp(p, n);

  // D2x u
  EL_LOOP(nL) XDEC;
  Wt[p * LS + l] = ok ? 12. * s.dx1q * (F.u[b + pm] - 2. * F.u[b + p] + F.u[b + pp]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s10);
  EL_LOOP(nL) XDEC;
  (void)ok; (void)b; (void)pm; (void)pp;
  acc[m] = s.rinv * Wt[p * LS + l];
  EL_END;
  __syncthreads();
  // D1x(P), centre -> face
  EL_LOOP(nL) XDEC;
  Wt[p * LS + l] = ok ? 24. * s.dx1 * (-F.p[b + pm] + F.p[b + p]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s22);
  EL_LOOP(nL) XDEC;
  if (ok) F.ruh[0][b + p] = acc[m] - Wt[p * LS + l] - presg;
  EL_END;
  __syncthreads();
  // UU1 = Ix u (face -> centre), then D1x(UU1^2)
  EL_LOOP(nL) XDEC;
  Wt[p * LS + l] = ok ? 4.0 * (F.u[b + p] + F.u[b + pp]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s6);
  EL_LOOP(nL) XDEC;
  double a0 = Wt[p * LS + l], am = Wt[pm * LS + l];
  acc[m] = 24. * s.dx1 * (a0 * a0 - am * am);
  EL_END;
  __syncthreads();
  EL_LOOP(nL) XDEC;
  Wt[p * LS + l] = acc[m];
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s22);
  EL_LOOP(nL) XDEC;
  if (ok) F.hh[0][b + p] = Wt[p * LS + l];
  EL_END;
  __syncthreads();
  // D2x v
  EL_LOOP(nL) XDEC;
  Wt[p * LS + l] = ok ? 12. * s.dx1q * (F.v[b + pm] - 2. * F.v[b + p] + F.v[b + pp]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s10);
  EL_LOOP(nL) XDEC;
  if (ok) F.ruh[1][b + p] = (j >= 2) ? s.rinv * Wt[p * LS + l] : 0.0;
  EL_END;
  __syncthreads();
  // VV1 = Ix v (centre -> face)
  EL_LOOP(nL) XDEC;
  Wt[p * LS + l] = ok ? 4. * (F.v[b + pm] + F.v[b + p]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s6);
  EL_LOOP(nL) XDEC;
  double vv = Wt[p * LS + l];
  if (ok) F.vv1[b + p] = vv;
  // UU2 * VV1, UU2 = u interpolated to the v face j (:6966)
  double uu2 = 0.0;
  if (ok && j >= 2) {
    const long long bjm = gidx(g, 0, j - 1, k);
    uu2 = 1. / M.h[j] * (M.dy[j] / 2. * F.u[bjm + p] + M.dy[j - 1] / 2. * F.u[b + p]);
  }
  Kt[p * LS + l] = uu2 * vv;
  EL_END;
  __syncthreads();
  EL_LOOP(nL) XDEC;
  Wt[p * LS + l] = ok ? 24. * s.dx1 * (-Kt[p * LS + l] + Kt[pp * LS + l]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s22);
  EL_LOOP(nL) XDEC;
  if (ok) F.hh[1][b + p] = (j >= 2) ? Wt[p * LS + l] : 0.0;
  EL_END;
  __syncthreads();
  // D2x w
  EL_LOOP(nL) XDEC;
  Wt[p * LS + l] = ok ? 12. * s.dx1q * (F.w[b + pm] - 2. * F.w[b + p] + F.w[b + pp]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s10);
  EL_LOOP(nL) XDEC;
  if (ok) F.ruh[2][b + p] = s.rinv * Wt[p * LS + l];
  EL_END;
  __syncthreads();
  // WW1 = Ix w (centre -> face)
  EL_LOOP(nL) XDEC;
  Wt[p * LS + l] = ok ? 4. * (F.w[b + pm] + F.w[b + p]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s6);
  EL_LOOP(nL) XDEC;
  if (ok) F.ww1[b + p] = Wt[p * LS + l];
  EL_END;
#undef XDEC
}

// ------------------------------------------------------------ RHS z-pass
// per z-line (i,j), tile of L consecutive i at one j:
//   RUH1 += (1/RE) D2z u                                    (:6170-6185)
//   UW    = Iz(u) * WW1 (stored);  HH1 += D1z(UW)           (:6360-6433)
//   RUH2 += (1/RE) D2z v                          (j>=2)   (:6907-6923)
//   VV3   = Iz v (stored);  HH2 += D1z(VV3*WW2)   (j>=2)   (:7023-7084)
//   RUH3 += (1/RE) D2z w - D1z(P) - PRESG3                   (:7510-7550)
//   HH3   = D1z((Iz w)^2)                                    (:7633-7672)
__global__ __launch_bounds__(C3D_NT) void k_rhs_z(Ctx c) {
  extern __shared__ double sm[];
  const Geo g = c.g;
  const TileShape ts = c.o.tz;
  const int n = ts.n, L = ts.nl, LS = ts.ls, nL = n * L;
  double* Wt = sm;
  double* Kt = sm + n * LS;
  const int nit = (g.nx + L - 1) / L;
  const int it = blockIdx.x % nit;
  const int j = 1 + blockIdx.x / nit;
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  const double presg3 = F.sc[SC_PRESG3];
  const long long pl = g.plane * 1;  // k stride
  double acc[MAXE];

#define ZDEC                               \
  const int l = e % L, p = e / L;          \
  const int i = it * L + l;                \
  const bool ok = i < g.nx;                \
  const long long b = gidx(g, ok ? i : 0, j, 0); \
  const int pm = wrapm(p, n), pp = wrapp(p, n);

  // D2z u
  EL_LOOP(nL) ZDEC;
  Wt[p * LS + l] = ok ? 12. * s.dx3q * (F.u[b + pm * pl] - 2. * F.u[b + p * pl] + F.u[b + pp * pl]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s10);
  EL_LOOP(nL) ZDEC;
  if (ok) F.ruh[0][b + p * pl] = F.ruh[0][b + p * pl] + s.rinv * Wt[p * LS + l];
  EL_END;
  __syncthreads();
  // UU3 = Iz u (centre -> face); UW = UU3*WW1; HH1 += D1z(UW) face -> centre
  EL_LOOP(nL) ZDEC;
  Wt[p * LS + l] = ok ? 4. * (F.u[b + pm * pl] + F.u[b + p * pl]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s6);
  EL_LOOP(nL) ZDEC;
  double uw = ok ? Wt[p * LS + l] * F.ww1[b + p * pl] : 0.0;
  if (ok) F.uw[b + p * pl] = uw;
  Kt[p * LS + l] = uw;
  EL_END;
  __syncthreads();
  EL_LOOP(nL) ZDEC;
  Wt[p * LS + l] = ok ? 24. * s.dx3 * (-Kt[p * LS + l] + Kt[pp * LS + l]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s22);
  EL_LOOP(nL) ZDEC;
  if (ok) F.hh[0][b + p * pl] = F.hh[0][b + p * pl] + Wt[p * LS + l];
  EL_END;
  __syncthreads();
  // D2z v (j>=2)
  if (j >= 2) {
    EL_LOOP(nL) ZDEC;
    Wt[p * LS + l] = ok ? 12. * s.dx3q * (F.v[b + pm * pl] - 2. * F.v[b + p * pl] + F.v[b + pp * pl]) : 0.0;
    EL_END;
    __syncthreads();
    tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s10);
    EL_LOOP(nL) ZDEC;
    if (ok) F.ruh[1][b + p * pl] = F.ruh[1][b + p * pl] + s.rinv * Wt[p * LS + l];
    EL_END;
    __syncthreads();
  }
  // VV3 = Iz v (centre -> face), stored for all j; HH2 += D1z(VV3*WW2) for j>=2
  EL_LOOP(nL) ZDEC;
  Wt[p * LS + l] = ok ? 4. * (F.v[b + pm * pl] + F.v[b + p * pl]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s6);
  EL_LOOP(nL) ZDEC;
  double vv3 = Wt[p * LS + l];
  if (ok) F.vv3[b + p * pl] = vv3;
  double ww2 = 0.0;
  if (ok && j >= 2) {
    const long long bjm = gidx(g, i, j - 1, 0);
    ww2 = 1. / M.h[j] * (M.dy[j] / 2. * F.w[bjm + p * pl] + M.dy[j - 1] / 2. * F.w[b + p * pl]);
  }
  Kt[p * LS + l] = vv3 * ww2;
  EL_END;
  __syncthreads();
  if (j >= 2) {
    EL_LOOP(nL) ZDEC;
    Wt[p * LS + l] = ok ? 24. * s.dx3 * (-Kt[p * LS + l] + Kt[pp * LS + l]) : 0.0;
    EL_END;
    __syncthreads();
    tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s22);
    EL_LOOP(nL) ZDEC;
    if (ok) F.hh[1][b + p * pl] = F.hh[1][b + p * pl] + Wt[p * LS + l];
    EL_END;
    __syncthreads();
  }
  // D2z w, D1z(P): RUH3
  EL_LOOP(nL) ZDEC;
  Wt[p * LS + l] = ok ? 12. * s.dx3q * (F.w[b + pm * pl] - 2. * F.w[b + p * pl] + F.w[b + pp * pl]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s10);
  EL_LOOP(nL) ZDEC;
  (void)ok; (void)b;
  acc[m] = s.rinv * Wt[p * LS + l];
  EL_END;
  __syncthreads();
  EL_LOOP(nL) ZDEC;
  Wt[p * LS + l] = ok ? 24. * s.dx3 * (-F.p[b + pm * pl] + F.p[b + p * pl]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s22);
  EL_LOOP(nL) ZDEC;
  if (ok) F.ruh[2][b + p * pl] = F.ruh[2][b + p * pl] + acc[m] - Wt[p * LS + l] - presg3;
  EL_END;
  __syncthreads();
  // WW3 = Iz w (face -> centre), HH3 = D1z(WW3^2) centre -> face
  EL_LOOP(nL) ZDEC;
  Wt[p * LS + l] = ok ? 4. * (F.w[b + p * pl] + F.w[b + pp * pl]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s6);
  EL_LOOP(nL) ZDEC;
  (void)ok; (void)b;
  double a0 = Wt[p * LS + l], am = Wt[pm * LS + l];
  acc[m] = 24. * s.dx3 * (a0 * a0 - am * am);
  EL_END;
  __syncthreads();
  EL_LOOP(nL) ZDEC;
  (void)ok; (void)b; (void)pp;
  Wt[p * LS + l] = acc[m];
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s22);
  EL_LOOP(nL) ZDEC;
  if (ok) F.hh[2][b + p * pl] = Wt[p * LS + l];
  EL_END;
#undef ZDEC
}

// -------------------------------------------------------- RHS x-pass 2
// HH3 += D1x(UU3*WW1) face -> centre (:7734-7769)
__global__ __launch_bounds__(C3D_NT) void k_rhs_x2(Ctx c) {
  extern __shared__ double sm[];
  const Geo g = c.g;
  const TileShape ts = c.o.tx;
  const int n = ts.n, L = ts.nl, LS = ts.ls, nL = n * L;
  double* Wt = sm;
  const int nlines = g.n2m * g.nz;
  const int ln0 = blockIdx.x * L;
  const Scal s = c.s;
  const Fields F = c.d;
  EL_LOOP(nL);
  const int l = e / n, p = e % n;
  const int ln = ln0 + l;
  const bool ok = ln < nlines;
  double d = 0.0;
  if (ok) {
    const long long b = gidx(g, 0, 1 + ln % g.n2m, ln / g.n2m);
    d = 24. * s.dx1 * (-F.uw[b + p] + F.uw[b + wrapp(p, n)]);
  }
  Wt[p * LS + l] = d;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s22);
  EL_LOOP(nL);
  const int l = e / n, p = e % n;
  const int ln = ln0 + l;
  if (ln < nlines) {
    const long long b = gidx(g, 0, 1 + ln % g.n2m, ln / g.n2m);
    F.hh[2][b + p] = F.hh[2][b + p] + Wt[p * LS + l];
  }
  EL_END;
}

// ---------------------------------------------------- RHS pointwise y part
// y-direction viscous terms, pressure gradient and convection of RHS1-3,
// then the AB2 combination and the wall-stress BC:
//   RHS1 :6119-6123, :6255-6267, :6765-6773
//   RHS2 :6840-6863, :7386-7389
//   RHS3 :7457-7463, :7594-7605, :8096-8103
__global__ void k_rhs_y(Ctx c) {
  const Geo g = c.g;
  const long long tot = (long long)g.nx * g.n2m * g.nz;
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= tot) return;
  const int i = t % g.nx;
  const int j = 1 + (int)((t / g.nx) % g.n2m);
  const int k = (int)(t / ((long long)g.nx * g.n2m));
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  const long long b = gidx(g, i, j, k);
  const long long bp = b + g.nx, bm = b - g.nx;
  const long long bN2 = gidx(g, i, g.n2, k), b0 = gidx(g, i, 0, k);
  const int jp = j + 1, jm = j - 1;
  const double fjum = M.fjmu[j], fjup = M.fjpa[j];
  const double rinv = s.rinv;
  const long long t2 = (long long)k * g.nx + i;
  // RHS1
  {
    double ru = F.ruh[0][b] + rinv * (M.hp[j] * F.u[bp] - M.hc[j] * F.u[b] + M.hm[j] * F.u[bm]);
    double u2 = fjup * 1.0 / M.h[jp] * (M.dy[j] / 2. * F.u[bp] + M.dy[jp] / 2. * F.u[b]) +
                (1. - fjup) * F.u[bN2];
    double u1 = fjum * 1.0 / M.h[j] * (M.dy[jm] / 2. * F.u[b] + M.dy[j] / 2. * F.u[bm]) +
                (1. - fjum) * F.u[b0];
    double v2 = (j == g.n2m) ? 0.0 : F.vv1[bp];
    double v1 = (j == 1) ? 0.0 : F.vv1[b];
    double hh = F.hh[0][b] + 1. / M.dy[j] * (v2 * u2 - v1 * u1);
    if (s.files) {
      // D(SGS*(DU/DY))/DY and D(SGS*(DV/DX))/DY (:6551-6680); nux outside
      // J=1..N2M enters only multiplied by a zero wall flag and is taken as 0
      const int im = wrapm(i, g.nx);
      const double n0 = F.nux[b], np1 = (j < g.n2m) ? F.nux[bp] : 0.0, nm1 = (j > 1) ? F.nux[bm] : 0.0;
      const double sgsj2 = fjup * 1. / M.h[jp] * (M.dy[j] / 2. * np1 + M.dy[jp] / 2. * n0) +
                           (1. - fjup) * 0.5 * (F.sgs[bN2] + F.sgs[gidx(g, im, g.n2, k)]);
      const double sgsj1 = fjum * 1. / M.h[j] * (M.dy[jm] / 2. * n0 + M.dy[j] / 2. * nm1) +
                           (1. - fjum) * 0.5 * (F.sgs[b0] + F.sgs[gidx(g, im, 0, k)]);
      const double wfu = 1. - s.ifwall * (1. - fjup) * (s.aly - 1.);
      const double wfd = 1. - s.ifwall * (1. - fjum);
      const double uu22 = (F.u[bp] - F.u[b]) / M.h[jp];
      const double uu11 = (F.u[b] - F.u[bm]) / M.h[j];
      hh = hh - 1. / M.dy[j] * (wfu * sgsj2 * uu22 - wfd * sgsj1 * uu11);
      const double dvp = (j < g.n2m) ? F.dvdx[bp] : 0.0;
      hh = hh - 1.0 / M.dy[j] * (wfu * sgsj2 * dvp - wfd * sgsj1 * F.dvdx[b]);
    }
    F.hh[0][b] = hh;
    double bc_down = -F.tauw[t2] / M.dy[1] * s.ifwall;
    double bc = (1. - fjum) * bc_down;
    F.ruh[0][b] = ru - 3. / 2. * hh + 1. / 2. * F.hhp[0][b] + bc;
  }
  // RHS2 (J = 2..N2M)
  if (j >= 2) {
    double viscos = rinv * (M.dyp[j] * F.v[bp] - M.dyc[j] * F.v[b] + M.dym[j] * F.v[bm]);
    double pressg2 = (F.p[b] - F.p[bm]) / M.h[j];
    double ru = F.ruh[1][b] + (-pressg2 + viscos);
    double v2 = 0.5 * (F.v[bp] + F.v[b]);
    double v1 = 0.5 * (F.v[b] + F.v[bm]);
    double hh = F.hh[1][b] + 1. / M.h[j] * (v2 * v2 - v1 * v1);
    if (s.files) {
      double g2 = 1. / M.dy[j] * (F.v[bp] - F.v[b]);
      double g1 = 1. / M.dy[jm] * (F.v[b] - F.v[bm]);
      double vis22 = 1.0 / M.h[j] * (F.sgs[b] * g2 - F.sgs[bm] * g1);
      hh = hh - 2. * vis22;
    }
    F.hh[1][b] = hh;
    F.ruh[1][b] = ru - 3. / 2. * hh + 1. / 2. * F.hhp[1][b];
  }
  // RHS3
  {
    double ru = F.ruh[2][b] + rinv * (M.hp[j] * F.w[bp] - M.hc[j] * F.w[b] + M.hm[j] * F.w[bm]);
    double w2 = fjup * 1.0 / M.h[jp] * (M.dy[jp] / 2. * F.w[b] + M.dy[j] / 2. * F.w[bp]) +
                (1. - fjup) * F.w[bN2];
    double w1 = fjum * 1.0 / M.h[j] * (M.dy[j] / 2. * F.w[bm] + M.dy[jm] / 2. * F.w[b]) +
                (1. - fjum) * F.w[b0];
    double v2 = (j == g.n2m) ? 0.0 : F.vv3[bp];
    double v1 = (j == 1) ? 0.0 : F.vv3[b];
    double hh = F.hh[2][b] + 1. / M.dy[j] * (v2 * w2 - v1 * w1);
    if (s.files) {
      // D(SGS*(DW/DY))/DY and D(SGS*(DV/DZ))/DY (:7858-7908, :8042-8094)
      const int km = wrapm(k, g.nz);
      const double z0 = F.nuz[b], zp1 = (j < g.n2m) ? F.nuz[bp] : 0.0, zm1 = (j > 1) ? F.nuz[bm] : 0.0;
      const double sgsj1 = fjum * 1. / M.h[j] * (M.dy[jm] / 2. * z0 + M.dy[j] / 2. * zm1) +
                           (1. - fjum) * 0.5 * (F.sgs[b0] + F.sgs[gidx(g, i, 0, km)]);
      const double sgsj2 = fjup * 1.0 / M.h[jp] * (M.dy[j] / 2. * zp1 + M.dy[jp] / 2. * z0) +
                           (1. - fjup) * 0.5 * (F.sgs[bN2] + F.sgs[gidx(g, i, g.n2, km)]);
      const double wfu = 1. - s.ifwall * (1. - fjup) * (s.aly - 1.);
      const double wfd = 1. - s.ifwall * (1. - fjum);
      const double ww22 = 1. / M.h[jp] * (F.w[bp] - F.w[b]);
      const double ww11 = 1. / M.h[j] * (F.w[b] - F.w[bm]);
      hh = hh - 1. / M.dy[j] * (wfu * sgsj2 * ww22 - wfd * sgsj1 * ww11);
      const double dzp = (j < g.n2m) ? F.dvdz[bp] : 0.0;
      hh = hh - 1.0 / M.dy[j] * (wfu * sgsj2 * dzp - wfd * sgsj1 * F.dvdz[b]);
    }
    F.hh[2][b] = hh;
    double bc_down = -F.taub[t2] / M.dy[1] * s.ifwall;
    double bc = (1. - fjum) * bc_down;
    F.ruh[2][b] = ru - 3. / 2. * hh + 1. / 2. * F.hhp[2][b] + bc;
  }
}

// ------------------------------------------------------------- GETUH y
// TDMAI along y (compact3d.f:8630-8659) with the J-only coefficient chains
// (A, BET, GAM) precomputed by the host in the reference's order.  Tile of
// TL consecutive i at one k; blockIdx.y = component.
__global__ __launch_bounds__(C3D_NT) void k_adi_y(Ctx c, int tl) {
  extern __shared__ double sm[];
  const Geo g = c.g;
  const int comp = blockIdx.y;
  const int ns = (comp == 1) ? 2 : 1;
  const int nit = (g.nx + tl - 1) / tl;
  const int it = blockIdx.x % nit;
  const int k = blockIdx.x / nit;
  const Fields F = c.d;
  const Metrics M = c.m;
  const double* R = F.ruh[comp];
  double* X = F.uh[comp];
  const double dt = c.s.dt;
  const int nj = g.n2m + 1;
  const int tot = nj * tl;
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    const int l = e % tl, j = e / tl;
    const int i = it * tl + l;
    double r = 0.0;
    if (i < g.nx && j >= ns) r = R[gidx(g, i, j, k)] * dt;
    sm[j * tl + l] = r;
  }
  __syncthreads();
  if (threadIdx.x < tl) {
    const int l = threadIdx.x;
    const double* A = M.ya[comp];
    const double* BET = M.ybet[comp];
    const double* GAM = M.ygam[comp];
    double x = sm[ns * tl + l] * BET[ns];
    sm[ns * tl + l] = x;
    for (int j = ns + 1; j <= g.n2m; j++) {
      x = (sm[j * tl + l] - A[j] * x) * BET[j];
      sm[j * tl + l] = x;
    }
    for (int j = g.n2m - 1; j >= ns; j--) {
      x = sm[j * tl + l] - x * GAM[j];
      sm[j * tl + l] = x;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    const int l = e % tl, j = e / tl;
    const int i = it * tl + l;
    if (i < g.nx && j >= ns) X[gidx(g, i, j, k)] = sm[j * tl + l];
  }
}

// --------------------------------------------------- GETUH x / z sweeps
// CTDMA1J (x, compact3d.f:8285-8311) / CTDMA3I (z, :8314-8338): solve
// (A, B, A) x = UH(p-1) + 10 UH(p) + UH(p+1); `fin` adds U (UH = U + delta,
// :8610-8624).  blockIdx.y indexes the component list.
struct CompList {
  int n;
  int comp[3];
  int fin[3];
};

template <int DIR>
__global__ __launch_bounds__(C3D_NT) void k_adi_line(Ctx c, CompList cl) {
  extern __shared__ double sm[];
  const Geo g = c.g;
  const TileShape ts = DIR == 1 ? c.o.tx : c.o.tz;
  const int n = ts.n, L = ts.nl, LS = ts.ls, nL = n * L;
  double* Wt = sm;
  const int comp = cl.comp[blockIdx.y];
  const bool fin = cl.fin[blockIdx.y];
  const int js = (comp == 1) ? 2 : 1;
  const Fields F = c.d;
  double* X = F.uh[comp];
  const double* U = comp == 0 ? F.u : (comp == 1 ? F.v : F.w);
  const LineSolve op = DIR == 1 ? c.o.adix : c.o.adiz;
  long long stride;
  int nlines, ln0 = 0, it = 0, jz = 0, nit = 1;
  if (DIR == 1) {
    nlines = g.n2m * g.nz;
    ln0 = blockIdx.x * L;
    stride = 1;
  } else {
    nit = (g.nx + L - 1) / L;
    it = blockIdx.x % nit;
    jz = 1 + blockIdx.x / nit;
    nlines = g.nx;
    stride = g.plane;
  }
  auto decode = [&](int e, int& l, int& p, long long& base, bool& ok) {
    if (DIR == 1) {
      l = e / n;
      p = e % n;
      const int ln = ln0 + l;
      ok = ln < nlines;
      const int j = 1 + (ok ? ln % g.n2m : 0);
      ok = ok && j >= js;
      base = gidx(g, 0, j, ok ? ln / g.n2m : 0);
    } else {
      l = e % L;
      p = e / L;
      const int i = it * L + l;
      ok = i < g.nx && jz >= js;
      base = gidx(g, ok ? i : 0, jz, 0);
    }
  };
  EL_LOOP(nL);
  int l, p;
  long long b;
  bool ok;
  decode(e, l, p, b, ok);
  double d = 0.0;
  if (ok)
    d = X[b + wrapm(p, n) * stride] + 10.0 * X[b + p * stride] + X[b + wrapp(p, n) * stride];
  Wt[p * LS + l] = d;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, op);
  EL_LOOP(nL);
  int l, p;
  long long b;
  bool ok;
  decode(e, l, p, b, ok);
  if (ok) {
    const double x = Wt[p * LS + l];
    X[b + p * stride] = fin ? U[b + p * stride] + x : x;
  }
  EL_END;
}

// ------------------------------------------------------------------ RHSDP
// x-lines: RDP = (FJUP*UH2(J+1) - FJUM*UH2(J))/DY/DT + D1x(UH1)/DT  (:8138-8186)
// (the CBC term is zero: UBC(.,.,2,.) = 0, :978, :986)
__global__ __launch_bounds__(C3D_NT) void k_rhsdp_x(Ctx c) {
  extern __shared__ double sm[];
  const Geo g = c.g;
  const TileShape ts = c.o.tx;
  const int n = ts.n, L = ts.nl, LS = ts.ls, nL = n * L;
  double* Wt = sm;
  const int nlines = g.n2m * g.nz;
  const int ln0 = blockIdx.x * L;
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  EL_LOOP(nL);
  const int l = e / n, p = e % n;
  const int ln = ln0 + l;
  double d = 0.0;
  if (ln < nlines) {
    const long long b = gidx(g, 0, 1 + ln % g.n2m, ln / g.n2m);
    d = 24. * s.dx1 * (F.uh[0][b + wrapp(p, n)] - F.uh[0][b + p]);
  }
  Wt[p * LS + l] = d;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s22);
  EL_LOOP(nL);
  const int l = e / n, p = e % n;
  const int ln = ln0 + l;
  if (ln < nlines) {
    const int j = 1 + ln % g.n2m;
    const long long b = gidx(g, 0, j, ln / g.n2m);
    const double divuh = (M.fjpa[j] * F.uh[1][b + g.nx + p] - M.fjmu[j] * F.uh[1][b + p]) / M.dy[j];
    F.rdp[b + p] = divuh / s.dt + Wt[p * LS + l] / s.dt;
  }
  EL_END;
}

// z-lines: RDP += D1z(UH3)/DT (:8196-8211)
__global__ __launch_bounds__(C3D_NT) void k_rhsdp_z(Ctx c) {
  extern __shared__ double sm[];
  const Geo g = c.g;
  const TileShape ts = c.o.tz;
  const int n = ts.n, L = ts.nl, LS = ts.ls, nL = n * L;
  double* Wt = sm;
  const int nit = (g.nx + L - 1) / L;
  const int it = blockIdx.x % nit;
  const int j = 1 + blockIdx.x / nit;
  const Scal s = c.s;
  const Fields F = c.d;
  const long long pl = g.plane;
  EL_LOOP(nL);
  const int l = e % L, p = e / L;
  const int i = it * L + l;
  double d = 0.0;
  if (i < g.nx) {
    const long long b = gidx(g, i, j, 0);
    d = 24. * s.dx3 * (F.uh[2][b + wrapp(p, n) * pl] - F.uh[2][b + p * pl]);
  }
  Wt[p * LS + l] = d;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s22);
  EL_LOOP(nL);
  const int l = e % L, p = e / L;
  const int i = it * L + l;
  if (i < g.nx) {
    const long long b = gidx(g, i, j, 0) + p * pl;
    F.rdp[b] = F.rdp[b] + Wt[p * LS + l] / s.dt;
  }
  EL_END;
}

// ------------------------------------------------------------------ GETDP
typedef cplx_t<double> zc;

// forward x-DFT of RDP lines (J=1..N2M, K), kept half spectrum kx = 0..NX/2
// (the plane DFT of :9325-9346 factored as x then z)
__global__ __launch_bounds__(C3D_NT) void k_fft_x(Ctx c) {
  extern __shared__ double sm[];
  const Geo g = c.g;
  const FFTPlan f = c.f;
  const int n = g.nx, L = f.tlx, LS = L;
  zc* A = (zc*)sm;
  zc* B = A + n * LS;
  const int nlines = g.n2m * g.nz;
  const int ln0 = blockIdx.x * L;
  const Fields F = c.d;
  for (int e = threadIdx.x; e < n * L; e += blockDim.x) {
    const int l = e / n, p = e % n;
    const int ln = ln0 + l;
    double v = 0.0;
    if (ln < nlines) v = F.rdp[gidx(g, p, 1 + ln % g.n2m, ln / g.n2m)];
    A[p * LS + l] = {v, 0.0};
  }
  __syncthreads();
  zc* R = tile_fft<double>(A, B, n, L, LS, f.radx, f.nradx, (const zc*)f.twxf, -1);
  for (int e = threadIdx.x; e < g.nxh * L; e += blockDim.x) {
    const int l = e / g.nxh, kx = e % g.nxh;
    const int ln = ln0 + l;
    if (ln < nlines) {
      const int jm = ln % g.n2m, k = ln / g.n2m;
      zc v = R[kx * LS + l];
      F.spec[((long long)jm * g.nz + k) * g.nxh + kx] = make_double2(v.x, v.y);
    }
  }
}

// z-DFT of the half spectrum, columns (jm, kx); sign -1 forward, +1 backward
__global__ __launch_bounds__(C3D_NT) void k_fft_z(Ctx c, int sign) {
  extern __shared__ double sm[];
  const Geo g = c.g;
  const FFTPlan f = c.f;
  const int n = g.nz, L = f.tlz, LS = L;
  zc* A = (zc*)sm;
  zc* B = A + n * LS;
  const int nit = (g.nxh + L - 1) / L;
  const int it = blockIdx.x % nit;
  const int jm = blockIdx.x / nit;
  const Fields F = c.d;
  double2* S = F.spec + (long long)jm * g.nz * g.nxh;
  for (int e = threadIdx.x; e < n * L; e += blockDim.x) {
    const int l = e % L, p = e / L;
    const int kx = it * L + l;
    double2 v = make_double2(0.0, 0.0);
    if (kx < g.nxh) v = S[(long long)p * g.nxh + kx];
    A[p * LS + l] = {v.x, v.y};
  }
  __syncthreads();
  zc* R = tile_fft<double>(A, B, n, L, LS, f.radz, f.nradz,
                           (const zc*)(sign < 0 ? f.twzf : f.twzb), sign);
  for (int e = threadIdx.x; e < n * L; e += blockDim.x) {
    const int l = e % L, p = e / L;
    const int kx = it * L + l;
    if (kx < g.nxh) {
      zc v = R[p * LS + l];
      S[(long long)p * g.nxh + kx] = make_double2(v.x, v.y);
    }
  }
}

// Thomas solve in y per Fourier mode (TDMAP, compact3d.f:9352-9406, 9444-9459)
// with the reference's pinning: real part of mode (0,0) has row 1 replaced
// by the identity with zero RHS; imaginary parts of the self-conjugate modes
// (kx in {0, NX/2}) x (kz in {0, NZ/2}) are zero.  Tile: TL consecutive kx at
// one kz, all J in LDS.
__global__ __launch_bounds__(C3D_NT) void k_pois_y(Ctx c) {
  extern __shared__ double sm[];
  const Geo g = c.g;
  const int L = c.f.tly;
  const int nit = (g.nxh + L - 1) / L;
  const int it = blockIdx.x % nit;
  const int kz = blockIdx.x / nit;
  const Fields F = c.d;
  const Metrics M = c.m;
  const int n2m = g.n2m;
  zc* X = (zc*)sm;                       // [jm][l]
  double* GM = (double*)(X + n2m * L);   // [jm][l]
  const long long jstr = (long long)g.nz * g.nxh;
  for (int e = threadIdx.x; e < n2m * L; e += blockDim.x) {
    const int l = e % L, jm = e / L;
    const int kx = it * L + l;
    double2 v = make_double2(0.0, 0.0);
    if (kx < g.nxh) v = F.spec[jm * jstr + (long long)kz * g.nxh + kx];
    X[jm * L + l] = {v.x, v.y};
  }
  __syncthreads();
  if (threadIdx.x < L) {
    const int l = threadIdx.x;
    const int kx = it * L + l;
    if (kx < g.nxh) {
      const int kk = (kz <= g.nz / 2) ? kz : g.nz - kz;  // AK3 is even in kz
      const double ak = M.ak3[kk + 1];
      const double a1 = M.ak1[kx + 1];
      const bool selfconj = (kx == 0 || 2 * kx == g.nx) && (kz == 0 || 2 * kz == g.nz);
      const bool pin = (kx == 0 && kz == 0);
      // TDMAP: one chain (BET, GAM) serves the real and imaginary parts; the
      // pinned row of mode (0,0) is B(1)=1, C(1)=0, R(1)=0
      double bet = pin ? 1.0 : (M.pcj[1] - ak - a1);
      zc r = X[l];
      double xr = (pin ? 0.0 : r.x) / bet;
      double xi = selfconj ? 0.0 : r.y / bet;
      X[l] = {xr, xi};
      for (int jm = 1; jm < n2m; jm++) {
        const int J = jm + 1;
        const double cprev = (pin && J == 2) ? 0.0 : M.ppj[J - 1];
        const double gam = cprev / bet;
        bet = (M.pcj[J] - ak - a1) - M.pmj[J] * gam;
        r = X[jm * L + l];
        xr = (r.x - M.pmj[J] * xr) / bet;
        xi = selfconj ? 0.0 : (r.y - M.pmj[J] * xi) / bet;
        X[jm * L + l] = {xr, xi};
        GM[jm * L + l] = gam;
      }
      for (int jm = n2m - 2; jm >= 0; jm--) {
        const double gm = GM[(jm + 1) * L + l];
        zc xc = X[jm * L + l], xn = X[(jm + 1) * L + l];
        X[jm * L + l] = {xc.x - gm * xn.x, xc.y - gm * xn.y};
      }
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n2m * L; e += blockDim.x) {
    const int l = e % L, jm = e / L;
    const int kx = it * L + l;
    if (kx < g.nxh) {
      zc v = X[jm * L + l];
      F.spec[jm * jstr + (long long)kz * g.nxh + kx] = make_double2(v.x, v.y);
    }
  }
}

// inverse x-DFT from the half spectrum (Hermitian fill, :9419-9425), real part
// / (N1M*N3M) into DP (:9429-9433)
__global__ __launch_bounds__(C3D_NT) void k_fft_xinv(Ctx c) {
  extern __shared__ double sm[];
  const Geo g = c.g;
  const FFTPlan f = c.f;
  const int n = g.nx, L = f.tlx, LS = L;
  zc* A = (zc*)sm;
  zc* B = A + n * LS;
  const int nlines = g.n2m * g.nz;
  const int ln0 = blockIdx.x * L;
  const Fields F = c.d;
  for (int e = threadIdx.x; e < n * L; e += blockDim.x) {
    const int l = e / n, kx = e % n;
    const int ln = ln0 + l;
    zc v = {0.0, 0.0};
    if (ln < nlines) {
      const int jm = ln % g.n2m, k = ln / g.n2m;
      const double2* S = F.spec + ((long long)jm * g.nz + k) * g.nxh;
      if (kx < g.nxh) {
        double2 q = S[kx];
        v = {q.x, q.y};
      } else {
        double2 q = S[n - kx];
        v = {q.x, -q.y};
      }
    }
    A[kx * LS + l] = v;
  }
  __syncthreads();
  zc* R = tile_fft<double>(A, B, n, L, LS, f.radx, f.nradx, (const zc*)f.twxb, +1);
  const double norm = (double)(g.nx * g.nz);
  for (int e = threadIdx.x; e < n * L; e += blockDim.x) {
    const int l = e / n, p = e % n;
    const int ln = ln0 + l;
    if (ln < nlines) F.dp[gidx(g, p, 1 + ln % g.n2m, ln / g.n2m)] = R[p * LS + l].x / norm;
  }
}

// ----------------------------------------------------------------- UPCALC
// flow-rate sums of UPCALC (:1079-1093, :1160-1172), all before any update.
// Fixed grid => fixed reduction tree => bit-repeatable.
#define RED_BLOCKS 1024
__global__ __launch_bounds__(C3D_NT) void k_flow_red(Ctx c) {
  __shared__ double red[4][C3D_NT];
  const Geo g = c.g;
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  const long long tot = (long long)g.nx * g.n2m * g.nz;
  double f1 = 0, d1 = 0, f3 = 0, d3 = 0;
  for (long long t = blockIdx.x * (long long)C3D_NT + threadIdx.x; t < tot;
       t += (long long)RED_BLOCKS * C3D_NT) {
    const int i = t % g.nx;
    const int j = 1 + (int)((t / g.nx) % g.n2m);
    const int k = (int)(t / ((long long)g.nx * g.n2m));
    const long long b = gidx(g, i, j, k);
    const double dy = M.dy[j];
    const double dp = F.dp[b];
    const double dpm = F.dp[gidx(g, wrapm(i, g.nx), j, k)];
    const double dpk = F.dp[gidx(g, i, j, wrapm(k, g.nz))];
    f1 += F.u[b] * dy / s.dx3 / s.dx1;
    d1 += F.uh[0][b] * dy / s.dx3 / s.dx1 - s.dt * (dp - dpm) * dy / s.dx3;
    f3 += F.w[b] * dy / s.dx1 / s.dx3;
    d3 += F.uh[2][b] * dy / s.dx3 / s.dx1 - s.dt * (dp - dpk) * dy / s.dx1;
  }
  const int t = threadIdx.x;
  red[0][t] = f1;
  red[1][t] = d1;
  red[2][t] = f3;
  red[3][t] = d3;
  __syncthreads();
  for (int h = C3D_NT / 2; h > 0; h >>= 1) {
    if (t < h)
      for (int q = 0; q < 4; q++) red[q][t] += red[q][t + h];
    __syncthreads();
  }
  if (t < 4) F.red[(long long)blockIdx.x * 4 + t] = red[t][0];
}

__global__ __launch_bounds__(C3D_NT) void k_flow_fin(Ctx c, int update) {
  __shared__ double red[4][C3D_NT];
  const Fields F = c.d;
  const int t = threadIdx.x;
  double a[4] = {0, 0, 0, 0};
  for (int b = t; b < RED_BLOCKS; b += C3D_NT)
    for (int q = 0; q < 4; q++) a[q] += F.red[(long long)b * 4 + q];
  for (int q = 0; q < 4; q++) red[q][t] = a[q];
  __syncthreads();
  for (int h = C3D_NT / 2; h > 0; h >>= 1) {
    if (t < h)
      for (int q = 0; q < 4; q++) red[q][t] += red[q][t + h];
    __syncthreads();
  }
  if (t == 0) {
    const Scal s = c.s;
    const double flow1 = red[0][0], flow3 = red[2][0];
    const double dmp = (red[1][0] - flow1) / s.vol / s.dt;
    const double dmp3 = (red[3][0] - flow3) / s.vol / s.dt;
    F.sc[SC_FLOW1] = flow1;
    F.sc[SC_FLOW3] = flow3;
    if (update) {
      F.sc[SC_DMPRESG] = dmp;
      F.sc[SC_DMPRESG3] = dmp3;
      F.sc[SC_PRESG] = F.sc[SC_PRESG] + dmp;
      F.sc[SC_PRESG3] = F.sc[SC_PRESG3] + dmp3;
    }
  }
}

// x-lines: U1 = UH1 - DT*D1x(DP) - DT*DMPRESG (:1113-1128); U2 = UH2 -
// DT*(DP(J)-DP(J-1))/H(J) (:1139-1145); with `chk`, the x part of the
// divergence of the new field, (V(J+1)-V(J))/DY + D1x(U) (:9066-9102).
__global__ __launch_bounds__(C3D_NT) void k_upd_x(Ctx c, int chk) {
  extern __shared__ double sm[];
  const Geo g = c.g;
  const TileShape ts = c.o.tx;
  const int n = ts.n, L = ts.nl, LS = ts.ls, nL = n * L;
  double* Wt = sm;
  double* Kt = sm + n * LS;
  const int nlines = g.n2m * g.nz;
  const int ln0 = blockIdx.x * L;
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  const double dmp = F.sc[SC_DMPRESG];
#define UDEC                                  \
  const int l = e / n, p = e % n;             \
  const int ln = ln0 + l;                     \
  const bool ok = ln < nlines;                \
  const int j = 1 + (ok ? ln % g.n2m : 0);    \
  const long long b = gidx(g, 0, j, ok ? ln / g.n2m : 0);
  EL_LOOP(nL) UDEC;
  Wt[p * LS + l] = ok ? 24. * s.dx1 * (-F.dp[b + wrapm(p, n)] + F.dp[b + p]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s22);
  EL_LOOP(nL) UDEC;
  double un = 0.0;
  if (ok) {
    un = F.uh[0][b + p] - s.dt * Wt[p * LS + l] - s.dt * dmp;
    F.u[b + p] = un;
    if (j >= 2) F.v[b + p] = F.uh[1][b + p] - s.dt * (F.dp[b + p] - F.dp[b - g.nx + p]) / M.h[j];
  }
  Kt[p * LS + l] = un;
  EL_END;
  if (!chk) return;
  __syncthreads();
  EL_LOOP(nL) UDEC;
  Wt[p * LS + l] = ok ? 24. * s.dx1 * (Kt[wrapp(p, n) * LS + l] - Kt[p * LS + l]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s22);
  EL_LOOP(nL) UDEC;
  if (ok) {
    // new V at J and J+1 (walls J=1 and J=N2 hold UBC = 0)
    const double vj = (j >= 2) ? F.uh[1][b + p] - s.dt * (F.dp[b + p] - F.dp[b - g.nx + p]) / M.h[j] : 0.0;
    const double vp = (j + 1 <= g.n2m)
                          ? F.uh[1][b + g.nx + p] - s.dt * (F.dp[b + g.nx + p] - F.dp[b + p]) / M.h[j + 1]
                          : 0.0;
    F.div[b + p] = (vp - vj) / M.dy[j] + Wt[p * LS + l];
  }
  EL_END;
#undef UDEC
}

// ghost planes after the update (:1130-1135, :1147-1152, :1210-1215)
__global__ void k_ghosts(Ctx c) {
  const Geo g = c.g;
  const long long n = (long long)g.nx * g.nz;
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = t % g.nx, k = t / g.nx;
  const Fields F = c.d;
  F.u[gidx(g, i, 0, k)] = 0.0;
  F.u[gidx(g, i, g.n2, k)] = F.ubc_u[t];
  F.v[gidx(g, i, 1, k)] = 0.0;
  F.v[gidx(g, i, g.n2, k)] = 0.0;
  F.w[gidx(g, i, 0, k)] = 0.0;
  F.w[gidx(g, i, g.n2, k)] = F.ubc_w[t];
}

// z-lines: U3 = UH3 - DT*D1z(DP) - DT*DMPRESG3 (:1193-1208); P += DP
// (:1218-1230); with `chk`: DIV += D1z(U3), DIVMAX, and the CFL maximum of
// the new field (:9112-9128, :9150-9162).
__global__ __launch_bounds__(C3D_NT) void k_upd_z(Ctx c, int chk) {
  extern __shared__ double sm[];
  __shared__ double red[C3D_NT];
  const Geo g = c.g;
  const TileShape ts = c.o.tz;
  const int n = ts.n, L = ts.nl, LS = ts.ls, nL = n * L;
  double* Wt = sm;
  double* Kt = sm + n * LS;
  const int nit = (g.nx + L - 1) / L;
  const int it = blockIdx.x % nit;
  const int j = 1 + blockIdx.x / nit;
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  const double dmp3 = F.sc[SC_DMPRESG3];
  const long long pl = g.plane;
#define WDEC                              \
  const int l = e % L, p = e / L;         \
  const int i = it * L + l;               \
  const bool ok = i < g.nx;               \
  const long long b = gidx(g, ok ? i : 0, j, 0);
  EL_LOOP(nL) WDEC;
  Wt[p * LS + l] = ok ? 24. * s.dx3 * (-F.dp[b + wrapm(p, n) * pl] + F.dp[b + p * pl]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s22);
  EL_LOOP(nL) WDEC;
  double wn = 0.0;
  if (ok) {
    wn = F.uh[2][b + p * pl] - s.dt * Wt[p * LS + l] - s.dt * dmp3;
    F.w[b + p * pl] = wn;
    F.p[b + p * pl] = F.p[b + p * pl] + F.dp[b + p * pl];
  }
  Kt[p * LS + l] = wn;
  EL_END;
  if (!chk) return;
  __syncthreads();
  EL_LOOP(nL) WDEC;
  Wt[p * LS + l] = ok ? 24. * s.dx3 * (Kt[wrapp(p, n) * LS + l] - Kt[p * LS + l]) : 0.0;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s22);
  double dmax = 0.0, cmax = 0.0;
  EL_LOOP(nL) WDEC;
  if (ok) {
    const long long q = b + p * pl;
    const double dv = F.div[q] + Wt[p * LS + l];
    dmax = (dv != dv || dmax != dmax) ? NAN : fmax(dmax, fabs(dv));
    const double cfll = fabs(F.u[q] + F.u[gidx(g, wrapp(i, g.nx), j, p)]) * 0.5 * s.dx1 +
                        fabs(F.v[q] + F.v[q + g.nx]) * 0.5 / M.dy[j] +
                        fabs(Kt[p * LS + l] + Kt[wrapp(p, n) * LS + l]) * 0.5 * s.dx3;
    cmax = (cfll != cfll || cmax != cmax) ? NAN : fmax(cmax, cfll);
  }
  EL_END;
  dmax = block_max(dmax, red);
  cmax = block_max(cmax, red);
  if (threadIdx.x == 0) {
    atomic_max_pos(&F.sc[SC_DIVMAX], dmax);
    atomic_max_pos(&F.sc[SC_CFLM], cmax);
  }
#undef WDEC
}

// ------------------------------------------------ DIVCHECK / CFL (alone)
// DIVCHECK (:9049-9131) as two passes over the resident U
__global__ __launch_bounds__(C3D_NT) void k_div_x(Ctx c) {
  extern __shared__ double sm[];
  const Geo g = c.g;
  const TileShape ts = c.o.tx;
  const int n = ts.n, L = ts.nl, LS = ts.ls, nL = n * L;
  double* Wt = sm;
  const int nlines = g.n2m * g.nz;
  const int ln0 = blockIdx.x * L;
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  EL_LOOP(nL);
  const int l = e / n, p = e % n;
  const int ln = ln0 + l;
  double d = 0.0;
  if (ln < nlines) {
    const long long b = gidx(g, 0, 1 + ln % g.n2m, ln / g.n2m);
    d = 24. * s.dx1 * (F.u[b + wrapp(p, n)] - F.u[b + p]);
  }
  Wt[p * LS + l] = d;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s22);
  EL_LOOP(nL);
  const int l = e / n, p = e % n;
  const int ln = ln0 + l;
  if (ln < nlines) {
    const int j = 1 + ln % g.n2m;
    const long long b = gidx(g, 0, j, ln / g.n2m);
    F.div[b + p] = (F.v[b + g.nx + p] - F.v[b + p]) / M.dy[j] + Wt[p * LS + l];
  }
  EL_END;
}

__global__ __launch_bounds__(C3D_NT) void k_div_z(Ctx c) {
  extern __shared__ double sm[];
  __shared__ double red[C3D_NT];
  const Geo g = c.g;
  const TileShape ts = c.o.tz;
  const int n = ts.n, L = ts.nl, LS = ts.ls, nL = n * L;
  double* Wt = sm;
  const int nit = (g.nx + L - 1) / L;
  const int it = blockIdx.x % nit;
  const int j = 1 + blockIdx.x / nit;
  const Scal s = c.s;
  const Fields F = c.d;
  const long long pl = g.plane;
  EL_LOOP(nL);
  const int l = e % L, p = e / L;
  const int i = it * L + l;
  double d = 0.0;
  if (i < g.nx) {
    const long long b = gidx(g, i, j, 0);
    d = 24. * s.dx3 * (F.w[b + wrapp(p, n) * pl] - F.w[b + p * pl]);
  }
  Wt[p * LS + l] = d;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, c.o.s22);
  double dmax = 0.0;
  EL_LOOP(nL);
  const int l = e % L, p = e / L;
  const int i = it * L + l;
  if (i < g.nx) {
    const long long q = gidx(g, i, j, 0) + p * pl;
    const double dv = F.div[q] + Wt[p * LS + l];
    F.div[q] = dv;
    dmax = (dv != dv || dmax != dmax) ? NAN : fmax(dmax, fabs(dv));
  }
  EL_END;
  dmax = block_max(dmax, red);
  if (threadIdx.x == 0) atomic_max_pos(&F.sc[SC_DIVMAX], dmax);
}

__global__ __launch_bounds__(C3D_NT) void k_cfl(Ctx c) {
  __shared__ double red[C3D_NT];
  const Geo g = c.g;
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  const long long tot = (long long)g.nx * g.n2m * g.nz;
  double cm = 0.0;
  for (long long t = blockIdx.x * (long long)C3D_NT + threadIdx.x; t < tot;
       t += (long long)gridDim.x * C3D_NT) {
    const int i = t % g.nx;
    const int j = 1 + (int)((t / g.nx) % g.n2m);
    const int k = (int)(t / ((long long)g.nx * g.n2m));
    const long long b = gidx(g, i, j, k);
    const double cfll = fabs(F.u[b] + F.u[gidx(g, wrapp(i, g.nx), j, k)]) * 0.5 * s.dx1 +
                        fabs(F.v[b] + F.v[b + g.nx]) * 0.5 / M.dy[j] +
                        fabs(F.w[b] + F.w[gidx(g, i, j, wrapp(k, g.nz))]) * 0.5 * s.dx3;
    cm = (cfll != cfll || cm != cm) ? NAN : fmax(cm, cfll);
  }
  cm = block_max(cm, red);
  if (threadIdx.x == 0) atomic_max_pos(&F.sc[SC_CFLM], cm);
}

// CHKMF (:9478-9497).  `ref`: the reference's sums over U(1,J,K,1) and
// U(I,J,1,3); otherwise the full-field flow rates.
__global__ __launch_bounds__(C3D_NT) void k_chkmf(Ctx c, int ref) {
  __shared__ double red[2][C3D_NT];
  const Geo g = c.g;
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  const long long tot = (long long)g.nx * g.n2m * g.nz;
  double f1 = 0, f3 = 0;
  for (long long t = blockIdx.x * (long long)C3D_NT + threadIdx.x; t < tot;
       t += (long long)RED_BLOCKS * C3D_NT) {
    const int i = t % g.nx;
    const int j = 1 + (int)((t / g.nx) % g.n2m);
    const int k = (int)(t / ((long long)g.nx * g.n2m));
    f1 += F.u[gidx(g, ref ? 0 : i, j, k)] * M.dy[j] / s.dx3 / s.dx1;
    f3 += F.w[gidx(g, i, j, ref ? 0 : k)] * M.dy[j] / s.dx1 / s.dx3;
  }
  const int t = threadIdx.x;
  red[0][t] = f1;
  red[1][t] = f3;
  __syncthreads();
  for (int h = C3D_NT / 2; h > 0; h >>= 1) {
    if (t < h) {
      red[0][t] += red[0][t + h];
      red[1][t] += red[1][t + h];
    }
    __syncthreads();
  }
  if (t < 2) F.red[(long long)blockIdx.x * 4 + t] = red[t][0];
}

__global__ __launch_bounds__(C3D_NT) void k_chkmf_fin(Ctx c) {
  __shared__ double red[2][C3D_NT];
  const Fields F = c.d;
  const int t = threadIdx.x;
  double a = 0, b2 = 0;
  for (int b = t; b < RED_BLOCKS; b += C3D_NT) {
    a += F.red[(long long)b * 4];
    b2 += F.red[(long long)b * 4 + 1];
  }
  red[0][t] = a;
  red[1][t] = b2;
  __syncthreads();
  for (int h = C3D_NT / 2; h > 0; h >>= 1) {
    if (t < h) {
      red[0][t] += red[0][t + h];
      red[1][t] += red[1][t + h];
    }
    __syncthreads();
  }
  if (t == 0) {
    F.sc[SC_CHK1] = red[0][0] / c.s.vol * c.s.aly;  // /ALX/ALZ
    F.sc[SC_CHK3] = red[1][0] / c.s.vol * c.s.aly;
  }
}

__global__ void k_zero(double* p, long long n) {
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

// --------------------------------------------------------- operator hook
// one compact operator along x (dir 1) or z (dir 3) of a resident array
__global__ __launch_bounds__(C3D_NT) void k_apply_op(Ctx c, int op, int dir, const double* in,
                                                     double* out) {
  extern __shared__ double sm[];
  const Geo g = c.g;
  const TileShape ts = dir == 1 ? c.o.tx : c.o.tz;
  const int n = ts.n, L = ts.nl, LS = ts.ls, nL = n * L;
  double* Wt = sm;
  const Scal s = c.s;
  long long stride;
  int nlines, ln0 = 0, it = 0, jz = 0, nit = 1;
  if (dir == 1) {
    nlines = g.n2m * g.nz;
    ln0 = blockIdx.x * L;
    stride = 1;
  } else {
    nit = (g.nx + L - 1) / L;
    it = blockIdx.x % nit;
    jz = 1 + blockIdx.x / nit;
    nlines = g.nx;
    stride = g.plane;
  }
  const double dx = dir == 1 ? s.dx1 : s.dx3, dxq = dir == 1 ? s.dx1q : s.dx3q;
  const LineSolve ls = op == 0 ? c.o.s10 : (op <= 2 ? c.o.s22 : c.o.s6);
  auto decode = [&](int e, int& l, int& p, long long& base, bool& ok) {
    if (dir == 1) {
      l = e / n;
      p = e % n;
      const int ln = ln0 + l;
      ok = ln < nlines;
      base = gidx(g, 0, 1 + (ok ? ln % g.n2m : 0), ok ? ln / g.n2m : 0);
    } else {
      l = e % L;
      p = e / L;
      const int i = it * L + l;
      ok = i < g.nx;
      base = gidx(g, ok ? i : 0, jz, 0);
    }
  };
  EL_LOOP(nL);
  int l, p;
  long long b;
  bool ok;
  decode(e, l, p, b, ok);
  double d = 0.0;
  if (ok) {
    const double fm = in[b + wrapm(p, n) * stride], f0 = in[b + p * stride],
                 fp = in[b + wrapp(p, n) * stride];
    switch (op) {
      case 0: d = 12. * dxq * (fm - 2. * f0 + fp); break;
      case 1: d = 24. * dx * (-fm + f0); break;
      case 2: d = 24. * dx * (fp - f0); break;
      case 3: d = 4. * (fm + f0); break;
      default: d = 4. * (f0 + fp); break;
    }
  }
  Wt[p * LS + l] = d;
  EL_END;
  __syncthreads();
  tile_solve(Wt, n, L, LS, ts.g, ts.seg, ls);
  EL_LOOP(nL);
  int l, p;
  long long b;
  bool ok;
  decode(e, l, p, b, ok);
  if (ok) out[b + p * stride] = Wt[p * LS + l];
  EL_END;
}

// =================================================================== host
namespace c3dk {

static size_t tile_lds(const TileShape& t, int ntiles) { return sizeof(double) * (size_t)t.n * t.ls * ntiles; }

static int xgrid(const Ctx& c) { return (c.g.n2m * c.g.nz + c.o.tx.nl - 1) / c.o.tx.nl; }
static int zgrid(const Ctx& c) { return ((c.g.nx + c.o.tz.nl - 1) / c.o.tz.nl) * c.g.n2m; }
static int pgrid(long long n) { return (int)((n + C3D_NT - 1) / C3D_NT); }

void init_attrs() {
  // every kernel keeps its dynamic LDS within 64 KB; clear any stale error
  (void)hipGetLastError();
}

void zero(const Ctx& c, double* p, long long n) {
  { KScope ks_(c, KID_ZERO); hipLaunchKernelGGL(k_zero, dim3(pgrid(n)), dim3(C3D_NT), 0, c.st, p, n); }
}

void bcond(const Ctx& c) {
  { KScope ks_(c, KID_BCOND); hipLaunchKernelGGL(k_bcond, dim3(pgrid((long long)c.g.nx * c.g.nz)), dim3(C3D_NT), 0, c.st, c); }
}

void rhs(const Ctx& c) {
  { KScope ks_(c, KID_RHS_X); hipLaunchKernelGGL(k_rhs_x, dim3(xgrid(c)), dim3(C3D_NT), tile_lds(c.o.tx, 2), c.st, c); }
  { KScope ks_(c, KID_RHS_Z); hipLaunchKernelGGL(k_rhs_z, dim3(zgrid(c)), dim3(C3D_NT), tile_lds(c.o.tz, 2), c.st, c); }
  { KScope ks_(c, KID_RHS_X2); hipLaunchKernelGGL(k_rhs_x2, dim3(xgrid(c)), dim3(C3D_NT), tile_lds(c.o.tx, 1), c.st, c); }
  if (c.s.files) sgs_terms(c);
  { KScope ks_(c, KID_RHS_Y); hipLaunchKernelGGL(k_rhs_y, dim3(pgrid((long long)c.g.nx * c.g.n2m * c.g.nz)), dim3(C3D_NT), 0,
                     c.st, c); }
}

void getuh(const Ctx& c) {
  int tl = 32;  // i-columns per y tile; (N2M+1)*tl doubles of LDS <= 64 KB
  while (tl > 1 && sizeof(double) * (size_t)(c.g.n2m + 1) * tl > 65536) tl /= 2;
  const int nit = (c.g.nx + tl - 1) / tl;
  { KScope ks_(c, KID_ADI_Y); hipLaunchKernelGGL(k_adi_y, dim3(nit * c.g.nz, 3), dim3(C3D_NT),
                     sizeof(double) * (size_t)(c.g.n2m + 1) * tl, c.st, c, tl); }
  // GETUH1/2: y, x, z;  GETUH3: y, z, x (compact3d.f:8508-8607)
  CompList x12{2, {0, 1, 0}, {0, 0, 0}};
  { KScope ks_(c, KID_ADI_X); hipLaunchKernelGGL(k_adi_line<1>, dim3(xgrid(c), 2), dim3(C3D_NT), tile_lds(c.o.tx, 1), c.st, c, x12); }
  CompList z123{3, {0, 1, 2}, {1, 1, 0}};
  { KScope ks_(c, KID_ADI_Z); hipLaunchKernelGGL(k_adi_line<3>, dim3(zgrid(c), 3), dim3(C3D_NT), tile_lds(c.o.tz, 1), c.st, c, z123); }
  CompList x3{1, {2, 0, 0}, {1, 0, 0}};
  { KScope ks_(c, KID_ADI_X); hipLaunchKernelGGL(k_adi_line<1>, dim3(xgrid(c), 1), dim3(C3D_NT), tile_lds(c.o.tx, 1), c.st, c, x3); }
}

void rhsdp(const Ctx& c) {
  { KScope ks_(c, KID_RHSDP_X); hipLaunchKernelGGL(k_rhsdp_x, dim3(xgrid(c)), dim3(C3D_NT), tile_lds(c.o.tx, 1), c.st, c); }
  { KScope ks_(c, KID_RHSDP_Z); hipLaunchKernelGGL(k_rhsdp_z, dim3(zgrid(c)), dim3(C3D_NT), tile_lds(c.o.tz, 1), c.st, c); }
}

void getdp(const Ctx& c) {
  const Geo& g = c.g;
  const FFTPlan& f = c.f;
  const size_t ldsx = 2 * sizeof(zc) * (size_t)g.nx * f.tlx;
  const size_t ldsz = 2 * sizeof(zc) * (size_t)g.nz * f.tlz;
  const int gx = (g.n2m * g.nz + f.tlx - 1) / f.tlx;
  const int gz = ((g.nxh + f.tlz - 1) / f.tlz) * g.n2m;
  const int gy = ((g.nxh + f.tly - 1) / f.tly) * g.nz;
  const size_t ldsy = (sizeof(zc) + sizeof(double)) * (size_t)g.n2m * f.tly;
  { KScope ks_(c, KID_FFT_X); hipLaunchKernelGGL(k_fft_x, dim3(gx), dim3(C3D_NT), ldsx, c.st, c); }
  { KScope ks_(c, KID_FFT_Z); hipLaunchKernelGGL(k_fft_z, dim3(gz), dim3(C3D_NT), ldsz, c.st, c, -1); }
  { KScope ks_(c, KID_POIS_Y); hipLaunchKernelGGL(k_pois_y, dim3(gy), dim3(C3D_NT), ldsy, c.st, c); }
  { KScope ks_(c, KID_FFT_Z); hipLaunchKernelGGL(k_fft_z, dim3(gz), dim3(C3D_NT), ldsz, c.st, c, +1); }
  { KScope ks_(c, KID_FFT_XINV); hipLaunchKernelGGL(k_fft_xinv, dim3(gx), dim3(C3D_NT), ldsx, c.st, c); }
}

void upcalc(const Ctx& c, bool chk) {
  if (chk) zero(c, c.d.sc + SC_DIVMAX, 2);
  { KScope ks_(c, KID_FLOW_RED); hipLaunchKernelGGL(k_flow_red, dim3(RED_BLOCKS), dim3(C3D_NT), 0, c.st, c); }
  { KScope ks_(c, KID_FLOW_FIN); hipLaunchKernelGGL(k_flow_fin, dim3(1), dim3(C3D_NT), 0, c.st, c, 1); }
  { KScope ks_(c, KID_UPD_X); hipLaunchKernelGGL(k_upd_x, dim3(xgrid(c)), dim3(C3D_NT), tile_lds(c.o.tx, 2), c.st, c, chk ? 1 : 0); }
  { KScope ks_(c, KID_GHOSTS); hipLaunchKernelGGL(k_ghosts, dim3(pgrid((long long)c.g.nx * c.g.nz)), dim3(C3D_NT), 0, c.st, c); }
  { KScope ks_(c, KID_UPD_Z); hipLaunchKernelGGL(k_upd_z, dim3(zgrid(c)), dim3(C3D_NT), tile_lds(c.o.tz, 2), c.st, c, chk ? 1 : 0); }
}

void divcheck(const Ctx& c) {
  zero(c, c.d.sc + SC_DIVMAX, 1);
  { KScope ks_(c, KID_DIV_X); hipLaunchKernelGGL(k_div_x, dim3(xgrid(c)), dim3(C3D_NT), tile_lds(c.o.tx, 1), c.st, c); }
  { KScope ks_(c, KID_DIV_Z); hipLaunchKernelGGL(k_div_z, dim3(zgrid(c)), dim3(C3D_NT), tile_lds(c.o.tz, 1), c.st, c); }
}

void cfl(const Ctx& c) {
  zero(c, c.d.sc + SC_CFLM, 1);
  { KScope ks_(c, KID_CFL); hipLaunchKernelGGL(k_cfl, dim3(RED_BLOCKS), dim3(C3D_NT), 0, c.st, c); }
}

void chkmf(const Ctx& c, bool ref) {
  { KScope ks_(c, KID_CHKMF); hipLaunchKernelGGL(k_chkmf, dim3(RED_BLOCKS), dim3(C3D_NT), 0, c.st, c, ref ? 1 : 0); }
  { KScope ks_(c, KID_CHKMF); hipLaunchKernelGGL(k_chkmf_fin, dim3(1), dim3(C3D_NT), 0, c.st, c); }
}

void apply_op(const Ctx& c, int op, int dir, const double* in, double* out) {
  if (dir == 1)
    { KScope ks_(c, KID_APPLY); hipLaunchKernelGGL(k_apply_op, dim3(xgrid(c)), dim3(C3D_NT), tile_lds(c.o.tx, 1), c.st, c, op, dir, in,
                       out); }
  else
    { KScope ks_(c, KID_APPLY); hipLaunchKernelGGL(k_apply_op, dim3(zgrid(c)), dim3(C3D_NT), tile_lds(c.o.tz, 1), c.st, c, op, dir, in,
                       out); }
}


}  // namespace c3dk
