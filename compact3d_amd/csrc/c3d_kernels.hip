// c3d_kernels.hip -- pointwise kernels of the step and the stand-alone
// checks of the compact3d hot path for gfx950 (FP64).
//
//   k_bcond                 BCOND                                 :963-1001
//   k_rhs_y                 RHS1/2/3 pointwise y parts, AB2, wall BC :6063-8114
//   k_div_x/k_div_z, k_cfl  DIVCHECK and CFL on the resident field  :9049-9169
//   k_chkmf                 CHKMF                                   :9463-9503
//   k_apply_op              one compact operator (test hook)
// The line-solve passes of the step are in c3d_step.hip.  The stand-alone
// checks and the LES kernels solve their compact operators with
// tile_solve() (c3d_dev.h) on LDS tiles.
#include <math.h>

#include <algorithm>

#include "c3d_impl.h"
#include "c3d_yterms.h"

// blocks of the grid-stride reductions (fixed, so every sum has a fixed order)
#define RED_BLOCKS 1024

// ------------------------------------------------------------------ BCOND
template <bool CB_>
__global__ void k_bcond(CtxArg<CB_> ca_) {
  const Ctx c = ctx_get(ca_);
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
  if (c.s.mixed) {
    F.uf[b0] = 0.f;
    F.vf[b0] = 0.f;
    F.wf[b0] = 0.f;
    F.uf[bn] = (float)ut;
    F.vf[bn] = 0.f;
    F.wf[bn] = (float)wt;
  }
  if (c.s.files) {
    F.sgs[b0] = 0.0;
    F.sgs[bn] = F.sgs[bm] * (2.0 - aly);
    if (F.sgsf) {
      F.sgsf[b0] = 0.f;
      F.sgsf[bn] = (float)F.sgs[bn];
    }
  }
}

// ---------------------------------------------------- RHS pointwise y part
// the y terms, AB2 and wall BC of RHS1-3 (c3d_yterms.h) as a pass of its own;
// the step fuses them into the ADI y-sweep instead (c3d_step.hip).  Grid
// (i-block, J, K-block); each thread does YK consecutive K of one (I, J).
#define YK 4
template <typename T>
__global__ __launch_bounds__(256) void k_rhs_y(Ctx c) {
  const Geo g = c.g;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.nx) return;
  const int j = 1 + blockIdx.y;
  for (int kk = 0; kk < YK; kk++) {
    const int k = blockIdx.z * YK + kk;
    if (k >= g.nz) break;
    const YIdx x = yidx(g, i, j, k);
    ((T*)c.d.ruh[0])[x.b] = yterm_u<T>(c, x, i, j, k);
    if (j >= 2) ((T*)c.d.ruh[1])[x.b] = yterm_v<T>(c, x, i, j, k);
    ((T*)c.d.ruh[2])[x.b] = yterm_w<T>(c, x, i, j, k);
  }
}

// ------------------------------------------------------------ statistics
// The plane statistics of ENERGY (compact3d.f:9902): the cell-centre
// vorticity of loop 10 (:10005-10069) and the means and moments of loop 20
// (:10073-10158), one block per J plane; thread t sums the points
// t, t+blockDim, ... of the plane, then a tree over the block (fixed
// order).  out[s*N2M + J-1] for the 20 quantities: VM1-3, PM, WM1-3,
// VRMS1-6 (uu vv ww uv uw vw), PRMS, PV, WRMS1-3, U2V, VW2.
#define ST_NQ 20
__global__ __launch_bounds__(256) void k_plane_stats(Ctx c, double* out) {
  __shared__ double red[ST_NQ][256];
  const Geo g = c.g;
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  const int j = 1 + blockIdx.x;
  const int jp = j + 1, jm = j - 1, N2 = g.n2;
  const double fjum = M.fjmu[j], fjup = M.fjpa[j];
  const double* DY = M.dy;
  const double* H = M.h;
  double acc[ST_NQ];
#pragma unroll
  for (int q = 0; q < ST_NQ; q++) acc[q] = 0.0;
  const int npl = g.nx * g.nz;
  for (int e = threadIdx.x; e < npl; e += blockDim.x) {
    const int i = e % g.nx, k = e / g.nx;
    const int ip = wrapp(i, g.nx), im = wrapm(i, g.nx), kp = wrapp(k, g.nz), km = wrapm(k, g.nz);
    auto U = [&](int ii, int jj, int kk) { return F.u[gidx(g, ii, jj, kk)]; };
    auto V = [&](int ii, int jj, int kk) { return F.v[gidx(g, ii, jj, kk)]; };
    auto W = [&](int ii, int jj, int kk) { return F.w[gidx(g, ii, jj, kk)]; };
    const double u1kp = 0.5 * (U(i, j, kp) + U(ip, j, kp));
    const double u1km = 0.5 * (U(i, j, km) + U(ip, j, km));
    const double u2kp = 0.5 * (V(i, j, kp) + V(i, jp, kp));
    const double u2km = 0.5 * (V(i, j, km) + V(i, jp, km));
    const double u2ip = 0.5 * (V(ip, j, k) + V(ip, jp, k));
    const double u2im = 0.5 * (V(im, j, k) + V(im, jp, k));
    const double u3ip = 0.5 * (W(ip, j, k) + W(ip, j, kp));
    const double u3im = 0.5 * (W(im, j, k) + W(im, j, kp));
    const double u1jp = 0.5 * (U(i, jp, k) + U(ip, jp, k));
    const double u1jc = 0.5 * (U(i, j, k) + U(ip, j, k));
    const double u1jm = 0.5 * (U(i, jm, k) + U(ip, jm, k));
    double u12 = 0.5 / H[jp] * (DY[jp] * u1jc + DY[j] * u1jp);
    double u11 = 0.5 / H[j] * (DY[j] * u1jm + DY[jm] * u1jc);
    u12 = u12 * fjup + (1. - fjup) * (U(i, N2, k) + U(ip, N2, k)) * 0.5;
    u11 = u11 * fjum + (1. - fjum) * (U(i, 0, k) + U(ip, 0, k)) * 0.5;
    const double u3jp = 0.5 * (W(i, jp, k) + W(i, jp, kp));
    const double u3jc = 0.5 * (W(i, j, k) + W(i, j, kp));
    const double u3jm = 0.5 * (W(i, jm, k) + W(i, jm, kp));
    double u32 = 0.5 / H[jp] * (DY[jp] * u3jc + DY[j] * u3jp);
    double u31 = 0.5 / H[j] * (DY[j] * u3jm + DY[jm] * u3jc);
    u32 = u32 * fjup + (1. - fjup) * (W(i, N2, k) + W(i, N2, kp)) * 0.5;
    u31 = u31 * fjum + (1. - fjum) * (W(i, 0, k) + W(i, 0, kp)) * 0.5;
    const double dv3dx2 = (u32 - u31) / DY[j];
    const double dv2dx3 = (u2kp - u2km) * 0.5 * s.dx3;
    const double dv3dx1 = (u3ip - u3im) * 0.5 * s.dx1;
    const double dv1dx3 = (u1kp - u1km) * 0.5 * s.dx3;
    const double dv2dx1 = (u2ip - u2im) * 0.5 * s.dx1;
    const double dv1dx2 = (u12 - u11) / DY[j];
    const double w1 = dv3dx2 - dv2dx3, w2 = dv1dx3 - dv3dx1, w3 = dv2dx1 - dv1dx2;
    const double v1 = (U(i, j, k) + U(ip, j, k)) * 0.5;
    const double v2 = (V(i, j, k) + V(i, jp, k)) * 0.5;
    const double v3 = (W(i, j, k) + W(i, j, kp)) * 0.5;
    const double p = F.p[gidx(g, i, j, k)];
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
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < ST_NQ; q++) red[q][t] = acc[q];
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (t < h)
#pragma unroll
      for (int q = 0; q < ST_NQ; q++) red[q][t] += red[q][t + h];
    __syncthreads();
  }
  const double nxz = 1.0 / (double)npl;
  if (t < ST_NQ) out[(long long)t * g.n2m + (j - 1)] = red[t][0] * nxz;
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

// precision conversions of whole fields (set-up and host transfers)
__global__ void k_to_f32(const double* in, float* out, long long n) {
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t < n) out[t] = (float)in[t];
}
__global__ void k_to_f64(const float* in, double* out, long long n) {
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t < n) out[t] = (double)in[t];
}
__global__ void k_add_f32(const double* a, const float* b, double* out, long long n) {
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t < n) out[t] = a[t] + (double)b[t];
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
  fft_attrs();
  // clear any stale error
  (void)hipGetLastError();
}

void to_f32(const Ctx& c, const double* in, float* out, long long n) {
  hipLaunchKernelGGL(k_to_f32, dim3(pgrid(n)), dim3(C3D_NT), 0, c.st, in, out, n);
}
void to_f64(const Ctx& c, const float* in, double* out, long long n) {
  hipLaunchKernelGGL(k_to_f64, dim3(pgrid(n)), dim3(C3D_NT), 0, c.st, in, out, n);
}
void add_f32(const Ctx& c, const double* a, const float* b, double* out, long long n) {
  hipLaunchKernelGGL(k_add_f32, dim3(pgrid(n)), dim3(C3D_NT), 0, c.st, a, b, out, n);
}

void zero(const Ctx& c, double* p, long long n) {
  { KScope ks_(c, KID_ZERO); hipLaunchKernelGGL(k_zero, dim3(pgrid(n)), dim3(C3D_NT), 0, c.st, p, n); }
}

void bcond(const Ctx& c) {
  { KScope ks_(c, KID_BCOND); C3D_LAUNCH0(k_bcond, dim3(pgrid((long long)c.g.nx * c.g.nz)), dim3(C3D_NT), 0, c); }
}

void rhs_xz(const Ctx& c) {
  rhs_lines(c);
  if (c.s.files) sgs_terms(c);
}

void rhs(const Ctx& c) {
  rhs_xz(c);
  {
    const int bx = std::min(256, c.g.nx);
    KScope ks_(c, KID_RHS_Y);
    const dim3 grid((c.g.nx + bx - 1) / bx, c.g.n2m, (c.g.nz + YK - 1) / YK);
    if (c.s.mixed)
      hipLaunchKernelGGL(k_rhs_y<float>, grid, dim3(bx), 0, c.st, c);
    else
      hipLaunchKernelGGL(k_rhs_y<double>, grid, dim3(bx), 0, c.st, c);
  }
}

void upcalc(const Ctx& c, bool chk) {
  // mixed mode always forms DIV of the new field: the next RDP starts from it
  chk = chk || c.s.mixed;
  flow_fin(c);
  upd_x(c, chk);  // also the ghost planes
  upd_z(c, chk);
}

void plane_stats(const Ctx& c, double* out) {
  KScope ks_(c, KID_CHKMF);
  hipLaunchKernelGGL(k_plane_stats, dim3(c.g.n2m), dim3(256), 0, c.st, c, out);
}

// WRITEUP's cell-centre values (compact3d.f:796-808): quantity q (0 UF,
// 1 VF, 2 WF, 3 P) of every cell, I fastest, then J, then K.  VF averages
// V(J) with V(JPA(J)), and JPA(N2M) = N2M (:473), so the top row is V(N2M)
// itself.  TERMP = 1 (:794).
__global__ __launch_bounds__(C3D_NT) void k_cellcentre(Ctx c, int q, double* out) {
  const Geo g = c.g;
  const Fields F = c.d;
  const long long n = (long long)g.nx * g.n2m * g.nz;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e % g.nx);
    const long long r = e / g.nx;
    const int j = 1 + (int)(r % g.n2m), k = (int)(r / g.n2m);
    double v;
    if (q == 0) {
      v = 0.5 * (F.u[gidx(g, i, j, k)] + F.u[gidx(g, wrapp(i, g.nx), j, k)]);
    } else if (q == 1) {
      const int jp = j < g.n2m ? j + 1 : g.n2m;
      v = 0.5 * (F.v[gidx(g, i, j, k)] + F.v[gidx(g, i, jp, k)]);
    } else if (q == 2) {
      v = 0.5 * (F.w[gidx(g, i, j, k)] + F.w[gidx(g, i, j, wrapp(k, g.nz))]);
    } else {
      v = F.p[gidx(g, i, j, k)];
    }
    out[e] = v;
  }
}

void cellcentre(const Ctx& c, int q, double* out) {
  KScope ks_(c, KID_CHKMF);
  hipLaunchKernelGGL(k_cellcentre, dim3(pgrid((long long)c.g.nx * c.g.n2m * c.g.nz)), dim3(C3D_NT), 0, c.st, c,
                     q, out);
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
