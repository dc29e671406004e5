// c3d_yterms.h -- the pointwise y parts of RHS1/2/3 for one component at one
// point: y-direction viscous term, y convection, the SGS y terms when FILES
// is set, then the AB2 combination and the wall-stress boundary term.
//   RHS1 :6119-6123, :6255-6267, :6551-6680, :6765-6773
//   RHS2 :6840-6863, :7386-7389
//   RHS3 :7457-7463, :7594-7605, :7858-7908, :8042-8094, :8096-8103
// On entry RUH and HH hold the x and z parts; yterm_c() stores the final HH
// and returns the final RUH (it does not store RUH).  Reciprocals and
// products of J-only metrics come from the per-J vectors the host formed
// with the same expressions (Metrics: cu1, cu2, dyh, rdy, rh), so each term
// evaluates as the reference's does without a division per point.
#pragma once
#include "c3d_impl.h"

struct YIdx {
  long long b, bp, bm, bN2, b0, t2;
};

__device__ __forceinline__ YIdx yidx(const Geo& g, int i, int j, int k) {
  YIdx x;
  x.b = gidx(g, i, j, k);
  x.bp = x.b + g.nx;
  x.bm = x.b - g.nx;
  x.bN2 = gidx(g, i, g.n2, k);
  x.b0 = gidx(g, i, 0, k);
  x.t2 = (long long)k * g.nx + i;
  return x;
}

// u component (RHS1)
template <typename T>
__device__ __forceinline__ T yterm_u(const Ctx& c, const YIdx& x, int i, int j, int k) {
  const Geo& g = c.g;
  const Metrics& M = c.m;
  const Fields& F = c.d;
  const Scal& s = c.s;
  const T* su = (const T*)F.su;
  const T* vv1 = (const T*)F.vv1;
  const T u0 = su[x.b], up = su[x.bp], um = su[x.bm];
  const T ru = ((const T*)F.ruh[0])[x.b] +
               (T)s.rinv * ((T)M.hp[j] * up - (T)M.hc[j] * u0 + (T)M.hm[j] * um);
  const double fjup = M.fjpa[j], fjum = M.fjmu[j];
  const T u2 = (T)M.cu2[j] * ((T)M.dyh[j] * up + (T)M.dyh[j + 1] * u0) + (T)(1. - fjup) * su[x.bN2];
  const T u1 = (T)M.cu1[j] * ((T)M.dyh[j - 1] * u0 + (T)M.dyh[j] * um) + (T)(1. - fjum) * su[x.b0];
  const T v2 = (j == g.n2m) ? (T)0 : vv1[x.bp];
  const T v1 = (j == 1) ? (T)0 : vv1[x.b];
  T hh = ((const T*)F.hh[0])[x.b] + (T)M.rdy[j] * (v2 * u2 - v1 * u1);
  if (s.files) {
    // D(SGS*(DU/DY))/DY and D(SGS*(DV/DX))/DY (:6551-6680); nux outside
    // J=1..N2M enters only multiplied by a zero wall flag and is taken as 0
    const int jp = j + 1, jm = j - 1;
    const int im = wrapm(i, g.nx);
    const double n0 = F.nux[x.b], np1 = (j < g.n2m) ? F.nux[x.bp] : 0.0, nm1 = (j > 1) ? F.nux[x.bm] : 0.0;
    const double sgsj2 = M.cu2[j] * (M.dyh[j] * np1 + M.dyh[jp] * n0) +
                         (1. - fjup) * 0.5 * (F.sgs[x.bN2] + F.sgs[gidx(g, im, g.n2, k)]);
    const double sgsj1 = M.cu1[j] * (M.dyh[jm] * n0 + M.dyh[j] * nm1) +
                         (1. - fjum) * 0.5 * (F.sgs[x.b0] + F.sgs[gidx(g, im, 0, k)]);
    const double wfu = 1. - s.ifwall * (1. - fjup) * (s.aly - 1.);
    const double wfd = 1. - s.ifwall * (1. - fjum);
    const double uu22 = (F.u[x.bp] - F.u[x.b]) / M.h[jp];
    const double uu11 = (F.u[x.b] - F.u[x.bm]) / M.h[j];
    double hd = (double)hh - M.rdy[j] * (wfu * sgsj2 * uu22 - wfd * sgsj1 * uu11);
    const double dvp = (j < g.n2m) ? F.dvdx[x.bp] : 0.0;
    hd = hd - M.rdy[j] * (wfu * sgsj2 * dvp - wfd * sgsj1 * F.dvdx[x.b]);
    hh = (T)hd;
  }
  ((T*)F.hh[0])[x.b] = hh;
  // wall stresses of the bottom and the top wall (:6765-6768): the flags
  // make the term zero except at J = 1 and J = N2M; the top layer of TAUW is
  // zero unless a two-wall model set it
  T bc = (T)0;
  if (j == 1) bc = (T)(1. - fjum) * (-(T)F.tauw[x.t2] / (T)M.dy[1] * (T)s.ifwall);
  if (j == g.n2m)
    bc = bc + (T)(1. - fjup) * (-(T)F.tauw[(unsigned)(g.nx * g.nz) + x.t2] / (T)M.dy[g.n2m] * (T)s.ifwall *
                               (T)(s.aly - 1.));
  return ru - (T)(3. / 2.) * hh + (T)(1. / 2.) * ((const T*)F.hhp[0])[x.b] + bc;
}

// v component (RHS2), J = 2..N2M
template <typename T>
__device__ __forceinline__ T yterm_v(const Ctx& c, const YIdx& x, int i, int j, int k) {
  const Metrics& M = c.m;
  const Fields& F = c.d;
  const Scal& s = c.s;
  const T* sv = (const T*)F.sv;
  const T* sp = (const T*)F.sp;
  const T v0 = sv[x.b], vp = sv[x.bp], vm = sv[x.bm];
  const T viscos = (T)s.rinv * ((T)M.dyp[j] * vp - (T)M.dyc[j] * v0 + (T)M.dym[j] * vm);
  const T pressg2 = (sp[x.b] - sp[x.bm]) * (T)M.rh[j];
  const T ru = ((const T*)F.ruh[1])[x.b] + (-pressg2 + viscos);
  const T v2 = (T)0.5 * (vp + v0);
  const T v1 = (T)0.5 * (v0 + vm);
  T hh = ((const T*)F.hh[1])[x.b] + (T)M.rh[j] * (v2 * v2 - v1 * v1);
  if (s.files) {
    const int jm = j - 1;
    const double g2 = M.rdy[j] * (F.v[x.bp] - F.v[x.b]);
    const double g1 = M.rdy[jm] * (F.v[x.b] - F.v[x.bm]);
    const double vis22 = M.rh[j] * (F.sgs[x.b] * g2 - F.sgs[x.bm] * g1);
    hh = (T)((double)hh - 2. * vis22);
  }
  ((T*)F.hh[1])[x.b] = hh;
  return ru - (T)(3. / 2.) * hh + (T)(1. / 2.) * ((const T*)F.hhp[1])[x.b];
}

// w component (RHS3)
template <typename T>
__device__ __forceinline__ T yterm_w(const Ctx& c, const YIdx& x, int i, int j, int k) {
  const Geo& g = c.g;
  const Metrics& M = c.m;
  const Fields& F = c.d;
  const Scal& s = c.s;
  const T* sw = (const T*)F.sw;
  const T* vv3 = (const T*)F.vv3;
  const T w0 = sw[x.b], wp = sw[x.bp], wm = sw[x.bm];
  const T ru = ((const T*)F.ruh[2])[x.b] +
               (T)s.rinv * ((T)M.hp[j] * wp - (T)M.hc[j] * w0 + (T)M.hm[j] * wm);
  const double fjup = M.fjpa[j], fjum = M.fjmu[j];
  const T w2 = (T)M.cu2[j] * ((T)M.dyh[j + 1] * w0 + (T)M.dyh[j] * wp) + (T)(1. - fjup) * sw[x.bN2];
  const T w1 = (T)M.cu1[j] * ((T)M.dyh[j] * wm + (T)M.dyh[j - 1] * w0) + (T)(1. - fjum) * sw[x.b0];
  const T v2 = (j == g.n2m) ? (T)0 : vv3[x.bp];
  const T v1 = (j == 1) ? (T)0 : vv3[x.b];
  T hh = ((const T*)F.hh[2])[x.b] + (T)M.rdy[j] * (v2 * w2 - v1 * w1);
  if (s.files) {
    // D(SGS*(DW/DY))/DY and D(SGS*(DV/DZ))/DY (:7858-7908, :8042-8094)
    const int jp = j + 1, jm = j - 1;
    const int km = wrapm(k, g.nz);
    const double z0 = F.nuz[x.b], zp1 = (j < g.n2m) ? F.nuz[x.bp] : 0.0, zm1 = (j > 1) ? F.nuz[x.bm] : 0.0;
    const double sgsj1 = M.cu1[j] * (M.dyh[jm] * z0 + M.dyh[j] * zm1) +
                         (1. - fjum) * 0.5 * (F.sgs[x.b0] + F.sgs[gidx(g, i, 0, km)]);
    const double sgsj2 = M.cu2[j] * (M.dyh[j] * zp1 + M.dyh[jp] * z0) +
                         (1. - fjup) * 0.5 * (F.sgs[x.bN2] + F.sgs[gidx(g, i, g.n2, km)]);
    const double wfu = 1. - s.ifwall * (1. - fjup) * (s.aly - 1.);
    const double wfd = 1. - s.ifwall * (1. - fjum);
    const double ww22 = M.rh[jp] * (F.w[x.bp] - F.w[x.b]);
    const double ww11 = M.rh[j] * (F.w[x.b] - F.w[x.bm]);
    double hd = (double)hh - M.rdy[j] * (wfu * sgsj2 * ww22 - wfd * sgsj1 * ww11);
    const double dzp = (j < g.n2m) ? F.dvdz[x.bp] : 0.0;
    hd = hd - M.rdy[j] * (wfu * sgsj2 * dzp - wfd * sgsj1 * F.dvdz[x.b]);
    hh = (T)hd;
  }
  ((T*)F.hh[2])[x.b] = hh;
  T bc = (T)0;
  if (j == 1) bc = (T)(1. - fjum) * (-(T)F.taub[x.t2] / (T)M.dy[1] * (T)s.ifwall);
  if (j == g.n2m)
    bc = bc + (T)(1. - fjup) * (-(T)F.taub[(unsigned)(g.nx * g.nz) + x.t2] / (T)M.dy[g.n2m] * (T)s.ifwall *
                               (T)(s.aly - 1.));
  return ru - (T)(3. / 2.) * hh + (T)(1. / 2.) * ((const T*)F.hhp[2])[x.b] + bc;
}
