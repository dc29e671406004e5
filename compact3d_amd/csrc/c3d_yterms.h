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
// evaluates as the reference's does without a division per point.  The SGS
// terms use the compute type T too: in mixed mode the FP32 shadows of the
// state and of SGSVIS and the FP32 SGS intermediates.
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
    const T* nux = (const T*)F.nux;
    const T* sg = (const T*)F.ssgs;
    const T* dvdx = (const T*)F.dvdx;
    const T n0 = nux[x.b], np1 = (j < g.n2m) ? nux[x.bp] : (T)0, nm1 = (j > 1) ? nux[x.bm] : (T)0;
    const T sgsj2 = (T)M.cu2[j] * ((T)M.dyh[j] * np1 + (T)M.dyh[jp] * n0) +
                    (T)(1. - fjup) * (T)0.5 * (sg[x.bN2] + sg[gidx(g, im, g.n2, k)]);
    const T sgsj1 = (T)M.cu1[j] * ((T)M.dyh[jm] * n0 + (T)M.dyh[j] * nm1) +
                    (T)(1. - fjum) * (T)0.5 * (sg[x.b0] + sg[gidx(g, im, 0, k)]);
    const T wfu = (T)(1. - s.ifwall * (1. - fjup) * (s.aly - 1.));
    const T wfd = (T)(1. - s.ifwall * (1. - fjum));
    const T uu22 = (up - u0) / (T)M.h[jp];
    const T uu11 = (u0 - um) / (T)M.h[j];
    const T rdy = (T)M.rdy[j];
    T hd = hh - rdy * (wfu * sgsj2 * uu22 - wfd * sgsj1 * uu11);
    const T dvp = (j < g.n2m) ? dvdx[x.bp] : (T)0;
    hh = hd - rdy * (wfu * sgsj2 * dvp - wfd * sgsj1 * dvdx[x.b]);
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
    const T* sg = (const T*)F.ssgs;
    const T g2 = (T)M.rdy[j] * (vp - v0);
    const T g1 = (T)M.rdy[jm] * (v0 - vm);
    const T vis22 = (T)M.rh[j] * (sg[x.b] * g2 - sg[x.bm] * g1);
    hh = hh - (T)2. * vis22;
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
    const T* nuz = (const T*)F.nuz;
    const T* sg = (const T*)F.ssgs;
    const T* dvdz = (const T*)F.dvdz;
    const T z0 = nuz[x.b], zp1 = (j < g.n2m) ? nuz[x.bp] : (T)0, zm1 = (j > 1) ? nuz[x.bm] : (T)0;
    const T sgsj1 = (T)M.cu1[j] * ((T)M.dyh[jm] * z0 + (T)M.dyh[j] * zm1) +
                    (T)(1. - fjum) * (T)0.5 * (sg[x.b0] + sg[gidx(g, i, 0, km)]);
    const T sgsj2 = (T)M.cu2[j] * ((T)M.dyh[j] * zp1 + (T)M.dyh[jp] * z0) +
                    (T)(1. - fjup) * (T)0.5 * (sg[x.bN2] + sg[gidx(g, i, g.n2, km)]);
    const T wfu = (T)(1. - s.ifwall * (1. - fjup) * (s.aly - 1.));
    const T wfd = (T)(1. - s.ifwall * (1. - fjum));
    const T ww22 = (T)M.rh[jp] * (wp - w0);
    const T ww11 = (T)M.rh[j] * (w0 - wm);
    const T rdy = (T)M.rdy[j];
    T hd = hh - rdy * (wfu * sgsj2 * ww22 - wfd * sgsj1 * ww11);
    const T dzp = (j < g.n2m) ? dvdz[x.bp] : (T)0;
    hh = hd - rdy * (wfu * sgsj2 * dzp - wfd * sgsj1 * dvdz[x.b]);
  }
  ((T*)F.hh[2])[x.b] = hh;
  T bc = (T)0;
  if (j == 1) bc = (T)(1. - fjum) * (-(T)F.taub[x.t2] / (T)M.dy[1] * (T)s.ifwall);
  if (j == g.n2m)
    bc = bc + (T)(1. - fjup) * (-(T)F.taub[(unsigned)(g.nx * g.nz) + x.t2] / (T)M.dy[g.n2m] * (T)s.ifwall *
                               (T)(s.aly - 1.));
  return ru - (T)(3. / 2.) * hh + (T)(1. / 2.) * ((const T*)F.hhp[2])[x.b] + bc;
}
