// c3d_step.hip -- the time step of the channel solver on gfx950 (GETUP,
// compact3d.f:879-960), written on the register line engine of c3d_dev.h.
//
// Pass structure of one step:
//   k_rhs_x, k_rhs_z, k_rhs_x2   RHS1/2/3 x- and z-line terms        :6063-8114
//   k_rhs_y (c3d_kernels.hip)    pointwise y terms, AB2, wall BC
//   k_adi_y, k_adi_x, k_adi_z    GETUH1/2/3 (TDMAI, CTDMA1J/3I)      :8218-8627
//   k_rhsdp_x, k_rhsdp_z         RHSDP                               :8117-8214
//   (mixed precision: k_adi_y, k_adi_z, then k_adi_xr and k_adi_zr,
//    the last sweeps, which also form RHSDP)
//   k_fft_x, k_fft_z, k_pois_y,  GETDP: real-to-complex x-DFT (two  :9271-9441
//     k_fft_z, k_fft_xinv          rows per complex transform), z-DFT, one
//                                  partitioned Thomas solve per mode (FP64),
//                                  inverse z- and x-DFTs (FP64)
//   k_flow_fin, k_upd_x (and     UPCALC (+ DIVCHECK and CFL fused)   :1052-1238
//     the ghost planes), k_upd_z
//
// Thread layouts.  z-lines: a tile is L consecutive i at one j; thread
// (l, g) keeps points k in [g*seg, (g+1)*seg) of column i = i0+l in
// registers and reads/writes HBM directly (16 lanes = one 128-byte row
// piece).  x-lines: a tile is L consecutive (j,k) rows; rows are staged
// through LDS with coalesced row loads, then each thread takes its segment.
// y-lines (ADI TDMAI, Poisson TDMAP): thread (l, c) owns chunk c of column
// l; the recurrences are split into chunk-local solves plus an exact
// carry pass (partitioned Thomas).
#include <math.h>

#include "c3d_impl.h"
#include "c3d_lines.h"
#include "c3d_yterms.h"

// ------------------------------------------------------------ RHS x-pass
// per x-line (j,k), j = 1..N2M:
//   RUH2  = (1/RE) D2x v                         (j>=2)   (:6879-6895)
//   VV1   = Ix v                 (stored)                  (:6228-6242, :6946-6959)
//   HH2   = D1x(UU2 * VV1)                       (j>=2)   (:6961-7004)
//   RUH1  = (1/RE) D2x u - D1x(P) - PRESG                 (:6142-6157, :6196-6210)
//   HH1   = D1x((Ix u)^2)                                  (:6295-6334)
//   RUH3  = (1/RE) D2x w                                   (:7483-7498)
//   WW1   = Ix w                 (stored)                  (:6380-6394, :7718-7732)
// With SGS (FILES != 0) the x-line parts of the SGS terms join, from the
// same staged tiles: nux = Ix(SGSVIS), dvdx = D1x(v), dwdx = D1x(w) (the
// intermediates of the other SGS line solves) and HH1 -= 2 D1x(SGSVIS *
// D1x(u)) (:6445-6466), which HH1 takes before it is stored.
template <typename T, int SEG, bool SGS, bool CB_>
__global__ __launch_bounds__(256) void k_rhs_x(CtxArg<CB_> ca_) {
  const Ctx c = ctx_get(ca_);
  XSETUP;
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  const PSolve s6 = c.o.px.s6, s10 = c.o.px.s10, s22 = c.o.px.s22;
  const T dx1 = (T)s.dx1, dx1q = (T)s.dx1q, rinv = (T)s.rinv;
  const bool j2 = j >= 2;
  T d[SEG], acc[SEG], vv[SEG], pf[SEG];
  T sv[SGS ? SEG : 1], e[SGS ? SEG : 1];  // SGSVIS of the segment; the SGS solves
  const T* B;
  if constexpr (SGS) {
    // SGSVIS: nux = Ix(SGSVIS), centre -> x-face
    X.pre((const T*)F.ssgs, 0, pf);
    B = X.put(pf);
    X.pre((const T*)F.sv, 0, pf);
    SQ sv[q] = XA(B, q);
    SQ d[q] = (T)4. * (XA(B, q) + XA(B, q - 1));
    psolve<T, SEG>(d, G, L, l, gs, s6, bf.xf, bf.xb);
    X.out((T*)F.nux, d);
  } else {
    X.pre((const T*)F.sv, 0, pf);
  }
  // v: RUH2, VV1 (and dvdx)
  B = X.put(pf);
  X.pre((const T*)F.su, -1, pf);  // next tile in flight during the v work
  SQ d[q] = (T)12. * dx1q * (XA(B, q - 1) - (T)2. * XA(B, q) + XA(B, q + 1));
  psolve<T, SEG>(d, G, L, l, gs, s10, bf.xf, bf.xb);
  SQ d[q] = j2 ? rinv * d[q] : (T)0;
  X.out((T*)F.ruh[1], d);
  SQ vv[q] = (T)4. * (XA(B, q - 1) + XA(B, q));
  if constexpr (SGS) SQ e[q] = (T)24. * dx1 * (XA(B, q) - XA(B, q - 1));
  psolve<T, SEG>(vv, G, L, l, gs, s6, bf.xf, bf.xb);
  X.out((T*)F.vv1, vv);
  if constexpr (SGS) {
    psolve<T, SEG>(e, G, L, l, gs, s22, bf.xf, bf.xb);
    X.out((T*)F.dvdx, e);
  }
  // UU2 = (DY(J)/2 U(J-1) + DY(J-1)/2 U(J)) / H(J)  (:6966); HH2 = D1x(UU2 VV1)
  const T hj = (T)(1. / M.h[j]), dyj = (T)(M.dy[j] / 2.), dyjm = (T)(M.dy[j - 1] / 2.);
  B = X.put(pf);
  X.pre((const T*)F.su, 0, pf);
  SQ acc[q] = dyj * XA(B, q);
  B = X.put(pf);  // u, readable until the second staging op from here
  X.pre((const T*)F.sp, 0, pf);
  SQ acc[q] = j2 ? hj * (acc[q] + dyjm * XA(B, q)) * vv[q] : (T)0;
  T rt = xch_right(acc[0], G, L, l, gs, bf.xr);
  SQ d[q] = (T)24. * dx1 * (-acc[q] + nxt(acc, q, rt));
  psolve<T, SEG>(d, G, L, l, gs, s22, bf.xf, bf.xb);
  SQ d[q] = j2 ? d[q] : (T)0;
  X.out((T*)F.hh[1], d);
  // u: viscous part of RUH1, HH1 = D1x((Ix u)^2)
  SQ acc[q] = (T)12. * dx1q * (XA(B, q - 1) - (T)2. * XA(B, q) + XA(B, q + 1));
  psolve<T, SEG>(acc, G, L, l, gs, s10, bf.xf, bf.xb);
  SQ d[q] = (T)4. * (XA(B, q) + XA(B, q + 1));
  if constexpr (SGS) SQ e[q] = (T)24. * dx1 * (XA(B, q + 1) - XA(B, q));
  psolve<T, SEG>(d, G, L, l, gs, s6, bf.xf, bf.xb);
  SQ d[q] = d[q] * d[q];
  const T lf = xch_left(d[SEG - 1], G, L, l, gs, bf.xl);
  SQR d[q] = (T)24. * dx1 * (d[q] - prv(d, q, lf));
  psolve<T, SEG>(d, G, L, l, gs, s22, bf.xf, bf.xb);
  if constexpr (SGS) {
    // HH1 -= 2 D1x(SGSVIS * D1x(u)), u face -> centre -> face
    psolve<T, SEG>(e, G, L, l, gs, s22, bf.xf, bf.xb);
    SQ e[q] = sv[q] * e[q];
    const T le = xch_left(e[SEG - 1], G, L, l, gs, bf.xl);
    SQR e[q] = (T)24. * dx1 * (e[q] - prv(e, q, le));
    psolve<T, SEG>(e, G, L, l, gs, s22, bf.xf, bf.xb);
    SQ d[q] = d[q] - (T)2. * e[q];
  }
  X.out((T*)F.hh[0], d);
  // P: RUH1 = (1/RE) D2x u - D1x P - PRESG
  B = X.put(pf);
  X.pre((const T*)F.sw, 0, pf);
  SQ d[q] = (T)24. * dx1 * (-XA(B, q - 1) + XA(B, q));
  psolve<T, SEG>(d, G, L, l, gs, s22, bf.xf, bf.xb);
  const T presg = (T)F.sc[SC_PRESG];
  SQ d[q] = rinv * acc[q] - d[q] - presg;
  X.out((T*)F.ruh[0], d);
  // w: RUH3, WW1 (and dwdx)
  B = X.put(pf);
  SQ d[q] = (T)12. * dx1q * (XA(B, q - 1) - (T)2. * XA(B, q) + XA(B, q + 1));
  psolve<T, SEG>(d, G, L, l, gs, s10, bf.xf, bf.xb);
  SQ d[q] = rinv * d[q];
  X.out((T*)F.ruh[2], d);
  SQ d[q] = (T)4. * (XA(B, q - 1) + XA(B, q));
  if constexpr (SGS) SQ e[q] = (T)24. * dx1 * (XA(B, q) - XA(B, q - 1));
  psolve<T, SEG>(d, G, L, l, gs, s6, bf.xf, bf.xb);
  X.out((T*)F.ww1, d);
  if constexpr (SGS) {
    psolve<T, SEG>(e, G, L, l, gs, s22, bf.xf, bf.xb);
    X.out((T*)F.dwdx, e);
  }
}

// ------------------------------------------------------------ RHS z-pass
// per z-line (i,j), u, v, then w and P:
//   RUH1 += (1/RE) D2z u                                    (:6170-6185)
//   UW    = Iz(u) * WW1 (stored);  HH1 += D1z(UW)           (:6360-6433)
//   RUH2 += (1/RE) D2z v                          (j>=2)   (:6907-6923)
//   VV3   = Iz v (stored);  HH2 += D1z(VV3*WW2)   (j>=2)   (:7023-7084)
//   RUH3 += (1/RE) D2z w - D1z(P) - PRESG3                  (:7510-7550)
//   HH3   = D1z((Iz w)^2)                                   (:7633-7672)
template <typename T, int SEG, bool CB_>
__global__ __launch_bounds__(512) void k_rhs_z(CtxArg<CB_> ca_) {
  const Ctx c = ctx_get(ca_);
  ZSETUP;
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  const PSolve s6 = PZ.s6, s10 = PZ.s10, s22 = PZ.s22;
  const T* sw = (const T*)F.sw;
  T* ruh0 = (T*)F.ruh[0];
  T* ruh1 = (T*)F.ruh[1];
  T* ruh2 = (T*)F.ruh[2];
  T* hh0 = (T*)F.hh[0];
  T* hh1 = (T*)F.hh[1];
  T* hh2 = (T*)F.hh[2];
  const T dx3 = (T)s.dx3, dx3q = (T)s.dx3q, rinv = (T)s.rinv;
  // each read-modify-write operand is loaded one solve ahead of its use,
  // nontemporally (ld_nt), like P and WW1: read once here
  T a[SEG], d[SEG], e[SEG], pr[SEG], lf, rt;
  // u
  zload((const T*)F.su, b, pl, p0, bm, bp, a, lf, rt);
  SQ d[q] = (T)12. * dx3q * (prv(a, q, lf) - (T)2. * a[q] + nxt(a, q, rt));
  SQ pr[q] = ld_nt(&ruh0[ZI(q)]);
  psolve<T, SEG>(d, G, L, l, gs, s10, bf.xf, bf.xb);
  if (ok) SQ st_nt(&ruh0[ZI(q)], (T)(pr[q] + rinv * d[q]));
  SQ d[q] = (T)4. * (prv(a, q, lf) + a[q]);
  SQ e[q] = ld_nt(&((const T*)F.ww1)[ZI(q)]);
  psolve<T, SEG>(d, G, L, l, gs, s6, bf.xf, bf.xb);
  SQ d[q] = d[q] * e[q];  // UW
  if (ok) SQ st_nt(&((T*)F.uw)[ZI(q)], (T)(d[q]));
  SQ pr[q] = ld_nt(&hh0[ZI(q)]);
  rt = xch_right(d[0], G, L, l, gs, bf.xr);
  SQ d[q] = (T)24. * dx3 * (-d[q] + nxt(d, q, rt));
  zload((const T*)F.sv, b, pl, p0, bm, bp, a, lf, rt);
  psolve<T, SEG>(d, G, L, l, gs, s22, bf.xf, bf.xb);
  if (ok) SQ st_nt(&hh0[ZI(q)], (T)(pr[q] + d[q]));
  // v
  if (j >= 2) {
    SQ d[q] = (T)12. * dx3q * (prv(a, q, lf) - (T)2. * a[q] + nxt(a, q, rt));
    SQ pr[q] = ld_nt(&ruh1[ZI(q)]);
    psolve<T, SEG>(d, G, L, l, gs, s10, bf.xf, bf.xb);
    if (ok) SQ st_nt(&ruh1[ZI(q)], (T)(pr[q] + rinv * d[q]));
  }
  SQ d[q] = (T)4. * (prv(a, q, lf) + a[q]);
  const T hj = (T)(1. / M.h[j]), dyj = (T)(M.dy[j] / 2.), dyjm = (T)(M.dy[j - 1] / 2.);
  if (j >= 2) SQ e[q] = hj * (dyj * sw[ZI(q) - g.nx] + dyjm * sw[ZI(q)]);  // WW2 (:7028)
  psolve<T, SEG>(d, G, L, l, gs, s6, bf.xf, bf.xb);
  if (ok) SQ st_nt(&((T*)F.vv3)[ZI(q)], (T)(d[q]));
  if (j >= 2) {
    SQ d[q] = d[q] * e[q];
    SQ pr[q] = ld_nt(&hh1[ZI(q)]);
    rt = xch_right(d[0], G, L, l, gs, bf.xr);
    SQ d[q] = (T)24. * dx3 * (-d[q] + nxt(d, q, rt));
    psolve<T, SEG>(d, G, L, l, gs, s22, bf.xf, bf.xb);
    if (ok) SQ st_nt(&hh1[ZI(q)], (T)(pr[q] + d[q]));
  }
  // w and P (a phase of its own, so its registers are those the u and v
  // phases used): RUH3 += (1/RE) D2z w - D1z(P) - PRESG3 (:7510-7550),
  // HH3 = D1z((Iz w)^2) (:7633-7672); w, P and the RUH3 operand are all
  // loaded before the first solve
  {
    T r[SEG];
    zload(sw, b, pl, p0, bm, bp, a, lf, rt);
    const T* sp = (const T*)F.sp;
    const T pv0 = sp[bm];
    SQ d[q] = ld_nt(&sp[ZI(q)]);
    SQ r[q] = ld_nt(&ruh2[ZI(q)]);
    SQ e[q] = (T)12. * dx3q * (prv(a, q, lf) - (T)2. * a[q] + nxt(a, q, rt));
    psolve<T, SEG>(e, G, L, l, gs, s10, bf.xf, bf.xb);
    SQR d[q] = (T)24. * dx3 * (-(q == 0 ? pv0 : d[q > 0 ? q - 1 : 0]) + d[q]);
    psolve<T, SEG>(d, G, L, l, gs, s22, bf.xf, bf.xb);
    const T presg3 = (T)F.sc[SC_PRESG3];
    if (ok) SQ st_nt(&ruh2[ZI(q)], (T)(r[q] + rinv * e[q] - d[q] - presg3));
    // HH3 = D1z((Iz w)^2), w face -> centre -> face
    SQ d[q] = (T)4. * (a[q] + nxt(a, q, rt));
    psolve<T, SEG>(d, G, L, l, gs, s6, bf.xf, bf.xb);
    SQ d[q] = d[q] * d[q];
    lf = xch_left(d[SEG - 1], G, L, l, gs, bf.xl);
    SQR d[q] = (T)24. * dx3 * (d[q] - prv(d, q, lf));
    psolve<T, SEG>(d, G, L, l, gs, s22, bf.xf, bf.xb);
    if (ok) SQ st_nt(&hh2[ZI(q)], (T)(d[q]));
  }
}


// -------------------------------------------------------- RHS x-pass 2
// HH3 += D1x(UU3*WW1) face -> centre (:7734-7769)
template <typename T, int SEG, bool CB_>
__global__ __launch_bounds__(256) void k_rhs_x2(CtxArg<CB_> ca_) {
  const Ctx c = ctx_get(ca_);
  XSETUP_NT;
  const T dx1 = (T)c.s.dx1;
  T d[SEG], pf[SEG];
  X.pre((const T*)c.d.uw, 0, pf);
  const T* B = X.put(pf);
  X.pre((const T*)c.d.hh[2], 0, pf);
  SQ d[q] = (T)24. * dx1 * (-XA(B, q) + XA(B, q + 1));
  psolve<T, SEG>(d, G, L, l, gs, c.o.px.s22, bf.xf, bf.xb);
  B = X.put(pf);
  SQ d[q] = XA(B, q) + d[q];
  X.out((T*)c.d.hh[2], d);
}

// ------------------------------------------------------------- GETUH y
// TDMAI along y (compact3d.f:8630-8659) with the J-only chains (A, BET, GAM)
// of the host, as a partitioned recurrence: thread (l, c) holds chunk c of
// column i = i0 + l.  Forward x(J) = BET(J)(R(J) - A(J) x(J-1)) is
// y = c + m y_prev with m = -A BET; backward x(J) -= GAM(J) x(J+1).
#define CHMAX 12
struct YPos {
  int l, c, j0, len;
};

// Column values of the y-sweeps.  A thread holds one column (V = T) or, in
// FP32, two neighbouring columns (V = f2v, one 8-byte load or store per
// pair of points: a wave's 4-byte loads carry half the bytes of its 8-byte
// ones).  The J-only coefficients stay scalars (S) and multiply both.
typedef float f2v __attribute__((ext_vector_type(2)));
template <typename V>
struct YCol {
  typedef V S;
  static constexpr int P = 1;
};
template <>
struct YCol<f2v> {
  typedef float S;
  static constexpr int P = 2;
};
template <typename V>
__device__ __forceinline__ V vsplat(typename YCol<V>::S x) {
  return (V)x;
}
// element idx (and idx+1 for a pair) of a field of the scalar type
template <typename V>
__device__ __forceinline__ V vld(const typename YCol<V>::S* p, unsigned idx) {
  return *(const V*)(p + idx);
}
// nontemporal: the sweep's read-once streams (RUH, HH, HH2) and its last
// reads of the step's state shadows and VV1/VV3 (measured: the z-sweep
// after it 41 -> 35 us mixed)
template <typename V>
__device__ __forceinline__ V vld_nt(const typename YCol<V>::S* p, unsigned idx) {
  return __builtin_nontemporal_load((const V*)(p + idx));
}
template <typename V>
__device__ __forceinline__ void vst(typename YCol<V>::S* p, unsigned idx, V x) {
  *(V*)(p + idx) = x;
}
template <typename V>
__device__ __forceinline__ void vst_nt(typename YCol<V>::S* p, unsigned idx, V x) {
  __builtin_nontemporal_store(x, (V*)(p + idx));
}
// FP64 plane values (wall stresses) of the thread's columns
template <typename V>
__device__ __forceinline__ V vld_d(const double* p, unsigned idx) {
  typedef typename YCol<V>::S S;
  if constexpr (YCol<V>::P == 2)
    return V{(S)p[idx], (S)p[idx + 1]};
  else
    return (V)p[idx];
}

// chunk-parallel linear recurrence v[q] <- v[q] + m(q) * v[q-1] (forward):
// each chunk runs it from zero, the chunks' end values and multiplier
// products are combined across the block through LDS, and the incoming
// value Y reaches point q of the chunk as Y * m(0) * ... * m(q).  m(q) is
// evaluated twice, so a multiplier held in LDS or recomputed costs no
// registers across the barrier.  V: the column value type (one column, or
// a pair sharing the multipliers).
template <typename V, typename T, int N, typename MF>
__device__ __forceinline__ void ychain_fwd(V (&v)[N], MF m, const YPos& y, int L, V* xa, T* xm) {
  V r = vsplat<V>((T)0);
  T P = (T)1;
#pragma unroll
  for (int q = 0; q < N; q++)
    if (q < y.len) {
      const T mq = m(q);
      r = v[q] + mq * r;
      v[q] = r;
      P *= mq;
    }
  xa[y.c * L + y.l] = r;
  xm[y.c * L + y.l] = P;
  __syncthreads();
  V Y = vsplat<V>((T)0);
  for (int cc = 0; cc < y.c; cc++) Y = xa[cc * L + y.l] + xm[cc * L + y.l] * Y;
#pragma unroll
  for (int q = 0; q < N; q++)
    if (q < y.len) {
      Y *= m(q);
      v[q] += Y;
    }
}
// backward: v[q] <- v[q] + m(q) * v[q+1]
template <typename V, typename T, int N, typename MF>
__device__ __forceinline__ void ychain_bwd(V (&v)[N], MF m, const YPos& y, int C, int L, V* xa, T* xm) {
  V r = vsplat<V>((T)0);
  T P = (T)1;
#pragma unroll
  for (int q = N - 1; q >= 0; q--)
    if (q < y.len) {
      const T mq = m(q);
      r = v[q] + mq * r;
      v[q] = r;
      P *= mq;
    }
  xa[y.c * L + y.l] = r;
  xm[y.c * L + y.l] = P;
  __syncthreads();
  V Z = vsplat<V>((T)0);
  for (int cc = C - 1; cc > y.c; cc--) Z = xa[cc * L + y.l] + xm[cc * L + y.l] * Z;
#pragma unroll
  for (int q = N - 1; q >= 0; q--)
    if (q < y.len) {
      Z *= m(q);
      v[q] += Z;
    }
}

// the same two recurrences on two arrays with one multiplier sequence (the
// real and imaginary parts of a Poisson mode): one exchange, one product per
// chunk, and the two cross-chunk chains interleaved
template <typename T, int N, typename MF>
__device__ __forceinline__ void ychain_fwd2(T (&v)[N], T (&w)[N], MF m, const YPos& y, int L, T* xa, T* xb,
                                            T* xm) {
  T r = (T)0, s = (T)0, P = (T)1;
#pragma unroll
  for (int q = 0; q < N; q++)
    if (q < y.len) {
      const T mq = m(q);
      r = v[q] + mq * r;
      s = w[q] + mq * s;
      v[q] = r;
      w[q] = s;
      P *= mq;
    }
  xa[y.c * L + y.l] = r;
  xb[y.c * L + y.l] = s;
  xm[y.c * L + y.l] = P;
  __syncthreads();
  T Y = (T)0, Z = (T)0;
  for (int cc = 0; cc < y.c; cc++) {
    const T mc = xm[cc * L + y.l];
    Y = xa[cc * L + y.l] + mc * Y;
    Z = xb[cc * L + y.l] + mc * Z;
  }
#pragma unroll
  for (int q = 0; q < N; q++)
    if (q < y.len) {
      const T mq = m(q);
      Y *= mq;
      Z *= mq;
      v[q] += Y;
      w[q] += Z;
    }
}
template <typename T, int N, typename MF>
__device__ __forceinline__ void ychain_bwd2(T (&v)[N], T (&w)[N], MF m, const YPos& y, int C, int L, T* xa,
                                            T* xb, T* xm) {
  T r = (T)0, s = (T)0, P = (T)1;
#pragma unroll
  for (int q = N - 1; q >= 0; q--)
    if (q < y.len) {
      const T mq = m(q);
      r = v[q] + mq * r;
      s = w[q] + mq * s;
      v[q] = r;
      w[q] = s;
      P *= mq;
    }
  xa[y.c * L + y.l] = r;
  xb[y.c * L + y.l] = s;
  xm[y.c * L + y.l] = P;
  __syncthreads();
  T Y = (T)0, Z = (T)0;
  for (int cc = C - 1; cc > y.c; cc--) {
    const T mc = xm[cc * L + y.l];
    Y = xa[cc * L + y.l] + mc * Y;
    Z = xb[cc * L + y.l] + mc * Z;
  }
#pragma unroll
  for (int q = N - 1; q >= 0; q--)
    if (q < y.len) {
      const T mq = m(q);
      Y *= mq;
      Z *= mq;
      v[q] += Y;
      w[q] += Z;
    }
}

// ys: the tile shape (one column per thread)
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_adi_y(Ctx c, YShape ys) {
  __shared__ T xs[4][NT];  // C*L <= NT entries each
  const Geo g = c.g;
  const int comp = blockIdx.y;
  const int ns = (comp == 1) ? 2 : 1;
  const int L = ys.L, C = ys.C;
  const int t = threadIdx.x, l = t % L;
  YPos y;
  y.l = l;
  y.c = t / L;
  const int m = g.n2m - ns + 1;
  const int ch = (m + C - 1) / C;
  y.j0 = ns + y.c * ch;
  y.len = max(0, min(ch, g.n2m + 1 - y.j0));
  const int nit = (g.nx + L - 1) / L;
  const int it = blockIdx.x % nit;
  const int k = blockIdx.x / nit;
  const int i = it * L + l;
  const bool ok = i < g.nx;
  const Fields F = c.d;
  const Metrics M = c.m;
  const T* R = (const T*)(comp == 0 ? F.ruh[0] : comp == 1 ? F.ruh[1] : F.ruh[2]);
  T* X = (T*)(comp == 0 ? F.uh[0] : comp == 1 ? F.uh[1] : F.uh[2]);
  const double* A = comp == 0 ? M.ya[0] : comp == 1 ? M.ya[1] : M.ya[2];
  const double* BET = comp == 0 ? M.ybet[0] : comp == 1 ? M.ybet[1] : M.ybet[2];
  const double* GAM = comp == 0 ? M.ygam[0] : comp == 1 ? M.ygam[1] : M.ygam[2];
  const T dt = (T)c.s.dt;
  const long long b = gidx(g, ok ? i : 0, 0, k);
  T v[CHMAX];
#pragma unroll
  for (int q = 0; q < CHMAX; q++)
    if (q < y.len) {
      const int J = y.j0 + q;
      v[q] = (T)BET[J] * (R[b + (long long)J * g.nx] * dt);
    }
  ychain_fwd(
      v, [&](int q) { const int J = y.j0 + q; return (J == ns) ? (T)0 : -(T)A[J] * (T)BET[J]; }, y, L, xs[0], xs[1]);
  ychain_bwd(
      v, [&](int q) { const int J = y.j0 + q; return (J == g.n2m) ? (T)0 : -(T)GAM[J]; }, y, C, L, xs[2], xs[3]);
  if (ok)
#pragma unroll
    for (int q = 0; q < CHMAX; q++)
      if (q < y.len) X[b + (long long)(y.j0 + q) * g.nx] = v[q];
}

// RHS y parts fused into the ADI y-sweep: each thread forms the final RUH of
// its chunk points and feeds it straight into the TDMAI recurrences; RUH is
// never stored, HH is.  Without SGS terms every operand of the thread's
// chunk (its component's field at J0-1..J0+len, the VV1/P/VV3 rows, RUH, HH
// and HH2) is loaded into registers before any arithmetic, so a thread has
// all of its loads in flight at once; the expressions are those of
// yterm_u/v/w.  With SGS terms it calls yterm_*.  Returns TDMAI's right-hand
// side BET(J) * RUH(J) * DT.
// the wall-stress term of RHS1/RHS3 (:6765-6768, :8096-8103):
// (1-FJMU(J)) (-TAU(I,K,1)/DY(1)) IFWALL + (1-FJPA(J)) (-TAU(I,K,2)/DY(N2M))
// IFWALL (ALY-1).  The flags are 0 except FJMU(1) and FJPA(N2M), so only the
// rows J = 1 and J = N2M load a stress; t: the row's coefficients.
template <typename V, typename T>
__device__ __forceinline__ V wall_bc(const Ctx& c, const double* tau, unsigned t2, int J, const T* t) {
  V bc = vsplat<V>((T)0);
  if (J == 1) bc = (T)t[10] * (-vld_d<V>(tau, t2) / (T)c.m.dy[1] * (T)c.s.ifwall);
  if (J == c.g.n2m)
    bc = bc + (T)t[9] * (-vld_d<V>(tau, (unsigned)(c.g.nx * c.g.nz) + t2) / (T)c.m.dy[c.g.n2m] * (T)c.s.ifwall *
                         (T)(c.s.aly - 1.));
  return bc;
}
// the value at column I-1 of each of the thread's columns (di: the offset
// from column I to its periodic left neighbour)
template <typename V>
__device__ __forceinline__ V vld_im(const typename YCol<V>::S* p, unsigned idx, int di) {
  if constexpr (YCol<V>::P == 2)
    return V{p[idx + di], p[idx]};
  else
    return p[idx + di];
}

// RUH, HH and HH2 are read once and HH written once per step: their loads
// and stores carry the nontemporal hint (measured 1-1.5 % faster than plain).
// With SGS (FILES != 0) the SGS y terms of c3d_yterms.h join HH before it is
// stored, their operands loaded with the chunk's (SGSVIS at the wall rows
// only where the chunk holds J = 1 or N2M, the only rows whose weight is
// not zero); di: the offset of column I-1 (periodic) for SGSVIS's wall rows.
// V: one column (T) or a pair of FP32 columns (f2v), with the same
// arithmetic per column.
template <typename T, typename V, int CH, bool SGS = false>
__device__ __forceinline__ void yfuse_u(const Ctx& c, const T* tb, unsigned b, unsigned t2, const YPos& y,
                                        V (&v)[CH], int di = 0) {
  const Geo& g = c.g;
  const Fields& F = c.d;
  const T* __restrict__ su = (const T*)F.su;
  const T* __restrict__ vv1 = (const T*)F.vv1;
  const T* __restrict__ ruh = (const T*)F.ruh[0];
  T* __restrict__ hh = (T*)F.hh[0];
  const T* __restrict__ hp = (const T*)F.hhp[0];
  const T rinv = (T)c.s.rinv, dt = (T)c.s.dt;
  const unsigned nx = g.nx;
  const V zero = vsplat<V>((T)0);
  const V uN2 = vld<V>(su, b + g.n2 * nx), uW = vld<V>(su, b);

  const unsigned a0 = b + (unsigned)y.j0 * nx;  // row J0
  V us[CH + 2], w[CH + 1], ru[CH], hx[CH], hq[CH];
#pragma unroll
  for (int q = 0; q < CH + 2; q++) us[q] = (q < y.len + 2) ? vld_nt<V>(su, a0 - nx + q * nx) : zero;
#pragma unroll
  for (int q = 0; q < CH + 1; q++) w[q] = (q < y.len + 1) ? vld_nt<V>(vv1, a0 + q * nx) : zero;
#pragma unroll
  for (int q = 0; q < CH; q++) {
    const bool in = q < y.len;
    ru[q] = in ? vld_nt<V>(ruh, a0 + q * nx) : zero;
    hx[q] = in ? vld_nt<V>(hh, a0 + q * nx) : zero;
    hq[q] = in ? vld_nt<V>(hp, a0 + q * nx) : zero;
  }
  if constexpr (!SGS) {
#pragma unroll
    for (int q = 0; q < CH; q++)
      if (q < y.len) {
        const int J = y.j0 + q;
        const V um = us[q], u0 = us[q + 1], up = us[q + 2];
        const T* t = tb + J * YTAB_W;
        const V r0 = ru[q] + rinv * ((T)t[0] * up - (T)t[1] * u0 + (T)t[2] * um);
        const V u2 = (T)t[3] * ((T)t[6] * up + (T)t[7] * u0) + (T)t[9] * uN2;
        const V u1 = (T)t[4] * ((T)t[5] * u0 + (T)t[6] * um) + (T)t[10] * uW;
        const V v2 = (J == g.n2m) ? zero : w[q + 1];
        const V v1 = (J == 1) ? zero : w[q];
        const V h = hx[q] + (T)t[8] * (v2 * u2 - v1 * u1);
        vst_nt<V>(hh, a0 + q * nx, h);
        const V r = r0 - (T)(3. / 2.) * h + (T)(1. / 2.) * hq[q] + wall_bc<V>(c, F.tauw, t2, J, t);
        v[q] = (T)t[YT_BET] * (r * dt);
      }
  } else if constexpr (sizeof(T) == 8) {
    // FP64 with SGS terms: every operand of the chunk loaded before the
    // arithmetic (the two phases below measured 11 % slower in FP64): nux at
    // J0-1..J0+len, dv/dx at J0..J0+len, and the sums of SGSVIS at
    // (I,0,K) + (I-1,0,K) and (I,N2,K) + (I-1,N2,K)
    const T* __restrict__ nux = (const T*)F.nux;
    const T* __restrict__ dvdx = (const T*)F.dvdx;
    const T* __restrict__ sg = (const T*)F.ssgs;
    V nu[CH + 2], dv[CH + 1], sgb = zero, sgt = zero;
#pragma unroll
    for (int q = 0; q < CH + 2; q++) nu[q] = (q < y.len + 2) ? vld<V>(nux, a0 - nx + q * nx) : zero;
#pragma unroll
    for (int q = 0; q < CH + 1; q++) dv[q] = (q < y.len + 1) ? vld<V>(dvdx, a0 + q * nx) : zero;
    if (y.j0 == 1) sgb = vld<V>(sg, b) + vld_im<V>(sg, b, di);
    if (y.j0 + y.len - 1 == g.n2m) sgt = vld<V>(sg, b + g.n2 * nx) + vld_im<V>(sg, b + g.n2 * nx, di);
#pragma unroll
    for (int q = 0; q < CH; q++)
      if (q < y.len) {
        const int J = y.j0 + q;
        const V um = us[q], u0 = us[q + 1], up = us[q + 2];
        const T* t = tb + J * YTAB_W;
        const V r0 = ru[q] + rinv * ((T)t[0] * up - (T)t[1] * u0 + (T)t[2] * um);
        const V u2 = (T)t[3] * ((T)t[6] * up + (T)t[7] * u0) + (T)t[9] * uN2;
        const V u1 = (T)t[4] * ((T)t[5] * u0 + (T)t[6] * um) + (T)t[10] * uW;
        const V v2 = (J == g.n2m) ? zero : w[q + 1];
        const V v1 = (J == 1) ? zero : w[q];
        const V h = hx[q] + (T)t[8] * (v2 * u2 - v1 * u1);
        // D(SGS*(DU/DY))/DY and D(SGS*(DV/DX))/DY (:6551-6680), as yterm_u
        const V n0 = nu[q + 1], np1 = (J < g.n2m) ? nu[q + 2] : zero, nm1 = (J > 1) ? nu[q] : zero;
        const V sgsj2 = (T)t[3] * ((T)t[6] * np1 + (T)t[7] * n0) + (T)t[9] * (T)0.5 * sgt;
        const V sgsj1 = (T)t[4] * ((T)t[5] * n0 + (T)t[6] * nm1) + (T)t[10] * (T)0.5 * sgb;
        const T wfu = (T)1 - (T)c.s.ifwall * (T)t[9] * (T)(c.s.aly - 1.);
        const T wfd = (T)1 - (T)c.s.ifwall * (T)t[10];
        const V uu22 = (up - u0) * (T)t[15];
        const V uu11 = (u0 - um) * (T)t[14];
        const V hd = h - (T)t[8] * (wfu * sgsj2 * uu22 - wfd * sgsj1 * uu11);
        const V dvp = (J < g.n2m) ? dv[q + 1] : zero;
        const V hs = hd - (T)t[8] * (wfu * sgsj2 * dvp - wfd * sgsj1 * dv[q]);
        vst_nt<V>(hh, a0 + q * nx, hs);
        const V r = r0 - (T)(3. / 2.) * hs + (T)(1. / 2.) * hq[q] + wall_bc<V>(c, F.tauw, t2, J, t);
        v[q] = (T)t[YT_BET] * (r * dt);
      }
  } else {
    // FP32 with SGS terms: the chunk's operands come in two phases, so that the
    // registers of the first are free for the second: the convective and
    // viscous terms (h, r0), then the SGS terms' operands (nux at
    // J0-1..J0+len, dv/dx at J0..J0+len, and the sums of SGSVIS at
    // (I,0,K) + (I-1,0,K) and (I,N2,K) + (I-1,N2,K)); each point's
    // expressions are the same.
    V h[CH], r0[CH];
#pragma unroll
    for (int q = 0; q < CH; q++)
      if (q < y.len) {
        const int J = y.j0 + q;
        const V um = us[q], u0 = us[q + 1], up = us[q + 2];
        const T* t = tb + J * YTAB_W;
        r0[q] = ru[q] + rinv * ((T)t[0] * up - (T)t[1] * u0 + (T)t[2] * um);
        const V u2 = (T)t[3] * ((T)t[6] * up + (T)t[7] * u0) + (T)t[9] * uN2;
        const V u1 = (T)t[4] * ((T)t[5] * u0 + (T)t[6] * um) + (T)t[10] * uW;
        const V v2 = (J == g.n2m) ? zero : w[q + 1];
        const V v1 = (J == 1) ? zero : w[q];
        h[q] = hx[q] + (T)t[8] * (v2 * u2 - v1 * u1);
      }
    // the barrier keeps the second phase's loads after the first phase's
    // arithmetic (512-thread pair tiles 165 -> 153 VGPRs, 3.6 % faster)
    __builtin_amdgcn_sched_barrier(0);
    const T* __restrict__ nux = (const T*)F.nux;
    const T* __restrict__ dvdx = (const T*)F.dvdx;
    const T* __restrict__ sg = (const T*)F.ssgs;
    V nu[CH + 2], dv[CH + 1], sgb = zero, sgt = zero;
#pragma unroll
    for (int q = 0; q < CH + 2; q++) nu[q] = (q < y.len + 2) ? vld<V>(nux, a0 - nx + q * nx) : zero;
#pragma unroll
    for (int q = 0; q < CH + 1; q++) dv[q] = (q < y.len + 1) ? vld<V>(dvdx, a0 + q * nx) : zero;
    if (y.j0 == 1) sgb = vld<V>(sg, b) + vld_im<V>(sg, b, di);
    if (y.j0 + y.len - 1 == g.n2m) sgt = vld<V>(sg, b + g.n2 * nx) + vld_im<V>(sg, b + g.n2 * nx, di);
#pragma unroll
    for (int q = 0; q < CH; q++)
      if (q < y.len) {
        const int J = y.j0 + q;
        const V um = us[q], u0 = us[q + 1], up = us[q + 2];
        const T* t = tb + J * YTAB_W;
        // D(SGS*(DU/DY))/DY and D(SGS*(DV/DX))/DY (:6551-6680), as yterm_u
        const V n0 = nu[q + 1], np1 = (J < g.n2m) ? nu[q + 2] : zero, nm1 = (J > 1) ? nu[q] : zero;
        const V sgsj2 = (T)t[3] * ((T)t[6] * np1 + (T)t[7] * n0) + (T)t[9] * (T)0.5 * sgt;
        const V sgsj1 = (T)t[4] * ((T)t[5] * n0 + (T)t[6] * nm1) + (T)t[10] * (T)0.5 * sgb;
        const T wfu = (T)1 - (T)c.s.ifwall * (T)t[9] * (T)(c.s.aly - 1.);
        const T wfd = (T)1 - (T)c.s.ifwall * (T)t[10];
        const V uu22 = (up - u0) * (T)t[15];
        const V uu11 = (u0 - um) * (T)t[14];
        const V hd = h[q] - (T)t[8] * (wfu * sgsj2 * uu22 - wfd * sgsj1 * uu11);
        const V dvp = (J < g.n2m) ? dv[q + 1] : zero;
        const V hs = hd - (T)t[8] * (wfu * sgsj2 * dvp - wfd * sgsj1 * dv[q]);
        vst_nt<V>(hh, a0 + q * nx, hs);
        const V r = r0[q] - (T)(3. / 2.) * hs + (T)(1. / 2.) * hq[q] + wall_bc<V>(c, F.tauw, t2, J, t);
        v[q] = (T)t[YT_BET] * (r * dt);
      }
  }
}

template <typename T, typename V, int CH, bool SGS = false>
__device__ __forceinline__ void yfuse_v(const Ctx& c, const T* tb, unsigned b, const YPos& y, V (&v)[CH]) {
  const Fields& F = c.d;
  const T* __restrict__ sv = (const T*)F.sv;
  const T* __restrict__ sp = (const T*)F.sp;
  const T* __restrict__ ruh = (const T*)F.ruh[1];
  T* __restrict__ hh = (T*)F.hh[1];
  const T* __restrict__ hp = (const T*)F.hhp[1];
  const T rinv = (T)c.s.rinv, dt = (T)c.s.dt;
  const unsigned nx = c.g.nx;
  const V zero = vsplat<V>((T)0);
  const unsigned a0 = b + (unsigned)y.j0 * nx;
  V vs[CH + 2], ps[CH + 1], ru[CH], hx[CH], hq[CH];
#pragma unroll
  for (int q = 0; q < CH + 2; q++) vs[q] = (q < y.len + 2) ? vld_nt<V>(sv, a0 - nx + q * nx) : zero;
#pragma unroll
  for (int q = 0; q < CH + 1; q++) ps[q] = (q < y.len + 1) ? vld_nt<V>(sp, a0 - nx + q * nx) : zero;
#pragma unroll
  for (int q = 0; q < CH; q++) {
    const bool in = q < y.len;
    ru[q] = in ? vld_nt<V>(ruh, a0 + q * nx) : zero;
    hx[q] = in ? vld_nt<V>(hh, a0 + q * nx) : zero;
    hq[q] = in ? vld_nt<V>(hp, a0 + q * nx) : zero;
  }
  // SGS operand: SGSVIS at J0-1..J0+len-1
  V sgv[SGS ? CH + 1 : 1];
  if constexpr (SGS) {
    const T* __restrict__ sg = (const T*)c.d.ssgs;
#pragma unroll
    for (int q = 0; q < CH + 1; q++) sgv[q] = (q < y.len + 1) ? vld<V>(sg, a0 - nx + q * nx) : zero;
  }
#pragma unroll
  for (int q = 0; q < CH; q++)
    if (q < y.len) {
      const int J = y.j0 + q;
      const V vm = vs[q], v0 = vs[q + 1], vp = vs[q + 2];
      const T* t = tb + J * YTAB_W;
      const V viscos = rinv * ((T)t[0] * vp - (T)t[1] * v0 + (T)t[2] * vm);
      const V pressg2 = (ps[q + 1] - ps[q]) * (T)t[4];  // t[4] = 1/H(J)
      const V r0 = ru[q] + (-pressg2 + viscos);
      const V v2 = (T)0.5 * (vp + v0);
      const V v1 = (T)0.5 * (v0 + vm);
      V h = hx[q] + (T)t[4] * (v2 * v2 - v1 * v1);
      if constexpr (SGS) {
        // 2 D(SGS*(DV/DY))/DY (as yterm_v); t[5] = 1/DY(J), t[6] = 1/DY(J-1)
        const V g2 = (T)t[5] * (vp - v0);
        const V g1 = (T)t[6] * (v0 - vm);
        const V vis22 = (T)t[4] * (sgv[q + 1] * g2 - sgv[q] * g1);
        h = h - (T)2. * vis22;
      }
      vst_nt<V>(hh, a0 + q * nx, h);
      const V r = r0 - (T)(3. / 2.) * h + (T)(1. / 2.) * hq[q];
      v[q] = (T)t[YT_BET] * (r * dt);
    }
}

template <typename T, typename V, int CH, bool SGS = false>
__device__ __forceinline__ void yfuse_w(const Ctx& c, const T* tb, unsigned b, unsigned t2, const YPos& y,
                                        V (&v)[CH], int dk = 0) {
  const Geo& g = c.g;
  const Fields& F = c.d;
  const T* __restrict__ sw = (const T*)F.sw;
  const T* __restrict__ vv3 = (const T*)F.vv3;
  const T* __restrict__ ruh = (const T*)F.ruh[2];
  T* __restrict__ hh = (T*)F.hh[2];
  const T* __restrict__ hp = (const T*)F.hhp[2];
  const T rinv = (T)c.s.rinv, dt = (T)c.s.dt;
  const unsigned nx = g.nx;
  const V zero = vsplat<V>((T)0);
  const V wN2 = vld<V>(sw, b + g.n2 * nx), wW = vld<V>(sw, b);

  const unsigned a0 = b + (unsigned)y.j0 * nx;
  V ws[CH + 2], s3[CH + 1], ru[CH], hx[CH], hq[CH];
#pragma unroll
  for (int q = 0; q < CH + 2; q++) ws[q] = (q < y.len + 2) ? vld_nt<V>(sw, a0 - nx + q * nx) : zero;
#pragma unroll
  for (int q = 0; q < CH + 1; q++) s3[q] = (q < y.len + 1) ? vld_nt<V>(vv3, a0 + q * nx) : zero;
#pragma unroll
  for (int q = 0; q < CH; q++) {
    const bool in = q < y.len;
    ru[q] = in ? vld_nt<V>(ruh, a0 + q * nx) : zero;
    hx[q] = in ? vld_nt<V>(hh, a0 + q * nx) : zero;
    hq[q] = in ? vld_nt<V>(hp, a0 + q * nx) : zero;
  }
  if constexpr (!SGS) {
#pragma unroll
    for (int q = 0; q < CH; q++)
      if (q < y.len) {
        const int J = y.j0 + q;
        const V wm = ws[q], w0 = ws[q + 1], wp = ws[q + 2];
        const T* t = tb + J * YTAB_W;
        const V r0 = ru[q] + rinv * ((T)t[0] * wp - (T)t[1] * w0 + (T)t[2] * wm);
        const V w2 = (T)t[3] * ((T)t[7] * w0 + (T)t[6] * wp) + (T)t[9] * wN2;
        const V w1 = (T)t[4] * ((T)t[6] * wm + (T)t[5] * w0) + (T)t[10] * wW;
        const V v2 = (J == g.n2m) ? zero : s3[q + 1];
        const V v1 = (J == 1) ? zero : s3[q];
        const V h = hx[q] + (T)t[8] * (v2 * w2 - v1 * w1);
        vst_nt<V>(hh, a0 + q * nx, h);
        const V r = r0 - (T)(3. / 2.) * h + (T)(1. / 2.) * hq[q] + wall_bc<V>(c, F.taub, t2, J, t);
        v[q] = (T)t[YT_BET] * (r * dt);
      }
  } else if constexpr (sizeof(T) == 8) {
    // FP64 with SGS terms: every operand of the chunk loaded before the
    // arithmetic (as yfuse_u): nuz at J0-1..J0+len, dv/dz at J0..J0+len,
    // and the sums of SGSVIS at (I,0,K) + (I,0,K-1) and (I,N2,K) + (I,N2,K-1)
    const T* __restrict__ nuz = (const T*)F.nuz;
    const T* __restrict__ dvdz = (const T*)F.dvdz;
    const T* __restrict__ sg = (const T*)F.ssgs;
    V nz[CH + 2], dz[CH + 1], sgb = zero, sgt = zero;
#pragma unroll
    for (int q = 0; q < CH + 2; q++) nz[q] = (q < y.len + 2) ? vld<V>(nuz, a0 - nx + q * nx) : zero;
#pragma unroll
    for (int q = 0; q < CH + 1; q++) dz[q] = (q < y.len + 1) ? vld<V>(dvdz, a0 + q * nx) : zero;
    const unsigned bkm = b + dk;
    if (y.j0 == 1) sgb = vld<V>(sg, b) + vld<V>(sg, bkm);
    if (y.j0 + y.len - 1 == g.n2m) sgt = vld<V>(sg, b + g.n2 * nx) + vld<V>(sg, bkm + g.n2 * nx);
#pragma unroll
    for (int q = 0; q < CH; q++)
      if (q < y.len) {
        const int J = y.j0 + q;
        const V wm = ws[q], w0 = ws[q + 1], wp = ws[q + 2];
        const T* t = tb + J * YTAB_W;
        const V r0 = ru[q] + rinv * ((T)t[0] * wp - (T)t[1] * w0 + (T)t[2] * wm);
        const V w2 = (T)t[3] * ((T)t[7] * w0 + (T)t[6] * wp) + (T)t[9] * wN2;
        const V w1 = (T)t[4] * ((T)t[6] * wm + (T)t[5] * w0) + (T)t[10] * wW;
        const V v2 = (J == g.n2m) ? zero : s3[q + 1];
        const V v1 = (J == 1) ? zero : s3[q];
        const V h = hx[q] + (T)t[8] * (v2 * w2 - v1 * w1);
        // D(SGS*(DW/DY))/DY and D(SGS*(DV/DZ))/DY (:7858-7908, :8042-8094),
        // as yterm_w
        const V z0 = nz[q + 1], zp1 = (J < g.n2m) ? nz[q + 2] : zero, zm1 = (J > 1) ? nz[q] : zero;
        const V sgsj1 = (T)t[4] * ((T)t[5] * z0 + (T)t[6] * zm1) + (T)t[10] * (T)0.5 * sgb;
        const V sgsj2 = (T)t[3] * ((T)t[6] * zp1 + (T)t[7] * z0) + (T)t[9] * (T)0.5 * sgt;
        const T wfu = (T)1 - (T)c.s.ifwall * (T)t[9] * (T)(c.s.aly - 1.);
        const T wfd = (T)1 - (T)c.s.ifwall * (T)t[10];
        const V ww22 = (T)t[15] * (wp - w0);
        const V ww11 = (T)t[14] * (w0 - wm);
        const V hd = h - (T)t[8] * (wfu * sgsj2 * ww22 - wfd * sgsj1 * ww11);
        const V dzp = (J < g.n2m) ? dz[q + 1] : zero;
        const V hs = hd - (T)t[8] * (wfu * sgsj2 * dzp - wfd * sgsj1 * dz[q]);
        vst_nt<V>(hh, a0 + q * nx, hs);
        const V r = r0 - (T)(3. / 2.) * hs + (T)(1. / 2.) * hq[q] + wall_bc<V>(c, F.taub, t2, J, t);
        v[q] = (T)t[YT_BET] * (r * dt);
      }
  } else {
    // the two operand phases of yfuse_u: the convective and viscous terms,
    // then the SGS operands (nuz at J0-1..J0+len, dv/dz at J0..J0+len, and
    // the sums of SGSVIS at (I,0,K) + (I,0,K-1) and (I,N2,K) + (I,N2,K-1))
    V h[CH], r0[CH];
#pragma unroll
    for (int q = 0; q < CH; q++)
      if (q < y.len) {
        const int J = y.j0 + q;
        const V wm = ws[q], w0 = ws[q + 1], wp = ws[q + 2];
        const T* t = tb + J * YTAB_W;
        r0[q] = ru[q] + rinv * ((T)t[0] * wp - (T)t[1] * w0 + (T)t[2] * wm);
        const V w2 = (T)t[3] * ((T)t[7] * w0 + (T)t[6] * wp) + (T)t[9] * wN2;
        const V w1 = (T)t[4] * ((T)t[6] * wm + (T)t[5] * w0) + (T)t[10] * wW;
        const V v2 = (J == g.n2m) ? zero : s3[q + 1];
        const V v1 = (J == 1) ? zero : s3[q];
        h[q] = hx[q] + (T)t[8] * (v2 * w2 - v1 * w1);
      }
    // the barrier keeps the second phase's loads after the first phase's
    // arithmetic (512-thread pair tiles 165 -> 153 VGPRs, 3.6 % faster)
    __builtin_amdgcn_sched_barrier(0);
    const T* __restrict__ nuz = (const T*)F.nuz;
    const T* __restrict__ dvdz = (const T*)F.dvdz;
    const T* __restrict__ sg = (const T*)F.ssgs;
    V nz[CH + 2], dz[CH + 1], sgb = zero, sgt = zero;
#pragma unroll
    for (int q = 0; q < CH + 2; q++) nz[q] = (q < y.len + 2) ? vld<V>(nuz, a0 - nx + q * nx) : zero;
#pragma unroll
    for (int q = 0; q < CH + 1; q++) dz[q] = (q < y.len + 1) ? vld<V>(dvdz, a0 + q * nx) : zero;
    const unsigned bkm = b + dk;
    if (y.j0 == 1) sgb = vld<V>(sg, b) + vld<V>(sg, bkm);
    if (y.j0 + y.len - 1 == g.n2m) sgt = vld<V>(sg, b + g.n2 * nx) + vld<V>(sg, bkm + g.n2 * nx);
#pragma unroll
    for (int q = 0; q < CH; q++)
      if (q < y.len) {
        const int J = y.j0 + q;
        const V wm = ws[q], w0 = ws[q + 1], wp = ws[q + 2];
        const T* t = tb + J * YTAB_W;
        // D(SGS*(DW/DY))/DY and D(SGS*(DV/DZ))/DY (:7858-7908, :8042-8094),
        // as yterm_w
        const V z0 = nz[q + 1], zp1 = (J < g.n2m) ? nz[q + 2] : zero, zm1 = (J > 1) ? nz[q] : zero;
        const V sgsj1 = (T)t[4] * ((T)t[5] * z0 + (T)t[6] * zm1) + (T)t[10] * (T)0.5 * sgb;
        const V sgsj2 = (T)t[3] * ((T)t[6] * zp1 + (T)t[7] * z0) + (T)t[9] * (T)0.5 * sgt;
        const T wfu = (T)1 - (T)c.s.ifwall * (T)t[9] * (T)(c.s.aly - 1.);
        const T wfd = (T)1 - (T)c.s.ifwall * (T)t[10];
        const V ww22 = (T)t[15] * (wp - w0);
        const V ww11 = (T)t[14] * (w0 - wm);
        const V hd = h[q] - (T)t[8] * (wfu * sgsj2 * ww22 - wfd * sgsj1 * ww11);
        const V dzp = (J < g.n2m) ? dz[q + 1] : zero;
        const V hs = hd - (T)t[8] * (wfu * sgsj2 * dzp - wfd * sgsj1 * dz[q]);
        vst_nt<V>(hh, a0 + q * nx, hs);
        const V r = r0[q] - (T)(3. / 2.) * hs + (T)(1. / 2.) * hq[q] + wall_bc<V>(c, F.taub, t2, J, t);
        v[q] = (T)t[YT_BET] * (r * dt);
      }
  }
}

// NT: the block size bound, 512 or 1024 (tiles of L*C threads); V: T, or
// f2v for FP32 tiles of column pairs (YShape.P = 2: L threads cover 2L
// columns)
template <typename T, typename V, bool SGS, int CH, int NT, bool CB_>
__global__ __launch_bounds__(NT) void k_rhsy_adiy(CtxArg<CB_> ca_, int staged) {
  constexpr int PR = YCol<V>::P;
  const Ctx c = ctx_get(ca_);
  __shared__ V xs[2][NT];  // C*L <= NT entries each
  __shared__ T xp[2][NT];
  extern __shared__ __align__(16) unsigned char smem[];
  const Geo g = c.g;
  const YShape ys = sizeof(T) == 4 ? c.o.ys32 : c.o.ys;
  const int comp = blockIdx.y;
  const int ns = (comp == 1) ? 2 : 1;
  const int L = ys.L, C = ys.C;
  const int t = threadIdx.x, l = t % L;
  YPos y;
  y.l = l;
  y.c = t / L;
  const int m = g.n2m - ns + 1;
  const int ch = (m + C - 1) / C;
  y.j0 = ns + y.c * ch;
  y.len = max(0, min(ch, g.n2m + 1 - y.j0));
  const int nit = (g.nx + L * PR - 1) / (L * PR);
  const int it = blockIdx.x % nit;
  const int k = blockIdx.x / nit;
  const int i = (it * L + l) * PR;
  const bool ok = i < g.nx;  // NX is even, so a pair is whole
  const Fields F = c.d;
  const Metrics M = c.m;
  T* X = (T*)(comp == 0 ? F.uh[0] : comp == 1 ? F.uh[1] : F.uh[2]);
  const T dt = (T)c.s.dt;
  const unsigned b = (unsigned)gidx(g, ok ? i : 0, 0, k);
  const unsigned t2 = (unsigned)k * g.nx + (ok ? i : 0);
  // this component's coefficient rows (in the sweep's precision), in LDS
  // when they fit
  const int ntab = (g.n2 + 2) * YTAB_W;
  const T* tb = (sizeof(T) == 4 ? (const T*)M.ytab32 : (const T*)M.ytab) + (long long)comp * ntab;
  if (staged) {
    T* st = (T*)smem;
    for (int e = threadIdx.x; e < ntab; e += blockDim.x) st[e] = tb[e];
    __syncthreads();
    tb = st;
  }
  V v[CH];
#pragma unroll
  for (int q = 0; q < CH; q++) v[q] = vsplat<V>((T)0);
  if (ok && y.len > 0) {
    if (comp == 0) {
      yfuse_u<T, V, CH, SGS>(c, tb, b, t2, y, v, (i == 0 ? g.nx - 1 : i - 1) - i);
    } else if (comp == 1) {
      yfuse_v<T, V, CH, SGS>(c, tb, b, y, v);
    } else {
      yfuse_w<T, V, CH, SGS>(c, tb, b, t2, y, v, ((k == 0 ? g.nz - 1 : k - 1) - k) * (int)g.plane);
    }
  }
  ychain_fwd(v, [&](int q) { return (T)tb[(y.j0 + q) * YTAB_W + YT_M]; }, y, L, xs[0], xp[0]);
  ychain_bwd(v, [&](int q) { return (T)tb[(y.j0 + q) * YTAB_W + YT_G]; }, y, C, L, xs[1], xp[1]);
  if (ok)
#pragma unroll
    for (int q = 0; q < CH; q++)
      if (q < y.len) vst<V>(X, b + (unsigned)(y.j0 + q) * g.nx, v[q]);
}

// --------------------------------------------------- GETUH x / z sweeps
// CTDMA1J (x, :8285-8311) / CTDMA3I (z, :8314-8338): solve
// (A, B, A) x = UH(p-1) + 10 UH(p) + UH(p+1); `fin` adds U (UH = U + delta,
// :8610-8624).  blockIdx.y indexes the component list.
struct CompList {
  int n;
  int comp[3];
  int fin[3];
};

template <typename T, int SEG, bool CB_>
__global__ __launch_bounds__(256) void k_adi_x(CtxArg<CB_> ca_, CompList cl) {
  const Ctx c = ctx_get(ca_);
  XSETUP;
  const int sel = blockIdx.y;
  const int comp = sel == 0 ? cl.comp[0] : sel == 1 ? cl.comp[1] : cl.comp[2];
  const bool fin = (sel == 0 ? cl.fin[0] : sel == 1 ? cl.fin[1] : cl.fin[2]) != 0;
  const Fields F = c.d;
  T* UH = (T*)(comp == 0 ? F.uh[0] : comp == 1 ? F.uh[1] : F.uh[2]);
  const double* U = comp == 0 ? F.u : comp == 1 ? F.v : F.w;
  T d[SEG], pf[SEG];
  // mixed mode keeps the increment DELTA in UH (no U added)
  const bool addu = fin && sizeof(T) == 8;
  X.pre((const T*)UH, 0, pf);
  const T* B = X.put(pf);
  if (addu) X.pre(U, 0, pf);
  SQ d[q] = XA(B, q - 1) + (T)10. * XA(B, q) + XA(B, q + 1);
  psolve<T, SEG>(d, G, L, l, gs, c.o.px.adi, bf.xf, bf.xb);
  // the flow-rate terms of UPCALC for w (:1160-1172): sum DELTA*DY and U*DY
  double sd = 0.0, su = 0.0;
  SQ sd += (double)d[q];
  if (addu) {
    B = X.put(pf);
    SQ {
      const T uq = XA(B, q);
      su += (double)uq;
      d[q] = uq + d[q];
    }
  }
  X.out(UH, d, comp == 1 ? 2 : 1);  // V rows J=1 are not part of the system
  if (fin && comp == 2) {
    const double dy = ok ? c.m.dy[j] : 0.0;
    block_sum2(sd * dy, su * dy, F.flowp + FP_DW * F.flowr + blockIdx.x, F.flowp + FP_W * F.flowr + blockIdx.x);
  }
}

template <typename T, int SEG, bool CB_>
__global__ __launch_bounds__(512) void k_adi_z(CtxArg<CB_> ca_, CompList cl) {
  const Ctx c = ctx_get(ca_);
  ZSETUP;
  const int sel = blockIdx.y;
  const int comp = sel == 0 ? cl.comp[0] : sel == 1 ? cl.comp[1] : cl.comp[2];
  const bool fin = (sel == 0 ? cl.fin[0] : sel == 1 ? cl.fin[1] : cl.fin[2]) != 0;
  if (comp == 1 && j < 2) return;  // uniform per block
  const Fields F = c.d;
  T* UH = (T*)(comp == 0 ? F.uh[0] : comp == 1 ? F.uh[1] : F.uh[2]);
  const double* U = comp == 0 ? F.u : comp == 1 ? F.v : F.w;
  T a[SEG], d[SEG], lf, rt;
  zload((const T*)UH, b, pl, p0, bm, bp, a, lf, rt);
  SQ d[q] = prv(a, q, lf) + (T)10. * a[q] + nxt(a, q, rt);
  psolve<T, SEG>(d, G, L, l, gs, PZ.adi, bf.xf, bf.xb);
  // the flow-rate terms of UPCALC for u (:1079-1090): sum DELTA*DY and U*DY
  double sd = 0.0, su = 0.0;
  if (ok) {
    if (fin && sizeof(T) == 8) {  // mixed mode keeps the increment DELTA
      SQ {
        const T uq = (T)U[ZI(q)];
        sd += (double)d[q];
        su += (double)uq;
        st_nt(&UH[ZI(q)], (T)(uq + d[q]));
      }
    } else {
      SQ {
        sd += (double)d[q];
        st_nt(&UH[ZI(q)], (T)(d[q]));
      }
    }
  }
  if (fin && comp == 0) {
    const double dy = c.m.dy[j];
    block_sum2(sd * dy, su * dy, F.flowp + FP_DU * F.flowr + blockIdx.x, F.flowp + FP_U * F.flowr + blockIdx.x);
  }
}

// Mixed mode (FP32 sweeps): the x and z ADI factors act on different indices
// with coefficients that depend on that index only, so they commute, and
// the sweeps run in the order y, z, x for u and v and y, x, z for w.  The
// last sweep of u and of w then holds the final increments DELTA in
// registers, and RHSDP (:8117-8214) is formed there instead of in two passes
// of its own: the x-sweep forms RDP = D1x(DELTA_u)/DT, the z-sweep of w adds
// the y term, DIV(U)/DT of the current state and D1z(DELTA_w)/DT, in FP64
// arithmetic from the FP32 increments.  The flow-rate sums of UPCALC come
// from the same sweeps (u: x-blocks, w: z-blocks).
//
// x-lines, the last sweep of u and v and the first of w (blockIdx.y: u, v,
// w).  The FP64 D1x of u reuses the block's LDS: the two FP32 staging tiles
// as one FP64 tile and the four FP32 exchange arrays as two FP64 ones.
template <int SEG, bool CB_>
__global__ __launch_bounds__(256) void k_adi_xr(CtxArg<CB_> ca_) {
  typedef float T;
  const Ctx c = ctx_get(ca_);
  XSETUP_NT;
  const int comp = blockIdx.y;
  const Fields F = c.d;
  T* UH = (T*)(comp == 0 ? F.uh[0] : comp == 1 ? F.uh[1] : F.uh[2]);
  T d[SEG], pf[SEG];
  X.pre((const T*)UH, 0, pf);
  const T* B = X.put(pf);
  SQ d[q] = XA(B, q - 1) + (T)10. * XA(B, q) + XA(B, q + 1);
  psolve<T, SEG>(d, G, L, l, gs, c.o.px.adi, bf.xf, bf.xb);
  X.out(UH, d, comp == 1 ? 2 : 1);  // V rows J=1 are not part of the system
  if (comp != 0) return;  // uniform per block
  // u: the flow-rate term of UPCALC (:1079-1090), sum DELTA*DY
  double sd = 0.0;
  SQ sd += (double)d[q];
  block_sum2(sd * (ok ? c.m.dy[j] : 0.0), 0.0, F.flowp + FP_DU * F.flowr + blockIdx.x, nullptr);
  // RDP = D1x(DELTA_u)/DT, face -> centre (:8138-8186)
  double* const x8f = (double*)bf.xf;  // G*L doubles over xf, xb
  double* const x8b = (double*)bf.xl;  // G*L doubles over xl, xr
  double e[SEG];
  __syncthreads();  // the FP32 exchange arrays are read no more
  const double rt = xch_right((double)d[0], G, L, l, gs, x8b);
  SQ e[q] = 24. * c.s.dx1 * ((double)nxt(d, q, (T)rt) - (double)d[q]);
  psolve<double, SEG>(e, G, L, l, gs, c.o.px.s22, x8f, x8b);
  const double rdt = 1. / c.s.dt;
  SQ e[q] = e[q] * rdt;
  XStage<double, SEG> X8;
  const LBuf<double> b8{(double*)bf.st0, (double*)bf.st0, x8f, x8b, nullptr, nullptr};
  xstage(X8, b8, g, sh, l, p0);
  X8.out(F.rdp, e);
}

// z-lines: the last sweep of w, then RDP += ((FJPA DELTA_v(J+1) -
// FJMU DELTA_v(J))/DY)/DT + DIV(U)/DT + D1z(DELTA_w)/DT (:8138-8211; the
// CBC term is zero, UBC(.,.,2,.) = 0)
template <int SEG, bool CB_>
__global__ __launch_bounds__(512) void k_adi_zr(CtxArg<CB_> ca_) {
  typedef float T;
  const Ctx c = ctx_get(ca_);
  ZSETUP;
  const Fields F = c.d;
  const Metrics M = c.m;
  T* UH = (T*)F.uh[2];
  T a[SEG], d[SEG], lf, rt;
  zload((const T*)UH, b, pl, p0, bm, bp, a, lf, rt);
  SQ d[q] = prv(a, q, lf) + (T)10. * a[q] + nxt(a, q, rt);
  psolve<T, SEG>(d, G, L, l, gs, PZ.adi, bf.xf, bf.xb);
  double sd = 0.0;
  if (ok) SQ {
      sd += (double)d[q];
      st_nt(&UH[ZI(q)], (T)(d[q]));
    }
  // w: the flow-rate term of UPCALC (:1160-1172), sum DELTA*DY
  block_sum2(sd * M.dy[j], 0.0, F.flowp + FP_DW * F.flowr + blockIdx.x, nullptr);
  double* const x8f = (double*)bf.xf;
  double* const x8b = (double*)bf.xl;
  double e[SEG];
  __syncthreads();  // the FP32 exchange arrays are read no more
  const double r8 = xch_right((double)d[0], G, L, l, gs, x8b);
  SQ e[q] = 24. * c.s.dx3 * ((double)nxt(d, q, (T)r8) - (double)d[q]);
  psolve<double, SEG>(e, G, L, l, gs, PZ.s22, x8f, x8b);
  if (!ok) return;
  const T* dv = (const T*)F.uh[1];
  double* rdp = F.rdp;
  const float* dvu = F.divf;
  const double rdt = 1. / c.s.dt, rdy = M.rdy[j], fjpa = M.fjpa[j], fjmu = M.fjmu[j];
  const unsigned nx = g.nx;
  SQ {
    const unsigned a0 = ZI(q);
    const double ty = ((fjpa * (double)dv[a0 + nx] - fjmu * (double)dv[a0]) * rdy) * rdt;
    rdp[a0] = ((rdp[a0] + ty) + (double)dvu[a0] * rdt) + e[q] * rdt;
  }
}

// ------------------------------------------------------------------ RHSDP
// FP64 mode (mixed mode forms RDP in its last ADI sweeps, k_adi_xr and
// k_adi_zr).  x-lines: RDP = (FJUP*UH2(J+1) - FJUM*UH2(J))/DY/DT +
// D1x(UH1)/DT (:8138-8186; the CBC term is zero: UBC(.,.,2,.) = 0, :978,
// :986).
template <typename T, int SEG, bool CB_>
__global__ __launch_bounds__(256) void k_rhsdp_x(CtxArg<CB_> ca_) {
  const Ctx c = ctx_get(ca_);
  XSETUP;
  const Fields F = c.d;
  const Metrics M = c.m;
  const T dx1 = (T)c.s.dx1, dt = (T)c.s.dt;
  T d[SEG], e[SEG], pf[SEG];
  X.pre_raw((const T*)F.uh[0], 0, pf);
  const T* B = X.put(pf);
  X.pre_raw((const T*)F.uh[1], 1, pf);
  SQ d[q] = (T)24. * dx1 * (XA(B, q + 1) - XA(B, q));
  psolve<T, SEG>(d, G, L, l, gs, c.o.px.s22, bf.xf, bf.xb);
  const T fjpa = (T)M.fjpa[j], fjmu = (T)M.fjmu[j];
  B = X.put(pf);
  X.pre_raw((const T*)F.uh[1], 0, pf);
  SQ e[q] = fjpa * XA(B, q);
  B = X.put(pf);
  // reciprocals instead of three divisions per point
  const T rdy = (T)M.rdy[j], rdt = (T)1. / dt;
  SQ d[q] = ((e[q] - fjmu * XA(B, q)) * rdy) * rdt + d[q] * rdt;
  X.out(F.rdp, d);
}

// z-lines: RDP += D1z(UH3)/DT (:8196-8211)
template <typename T, int SEG, bool CB_>
__global__ __launch_bounds__(512) void k_rhsdp_z(CtxArg<CB_> ca_) {
  const Ctx c = ctx_get(ca_);
  ZSETUP;
  const T dx3 = (T)c.s.dx3, dt = (T)c.s.dt;
  T a[SEG], d[SEG], lf, rt;
  zload((const T*)c.d.uh[2], b, pl, p0, bm, bp, a, lf, rt);
  SQ d[q] = (T)24. * dx3 * (nxt(a, q, rt) - a[q]);
  psolve<T, SEG>(d, G, L, l, gs, PZ.s22, bf.xf, bf.xb);
  double* rdp = c.d.rdp;
  const T rdt = (T)1. / dt;
  if (ok) SQ st_nt(&rdp[ZI(q)], rdp[ZI(q)] + d[q] * rdt);
}

// ------------------------------------------------------------------ GETDP
// The plane DFT of :9325-9346 is factored as a real x-DFT (two rows packed
// into one complex transform: z = a + i b, A[k] = (Z[k] + conj Z[n-k])/2,
// B[k] = (Z[k] - conj Z[n-k])/(2i)), keeping kx = 0..NX/2, then complex
// z-DFTs of that half spectrum.  The x-modes 0 and NX/2 of a real row are
// real, so they share one stored column: X(0) + i X(NX/2) at kx = 0, and the
// spectrum has NX/2 columns (the tiles of the z-DFT and the Poisson solve
// divide them evenly).  Spectrum layout [jm][kz][kx], rows of NXSP >= NX/2
// (an odd multiple of 256 bytes).
template <typename T, int R>
__device__ __forceinline__ void dft_r(cplx_t<T> (&v)[R], int sign) {
  if (R == 2) {
    cplx_t<T> a = v[0], b = v[1];
    v[0] = {a.x + b.x, a.y + b.y};
    v[1] = {a.x - b.x, a.y - b.y};
  } else if (R == 4) {
    cplx_t<T> a0 = {v[0].x + v[2].x, v[0].y + v[2].y};
    cplx_t<T> a1 = {v[0].x - v[2].x, v[0].y - v[2].y};
    cplx_t<T> a2 = {v[1].x + v[3].x, v[1].y + v[3].y};
    cplx_t<T> a3 = {v[1].x - v[3].x, v[1].y - v[3].y};
    cplx_t<T> a3r = (sign < 0) ? cplx_t<T>{a3.y, -a3.x} : cplx_t<T>{-a3.y, a3.x};
    v[0] = {a0.x + a2.x, a0.y + a2.y};
    v[2] = {a0.x - a2.x, a0.y - a2.y};
    v[1] = {a1.x + a3r.x, a1.y + a3r.y};
    v[3] = {a1.x - a3r.x, a1.y - a3r.y};
  } else if (R == 3) {
    const T c1 = (T)-0.5;
    const T s1 = (T)(sign < 0 ? -0.86602540378443864676 : 0.86602540378443864676);
    cplx_t<T> a = v[0], b = v[1], cc = v[2];
    cplx_t<T> t1 = {b.x + cc.x, b.y + cc.y};
    cplx_t<T> t2 = {a.x + c1 * t1.x, a.y + c1 * t1.y};
    cplx_t<T> t3 = {s1 * (b.x - cc.x), s1 * (b.y - cc.y)};
    v[0] = {a.x + t1.x, a.y + t1.y};
    v[1] = {t2.x - t3.y, t2.y + t3.x};
    v[2] = {t2.x + t3.y, t2.y - t3.x};
  } else if (R == 5) {
    const T cs[5] = {(T)1.0, (T)0.30901699437494742410, (T)-0.80901699437494742410,
                     (T)-0.80901699437494742410, (T)0.30901699437494742410};
    const T sn[5] = {(T)0.0, (T)0.95105651629515357212, (T)0.58778525229247312917,
                     (T)-0.58778525229247312917, (T)-0.95105651629515357212};
    cplx_t<T> in[5];
#pragma unroll
    for (int q = 0; q < 5; q++) in[q] = v[q];
#pragma unroll
    for (int p = 0; p < 5; p++) {
      T ax = 0, ay = 0;
#pragma unroll
      for (int q = 0; q < 5; q++) {
        const int mq = (p * q) % 5;
        const T cq = cs[mq], sq = (sign < 0 ? -sn[mq] : sn[mq]);
        ax += in[q].x * cq - in[q].y * sq;
        ay += in[q].x * sq + in[q].y * cq;
      }
      v[p] = {ax, ay};
    }
  }
}

// one Stockham stage of radix R over nl lines of n points held in LDS at
// a[p * ls + l], in place: every thread first gathers its butterflies into
// registers (n*nl <= NE complex values per block of NT
// threads), then all write back after a barrier.  One buffer instead of two
// halves the LDS of a block.
template <typename T, int R, int NE, int NT>
__device__ __forceinline__ void fft_stage(cplx_t<T>* a, int n, int nl, int ls, int ns, const cplx_t<T>* tw,
                                          int sign) {
  constexpr int IT = (NE / R + NT - 1) / NT;
  const int nb = n / R;
  const int tot = nb * nl;
  const int tq = n / (ns * R);
  cplx_t<T> v[IT][R];
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int e = threadIdx.x + it * NT;
    if (e < tot) {
      const int l = e % nl;
      const int jb = e / nl;
      const int ts = (jb % ns) * tq;  // twiddle step; ts*r < n
#pragma unroll
      for (int r = 0; r < R; r++) {
        const cplx_t<T> x = a[(jb + r * nb) * ls + l];
        v[it][r] = (r == 0) ? x : cmul(x, tw[ts * r]);
      }
      dft_r<T, R>(v[it], sign);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int e = threadIdx.x + it * NT;
    if (e < tot) {
      const int l = e % nl;
      const int jb = e / nl;
      const int d = (jb / ns) * ns * R + jb % ns;
#pragma unroll
      for (int r = 0; r < R; r++) a[(d + r * ns) * ls + l] = v[it][r];
    }
  }
  __syncthreads();
}

// mixed-radix transform in place; radices packed 3 bits each in `code`,
// `nrad` stages
template <typename T, int NE, int NT>
__device__ __forceinline__ void fft_lines(cplx_t<T>* a, int n, int nl, int ls, int code, int nrad,
                                          const cplx_t<T>* tw, int sign) {
  int ns = 1;
  for (int s = 0; s < nrad; s++) {
    const int R = (code >> (3 * s)) & 7;
    if (R == 4)
      fft_stage<T, 4, NE, NT>(a, n, nl, ls, ns, tw, sign);
    else if (R == 2)
      fft_stage<T, 2, NE, NT>(a, n, nl, ls, ns, tw, sign);
    else if (R == 3)
      fft_stage<T, 3, NE, NT>(a, n, nl, ls, ns, tw, sign);
    else
      fft_stage<T, 5, NE, NT>(a, n, nl, ls, ns, tw, sign);
    ns *= R;
  }
}

// ------------------------------------------- two-pass register FFT
// For n = N1*N2 in {64 = 8*8, 128 = 8*16, 256 = 16*16} a line is split as
// p = a + N1*b: pass A (thread a of the line) takes the N2 points
// x[a + N1 b] from HBM (or the caller) into registers, transforms over b,
// multiplies by W_n^(a k2), and writes Y[a][k2] to LDS; after one barrier
// pass B (thread c) reads Y[.][k2] for its k2 = c, c + N1, ... and
// transforms over a, giving X[k2 + N2 k1].  One LDS exchange per transform
// (the Stockham kernels above make a round trip per radix stage).

// DFT of 4 values in place (x[m] at v[s*m], natural order out)
template <typename T>
__device__ __forceinline__ void rdft4(cplx_t<T>& v0, cplx_t<T>& v1, cplx_t<T>& v2, cplx_t<T>& v3, int sign) {
  const cplx_t<T> a0 = {v0.x + v2.x, v0.y + v2.y};
  const cplx_t<T> a1 = {v0.x - v2.x, v0.y - v2.y};
  const cplx_t<T> a2 = {v1.x + v3.x, v1.y + v3.y};
  const cplx_t<T> a3 = {v1.x - v3.x, v1.y - v3.y};
  const cplx_t<T> a3r = (sign < 0) ? cplx_t<T>{a3.y, -a3.x} : cplx_t<T>{-a3.y, a3.x};
  v0 = {a0.x + a2.x, a0.y + a2.y};
  v2 = {a0.x - a2.x, a0.y - a2.y};
  v1 = {a1.x + a3r.x, a1.y + a3r.y};
  v3 = {a1.x - a3r.x, a1.y - a3r.y};
}
// exp(sign 2 pi i k / 16), k = 0..15
template <typename T>
__device__ __forceinline__ cplx_t<T> w16(int k, int sign) {
  const T c[5] = {(T)1.0, (T)0.92387953251128675613, (T)0.70710678118654752440, (T)0.38268343236508977173,
                  (T)0.0};
  // cos and sin of k*pi/8 from the first quadrant
  const int q = k & 3, h = (k >> 2) & 3;
  T cr = c[q], si = c[4 - q];
  T x, y;
  if (h == 0) { x = cr; y = si; }
  else if (h == 1) { x = -si; y = cr; }
  else if (h == 2) { x = -cr; y = -si; }
  else { x = si; y = -cr; }
  return {x, sign < 0 ? -y : y};
}
// in-register DFT of N = 8 or 16 values, natural order in and out
template <typename T, int N>
__device__ __forceinline__ void rdft(cplx_t<T> (&v)[N], int sign) {
  static_assert(N == 8 || N == 16, "register DFT of 8 or 16 points");
  constexpr int R = N / 4;  // n = r + R m: R DFT4s over m, then R-point DFTs over r
  cplx_t<T> y[N];
#pragma unroll
  for (int r = 0; r < R; r++) {
    cplx_t<T> t0 = v[r], t1 = v[r + R], t2 = v[r + 2 * R], t3 = v[r + 3 * R];
    rdft4(t0, t1, t2, t3, sign);
    y[r * 4 + 0] = t0;
    y[r * 4 + 1] = t1;
    y[r * 4 + 2] = t2;
    y[r * 4 + 3] = t3;
  }
  // twiddles W_N^(r s) = W_16^(r s 16/N)
#pragma unroll
  for (int r = 1; r < R; r++)
#pragma unroll
    for (int sx = 1; sx < 4; sx++) y[r * 4 + sx] = cmul(y[r * 4 + sx], w16<T>((r * sx * (16 / N)) & 15, sign));
  // X[s + 4 t] = sum_r y[r][s] W_R^(r t)
#pragma unroll
  for (int sx = 0; sx < 4; sx++) {
    if (R == 4) {
      cplx_t<T> t0 = y[sx], t1 = y[4 + sx], t2 = y[8 + sx], t3 = y[12 + sx];
      rdft4(t0, t1, t2, t3, sign);
      v[sx] = t0;
      v[sx + 4] = t1;
      v[sx + 8] = t2;
      v[sx + 12] = t3;
    } else {
      const cplx_t<T> a = y[sx], b = y[4 + sx];
      v[sx] = {a.x + b.x, a.y + b.y};
      v[sx + 4] = {a.x - b.x, a.y - b.y};
    }
  }
}

template <int N>
struct RFFT;
template <>
struct RFFT<64> {
  static constexpr int N1 = 8, N2 = 8;
};
template <>
struct RFFT<128> {
  static constexpr int N1 = 8, N2 = 16;
};
template <>
struct RFFT<256> {
  static constexpr int N1 = 16, N2 = 16;
};

// LDS slot of point p of line l.  Column layout (z-lines, the line index l
// fastest across a wavefront): p*(L+1) + l.  Row layout (x-lines, the
// point index fastest): l*(N + N/N1) + p + p/N1, one pad slot per N1
// points so that pass B's lanes (stride N1 points) fall on distinct banks.
template <int N, bool ROW>
__device__ __forceinline__ int rslot(int p, int l) {
  constexpr int N1 = RFFT<N>::N1, L = 256 / N1;
  return ROW ? l * (N + N / N1) + p + p / N1 : p * (L + 1) + l;
}
template <int N, bool ROW>
constexpr size_t rfft_lds() {
  return sizeof(double) * 2 * (size_t)(256 / RFFT<N>::N1) * (size_t)(N + N / RFFT<N>::N1 + 1) +
         sizeof(double) * 2 * N;
}

// pass A over the values v[b] = x[a + N1 b] of this thread: transform,
// twiddle, store Y[a][k2] at the slot of point a + N1 k2
template <typename T, int N, bool ROW>
__device__ __forceinline__ void rfft_a(cplx_t<T> (&v)[RFFT<N>::N2], int a, int l, cplx_t<T>* A,
                                       const cplx_t<T>* TW, int sign) {
  constexpr int N1 = RFFT<N>::N1, N2 = RFFT<N>::N2;
  rdft<T, N2>(v, sign);
#pragma unroll
  for (int k2 = 1; k2 < N2; k2++) v[k2] = cmul(v[k2], TW[a * k2]);
#pragma unroll
  for (int k2 = 0; k2 < N2; k2++) A[rslot<N, ROW>(a + N1 * k2, l)] = v[k2];
}
// pass B for k2 = c + N1 m (m < N2/N1): w[m][k1] = X[k2 + N2 k1]
template <typename T, int N, bool ROW>
__device__ __forceinline__ void rfft_b(cplx_t<T> (&w)[RFFT<N>::N2 / RFFT<N>::N1][RFFT<N>::N1], int c, int l,
                                       const cplx_t<T>* A, int sign) {
  constexpr int N1 = RFFT<N>::N1, N2 = RFFT<N>::N2, M = N2 / N1;
#pragma unroll
  for (int m = 0; m < M; m++) {
    const int k2 = c + N1 * m;
#pragma unroll
    for (int a = 0; a < N1; a++) w[m][a] = A[rslot<N, ROW>(a + N1 * k2, l)];
    rdft<T, N1>(w[m], sign);
  }
}

// z-DFT of the half spectrum (as k_fft_z) with the register FFT: L columns
// kx per block, N1 threads per column, the N2 values of pass A read from
// HBM and the results of pass B written to HBM directly
template <typename T, int N, bool CB_>
__global__ __launch_bounds__(256) void k_rfft_z(CtxArg<CB_> ca_, int sign) {
  const Ctx c = ctx_get(ca_);
  constexpr int N1 = RFFT<N>::N1, N2 = RFFT<N>::N2, M = N2 / N1;
  constexpr int L = 256 / N1;
  extern __shared__ __align__(16) unsigned char smem[];
  typedef cplx_t<T> zc;
  const Geo g = c.g;
  zc* A = (zc*)smem;
  zc* TW = A + (size_t)L * (N + N / N1 + 1);
  {
    const zc* twg = (const zc*)(sign < 0 ? c.f.twzf : c.f.twzb);
    for (int e = threadIdx.x; e < N; e += 256) TW[e] = twg[e];
  }
  const int nit = (g.nxs + L - 1) / L;
  const int it = blockIdx.x % nit;
  const int jm = blockIdx.x / nit;
  const int l = threadIdx.x % L, a = threadIdx.x / L;
  const int kx = it * L + l;
  const bool ok = kx < g.nxs;
  zc* S = (zc*)c.d.spec + (long long)jm * g.nz * g.nxsp + (ok ? kx : 0);
  const unsigned rs = (unsigned)g.nxsp;
  zc v[N2];
#pragma unroll
  for (int b = 0; b < N2; b++) v[b] = S[(unsigned)(a + N1 * b) * rs];
  __syncthreads();  // the twiddle table
  rfft_a<T, N, false>(v, a, l, A, TW, sign);
  __syncthreads();
  zc w[M][N1];
  rfft_b<T, N, false>(w, a, l, A, sign);
  if (ok)
#pragma unroll
    for (int m = 0; m < M; m++)
#pragma unroll
      for (int k1 = 0; k1 < N1; k1++) S[(unsigned)(a + N1 * m + N2 * k1) * rs] = w[m][k1];
}

// forward real x-DFT of RDP rows (as k_fft_x) with the register FFT: L
// complex lines (2L rows) per block; pass A reads x[a + N1 b] of both rows
// straight from HBM, the post-processing pairs X[kx] with X[n-kx] through
// LDS
template <typename T, int N, bool CB_>
__global__ __launch_bounds__(256) void k_rfft_x(CtxArg<CB_> ca_) {
  const Ctx c = ctx_get(ca_);
  constexpr int N1 = RFFT<N>::N1, N2 = RFFT<N>::N2, M = N2 / N1;
  constexpr int L = 256 / N1;
  extern __shared__ __align__(16) unsigned char smem[];
  typedef cplx_t<T> zc;
  const Geo g = c.g;
  zc* A = (zc*)smem;
  zc* TW = A + (size_t)L * (N + N / N1 + 1);
  for (int e = threadIdx.x; e < N; e += 256) TW[e] = ((const zc*)c.f.twxf)[e];
  const int nlines = g.n2m * g.nz;
  const int ln0 = blockIdx.x * 2 * L;
  // thread (a, l): a fastest, so a wavefront reads 16 or 8 consecutive points
  const int a = threadIdx.x % N1, l = threadIdx.x / N1;
  const int la = ln0 + l, lb = ln0 + L + l;
  const T* src = (const T*)c.d.rdp;
  const unsigned oa = la < nlines ? (unsigned)gidx(g, 0, 1 + la % g.n2m, la / g.n2m) : 0u;
  const unsigned ob = lb < nlines ? (unsigned)gidx(g, 0, 1 + lb % g.n2m, lb / g.n2m) : 0u;
  zc v[N2];
#pragma unroll
  for (int b = 0; b < N2; b++) {
    const unsigned p = (unsigned)(a + N1 * b);
    v[b] = {la < nlines ? src[oa + p] : (T)0, lb < nlines ? src[ob + p] : (T)0};
  }
  __syncthreads();  // the twiddle table
  rfft_a<T, N, true>(v, a, l, A, TW, -1);
  __syncthreads();
  zc w[M][N1];
  rfft_b<T, N, true>(w, a, l, A, -1);
  __syncthreads();  // every pass-B read is done before the results overwrite A
#pragma unroll
  for (int m = 0; m < M; m++)
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) A[rslot<N, true>(a + N1 * m + N2 * k1, l)] = w[m][k1];
  __syncthreads();
  const int nxs = g.nxs;
  zc* S = (zc*)c.d.spec;
  for (int e = threadIdx.x; e < L * nxs; e += 256) {
    const int ll = e / nxs, kk = e - ll * nxs;
    const int ea = ln0 + ll, eb = ln0 + L + ll;
    const zc z1 = A[rslot<N, true>(kk, ll)];
    const zc z2 = A[rslot<N, true>(kk == 0 ? N / 2 : N - kk, ll)];
    // column 0: the real modes 0 and N/2 of each row, packed (Z(0) and Z(N/2)
    // hold X_a + i X_b there, all four real)
    const zc va = kk == 0 ? zc{z1.x, z2.x} : zc{(T)0.5 * (z1.x + z2.x), (T)0.5 * (z1.y - z2.y)};
    const zc vb = kk == 0 ? zc{z1.y, z2.y} : zc{(T)0.5 * (z1.y + z2.y), (T)-0.5 * (z1.x - z2.x)};
    if (ea < nlines) S[((long long)(ea % g.n2m) * g.nz + ea / g.n2m) * g.nxsp + kk] = va;
    if (eb < nlines) S[((long long)(eb % g.n2m) * g.nz + eb / g.n2m) * g.nxsp + kk] = vb;
  }
}

// inverse x-DFT (as k_fft_xinv) with the register FFT: pass A forms its
// inputs (Hermitian fill, two rows per complex line) from the half spectrum
// in HBM, pass B writes the real and imaginary parts / (N1M N3M) to DP
template <typename T, int N, bool CB_>
__global__ __launch_bounds__(256) void k_rfft_xinv(CtxArg<CB_> ca_) {
  const Ctx c = ctx_get(ca_);
  constexpr int N1 = RFFT<N>::N1, N2 = RFFT<N>::N2, M = N2 / N1;
  constexpr int L = 256 / N1;
  extern __shared__ __align__(16) unsigned char smem[];
  typedef cplx_t<T> zc;
  const Geo g = c.g;
  zc* A = (zc*)smem;
  zc* TW = A + (size_t)L * (N + N / N1 + 1);
  for (int e = threadIdx.x; e < N; e += 256) TW[e] = ((const zc*)c.f.twxb)[e];
  const int nlines = g.n2m * g.nz;
  const int ln0 = blockIdx.x * 2 * L;
  const int a = threadIdx.x % N1, l = threadIdx.x / N1;
  const int la = ln0 + l, lb = ln0 + L + l;
  const zc* S = (const zc*)c.d.spec;
  const long long sa = la < nlines ? ((long long)(la % g.n2m) * g.nz + la / g.n2m) * g.nxsp : -1;
  const long long sb = lb < nlines ? ((long long)(lb % g.n2m) * g.nz + lb / g.n2m) * g.nxsp : -1;
  zc v[N2];
#pragma unroll
  for (int b = 0; b < N2; b++) {
    const int kx = a + N1 * b;
    const bool rl = kx == 0 || 2 * kx == N;  // the real modes, packed in column 0
    const int kk = rl ? 0 : (2 * kx < N ? kx : N - kx);
    zc xa = sa >= 0 ? S[sa + kk] : zc{(T)0, (T)0};
    zc xb = sb >= 0 ? S[sb + kk] : zc{(T)0, (T)0};
    if (rl) {
      xa = {kx == 0 ? xa.x : xa.y, (T)0};
      xb = {kx == 0 ? xb.x : xb.y, (T)0};
    } else if (2 * kx > N) {
      xa.y = -xa.y;
      xb.y = -xb.y;
    }
    v[b] = {xa.x - xb.y, xa.y + xb.x};
  }
  __syncthreads();  // the twiddle table
  rfft_a<T, N, true>(v, a, l, A, TW, +1);
  __syncthreads();
  zc w[M][N1];
  rfft_b<T, N, true>(w, a, l, A, +1);
  const T norm = (T)(1.0 / ((double)g.nx * (double)g.nz));
  double* dp = c.d.dp;
  const unsigned oa = la < nlines ? (unsigned)gidx(g, 0, 1 + la % g.n2m, la / g.n2m) : 0u;
  const unsigned ob = lb < nlines ? (unsigned)gidx(g, 0, 1 + lb % g.n2m, lb / g.n2m) : 0u;
#pragma unroll
  for (int m = 0; m < M; m++)
#pragma unroll
    for (int k1 = 0; k1 < N1; k1++) {
      const unsigned p = (unsigned)(a + N1 * m + N2 * k1);
      const zc z = w[m][k1];
      if (la < nlines) dp[oa + p] = z.x * norm;
      if (lb < nlines) dp[ob + p] = z.y * norm;
    }
}

// forward x-DFT of RDP rows (J=1..N2M, K), 2*TL rows per block; line
// stride TL+1 in LDS so that the row-wise fill and drain hit distinct banks.
// The FFT kernels copy their twiddle table (n complex values) into LDS
// behind the tile, so the stages read twiddles from LDS.
template <typename T, int NE, int NT, bool CB_>
__global__ __launch_bounds__(NT) void k_fft_x(CtxArg<CB_> ca_) {
  const Ctx c = ctx_get(ca_);
  static_assert(sizeof(T) == 8, "the twiddle tables are FP64");
  extern __shared__ __align__(16) unsigned char smem[];
  typedef cplx_t<T> zc;
  const Geo g = c.g;
  const int n = g.nx, NL = c.f.tlx, LS = NL + 1;  // NL complex lines = 2*NL rows
  zc* A = (zc*)smem;
  zc* TW = A + (size_t)n * LS;
  for (int e = threadIdx.x; e < n; e += NT) TW[e] = ((const zc*)c.f.twxf)[e];
  const int nlines = g.n2m * g.nz;
  const int ln0 = blockIdx.x * 2 * NL;
  const T* src = (const T*)c.d.rdp;
#pragma unroll 4
  for (int e = threadIdx.x; e < NL * n; e += NT) {
    const int l = e / n, p = e - l * n;
    const int la = ln0 + l, lb = ln0 + NL + l;
    const T xa = la < nlines ? src[gidx(g, p, 1 + la % g.n2m, la / g.n2m)] : (T)0;
    const T xb = lb < nlines ? src[gidx(g, p, 1 + lb % g.n2m, lb / g.n2m)] : (T)0;
    A[p * LS + l] = {xa, xb};
  }
  __syncthreads();
  fft_lines<T, NE, NT>(A, n, NL, LS, c.f.codex, c.f.nradx, TW, -1);
  const int nxs = g.nxs;
  zc* S = (zc*)c.d.spec;
#pragma unroll 2
  for (int e = threadIdx.x; e < NL * nxs; e += NT) {
    const int l = e / nxs, kx = e - l * nxs;
    const int la = ln0 + l, lb = ln0 + NL + l;
    const zc z1 = A[kx * LS + l];
    const zc z2 = A[(kx == 0 ? n / 2 : n - kx) * LS + l];
    // A = (z1 + conj z2)/2, B = (z1 - conj z2)/(2i); column 0 packs the real
    // modes 0 and n/2 of each row: A(0) + i A(n/2) and B(0) + i B(n/2)
    const zc va = kx == 0 ? zc{z1.x, z2.x} : zc{(T)0.5 * (z1.x + z2.x), (T)0.5 * (z1.y - z2.y)};
    const zc vb = kx == 0 ? zc{z1.y, z2.y} : zc{(T)0.5 * (z1.y + z2.y), (T)-0.5 * (z1.x - z2.x)};
    if (la < nlines) S[((long long)(la % g.n2m) * g.nz + la / g.n2m) * g.nxsp + kx] = va;
    if (lb < nlines) S[((long long)(lb % g.n2m) * g.nz + lb / g.n2m) * g.nxsp + kx] = vb;
  }
}

// z-DFT of the half spectrum, columns (jm, kx); sign -1 forward, +1 backward
template <typename T, int NE, int NT, bool CB_>
__global__ __launch_bounds__(NT) void k_fft_z(CtxArg<CB_> ca_, int sign) {
  const Ctx c = ctx_get(ca_);
  void* spec = (void*)c.d.spec;
  static_assert(sizeof(T) == 8, "the twiddle tables are FP64");
  extern __shared__ __align__(16) unsigned char smem[];
  typedef cplx_t<T> zc;
  const Geo g = c.g;
  const int n = g.nz, L = c.f.tlz, LS = L;
  zc* A = (zc*)smem;
  zc* TW = A + (size_t)n * LS;
  {
    const zc* twg = (const zc*)(sign < 0 ? c.f.twzf : c.f.twzb);
    for (int e = threadIdx.x; e < n; e += NT) TW[e] = twg[e];
  }
  const int nit = (g.nxs + L - 1) / L;
  const int it = blockIdx.x % nit;
  const int jm = blockIdx.x / nit;
  zc* S = (zc*)spec + (long long)jm * g.nz * g.nxsp;
  const int tot = n * L;
#pragma unroll 4
  for (int e = threadIdx.x; e < tot; e += NT) {
    const int l = e % L, p = e / L;
    const int kx = it * L + l;
    A[p * LS + l] = kx < g.nxs ? S[(long long)p * g.nxsp + kx] : zc{(T)0, (T)0};
  }
  __syncthreads();
  fft_lines<T, NE, NT>(A, n, L, LS, c.f.codez, c.f.nradz, TW, sign);
#pragma unroll 4
  for (int e = threadIdx.x; e < tot; e += NT) {
    const int l = e % L, p = e / L;
    const int kx = it * L + l;
    if (kx < g.nxs) S[(long long)p * g.nxsp + kx] = A[p * LS + l];
  }
}

// Thomas solve in y per Fourier mode (TDMAP, :9352-9406, :9444-9459) with
// the reference's pinning: real part of mode (0,0) has row 1 replaced by the
// identity with zero RHS; imaginary parts of the self-conjugate modes
// (kx in {0, NX/2}) x (kz in {0, NZ/2}) are zero.  The elimination factors
// depend on the mode only: IBET(J) = 1/BET(J) and GAM(J+1) are tabulated at
// set-up (k_pois_table) for kx = 0..NX/2 and kz = 0..NZ/2, since AK3 is even
// in kz and the factors of kz and NZ-kz are the same numbers.  A block solves
// L/2 stored columns kx at the two kz of a pair (kz, NZ-kz) (pair 0: kz = 0
// and NZ/2), so the lanes of the two halves read each factor row in the same
// load; thread (l, c) solves chunk c of column l by the partitioned
// recurrences.
// The stored column 0 holds Z = A + iB: A and B are the z-spectra of the
// real x-modes 0 and NX/2, so A(-kz) = conj A(kz), likewise B.  In the tile
// of column 0 the two threads of a chunk (halves 0 and 1) swap their values:
// for pairs pz > 0 half 0 solves A(pz) = (Z(pz) + conj Z(NZ-pz))/2 and half 1
// B(NZ-pz) = (Z(NZ-pz) - conj Z(pz))/(2i) with the kx = NX/2 factors; they
// swap again to store Z(pz) = A(pz) + i conj B(NZ-pz) and Z(NZ-pz) =
// conj A(pz) + i B(NZ-pz).  In pair 0 the four modes are real (their
// imaginary parts are the zeros above): the tile of column 0 solves A(0) and
// A(NZ/2) in the real parts, and B(0) and B(NZ/2) are solved with the kx =
// NX/2 factors either (sep = 1) by block 0 of the launch, an extra block that
// streams next to nothing and is dispatched first, which stores the
// imaginary parts, or (sep = 0) by that tile in a second pass over the
// imaginary parts, with the kx = NX/2 multipliers staged in LDS (dynamic,
// 4 N2M TO) and multipliers 0 in its other threads, which leaves their
// values unchanged.  The extra block lengthens the launch by its own time
// when the tiles fill whole rounds of resident blocks, the second pass by
// its time in any case; the host picks.
// Spectrum in TI, recurrences and result in TO (FP64 in both modes); chunks
// of at most CH points.  Grid: sep + the tiles.
template <typename TI, typename TO, int CH, int NT, bool CB_>
__global__ __launch_bounds__(NT, CH <= 6 ? 6 : 4) void k_pois_y(CtxArg<CB_> ca_, int sep) {
  const Ctx c = ctx_get(ca_);
  const void* spec_in = (const void*)c.d.spec;
  void* spec_out = (void*)c.d.spec;
  __shared__ TO xs[6][NT];
  extern __shared__ __align__(16) unsigned char smem[];
  typedef cplx_t<TI> zi;
  typedef cplx_t<TO> zo;
  const Geo g = c.g;
  const YShape ys = c.o.yp;
  const int L = ys.L, C = ys.C, LH = L >> 1;  // L columns: LH modes x 2 kz
  const int t = threadIdx.x, l = t % L;
  YPos y;
  y.l = l;
  y.c = t / L;
  const int ch = (g.n2m + C - 1) / C;
  y.j0 = y.c * ch;  // jm index (J - 1)
  y.len = max(0, min(ch, g.n2m - y.j0));
  // block 0 solves B(0) and B(NZ/2) in column 0 at kz = 0 and NZ/2 (threads
  // l = 0 and LH, the kx = NX/2 factors; its other threads load the same
  // addresses and store nothing); blocks 1.. are the tiles, in XCD-aware
  // order: the FP64 factor rows of 8 modes are half a 128-byte line, shared
  // with the neighbouring tile, which then runs on the same L2
  const bool nyq = sep && blockIdx.x == 0;
  const int nit = (g.nxs + LH - 1) / LH;
  const int tile = nyq ? 0 : xcd_tile(blockIdx.x - sep, gridDim.x - sep);
  const int it = nyq ? -1 : tile % nit;
  const int pz = nyq ? 0 : tile / nit;  // the pair (pz, NZ - pz); pair 0 is (0, NZ/2)
  const int half = l >= LH;
  const int kz = half ? (pz == 0 ? g.nz / 2 : g.nz - pz) : pz;
  const int kk = (half && pz == 0) ? g.nz / 2 : pz;  // factor row
  const int kx = nyq ? ((l == 0 || l == LH) ? 0 : g.nxs) : it * LH + l - half * LH;
  const bool ok = kx < g.nxs;
  const bool c0 = it == 0;          // the tile of column 0 (block-uniform)
  const bool pk = c0 && kx == 0;    // a thread of column 0
  const bool dual = !sep && c0 && pz == 0;  // the second pass (block-uniform)
  // factor column: kx, or NX/2 in block 0 and for the B half of column 0 in
  // pairs pz > 0
  const int fx = (nyq || (pk && half && pz != 0)) ? g.nxh - 1 : (ok ? kx : 0);
  const long long jstr = (long long)g.nz * g.nxsp;
  const long long tstr = (long long)(g.nz / 2 + 1) * g.nxhp;
  const long long b = (long long)kz * g.nxsp + (ok ? kx : 0);
  const long long bt = (long long)kk * g.nxhp + fx;
  const zi* Si = (const zi*)spec_in;
  zo* So = (zo*)spec_out;
  const double* IB = c.d.pib;
  const double* GT = c.d.pgt;
  const double* pmj = c.m.pmj;
  const bool pin = pk && kz == 0;  // the real part of mode (0,0)
  // the second pass: the kx = NX/2 multipliers of the factor rows 0 and NZ/2
  // (nm and ng, 2 x N2M each, row r = half)
  TO* nm = (TO*)smem;
  TO* ng = nm + 2 * g.n2m;
  if (dual) {
    for (int e = t; e < 2 * g.n2m; e += NT) {
      const int r = e / g.n2m, jm = e - r * g.n2m;
      const long long a = jm * tstr + (long long)(r ? g.nz / 2 : 0) * g.nxhp + (g.nxh - 1);
      nm[e] = jm == 0 ? (TO)0 : -(TO)pmj[jm + 1] * (TO)IB[a];
      ng[e] = jm == g.n2m - 1 ? (TO)0 : -(TO)GT[a];
    }
    __syncthreads();
  }
  // every load of the chunk (spectrum, both factor rows) is issued first
  TO vr[CH], vi[CH], mm[CH], gt[CH];
#pragma unroll
  for (int q = 0; q < CH; q++)
    if (q < y.len) {
      const int jm = y.j0 + q;
      const zi r = Si[jm * jstr + b];
      const TO ib = (TO)IB[jm * tstr + bt];
      gt[q] = (jm == g.n2m - 1) ? (TO)0 : -(TO)GT[jm * tstr + bt];
      if (c0) {
        // scaled once column 0 is split (mm holds IBET until then)
        vr[q] = (TO)r.x;
        vi[q] = (TO)r.y;
        mm[q] = ib;
      } else {
        vr[q] = (TO)r.x * ib;
        vi[q] = (TO)r.y * ib;
        mm[q] = (jm == 0) ? (TO)0 : -(TO)pmj[jm + 1] * ib;
      }
    }
  if (c0) {
#pragma unroll
    for (int q = 0; q < CH; q++)
      if (q < y.len) {
        const int jm = y.j0 + q;
        // the other half's thread of this chunk and column (same wavefront)
        const TO pr = __shfl_xor(vr[q], LH, 64), pi = __shfl_xor(vi[q], LH, 64);
        TO ar = vr[q], ai = vi[q];
        if (pk && pz != 0) {
          if (!half) {
            ar = (TO)0.5 * (vr[q] + pr);  // A(pz)
            ai = (TO)0.5 * (vi[q] - pi);
          } else {
            ar = (TO)0.5 * (vi[q] + pi);  // B(NZ - pz)
            ai = (TO)0.5 * (pr - vr[q]);
          }
        }
        const TO ib = mm[q];
        vr[q] = (pin && jm == 0) ? (TO)0 : ar * ib;
        vi[q] = ai * ib;
        mm[q] = (jm == 0) ? (TO)0 : -(TO)pmj[jm + 1] * ib;
      }
  }
  ychain_fwd2(vr, vi, [&](int q) { return mm[q]; }, y, L, xs[0], xs[1], xs[2]);
  ychain_bwd2(vr, vi, [&](int q) { return gt[q]; }, y, C, L, xs[3], xs[4], xs[5]);
  if (dual) {
    // B(0), B(NZ/2): the imaginary parts of column 0 again, from the
    // spectrum, with the kx = NX/2 factors.  Every thread is past the last
    // barrier of the first pass, so the exchange arrays are free.
#pragma unroll
    for (int q = 0; q < CH; q++)
      if (q < y.len && pk) {
        const int jm = y.j0 + q;
        vi[q] = (TO)Si[jm * jstr + b].y * (TO)IB[jm * tstr + (long long)kk * g.nxhp + (g.nxh - 1)];
      }
    const TO* mr = nm + half * g.n2m + y.j0;
    const TO* gr = ng + half * g.n2m + y.j0;
    ychain_fwd(vi, [&](int q) { return pk ? mr[q] : (TO)0; }, y, L, xs[0], xs[2]);
    ychain_bwd(vi, [&](int q) { return pk ? gr[q] : (TO)0; }, y, C, L, xs[3], xs[5]);
  }
  if (c0) {
#pragma unroll
    for (int q = 0; q < CH; q++)
      if (q < y.len) {
        const TO pr = __shfl_xor(vr[q], LH, 64), pi = __shfl_xor(vi[q], LH, 64);
        if (pk && pz != 0) {
          const TO sr = vr[q], si = vi[q];
          if (!half) {
            vr[q] = sr + pi;  // A(pz) + i conj B(NZ - pz)
            vi[q] = si + pr;
          } else {
            vr[q] = pr - si;  // conj A(pz) + i B(NZ - pz)
            vi[q] = sr - pi;
          }
        }
      }
  }
  // with sep, pair 0's column 0: the real parts here, the imaginary parts in
  // block 0
  const bool sx = !nyq, sy = !(sep && pk && pz == 0);
  if (ok)
#pragma unroll
    for (int q = 0; q < CH; q++)
      if (q < y.len) {
        zo* o = So + (y.j0 + q) * jstr + b;
        if (sx && sy)
          *o = zo{vr[q], vi[q]};
        else if (sx)
          o->x = vr[q];
        else
          o->y = vi[q];
      }
}

// set-up: per mode, BET(J) and GAM(J) of TDMAP on (PMJ, PCJ - AK1 - AK3, PPJ)
// with the pinned first row of mode (0,0); IB = 1/BET(J), GT(J) = GAM(J+1),
// for kz = 0..NZ/2 (AK3 is even in kz), rows of the spectrum's pitch
__global__ void k_pois_table(Ctx c) {
  const Geo g = c.g;
  const long long nm = (long long)g.nxh * (g.nz / 2 + 1);
  const long long tm = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (tm >= nm) return;
  const int kx = tm % g.nxh, kk = tm / g.nxh;
  const long long t = (long long)kk * g.nxhp + kx;
  const Metrics M = c.m;
  const double ak = M.ak3[kk + 1];
  const double a1 = M.ak1[kx + 1];
  const bool pin = (kx == 0 && kk == 0);
  const long long jstr = (long long)(g.nz / 2 + 1) * g.nxhp;
  double bet = pin ? 1.0 : (M.pcj[1] - ak - a1);
  c.d.pib[t] = 1.0 / bet;
  for (int jm = 1; jm < g.n2m; jm++) {
    const int J = jm + 1;
    const double cprev = (pin && J == 2) ? 0.0 : M.ppj[J - 1];
    const double gam = cprev / bet;
    bet = (M.pcj[J] - ak - a1) - M.pmj[J] * gam;
    c.d.pib[jm * jstr + t] = 1.0 / bet;
    c.d.pgt[(jm - 1) * jstr + t] = gam;
  }
  c.d.pgt[(g.n2m - 1) * jstr + t] = 0.0;
}

// inverse x-DFT from the half spectrum (Hermitian fill, :9419-9425), real
// part / (N1M*N3M) into DP (:9429-9433); two rows per complex transform:
// Z = A + iB, with A and B at kx = 0 and NX/2 the real and imaginary parts of
// the stored column 0 (their own imaginary parts do not reach the real part
// of the reference's transform)
template <typename T, int NE, int NT, bool CB_>
__global__ __launch_bounds__(NT) void k_fft_xinv(CtxArg<CB_> ca_) {
  const Ctx c = ctx_get(ca_);
  const void* spec = (const void*)c.d.spec;
  static_assert(sizeof(T) == 8, "the twiddle tables are FP64");
  extern __shared__ __align__(16) unsigned char smem[];
  typedef cplx_t<T> zc;
  const Geo g = c.g;
  const int n = g.nx, NL = c.f.tlx, LS = NL + 1;
  zc* A = (zc*)smem;
  zc* TW = A + (size_t)n * LS;
  for (int e = threadIdx.x; e < n; e += NT) TW[e] = ((const zc*)c.f.twxb)[e];
  const int nlines = g.n2m * g.nz;
  const int ln0 = blockIdx.x * 2 * NL;
  const zc* S = (const zc*)spec;
#pragma unroll 4
  for (int e = threadIdx.x; e < NL * n; e += NT) {
    const int l = e / n, kx = e - l * n;
    const int la = ln0 + l, lb = ln0 + NL + l;
    const long long sa = la < nlines ? ((long long)(la % g.n2m) * g.nz + la / g.n2m) * g.nxsp : -1;
    const long long sb = lb < nlines ? ((long long)(lb % g.n2m) * g.nz + lb / g.n2m) * g.nxsp : -1;
    const bool rl = kx == 0 || 2 * kx == n;  // the real modes, packed in column 0
    const int kk = rl ? 0 : (2 * kx < n ? kx : n - kx);
    zc xa = sa >= 0 ? S[sa + kk] : zc{(T)0, (T)0};
    zc xb = sb >= 0 ? S[sb + kk] : zc{(T)0, (T)0};
    if (rl) {
      xa = {kx == 0 ? xa.x : xa.y, (T)0};
      xb = {kx == 0 ? xb.x : xb.y, (T)0};
    } else if (2 * kx > n) {
      xa.y = -xa.y;
      xb.y = -xb.y;
    }
    A[kx * LS + l] = {xa.x - xb.y, xa.y + xb.x};
  }
  __syncthreads();
  fft_lines<T, NE, NT>(A, n, NL, LS, c.f.codex, c.f.nradx, TW, +1);
  const T norm = (T)(1.0 / ((double)g.nx * (double)g.nz));
  double* dp = c.d.dp;
#pragma unroll 4
  for (int e = threadIdx.x; e < NL * n; e += NT) {
    const int l = e / n, p = e - l * n;
    const int la = ln0 + l, lb = ln0 + NL + l;
    const zc z = A[p * LS + l];
    if (la < nlines) dp[gidx(g, p, 1 + la % g.n2m, la / g.n2m)] = z.x * norm;
    if (lb < nlines) dp[gidx(g, p, 1 + lb % g.n2m, lb / g.n2m)] = z.y * norm;
  }
}

// ----------------------------------------------------------------- UPCALC
// The ghost planes after the update (:1130-1135, :1147-1152, :1210-1215),
// written by the x-line update blocks that hold the rows J = 1 and J = N2M,
// in the staging's pair layout: U(0) = W(0) = V(1) = V(N2) = 0, U(N2) and
// W(N2) from BCOND's UBC, and their FP32 shadows in mixed mode.
template <int SEG, bool NTL>
__device__ __forceinline__ void upd_ghosts(const Ctx& c, const XStage<double, SEG, NTL>& X, int L, int n) {
  const Geo& g = c.g;
  const Fields& F = c.d;
  const unsigned nx = (unsigned)g.nx;
  const int nh = n >> 1, pq = threadIdx.x % nh, lq = threadIdx.x / nh;
  const bool mix = c.s.mixed != 0;
#pragma unroll
  for (int m = 0; m < SEG / 2; m++) {
    if (!((X.vm2 >> m) & 1u)) continue;
    const int ln = blockIdx.x * L + lq + m * X.lstep2;
    const int jj = 1 + ln % g.n2m;
    const unsigned o = X.go2[m];
    if (jj == 1) {
      const double2 z{0.0, 0.0};
      *(double2*)(F.u + (o - nx)) = z;
      *(double2*)(F.w + (o - nx)) = z;
      *(double2*)(F.v + o) = z;
      if (mix) {
        const float2 zf{0.f, 0.f};
        *(float2*)(F.uf + (o - nx)) = zf;
        *(float2*)(F.wf + (o - nx)) = zf;
        *(float2*)(F.vf + o) = zf;
      }
    }
    if (jj == g.n2m) {
      const unsigned t2 = (unsigned)(ln / g.n2m) * nx + 2u * (unsigned)pq;
      const double2 ub{F.ubc_u[t2], F.ubc_u[t2 + 1]}, wb{F.ubc_w[t2], F.ubc_w[t2 + 1]};
      *(double2*)(F.u + (o + nx)) = ub;
      *(double2*)(F.w + (o + nx)) = wb;
      *(double2*)(F.v + (o + nx)) = double2{0.0, 0.0};
      if (mix) {
        *(float2*)(F.uf + (o + nx)) = float2{(float)ub.x, (float)ub.y};
        *(float2*)(F.wf + (o + nx)) = float2{(float)wb.x, (float)wb.y};
        *(float2*)(F.vf + (o + nx)) = float2{0.f, 0.f};
      }
    }
  }
}

// x-lines (FP64 mode): U1 = UH1 - DT*D1x(DP) - DT*DMPRESG (:1113-1128); U2 =
// UH2 - DT*(DP(J)-DP(J-1))/H(J) (:1139-1145); with `chk`, the x part of the
// divergence of the new field, (V(J+1)-V(J))/DY + D1x(U) (:9066-9102).
// Mixed mode updates U and V with k_upd_xu below.
template <typename T, int SEG, bool CB_>
__global__ __launch_bounds__(256) void k_upd_x(CtxArg<CB_> ca_, int chk) {
  const Ctx c = ctx_get(ca_);
  XSETUP_NT;
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  const T dx1 = (T)s.dx1, dt = (T)s.dt;
  const T dmp = (T)F.sc[SC_DMPRESG];
  // V(J) and V(J+1) read DP at J-1, J, J+1 and UH2 at J, J+1: rows J-1 and
  // J+1 of tile line l are lines l-1 and l+1 of the same tiles (in one
  // k-plane wherever they are used), the row before line 0 (top halo) sits
  // in ht, the row after line L-1 (bottom halo) in the tiles' pad slots
  // p*LS + L.  So DP and UH2 are read from HBM once, plus a row or two in L.
  __shared__ double ht[512];
  const int nh = sh.n >> 1, pq = threadIdx.x % nh, lq = threadIdx.x / nh;
  const int nlines = g.n2m * g.nz, ln0 = blockIdx.x * L;
  const int lnl = min(ln0 + L, nlines) - 1;  // the tile's last line
  const int j0 = 1 + ln0 % g.n2m, jl = 1 + lnl % g.n2m;
  const bool htop = lq == 0 && j0 >= 2, hbot = chk && lq == 0 && ln0 + L <= nlines && jl + 1 <= g.n2m;
  const unsigned otop = (unsigned)gidx(g, 2 * pq, j0 - 1, ln0 / g.n2m);
  const unsigned obot = (unsigned)gidx(g, 2 * pq, jl + 1, lnl / g.n2m);
  T d[SEG], un[SEG], pf[SEG], vj[SEG], vd[SEG];
  double2 h0 = {0.0, 0.0}, h1 = {0.0, 0.0};
  X.pre(F.dp, 0, pf);
  if (htop) h0 = *(const double2*)(F.dp + otop);
  if (hbot) h1 = *(const double2*)(F.dp + obot);
  const T* A = X.put(pf);  // DP
  if (htop) {
    ht[2 * pq] = h0.x;
    ht[2 * pq + 1] = h0.y;
  }
  if (hbot) {
    T* Aw = (T*)A;
    Aw[2 * pq * X.LS + L] = h1.x;
    Aw[(2 * pq + 1) * X.LS + L] = h1.y;
  }
  X.pre((const T*)F.uh[1], 0, pf);
  if (hbot) h1 = *(const double2*)((const T*)F.uh[1] + obot);
  SQ d[q] = (T)24. * dx1 * (-XA(A, q - 1) + XA(A, q));
  psolve<T, SEG>(d, G, L, l, gs, c.o.px.s22, bf.xf, bf.xb);  // its barriers publish the halos
  const T* B = X.put(pf);  // UH2
  if (hbot) {
    T* Bw = (T*)B;
    Bw[2 * pq * X.LS + L] = h1.x;
    Bw[(2 * pq + 1) * X.LS + L] = h1.y;
  }
  if (chk) __syncthreads();
  X.pre((const T*)F.uh[0], 0, pf);
  // V(J) = UH2(J) - DT*(DP(J) - DP(J-1))/H(J) (:1139-1145); with chk also
  // (V(J+1) - V(J))/DY, V(N2) = 0
  const T rhj = (T)M.rh[j];  // 1/H(J)
  SQ {
    const int p = p0 + q;
    const T dm = l == 0 ? ht[p] : A[p * X.LS + l - 1];
    vj[q] = (j >= 2) ? B[p * X.LS + l] - dt * (A[p * X.LS + l] - dm) * rhj : (T)0;
  }
  if (chk) {
    const T rhjp = (T)M.rh[j + 1], rdy = (T)M.rdy[j];
    SQ {
      const int p = p0 + q;
      const T e = B[p * X.LS + l + 1] - dt * (A[p * X.LS + l + 1] - A[p * X.LS + l]) * rhjp;
      vd[q] = (((j + 1 <= g.n2m) ? e : (T)0) - vj[q]) * rdy;
    }
  }
  X.out(F.v, vj, 2);  // V(I,1,K) is not written by UPCALC
  B = X.put(pf);      // UH1
  SQ un[q] = XA(B, q) - dt * d[q] - dt * dmp;
  X.out(F.u, un);
  upd_ghosts<SEG>(c, X, L, sh.n);
  if (!chk) return;
  const T rt = xch_right(un[0], G, L, l, gs, bf.xr);
  SQ d[q] = (T)24. * dx1 * (nxt(un, q, rt) - un[q]);
  psolve<T, SEG>(d, G, L, l, gs, c.o.px.s22, bf.xf, bf.xb);
  SQ d[q] = vd[q] + d[q];
  X.out(F.div, d);
}

// Mixed mode: k_upd_xu updates V pointwise from its DP tile, then U on
// x-lines, and stores D1x of the new U (FP64, in RDP's buffer); k_upd_z adds
// (V(J+1) - V(J))/DY, from the V(J), V(J+1) its CFL term reads anyway, and
// D1z of the new W.  Same expressions and operation order as k_upd_x.
template <int SEG, bool CB_>
__global__ __launch_bounds__(256) void k_upd_xu(CtxArg<CB_> ca_, int refine) {
  const Ctx c = ctx_get(ca_);
  typedef double T;
  XSETUP_NT;
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  const T dx1 = (T)s.dx1, dt = (T)s.dt;
  const T dmp = refine ? (T)0 : (T)F.sc[SC_DMPRESG];
  const T ku = refine ? (T)0 : (T)1;
  T d[SEG], un[SEG], pf[SEG], pg[SEG];
  float ph[SEG];
  // V (J >= 2; V(1) is the wall's, left to the ghost pass), pointwise in the
  // staging's pair layout: V = VOLD + DELTA_v - DT*(DP(J) - DP(J-1))/H(J)
  // (:1139-1145), with DP(J) and DP(J-1) from the DP tile in LDS: row J-1 of
  // tile line ll >= 1 is line ll-1 (when J >= 2 both lie in one k-plane),
  // and the row before line 0 (the halo) goes to the tile's pad slots
  // p*LS + L.  So DP is read from HBM once, plus one row in L.
  const int nh = sh.n >> 1, pq = threadIdx.x % nh, lq = threadIdx.x / nh;
  const int jh = 1 + (blockIdx.x * L) % g.n2m;  // J of line 0
  const bool hl = lq == 0 && jh >= 2 && blockIdx.x * L < g.n2m * g.nz;
  double2 hv = {0.0, 0.0};
  if (hl) hv = *(const double2*)(F.dp + (X.go2[0] - (unsigned)g.nx));
  X.pre(F.dp, 0, pf);
  double vo[SEG];
  float dv[SEG];
  X.pre_pairs(F.vold, 0u, vo);
  X.pre_pairs((const float*)F.uh[1], 0u, dv);
  const T* B = X.put(pf);
  if (hl) {
    T* Bw = (T*)B;
    Bw[2 * pq * X.LS + L] = hv.x;
    Bw[(2 * pq + 1) * X.LS + L] = hv.y;
  }
  __syncthreads();
  X.pre_raw((const float*)F.uh[0], 0, ph);
#pragma unroll
  for (int m = 0; m < SEG / 2; m++) {
    const int ll = lq + m * X.lstep2;
    const int jm = 1 + (blockIdx.x * L + ll) % g.n2m;
    if (((X.vm2 >> m) & 1u) && jm >= 2) {
      const int ij = X.li2 + m * X.lstep2, im = ll == 0 ? 2 * pq * X.LS + L : ij - 1;
      const double rh = M.rh[jm];
      const double v0 = vo[2 * m] + (ku * (double)dv[2 * m] - dt * (B[ij] - B[im]) * rh);
      const double v1 = vo[2 * m + 1] + (ku * (double)dv[2 * m + 1] - dt * (B[ij + X.LS] - B[im + X.LS]) * rh);
      *(double2*)(F.v + X.go2[m]) = double2{v0, v1};
      *(float2*)(F.vf + X.go2[m]) = float2{(float)v0, (float)v1};
    }
  }
  SQ d[q] = (T)24. * dx1 * (-XA(B, q - 1) + XA(B, q));
  psolve<T, SEG>(d, G, L, l, gs, c.o.px.s22, bf.xf, bf.xb);
  X.pre(F.u, 0, pg);  // after the solve: the old U tile is not live across it
  B = X.put(ph);
  SQ un[q] = ku * XA(B, q) - dt * d[q] - dt * dmp;
  // the flow rate of U before the update (FLOW1, :1081)
  double so = 0.0;
  B = X.put(pg);
  SQ {
    const T uo = XA(B, q);
    so += (double)uo;
    un[q] = uo + un[q];
  }
  X.out2(F.u, F.uf, un);
  block_sum2(so * (ok ? M.dy[j] : 0.0), 0.0, F.flowp + FP_U * F.flowr + blockIdx.x, nullptr);
  const T rt = xch_right(un[0], G, L, l, gs, bf.xr);
  SQ d[q] = (T)24. * dx1 * (nxt(un, q, rt) - un[q]);
  psolve<T, SEG>(d, G, L, l, gs, c.o.px.s22, bf.xf, bf.xb);
  X.out(F.rdp, d);  // D1x U of the new field, for k_upd_z
  upd_ghosts<SEG>(c, X, L, sh.n);
}

// z-lines: U3 = UH3 - DT*D1z(DP) - DT*DMPRESG3 (:1193-1208); P += DP
// (:1218-1230); with `chk`: DIV += D1z(U3), DIVMAX, and the CFL maximum of
// the new field (:9112-9128, :9150-9162).  Mixed mode also stores the full
// DIV (the next RDP starts from it) and the FP32 shadows of W and P.
template <typename T, int SEG, typename TS, int NT, bool CB_>
__global__ __launch_bounds__(NT) void k_upd_z(CtxArg<CB_> ca_, int chk, int refine) {
  const Ctx c = ctx_get(ca_);
  ZSETUP;
  constexpr bool MIX = sizeof(TS) == 4;
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  const T dx3 = (T)s.dx3, dt = (T)s.dt;
  const T dmp3 = refine ? (T)0 : (T)F.sc[SC_DMPRESG3];
  const T ku = refine ? (T)0 : (T)1;
  T d[SEG], wn[SEG], lf, rt;
  {
    // P += DP first, so that DP's segment is dead during the solve
    T dpr[SEG];
    zload_nt(F.dp, b, pl, p0, bm, bp, dpr, lf, rt);
    SQ d[q] = (T)24. * dx3 * (-prv(dpr, q, lf) + dpr[q]);
    if (ok) {
      SQ {
        const double pn = ld_nt(&F.p[ZI(q)]) + dpr[q];
        st_nt(&F.p[ZI(q)], pn);
        if (MIX) st_nt(&F.pf[ZI(q)], (float)pn);
      }
    }
  }
  psolve<T, SEG>(d, G, L, l, gs, PZ.s22, bf.xf, bf.xb);
  SQ wn[q] = ku * (T)ld_nt(&((const TS*)F.uh[2])[ZI(q)]) - dt * d[q] - dt * dmp3;
  // mixed mode: the flow rate of W before the update (FLOW3, :1168), summed
  // over the block with the maxima at the end
  double so = 0.0;
  if (MIX) {
    SQ {
      const T wo = (T)ld_nt(&F.w[ZI(q)]);
      so += (double)wo;
      wn[q] = wo + wn[q];
    }
    so = ok ? so * M.dy[j] : 0.0;
  }
  if (ok) {
    SQ {
      st_nt(&F.w[ZI(q)], (double)wn[q]);
      if (MIX) st_nt(&F.wf[ZI(q)], (float)wn[q]);
    }
  }
  if (!chk) {
    if (MIX) block_sum2(so, 0.0, F.flowp + FP_W * F.flowr + blockIdx.x, nullptr);
    return;
  }
  rt = xch_right(wn[0], G, L, l, gs, bf.xr);
  SQ d[q] = (T)24. * dx3 * (nxt(wn, q, rt) - wn[q]);
  psolve<T, SEG>(d, G, L, l, gs, PZ.s22, bf.xf, bf.xb);
  double dmax = 0.0, cmax = 0.0;
  if (ok) {
    const double hdy = 0.5 / M.dy[j];
    const unsigned bi = (unsigned)gidx(g, i + 1 == g.nx ? 0 : i + 1, j, 0);
    const double rdyj = M.rdy[j];
    SQ {
      const unsigned a = ZI(q);
      const double v0 = F.v[a], v1 = F.v[a + g.nx];
      double dv;
      if (MIX) {
        // DIV of the new field: D1x U (k_upd_xu, in RDP's buffer) +
        // (V(J+1) - V(J))/DY (the wall values V(1) = V(N2) = 0) + D1z W,
        // kept in FP32 for the next RDP
        const double vj = j >= 2 ? v0 : 0.0, vp = j + 1 <= g.n2m ? v1 : 0.0;
        dv = ((vp - vj) * rdyj + ld_nt(&F.rdp[a])) + (double)d[q];
        st_nt(&F.divf[a], (float)dv);
      } else {
        dv = F.div[a] + (double)d[q];
      }
      dmax = (dv != dv || dmax != dmax) ? NAN : fmax(dmax, fabs(dv));
      const double wq = (double)wn[q], wq1 = (double)nxt(wn, q, rt);
      const double cfll = fabs(F.u[a] + F.u[bi + (unsigned)(p0 + q) * pl]) * 0.5 * s.dx1 +
                          fabs(v0 + v1) * hdy + fabs(wq + wq1) * 0.5 * s.dx3;
      cmax = (cfll != cfll || cmax != cmax) ? NAN : fmax(cmax, cfll);
    }
  }
  // block maxima into 64 spread slots (bit patterns of non-negative
  // doubles order like the values); k_max_fin reduces the slots.  Mixed
  // mode: the block's flow-rate sum in block_sum2's order (wavefront sums,
  // then the wavefronts in index order)
  __shared__ double red[3][NT / 64];
  dmax = wave_max(dmax);
  cmax = wave_max(cmax);
  if (MIX) so = wave_sum(so);
  const int wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][wv] = dmax;
    red[1][wv] = cmax;
    if (MIX) red[2][wv] = so;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < nw; q++) {
      dmax = (dmax != dmax || red[0][q] != red[0][q]) ? NAN : fmax(dmax, red[0][q]);
      cmax = (cmax != cmax || red[1][q] != red[1][q]) ? NAN : fmax(cmax, red[1][q]);
    }
    atomic_max_pos(&F.mxs[blockIdx.x & 63], dmax);
    atomic_max_pos(&F.mxs[64 + (blockIdx.x & 63)], cmax);
    if (MIX) {
      double sa = 0.0;
      for (int q = 0; q < nw; q++) sa += red[2][q];
      F.flowp[FP_W * F.flowr + blockIdx.x] = sa;
    }
  }
}

// Sums of the first n[r] entries of NR partial-sum regions (region r at
// F.flowp + reg[r]*F.flowr) by one block of FIN_THREADS threads, in a fixed
// order: thread t adds entries t, t+FIN_THREADS, ... (FIN_PER loads of each
// region in flight before their adds), then the wavefronts' butterflies,
// then the wavefronts in index order.  The results are valid in thread 0.
#define FIN_THREADS 1024
#define FIN_PER 8
template <int NR>
__device__ __forceinline__ void region_sums(const Fields& F, const int (&reg)[NR], const int (&n)[NR],
                                            double (&out)[NR]) {
  __shared__ double rs[NR][FIN_THREADS / 64];
  const int t = threadIdx.x;
  int nmax = 0;
#pragma unroll
  for (int r = 0; r < NR; r++) nmax = max(nmax, n[r]);
  double acc[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) acc[r] = 0.0;
  for (int base = 0; base < nmax; base += FIN_THREADS * FIN_PER) {
    double v[NR][FIN_PER];
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
      for (int m = 0; m < FIN_PER; m++) {
        const int e = base + t + m * FIN_THREADS;
        v[r][m] = e < n[r] ? F.flowp[(long long)reg[r] * F.flowr + e] : 0.0;
      }
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
      for (int m = 0; m < FIN_PER; m++) acc[r] += v[r][m];
  }
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const double a = wave_sum(acc[r]);
    if ((t & 63) == 0) rs[r][t >> 6] = a;
  }
  __syncthreads();
  if (t == 0)
#pragma unroll
    for (int r = 0; r < NR; r++) {
      double a = 0.0;
      for (int w = 0; w < FIN_THREADS / 64; w++) a += rs[r][w];
      out[r] = a;
    }
}

// UPCALC's mean-pressure-gradient corrections (:1079-1093, :1160-1172) from
// the ADI sweeps' partial sums.  The reference's sums are
//   DMPRESG = [sum(UH*DY/DX3/DX1 - DT*(DP(I)-DP(I-1))*DY/DX3) - FLOW1]/VOL/DT
// with FLOW1 = sum(U*DY/DX3/DX1).  The DP differences sum to zero over each
// periodic x-line, and UH - U is the ADI increment DELTA, so
//   DMPRESG = sum(DELTA*DY)/DX3/DX1/VOL/DT
// (likewise DMPRESG3 along z): the same quantity without the cancellation of
// the two flow sums.  nu / nw: partial counts of the u and w regions; with
// `flows` the flow rates FLOW1/FLOW3 of U are formed too (FP64 mode, where
// the ADI sweeps read U).
template <bool CB_>
__global__ __launch_bounds__(FIN_THREADS) void k_flow_fin(CtxArg<CB_> ca_, int nu, int nw, int flows) {
  const Ctx c = ctx_get(ca_);
  const Fields F = c.d;
  const Scal s = c.s;
  const int reg[4] = {FP_DU, FP_DW, FP_U, FP_W};
  const int n[4] = {nu, nw, flows ? nu : 0, flows ? nw : 0};
  double a[4];
  region_sums<4>(F, reg, n, a);
  if (threadIdx.x == 0) {
    const double dmp = a[0] / s.dx3 / s.dx1 / s.vol / s.dt;
    const double dmp3 = a[1] / s.dx3 / s.dx1 / s.vol / s.dt;
    F.sc[SC_DMPRESG] = dmp;
    F.sc[SC_DMPRESG3] = dmp3;
    F.sc[SC_PRESG] = F.sc[SC_PRESG] + dmp;
    F.sc[SC_PRESG3] = F.sc[SC_PRESG3] + dmp3;
    if (flows) {
      F.sc[SC_FLOW1] = a[2] / s.dx3 / s.dx1;
      F.sc[SC_FLOW3] = a[3] / s.dx1 / s.dx3;
    }
  }
}

// final maxima of the 64 max slots of k_upd_z into SC_DIVMAX, SC_CFLM, which
// are cleared again for the next k_upd_z; with `flows` (mixed mode, where the
// update kernels read the old U) also the flow rates FLOW1/FLOW3 from nu /
// nw partial sums
template <bool CB_>
__global__ __launch_bounds__(FIN_THREADS) void k_max_fin(CtxArg<CB_> ca_, int nu, int nw, int flows) {
  const Ctx c = ctx_get(ca_);
  const Fields F = c.d;
  if (flows) {
    const int reg[2] = {FP_U, FP_W};
    const int n[2] = {nu, nw};
    double a[2];
    region_sums<2>(F, reg, n, a);
    if (threadIdx.x == 0) {
      F.sc[SC_FLOW1] = a[0] / c.s.dx3 / c.s.dx1;
      F.sc[SC_FLOW3] = a[1] / c.s.dx1 / c.s.dx3;
    }
  }
  if (threadIdx.x >= 64) return;
  unsigned long long* r = (unsigned long long*)F.mxs;
  unsigned long long a = r[threadIdx.x], b = r[64 + threadIdx.x];
  r[threadIdx.x] = 0ull;
  r[64 + threadIdx.x] = 0ull;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long x = __shfl_xor(a, o, 64), y = __shfl_xor(b, o, 64);
    a = x > a ? x : a;
    b = y > b ? y : b;
  }
  if (threadIdx.x == 0) {
    ((unsigned long long*)F.sc)[SC_DIVMAX] = a;
    ((unsigned long long*)F.sc)[SC_CFLM] = b;
  }
}

// =================================================================== host
namespace c3dk {

// compile-time segment lengths: 8, or 2 when the line length is not a
// multiple of 8
#define XSEG(seg, ...)                \
  if ((seg) == 8) {                   \
    constexpr int S_ = 8;             \
    __VA_ARGS__;                      \
  } else {                            \
    constexpr int S_ = 2;             \
    __VA_ARGS__;                      \
  }
#define ZSEG(seg, ...) XSEG(seg, __VA_ARGS__)

// LDS bytes of a line kernel
static size_t lds_line(const LShape& sh, bool staged, size_t es) {
  const size_t tile = staged ? (size_t)sh.n * (sh.L + 1) : 0;
  return es * (2 * tile + 4 * (size_t)sh.G * sh.L);
}
static int xblocks(const Ctx& c) { return (c.g.n2m * c.g.nz + c.o.lx.L - 1) / c.o.lx.L; }
static int zblocks(const Ctx& c, const LShape& z) { return ((c.g.nx + z.L - 1) / z.L) * c.g.n2m; }
static int zblocks(const Ctx& c) { return zblocks(c, c.o.lz); }
static dim3 xthreads(const Ctx& c) { return dim3(c.o.lx.L * c.o.lx.G); }
static dim3 zthreads(const Ctx& c) { return dim3(c.o.lz.L * c.o.lz.G); }
// the z-line shape of compute type T
template <typename T>
static const LShape& zshape(const Ctx& c) {
  return sizeof(T) == 4 ? c.o.lz32 : c.o.lz;
}
// compile-time segment of the FP32 z-line kernels: 16, 8 or 2
#define ZSEG32(seg, ...)              \
  if ((seg) == 16) {                  \
    constexpr int S_ = 16;            \
    __VA_ARGS__;                      \
  } else if ((seg) == 8) {            \
    constexpr int S_ = 8;             \
    __VA_ARGS__;                      \
  } else {                            \
    constexpr int S_ = 2;             \
    __VA_ARGS__;                      \
  }

// z-line launches of compute type T (the FP32 kernels also have SEG 16)
#define ZLAUNCH(T, KID, KERNEL, GY, ...)                                                              \
  {                                                                                                   \
    const LShape& z_ = zshape<T>(c);                                                                  \
    const size_t lz_ = lds_line(z_, false, sizeof(T));                                                \
    const dim3 gr_(zblocks(c, z_), GY), th_(z_.L * z_.G);                                             \
    if (sizeof(T) == 4) {                                                                             \
      ZSEG32(z_.seg, { KScope ks_(c, KID); C3D_LAUNCH(KERNEL, (float, S_), gr_, th_, lz_, c, ##__VA_ARGS__); }); \
    } else {                                                                                          \
      ZSEG(z_.seg, { KScope ks_(c, KID); C3D_LAUNCH(KERNEL, (double, S_), gr_, th_, lz_, c, ##__VA_ARGS__); }); \
    }                                                                                                 \
  }

template <typename T>
static void rhs_lines(const Ctx& c) {
  const size_t lx = lds_line(c.o.lx, true, sizeof(T));
  // with SGS the x-line SGS terms run in the same kernel (k_rhs_x)
  if (c.s.files) {
    XSEG(c.o.lx.seg, { KScope ks_(c, KID_RHS_X); C3D_LAUNCH(k_rhs_x, (T, S_, true), dim3(xblocks(c)), xthreads(c), lx, c); });
  } else {
    XSEG(c.o.lx.seg, { KScope ks_(c, KID_RHS_X); C3D_LAUNCH(k_rhs_x, (T, S_, false), dim3(xblocks(c)), xthreads(c), lx, c); });
  }
  ZLAUNCH(T, KID_RHS_Z, k_rhs_z, 1);
  // with SGS, HH3 += D1x(UW) joins the last SGS x-pass (k_sgs_x2)
  if (!c.s.files) {
    XSEG(c.o.lx.seg, { KScope ks_(c, KID_RHS_X2); C3D_LAUNCH(k_rhs_x2, (T, S_), dim3(xblocks(c)), xthreads(c), lx, c); });
  }
}
void rhs_lines(const Ctx& c) {
  if (c.s.mixed)
    rhs_lines<float>(c);
  else
    rhs_lines<double>(c);
}

template <typename T>
static void getuh_t(const Ctx& c, bool fused) {
  const YShape& ys = sizeof(T) == 4 ? c.o.ys32 : c.o.ys;
  // FP32 tiles of column pairs (P = 2: L threads cover 2L columns) exist
  // with chunks of at most 6 points only (the shape candidates ensure it)
  const bool pair = sizeof(T) == 4 && ys.P == 2;
  const int ylx = (c.g.nx + ys.L * ys.P - 1) / (ys.L * ys.P);
  // tiles of more than 512 threads run the NT = 1024 instances (rows of
  // 256 bytes: 64 FP32 or 32 FP64 columns)
  const bool big = ys.L * ys.C > 512;
  if (fused) {
    const size_t tbytes = sizeof(T) * (size_t)(c.g.n2 + 2) * YTAB_W;
    // the coefficient table goes to LDS when it fits beside the chain arrays
    // within 64 KB
    const size_t xsb = 2 * (sizeof(T) * ys.P + sizeof(T)) * (big ? 1024 : 512);
    const int staged = tbytes + xsb <= 64 * 1024 && tbytes <= 40 * 1024 ? 1 : 0;
    // chunk length of the shape (u and w have the most points, N2M)
    const int ch = (c.g.n2m + ys.C - 1) / ys.C;
    const dim3 gr(ylx * c.g.nz, 3), th(ys.L * ys.C);
    const size_t lb = staged ? tbytes : 0;
    KScope ks_(c, KID_ADI_Y);
#define YF_(V, SGS, CH, NT) C3D_LAUNCH(k_rhsy_adiy, (T, V, SGS, CH, NT), gr, th, lb, c, staged)
#define YFN_(V, SGS, CH) \
  if (big)               \
    YF_(V, SGS, CH, 1024); \
  else                   \
    YF_(V, SGS, CH, 512)
    bool done = false;
    if constexpr (sizeof(T) == 4) {
      if (pair) {
        if (c.s.files) { YFN_(f2v, true, 6); } else { YFN_(f2v, false, 6); }
        done = true;
      }
    }
    if (done) {
    } else if (c.s.files) {
      if (ch <= 6) { YFN_(T, true, 6); } else { YFN_(T, true, CHMAX); }
    } else {
      if (ch <= 6) { YFN_(T, false, 6); } else { YFN_(T, false, CHMAX); }
    }
#undef YFN_
#undef YF_
  } else {
    // the stand-alone sweep holds one column per thread: a pair shape runs
    // as L columns of the same chunks
    KScope ks_(c, KID_ADI_Y);
    const YShape y1{ys.L, ys.C, 1};
    const dim3 gr((c.g.nx + y1.L - 1) / y1.L * c.g.nz, 3), th(y1.L * y1.C);
    if (big)
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adi_y<T, 1024>), gr, th, 0, c.st, c, y1);
    else
      hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adi_y<T, 512>), gr, th, 0, c.st, c, y1);
  }
  const size_t lx = lds_line(c.o.lx, true, sizeof(T));
  if constexpr (sizeof(T) == 4) {
    // mixed mode: u and v sweep y, z, x and w y, x, z; the last sweeps form
    // RDP and the flow-rate sums (k_adi_xr, k_adi_zr)
    const CompList z12{2, {0, 1, 0}, {0, 0, 0}};
    ZLAUNCH(T, KID_ADI_Z, k_adi_z, 2, z12);
    XSEG(c.o.lx.seg, { KScope ks_(c, KID_ADI_XR); C3D_LAUNCH(k_adi_xr, (S_), dim3(xblocks(c), 3), xthreads(c), lx, c); });
    const LShape& z_ = c.o.lz32;
    const size_t lz_ = lds_line(z_, false, sizeof(T));
    ZSEG32(z_.seg, { KScope ks_(c, KID_ADI_ZR); C3D_LAUNCH(k_adi_zr, (S_), dim3(zblocks(c, z_)), dim3(z_.L * z_.G), lz_, c); });
  } else {
    // GETUH1/2: y, x, z;  GETUH3: y, z, x (compact3d.f:8508-8607)
    const CompList x12{2, {0, 1, 0}, {0, 0, 0}};
    const CompList z123{3, {0, 1, 2}, {1, 1, 0}};
    const CompList x3{1, {2, 0, 0}, {1, 0, 0}};
    XSEG(c.o.lx.seg, { KScope ks_(c, KID_ADI_X); C3D_LAUNCH(k_adi_x, (T, S_), dim3(xblocks(c), 2), xthreads(c), lx, c, x12); });
    ZLAUNCH(T, KID_ADI_Z, k_adi_z, 3, z123);
    XSEG(c.o.lx.seg, { KScope ks_(c, KID_ADI_X); C3D_LAUNCH(k_adi_x, (T, S_), dim3(xblocks(c), 1), xthreads(c), lx, c, x3); });
  }
}
void getuh(const Ctx& c) {
  if (c.s.mixed)
    getuh_t<float>(c, false);
  else
    getuh_t<double>(c, false);
}
void getuh_fused(const Ctx& c) {
  if (c.s.mixed)
    getuh_t<float>(c, true);
  else
    getuh_t<double>(c, true);
}

// FP64 mode; mixed mode: the last ADI sweeps formed RDP already (k_adi_xr,
// k_adi_zr)
void rhsdp(const Ctx& c) {
  if (c.s.mixed) return;
  const size_t lx = lds_line(c.o.lx, true, sizeof(double)), lz = lds_line(c.o.lz, false, sizeof(double));
  XSEG(c.o.lx.seg, { KScope ks_(c, KID_RHSDP_X); C3D_LAUNCH(k_rhsdp_x, (double, S_), dim3(xblocks(c)), xthreads(c), lx, c); });
  ZSEG(c.o.lz.seg, { KScope ks_(c, KID_RHSDP_Z); C3D_LAUNCH(k_rhsdp_z, (double, S_), dim3(zblocks(c)), zthreads(c), lz, c); });
}

// GETDP in FP64 in both modes: rounding RDP's transform or the solution to
// FP32 would be amplified ~1/DY^2 by the projection's divergence
// k_pois_y's dynamic LDS in its second-pass form: pair 0's kx = NX/2
// multipliers
static size_t pois_lds(const Geo& g) { return sizeof(double) * 4 * (size_t)g.n2m; }
// resident blocks of the k_pois_y instantiations (6, 1024), (6, 512),
// (CHMAX, 512), (6, 256), (CHMAX, 256) on the whole GPU, from fft_attrs()
static int g_pois_slots[5];

static void getdp_fp64(const Ctx& c) {
  typedef double T;
  const Geo& g = c.g;
  const FFTPlan& f = c.f;
  // one in-place buffer per block; x lines padded to a stride of TL+1
  const size_t ldsx = sizeof(cplx_t<T>) * (size_t)g.nx * (f.tlx + 2);
  const size_t ldsz = sizeof(cplx_t<T>) * (size_t)g.nz * (f.tlz + 1);
  const int gx = (g.n2m * g.nz + 2 * f.tlx - 1) / (2 * f.tlx);
  const int gz = ((g.nxs + f.tlz - 1) / f.tlz) * g.n2m;
  const int gy = ((g.nxs + c.o.yp.L / 2 - 1) / (c.o.yp.L / 2)) * (g.nz / 2);
  // tile sizes (complex values per block): up to 1024 on 256 threads, up
  // to 2048 or 4096 on 512 (the autotuner picks the lines per tile)
  const int ex = g.nx * f.tlx, ez = g.nz * f.tlz;
  const int nx_t = ex <= 1024 ? 256 : 512, nz_t = ez <= 1024 ? 256 : 512;
#define FFTW_(K, e, nt, ...)                                                            \
  do {                                                                                  \
    if ((e) > 2048)                                                                     \
      C3D_LAUNCH(K, (T, 4096, 512), __VA_ARGS__);                                       \
    else if ((e) > 1024)                                                                \
      C3D_LAUNCH(K, (T, 2048, 512), __VA_ARGS__);                                       \
    else                                                                                \
      C3D_LAUNCH(K, (T, 1024, 256), __VA_ARGS__);                                       \
  } while (0)
  // the register FFT for line lengths 64, 128, 256 (256 threads, L lines or
  // columns per block)
  const int rx = (g.nx == 64 || g.nx == 128 || g.nx == 256) ? g.nx : 0;
  const int rz = (g.nz == 64 || g.nz == 128 || g.nz == 256) ? g.nz : 0;
  const int lx = rx ? 256 / (rx == 256 ? 16 : 8) : 0, lzr = rz ? 256 / (rz == 256 ? 16 : 8) : 0;
  const int gxr = rx ? (g.n2m * g.nz + 2 * lx - 1) / (2 * lx) : 0;
  const int gzr = rz ? ((g.nxs + lzr - 1) / lzr) * g.n2m : 0;
#define RF_(K, n, GR, ...)                                                                                  \
  do {                                                                                                      \
    if ((n) == 256)                                                                                         \
      C3D_LAUNCH(K, (T, 256), dim3(GR), dim3(256), (rfft_lds<256, true>()), c, ##__VA_ARGS__);            \
    else if ((n) == 128)                                                                                    \
      C3D_LAUNCH(K, (T, 128), dim3(GR), dim3(256), (rfft_lds<128, true>()), c, ##__VA_ARGS__);            \
    else                                                                                                    \
      C3D_LAUNCH(K, (T, 64), dim3(GR), dim3(256), (rfft_lds<64, true>()), c, ##__VA_ARGS__);              \
  } while (0)
  {
    KScope ks_(c, KID_FFT_X);
    if (rx)
      RF_(k_rfft_x, rx, gxr);
    else
      FFTW_(k_fft_x, ex, nx_t, dim3(gx), dim3(nx_t), ldsx, c);
  }
  {
    KScope ks_(c, KID_FFT_Z);
    if (rz)
      RF_(k_rfft_z, rz, gzr, -1);
    else
      FFTW_(k_fft_z, ez, nz_t, dim3(gz), dim3(nz_t), ldsz, c, -1);
  }
  {
    // tiles of up to 512 threads, or 1024 with 6-point chunks
    KScope ks_(c, KID_POIS_Y);
    const int npt = c.o.yp.L * c.o.yp.C;
    const bool six = (g.n2m + c.o.yp.C - 1) / c.o.yp.C <= 6;
    // the extra block for pair 0's B modes, unless the tiles fill whole
    // rounds of resident blocks (then the second pass)
    const long long tiles = (long long)gy * (c.bdev ? c.bn : 1);
#define PY_(CH, NT, S)                                                                             \
  do {                                                                                             \
    const int sep = (g_pois_slots[S] > 0 && tiles % g_pois_slots[S] == 0) ? 0 : 1;               \
    C3D_LAUNCH(k_pois_y, (T, T, CH, NT), dim3(gy + sep), dim3(npt), sep ? 0 : pois_lds(g), c, sep); \
  } while (0)
    if (npt > 512)
      PY_(6, 1024, 0);
    else if (npt > 256) {
      if (six) PY_(6, 512, 1); else PY_(CHMAX, 512, 2);
    } else {
      if (six) PY_(6, 256, 3); else PY_(CHMAX, 256, 4);
    }
#undef PY_
  }
  {
    KScope ks_(c, KID_FFT_Z);
    if (rz)
      RF_(k_rfft_z, rz, gzr, +1);
    else
      FFTW_(k_fft_z, ez, nz_t, dim3(gz), dim3(nz_t), ldsz, c, +1);
  }
  {
    KScope ks_(c, KID_FFT_XINV);
    if (rx)
      RF_(k_rfft_xinv, rx, gxr);
    else
      FFTW_(k_fft_xinv, ex, nx_t, dim3(gx), dim3(nx_t), ldsx, c);
  }
#undef RF_
#undef FFTW_
}
void getdp(const Ctx& c) { getdp_fp64(c); }

// FFT tiles may use more than 64 KB of LDS (up to the 160 KB a workgroup
// can address on gfx950)
void fft_attrs() {
  const int lim = 160 * 1024;
#define FA_(K)                                                                                                   \
  (void)hipFuncSetAttribute((const void*)&K<double, 4096, 512, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
  (void)hipFuncSetAttribute((const void*)&K<double, 2048, 512, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
  (void)hipFuncSetAttribute((const void*)&K<double, 1024, 256, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
  (void)hipFuncSetAttribute((const void*)&K<double, 4096, 512, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
  (void)hipFuncSetAttribute((const void*)&K<double, 2048, 512, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
  (void)hipFuncSetAttribute((const void*)&K<double, 1024, 256, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim)
  FA_(k_fft_x);
  FA_(k_fft_z);
  FA_(k_fft_xinv);
#undef FA_
#define FR_(K)                                                                                          \
  (void)hipFuncSetAttribute((const void*)&K<double, 256, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
  (void)hipFuncSetAttribute((const void*)&K<double, 128, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim); \
  (void)hipFuncSetAttribute((const void*)&K<double, 64, false>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);  \
  (void)hipFuncSetAttribute((const void*)&K<double, 256, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);  \
  (void)hipFuncSetAttribute((const void*)&K<double, 128, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim);  \
  (void)hipFuncSetAttribute((const void*)&K<double, 64, true>, hipFuncAttributeMaxDynamicSharedMemorySize, lim)
  FR_(k_rfft_x);
  FR_(k_rfft_z);
  FR_(k_rfft_xinv);
#undef FR_
  // k_pois_y: its second-pass LDS (4 N2M doubles) beside the static exchange
  // arrays (6 NT doubles), and its resident blocks per GPU
  int dev = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
#define PA_(CH, NT, S)                                                                                \
  do {                                                                                                \
    (void)hipFuncSetAttribute((const void*)&k_pois_y<double, double, CH, NT, false>,                 \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim - 48 * (NT));           \
    (void)hipFuncSetAttribute((const void*)&k_pois_y<double, double, CH, NT, true>,                  \
                              hipFuncAttributeMaxDynamicSharedMemorySize, lim - 48 * (NT));           \
    int nb = 0;                                                                                       \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)&k_pois_y<double, double, CH, NT, false>, \
                                                     NT, 0) != hipSuccess)                            \
      nb = 0;                                                                                         \
    g_pois_slots[S] = nb * ncu;                                                                       \
  } while (0)
  PA_(6, 1024, 0);
  PA_(6, 512, 1);
  PA_(CHMAX, 512, 2);
  PA_(6, 256, 3);
  PA_(CHMAX, 256, 4);
#undef PA_
}

void pois_table(const Ctx& c) {
  const long long nm = (long long)c.g.nxh * (c.g.nz / 2 + 1);
  hipLaunchKernelGGL(k_pois_table, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, c.st, c);
}

static void upd_x_fp64(const Ctx& c, bool chk) {
  const size_t lx = lds_line(c.o.lx, true, sizeof(double));
  XSEG(c.o.lx.seg, { KScope ks_(c, KID_UPD_X); C3D_LAUNCH(k_upd_x, (double, S_), dim3(xblocks(c)), xthreads(c), lx, c, chk ? 1 : 0); });
}
template <typename TS>
static void upd_z_t(const Ctx& c, bool chk, int refine) {
  const size_t lz = lds_line(c.o.lz, false, sizeof(double));
  // tiles of more than 512 threads (mixed mode only, 8-point segments: the
  // candidate sets offer no other) run the NT = 1024 form
  bool done = false;
  if constexpr (sizeof(TS) == 4) {
    if (c.o.lz.L * c.o.lz.G > 512) {
      KScope ks_(c, KID_UPD_Z);
      C3D_LAUNCH(k_upd_z, (double, 8, TS, 1024), dim3(zblocks(c)), zthreads(c), lz, c, chk ? 1 : 0, refine);
      done = true;
    }
  }
  if (!done) {
    ZSEG(c.o.lz.seg, { KScope ks_(c, KID_UPD_Z); C3D_LAUNCH(k_upd_z, (double, S_, TS, 512), dim3(zblocks(c)), zthreads(c), lz, c, chk ? 1 : 0, refine); });
  }
  if (chk) {
    KScope ks_(c, KID_FLOW_FIN);
    C3D_LAUNCH0(k_max_fin, dim3(1), dim3(FIN_THREADS), 0, c, xblocks(c), zblocks(c), c.s.mixed ? 1 : 0);
  }
}
// DMPRESG/DMPRESG3 (and PRESG += ...) from the partial sums of the last
// ADI sweeps: u in the z-sweep (blocks of the compute type's z shape), w in
// the final x-sweep
void flow_fin(const Ctx& c) {
  // FP64: u from the z-sweep, w from the x-sweep; mixed: u from the x-sweep,
  // w from the z-sweep (FP32 z tiles)
  const int nu = c.s.mixed ? xblocks(c) : zblocks(c, c.o.lz);
  const int nw = c.s.mixed ? zblocks(c, c.o.lz32) : xblocks(c);
  KScope ks_(c, KID_FLOW_FIN);
  C3D_LAUNCH0(k_flow_fin, dim3(1), dim3(FIN_THREADS), 0, c, nu, nw, c.s.mixed ? 0 : 1);
}
void upd_x(const Ctx& c, bool chk, bool refine) {
  if (c.s.mixed) {
    // U on x-lines, then V pointwise in the same blocks
    const size_t lx = lds_line(c.o.lx, true, sizeof(double));
    XSEG(c.o.lx.seg, { KScope ks_(c, KID_UPD_X); C3D_LAUNCH(k_upd_xu, (S_), dim3(xblocks(c)), xthreads(c), lx, c, refine ? 1 : 0); });
  } else {
    upd_x_fp64(c, chk);
  }
}
void upd_z(const Ctx& c, bool chk, bool refine) {
  if (c.s.mixed)
    upd_z_t<float>(c, chk, refine ? 1 : 0);
  else
    upd_z_t<double>(c, chk, 0);
}

// RDP = DIV/DT for the refinement solve
__global__ void k_div_rdp(Ctx c) {
  const Geo g = c.g;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.nx) return;
  const long long b = gidx(g, i, 1 + blockIdx.y, blockIdx.z);
  c.d.rdp[b] = (c.s.mixed ? (double)c.d.divf[b] : c.d.div[b]) / c.s.dt;
}
void div_rdp(const Ctx& c) {
  const int bx = c.g.nx < 256 ? c.g.nx : 256;
  KScope ks_(c, KID_RHSDP_X);
  hipLaunchKernelGGL(k_div_rdp, dim3((c.g.nx + bx - 1) / bx, c.g.n2m, c.g.nz), dim3(bx), 0, c.st, c);
}

}  // namespace c3dk
