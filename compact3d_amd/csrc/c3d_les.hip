// c3d_les.hip -- LES parts of the hot path: Smagorinsky eddy viscosity
// (SGSCS + STRAIN), the instantaneous log-law wall model, WALLSS, the SGS
// stress-divergence line solves of RHS1-3 and the explicit test filters.
#include <math.h>

#include <algorithm>

#include "c3d_impl.h"
#include "c3d_lines.h"

// ------------------------------------------------------------ STRAIN
// STRAIN (compact3d.f:5558-5689) at one cell centre: the six components
// S11 S12 S13 S22 S23 S33 from second-order differences (wall rows by the
// FJPA/FJMU flags, WMODEL != 0 doubling the y differences at J = 1, N2M).
// The factors that depend on J only are formed once per plane (StrainJ),
// each as the reference's expression evaluates it left to right, so the
// per-point arithmetic is the same.
struct StrainJ {
  double up, um;    // FJPA(J)*1/H(J+1), FJMU(J)*1/H(J)
  double cp, cm;    // wall weights: (1-FJPA)*0.5, (1-FJMU)*0.5 (WMODEL = 0); 1-FJPA, 1-FJMU otherwise
  double yf, rdy;   // YFAC/DY(J), 1/DY(J)
  double dyj, dyjp, dyjm;  // DY(J)/2, DY(J+1)/2, DY(J-1)/2
  bool wm0;
};
__device__ __forceinline__ StrainJ strain_j(const Ctx& c, int j) {
  const Metrics& M = c.m;
  const double* DY = M.dy;
  const double* H = M.h;
  const int jp = j + 1, jm = j - 1;
  const int fjum = (int)M.fjmu[j], fjup = (int)M.fjpa[j];
  StrainJ f;
  f.wm0 = (c.s.wmodel == 0.0);
  f.up = fjup * 1.0 / H[jp];
  f.um = fjum * 1.0 / H[j];
  f.cp = f.wm0 ? (1. - fjup) * 0.5 : (1. - fjup);
  f.cm = f.wm0 ? (1. - fjum) * 0.5 : (1. - fjum);
  const double yfac = (!f.wm0 && (j == 1 || j == c.g.n2m)) ? 2.0 : 1.0;
  f.yf = yfac / DY[j];
  f.rdy = 1.0 / DY[j];
  f.dyj = DY[j] / 2.0;
  f.dyjp = DY[jp] / 2.0;
  f.dyjm = DY[jm] / 2.0;
  return f;
}
// T = float (mixed mode): the FP32 shadows of the state, FP32 arithmetic
template <typename T>
__device__ __forceinline__ void strain_core(const Ctx& c, const StrainJ& f, int i, int j, int k, T sr[6]) {
  const Geo& g = c.g;
  const T* U = sizeof(T) == 4 ? (const T*)c.d.su : (const T*)c.d.u;
  const T* V = sizeof(T) == 4 ? (const T*)c.d.sv : (const T*)c.d.v;
  const T* W = sizeof(T) == 4 ? (const T*)c.d.sw : (const T*)c.d.w;
  const int ip = wrapp(i, g.nx), im = wrapm(i, g.nx), kp = wrapp(k, g.nz), km = wrapm(k, g.nz);
  const int jp = j + 1, jm = j - 1;
  const T DX1 = (T)c.s.dx1, DX3 = (T)c.s.dx3;
  const T up = (T)f.up, um = (T)f.um, cp = (T)f.cp, cm = (T)f.cm, yf = (T)f.yf, rdy = (T)f.rdy;
  const T dyj = (T)f.dyj, dyjp = (T)f.dyjp, dyjm = (T)f.dyjm, h = (T)0.5;
  // 32-bit element offsets (fields hold fewer than 2^31 elements)
  const unsigned nx = (unsigned)g.nx, pl = (unsigned)g.plane;
  auto o_ = [&](int ii, int jj, int kk) { return (unsigned)kk * pl + (unsigned)jj * nx + (unsigned)ii; };
  auto u_ = [&](int ii, int jj, int kk) { return U[o_(ii, jj, kk)]; };
  auto v_ = [&](int ii, int jj, int kk) { return V[o_(ii, jj, kk)]; };
  auto w_ = [&](int ii, int jj, int kk) { return W[o_(ii, jj, kk)]; };
  const int N2 = g.n2;
  T u2 = u_(ip, j, k), u1 = u_(i, j, k);
  sr[0] = DX1 * (u2 - u1);
  T ujp = h * (u_(i, jp, k) + u_(ip, jp, k));
  T ujc = h * (u_(i, j, k) + u_(ip, j, k));
  T ujm = h * (u_(i, jm, k) + u_(ip, jm, k));
  if (f.wm0) {
    u2 = up * (dyj * ujp + dyjp * ujc) + cp * (u_(i, N2, k) + u_(ip, N2, k));
    u1 = um * (dyj * ujm + dyjm * ujc) + cm * (u_(i, 0, k) + u_(ip, 0, k));
  } else {
    u2 = up * (dyj * ujp + dyjp * ujc) + cp * ujc;
    u1 = um * (dyj * ujm + dyjm * ujc) + cm * ujc;
  }
  T vip = h * (v_(ip, j, k) + v_(ip, jp, k));
  T vic = h * (v_(i, j, k) + v_(i, jp, k));
  T vim = h * (v_(im, j, k) + v_(im, jp, k));
  T v2 = h * (vip + vic), v1 = h * (vim + vic);
  sr[1] = h * (DX1 * (v2 - v1) + yf * (u2 - u1));
  T ukp = h * (u_(i, j, kp) + u_(ip, j, kp));
  T ukc = h * (u_(i, j, k) + u_(ip, j, k));
  T ukm = h * (u_(i, j, km) + u_(ip, j, km));
  u2 = h * (ukp + ukc);
  u1 = h * (ukm + ukc);
  T wip = h * (w_(ip, j, k) + w_(ip, j, kp));
  T wic = h * (w_(i, j, k) + w_(i, j, kp));
  T wim = h * (w_(im, j, k) + w_(im, j, kp));
  T w2 = h * (wip + wic), w1 = h * (wim + wic);
  sr[2] = h * (DX1 * (w2 - w1) + DX3 * (u2 - u1));
  v2 = v_(i, jp, k);
  v1 = v_(i, j, k);
  sr[3] = rdy * (v2 - v1);
  T wjp = h * (w_(i, jp, k) + w_(i, jp, kp));
  T wjc = h * (w_(i, j, k) + w_(i, j, kp));
  T wjm = h * (w_(i, jm, k) + w_(i, jm, kp));
  if (f.wm0) {
    w2 = up * (dyj * wjp + dyjp * wjc) + cp * (w_(i, N2, k) + w_(i, N2, kp));
    w1 = um * (dyj * wjm + dyjm * wjc) + cm * (w_(i, 0, k) + w_(i, 0, kp));
  } else {
    w2 = up * (dyj * wjp + dyjp * wjc) + cp * wjc;
    w1 = um * (dyj * wjm + dyjm * wjc) + cm * wjc;
  }
  T vkp = h * (v_(i, j, kp) + v_(i, jp, kp));
  T vkc = h * (v_(i, j, k) + v_(i, jp, k));
  T vkm = h * (v_(i, j, km) + v_(i, jp, km));
  v2 = h * (vkp + vkc);
  v1 = h * (vkm + vkc);
  sr[4] = h * (yf * (w2 - w1) + DX3 * (v2 - v1));
  w2 = w_(i, j, kp);
  w1 = w_(i, j, k);
  sr[5] = DX3 * (w2 - w1);
}
__device__ __forceinline__ void strain_at(const Ctx& c, int i, int j, int k, double sr[6]) {
  strain_core<double>(c, strain_j(c, j), i, j, k, sr);
}

// 2 S:S with the off-diagonal components counted twice, in STRAIN's order
template <typename T>
__device__ __forceinline__ T sdots(const T s[6]) {
  return (T)2 * (s[1] * s[1]) + s[0] * s[0] + (T)2 * (s[2] * s[2]) + s[3] * s[3] + (T)2 * (s[4] * s[4]) + s[5] * s[5];
}

// ------------------------------------------------------------ SGSCS
// SGSCS (compact3d.f:1413-1488) with STRAIN: |S| = sqrt(2 S:S),
// nu_t = max((0.1 (1-exp(-y+/26)))^2 (DY/(DX1 DX3))^(2/3) |S|, -1/RE),
// y+ = (Y(J)+DY(J)/2) u_tau RE, u_tau = (WSM1^2 + WSM3^2)^(1/4) (previous step).
// One block per (tile of 64 i x 4 runs of kpt consecutive k, plane J): the
// plane's factors (STRAIN's and the model constant times the filter width)
// once per thread; each thread walks its run of k with the stencil values
// of the planes K-1, K and K+1 in registers (SPlane), so every value of a
// plane is loaded once instead of once per point that reads it.  The
// arithmetic of each point is strain_core's, term for term.
#define SG_TI 64
// the values of one k-plane that STRAIN reads at (I, J) (as plane K or K+1;
// plane K-1 uses the U and V entries at (I, J) and (I, J+1) only)
template <typename T>
struct SPlane {
  T u_i, u_ip, u_ijp, u_ipjp, u_ijm, u_ipjm, u_in2, u_ipn2, u_i0, u_ip0;
  T v_ip, v_ipjp, v_i, v_ijp, v_im, v_imjp;
  T w_ip, w_i, w_im, w_ijp, w_ijm, w_in2, w_i0;
};
template <typename T>
__device__ __forceinline__ SPlane<T> splane(const Ctx& c, bool wm0, int i, int ip, int im, int j, int k) {
  const Geo& g = c.g;
  const T* U = sizeof(T) == 4 ? (const T*)c.d.su : (const T*)c.d.u;
  const T* V = sizeof(T) == 4 ? (const T*)c.d.sv : (const T*)c.d.v;
  const T* W = sizeof(T) == 4 ? (const T*)c.d.sw : (const T*)c.d.w;
  const unsigned nx = (unsigned)g.nx;
  const unsigned r0 = (unsigned)k * (unsigned)g.plane, rj = r0 + (unsigned)j * nx;
  const unsigned rjp = rj + nx, rjm = rj - nx, rn2 = r0 + (unsigned)g.n2 * nx;
  SPlane<T> p;
  p.u_i = U[rj + i];
  p.u_ip = U[rj + ip];
  p.u_ijp = U[rjp + i];
  p.u_ipjp = U[rjp + ip];
  p.u_ijm = U[rjm + i];
  p.u_ipjm = U[rjm + ip];
  p.v_ip = V[rj + ip];
  p.v_ipjp = V[rjp + ip];
  p.v_i = V[rj + i];
  p.v_ijp = V[rjp + i];
  p.v_im = V[rj + im];
  p.v_imjp = V[rjp + im];
  p.w_ip = W[rj + ip];
  p.w_i = W[rj + i];
  p.w_im = W[rj + im];
  p.w_ijp = W[rjp + i];
  p.w_ijm = W[rjm + i];
  if (wm0) {
    p.u_in2 = U[rn2 + i];
    p.u_ipn2 = U[rn2 + ip];
    p.u_i0 = U[r0 + i];
    p.u_ip0 = U[r0 + ip];
    p.w_in2 = W[rn2 + i];
    p.w_i0 = W[r0 + i];
  } else {
    p.u_in2 = p.u_ipn2 = p.u_i0 = p.u_ip0 = p.w_in2 = p.w_i0 = (T)0;
  }
  return p;
}
// strain_core's six components from the planes K-1 (pm), K (p0), K+1 (pp)
template <typename T>
__device__ __forceinline__ void strain_win(const Ctx& c, const StrainJ& f, const SPlane<T>& pm, const SPlane<T>& p0,
                                           const SPlane<T>& pp, T sr[6]) {
  const T DX1 = (T)c.s.dx1, DX3 = (T)c.s.dx3;
  const T up = (T)f.up, um = (T)f.um, cp = (T)f.cp, cm = (T)f.cm, yf = (T)f.yf, rdy = (T)f.rdy;
  const T dyj = (T)f.dyj, dyjp = (T)f.dyjp, dyjm = (T)f.dyjm, h = (T)0.5;
  T u2 = p0.u_ip, u1 = p0.u_i;
  sr[0] = DX1 * (u2 - u1);
  T ujp = h * (p0.u_ijp + p0.u_ipjp);
  T ujc = h * (p0.u_i + p0.u_ip);
  T ujm = h * (p0.u_ijm + p0.u_ipjm);
  if (f.wm0) {
    u2 = up * (dyj * ujp + dyjp * ujc) + cp * (p0.u_in2 + p0.u_ipn2);
    u1 = um * (dyj * ujm + dyjm * ujc) + cm * (p0.u_i0 + p0.u_ip0);
  } else {
    u2 = up * (dyj * ujp + dyjp * ujc) + cp * ujc;
    u1 = um * (dyj * ujm + dyjm * ujc) + cm * ujc;
  }
  T vip = h * (p0.v_ip + p0.v_ipjp);
  T vic = h * (p0.v_i + p0.v_ijp);
  T vim = h * (p0.v_im + p0.v_imjp);
  T v2 = h * (vip + vic), v1 = h * (vim + vic);
  sr[1] = h * (DX1 * (v2 - v1) + yf * (u2 - u1));
  T ukp = h * (pp.u_i + pp.u_ip);
  T ukc = h * (p0.u_i + p0.u_ip);
  T ukm = h * (pm.u_i + pm.u_ip);
  u2 = h * (ukp + ukc);
  u1 = h * (ukm + ukc);
  T wip = h * (p0.w_ip + pp.w_ip);
  T wic = h * (p0.w_i + pp.w_i);
  T wim = h * (p0.w_im + pp.w_im);
  T w2 = h * (wip + wic), w1 = h * (wim + wic);
  sr[2] = h * (DX1 * (w2 - w1) + DX3 * (u2 - u1));
  v2 = p0.v_ijp;
  v1 = p0.v_i;
  sr[3] = rdy * (v2 - v1);
  T wjp = h * (p0.w_ijp + pp.w_ijp);
  T wjc = h * (p0.w_i + pp.w_i);
  T wjm = h * (p0.w_ijm + pp.w_ijm);
  if (f.wm0) {
    w2 = up * (dyj * wjp + dyjp * wjc) + cp * (p0.w_in2 + pp.w_in2);
    w1 = um * (dyj * wjm + dyjm * wjc) + cm * (p0.w_i0 + pp.w_i0);
  } else {
    w2 = up * (dyj * wjp + dyjp * wjc) + cp * wjc;
    w1 = um * (dyj * wjm + dyjm * wjc) + cm * wjc;
  }
  T vkp = h * (pp.v_i + pp.v_ijp);
  T vkc = h * (p0.v_i + p0.v_ijp);
  T vkm = h * (pm.v_i + pm.v_ijp);
  v2 = h * (vkp + vkc);
  v1 = h * (vkm + vkc);
  sr[4] = h * (yf * (w2 - w1) + DX3 * (v2 - v1));
  w2 = pp.w_i;
  w1 = p0.w_i;
  sr[5] = DX3 * (w2 - w1);
}

// T = float (mixed mode): STRAIN from the FP32 shadows in FP32; SGSVIS is
// stored in FP64 and as its FP32 shadow
template <typename T, bool CB_>
__global__ __launch_bounds__(256) void k_sgscs(CtxArg<CB_> ca_, int kpt) {
  const Ctx c = ctx_get(ca_);
  const Geo g = c.g;
  const Scal s = c.s;
  const Metrics M = c.m;
  const Fields F = c.d;
  const int ntx = (g.nx + SG_TI - 1) / SG_TI;
  const int ti = blockIdx.x % ntx, tk = blockIdx.x / ntx;
  const int j = 1 + blockIdx.y;
  const int i = ti * SG_TI + (threadIdx.x % SG_TI);
  if (i >= g.nx) return;
  const double* DY = M.dy;
  const StrainJ fj = strain_j(c, j);
  const double w1s = F.sc[SC_WSM1], w3s = F.sc[SC_WSM3];
  const double utau = sqrt(sqrt(w1s * w1s + w3s * w3s));
  double cs = 0.1;
  const double yplus = (M.y[j] + DY[j] * 0.5) * utau * s.re;
  cs = cs * (1 - exp(-yplus / 26.0));
  const double smagc = cs * cs;
  const double dxyz = 1.0 / s.dx1 * DY[j] / s.dx3;
  const T coef = (T)(smagc * pow(dxyz, 2. / 3.));
  const T nmin = (T)(-s.rinv);
  const int kb = (tk * 4 + threadIdx.x / SG_TI) * kpt;
  if (kb >= g.nz) return;
  const int ke = min(kb + kpt, g.nz);
  const int ip = wrapp(i, g.nx), im = wrapm(i, g.nx);
  SPlane<T> pm = splane<T>(c, fj.wm0, i, ip, im, j, wrapm(kb, g.nz));
  SPlane<T> p0 = splane<T>(c, fj.wm0, i, ip, im, j, kb);
  for (int k = kb; k < ke; k++) {
    const SPlane<T> pp = splane<T>(c, fj.wm0, i, ip, im, j, wrapp(k, g.nz));
    T sr[6];
    strain_win<T>(c, fj, pm, p0, pp, sr);
    const T sumsr = sdots(sr);
    const T abssr = sqrt((T)2 * sumsr);
    const T turbv = coef * abssr;
    const T nu = fmax(turbv, nmin);
    const unsigned a = (unsigned)k * (unsigned)g.plane + (unsigned)j * (unsigned)g.nx + (unsigned)i;
    F.sgs[a] = (double)nu;
    if (sizeof(T) == 4) F.sgsf[a] = (float)nu;
    pm = p0;
    p0 = pp;
  }
}

// ------------------------------------------------------------ SGSDY
// SGSDY (compact3d.f:1250-1400), the dynamic model, per J plane in three
// kernels.  k_sgsdy_tile: a block takes a DT x DT tile of (I,K) with a
// one-point periodic halo and forms at every point of it STRAIN, |S| (kept
// in SGSVIS), COM1 = DELTA^2 |S| S_ij, the cell-centre velocities UC and
// their products COM2; after a barrier each interior point applies the
// test filter FILTER (Simpson (1,4,1)/6 in x, then in z, :5190-5228) to
// the 21 quantities from LDS, forms M_ij = 4 DELTA^2 |S~| S~_ij - COM1~_ij
// and L_ij = COM2~_ij - UC~_i UC~_j, and the block's sums of L:M and M:M go
// to a per-plane partial array.  k_sgsdy_coef: C(J) = -0.5 <L:M>/<M:M> from
// the partials in a fixed order.  k_sgsdy_clip: SGSVIS = max(C(J)
// DELTA^2 |S|, -1/RE).  SGSSD's last stage is the same with the factor
// 4 BETA(J) in place of 4 (beta != NULL).
#define DT_TILE 16
#define DT_HALO (DT_TILE + 2)
#define DT_NQ 21
__global__ __launch_bounds__(DT_TILE* DT_TILE) void k_sgsdy_tile(Ctx c, const double* beta) {
  __shared__ double Q[DT_NQ][DT_HALO * DT_HALO];
  const Geo g = c.g;
  const Scal s = c.s;
  const Fields F = c.d;
  const int tix = (g.nx + DT_TILE - 1) / DT_TILE, tiz = (g.nz + DT_TILE - 1) / DT_TILE;
  const int ntile = tix * tiz;
  const int j = 1 + blockIdx.x / ntile;
  const int tile = blockIdx.x % ntile;
  const int i0 = (tile % tix) * DT_TILE, k0 = (tile / tix) * DT_TILE;
  const double dxyz = 1.0 / s.dx1 * c.m.dy[j] / s.dx3;
  const double delta2 = pow(dxyz, 2. / 3.);
  // every point of the tile and its halo
  for (int e = threadIdx.x; e < DT_HALO * DT_HALO; e += blockDim.x) {
    const int px = e % DT_HALO, pz = e / DT_HALO;
    int i = i0 - 1 + px, k = k0 - 1 + pz;
    i = ((i % g.nx) + g.nx) % g.nx;
    k = ((k % g.nz) + g.nz) % g.nz;
    double sr[6];
    strain_at(c, i, j, k, sr);
    const double abssr = sqrt(2. * sdots(sr));
    const bool inner = px >= 1 && px <= DT_TILE && pz >= 1 && pz <= DT_TILE && i0 - 1 + px < g.nx &&
                       k0 - 1 + pz < g.nz;
    if (inner) F.sgs[gidx(g, i, j, k)] = abssr;
    const double delabs = delta2 * abssr;
#pragma unroll
    for (int m = 0; m < 6; m++) {
      Q[m][e] = delabs * sr[m];  // COM1
      Q[6 + m][e] = sr[m];       // S
    }
    const int ip = wrapp(i, g.nx), kp = wrapp(k, g.nz);
    const double u1 = (F.u[gidx(g, i, j, k)] + F.u[gidx(g, ip, j, k)]) * 0.5;
    const double u2 = (F.v[gidx(g, i, j, k)] + F.v[gidx(g, i, j + 1, k)]) * 0.5;
    const double u3 = (F.w[gidx(g, i, j, k)] + F.w[gidx(g, i, j, kp)]) * 0.5;
    Q[12][e] = u1;
    Q[13][e] = u2;
    Q[14][e] = u3;
    Q[15][e] = u1 * u1;
    Q[16][e] = u1 * u2;
    Q[17][e] = u1 * u3;
    Q[18][e] = u2 * u2;
    Q[19][e] = u2 * u3;
    Q[20][e] = u3 * u3;
  }
  __syncthreads();
  const int tx = threadIdx.x % DT_TILE, tz = threadIdx.x / DT_TILE;
  const bool ok = i0 + tx < g.nx && k0 + tz < g.nz;
  // FILTER: x pass on rows z-1, z, z+1, then the z pass
  auto filt = [&](int q) {
    const double* A = Q[q];
    double fx[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
      const int row = (tz + r) * DT_HALO + tx + 1;
      fx[r] = (A[row - 1] + 4.0 * A[row] + A[row + 1]) / 6.0;
    }
    return (fx[0] + 4.0 * fx[1] + fx[2]) / 6.0;
  };
  double num = 0.0, den = 0.0;
  if (ok) {
    double fc[6], fs[6];
#pragma unroll
    for (int m = 0; m < 6; m++) {
      fc[m] = filt(m);
      fs[m] = filt(6 + m);
    }
    const double absfsr = sqrt(2. * sdots(fs));
    double fdxyz = 4.0 * pow(dxyz, 2. / 3.);
    if (beta) fdxyz = fdxyz * beta[j];  // SGSSD's scale-dependent factor (:1823)
    const double fdabs = fdxyz * absfsr;
    double mm[6];
#pragma unroll
    for (int m = 0; m < 6; m++) mm[m] = fdabs * fs[m] - fc[m];
    const double c1 = filt(12), c2 = filt(13), c3 = filt(14);
    const double xl11 = filt(15) - c1 * c1, xl12 = filt(16) - c1 * c2, xl13 = filt(17) - c1 * c3;
    const double xl22 = filt(18) - c2 * c2, xl23 = filt(19) - c2 * c3, xl33 = filt(20) - c3 * c3;
    den = sdots(mm);
    num = 2. * xl12 * mm[1] + xl11 * mm[0] + 2. * xl13 * mm[2] + xl22 * mm[3] + 2. * xl23 * mm[4] + xl33 * mm[5];
  }
  block_sum2(num, den, F.dyp + 2 * ((long long)(j - 1) * ntile + tile),
             F.dyp + 2 * ((long long)(j - 1) * ntile + tile) + 1);
}

// C(J) per plane from the tile partials, summed in tile order
__global__ __launch_bounds__(256) void k_sgsdy_coef(Ctx c, int ntile) {
  __shared__ double red[2][256];
  const Geo g = c.g;
  const int j = 1 + blockIdx.x;
  const double* P = c.d.dyp + 2 * (long long)(j - 1) * ntile;
  double a = 0.0, b = 0.0;
  for (int t = threadIdx.x; t < ntile; t += blockDim.x) {
    a += P[2 * t];
    b += P[2 * t + 1];
  }
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      red[0][threadIdx.x] += red[0][threadIdx.x + h];
      red[1][threadIdx.x] += red[1][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double nxz = 1.0 / (double)(g.nx * g.nz);
    const double sumnum = red[0][0] * nxz, sumden = red[1][0] * nxz;
    c.d.smagc[j] = -0.5 * sumnum / sumden;
  }
}

__global__ void k_sgsdy_clip(Ctx c) {
  const Geo g = c.g;
  const long long tot = (long long)g.nx * g.n2m * g.nz;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= tot) return;
  const int i = t % g.nx;
  const int j = 1 + (int)((t / g.nx) % g.n2m);
  const int k = (int)(t / ((long long)g.nx * g.n2m));
  const double dxyz = 1.0 / c.s.dx1 * c.m.dy[j] / c.s.dx3;
  const long long a = gidx(g, i, j, k);
  const double turbv = c.d.smagc[j] * pow(dxyz, 2. / 3.) * c.d.sgs[a];
  const double nu = fmax(turbv, -c.s.rinv);
  c.d.sgs[a] = nu;
  if (c.d.sgsf) c.d.sgsf[a] = (float)nu;
}

// ------------------------------------------------------------ SGSSD
// SGSSD (compact3d.f:1502-1948), the scale-dependent dynamic model, per J
// plane: the sums A..E of the Germano terms at the test-filter levels
// FILTER (2 DELTA) and FILTER4 (4 DELTA) (:1544-1800), the quintic of
// COEFFICIENT (:2089-2093) and its root BETA (DQRRT/CHHQR), then SGSDY's
// stage with the factor 4 BETA (:1804-1930).
//
// k_sgssd_tile<H>: a block takes a 16 x 16 tile of (I,K) with an H-point
// periodic halo (H = 1 for FILTER, 2 for FILTER4) in two phases that share
// one LDS buffer: STRAIN -> COM1 = DELTA^2 |S| S_ij and S_ij at every point,
// filtered at the interior points to COM1~ and SR1 = DELTA^2 |S~| S~_ij;
// then the cell-centre velocities UC and their products, filtered, give
// L_ij.  The block's five sums go to a per-plane partial array.
template <int H>
__global__ __launch_bounds__(DT_TILE* DT_TILE) void k_sgssd_tile(Ctx c, double* part) {
  constexpr int HD = DT_TILE + 2 * H;
  __shared__ double Q[12][HD * HD];
  const Geo g = c.g;
  const Scal s = c.s;
  const Fields F = c.d;
  const int tix = (g.nx + DT_TILE - 1) / DT_TILE, tiz = (g.nz + DT_TILE - 1) / DT_TILE;
  const int ntile = tix * tiz;
  const int j = 1 + blockIdx.x / ntile;
  const int tile = blockIdx.x % ntile;
  const int i0 = (tile % tix) * DT_TILE, k0 = (tile / tix) * DT_TILE;
  const double dxyz = 1.0 / s.dx1 * c.m.dy[j] / s.dx3;
  const double delta2 = pow(dxyz, 2. / 3.);
  auto wrap_ik = [&](int e, int& i, int& k) {
    i = i0 - H + e % HD;
    k = k0 - H + e / HD;
    i = ((i % g.nx) + g.nx) % g.nx;
    k = ((k % g.nz) + g.nz) % g.nz;
  };
  for (int e = threadIdx.x; e < HD * HD; e += blockDim.x) {
    int i, k;
    wrap_ik(e, i, k);
    double sr[6];
    strain_at(c, i, j, k, sr);
    const double delabs = delta2 * sqrt(2. * sdots(sr));
#pragma unroll
    for (int m = 0; m < 6; m++) {
      Q[m][e] = delabs * sr[m];
      Q[6 + m][e] = sr[m];
    }
  }
  __syncthreads();
  const int tx = threadIdx.x % DT_TILE, tz = threadIdx.x / DT_TILE;
  const bool ok = i0 + tx < g.nx && k0 + tz < g.nz;
  // FILTER (:5190-5228) / FILTER4 (:5481-5546): x pass on the 2H+1 rows
  // around the point, then the z pass
  auto filt = [&](int q) {
    const double* A = Q[q];
    if constexpr (H == 1) {
      double fx[3];
#pragma unroll
      for (int r = 0; r < 3; r++) {
        const int row = (tz + r) * HD + tx + 1;
        fx[r] = (A[row - 1] + 4.0 * A[row] + A[row + 1]) / 6.0;
      }
      return (fx[0] + 4.0 * fx[1] + fx[2]) / 6.0;
    } else {
      double fx[5];
#pragma unroll
      for (int r = 0; r < 5; r++) {
        const int row = (tz + r) * HD + tx + 2;
        fx[r] = (A[row - 2] + 4.0 * A[row - 1] + 2.0 * A[row] + 4.0 * A[row + 1] + A[row + 2]) / 12.0;
      }
      return (fx[0] + 4.0 * fx[1] + 2.0 * fx[2] + 4.0 * fx[3] + fx[4]) / 12.0;
    }
  };
  double fc[6], sr1[6];
  if (ok) {
    double fs[6];
#pragma unroll
    for (int m = 0; m < 6; m++) {
      fc[m] = filt(m);
      fs[m] = filt(6 + m);
    }
    const double fdabs = pow(dxyz, 2. / 3.) * sqrt(2. * sdots(fs));
#pragma unroll
    for (int m = 0; m < 6; m++) sr1[m] = fdabs * fs[m];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < HD * HD; e += blockDim.x) {
    int i, k;
    wrap_ik(e, i, k);
    const int ip = wrapp(i, g.nx), kp = wrapp(k, g.nz);
    const double u1 = (F.u[gidx(g, i, j, k)] + F.u[gidx(g, ip, j, k)]) * 0.5;
    const double u2 = (F.v[gidx(g, i, j, k)] + F.v[gidx(g, i, j + 1, k)]) * 0.5;
    const double u3 = (F.w[gidx(g, i, j, k)] + F.w[gidx(g, i, j, kp)]) * 0.5;
    Q[0][e] = u1;
    Q[1][e] = u2;
    Q[2][e] = u3;
    Q[3][e] = u1 * u1;
    Q[4][e] = u1 * u2;
    Q[5][e] = u1 * u3;
    Q[6][e] = u2 * u2;
    Q[7][e] = u2 * u3;
    Q[8][e] = u3 * u3;
  }
  __syncthreads();
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (ok) {
    const double c1 = filt(0), c2 = filt(1), c3 = filt(2);
    const double xl11 = filt(3) - c1 * c1, xl12 = filt(4) - c1 * c2, xl13 = filt(5) - c1 * c3;
    const double xl22 = filt(6) - c2 * c2, xl23 = filt(7) - c2 * c3, xl33 = filt(8) - c3 * c3;
    v[0] = 2. * xl12 * sr1[1] + xl11 * sr1[0] + 2. * xl13 * sr1[2] + xl22 * sr1[3] + 2. * xl23 * sr1[4] +
           xl33 * sr1[5];
    v[1] = 2. * xl12 * fc[1] + xl11 * fc[0] + 2. * xl13 * fc[2] + xl22 * fc[3] + 2. * xl23 * fc[4] + xl33 * fc[5];
    v[2] = 2. * fc[1] * fc[1] + fc[0] * fc[0] + 2. * fc[2] * fc[2] + fc[3] * fc[3] + 2. * fc[4] * fc[4] +
           fc[5] * fc[5];
    v[3] = 2. * sr1[1] * sr1[1] + sr1[0] * sr1[0] + 2. * sr1[2] * sr1[2] + sr1[3] * sr1[3] +
           2. * sr1[4] * sr1[4] + sr1[5] * sr1[5];
    v[4] = 2. * sr1[1] * fc[1] + sr1[0] * fc[0] + 2. * sr1[2] * fc[2] + sr1[3] * fc[3] + 2. * sr1[4] * fc[4] +
           sr1[5] * fc[5];
  }
  // block sums in a fixed order: wavefront butterflies, then the
  // wavefronts in index order (the block is 4 full wavefronts)
  __shared__ double rs[5][DT_TILE * DT_TILE / 64];
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const double a = wave_sum(v[q]);
    if ((threadIdx.x & 63) == 0) rs[q][threadIdx.x >> 6] = a;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    double a = 0.0;
    for (int w = 0; w < DT_TILE * DT_TILE / 64; w++) a += rs[threadIdx.x][w];
    part[5 * ((long long)(j - 1) * ntile + tile) + threadIdx.x] = a;
  }
}

// CHHQR (compact3d.f:11571-11677): eigenvalues of an upper Hessenberg
// matrix by the double-shift QR iteration, deflating where a subdiagonal
// entry is at most eps times its diagonal neighbours.  With eps = 1e-5 the
// roots carry that deflation's error and depend on the iteration path, so
// the steps are the reference's: no exceptional shifts, its reflector
// formulas, at most 60 iterations per eigenvalue (returns 0 when reached,
// leaving the rest unset).  A: n x n column-major, n <= 5.
__device__ int sd_chhqr(double* A, int n, double* wr, double* wi, double eps) {
#pragma clang fp contract(off)
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
      const double cc = HA(m, m) * HA(m - 1, m - 1) - HA(m, m - 1) * HA(m - 1, m);
      const double w = b * b - 4 * cc;
      const double y = sqrt(fabs(w));
      if (w > 0.0) {
        const double xy = (b < 0.0) ? -1.0 : 1.0;
        wr[m - 1] = (-b - xy * y) / 2.0;
        wr[m - 2] = cc / wr[m - 1];
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
    for (int jj = l + 2; jj <= m; jj++) HA(jj, jj - 2) = 0.0;
    for (int jj = l + 3; jj <= m; jj++) HA(jj, jj - 3) = 0.0;
    for (int k = l; k <= m - 1; k++) {
      double p, q, r, x = 0.0, y = 0.0;
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
        const double sg = xy * sqrt(p * p + q * q + r * r);
        if (k != l) HA(k, k - 1) = -sg;
        const double e = -q / sg, f = -r / sg;
        x = -p / sg;
        y = -x - f * r / (p + sg);
        const double gg = e * r / (p + sg);
        const double z = -x - e * q / (p + sg);
        for (int jj = k; jj <= m; jj++) {
          p = x * HA(k, jj) + e * HA(k + 1, jj);
          q = e * HA(k, jj) + y * HA(k + 1, jj);
          r = f * HA(k, jj) + gg * HA(k + 1, jj);
          if (k != m - 1) {
            p = p + f * HA(k + 2, jj);
            q = q + gg * HA(k + 2, jj);
            r = r + z * HA(k + 2, jj);
            HA(k + 2, jj) = r;
          }
          HA(k + 1, jj) = q;
          HA(k, jj) = p;
        }
        const int imax = (k + 3 >= m) ? m : k + 3;
        for (int i = l; i <= imax; i++) {
          p = x * HA(i, k) + e * HA(i, k + 1);
          q = e * HA(i, k) + y * HA(i, k + 1);
          r = f * HA(i, k) + gg * HA(i, k + 1);
          if (k != m - 1) {
            p = p + f * HA(i, k + 2);
            q = q + gg * HA(i, k + 2);
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

// per plane: the two levels' sums from the tile partials (tile order),
// scaled as :1643-1667 and :1772-1800; COEFFICIENT's quintic; BETA = the
// largest real part of a root in (0, 2) (DQRRT :11548-11570), else 0
__global__ __launch_bounds__(256) void k_sgssd_coef(Ctx c, int ntile, const double* p1, const double* p2) {
  __shared__ double red[10][256];
  const Geo g = c.g;
  const int j = 1 + blockIdx.x;
  double a[10];
#pragma unroll
  for (int q = 0; q < 10; q++) a[q] = 0.0;
  for (int t = threadIdx.x; t < ntile; t += blockDim.x) {
    const long long o = 5 * ((long long)(j - 1) * ntile + t);
#pragma unroll
    for (int q = 0; q < 5; q++) {
      a[q] += p1[o + q];
      a[5 + q] += p2[o + q];
    }
  }
#pragma unroll
  for (int q = 0; q < 10; q++) red[q][threadIdx.x] = a[q];
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h)
#pragma unroll
      for (int q = 0; q < 10; q++) red[q][threadIdx.x] += red[q][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double nxz = 1.0 / (double)(g.nx * g.nz);
  const double A1 = -8.0 * red[0][0] * nxz, B1 = -2.0 * red[1][0] * nxz, C1 = 4.0 * red[2][0] * nxz;
  const double D1 = 64.0 * red[3][0] * nxz, E1 = 32.0 * red[4][0] * nxz;
  const double A2 = -32.0 * red[5][0] * nxz, B2 = -2.0 * red[6][0] * nxz, C2 = 4.0 * red[7][0] * nxz;
  const double D2 = 1024.0 * red[8][0] * nxz, E2 = 128.0 * red[9][0] * nxz;
  double pa[5];
  pa[4] = (B2 * C1 - B1 * C2) / (A1 * D2);
  pa[3] = (A1 * C2 - B2 * E1) / (A1 * D2);
  pa[2] = (B2 * D1 + B1 * E2 - A2 * C1) / (A1 * D2);
  pa[1] = (A2 * E1 - A1 * E2) / (A1 * D2);
  pa[0] = -(A2 * D1 + B1 * D2) / (A1 * D2);
  // companion matrix of x^5 + pa[0] x^4 + ... + pa[4]
  double B[25], wr[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, wi[5];
  for (int q = 0; q < 25; q++) B[q] = 0.0;
  for (int q = 0; q < 5; q++) B[q * 5] = -pa[q];
  for (int q = 0; q < 4; q++) B[q * 5 + q + 1] = 1.0;
  sd_chhqr(B, 5, wr, wi, 1.0E-5);
  double beta = 0.0;
  for (int q = 0; q < 5; q++)
    if (wr[q] > beta && wr[q] < 2.0) beta = wr[q];
  c.d.sdbeta[j] = beta;
}

// --------------------------------------------------------- wall model
// WALLMODEL_INSTANT (compact3d.f:2319-2493): per wall point the damped Newton
// iteration of U/u_tau = ln(u_tau*DELTAP*RE)/kappa + B on the third cell,
// tau_w = u_tau^2 of the last accepted iterate, TAUW = tau_w u/U,
// TAUB = tau_w w/U; WSM(1) = TAOW = mean tau_w.  The reference iterates
// without a bound; here the loop also stops after 10000 iterations (the
// damped iteration converges in tens).
#define WM_MAXIT 10000
// `top` (WMODEL = 6, a two-wall extension with no reference counterpart): the same iteration
// at the top wall, on the third cell from it (J = N2M-2, wall distance
// DY(N2M) + DY(N2M-1) + DY(N2M-2)/2), into TAUW(I,K,2), TAUB(I,K,2); its
// plane sums are not kept (WSM(1) is the bottom wall's).
template <bool CB_>
__global__ __launch_bounds__(C3D_NT) void k_wallmodel(CtxArg<CB_> ca_, int top) {
  const Ctx c = ctx_get(ca_);
  __shared__ double red[C3D_NT];
  const Geo g = c.g;
  const Metrics M = c.m;
  const Fields F = c.d;
  const long long n = (long long)g.nx * g.nz;
  const double RE = c.s.re, KAPA = 0.41, B = 5.2;
  const int jw = top ? g.n2m - 2 : 3;
  const double deltap = top ? M.dy[g.n2m] + M.dy[g.n2m - 1] + M.dy[g.n2m - 2] / 2.0
                            : M.dy[1] + M.dy[2] + M.dy[3] / 2.0;
  double* TW = F.tauw + (top ? n : 0);
  double* TB = F.taub + (top ? n : 0);
  double acc = 0.0;
  for (long long t = blockIdx.x * (long long)C3D_NT + threadIdx.x; t < n; t += (long long)gridDim.x * C3D_NT) {
    const int i = t % g.nx, k = t / g.nx;
    const double u3 = F.u[gidx(g, i, jw, k)], w3 = F.w[gidx(g, i, jw, k)];
    const double u_ave = sqrt(u3 * u3 + w3 * w3);
    double ut1 = 1.0, ut2 = 0.000001, f1, f2, f3, lam = 1.0;
    int it = 0;
    while (fabs(ut2 - ut1) > 1.0E-6 && it < WM_MAXIT) {
      ut1 = ut2;
      f1 = u_ave / ut1 - log(ut1 * deltap * RE) / KAPA - B;
      f2 = -u_ave / (ut1 * ut1) - 1. / KAPA / ut1;
      ut2 = ut1 - lam * f1 / f2;
      f3 = u_ave / ut2 - log(ut2 * deltap * RE) / KAPA - B;
      if (fabs(f3) > fabs(f1)) lam = 0.5 * lam;
      it++;
    }
    const double tauw1 = ut1 * ut1;
    acc += tauw1;
    TW[t] = tauw1 * u3 / u_ave;
    TB[t] = tauw1 * w3 / u_ave;
  }
  if (top) return;
  const int t = threadIdx.x;
  red[t] = acc;
  __syncthreads();
  for (int h = C3D_NT / 2; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  if (t == 0) F.red[blockIdx.x * 4] = red[0];
}

template <bool CB_>
__global__ __launch_bounds__(C3D_NT) void k_wallmodel_fin(CtxArg<CB_> ca_, int nblk, int zero23) {
  const Ctx c = ctx_get(ca_);
  __shared__ double red[C3D_NT];
  const Fields F = c.d;
  const int t = threadIdx.x;
  double a = 0.0;
  for (int b = t; b < nblk; b += C3D_NT) a += F.red[b * 4];
  red[t] = a;
  __syncthreads();
  for (int h = C3D_NT / 2; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  if (t == 0) {
    const double w = red[0] / (double)((long long)c.g.nx * c.g.nz);
    F.sc[SC_WSM1] = w;
    F.sc[SC_TAOW] = w;
    if (zero23) {
      F.sc[SC_WSM2] = 0.0;
      F.sc[SC_WSM3] = 0.0;
    }
  }
}

// WALLMODEL_IWMLES_SMOOTH (compact3d.f:3386-4329), WMODEL = 2: the integral
// wall model of Yang et al. (2015) at the bottom wall.  Every fifth step
// (NTIME a multiple of 5) the third cell's velocity and the first cell's
// pressure fluctuation are time-filtered (k_iwm_filter), the integrated
// momentum balance of the wall layer is formed from the integrals of the
// last compute step and their x/z differences (k_iwm_lhs), and each point
// solves it for its friction velocities (utx, uty) by Newton's method with a
// finite-difference Jacobian, then stores the new integrals, the diffusion
// term, the stresses and the filter weight (k_iwm_solve).  Every step the
// stored stresses are imposed and WSM(1) = TAOW = mean |tau| (k_iwm_apply,
// k_wallmodel_fin).  The kernels read NTIME from the device (k_iwm_begin
// advances it), so a captured step replays correctly.  As in the
// reference: x**2.0 and x**3.0 are products, x**4.0 and x**0.25 pow.
// Where the reference is undefined: the spanwise stress goes to
// TAUB(I,K,1), the element its TAUW(I,K,3) addresses through COMMON
// /TAUXYZ/ (:3549); the excessive-linear-correction branch (:4139) uses the
// point's own equilibrium stress, computed there when the Newton iteration
// had not (the reference reads a stale local); the equilibrium iterations
// stop after WM_MAXIT iterations.  No multiply-add contraction: the
// arithmetic is the reference's operation by operation.
#define IWM_SKIP 5

__device__ double iwm_equil(double u_ave, double deltap, double re) {
#pragma clang fp contract(off)
  const double KAPA = 0.4, B = 5.0;
  double ut1 = 1.0, ut2 = 0.000001, f1, f2, f3, lam = 1.0;
  int it = 0;
  while (fabs(ut2 - ut1) > 1.0E-6 && it < WM_MAXIT) {
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

// iwm_slv_smooth (:3881-3926)
__device__ void iwm_slv(double lhsx, double lhsy, double Ux, double Uy, double Dz, double utx, double uty,
                        double re, double* fx, double* fy) {
#pragma clang fp contract(off)
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
  *fx = 0.5 * utx * (deli * deli) / delv +
        us * Dz * (0.5 * Ax * (1.0 - (deli * deli) / (Dz * Dz)) + Cx * (1.0 - deli / Dz) -
                   1.0 / vonk * Ux / Vel * (1.0 - deli / Dz + deli / Dz * lg)) +
        lhsx;
  *fy = 0.5 * uty * (deli * deli) / delv +
        us * Dz * (0.5 * Ay * (1.0 - (deli * deli) / (Dz * Dz)) + Cy * (1.0 - deli / Dz) -
                   1.0 / vonk * Uy / Vel * (1.0 - deli / Dz + deli / Dz * lg)) +
        lhsy;
}

__device__ __forceinline__ double iwm_dzp(const Ctx& c) { return c.m.dy[1] + c.m.dy[2] + c.m.dy[3] / 2.0; }

// NTIME and wm_dt (:3528-3533); at NTIME = 1 the mean third-cell velocity
// and its friction velocity (iwm_init_smooth, :3598-3628); at compute steps
// the mean first-cell pressure (:3786-3797).  One block.
__global__ __launch_bounds__(C3D_NT) void k_iwm_begin(Ctx c, int ntime) {
#pragma clang fp contract(off)
  __shared__ double red[C3D_NT];
  __shared__ int snt;
  const Geo g = c.g;
  const Fields F = c.d;
  const int t = threadIdx.x;
  if (t == 0) {
    const int nt = ntime > 0 ? ntime : (int)F.sc[SC_NTIME] + 1;
    F.sc[SC_NTIME] = (double)nt;
    F.sc[SC_WMDT] = (nt % IWM_SKIP == 1) ? c.s.dt : F.sc[SC_WMDT] + c.s.dt;
    snt = nt;
  }
  __syncthreads();
  const int nt = snt;
  const int n = g.nx * g.nz;
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 0 && nt != 1) continue;
    if (pass == 1 && nt % IWM_SKIP != 0) continue;
    double a = 0.0;
    for (int q = t; q < n; q += C3D_NT) {
      const int i = q % g.nx, k = q / g.nx;
      a += pass == 0 ? F.u[gidx(g, i, 3, k)] : F.p[gidx(g, i, 1, k)];
    }
    red[t] = a;
    __syncthreads();
    for (int h = C3D_NT / 2; h > 0; h >>= 1) {
      if (t < h) red[t] += red[t + h];
      __syncthreads();
    }
    if (t == 0) {
      if (pass == 0) {
        const double uinit = red[0] / (double)n;
        F.sc[SC_UINIT] = uinit;
        F.sc[SC_USINIT] = iwm_equil(uinit, iwm_dzp(c), c.s.re);
      } else {
        F.sc[SC_PBAR] = red[0] / (double)g.nx / (double)g.nz;
      }
    }
    __syncthreads();
  }
}

// iwm_init_smooth (:3629-3710), at NTIME = 1
__global__ __launch_bounds__(C3D_NT) void k_iwm_init(Ctx c) {
#pragma clang fp contract(off)
  const Fields F = c.d;
  if ((int)F.sc[SC_NTIME] != 1) return;
  const unsigned n = (unsigned)(c.g.nx * c.g.nz);
  const unsigned t = blockIdx.x * C3D_NT + threadIdx.x;
  if (t >= n) return;
  const double vonk = 0.4, REINV = 1.0 / c.s.re, Dzp = iwm_dzp(c);
  const double uinit = F.sc[SC_UINIT], usinit = F.sc[SC_USINIT];
  double* W = F.iwm;
  for (int q = 0; q < IW_NPL; q++) W[q * n + t] = 0.0;
  W[IW_UTX * n + t] = usinit;
  W[IW_TAUWX * n + t] = usinit * usinit;
  W[IW_FLT * n + t] = uinit;
  W[IW_FLTM * n + t] = uinit;
  W[(IW_INTE + 0) * n + t] = uinit * Dzp;
  W[(IW_INTE + 1) * n + t] = uinit * uinit * Dzp;
  W[(IW_INTEM + 0) * n + t] = uinit * Dzp;
  W[(IW_INTEM + 1) * n + t] = uinit * uinit * Dzp;
  W[IW_LHS * n + t] = -uinit * Dzp;
  W[(IW_LHS + 1) * n + t] = -uinit * Dzp;
  W[IW_DUDZT * n + t] = usinit / vonk / Dzp;
  W[IW_FLTUS * n + t] = usinit;
  W[IW_TR * n + t] = (F.sc[SC_CFLMAX] * c.s.alx / c.g.nx / uinit) / (Dzp / vonk / usinit);
  W[IW_DZ * n + t] = Dzp;
  W[IW_DELI * n + t] = fmin(11.0 * REINV / usinit, Dzp);
  W[IW_DELV * n + t] = REINV / usinit;
}

// the temporal filter of iwm_calc_lhs_smooth (:3766-3817)
__global__ __launch_bounds__(C3D_NT) void k_iwm_filter(Ctx c) {
#pragma clang fp contract(off)
  const Fields F = c.d;
  if ((int)F.sc[SC_NTIME] % IWM_SKIP != 0) return;
  const Geo g = c.g;
  const unsigned n = (unsigned)(g.nx * g.nz);
  const unsigned t = blockIdx.x * C3D_NT + threadIdx.x;
  if (t >= n) return;
  const int i = t % g.nx, k = t / g.nx;
  double* W = F.iwm;
  const double tr = W[IW_TR * n + t];
  const double fu = W[IW_FLT * n + t], fw = W[(IW_FLT + 1) * n + t];
  W[IW_FLTM * n + t] = fu;
  W[(IW_FLTM + 1) * n + t] = fw;
  W[IW_FLT * n + t] = fu * (1.0 - tr) + F.u[gidx(g, i, 3, k)] * tr;
  W[(IW_FLT + 1) * n + t] = fw * (1.0 - tr) + F.w[gidx(g, i, 3, k)] * tr;
  W[IW_FLTP * n + t] = W[IW_FLTP * n + t] * (1.0 - tr) + (F.p[gidx(g, i, 1, k)] - F.sc[SC_PBAR]) * tr;
}

// the balance of iwm_calc_lhs_smooth (:3821-3875): unsteady, convective
// and pressure-gradient terms and the left-hand side
__global__ __launch_bounds__(C3D_NT) void k_iwm_lhs(Ctx c) {
#pragma clang fp contract(off)
  const Fields F = c.d;
  if ((int)F.sc[SC_NTIME] % IWM_SKIP != 0) return;
  const Geo g = c.g;
  const unsigned n = (unsigned)(g.nx * g.nz);
  const unsigned t = blockIdx.x * C3D_NT + threadIdx.x;
  if (t >= n) return;
  const int i = t % g.nx, k = t / g.nx;
  const unsigned ip = k * g.nx + (i + 1 < g.nx ? i + 1 : 0), im = k * g.nx + (i > 0 ? i - 1 : g.nx - 1);
  const unsigned kp = (k + 1 < g.nz ? k + 1 : 0) * g.nx + i, km = (k > 0 ? k - 1 : g.nz - 1) * g.nx + i;
  const double DX1 = c.s.dx1, DX3 = c.s.dx3, wm_dt = F.sc[SC_WMDT];
  const double presg = F.sc[SC_PRESG], presg3 = F.sc[SC_PRESG3];
  double* W = F.iwm;
  const double* I0 = W + IW_INTE * n;  // integrals: u, uu, w, ww, uw
  const double Dz = W[IW_DZ * n + t];
  W[IW_UNSDY * n + t] = (I0[t] - W[IW_INTEM * n + t]) / wm_dt;
  W[(IW_UNSDY + 1) * n + t] = (I0[2 * n + t] - W[(IW_INTEM + 2) * n + t]) / wm_dt;
  const double uux = (I0[n + ip] - I0[n + im]) * DX1 / 2.0;
  const double uvy = (I0[4 * n + kp] - I0[4 * n + km]) * DX3 / 2.0;
  const double uvx = (I0[4 * n + ip] - I0[4 * n + im]) * DX1 / 2.0;
  const double vvy = (I0[3 * n + kp] - I0[3 * n + km]) * DX3 / 2.0;
  const double ux = (I0[ip] - I0[im]) * DX1 / 2.0;
  const double vy = (I0[2 * n + kp] - I0[2 * n + km]) * DX3 / 2.0;
  const double cx = uux + uvy - W[IW_FLTM * n + t] * (ux + vy);
  const double cy = uvx + vvy - W[(IW_FLTM + 1) * n + t] * (ux + vy);
  const double* FP = W + IW_FLTP * n;
  const double px = (FP[ip] - FP[im]) * DX1 / 2.0 * Dz + presg * Dz;
  const double py = (FP[kp] - FP[km]) * DX3 / 2.0 * Dz + presg3 * Dz;
  W[IW_CONV * n + t] = cx;
  W[(IW_CONV + 1) * n + t] = cy;
  W[IW_PRSG * n + t] = px;
  W[(IW_PRSG + 1) * n + t] = py;
  W[IW_LHS * n + t] = -I0[t] + wm_dt * (cx + px - W[IW_DIFF * n + t]);
  W[(IW_LHS + 1) * n + t] = -I0[2 * n + t] + wm_dt * (cy + py - W[(IW_DIFF + 1) * n + t]);
}

// iwm_calc_wallstress_smooth (:3930-4279) at one point per thread
__global__ __launch_bounds__(C3D_NT) void k_iwm_solve(Ctx c) {
#pragma clang fp contract(off)
  const Fields F = c.d;
  if ((int)F.sc[SC_NTIME] % IWM_SKIP != 0) return;
  const unsigned n = (unsigned)(c.g.nx * c.g.nz);
  const unsigned t = blockIdx.x * C3D_NT + threadIdx.x;
  if (t >= n) return;
  double* W = F.iwm;
  const double vonk = 0.4, re = c.s.re, reinv = 1.0 / re;
  const double deltap = iwm_dzp(c);
  const int MaxIter = 1500;
  const double wm_tol = 0.000001, wm_eps = 0.000000001;
  const double lx = W[IW_LHS * n + t], ly = W[(IW_LHS + 1) * n + t];
  const double Ux = W[IW_FLT * n + t], Uy = W[(IW_FLT + 1) * n + t], Dz = W[IW_DZ * n + t];
  double utx = W[IW_UTX * n + t], uty = W[IW_UTY * n + t];
  double fx, fy, fxp, fyp;
  iwm_slv(lx, ly, Ux, Uy, Dz, utx, uty, re, &fx, &fy);
  int iter = 0;
  bool equil = false, have_eq = false;
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
      equil = true;
      break;
    }
  }
  if (utx - 1.0 == utx || uty - 1.0 == uty) equil = true;
  double equilutx = 0.0, equiluty = 0.0;
  if (equil) {
    const double U_ave = sqrt(Ux * Ux + Uy * Uy);
    const double ut = iwm_equil(U_ave, deltap, re);
    const double TAUW1 = ut * ut;
    equilutx = TAUW1 * Ux / U_ave;
    equiluty = TAUW1 * Uy / U_ave;
    utx = sqrt(fabs(equilutx));
    uty = sqrt(fabs(equiluty));
    have_eq = true;
  }
  // the Newton (or equilibrium) friction velocities: everything below uses
  // these, also where the correction branch stores other values (:4110-4147)
  const double putx = utx, puty = uty;
  const double utaup = pow(pow(putx, 4.0) + pow(puty, 4.0), 0.25);
  const double delv = reinv / utaup;
  const double deli = fmin(11.0 * reinv / utaup, Dz);
  const double Vel = sqrt(Ux * Ux + Uy * Uy);
  const double lg = log(deli / Dz);
  double Ax = 0.0, Ay = 0.0, Cx = 0.0, Cy = 0.0;
  if (!equil) {
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
    }
    equil = true;
    utx = sqrt(fabs(equilutx));
    uty = sqrt(fabs(equiluty));
    Ax = Ay = Cx = Cy = 0.0;
  }
  W[IW_UTX * n + t] = utx;
  W[IW_UTY * n + t] = uty;
  W[IW_AX * n + t] = Ax;
  W[IW_AY * n + t] = Ay;
  W[IW_CX * n + t] = Cx;
  W[IW_CY * n + t] = Cy;
  for (int q = 0; q < 5; q++) W[(IW_INTEM + q) * n + t] = W[(IW_INTE + q) * n + t];
  const double d2 = deli * deli, d3 = deli * deli * deli, z2 = Dz * Dz, z3 = Dz * Dz * Dz;
  const double dv2 = delv * delv, us2 = utaup * utaup, V2 = Vel * Vel, k2 = vonk * vonk;
  W[IW_INTE * n + t] = 0.5 * putx * d2 / delv +
                       utaup * Dz * (0.5 * Ax * (1.0 - d2 / z2) + Cx * (1.0 - deli / Dz) -
                                     1.0 / vonk * Ux / Vel * (1.0 - deli / Dz + deli / Dz * lg));
  W[(IW_INTE + 2) * n + t] = 0.5 * puty * d2 / delv +
                             utaup * Dz * (0.5 * Ay * (1.0 - d2 / z2) + Cy * (1.0 - deli / Dz) -
                                           1.0 / vonk * Uy / Vel * (1.0 - deli / Dz + deli / Dz * lg));
  W[(IW_INTE + 4) * n + t] =
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
    W[(IW_INTE + 1) * n + t] =
        1.0 / 3.0 * (putx * putx) * d3 / dv2 +
        us2 * Dz *
            (-Ax * Ux / Vel / vonk * d2 / z2 * lg + Ax * (Cx - 1.0 / 2.0 / vonk * Ux / Vel) * (1.0 - d2 / z2) +
             Ax * Ax / 3.0 * (1.0 - d3 / z3) + tx * tx - deli / Dz * (sx * sx) +
             1.0 / k2 * (Ux * Ux) / V2 * (1.0 - deli / Dz));
    W[(IW_INTE + 3) * n + t] =
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
  W[IW_DUDZT * n + t] = tx;
  W[(IW_DUDZT + 1) * n + t] = ty;
  W[IW_DUDZB * n + t] = bx;
  W[(IW_DUDZB + 1) * n + t] = by;
  const double dVelzT = fabs(Ux / Vel * tx + Uy / Vel * ty);
  if (deli < Dz) {
    W[IW_DIFF * n + t] = (reinv + (vonk * Dz) * (vonk * Dz) * dVelzT) * tx - reinv * bx;
    W[(IW_DIFF + 1) * n + t] = (reinv + (vonk * Dz) * (vonk * Dz) * dVelzT) * ty - reinv * by;
  } else {
    W[IW_DIFF * n + t] = 0.0;
    W[(IW_DIFF + 1) * n + t] = 0.0;
  }
  double twx, twy;
  if (equil) {
    twx = equilutx;
    twy = equiluty;
  } else if (deli < Dz) {
    twx = putx * putx;
    twy = puty * puty;
  } else {
    twx = reinv * Ux / Dz;
    twy = reinv * Uy / Dz;
  }
  W[IW_TAUWX * n + t] = twx;
  W[IW_TAUWY * n + t] = twy;
  const double tR = W[IW_TR * n + t];
  const double fus = W[IW_FLTUS * n + t] * (1.0 - tR) + utaup * tR;
  W[IW_FLTUS * n + t] = fus;
  double ntr = F.sc[SC_WMDT] / (Dz / fus / vonk);
  if (ntr > 1.0) ntr = 1.0;
  W[IW_TR * n + t] = ntr;
}

// iwm_wallstress_smooth's imposed stresses (:3541-3553) and the block sums
// of |tau| for WSM(1) (:3463-3476); k_wallmodel_fin completes WSM(1)
__global__ __launch_bounds__(C3D_NT) void k_iwm_apply(Ctx c) {
#pragma clang fp contract(off)
  __shared__ double red[C3D_NT];
  const Fields F = c.d;
  const unsigned n = (unsigned)(c.g.nx * c.g.nz);
  double acc = 0.0;
  for (unsigned t = blockIdx.x * C3D_NT + threadIdx.x; t < n; t += gridDim.x * C3D_NT) {
    const double a = F.iwm[IW_TAUWX * n + t], b = F.iwm[IW_TAUWY * n + t];
    F.tauw[t] = a;
    F.taub[t] = b;
    acc += sqrt(a * a + b * b);
  }
  const int t = threadIdx.x;
  red[t] = acc;
  __syncthreads();
  for (int h = C3D_NT / 2; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  if (t == 0) F.red[blockIdx.x * 4] = red[0];
}

// WALLMODEL_WERNER (compact3d.f:4815-4958), WMODEL = 4: the Werner-Wengle
// power law at both walls from the first cell's centre velocity
// (|u| <= AB3 sqrt|WSM(1)|/11.81: the viscous branch 2|u|/DELTAP/RE), WSM(1)
// being the previous call's.  k_werner: the cell-centre stresses TAUWE of
// both walls (:4840-4909); the I neighbour is IPA(K), as the reference has
// it (:4849).  k_werner_avg: the face stresses TAUW, TAUB of both walls
// (neighbours IMA(K) and KMA(K), :4918-4931) and the block sums of
// TAUWE(.,.,1) for WSM(1) (:4934-4951).
__global__ __launch_bounds__(C3D_NT) void k_werner(Ctx c) {
  const Geo g = c.g;
  const Metrics M = c.m;
  const Fields F = c.d;
  const long long n2d = (long long)g.nx * g.nz;
  const double RE = c.s.re;
  const double B = 0.1428571429;
  const double AB1 = 7.202125274, AB3 = 69.74057973, AB4 = 0.1376936317, AB5 = 1.75;
  const double temp = sqrt(fabs(F.sc[SC_WSM1])) / 11.81;
  const double vcase = AB3 * temp;
  for (long long t = blockIdx.x * (long long)C3D_NT + threadIdx.x; t < 2 * n2d; t += (long long)gridDim.x * C3D_NT) {
    const int wall = (int)(t / n2d);
    const long long e = t - wall * n2d;
    const int i = (int)(e % g.nx), k = (int)(e / g.nx);
    const int j = wall == 0 ? 1 : g.n2m;
    const double deltap = M.dy[j];
    const int ip = (k + 1 < g.nx) ? k + 1 : 0;  // IPA(K), K = k+1
    const int kp = wrapp(k, g.nz);
    const double u1 = 0.5 * (F.u[gidx(g, i, j, k)] + F.u[gidx(g, ip, j, k)]);
    const double u2 = 0.5 * (F.w[gidx(g, i, j, k)] + F.w[gidx(g, i, j, kp)]);
    const double velocp = sqrt(u1 * u1 + u2 * u2);
    double tauw1 = 0.0;
    if (velocp <= vcase) {
      tauw1 = 2. * velocp / deltap / RE;
    } else if (velocp > vcase) {
      tauw1 = AB1 * pow(1.0 / deltap / RE, 1. + B);
      tauw1 = tauw1 + AB4 * pow(1.0 / deltap / RE, B) * velocp;
      tauw1 = pow(tauw1, AB5);
    }
    F.tauwe[(2 * wall) * n2d + e] = tauw1 * (u1 / velocp);
    F.tauwe[(2 * wall + 1) * n2d + e] = tauw1 * (u2 / velocp);
  }
}

__global__ __launch_bounds__(C3D_NT) void k_werner_avg(Ctx c) {
  __shared__ double red[C3D_NT];
  const Geo g = c.g;
  const Fields F = c.d;
  const long long n2d = (long long)g.nx * g.nz;
  const double* T1 = F.tauwe;
  const double* T2 = F.tauwe + n2d;
  const double* T3 = F.tauwe + 2 * n2d;
  const double* T4 = F.tauwe + 3 * n2d;
  double acc = 0.0;
  for (long long e = blockIdx.x * (long long)C3D_NT + threadIdx.x; e < n2d; e += (long long)gridDim.x * C3D_NT) {
    const int i = (int)(e % g.nx), k = (int)(e / g.nx);
    const long long im = (long long)k * g.nx + (k > 0 ? k - 1 : g.nx - 1);  // (IMA(K), K)
    const long long km = (long long)wrapm(k, g.nz) * g.nx + i;              // (I, KMA(K))
    F.tauw[e] = 0.5 * (T1[e] + T1[im]);
    F.taub[e] = 0.5 * (T2[e] + T2[km]);
    F.tauw[n2d + e] = 0.5 * (T3[e] + T3[im]);
    F.taub[n2d + e] = 0.5 * (T4[e] + T4[km]);
    acc += T1[e];
  }
  const int t = threadIdx.x;
  red[t] = acc;
  __syncthreads();
  for (int h = C3D_NT / 2; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  if (t == 0) F.red[blockIdx.x * 4] = red[0];
}

// WALLSS (compact3d.f:11077-11143), including its use of WST(.,.,1) in WSM(2)
// and WSM(3) (:11130-11132)
template <bool CB_>
__global__ __launch_bounds__(C3D_NT) void k_wallss(CtxArg<CB_> ca_) {
  const Ctx c = ctx_get(ca_);
  __shared__ double red[3][C3D_NT];
  const Geo g = c.g;
  const Metrics M = c.m;
  const Fields F = c.d;
  const long long n = (long long)g.nx * g.nz;
  const double aly = c.s.aly;
  double a1 = 0, a2 = 0, a3 = 0;
  for (long long t = blockIdx.x * (long long)C3D_NT + threadIdx.x; t < n; t += (long long)gridDim.x * C3D_NT) {
    const int i = t % g.nx, k = t / g.nx;
    const int ip = wrapp(i, g.nx), kp = wrapp(k, g.nz);
    double uhc = 0.5 * (F.u[gidx(g, i, 1, k)] + F.u[gidx(g, ip, 1, k)]);
    double uhw = 0.5 * (F.u[gidx(g, i, 0, k)] + F.u[gidx(g, ip, 0, k)]);
    double vhc = F.v[gidx(g, i, 2, k)], vhw = F.v[gidx(g, i, 1, k)];
    double whc = 0.5 * (F.w[gidx(g, i, 1, k)] + F.w[gidx(g, i, 1, kp)]);
    double whw = 0.5 * (F.w[gidx(g, i, 0, k)] + F.w[gidx(g, i, 0, kp)]);
    const double wsb1 = (uhc - uhw) / (M.dy[1] * 0.5);
    const double wsb2 = (vhc - vhw) / M.dy[1];
    const double wsb3 = (whc - whw) / (M.dy[1] * 0.5);
    uhc = 0.5 * (F.u[gidx(g, i, g.n2m, k)] + F.u[gidx(g, ip, g.n2m, k)]);
    uhw = 0.5 * (F.u[gidx(g, i, g.n2, k)] + F.u[gidx(g, ip, g.n2, k)]);
    const double wst1 = (uhc - uhw) / (M.dy[g.n2m] * 0.5);
    a1 = a1 + wsb1 + wst1 * (aly - 1.0);
    a2 = a2 + wsb2 + wst1 * (aly - 1.0);
    a3 = a3 + wsb3 + wst1 * (aly - 1.0);
  }
  const int t = threadIdx.x;
  red[0][t] = a1;
  red[1][t] = a2;
  red[2][t] = a3;
  __syncthreads();
  for (int h = C3D_NT / 2; h > 0; h >>= 1) {
    if (t < h)
      for (int q = 0; q < 3; q++) red[q][t] += red[q][t + h];
    __syncthreads();
  }
  if (t < 3) F.red[blockIdx.x * 4 + t] = red[t][0];
}

template <bool CB_>
__global__ __launch_bounds__(C3D_NT) void k_wallss_fin(CtxArg<CB_> ca_, int nblk) {
  const Ctx c = ctx_get(ca_);
  __shared__ double red[3][C3D_NT];
  const Fields F = c.d;
  const int t = threadIdx.x;
  double a[3] = {0, 0, 0};
  for (int b = t; b < nblk; b += C3D_NT)
    for (int q = 0; q < 3; q++) a[q] += F.red[b * 4 + q];
  for (int q = 0; q < 3; q++) red[q][t] = a[q];
  __syncthreads();
  for (int h = C3D_NT / 2; h > 0; h >>= 1) {
    if (t < h)
      for (int q = 0; q < 3; q++) red[q][t] += red[q][t + h];
    __syncthreads();
  }
  if (t == 0) {
    const double nxz = 1.0 / (double)((long long)c.g.nx * c.g.nz);
    for (int q = 0; q < 3; q++) F.sc[SC_WSM1 + q] = red[q][0] * nxz / c.s.re / c.s.aly;
  }
}

// --------------------------------------------------- SGS line-set terms
// Intermediates shared by the 36 SGS line solves of RHS1-3:
//   nux = Ix(SGSVIS)   (centre -> x-face)    SGSVIS1/12/13, SGSVISV1/V12
//   nuz = Iz(SGSVIS)   (centre -> z-face)    SGSVIS3/V3/W2/W3/31/32
//   nue = Iz(Ix(SGSVIS)) (x-z edge)          the (1,6,1) pair of :6512-6517,
//         :6747-6752, :7819-7823, :8005-8009 (Ix Iz and Iz Ix commute)
//   dvdx = D1x(v), dwdx = D1x(w) (centre -> x-face)   UU12, VV12, WW13, X11
//   dudz = D1z(u), dvdz = D1z(v) (centre -> z-face)   X33 of :6519-6524, UU31,
//                                                     VV23, VV32
// These kernels (and the x-line part inside k_rhs_x) run on the register
// line engine (c3d_lines.h) in the
// compute type T of the step's line kernels: FP64, or FP32 in mixed mode,
// where they read the FP32 shadows of the state and of SGSVIS and keep the
// intermediates in FP32 (HH is FP32 there too).  Each HH update is a
// separate rounded store of a term, as the reference's sequence of updates.

// x-lines (j = 1..N2M): nux, dvdx, dwdx, and D(SGS*(DU/DX))/DX of RHS1
// (:6445-6466) run inside the RHS x-pass (k_rhs_x, c3d_step.hip), which
// stages u, v and w on the same lines.

// z-lines (j = 1..N2M): nuz, nue, dudz, dvdz, and
//   D(SGS*(DW/DZ))/DZ of RHS3 (:7919-7940): HH3 -= 2 D1z(SGSVIS * D1z(w))
//   D(SGS*(DU/DZ))/DZ of RHS1 (:6491-6540): HH1 -= D1z(dudz * nue)
//   D(SGS*(DW/DX))/DZ of RHS1 (:6713-6764): HH1 -= D1z(dwdx * nue)
// and, at J >= 2, the RHS2 terms D(SGS*(DW/DY))/DZ and D(SGS*(DV/DZ))/DZ
template <typename T, int SEG, bool CB_>
__global__ __launch_bounds__(512) void k_sgs_z(CtxArg<CB_> ca_) {
  const Ctx c = ctx_get(ca_);
  ZSETUP;
  const Fields F = c.d;
  const PSolve s6 = PZ.s6, s22 = PZ.s22;
  const T dx3 = (T)c.s.dx3;
  T* nuz = (T*)F.nuz;
  T* nue = (T*)F.nue;
  T* dudz = (T*)F.dudz;
  T* dvdz = (T*)F.dvdz;
  T* hh0 = (T*)F.hh[0];
  T* hh2 = (T*)F.hh[2];
  T a[SEG], d[SEG], sv[SEG], e[SEG], dz[SEG], lf, rt;
  // nuz = Iz(SGSVIS)
  zload((const T*)F.ssgs, b, pl, p0, bm, bp, sv, lf, rt);
  SQ d[q] = (T)4. * (prv(sv, q, lf) + sv[q]);
  zload((const T*)F.nux, b, pl, p0, bm, bp, a, lf, rt);
  psolve<T, SEG>(d, G, L, l, gs, s6, bf.xf, bf.xb);
  if (ok) SQ st_nt(&nuz[ZI(q)], (T)(d[q]));
  // nue = Iz(nux)
  SQ e[q] = (T)4. * (a[q] + prv(a, q, lf));
  zload((const T*)F.su, b, pl, p0, bm, bp, a, lf, rt);
  psolve<T, SEG>(e, G, L, l, gs, s6, bf.xf, bf.xb);
  if (ok) SQ st_nt(&nue[ZI(q)], (T)(e[q]));
  // dudz = D1z(u)
  SQ dz[q] = (T)24. * dx3 * (a[q] - prv(a, q, lf));
  zload((const T*)F.sv, b, pl, p0, bm, bp, a, lf, rt);
  psolve<T, SEG>(dz, G, L, l, gs, s22, bf.xf, bf.xb);
  if (ok) SQ st_nt(&dudz[ZI(q)], (T)(dz[q]));
  // dvdz = D1z(v)
  SQ d[q] = (T)24. * dx3 * (a[q] - prv(a, q, lf));
  zload((const T*)F.sw, b, pl, p0, bm, bp, a, lf, rt);
  psolve<T, SEG>(d, G, L, l, gs, s22, bf.xf, bf.xb);
  if (ok) SQ st_nt(&dvdz[ZI(q)], (T)(d[q]));
  // HH3 -= 2 D1z(SGSVIS * D1z(w))
  SQ d[q] = (T)24. * dx3 * (nxt(a, q, rt) - a[q]);
  psolve<T, SEG>(d, G, L, l, gs, s22, bf.xf, bf.xb);
  SQ d[q] = d[q] * sv[q];
  lf = xch_left(d[SEG - 1], G, L, l, gs, bf.xl);
  SQR d[q] = (T)24. * dx3 * (d[q] - prv(d, q, lf));
  SQ a[q] = hh2[ZI(q)];
  psolve<T, SEG>(d, G, L, l, gs, s22, bf.xf, bf.xb);
  if (ok) SQ st_nt(&hh2[ZI(q)], (T)(a[q] - (T)2. * d[q]));
  // HH1 -= D1z(dudz * nue)
  SQ d[q] = dz[q] * e[q];
  rt = xch_right(d[0], G, L, l, gs, bf.xr);
  SQ d[q] = (T)24. * dx3 * (nxt(d, q, rt) - d[q]);
  SQ a[q] = hh0[ZI(q)];
  psolve<T, SEG>(d, G, L, l, gs, s22, bf.xf, bf.xb);
  SQ a[q] = a[q] - d[q];
  // HH1 -= D1z(dwdx * nue)
  SQ d[q] = ((const T*)F.dwdx)[ZI(q)] * e[q];
  rt = xch_right(d[0], G, L, l, gs, bf.xr);
  SQ d[q] = (T)24. * dx3 * (nxt(d, q, rt) - d[q]);
  psolve<T, SEG>(d, G, L, l, gs, s22, bf.xf, bf.xb);
  if (ok) SQ st_nt(&hh0[ZI(q)], (T)(a[q] - d[q]));
  if (j < 2) return;  // uniform per block
  // J >= 2: the two RHS2 terms that need nuz at J-1 and J,
  //   D(SGS*(DW/DY))/DZ (:7177-7226), D(SGS*(DV/DZ))/DZ (:7245-7304).
  // nuz(J-1) = Iz(SGSVIS(J-1)) is formed here by the operations the tile at
  // J-1 uses for its nuz (the same values, bit for bit); nuz(J), dvdz(J)
  // and w come back from L2 (this block wrote or read them above).
  {
    const Metrics M = c.m;
    const T hinv = (T)(1. / M.h[j]), dyj = (T)(M.dy[j] / 2.), dyjm = (T)(M.dy[j - 1] / 2.);
    const T* sw = (const T*)F.sw;
    T* hh1 = (T*)F.hh[1];
    const unsigned nx = (unsigned)g.nx;
    T yn[SEG];
    zload((const T*)F.ssgs, b - nx, pl, p0, bm - nx, bp - nx, sv, lf, rt);
    SQ yn[q] = (T)4. * (prv(sv, q, lf) + sv[q]);
    psolve<T, SEG>(yn, G, L, l, gs, s6, bf.xf, bf.xb);
    SQ {
      yn[q] = hinv * (dyj * yn[q] + dyjm * nuz[ZI(q)]);
      d[q] = yn[q] * (hinv * (sw[ZI(q)] - sw[ZI(q) - nx]));
    }
    rt = xch_right(d[0], G, L, l, gs, bf.xr);
    SQ d[q] = (T)24. * dx3 * (nxt(d, q, rt) - d[q]);
    SQ a[q] = hh1[ZI(q)];
    psolve<T, SEG>(d, G, L, l, gs, s22, bf.xf, bf.xb);
    SQ a[q] = a[q] - d[q];
    SQ d[q] = yn[q] * dvdz[ZI(q)];
    rt = xch_right(d[0], G, L, l, gs, bf.xr);
    SQ d[q] = (T)24. * dx3 * (nxt(d, q, rt) - d[q]);
    psolve<T, SEG>(d, G, L, l, gs, s22, bf.xf, bf.xb);
    if (ok) SQ st_nt(&hh1[ZI(q)], (T)(a[q] - d[q]));
  }
}

// x-lines: D(SGS*(DU/DY))/DX and D(SGS*(DV/DX))/DX of RHS2 (:7104-7158,
// :7323-7384, J >= 2), D(SGS*(DW/DX))/DX and D(SGS*(DU/DZ))/DX of RHS3
// (:7797-7847, :7972-8024), after the RHS x-pass 2's HH3 += D1x(UW)
// (:7734-7769, k_rhs_x2 without SGS), which joins here: HH3 is read and
// written once for both.
template <typename T, int SEG, bool CB_>
__global__ __launch_bounds__(256) void k_sgs_x2(CtxArg<CB_> ca_) {
  const Ctx c = ctx_get(ca_);
  XSETUP;
  const Metrics M = c.m;
  const Fields F = c.d;
  const PSolve s22 = c.o.px.s22;
  const T dx1 = (T)c.s.dx1;
  const bool j2 = ok && j >= 2;
  const T hinv = (T)(1. / M.h[j]);
  const T dyjm = (T)(M.dy[j - 1] / 2.), dyj = (T)(M.dy[j] / 2.);
  T d[SEG], yn[SEG], e[SEG], ux[SEG], pf[SEG], rt;
  // D1x(UW), face -> centre, as k_rhs_x2
  X.pre((const T*)F.uw, 0, pf);
  const T* B = X.put(pf);  // UW
  X.pre((const T*)F.nux, 0, pf);
  SQ ux[q] = (T)24. * dx1 * (-XA(B, q) + XA(B, q + 1));
  psolve<T, SEG>(ux, G, L, l, gs, s22, bf.xf, bf.xb);
  B = X.put(pf);  // nux(J)
  X.pre((const T*)F.nux, -1, pf);
  SQ yn[q] = dyjm * XA(B, q);
  B = X.put(pf);  // nux(J-1)
  X.pre((const T*)F.su, 0, pf);
  SQ yn[q] = j2 ? hinv * (yn[q] + dyj * XA(B, q)) : (T)0;
  B = X.put(pf);  // u(J)
  X.pre((const T*)F.su, -1, pf);
  SQ d[q] = XA(B, q);
  B = X.put(pf);  // u(J-1)
  X.pre((const T*)F.dvdx, 0, pf);
  SQ d[q] = j2 ? yn[q] * (hinv * (d[q] - XA(B, q))) : (T)0;
  rt = xch_right(d[0], G, L, l, gs, bf.xr);
  SQ d[q] = (T)24. * dx1 * (nxt(d, q, rt) - d[q]);
  psolve<T, SEG>(d, G, L, l, gs, s22, bf.xf, bf.xb);
  B = X.put(pf);  // dvdx
  X.pre((const T*)F.hh[1], 0, pf);
  SQ e[q] = yn[q] * XA(B, q);
  rt = xch_right(e[0], G, L, l, gs, bf.xr);
  SQ e[q] = (T)24. * dx1 * (nxt(e, q, rt) - e[q]);
  psolve<T, SEG>(e, G, L, l, gs, s22, bf.xf, bf.xb);
  B = X.put(pf);  // HH2
  X.pre((const T*)F.nue, 0, pf);
  SQ d[q] = (XA(B, q) - d[q]) - e[q];
  X.out((T*)F.hh[1], d, 2);
  B = X.put(pf);  // nue
  X.pre((const T*)F.dwdx, 0, pf);
  SQ yn[q] = XA(B, q);
  B = X.put(pf);  // dwdx
  X.pre((const T*)F.dudz, 0, pf);
  SQ d[q] = yn[q] * XA(B, q);
  rt = xch_right(d[0], G, L, l, gs, bf.xr);
  SQ d[q] = (T)24. * dx1 * (nxt(d, q, rt) - d[q]);
  psolve<T, SEG>(d, G, L, l, gs, s22, bf.xf, bf.xb);
  B = X.put(pf);  // dudz
  X.pre((const T*)F.hh[2], 0, pf);
  SQ e[q] = yn[q] * XA(B, q);
  rt = xch_right(e[0], G, L, l, gs, bf.xr);
  SQ e[q] = (T)24. * dx1 * (nxt(e, q, rt) - e[q]);
  psolve<T, SEG>(e, G, L, l, gs, s22, bf.xf, bf.xb);
  B = X.put(pf);  // HH3
  SQ d[q] = ((XA(B, q) + ux[q]) - d[q]) - e[q];
  X.out((T*)F.hh[2], d);
}

// ------------------------------------------------------- test filters
// FILTER (compact3d.f:5190-5228, width 2) and FILTER4 (:5481-5546, width 4)
// on a stack of x-z planes FILV(ld1, ld3, n): x pass then z pass, each output
// recomputing the x-filtered neighbours it needs (same expressions).
__global__ void k_filter(int nx, int nz, int ld1, int ld3, int np, int width, const double* in,
                         double* out) {
  const long long tot = (long long)nx * nz * np;
  long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= tot) return;
  const int i = t % nx, k = (t / nx) % nz, q = t / ((long long)nx * nz);
  const double* f = in + (long long)q * ld1 * ld3;
  auto F = [&](int ii, int kk) { return f[(long long)kk * ld1 + ii]; };
  auto wx = [&](int n, int a) { int r = (a + n) % n; return r; };
  auto tx = [&](int kk) {
    if (width == 2)
      return (F(wx(nx, i - 1), kk) + 4.0 * F(i, kk) + F(wx(nx, i + 1), kk)) / 6.0;
    return (F(wx(nx, i - 2), kk) + 4.0 * F(wx(nx, i - 1), kk) + 2.0 * F(i, kk) + 4.0 * F(wx(nx, i + 1), kk) +
            F(wx(nx, i + 2), kk)) / 12.0;
  };
  double r;
  if (width == 2)
    r = (tx(wx(nz, k - 1)) + 4.0 * tx(k) + tx(wx(nz, k + 1))) / 6.0;
  else
    r = (tx(wx(nz, k - 2)) + 4.0 * tx(wx(nz, k - 1)) + 2.0 * tx(k) + 4.0 * tx(wx(nz, k + 1)) + tx(wx(nz, k + 2))) /
        12.0;
  out[(long long)q * ld1 * ld3 + (long long)k * ld1 + i] = r;
}

// ================================================================ host
namespace c3dk {

static int pgrid(long long n) { return (int)((n + C3D_NT - 1) / C3D_NT); }
static int xgrid(const Ctx& c, const TileShape& t) { return (c.g.n2m * c.g.nz + t.nl - 1) / t.nl; }
static int zgrid(const Ctx& c, const TileShape& t, int jlo) {
  return ((c.g.nx + t.nl - 1) / t.nl) * (c.g.n2m - jlo + 1);
}
static size_t lds(const TileShape& t, int ntiles) { return sizeof(double) * (size_t)t.n * t.ls * ntiles; }

void sgscs(const Ctx& c) {
  KScope ks_(c, KID_SGS);
  const int ntx = (c.g.nx + SG_TI - 1) / SG_TI;
  // runs of up to 16 consecutive k (the window loads kpt + 2 planes for
  // kpt points), shorter when the grid would have fewer than ~2048 blocks
  int kpt = 16;
  while (kpt > 1 && (long long)ntx * ((c.g.nz + 4 * kpt - 1) / (4 * kpt)) * c.g.n2m < 2048) kpt /= 2;
  const dim3 gr(ntx * ((c.g.nz + 4 * kpt - 1) / (4 * kpt)), c.g.n2m);
  if (c.s.mixed)
    C3D_LAUNCH(k_sgscs, (float), gr, dim3(256), 0, c, kpt);
  else
    C3D_LAUNCH(k_sgscs, (double), gr, dim3(256), 0, c, kpt);
}

int sgsdy_tiles(const Geo& g) {
  return ((g.nx + DT_TILE - 1) / DT_TILE) * ((g.nz + DT_TILE - 1) / DT_TILE);
}

void sgsdy(const Ctx& c) {
  const int nt = sgsdy_tiles(c.g);
  KScope ks_(c, KID_SGS);
  hipLaunchKernelGGL(k_sgsdy_tile, dim3(nt * c.g.n2m), dim3(DT_TILE * DT_TILE), 0, c.st, c, nullptr);
  hipLaunchKernelGGL(k_sgsdy_coef, dim3(c.g.n2m), dim3(256), 0, c.st, c, nt);
  hipLaunchKernelGGL(k_sgsdy_clip, dim3(pgrid((long long)c.g.nx * c.g.n2m * c.g.nz)), dim3(C3D_NT), 0, c.st, c);
}

// the partials of the two filter levels sit at dyp and dyp + 5*N2M*tiles;
// the last stage's L:M, M:M partials reuse the start of dyp once BETA is set
void sgssd(const Ctx& c) {
  const int nt = sgsdy_tiles(c.g);
  double* p1 = c.d.dyp;
  double* p2 = c.d.dyp + 5 * (size_t)c.g.n2m * nt;
  KScope ks_(c, KID_SGS);
  hipLaunchKernelGGL(k_sgssd_tile<1>, dim3(nt * c.g.n2m), dim3(DT_TILE * DT_TILE), 0, c.st, c, p1);
  hipLaunchKernelGGL(k_sgssd_tile<2>, dim3(nt * c.g.n2m), dim3(DT_TILE * DT_TILE), 0, c.st, c, p2);
  hipLaunchKernelGGL(k_sgssd_coef, dim3(c.g.n2m), dim3(256), 0, c.st, c, nt, (const double*)p1, (const double*)p2);
  hipLaunchKernelGGL(k_sgsdy_tile, dim3(nt * c.g.n2m), dim3(DT_TILE * DT_TILE), 0, c.st, c,
                     (const double*)c.d.sdbeta);
  hipLaunchKernelGGL(k_sgsdy_coef, dim3(c.g.n2m), dim3(256), 0, c.st, c, nt);
  hipLaunchKernelGGL(k_sgsdy_clip, dim3(pgrid((long long)c.g.nx * c.g.n2m * c.g.nz)), dim3(C3D_NT), 0, c.st, c);
}

void wallmodel(const Ctx& c, bool both) {
  const int nb = std::min(1024, pgrid((long long)c.g.nx * c.g.nz));
  {
    KScope ks_(c, KID_WALL);
    C3D_LAUNCH0(k_wallmodel, dim3(nb), dim3(C3D_NT), 0, c, 0);
  }
  KScope ks_(c, KID_WALL);
  C3D_LAUNCH0(k_wallmodel_fin, dim3(1), dim3(C3D_NT), 0, c, nb, 0);
  if (both) C3D_LAUNCH0(k_wallmodel, dim3(nb), dim3(C3D_NT), 0, c, 1);
}

void wallmodel_iwmles(const Ctx& c, int ntime) {
  const long long n2d = (long long)c.g.nx * c.g.nz;
  const int nb = std::min(1024, pgrid(n2d));
  KScope ks_(c, KID_WALL);
  hipLaunchKernelGGL(k_iwm_begin, dim3(1), dim3(C3D_NT), 0, c.st, c, ntime);
  hipLaunchKernelGGL(k_iwm_init, dim3(pgrid(n2d)), dim3(C3D_NT), 0, c.st, c);
  hipLaunchKernelGGL(k_iwm_filter, dim3(pgrid(n2d)), dim3(C3D_NT), 0, c.st, c);
  hipLaunchKernelGGL(k_iwm_lhs, dim3(pgrid(n2d)), dim3(C3D_NT), 0, c.st, c);
  hipLaunchKernelGGL(k_iwm_solve, dim3(pgrid(n2d)), dim3(C3D_NT), 0, c.st, c);
  hipLaunchKernelGGL(k_iwm_apply, dim3(nb), dim3(C3D_NT), 0, c.st, c);
  C3D_LAUNCH0(k_wallmodel_fin, dim3(1), dim3(C3D_NT), 0, c, nb, 0);
}

void wallmodel_werner(const Ctx& c) {
  const long long n2d = (long long)c.g.nx * c.g.nz;
  const int nb = std::min(1024, pgrid(n2d));
  KScope ks_(c, KID_WALL);
  hipLaunchKernelGGL(k_werner, dim3(std::min(2048, pgrid(2 * n2d))), dim3(C3D_NT), 0, c.st, c);
  hipLaunchKernelGGL(k_werner_avg, dim3(nb), dim3(C3D_NT), 0, c.st, c);
  C3D_LAUNCH0(k_wallmodel_fin, dim3(1), dim3(C3D_NT), 0, c, nb, 1);
}

void wallss(const Ctx& c) {
  const int nb = std::min(1024, pgrid((long long)c.g.nx * c.g.nz));
  {
    KScope ks_(c, KID_WALLSS);
    C3D_LAUNCH0(k_wallss, dim3(nb), dim3(C3D_NT), 0, c);
  }
  KScope ks_(c, KID_WALLSS);
  C3D_LAUNCH0(k_wallss_fin, dim3(1), dim3(C3D_NT), 0, c, nb);
}

// TH: storage type of HH (FP32 in mixed mode); the SGS terms are formed in
// FP64 from the FP64 state and added to HH in FP64 arithmetic
// the SGS line kernels in the compute type T of the step's line kernels:
// x-lines on the x shape (4-byte fields staged in pairs), z-lines on the
// z shape of T (FP32 tiles: 32 columns, 16-point segments)
template <typename T>
static void sgs_terms_t(const Ctx& c) {
  const LShape &sx = c.o.lx, &sz = sizeof(T) == 4 ? c.o.lz32 : c.o.lz;
  const size_t lx = sizeof(T) * (2 * (size_t)sx.n * (sx.L + 1) + 4 * (size_t)sx.G * sx.L);
  const size_t lz = sizeof(T) * 4 * (size_t)sz.G * sz.L;
  const int nxb = (c.g.n2m * c.g.nz + sx.L - 1) / sx.L;
  const int nit = (c.g.nx + sz.L - 1) / sz.L;
  const dim3 tx(sx.L * sx.G), tz(sz.L * sz.G);
#define SEGX_(seg, ...)        \
  if ((seg) == 8) {            \
    constexpr int S_ = 8;      \
    __VA_ARGS__;               \
  } else {                     \
    constexpr int S_ = 2;      \
    __VA_ARGS__;               \
  }
#define SEGZ_(seg, ...)                       \
  if (sizeof(T) == 4 && (seg) == 16) {        \
    constexpr int S_ = sizeof(T) == 4 ? 16 : 8; \
    __VA_ARGS__;                              \
  } else if ((seg) == 8) {                    \
    constexpr int S_ = 8;                     \
    __VA_ARGS__;                              \
  } else {                                    \
    constexpr int S_ = 2;                     \
    __VA_ARGS__;                              \
  }
  {
    KScope ks_(c, KID_SGS_Z);
    SEGZ_(sz.seg, C3D_LAUNCH(k_sgs_z, (T, S_), dim3(nit * c.g.n2m), tz, lz, c));
  }
  KScope ks_(c, KID_SGS_X2);
  SEGX_(sx.seg, C3D_LAUNCH(k_sgs_x2, (T, S_), dim3(nxb), tx, lx, c));
#undef SEGX_
#undef SEGZ_
}
void sgs_terms(const Ctx& c) {
  if (c.s.mixed)
    sgs_terms_t<float>(c);
  else
    sgs_terms_t<double>(c);
}

void filter(const Ctx& c, int width, const double* in, double* out, int np, int ld1, int ld3) {
  KScope ks_(c, KID_FILTER);
  const long long tot = (long long)c.g.nx * c.g.nz * np;
  hipLaunchKernelGGL(k_filter, dim3(pgrid(tot)), dim3(C3D_NT), 0, c.st, c.g.nx, c.g.nz, ld1, ld3, np, width, in,
                     out);
}

}  // namespace c3dk
