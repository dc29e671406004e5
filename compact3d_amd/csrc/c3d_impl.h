// c3d_impl.h -- host/device shared structures of libc3d.
#pragma once
#include <hip/hip_runtime.h>
#include "c3d_dev.h"

#define YTAB_W 16
enum { YT_BET = 11, YT_M = 12, YT_G = 13 };

// device scalar slots (double unless noted)
enum {
  SC_PRESG = 0, SC_PRESG3, SC_DMPRESG, SC_DMPRESG3, SC_FLOW1, SC_FLOW3,
  SC_WSM1, SC_WSM2, SC_WSM3, SC_TAOW,
  SC_DIVMAX,  // stored as uint64 bit pattern (atomicMax on non-negative doubles)
  SC_CFLM,    // same
  SC_CHK1, SC_CHK3,
  // WMODEL = 2: NTIME of the last wall-model call, wm_dt, CFLMAX, the
  // initial mean velocity and friction velocity, the mean first-cell pressure
  SC_NTIME, SC_WMDT, SC_CFLMAX, SC_UINIT, SC_USINIT, SC_PBAR,
  SC_COUNT = 24
};

// WMODEL = 2 (the integral wall model): its state planes, NX*NZ doubles
// each ([(K-1)*NX + I-1]), in the order of the reference's COMMON /IWMLES/
// (compact3d.f:3399-3408); the five integrals are those of u, uu, w, ww, uw
enum {
  IW_UTX = 0, IW_UTY, IW_TAUWX, IW_TAUWY, IW_FLT, IW_FLTM = 6, IW_FLTP = 8, IW_INTE = 9, IW_INTEM = 14,
  IW_UNSDY = 19, IW_CONV = 21, IW_PRSG = 23, IW_DIFF = 25, IW_LHS = 27, IW_DUDZT = 29, IW_DUDZB = 31,
  IW_FLTUS = 33, IW_TR, IW_DZ, IW_Z0, IW_AX, IW_AY, IW_DELV, IW_DELI, IW_CX, IW_CY, IW_NPL
};

// metric arrays on the device, indexed by the reference's J (0..N2+1)
struct Metrics {
  const double *y, *dy, *h, *hp, *hc, *hm, *dyp, *dyc, *dym, *fjpa, *fjmu, *fjmv;
  const double *pmj, *pcj, *ppj;
  // J-only factors of the y terms, formed once on the host exactly as the
  // reference's expressions evaluate them: FJPA(J)*1/H(J+1), FJMU(J)*1/H(J),
  // DY(J)/2, 1/DY(J), 1/H(J)
  const double *cu2, *cu1, *dyh, *rdy, *rh;
  // per-component, per-J coefficients of the fused RHS-y / ADI-y sweep,
  // YTAB_W doubles per J (index comp*(N2+2)*YTAB_W + J*YTAB_W + slot):
  //   u, w: HP HC HM CU2 CU1 DYH(J-1) DYH(J) DYH(J+1) RDY 1-FJPA 1-FJMU,
  //         slot 14 1/H(J), 15 1/H(J+1)
  //   v:    DYP DYC DYM H RH 1/DY(J) 1/DY(J-1)
  //   all:  slot 11 BET, 12 -A*BET (0 at the first row), 13 -GAM (0 at N2M)
  const double* ytab;
  const float* ytab32;  // ytab rounded to FP32, for the FP32 sweep
  const double *ak1;  // AK1(I), I = 1..N1M  (index I)
  const double *ak3;  // AK3(K), K = 1..N3M/2+1
  // TDMAI chains of the ADI y-sweeps (component c = 0,1,2): A(J), BET(J), GAM(J)
  const double *ya[3], *ybet[3], *ygam[3];
};

struct Scal {
  double re, rinv, dt, dx1, dx1q, dx3, dx3q, aly, vol, alx;
  int ifwall;
  int mixed;  // C3D_MIXED: FP32 intermediates (see Fields)
  int files;  // FILES != 0
  double wmodel;
};

// the four periodic solves of one direction, for that direction's segment
struct OpSet {
  PSolve s6, s10, s22, adi;
};

struct Ops {
  LShape lx, lz;  // register line engine tiles (x-lines staged, z-lines direct)
  OpSet px, pz;
  // FP32 z-line tiles: 32 columns (128-byte rows), 16-point segments
  LShape lz32;
  OpSet pz32;
  YShape ys;      // ADI y-sweep tiles (16 columns of FP64: 128-byte rows)
  YShape ys32;    // the same for FP32 (32 columns)
  YShape yp;      // Poisson y-solve tiles
  LineSolve s6, s10, s22, adix, adiz;
  TileShape tx, tz;    // x-line and z-line tile shapes (2 tiles in 64 KB)
  TileShape tx4, tz4;  // shapes for kernels holding 3-4 tiles
  double adix_a, adix_b, adiz_a, adiz_b;
};

struct FFTPlan {
  int nx, nz;
  int radx[12], nradx, radz[12], nradz;
  const double2 *twxf, *twxb, *twzf, *twzb;  // exp(-+2 pi i m/n)
  int codex, codez;  // radices, 3 bits per stage
  int tlx;  // complex lines per x-FFT tile (2 rows each)
  int tlz;  // kx columns per z-FFT tile
  int tly;  // modes per y-solve tile
};

// Precision.  The state U, V, W, P, DIV and the whole Poisson solve (RDP,
// spectrum, DP) are FP64 in both modes.  FP64 mode: every other field is
// FP64 and the line kernels read the state itself (su..sp = u..p).  Mixed
// mode: RUH, HH, UH, VV1, WW1, UW, VV3 are FP32; the line kernels read FP32
// shadows of the state (su..sp = uf..pf, refreshed wherever the state is
// written); UH holds the ADI increment DELTA = UH - U, RDP is formed in FP64
// from DIV(U) and DELTA, and the update forms U + DELTA - DT*grad(DP) in
// FP64.
struct Fields {
  double *u, *v, *w, *p;
  float *uf, *vf, *wf, *pf;  // FP32 shadows (mixed mode)
  // mixed mode: DIV of the state, which the next RDP starts from, in FP32
  // (after the FP64 projection its values are at the rounding level); the
  // update passes the FP64 D1x U of the new field to k_upd_z in RDP's
  // buffer, which GETDP has consumed by then
  float* divf;
  // mixed mode double-buffers V: the update reads V(J+1) of the old field
  // while other blocks write the new one
  const double* vold;
  // the state as read by the line kernels
  const void *su, *sv, *sw, *sp;
  double *hh[3], *hhp[3];
  double *ruh[3];
  double *uh[3];
  double *vv1, *ww1, *uw, *vv3;
  double *rdp, *dp;
  double *sgs;
  float *sgsf;        // FP32 shadow of SGSVIS (mixed mode, FILES != 0)
  const void *ssgs;   // SGSVIS as read by the SGS line kernels and y terms
  // LES intermediates (allocated when FILES != 0): nu_t interpolated to
  // x-faces (nux), z-faces (nuz), x-z edges (nue), and the compact
  // derivatives dv/dx, dw/dx (at x-faces), du/dz, dv/dz (at z-faces); in
  // the compute type of the line kernels (FP32 in mixed mode)
  double *nux, *nuz, *nue, *dvdx, *dwdx, *dudz, *dvdz;
  double *div;
  double *ubc_u, *ubc_w;  // top ghost values U(I,N2,K,1), U(I,N2,K,3) from BCOND (NX*NZ)
  // wall stresses TAUW(I,K,L), TAUB(I,K,L) at [(L-1)*NX*NZ + (K-1)*NX + I-1]:
  // L = 1 bottom, 2 top (zero unless a two-wall model sets it)
  double *tauw, *taub;
  double *tauwe;          // WMODEL = 4: the cell-centre stresses TAUWE(I,K,1:4)
  double *iwm;            // WMODEL = 2: IW_NPL state planes
  double2 *spec;          // Poisson spectrum [jm][kz][kx], rows of nxsp: kx < NX/2, column 0 = X(0) + i X(NX/2)
  double *pib, *pgt;      // Poisson TDMAP factors per mode: 1/BET(J), GAM(J+1) [jm][kz][kx]
  double *sc;             // device scalars (SC_*)
  double *red;            // reduction partials
  int nred;               // number of partial slots
  // per-block partial sums of UPCALC's flow-rate terms, four regions of
  // flowr entries (flowr >= the block count of any line kernel):
  //   FP_DU: sum DELTA_u*DY (ADI z-sweep of u)   FP_U: sum U_u*DY (before the update)
  //   FP_DW: sum DELTA_w*DY (ADI x-sweep of w)   FP_W: sum U_w*DY
  double *flowp;
  int flowr;
  // dynamic SGS models (FILES = 1, 3): per-plane tile partial sums (L:M and
  // M:M; FILES = 3 also the five sums of each of its two filter levels), the
  // coefficient C(J) and, for FILES = 3, BETA(J) at [J]
  double *dyp, *smagc, *sdbeta;
  // the 2 x 64 max slots (DIVMAX, CFL bit patterns) of the update's checks;
  // zero between steps (cleared at set-up and by the kernel that reads them)
  double *mxs;
};
enum { FP_DU = 0, FP_U = 1, FP_DW = 2, FP_W = 3 };

// kernel ids for the per-kernel event timer
// kernel ids for the per-kernel event timer (KID_COUNT <= C3D_NKID)
enum {
  KID_BCOND = 0, KID_RHS_X, KID_RHS_Z, KID_RHS_X2, KID_RHS_Y, KID_ADI_Y, KID_ADI_X, KID_ADI_Z,
  KID_RHSDP_X, KID_RHSDP_Z, KID_FFT_X, KID_FFT_Z, KID_POIS_Y, KID_FFT_XINV, KID_FLOW_RED,
  KID_FLOW_FIN, KID_UPD_X, KID_GHOSTS, KID_UPD_Z, KID_DIV_X, KID_DIV_Z, KID_CFL, KID_CHKMF,
  KID_ZERO, KID_APPLY, KID_SGS, KID_WALL, KID_SGS_X, KID_SGS_Z, KID_SGS_Z2, KID_SGS_X2,
  KID_WALLSS, KID_FILTER, KID_RHS_ZW, KID_UPD_V, KID_ADI_XR, KID_ADI_ZR, KID_COUNT
};

// host-side recorder: when `on`, every launch is bracketed by two events on
// the launch stream; times are summed per kernel id on request
struct KTimer {
  bool on = false;
  int only = -1;  // >= 0: bracket only launches of this kernel id
  int used = 0;
  hipEvent_t* ev = nullptr;  // 2*cap events
  int* id = nullptr;         // kernel id per pair
  int cap = 0;
};

struct Ctx {
  Geo g;
  Metrics m;
  Scal s;
  Ops o;
  FFTPlan f;
  Fields d;
  hipStream_t st;
  KTimer* kt;
  // batched launches (ensembles of same-shape members): when bdev is set,
  // the launchers run every kernel once for bn members, blockIdx.z selecting
  // the member's Ctx in the device array bdev (this Ctx supplies the shapes,
  // stream and timer)
  const Ctx* bdev = nullptr;
  int bn = 0;
};

// kernel argument carrying the context: by value (one channel) or as a
// device array indexed by blockIdx.z (a batch of same-shape channels)
template <bool B>
struct CtxArg;
template <>
struct CtxArg<false> {
  Ctx c;
};
template <>
struct CtxArg<true> {
  const Ctx* cs;
};
__device__ __forceinline__ const Ctx& ctx_get(const CtxArg<false>& a) { return a.c; }
__device__ __forceinline__ const Ctx& ctx_get(const CtxArg<true>& a) { return a.cs[blockIdx.z]; }

#define C3D_TARGS(...) __VA_ARGS__
// launches kernel template K<TA..., B> for the context C (one channel, or
// C.bn channels along grid z)
#define C3D_LAUNCH(K, TA, GRID, BLOCK, LDS, C, ...)                                                      \
  do {                                                                                                   \
    if ((C).bdev) {                                                                                      \
      dim3 g_ = (GRID);                                                                                  \
      g_.z = (unsigned)(C).bn;                                                                           \
      hipLaunchKernelGGL(HIP_KERNEL_NAME(K<C3D_TARGS TA, true>), g_, BLOCK, LDS, (C).st,                 \
                         CtxArg<true>{(C).bdev}, ##__VA_ARGS__);                                         \
    } else {                                                                                             \
      hipLaunchKernelGGL(HIP_KERNEL_NAME(K<C3D_TARGS TA, false>), GRID, BLOCK, LDS, (C).st,              \
                         CtxArg<false>{(C)}, ##__VA_ARGS__);                                             \
    }                                                                                                    \
  } while (0)
// the same for kernels templated on the context argument only
#define C3D_LAUNCH0(K, GRID, BLOCK, LDS, C, ...)                                                         \
  do {                                                                                                   \
    if ((C).bdev) {                                                                                      \
      dim3 g_ = (GRID);                                                                                  \
      g_.z = (unsigned)(C).bn;                                                                           \
      hipLaunchKernelGGL(K<true>, g_, BLOCK, LDS, (C).st, CtxArg<true>{(C).bdev}, ##__VA_ARGS__);       \
    } else {                                                                                             \
      hipLaunchKernelGGL(K<false>, GRID, BLOCK, LDS, (C).st, CtxArg<false>{(C)}, ##__VA_ARGS__);         \
    }                                                                                                    \
  } while (0)

// brackets a launch with timing events when the recorder is on
struct KScope {
  const Ctx& c;
  int slot;
  KScope(const Ctx& cc, int kid) : c(cc), slot(-1) {
    KTimer* t = c.kt;
    if (t && t->on && t->used < t->cap && (t->only < 0 || t->only == kid)) {
      slot = t->used++;
      t->id[slot] = kid;
      (void)hipEventRecord(t->ev[2 * slot], c.st);
    }
  }
  ~KScope() {
    if (slot >= 0) (void)hipEventRecord(c.kt->ev[2 * slot + 1], c.st);
  }
};

// host launchers (c3d_kernels.hip)
namespace c3dk {
void bcond(const Ctx& c);
void rhs(const Ctx& c);      // RHS1/2/3 complete (x, z and y parts)
void rhs_xz(const Ctx& c);   // x and z parts only
void rhs_lines(const Ctx& c);
void getuh(const Ctx& c);    // GETUH1/2/3 from the complete RUH
void getuh_fused(const Ctx& c);  // y parts of RHS1/2/3 fused into GETUH1/2/3
void rhsdp(const Ctx& c);
void getdp(const Ctx& c);
void upcalc(const Ctx& c, bool fuse_checks);
void flow_fin(const Ctx& c);  // DMPRESG/DMPRESG3 from the ADI sweeps' partial sums
void upd_x(const Ctx& c, bool chk, bool refine = false);
void upd_z(const Ctx& c, bool chk, bool refine = false);
void div_rdp(const Ctx& c);
void pois_table(const Ctx& c);
void divcheck(const Ctx& c);
void cfl(const Ctx& c);
void chkmf(const Ctx& c, bool reference_sums);
void plane_stats(const Ctx& c, double* out);  // ENERGY's plane statistics, 20 x N2M on the device
void cellcentre(const Ctx& c, int q, double* out);  // WRITEUP's cell-centre quantity q, N1M x N2M x N3M
void sgscs(const Ctx& c);
void sgsdy(const Ctx& c);            // SGSDY, the dynamic model (FILES = 1)
void sgssd(const Ctx& c);            // SGSSD, the scale-dependent dynamic model (FILES = 3)
int sgsdy_tiles(const Geo& g);      // tiles per plane of its partial sums
void wallmodel(const Ctx& c, bool both = false);  // WALLMODEL_INSTANT; both: also the top wall (WMODEL = 6)
void wallmodel_werner(const Ctx& c);  // WALLMODEL_WERNER (WMODEL = 4), both walls
// WALLMODEL_IWMLES_SMOOTH (WMODEL = 2); ntime > 0: this call's NTIME, else
// the last call's + 1 (graph replays)
void wallmodel_iwmles(const Ctx& c, int ntime);
void wallss(const Ctx& c);
void sgs_terms(const Ctx& c);
void init_attrs();
void fft_attrs();  // LDS limits of the FFT kernels
void filter(const Ctx& c, int width, const double* in, double* out, int nplanes, int ld1, int ld3);
void apply_op(const Ctx& c, int op, int dir, const double* in, double* out);
void zero(const Ctx& c, double* p, long long n);
void to_f32(const Ctx& c, const double* in, float* out, long long n);
void to_f64(const Ctx& c, const float* in, double* out, long long n);
void add_f32(const Ctx& c, const double* a, const float* b, double* out, long long n);
}  // namespace c3dk
