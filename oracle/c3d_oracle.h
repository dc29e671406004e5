/*
 * c3d_oracle.h -- CPU restatement (FP64) of the compact3d per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP library is
 * compared against; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never does.
 *
 * Every function restates one reference subroutine of compact3d.f
 * (file:line given at each definition in c3d_oracle.c) and keeps the
 * reference's operation order so that results agree to rounding.  Arrays use
 * the reference's Fortran layouts with M1=N1, M2=N2, M3=N3:
 *   U-like   (0:N1,0:N2,0:N3[,3])  column-major, I fastest
 *   P-like   (1:N1,1:N2,1:N3)
 */
#ifndef C3D_ORACLE_H
#define C3D_ORACLE_H
#include <stddef.h>

typedef struct c3o c3o;

/* setup: SETUP/MESH/INDICES/INIWAVE/METRICPOISSON (compact3d.f:192-494, 9182-9263) */
c3o *c3o_new(int n1, int n2, int n3, double re, double alx, double aly,
             double alz, double dt, double files, double wmodel, double cks);
void c3o_free(c3o *o);
void c3o_set_threads(int n);
/* the planes of COMMON /IWMLES/ (compact3d.f:3399-3408) in their order */
enum {
  IW_UTX = 0, IW_UTY, IW_TAUWX, IW_TAUWY, IW_FLT, IW_FLTM = 6, IW_FLTP = 8, IW_INTE = 9, IW_INTEM = 14,
  IW_UNSDY = 19, IW_CONV = 21, IW_PRSG = 23, IW_DIFF = 25, IW_LHS = 27, IW_DUDZT = 29, IW_DUDZB = 31,
  IW_FLTUS = 33, IW_TR, IW_DZ, IW_Z0, IW_AX, IW_AY, IW_DELV, IW_DELI, IW_CX, IW_CY, IW_NPL
};
void c3o_set_dt(c3o *o, double dt);
void c3o_set_ifwall(c3o *o, int ifwall);
/* which: 0 Y(0:N2) 1 DY(0:N2) 2 H 3 HM 4 HC 5 HP 6 DYM 7 DYC 8 DYP (1..N2 stored at
 * [j]) 9 AK1(1:N1) 10 AK3(1:N3/2+1) 11 PMJ 12 PCJ 13 PPJ 14 FJPA 15 FJMU 16 FJMV;
 * scalars: 100 DX1 101 DX1Q 102 DX3 103 DX3Q 104 ALY 105 VOL */
int c3o_get_vec(const c3o *o, int which, double *out);
double c3o_get_scalar(const c3o *o, int which);
void c3o_get_tau(const c3o *o, double *tauw, double *taub);
void c3o_set_tau(c3o *o, const double *tauw, const double *taub);
void c3o_get_wsm(const c3o *o, double *wsm3);
void c3o_set_wsm(c3o *o, const double *wsm3);

/* line solvers (compact3d.f:8952-9046, 8630-8809, 9444-9459) */
void c3o_ctdma(int n, double b, const double *d, double *x);
void c3o_ctdma_ref_general(int n, const double *a, const double *b,
                           const double *c, const double *d, double *x);

/* one step phases */
void c3o_bcond(c3o *o, double *U, double *SGSVIS);
void c3o_rhs1(c3o *o, const double *U, const double *P, double *HH,
              const double *HH2, double *RUH1, double presg,
              const double *SGSVIS);
void c3o_rhs2(c3o *o, const double *U, const double *P, double *HH,
              const double *HH2, double *RUH2, const double *SGSVIS);
void c3o_rhs3(c3o *o, const double *U, const double *P, double *HH,
              const double *HH2, double *RUH3, double presg3,
              const double *SGSVIS);
void c3o_getuh1(c3o *o, const double *U, double *UH, const double *RUH1);
void c3o_getuh2(c3o *o, const double *U, double *UH, const double *RUH2);
void c3o_getuh3(c3o *o, const double *U, double *UH, const double *RUH3);
void c3o_rhsdp(c3o *o, double *RDP, const double *UH);
void c3o_getdp(c3o *o, double *DP, const double *RDP);
void c3o_upcalc(c3o *o, double *U, double *P, const double *UH,
                const double *DP, double *presg, double *presg3);
double c3o_divcheck(c3o *o, const double *U, double *DIV);
double c3o_cfl(c3o *o, const double *U);
void c3o_chkmf(c3o *o, const double *U, double *flow1, double *flow3);
void c3o_wallss(c3o *o, const double *U);
void c3o_strain(c3o *o, const double *U, double *SR, int j);
void c3o_sgscs(c3o *o, const double *U, double *SGSVIS);
/* SGSDY (compact3d.f:1250-1400); smagc_out (may be NULL) receives the
 * dynamic coefficient C(J) at [J], J = 1..N2M */
void c3o_sgsdy(c3o *o, const double *U, double *SGSVIS, double *smagc_out);
/* SGSSD compact3d.f:1502-1948 (scale-dependent dynamic model, FILES=3) */
void c3o_sgssd(c3o *o, const double *U, double *SGSVIS, double *smagc_out, double *beta_out);
/* COEFFICIENT :2079-2097 and DQRRT/CHHQR :11548-11677 */
void c3o_sd_coefficient(const double s1[5], const double s2[5], double pa[5]);
double c3o_dqrrt(const double *pa, int n);
/* one plane of SGSSD: its two levels' sums s1, s2 (A..E), PA; returns BETA */
double c3o_sgssd_plane(c3o *o, const double *U, int j, double *s1, double *s2, double *pa);
void c3o_wallmodel_instant(c3o *o, const double *U);
/* WMODEL = 6 (an extension with no reference counterpart): WALLMODEL_INSTANT at both walls */
void c3o_wallmodel_instant2(c3o *o, const double *U);
/* WALLMODEL_IWMLES_SMOOTH compact3d.f:3386-4329 (WMODEL = 2, Y0 = 0), called
 * with GETUP's NTIME; P is (1:N1,1:N2,1:N3) */
void c3o_wallmodel_iwmles(c3o *o, const double *U, const double *P, double presg, double presg3, int ntime);
/* its state: the IW_NPL planes of COMMON /IWMLES/ (N1M x N3M each, I
 * fastest), wm_dt, and the points of the last call whose excessive linear
 * correction needed an equilibrium stress of their own (any may be NULL) */
void c3o_iwm_state(const c3o *o, double *planes, double *wm_dt, int *fixed);
void c3o_iwm_set(c3o *o, const double *planes, double wm_dt);
void c3o_set_cflmax(c3o *o, double cflmax);
/* WALLMODEL_WERNER compact3d.f:4815-4958 (WMODEL = 4; needs N3 <= N1) */
void c3o_wallmodel_werner(c3o *o, const double *U);
void c3o_filter(c3o *o, double *FILV, int n);
void c3o_filter4(c3o *o, double *FILV, int n);

/* the plane statistics of ENERGY (compact3d.f:9902, loops 10 and 20 at
 * :10005-10158): out[s*N2M + J-1] for the 20 quantities listed at the
 * definition */
void c3o_plane_stats(c3o *o, const double *U, const double *P, double *out);

/* GETUP (compact3d.f:879-960): one full step, scratch allocated inside */
void c3o_getup(c3o *o, double *U, double *P, double *HH, double *SGSVIS,
               double *presg, double *presg3, int ntime);

/* reference-style complex DFT of an n1 x n3 column-major plane, sign -1/+1,
 * unnormalised (the FFTW contract GETDP relies on) */
void c3o_dft2d(int n1, int n3, double *re_im_interleaved, int sign);

#endif
