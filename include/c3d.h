/*
 * c3d.h -- C-ABI of the MI355X-native compact3d hot-path library (libc3d.so).
 *
 * The reference (compact3d.f) has no plugin or operator API: its boundary is
 * the set of Fortran subroutines the time loop calls (MAIN, compact3d.f:101-172)
 * and the COMMON blocks they share.  Each entry point below replaces one of
 * them; the project's Fortran interface module (compact3d_amd/fortran/) binds
 * them through ISO_C_BINDING under the reference's subroutine names.
 *
 * Conventions
 *  - Status codes: 0 = success, < 0 = error; the message is available from
 *    c3d_last_error().  No exception or longjmp crosses the ABI.
 *    C3D_ERR_NONFINITE (c3d_step, c3d_sync, c3d_ensemble_sync) says that
 *    the step ran but its DIVMAX or CFLM is NaN or infinite: the state has
 *    diverged (the reference has no status returns; its runs go on with
 *    NaNs, compact3d.f:130-163).
 *  - Host arrays use the reference's Fortran layouts with M1=N1, M2=N2,
 *    M3=N3 (compact3d.f:55-60):
 *      U, HH  (0:N1, 0:N2, 0:N3, 3)   column-major, I fastest
 *      SGSVIS (0:N1, 0:N2, 0:N3)
 *      P      (1:N1, 1:N2, 1:N3)
 *    The state lives in HBM between calls; upload/download are the only
 *    host<->device copies.
 *  - One host thread per handle.  Handles are independent (an ensemble or a
 *    server keeps many).
 */
#ifndef C3D_H
#define C3D_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct c3d_handle c3d_handle;

#define C3D_ERR (-1)
#define C3D_ERR_NONFINITE (-3)

enum c3d_precision {
  C3D_FP64 = 0,  /* every operator in FP64 */
  /* FP32 compact derivative and interpolation operators, RHS, ADI sweeps
   * and their intermediates (the line kernels read FP32 shadows of the
   * state); FP64 state (U, P), RDP, the whole Poisson solve (FFTs and
   * Thomas solves), the pressure correction, the flow-rate constraint,
   * DIVCHECK, CFL and the LES terms; an FP64-residual refinement of the
   * projection when its DIVMAX exceeds 1e-10. */
  C3D_MIXED = 1
};

/* Run configuration.  Mirrors what SETUP reads from parame.ter and what MESH,
 * INDICES, INIWAVE and METRICPOISSON derive from it (compact3d.f:192-494,
 * 9182-9263). */
typedef struct c3d_config {
  int n1, n2, n3;     /* N1, N2, N3 (= M1, M2, M3 of PARAM.H); cells N1-1 x N2-1 x N3-1 */
  double re;          /* RE = 1/nu in the reference's units */
  double alx, aly, alz; /* box; ALY = 2 full channel, 1 half channel */
  double dt;          /* DT */
  double files;       /* SGS model switch FILES: 0 none (DNS), 1 dynamic (SGSDY), 2 Smagorinsky (SGSCS),
                         3 scale-dependent dynamic (SGSSD) */
  double wmodel;      /* WMODEL: 0 resolved walls, 1 instantaneous log-law wall model (bottom wall),
                         2 the integral wall model IWMLES for a smooth wall (bottom wall; needs
                           cflmax > 0),
                         4 Werner-Wengle wall model (both walls; needs N3 <= N1),
                         6 the log-law model at both walls (no counterpart in the
                         reference, whose model 1 leaves the top wall unmodelled) */
  double cks;         /* tanh stretching of y (the reference hard-codes 2.5, :338) */
  const double *y;    /* optional Y(0:N2) to use instead of the tanh grid, or NULL */
  int precision;      /* enum c3d_precision */
  int device;         /* HIP device ordinal */
  void *stream;       /* hipStream_t to run on, or NULL for the library's own */
  double cflmax;      /* CFLMAX of /TSTEP/ (compact3d.f:247): WMODEL = 2 sets its first
                         time-filter weight from it (:3700) and needs it > 0; unused otherwise */
} c3d_config;

/* SETUP/MESH/INDICES/INIWAVE/METRICPOISSON (compact3d.f:192-299, 302-494,
 * 9182-9263): allocates the resident state in HBM and precomputes metrics,
 * line-solve factors and FFT twiddles on the device. */
int c3d_create(const c3d_config *cfg, c3d_handle **out);
/* frees the handle and all device memory */
int c3d_destroy(c3d_handle *h);
/* last error message of this handle (or of the last failed c3d_create) */
const char *c3d_last_error(const c3d_handle *h);

/* READUP (compact3d.f:729-749) followed by the in-memory state of MAIN:
 * copies U, P, HH (AB2 history), SGSVIS and PRESG/PRESG3 to the device.
 * HH and SGSVIS may be NULL (zero). */
int c3d_upload_state(c3d_handle *h, const double *U, const double *P, const double *HH,
                     const double *SGSVIS, double presg, double presg3);
/* copies the resident state back (any pointer may be NULL) */
int c3d_download_state(c3d_handle *h, double *U, double *P, double *HH, double *SGSVIS,
                       double *presg, double *presg3);

/* GETUP(U,P,HH,SS,NTIME,PRESG,PRESG3,SGSVIS) (compact3d.f:879-960): one full
 * time step on the resident state. */
int c3d_getup(c3d_handle *h, int ntime);
/* GETUP + DIVCHECK + CFL as MAIN calls them after the step
 * (compact3d.f:130-138); divmax / cflm may be NULL. */
int c3d_step(c3d_handle *h, int ntime, double *divmax, double *cflm);
/* nsteps of MAIN's loop (GETUP with DIV, DIVMAX and CFL formed) queued on
 * the handle's stream without waiting; use_graph != 0 replays two captured
 * HIP graphs (even and odd steps) instead of launching kernel by kernel.
 * No counterpart in the reference: it is the batched driver the ensemble
 * and serving loops use.  Mixed mode skips the per-step refinement check;
 * c3d_sync applies it to the last step. */
int c3d_advance(c3d_handle *h, int nsteps, int use_graph);
/* waits for the handle's stream and returns DIVMAX and CFLM of the last
 * step (mixed mode: after the refinement check); either may be NULL.
 * C3D_ERR_NONFINITE when either is NaN or infinite (both are still
 * returned). */
int c3d_sync(c3d_handle *h, double *divmax, double *cflm);

/* Ensembles (no counterpart in the reference, which runs one channel per
 * process): independent handles of any grids, precisions and models whose
 * steps are captured together as the parallel branches of one HIP graph, so
 * one launch advances every member by one step (GETUP with DIV, DIVMAX and
 * CFL, as c3d_advance) and the device overlaps their kernels.  Members stay
 * owned by the caller and usable on their own; create an ensemble once for
 * a fixed set of members and keep it while they run (a member must not be
 * in two ensembles at the same time).  c3d_ensemble_advance queues nsteps
 * steps without waiting (each member's stream is ordered before and after
 * them); c3d_ensemble_sync waits and returns every member's DIVMAX and CFLM
 * (arrays of n, either may be NULL; mixed-mode members apply their
 * refinement check as in c3d_sync). */
typedef struct c3d_ensemble c3d_ensemble;
int c3d_ensemble_create(c3d_handle *const *members, int n, c3d_ensemble **out);
int c3d_ensemble_advance(c3d_ensemble *e, int nsteps);
int c3d_ensemble_sync(c3d_ensemble *e, double *divmax, double *cflm);
int c3d_ensemble_destroy(c3d_ensemble *e);
const char *c3d_ensemble_last_error(const c3d_ensemble *e);

/* Phase entry points of GETUP, for operator-level parity.  They act on the
 * resident state exactly as the named reference routines do. */
int c3d_bcond(c3d_handle *h);     /* BCOND  compact3d.f:963-1001 */
int c3d_rhs(c3d_handle *h);       /* RHS1/RHS2/RHS3 compact3d.f:6063, 6784, 7399 (after the
                                     GETUP prologue :895-923 and BCOND) */
int c3d_getuh(c3d_handle *h);     /* GETUH1/2/3 compact3d.f:8218, 8344, 8478; in
                                     mixed precision also RHSDP (:8117), formed by
                                     the last ADI sweeps */
int c3d_dpcalc(c3d_handle *h);    /* DPCALC = RHSDP + GETDP compact3d.f:1037-1049
                                     (mixed precision: GETDP only) */
int c3d_upcalc(c3d_handle *h);    /* UPCALC compact3d.f:1052-1238 */
int c3d_sgs(c3d_handle *h);       /* SGSDY (FILES=1) compact3d.f:1250 / SGSCS (FILES=2) :1413 /
                                     SGSSD (FILES=3) :1502 */
int c3d_wall_model(c3d_handle *h); /* WALL_MODEL compact3d.f:1952-2013 -> WALLMODEL_INSTANT :2319,
                                     WALLMODEL_IWMLES_SMOOTH :3386 (NTIME = the last call's + 1),
                                     WALLMODEL_WERNER :4815 */
int c3d_wallss(c3d_handle *h);    /* WALLSS compact3d.f:11077 (sets WSM(1:3)) */

/* FILTER (width 2, compact3d.f:5190) / FILTER4 (width 4, :5481) of a host
 * stack of x-z planes FILV(1:N1,1:N3,n); in and out may alias */
int c3d_filter(c3d_handle *h, int width, const double *filv_in, double *filv_out, int n);

/* DIVCHECK(U,TIME,DIVMAX,NTIME) compact3d.f:9049 */
int c3d_divcheck(c3d_handle *h, double *divmax);
/* CFL(U,CFLM) compact3d.f:9138 */
int c3d_cfl(c3d_handle *h, double *cflm);
/* CHKMF(U,TIME) compact3d.f:9463 (reproduces its sums over U(1,J,K,1) and
 * U(I,J,1,3), :9484 and :9495) */
int c3d_chkmf(c3d_handle *h, double *flow1, double *flow3);
/* the flow rates CHKMF intends: sum over all cells */
int c3d_flowrate(c3d_handle *h, double *flow1, double *flow3);
/* the plane statistics ENERGY forms each output interval (compact3d.f:9902,
 * loops 10 and 20 at :10005-10158) of the resident field:
 * out[s*(N2-1) + J-1], J = 1..N2-1, s = 0 VM(J,1..3) 3 PM 4 WM(J,1..3)
 * (mean vorticity) 7 VRMS(J,1..6) (uu vv ww uv uw vw) 13 PRMS 14 PV
 * 15 WRMS(J,1..3) 18 U2V 19 VW2; 20*(N2-1) values */
int c3d_stats(c3d_handle *h, double *out);
/* c3d_stats without any allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops).  Its kernel stages the statistics in a library-owned device array
 * of C3D_STATS_NOALLOC_MAX_PLANES * 20 doubles per GPU (held until the
 * process ends) instead of the handle's scratch field; grids with N2-1
 * above that limit are refused (use c3d_stats).  Calls that overlap on one
 * GPU share that array and race. */
#define C3D_STATS_NOALLOC_MAX_PLANES 2048
int c3d_stats_noalloc(c3d_handle *h, double *out);
/* the dynamic models' per-plane profiles of the last SGS evaluation:
 * smagc[J-1] = SMAGC(J) (COMMON/SGSMEAN, FILES = 1 or 3) and beta[J-1] =
 * SGSSD's BETA(J) (the BETAZ(J) of BETA.plt before its 0.125 floor,
 * compact3d.f:1802-1804; FILES = 3), J = 1..N2-1; either may be NULL */
int c3d_sgs_profiles(c3d_handle *h, double *smagc, double *beta);

/* WRITEUP's Tecplot output (compact3d.f:757-835).
 * c3d_cellcentre: the cell-centre values WRITEUP writes (:796-808),
 * out[q*n + (K-1)*(N2-1)*(N1-1) + (J-1)*(N1-1) + I-1], n = (N1-1)(N2-1)(N3-1),
 * q = 0 UF 1 VF 2 WF 3 P.
 * c3d_write_flow: RESULT/FLOW-*.plt of the resident field (:791-812) to
 * `path`, text as the reference's Fortran runtime writes it.
 * c3d_write_strain: RESULT/STRAIN-*.plt (:814-832) from the bottom-wall
 * TAUW; utau = SQRT(WSMO(1)) (WMODEL=0) or SQRT(TAOWO) as WRITEUP forms it.
 * c3d_format_flow / c3d_format_strain: the same files from host arrays
 * (no handle, no GPU): xc[N1-1], yc[N2-1], zc[N3-1] = XC(1:), YC(1:),
 * ZC(1:); cell as c3d_cellcentre's out; tauw[(K-1)*(N1-1) + I-1] =
 * TAUW(I,K,1).  Errors: c3d_last_error(NULL). */
int c3d_cellcentre(c3d_handle *h, double *out);
int c3d_write_flow(c3d_handle *h, const char *path, int ntime);
int c3d_write_strain(c3d_handle *h, const char *path, int ntime, double utau);
int c3d_format_flow(const char *path, int ntime, int n1m, int n2m, int n3m, const double *xc, const double *yc,
                    const double *zc, const double *cell);
int c3d_format_strain(const char *path, int ntime, int n1m, int n3m, const double *xc, const double *zc,
                      const double *tauw, double utau);

/* Intermediate fields of the last step, for operator-level parity. */
enum c3d_field {
  C3D_F_RUH1 = 1, C3D_F_RUH2, C3D_F_RUH3, /* U-like single component */
  C3D_F_UH1, C3D_F_UH2, C3D_F_UH3,
  C3D_F_RDP, C3D_F_DP,                    /* P-like */
  C3D_F_HH1, C3D_F_HH2, C3D_F_HH3,        /* current HH, U-like single component */
  C3D_F_TAUW, C3D_F_TAUB,                 /* (0:N1,0:N3,2) */
  C3D_F_IWM, /* WMODEL = 2: the 43 planes of the reference's COMMON /IWMLES/ (compact3d.f:3399-3408),
                N1-1 x N3-1 each (I fastest), in that order */
  C3D_F_DIV  /* P-like: the divergence DIVCHECK forms (compact3d.f:9112) of the resident field (mixed
                mode: the FP64 DIV the last step's update formed, which the next step's RDP starts
                from; FP64 mode: formed by DIVCHECK when read) */
};
/* copies one device field into a host array in its Fortran layout */
int c3d_get_field(c3d_handle *h, int which, double *out);
/* device scalars: 0 PRESG 1 PRESG3 2 DMPRESG 3 DMPRESG3 4 FLOW1 5 FLOW3
 * 6 WSM(1) 7 WSM(2) 8 WSM(3); get only (mixed mode): 9 refinement passes
 * run, 10 DIVMAX of the last step's first projection; WMODEL = 2: 11 NTIME
 * of the last wall-model call (c3d_advance and c3d_wall_model continue
 * from it), 12 CFLMAX (/TSTEP/, sets the first filter weight; the config's
 * cflmax), 13 wm_dt; 14 TAOW (/WALLTAO/); WMODEL = 2: 15 the initial mean
 * velocity, 16 the initial friction velocity, 17 the mean first-cell
 * pressure of the last compute step.  Slots 11-17 with 6-8 and the fields
 * C3D_F_TAUW, C3D_F_TAUB, C3D_F_IWM (and C3D_F_DIV in mixed mode) are the
 * state a step carries beyond U, P, HH and PRESG: a checkpoint saves them. */
#define C3D_NSCALAR 18
int c3d_get_scalar(c3d_handle *h, int which, double *out);
int c3d_set_scalar(c3d_handle *h, int which, double value);
/* the inverse of c3d_get_field for the carried state: C3D_F_TAUW and
 * C3D_F_TAUB ((0:N1,0:N3,2), entries outside I = 1..N1-1, K = 1..N3-1
 * ignored), C3D_F_IWM (a saved or constructed state of the integral wall
 * model) and C3D_F_DIV (P-like) */
int c3d_set_field(c3d_handle *h, int which, const double *in);
int c3d_set_dt(c3d_handle *h, double dt);
/* the grid MESH built (compact3d.f:302-443): Y(0:N2) and {DX1, DX3, ALY, VOL} */
int c3d_get_mesh(c3d_handle *h, double *y, double *scal4);

/* Operator test hook: applies one compact operator (CTDMA11/CTDMA33 with the
 * reference's RHS builders, e.g. compact3d.f:6146-6151, 6200-6204, 6232-6236)
 * along x (dir 1) or z (dir 3) to a host P-like array in (1:N1,1:N2,1:N3)
 * layout, on every line J=1..N2-1.
 * op: 0 D2 (b=10), 1 D1 centre->face (b=22), 2 D1 face->centre (b=22),
 *     3 I centre->face (b=6), 4 I face->centre (b=6) */
int c3d_apply_op(c3d_handle *h, int op, int dir, const double *in, double *out);

/* Per-kernel timing (no counterpart in the reference; it replaces the
 * reference's wall-clock print of MAIN, compact3d.f:101-103, 160-163).
 * c3d_timing(h, 1, cap) brackets every following launch with two HIP events
 * on the handle's stream (up to cap launches); c3d_kernel_times()
 * synchronises, returns per-kernel-id summed milliseconds and launch counts
 * (arrays of C3D_NKID entries) and resets the recorder. */
#define C3D_NKID 48
int c3d_timing(c3d_handle *h, int on, int capacity);
/* restricts the recorder to launches of one kernel id (kid < 0: all), so a
 * timed loop carries two events per launch of that kernel only */
int c3d_timing_select(c3d_handle *h, int kid);
int c3d_kernel_times(c3d_handle *h, double *ms, int *count);
const char *c3d_kernel_name(int id);

/* Tile-shape autotuner (no counterpart in the reference).  Times candidate
 * line-group and tile shapes of the step's kernels (x-, z- and y-line
 * tiles, Poisson y tiles, FFT tiles) on a scratch instance of this
 * configuration, `steps` steps per candidate, one knob after another from
 * the rule-derived shapes, and records the fastest set in a process-wide
 * table keyed by (N1, N2, N3, precision, FILES != 0).  Handles created
 * afterwards with that key use it; existing handles keep theirs.  Shapes
 * change only the rounding order of the line solves.  shapes
 * (C3D_NSHAPES ints, may be NULL): x lines per tile, FP64 z columns, FP32 z
 * columns, FP32 z segment, FP64 y columns and chunks, FP32 y threads per
 * chunk row and chunks, Poisson y columns and chunks, x- and z-FFT lines
 * per tile, FP32 y columns per thread (1, or 2: neighbouring column pairs
 * moved as 8-byte values, so a tile covers twice its threads per row in
 * columns); ms (2, may be NULL): ms per step with the rule-derived and with
 * the chosen shapes. */
#define C3D_NSHAPES 13
int c3d_autotune(const c3d_config *cfg, int steps, int *shapes, double *ms);
/* records given shapes (c3d_autotune's C3D_NSHAPES-value layout, e.g. from an
 * earlier c3d_autotune) for this configuration without timing anything;
 * each must be the rule-derived shape or one of the autotuner's candidates,
 * else the call fails (c3d_last_error(NULL)) and records nothing */
int c3d_autotune_set(const c3d_config *cfg, const int *shapes);
/* forgets every tuned shape set */
void c3d_autotune_reset(void);

/* device memory held by one handle, bytes (for admission by footprint) */
long long c3d_footprint_bytes(const c3d_config *cfg);

#ifdef __cplusplus
}
#endif
#endif
