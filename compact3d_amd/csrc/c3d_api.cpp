// c3d_api.cpp -- host side of libc3d: the C-ABI declared in include/c3d.h.
//
// Setup reproduces SETUP/MESH/INDICES/INIWAVE/METRICPOISSON
// (compact3d.f:192-494, 9182-9263) on the host and uploads the resulting
// vectors; the step runs entirely on the device from then on.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/c3d.h"
#include "c3d_impl.h"

struct c3d_handle {
  c3d_config cfg;
  Ctx c;
  bool own_stream = false;
  std::string err;
  std::vector<void*> allocs;
  int hh_cur = 0;               // which of the two HH sets is "current"
  double* vbuf[2] = {nullptr, nullptr};  // mixed mode: the two V buffers
  int vcur = 0;
  long long refines = 0;        // mixed mode: refinement passes run so far
  // captured steps for c3d_advance: graph p was captured with the buffer
  // state gkey[p] (HH set, V buffer); -1 = none
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  int gkey[2] = {-1, -1};
  double first_divmax = 0.0;    // mixed mode: DIVMAX of the last first projection
  double* hhbuf[2][3] = {{nullptr}};
  // host copies of the 1-D vectors (reference J indexing)
  std::vector<double> Y, DY, H, HM, HC, HP, DYM, DYC, DYP, FJPA, FJMU, FJMV, PMJ, PCJ, PPJ, AK1, AK3;
  double* dvec = nullptr;       // device block holding all vectors
  double* dytab = nullptr;      // fused y-sweep coefficient table
  float* dytab32 = nullptr;     // the same rounded to FP32 (mixed mode's FP32 sweep)
  double* dtmp = nullptr;       // scratch field for apply_op
  double* dtmp2 = nullptr;
  std::vector<double> stage;    // host staging buffer
  double* hread = nullptr;      // pinned host slots for the per-step DIVMAX/CFLM read-back
  KTimer kt;
};

static std::string g_create_err;

#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) {                                                               \
      h->err = std::string(#x) + ": " + hipGetErrorString(e_);                            \
      return -1;                                                                          \
    }                                                                                     \
  } while (0)

static int fail(c3d_handle* h, const std::string& m) {
  h->err = m;
  return -1;
}

// ---------------------------------------------- the allocation-free scratch
// c3d_stats_noalloc stages its 20 x (N2-1) plane statistics in this
// file-scope device array instead of the handle's scratch field.  Every GPU
// the library runs on holds its own copy (C3D_STATS_NOALLOC_MAX_PLANES * 20
// doubles, 320 KiB) until the process ends.  In host code `g_scratch` names a
// host-side stand-in, not a device address: only the per-device pointers
// from hipGetSymbolAddress (g_scratch_ptr) are passed to kernels and copies.
static __device__ __align__(256) double g_scratch[20 * C3D_STATS_NOALLOC_MAX_PLANES];
static const int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// once per device, with that device current (c3d_create calls it)
static int scratch_init(int device, std::string& err) {
  if (device < 0 || device >= kMaxDevices) {
    err = "device ordinal out of range";
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return 0;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(g_scratch));
  if (e != hipSuccess || !p) {
    err = std::string("hipGetSymbolAddress of the stats scratch: ") + hipGetErrorString(e);
    return -1;
  }
  g_scratch_ptr[device] = (double*)p;
  return 0;
}

static double* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
}

// ------------------------------------------------------------ line solves
static LineSolve make_solve(double a, double b) {
  // a*x[p-1] + b*x[p] + a*x[p+1] = d[p]  ->  alpha, scale, warm-up (c3d_dev.h)
  LineSolve s;
  if (a == 0.0) {
    s.alpha = 0.0;
    s.scale = 1.0 / b;
    s.w = 0;
    return s;
  }
  long double beta = (long double)b / (long double)a;
  long double sq = sqrtl(beta * beta - 4.0L);
  long double al = 2.0L / (beta + (beta > 0 ? sq : -sq));
  s.alpha = (double)al;
  s.scale = (double)(al / (long double)a);
  long double aa = fabsl(al);
  int w = (aa > 0) ? (int)ceill(logl(0x1p-60L) / logl(aa)) + 2 : 0;
  s.w = w;
  return s;
}

// carry-solve parameters of a periodic (a, b, a) system for segments of
// `seg` points: hop = (-alpha)^seg, nhop terms with |alpha|^(nhop*seg) <= 2^-60
static PSolve make_psolve(double a, double b, int seg) {
  const LineSolve ls = make_solve(a, b);
  PSolve p;
  p.alpha = ls.alpha;
  p.scale = ls.scale;
  long double h = 1.0L;
  for (int q = 0; q < seg; q++) h *= -(long double)ls.alpha;
  p.hop = (double)h;
  const long double aa = fabsl((long double)ls.alpha);
  if (aa == 0.0L)
    p.nhop = 0;
  else
    p.nhop = std::max(1, (int)ceill(logl(0x1p-60L) / (seg * logl(aa))));
  return p;
}

// x-line tiles: the staging's pair index arithmetic needs the block, L*G =
// (2L/seg)*(n/2) threads, to be a multiple of n/2, i.e. L a multiple of
// seg/2; L is kept a multiple of seg where L*G <= 256 allows it, and lines
// of 512 points (G = 64) take 4 lines per 256-thread tile
static int xline_step(int seg, int G) { return seg * G <= 256 ? seg : std::max(1, seg / 2); }

// row stride of the stored spectrum (complex values): the smallest odd
// multiple of 16 (256 bytes) that holds n columns, so that rows written or
// read together do not sit a power of two apart
static int spec_pitch(int n) {
  int p = (n + 15) / 16;
  if (p % 2 == 0) p++;
  return 16 * p;
}

// line-engine tile: seg = 8 when it divides n (else 2, the other compiled
// length), G = n/seg segments, L lines per tile.  x-lines: L*G <= 256 with L
// a multiple of xline_step, 2 LDS row tiles of n*(L+1) plus the exchange
// arrays within 64 KB (so n <= 512).  z-lines: L*G <= 512 and L a multiple
// of 16 (128-byte rows) up to 64.
static bool make_lshape(int n, bool staged, size_t es, LShape& sh) {
  const int seg = (n % 8 == 0) ? 8 : 2;
  if (n % seg) return false;
  const int G = n / seg;
  const int tmax = staged ? 256 : 512;
  if (G > tmax) return false;
  int L = std::min(64, tmax / G);
  if (staged) {
    const int step = xline_step(seg, G);
    L -= L % step;
    while (L > step && es * (2 * (size_t)n * (L + 1) + 4 * (size_t)G * L) > 65536) L -= step;
    if (L < step || es * (2 * (size_t)n * (L + 1) + 4 * (size_t)G * L) > 65536) return false;
  } else if (L >= 16) {
    L -= L % 16;
  }
  sh.n = n;
  sh.seg = seg;
  sh.G = G;
  sh.L = L;
  return true;
}

// FP32 z-line tile: 128-byte rows need 32 columns, so 16-point segments
// (8 or 2 when they do not divide n) keep L*G <= 512
static bool make_lshape32(int n, LShape& sh) {
  const int seg = (n % 16 == 0) ? 16 : (n % 8 == 0) ? 8 : 2;
  const int G = n / seg;
  if (G > 512) return false;
  int L = std::min(64, 512 / G);
  if (L >= 32)
    L -= L % 32;
  else if (L >= 16)
    L -= L % 16;
  sh.n = n;
  sh.seg = seg;
  sh.G = G;
  sh.L = L;
  return true;
}

static OpSet make_opset(int seg, double adi_a, double adi_b) {
  OpSet o;
  o.s6 = make_psolve(1.0, 6.0, seg);
  o.s10 = make_psolve(1.0, 10.0, seg);
  o.s22 = make_psolve(1.0, 22.0, seg);
  o.adi = make_psolve(adi_a, adi_b, seg);
  return o;
}

// y tiles: C chunks of at most `ch` points (C a power of two) and L columns,
// L*C <= nt; L a multiple of `row` columns where possible
static YShape make_yshape(int m, int nt, int row, int ch = 12) {
  YShape y;
  y.C = 1;
  while (y.C * ch < m) y.C *= 2;
  y.L = std::max(1, std::min(64, nt / y.C));
  if (y.L >= row) y.L -= y.L % row;
  return y;
}

// FP32 y tiles of column pairs: chunks of at most 6 points, L threads (a
// multiple of 16, i.e. 32 columns = 128-byte rows, where possible)
static YShape make_yshape_pair(int m, int nt) {
  YShape y = make_yshape(m, nt, 16, 6);
  y.P = 2;
  return y;
}

static TileShape make_tile(int n, int ntiles) {
  TileShape t;
  t.n = n;
  int g = (n + 7) / 8;  // segments of ~8 points
  if (g > C3D_NT) g = C3D_NT;
  int l = C3D_NT / g;
  int lmax = (int)(65536 / (sizeof(double) * (size_t)n * ntiles)) - 1;
  if (l > lmax) l = lmax;
  if (l < 1) l = 1;
  // n*l must fit MAXE*C3D_NT = 4096 elements
  while (n * l > 4096 && l > 1) l--;
  t.nl = l;
  t.ls = l + 1;
  t.g = g;
  t.seg = (n + g - 1) / g;
  return t;
}

static bool factor_fft(int n, int* rad, int& nr) {
  nr = 0;
  while (n % 4 == 0 && n > 0) {
    rad[nr++] = 4;
    n /= 4;
  }
  const int ps[3] = {2, 3, 5};
  for (int q = 0; q < 3; q++)
    while (n % ps[q] == 0) {
      rad[nr++] = ps[q];
      n /= ps[q];
    }
  return n == 1 && nr <= 12;
}

static std::vector<double> twiddles(int n, int sign) {
  std::vector<double> t(2 * (size_t)n);
  const long double tp = 6.283185307179586476925286766559L;
  for (int m = 0; m < n; m++) {
    long double a = sign * tp * (long double)m / (long double)n;
    t[2 * m] = (double)cosl(a);
    t[2 * m + 1] = (double)sinl(a);
  }
  return t;
}

static int pow2_floor(int x) {
  int p = 1;
  while (p * 2 <= x) p *= 2;
  return p;
}

// ------------------------------------------------------------------ mesh
static void build_mesh(c3d_handle* h) {
  const c3d_config& cf = h->cfg;
  const int N2 = cf.n2, N2M = N2 - 1, N1M = cf.n1 - 1, N3M = cf.n3 - 1;
  auto z = [&](std::vector<double>& v, int n) { v.assign(n, 0.0); };
  z(h->Y, N2 + 2); z(h->DY, N2 + 2); z(h->H, N2 + 2); z(h->HM, N2 + 2); z(h->HC, N2 + 2);
  z(h->HP, N2 + 2); z(h->DYM, N2 + 2); z(h->DYC, N2 + 2); z(h->DYP, N2 + 2); z(h->FJPA, N2 + 2);
  z(h->FJMU, N2 + 2); z(h->FJMV, N2 + 2); z(h->PMJ, N2 + 2); z(h->PCJ, N2 + 2); z(h->PPJ, N2 + 2);
  z(h->AK1, N1M + 2); z(h->AK3, N3M / 2 + 3);
  std::vector<double>&Y = h->Y, &DY = h->DY, &H = h->H;
  double aly = cf.aly;
  const double CKS = cf.cks, Y0 = 0.0;
  // MESH (:338-361)
  if (cf.y) {
    for (int j = 0; j <= N2; j++) Y[j] = cf.y[j];
  } else {
    Y[0] = (cf.wmodel == 0.0) ? 0.0 : Y0;
    for (int j = 1; j <= N2; j++) {
      if (cf.wmodel == 0.0) {
        if (aly == 2.0)
          Y[j] = 1 - tanh(CKS * (1 - aly * (j - 1) / (N2 - 1))) / tanh(CKS);
        else
          Y[j] = 1 - tanh(CKS * (1 - 2. * aly * (j - 1) / (2 * N2 - 1))) / tanh(CKS);
      } else {
        Y[j] = (aly - Y0) * (j - 1) / (N2 - 1) + Y0;
      }
    }
  }
  aly = Y[N2];
  Scal& s = h->c.s;
  s.aly = aly;
  s.re = cf.re;
  s.rinv = 1. / cf.re;
  s.dt = cf.dt;
  s.dx1 = (double)N1M / cf.alx;
  s.dx3 = (double)N3M / cf.alz;
  s.dx1q = s.dx1 * s.dx1;
  s.dx3q = s.dx3 * s.dx3;
  s.vol = cf.alx * aly * cf.alz;
  s.alx = cf.alx;
  s.ifwall = 0;
  s.files = cf.files != 0.0;
  s.wmodel = cf.wmodel;
  DY[1] = Y[2];
  for (int j = 2; j <= N2M; j++) {
    DY[j] = Y[j + 1] - Y[j];
    H[j] = 0.5 * (DY[j] + DY[j - 1]);
  }
  H[1] = 0.5 * DY[1];
  H[N2] = 0.5 * DY[N2M];
  std::vector<double>&HP = h->HP, &HC = h->HC, &HM = h->HM;
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
  if (cf.wmodel != 0.0) {
    HP[1] = 1.0 / H[2] / DY[1];
    HC[1] = 1.0 / H[2] / DY[1];
    HM[1] = 0.0;
    HP[N2M] = 0.0;
    HC[N2M] = 1.0 / H[N2M] / DY[N2M];
    HM[N2M] = 1.0 / H[N2M] / DY[N2M];
  }
  for (int j = 2; j <= N2M; j++) {
    h->DYP[j] = 1.0 / H[j] / DY[j];
    h->DYC[j] = 1.0 / H[j] * (1.0 / DY[j] + 1.0 / DY[j - 1]);
    h->DYM[j] = 1.0 / H[j] / DY[j - 1];
  }
  // INDICES (:468-481): wall flags
  for (int j = 1; j <= N2M; j++) {
    int jpa = (j == N2M) ? N2M : j + 1;
    int jmu = (j == 1) ? 1 : j - 1;
    int jmv = (j == 2) ? 2 : j - 1;
    h->FJPA[j] = jpa - j;
    h->FJMU[j] = j - jmu;
    h->FJMV[j] = j - jmv;
  }
  // INIWAVE (:9197-9222)
  const double PI = acos(-1.0);
  for (int i = 1; i <= N1M; i++) {
    int ii = i;
    if (i > N1M / 2 + 1) ii = i - N1M;
    double c = cos(((ii - 1) * 2.0 * PI) / N1M);
    double den = 11. + c;
    h->AK1[i] = 288. * s.dx1q * (1. - c) / (den * den);
  }
  for (int k = 1; k <= N3M / 2 + 1; k++) {
    double c = cos(((k - 1) * 2. * PI) / N3M);
    double den = 11. + c;
    h->AK3[k] = 288. * s.dx3q * (1. - c) / (den * den);
  }
  // METRICPOISSON (:9249-9260)
  for (int jc = 1; jc <= N2M; jc++) {
    int jp = jc + 1;
    double fjum = h->FJMU[jc], fjup = h->FJPA[jc];
    h->PMJ[jc] = fjum * fjup * (1.0 / DY[jc] / H[jc]) + (1. - fjup) * (1.0 / DY[N2M] / H[N2M]);
    h->PCJ[jc] = fjum * fjup * (-1.0 / DY[jc] / H[jc] - 1.0 / DY[jc] / H[jp]) +
                 (1. - fjup) * (-1.0 / DY[N2M] / H[N2M]) + (1. - fjum) * (-1.0 / DY[1] / H[2]);
    h->PPJ[jc] = fjum * fjup * (1.0 / DY[jc] / H[jp]) + (1. - fjum) * (1.0 / DY[1] / H[2]);
  }
}

// ADI y chains (GETUH1/2/3 coefficient loops :8261-8267, :8389-8397,
// :8524-8529 and TDMAI :8637-8650), all J-only; plus the x/z ADI factors
// (:8294-8296, :8324-8326).  Depend on DT.
static void build_dt_terms(c3d_handle* h, std::vector<double>& ych) {
  const int N2 = h->cfg.n2, N2M = N2 - 1;
  const double RE = h->c.s.re, DT = h->c.s.dt;
  ych.assign(9 * (size_t)(N2 + 2), 0.0);
  for (int comp = 0; comp < 3; comp++) {
    double* A = &ych[(3 * comp + 0) * (size_t)(N2 + 2)];
    double* BET = &ych[(3 * comp + 1) * (size_t)(N2 + 2)];
    double* GAM = &ych[(3 * comp + 2) * (size_t)(N2 + 2)];
    const int ns = (comp == 1) ? 2 : 1;
    std::vector<double> ap(N2 + 2), ac(N2 + 2), am(N2 + 2);
    for (int j = ns; j <= N2M; j++) {
      double fjup = h->FJPA[j];
      if (comp == 1) {
        double fjum = h->FJMV[j];
        ap[j] = fjup * (-(0.5 * h->DYP[j] * (1. / RE))) * DT;
        ac[j] = 1.0 + (+0.5 * h->DYC[j] * (1. / RE)) * DT;
        am[j] = fjum * (-(0.5 * h->DYM[j] * (1. / RE))) * DT;
      } else {
        double fjum = h->FJMU[j];
        ap[j] = fjup * (-(0.5 * h->HP[j] * (1. / RE))) * DT;
        ac[j] = 1.0 + (0.5 * h->HC[j] * (1. / RE)) * DT;
        am[j] = fjum * (-(0.5 * h->HM[j] * (1. / RE))) * DT;
      }
    }
    double bet = 1.0 / ac[ns];
    BET[ns] = bet;
    GAM[ns] = ap[ns] * bet;
    A[ns] = am[ns];
    for (int j = ns + 1; j <= N2M; j++) {
      bet = 1.0 / (ac[j] - am[j] * GAM[j - 1]);
      BET[j] = bet;
      GAM[j] = ap[j] * bet;
      A[j] = am[j];
    }
  }
  Ops& o = h->c.o;
  const double dx1q = h->c.s.dx1q, dx3q = h->c.s.dx3q;
  o.adix_a = 1.0 - DT / 2.0 * (1. / RE) * 12.0 * dx1q * (1.0);
  o.adix_b = 10.0 - DT / 2.0 * (1. / RE) * 12.0 * dx1q * (-2.0);
  o.adiz_a = 1.0 - DT / 2.0 * (1. / RE) * 12.0 * dx3q * (1.0);
  o.adiz_b = 10.0 - DT / 2.0 * (1. / RE) * 12.0 * dx3q * (-2.0);
  o.adix = make_solve(o.adix_a, o.adix_b);
  o.adiz = make_solve(o.adiz_a, o.adiz_b);
  o.px = make_opset(o.lx.seg, o.adix_a, o.adix_b);
  o.pz = make_opset(o.lz.seg, o.adiz_a, o.adiz_b);
  o.pz32 = make_opset(o.lz32.seg, o.adiz_a, o.adiz_b);
}

// coefficient table of the fused RHS-y / ADI-y sweep (layout in c3d_impl.h)
static void build_ytab(c3d_handle* h, const std::vector<double>& ych, std::vector<double>& tab) {
  const int N2 = h->cfg.n2, N2M = N2 - 1;
  const size_t nv = (size_t)(N2 + 2);
  tab.assign(3 * nv * YTAB_W, 0.0);
  for (int comp = 0; comp < 3; comp++) {
    const double* A = &ych[(3 * comp + 0) * nv];
    const double* BET = &ych[(3 * comp + 1) * nv];
    const double* GAM = &ych[(3 * comp + 2) * nv];
    const int ns = (comp == 1) ? 2 : 1;
    for (int J = 1; J <= N2M; J++) {
      double* t = &tab[(comp * nv + J) * YTAB_W];
      if (comp == 1) {
        t[0] = h->DYP[J];
        t[1] = h->DYC[J];
        t[2] = h->DYM[J];
        t[3] = h->H[J];
        t[4] = 1. / h->H[J];
        t[5] = 1. / h->DY[J];  // the SGS term's 1/DY(J), 1/DY(J-1)
        t[6] = J >= 2 ? 1. / h->DY[J - 1] : 0.0;
      } else {
        t[0] = h->HP[J];
        t[1] = h->HC[J];
        t[2] = h->HM[J];
        t[3] = h->FJPA[J] * 1.0 / h->H[J + 1];
        t[4] = h->FJMU[J] * 1.0 / h->H[J];
        t[5] = h->DY[J - 1] / 2.;
        t[6] = h->DY[J] / 2.;
        t[7] = h->DY[J + 1] / 2.;
        t[8] = 1. / h->DY[J];
        t[9] = 1. - h->FJPA[J];
        t[10] = 1. - h->FJMU[J];
        t[14] = 1. / h->H[J];  // the SGS terms' 1/H(J), 1/H(J+1)
        t[15] = 1. / h->H[J + 1];
      }
      t[YT_BET] = BET[J];
      t[YT_M] = (J == ns) ? 0.0 : -A[J] * BET[J];
      t[YT_G] = (J == N2M) ? 0.0 : -GAM[J];
    }
  }
}

static int upload_dt_terms(c3d_handle* h) {
  std::vector<double> ych, tab;
  build_dt_terms(h, ych);
  double* base = (double*)h->c.m.ya[0];
  HIPCHK(hipMemcpyAsync(base, ych.data(), ych.size() * sizeof(double), hipMemcpyHostToDevice, h->c.st));
  build_ytab(h, ych, tab);
  HIPCHK(hipMemcpyAsync(h->dytab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, h->c.st));
  std::vector<float> tab32(tab.begin(), tab.end());
  HIPCHK(hipMemcpyAsync(h->dytab32, tab32.data(), tab32.size() * sizeof(float), hipMemcpyHostToDevice, h->c.st));
  // the host vectors go out of scope: wait for the copies
  HIPCHK(hipStreamSynchronize(h->c.st));
  return 0;
}

template <typename T>
static int dalloc(c3d_handle* h, T** p, size_t n) {
  // Successive allocations start at different 256-byte phases (0..63 x 256
  // bytes, stepping by 37): fields of equal size would otherwise all begin
  // at the same large alignment, and the same element of every field (or of
  // every channel of an ensemble) would fall on the same HBM channel when
  // they are streamed together.
  static std::atomic<unsigned> seq{0};
  const size_t pad = 256 * (size_t)((seq.fetch_add(1) * 37u) % 64u);
  void* q = nullptr;
  HIPCHK(hipMalloc(&q, n * sizeof(T) + 64 + pad));
  // on the handle's stream: it is non-blocking, so the null stream would not
  // order the fill before the kernels and copies that follow on it
  HIPCHK(hipMemsetAsync(q, 0, n * sizeof(T) + 64 + pad, h->c.st));
  h->allocs.push_back(q);
  *p = (T*)((char*)q + pad);
  return 0;
}

long long c3d_footprint_bytes(const c3d_config* cfg) {
  if (!cfg) return -1;
  long long nx = cfg->n1 - 1, nz = cfg->n3 - 1, nj = cfg->n2 + 1, n2m = cfg->n2 - 1;
  long long fsz = nx * nj * nz;
  const bool mixed = cfg->precision == C3D_MIXED;
  const long long es = mixed ? 4 : 8;
  // FP64: state 4, DP, RDP, SGSVIS, DIV, 2 scratch (+ second V in mixed
  // mode); intermediates 10 + HH 6; FP32 shadows 4 and DIV in mixed mode;
  // LES 7
  long long bytes = (10 + (mixed ? 1 : 0)) * fsz * 8 + 16 * fsz * es + (mixed ? 5 * fsz * 4 : 0);
  if (cfg->files != 0.0) bytes += 7 * fsz * es + (mixed ? fsz * 4 : 0);
  if (cfg->files == 1.0) bytes += (2 * n2m * ((nx + 15) / 16) * ((nz + 15) / 16) + n2m + 3) * 8;
  if (cfg->files == 3.0) bytes += (10 * n2m * ((nx + 15) / 16) * ((nz + 15) / 16) + 2 * (n2m + 3)) * 8;
  if (cfg->wmodel == 2.0) bytes += (long long)IW_NPL * nx * nz * 8;
  const long long nxhp = (nx / 2 + 1 + 15) / 16 * 16, nxsp = spec_pitch(nx / 2);
  bytes += nxsp * nz * n2m * 16 + nxhp * (nz / 2 + 1) * n2m * 16;  // spectrum + TDMAP factors
  bytes += 4 * std::max(n2m * nz, nx * n2m) * 8 + 128 * 8;  // flow-rate partial sums, max slots
  return bytes + 4 * nx * nz * 8 + (1 << 20);
}

// ------------------------------------------------------ tuned tile shapes
// The line-group and tile shapes of the step's kernels: x-line tiles (lx),
// FP64 and FP32 z-line tiles (lz, lz32), FP64/FP32 fused y-sweep tiles (ys,
// ys32), Poisson y-solve tiles (yp) and the FFT tiles (tlx, tlz).  c3d_create
// derives them from the grid by rule; c3d_autotune times candidates and
// records the fastest in a process-wide table keyed by the configuration,
// which later c3d_create calls of the same key apply.
struct Shapes {
  LShape lx, lz, lz32;
  YShape ys, ys32, yp;
  int tlx, tlz;
};
static std::mutex g_tune_mu;
static std::map<std::string, Shapes> g_tuned;

static std::string tune_key(const c3d_config& c) {
  char b[96];
  snprintf(b, sizeof b, "%d:%d:%d:%d:%d", c.n1, c.n2, c.n3, c.precision, c.files != 0.0 ? 1 : 0);
  return b;
}

static Shapes get_shapes(const c3d_handle* h) {
  const Ops& o = h->c.o;
  return Shapes{o.lx, o.lz, o.lz32, o.ys, o.ys32, o.yp, h->c.f.tlx, h->c.f.tlz};
}

// installs shapes and the segment-dependent periodic solves that go with them
static void apply_shapes(c3d_handle* h, const Shapes& t) {
  Ops& o = h->c.o;
  o.lx = t.lx;
  o.lz = t.lz;
  o.lz32 = t.lz32;
  o.ys = t.ys;
  o.ys32 = t.ys32;
  o.yp = t.yp;
  h->c.f.tlx = t.tlx;
  h->c.f.tlz = t.tlz;
  o.px = make_opset(o.lx.seg, o.adix_a, o.adix_b);
  o.pz = make_opset(o.lz.seg, o.adiz_a, o.adiz_b);
  o.pz32 = make_opset(o.lz32.seg, o.adiz_a, o.adiz_b);
}

extern "C" int c3d_create(const c3d_config* cfg, c3d_handle** out) {
  g_create_err.clear();
  if (!cfg || !out) {
    g_create_err = "null argument";
    return -1;
  }
  *out = nullptr;
  c3d_handle* h = new c3d_handle();
  h->cfg = *cfg;
  auto bad = [&](const std::string& m) {
    g_create_err = m;
    for (void* p : h->allocs) (void)hipFree(p);
    if (h->hread) (void)hipHostFree(h->hread);
    delete h;
    return -1;
  };
  const c3d_config& cf = h->cfg;
  if (cf.n1 < 9 || cf.n2 < 5 || cf.n3 < 9) return bad("grid too small (need N1,N3 >= 9, N2 >= 5)");
  if (cf.precision != C3D_FP64 && cf.precision != C3D_MIXED) return bad("precision must be C3D_FP64 or C3D_MIXED");
  if (cf.files != 0.0 && cf.files != 1.0 && cf.files != 2.0 && cf.files != 3.0)
    return bad("FILES must be 0 (DNS), 1 (dynamic model), 2 (Smagorinsky) or 3 (scale-dependent dynamic model)");
  if (cf.wmodel != 0.0 && cf.wmodel != 1.0 && cf.wmodel != 2.0 && cf.wmodel != 4.0 && cf.wmodel != 6.0)
    return bad("WMODEL must be 0 (resolved walls), 1 (log law, bottom wall), 2 (integral wall model, bottom "
               "wall), 4 (Werner-Wengle, both walls) or 6 (log law, both walls)");
  // WALLMODEL_WERNER takes its I neighbour from IPA(K) (compact3d.f:4849)
  if (cf.wmodel == 4.0 && cf.n3 > cf.n1) return bad("WMODEL = 4 needs N3 <= N1");
  // the integral wall model's first time-filter weight is CFLMAX*ALX/N1M/UINIT
  // over its viscous time scale (compact3d.f:3700): a zero CFLMAX would
  // freeze its filtered inputs at their initial values
  if (cf.wmodel == 2.0 && !(cf.cflmax > 0)) return bad("WMODEL = 2 needs cflmax > 0 (CFLMAX of /TSTEP/)");
  if (cf.re <= 0 || cf.dt <= 0 || cf.alx <= 0 || cf.alz <= 0) return bad("RE, DT, ALX, ALZ must be > 0");
  if (hipSetDevice(cf.device) != hipSuccess) return bad("hipSetDevice failed");
  {
    std::string serr;
    if (scratch_init(cf.device, serr)) return bad(serr);
  }
  Ctx& c = h->c;
  c.kt = &h->kt;
  Geo& g = c.g;
  g.nx = cf.n1 - 1;
  g.nz = cf.n3 - 1;
  g.n2 = cf.n2;
  g.n2m = cf.n2 - 1;
  g.nj = cf.n2 + 1;
  g.nxh = g.nx / 2 + 1;
  g.nxhp = (g.nxh + 15) / 16 * 16;
  g.nxs = g.nx / 2;
  g.nxsp = spec_pitch(g.nxs);
  g.plane = (long long)g.nj * g.nx;
  g.fsz = g.plane * g.nz;
  if (g.nx % 2 || g.nz % 2) return bad("N1-1 and N3-1 must be even");
  // x-line staging addresses elements with 32-bit offsets
  if (g.fsz + 2 * g.plane >= (1ll << 31)) return bad("grid too large: a field must hold fewer than 2^31 elements");
  FFTPlan& f = c.f;
  f.nx = g.nx;
  f.nz = g.nz;
  if (!factor_fft(g.nx, f.radx, f.nradx) || !factor_fft(g.nz, f.radz, f.nradz))
    return bad("N1-1 and N3-1 must factor into 2, 3 and 5");
  f.codex = f.codez = 0;
  for (int q = 0; q < f.nradx; q++) f.codex |= f.radx[q] << (3 * q);
  for (int q = 0; q < f.nradz; q++) f.codez |= f.radz[q] << (3 * q);
  f.tlx = std::max(1, 1024 / g.nx);  // x-FFT tiles of <= 1024 values on 256 threads
  f.tlz = std::max(1, 2048 / g.nz);
  f.tly = std::max(1, std::min(32, pow2_floor((int)(65536 / (24 * (size_t)g.n2m)))));
  if (g.n2m > 3072) return bad("N2 too large for the y tiles (N2-1 <= 3072)");
  build_mesh(h);
  Ops& o = c.o;
  if (!make_lshape(g.nx, true, sizeof(double), o.lx) || !make_lshape(g.nz, false, sizeof(double), o.lz))
    return bad("N1-1 (N3-1) too long for the line tiles: at most 512 (4096) points");
  if (!make_lshape32(g.nz, o.lz32)) return bad("N3-1 too long for the FP32 line tiles");
  // FP64 y-sweep: 6-point chunks (every operand of a chunk is loaded before
  // its arithmetic), 16 columns (128-byte rows), up to 512 threads
  o.ys = make_yshape(g.n2m, 512, 16, 6);
  // FP32 y-sweep: the same tiles over column pairs (8-byte loads, 32
  // columns per 16 threads: 128-byte rows)
  o.ys32 = make_yshape_pair(g.n2m, 512);
  // Poisson: 16 columns (8 modes of 16 bytes per row at the two kz of a
  // pair) and chunks of 6 points keep its three FP64 chunk arrays small;
  // 12-point chunks beyond N2-1 = 1536 (at least 2 columns up to 3072)
  o.yp = g.n2m <= 6 * 256 ? make_yshape(g.n2m, 512, 16, 6) : make_yshape(g.n2m, 512, 16);
  o.s6 = make_solve(1.0, 6.0);
  o.s10 = make_solve(1.0, 10.0);
  o.s22 = make_solve(1.0, 22.0);
  o.tx = make_tile(g.nx, 2);
  o.tz = make_tile(g.nz, 2);
  o.tx4 = make_tile(g.nx, 4);
  o.tz4 = make_tile(g.nz, 4);

  if (cf.stream) {
    c.st = (hipStream_t)cf.stream;
  } else {
    if (hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking) != hipSuccess) return bad("stream create failed");
    h->own_stream = true;
  }
  // device vectors: 15 J-vectors + AK1 + AK3 + 9 ADI chains
  const size_t nv = (size_t)(g.n2 + 2);
  const size_t nvec = 15 * nv + (g.nx + 2) + (g.nz / 2 + 3) + 9 * nv + 5 * nv;
  if (dalloc(h, &h->dvec, nvec)) return bad(h->err);
  {
    std::vector<double> hv(nvec, 0.0);
    std::vector<double>* src[15] = {&h->Y,   &h->DY,  &h->H,    &h->HP,   &h->HC,
                                    &h->HM,  &h->DYP, &h->DYC,  &h->DYM,  &h->FJPA,
                                    &h->FJMU, &h->FJMV, &h->PMJ, &h->PCJ, &h->PPJ};
    for (int q = 0; q < 15; q++) std::copy(src[q]->begin(), src[q]->end(), hv.begin() + q * nv);
    std::copy(h->AK1.begin(), h->AK1.end(), hv.begin() + 15 * nv);
    std::copy(h->AK3.begin(), h->AK3.end(), hv.begin() + 15 * nv + (g.nx + 2));
    {
      double* yt = hv.data() + 15 * nv + (g.nx + 2) + (g.nz / 2 + 3) + 9 * nv;
      for (int J = 1; J <= g.n2m; J++) {
        yt[0 * nv + J] = h->FJPA[J] * 1.0 / h->H[J + 1];
        yt[1 * nv + J] = h->FJMU[J] * 1.0 / h->H[J];
        yt[3 * nv + J] = 1. / h->DY[J];
      }
      for (int J = 0; J <= g.n2; J++) yt[2 * nv + J] = h->DY[J] / 2.;
      for (int J = 1; J <= g.n2; J++) yt[4 * nv + J] = 1. / h->H[J];
    }
    if (hipMemcpyAsync(h->dvec, hv.data(), nvec * sizeof(double), hipMemcpyHostToDevice, c.st) != hipSuccess ||
        hipStreamSynchronize(c.st) != hipSuccess)
      return bad("vector upload failed");
    const double* d = h->dvec;
    Metrics& m = c.m;
    m.y = d + 0 * nv;
    m.dy = d + 1 * nv;
    m.h = d + 2 * nv;
    m.hp = d + 3 * nv;
    m.hc = d + 4 * nv;
    m.hm = d + 5 * nv;
    m.dyp = d + 6 * nv;
    m.dyc = d + 7 * nv;
    m.dym = d + 8 * nv;
    m.fjpa = d + 9 * nv;
    m.fjmu = d + 10 * nv;
    m.fjmv = d + 11 * nv;
    m.pmj = d + 12 * nv;
    m.pcj = d + 13 * nv;
    m.ppj = d + 14 * nv;
    m.ak1 = d + 15 * nv;
    m.ak3 = d + 15 * nv + (g.nx + 2);
    const double* yc = d + 15 * nv + (g.nx + 2) + (g.nz / 2 + 3);
    const double* yt = yc + 9 * nv;
    m.cu2 = yt;
    m.cu1 = yt + nv;
    m.dyh = yt + 2 * nv;
    m.rdy = yt + 3 * nv;
    m.rh = yt + 4 * nv;
    for (int q = 0; q < 3; q++) {
      m.ya[q] = yc + (3 * q + 0) * nv;
      m.ybet[q] = yc + (3 * q + 1) * nv;
      m.ygam[q] = yc + (3 * q + 2) * nv;
    }
  }
  if (dalloc(h, &h->dytab, 3 * (size_t)(g.n2 + 2) * YTAB_W) ||
      dalloc(h, &h->dytab32, 3 * (size_t)(g.n2 + 2) * YTAB_W))
    return bad(h->err);
  c.m.ytab = h->dytab;
  c.m.ytab32 = h->dytab32;
  if (upload_dt_terms(h)) return bad(h->err);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(tune_key(cf));
    if (it != g_tuned.end()) apply_shapes(h, it->second);
  }
  // FFT twiddles
  {
    std::vector<double> a = twiddles(g.nx, -1), b = twiddles(g.nx, +1), cz = twiddles(g.nz, -1),
                        dz = twiddles(g.nz, +1);
    double2 *p1, *p2, *p3, *p4;
    if (dalloc(h, &p1, g.nx) || dalloc(h, &p2, g.nx) || dalloc(h, &p3, g.nz) || dalloc(h, &p4, g.nz))
      return bad(h->err);
    if (hipMemcpyAsync(p1, a.data(), a.size() * 8, hipMemcpyHostToDevice, c.st) != hipSuccess ||
        hipMemcpyAsync(p2, b.data(), b.size() * 8, hipMemcpyHostToDevice, c.st) != hipSuccess ||
        hipMemcpyAsync(p3, cz.data(), cz.size() * 8, hipMemcpyHostToDevice, c.st) != hipSuccess ||
        hipMemcpyAsync(p4, dz.data(), dz.size() * 8, hipMemcpyHostToDevice, c.st) != hipSuccess ||
        hipStreamSynchronize(c.st) != hipSuccess)
      return bad("twiddle upload failed");
    f.twxf = p1;
    f.twxb = p2;
    f.twzf = p3;
    f.twzb = p4;
  }
  // fields
  Fields& F = c.d;
  const size_t fs = (size_t)g.fsz;
  // FP64 always: the state, SGSVIS, DIV and host-transfer scratch
  double** f64[] = {&F.u, &F.v, &F.w, &F.p, &F.sgs, &F.div, &h->dtmp, &h->dtmp2};
  for (auto pp : f64)
    if (dalloc(h, pp, fs)) return bad(h->err);
  // intermediates: FP32 in mixed mode (allocated as doubles of that size)
  const bool mixed = cf.precision == C3D_MIXED;
  c.s.mixed = mixed ? 1 : 0;
  const size_t fsi = mixed ? (fs + 1) / 2 : fs;
  if (dalloc(h, &F.dp, fs) || dalloc(h, &F.rdp, fs)) return bad(h->err);
  double** fin[] = {&F.ruh[0], &F.ruh[1], &F.ruh[2], &F.uh[0], &F.uh[1], &F.uh[2],
                    &F.vv1,    &F.ww1,    &F.uw,     &F.vv3};
  for (auto pp : fin)
    if (dalloc(h, pp, fsi)) return bad(h->err);
  F.vold = nullptr;
  if (mixed) {
    h->vbuf[0] = F.v;
    if (dalloc(h, &h->vbuf[1], fs)) return bad(h->err);
    float** f32[] = {&F.uf, &F.vf, &F.wf, &F.pf, &F.divf};
    for (auto pp : f32)
      if (dalloc(h, pp, fs)) return bad(h->err);
    F.su = F.uf;
    F.sv = F.vf;
    F.sw = F.wf;
    F.sp = F.pf;
  } else {
    F.uf = F.vf = F.wf = F.pf = nullptr;
    F.su = F.u;
    F.sv = F.v;
    F.sw = F.w;
    F.sp = F.p;
  }
  for (int s2 = 0; s2 < 2; s2++)
    for (int q = 0; q < 3; q++)
      if (dalloc(h, &h->hhbuf[s2][q], fsi)) return bad(h->err);
  for (int q = 0; q < 3; q++) {
    F.hh[q] = h->hhbuf[0][q];
    F.hhp[q] = h->hhbuf[1][q];
  }
  const size_t n2d = (size_t)g.nx * g.nz;
  if (dalloc(h, &F.ubc_u, n2d) || dalloc(h, &F.ubc_w, n2d) || dalloc(h, &F.tauw, 2 * n2d) ||
      dalloc(h, &F.taub, 2 * n2d))
    return bad(h->err);
  F.tauwe = nullptr;
  if (cf.wmodel == 4.0 && dalloc(h, &F.tauwe, 4 * n2d)) return bad(h->err);
  F.iwm = nullptr;
  if (cf.wmodel == 2.0 && dalloc(h, &F.iwm, (size_t)IW_NPL * n2d)) return bad(h->err);
  if (dalloc(h, &F.spec, (size_t)g.nxsp * g.nz * g.n2m)) return bad(h->err);
  F.sgsf = nullptr;
  F.ssgs = F.sgs;
  if (cf.files != 0.0) {
    double** les[] = {&F.nux, &F.nuz, &F.nue, &F.dvdx, &F.dwdx, &F.dudz, &F.dvdz};
    for (auto pp : les)
      if (dalloc(h, pp, fsi)) return bad(h->err);
    if (mixed) {
      if (dalloc(h, &F.sgsf, fs)) return bad(h->err);
      F.ssgs = F.sgsf;
    }
  }
  F.dyp = F.smagc = F.sdbeta = nullptr;
  if (cf.files == 1.0 || cf.files == 3.0) {
    const size_t per = cf.files == 3.0 ? 10 : 2;
    if (dalloc(h, &F.dyp, per * (size_t)g.n2m * c3dk::sgsdy_tiles(g)) || dalloc(h, &F.smagc, (size_t)g.n2 + 2))
      return bad(h->err);
    if (cf.files == 3.0 && dalloc(h, &F.sdbeta, (size_t)g.n2 + 2)) return bad(h->err);
  }
  c3dk::init_attrs();
  {
    const size_t nmj = (size_t)g.nxhp * (g.nz / 2 + 1) * g.n2m;
    if (dalloc(h, &F.pib, nmj) || dalloc(h, &F.pgt, nmj)) return bad(h->err);
    c3dk::pois_table(c);
    if (hipStreamSynchronize(c.st) != hipSuccess || hipGetLastError() != hipSuccess)
      return bad("Poisson factor set-up failed");
  }
  if (dalloc(h, &F.sc, SC_COUNT)) return bad(h->err);
  if (hipMemcpyAsync(F.sc + SC_CFLMAX, &cf.cflmax, sizeof(double), hipMemcpyHostToDevice, c.st) != hipSuccess)
    return bad("scalar upload failed");
  if (hipHostMalloc((void**)&h->hread, 2 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    h->hread = nullptr;
    return bad("hipHostMalloc of the read-back slots failed");
  }
  F.nred = 1024;
  if (dalloc(h, &F.red, 4 * (size_t)F.nred)) return bad(h->err);
  // flow-rate partial sums: one entry per block of any line kernel (x tiles
  // hold >= 1 line of N2M*N3M, z tiles >= 1 column of N1M*N2M)
  F.flowr = std::max(g.n2m * g.nz, g.nx * g.n2m);
  if (dalloc(h, &F.flowp, 4 * (size_t)F.flowr)) return bad(h->err);
  if (dalloc(h, &F.mxs, 128)) return bad(h->err);  // zeroed
  if (hipStreamSynchronize(c.st) != hipSuccess) return bad("set-up failed");
  *out = h;
  return 0;
}

static void graphs_free(c3d_handle* h);

static void kt_free(c3d_handle* h) {
  KTimer& t = h->kt;
  for (int q = 0; q < 2 * t.cap; q++) (void)hipEventDestroy(t.ev[q]);
  delete[] t.ev;
  delete[] t.id;
  t.ev = nullptr;
  t.id = nullptr;
  t.cap = t.used = 0;
  t.on = false;
}

extern "C" int c3d_destroy(c3d_handle* h) {
  if (!h) return 0;
  (void)hipStreamSynchronize(h->c.st);
  graphs_free(h);
  kt_free(h);
  for (void* p : h->allocs) (void)hipFree(p);
  if (h->hread) (void)hipHostFree(h->hread);
  if (h->own_stream) (void)hipStreamDestroy(h->c.st);
  delete h;
  return 0;
}

extern "C" const char* c3d_last_error(const c3d_handle* h) {
  if (!h) return g_create_err.c_str();
  return h->err.c_str();
}

// ------------------------------------------------------- layout transfers
// Fortran U-like (0:N1,0:N2,0:N3) index
static inline size_t fu(const c3d_handle* h, int I, int J, int K) {
  return (size_t)I + (size_t)(h->cfg.n1 + 1) * ((size_t)J + (size_t)(h->cfg.n2 + 1) * (size_t)K);
}
static inline size_t fp(const c3d_handle* h, int I, int J, int K) {
  return (size_t)(I - 1) + (size_t)h->cfg.n1 * ((size_t)(J - 1) + (size_t)h->cfg.n2 * (size_t)(K - 1));
}

enum { LAY_U = 0, LAY_P = 1 };

static int to_dev(c3d_handle* h, double* dst, const double* src, int lay) {
  const Geo& g = h->c.g;
  h->stage.assign((size_t)g.fsz, 0.0);
  double* s = h->stage.data();
  if (src) {
    for (int k = 0; k < g.nz; k++)
      for (int j = 0; j <= g.n2; j++) {
        if (lay == LAY_P && (j < 1 || j > g.n2m)) continue;
        double* row = s + ((size_t)k * g.nj + j) * g.nx;
        if (lay == LAY_U)
          for (int i = 0; i < g.nx; i++) row[i] = src[fu(h, i + 1, j, k + 1)];
        else
          for (int i = 0; i < g.nx; i++) row[i] = src[fp(h, i + 1, j, k + 1)];
      }
  }
  HIPCHK(hipMemcpyAsync(dst, s, (size_t)g.fsz * 8, hipMemcpyHostToDevice, h->c.st));
  HIPCHK(hipStreamSynchronize(h->c.st));
  return 0;
}

static int from_dev(c3d_handle* h, double* dst, const double* src, int lay) {
  const Geo& g = h->c.g;
  h->stage.assign((size_t)g.fsz, 0.0);
  double* s = h->stage.data();
  HIPCHK(hipMemcpyAsync(s, src, (size_t)g.fsz * 8, hipMemcpyDeviceToHost, h->c.st));
  HIPCHK(hipStreamSynchronize(h->c.st));
  for (int k = 0; k < g.nz; k++)
    for (int j = 0; j <= g.n2; j++) {
      if (lay == LAY_P && (j < 1 || j > g.n2m)) continue;
      const double* row = s + ((size_t)k * g.nj + j) * g.nx;
      if (lay == LAY_U)
        for (int i = 0; i < g.nx; i++) dst[fu(h, i + 1, j, k + 1)] = row[i];
      else
        for (int i = 0; i < g.nx; i++) dst[fp(h, i + 1, j, k + 1)] = row[i];
    }
  return 0;
}

extern "C" int c3d_upload_state(c3d_handle* h, const double* U, const double* P, const double* HH,
                                const double* SGSVIS, double presg, double presg3) {
  if (!h) return -1;
  if (!U || !P) return fail(h, "U and P are required");
  const size_t nu = (size_t)(h->cfg.n1 + 1) * (h->cfg.n2 + 1) * (h->cfg.n3 + 1);
  Fields& F = h->c.d;
  if (to_dev(h, F.u, U, LAY_U) || to_dev(h, F.v, U + nu, LAY_U) || to_dev(h, F.w, U + 2 * nu, LAY_U))
    return -1;
  if (to_dev(h, F.p, P, LAY_P)) return -1;
  const long long fsz = h->c.g.fsz;
  for (int q = 0; q < 3; q++) {
    if (h->c.s.mixed) {
      if (to_dev(h, h->dtmp, HH ? HH + q * nu : nullptr, LAY_U)) return -1;
      c3dk::to_f32(h->c, h->dtmp, (float*)F.hh[q], fsz);
    } else if (to_dev(h, F.hh[q], HH ? HH + q * nu : nullptr, LAY_U)) {
      return -1;
    }
  }
  if (to_dev(h, F.sgs, SGSVIS, LAY_U)) return -1;
  double sc[2] = {presg, presg3};
  HIPCHK(hipMemcpyAsync(F.sc + SC_PRESG, sc, 2 * sizeof(double), hipMemcpyHostToDevice, h->c.st));
  HIPCHK(hipStreamSynchronize(h->c.st));
  if (h->c.s.mixed) {
    // FP32 shadows of the state, and DIV(U) of the uploaded field (the
    // next RDP starts from it)
    c3dk::to_f32(h->c, F.u, F.uf, fsz);
    c3dk::to_f32(h->c, F.v, F.vf, fsz);
    c3dk::to_f32(h->c, F.w, F.wf, fsz);
    c3dk::to_f32(h->c, F.p, F.pf, fsz);
    if (F.sgsf) c3dk::to_f32(h->c, F.sgs, F.sgsf, fsz);
    c3dk::divcheck(h->c);
    c3dk::to_f32(h->c, F.div, F.divf, fsz);
    HIPCHK(hipStreamSynchronize(h->c.st));
    HIPCHK(hipGetLastError());
  }
  return 0;
}

extern "C" int c3d_download_state(c3d_handle* h, double* U, double* P, double* HH, double* SGSVIS,
                                  double* presg, double* presg3) {
  if (!h) return -1;
  const size_t nu = (size_t)(h->cfg.n1 + 1) * (h->cfg.n2 + 1) * (h->cfg.n3 + 1);
  Fields& F = h->c.d;
  if (U && (from_dev(h, U, F.u, LAY_U) || from_dev(h, U + nu, F.v, LAY_U) ||
            from_dev(h, U + 2 * nu, F.w, LAY_U)))
    return -1;
  if (P && from_dev(h, P, F.p, LAY_P)) return -1;
  if (HH)
    for (int q = 0; q < 3; q++) {
      const double* src = F.hh[q];
      if (h->c.s.mixed) {
        c3dk::to_f64(h->c, (const float*)F.hh[q], h->dtmp, h->c.g.fsz);
        src = h->dtmp;
      }
      if (from_dev(h, HH + q * nu, src, LAY_U)) return -1;
    }
  if (SGSVIS && from_dev(h, SGSVIS, F.sgs, LAY_U)) return -1;
  double sc[2];
  HIPCHK(hipMemcpyAsync(sc, F.sc + SC_PRESG, 2 * sizeof(double), hipMemcpyDeviceToHost, h->c.st));
  HIPCHK(hipStreamSynchronize(h->c.st));
  if (presg) *presg = sc[0];
  if (presg3) *presg3 = sc[1];
  return 0;
}

// ------------------------------------------------------------------ phases
static int launched(c3d_handle* h) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(h, std::string("kernel launch: ") + hipGetErrorString(e));
  return 0;
}

// GETUP prologue (:915-923): the current HH becomes HH2, the new HH is
// written from scratch by the RHS passes
static void swap_hh(c3d_handle* h) {
  h->hh_cur ^= 1;
  for (int q = 0; q < 3; q++) {
    h->c.d.hh[q] = h->hhbuf[h->hh_cur][q];
    h->c.d.hhp[q] = h->hhbuf[h->hh_cur ^ 1][q];
  }
}

extern "C" int c3d_bcond(c3d_handle* h) {
  if (!h) return -1;
  c3dk::bcond(h->c);
  return launched(h);
}

extern "C" int c3d_rhs(c3d_handle* h) {
  if (!h) return -1;
  swap_hh(h);
  c3dk::rhs(h->c);
  return launched(h);
}

extern "C" int c3d_getuh(c3d_handle* h) {
  if (!h) return -1;
  c3dk::getuh(h->c);
  return launched(h);
}

extern "C" int c3d_dpcalc(c3d_handle* h) {
  if (!h) return -1;
  c3dk::rhsdp(h->c);
  c3dk::getdp(h->c);
  return launched(h);
}

// mixed mode: the update writes V into the other buffer
static void swap_v(c3d_handle* h) {
  h->c.d.vold = h->c.d.v;
  h->vcur ^= 1;
  h->c.d.v = h->vbuf[h->vcur];
}

// DIVMAX above which mixed mode refines the projection (the FP64 path
// reaches ~1e-13 on the same fields)
static const double kRefineTol = 1e-10;

static int read_sc(c3d_handle* h, int slot, double* v);

// UPCALC (in mixed mode V goes to the other buffer)
static void upcalc_impl(c3d_handle* h, bool chk) {
  if (h->c.s.mixed) swap_v(h);
  c3dk::upcalc(h->c, chk);
}

// DIVMAX and CFLM of the last step (adjacent device slots) in one copy to
// pinned host memory; either pointer may be null
static int read_checks(c3d_handle* h, double* divmax, double* cflm) {
  static_assert(SC_CFLM == SC_DIVMAX + 1, "DIVMAX and CFLM slots must be adjacent");
  HIPCHK(hipMemcpyAsync(h->hread, h->c.d.sc + SC_DIVMAX, 2 * sizeof(double), hipMemcpyDeviceToHost, h->c.st));
  HIPCHK(hipStreamSynchronize(h->c.st));
  if (divmax) *divmax = h->hread[0];
  if (cflm) *cflm = h->hread[1];
  return 0;
}

// C3D_ERR_NONFINITE when a step's DIVMAX or CFLM is NaN or infinite: the
// step ran, but the state has diverged
static int check_finite(c3d_handle* h, double dm, double cf) {
  if (std::isfinite(dm) && std::isfinite(cf)) return 0;
  fail(h, "the state diverged: DIVMAX = " + std::to_string(dm) + ", CFLM = " + std::to_string(cf));
  return C3D_ERR_NONFINITE;
}

// Mixed mode, after UPCALC, given the DIVMAX dm of the projection (FP64, on
// the new field): above kRefineTol one FP64-residual refinement follows: DIV
// is the residual of the Poisson solve, DP2 solves L DP2 = DIV/DT the same
// way and the field takes U -= DT grad(DP2), P += DP2 (flow rate unchanged:
// the periodic gradients have zero line means).  *divmax / *cflm return the
// final DIVMAX and CFLM (unchanged when no refinement ran).
static int refine_from(c3d_handle* h, double dm, double* divmax, double* cflm) {
  h->first_divmax = dm;
  // a NaN or infinite DIVMAX is reported, not refined
  if (dm > kRefineTol && std::isfinite(dm)) {
    h->refines++;
    c3dk::div_rdp(h->c);
    c3dk::getdp(h->c);
    swap_v(h);
    c3dk::upd_x(h->c, true, true);
    c3dk::upd_z(h->c, true, true);
    if (read_checks(h, &dm, cflm)) return -1;
  }
  if (divmax) *divmax = dm;
  return 0;
}

// the same, reading DIVMAX first
static int refine_impl(c3d_handle* h, double* divmax) {
  double dm = 0.0;
  if (read_checks(h, &dm, nullptr)) return -1;
  return refine_from(h, dm, divmax, nullptr);
}

extern "C" int c3d_upcalc(c3d_handle* h) {
  if (!h) return -1;
  upcalc_impl(h, false);
  if (launched(h)) return -1;
  return h->c.s.mixed ? refine_impl(h, nullptr) : 0;
}

extern "C" int c3d_sgs(c3d_handle* h) {
  if (!h) return -1;
  if (!h->c.s.files) return fail(h, "the SGS model needs FILES != 0");
  if (h->cfg.files == 1.0)
    c3dk::sgsdy(h->c);
  else if (h->cfg.files == 3.0)
    c3dk::sgssd(h->c);
  else
    c3dk::sgscs(h->c);
  return launched(h);
}

// WALL_MODEL (compact3d.f:1981-2013): WMODEL=0 -> WALLSS, IFWALL=0;
// WMODEL=1 (Y0=0) -> WALLMODEL_INSTANT, 2 (Y0=0) -> WALLMODEL_IWMLES_SMOOTH,
// 4 -> WALLMODEL_WERNER, IFWALL=1.  ntime: GETUP's NTIME (IWMLES counts
// its steps by it); <= 0 continues from the last call's.
static void wall_model_impl(c3d_handle* h, int ntime) {
  if (h->cfg.wmodel == 2.0) {
    c3dk::wallmodel_iwmles(h->c, ntime);
    h->c.s.ifwall = 1;
  } else if (h->cfg.wmodel == 1.0 || h->cfg.wmodel == 6.0) {
    c3dk::wallmodel(h->c, h->cfg.wmodel == 6.0);
    h->c.s.ifwall = 1;
  } else if (h->cfg.wmodel == 4.0) {
    c3dk::wallmodel_werner(h->c);
    h->c.s.ifwall = 1;
  } else {
    c3dk::wallss(h->c);
    h->c.s.ifwall = 0;
  }
}

extern "C" int c3d_wall_model(c3d_handle* h) {
  if (!h) return -1;
  wall_model_impl(h, 0);
  return launched(h);
}

extern "C" int c3d_wallss(c3d_handle* h) {
  if (!h) return -1;
  c3dk::wallss(h->c);
  return launched(h);
}

// the kernels of GETUP (no host synchronisation)
static int step_impl(c3d_handle* h, bool chk, int ntime) {
  swap_hh(h);
  // GETUP (:925-929): SGS model, then the wall model
  if (h->cfg.files == 1.0) c3dk::sgsdy(h->c);
  if (h->cfg.files == 2.0) c3dk::sgscs(h->c);
  if (h->cfg.files == 3.0) c3dk::sgssd(h->c);
  if (h->cfg.files != 0.0) wall_model_impl(h, ntime);
  c3dk::bcond(h->c);
  c3dk::rhs_xz(h->c);
  c3dk::getuh_fused(h->c);
  c3dk::rhsdp(h->c);
  c3dk::getdp(h->c);
  upcalc_impl(h, chk);
  return launched(h);
}

extern "C" int c3d_getup(c3d_handle* h, int ntime) {
  if (!h) return -1;
  if (step_impl(h, false, ntime)) return -1;
  return h->c.s.mixed ? refine_impl(h, nullptr) : 0;
}

static int read_sc(c3d_handle* h, int slot, double* v) {
  HIPCHK(hipMemcpyAsync(v, h->c.d.sc + slot, sizeof(double), hipMemcpyDeviceToHost, h->c.st));
  HIPCHK(hipStreamSynchronize(h->c.st));
  return 0;
}

// ------------------------------------------------------- multi-step driving
// The kernels of one MAIN-loop step (WALLSS first when WMODEL=0 and
// FILES != 0, then GETUP with DIV, DIVMAX and CFL fused), with no host
// synchronisation and no mixed-mode refinement check.
static void step_launch(c3d_handle* h) {
  if (h->cfg.wmodel == 0.0 && h->cfg.files != 0.0) c3dk::wallss(h->c);
  swap_hh(h);
  if (h->cfg.files == 1.0) c3dk::sgsdy(h->c);
  if (h->cfg.files == 2.0) c3dk::sgscs(h->c);
  if (h->cfg.files == 3.0) c3dk::sgssd(h->c);
  if (h->cfg.files != 0.0) wall_model_impl(h, 0);
  c3dk::bcond(h->c);
  c3dk::rhs_xz(h->c);
  c3dk::getuh_fused(h->c);
  c3dk::rhsdp(h->c);
  c3dk::getdp(h->c);
  upcalc_impl(h, true);
}

static int buf_key(const c3d_handle* h) { return h->hh_cur * 2 + h->vcur; }

static void graphs_free(c3d_handle* h) {
  for (int p = 0; p < 2; p++) {
    if (h->gexec[p]) (void)hipGraphExecDestroy(h->gexec[p]);
    h->gexec[p] = nullptr;
    h->gkey[p] = -1;
  }
}

// Each step swaps the HH sets (and, in mixed mode, the V buffers), so the
// kernel arguments repeat every two steps: two graphs, captured from the
// current state; capturing advances the host-side buffer state by two steps,
// which returns it to where it was.
static int graphs_capture(c3d_handle* h) {
  graphs_free(h);
  const bool timing = h->kt.on;
  h->kt.on = false;  // no event nodes in the graphs
  for (int p = 0; p < 2; p++) {
    const int key = buf_key(h);
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(h->c.st, hipStreamCaptureModeThreadLocal));
    step_launch(h);
    HIPCHK(hipStreamEndCapture(h->c.st, &g));
    const hipError_t e = hipGraphInstantiate(&h->gexec[p], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      h->kt.on = timing;
      graphs_free(h);
      return fail(h, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    h->gkey[p] = key;
  }
  h->kt.on = timing;
  return 0;
}

extern "C" int c3d_advance(c3d_handle* h, int nsteps, int use_graph) {
  if (!h) return -1;
  if (nsteps < 0) return fail(h, "nsteps must be >= 0");
  for (int n = 0; n < nsteps; n++) {
    if (!use_graph) {
      step_launch(h);
      continue;
    }
    int p = (h->gkey[0] == buf_key(h)) ? 0 : (h->gkey[1] == buf_key(h)) ? 1 : -1;
    if (p < 0) {
      if (graphs_capture(h)) return -1;
      p = 0;
    }
    HIPCHK(hipGraphLaunch(h->gexec[p], h->c.st));
    // the host-side buffer state follows the step the graph performed
    swap_hh(h);
    if (h->c.s.mixed) swap_v(h);
  }
  return launched(h);
}

extern "C" int c3d_sync(c3d_handle* h, double* divmax, double* cflm) {
  if (!h) return -1;
  double dm = 0.0, cf = 0.0;
  if (read_checks(h, &dm, &cf)) return -1;
  if (h->c.s.mixed && refine_from(h, dm, &dm, &cf)) return -1;
  if (divmax) *divmax = dm;
  if (cflm) *cflm = cf;
  return check_finite(h, dm, cf);
}

// ------------------------------------------------------------- autotuner
// Candidate shapes per knob, each within what its kernels assume: x tiles a
// multiple of xline_step with L*G <= 256 and two FP64 staging tiles in
// 64 KB; z tiles L*G <= 512 with a compiled segment (FP64 8/2, FP32
// 16/8/2); y tiles of <= 12-point chunks and <= 512 threads (Poisson
// <= 256); FFT tiles of <= 4096 complex values and <= 64 KB of LDS (the
// tile, its padding and the twiddle table), on 256 threads up to 1024
// values.
static std::vector<LShape> cand_lx(int n) {
  std::vector<LShape> v;
  const int seg = (n % 8 == 0) ? 8 : 2, G = n / seg, step = xline_step(seg, G);
  for (int L = step; L * G <= 256 && L <= 64; L += step)
    if (8 * (2 * (size_t)n * (L + 1) + 4 * (size_t)G * L) <= 65536) v.push_back(LShape{n, L, G, seg});
  return v;
}
// wide: the FP64 z tiles of mixed mode, where only k_upd_z runs them; its
// 128 VGPRs also fit 1024-thread tiles (8-point segments, 32 or 64
// columns: 256-byte FP64 and 128-byte FP32 rows)
static std::vector<LShape> cand_lz(int n, bool fp32, bool wide = false) {
  std::vector<LShape> v;
  const int segs64[] = {8, 2}, segs32[] = {16, 8, 2};
  const int* segs = fp32 ? segs32 : segs64;
  const int ns = fp32 ? 3 : 2;
  for (int q = 0; q < ns; q++) {
    const int seg = segs[q];
    if (seg == 2 && !v.empty()) break;  // 2-point segments only when nothing longer fits
    if (n % seg) continue;
    const int G = n / seg;
    for (int L : {16, 32, 64})
      if (L * G <= 512 || (wide && !fp32 && seg == 8 && L * G <= 1024)) v.push_back(LShape{n, L, G, seg});
  }
  return v;
}
// pair: also the column-pair shapes (FP32)
static std::vector<YShape> cand_y(int m, int row, int ntmax, bool pair = false) {
  std::vector<YShape> v;
  auto add = [&](const YShape& y, int chmax) {
    if (y.L * y.C > ntmax || (m + y.C - 1) / y.C > chmax) return;
    bool dup = false;
    for (const YShape& w : v) dup = dup || (w.L == y.L && w.C == y.C && w.P == y.P);
    if (!dup) v.push_back(y);
  };
  for (int ch : {6, 12})
    for (int nt : {256, 512, 1024}) {
      if (nt > ntmax) continue;
      add(make_yshape(m, nt, row, ch), 12);
    }
  if (pair)
    for (int nt : {512, 1024})
      if (nt <= ntmax) add(make_yshape_pair(m, nt), 6);
  return v;
}
// Poisson y-solve tiles: an even number of columns (modes x the two kz of
// a pair), <= 12-point chunks in <= 512 threads or 6-point chunks in 1024
static std::vector<YShape> cand_yp(int m) {
  std::vector<YShape> yp;
  for (int ch : {6, 12})
    for (int nt : {512, 1024})
      for (int row : {16, 32}) {
        if (ch > 6 && nt > 512) continue;
        const YShape y = make_yshape(m, nt, row, ch);
        bool dup = false;
        for (const YShape& w : yp) dup = dup || (w.L == y.L && w.C == y.C);
        if (!dup && y.L % 2 == 0 && y.L * y.C <= nt && (m + y.C - 1) / y.C <= ch) yp.push_back(y);
      }
  return yp;
}
// LDS a workgroup can address on gfx950 (160 KB)
static const size_t kLdsMax = 160 * 1024;
static std::vector<int> cand_tl(int n, bool padded) {
  std::vector<int> v;
  for (int t = 1; t <= 16; t *= 2)
    if (n * t <= 4096 && 16 * (size_t)n * (t + (padded ? 2 : 1)) <= kLdsMax) v.push_back(t);
  return v;
}

// ms per step of n steps (MAIN's loop: GETUP with DIV, DIVMAX and CFL),
// queued back to back on the handle's stream
static double time_steps(c3d_handle* h, int n) {
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return 1e30;
  (void)hipEventRecord(a, h->c.st);
  for (int q = 0; q < n; q++) step_launch(h);
  (void)hipEventRecord(b, h->c.st);
  float ms = 0.f;
  const bool ok = hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess &&
                  hipGetLastError() == hipSuccess;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ok ? (double)ms / n : 1e30;
}

// best of two timings after one warm-up step
static double trial(c3d_handle* h, const Shapes& s, int n) {
  apply_shapes(h, s);
  step_launch(h);
  return std::min(time_steps(h, n), time_steps(h, n));
}

extern "C" int c3d_autotune(const c3d_config* cfg, int steps, int* shapes, double* ms) {
  g_create_err.clear();
  if (!cfg) return -1;
  if (steps < 1) steps = 1;
  // a scratch instance of the configuration, from the rule-derived shapes
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.erase(tune_key(*cfg));
  }
  c3d_handle* h = nullptr;
  if (c3d_create(cfg, &h)) return -1;
  // laminar Poiseuille state u = 1 - y^2 (the step's work does not depend
  // on the values; this keeps every model's iterations in their usual range)
  {
    const int n1 = cfg->n1, n2 = cfg->n2, n3 = cfg->n3;
    const size_t nu = (size_t)(n1 + 1) * (n2 + 1) * (n3 + 1);
    std::vector<double> U(3 * nu, 0.0), P((size_t)n1 * n2 * n3, 0.0);
    const double aly = h->c.s.aly;
    for (int K = 1; K < n3; K++)
      for (int J = 1; J < n2; J++)
        for (int I = 1; I < n1; I++) {
          const double yc = 0.5 * (h->Y[J] + h->Y[J + 1]) - aly / 2.0;
          U[fu(h, I, J, K)] = 1.0 - yc * yc;
        }
    if (c3d_upload_state(h, U.data(), P.data(), nullptr, nullptr, 0.0, 0.0)) {
      g_create_err = "autotune: " + h->err;
      c3d_destroy(h);
      return -1;
    }
  }
  const Geo& g = h->c.g;
  const bool mixed = h->c.s.mixed != 0;
  Shapes best = get_shapes(h);
  const double t0 = trial(h, best, steps);
  double tb = t0;
  auto knob = [&](auto cands, auto set) {
    for (const auto& cnd : cands) {
      Shapes s = best;
      set(s, cnd);
      const double t = trial(h, s, steps);
      if (t < tb) {
        tb = t;
        best = s;
      }
    }
  };
  if (mixed)
    knob(cand_y(g.n2m, 32, 1024, true), [](Shapes& s, const YShape& y) { s.ys32 = y; });
  else
    knob(cand_y(g.n2m, 16, 1024), [](Shapes& s, const YShape& y) { s.ys = y; });
  if (mixed) knob(cand_lz(g.nz, true), [](Shapes& s, const LShape& z) { s.lz32 = z; });
  knob(cand_lz(g.nz, false, mixed), [](Shapes& s, const LShape& z) { s.lz = z; });
  knob(cand_lx(g.nx), [](Shapes& s, const LShape& x) { s.lx = x; });
  knob(cand_yp(g.n2m), [](Shapes& s, const YShape& y) { s.yp = y; });
  knob(cand_tl(g.nx, true), [](Shapes& s, int t) { s.tlx = t; });
  knob(cand_tl(g.nz, false), [](Shapes& s, int t) { s.tlz = t; });
  c3d_destroy(h);
  if (!(tb < 1e29)) {
    g_create_err = "autotune: timing failed";
    return -1;
  }
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[tune_key(*cfg)] = best;
  }
  if (shapes) {
    const int v[C3D_NSHAPES] = {best.lx.L, best.lz.L, best.lz32.L, best.lz32.seg, best.ys.L, best.ys.C,
                                best.ys32.L, best.ys32.C, best.yp.L, best.yp.C, best.tlx, best.tlz,
                                best.ys32.P};
    for (int q = 0; q < C3D_NSHAPES; q++) shapes[q] = v[q];
  }
  if (ms) {
    ms[0] = t0;
    ms[1] = tb;
  }
  return 0;
}

// Registers given tile shapes (c3d_autotune's 12-value layout) for a
// configuration, as if the autotuner had chosen them, after checking that
// each is the rule-derived shape or one of the autotuner's candidates.
extern "C" int c3d_autotune_set(const c3d_config* cfg, const int* v) {
  g_create_err.clear();
  if (!cfg || !v) return -1;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.erase(tune_key(*cfg));
  }
  c3d_handle* h = nullptr;
  if (c3d_create(cfg, &h)) return -1;
  Shapes s = get_shapes(h);
  const Geo g = h->c.g;
  c3d_destroy(h);
  bool ok = true;
  auto lsh = [&](std::vector<LShape> c, const LShape& rule, int L, int seg, LShape& out) {
    c.push_back(rule);
    for (const LShape& x : c)
      if (x.L == L && (seg <= 0 || x.seg == seg)) {
        out = x;
        return;
      }
    ok = false;
  };
  auto ysh = [&](std::vector<YShape> c, const YShape& rule, int L, int C, int P, YShape& out) {
    c.push_back(rule);
    for (const YShape& x : c)
      if (x.L == L && x.C == C && x.P == P) {
        out = x;
        return;
      }
    ok = false;
  };
  auto tl = [&](std::vector<int> c, int rule, int t, int& out) {
    c.push_back(rule);
    for (int x : c)
      if (x == t) {
        out = x;
        return;
      }
    ok = false;
  };
  lsh(cand_lx(g.nx), s.lx, v[0], 0, s.lx);
  lsh(cand_lz(g.nz, false, cfg->precision == C3D_MIXED), s.lz, v[1], 0, s.lz);
  lsh(cand_lz(g.nz, true), s.lz32, v[2], v[3], s.lz32);
  ysh(cand_y(g.n2m, 16, 1024), s.ys, v[4], v[5], 1, s.ys);
  ysh(cand_y(g.n2m, 32, 1024, true), s.ys32, v[6], v[7], v[12], s.ys32);
  ysh(cand_yp(g.n2m), s.yp, v[8], v[9], 1, s.yp);
  tl(cand_tl(g.nx, true), s.tlx, v[10], s.tlx);
  tl(cand_tl(g.nz, false), s.tlz, v[11], s.tlz);
  if (!ok) {
    g_create_err = "c3d_autotune_set: a shape outside the candidate sets of this configuration";
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[tune_key(*cfg)] = s;
  return 0;
}

extern "C" void c3d_autotune_reset(void) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned.clear();
}

// ------------------------------------------------------------- ensembles
// Independent channels (any grids, precisions and models) advanced by one
// graph launch per step: the step of every member is captured as one branch
// of a single HIP graph (fork from the ensemble's stream, each branch on its
// member's stream, join back), so the device overlaps the members' kernels
// and the host issues one launch for all of them.  As for c3d_advance, two
// graphs cover the two buffer states a step alternates between.
// Members of one shape (grid, precision, models, tile shapes, buffer
// state) form a batch: each kernel of their step is launched once with the
// members along grid z, every block reading its member's context from a
// device array (CtxArg<true>).  Members of the models the batched launchers
// cover (FILES 0 or 2 with WMODEL 0, 1 or 6) are batched; the others run
// their own step as a branch of the graph, as before.
struct EnsBatch {
  std::vector<int> mem;              // member indices
  hipStream_t st = nullptr;          // capture branch stream (batches only)
  hipEvent_t join = nullptr;
  Ctx* dsnap[2] = {nullptr, nullptr};  // device context snapshots per graph parity
  std::vector<Ctx> hsnap[2];         // their host copies, written during capture
};
static const int kEnsSnapMax = 8;    // context snapshots per step (the step needs 3-4)

struct c3d_ensemble {
  std::vector<c3d_handle*> m;
  hipStream_t st = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  std::vector<int> gkey[2];
  std::vector<EnsBatch> batches;     // the grouping the graphs were captured with
  std::string err;
};

static std::vector<int> ens_key(const c3d_ensemble* e) {
  std::vector<int> k;
  for (const c3d_handle* h : e->m) k.push_back(buf_key(h));
  return k;
}

static void ens_batches_free(c3d_ensemble* e) {
  for (EnsBatch& b : e->batches) {
    for (int p = 0; p < 2; p++)
      if (b.dsnap[p]) (void)hipFree(b.dsnap[p]);
    if (b.join) (void)hipEventDestroy(b.join);
    if (b.st) (void)hipStreamDestroy(b.st);
  }
  e->batches.clear();
}

static void ens_graphs_free(c3d_ensemble* e) {
  for (int p = 0; p < 2; p++) {
    if (e->gexec[p]) (void)hipGraphExecDestroy(e->gexec[p]);
    e->gexec[p] = nullptr;
    e->gkey[p].clear();
  }
  ens_batches_free(e);
}

#define EHIPCHK(x)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) {                                                               \
      e->err = std::string(#x) + ": " + hipGetErrorString(e_);                            \
      return -1;                                                                          \
    }                                                                                     \
  } while (0)

// the batching key of a member ("" = not batchable): everything that fixes
// its kernel instances and launch geometry, and its buffer state
static std::string batch_key(const c3d_handle* h) {
  const c3d_config& f = h->cfg;
  const bool ok = (f.files == 0.0 || f.files == 2.0) && (f.wmodel == 0.0 || f.wmodel == 1.0 || f.wmodel == 6.0);
  if (!ok) return "";
  const Shapes sh = get_shapes(h);
  const Ops& o = h->c.o;
  std::string k = std::to_string(f.n1) + "x" + std::to_string(f.n2) + "x" + std::to_string(f.n3) + "/" +
                  std::to_string(f.precision) + "/" + std::to_string(f.files) + "/" + std::to_string(f.wmodel) +
                  "/" + std::to_string(buf_key(h)) + "/";
  k.append((const char*)&sh, sizeof(sh));
  const TileShape ts[4] = {o.tx, o.tz, o.tx4, o.tz4};
  k.append((const char*)ts, sizeof(ts));
  return k;
}

// the members grouped into batches (members of one key) and single members
static int ens_group(c3d_ensemble* e) {
  ens_batches_free(e);
  std::map<std::string, int> at;
  for (size_t q = 0; q < e->m.size(); q++) {
    const std::string k = batch_key(e->m[q]);
    if (!k.empty() && at.count(k)) {
      e->batches[at[k]].mem.push_back((int)q);
      continue;
    }
    if (!k.empty()) at[k] = (int)e->batches.size();
    e->batches.emplace_back();
    e->batches.back().mem.push_back((int)q);
  }
  for (EnsBatch& b : e->batches) {
    if (b.mem.size() < 2) continue;
    EHIPCHK(hipStreamCreateWithFlags(&b.st, hipStreamNonBlocking));
    EHIPCHK(hipEventCreateWithFlags(&b.join, hipEventDisableTiming));
    for (int p = 0; p < 2; p++) {
      EHIPCHK(hipMalloc((void**)&b.dsnap[p], sizeof(Ctx) * kEnsSnapMax * b.mem.size()));
      b.hsnap[p].clear();
    }
  }
  return 0;
}

// the context of batch b for a batched launch in graph parity p: the
// members' current host contexts as a device snapshot (a new one when they
// changed since the last), with member 0's shapes, the batch's stream and
// member 0's timer
static int batch_ctx(c3d_ensemble* e, EnsBatch& b, int p, Ctx& out) {
  const size_t n = b.mem.size();
  std::vector<Ctx>& hs = b.hsnap[p];
  const size_t nsnap = hs.size() / n;
  bool same = nsnap > 0;
  for (size_t q = 0; q < n && same; q++)
    same = std::memcmp(&hs[(nsnap - 1) * n + q], &e->m[b.mem[q]]->c, sizeof(Ctx)) == 0;
  size_t at = nsnap - 1;
  if (!same) {
    if (nsnap >= (size_t)kEnsSnapMax) {
      e->err = "ensemble capture: too many context snapshots";
      return -1;
    }
    for (size_t q = 0; q < n; q++) hs.push_back(e->m[b.mem[q]]->c);
    at = nsnap;
  }
  out = e->m[b.mem[0]]->c;
  out.bdev = b.dsnap[p] + at * n;
  out.bn = (int)n;
  out.st = b.st;
  return 0;
}

// one MAIN-loop step of every member of batch b (step_launch's sequence,
// each kernel launched once for the batch)
static int step_launch_batch(c3d_ensemble* e, EnsBatch& b, int p) {
  c3d_handle* h0 = e->m[b.mem[0]];
  const c3d_config& f = h0->cfg;
  Ctx c;
  if (f.wmodel == 0.0 && f.files != 0.0) {
    if (batch_ctx(e, b, p, c)) return -1;
    c3dk::wallss(c);
  }
  for (int q : b.mem) swap_hh(e->m[q]);
  if (f.files == 2.0) {
    if (batch_ctx(e, b, p, c)) return -1;
    c3dk::sgscs(c);
  }
  if (f.files != 0.0) {
    if (batch_ctx(e, b, p, c)) return -1;
    if (f.wmodel == 0.0)
      c3dk::wallss(c);
    else
      c3dk::wallmodel(c, f.wmodel == 6.0);
    for (int q : b.mem) e->m[q]->c.s.ifwall = f.wmodel == 0.0 ? 0 : 1;
  }
  if (batch_ctx(e, b, p, c)) return -1;
  c3dk::bcond(c);
  c3dk::rhs_xz(c);
  c3dk::getuh_fused(c);
  c3dk::rhsdp(c);
  c3dk::getdp(c);
  if (h0->c.s.mixed)
    for (int q : b.mem) swap_v(e->m[q]);
  if (batch_ctx(e, b, p, c)) return -1;
  c3dk::upcalc(c, true);
  return 0;
}

static int ens_capture(c3d_ensemble* e) {
  ens_graphs_free(e);
  if (ens_group(e)) return -1;
  std::vector<bool> timing;
  for (c3d_handle* h : e->m) {
    timing.push_back(h->kt.on);
    h->kt.on = false;  // no event nodes in the graphs
  }
  int rc = 0;
  for (int p = 0; p < 2 && rc == 0; p++) {
    e->gkey[p] = ens_key(e);
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      e->err = "hipStreamBeginCapture failed";
      rc = -1;
      break;
    }
    hipError_t r = hipEventRecord(e->fork, e->st);
    for (size_t q = 0; q < e->batches.size() && r == hipSuccess && rc == 0; q++) {
      EnsBatch& b = e->batches[q];
      if (b.mem.size() < 2) {
        c3d_handle* h = e->m[b.mem[0]];
        r = hipStreamWaitEvent(h->c.st, e->fork, 0);
        if (r != hipSuccess) break;
        step_launch(h);
        r = hipEventRecord(e->join[b.mem[0]], h->c.st);
      } else {
        r = hipStreamWaitEvent(b.st, e->fork, 0);
        if (r != hipSuccess) break;
        rc = step_launch_batch(e, b, p);
        if (r == hipSuccess) r = hipGetLastError();
        if (r == hipSuccess) r = hipEventRecord(b.join, b.st);
      }
      if (r == hipSuccess) r = hipStreamWaitEvent(e->st, b.mem.size() < 2 ? e->join[b.mem[0]] : b.join, 0);
    }
    const hipError_t r2 = hipStreamEndCapture(e->st, &g);
    if (r == hipSuccess) r = r2;
    if (r == hipSuccess && rc == 0) r = hipGraphInstantiate(&e->gexec[p], g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (r != hipSuccess) {
      e->err = std::string("ensemble capture: ") + hipGetErrorString(r);
      rc = -1;
    }
  }
  // the graphs read the context snapshots written during capture
  for (EnsBatch& b : e->batches)
    for (int p = 0; p < 2 && rc == 0; p++)
      if (b.dsnap[p] && !b.hsnap[p].empty() &&
          hipMemcpy(b.dsnap[p], b.hsnap[p].data(), sizeof(Ctx) * b.hsnap[p].size(), hipMemcpyHostToDevice) !=
              hipSuccess) {
        e->err = "ensemble context upload failed";
        rc = -1;
      }
  for (size_t q = 0; q < e->m.size(); q++) e->m[q]->kt.on = timing[q];
  if (rc) ens_graphs_free(e);
  return rc;
}

extern "C" int c3d_ensemble_create(c3d_handle* const* members, int n, c3d_ensemble** out) {
  if (!out) return -1;
  *out = nullptr;
  if (!members || n <= 0) {
    g_create_err = "an ensemble needs at least one member";
    return -1;
  }
  c3d_ensemble* e = new c3d_ensemble();
  for (int q = 0; q < n; q++) {
    if (!members[q]) {
      delete e;
      g_create_err = "null ensemble member";
      return -1;
    }
    for (int r = 0; r < q; r++)
      if (members[r] == members[q]) {
        delete e;
        g_create_err = "an ensemble member appears twice";
        return -1;
      }
    e->m.push_back(members[q]);
  }
  bool ok = hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) == hipSuccess;
  e->join.assign(n, nullptr);
  for (int q = 0; q < n && ok; q++) ok = hipEventCreateWithFlags(&e->join[q], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    (void)c3d_ensemble_destroy(e);
    g_create_err = "ensemble stream/event creation failed";
    return -1;
  }
  *out = e;
  return 0;
}

extern "C" int c3d_ensemble_destroy(c3d_ensemble* e) {
  if (!e) return 0;
  if (e->st) (void)hipStreamSynchronize(e->st);
  ens_graphs_free(e);
  for (hipEvent_t v : e->join)
    if (v) (void)hipEventDestroy(v);
  if (e->fork) (void)hipEventDestroy(e->fork);
  if (e->st) (void)hipStreamDestroy(e->st);
  delete e;
  return 0;
}

extern "C" const char* c3d_ensemble_last_error(const c3d_ensemble* e) {
  if (!e) return g_create_err.c_str();
  return e->err.c_str();
}

extern "C" int c3d_ensemble_advance(c3d_ensemble* e, int nsteps) {
  if (!e) return -1;
  if (nsteps < 0) {
    e->err = "nsteps must be >= 0";
    return -1;
  }
  if (nsteps == 0) return 0;
  // the ensemble's stream follows whatever each member has queued
  for (size_t q = 0; q < e->m.size(); q++) {
    EHIPCHK(hipEventRecord(e->join[q], e->m[q]->c.st));
    EHIPCHK(hipStreamWaitEvent(e->st, e->join[q], 0));
  }
  for (int n = 0; n < nsteps; n++) {
    const std::vector<int> key = ens_key(e);
    int p = (e->gexec[0] && e->gkey[0] == key) ? 0 : (e->gexec[1] && e->gkey[1] == key) ? 1 : -1;
    if (p < 0) {
      if (ens_capture(e)) return -1;
      p = 0;
    }
    EHIPCHK(hipGraphLaunch(e->gexec[p], e->st));
    for (c3d_handle* h : e->m) {
      swap_hh(h);
      if (h->c.s.mixed) swap_v(h);
    }
  }
  // and every member's stream follows the ensemble (for c3d_sync, c3d_step, ...)
  EHIPCHK(hipEventRecord(e->fork, e->st));
  for (c3d_handle* h : e->m) EHIPCHK(hipStreamWaitEvent(h->c.st, e->fork, 0));
  const hipError_t r = hipGetLastError();
  if (r != hipSuccess) {
    e->err = std::string("ensemble launch: ") + hipGetErrorString(r);
    return -1;
  }
  return 0;
}

extern "C" int c3d_ensemble_sync(c3d_ensemble* e, double* divmax, double* cflm) {
  if (!e) return -1;
  EHIPCHK(hipStreamSynchronize(e->st));
  // every member's values are returned; a diverged member gives
  // C3D_ERR_NONFINITE naming the first one
  int out = 0;
  for (size_t q = 0; q < e->m.size(); q++) {
    double dm = 0.0, cf = 0.0;
    const int rc = c3d_sync(e->m[q], &dm, &cf);
    if (divmax) divmax[q] = dm;
    if (cflm) cflm[q] = cf;
    if (rc == C3D_ERR_NONFINITE) {
      if (out == 0) {
        e->err = std::string("member ") + std::to_string(q) + ": " + e->m[q]->err;
        out = C3D_ERR_NONFINITE;
      }
    } else if (rc) {
      e->err = std::string("member ") + std::to_string(q) + ": " + e->m[q]->err;
      return -1;
    }
  }
  return out;
}

extern "C" int c3d_step(c3d_handle* h, int ntime, double* divmax, double* cflm) {
  if (!h) return -1;
  // MAIN (:125) evaluates WALLSS before GETUP when WMODEL=0; only the SGS
  // model consumes it
  if (h->cfg.wmodel == 0.0 && h->cfg.files != 0.0) c3dk::wallss(h->c);
  if (step_impl(h, true, ntime)) return -1;
  // one read-back of DIVMAX and CFLM per step (mixed mode: the refinement
  // check uses the same DIVMAX)
  double dm = 0.0, cf = 0.0;
  if (read_checks(h, &dm, &cf)) return -1;
  if (h->c.s.mixed && refine_from(h, dm, &dm, &cf)) return -1;
  if (divmax) *divmax = dm;
  if (cflm) *cflm = cf;
  return check_finite(h, dm, cf);
}

extern "C" int c3d_divcheck(c3d_handle* h, double* divmax) {
  if (!h) return -1;
  c3dk::divcheck(h->c);
  if (launched(h)) return -1;
  if (divmax) return read_sc(h, SC_DIVMAX, divmax);
  return 0;
}

extern "C" int c3d_cfl(c3d_handle* h, double* cflm) {
  if (!h) return -1;
  c3dk::cfl(h->c);
  if (launched(h)) return -1;
  if (cflm) return read_sc(h, SC_CFLM, cflm);
  return 0;
}

extern "C" int c3d_chkmf(c3d_handle* h, double* flow1, double* flow3) {
  if (!h) return -1;
  c3dk::chkmf(h->c, true);
  if (launched(h)) return -1;
  if (flow1 && read_sc(h, SC_CHK1, flow1)) return -1;
  if (flow3 && read_sc(h, SC_CHK3, flow3)) return -1;
  return 0;
}

// the same statistics with the scratch in g_scratch: no allocation of any
// kind; calls that overlap on one GPU share (and race on) that GPU's copy
extern "C" int c3d_stats_noalloc(c3d_handle* h, double* out) {
  if (!h) return -1;
  if (!out) return fail(h, "c3d_stats_noalloc needs an output array of 20*(N2-1) values");
  if (h->c.g.n2m > C3D_STATS_NOALLOC_MAX_PLANES)
    return fail(h, "c3d_stats_noalloc: N2-1 = " + std::to_string(h->c.g.n2m) + " exceeds its scratch of " +
                       std::to_string(C3D_STATS_NOALLOC_MAX_PLANES) + " planes (use c3d_stats)");
  double* sp = scratch_of(h->cfg.device);
  if (!sp) return fail(h, "c3d_stats_noalloc: no scratch address for this device");
  const size_t n = 20 * (size_t)h->c.g.n2m;
  c3dk::plane_stats(h->c, sp);
  if (launched(h)) return -1;
  HIPCHK(hipMemcpyAsync(out, sp, n * sizeof(double), hipMemcpyDeviceToHost, h->c.st));
  HIPCHK(hipStreamSynchronize(h->c.st));
  return 0;
}

extern "C" int c3d_stats(c3d_handle* h, double* out) {
  if (!h) return -1;
  if (!out) return fail(h, "c3d_stats needs an output array of 20*(N2-1) values");
  const size_t n = 20 * (size_t)h->c.g.n2m;
  c3dk::plane_stats(h->c, h->dtmp);
  if (launched(h)) return -1;
  HIPCHK(hipMemcpyAsync(out, h->dtmp, n * sizeof(double), hipMemcpyDeviceToHost, h->c.st));
  HIPCHK(hipStreamSynchronize(h->c.st));
  return 0;
}

// ------------------------------------------------------- WRITEUP output
// The Tecplot files of WRITEUP (compact3d.f:791-832), text as the
// reference's Fortran runtime writes it: list-directed header lines (a
// leading blank, items separated by one blank) and '(7F20.10)' records.

// F20.10: 10 decimals right-justified in 20 characters; 20 asterisks when
// the value does not fit; NaN, Inf, -Inf right-justified
static void f20_10(char* o, double x) {
  char b[400];
  int n;
  if (std::isnan(x))
    n = snprintf(b, sizeof b, "NaN");
  else if (std::isinf(x))
    n = snprintf(b, sizeof b, x < 0 ? "-Inf" : "Inf");
  else
    n = snprintf(b, sizeof b, "%.10f", x);
  if (n < 0 || n > 20) {
    memset(o, '*', 20);
    return;
  }
  memset(o, ' ', 20 - n);
  memcpy(o + 20 - n, b, n);
}

// records of nv F20.10 values, one per row, formatted by up to 16 threads
template <typename RowFn>
static bool write_rows(FILE* f, long long nrows, int nv, RowFn row) {
  const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const long long w = 20LL * nv + 1, batch = 1 << 16;
  std::vector<char> buf((size_t)(batch * nth * w));
  for (long long r0 = 0; r0 < nrows; r0 += batch * nth) {
    const long long nr = std::min(nrows - r0, batch * nth);
    std::vector<std::thread> th;
    for (int t = 0; t < nth; t++)
      th.emplace_back([&, t] {
        double v[8];
        for (long long r = t; r < nr; r += nth) {
          char* o = buf.data() + r * w;
          row(r0 + r, v);
          for (int q = 0; q < nv; q++) f20_10(o + 20 * q, v[q]);
          o[20 * nv] = '\n';
        }
      });
    for (auto& x : th) x.join();
    if (fwrite(buf.data(), 1, (size_t)(nr * w), f) != (size_t)(nr * w)) return false;
  }
  return true;
}

extern "C" int c3d_format_flow(const char* path, int ntime, int n1m, int n2m, int n3m, const double* xc,
                               const double* yc, const double* zc, const double* cell) {
  if (!path || !xc || !yc || !zc || !cell || n1m < 1 || n2m < 1 || n3m < 1) {
    g_create_err = "c3d_format_flow: bad arguments";
    return -1;
  }
  FILE* f = fopen(path, "wb");
  if (!f) {
    g_create_err = std::string("c3d_format_flow: cannot open ") + path;
    return -1;
  }
  fprintf(f, " VARIABLES=\"x\",\"y\",\"z\",\"U\",\"V\",\"W\",\"P\" \n");
  fprintf(f, " ZONE T=\" %d \"I=  %d J= %d K= %d\n", ntime, n1m, n2m, n3m);
  const long long nc = (long long)n1m * n2m * n3m;
  const bool ok = write_rows(f, nc, 7, [&](long long e, double* v) {
    const int i = (int)(e % n1m);
    const long long r = e / n1m;
    const int j = (int)(r % n2m), k = (int)(r / n2m);
    v[0] = xc[i];
    v[1] = yc[j];
    v[2] = zc[k];
    for (int q = 0; q < 4; q++) v[3 + q] = cell[q * nc + e];
  });
  if (fclose(f) != 0 || !ok) {
    g_create_err = std::string("c3d_format_flow: write failed: ") + path;
    return -1;
  }
  return 0;
}

extern "C" int c3d_format_strain(const char* path, int ntime, int n1m, int n3m, const double* xc, const double* zc,
                                 const double* tauw, double utau) {
  if (!path || !xc || !zc || !tauw || n1m < 1 || n3m < 1) {
    g_create_err = "c3d_format_strain: bad arguments";
    return -1;
  }
  // TAUDOWN: the plane mean of TAUW(I,K,1), summed K outer, I inner (:815-821)
  double td = 0.0;
  for (int k = 0; k < n3m; k++)
    for (int i = 0; i < n1m; i++) td += tauw[(size_t)k * n1m + i];
  td = td / (double)(n1m * n3m);
  const double u2 = utau * utau;
  FILE* f = fopen(path, "wb");
  if (!f) {
    g_create_err = std::string("c3d_format_strain: cannot open ") + path;
    return -1;
  }
  fprintf(f, " VARIABLES=\"x\",\"z\",\"strain\"\n");
  fprintf(f, " ZONE T=\" %d \"I=  %d K= %d\n", ntime, n1m, n3m);
  const bool ok = write_rows(f, (long long)n1m * n3m, 3, [&](long long e, double* v) {
    const int i = (int)(e % n1m), k = (int)(e / n1m);
    v[0] = xc[i];
    v[1] = zc[k];
    v[2] = (tauw[e] - td) / u2;
  });
  if (fclose(f) != 0 || !ok) {
    g_create_err = std::string("c3d_format_strain: write failed: ") + path;
    return -1;
  }
  return 0;
}

// XC(1:N1M), YC(1:N2M), ZC(1:N3M) of MESH (:322-333, :363-381)
static void cell_coords(const c3d_handle* h, std::vector<double>& xc, std::vector<double>& yc,
                        std::vector<double>& zc) {
  const int n1 = h->cfg.n1, n2 = h->cfg.n2, n3 = h->cfg.n3;
  auto grid = [](int i, int n, double al) { return i == n ? al : (double)(i - 1) * al / (double)(n - 1); };
  xc.resize(n1 - 1);
  zc.resize(n3 - 1);
  yc.resize(n2 - 1);
  for (int i = 1; i < n1; i++) xc[i - 1] = 0.5 * (grid(i + 1, n1, h->cfg.alx) + grid(i, n1, h->cfg.alx));
  for (int k = 1; k < n3; k++) zc[k - 1] = 0.5 * (grid(k + 1, n3, h->cfg.alz) + grid(k, n3, h->cfg.alz));
  for (int j = 1; j < n2; j++) yc[j - 1] = 0.5 * (h->Y[j + 1] + h->Y[j]);
}

extern "C" int c3d_cellcentre(c3d_handle* h, double* out) {
  if (!h) return -1;
  if (!out) return fail(h, "c3d_cellcentre needs an output array of 4*(N1-1)*(N2-1)*(N3-1) values");
  const Geo& g = h->c.g;
  const size_t nc = (size_t)g.nx * g.n2m * g.nz;
  for (int q = 0; q < 4; q++) {
    c3dk::cellcentre(h->c, q, h->dtmp);
    if (launched(h)) return -1;
    HIPCHK(hipMemcpyAsync(out + q * nc, h->dtmp, nc * sizeof(double), hipMemcpyDeviceToHost, h->c.st));
    HIPCHK(hipStreamSynchronize(h->c.st));
  }
  return 0;
}

extern "C" int c3d_write_flow(c3d_handle* h, const char* path, int ntime) {
  if (!h) return -1;
  const Geo& g = h->c.g;
  std::vector<double> cell(4 * (size_t)g.nx * g.n2m * g.nz), xc, yc, zc;
  if (c3d_cellcentre(h, cell.data())) return -1;
  cell_coords(h, xc, yc, zc);
  if (c3d_format_flow(path, ntime, g.nx, g.n2m, g.nz, xc.data(), yc.data(), zc.data(), cell.data()))
    return fail(h, g_create_err.c_str());
  return 0;
}

extern "C" int c3d_write_strain(c3d_handle* h, const char* path, int ntime, double utau) {
  if (!h) return -1;
  const Geo& g = h->c.g;
  std::vector<double> t((size_t)g.nx * g.nz), xc, yc, zc;
  HIPCHK(hipMemcpyAsync(t.data(), h->c.d.tauw, t.size() * sizeof(double), hipMemcpyDeviceToHost, h->c.st));
  HIPCHK(hipStreamSynchronize(h->c.st));
  cell_coords(h, xc, yc, zc);
  if (c3d_format_strain(path, ntime, g.nx, g.nz, xc.data(), zc.data(), t.data(), utau))
    return fail(h, g_create_err.c_str());
  return 0;
}

extern "C" int c3d_sgs_profiles(c3d_handle* h, double* smagc, double* beta) {
  if (!h) return -1;
  const Fields& F = h->c.d;
  const int n2m = h->c.g.n2m;
  if ((smagc && !F.smagc) || (beta && !F.sdbeta))
    return fail(h, "c3d_sgs_profiles: C(J) exists for FILES = 1 or 3, BETA(J) for FILES = 3");
  if (smagc) HIPCHK(hipMemcpyAsync(smagc, F.smagc + 1, n2m * sizeof(double), hipMemcpyDeviceToHost, h->c.st));
  if (beta) HIPCHK(hipMemcpyAsync(beta, F.sdbeta + 1, n2m * sizeof(double), hipMemcpyDeviceToHost, h->c.st));
  HIPCHK(hipStreamSynchronize(h->c.st));
  return 0;
}

extern "C" int c3d_flowrate(c3d_handle* h, double* flow1, double* flow3) {
  if (!h) return -1;
  c3dk::chkmf(h->c, false);
  if (launched(h)) return -1;
  if (flow1 && read_sc(h, SC_CHK1, flow1)) return -1;
  if (flow3 && read_sc(h, SC_CHK3, flow3)) return -1;
  return 0;
}

extern "C" int c3d_get_field(c3d_handle* h, int which, double* out) {
  if (!h || !out) return -1;
  Fields& F = h->c.d;
  double* src = nullptr;
  int lay = LAY_U;
  switch (which) {
    case C3D_F_RUH1: src = F.ruh[0]; break;
    case C3D_F_RUH2: src = F.ruh[1]; break;
    case C3D_F_RUH3: src = F.ruh[2]; break;
    case C3D_F_UH1: src = F.uh[0]; break;
    case C3D_F_UH2: src = F.uh[1]; break;
    case C3D_F_UH3: src = F.uh[2]; break;
    case C3D_F_RDP: src = F.rdp; lay = LAY_P; break;  // FP64 in both modes
    case C3D_F_DP: src = F.dp; lay = LAY_P; break;  // FP64 in both modes
    case C3D_F_DIV:
      // mixed mode keeps the full DIV of the state (the next RDP starts from
      // it); the FP64 step's update kernels keep only its x and y parts, so
      // DIVCHECK forms it here
      if (!h->c.s.mixed) {
        c3dk::divcheck(h->c);
        if (launched(h)) return -1;
        return from_dev(h, out, F.div, LAY_P);
      }
      c3dk::to_f64(h->c, F.divf, h->dtmp, h->c.g.fsz);
      if (launched(h)) return -1;
      return from_dev(h, out, h->dtmp, LAY_P);
    case C3D_F_HH1: src = F.hh[0]; break;
    case C3D_F_HH2: src = F.hh[1]; break;
    case C3D_F_HH3: src = F.hh[2]; break;
    case C3D_F_TAUW:
    case C3D_F_TAUB: {
      const Geo& g = h->c.g;
      const size_t n2d = (size_t)g.nx * g.nz;
      std::vector<double> t(2 * n2d);
      HIPCHK(hipMemcpyAsync(t.data(), which == C3D_F_TAUW ? F.tauw : F.taub, t.size() * 8, hipMemcpyDeviceToHost,
                            h->c.st));
      HIPCHK(hipStreamSynchronize(h->c.st));
      const int n1 = h->cfg.n1, n3 = h->cfg.n3;
      const size_t lay = (size_t)(n1 + 1) * (n3 + 1);
      for (size_t q = 0; q < 2 * lay; q++) out[q] = 0.0;
      for (int l = 0; l < 2; l++)
        for (int k = 0; k < g.nz; k++)
          for (int i = 0; i < g.nx; i++)
            out[l * lay + (size_t)(i + 1) + (size_t)(n1 + 1) * (k + 1)] = t[l * n2d + (size_t)k * g.nx + i];
      return 0;
    }
    case C3D_F_IWM: {
      if (!F.iwm) return fail(h, "the integral wall model's state exists only with WMODEL = 2");
      const size_t nb = (size_t)IW_NPL * h->c.g.nx * h->c.g.nz * sizeof(double);
      HIPCHK(hipMemcpyAsync(out, F.iwm, nb, hipMemcpyDeviceToHost, h->c.st));
      HIPCHK(hipStreamSynchronize(h->c.st));
      return 0;
    }
    default: return fail(h, "unknown field");
  }
  if (h->c.s.mixed && which != C3D_F_DP && which != C3D_F_RDP) {
    // FP32 intermediates; UH is U + DELTA
    const long long n = h->c.g.fsz;
    if (which == C3D_F_UH1 || which == C3D_F_UH2 || which == C3D_F_UH3) {
      const double* U = which == C3D_F_UH1 ? F.u : which == C3D_F_UH2 ? F.v : F.w;
      c3dk::add_f32(h->c, U, (const float*)src, h->dtmp, n);
    } else {
      c3dk::to_f64(h->c, (const float*)src, h->dtmp, n);
    }
    src = h->dtmp;
  }
  return from_dev(h, out, src, lay);
}

extern "C" int c3d_set_field(c3d_handle* h, int which, const double* in) {
  if (!h || !in) return -1;
  Fields& F = h->c.d;
  switch (which) {
    case C3D_F_IWM: {
      if (!F.iwm) return fail(h, "the integral wall model's state exists only with WMODEL = 2");
      const size_t nb = (size_t)IW_NPL * h->c.g.nx * h->c.g.nz * sizeof(double);
      HIPCHK(hipMemcpyAsync(F.iwm, in, nb, hipMemcpyHostToDevice, h->c.st));
      HIPCHK(hipStreamSynchronize(h->c.st));
      return 0;
    }
    case C3D_F_TAUW:
    case C3D_F_TAUB: {
      // the inverse of c3d_get_field's (0:N1,0:N3,2) layout
      const Geo& g = h->c.g;
      const size_t n2d = (size_t)g.nx * g.nz;
      const int n1 = h->cfg.n1, n3 = h->cfg.n3;
      const size_t lay = (size_t)(n1 + 1) * (n3 + 1);
      std::vector<double> t(2 * n2d);
      for (int l = 0; l < 2; l++)
        for (int k = 0; k < g.nz; k++)
          for (int i = 0; i < g.nx; i++)
            t[l * n2d + (size_t)k * g.nx + i] = in[l * lay + (size_t)(i + 1) + (size_t)(n1 + 1) * (k + 1)];
      HIPCHK(hipMemcpyAsync(which == C3D_F_TAUW ? F.tauw : F.taub, t.data(), t.size() * 8, hipMemcpyHostToDevice,
                            h->c.st));
      HIPCHK(hipStreamSynchronize(h->c.st));
      return 0;
    }
    case C3D_F_DIV:
      if (to_dev(h, F.div, in, LAY_P)) return -1;
      if (h->c.s.mixed) {
        c3dk::to_f32(h->c, F.div, F.divf, h->c.g.fsz);
        if (launched(h)) return -1;
      }
      return 0;
    default: return fail(h, "only C3D_F_TAUW, C3D_F_TAUB, C3D_F_IWM and C3D_F_DIV can be set");
  }
}

// c3d_get_scalar / c3d_set_scalar index -> device slot (-1: not a slot)
static int scalar_slot(int which) {
  static const int map[C3D_NSCALAR] = {SC_PRESG, SC_PRESG3, SC_DMPRESG, SC_DMPRESG3, SC_FLOW1,  SC_FLOW3,
                                       SC_WSM1,  SC_WSM2,   SC_WSM3,    -1,          -1,        SC_NTIME,
                                       SC_CFLMAX, SC_WMDT,  SC_TAOW,    SC_UINIT,    SC_USINIT, SC_PBAR};
  return which >= 0 && which < C3D_NSCALAR ? map[which] : -1;
}

extern "C" int c3d_get_scalar(c3d_handle* h, int which, double* out) {
  if (!h || !out) return -1;
  if (which == 9 || which == 10) {  // mixed mode diagnostics
    *out = which == 9 ? (double)h->refines : h->first_divmax;
    return 0;
  }
  const int slot = scalar_slot(which);
  if (slot < 0) return fail(h, "unknown scalar");
  return read_sc(h, slot, out);
}

extern "C" int c3d_set_scalar(c3d_handle* h, int which, double v) {
  if (!h) return -1;
  const int slot = scalar_slot(which);
  if (slot < 0) return fail(h, "unknown scalar");
  HIPCHK(hipMemcpyAsync(h->c.d.sc + slot, &v, sizeof(double), hipMemcpyHostToDevice, h->c.st));
  HIPCHK(hipStreamSynchronize(h->c.st));
  return 0;
}

extern "C" int c3d_set_dt(c3d_handle* h, double dt) {
  if (!h) return -1;
  if (!(dt > 0)) return fail(h, "dt must be > 0");
  HIPCHK(hipStreamSynchronize(h->c.st));
  h->c.s.dt = dt;
  h->cfg.dt = dt;
  graphs_free(h);
  return upload_dt_terms(h);
}

extern "C" int c3d_filter(c3d_handle* h, int width, const double* in, double* out, int n) {
  if (!h || !in || !out) return -1;
  if (width != 2 && width != 4) return fail(h, "width must be 2 or 4");
  if (n < 1) return fail(h, "n must be >= 1");
  const size_t plane = (size_t)h->cfg.n1 * h->cfg.n3;
  const size_t bytes = plane * n * sizeof(double);
  if (bytes > (size_t)h->c.g.fsz * sizeof(double)) return fail(h, "too many planes for the scratch field");
  HIPCHK(hipMemcpyAsync(h->dtmp, in, bytes, hipMemcpyHostToDevice, h->c.st));
  // entries outside I<N1, K<N3 keep their input values, as in the reference
  HIPCHK(hipMemcpyAsync(h->dtmp2, h->dtmp, bytes, hipMemcpyDeviceToDevice, h->c.st));
  c3dk::filter(h->c, width, h->dtmp, h->dtmp2, n, h->cfg.n1, h->cfg.n3);
  if (launched(h)) return -1;
  HIPCHK(hipMemcpyAsync(out, h->dtmp2, bytes, hipMemcpyDeviceToHost, h->c.st));
  HIPCHK(hipStreamSynchronize(h->c.st));
  return 0;
}

extern "C" int c3d_get_mesh(c3d_handle* h, double* y, double* scal4) {
  if (!h) return -1;
  if (y)
    for (int j = 0; j <= h->cfg.n2; j++) y[j] = h->Y[j];
  if (scal4) {
    scal4[0] = h->c.s.dx1;
    scal4[1] = h->c.s.dx3;
    scal4[2] = h->c.s.aly;
    scal4[3] = h->c.s.vol;
  }
  return 0;
}

extern "C" int c3d_apply_op(c3d_handle* h, int op, int dir, const double* in, double* out) {
  if (!h || !in || !out) return -1;
  if (op < 0 || op > 4 || (dir != 1 && dir != 3)) return fail(h, "bad op/dir");
  if (to_dev(h, h->dtmp, in, LAY_P)) return -1;
  c3dk::apply_op(h->c, op, dir, h->dtmp, h->dtmp2);
  if (launched(h)) return -1;
  return from_dev(h, out, h->dtmp2, LAY_P);
}

static const char* kid_names[C3D_NKID] = {
    "bcond", "rhs_x", "rhs_z", "rhs_x2", "rhs_y", "adi_y", "adi_x", "adi_z", "rhsdp_x", "rhsdp_z",
    "fft_x", "fft_z", "pois_y", "fft_xinv", "flow_red", "flow_fin", "upd_x", "ghosts", "upd_z",
    "div_x", "div_z", "cfl", "chkmf", "zero", "apply_op", "sgscs", "wallmodel", "sgs_x", "sgs_z",
    "sgs_z2", "sgs_x2", "wallss", "filter", "rhs_zw", "upd_v", "adi_xr", "adi_zr"};
static_assert(KID_COUNT <= C3D_NKID, "kernel id table too small");

extern "C" const char* c3d_kernel_name(int id) {
  if (id < 0 || id >= KID_COUNT) return "";
  return kid_names[id];
}

extern "C" int c3d_timing(c3d_handle* h, int on, int capacity) {
  if (!h) return -1;
  HIPCHK(hipStreamSynchronize(h->c.st));
  KTimer& t = h->kt;
  if (on && capacity > t.cap) {
    kt_free(h);
    t.ev = new hipEvent_t[2 * (size_t)capacity];
    t.id = new int[capacity];
    for (int q = 0; q < 2 * capacity; q++) HIPCHK(hipEventCreate(&t.ev[q]));
    t.cap = capacity;
  }
  t.used = 0;
  t.on = on != 0;
  return 0;
}

extern "C" int c3d_timing_select(c3d_handle* h, int kid) {
  if (!h) return -1;
  if (kid >= KID_COUNT) return fail(h, "no such kernel id");
  h->kt.only = kid < 0 ? -1 : kid;
  return 0;
}

extern "C" int c3d_kernel_times(c3d_handle* h, double* ms, int* count) {
  if (!h) return -1;
  HIPCHK(hipStreamSynchronize(h->c.st));
  KTimer& t = h->kt;
  for (int q = 0; q < C3D_NKID; q++) {
    if (ms) ms[q] = 0.0;
    if (count) count[q] = 0;
  }
  for (int s2 = 0; s2 < t.used; s2++) {
    float e = 0.f;
    HIPCHK(hipEventElapsedTime(&e, t.ev[2 * s2], t.ev[2 * s2 + 1]));
    if (ms) ms[t.id[s2]] += e;
    if (count) count[t.id[s2]] += 1;
  }
  t.used = 0;
  return 0;
}
