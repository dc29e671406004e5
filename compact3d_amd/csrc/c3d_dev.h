// c3d_dev.h -- device-side layout, parameters and the LDS line engines.
//
// Field layout in HBM (one array per field): f[(k*NJ + j)*NX + i]
//   i = I-1 in [0,NX), NX = N1M      (x, periodic, contiguous)
//   j = J   in [0,N2], NJ = N2+1     (y, wall-normal, ghost planes 0 and N2)
//   k = K-1 in [0,NZ), NZ = N3M      (z, periodic)
// x-lines are contiguous rows; z-lines have stride NJ*NX; y-lines stride NX.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef C3D_NT
#define C3D_NT 256  // threads per workgroup of the line kernels (4 wavefronts)
#endif

struct Geo {
  int nx, nz, n2, n2m, nj, nxh;  // nxh = NX/2+1 (the x-modes kx = 0..NX/2)
  int nxhp;                      // row stride of the Poisson factor tables: nxh rounded up to 16
  // the stored half spectrum: nxs = NX/2 columns, column 0 packing the two
  // real x-modes X(0) + i X(NX/2); row stride nxsp >= nxs, an odd multiple
  // of 16 complex values (256 bytes)
  int nxs, nxsp;
  long long plane;               // NJ*NX
  long long fsz;                 // NZ*NJ*NX
};

__device__ __forceinline__ long long gidx(const Geo& g, int i, int j, int k) {
  return ((long long)k * g.nj + j) * g.nx + i;
}

// Store with the nontemporal hint, for the line kernels' write-once fields
// (read again only by later kernels): their lines do not crowd the L2 that
// holds the launch's re-read neighbour rows and halos.  Measured over the
// z-line and x-line kernels at 256x193x256: ~35 us per step in each
// precision.
template <typename T>
__device__ __forceinline__ void st_nt(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}
// Load with the nontemporal hint, for operands a kernel reads once and no
// later kernel of the step reads soon (the z-line RHS pass's accumulators:
// 176 -> 160 us mixed, and the y-sweep after it 228 -> 220 us).
template <typename T>
__device__ __forceinline__ T ld_nt(const T* p) {
  return __builtin_nontemporal_load(p);
}

// Compact-operator line solve.  Every periodic system of the path has the
// constant symmetric form a*x[p-1] + b*x[p] + a*x[p+1] = d[p] (CTDMA11/CTDMA33
// with (1,b,1), CTDMA1J/CTDMA3I with the ADI factors).  Writing
// beta = b/a and alpha = the root of alpha^2 - beta*alpha + 1 = 0 with
// |alpha| < 1, the circulant factors exactly as
//   a*(S^-1 + beta + S) = (a/alpha) (1 + alpha S^-1)(1 + alpha S),
// so x = (alpha/a) * B^-1 F^-1 d with the first-order recursions
//   F: y[p] = d[p] - alpha*y[p-1]      (forward)
//   B: z[p] = y[p] - alpha*z[p+1]      (backward).
// Each recursion forgets its start value as |alpha|^W; starting W points
// early (wrapping periodically) with zero state gives the periodic solution
// to |alpha|^W <= 2^-60.  That lets G threads share one line: each owns a
// segment and warms up from its neighbours' data.
struct LineSolve {
  double alpha;  // decay factor
  double scale;  // alpha / a
  int w;         // warm-up length
};

// Work tile in LDS: element (p, l) at buf[p*ls + l]; p along the line, l the
// line within the tile.  Threads t = l + L*g, g = segment.
template <typename T>
__device__ __forceinline__ void tile_solve(T* buf, int n, int nl, int ls, int g_cnt, int seg,
                                           const LineSolve op) {
  const int t = threadIdx.x;
  const int l = t % nl;
  const int g = t / nl;
  const bool act = (g < g_cnt);
  const int p0 = g * seg;
  const int p1 = min(p0 + seg, n);
  const T a = (T)op.alpha;
  const int w = op.w;
  T y = (T)0;
  if (act && p0 < n) {
    int q = p0 - w;
    q %= n;
    if (q < 0) q += n;
    for (int s = 0; s < w; s++) {
      y = buf[q * ls + l] - a * y;
      q = (q + 1 == n) ? 0 : q + 1;
    }
  }
  __syncthreads();
  if (act) {
    for (int p = p0; p < p1; p++) {
      y = buf[p * ls + l] - a * y;
      buf[p * ls + l] = y;
    }
  }
  __syncthreads();
  T z = (T)0;
  if (act && p0 < n) {
    int q = (p1 + w - 1) % n;
    for (int s = 0; s < w; s++) {
      z = buf[q * ls + l] - a * z;
      q = (q == 0) ? n - 1 : q - 1;
    }
  }
  __syncthreads();
  if (act) {
    const T sc = (T)op.scale;
    for (int p = p1 - 1; p >= p0; p--) {
      z = buf[p * ls + l] - a * z;
      buf[p * ls + l] = sc * z;
    }
  }
  __syncthreads();
}

// Tile shape: L lines x N points, G segments per line, L*G <= C3D_NT.
struct TileShape {
  int n;    // line length
  int nl;   // lines per tile (L)
  int ls;   // LDS line stride (L, padded)
  int g;    // segments per line
  int seg;  // segment length
};

// ------------------------------------------------- element loops, reductions
#define MAXE 16  // elements per thread of a line tile (n*L <= 16*C3D_NT)

#define EL_LOOP(nL)                                    \
  _Pragma("unroll") for (int m = 0; m < MAXE; m++) {   \
    const int e = threadIdx.x + m * C3D_NT;            \
    if (e < (nL)) {
#define EL_END \
  }            \
  }

__device__ __forceinline__ int wrapm(int p, int n) { return p == 0 ? n - 1 : p - 1; }
__device__ __forceinline__ int wrapp(int p, int n) { return p == n - 1 ? 0 : p + 1; }

// atomic max of a non-negative double (NaN compares largest, so it propagates)
__device__ __forceinline__ void atomic_max_pos(double* slot, double v) {
  unsigned long long b = __double_as_longlong(fabs(v)) ;
  if (v != v) b = 0x7ff8000000000000ull;
  atomicMax((unsigned long long*)slot, b);
}

__device__ __forceinline__ double block_max(double v, double* red) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = C3D_NT / 2; s > 0; s >>= 1) {
    if (t < s) {
      double a = red[t], b = red[t + s];
      red[t] = (a != a || b != b) ? NAN : fmax(a, b);
    }
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}


// ---------------------------------------------------------------- FFT
struct cpx {
  double x, y;
};
template <typename T>
struct cplx_t {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cplx_t<T> cmul(cplx_t<T> a, cplx_t<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// small DFT of size R in place, sign = -1 forward / +1 backward
template <typename T>
__device__ __forceinline__ void dft_small(cplx_t<T>* v, int R, int sign) {
  if (R == 2) {
    cplx_t<T> a = v[0], b = v[1];
    v[0] = {a.x + b.x, a.y + b.y};
    v[1] = {a.x - b.x, a.y - b.y};
  } else if (R == 4) {
    cplx_t<T> a0 = {v[0].x + v[2].x, v[0].y + v[2].y};
    cplx_t<T> a1 = {v[0].x - v[2].x, v[0].y - v[2].y};
    cplx_t<T> a2 = {v[1].x + v[3].x, v[1].y + v[3].y};
    cplx_t<T> a3 = {v[1].x - v[3].x, v[1].y - v[3].y};
    // multiply a3 by sign*i
    cplx_t<T> a3r = (sign < 0) ? cplx_t<T>{a3.y, -a3.x} : cplx_t<T>{-a3.y, a3.x};
    v[0] = {a0.x + a2.x, a0.y + a2.y};
    v[2] = {a0.x - a2.x, a0.y - a2.y};
    v[1] = {a1.x + a3r.x, a1.y + a3r.y};
    v[3] = {a1.x - a3r.x, a1.y - a3r.y};
  } else if (R == 3) {
    const T c1 = (T)-0.5;
    const T s1 = (T)(sign < 0 ? -0.86602540378443864676 : 0.86602540378443864676);
    cplx_t<T> a = v[0], b = v[1], c = v[2];
    cplx_t<T> t1 = {b.x + c.x, b.y + c.y};
    cplx_t<T> t2 = {a.x + c1 * t1.x, a.y + c1 * t1.y};
    cplx_t<T> t3 = {s1 * (b.x - c.x), s1 * (b.y - c.y)};
    v[0] = {a.x + t1.x, a.y + t1.y};
    v[1] = {t2.x - t3.y, t2.y + t3.x};
    v[2] = {t2.x + t3.y, t2.y - t3.x};
  } else {  // R == 5, direct
    const T cs[5] = {(T)1.0, (T)0.30901699437494742410, (T)-0.80901699437494742410,
                     (T)-0.80901699437494742410, (T)0.30901699437494742410};
    const T sn0[5] = {(T)0.0, (T)0.95105651629515357212, (T)0.58778525229247312917,
                      (T)-0.58778525229247312917, (T)-0.95105651629515357212};
    cplx_t<T> in[5];
    for (int q = 0; q < 5; q++) in[q] = v[q];
    for (int p = 0; p < 5; p++) {
      T ax = 0, ay = 0;
      for (int q = 0; q < 5; q++) {
        int m = (p * q) % 5;
        T c = cs[m], s = (sign < 0 ? -sn0[m] : sn0[m]);
        ax += in[q].x * c - in[q].y * s;
        ay += in[q].x * s + in[q].y * c;
      }
      v[p] = {ax, ay};
    }
  }
}

// Mixed-radix Stockham FFT of nl lines of length n held in LDS as
// a[p*ls + l] (complex).  b is a second buffer of the same shape.  Radices
// (4,2,3,5) are listed in rad[0..nrad).  tw[m] = exp(sign*2*pi*i*m/n) for the
// sign of this transform.  Returns the buffer holding the result.
template <typename T>
__device__ cplx_t<T>* tile_fft(cplx_t<T>* a, cplx_t<T>* b, int n, int nl, int ls,
                               const int* rad, int nrad, const cplx_t<T>* tw, int sign) {
  int ns = 1;
  for (int s = 0; s < nrad; s++) {
    const int R = rad[s];
    const int nb = n / R;  // butterflies per line
    const int tot = nb * nl;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
      const int l = e % nl;
      const int jb = e / nl;
      const int k = jb % ns;
      const int tstep = k * (n / (ns * R));
      cplx_t<T> v[5];
      for (int r = 0; r < R; r++) {
        cplx_t<T> x = a[(jb + r * nb) * ls + l];
        v[r] = (r == 0) ? x : cmul(x, tw[(tstep * r) % n]);
      }
      dft_small(v, R, sign);
      const int d = (jb / ns) * ns * R + k;
      for (int r = 0; r < R; r++) b[(d + r * ns) * ls + l] = v[r];
    }
    __syncthreads();
    cplx_t<T>* t = a;
    a = b;
    b = t;
    ns *= R;
  }
  return a;
}

// ===================================================== register line engine
// A line tile is L lines x n points handled by L*G threads, thread
// t = l + L*g owning the contiguous segment p in [g*SEG, (g+1)*SEG) of line
// l in registers (n = G*SEG).
//
// The periodic system a x[p-1] + b x[p] + a x[p+1] = d[p] factors as
// (a/alpha)(1 + alpha S^-1)(1 + alpha S) (see LineSolve).  Each recursion
// y[p] = d[p] - alpha y[p-1] is solved per segment from a zero start; the
// true start value is the previous segment's end value, which itself is its
// local end value plus hop = (-alpha)^SEG times the start of the segment
// before.  Summing nhop terms of that geometric series, with
// |alpha|^(nhop*SEG) <= 2^-60, gives the periodic solution, after which
// y[p0+q] = ylocal[q] + (-alpha)^(q+1) * carry.  The backward recursion is
// the mirror image.  Two barriers per solve; no LDS traffic in the chains.
#define SEGMAX 16

struct PSolve {
  double alpha;  // decay factor, |alpha| < 1
  double scale;  // alpha / a
  double hop;    // (-alpha)^seg
  int nhop;      // carry terms
};

struct LShape {
  int n;    // line length
  int L;    // lines per tile
  int G;    // segments per line
  int seg;  // points per segment
};

template <typename T, int SEG>
__device__ __forceinline__ void psolve(T (&v)[SEG], const int G, const int L, const int l, const int g,
                                       const PSolve& op, T* xf, T* xb) {
  const T a = (T)op.alpha;
  const T na = -a;
  T y = (T)0;
#pragma unroll
  for (int q = 0; q < SEG; q++) {
    y = v[q] - a * y;
    v[q] = y;
  }
  xf[g * L + l] = y;
  __syncthreads();
  const T hop = (T)op.hop;
  T cy = (T)0, f = (T)1;
  int gg = g;
  for (int h = 0; h < op.nhop; h++) {
    gg = (gg == 0) ? G - 1 : gg - 1;
    cy += f * xf[gg * L + l];
    f *= hop;
  }
  T pw = na;
#pragma unroll
  for (int q = 0; q < SEG; q++) {
    v[q] += pw * cy;
    pw *= na;
  }
  T z = (T)0;
#pragma unroll
  for (int q = SEG - 1; q >= 0; q--) {
    z = v[q] - a * z;
    v[q] = z;
  }
  xb[g * L + l] = z;
  __syncthreads();
  T cz = (T)0;
  f = (T)1;
  gg = g;
  for (int h = 0; h < op.nhop; h++) {
    gg = (gg == G - 1) ? 0 : gg + 1;
    cz += f * xb[gg * L + l];
    f *= hop;
  }
  const T sc = (T)op.scale;
  pw = na;
#pragma unroll
  for (int q = SEG - 1; q >= 0; q--) {
    v[q] = sc * (v[q] + pw * cz);
    pw *= na;
  }
}

// value at p0-1 (end of the left neighbour segment) / at p0+SEG (start of
// the right one).  One barrier; the buffer must not be reused before the
// next barrier.
template <typename T>
__device__ __forceinline__ T xch_left(const T last, const int G, const int L, const int l, const int g,
                                      T* x) {
  x[g * L + l] = last;
  __syncthreads();
  return x[(g == 0 ? G - 1 : g - 1) * L + l];
}
template <typename T>
__device__ __forceinline__ T xch_right(const T first, const int G, const int L, const int l,
                                       const int g, T* x) {
  x[g * L + l] = first;
  __syncthreads();
  return x[(g == G - 1 ? 0 : g + 1) * L + l];
}

// NaN-propagating max over one wavefront (64 lanes)
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o, 64);
    v = (v != v || w != w) ? NAN : fmax(v, w);
  }
  return v;
}

// sum over one wavefront; the butterfly leaves the same value in every lane
// (a + b == b + a), so the result does not depend on the lane
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// block sums of a and b in a fixed order (lanes, then wavefronts in index
// order); thread 0 stores them to *oa and *ob.  Every thread of the block
// must call it.
__device__ __forceinline__ void block_sum2(double a, double b, double* oa, double* ob) {
  __shared__ double rs[2][16];
  a = wave_sum(a);
  b = wave_sum(b);
  const int wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) {
    rs[0][wv] = a;
    rs[1][wv] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0.0, sb = 0.0;
    for (int q = 0; q < nw; q++) {
      sa += rs[0][q];
      sb += rs[1][q];
    }
    *oa = sa;
    if (ob) *ob = sb;
  }
}

// y-line tile: L threads x C chunks, thread (l, c) = l + L*c; each thread
// holds P neighbouring columns (P = 2 only for FP32 sweeps), so a tile
// covers L*P columns
struct YShape {
  int L, C;
  int P = 1;
};
