// c3d_lines.h -- the register line engine's tile layouts (c3d_dev.h has
// the periodic solve psolve): x-line staging through LDS, z-line tiles read
// directly from HBM, and the helpers the line kernels share.
//
// z-lines: a tile is L consecutive i at one j; thread (l, g) keeps points k
// in [g*seg, (g+1)*seg) of column i = i0+l in registers and reads/writes HBM
// directly (16 FP64 lanes = one 128-byte row piece).  x-lines: a tile is L
// consecutive (j,k) rows; rows are staged through LDS with coalesced row
// loads, then each thread takes its segment.
#pragma once
#include "c3d_impl.h"

// ------------------------------------------------------------ helpers
template <typename T, int SEG>
__device__ __forceinline__ T prv(const T (&v)[SEG], const int q, const T left) {
  return q == 0 ? left : v[q > 0 ? q - 1 : 0];
}
template <typename T, int SEG>
__device__ __forceinline__ T nxt(const T (&v)[SEG], const int q, const T right) {
  return (q + 1 < SEG) ? v[q + 1 < SEG ? q + 1 : SEG - 1] : right;
}

#define SQ _Pragma("unroll") for (int q = 0; q < SEG; q++)
#define SQR _Pragma("unroll") for (int q = SEG - 1; q >= 0; q--)

// LDS carve-up of a line kernel: two staging tiles (x-lines only) and four
// exchange arrays of G*L entries
template <typename T>
struct LBuf {
  T *st0, *st1, *xf, *xb, *xl, *xr;
};
template <typename T>
__device__ __forceinline__ LBuf<T> lbuf(unsigned char* sm, const LShape& sh, bool staged) {
  LBuf<T> b;
  const int tile = staged ? sh.n * (sh.L + 1) : 0;
  b.st0 = (T*)sm;
  b.st1 = b.st0 + tile;
  b.xf = b.st1 + tile;
  b.xb = b.xf + sh.G * sh.L;
  b.xl = b.xb + sh.G * sh.L;
  b.xr = b.xl + sh.G * sh.L;
  return b;
}

// x-line staging.  Lines ln = ln0 + l (l < L) are rows (j, k) with
// j = 1 + ln % N2M, k = ln / N2M.  A tile holds n*L = SEG*blockDim values, so
// each thread moves exactly SEG of them, as SEG/2 pairs of neighbouring
// points: pair m of thread t holds points 2*(t % (n/2)) and the next one of
// tile row t/(n/2) + m*(2*blockDim/n) (the host makes blockDim a multiple of
// n/2 and n is even), one 8-byte (FP32) or 16-byte (FP64) load or store per
// pair.  A wave's load then carries at least 512 bytes: single 4-byte loads
// halved the bytes in flight per instruction and measured ~20 % slower in
// the FP32 line kernels, and 16-byte FP64 pairs beat single 8-byte loads by
// 8-13 %.  pre() issues the global loads of a tile into registers (so they
// can be in flight while the previous tile is used); put() writes them to
// LDS, converting to the tile's type; pre_raw() keeps a field's own type in
// the registers, so that the conversion waits for put().  Two LDS tiles
// alternate and every put()/out() starts with a barrier, so a tile stays
// readable until the second staging operation after the one that filled it
// begins.  Offsets are 32-bit (fields hold fewer than 2^31 elements).
// NTL: the loads carry the nontemporal hint, for kernels whose fields no
// later kernel reads soon again (measured: the x-line updates 9-13 us
// faster; the RHS x-pass slower, since the z-pass re-reads its fields).
template <typename T, int SEG, bool NTL = false>
struct XStage {
  T *st0, *st1;
  int sb;
  int nx, LS;
  int l, p0, pm, pp;      // this thread's line, segment start, and periodic neighbours
  unsigned go2[SEG / 2];  // element offset of pair m's first point (row j)
  unsigned vm2;           // bit m set: pair m exists (its line is in the grid)
  unsigned j12;           // bit m set: pair m lies in row j = 1
  int li2, lstep2;        // LDS index of pair m's first point: li2 + m*lstep2 (second: + LS)
  __device__ __forceinline__ T* next() {
    sb ^= 1;
    return sb ? st1 : st0;
  }
  // pair m of r holds r[2m], r[2m+1]
  template <typename S>
  __device__ __forceinline__ void pre_pairs(const S* F, unsigned off, S (&r)[SEG]) const {
    static_assert(sizeof(S) == 4 || sizeof(S) == 8, "4- or 8-byte fields");
#pragma unroll
    for (int m = 0; m < SEG / 2; m++) {
      S x0 = (S)0, x1 = (S)0;
      if ((vm2 >> m) & 1u) {
        // the offset is summed in 32 bits first: off wraps for joff < 0
        typedef float nf2 __attribute__((ext_vector_type(2)));
        typedef double nd2 __attribute__((ext_vector_type(2)));
        if constexpr (sizeof(S) == 4) {
          const nf2* a = (const nf2*)(F + (go2[m] + off));
          const nf2 v = NTL ? __builtin_nontemporal_load(a) : *a;
          x0 = (S)v.x;
          x1 = (S)v.y;
        } else {
          const nd2* a = (const nd2*)(F + (go2[m] + off));
          const nd2 v = NTL ? __builtin_nontemporal_load(a) : *a;
          x0 = (S)v.x;
          x1 = (S)v.y;
        }
      }
      r[2 * m] = x0;
      r[2 * m + 1] = x1;
    }
  }
  template <typename S>
  __device__ __forceinline__ void pre(const S* F, int joff, T (&r)[SEG]) const {
    const unsigned off = (unsigned)(joff * nx);  // row j+joff exists for every j (ghost planes)
    S t[SEG];
    pre_pairs(F, off, t);
    SQ r[q] = (T)t[q];
  }
  template <typename S>
  __device__ __forceinline__ void pre_raw(const S* F, int joff, S (&r)[SEG]) const {
    pre_pairs(F, (unsigned)(joff * nx), r);
  }
  template <typename S>
  __device__ __forceinline__ const T* put(const S (&r)[SEG]) {
    T* B = next();
    __syncthreads();
#pragma unroll
    for (int m = 0; m < SEG / 2; m++) {
      B[li2 + m * lstep2] = (T)r[2 * m];
      B[li2 + LS + m * lstep2] = (T)r[2 * m + 1];
    }
    __syncthreads();
    return B;
  }
  // element at segment offset q in [-1, SEG] of this thread's line
  __device__ __forceinline__ T at(const T* B, int q) const {
    const int p = q < 0 ? pm : (q >= SEG ? pp : p0 + q);
    return B[p * LS + l];
  }
  __device__ __forceinline__ bool keep2(int m, int jmin) const {
    return ((vm2 >> m) & 1u) && (jmin <= 1 || !((j12 >> m) & 1u));
  }
  template <typename S>
  __device__ __forceinline__ void out_pairs(S* F, const T* B, int jmin) const {
#pragma unroll
    for (int m = 0; m < SEG / 2; m++)
      if (keep2(m, jmin)) {
        const T x0 = B[li2 + m * lstep2], x1 = B[li2 + LS + m * lstep2];
        // pair stores with the nontemporal hint (st_nt); the builtin takes
        // clang vector types, not HIP's float2/double2 structs
        typedef float nf2 __attribute__((ext_vector_type(2)));
        typedef double nd2 __attribute__((ext_vector_type(2)));
        if constexpr (sizeof(S) == 4)
          st_nt((nf2*)(F + go2[m]), nf2{(float)x0, (float)x1});
        else
          st_nt((nd2*)(F + go2[m]), nd2{(double)x0, (double)x1});
      }
  }
  // the same tile to two arrays (a field and its FP32 shadow); rows with
  // j < jmin are left untouched
  template <typename S1, typename S2>
  __device__ __forceinline__ void out2(S1* F1, S2* F2, const T (&v)[SEG], int jmin = 1) {
    T* B = next();
    __syncthreads();
    SQ B[(p0 + q) * LS + l] = v[q];
    __syncthreads();
    out_pairs(F1, B, jmin);
    out_pairs(F2, B, jmin);
  }
  template <typename S>
  __device__ __forceinline__ void out(S* F, const T (&v)[SEG], int jmin = 1) {
    T* B = next();
    __syncthreads();
    SQ B[(p0 + q) * LS + l] = v[q];
    __syncthreads();
    out_pairs(F, B, jmin);
  }
};

// Tile order across the 8 XCDs: the dispatcher hands workgroup b to XCD
// b % 8, and each XCD has its own L2.  Tile t = xcd_tile(b, nb) gives XCD x
// the contiguous tile range [x*per + min(x, rem), ...), so the z tiles that
// share neighbour planes (j-1 and j+1, nit tiles apart) run on one L2.  A
// bijection of [0, nb) for any nb.  (x tiles hold 8 consecutive j of one k
// and keep the plain order, which measured faster for them.)
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int per = nb >> 3, rem = nb & 7, x = b & 7, r = b >> 3;
  return x * per + min(x, rem) + r;
}

template <typename T, int SEG, bool NTL>
__device__ __forceinline__ void xstage(XStage<T, SEG, NTL>& x, const LBuf<T>& b, const Geo& g, const LShape& sh,
                                       int l, int p0) {
  x.st0 = b.st0;
  x.st1 = b.st1;
  x.sb = 1;
  x.nx = g.nx;
  x.LS = sh.L + 1;
  x.l = l;
  x.p0 = p0;
  x.pm = p0 == 0 ? sh.n - 1 : p0 - 1;
  x.pp = p0 + SEG == sh.n ? 0 : p0 + SEG;
  const int ln0 = blockIdx.x * sh.L;
  const int nlines = g.n2m * g.nz;
  const int nt = blockDim.x;
  const int t = threadIdx.x;
  const int nh = sh.n >> 1;
  const int pq = t % nh, lq = t / nh;
  x.lstep2 = nt / nh;
  x.li2 = 2 * pq * x.LS + lq;
  x.j12 = 0u;
  x.vm2 = 0u;
#pragma unroll
  for (int m = 0; m < SEG / 2; m++) {
    const int ll = lq + m * x.lstep2;
    const int ln = ln0 + ll;
    const int jj = 1 + ln % g.n2m;
    const bool v = ll < sh.L && ln < nlines;
    x.go2[m] = v ? (unsigned)gidx(g, 2 * pq, jj, ln / g.n2m) : 0u;
    if (v) x.vm2 |= 1u << m;
    if (jj == 1) x.j12 |= 1u << m;
  }
}

#define XSETUP XSETUP_NTL(false)
// the same with nontemporal staging loads (XStage's NTL)
#define XSETUP_NT XSETUP_NTL(true)
#define XSETUP_NTL(NTL_)                                              \
  extern __shared__ __align__(16) unsigned char smem[];               \
  const Geo g = c.g;                                                  \
  const LShape sh = c.o.lx;                                           \
  const int L = sh.L, G = sh.G;                                       \
  const LBuf<T> bf = lbuf<T>(smem, sh, true);                         \
  const int l = threadIdx.x % L, gs = threadIdx.x / L, p0 = gs * SEG; \
  XStage<T, SEG, NTL_> X;                                             \
  xstage(X, bf, g, sh, l, p0);                                        \
  const int ln = blockIdx.x * L + l;                                  \
  const bool ok = ln < g.n2m * g.nz;                                  \
  const int j = 1 + (ok ? ln % g.n2m : 0);                            \
  (void)j;
#define XA(B, q) X.at(B, q)

// z-lines: a tile is L consecutive i at one j (up to 512 threads); column
// i = i0 + l, k = p
#define ZSETUP ZSETUP_J0(1)
// the same for tiles whose first plane is J0 (tile / nit = J - J0)
#define ZSETUP_J0(J0)                                                 \
  extern __shared__ __align__(16) unsigned char smem[];               \
  const Geo g = c.g;                                                  \
  const LShape sh = sizeof(T) == 4 ? c.o.lz32 : c.o.lz;             \
  const int n = sh.n, L = sh.L, G = sh.G;                             \
  const LBuf<T> bf = lbuf<T>(smem, sh, false);                        \
  const int l = threadIdx.x % L, gs = threadIdx.x / L, p0 = gs * SEG; \
  const int nit = (g.nx + L - 1) / L;                                 \
  const int tile = xcd_tile(blockIdx.x, gridDim.x);                   \
  const int j = (J0) + tile / nit;                                    \
  const int i = (tile % nit) * L + l;                                 \
  const bool ok = i < g.nx;                                           \
  const unsigned pl = (unsigned)g.plane;                              \
  const unsigned b = (unsigned)gidx(g, ok ? i : 0, j, 0);             \
  const unsigned bm = b + (unsigned)(p0 == 0 ? n - 1 : p0 - 1) * pl;  \
  const unsigned bp = b + (unsigned)(p0 + SEG == n ? 0 : p0 + SEG) * pl;
// element offsets are 32-bit unsigned (fields hold fewer than 2^31
// elements), which lets loads use a uniform base plus a 32-bit lane offset
#define ZI(q) (b + (unsigned)(p0 + (q)) * pl)
#define PZ (sizeof(T) == 4 ? c.o.pz32 : c.o.pz)

template <typename T, int SEG, typename S>
__device__ __forceinline__ void zload(const S* F, unsigned b, unsigned pl, int p0, unsigned bm, unsigned bp,
                                      T (&v)[SEG], T& left, T& right) {
  SQ v[q] = (T)F[b + (unsigned)(p0 + q) * pl];
  left = (T)F[bm];
  right = (T)F[bp];
}
// the same with the segment's own points loaded nontemporally (ld_nt); the
// two neighbour points, which the adjacent segments also read, plainly
template <typename T, int SEG, typename S>
__device__ __forceinline__ void zload_nt(const S* F, unsigned b, unsigned pl, int p0, unsigned bm, unsigned bp,
                                         T (&v)[SEG], T& left, T& right) {
  SQ v[q] = (T)ld_nt(F + (b + (unsigned)(p0 + q) * pl));
  left = (T)F[bm];
  right = (T)F[bp];
}

