// ekm_mgt.hip -- the consistent pressure projection (proj_mode 1 and 2; no
// counterpart in the reference).
//
// The Poisson rows of the cells the divergence diagnostic covers (J=2..NY-1,
// K=2..NZ-1, src/duct.F90:2790-2800) are the divergence (:2754-2760) of the
// projection's own face gradient (:2713-2724) and of the CU1 centre
// correction (:2731), so J div u after the projection equals the residual
// of the solve on those cells.  Every other row keeps MUDPACK's
// discretisation (host_mg_level), so the boundary closure and the coarse
// operators are the reference's.
//
// Solver: residual iterations in FP64 on the finest level; each correction
// is one V(1,1) cycle in T (double: proj_mode 1; float: proj_mode 2, the
// mixed-precision form) with alternating zebra-line Gauss-Seidel smoothing
// (z lines of both parities, then y lines of both parities), full weighting
// and bilinear prolongation; the coarsest level gets 20 smoothing steps.
// Line smoothing handles the index-space anisotropy of the stretched grids
// that makes MUDPACK's point relaxation stall.
//
// Layout and work split: one workgroup per real solve (2 per live kx mode),
// as for the MUDPACK replica.  The tridiagonal line matrices depend only on
// the level and kx, so their Thomas factors (1/pivot, -lower/pivot, upper/
// pivot) are computed once at set-up per live mode.  A z line (contiguous in
// memory) is solved by one wavefront: each lane owns a chunk of CM points and
// the two first-order recurrences of the substitution are composed as affine
// maps with a wave-level scan.  y lines (stride NZ+2) are solved by lanes over
// columns, each lane running the substitution down its column, so every load
// of a wavefront is contiguous across lanes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ekm_host.h"
#include "ekm_internal.h"

namespace {

constexpr int kMaxLevT = 16;
// threads per workgroup of the solve kernel: 16 wavefronts for the FP32
// cycles without cross terms (their registers fit the 128 a 1024-thread
// workgroup allows; long z lines stream their loads), else 8.  On the
// 1024x385x769 case the 16-wavefront form takes the solve from 110 to 91 ms
// per launch (the workgroup is bound by its own memory round trips, which
// twice the wavefronts overlap)
constexpr int tight_block(bool fp32, int cm, bool cross) {
  (void)cm;
  return (fp32 && !cross) ? 1024 : 512;
}
constexpr int kDirectMax = 128;  // the coarse level solved directly has at most this many points
constexpr int kMaxProg = 4 * kMaxLevT + 4;

template <typename T>
struct TLev {
  int nx, ny, n;
  int e_off, r_off, res_off;  // per-solve workspace offsets (elements of T)
  int t2_off;                 // second line-solve scratch (the residual array is the first)
  int lds;                    // 1: the four arrays above live in the workgroup's LDS
  int f_off;                  // per-mode offset: z factors [3][n], y factors [3][n] (column-parity layout), c9 [n],
                              // y-chunk products A [n] and Be [n] (column-parity layout); periodic ends
                              // add the z lines' wrap vectors W [2][n] and their 2x2 matrices [ny][4]
  const T* cf;                // c1..c8 planes [8][n]; rows of specified points are 0 (c9 = 1)
  const T* cfy;               // the same planes in the column-parity layout of the y-line smoother
  const unsigned char* spec;  // specified (Dirichlet) points
};

template <typename T>
struct TDesc {
  int ngrid, nfx, nfy, maxit;
  double tol;
  long ws;          // T elements per solve
  long ws64;        // doubles per solve: phi (ring), b; with krylov also r, q, Qp [n] and Zp (ring)
  long fstride;     // factor elements per mode
  int lc;           // coarsest level of the cycle, solved with its inverse
  long ainv_off;    // per-mode offset of that inverse (n_lc x n_lc, row-major)
  const T* fac;     // [mode][fstride]
  const double* cf64;  // finest level c1..c8 in double [8][n]
  const double* c964;  // finest level c9 in double, [mode][n]
  // singular kx = 0 problem (no specified side): the left null vector of the
  // finest operator scaled to 1 at the pin point (finest index `pin`), or null
  const double* wnull;
  int pin;
  // the V(1,1) cycle as a program, one step per byte: (level << 2) | op,
  // op 0 = smooth, residual and restriction to the next coarser level,
  // 1 = prolongation from the next coarser level and smooth, 2 = direct solve
  int krylov;  // 1: GCR(1) outer iteration (MG-preconditioned minimal residual with one kept direction)
  int nprog;
  unsigned char prog[kMaxProg];
  TLev<T> lev[kMaxLevT + 1];
};

// The per-solve workspace: global memory, and the LDS that holds the coarse
// levels; a level's arrays are at its offsets in one of the two
template <typename T>
struct Ws {
  T* g;
  T* s;
};
template <typename T>
__device__ inline T* wbase(const TLev<T>& L, const Ws<T>& W) {
  return L.lds ? W.s : W.g;
}

// affine map x -> a*x + b; compose(outer, inner) = outer(inner(x))
template <typename T>
struct Aff {
  T a, b;
};
template <typename T>
__device__ inline Aff<T> compose(const Aff<T>& o, const Aff<T>& i) {
  return Aff<T>{o.a * i.a, o.a * i.b + o.b};
}

// Cross-lane moves of the wavefront scans.  dpp<CTRL, RM>(old, v): v of the
// lane that the DPP control CTRL names, or `old` where that lane lies outside
// the row or the row is masked out by RM (row_shr:n = 0x110+n, row_shl:n =
// 0x100+n, row_bcast:15 = 0x142, row_bcast:31 = 0x143; 64-bit values move as
// two 32-bit halves).
template <int CTRL, int RM, typename T>
__device__ inline T dpp(T old, T v) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL, RM, 0xf, false));
  } else {
    const long long o = __double_as_longlong(old), x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)x, CTRL, RM, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(x >> 32), CTRL, RM, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}
template <typename T>
__device__ inline T rdlane(T v, int l) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
  } else {
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)x, l), hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}
template <int CTRL, int RM, typename T>
__device__ inline Aff<T> dpp_aff(const Aff<T>& v) {
  return Aff<T>{dpp<CTRL, RM>((T)1, v.a), dpp<CTRL, RM>((T)0, v.b)};
}

// inclusive scan over the lanes in increasing order: lane i gets
// map_i o map_{i-1} o ... o map_0 (in-row doubling steps, then the row
// prefixes carried by row_bcast:15 to rows 1 and 3 and row_bcast:31 to rows
// 2 and 3)
template <typename T>
__device__ inline Aff<T> wave_scan_up(Aff<T> v) {
  v = compose(v, dpp_aff<0x111, 0xf>(v));
  v = compose(v, dpp_aff<0x112, 0xf>(v));
  v = compose(v, dpp_aff<0x114, 0xf>(v));
  v = compose(v, dpp_aff<0x118, 0xf>(v));
  v = compose(v, dpp_aff<0x142, 0xa>(v));
  v = compose(v, dpp_aff<0x143, 0xc>(v));
  return v;
}

// inclusive scan in decreasing order: lane i gets map_i o ... o map_63
// (in-row doubling steps, then the suffix maps of the rows above, read from
// the first lane of each row)
template <typename T>
__device__ inline Aff<T> wave_scan_down(Aff<T> v) {
  v = compose(v, dpp_aff<0x101, 0xf>(v));
  v = compose(v, dpp_aff<0x102, 0xf>(v));
  v = compose(v, dpp_aff<0x104, 0xf>(v));
  v = compose(v, dpp_aff<0x108, 0xf>(v));
  const Aff<T> r1{rdlane(v.a, 16), rdlane(v.b, 16)};
  const Aff<T> r2{rdlane(v.a, 32), rdlane(v.b, 32)};
  const Aff<T> r3{rdlane(v.a, 48), rdlane(v.b, 48)};
  const Aff<T> s2 = r3, s1 = compose(r2, r3), s0 = compose(r1, s1);
  const int row = (threadIdx.x & 63) >> 4;
  const Aff<T> tail = row == 0 ? s0 : row == 1 ? s1 : row == 2 ? s2 : Aff<T>{(T)1, (T)0};
  return compose(v, tail);
}

// Zebra line relaxation of level L.  The Thomas factors of every line are
// precomputed (1/piv, -lower/piv, upper/piv), so a line solve is the forward
// recurrence dp_k = a_k dp_{k-1} + s_k/piv_k and the back substitution
// x_k = dp_k - c_k x_{k+1}, both first-order affine recurrences.
//
// z lines (rows j of parity par, contiguous in memory): one wavefront per row;
// point k = 64 c + lane sits in lane `lane` of stripe c, so every load and
// store of a stripe is one contiguous run; within a stripe the recurrence is a
// 64-lane scan of affine maps, carried from stripe to stripe.  ZERO: the
// correction is zero on the neighbouring rows (the first sweep of a cycle's
// down leg), so they are not read.
//
// Lines of more than kZRegStripes stripes are streamed: each stripe's loads
// are issued while the previous stripe is scanned and only the forward
// results stay in registers, which keeps the kernel's register count (and so
// its occupancy) independent of the line length.  The arithmetic is the same
// in both forms.
constexpr int kZRegStripes = 5;

// Periodic z ends (MUDPACK nxa = nxb = 0): every column of a level is an
// unknown with its own row, and the virtual neighbours are e(-1) = e(nx-2)
// and e(nx) = e(1) (MUDPACK's periodic points, src/mud2cr.f:1693-1795).
// They are held in the ring columns of the correction, so every writer of
// columns 1 and nx-2 also writes the ring copy; a value at index q of
// column i goes to q + ring_off_hi(i) / q + ring_off_lo(i) as well.
__device__ inline void ring_store_d(double* a, int q, int i, int nx, double v) {
  if (i == 1) a[q + nx - 1] = v;
  if (i == nx - 2) a[q - (nx - 1)] = v;
}
template <typename T>
__device__ inline void ring_store(T* a, int q, int i, int nx, T v) {
  if (i == 1) a[q + nx - 1] = v;
  if (i == nx - 2) a[q - (nx - 1)] = v;
}

// A z line of a periodic level is the chain 0..nx-1 (the precomputed Thomas
// factors) plus the couplings of point 0 to nx-2 and of point nx-1 to 1.
// With T the chain, U = [c5(0) e_0, c1(nx-1) e_{nx-1}] and V = [e_{nx-2}, e_1],
// Woodbury gives x = y - W Minv (y_{nx-2}, y_1) for y = T^-1 s, W = T^-1 U and
// Minv = (I + V^T W)^-1, both W and Minv precomputed per row and mode.  y is
// the chain solution the wavefront holds in x[] (stripe c, lane l: point
// 64 c + l); this writes the line with its ring copies.
template <typename T, int CM>
__device__ inline void zline_periodic(const TLev<T>& L, T* __restrict__ e, const T* __restrict__ fz, const T* x,
                                      int j, int lane) {
  const int nx = L.nx, n = L.n, ld = nx + 2;
  const T* __restrict__ w1 = fz + 9 * n;
  const T* __restrict__ w2 = w1 + n;
  const T* mi = w2 + n + 4 * j;
  const int im = nx - 2, cs = im >> 6, ls = im & 63;
  T xm = (T)0;
#pragma unroll
  for (int c = 0; c < CM; ++c)
    if (c == cs) xm = x[c];
  const T y1 = rdlane(x[0], 1), ym = rdlane(xm, ls);
  const T g0 = mi[0] * ym + mi[1] * y1, g1 = mi[2] * ym + mi[3] * y1;
#pragma unroll
  for (int c = 0; c < CM; ++c) {
    const int i = 64 * c + lane;
    if (i < nx) {
      const int p = i + nx * j;
      const T v = x[c] - (g0 * w1[p] + g1 * w2[p]);
      const int q = (i + 1) + ld * (j + 1);
      e[q] = v;
      ring_store(e, q, i, nx, v);
    }
  }
}

template <typename T, bool X, bool ZERO>
__device__ inline void zline_point(const T* __restrict__ e, const T* __restrict__ r, const T* __restrict__ cf,
                                   const T* fa, const T* f1, int nx, int n, int ld, int i, int j, T& a, T& b) {
  a = (T)1;
  b = (T)0;
  if (i < nx) {
    const int p = i + nx * j;
    const int q = (i + 1) + ld * (j + 1);
    // every load of the point is issued before any is used
    const T rv = r[p], av = fa[p], fv = f1[p];
    T c2 = (T)0, c6 = (T)0, eu = (T)0, ed = (T)0;
    T c1 = (T)0, c3 = (T)0, c5 = (T)0, c7 = (T)0, e1 = (T)0, e3 = (T)0, e5 = (T)0, e7 = (T)0;
    if (!ZERO) {
      c2 = cf[2 * n + p];
      c6 = cf[6 * n + p];
      eu = e[q + ld];
      ed = e[q - ld];
      if (X) {
        c1 = cf[1 * n + p];
        c3 = cf[3 * n + p];
        c5 = cf[5 * n + p];
        c7 = cf[7 * n + p];
        e1 = e[q + 1 + ld];
        e3 = e[q - 1 + ld];
        e5 = e[q - 1 - ld];
        e7 = e[q + 1 - ld];
      }
    }
    T s = rv;
    if (!ZERO) {
      s -= c2 * eu + c6 * ed;
      if (X) s -= (c1 * e1 + c3 * e3) + (c5 * e5 + c7 * e7);
    }
    a = av;
    b = s * fv;
  }
}

template <typename T, int CM, bool X, bool PER, bool ZERO = false>
__device__ void zlines(const TLev<T>& L, const Ws<T>& Wp, const T* __restrict__ fac, int par) {
  T* W = wbase(L, Wp);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int nx = L.nx, ny = L.ny, ld = nx + 2, n = L.n;
  T* __restrict__ e = W + L.e_off;
  const T* __restrict__ r = W + L.r_off;
  const T* __restrict__ cf = L.cf;
  const T* f1 = fac + L.f_off;
  const T* fa = f1 + n;
  const T* fc = fa + n;
  if constexpr (CM > kZRegStripes) {
    for (int j = par + 2 * wave; j < ny; j += 2 * nw) {
      T dk[CM];
      T an, bn;
      zline_point<T, X, ZERO>(e, r, cf, fa, f1, nx, n, ld, lane, j, an, bn);
      T carry = (T)0;
#pragma unroll
      for (int c = 0; c < CM; ++c) {
        const T a = an, b = bn;
        if (c + 1 < CM) zline_point<T, X, ZERO>(e, r, cf, fa, f1, nx, n, ld, 64 * (c + 1) + lane, j, an, bn);
        const Aff<T> inc = wave_scan_up(Aff<T>{a, b});
        dk[c] = inc.a * carry + inc.b;
        carry = rdlane(dk[c], 63);
      }
      T cn = (T)0;
      {
        const int i = 64 * (CM - 1) + lane;
        if (i < nx) cn = fc[i + nx * j];
      }
      T xo = (T)0;
#pragma unroll
      for (int c = CM - 1; c >= 0; --c) {
        const T ck = cn;
        if (c > 0) {
          const int i = 64 * (c - 1) + lane;
          cn = i < nx ? fc[i + nx * j] : (T)0;
        }
        const Aff<T> inc = wave_scan_down(Aff<T>{-ck, dk[c]});
        const T x = inc.a * xo + inc.b;
        const int i = 64 * c + lane;
        if (!PER && i < nx) e[(i + 1) + ld * (j + 1)] = x;
        dk[c] = x;
        xo = rdlane(x, 0);
      }
      if constexpr (PER) zline_periodic<T, CM>(L, e, f1, dk, j, lane);
    }
    return;
  }
  for (int j = par + 2 * wave; j < ny; j += 2 * nw) {
    T ak[CM], bk[CM], ck[CM], dk[CM];
#pragma unroll
    for (int c = 0; c < CM; ++c) {
      const int i = 64 * c + lane;
      ak[c] = (T)1;
      bk[c] = (T)0;
      ck[c] = (T)0;
      if (i < nx) {
        const int p = i + nx * j;
        const int q = (i + 1) + ld * (j + 1);
        T s = r[p];
        if (!ZERO) {
          s -= cf[2 * n + p] * e[q + ld] + cf[6 * n + p] * e[q - ld];
          if (X)
            s -= (cf[1 * n + p] * e[q + 1 + ld] + cf[3 * n + p] * e[q - 1 + ld]) +
                 (cf[5 * n + p] * e[q - 1 - ld] + cf[7 * n + p] * e[q + 1 - ld]);
        }
        ak[c] = fa[p];
        bk[c] = s * f1[p];
        ck[c] = fc[p];
      }
    }
    T carry = (T)0;  // dp at the last point of the previous stripe
#pragma unroll
    for (int c = 0; c < CM; ++c) {
      const Aff<T> inc = wave_scan_up(Aff<T>{ak[c], bk[c]});
      dk[c] = inc.a * carry + inc.b;
      carry = rdlane(dk[c], 63);
    }
    T xo = (T)0;  // x at the first point of the next stripe
#pragma unroll
    for (int c = CM - 1; c >= 0; --c) {
      const Aff<T> inc = wave_scan_down(Aff<T>{-ck[c], dk[c]});
      const T x = inc.a * xo + inc.b;
      const int i = 64 * c + lane;
      if (!PER && i < nx) e[(i + 1) + ld * (j + 1)] = x;
      dk[c] = x;
      xo = rdlane(x, 0);
    }
    if constexpr (PER) zline_periodic<T, CM>(L, e, f1, dk, j, lane);
  }
}

// y lines (columns i of parity par, stride nx+2): lanes over 64 columns of a
// group, so every load of a row is contiguous across lanes; the rows are split
// into chunks worked by different wavefronts.  Each wavefront runs its
// chunk's recurrences as affine maps of the unknown value entering the chunk
// (stored per point in the two scratch arrays), the chunk maps are combined
// through LDS, and a last pass applies the entering values.
//
// The line factors, the coefficient planes the smoother reads and the two
// scratch arrays are kept in a column-parity layout: the columns of parity 0
// (i = 2c) first, row by row (c + ncol0 j), then those of parity 1
// (ncol0 ny + c + ncol1 j).  A wavefront's row load then covers whole cache
// lines; in the natural layout (i + nx j) every load of one parity fetched
// the lines of both.  The right-hand side and the correction keep the natural
// layout the z lines and the transfers use.
__host__ __device__ inline int ysplit(int i, int j, int nx, int ny) {
  const int ncol0 = (nx + 1) >> 1;
  return (i & 1) ? ncol0 * ny + (i >> 1) + (nx >> 1) * j : (i >> 1) + ncol0 * j;
}
constexpr int kYMaxTasks = 32;
// the rows of a y-line pass split into chunks: nch chunks of cl rows per
// 64-column group (ngrp groups), one task per (group, chunk)
struct YChunks {
  int ncol, ngrp, nch, cl;
};
__host__ __device__ inline YChunks ychunks(int nx, int ny, int par, int nw) {
  YChunks c;
  c.ncol = (nx - par + 1) / 2;
  c.ngrp = (c.ncol + 63) / 64;
  c.nch = nw / c.ngrp;
  if (c.nch < 1) c.nch = 1;
  if (c.nch > ny) c.nch = ny;
  if (c.ngrp * c.nch > kYMaxTasks) c.nch = kYMaxTasks / c.ngrp;
  c.cl = (ny + c.nch - 1) / c.nch;
  return c;
}
constexpr int kYUnroll = 4;
constexpr int kYApply = 8;  // rows per batch of the last pass
constexpr int kYBack = 6;   // rows per batch of the backward pass

template <typename T, bool X, bool PER>
__device__ void ylines(const TLev<T>& L, const Ws<T>& Wp, const T* __restrict__ fac, int par, T* lds) {
  T* W = wbase(L, Wp);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int nx = L.nx, ny = L.ny, ld = nx + 2, n = L.n;
  T* __restrict__ e = W + L.e_off;
  const T* __restrict__ r = W + L.r_off;
  T* __restrict__ tA = W + L.res_off;
  T* __restrict__ tB = W + L.t2_off;
  const T* __restrict__ cf = L.cfy;
  const T* f1 = fac + L.f_off + 3 * n;
  const T* fa = f1 + n;
  const T* fc = fa + n;
  // the data-independent parts of the chunk maps, precomputed per mode:
  // yA = products of fa from the chunk's first row, yBe = products of -fc
  // from its last row
  const T* yA = fac + L.f_off + 7 * n;
  const T* yBe = yA + n;
  const YChunks yc = ychunks(nx, ny, par, nw);
  const int ncol = yc.ncol, nch = yc.nch, cl = yc.cl;
  // column-parity layout: this parity's block starts at pb, rows ncol apart
  const int pb = par ? ((nx + 1) >> 1) * ny : 0;
  const int ntask = yc.ngrp * nch;
  const int ngrp = yc.ngrp;
  (void)tA;
  // LDS: per task and lane: chunk map (a, b), then the entering value
  T* sa = lds;
  T* sb = sa + kYMaxTasks * 64;
  T* sv = sb + kYMaxTasks * 64;
  // forward: dp_j = A_j dp_in + B_j within the chunk; only B is formed here
  for (int task = wave; task < ntask; task += nw) {
    const int grp = task / nch, ch = task - grp * nch;
    const int col = grp * 64 + lane;
    const int i = par + 2 * col;
    const int j0 = ch * cl, j1 = min(j0 + cl, ny);
    T A = (T)1, B = (T)0;
    if (i < nx) {
      if (j1 > j0) A = yA[pb + col + ncol * (j1 - 1)];
      for (int jb = j0; jb < j1; jb += kYUnroll) {
        T sv_[kYUnroll], av[kYUnroll], fv[kYUnroll];
#pragma unroll
        for (int u = 0; u < kYUnroll; ++u) {
          const int j = jb + u;
          sv_[u] = (T)0;
          av[u] = (T)1;
          fv[u] = (T)0;
          if (j < j1) {
            const int p = i + nx * j;
            const int ps = pb + col + ncol * j;
            const int q = (i + 1) + ld * (j + 1);
            sv_[u] = r[p] - (cf[0 * n + ps] * e[q + 1] + cf[4 * n + ps] * e[q - 1]);
            if (X)
              sv_[u] -= (cf[1 * n + ps] * e[q + 1 + ld] + cf[3 * n + ps] * e[q - 1 + ld]) +
                        (cf[5 * n + ps] * e[q - 1 - ld] + cf[7 * n + ps] * e[q + 1 - ld]);
            av[u] = fa[ps];
            fv[u] = f1[ps];
          }
        }
#pragma unroll
        for (int u = 0; u < kYUnroll; ++u) {
          const int j = jb + u;
          if (j < j1) {
            B = av[u] * B + sv_[u] * fv[u];
            tB[pb + col + ncol * j] = B;
          }
        }
      }
    }
    sa[task * 64 + lane] = A;
    sb[task * 64 + lane] = B;
  }
  __syncthreads();
  // entering dp of every chunk, chunks in order
  for (int grp = wave; grp < ngrp; grp += nw) {
    T dp = (T)0;
    for (int ch = 0; ch < nch; ++ch) {
      const int t = (grp * nch + ch) * 64 + lane;
      sv[t] = dp;
      dp = sa[t] * dp + sb[t];
    }
  }
  __syncthreads();
  // backward: x_j = alpha_j + beta_j x_out within the chunk.  The first
  // chunk enters with dp = 0 (its dp_j are the B_j), and the last chunk
  // leaves with x_out = 0, so its alpha_j are the solution and go straight
  // to e (the last pass skips that chunk).
  for (int task = wave; task < ntask; task += nw) {
    const int grp = task / nch, ch = task - grp * nch;
    const int col = grp * 64 + lane;
    const int i = par + 2 * col;
    const int j0 = ch * cl, j1 = min(j0 + cl, ny);
    const bool first = ch == 0, last = ch == nch - 1;
    const T din = sv[task * 64 + lane];
    T al = (T)0, be = (T)1;
    if (i < nx) {
      if (j1 > j0) be = yBe[pb + col + ncol * j0];
      for (int jb = j1 - 1; jb >= j0; jb -= kYBack) {
        T a_[kYBack], b_[kYBack], c_[kYBack];
#pragma unroll
        for (int u = 0; u < kYBack; ++u) {
          const int j = jb - u;
          a_[u] = (T)0;
          if (j >= j0) {
            const int ps = pb + col + ncol * j;
            if (!first) a_[u] = yA[ps];
            b_[u] = tB[ps];
            c_[u] = fc[ps];
          }
        }
#pragma unroll
        for (int u = 0; u < kYBack; ++u) {
          const int j = jb - u;
          if (j >= j0) {
            const T dp = a_[u] * din + b_[u];
            al = dp - c_[u] * al;
            if (last) {
              const int q = (i + 1) + ld * (j + 1);
              e[q] = al;
              if constexpr (PER) ring_store(e, q, i, nx, al);
            } else {
              tB[pb + col + ncol * j] = al;
            }
          }
        }
      }
    }
    sa[task * 64 + lane] = al;
    sb[task * 64 + lane] = be;
  }
  __syncthreads();
  // x entering each chunk from above, chunks in reverse order
  for (int grp = wave; grp < ngrp; grp += nw) {
    T x = (T)0;
    for (int ch = nch - 1; ch >= 0; --ch) {
      const int t = (grp * nch + ch) * 64 + lane;
      sv[t] = x;
      x = sa[t] + sb[t] * x;
    }
  }
  __syncthreads();
  for (int task = wave; task < ntask; task += nw) {
    const int grp = task / nch, ch = task - grp * nch;
    const int col = grp * 64 + lane;
    const int i = par + 2 * col;
    if (i >= nx || ch == nch - 1) continue;
    const int j0 = ch * cl, j1 = min(j0 + cl, ny);
    const T xo = sv[task * 64 + lane];
    for (int jb = j0; jb < j1; jb += kYApply) {
      T a_[kYApply], b_[kYApply];
#pragma unroll
      for (int u = 0; u < kYApply; ++u) {
        const int j = min(jb + u, j1 - 1);
        a_[u] = tB[pb + col + ncol * j];
        b_[u] = yBe[pb + col + ncol * j];
      }
#pragma unroll
      for (int u = 0; u < kYApply; ++u)
        if (jb + u < j1) {
          const int q = (i + 1) + ld * (jb + u + 1);
          const T v = a_[u] + b_[u] * xo;
          e[q] = v;
          if constexpr (PER) ring_store(e, q, i, nx, v);
        }
    }
  }
  __syncthreads();
}

// One alternating zebra sweep.  The rows of parity 0 are solved first and
// read only the rows of parity 1, so a correction that is zero on entry (the
// down leg: nothing has been written since the level was restricted to) is
// not read at all, and the prolongation of the up leg need only have updated
// the rows of parity 1.  The interior of the level's correction is fully
// rewritten by the two z sweeps; its ring stays zero (with periodic ends its
// ring columns are rewritten with columns 1 and nx-2).
template <typename T, int CM, bool X, bool PER>
__device__ void smooth(const TDesc<T>& g, const TLev<T>& L, const Ws<T>& W, const T* fac, T* lds, bool down) {
  if (down)
    zlines<T, CM, X, PER, true>(L, W, fac, 0);
  else
    zlines<T, CM, X, PER>(L, W, fac, 0);
  __syncthreads();
  zlines<T, CM, X, PER>(L, W, fac, 1);
  __syncthreads();
  ylines<T, X, PER>(L, W, fac, 0, lds);
  ylines<T, X, PER>(L, W, fac, 1, lds);
}

// res = r - A e on level L (specified points have the row e = r = 0); each
// thread issues the loads of kEw points before it uses any
constexpr int kEw = 4;
constexpr int kEwR = 2;   // the same for the residual (more loads per point)
constexpr int kEw64 = 2;
constexpr int kEwQ = 1;   // the same for A z of the Krylov step (2 needs scratch)
constexpr int kEwG = 2;   // the same for the Krylov step's update pass

template <typename T, bool X>
__device__ void residual(const TLev<T>& L, const Ws<T>& Wp, const T* __restrict__ fac) {
  T* W = wbase(L, Wp);
  const int nx = L.nx, ld = nx + 2, n = L.n;
  const T* __restrict__ e = W + L.e_off;
  const T* __restrict__ r = W + L.r_off;
  const T* __restrict__ cf = L.cf;
  const T* __restrict__ c9 = fac + L.f_off + 6 * n;
  T* __restrict__ res = W + L.res_off;
  const int nt = blockDim.x;
  for (int p0 = threadIdx.x; p0 < n; p0 += kEwR * nt) {
    T v[kEwR];
#pragma unroll
    for (int u = 0; u < kEwR; ++u) {
      const int p = min(p0 + u * nt, n - 1);
      const int j = p / nx, i = p - j * nx;
      const int q = (i + 1) + ld * (j + 1);
      const T k0 = cf[0 * n + p], k2 = cf[2 * n + p], k4 = cf[4 * n + p], k6 = cf[6 * n + p], k9 = c9[p];
      const T e0 = e[q + 1], e2 = e[q + ld], e4 = e[q - 1], e6 = e[q - ld], e9 = e[q], rv = r[p];
      T acc = (k0 * e0 + k2 * e2) + (k4 * e4 + k6 * e6);
      if (X)
        acc += (cf[1 * n + p] * e[q + 1 + ld] + cf[3 * n + p] * e[q - 1 + ld]) +
               (cf[5 * n + p] * e[q - 1 - ld] + cf[7 * n + p] * e[q + 1 - ld]);
      acc += k9 * e9;
      v[u] = rv - acc;
    }
#pragma unroll
    for (int u = 0; u < kEwR; ++u)
      if (p0 + u * nt < n) res[p0 + u * nt] = v[u];
  }
  __syncthreads();
}

// full weighting of the residual of F onto the right-hand side of C (zero
// outside the grid, or the periodic virtual columns; a direction that is not
// coarsened keeps its index).  C's
// correction is not cleared: the down leg's first sweep treats it as zero.
template <typename T, bool PER>
__device__ void restrict_fw(const TLev<T>& F, const TLev<T>& C, const Ws<T>& W) {
  const int sx = C.nx < F.nx ? 2 : 1, sy = C.ny < F.ny ? 2 : 1;
  const T* __restrict__ res = wbase(F, W) + F.res_off;
  T* __restrict__ rc = wbase(C, W) + C.r_off;
  const T w[3] = {(T)0.25, (T)0.5, (T)0.25};
  const int nt = blockDim.x;
  for (int p0 = threadIdx.x; p0 < C.n; p0 += kEwR * nt) {
    T v[kEwR];
#pragma unroll
    for (int u = 0; u < kEwR; ++u) {
      const int pc = min(p0 + u * nt, C.n - 1);
      const int jc = pc / C.nx, ic = pc - jc * C.nx;
      const int i = sx * ic, j = sy * jc;
      // the nine loads are issued before the weighted sum uses any
      T x[9];
      bool in[9];
#pragma unroll
      for (int b = -1; b <= 1; ++b)
#pragma unroll
        for (int a = -1; a <= 1; ++a) {
          int ii = i + a;
          const int jj = j + b;
          // periodic ends: the virtual columns -1 -> nx-2 and nx -> 1
          if constexpr (PER) ii = ii < 0 ? F.nx - 2 : ii >= F.nx ? 1 : ii;
          const int k = 3 * (b + 1) + (a + 1);
          in[k] = ii >= 0 && ii < F.nx && jj >= 0 && jj < F.ny;
          x[k] = res[in[k] ? ii + F.nx * jj : 0];
        }
      const bool sp = C.spec[pc];
      T acc = (T)0;
#pragma unroll
      for (int b = -1; b <= 1; ++b)
#pragma unroll
        for (int a = -1; a <= 1; ++a) {
          const int k = 3 * (b + 1) + (a + 1);
          acc += in[k] ? (w[a + 1] * w[b + 1]) * x[k] : (T)0;
        }
      v[u] = sp ? (T)0 : acc;
    }
#pragma unroll
    for (int u = 0; u < kEwR; ++u)
      if (p0 + u * nt < C.n) rc[p0 + u * nt] = v[u];
  }
  __syncthreads();
}

// e_F += bilinear interpolation of e_C (not at specified points), on the
// rows of parity 1 only: the smoothing that follows solves the rows of
// parity 0 first, reading only those of parity 1, and overwrites every row
// before anything else reads the correction
template <typename T, bool PER>
__device__ void prolong_add(const TLev<T>& C, const TLev<T>& F, const Ws<T>& W) {
  const int sx = C.nx < F.nx ? 2 : 1, sy = C.ny < F.ny ? 2 : 1;
  const T* __restrict__ ec = wbase(C, W) + C.e_off;
  T* __restrict__ ef = wbase(F, W) + F.e_off;
  const int lc = C.nx + 2, lf = F.nx + 2;
  const int nt = blockDim.x;
  const int nodd = F.nx * (F.ny >> 1);  // points of rows 1, 3, ...
  for (int p0 = threadIdx.x; p0 < nodd; p0 += kEw * nt) {
    T v[kEw];
    int qf[kEw];
    int pf[kEw];
#pragma unroll
    for (int u = 0; u < kEw; ++u) {
      const int po = min(p0 + u * nt, nodd - 1);
      const int jr = po / F.nx, i = po - jr * F.nx;
      const int j = 2 * jr + 1;
      const int p = i + F.nx * j;
      pf[u] = p;
      const int i0 = i / sx, i1 = (i + sx - 1) / sx, j0 = j / sy, j1 = (j + sy - 1) / sy;
      qf[u] = (i + 1) + lf * (j + 1);
      // the six loads are issued before any is used
      const T c00 = ec[(i0 + 1) + lc * (j0 + 1)], c10 = ec[(i1 + 1) + lc * (j0 + 1)];
      const T c01 = ec[(i0 + 1) + lc * (j1 + 1)], c11 = ec[(i1 + 1) + lc * (j1 + 1)];
      const T fo = ef[qf[u]];
      const bool sp = F.spec[p];
      const T x = (T)0.25 * ((c00 + c10) + (c01 + c11));
      v[u] = sp ? (T)0 : x + fo;
    }
#pragma unroll
    for (int u = 0; u < kEw; ++u)
      if (p0 + u * nt < nodd && !F.spec[pf[u]]) {
        ef[qf[u]] = v[u];
        if constexpr (PER) {
          const int i = pf[u] % F.nx;
          ring_store(ef, qf[u], i, F.nx, v[u]);
        }
      }
  }
  __syncthreads();
}

// e = A^-1 r on the coarsest level of the cycle (A^-1 precomputed per mode)
template <typename T, bool PER>
__device__ void direct_solve(const TLev<T>& L, const Ws<T>& Wp, const T* __restrict__ ainv) {
  T* W = wbase(L, Wp);
  const int n = L.n, nx = L.nx, ld = nx + 2;
  const T* __restrict__ r = W + L.r_off;
  T* e = W + L.e_off;
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    const T* row = ainv + (size_t)p * n;
    T acc = (T)0;
    for (int q = 0; q < n; ++q) acc += row[q] * r[q];
    const int j = p / nx, i = p - j * nx;
    const int q = (i + 1) + ld * (j + 1);
    e[q] = acc;
    if constexpr (PER) ring_store(e, q, i, nx, acc);
  }
  __syncthreads();
}

// sums of four values over the workgroup (wavefront sums, then one value per
// wavefront through `red`); every thread gets the totals
__device__ inline void block_sum4(double* v, double* red) {
  for (int k = 0; k < 4; ++k)
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0)
    for (int k = 0; k < 4; ++k) red[4 * wave + k] = v[k];
  __syncthreads();
  for (int k = 0; k < 4; ++k) {
    double a = 0.0;
    for (int w = 0; w < nw; ++w) a += red[4 * w + k];
    v[k] = a;
  }
  __syncthreads();
}

// One GCR(1) step after a cycle (the cycle's correction z sits in the finest
// level's e): q = A z in FP64; q and z are made A-orthogonal to the kept
// direction (beta), then the minimal residual step alpha = (r.q')/(q'.q')
// updates phi, r and the kept direction, and r seeds the next cycle in the
// cycle's precision.  Returns this thread's max |r|.  Kept out of line so its
// pointers and scalars are not live across the cycle.
template <typename T, int NT, bool X, bool PER>
__device__ __noinline__ double gcr_step(const TDesc<T>& g, const TLev<T>& F, double* __restrict__ phi, T* W,
                                        const double* __restrict__ c964, double* red, int it) {
  __shared__ double sQQ;  // Qp . Qp of the kept direction
  const int nx = F.nx, ny = F.ny, ld = nx + 2, n = F.n, P = (nx + 2) * (ny + 2);
  double* __restrict__ rr = phi + P + n;  // after phi (ring) and b
  double* __restrict__ qv = rr + n;       // A z
  double* __restrict__ Qp = qv + n;       // kept direction's A z
  double* __restrict__ Zp = Qp + n;       // kept direction (ring layout)
  T* __restrict__ rf = W + F.r_off;
  T* __restrict__ ef = W + F.e_off;
  const double* __restrict__ c = g.cf64;
  double rmax = 0.0;
  const bool hasq = it > 1;
  double sm[4] = {0.0, 0.0, 0.0, 0.0};  // q.Qp, q.q, r.q, r.Qp
  for (int p0 = threadIdx.x; p0 < n; p0 += kEwQ * NT) {
    double v[kEwQ];
#pragma unroll
    for (int u = 0; u < kEwQ; ++u) {
      const int p = min(p0 + u * NT, n - 1);
      const int j = p / nx, i = p - j * nx;
      const int q = (i + 1) + ld * (j + 1);
      double acc = (c[0 * n + p] * (double)ef[q + 1] + c[2 * n + p] * (double)ef[q + ld]) +
                   (c[4 * n + p] * (double)ef[q - 1] + c[6 * n + p] * (double)ef[q - ld]);
      if (X)
        acc += (c[1 * n + p] * (double)ef[q + 1 + ld] + c[3 * n + p] * (double)ef[q - 1 + ld]) +
               (c[5 * n + p] * (double)ef[q - 1 - ld] + c[7 * n + p] * (double)ef[q + 1 - ld]);
      v[u] = acc + c964[p] * (double)ef[q];
    }
#pragma unroll
    for (int u = 0; u < kEwQ; ++u) {
      const int p = p0 + u * NT;
      if (p < n) {
        const double qp = v[u], rp = rr[p], qk = hasq ? Qp[p] : 0.0;
        qv[p] = qp;
        sm[0] += qp * qk;
        sm[1] += qp * qp;
        sm[2] += rp * qp;
        sm[3] += rp * qk;
      }
    }
  }
  block_sum4(sm, red);
  const double QQ = hasq ? sQQ : 0.0;
  const double beta = (hasq && QQ > 0.0) ? sm[0] / QQ : 0.0;
  const double qq = sm[1] - 2.0 * beta * sm[0] + beta * beta * QQ;
  const double alpha = qq > 0.0 ? (sm[2] - beta * sm[3]) / qq : 0.0;
  __syncthreads();  // every thread has read the previous value
  if (threadIdx.x == 0) sQQ = qq;
  // one pass over the interior: phi and the kept direction (ring layout),
  // then A z and the residual.  The rings of phi and of the direction are
  // never visited: ef is zero there, so the update would keep them zero
  // (phi's ring is cleared at the start; the direction's ring is never read);
  // with periodic ends phi's ring holds the virtual columns, written with
  // columns 1 and nx-2.
  for (int p0 = threadIdx.x; p0 < n; p0 += kEwG * NT) {
    double zv[kEwG], zk[kEwG], ph[kEwG], qn[kEwG], qk[kEwG], rk[kEwG];
    int qi[kEwG];
#pragma unroll
    for (int u = 0; u < kEwG; ++u) {
      const int p = min(p0 + u * NT, n - 1);
      const int j = p / nx, i = p - j * nx;
      const int q = (i + 1) + ld * (j + 1);
      qi[u] = q;
      zv[u] = (double)ef[q];
      zk[u] = hasq ? Zp[q] : 0.0;
      ph[u] = phi[q];
      qn[u] = qv[p];
      qk[u] = hasq ? Qp[p] : 0.0;
      rk[u] = rr[p];
    }
#pragma unroll
    for (int u = 0; u < kEwG; ++u) {
      const int p = p0 + u * NT;
      if (p < n) {
        const double z = zv[u] - (hasq ? beta * zk[u] : 0.0);
        const double np = ph[u] + alpha * z;
        phi[qi[u]] = np;
        if constexpr (PER) ring_store_d(phi, qi[u], p % nx, nx, np);
        Zp[qi[u]] = z;
        const double q2 = qn[u] - (hasq ? beta * qk[u] : 0.0);
        const double r2 = rk[u] - alpha * q2;
        Qp[p] = q2;
        rr[p] = r2;
        rf[p] = (T)r2;
        rmax = fmax(rmax, fabs(r2));
      }
    }
  }
  return rmax;
}

template <typename T, int CM, int NT, bool X, bool PER>
__global__ void __launch_bounds__(NT) k_tight_solve(const TDesc<T>* __restrict__ gp, T* __restrict__ ws,
                                                     double* __restrict__ ws64, const double* __restrict__ rhs_tmp,
                                                     double* __restrict__ phi_out, const double* __restrict__ kx2v,
                                                     int* __restrict__ cycles, int* __restrict__ unconv,
                                                     size_t plane, const double* __restrict__ guessp, int nsolve) {
  __shared__ double red[NT];
  __shared__ T ylds[3 * kYMaxTasks * 64];
  extern __shared__ __align__(16) unsigned char tight_lds[];
  const TDesc<T>& g = *gp;
  const double guess = guessp ? *guessp : 0.0;
  // workgroups are dealt to the 8 XCDs round robin: solves 2m and 2m+1 (the
  // two parts of mode m, which read the same per-mode tables) are placed on
  // one XCD so the second reader finds the tables in that XCD's L2.  The grid
  // is a multiple of 16; the map is a bijection on it.
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int s = 2 * ((slot >> 1) * 8 + xcd) + (slot & 1);
  if (s >= nsolve) return;
  (void)kx2v;  // the wavenumber enters through the per-mode tables
  T* W = ws + (size_t)s * g.ws;
  const Ws<T> Wp{W, reinterpret_cast<T*>(tight_lds)};
  double* __restrict__ phi = ws64 + (size_t)s * g.ws64;
  const int ng = g.ngrid;
  const TLev<T>& F = g.lev[ng];
  const int nx = F.nx, ny = F.ny, ld = nx + 2, n = F.n;
  double* __restrict__ b = phi + (size_t)(nx + 2) * (ny + 2);
  const T* fac = g.fac + (size_t)(s >> 1) * g.fstride;
  const double* __restrict__ c964 = g.c964 + (size_t)(s >> 1) * n;
  const double* rt = rhs_tmp + (size_t)s * plane;
  // phi starts from zero, or from the previous solution of this solve scaled
  // by `guess` (phi_out still holds it)
  const double* pg = phi_out + (size_t)s * plane;
  for (int t = threadIdx.x; t < (nx + 2) * (ny + 2); t += blockDim.x) {
    const int jj = t / ld - 1, ii = t - (jj + 1) * ld - 1;
    double v = 0.0;
    // periodic ends: the ring columns start as the virtual columns
    const int iv = PER ? (ii < 0 ? nx - 2 : ii >= nx ? 1 : ii) : ii;
    if (guess != 0.0 && iv >= 0 && iv < nx && jj >= 0 && jj < ny && !F.spec[iv + nx * jj]) v = guess * pg[iv + nx * jj];
    phi[t] = v;
  }
  const int P = (nx + 2) * (ny + 2);
  double* __restrict__ rr = b + n;  // krylov: the recursive residual (first iteration)
  T* __restrict__ rf = W + F.r_off;
  T* __restrict__ ef = W + F.e_off;
  // From phi = 0 the first residual is b itself (A 0 adds only signed zeros),
  // so it is formed here instead of by the residual pass; b is kept only
  // where a later residual reads it.
  const bool zero_start = guess == 0.0;
  // the singular kx = 0 solve: b_pin -= w.b makes the right-hand side
  // compatible (w the left null vector, w_pin = 1)
  const bool sing = g.wnull != nullptr && (s >> 1) == 0;
  double rmax0 = 0.0;
  double wb[4] = {0.0, 0.0, 0.0, 0.0};
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    const double bv = F.spec[p] ? 0.0 : -rt[p];
    if (!zero_start || !g.krylov) b[p] = bv;
    if (zero_start) {
      rmax0 = fmax(rmax0, fabs(bv));
      rf[p] = (T)bv;
      if (g.krylov) rr[p] = bv;
    }
    if (sing) wb[0] += g.wnull[p] * bv;
  }
  if (sing) {
    block_sum4(wb, red);
    if (threadIdx.x == 0) {
      const int p = g.pin;
      const double bv = (F.spec[p] ? 0.0 : -rt[p]) - wb[0];
      if (!zero_start || !g.krylov) b[p] = bv;
      if (zero_start) {
        rmax0 = fmax(rmax0, fabs(bv));
        rf[p] = (T)bv;
        if (g.krylov) rr[p] = bv;
      }
    }
  }
  // every level's correction starts at zero: the rings stay zero, the interiors
  // are rewritten by every sweep
  for (int l = 1; l <= ng; ++l) {
    const TLev<T>& L = g.lev[l];
    T* el = wbase(L, Wp) + L.e_off;
    for (int t = threadIdx.x; t < (L.nx + 2) * (L.ny + 2); t += blockDim.x) el[t] = (T)0;
  }
  __syncthreads();
  const double* __restrict__ c = g.cf64;
  int it = 0;
  bool conv = false;  // the residual met the tolerance (uniform over the workgroup)
  for (;;) {
    double rmax = 0.0;
    if (it == 0 && zero_start) {
      rmax = rmax0;
    } else if (!g.krylov || it == 0) {
      // FP64 residual of the finest level; its max decides; it seeds the cycle
      for (int p0 = threadIdx.x; p0 < n; p0 += kEw64 * NT) {
        double v[kEw64];
#pragma unroll
        for (int u = 0; u < kEw64; ++u) {
          const int p = min(p0 + u * NT, n - 1);
          const int j = p / nx, i = p - j * nx;
          const int q = (i + 1) + ld * (j + 1);
          double acc = (c[0 * n + p] * phi[q + 1] + c[2 * n + p] * phi[q + ld]) +
                       (c[4 * n + p] * phi[q - 1] + c[6 * n + p] * phi[q - ld]);
          if (X)
            acc += (c[1 * n + p] * phi[q + 1 + ld] + c[3 * n + p] * phi[q - 1 + ld]) +
                   (c[5 * n + p] * phi[q - 1 - ld] + c[7 * n + p] * phi[q + 1 - ld]);
          acc += c964[p] * phi[q];
          v[u] = b[p] - acc;
        }
#pragma unroll
        for (int u = 0; u < kEw64; ++u)
          if (p0 + u * NT < n) {
            rmax = fmax(rmax, fabs(v[u]));
            rf[p0 + u * NT] = (T)v[u];
            if (g.krylov) rr[p0 + u * NT] = v[u];
          }
      }
    } else {
      rmax = gcr_step<T, NT, X, PER>(g, F, phi, W, c964, red, it);
    }
    // max over the workgroup: wavefront maxima, then one value per wavefront
    for (int o = 32; o > 0; o >>= 1) rmax = fmax(rmax, __shfl_xor(rmax, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = rmax;
    __syncthreads();
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) rmax = fmax(rmax, red[w]);
    __syncthreads();
    if (rmax <= g.tol) {
      conv = true;
      break;
    }
    if (it >= g.maxit) break;
    // one cycle on the correction equation: the steps of the mode's program
    // in order, one call site of the smoother
    const T* ainv = fac + g.ainv_off;
    for (int k = 0; k < g.nprog; ++k) {
      const int step = g.prog[k];
      const int l = step >> 2, op = step & 3;
      if (op == 2) {
        direct_solve<T, PER>(g.lev[l], Wp, ainv);
        continue;
      }
      if (op == 1) prolong_add<T, PER>(g.lev[l - 1], g.lev[l], Wp);
      smooth<T, CM, X, PER>(g, g.lev[l], Wp, fac, ylds, op == 0);
      if (op == 0) {
        residual<T, X>(g.lev[l], Wp, fac);
        restrict_fw<T, PER>(g.lev[l], g.lev[l - 1], Wp);
      }
    }
    if (!g.krylov) {
      for (int t0 = threadIdx.x; t0 < P; t0 += 2 * kEw * NT) {
        double v[2 * kEw];
#pragma unroll
        for (int u = 0; u < 2 * kEw; ++u) {
          const int t = min(t0 + u * NT, P - 1);
          v[u] = phi[t] + (double)ef[t];
        }
#pragma unroll
        for (int u = 0; u < 2 * kEw; ++u)
          if (t0 + u * NT < P) phi[t0 + u * NT] = v[u];
      }
    }
    __syncthreads();
    ++it;
  }
  double* po = phi_out + (size_t)s * plane;
  // the singular solve: the solution with phi = 0 at the pin
  const double shift = sing ? phi[(g.pin % nx + 1) + ld * (g.pin / nx + 1)] : 0.0;
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    const int j = p / nx, i = p - j * nx;
    po[p] = phi[(i + 1) + ld * (j + 1)] - shift;
  }
  if (threadIdx.x == 0) {
    cycles[s] = it;
    unconv[s] = conv ? 0 : 1;
  }
}

// Gauss-Jordan inverse with partial pivoting of a dense n x n matrix
bool invert_dense(int n, std::vector<double> A, std::vector<double>& inv) {
  inv.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int pr = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)pr * n + c])) pr = r;
    if (A[(size_t)pr * n + c] == 0.0) return false;
    if (pr != c)
      for (int k = 0; k < n; ++k) {
        std::swap(A[(size_t)pr * n + k], A[(size_t)c * n + k]);
        std::swap(inv[(size_t)pr * n + k], inv[(size_t)c * n + k]);
      }
    const double rp = 1.0 / A[(size_t)c * n + c];
    for (int k = 0; k < n; ++k) {
      A[(size_t)c * n + k] *= rp;
      inv[(size_t)c * n + k] *= rp;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = A[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        A[(size_t)r * n + k] -= f * A[(size_t)c * n + k];
        inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
      }
    }
  }
  return true;
}

// The left null vector w of a singular finest operator (kx = 0, no specified
// side), scaled to w[pin] = 1: the equations of A^T w = 0 but the pin's,
// which becomes w[pin] = 1 (the rows of A^T sum to zero, since A annihilates
// constants, so any one of them follows from the others and the system is
// nonsingular).  Points ordered i + nx j, so every coupling, the periodic
// virtual columns included, lies within a band of half-width nx+1 (2nx-2
// when the periodic wrap has cross terms); banded Gaussian elimination with
// partial pivoting, O(n band^2) once per solver.
bool left_null_vector(const HostMgLevel& H, bool per, int pin, std::vector<double>& w) {
  const int nx = H.nx, ny = H.ny, n = nx * ny;
  static const int di8[8] = {1, 1, 0, -1, -1, -1, 0, 1}, dj8[8] = {0, 1, 1, 1, 0, -1, -1, -1};
  // the couplings: A(q, qq) = coefficient of point qq in row q, A^T(qq, q) = A(q, qq)
  auto each = [&](auto&& fn) {
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int q = i + nx * j;
        for (int c = 0; c < 8; ++c) {
          const double v = H.c[(size_t)c * n + q];
          if (v == 0.0) continue;
          int ii = i + di8[c];
          const int jj = j + dj8[c];
          if (per) ii = ii < 0 ? nx - 2 : ii >= nx ? 1 : ii;
          if (ii < 0 || ii >= nx || jj < 0 || jj >= ny) continue;
          fn(q, ii + nx * jj, v);
        }
      }
  };
  int kl = 1;
  each([&](int q, int qq, double) { kl = std::max(kl, std::abs(qq - q)); });
  const int ku = kl, wd = 2 * kl + ku + 1;
  std::vector<double> band((size_t)n * wd, 0.0), rhs(n, 0.0);
  auto at = [&](int r, int c) -> double& { return band[(size_t)r * wd + (c - r + kl)]; };
  for (int q = 0; q < n; ++q)
    if (q != pin) at(q, q) += (-H.s9[q]) + H.corr9[q];
  each([&](int q, int qq, double v) {
    if (qq != pin) at(qq, q) += v;
  });
  at(pin, pin) = 1.0;
  rhs[pin] = 1.0;
  for (int k = 0; k < n; ++k) {
    const int last = std::min(n - 1, k + kl);
    int pr = k;
    for (int r = k + 1; r <= last; ++r)
      if (std::fabs(at(r, k)) > std::fabs(at(pr, k))) pr = r;
    if (at(pr, k) == 0.0) return false;
    const int cend = std::min(n - 1, k + kl + ku);
    if (pr != k) {
      for (int c = k; c <= cend; ++c) std::swap(at(k, c), at(pr, c));
      std::swap(rhs[k], rhs[pr]);
    }
    const double inv = 1.0 / at(k, k);
    for (int r = k + 1; r <= last; ++r) {
      const double f = at(r, k) * inv;
      if (f == 0.0) continue;
      double* rr = &at(r, k);
      const double* rk = &at(k, k);
      for (int c = 0; c <= cend - k; ++c) rr[c] -= f * rk[c];
      rhs[r] -= f * rhs[k];
    }
  }
  w.assign(n, 0.0);
  for (int k = n - 1; k >= 0; --k) {
    double acc = rhs[k];
    const int cend = std::min(n - 1, k + kl + ku);
    for (int c = k + 1; c <= cend; ++c) acc -= at(k, c) * w[c];
    w[k] = acc / at(k, k);
  }
  return std::isfinite(w[0]);
}

// Thomas factors of one line: lower/diag/upper -> (1/piv, -lower/piv, upper/piv)
void line_factors(int len, const double* lo, const double* di, const double* up, double* f1, double* fa,
                  double* fc) {
  double cprev = 0.0;
  for (int k = 0; k < len; ++k) {
    const double piv = di[k] - lo[k] * cprev;
    f1[k] = 1.0 / piv;
    fa[k] = -lo[k] * f1[k];
    fc[k] = up[k] * f1[k];
    cprev = fc[k];
  }
}

template <typename T>
int tight_build_typed(TightPlan& tp, const ekm_params& p, const Dims& d, const HostMetrics& hm, const Ranges& rv,
                      const double* kx2_mode, int nmode) {
  const int ng = p.mg_iex > p.mg_jey ? p.mg_iex : p.mg_jey;
  if (ng > kMaxLevT) return EKM_ERR_UNSUPPORTED;
  TDesc<T> g;
  std::memset(&g, 0, sizeof(g));
  g.ngrid = ng;
  g.nfx = d.NZ2;
  g.nfy = d.NY2;
  g.maxit = p.proj_maxit > 0 ? p.proj_maxit : 50;
  g.tol = p.proj_tol;
  std::vector<HostMgLevel> hl(ng + 1);
  long ws = 0, fstride = 0;
  // periodic z ends; with no specified side the kx = 0 problem is singular
  // (constants are in the null space of every row).  Its solution is the one
  // with phi = 0 at the point p = (K = 0, J = NY+1), a MUDPACK boundary row,
  // that satisfies every other row.  A pin on every level (p's diagonal
  // scaled by kPin) gives that solution but slows the cycles several-fold
  // (a point constraint is not a smooth error), so the finest level stays
  // singular: its right-hand side is made compatible, b_p -= w.b with w the
  // operator's left null vector scaled to w_p = 1 (which changes row p only),
  // the cycles' coarse levels carry the pin, and the converged solution is
  // shifted to phi_p = 0.  The coarse pins only move the correction by
  // constants near p, which the finest residual sees and removes.
  const bool per = p.mudpack_bc[0] == 0;
  bool anyspec = false;
  for (int s = 0; s < 4; ++s) anyspec = anyspec || p.mudpack_bc[s] == 1;
  constexpr double kPin = 1e20;
  for (int k = 1; k <= ng; ++k) {
    const int ex = k + p.mg_iex - ng, ey = k + p.mg_jey - ng;
    const int nx = p.mg_ixp * (1 << ((ex > 1 ? ex : 1) - 1)) + 1;
    const int ny = p.mg_jyq * (1 << ((ey > 1 ? ey : 1) - 1)) + 1;
    // the y-line smoother keeps one LDS slot per (64-column group, chunk)
    if ((nx + 1) / 2 > 64 * kYMaxTasks) return EKM_ERR_UNSUPPORTED;
    int st = host_mg_level(hm, d.NZ, d.NY, p.mudpack_bc, nx, ny, hl[k]);
    if (st) return st;
    TLev<T>& L = g.lev[k];
    L.nx = nx;
    L.ny = ny;
    L.n = nx * ny;
    L.f_off = (int)fstride;
    fstride += 9L * L.n;  // z factors, y factors, c9, y-chunk products
    if (per) fstride += 2L * L.n + 4L * ny;  // wrap vectors and matrices of the z lines
  }
  int lc = 1;
  for (int k = 1; k <= ng; ++k)
    if (g.lev[k].n <= kDirectMax) lc = k;
  g.lc = lc;
  {
    // the cycle program: down ng..lc+1, direct solve on lc, up lc+1..ng
    auto put = [&](int l, int op) { g.prog[g.nprog++] = (unsigned char)((l << 2) | op); };
    for (int l = ng; l > lc; --l) put(l, 0);
    put(lc, 2);
    for (int l = lc + 1; l <= ng; ++l) put(l, 1);
  }
  g.ainv_off = fstride;
  fstride += (long)g.lev[lc].n * g.lev[lc].n;
  fstride = (fstride + 63) & ~63L;
  g.krylov = p.proj_krylov;
  g.ws64 = (long)(d.NZ2 + 2) * (d.NY2 + 2) + (long)d.NZ2 * d.NY2;
  if (g.krylov) g.ws64 += 3L * d.NZ2 * d.NY2 + (long)(d.NZ2 + 2) * (d.NY2 + 2);
  g.ws64 = (g.ws64 + 31) & ~31L;
  g.fstride = fstride;
  // consistent rows of the finest level (cells J=2..NY-1, K=2..NZ-1)
  {
    HostMgLevel& L = hl[ng];
    const int nx = L.nx;
    const size_t n = (size_t)L.nx * L.ny;
    const int NZ3 = d.NZ + 3, NY3 = d.NY + 3;
    auto G = [&](int w, int k, int j, int f) { return hm.m[w][(size_t)k + (size_t)NZ3 * ((size_t)j + (size_t)NY3 * (f - 1))]; };
    for (int J = 2; J <= d.NY - 1; ++J)
      for (int K = 2; K <= d.NZ - 1; ++K) {
        const double tm = (J + 1 <= rv.je) ? 1.0 : 0.0, bm = (J >= rv.js) ? 1.0 : 0.0;
        const double rm = (K + 1 <= rv.ze) ? 1.0 : 0.0, lm = (K >= rv.zs) ? 1.0 : 0.0;
        const double a = G(2, K, J + 1, 1) * tm, bb = G(2, K, J, 1) * bm;
        const double cc = G(0, K + 1, J, 2) * rm, dd = G(0, K, J, 2) * lm;
        const double e = 0.25 * G(1, K, J + 1, 1) * tm, f = 0.25 * G(1, K, J, 1) * bm;
        const double gg = 0.25 * G(1, K + 1, J, 2) * rm, h = 0.25 * G(1, K, J, 2) * lm;
        const size_t q = (size_t)K + (size_t)nx * J;
        L.c[0 * n + q] = (e - f) + cc;
        L.c[1 * n + q] = e + gg;
        L.c[2 * n + q] = (a + gg) - h;
        L.c[3 * n + q] = -e - h;
        L.c[4 * n + q] = (-e + f) + dd;
        L.c[5 * n + q] = f + h;
        L.c[6 * n + q] = (bb - gg) + h;
        L.c[7 * n + q] = -f - gg;
        L.s9[q] = ((a + bb) + cc) + dd;
        L.corr9[q] = 0.0;
        L.jl[q] = hm.jac[(size_t)K + (size_t)NZ3 * J];
      }
  }
  // specified points: identity rows (c1..c8 = 0, c9 = 1); cross terms present?
  size_t tot_n = 0;
  std::vector<size_t> loff(ng + 1);
  for (int k = 1; k <= ng; ++k) {
    loff[k] = tot_n;
    tot_n += (size_t)g.lev[k].n;
  }
  std::vector<unsigned char> spec(tot_n);
  bool cross = false;
  for (int k = 1; k <= ng; ++k) {
    HostMgLevel& H = hl[k];
    const int nx = H.nx, ny = H.ny;
    const size_t n = (size_t)nx * ny;
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t q = i + (size_t)nx * j;
        const bool sp = (p.mudpack_bc[0] == 1 && i == 0) || (p.mudpack_bc[1] == 1 && i == nx - 1) ||
                        (p.mudpack_bc[2] == 1 && j == 0) || (p.mudpack_bc[3] == 1 && j == ny - 1);
        spec[loff[k] + q] = sp;
        if (sp) {
          for (int c = 0; c < 8; ++c) H.c[c * n + q] = 0.0;
          H.s9[q] = -1.0;
          H.corr9[q] = 0.0;
          H.jl[q] = 0.0;
        }
        for (int c = 1; c < 8; c += 2)
          if (H.c[c * n + q] != 0.0) cross = true;
      }
  }
  tp.cross = cross;
  tp.per = per;
  // workspace: from the coarsest level up, the levels whose correction,
  // right-hand side and two scratch arrays fit in the LDS left beside the
  // kernel's static arrays live there; the others (always the finest) in
  // global memory, one slice per solve
  {
    const int nt = tight_block(sizeof(T) == 4, tp.cm, cross);
    const long stat = (long)nt * 8 + 3L * kYMaxTasks * 64 * (long)sizeof(T);
    int dev = 0, maxsh = 0;
    EKM_CHECK(hipGetDevice(&dev));
    EKM_CHECK(hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    long budget = (long)maxsh - stat - 256;
    if (tp.lds_budget > 0) budget = std::min(budget, tp.lds_budget);
    long lds = 0;
    bool fits = true;
    for (int k = 1; k <= ng; ++k) {
      TLev<T>& L = g.lev[k];
      const long need = (((long)(L.nx + 2) * (L.ny + 2) + 3L * L.n + 31) & ~31L) * (long)sizeof(T);
      fits = fits && k < ng && lds + need <= budget;
      long& off = fits ? lds : ws;
      const long base = fits ? off / (long)sizeof(T) : off;
      L.lds = fits ? 1 : 0;
      L.e_off = (int)base;
      L.r_off = (int)(base + (long)(L.nx + 2) * (L.ny + 2));
      L.res_off = L.r_off + L.n;
      L.t2_off = L.res_off + L.n;
      const long elems = ((long)(L.nx + 2) * (L.ny + 2) + 3L * L.n + 31) & ~31L;
      off += fits ? elems * (long)sizeof(T) : elems;
    }
    tp.lds_bytes = (int)lds;
    g.ws = ws;
  }
  // coefficient tables, and per live mode: line factors, c9, y-chunk
  // products (for the kernel's wavefront count), coarsest inverse
  const int nw = tight_block(sizeof(T) == 4, tp.cm, cross) / 64;
  const size_t nfine = (size_t)g.lev[ng].n;
  std::vector<T> cfT(8 * tot_n), cfyT(8 * tot_n);
  std::vector<T> fac((size_t)nmode * fstride, (T)0);
  std::vector<double> c964((size_t)nmode * nfine);
  for (int k = 1; k <= ng; ++k) {
    const HostMgLevel& H = hl[k];
    const int nx = H.nx, ny = H.ny;
    const size_t n = (size_t)nx * ny, off = loff[k];
    for (size_t t = 0; t < 8 * n; ++t) cfT[8 * off + t] = (T)H.c[t];
    for (int c = 0; c < 8; ++c)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
          cfyT[8 * off + c * n + (size_t)ysplit(i, j, nx, ny)] = (T)H.c[c * n + i + (size_t)nx * j];
    std::vector<double> lo(std::max(nx, ny)), di(lo.size()), up(lo.size()), f1(lo.size()), fa(lo.size()), fc(lo.size());
    for (int m = 0; m < nmode; ++m) {
      const double kx2 = kx2_mode[m];
      T* fz = fac.data() + (size_t)m * fstride + g.lev[k].f_off;
      T* fy = fz + 3 * n;
      T* fc9 = fz + 6 * n;
      // the pin: every level below the finest (the finest operator stays
      // singular and its right-hand side is made compatible, see below), or
      // the finest too when it is the directly solved level
      const size_t qpin = (!anyspec && kx2 == 0.0 && (k < ng || lc == ng)) ? (size_t)nx * (ny - 1) : (size_t)-1;
      auto c9 = [&](size_t q) {
        const double v = ((-(kx2 * H.jl[q])) - H.s9[q]) + H.corr9[q];
        return q == qpin ? kPin * v : v;
      };
      for (size_t q = 0; q < n; ++q) fc9[q] = (T)c9(q);
      if (k == ng)
        for (size_t q = 0; q < n; ++q) c964[(size_t)m * nfine + q] = c9(q);
      for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
          const size_t q = i + (size_t)nx * j;
          lo[i] = H.c[4 * n + q];
          di[i] = c9(q);
          up[i] = H.c[0 * n + q];
        }
        line_factors(nx, lo.data(), di.data(), up.data(), f1.data(), fa.data(), fc.data());
        // the chain of a line ends at its last point: the scans carry the
        // values of the lanes past it into the back substitution, so its
        // upper factor must be zero (with periodic ends c1 there is the wrap
        // coupling, which the Woodbury step handles)
        fa[0] = 0.0;
        fc[nx - 1] = 0.0;
        for (int i = 0; i < nx; ++i) {
          const size_t q = i + (size_t)nx * j;
          // a specified point's correction is zero: 1/piv is stored as 0
          fz[q] = spec[off + q] ? (T)0 : (T)f1[i];
          fz[n + q] = (T)fa[i];
          fz[2 * n + q] = (T)fc[i];
        }
        if (per) {
          // Woodbury terms of the line's wrap couplings (see zline_periodic):
          // W = T^-1 [c5(0) e_0, c1(nx-1) e_{nx-1}] with the chain's factors
          std::vector<double> w(2 * (size_t)nx);
          for (int v = 0; v < 2; ++v) {
            double* wv = w.data() + (size_t)v * nx;
            double dp = 0.0;
            for (int i = 0; i < nx; ++i) {
              const size_t q = i + (size_t)nx * j;
              const double sv = (v == 0 && i == 0) ? lo[0] : (v == 1 && i == nx - 1) ? up[nx - 1] : 0.0;
              dp = spec[off + q] ? 0.0 : fa[i] * dp + sv * f1[i];
              wv[i] = dp;
            }
            for (int i = nx - 2; i >= 0; --i) wv[i] -= fc[i] * wv[i + 1];
          }
          const double m00 = 1.0 + w[nx - 2], m01 = w[nx + nx - 2], m10 = w[1], m11 = 1.0 + w[nx + 1];
          const double det = m00 * m11 - m01 * m10;
          T* fw = fz + 9 * n;
          for (int i = 0; i < nx; ++i) {
            fw[i + (size_t)nx * j] = (T)w[i];
            fw[n + i + (size_t)nx * j] = (T)w[nx + i];
          }
          T* mi = fz + 11 * n + 4 * (size_t)j;
          mi[0] = (T)(m11 / det);
          mi[1] = (T)(-m01 / det);
          mi[2] = (T)(-m10 / det);
          mi[3] = (T)(m00 / det);
        }
      }
      for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) {
          const size_t q = i + (size_t)nx * j;
          lo[j] = H.c[6 * n + q];
          di[j] = c9(q);
          up[j] = H.c[2 * n + q];
        }
        line_factors(ny, lo.data(), di.data(), up.data(), f1.data(), fa.data(), fc.data());
        for (int j = 0; j < ny; ++j) {
          const size_t q = i + (size_t)nx * j;
          const size_t qs = (size_t)ysplit(i, j, nx, ny);
          fy[qs] = spec[off + q] ? (T)0 : (T)f1[j];
          fy[n + qs] = (T)fa[j];
          fy[2 * n + qs] = (T)fc[j];
        }
        // the chunk products of this column, in the cycle precision and the
        // order the smoother would form them
        const YChunks yc = ychunks(nx, ny, i & 1, nw);
        T* yA = fz + 7 * n;
        T* yBe = fz + 8 * n;
        for (int ch = 0; ch < yc.nch; ++ch) {
          const int j0 = ch * yc.cl, j1 = std::min(j0 + yc.cl, ny);
          T A = (T)1;
          for (int j = j0; j < j1; ++j) {
            const size_t qs = (size_t)ysplit(i, j, nx, ny);
            A = fy[n + qs] * A;
            yA[qs] = A;
          }
          T be = (T)1;
          for (int j = j1 - 1; j >= j0; --j) {
            const size_t qs = (size_t)ysplit(i, j, nx, ny);
            be = -fy[2 * n + qs] * be;
            yBe[qs] = be;
          }
        }
      }
      if (k == lc) {
        std::vector<double> A(n * n, 0.0);
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i) {
            const size_t q = i + (size_t)nx * j;
            A[q * n + q] = c9(q);
            if (spec[off + q]) continue;
            static const int di8[8] = {1, 1, 0, -1, -1, -1, 0, 1}, dj8[8] = {0, 1, 1, 1, 0, -1, -1, -1};
            for (int c = 0; c < 8; ++c) {
              int ii = i + di8[c];
              const int jj = j + dj8[c];
              if (per) ii = ii < 0 ? nx - 2 : ii >= nx ? 1 : ii;  // periodic virtual columns
              if (ii < 0 || ii >= nx || jj < 0 || jj >= ny) continue;
              const size_t qq = ii + (size_t)nx * jj;
              if (spec[off + qq]) continue;  // the correction is zero there
              A[q * n + qq] += H.c[c * n + q];
            }
          }
        std::vector<double> inv;
        if (!invert_dense((int)n, A, inv)) return EKM_ERR_UNSUPPORTED;
        T* ai = fac.data() + (size_t)m * fstride + g.ainv_off;
        for (size_t t = 0; t < n * n; ++t) ai[t] = (T)inv[t];
      }
    }
  }
  // the singular kx = 0 problem's left null vector (see the pin above)
  std::vector<double> wnull;
  g.pin = (int)((size_t)g.lev[ng].nx * (g.lev[ng].ny - 1));
  if (!anyspec && lc < ng) {
    if (nmode < 1 || kx2_mode[0] != 0.0) return EKM_ERR_UNSUPPORTED;
    if (!left_null_vector(hl[ng], per, g.pin, wnull)) return EKM_ERR_UNSUPPORTED;
  }
  // device copies: descriptor, spec, c1..c8 (T), finest c1..c8 (double), c9 per mode (double),
  // the left null vector, factors
  const size_t b_desc = (sizeof(TDesc<T>) + 255) & ~(size_t)255;
  const size_t b_spec = (tot_n + 255) & ~(size_t)255;
  const size_t b_cf = (8 * tot_n * sizeof(T) + 255) & ~(size_t)255;
  const size_t b_cf64 = (8 * nfine * sizeof(double) + 255) & ~(size_t)255;
  const size_t b_c964 = (c964.size() * sizeof(double) + 255) & ~(size_t)255;
  char* mem = nullptr;
  const size_t b_w = (wnull.size() * sizeof(double) + 255) & ~(size_t)255;
  EKM_CHECK(hipMalloc(&mem, b_desc + b_spec + 2 * b_cf + b_cf64 + b_c964 + b_w));
  tp.mem = mem;
  char* cur = mem + b_desc;
  unsigned char* d_spec = (unsigned char*)cur;
  cur += b_spec;
  T* d_cf = (T*)cur;
  cur += b_cf;
  T* d_cfy = (T*)cur;
  cur += b_cf;
  double* d_cf64 = (double*)cur;
  cur += b_cf64;
  double* d_c964 = (double*)cur;
  for (int k = 1; k <= ng; ++k) {
    g.lev[k].spec = d_spec + loff[k];
    g.lev[k].cf = d_cf + 8 * loff[k];
    g.lev[k].cfy = d_cfy + 8 * loff[k];
  }
  g.cf64 = d_cf64;
  g.c964 = d_c964;
  double* d_w = (double*)((char*)d_c964 + b_c964);
  g.wnull = wnull.empty() ? nullptr : d_w;
  if (!wnull.empty()) EKM_CHECK(hipMemcpy(d_w, wnull.data(), wnull.size() * sizeof(double), hipMemcpyHostToDevice));
  EKM_CHECK(hipMalloc(&tp.fac, fac.size() * sizeof(T)));
  g.fac = (const T*)tp.fac;
  EKM_CHECK(hipMemcpy(tp.fac, fac.data(), fac.size() * sizeof(T), hipMemcpyHostToDevice));
  EKM_CHECK(hipMemcpy(mem, &g, sizeof(g), hipMemcpyHostToDevice));
  EKM_CHECK(hipMemcpy(d_spec, spec.data(), tot_n, hipMemcpyHostToDevice));
  EKM_CHECK(hipMemcpy(d_cf, cfT.data(), 8 * tot_n * sizeof(T), hipMemcpyHostToDevice));
  EKM_CHECK(hipMemcpy(d_cfy, cfyT.data(), 8 * tot_n * sizeof(T), hipMemcpyHostToDevice));
  EKM_CHECK(hipMemcpy(d_cf64, hl[ng].c.data(), 8 * nfine * sizeof(double), hipMemcpyHostToDevice));
  EKM_CHECK(hipMemcpy(d_c964, c964.data(), c964.size() * sizeof(double), hipMemcpyHostToDevice));
  tp.ws_elems = g.ws;
  tp.ws64_elems = g.ws64;
  return EKM_OK;
}

}  // namespace

int tight_plan_build(TightPlan& tp, const ekm_params& p, const Dims& d, const HostMetrics& hm, const Ranges& rv,
                     const double* kx2_mode, int nmode, int nsolve_max, long lds_budget) {
  tp.fp32 = p.proj_mode == 2;
  tp.lds_budget = lds_budget;
  const int nf = d.NZ2 > d.NY2 ? d.NZ2 : d.NY2;  // longest line of the finest level
  tp.cm = nf <= 320 ? 5 : nf <= 576 ? 9 : nf <= 832 ? 13 : 0;
  if (!tp.cm) return EKM_ERR_UNSUPPORTED;
  tp.nsolve_max = nsolve_max;
  int st = tp.fp32 ? tight_build_typed<float>(tp, p, d, hm, rv, kx2_mode, nmode)
                   : tight_build_typed<double>(tp, p, d, hm, rv, kx2_mode, nmode);
  if (st) return st;
  const size_t esz = tp.fp32 ? sizeof(float) : sizeof(double);
  EKM_CHECK(hipMalloc(&tp.ws, (size_t)nsolve_max * tp.ws_elems * esz));
  EKM_CHECK(hipMalloc(&tp.ws64, (size_t)nsolve_max * tp.ws64_elems * sizeof(double)));
  return EKM_OK;
}

void tight_plan_free(TightPlan& tp) {
  if (tp.mem) (void)hipFree(tp.mem);
  if (tp.fac) (void)hipFree(tp.fac);
  if (tp.ws) (void)hipFree(tp.ws);
  if (tp.ws64) (void)hipFree(tp.ws64);
  tp = TightPlan();
}

int tight_solve_ex(TightPlan& tp, MgPlan& mg, const Dims& d, int nsolve, const double* guess, hipStream_t st) {
  if (nsolve <= 0) return EKM_OK;
  if (nsolve > tp.nsolve_max) return EKM_ERR_ARG;
  // a multiple of 16 workgroups for the XCD placement of the solve pairs
  const dim3 grid((unsigned)((nsolve + 15) & ~15));
  // workgroup size: tight_block
#define LAUNCH3(T, CM, X, PER)                                                                           \
  do {                                                                                                   \
    constexpr int NT = tight_block(sizeof(T) == 4, CM, X);                                                \
    EKM_CHECK(hipFuncSetAttribute((const void*)k_tight_solve<T, CM, NT, X, PER>,                        \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, tp.lds_bytes));            \
    hipLaunchKernelGGL((k_tight_solve<T, CM, NT, X, PER>), grid, dim3(NT), tp.lds_bytes, st, (const TDesc<T>*)tp.mem, \
                       (T*)tp.ws, tp.ws64, (const double*)mg.rhs, mg.phi, (const double*)mg.kx2, mg.cycles, \
                       mg.cycles + mg.nsolve_max, d.plane, guess, nsolve);                                                          \
  } while (0)
#define LAUNCH2(T, CM, X)          \
  do {                             \
    if (tp.per)                    \
      LAUNCH3(T, CM, X, true);     \
    else                           \
      LAUNCH3(T, CM, X, false);    \
  } while (0)
#define LAUNCH(T, CM)           \
  do {                          \
    if (tp.cross)               \
      LAUNCH2(T, CM, true);     \
    else                        \
      LAUNCH2(T, CM, false);    \
  } while (0)
  if (tp.fp32) {
    if (tp.cm == 5) LAUNCH(float, 5);
    else if (tp.cm == 9) LAUNCH(float, 9);
    else LAUNCH(float, 13);
  } else {
    if (tp.cm == 5) LAUNCH(double, 5);
    else if (tp.cm == 9) LAUNCH(double, 9);
    else LAUNCH(double, 13);
  }
#undef LAUNCH
#undef LAUNCH2
#undef LAUNCH3
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
