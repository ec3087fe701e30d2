// ekm_kernels.hip -- explicit RHS, nonlinear fluxes, ADI sweeps and the
// pressure projection of one RK3 substep (reference: src/duct.F90 RK_DUCT
// :237-2469 and REM_DIV_CURV :2472-2818).
//
// All kernels are memory bound FP64 passes over the device-resident fields.
// Physical kernels map consecutive threads to consecutive x (i) so every
// load of a (k,j) row is one coalesced run; Fourier kernels map threads to
// consecutive kx modes the same way.  The ADI sweeps give every thread one
// (i,k) or (i,j) line: for a fixed row the threads of a wavefront touch
// consecutive i, so each step of the sequential recurrence is a coalesced
// row load.  Without eddy viscosity the tridiagonal factors do not depend on
// i, so they are computed once per line by k_lu and shared; with it every
// point's system is formed inside its sweep.
#include "ekm_kernels.h"
#include <algorithm>

namespace {

__device__ inline size_t ph(const Dims& d, int i, int k, int j) {
  return (size_t)i + (size_t)d.NX2 * ((size_t)k + (size_t)d.NZ2 * (size_t)j);
}
__device__ inline size_t fo(const Dims& d, int m, int k, int j) {
  return (size_t)m + (size_t)d.NXC * ((size_t)k + (size_t)d.NZ2 * (size_t)j);
}

#define G11(k, j, f) M.g11[mf(d, k, j, f)]
#define G12(k, j, f) M.g12[mf(d, k, j, f)]
#define G22(k, j, f) M.g22[mf(d, k, j, f)]
#define C11(k, j, f) M.c11[mf(d, k, j, f)]
#define C12(k, j, f) M.c12[mf(d, k, j, f)]
#define C21(k, j, f) M.c21[mf(d, k, j, f)]
#define C22(k, j, f) M.c22[mf(d, k, j, f)]
#define JAC(k, j) M.jac[mi(d, k, j)]

__device__ inline bool in(int v, int lo, int hi) { return v >= lo && v <= hi; }

// ---------------------------------------------------------------------------
// explicit Fourier-space RHS (src/duct.F90:298-632)
// ---------------------------------------------------------------------------
__global__ void k_rhs_fourier(Dims d, Metrics M, RhsArgs a) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t n = (size_t)d.NXC * d.NZ2 * d.NY2;
  if (t >= n) return;
  int m = (int)(t % d.NXC);
  size_t r = t / d.NXC;
  int k = (int)(r % d.NZ2), j = (int)(r / d.NZ2);
  size_t q = t;
  const size_t dk = d.NXC, dj = (size_t)d.NXC * d.NZ2;
  const double temp3 = a.ss->temp3, temp4 = a.ss->temp4;
  if (in(j, a.rv.js, a.rv.je) && in(k, a.rv.zs, a.rv.ze)) {
    double Jc = JAC(k, j);
    const double2* CP = a.cp;
    double2 p = CP[q], pkp = CP[q + dk], pkm = CP[q - dk], pjp = CP[q + dj], pjm = CP[q - dj];
    double2 cr[3], cf[3], cu[3];
    for (int f = 0; f < 3; ++f) {
      cu[f] = a.cu[f][q];
      cr[f] = make_double2(Jc * cu[f].x, Jc * cu[f].y);
      if (a.rk > 1) {
        double2 o = a.cf[f][q];
        cr[f].x = cr[f].x + temp3 * o.x;
        cr[f].y = cr[f].y + temp3 * o.y;
      }
    }
    // pressure gradient as explicit Euler (:366-409)
    double a11p = 0.5 * C11(k + 1, j, 2), a11 = 0.5 * C11(k, j, 2);
    double a21p = 0.5 * C21(k, j + 1, 1), a21 = 0.5 * C21(k, j, 1);
    double a12p = 0.5 * C12(k + 1, j, 2), a12 = 0.5 * C12(k, j, 2);
    double a22p = 0.5 * C22(k, j + 1, 1), a22 = 0.5 * C22(k, j, 1);
    double2 cs1, cs2;
    cs1.x = a11p * (p.x + pkp.x) - a11 * (p.x + pkm.x) + a21p * (p.x + pjp.x) - a21 * (p.x + pjm.x);
    cs1.y = a11p * (p.y + pkp.y) - a11 * (p.y + pkm.y) + a21p * (p.y + pjp.y) - a21 * (p.y + pjm.y);
    cs2.x = a12p * (p.x + pkp.x) - a12 * (p.x + pkm.x) + a22p * (p.x + pjp.x) - a22 * (p.x + pjm.x);
    cs2.y = a12p * (p.y + pkp.y) - a12 * (p.y + pkm.y) + a22p * (p.y + pjp.y) - a22 * (p.y + pjm.y);
    double b = temp4 * a.kx[m] * Jc;
    cr[0].x = cr[0].x - (0.0 * p.x - b * p.y);
    cr[0].y = cr[0].y - (0.0 * p.y + b * p.x);
    cr[1].x = cr[1].x - temp4 * cs2.x;
    cr[1].y = cr[1].y - temp4 * cs2.y;
    cr[2].x = cr[2].x - temp4 * cs1.x;
    cr[2].y = cr[2].y - temp4 * cs1.y;
    if (m == 0 && a.px_on) cr[2].x = cr[2].x - temp4 * a.ss->px * Jc;  // (:431-453)
    // horizontal viscous terms and rotation (:461-475)
    double fv = a.nu * a.kx2[m] * Jc;
    for (int f = 0; f < 3; ++f) cf[f] = make_double2(-(fv * cu[f].x), -(fv * cu[f].y));
    double og = (a.cor_mean_only && m > 0) ? 0.0 : a.ug;
    double ow = (a.cor_mean_only && m > 0) ? 0.0 : a.wg;
    cf[2].x = cf[2].x + a.f0 * (cu[0].x - og) * Jc;
    cf[2].y = cf[2].y + a.f0 * cu[0].y * Jc;
    cf[0].x = cf[0].x - a.f0 * (cu[2].x - ow) * Jc;
    cf[0].y = cf[0].y - a.f0 * cu[2].y * Jc;
    if (a.th) {  // buoyancy (:517-523)
      double2 c = a.cth[q];
      double bj = Jc * a.ri_tau;
      cf[1].x = cf[1].x - bj * c.x;
      cf[1].y = cf[1].y - bj * c.y;
    }
    for (int f = 0; f < 3; ++f) {
      a.cr[f][q] = cr[f];
      a.cf[f][q] = cf[f];
    }
  } else {
    for (int f = 0; f < 3; ++f) a.cr[f][q] = make_double2(0.0, 0.0);
  }
  if (!a.th) return;
  // scalar (:532-632)
  if (in(j, a.rt.js, a.rt.je) && in(k, a.rt.zs, a.rt.ze)) {
    double Jc = JAC(k, j);
    double2 c = a.cth[q];
    double2 crth = make_double2(Jc * c.x, Jc * c.y);
    if (a.rk > 1) {
      double2 o = a.cfth[q];
      crth.x = crth.x + temp3 * o.x;
      crth.y = crth.y + temp3 * o.y;
    }
    double fth = -a.nupr * Jc * a.kx2[m];
    double2 cfth = make_double2(fth * c.x, fth * c.y);
    double2 u2 = a.cu[1][q];
    if (a.cont_strat) {
      cfth.x = cfth.x + Jc * u2.x;
      cfth.y = cfth.y + Jc * u2.y;
    } else {
      const double* TB = a.thbar;
      size_t tt = (size_t)k + (size_t)d.NZ2 * j, tk = 1, tj = d.NZ2;
      double gb = 0.5 * C12(k + 1, j, 2) * (TB[tt] + TB[tt + tk]) - 0.5 * C12(k, j, 2) * (TB[tt] + TB[tt - tk]) +
                  0.5 * C22(k, j + 1, 1) * (TB[tt] + TB[tt + tj]) - 0.5 * C22(k, j, 1) * (TB[tt] + TB[tt - tj]);
      cfth.x = cfth.x - u2.x * gb;
      cfth.y = cfth.y - u2.y * gb;
      double dif = 0.25 * a.nupr *
                   (+G12(k + 1, j, 2) * (TB[tt + tk + tj] + TB[tt + tj] - TB[tt - tj] - TB[tt + tk - tj]) -
                    G12(k, j, 2) * (TB[tt - tk + tj] + TB[tt + tj] - TB[tt - tj] - TB[tt - tk - tj]) +
                    G12(k, j + 1, 1) * (TB[tt + tk + tj] + TB[tt + tk] - TB[tt - tk + tj] - TB[tt - tk]) -
                    G12(k, j, 1) * (TB[tt + tk] + TB[tt + tk - tj] - TB[tt - tk] - TB[tt - tk - tj]) +
                    4.0 * (G11(k + 1, j, 2) * (TB[tt + tk] - TB[tt]) - G11(k, j, 2) * (TB[tt] - TB[tt - tk])) +
                    4.0 * (G22(k, j + 1, 1) * (TB[tt + tj] - TB[tt]) - G22(k, j, 1) * (TB[tt] - TB[tt - tj])));
      cfth.x = cfth.x + dif;
    }
    a.crth[q] = crth;
    a.cfth[q] = cfth;
  } else {
    a.crth[q] = make_double2(0.0, 0.0);
  }
}

// ---------------------------------------------------------------------------
// contravariant face velocities (:721-732)
// ---------------------------------------------------------------------------
__global__ void k_contravariant(Dims d, Metrics M, Ranges r, const double* __restrict__ U2,
                                const double* __restrict__ U3, double* __restrict__ U2b,
                                double* __restrict__ U3b) {
  int nk = r.ze + 1 - r.zs + 1, nj = r.je + 1 - r.js + 1;
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t n = (size_t)d.NX2 * nk * nj;
  if (t >= n) return;
  int i = (int)(t % d.NX2);
  size_t rr = t / d.NX2;
  int k = r.zs + (int)(rr % nk), j = r.js + (int)(rr / nk);
  size_t q = ph(d, i, k, j), dj = (size_t)d.NX2 * d.NZ2, dk = d.NX2;
  U2b[q] = 0.5 * C22(k, j, 1) * (U2[q] + U2[q - dj]) + 0.5 * C21(k, j, 1) * (U3[q] + U3[q - dj]);
  U3b[q] = 0.5 * C11(k, j, 2) * (U3[q] + U3[q - dk]) + 0.5 * C12(k, j, 2) * (U2[q] + U2[q - dk]);
}

// ---------------------------------------------------------------------------
// ADI: explicit half before the y sweep (:1334-1371, :1640-1665, ...)
// S = R - J U;  R = S/2 + J U + t1 [G11(K+1)(U(K+1)-U) - G11(K)(U-U(K-1))]
// ---------------------------------------------------------------------------
__global__ void k_adi_pre(Dims d, Metrics M, const AdiField* __restrict__ fs, AdiT1 t1s) {
  const AdiField& F = fs[blockIdx.y];
  const Ranges r = F.r;
  const double t1 = t1s.ss->temp1 / t1s.div[blockIdx.y];
  const double* __restrict__ U = F.U;
  double* __restrict__ R = F.R;
  double* __restrict__ S = F.S;
  int nk = r.ze - r.zs + 1, nj = r.je - r.js + 1;
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t n = (size_t)d.NX2 * nk * nj;
  if (t >= n) return;
  int i = (int)(t % d.NX2);
  size_t rr = t / d.NX2;
  int k = r.zs + (int)(rr % nk), j = r.js + (int)(rr / nk);
  size_t q = ph(d, i, k, j), dk = d.NX2;
  double Jc = JAC(k, j), u = U[q];
  double s = R[q] - Jc * u;
  S[q] = s;
  double rv = 0.5 * s + Jc * u + t1 * (G11(k + 1, j, 2) * (U[q + dk] - u) - G11(k, j, 2) * (u - U[q - dk]));
  if (F.nut) {  // subgrid viscosity, z direction (:1667-1680 and analogues)
    const double* __restrict__ nut = F.nut;
    const double* __restrict__ gz = F.gz;
    const double t2 = t1s.ss->hb / 2.0, nc = nut[q];
    rv = rv + t2 * (0.5 * (nut[q + dk] + nc) * (gz[mf(d, k + 1, j, 2)] * (U[q + dk] - u)) -
                    0.5 * (nc + nut[q - dk]) * (gz[mf(d, k, j, 2)] * (u - U[q - dk])));
  }
  R[q] = rv;
}

// Per-line factorisation of one sweep (THOMAS_REAL forward elimination,
// src/boundary.F90:53-84, with the rows of the APPLY_BC_* routines).  dir 0:
// y sweep, lines K = zs..ze, rows J = 0..NY+1; dir 1: z sweep, lines J, rows K.
// Output per (line, row): mult (the reference's modified A), piv (modified B),
// up (C).  Rows that neither the interior loop nor a BC routine sets keep the
// [0 1 0 | 0] the reference initialises them to (:1399-1406, :1690-1697).
// One workgroup per line: the rows' coefficients are formed in parallel into
// LDS, one lane runs the elimination recurrence out of LDS, and the factors
// go back to memory with contiguous stores.
constexpr int kLuBlock = 256;
constexpr int kLuMaxRows = 1040;
__global__ void __launch_bounds__(kLuBlock) k_lu(Dims d, Metrics M, const AdiField* __restrict__ fs, AdiT1 t1s,
                                                  int dir) {
  __shared__ double sa[kLuMaxRows], sb[kLuMaxRows], sc[kLuMaxRows];
  const AdiField& F = fs[blockIdx.y];
  if (F.nut) return;  // point-dependent systems: factored inside the sweeps
  const Ranges r = F.r;
  const BcTable& bc = dir == 0 ? F.by : F.bz;
  const double t1 = t1s.ss->temp1 / t1s.div[blockIdx.y];
  const size_t ly = (size_t)d.NZ2 * d.NY2;
  double* __restrict__ mult = F.lu + (dir == 0 ? 0 : 3 * ly);
  double* __restrict__ piv = mult + ly;
  double* __restrict__ up = mult + 2 * ly;
  double* __restrict__ spa = F.lu + 6 * ly;
  const int line0 = dir == 0 ? r.zs : r.js;
  if ((int)blockIdx.x > (dir == 0 ? r.ze - r.zs : r.je - r.js)) return;
  const int L = line0 + (int)blockIdx.x;
  const int N = dir == 0 ? d.NY + 1 : d.NZ + 1;  // last row index
  const int lo = dir == 0 ? r.js : r.zs, hi = dir == 0 ? r.je : r.ze;
  for (int row = threadIdx.x; row <= N; row += blockDim.x) {
    double a = 0.0, b = 1.0, c = 0.0;
    if (row >= lo && row <= hi) {
      const int k = dir == 0 ? L : row, j = dir == 0 ? row : L;
      if (dir == 0) {
        a = -(t1 * G22(k, j, 1));
        b = JAC(k, j) + t1 * (G22(k, j + 1, 1) + G22(k, j, 1));
        c = -(t1 * G22(k, j + 1, 1));
      } else {
        a = -(t1 * G11(k, j, 2));
        b = JAC(k, j) + t1 * (G11(k + 1, j, 2) + G11(k, j, 2));
        c = -(t1 * G11(k + 1, j, 2));
      }
    }
    for (int e = 0; e < 4; ++e)
      if (bc.r[e].row == row) {
        a = bc.r[e].a;
        b = bc.r[e].b;
        c = bc.r[e].c;
      }
    sa[row] = a;
    sb[row] = b;
    sc[row] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (dir == 1 && bc.sp) {
      // THOMAS_REAL_SP (src/boundary.F90:17-24): row NZ+1 -= row NZ * D/A(NZ)
      const double D = bc.sp_metric ? G11(d.NZ, L, 2) : 1.0;
      const double a_nz = sa[N - 1];
      sa[N] = sa[N] - sb[N - 1] * D / a_nz;
      sb[N] = sb[N] - sc[N - 1] * D / a_nz;
      spa[2 * L] = D;
      spa[2 * L + 1] = a_nz;
    }
    double cprev = 0.0, pvprev = 1.0;
    for (int row = 0; row <= N; ++row) {
      const double a = sa[row], b = sb[row], c = sc[row];
      double mlt = 0.0, pv = b;
      if (row > 0) {
        mlt = -a / pvprev;
        pv = b + mlt * cprev;
      }
      sa[row] = mlt;
      sb[row] = pv;
      cprev = c;
      pvprev = pv;
    }
  }
  __syncthreads();
  const size_t base = (size_t)L * (N + 1);
  for (int row = threadIdx.x; row <= N; row += blockDim.x) {
    mult[base + row] = sa[row];
    piv[base + row] = sb[row];
    up[base + row] = sc[row];
  }
}

__device__ inline double bc_rhs(const Dims& d, const Metrics& M, const BcRow& e, int K, double lag) {
  switch (e.kind) {
    case BC_LAGGED:
    case BC_WALL: return lag;
    case BC_THFLUX_LO: return e.val / (0.5 * (C22(K, 1, 2) + C22(K + 1, 1, 1))) * JAC(K, 1);
    case BC_THFLUX_NY: return 0.5 * e.val * (JAC(K, d.NY) + JAC(K, d.NY - 1)) / C22(K, d.NY - 1, 1);
    case BC_THFLUX_NY1: return 0.5 * e.val * (JAC(K, d.NY + 1) + JAC(K, d.NY)) / C22(K, d.NY, 1);
    default: return e.val;
  }
}

// y sweep for lines (i,K): solves in place into U, then emits the z-sweep
// right-hand side R = S/2 + J U* + t1 [G22(J+1)(U*(J+1)-U*) - G22(J)(U*-U*(J-1))]
// (:1460-1478 for theta, :1744-1757 for U1, ...).
// The sweeps below run each line's recurrence in blocks of kSw rows: the
// right-hand sides and factors of a block are loaded before its recurrence
// steps, so the memory latency is paid once per block instead of per row.
// kSw and the workgroup size are tuning parameters (ekm_tuning); the
// arithmetic does not depend on them.
template <int kSw>
__global__ void __launch_bounds__(256) k_ysweep(Dims d, Metrics M, const AdiField* __restrict__ fs, AdiT1 t1s) {
  const AdiField& F = fs[blockIdx.y];
  if (F.nut) return;
  const Ranges r = F.r;
  const BcTable& bc = F.by;
  const double t1 = t1s.ss->temp1 / t1s.div[blockIdx.y];
  const size_t ly = (size_t)d.NZ2 * d.NY2;
  const double* __restrict__ mult = F.lu;
  const double* __restrict__ piv = F.lu + ly;
  const double* __restrict__ up = F.lu + 2 * ly;
  double* __restrict__ U = F.U;
  double* __restrict__ R = F.R;
  const double* __restrict__ S = F.S;
  int nk = r.ze - r.zs + 1;
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)d.NX2 * nk) return;
  int i = (int)(t % d.NX2), K = r.zs + (int)(t / d.NX2);
  const int N = d.NY + 1;
  const size_t dj = (size_t)d.NX2 * d.NZ2;
  const size_t q0 = ph(d, i, K, 0);
  const size_t base = (size_t)K * (N + 1);
  double lag[4];
  for (int e = 0; e < 4; ++e)
    lag[e] = bc.r[e].kind == BC_LAGGED ? U[q0 + dj * bc.r[e].src]
             : bc.r[e].kind == BC_WALL ? F.wall[(size_t)bc.r[e].src * dj + (size_t)K * d.NX2 + i]
                                       : 0.0;
  double g = 0.0;
  // the loads of the next batch of rows are issued before this batch is
  // eliminated
  double rn[kSw], mn[kSw];
#pragma unroll
  for (int u = 0; u < kSw; ++u) {
    const int row = u;
    rn[u] = (row <= N && row >= r.js && row <= r.je) ? R[q0 + dj * row] : 0.0;
    mn[u] = row <= N ? mult[base + row] : 0.0;
  }
  for (int row0 = 0; row0 <= N; row0 += kSw) {
    double rr[kSw], mm[kSw];
#pragma unroll
    for (int u = 0; u < kSw; ++u) {
      rr[u] = rn[u];
      mm[u] = mn[u];
    }
    if (row0 + kSw <= N) {
#pragma unroll
      for (int u = 0; u < kSw; ++u) {
        const int row = row0 + kSw + u;
        rn[u] = (row <= N && row >= r.js && row <= r.je) ? R[q0 + dj * row] : 0.0;
        mn[u] = row <= N ? mult[base + row] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < kSw; ++u) {
      const int row = row0 + u;
      if (row > N) break;
      double rhs = rr[u];
      for (int e = 0; e < 4; ++e)
        if (bc.r[e].row == row) rhs = bc_rhs(d, M, bc.r[e], K, lag[e]);
      g = (row == 0) ? rhs : rhs + mm[u] * g;
      U[q0 + dj * row] = g;
    }
  }
  double x2 = 0.0, x1 = g / piv[base + N];  // x1 = x(row+1), x2 = x(row+2)
  U[q0 + dj * N] = x1;
  for (int row0 = N - 1; row0 >= 0; row0 -= kSw) {
    double gg[kSw], uu[kSw], pp[kSw], ss[kSw], jc[kSw], gp[kSw], g0[kSw];
#pragma unroll
    for (int u = 0; u < kSw; ++u) {
      const int row = row0 - u, jr = row + 1;
      const bool ok = row >= 0;
      const bool in = ok && jr >= r.js && jr <= r.je;
      gg[u] = ok ? U[q0 + dj * row] : 0.0;
      uu[u] = ok ? up[base + row] : 0.0;
      pp[u] = ok ? piv[base + row] : 1.0;
      ss[u] = in ? S[q0 + dj * jr] : 0.0;
      jc[u] = in ? JAC(K, jr) : 0.0;
      gp[u] = in ? G22(K, jr + 1, 1) : 0.0;
      g0[u] = in ? G22(K, jr, 1) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kSw; ++u) {
      const int row = row0 - u;
      if (row < 0) break;
      const double x0 = (gg[u] - uu[u] * x1) / pp[u];
      U[q0 + dj * row] = x0;
      const int jr = row + 1;  // R at row+1 now has x(row), x(row+1), x(row+2)
      if (jr >= r.js && jr <= r.je)
        R[q0 + dj * jr] = 0.5 * ss[u] + jc[u] * x1 + t1 * (gp[u] * (x2 - x1) - g0[u] * (x1 - x0));
      x2 = x1;
      x1 = x0;
    }
  }
}

// z sweep for lines (i,J) (:1503-1596 and analogues)
template <int kSw>
__global__ void __launch_bounds__(256) k_zsweep(Dims d, Metrics M, const AdiField* __restrict__ fs) {
  const AdiField& F = fs[blockIdx.y];
  if (F.nut) return;
  const Ranges r = F.r;
  const BcTable& bc = F.bz;
  const size_t ly = (size_t)d.NZ2 * d.NY2;
  const double* __restrict__ mult = F.lu + 3 * ly;
  const double* __restrict__ piv = F.lu + 4 * ly;
  const double* __restrict__ up = F.lu + 5 * ly;
  const double* __restrict__ spa = F.lu + 6 * ly;
  double* __restrict__ U = F.U;
  const double* __restrict__ R = F.R;
  int nj = r.je - r.js + 1;
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)d.NX2 * nj) return;
  int i = (int)(t % d.NX2), J = r.js + (int)(t / d.NX2);
  const int N = d.NZ + 1;
  const size_t dk = d.NX2;
  const size_t q0 = ph(d, i, 0, J);
  const size_t base = (size_t)J * (N + 1);
  double lag[4];
  for (int e = 0; e < 4; ++e) lag[e] = (bc.r[e].kind == BC_LAGGED) ? U[q0 + dk * bc.r[e].src] : 0.0;
  double g = 0.0, rprev = 0.0;
  // the loads of the next batch of rows are issued before this batch is
  // eliminated
  double rn[kSw], mn[kSw];
#pragma unroll
  for (int u = 0; u < kSw; ++u) {
    const int row = u;
    rn[u] = (row <= N && row >= r.zs && row <= r.ze) ? R[q0 + dk * row] : 0.0;
    mn[u] = row <= N ? mult[base + row] : 0.0;
  }
  for (int row0 = 0; row0 <= N; row0 += kSw) {
    double rr[kSw], mm[kSw];
#pragma unroll
    for (int u = 0; u < kSw; ++u) {
      rr[u] = rn[u];
      mm[u] = mn[u];
    }
    if (row0 + kSw <= N) {
#pragma unroll
      for (int u = 0; u < kSw; ++u) {
        const int row = row0 + kSw + u;
        rn[u] = (row <= N && row >= r.zs && row <= r.ze) ? R[q0 + dk * row] : 0.0;
        mn[u] = row <= N ? mult[base + row] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < kSw; ++u) {
      const int row = row0 + u;
      if (row > N) break;
      double rhs = rr[u];
      for (int e = 0; e < 4; ++e)
        if (bc.r[e].row == row) rhs = bc_rhs(d, M, bc.r[e], J, lag[e]);
      if (bc.sp && row == N) rhs = rhs - rprev * spa[2 * J] / spa[2 * J + 1];
      g = (row == 0) ? rhs : rhs + mm[u] * g;
      U[q0 + dk * row] = g;
      rprev = rhs;
    }
  }
  double x1 = g / piv[base + N];
  U[q0 + dk * N] = x1;
  for (int row0 = N - 1; row0 >= 0; row0 -= kSw) {
    double gg[kSw], uu[kSw], pp[kSw];
#pragma unroll
    for (int u = 0; u < kSw; ++u) {
      const int row = row0 - u;
      const bool ok = row >= 0;
      gg[u] = ok ? U[q0 + dk * row] : 0.0;
      uu[u] = ok ? up[base + row] : 0.0;
      pp[u] = ok ? piv[base + row] : 1.0;
    }
#pragma unroll
    for (int u = 0; u < kSw; ++u) {
      const int row = row0 - u;
      if (row < 0) break;
      const double x0 = (gg[u] - uu[u] * x1) / pp[u];
      U[q0 + dk * row] = x0;
      x1 = x0;
    }
  }
}

// kSwL: rows per batch of the sweeps with a subgrid viscosity, which keep
// more operands per row in flight than the plain sweeps (a tuning parameter)

// ---------------------------------------------------------------------------
// Sweeps with the subgrid viscosity (fields with nut): every (i, line) has its
// own system, so each thread forms its rows as it eliminates -- the molecular
// row, then the eddy-viscosity face terms (:1718-1731, :1819-1832 and the U2,
// U3 analogues), then the boundary rows of APPLY_BC_* override -- and keeps
// its pivots in the scratch field pv for the back substitution, which forms
// the upper coefficients again.  Thomas arithmetic as THOMAS_REAL(_SP)
// (src/boundary.F90:1-84).
// ---------------------------------------------------------------------------
template <int kSwL>
__global__ void __launch_bounds__(256) k_ysweep_les(Dims d, Metrics M, const AdiField* __restrict__ fs, AdiT1 t1s) {
  const AdiField& F = fs[blockIdx.y];
  if (!F.nut) return;
  const Ranges r = F.r;
  const BcTable& bc = F.by;
  const double t1 = t1s.ss->temp1 / t1s.div[blockIdx.y], t2 = t1s.ss->hb / 2.0, t2h = t2 * 0.5;
  double* __restrict__ U = F.U;
  double* __restrict__ R = F.R;
  const double* __restrict__ S = F.S;
  const double* __restrict__ nut = F.nut;
  const double* __restrict__ gy = F.gy;
  double* __restrict__ pv = F.pv;
  int nk = r.ze - r.zs + 1;
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)d.NX2 * nk) return;
  int i = (int)(t % d.NX2), K = r.zs + (int)(t / d.NX2);
  const int N = d.NY + 1;
  const size_t dj = (size_t)d.NX2 * d.NZ2;
  const size_t q0 = ph(d, i, K, 0);
  double lag[4];
  for (int e = 0; e < 4; ++e)
    lag[e] = bc.r[e].kind == BC_LAGGED ? U[q0 + dj * bc.r[e].src]
             : bc.r[e].kind == BC_WALL ? F.wall[(size_t)bc.r[e].src * dj + (size_t)K * d.NX2 + i]
                                       : 0.0;
  double g = 0.0, pvp = 1.0, cp = 0.0;
  double nm = 0.0, nc = nut[q0];  // NU_T at row-1 and row
  // G22 and the subgrid metric at row: the row+1 values of the row before
  // (the rows js..je are consecutive)
  double g0c = G22(K, r.js, 1), l0c = gy[mf(d, K, r.js, 1)];
  for (int row0 = 0; row0 <= N; row0 += kSwL) {
    double rr[kSwL], nn[kSwL];
#pragma unroll
    for (int u = 0; u < kSwL; ++u) {
      const int row = row0 + u;
      rr[u] = (row <= N && row >= r.js && row <= r.je) ? R[q0 + dj * row] : 0.0;
      nn[u] = row + 1 <= N ? nut[q0 + dj * (row + 1)] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kSwL; ++u) {
      const int row = row0 + u;
      if (row > N) break;
      const double np = nn[u];
      double a = 0.0, b = 1.0, c = 0.0, rhs = rr[u];
      if (row >= r.js && row <= r.je) {
        const double gp = G22(K, row + 1, 1), g0 = g0c;
        const double lp = gy[mf(d, K, row + 1, 1)];
        a = -(t1 * g0);
        b = JAC(K, row) + t1 * (gp + g0);
        c = -(t1 * gp);
        const double wm = t2h * (nc + nm) * l0c;
        const double wp = t2h * (nc + np) * lp;
        g0c = gp;
        l0c = lp;
        a = a - wm;
        b = b + wp + wm;
        c = c - wp;
      }
      for (int e = 0; e < 4; ++e)
        if (bc.r[e].row == row) {
          a = bc.r[e].a;
          b = bc.r[e].b;
          c = bc.r[e].c;
          rhs = bc_rhs(d, M, bc.r[e], K, lag[e]);
        }
      double p = b;
      if (row == 0) {
        g = rhs;
      } else {
        const double mlt = -a / pvp;
        p = b + mlt * cp;
        g = rhs + mlt * g;
      }
      U[q0 + dj * row] = g;
      pv[q0 + dj * row] = p;
      pvp = p;
      cp = c;
      nm = nc;
      nc = np;
    }
  }
  double x2 = 0.0, x1 = g / pvp;  // x1 = x(row+1), x2 = x(row+2)
  U[q0 + dj * N] = x1;
  double n1 = nut[q0 + dj * N], n2 = 0.0;  // NU_T at row+1, row+2
  // G22 and the subgrid metric at row+2 are the row+1 values of the row
  // before; only the first row of the backward pass loads them
  const bool in_n = N >= r.js && N <= r.je;
  double gpc = in_n ? G22(K, N + 1, 1) : 0.0, lpc = in_n ? gy[mf(d, K, N + 1, 1)] : 0.0;
  for (int row0 = N - 1; row0 >= 0; row0 -= kSwL) {
    double gg[kSwL], pp[kSwL], nl[kSwL], ss[kSwL], jc[kSwL], g0[kSwL], l0[kSwL];
#pragma unroll
    for (int u = 0; u < kSwL; ++u) {
      const int row = row0 - u, jr = row + 1;
      const bool ok = row >= 0;
      const bool in = ok && jr >= r.js && jr <= r.je;
      gg[u] = ok ? U[q0 + dj * row] : 0.0;
      pp[u] = ok ? pv[q0 + dj * row] : 1.0;
      nl[u] = ok ? nut[q0 + dj * row] : 0.0;
      ss[u] = in ? S[q0 + dj * jr] : 0.0;
      jc[u] = in ? JAC(K, jr) : 0.0;
      g0[u] = ok ? G22(K, jr, 1) : 0.0;
      l0[u] = ok ? gy[mf(d, K, jr, 1)] : 0.0;
    }
    double gp[kSwL], lp[kSwL];
    gp[0] = gpc;
    lp[0] = lpc;
#pragma unroll
    for (int u = 1; u < kSwL; ++u) {
      gp[u] = g0[u - 1];
      lp[u] = l0[u - 1];
    }
    gpc = g0[kSwL - 1];
    lpc = l0[kSwL - 1];
#pragma unroll
    for (int u = 0; u < kSwL; ++u) {
      const int row = row0 - u;
      if (row < 0) break;
      const double n0 = nl[u];
      double c = 0.0;
      if (row >= r.js && row <= r.je) c = -(t1 * g0[u]) - t2h * (n0 + n1) * l0[u];
      for (int e = 0; e < 4; ++e)
        if (bc.r[e].row == row) c = bc.r[e].c;
      const double x0 = (gg[u] - c * x1) / pp[u];
      U[q0 + dj * row] = x0;
      const int jr = row + 1;  // R at row+1 now has x(row), x(row+1), x(row+2)
      if (jr >= r.js && jr <= r.je) {
        double rv = 0.5 * ss[u] + jc[u] * x1 + t1 * (gp[u] * (x2 - x1) - g0[u] * (x1 - x0));
        rv = rv + t2 * (0.5 * (n2 + n1) * (lp[u] * (x2 - x1)) - 0.5 * (n0 + n1) * (l0[u] * (x1 - x0)));
        R[q0 + dj * jr] = rv;
      }
      x2 = x1;
      x1 = x0;
      n2 = n1;
      n1 = n0;
    }
  }
}

template <int kSwL>
__global__ void __launch_bounds__(256) k_zsweep_les(Dims d, Metrics M, const AdiField* __restrict__ fs, AdiT1 t1s) {
  const AdiField& F = fs[blockIdx.y];
  if (!F.nut) return;
  const Ranges r = F.r;
  const BcTable& bc = F.bz;
  const double t1 = t1s.ss->temp1 / t1s.div[blockIdx.y], t2h = (t1s.ss->hb / 2.0) * 0.5;
  double* __restrict__ U = F.U;
  const double* __restrict__ R = F.R;
  const double* __restrict__ nut = F.nut;
  const double* __restrict__ gz = F.gz;
  double* __restrict__ pv = F.pv;
  int nj = r.je - r.js + 1;
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)d.NX2 * nj) return;
  int i = (int)(t % d.NX2), J = r.js + (int)(t / d.NX2);
  const int N = d.NZ + 1;
  const size_t dk = d.NX2;
  const size_t q0 = ph(d, i, 0, J);
  double lag[4];
  for (int e = 0; e < 4; ++e) lag[e] = (bc.r[e].kind == BC_LAGGED) ? U[q0 + dk * bc.r[e].src] : 0.0;
  const double D = bc.sp_metric ? G11(d.NZ, J, 2) : 1.0;
  double g = 0.0, pvp = 1.0, cp = 0.0;
  double ap = 0.0, bp = 1.0, rp = 0.0;  // previous row before elimination (THOMAS_REAL_SP)
  double nm = 0.0, nc = nut[q0];
  // G11 and the subgrid metric at row: the row+1 values of the row before
  // (the rows zs..ze are consecutive)
  double g0c = G11(r.zs, J, 2), l0c = gz[mf(d, r.zs, J, 2)];
  // the loads of the next batch of rows are issued before this batch is
  // eliminated
  double rn[kSwL], nx_[kSwL];
#pragma unroll
  for (int u = 0; u < kSwL; ++u) {
    const int row = u;
    rn[u] = (row <= N && row >= r.zs && row <= r.ze) ? R[q0 + dk * row] : 0.0;
    nx_[u] = row + 1 <= N ? nut[q0 + dk * (row + 1)] : 0.0;
  }
  for (int row0 = 0; row0 <= N; row0 += kSwL) {
    double rr[kSwL], nn[kSwL];
#pragma unroll
    for (int u = 0; u < kSwL; ++u) {
      rr[u] = rn[u];
      nn[u] = nx_[u];
    }
    if (row0 + kSwL <= N) {
#pragma unroll
      for (int u = 0; u < kSwL; ++u) {
        const int row = row0 + kSwL + u;
        rn[u] = (row <= N && row >= r.zs && row <= r.ze) ? R[q0 + dk * row] : 0.0;
        nx_[u] = row + 1 <= N ? nut[q0 + dk * (row + 1)] : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < kSwL; ++u) {
      const int row = row0 + u;
      if (row > N) break;
      const double np = nn[u];
      double a = 0.0, b = 1.0, c = 0.0, rhs = rr[u];
      if (row >= r.zs && row <= r.ze) {
        const double gp = G11(row + 1, J, 2), g0 = g0c;
        const double lp = gz[mf(d, row + 1, J, 2)];
        a = -(t1 * g0);
        b = JAC(row, J) + t1 * (gp + g0);
        c = -(t1 * gp);
        const double wm = t2h * (nc + nm) * l0c;
        const double wp = t2h * (np + nc) * lp;
        g0c = gp;
        l0c = lp;
        a = a - wm;
        b = b + wp + wm;
        c = c - wp;
      }
      for (int e = 0; e < 4; ++e)
        if (bc.r[e].row == row) {
          a = bc.r[e].a;
          b = bc.r[e].b;
          c = bc.r[e].c;
          rhs = bc_rhs(d, M, bc.r[e], J, lag[e]);
        }
      if (bc.sp && row == N) {
        a = a - bp * D / ap;
        b = b - cp * D / ap;
        rhs = rhs - rp * D / ap;
      }
      double p = b;
      if (row == 0) {
        g = rhs;
      } else {
        const double mlt = -a / pvp;
        p = b + mlt * cp;
        g = rhs + mlt * g;
      }
      U[q0 + dk * row] = g;
      pv[q0 + dk * row] = p;
      ap = a;
      bp = b;
      rp = rhs;
      pvp = p;
      cp = c;
      nm = nc;
      nc = np;
    }
  }
  double x1 = g / pvp;
  U[q0 + dk * N] = x1;
  double n1 = nut[q0 + dk * N];
  for (int row0 = N - 1; row0 >= 0; row0 -= kSwL) {
    double gg[kSwL], pp[kSwL], nl[kSwL], gk[kSwL], lk[kSwL];
#pragma unroll
    for (int u = 0; u < kSwL; ++u) {
      const int row = row0 - u;
      const bool ok = row >= 0;
      gg[u] = ok ? U[q0 + dk * row] : 0.0;
      pp[u] = ok ? pv[q0 + dk * row] : 1.0;
      nl[u] = ok ? nut[q0 + dk * row] : 0.0;
      gk[u] = ok ? G11(row + 1, J, 2) : 0.0;
      lk[u] = ok ? gz[mf(d, row + 1, J, 2)] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kSwL; ++u) {
      const int row = row0 - u;
      if (row < 0) break;
      const double n0 = nl[u];
      double c = 0.0;
      if (row >= r.zs && row <= r.ze) c = -(t1 * gk[u]) - t2h * (n1 + n0) * lk[u];
      for (int e = 0; e < 4; ++e)
        if (bc.r[e].row == row) c = bc.r[e].c;
      const double x0 = (gg[u] - c * x1) / pp[u];
      U[q0 + dk * row] = x0;
      x1 = x0;
      n1 = n0;
    }
  }
}

// scalar corner updates after the z sweep (:1606-1622)
__global__ void k_th_corners(Dims d, double* __restrict__ U, int mode) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.NX2) return;
  int NZ = d.NZ, NY = d.NY;
  auto P = [&](int k, int j) -> double& { return U[ph(d, i, k, j)]; };
  if (mode & 1) {
    P(NZ + 1, NY + 1) = P(NZ, NY + 1);
    P(NZ + 1, 0) = P(NZ, 0);
    P(0, NY + 1) = P(1, NY + 1);
    P(0, 0) = P(1, 0);
  }
  if (mode & 2) {
    P(NZ + 1, NY + 1) = P(1, NY + 1);
    P(NZ + 1, 0) = P(1, 0);
    P(0, NY + 1) = P(NZ, NY + 1);
    P(0, 0) = P(NZ, 0);
  }
}

// ---------------------------------------------------------------------------
// CFL time step (courant_curv_mpi, src/duct.F90:3605-3674): min over the
// interior of cfl dx/|U1|, cfl J/|U2b|, cfl J/|U3b|.  Positive doubles order
// like their bit patterns, so the minimum is an atomicMin on uint64.
// ---------------------------------------------------------------------------
__global__ void k_cfl(Dims d, Metrics M, Ranges r, const double* __restrict__ U1,
                      const double* __restrict__ U2, const double* __restrict__ U3, double cfl,
                      double dx, unsigned long long* __restrict__ out) {
  int nk = r.ze - r.zs + 1, nj = r.je - r.js + 1;
  size_t n = (size_t)d.NX2 * nk * nj;
  double dt = 999.0;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(t % d.NX2);
    size_t rr = t / d.NX2;
    int k = r.zs + (int)(rr % nk), j = r.js + (int)(rr / nk);
    size_t q = ph(d, i, k, j), dj = (size_t)d.NX2 * d.NZ2, dk = d.NX2;
    double u2b = 0.5 * C22(k, j, 1) * (U2[q] + U2[q - dj]) + 0.5 * C21(k, j, 1) * (U3[q] + U3[q - dj]);
    double u3b = 0.5 * C11(k, j, 2) * (U3[q] + U3[q - dk]) + 0.5 * C12(k, j, 2) * (U2[q] + U2[q - dk]);
    double Jc = JAC(k, j);
    double dtx = cfl * dx / fabs(U1[q]);
    double dty = cfl * Jc / fabs(u2b);
    double dtz = cfl * Jc / fabs(u3b);
    dt = fmin(dt, fmin(dtx, fmin(dty, dtz)));
  }
  __shared__ double red[256];
  red[threadIdx.x] = dt;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMin(out, (unsigned long long)__double_as_longlong(red[0]));
}

// ---------------------------------------------------------------------------
// projection (REM_DIV_CURV)
// ---------------------------------------------------------------------------
// face fluxes of the Fourier velocity (:2511-2522), zero outside their range
__global__ void k_faces(Dims d, Metrics M, Ranges r, const double2* __restrict__ CU2,
                        const double2* __restrict__ CU3, double2* __restrict__ F2,
                        double2* __restrict__ F3) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t n = (size_t)d.NXC * d.NZ2 * d.NY2;
  if (t >= n) return;
  size_t rr = t / d.NXC;
  int k = (int)(rr % d.NZ2), j = (int)(rr / d.NZ2);
  if (!(in(j, r.js, r.je + 1) && in(k, r.zs, r.ze + 1))) {
    F2[t] = make_double2(0.0, 0.0);
    F3[t] = make_double2(0.0, 0.0);
    return;
  }
  size_t dk = d.NXC, dj = (size_t)d.NXC * d.NZ2;
  double a22 = 0.5 * C22(k, j, 1), a21 = 0.5 * C21(k, j, 1);
  double a11 = 0.5 * C11(k, j, 2), a12 = 0.5 * C12(k, j, 2);
  double2 u2 = CU2[t], u3 = CU3[t], u2j = CU2[t - dj], u3j = CU3[t - dj], u2k = CU2[t - dk], u3k = CU3[t - dk];
  F2[t] = make_double2(a22 * (u2.x + u2j.x) + a21 * (u3.x + u3j.x), a22 * (u2.y + u2j.y) + a21 * (u3.y + u3j.y));
  F3[t] = make_double2(a11 * (u3.x + u3k.x) + a12 * (u2.x + u2k.x), a11 * (u3.y + u3k.y) + a12 * (u2.y + u2k.y));
}

// RHS_TMP planes of the per-mode solves (:2527-2535 and :2620-2656): solve
// s = 2m + part, part 0 the imaginary and part 1 the real component; the
// divergence is formed on the fly from the face fluxes.
// Both transfers between the Fourier fields (modes contiguous) and the solve
// planes (points contiguous) go through an LDS tile of 32 modes x 32 plane
// points, so the loads and the stores of a wavefront are both contiguous.
constexpr int kTileM = 32, kTileP = 32;
__global__ void __launch_bounds__(256) k_mg_gather(Dims d, Metrics M, Ranges r, const double* __restrict__ kx,
                                                   const double2* __restrict__ CU1, const double2* __restrict__ F2,
                                                   const double2* __restrict__ F3, double* __restrict__ rhs,
                                                   int nmode) {
  __shared__ double2 tile[kTileP][kTileM + 1];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int m0 = blockIdx.x * kTileM;
  const size_t p0 = (size_t)blockIdx.y * kTileP;
  for (int pp = ty; pp < kTileP; pp += 8) {  // lanes along the modes
    const size_t p = p0 + pp;
    const int m = m0 + tx;
    double vi = 0.0, vr = 0.0;
    if (p < d.plane && m < nmode) {
      const int k = (int)(p % d.NZ2), j = (int)(p / d.NZ2);
      if (in(j, r.js, r.je) && in(k, r.zs, r.ze)) {
        size_t q = fo(d, m, k, j), dk = d.NXC, dj = (size_t)d.NXC * d.NZ2;
        double kj = kx[m] * JAC(k, j);
        double2 u = CU1[q], f2 = F2[q], f2p = F2[q + dj], f3 = F3[q], f3p = F3[q + dk];
        double re = (0.0 * u.x - kj * u.y) + (f2p.x - f2.x) + (f3p.x - f3.x);
        double im = (0.0 * u.y + kj * u.x) + (f2p.y - f2.y) + (f3p.y - f3.y);
        vi = -im;
        vr = -re;
      }
    }
    tile[pp][tx] = make_double2(vi, vr);
  }
  __syncthreads();
  for (int mm = ty; mm < kTileM; mm += 8) {  // lanes along the plane points
    const int m = m0 + mm;
    const size_t p = p0 + tx;
    if (m < nmode && p < d.plane) {
      const double2 v = tile[tx][mm];
      rhs[(size_t)(2 * m) * d.plane + p] = v.x;
      rhs[(size_t)(2 * m + 1) * d.plane + p] = v.y;
    }
  }
}

// phi = CMPLX(P_TMP2, P_TMP) (:2662-2666), modes above NKX are zero
__global__ void __launch_bounds__(256) k_mg_scatter(Dims d, const double* __restrict__ phi, int nmode, int fp32,
                                                    double2* __restrict__ PHI) {
  __shared__ double2 tile[kTileM][kTileP + 1];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int m0 = blockIdx.x * kTileM;
  const size_t p0 = (size_t)blockIdx.y * kTileP;
  for (int mm = ty; mm < kTileM; mm += 8) {  // lanes along the plane points
    const int m = m0 + mm;
    const size_t p = p0 + tx;
    double2 v = make_double2(0.0, 0.0);
    if (m < nmode && p < d.plane) {
      v.x = phi[(size_t)(2 * m + 1) * d.plane + p];
      v.y = phi[(size_t)(2 * m) * d.plane + p];
      if (fp32) {
        v.x = (double)(float)v.x;
        v.y = (double)(float)v.y;
      }
    }
    tile[mm][tx] = v;
  }
  __syncthreads();
  for (int pp = ty; pp < kTileP; pp += 8) {  // lanes along the modes; modes above NKX get zero
    const int m = m0 + tx;
    const size_t p = p0 + pp;
    if (m < d.NXC && p < d.plane) PHI[(size_t)m + (size_t)d.NXC * p] = tile[tx][pp];
  }
}

__device__ inline double2 face2(const Dims& d, const Metrics& M, const Ranges& r, const double2* F2,
                                const double2* P, int m, int k, int j) {
  size_t q = fo(d, m, k, j);
  double2 f = F2[q];
  if (in(j, r.js, r.je) && in(k, r.zs, r.ze)) {
    size_t dk = d.NXC, dj = (size_t)d.NXC * d.NZ2;
    double g = G22(k, j, 1), h = 0.25 * G12(k, j, 1);
    double2 p0 = P[q], pj = P[q - dj], pk = P[q + dk], pkj = P[q + dk - dj], qk = P[q - dk], qkj = P[q - dk - dj];
    f.x = f.x - g * (p0.x - pj.x) - h * (pk.x + pkj.x - qk.x - qkj.x);
    f.y = f.y - g * (p0.y - pj.y) - h * (pk.y + pkj.y - qk.y - qkj.y);
  }
  return f;
}
__device__ inline double2 face3(const Dims& d, const Metrics& M, const Ranges& r, const double2* F3,
                                const double2* P, int m, int k, int j) {
  size_t q = fo(d, m, k, j);
  double2 f = F3[q];
  if (in(j, r.js, r.je) && in(k, r.zs, r.ze)) {
    size_t dk = d.NXC, dj = (size_t)d.NXC * d.NZ2;
    double g = G11(k, j, 2), h = 0.25 * G12(k, j, 2);
    double2 p0 = P[q], pk = P[q - dk], pj = P[q + dj], pkj = P[q - dk + dj], qj = P[q - dj], qkj = P[q - dk - dj];
    f.x = f.x - g * (p0.x - pk.x) - h * (pj.x + pkj.x - qj.x - qkj.x);
    f.y = f.y - g * (p0.y - pk.y) - h * (pj.y + pkj.y - qj.y - qkj.y);
  }
  return f;
}

// velocity correction (:2713-2750) and the divergence of the corrected face
// fluxes (:2754-2761); the corrected faces are formed on the fly and not kept.
__global__ void k_project(Dims d, Metrics M, Ranges r, const double* __restrict__ kx,
                          const double2* __restrict__ P, const double2* __restrict__ F2,
                          const double2* __restrict__ F3, double2* __restrict__ CU1,
                          double2* __restrict__ CU2, double2* __restrict__ CU3,
                          double2* __restrict__ DIV) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t n = (size_t)d.NXC * d.NZ2 * d.NY2;
  if (t >= n) return;
  int m = (int)(t % d.NXC);
  size_t rr = t / d.NXC;
  int k = (int)(rr % d.NZ2), j = (int)(rr / d.NZ2);
  if (!(in(j, r.js, r.je) && in(k, r.zs, r.ze))) {
    DIV[t] = make_double2(0.0, 0.0);
    return;
  }
  size_t q = t, dk = d.NXC, dj = (size_t)d.NXC * d.NZ2;
  double w = kx[m];
  double2 p0 = P[q], pjp = P[q + dj], pjm = P[q - dj], pkp = P[q + dk], pkm = P[q - dk];
  double2 u1 = CU1[q], u2 = CU2[q], u3 = CU3[q];
  double Jc = JAC(k, j);
  u1.x = u1.x - (0.0 * p0.x - w * p0.y);
  u1.y = u1.y - (0.0 * p0.y + w * p0.x);
  double b22p = C22(k, j + 1, 1), b22 = C22(k, j, 1), b12p = C12(k + 1, j, 2), b12 = C12(k, j, 2);
  double b11p = C11(k + 1, j, 2), b11 = C11(k, j, 2), b21p = C21(k, j + 1, 1), b21 = C21(k, j, 1);
  u2.x = u2.x - 0.5 * (b22p * (p0.x + pjp.x) - b22 * (p0.x + pjm.x) + b12p * (pkp.x + p0.x) - b12 * (p0.x + pkm.x)) / Jc;
  u2.y = u2.y - 0.5 * (b22p * (p0.y + pjp.y) - b22 * (p0.y + pjm.y) + b12p * (pkp.y + p0.y) - b12 * (p0.y + pkm.y)) / Jc;
  u3.x = u3.x - 0.5 * (b11p * (p0.x + pkp.x) - b11 * (p0.x + pkm.x) + b21p * (pjp.x + p0.x) - b21 * (p0.x + pjm.x)) / Jc;
  u3.y = u3.y - 0.5 * (b11p * (p0.y + pkp.y) - b11 * (p0.y + pkm.y) + b21p * (pjp.y + p0.y) - b21 * (p0.y + pjm.y)) / Jc;
  CU1[q] = u1;
  CU2[q] = u2;
  CU3[q] = u3;
  double2 f2 = face2(d, M, r, F2, P, m, k, j), f2p = face2(d, M, r, F2, P, m, k, j + 1);
  double2 f3 = face3(d, M, r, F3, P, m, k, j), f3p = face3(d, M, r, F3, P, m, k + 1, j);
  double kj = w * Jc;
  DIV[q] = make_double2((0.0 * u1.x - kj * u1.y) + (f2p.x - f2.x) + (f3p.x - f3.x),
                        (0.0 * u1.y + kj * u1.x) + (f2p.y - f2.y) + (f3p.y - f3.y));
}

// CPX += phi / TEMP4 over J = js-1..je+1, K = zs-1..ze+1 (:2460-2466)
__global__ void k_pupdate(Dims d, Ranges r, const double2* __restrict__ P, double2* __restrict__ CP,
                          const StepState* __restrict__ ss) {
  int nk = r.ze - r.zs + 3, nj = r.je - r.js + 3;
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t n = (size_t)d.NXC * nk * nj;
  if (t >= n) return;
  int m = (int)(t % d.NXC);
  size_t rr = t / d.NXC;
  int k = r.zs - 1 + (int)(rr % nk), j = r.js - 1 + (int)(rr / nk);
  size_t q = fo(d, m, k, j);
  const double temp4 = ss->temp4;
  double2 c = CP[q], p = P[q];
  c.x = c.x + p.x / temp4;
  c.y = c.y + p.y / temp4;
  CP[q] = c;
}

// x statistics of SAVE_STATS_CURVI (src/flow_statistics.F90:163-215, :420-495)
// at every (K,J), from the Fourier state by Parseval: with the forward
// transform scaled by 1/NX and the modes above NX/3 zero,
//   sum_i f'_i g'_i = 2 NX sum_{m>=1} Re(f_m conj g_m)   (i over the NX points).
// One wavefront per (K,J) pencil, lanes over the modes.  Planes of out
// ((NZ+2)*(NY+2) each): 0-4 the means of U1, U2, U3, P, TH (the real parts of
// mode 0); 5-8 the reference's urms, vrms, wrms, thrms (un-normalised sums of
// squared fluctuations); 9-15 uv, uw, wv, pu (= w'p'), pv, thv, thw (sums / NX).
constexpr int kXStats = 16;
__global__ void __launch_bounds__(256) k_xstats(Dims d, const double2* __restrict__ U1,
                                                const double2* __restrict__ U2, const double2* __restrict__ U3,
                                                const double2* __restrict__ P, const double2* __restrict__ TH,
                                                double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const size_t pen = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (pen >= d.plane) return;
  const size_t f0 = pen * d.NXC;
  double a[11];
  for (int q = 0; q < 11; ++q) a[q] = 0.0;
  for (int m = 1 + lane; m < d.NXC; m += 64) {
    const double2 u = U1[f0 + m], v = U2[f0 + m], w = U3[f0 + m], p = P[f0 + m];
    const double2 t = TH ? TH[f0 + m] : make_double2(0.0, 0.0);
    // the Nyquist mode has no conjugate partner
    const double wgt = (2 * m == d.NX) ? 1.0 : 2.0;
    auto re = [](double2 x, double2 y) { return x.x * y.x + x.y * y.y; };
    a[0] += wgt * re(u, u);
    a[1] += wgt * re(v, v);
    a[2] += wgt * re(w, w);
    a[3] += wgt * re(t, t);
    a[4] += wgt * re(u, v);
    a[5] += wgt * re(u, w);
    a[6] += wgt * re(w, v);
    a[7] += wgt * re(w, p);
    a[8] += wgt * re(v, p);
    a[9] += wgt * re(t, v);
    a[10] += wgt * re(t, w);
  }
  for (int o = 32; o > 0; o >>= 1)
    for (int q = 0; q < 11; ++q) a[q] += __shfl_xor(a[q], o, 64);
  if (lane == 0) {
    const size_t pl = d.plane;
    const double nx = (double)d.NX;
    out[0 * pl + pen] = U1[f0].x;
    out[1 * pl + pen] = U2[f0].x;
    out[2 * pl + pen] = U3[f0].x;
    out[3 * pl + pen] = P[f0].x;
    out[4 * pl + pen] = TH ? TH[f0].x : 0.0;
    for (int q = 0; q < 4; ++q) out[(5 + q) * pl + pen] = nx * a[q];
    for (int q = 4; q < 11; ++q) out[(5 + q) * pl + pen] = a[q];
  }
}

// flow rate of the F_TYPE 0 controller (:416-428): the sum over J = JSTART
// .. JEND-1, K = ZSTART+1 .. ZEND-1 of (W0(K,J+1) + W0(K,J)) dy(K,J) / 2 with
// W0 the real part of the mean (kx = 0) mode of U3 and dy(K,J) =
// ypoint(K,J+1) - ypoint(K,J); one workgroup, written to out[0]
__global__ void __launch_bounds__(256) k_flow_rate(Dims d, Ranges r, const double2* __restrict__ CU3,
                                                   const double* __restrict__ dy, double* __restrict__ out) {
  __shared__ double red[256];
  const int nk = r.ze - r.zs - 1, nj = r.je - r.js;
  double acc = 0.0;
  for (int t = threadIdx.x; t < nk * nj; t += blockDim.x) {
    const int k = r.zs + 1 + t % nk, j = r.js + t / nk;
    const double w1 = CU3[fo(d, 0, k, j + 1)].x, w0 = CU3[fo(d, 0, k, j)].x;
    acc += (w1 + w0) * dy[(size_t)k + (size_t)d.NZ2 * j] / 2;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// sponge_th (src/duct.F90:3196-3362).  Fluctuations, modes 1 and up: CF -=
// (Sponge_Amp SIGMA(J)) CU INT_JACOB for the three velocities and the scalar
// (:3228-3236), then the same with SIGMA_OUT(K) when the streamwise profile is
// on (:3240-3250).  Without it, SIGMA is zero below NY_S_T, where the update is
// the identity, so only rows from j0 are visited.
__global__ void k_sponge_fluct(Dims d, Metrics M, SpongeArgs a) {
  const int nm = d.NXC - 1, nk = a.r.ze - a.r.zs + 1;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)nm * nk * (a.r.je - a.j0 + 1)) return;
  const int m = 1 + (int)(t % nm);
  const size_t rr = t / nm;
  const int k = a.r.zs + (int)(rr % nk), j = a.j0 + (int)(rr / nk);
  const size_t q = fo(d, m, k, j);
  const double Jc = JAC(k, j), as = a.ss->amp * a.sigma[j];
  const double ao = a.sigma_out ? a.ss->amp * a.sigma_out[k] : 0.0;
  for (int f = 0; f < 4; ++f) {
    double2* cf = f < 3 ? a.cf[f] : a.cfth;
    const double2 u = f < 3 ? a.cu[f][q] : a.cth[q];
    double2 c = cf[q];
    c.x = c.x - (as * u.x) * Jc;
    c.y = c.y - (as * u.y) * Jc;
    if (a.sigma_out) {
      c.x = c.x - (ao * u.x) * Jc;
      c.y = c.y - (ao * u.y) * Jc;
    }
    cf[q] = c;
  }
}

// Mode 0 of sponge_th, one workgroup: the mean of Re CU3 over J = NY_sponge..
// JEND, K = ZSTART..ZEND (:3261-3272) decides the dynamic sponge (:3283-3305):
// beyond 0.1% of WBULK0 the three velocities relax to their targets and the
// scalar to zero, else only CU1 and CU2.  With type-6 z-max rows and
// stratification (:3324-3332), CF2 of mode 0 relaxes on the whole plane with
// the Jacobian at the exit indices of the loops above, and CU2X(0,:,:) is
// zeroed on the steps where MOD(2*TIME_STEP, SAVE_STATS_INT) = 0.  Last, the
// streamwise profile's inflow/outflow terms with twice the amplitude
// (:3350-3356), on the possibly zeroed CU2.
__global__ void __launch_bounds__(1024) k_sponge_mean(Dims d, Metrics M, SpongeArgs a) {
  __shared__ double red[1024];
  const int nk = a.r.ze - a.r.zs + 1, nj = a.r.je - a.ny_sponge + 1;
  double acc = 0.0;
  for (int t = threadIdx.x; t < nk * nj; t += blockDim.x)
    acc += a.cu[2][fo(d, 0, a.r.zs + t % nk, a.ny_sponge + t / nk)].x;
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double mean3 = red[0] / (double)(nk * nj);
  const bool dyn = fabs(mean3 - a.t3) > a.t3 * 0.001;
  const int nji = a.r.je - a.r.js + 1;
  for (int t = threadIdx.x; t < nk * nji; t += blockDim.x) {
    const int k = a.r.zs + t % nk, j = a.r.js + t / nk;
    const size_t q = fo(d, 0, k, j);
    const double Jc = JAC(k, j), as = a.ss->amp * a.sigma[j];
    const double2 u1 = a.cu[0][q], u2 = a.cu[1][q];
    double2 c1 = a.cf[0][q], c2 = a.cf[1][q];
    c1.x = c1.x - (as * (u1.x - a.t1)) * Jc;
    c1.y = c1.y - (as * u1.y) * Jc;
    c2.x = c2.x - (as * (u2.x - a.t2)) * Jc;
    c2.y = c2.y - (as * u2.y) * Jc;
    a.cf[0][q] = c1;
    a.cf[1][q] = c2;
    if (dyn) {
      const double2 u3 = a.cu[2][q], th = a.cth[q];
      double2 c3 = a.cf[2][q], ct = a.cfth[q];
      c3.x = c3.x - (as * (u3.x - a.t3)) * Jc;
      c3.y = c3.y - (as * u3.y) * Jc;
      ct.x = ct.x - (as * th.x) * Jc;
      ct.y = ct.y - (as * th.y) * Jc;
      a.cf[2][q] = c3;
      a.cfth[q] = ct;
    }
  }
  if (a.zmax6) {
    __syncthreads();
    for (int t = threadIdx.x; t < (int)d.plane; t += blockDim.x) {
      const size_t q = fo(d, 0, t % d.NZ2, t / d.NZ2);
      const double2 u2 = a.cu[1][q];
      double2 c2 = a.cf[1][q];
      c2.x = c2.x - (a.ss->amp * (u2.x - a.t2)) * a.jst;
      c2.y = c2.y - (a.ss->amp * u2.y) * a.jst;
      a.cf[1][q] = c2;
      if (a.zero_cu2) a.cu[1][q] = make_double2(0.0, 0.0);
    }
  }
  if (!a.sigma_out) return;
  __syncthreads();
  for (int t = threadIdx.x; t < nk * nji; t += blockDim.x) {
    const int k = a.r.zs + t % nk, j = a.r.js + t / nk;
    const size_t q = fo(d, 0, k, j);
    const double Jc = JAC(k, j), ao = (2.0 * a.ss->amp) * a.sigma_out[k];
    const double tg[4] = {0.0, 0.0, a.t3, 0.0};
    for (int f = 0; f < 4; ++f) {
      double2* cf = f < 3 ? a.cf[f] : a.cfth;
      const double2 u = f < 3 ? a.cu[f][q] : a.cth[q];
      double2 c = cf[q];
      c.x = c.x - (ao * (u.x - tg[f])) * Jc;
      c.y = c.y - (ao * u.y) * Jc;
      cf[q] = c;
    }
  }
}

// The mean-gradient statistics of SAVE_STATS_CURVI (src/flow_statistics.F90)
// from the x means (real parts of mode 0), one workgroup.  Planes (K fastest):
// 0 dudy, 1 dwdy over ZSTART..ZEND x JSTART..JEND (:228-240), 2 dudz,
// 3 dwdz (:250-260), 4 dthdy with the background gradient when CONT_STRAT
// is off (:464-480), 5 dthdz (:495-503), 6 Rig (:520-551), 7 the mean-flow
// kinetic energy TKE (:164-175); then THBULK (:517-553).  The derivatives are
// the contravariant face metrics averaged to the cell (0.125 times four faces)
// times index differences, over INT_JACOB; the reference's row and column
// fills follow in its order.  Points it never writes stay zero.
__device__ inline double cell_avg(const Dims& d, const double* c, int k, int j) {
  return 0.125 * (((c[mf(d, k, j, 1)] + c[mf(d, k, j + 1, 1)]) + c[mf(d, k, j, 2)]) + c[mf(d, k + 1, j, 2)]);
}
__global__ void __launch_bounds__(1024) k_mean_grads(Dims d, Metrics M, MeanGradArgs a) {
  __shared__ double red[1024];
  const int NZ = d.NZ, NY = d.NY, NZ2 = d.NZ2;
  const int pl = (int)d.plane;
  double* o = a.out;
  auto P = [&](int f, int k, int j) -> double& { return o[(size_t)f * pl + k + NZ2 * j]; };
  auto m0 = [&](const double2* f, int k, int j) { return f[fo(d, 0, k, j)].x; };
  for (int t = threadIdx.x; t < 8 * pl; t += blockDim.x) o[t] = 0.0;
  __syncthreads();
  for (int t = threadIdx.x; t < pl; t += blockDim.x) {
    const int k = t % NZ2, j = t / NZ2;
    const double u1 = m0(a.u1, k, j), u2 = m0(a.u2, k, j), u3 = m0(a.u3, k, j);
    P(7, k, j) = ((u1 * u1 + u2 * u2) + u3 * u3) * a.ke_scale;
    if (k >= 1 && k <= NZ && j >= 1 && j <= NY) {
      const double Jc = JAC(k, j);
      const double a11 = cell_avg(d, M.c11, k, j), a12 = cell_avg(d, M.c12, k, j);
      const double a21 = cell_avg(d, M.c21, k, j), a22 = cell_avg(d, M.c22, k, j);
      if (in(k, a.rv.zs, a.rv.ze) && in(j, a.rv.js, a.rv.je)) {
        P(0, k, j) = (a12 * (m0(a.u1, k + 1, j) - m0(a.u1, k - 1, j)) + a22 * (m0(a.u1, k, j + 1) - m0(a.u1, k, j - 1))) / Jc;
        P(1, k, j) = (a12 * (m0(a.u3, k + 1, j) - m0(a.u3, k - 1, j)) + a22 * (m0(a.u3, k, j + 1) - m0(a.u3, k, j - 1))) / Jc;
      }
      P(2, k, j) = (a21 * (m0(a.u1, k, j + 1) - m0(a.u1, k, j - 1)) + a11 * (m0(a.u1, k + 1, j) - m0(a.u1, k - 1, j))) / Jc;
      P(3, k, j) = (a21 * (m0(a.u3, k, j + 1) - m0(a.u3, k, j - 1)) + a11 * (m0(a.u3, k + 1, j) - m0(a.u3, k - 1, j))) / Jc;
      if (a.th) {
        double g = (a12 * (m0(a.th, k + 1, j) - m0(a.th, k - 1, j)) + a22 * (m0(a.th, k, j + 1) - m0(a.th, k, j - 1))) / Jc;
        if (a.thbar) {
          const double* T = a.thbar;
          const int q = k + NZ2 * j;
          g = g + (a12 * (T[q + 1] - T[q - 1]) + a22 * (T[q + NZ2] - T[q - NZ2])) / Jc;
        }
        P(4, k, j) = g;
        P(5, k, j) = (a21 * (m0(a.th, k, j + 1) - m0(a.th, k, j - 1)) + a11 * (m0(a.th, k + 1, j) - m0(a.th, k - 1, j))) / Jc;
      }
    }
  }
  __syncthreads();
  // rows 0 and NY+1 of the velocity gradients (:242-249, :262-267); columns 0
  // and NZ+1 of the scalar gradients (:482-486, :505-508)
  for (int k = 1 + (int)threadIdx.x; k <= NZ; k += blockDim.x)
    for (int f = 0; f < 4; ++f) {
      P(f, k, 0) = P(f, k, 1);
      P(f, k, NY + 1) = P(f, k, NY);
    }
  if (a.th)
    for (int j = 1 + (int)threadIdx.x; j <= NY; j += blockDim.x)
      for (int f = 4; f < 6; ++f) {
        P(f, 0, j) = P(f, 1, j);
        P(f, NZ + 1, j) = P(f, NZ, j);
      }
  __syncthreads();
  // column 1 of dudz, dwdz (:269-272); rows 0 and NY+1 of the scalar
  // gradients, extrapolated (:488-491, :510-513)
  for (int j = 1 + (int)threadIdx.x; j <= NY; j += blockDim.x) {
    P(3, 1, j) = P(3, 2, j);
    P(2, 1, j) = P(2, 2, j);
  }
  if (a.th)
    for (int k = threadIdx.x; k <= NZ + 1; k += blockDim.x)
      for (int f = 4; f < 6; ++f) {
        P(f, k, 0) = 2.0 * P(f, k, 1) - P(f, k, 2);
        P(f, k, NY + 1) = 2.0 * P(f, k, NY) - P(f, k, NY - 1);
      }
  __syncthreads();
  // gradient Richardson number and THBULK (:517-553)
  double acc = 0.0;
  if (a.th) {
    for (int t = threadIdx.x; t < NZ * NY; t += blockDim.x) {
      const int k = 1 + t % NZ, j = 1 + t / NZ;
      acc += 0.25 * (((m0(a.th, k, j) + m0(a.th, k - 1, j)) + m0(a.th, k - 1, j - 1)) + m0(a.th, k, j - 1));
      const double gth = a.cont_strat ? P(4, k, j) - 1.0 : P(4, k, j);
      const double dw = P(1, k, j), du = P(0, k, j);
      P(6, k, j) = dw == 0.0 ? -(a.ri_tau * gth) * 1.0e18 : -(a.ri_tau * gth) / (dw * dw + du * du);
    }
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) o[(size_t)8 * pl] = a.th ? red[0] / (double)(float)(NY * NZ) : 0.0;
}

// Mean streamwise vorticity of SAVE_STATS_CURVI (src/flow_statistics.F90:
// 301-327): omega_x = dW/dy - dV/dz of the x means (real parts of mode 0),
// each derivative the cell-averaged face metrics times central index
// differences over INT_JACOB, for K = 1..NZ, J = 1..NY; rows 0 and NY+1 copy
// rows 1 and NY (:323-326); columns 0 and NZ+1 stay zero.  One thread per
// point, each row fill recomputing its source row.
__global__ void __launch_bounds__(256) k_omega_x(Dims d, Metrics M, const double2* __restrict__ u2,
                                                 const double2* __restrict__ u3, double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int)d.plane) return;
  const int k = t % d.NZ2, jo = t / d.NZ2;
  if (k < 1 || k > d.NZ) {
    out[t] = 0.0;
    return;
  }
  const int j = jo < 1 ? 1 : jo > d.NY ? d.NY : jo;
  auto m0 = [&](const double2* f, int kk, int jj) { return f[fo(d, 0, kk, jj)].x; };
  const double Jc = JAC(k, j);
  const double a11 = cell_avg(d, M.c11, k, j), a12 = cell_avg(d, M.c12, k, j);
  const double a21 = cell_avg(d, M.c21, k, j), a22 = cell_avg(d, M.c22, k, j);
  const double dwdy = (a12 * (m0(u3, k + 1, j) - m0(u3, k - 1, j)) + a22 * (m0(u3, k, j + 1) - m0(u3, k, j - 1))) / Jc;
  const double dvdz = (a21 * (m0(u2, k, j + 1) - m0(u2, k, j - 1)) + a11 * (m0(u2, k + 1, j) - m0(u2, k - 1, j))) / Jc;
  out[t] = dwdy - dvdz;
}

// One plane of a physical field (the gather under the reference's plane
// writers, src/flow_statistics.F90:2091-2629): axis 0 the x-y plane at
// K = index, out[j][i] (NY+2 rows of NX); axis 1 the z-y plane at I = index,
// out[j][k]; axis 2 the x-z plane at J = index, out[k][i].
__global__ void __launch_bounds__(256) k_plane_gather(Dims d, int axis, int index, const double* __restrict__ U,
                                                      double* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (axis == 0) {
    if (t >= (size_t)d.NY2 * d.NX) return;
    const int i = (int)(t % d.NX), j = (int)(t / d.NX);
    out[t] = U[ph(d, i, index, j)];
  } else if (axis == 1) {
    if (t >= d.plane) return;
    const int k = (int)(t % d.NZ2), j = (int)(t / d.NZ2);
    out[t] = U[ph(d, index, k, j)];
  } else {
    if (t >= (size_t)d.NZ2 * d.NX) return;
    const int i = (int)(t % d.NX), k = (int)(t / d.NX);
    out[t] = U[ph(d, i, k, index)];
  }
}

// high_oder_momentums_curvi (src/flow_statistics.F90:1026-1095): one
// wavefront per (K,J) pencil sums, over the NX physical points, the powers of
// the velocity fluctuations about the mode-0 means (real parts of the Fourier
// state), divided by NX.  Planes: UU, VV, WW, UUU, VVV, WWW, UUUU, VVVV, WWWW,
// UUV, WWV, UUUV, WWWV.
__global__ void __launch_bounds__(256) k_xmoments(Dims d, const double* __restrict__ U1p,
                                                  const double* __restrict__ U2p, const double* __restrict__ U3p,
                                                  const double2* __restrict__ U1, const double2* __restrict__ U2,
                                                  const double2* __restrict__ U3, double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const size_t pen = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (pen >= d.plane) return;
  const int k = (int)(pen % d.NZ2), j = (int)(pen / d.NZ2);
  const double mu = U1[fo(d, 0, k, j)].x, mv = U2[fo(d, 0, k, j)].x, mw = U3[fo(d, 0, k, j)].x;
  double s[13];
#pragma unroll
  for (int q = 0; q < 13; ++q) s[q] = 0.0;
  for (int i = lane; i < d.NX; i += 64) {
    const size_t q = ph(d, i, k, j);
    const double u = U1p[q] - mu, v = U2p[q] - mv, w = U3p[q] - mw;
    s[0] += u * u; s[1] += v * v; s[2] += w * w;
    s[3] += u * u * u; s[4] += v * v * v; s[5] += w * w * w;
    s[6] += (u * u) * (u * u); s[7] += (v * v) * (v * v); s[8] += (w * w) * (w * w);
    s[9] += u * u * v; s[10] += w * w * v; s[11] += u * u * u * v; s[12] += w * w * w * v;
  }
#pragma unroll
  for (int q = 0; q < 13; ++q) {
    double v = s[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) out[(size_t)q * d.plane + pen] = v / (double)d.NX;
  }
}

// divergence diagnostic (:2790-2800): sum and max of |div| over
// J = 2..NY-1, K = 2..NZ-1, I = 1..NX+1 of the physical field.  A block takes
// one (K, J) line at a time (x contiguous, so the loads are coalesced), each
// thread issues the loads of kDivLd points of the line before using any, and
// the line index is 32-bit (no 64-bit division per point).
constexpr int kDivBlock = 256, kDivLd = 4;
__global__ void __launch_bounds__(kDivBlock) k_div_reduce(Dims d, const double* __restrict__ S,
                                                          double* __restrict__ out) {
  const int nk = d.NZ - 2, nj = d.NY - 2, ni = d.NX + 1;
  const int nl = nk * nj;
  double s = 0.0, mx = 0.0;
  for (int l = blockIdx.x; l < nl; l += gridDim.x) {
    const int k = 2 + l % nk, j = 2 + l / nk;
    const double* __restrict__ row = S + ph(d, 0, k, j);
    for (int i0 = 1; i0 <= ni; i0 += kDivBlock * kDivLd) {
      double v[kDivLd];
#pragma unroll
      for (int u = 0; u < kDivLd; ++u) {
        const int i = i0 + (int)threadIdx.x + kDivBlock * u;
        v[u] = i <= ni ? fabs(row[i]) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kDivLd; ++u) {
        s += v[u];
        mx = fmax(mx, v[u]);
      }
    }
  }
  __shared__ double rs[kDivBlock], rm[kDivBlock];
  rs[threadIdx.x] = s;
  rm[threadIdx.x] = mx;
  __syncthreads();
  for (int w = kDivBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      rs[threadIdx.x] += rs[threadIdx.x + w];
      rm[threadIdx.x] = fmax(rm[threadIdx.x], rm[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&out[0], rs[0]);
    atomicMax(reinterpret_cast<unsigned long long*>(&out[1]), (unsigned long long)__double_as_longlong(rm[0]));
  }
}

// generic THOMAS_REAL with per-system coefficients (the drop-in entry point):
// systems i = 0..nx, unknowns j = 0..ny, element (i,j) at i + (nx+1) j.
// With ksp >= 0 it is THOMAS_REAL_SP (src/boundary.F90:4-47): row ksp+1 is
// first eliminated by row ksp with the factor dd, on the fly (the inputs are
// not modified).
__global__ void k_thomas_generic(const double* __restrict__ A, const double* __restrict__ B,
                                 const double* __restrict__ C, double* __restrict__ G,
                                 double* __restrict__ W, int ny, int nx, int ksp, double dd) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nx) return;
  size_t ld = (size_t)nx + 1;
  double ak = 0.0, bk = 0.0, ck = 0.0, gk = 0.0;
  if (ksp >= 0) {
    const size_t p = i + ld * ksp;
    ak = A[p];
    bk = B[p];
    ck = C[p];
    gk = G[p];
  }
  double bp = B[i];
  W[i] = bp;
  double g = G[i];
  for (int j = 0; j < ny; ++j) {
    size_t qn = i + ld * (j + 1);
    double an = A[qn], bn = B[qn], gn = G[qn];
    if (j == ksp) {
      an = an - bk * dd / ak;
      bn = bn - ck * dd / ak;
      gn = gn - gk * dd / ak;
    }
    double mlt = -an / bp;
    bp = bn + mlt * C[qn - ld];
    g = gn + mlt * g;
    W[qn] = bp;
    G[qn] = g;
  }
  double x = g / bp;
  G[i + ld * ny] = x;
  for (int j = ny - 1; j >= 0; --j) {
    size_t q = i + ld * j;
    x = (G[q] - C[q] * x) / W[q];
    G[q] = x;
  }
}

// LES_WALL_MODEL, WALL_MODEL_TYPE 3 (src/les_chan.F90:1914-2144, one rank
// holding every x point, MIN(NXP,NXP_L) = NX-1), one workgroup:
// the x means of U1 and U3 at JSTART per K (sums over the NX points divided
// by NX-1, :1996-1999), their z means U_1, W_1 over K = 1..NZ, M = |(U_1,W_1)|,
// thirty fixed-point steps of U_STAR = M / (ln(Z U_STAR / NU)/KAPA + 5.2)
// (:2004-2007; 5.2 is a default real), TAO_WALL = U_STAR^2 cos(alpha),
// TAO_WALL_1 = U_STAR^2 sin(alpha) (:2012-2013); then per (I,K) the row
// values W = U3/W_1 TAO/NU / (C22 face mean) J and U = max(U1/W_1 TAO/NU,
// U1/U_1 TAO_1/NU) / (C22 face mean) J (:2101-2102).  On TIME_STEP 1 the U3
// value is U3X(I,K,2) and the U1 value is kept (:2115-2118).
__global__ void __launch_bounds__(1024) k_wall_model(Dims d, Metrics M, WallArgs a) {
  __shared__ double sh[4];
  __shared__ double tu[kWallMaxNz2], tw[kWallMaxNz2];
  const size_t pl = (size_t)d.NX2 * d.NZ2;
  double* __restrict__ wu = a.wall;
  double* __restrict__ ww = a.wall + pl;
  const int js = a.r.js, NX = d.NX, nk = a.r.ze - a.r.zs + 1;
  if (!a.first) {
    const double nxm = (double)(NX - 1);
    for (int k = a.r.zs + (int)threadIdx.x; k <= a.r.ze; k += blockDim.x) {
      double su = 0.0, sw = 0.0;
      for (int i = 0; i < NX; ++i) {
        su += a.u1[ph(d, i, k, js)];
        sw += a.u3[ph(d, i, k, js)];
      }
      tu[k] = su / nxm;
      tw[k] = sw / nxm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double su = 0.0, sw = 0.0;
      for (int k = 1; k <= d.NZ; ++k) {
        su += tu[k];
        sw += tw[k];
      }
      const double u1m = su / (double)d.NZ, w1m = sw / (double)d.NZ;
      const double m = sqrt(u1m * u1m + w1m * w1m);
      double us = m;
      for (int it = 0; it < 30; ++it) us = m / ((1.0 / 0.41) * log(a.z / a.nu * us) + (double)5.2f);
      sh[0] = u1m;
      sh[1] = w1m;
      sh[2] = us * us * a.cos_a;
      sh[3] = us * us * a.sin_a;
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < NX * nk; t += blockDim.x) {
    const int i = t % NX, k = a.r.zs + t / NX;
    const size_t q = (size_t)k * d.NX2 + i;
    if (a.first) {
      ww[q] = a.u3[ph(d, i, k, 2)];
      continue;
    }
    const double den = 0.5 * (C22(k, 1, 2) + C22(k + 1, 1, 1)), jk = JAC(k, 1);
    const double u = a.u1[ph(d, i, k, js)], w = a.u3[ph(d, i, k, js)];
    ww[q] = w / sh[1] * (sh[2] / a.nu) / den * jk;
    wu[q] = fmax(u / sh[1] * (sh[2] / a.nu), u / sh[0] * (sh[3] / a.nu)) / den * jk;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned nblk(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

int launch_rhs_fourier(const Dims& d, const Metrics& M, const RhsArgs& a, hipStream_t st) {
  size_t n = (size_t)d.NXC * d.NZ2 * d.NY2;
  hipLaunchKernelGGL(k_rhs_fourier, dim3(nblk(n, 256)), dim3(256), 0, st, d, M, a);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_contravariant(const Dims& d, const Metrics& M, const Ranges& r, const double* U2,
                         const double* U3, double* U2b, double* U3b, hipStream_t st) {
  size_t n = (size_t)d.NX2 * (r.ze - r.zs + 2) * (r.je - r.js + 2);
  hipLaunchKernelGGL(k_contravariant, dim3(nblk(n, 256)), dim3(256), 0, st, d, M, r, U2, U3, U2b, U3b);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_adi_batch(const Dims& d, const Metrics& M, const AdiField* fs, const AdiField* hf, int nf,
                     const AdiT1& t1, int th_field, int th_corners, hipStream_t st, const ekm_tuning& tn) {
  if (nf < 1 || nf > kAdiMaxFields) return EKM_ERR_ARG;
  const int bs = tn.adi_block;
  if (d.NY + 2 > kLuMaxRows || d.NZ + 2 > kLuMaxRows) return EKM_ERR_UNSUPPORTED;
  size_t npre = 0, nly = 0, nlz = 0, nys = 0, nzs = 0;
  bool plain = false, les = false;
  for (int f = 0; f < nf; ++f) {
    if (hf[f].nut) {
      if (!hf[f].gz || !hf[f].gy || !hf[f].pv) return EKM_ERR_ARG;
      les = true;
    } else {
      plain = true;
    }
    const Ranges& r = hf[f].r;
    const size_t nk = (size_t)(r.ze - r.zs + 1), nj = (size_t)(r.je - r.js + 1);
    npre = std::max(npre, (size_t)d.NX2 * nk * nj);
    nly = std::max(nly, nk);
    nlz = std::max(nlz, nj);
    nys = std::max(nys, (size_t)d.NX2 * nk);
    nzs = std::max(nzs, (size_t)d.NX2 * nj);
  }
  hipLaunchKernelGGL(k_adi_pre, dim3(nblk(npre, 256), nf), dim3(256), 0, st, d, M, fs, t1);
  EKM_CHECK(hipGetLastError());
  if (plain) {
    hipLaunchKernelGGL(k_lu, dim3((unsigned)nly, nf), dim3(kLuBlock), 0, st, d, M, fs, t1, 0);
    EKM_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_lu, dim3((unsigned)nlz, nf), dim3(kLuBlock), 0, st, d, M, fs, t1, 1);
    EKM_CHECK(hipGetLastError());
    const dim3 g(nblk(nys, bs), nf), b(bs);
    switch (tn.adi_rows) {
      case 4: hipLaunchKernelGGL(k_ysweep<4>, g, b, 0, st, d, M, fs, t1); break;
      case 12: hipLaunchKernelGGL(k_ysweep<12>, g, b, 0, st, d, M, fs, t1); break;
      default: hipLaunchKernelGGL(k_ysweep<8>, g, b, 0, st, d, M, fs, t1); break;
    }
    EKM_CHECK(hipGetLastError());
  }
  if (les) {
    const dim3 g(nblk(nys, 256), nf), b(256);
    switch (tn.adi_les_rows) {
      case 4: hipLaunchKernelGGL(k_ysweep_les<4>, g, b, 0, st, d, M, fs, t1); break;
      case 5: hipLaunchKernelGGL(k_ysweep_les<5>, g, b, 0, st, d, M, fs, t1); break;
      case 8: hipLaunchKernelGGL(k_ysweep_les<8>, g, b, 0, st, d, M, fs, t1); break;
      default: hipLaunchKernelGGL(k_ysweep_les<6>, g, b, 0, st, d, M, fs, t1); break;
    }
    EKM_CHECK(hipGetLastError());
  }
  if (plain) {
    const dim3 g(nblk(nzs, bs), nf), b(bs);
    switch (tn.adi_rows) {
      case 4: hipLaunchKernelGGL(k_zsweep<4>, g, b, 0, st, d, M, fs); break;
      case 12: hipLaunchKernelGGL(k_zsweep<12>, g, b, 0, st, d, M, fs); break;
      default: hipLaunchKernelGGL(k_zsweep<8>, g, b, 0, st, d, M, fs); break;
    }
    EKM_CHECK(hipGetLastError());
  }
  if (les) {
    const dim3 g(nblk(nzs, 256), nf), b(256);
    switch (tn.adi_les_rows) {
      case 4: hipLaunchKernelGGL(k_zsweep_les<4>, g, b, 0, st, d, M, fs, t1); break;
      case 5: hipLaunchKernelGGL(k_zsweep_les<5>, g, b, 0, st, d, M, fs, t1); break;
      case 8: hipLaunchKernelGGL(k_zsweep_les<8>, g, b, 0, st, d, M, fs, t1); break;
      default: hipLaunchKernelGGL(k_zsweep_les<6>, g, b, 0, st, d, M, fs, t1); break;
    }
    EKM_CHECK(hipGetLastError());
  }
  if (th_corners && th_field >= 0 && th_field < nf) {
    hipLaunchKernelGGL(k_th_corners, dim3(nblk(d.NX2, 256)), dim3(256), 0, st, d, hf[th_field].U, th_corners);
    EKM_CHECK(hipGetLastError());
  }
  return EKM_OK;
}
int launch_wall_model(const Dims& d, const Metrics& M, const WallArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_wall_model, dim3(1), dim3(1024), 0, st, d, M, a);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_cfl(const Dims& d, const Metrics& M, const Ranges& r, const double* U1, const double* U2,
               const double* U3, double cfl, double dx, unsigned long long* out, hipStream_t st) {
  hipLaunchKernelGGL(k_cfl, dim3(1024), dim3(256), 0, st, d, M, r, U1, U2, U3, cfl, dx, out);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_faces(const Dims& d, const Metrics& M, const Ranges& r, const double* CU2, const double* CU3,
                 double* F2, double* F3, hipStream_t st) {
  size_t n = (size_t)d.NXC * d.NZ2 * d.NY2;
  hipLaunchKernelGGL(k_faces, dim3(nblk(n, 256)), dim3(256), 0, st, d, M, r, (const double2*)CU2,
                     (const double2*)CU3, (double2*)F2, (double2*)F3);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_mg_gather(const Dims& d, const Metrics& M, const Ranges& r, const double* kx, const double* CU1,
                     const double* F2, const double* F3, double* rhs, int nmode, hipStream_t st) {
  const dim3 grid((unsigned)((nmode + kTileM - 1) / kTileM), (unsigned)((d.plane + kTileP - 1) / kTileP));
  hipLaunchKernelGGL(k_mg_gather, grid, dim3(256), 0, st, d, M, r, kx, (const double2*)CU1,
                     (const double2*)F2, (const double2*)F3, rhs, nmode);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_mg_scatter(const Dims& d, const double* phi, int nmode, int fp32, double* PHI, hipStream_t st) {
  const dim3 grid((unsigned)((d.NXC + kTileM - 1) / kTileM), (unsigned)((d.plane + kTileP - 1) / kTileP));
  hipLaunchKernelGGL(k_mg_scatter, grid, dim3(256), 0, st, d, phi, nmode, fp32, (double2*)PHI);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_project(const Dims& d, const Metrics& M, const Ranges& r, const double* kx, const double* PHI,
                   const double* F2, const double* F3, double* CU1, double* CU2, double* CU3, double* DIV,
                   hipStream_t st) {
  size_t n = (size_t)d.NXC * d.NZ2 * d.NY2;
  hipLaunchKernelGGL(k_project, dim3(nblk(n, 256)), dim3(256), 0, st, d, M, r, kx, (const double2*)PHI,
                     (const double2*)F2, (const double2*)F3, (double2*)CU1, (double2*)CU2, (double2*)CU3,
                     (double2*)DIV);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_pupdate(const Dims& d, const Ranges& r, const double* PHI, double* CP, const StepState* ss,
                   hipStream_t st) {
  size_t n = (size_t)d.NXC * (r.ze - r.zs + 3) * (r.je - r.js + 3);
  hipLaunchKernelGGL(k_pupdate, dim3(nblk(n, 256)), dim3(256), 0, st, d, r, (const double2*)PHI, (double2*)CP,
                     ss);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_flow_rate(const Dims& d, const Ranges& r, const double* CU3, const double* dy, double* out,
                     hipStream_t st) {
  if (r.ze - r.zs - 1 < 1 || r.je - r.js < 1) return EKM_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(k_flow_rate, dim3(1), dim3(256), 0, st, d, r, (const double2*)CU3, dy, out);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_xstats(const Dims& d, const double* U1, const double* U2, const double* U3, const double* P,
                  const double* TH, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_xstats, dim3(nblk(d.plane, 4)), dim3(256), 0, st, d, (const double2*)U1, (const double2*)U2,
                     (const double2*)U3, (const double2*)P, (const double2*)TH, out);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_sponge(const Dims& d, const Metrics& M, const SpongeArgs& a, hipStream_t st) {
  if (a.r.je >= a.j0 && d.NXC > 1) {
    const size_t n = (size_t)(d.NXC - 1) * (a.r.ze - a.r.zs + 1) * (a.r.je - a.j0 + 1);
    hipLaunchKernelGGL(k_sponge_fluct, dim3(nblk(n, 256)), dim3(256), 0, st, d, M, a);
    EKM_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_sponge_mean, dim3(1), dim3(1024), 0, st, d, M, a);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_mean_grads(const Dims& d, const Metrics& M, const MeanGradArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_mean_grads, dim3(1), dim3(1024), 0, st, d, M, a);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_omega_x(const Dims& d, const Metrics& M, const double* U2, const double* U3, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_omega_x, dim3(nblk(d.plane, 256)), dim3(256), 0, st, d, M, (const double2*)U2,
                     (const double2*)U3, out);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
size_t plane_elems(const Dims& d, int axis) {
  return axis == 0 ? (size_t)d.NY2 * d.NX : axis == 1 ? d.plane : (size_t)d.NZ2 * d.NX;
}
int launch_plane_gather(const Dims& d, int axis, int index, const double* U, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_plane_gather, dim3(nblk(plane_elems(d, axis), 256)), dim3(256), 0, st, d, axis, index, U, out);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_xmoments(const Dims& d, const double* U1p, const double* U2p, const double* U3p, const double* U1,
                    const double* U2, const double* U3, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_xmoments, dim3(nblk(d.plane, 4)), dim3(256), 0, st, d, U1p, U2p, U3p, (const double2*)U1,
                     (const double2*)U2, (const double2*)U3, out);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_div_reduce(const Dims& d, const double* S, double* out, hipStream_t st) {
  if (d.NZ < 3 || d.NY < 3) return EKM_OK;
  const int nl = (d.NZ - 2) * (d.NY - 2);
  hipLaunchKernelGGL(k_div_reduce, dim3((unsigned)std::min(nl, 4096)), dim3(kDivBlock), 0, st, d, S, out);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_thomas_generic(const double* A, const double* B, const double* C, double* G, double* W, int ny,
                          int nx, hipStream_t st, int ksp, double dd) {
  hipLaunchKernelGGL(k_thomas_generic, dim3(nblk((size_t)nx + 1, 256)), dim3(256), 0, st, A, B, C, G, W, ny, nx,
                     ksp, dd);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}

// ---------------------------------------------------------------------------
// step-state control kernels: one thread each, launched between the substep's
// data-parallel kernels, so the substep's scalar decisions stay on the device
// ---------------------------------------------------------------------------
namespace {

// constants of RK substep rk (src/duct.F90:283-287, :431-453; sponge_th's
// Sponge_Amp, src/duct.F90:3220)
__global__ void k_step_begin(StepState* ss, int rk, StepConsts c) {
  const double hb = ss->h_bar[rk - 1];
  ss->hb = hb;
  ss->temp1 = c.nu * hb / 2.0;
  ss->temp3 = c.zeta[rk - 1] * hb;
  ss->temp4 = hb;
  ss->temp5 = c.beta[rk - 1] * hb;
  if (c.f_type == 0)
    ss->px = ss->px0;  // the controller updates it before the RHS
  else if (c.f_type == 1)
    ss->px = c.px0;
  else
    ss->px = c.px0 + c.amp_omega0 * cos(c.omega0 * ss->time);
  ss->amp = 0.1 / ss->delta_t;
  ss->prev_ratio = ss->hb_prev > 0.0 ? hb / ss->hb_prev : 0.0;
  ss->guess = c.guess_on == 1 ? hb / ss->hb_prev : c.guess_on == 2 ? 1.0 : 0.0;
}

// F_TYPE 0: PX0 <- PX0 - (WEIGHT/TEMP4)(Qn + Q0 - 2 Qn1) with the flow rate of
// the state entering the substep (src/duct.F90:414-436); WEIGHT is the
// single-precision 0.1 the reference assigns (src/diablo.F90:361)
__global__ void k_flow_ctrl(StepState* ss, double q_den, int nzz, double wbulk0) {
  const double sum = ss->red[3];
  ss->red[3] = 0.0;
  const double qn1 = sum / q_den / (double)(float)nzz;
  const double weight = (double)0.1f;
  ss->px0 = ss->px0 - (weight / ss->temp4) * (ss->qn + wbulk0 - 2 * qn1);
  ss->qn = qn1;
  ss->px = ss->px0;
}

// DELTA_T from the CFL minimum, capped at DELTA_T_SET, and H_BAR
// (src/duct.F90:2417-2419, :3605-3674; src/diablo.F90:568-576)
__global__ void k_dt_update(StepState* ss) {
  const double dt = ss->red[2];
  ss->red[2] = 999.0;
  if (dt > 0) {
    const double d = dt >= ss->delta_t_set ? ss->delta_t_set : dt;
    ss->delta_t = d;
    ss->h_bar[0] = d * (8.0 / 15.0);
    ss->h_bar[1] = d * (2.0 / 15.0);
    ss->h_bar[2] = d * (5.0 / 15.0);
  }
}

// the projection's diagnostics: DIV (src/duct.F90:2790-2800), its max, the
// cycle totals, and the STOP conditions of src/duct.F90:2811-2812 as a sticky
// status (the first failure is kept)
__global__ void __launch_bounds__(64) k_div_final(StepState* ss, const int* __restrict__ cycles,
                                                  const int* __restrict__ unconv, int nsolve, double denom) {
  int sum = 0, mx = 0, nun = 0;
  for (int q = threadIdx.x; q < nsolve; q += 64) {
    const int c = cycles[q];
    sum += c;
    mx = c > mx ? c : mx;
    if (unconv) nun += unconv[q];
  }
  for (int o = 32; o > 0; o >>= 1) {
    sum += __shfl_xor(sum, o);
    nun += __shfl_xor(nun, o);
    const int m2 = __shfl_xor(mx, o);
    mx = m2 > mx ? m2 : mx;
  }
  if (threadIdx.x != 0) return;
  const double dr = ss->red[0] / denom;
  ss->div_ref = dr;
  ss->div_max = ss->red[1];
  ss->red[0] = 0.0;
  ss->red[1] = 0.0;
  ss->mg_solves = nsolve;
  ss->mg_cycles = sum;
  ss->mg_max_cycles = mx;
  ss->mg_unconverged = nun;
  ss->prof_launches += 1;
  ss->prof_solves += nsolve;
  ss->prof_cycles += sum;
  int st = EKM_OK;
  if (isnan(dr))
    st = EKM_ERR_DIV_NAN;
  else if (ss->delta_t < 1.0e-4)
    st = EKM_ERR_DT_SMALL;
  if (ss->status == EKM_OK) ss->status = st;
}

__global__ void k_step_end(StepState* ss, int rk) {
  ss->hb_prev = ss->hb;
  if (rk == 3) ss->time = ss->time + ss->delta_t;
}

}  // namespace

int launch_step_begin(StepState* ss, int rk, const StepConsts& c, hipStream_t st) {
  if (rk < 1 || rk > 3) return EKM_ERR_ARG;
  hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, st, ss, rk, c);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_flow_ctrl(StepState* ss, double q_den, int nzz, double wbulk0, hipStream_t st) {
  hipLaunchKernelGGL(k_flow_ctrl, dim3(1), dim3(1), 0, st, ss, q_den, nzz, wbulk0);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_dt_update(StepState* ss, hipStream_t st) {
  hipLaunchKernelGGL(k_dt_update, dim3(1), dim3(1), 0, st, ss);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_div_final(StepState* ss, const int* cycles, const int* unconv, int nsolve, double denom,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_div_final, dim3(1), dim3(64), 0, st, ss, cycles, unconv, nsolve, denom);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
int launch_step_end(StepState* ss, int rk, hipStream_t st) {
  hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1), 0, st, ss, rk);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
