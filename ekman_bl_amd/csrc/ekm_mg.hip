// ekm_mg.hip -- batched MUDPACK-replica multigrid for the per-kx pressure
// problems (MULTIGRID_MUDPACK, src/mg_solver2d.F90:75-201, solver mud2cr,
// src/mud2cr.f / src/mudcom.f) on gfx950.
//
// Every real solve of a projection (2 per live kx mode) runs in ONE launch:
// one workgroup owns one solve for its whole life -- full multigrid start,
// up to maxcy W(2,1) cycles and the relmax stopping test -- so no cycle
// needs a host round trip.  The order of operations of a cycle does not
// depend on the data, so the host expands kcymd2cr's kount state machine
// (src/mud2cr.f:340-499) once into a flat program of (operation, level)
// pairs; the kernel walks that program, which keeps exactly one copy of every
// primitive in the code object.
//
// Levels 1..lc (the coarse ones, where a W-cycle spends most of its sweeps)
// live in LDS and are worked by wave 0 alone: each grid point carries one
// 80-byte record (c1..c8, c9, rhs), every one of the 16 Gauss-Seidel phases of
// a sweep is a host-built list of at most 64 points (one per lane), and the
// phases are separated by a wave-level wait on LDS instead of a block barrier.
// The fine levels live in a per-solve global workspace and are worked by the
// whole block.  The relaxation keeps MUDPACK's point order exactly: four
// colours, each swept as four row phases (j = 1..4 mod 4) whose points are
// independent, so iterates match the reference solver's iterate for iterate,
// and so do the cycle counts.
//
// Stencil coefficients are shared by all solves; only c9 carries the
// wavenumber, through ce = -kx^2 J: c9 = ((ce - s9) + corr9) reproduces
// MUDPACK's c9 = ce - (c1+c3+c5+c7) plus its corner correction exactly.
#include <cstring>
#include <vector>

#include "ekm_host.h"
#include "ekm_internal.h"

namespace {

constexpr int kMaxLev = 16;
constexpr int kRec = 10;  // LDS record: c1..c8, c9, rhs

enum MgOp : int { OP_RELAX = 0, OP_RESTRICT = 1, OP_CORRECT = 2, OP_FMGPROL = 3 };

struct LevDev {
  int nx, ny, in_lds;
  int phi_off;           // phi(0:nx+1,0:ny+1), in LDS or the solve's workspace
  int rhs_off, rhs_st;   // rhs(p) at base + rhs_off + p*rhs_st
  int c9_off, c9_st;     // c9(p)
  int rec_off;           // LDS levels: [p][kRec] records
  int tab_off;           // LDS levels: phase table (ints) at this LDS double offset
  int ph_start[17];      // LDS levels: phase k lists tab[ph_start[k] .. ph_start[k+1])
  const double* coef;    // c1..c8 interleaved [p][8] (global)
  const double* s9;
  const double* corr9;
  const double* jl;
};

struct MgDev {
  int ngrid, nxa, nxb, nyc, nyd, maxcy, nfx, nfy;
  double tolmax;
  int ws_global;
  int resf_g, phcor_g;   // global scratch offsets
  int resf_l, phcor_l;   // LDS scratch offsets
  int red_l;             // LDS reduction scratch offset
  int tab_lds;           // LDS offset of all phase tables; tab_n ints
  int tab_n;
  int lc;
  int nops_fmg, nops_w;
  LevDev lev[kMaxLev + 1];
  const int* tab;        // phase tables (global copy)
  const int* ops;        // FMG program, then one W cycle: (op << 8) | level
};

// t = q*d + r for 0 <= t < 2^22 through a float reciprocal and one
// correction step (no integer division).
struct FDiv {
  float rcp;
  int d;
};
__device__ inline FDiv fdiv(int d) { return FDiv{1.0f / (float)d, d}; }
__device__ inline void divmod(int t, const FDiv& f, int& q, int& r) {
  q = (int)((float)t * f.rcp);
  r = t - q * f.d;
  if (r < 0) {
    --q;
    r += f.d;
  } else if (r >= f.d) {
    ++q;
    r -= f.d;
  }
}

// Executor of a primitive: the whole block (global levels) or wave 0 alone
// (LDS levels).  Wave 0 separates dependent steps with a workgroup-scope
// fence, which waits for its own LDS traffic.
struct Exec {
  int tid, nt;
  bool wave;
  __device__ void sync() const {
    if (wave) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
      __syncthreads();
    }
  }
};

struct Ctx {
  const MgDev& g;
  double* W;  // per-solve global workspace
  double* S;  // LDS base
  __device__ double* base(int l) const { return g.lev[l].in_lds ? S : W; }
  __device__ double* phi(int l) const { return base(l) + g.lev[l].phi_off; }
  __device__ double* rhs(int l) const { return base(l) + g.lev[l].rhs_off; }
  __device__ int rhs_st(int l) const { return g.lev[l].rhs_st; }
  __device__ double* c9(int l) const { return base(l) + g.lev[l].c9_off; }
  __device__ int c9_st(int l) const { return g.lev[l].c9_st; }
  __device__ double* resf(int l) const { return g.lev[l].in_lds ? S + g.resf_l : W + g.resf_g; }
  __device__ double* phcor(int l) const { return g.lev[l].in_lds ? S + g.phcor_l : W + g.phcor_g; }
  __device__ Exec ex(int l) const {
    bool w = g.lev[l].in_lds != 0;
    return Exec{w ? (int)(threadIdx.x & 63) : (int)threadIdx.x, w ? 64 : (int)blockDim.x, w};
  }
  __device__ bool active(const Exec& e) const { return !e.wave || threadIdx.x < 64; }
};

// c1..c8 of point p of level l: LDS levels from the point's record, global
// levels from the interleaved table (one 64-byte record).
struct Coef {
  double c[8];
};
__device__ inline Coef load_coef(const Ctx& C, int l, int p) {
  const LevDev& L = C.g.lev[l];
  Coef r;
  if (L.in_lds) {
    const double* rc = C.S + L.rec_off + kRec * p;
#pragma unroll
    for (int q = 0; q < 8; ++q) r.c[q] = rc[q];
  } else {
    const double2* cf = reinterpret_cast<const double2*>(L.coef + 8 * (size_t)p);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double2 v = cf[q];
      r.c[2 * q] = v.x;
      r.c[2 * q + 1] = v.y;
    }
  }
  return r;
}

// MUDPACK's point update (relmd2crp, src/mud2cr.f:1693-1795), same operation
// order: phi = (rhs - (c1 phi(i+1,j) + c2 phi(i+1,j+1) + ... + c8 phi(i+1,j-1))) / c9
__device__ inline double gs_point(const Coef& c, const double* phi, int q, int ld, double rhs, double c9) {
  return (rhs - (c.c[0] * phi[q + 1] + c.c[1] * phi[q + 1 + ld] + c.c[2] * phi[q + ld] + c.c[3] * phi[q - 1 + ld] +
                 c.c[4] * phi[q - 1] + c.c[5] * phi[q - 1 - ld] + c.c[6] * phi[q - ld] + c.c[7] * phi[q + 1 - ld])) /
         c9;
}

// periodic virtual points after each colour
__device__ inline void periodic_fix(const MgDev& g, const Exec& e, double* phi, int nx, int ny) {
  const int ld = nx + 2;
  int nt = (g.nxa == 0 ? ny : 0) + (g.nyc == 0 ? nx : 0);
  for (int t = e.tid; t < nt; t += e.nt) {
    if (g.nxa == 0 && t < ny) {
      int j = t + 1;
      phi[0 + ld * j] = phi[(nx - 1) + ld * j];
      phi[(nx + 1) + ld * j] = phi[2 + ld * j];
    } else {
      int i = t - (g.nxa == 0 ? ny : 0) + 1;
      phi[i] = phi[i + ld * (ny - 1)];
      phi[i + ld * (ny + 1)] = phi[i + ld * 2];
    }
  }
  e.sync();
}

// One relaxation sweep of an LDS level by wave 0: 16 phases, each a list of
// at most 64 points read from the phase table.
__device__ void relax_wave(const Ctx& C, int l) {
  const MgDev& g = C.g;
  const LevDev& L = g.lev[l];
  const Exec e{(int)threadIdx.x, 64, true};
  const int lane = threadIdx.x;
  double* phi = C.S + L.phi_off;
  const double* rec = C.S + L.rec_off;
  const int* tab = reinterpret_cast<const int*>(C.S + g.tab_lds);
  const int ld = L.nx + 2;
  const bool per = g.nxa == 0 || g.nyc == 0;
  int ent[16];
#pragma unroll
  for (int ph = 0; ph < 16; ++ph) {
    const int a = L.ph_start[ph], n = L.ph_start[ph + 1] - a;
    ent[ph] = lane < n ? tab[L.tab_off + a + lane] : -1;
  }
#pragma unroll
  for (int ph = 0; ph < 16; ++ph) {
    const int en = ent[ph];
    if (en >= 0) {
      const int q = en & 0xffff, p = en >> 16;
      const double* r = rec + kRec * p;
      Coef c;
#pragma unroll
      for (int k = 0; k < 8; ++k) c.c[k] = r[k];
      phi[q] = gs_point(c, phi, q, ld, r[9], r[8]);
    }
    e.sync();
    if (per && (ph & 3) == 3) periodic_fix(g, e, phi, L.nx, L.ny);
  }
}

// One relaxation sweep of a global level by the whole block; each thread
// takes two points of a phase at a time and issues all their loads before
// either store (points of one phase never neighbour each other).
__device__ void relax_block(const Ctx& C, int l) {
  const MgDev& g = C.g;
  const LevDev& L = g.lev[l];
  const Exec e{(int)threadIdx.x, (int)blockDim.x, false};
  __syncthreads();
  double* phi = C.W + L.phi_off;
  const double* rhs = C.W + L.rhs_off;
  const double* c9 = C.W + L.c9_off;
  const int nx = L.nx, ny = L.ny, ld = nx + 2;
  int ist[4] = {1, 4, 3, 2};
  for (int color = 0; color < 4; ++color) {
    for (int ph = 0; ph < 4; ++ph) {
      const int jr = 1 + ph, i0 = ist[ph];
      const int nj = jr <= ny ? (ny - jr) / 4 + 1 : 0, ni = i0 <= nx ? (nx - i0) / 4 + 1 : 0;
      const int np = ni * nj;
      const FDiv fd = fdiv(ni > 0 ? ni : 1);
      for (int t = e.tid; t < np; t += 2 * e.nt) {
        const int t2 = t + e.nt;
        const bool two = t2 < np;
        int a, b, a2, b2;
        divmod(t, fd, b, a);
        divmod(two ? t2 : t, fd, b2, a2);
        const int i = i0 + 4 * a, j = jr + 4 * b, i2 = i0 + 4 * a2, j2 = jr + 4 * b2;
        const int p = (i - 1) + nx * (j - 1), q = i + ld * j;
        const int p2 = (i2 - 1) + nx * (j2 - 1), q2 = i2 + ld * j2;
        const Coef c = load_coef(C, l, p);
        const Coef c2 = load_coef(C, l, p2);
        const double r1 = rhs[p], r2 = rhs[p2], d1 = c9[p], d2 = c9[p2];
        const double v1 = gs_point(c, phi, q, ld, r1, d1);
        const double v2 = gs_point(c2, phi, q2, ld, r2, d2);
        phi[q] = v1;
        if (two) phi[q2] = v2;
      }
      e.sync();
    }
    if (g.nxa == 0 || g.nyc == 0) periodic_fix(g, e, phi, nx, ny);
    int tt = ist[3];
    ist[3] = ist[2];
    ist[2] = ist[1];
    ist[1] = ist[0];
    ist[0] = tt;
  }
}

// resmd2cr + res2 (src/mud2cr.f:1618-1665, src/mudcom.f:120-258): fine level
// l -> coarse level l-1, on the executor of the fine level.
__device__ void restrict_(const Ctx& C, int l) {
  const Exec e = C.ex(l);
  if (!e.wave) __syncthreads();
  if (!C.active(e)) return;
  const MgDev& g = C.g;
  const LevDev &F = g.lev[l], &Cl = g.lev[l - 1];
  const int nx = F.nx, ny = F.ny, ncx = Cl.nx, ncy = Cl.ny;
  const double* phi = C.phi(l);
  const double* rhs = C.rhs(l);
  const double* c9 = C.c9(l);
  const int rst = C.rhs_st(l), cst = C.c9_st(l);
  double* resf = C.resf(l);
  double* phic = C.phi(l - 1);
  double* rhsc = C.rhs(l - 1);
  const int rstc = C.rhs_st(l - 1);
  const int ld = nx + 2;
  for (int t = e.tid; t < (ncx + 2) * (ncy + 2); t += e.nt) phic[t] = 0.0;
  const FDiv fx = fdiv(nx);
  for (int t = e.tid; t < nx * ny; t += e.nt) {
    int jj, ii;
    divmod(t, fx, jj, ii);
    const int q = (ii + 1) + ld * (jj + 1);
    const Coef c = load_coef(C, l, t);
    resf[t] = rhs[t * rst] - (c.c[0] * phi[q + 1] + c.c[1] * phi[q + 1 + ld] + c.c[2] * phi[q + ld] +
                              c.c[3] * phi[q - 1 + ld] + c.c[4] * phi[q - 1] + c.c[5] * phi[q - 1 - ld] +
                              c.c[6] * phi[q - ld] + c.c[7] * phi[q + 1 - ld] + c9[t * cst] * phi[q]);
  }
  e.sync();
  const int ix = (ncx == nx) ? 0 : 1, jy = (ncy == ny) ? 0 : 1;
  const FDiv fcx = fdiv(ncx);
  for (int t = e.tid; t < ncx * ncy; t += e.nt) {
    int jc0, ic0;
    divmod(t, fcx, jc0, ic0);
    const int ic = ic0 + 1, jc = jc0 + 1;
    const int i = ic + ix * (ic - 1), j = jc + jy * (jc - 1);
    int im1 = i - 1, ip1 = i + 1, jm1 = j - 1, jp1 = j + 1;
    if (jc == 1 || jc == ncy) {
      jm1 = j - 1 > 2 ? j - 1 : 2;
      jp1 = j + 1 < ny - 1 ? j + 1 : ny - 1;
      if (j == 1 && g.nyc == 0) jm1 = ny - 1;
      if (j == ny && g.nyc == 0) jp1 = 2;
    }
    if (ic == 1 || ic == ncx) {
      im1 = i - 1 > 2 ? i - 1 : 2;
      ip1 = i + 1 < nx - 1 ? i + 1 : nx - 1;
      if (i == 1 && g.nxa == 0) im1 = nx - 1;
      if (i == nx && g.nxa == 0) ip1 = 2;
    }
#define RF(a, b) resf[((a)-1) + nx * ((b)-1)]
    double v = (RF(im1, jm1) + RF(ip1, jm1) + RF(im1, jp1) + RF(ip1, jp1) +
                2. * (RF(im1, j) + RF(ip1, j) + RF(i, jm1) + RF(i, jp1)) + 4. * RF(i, j)) * .0625;
#undef RF
    if ((g.nxa == 1 && ic == 1) || (g.nxb == 1 && ic == ncx) || (g.nyc == 1 && jc == 1) || (g.nyd == 1 && jc == ncy))
      v = 0.0;
    rhsc[t * rstc] = v;
  }
  e.sync();
}

// prolon2 / prolon1 (src/mudcom.f:261-500): p(0:ncx+1,0:ncy+1) -> q(0:nx+1,0:ny+1)
__device__ void prolong(const MgDev& g, const Exec& e, int ncx, int ncy, const double* p, int nx, int ny, double* q) {
  const int nxa = g.nxa, nxb = g.nxb, nyc = g.nyc, nyd = g.nyd;
  const int lq = nx + 2, lp = ncx + 2;
  int ist = 1, ifn = nx, jst = 1, jfn = ny, joddst = 1, joddfn = ny;
  if (nxa == 1) ist = 2;
  if (nxb == 1) ifn = nx - 1;
  if (nyc == 1) {
    jst = 2;
    joddst = 3;
  }
  if (nyd == 1) {
    jfn = ny - 1;
    joddfn = ny - 2;
  }
  const int ioddst = nxa == 1 ? 3 : 1, ioddfn = nxb == 1 ? nx - 2 : nx;
  const bool ysub = ncy < ny;
  const int r0 = ysub ? joddst : jst, r1 = ysub ? joddfn : jfn, rs = ysub ? 2 : 1;
  const int nrow = r1 >= r0 ? (r1 - r0) / rs + 1 : 0;
  const bool xcub = ncx >= 4;  // intpol = 3
  if (ncx < nx) {
    const int ni = ioddfn >= ioddst ? (ioddfn - ioddst) / 2 + 1 : 0;
    const FDiv f = fdiv(ni > 0 ? ni : 1);
    for (int t = e.tid; t < nrow * ni; t += e.nt) {
      int a, b;
      divmod(t, f, b, a);
      const int j = r0 + rs * b, i = ioddst + 2 * a;
      const int jc = ysub ? j / 2 + 1 : j;
      q[i + lq * j] = p[(i + 1) / 2 + lp * jc];
    }
  } else {
    const int ni = ifn - ist + 1;
    const FDiv f = fdiv(ni);
    for (int t = e.tid; t < nrow * ni; t += e.nt) {
      int a, b;
      divmod(t, f, b, a);
      const int j = r0 + rs * b, i = ist + a;
      const int jc = ysub ? j / 2 + 1 : j;
      q[i + lq * j] = p[i + lp * jc];
    }
  }
  e.sync();
  if (ncx < nx) {
    if (xcub) {
      const int nint = (nx - 3 >= 4) ? (nx - 3 - 4) / 2 + 1 : 0;
      const int per = nint + 2;
      const FDiv f = fdiv(per);
      for (int t = e.tid; t < nrow * per; t += e.nt) {
        int k, b;
        divmod(t, f, b, k);
        const int j = r0 + rs * b;
        double* qr = q + lq * j;
        if (k < nint) {
          const int i = 4 + 2 * k;
          qr[i] = (-qr[i - 3] + 9. * (qr[i - 1] + qr[i + 1]) - qr[i + 3]) * .0625;
        } else if (k == nint) {
          qr[2] = (nxa != 0) ? (5. * qr[1] + 15. * qr[3] - 5. * qr[5] + qr[7]) * .0625
                             : (-qr[nx - 2] + 9. * (qr[1] + qr[3]) - qr[5]) * .0625;
        } else {
          qr[nx - 1] = (nxa != 0) ? (5. * qr[nx] + 15. * qr[nx - 2] - 5. * qr[nx - 4] + qr[nx - 6]) * .0625
                                  : (-qr[nx - 4] + 9. * (qr[nx - 2] + qr[nx]) - qr[3]) * .0625;
        }
      }
    } else {
      const int ne = ifn >= 2 ? (ifn - 2) / 2 + 1 : 0;
      const FDiv f = fdiv(ne > 0 ? ne : 1);
      for (int t = e.tid; t < nrow * ne; t += e.nt) {
        int a, b;
        divmod(t, f, b, a);
        const int j = r0 + rs * b, i = 2 + 2 * a;
        q[i + lq * j] = 0.5 * (q[i - 1 + lq * j] + q[i + 1 + lq * j]);
      }
    }
    e.sync();
  }
  if (nxa == 0) {
    for (int t = e.tid; t < nrow; t += e.nt) {
      const int j = r0 + rs * t;
      q[0 + lq * j] = q[nx - 1 + lq * j];
      q[nx + 1 + lq * j] = q[2 + lq * j];
    }
    e.sync();
  }
  const int ni = ifn - ist + 1;
  const FDiv fi = fdiv(ni);
  if (ysub) {
    if (ncy >= 4) {
      const int nint = (ny - 3 >= 4) ? (ny - 3 - 4) / 2 + 1 : 0;
      for (int t = e.tid; t < nint * ni; t += e.nt) {
        int a, b;
        divmod(t, fi, b, a);
        const int j = 4 + 2 * b, i = ist + a;
        q[i + lq * j] = (-q[i + lq * (j - 3)] + 9. * (q[i + lq * (j - 1)] + q[i + lq * (j + 1)]) - q[i + lq * (j + 3)]) * .0625;
      }
      for (int t = e.tid; t < ni; t += e.nt) {
        const int i = ist + t;
#define QQ(a) q[i + lq * (a)]
        if (nyc != 0) {
          QQ(2) = (5. * QQ(1) + 15. * QQ(3) - 5. * QQ(5) + QQ(7)) * .0625;
          QQ(ny - 1) = (5. * QQ(ny) + 15. * QQ(ny - 2) - 5. * QQ(ny - 4) + QQ(ny - 6)) * .0625;
        } else {
          QQ(2) = (-QQ(ny - 2) + 9. * (QQ(1) + QQ(3)) - QQ(5)) * .0625;
          QQ(ny - 1) = (-QQ(ny - 4) + 9. * (QQ(ny - 2) + QQ(ny)) - QQ(3)) * .0625;
          QQ(ny + 1) = QQ(2);
          QQ(0) = QQ(ny - 1);
        }
#undef QQ
      }
      e.sync();
    } else {
      const int ne = jfn >= 2 ? (jfn - 2) / 2 + 1 : 0;
      for (int t = e.tid; t < ne * ni; t += e.nt) {
        int a, b;
        divmod(t, fi, b, a);
        const int j = 2 + 2 * b, i = ist + a;
        q[i + lq * j] = 0.5 * (q[i + lq * (j - 1)] + q[i + lq * (j + 1)]);
      }
      e.sync();
      if (nyc == 0) {
        for (int t = e.tid; t < ni; t += e.nt) {
          const int i = ist + t;
          q[i] = q[i + lq * (ny - 1)];
          q[i + lq * (ny + 1)] = q[i + lq * 2];
        }
        e.sync();
      }
    }
  } else if (nyc == 0) {
    for (int t = e.tid; t < ni; t += e.nt) {
      const int i = ist + t;
      q[i] = q[i + lq * (ny - 1)];
      q[i + lq * (ny + 1)] = q[i + lq * 2];
    }
    e.sync();
  }
}

// cor2 (src/mudcom.f:502-553) when `add`: phi(l+1) += prolongated phi(l);
// otherwise the full-multigrid start's prolongation phi(l+1) = P phi(l)
// (src/mud2cr.f:281-290).  Runs on the executor of the fine level l+1.
__device__ void correct_or_prolong(const Ctx& C, int l, bool add) {
  const Exec e = C.ex(l + 1);
  if (!e.wave) __syncthreads();
  if (!C.active(e)) return;
  const MgDev& g = C.g;
  const LevDev &Cl = g.lev[l], &F = g.lev[l + 1];
  const int nx = F.nx, ny = F.ny;
  double* phif = C.phi(l + 1);
  double* dst = add ? C.phcor(l + 1) : phif;
  const int ld = nx + 2;
  if (add) {
    for (int t = e.tid; t < (nx + 2) * (ny + 2); t += e.nt) dst[t] = 0.0;
    e.sync();
  }
  prolong(g, e, Cl.nx, Cl.ny, C.phi(l), nx, ny, dst);
  if (!add) return;
  int ist = 1, ifn = nx, jst = 1, jfn = ny;
  if (g.nxa == 1) ist = 2;
  if (g.nxb == 1) ifn = nx - 1;
  if (g.nyc == 1) jst = 2;
  if (g.nyd == 1) jfn = ny - 1;
  const int ni = ifn - ist + 1, nj = jfn - jst + 1;
  const FDiv f = fdiv(ni);
  for (int t = e.tid; t < ni * nj; t += e.nt) {
    int a, b;
    divmod(t, f, b, a);
    const int q = (ist + a) + ld * (jst + b);
    phif[q] = phif[q] + dst[q];
  }
  e.sync();
  if (g.nyc == 0) {
    for (int t = e.tid; t < ni; t += e.nt) {
      const int i = ist + t;
      phif[i] = phif[i + ld * (ny - 1)];
      phif[i + ld * (ny + 1)] = phif[i + ld * 2];
    }
    e.sync();
  }
  if (g.nxa == 0) {
    for (int t = e.tid; t < nj; t += e.nt) {
      const int j = jst + t;
      phif[ld * j] = phif[nx - 1 + ld * j];
      phif[nx + 1 + ld * j] = phif[2 + ld * j];
    }
    e.sync();
  }
}

// Runs the host-expanded program ops[0..n): every thread of the block walks
// it; each primitive is carried out by the executor of its level.
__device__ void run_program(const Ctx& C, const int* ops, int n) {
  for (int k = 0; k < n; ++k) {
    const int op = ops[k] >> 8, l = ops[k] & 0xff;
    if (op == OP_RELAX) {
      if (C.g.lev[l].in_lds) {
        if (threadIdx.x < 64) relax_wave(C, l);
      } else {
        relax_block(C, l);
      }
    } else if (op == OP_RESTRICT) {
      restrict_(C, l);
    } else {
      correct_or_prolong(C, l, op == OP_CORRECT);
    }
  }
}

__global__ void __launch_bounds__(1024) k_mg_solve(const MgDev* __restrict__ gp, double* __restrict__ ws,
                                                   const double* __restrict__ rhs_tmp, double* __restrict__ phi_out,
                                                   const double* __restrict__ kx2, int* __restrict__ cycles,
                                                   size_t plane, int iguess) {
  extern __shared__ double lds[];
  const MgDev& g = *gp;
  const int s = blockIdx.x;
  Ctx C{g, ws + (size_t)s * g.ws_global, lds};
  const double k2 = kx2[s];
  const int nfx = g.nfx, nfy = g.nfy, ng = g.ngrid;
  const double* rt = rhs_tmp + (size_t)s * plane;
  double* pout = phi_out + (size_t)s * plane;
  // per-solve c9 = ((ce - s9) + corr9) with ce = -kx2 J; LDS levels also get
  // their c1..c8 records and the phase tables
  for (int l = 1; l <= ng; ++l) {
    const LevDev& L = g.lev[l];
    const int n = L.nx * L.ny;
    double* c9 = C.c9(l);
    const int st = L.c9_st;
    for (int t = threadIdx.x; t < n; t += blockDim.x) c9[t * st] = ((-(k2 * L.jl[t])) - L.s9[t]) + L.corr9[t];
    if (L.in_lds) {
      double* rec = lds + L.rec_off;
      for (int t = threadIdx.x; t < 8 * n; t += blockDim.x) {
        const int p = t >> 3, q = t & 7;
        rec[kRec * p + q] = L.coef[t];
      }
    }
  }
  {
    int* tab = reinterpret_cast<int*>(lds + g.tab_lds);
    for (int t = threadIdx.x; t < g.tab_n; t += blockDim.x) tab[t] = g.tab[t];
  }
  // swk2 + phi_bc: phi = initial guess (0 unless iguess), rhs = -RHS_TMP
  {
    double* phi = C.phi(ng);
    double* rhs = C.rhs(ng);
    const int rst = C.rhs_st(ng);
    for (int t = threadIdx.x; t < (nfx + 2) * (nfy + 2); t += blockDim.x) {
      const int j = t / (nfx + 2), i = t - j * (nfx + 2);
      const bool inner = i >= 1 && i <= nfx && j >= 1 && j <= nfy;
      phi[t] = (inner && iguess) ? pout[(i - 1) + nfx * (j - 1)] : 0.0;
    }
    for (int t = threadIdx.x; t < nfx * nfy; t += blockDim.x) rhs[t * rst] = -rt[t];
  }
  __syncthreads();
  // adjmd2cr: specified boundaries take phi
  auto adjust = [&](int l) {
    const LevDev& L = g.lev[l];
    const double* phi = C.phi(l);
    double* rhs = C.rhs(l);
    const int rst = L.rhs_st;
    const int nx = L.nx, ny = L.ny, ld = nx + 2;
    for (int t = threadIdx.x; t < nx * ny; t += blockDim.x) {
      const int j = 1 + t / nx, i = 1 + t - (j - 1) * nx;
      if ((g.nxa == 1 && i == 1) || (g.nxb == 1 && i == nx) || (g.nyc == 1 && j == 1) || (g.nyd == 1 && j == ny))
        rhs[t * rst] = phi[i + ld * j];
    }
    __syncthreads();
  };
  if (!iguess) {
    // trsfc2 down to every level (injection of phi and rhs)
    for (int l = ng - 1; l >= 1; --l) {
      const LevDev &F = g.lev[l + 1], &Cl = g.lev[l];
      const double* phi = C.phi(l + 1);
      const double* rhs = C.rhs(l + 1);
      const int rst = F.rhs_st, rstc = Cl.rhs_st;
      double* phic = C.phi(l);
      double* rhsc = C.rhs(l);
      const int nx = F.nx, ny = F.ny, ncx = Cl.nx, ncy = Cl.ny;
      for (int t = threadIdx.x; t < (ncx + 2) * (ncy + 2); t += blockDim.x) {
        const int jc = t / (ncx + 2), ic = t - jc * (ncx + 2);
        double v = 0.0;
        if (ic >= 1 && ic <= ncx && jc >= 1 && jc <= ncy) {
          const int i = (ncx < nx) ? ic + ic - 1 : ic, j = (ncy < ny) ? jc + jc - 1 : jc;
          v = phi[i + (nx + 2) * j];
          rhsc[((ic - 1) + ncx * (jc - 1)) * rstc] = rhs[((i - 1) + nx * (j - 1)) * rst];
        }
        phic[t] = v;
      }
      __syncthreads();
    }
    for (int l = 1; l <= ng; ++l) adjust(l);
  } else {
    adjust(ng);
  }
  // Program 0 is the full multigrid start (kcycle(k), then prolongation to
  // k+1, k = 1..ng-1; skipped with an initial guess); programs 1..maxcy are
  // the W cycles from the finest level with the relmax test (mud2cr1
  // :238-337).  One call site keeps one copy of the primitives.
  double* red = lds + g.red_l;
  double* phi = C.phi(ng);
  const int ld = nfx + 2;
  const int NT = blockDim.x, TID = threadIdx.x;
  int itero = 0;
  bool done = false;
  for (int iter = iguess ? 1 : 0; iter <= g.maxcy && !done; ++iter) {
    __syncthreads();
    run_program(C, iter == 0 ? g.ops : g.ops + g.nops_fmg, iter == 0 ? g.nops_fmg : g.nops_w);
    __syncthreads();
    if (iter == 0) {
      // the caller's phi is the relmax reference of the first cycle: zero
      for (int t = TID; t < nfx * nfy; t += NT) pout[t] = 0.0;
      continue;
    }
    itero = iter;
    if (g.tolmax > 0.0) {
      double rel = 0.0, pm = 0.0;
      for (int t = TID; t < nfx * nfy; t += NT) {
        const int j = 1 + t / nfx, i = 1 + t - (j - 1) * nfx;
        const double w = phi[i + ld * j];
        pm = fmax(pm, fabs(w));
        rel = fmax(rel, fabs(w - pout[t]));
        pout[t] = w;
      }
      red[TID] = rel;
      red[NT + TID] = pm;
      __syncthreads();
      for (int w = NT / 2; w > 0; w >>= 1) {
        if (TID < w) {
          red[TID] = fmax(red[TID], red[TID + w]);
          red[NT + TID] = fmax(red[NT + TID], red[NT + TID + w]);
        }
        __syncthreads();
      }
      double relmax = red[0], phmax = red[NT];
      if (phmax > 0.0) relmax = relmax / phmax;
      if (relmax <= g.tolmax) done = true;
    }
  }
  __syncthreads();
  if (!done) {
    for (int t = TID; t < nfx * nfy; t += NT) {
      const int j = 1 + t / nfx, i = 1 + t - (j - 1) * nfx;
      pout[t] = phi[i + ld * j];
    }
  }
  if (TID == 0) cycles[s] = itero;
}

// kcymd2cr's order of operations for kcycle(kcur) (src/mud2cr.f:340-499):
// kcyc = 2 (W), iprer = 2, ipost = 1.
void kcycle_program(int kcur, std::vector<int>& ops) {
  const int kcyc = 2, iprer = 2, ipost = 1;
  auto emit = [&](int op, int l) { ops.push_back((op << 8) | l); };
  int kount[kMaxLev + 1] = {0};
  for (int r = 0; r < iprer; ++r) emit(OP_RELAX, kcur);
  if (kcur != 1) {
    emit(OP_RESTRICT, kcur);
    int klevel = kcur - 1, nrel = iprer;
    while (klevel != kcur) {
      kount[klevel]++;
      for (int r = 0; r < nrel; ++r) emit(OP_RELAX, klevel);
      if (kount[klevel] == kcyc + 1) {
        emit(OP_CORRECT, klevel);
        kount[klevel] = 0;
        klevel++;
        nrel = ipost;
      } else if (klevel > 1) {
        emit(OP_RESTRICT, klevel);
        klevel--;
        nrel = iprer;
      } else {
        for (int r = 0; r < ipost; ++r) emit(OP_RELAX, klevel);
        emit(OP_CORRECT, 1);
        nrel = ipost;
        klevel = 2;
      }
    }
  }
  for (int r = 0; r < ipost; ++r) emit(OP_RELAX, kcur);
}

}  // namespace

// Host side of the plan: level sizes, LDS split, coefficient tables, phase
// tables and the cycle programs.
int mg_plan_build(MgPlan& mg, const ekm_params& p, const Dims& d, const HostMetrics& hm, int nsolve_max,
                  int block, size_t lds_budget_doubles) {
  mg.nxa = p.mudpack_bc[0];
  mg.nxb = p.mudpack_bc[1];
  mg.nyc = p.mudpack_bc[2];
  mg.nyd = p.mudpack_bc[3];
  mg.maxcy = p.mg_maxcy;
  mg.tolmax = p.mg_tolmax;
  mg.ngrid = p.mg_iex > p.mg_jey ? p.mg_iex : p.mg_jey;
  if (mg.ngrid > kMaxLev || mg.ngrid < 1) return EKM_ERR_UNSUPPORTED;
  mg.nsolve_max = nsolve_max;
  mg.block = block;
  std::vector<HostMgLevel> hl(mg.ngrid + 1);
  size_t coef_total = 0;
  for (int k = 1; k <= mg.ngrid; ++k) {
    int ex = k + p.mg_iex - mg.ngrid, ey = k + p.mg_jey - mg.ngrid;
    int nx = p.mg_ixp * (1 << ((ex > 1 ? ex : 1) - 1)) + 1;
    int ny = p.mg_jyq * (1 << ((ey > 1 ? ey : 1) - 1)) + 1;
    int st = host_mg_level(hm, d.NZ, d.NY, p.mudpack_bc, nx, ny, hl[k]);
    if (st) return st;
    mg.lev[k].nx = nx;
    mg.lev[k].ny = ny;
    coef_total += ((size_t)nx * ny * 11 + 7) & ~(size_t)7;  // each level 64-byte aligned
  }
  // phase tables of the candidate LDS levels: for colour c and row phase r
  // the points i = ist_c[r] (mod 4 steps), j = r+1 (mod 4 steps), packed as
  // (phi index) | (point index << 16)
  auto phase_table = [&](int l, std::vector<int>& tab, int* start) {
    const int nx = mg.lev[l].nx, ny = mg.lev[l].ny, ld = nx + 2;
    int ist[4] = {1, 4, 3, 2};
    int base = (int)tab.size();
    for (int c = 0; c < 4; ++c) {
      for (int r = 0; r < 4; ++r) {
        start[4 * c + r] = (int)tab.size() - base;
        for (int j = 1 + r; j <= ny; j += 4)
          for (int i = ist[r]; i <= nx; i += 4) tab.push_back((i + ld * j) | (((i - 1) + nx * (j - 1)) << 16));
      }
      int tt = ist[3];
      ist[3] = ist[2];
      ist[2] = ist[1];
      ist[1] = ist[0];
      ist[0] = tt;
    }
    start[16] = (int)tab.size() - base;
    for (int k = 0; k < 16; ++k)
      if (start[k + 1] - start[k] > 64) return false;
    return (nx + 2) * (ny + 2) < 65536 && nx * ny < 32768;
  };
  // finest level lc whose levels 1..lc (phi, records, phase tables) plus the
  // restriction/correction scratch of lc and the reduction buffer fit the
  // LDS budget, and whose phases hold at most 64 points each
  const size_t red = (size_t)2 * block;
  auto lev_lds = [&](int l) {
    size_t n = (size_t)mg.lev[l].nx * mg.lev[l].ny;
    return (size_t)(mg.lev[l].nx + 2) * (mg.lev[l].ny + 2) + kRec * n + (n + 1) / 2 + 8;
  };
  mg.lc = 0;
  {
    size_t tot = red;
    for (int k = 1; k <= mg.ngrid - 1; ++k) {
      std::vector<int> tmp;
      int st[17];
      if (!phase_table(k, tmp, st)) break;
      tot += lev_lds(k);
      size_t need = tot + (size_t)mg.lev[k].nx * mg.lev[k].ny + (size_t)(mg.lev[k].nx + 2) * (mg.lev[k].ny + 2);
      if (need > lds_budget_doubles) break;
      mg.lc = k;
    }
  }
  MgDev g;
  std::memset(&g, 0, sizeof(g));
  std::vector<int> tab;
  size_t go = 0, lo = 0;
  for (int k = 1; k <= mg.ngrid; ++k) {
    MgLevel& L = mg.lev[k];
    LevDev& D = g.lev[k];
    L.in_lds = k <= mg.lc;
    D.nx = L.nx;
    D.ny = L.ny;
    D.in_lds = L.in_lds;
    const size_t nphi = (size_t)(L.nx + 2) * (L.ny + 2), n = (size_t)L.nx * L.ny;
    size_t& o = L.in_lds ? lo : go;
    D.phi_off = (int)o;
    o += nphi;
    if (L.in_lds) {
      o = (o + 1) & ~(size_t)1;
      D.rec_off = (int)o;
      D.c9_off = (int)o + 8;
      D.rhs_off = (int)o + 9;
      D.c9_st = D.rhs_st = kRec;
      o += kRec * n;
      D.tab_off = (int)tab.size();
      phase_table(k, tab, D.ph_start);
    } else {
      D.rhs_off = (int)o;
      D.rhs_st = 1;
      o += n;
      D.c9_off = (int)o;
      D.c9_st = 1;
      o += n;
    }
    o = (o + 7) & ~(size_t)7;
  }
  const int nfx = d.NZ2, nfy = d.NY2;
  g.resf_g = (int)go;
  go += (size_t)nfx * nfy;
  g.phcor_g = (int)go;
  go += (size_t)(nfx + 2) * (nfy + 2);
  go = (go + 31) & ~(size_t)31;
  mg.ws_global = go;
  g.resf_l = (int)lo;
  if (mg.lc > 0) {
    lo += (size_t)mg.lev[mg.lc].nx * mg.lev[mg.lc].ny;
    g.phcor_l = (int)lo;
    lo += (size_t)(mg.lev[mg.lc].nx + 2) * (mg.lev[mg.lc].ny + 2);
  }
  g.tab_lds = (int)lo;
  g.tab_n = (int)tab.size();
  lo += (tab.size() + 1) / 2;
  lo = (lo + 7) & ~(size_t)7;
  g.red_l = (int)lo;
  lo += red;
  mg.ws_lds = lo;
  if (lo > lds_budget_doubles) return EKM_ERR_UNSUPPORTED;
  // cycle programs: full multigrid start, then one W cycle at the finest level
  std::vector<int> ops;
  for (int k = 1; k <= mg.ngrid - 1; ++k) {
    kcycle_program(k, ops);
    ops.push_back((OP_FMGPROL << 8) | k);
  }
  mg.nops_fmg = (int)ops.size();
  kcycle_program(mg.ngrid, ops);
  mg.nops_w = (int)ops.size() - mg.nops_fmg;
  // coefficient tables: c1..c8 interleaved per point ([p][8], one 64-byte
  // record), then s9, corr9 and the Jacobian of each level
  EKM_CHECK(hipMalloc(&mg.coefmem, coef_total * sizeof(double)));
  std::vector<double> host(coef_total);
  size_t off = 0;
  for (int k = 1; k <= mg.ngrid; ++k) {
    size_t n = (size_t)mg.lev[k].nx * mg.lev[k].ny;
    for (size_t q = 0; q < n; ++q)
      for (int c = 0; c < 8; ++c) host[off + 8 * q + c] = hl[k].c[c * n + q];
    g.lev[k].coef = mg.coefmem + off;
    off += 8 * n;
    std::memcpy(&host[off], hl[k].s9.data(), n * sizeof(double));
    g.lev[k].s9 = mg.coefmem + off;
    off += n;
    std::memcpy(&host[off], hl[k].corr9.data(), n * sizeof(double));
    g.lev[k].corr9 = mg.coefmem + off;
    off += n;
    std::memcpy(&host[off], hl[k].jl.data(), n * sizeof(double));
    g.lev[k].jl = mg.coefmem + off;
    off += n;
    off = (off + 7) & ~(size_t)7;
  }
  EKM_CHECK(hipMemcpy(mg.coefmem, host.data(), coef_total * sizeof(double), hipMemcpyHostToDevice));
  if (mg.ws_lds * sizeof(double) > 65536)
    EKM_CHECK(hipFuncSetAttribute((const void*)k_mg_solve, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(mg.ws_lds * sizeof(double))));
  // descriptor, phase tables and programs in one device block
  g.ngrid = mg.ngrid;
  g.nxa = mg.nxa;
  g.nxb = mg.nxb;
  g.nyc = mg.nyc;
  g.nyd = mg.nyd;
  g.maxcy = mg.maxcy;
  g.nfx = nfx;
  g.nfy = nfy;
  g.tolmax = mg.tolmax;
  g.ws_global = (int)mg.ws_global;
  g.lc = mg.lc;
  g.nops_fmg = mg.nops_fmg;
  g.nops_w = mg.nops_w;
  const size_t desc_bytes = (sizeof(MgDev) + 255) & ~(size_t)255;
  const size_t tab_bytes = ((tab.size() + 1) * sizeof(int) + 255) & ~(size_t)255;
  const size_t ops_bytes = ops.size() * sizeof(int);
  char* dmem = nullptr;
  EKM_CHECK(hipMalloc(&dmem, desc_bytes + tab_bytes + ops_bytes));
  mg.desc = dmem;
  g.tab = reinterpret_cast<const int*>(dmem + desc_bytes);
  g.ops = reinterpret_cast<const int*>(dmem + desc_bytes + tab_bytes);
  std::vector<char> hbuf(desc_bytes + tab_bytes + ops_bytes, 0);
  std::memcpy(hbuf.data(), &g, sizeof(MgDev));
  if (!tab.empty()) std::memcpy(hbuf.data() + desc_bytes, tab.data(), tab.size() * sizeof(int));
  std::memcpy(hbuf.data() + desc_bytes + tab_bytes, ops.data(), ops_bytes);
  EKM_CHECK(hipMemcpy(dmem, hbuf.data(), hbuf.size(), hipMemcpyHostToDevice));
  EKM_CHECK(hipMalloc(&mg.ws, (size_t)nsolve_max * mg.ws_global * sizeof(double)));
  EKM_CHECK(hipMalloc(&mg.cycles, (size_t)2 * nsolve_max * sizeof(int)));
  EKM_CHECK(hipMemset(mg.cycles, 0, (size_t)2 * nsolve_max * sizeof(int)));
  EKM_CHECK(hipMalloc(&mg.kx2, (size_t)nsolve_max * sizeof(double)));
  EKM_CHECK(hipMalloc(&mg.rhs, (size_t)nsolve_max * d.plane * sizeof(double)));
  EKM_CHECK(hipMalloc(&mg.phi, (size_t)nsolve_max * d.plane * sizeof(double)));
  EKM_CHECK(hipMemset(mg.phi, 0, (size_t)nsolve_max * d.plane * sizeof(double)));
  return EKM_OK;
}

void mg_plan_free(MgPlan& mg) {
  double* ptrs[] = {mg.ws, mg.coefmem, mg.kx2, mg.rhs, mg.phi};
  for (double* q : ptrs)
    if (q) (void)hipFree(q);
  if (mg.cycles) (void)hipFree(mg.cycles);
  if (mg.desc) (void)hipFree(mg.desc);
  mg.desc = nullptr;
  mg.ws = mg.coefmem = mg.kx2 = mg.rhs = mg.phi = nullptr;
  mg.cycles = nullptr;
}

int mg_solve_ex(MgPlan& mg, const Dims& d, int nsolve, int iguess, hipStream_t st) {
  if (nsolve <= 0) return EKM_OK;
  if (nsolve > mg.nsolve_max) return EKM_ERR_ARG;
  size_t lds = mg.ws_lds * sizeof(double);
  hipLaunchKernelGGL(k_mg_solve, dim3(nsolve), dim3(mg.block), lds, st, (const MgDev*)mg.desc, mg.ws,
                     (const double*)mg.rhs, mg.phi, (const double*)mg.kx2, mg.cycles, d.plane, iguess);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
