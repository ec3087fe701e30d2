// ekm_mg.hip -- batched MUDPACK-replica multigrid for the per-kx pressure
// problems (MULTIGRID_MUDPACK, src/mg_solver2d.F90:75-201, solver mud2cr,
// src/mud2cr.f / src/mudcom.f) on gfx950.
//
// Every real solve of a projection (2 per live kx mode) runs in ONE launch:
// one workgroup owns one solve for its whole life -- full multigrid start,
// up to maxcy W(2,1) cycles and the relmax stopping test -- so the
// reference's control flow (kcymd2cr's kount state machine) runs uniformly
// inside the block and no cycle needs a host round trip.  The coarse levels,
// where a W-cycle spends most of its sweeps, live in LDS; the fine levels
// live in a per-solve global workspace.  The relaxation keeps MUDPACK's
// point order exactly: four colours, each swept as four row phases
// (j = 1..4 mod 4), points of one phase are independent, phases are
// separated by a block barrier.  Iterates therefore match the reference
// solver's iterate for iterate, and so do the cycle counts.
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

struct LevDev {
  int nx, ny, in_lds, pad;
  size_t phi_off, rhs_off, c9_off, cf_off;  // per-solve offsets (LDS or global workspace)
  const double* coef;                       // global c1..c8 interleaved [p][8]
  const double* s9;
  const double* corr9;
  const double* jl;
};

struct MgDev {
  int ngrid, nxa, nxb, nyc, nyd, maxcy, nfx, nfy;
  double tolmax;
  size_t ws_global;
  size_t resf_g, phcor_g;  // global scratch offsets
  size_t resf_l, phcor_l;  // LDS scratch offsets
  size_t red_l;            // LDS reduction scratch offset
  int lc;
  LevDev lev[kMaxLev + 1];
};

// Who executes a primitive: the whole block (fine levels, data in the global
// workspace) or wave 0 alone (levels <= lc, data in LDS).  Wave 0 separates
// its dependent phases with a workgroup-scope fence, which waits for its own
// LDS traffic; a whole-block primitive starts and ends with a barrier.
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
  double kx2;
  __device__ double* base(int l) const { return g.lev[l].in_lds ? S : W; }
  __device__ double* phi(int l) const { return base(l) + g.lev[l].phi_off; }
  __device__ double* rhs(int l) const { return base(l) + g.lev[l].rhs_off; }
  __device__ double* c9(int l) const { return base(l) + g.lev[l].c9_off; }
  __device__ double* resf(int l) const { return g.lev[l].in_lds ? S + g.resf_l : W + g.resf_g; }
  __device__ double* phcor(int l) const { return g.lev[l].in_lds ? S + g.phcor_l : W + g.phcor_g; }
  // executor of a primitive whose data live on level l
  __device__ Exec ex(int l) const {
    bool w = g.lev[l].in_lds != 0;
    return Exec{w ? (int)(threadIdx.x & 63) : (int)threadIdx.x, w ? 64 : (int)blockDim.x, w};
  }
  __device__ bool active(const Exec& e) const { return !e.wave || threadIdx.x < 64; }
};

// c1..c8 of point p on level l: coarse levels keep eight planes in LDS,
// fine levels read the interleaved global table (one 64-byte record).
struct Coef {
  double c[8];
};
__device__ inline Coef load_coef(const Ctx& C, int l, size_t p, size_t n) {
  const LevDev& L = C.g.lev[l];
  Coef r;
  if (L.in_lds) {
    const double* cf = C.S + L.cf_off;
#pragma unroll
    for (int q = 0; q < 8; ++q) r.c[q] = cf[q * n + p];
  } else {
    const double2* cf = reinterpret_cast<const double2*>(L.coef + 8 * p);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double2 v = cf[q];
      r.c[2 * q] = v.x;
      r.c[2 * q + 1] = v.y;
    }
  }
  return r;
}

// relmd2crp (src/mud2cr.f:1693-1795)
__device__ void relax(const Ctx& C, int l) {
  const Exec e = C.ex(l);
  if (!e.wave) __syncthreads();
  if (C.active(e)) {
    const LevDev& L = C.g.lev[l];
    double* phi = C.phi(l);
    const double* rhs = C.rhs(l);
    const double* c9 = C.c9(l);
    const int nx = L.nx, ny = L.ny;
    const size_t ld = (size_t)nx + 2, n = (size_t)nx * ny;
    int ist[4] = {1, 4, 3, 2};
    for (int color = 0; color < 4; ++color) {
      for (int ph = 0; ph < 4; ++ph) {
        int jr = 1 + ph, i0 = ist[ph];
        int nj = jr <= ny ? (ny - jr) / 4 + 1 : 0, ni = i0 <= nx ? (nx - i0) / 4 + 1 : 0;
        int np = ni * nj;
        if (np == 0) continue;
        for (int t = e.tid; t < np; t += e.nt) {
          int i = i0 + 4 * (t % ni), j = jr + 4 * (t / ni);
          size_t p = (size_t)(i - 1) + (size_t)nx * (j - 1), q = (size_t)i + ld * j;
          Coef c = load_coef(C, l, p, n);
          phi[q] = (rhs[p] - (c.c[0] * phi[q + 1] + c.c[1] * phi[q + 1 + ld] + c.c[2] * phi[q + ld] +
                              c.c[3] * phi[q - 1 + ld] + c.c[4] * phi[q - 1] + c.c[5] * phi[q - 1 - ld] +
                              c.c[6] * phi[q - ld] + c.c[7] * phi[q + 1 - ld])) /
                   c9[p];
        }
        e.sync();
      }
      if (C.g.nxa == 0 || C.g.nyc == 0) {
        int nt = (C.g.nxa == 0 ? ny : 0) + (C.g.nyc == 0 ? nx : 0);
        for (int t = e.tid; t < nt; t += e.nt) {
          if (C.g.nxa == 0 && t < ny) {
            int j = t + 1;
            phi[0 + ld * j] = phi[(nx - 1) + ld * j];
            phi[(nx + 1) + ld * j] = phi[2 + ld * j];
          } else {
            int i = t - (C.g.nxa == 0 ? ny : 0) + 1;
            phi[i] = phi[i + ld * (ny - 1)];
            phi[i + ld * (ny + 1)] = phi[i + ld * 2];
          }
        }
        e.sync();
      }
      int tt = ist[3];
      ist[3] = ist[2];
      ist[2] = ist[1];
      ist[1] = ist[0];
      ist[0] = tt;
    }
  }
}

// resmd2cr + res2 (src/mud2cr.f:1618-1665, src/mudcom.f:120-258): fine level
// l -> coarse level l-1.  Runs on the executor of the fine level.
__device__ void restrict_(const Ctx& C, int l) {
  const Exec e = C.ex(l);
  if (!e.wave) __syncthreads();
  if (C.active(e)) {
    const MgDev& g = C.g;
    const LevDev &F = g.lev[l], &Cl = g.lev[l - 1];
    const int nx = F.nx, ny = F.ny, ncx = Cl.nx, ncy = Cl.ny;
    const double* phi = C.phi(l);
    const double* rhs = C.rhs(l);
    const double* c9 = C.c9(l);
    double* resf = C.resf(l);
    double* phic = C.phi(l - 1);
    double* rhsc = C.rhs(l - 1);
    const size_t ld = (size_t)nx + 2, n = (size_t)nx * ny;
    for (int t = e.tid; t < (ncx + 2) * (ncy + 2); t += e.nt) phic[t] = 0.0;
    for (int t = e.tid; t < nx * ny; t += e.nt) {
      int i = 1 + t % nx, j = 1 + t / nx;
      size_t q = (size_t)i + ld * j;
      Coef c = load_coef(C, l, t, n);
      resf[t] = rhs[t] - (c.c[0] * phi[q + 1] + c.c[1] * phi[q + 1 + ld] + c.c[2] * phi[q + ld] +
                          c.c[3] * phi[q - 1 + ld] + c.c[4] * phi[q - 1] + c.c[5] * phi[q - 1 - ld] +
                          c.c[6] * phi[q - ld] + c.c[7] * phi[q + 1 - ld] + c9[t] * phi[q]);
    }
    e.sync();
    const int ix = (ncx == nx) ? 0 : 1, jy = (ncy == ny) ? 0 : 1;
    for (int t = e.tid; t < ncx * ncy; t += e.nt) {
      int ic = 1 + t % ncx, jc = 1 + t / ncx;
      int i = ic + ix * (ic - 1), j = jc + jy * (jc - 1);
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
#define RF(a, b) resf[((a)-1) + (size_t)nx * ((b)-1)]
      double v = (RF(im1, jm1) + RF(ip1, jm1) + RF(im1, jp1) + RF(ip1, jp1) +
                  2. * (RF(im1, j) + RF(ip1, j) + RF(i, jm1) + RF(i, jp1)) + 4. * RF(i, j)) * .0625;
#undef RF
      if ((g.nxa == 1 && ic == 1) || (g.nxb == 1 && ic == ncx) || (g.nyc == 1 && jc == 1) || (g.nyd == 1 && jc == ncy))
        v = 0.0;
      rhsc[t] = v;
    }
    e.sync();
  }
}

// prolon2 / prolon1 (src/mudcom.f:261-500): p(0:ncx+1,0:ncy+1) -> q(0:nx+1,0:ny+1)
__device__ void prolong(const Ctx& C, const Exec& e, int ncx, int ncy, const double* p, int nx, int ny, double* q) {
  const MgDev& g = C.g;
  const int nxa = g.nxa, nxb = g.nxb, nyc = g.nyc, nyd = g.nyd;
  const size_t lq = (size_t)nx + 2, lp = (size_t)ncx + 2;
  int ist = 1, ifn = nx, jst = 1, jfn = ny, joddst = 1, joddfn = ny;
  if (nxa == 1) ist = 2;
  if (nxb == 1) ifn = nx - 1;
  if (nyc == 1) { jst = 2; joddst = 3; }
  if (nyd == 1) { jfn = ny - 1; joddfn = ny - 2; }
  int ioddst = nxa == 1 ? 3 : 1, ioddfn = nxb == 1 ? nx - 2 : nx;
  const bool ysub = ncy < ny;
  const int r0 = ysub ? joddst : jst, r1 = ysub ? joddfn : jfn, rs = ysub ? 2 : 1;
  const int nrow = r1 >= r0 ? (r1 - r0) / rs + 1 : 0;
  const bool xcub = ncx >= 4;  // intpol = 3
  if (ncx < nx) {
    int ni = ioddfn >= ioddst ? (ioddfn - ioddst) / 2 + 1 : 0;
    for (int t = e.tid; t < nrow * ni; t += e.nt) {
      int j = r0 + rs * (t / ni), i = ioddst + 2 * (t % ni);
      int jc = ysub ? j / 2 + 1 : j;
      q[i + lq * j] = p[(i + 1) / 2 + lp * jc];
    }
  } else {
    int ni = ifn - ist + 1;
    for (int t = e.tid; t < nrow * ni; t += e.nt) {
      int j = r0 + rs * (t / ni), i = ist + t % ni;
      int jc = ysub ? j / 2 + 1 : j;
      q[i + lq * j] = p[i + lp * jc];
    }
  }
  e.sync();
  if (ncx < nx) {
    if (xcub) {
      int nint = (nx - 3 >= 4) ? (nx - 3 - 4) / 2 + 1 : 0;
      int per = nint + 2;
      for (int t = e.tid; t < nrow * per; t += e.nt) {
        int j = r0 + rs * (t / per), k = t % per;
        double* qr = q + lq * j;
        if (k < nint) {
          int i = 4 + 2 * k;
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
      int ne = ifn >= 2 ? (ifn - 2) / 2 + 1 : 0;
      for (int t = e.tid; t < nrow * ne; t += e.nt) {
        int j = r0 + rs * (t / ne), i = 2 + 2 * (t % ne);
        q[i + lq * j] = 0.5 * (q[i - 1 + lq * j] + q[i + 1 + lq * j]);
      }
    }
    e.sync();
  }
  if (nxa == 0) {
    for (int t = e.tid; t < nrow; t += e.nt) {
      int j = r0 + rs * t;
      q[0 + lq * j] = q[nx - 1 + lq * j];
      q[nx + 1 + lq * j] = q[2 + lq * j];
    }
    e.sync();
  }
  const int ni = ifn - ist + 1;
  if (ysub) {
    if (ncy >= 4) {
      int nint = (ny - 3 >= 4) ? (ny - 3 - 4) / 2 + 1 : 0;
      for (int t = e.tid; t < nint * ni; t += e.nt) {
        int j = 4 + 2 * (t / ni), i = ist + t % ni;
        q[i + lq * j] = (-q[i + lq * (j - 3)] + 9. * (q[i + lq * (j - 1)] + q[i + lq * (j + 1)]) - q[i + lq * (j + 3)]) * .0625;
      }
      for (int t = e.tid; t < ni; t += e.nt) {
        int i = ist + t;
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
      int ne = jfn >= 2 ? (jfn - 2) / 2 + 1 : 0;
      for (int t = e.tid; t < ne * ni; t += e.nt) {
        int j = 2 + 2 * (t / ni), i = ist + t % ni;
        q[i + lq * j] = 0.5 * (q[i + lq * (j - 1)] + q[i + lq * (j + 1)]);
      }
      e.sync();
      if (nyc == 0) {
        for (int t = e.tid; t < ni; t += e.nt) {
          int i = ist + t;
          q[i] = q[i + lq * (ny - 1)];
          q[i + lq * (ny + 1)] = q[i + lq * 2];
        }
        e.sync();
      }
    }
  } else if (nyc == 0) {
    for (int t = e.tid; t < ni; t += e.nt) {
      int i = ist + t;
      q[i] = q[i + lq * (ny - 1)];
      q[i + lq * (ny + 1)] = q[i + lq * 2];
    }
    e.sync();
  }
}

// cor2 (src/mudcom.f:502-553): phi(l+1) += prolongated phi(l); runs on the
// executor of the fine level l+1.
__device__ void correct(const Ctx& C, int l) {
  const Exec e = C.ex(l + 1);
  if (!e.wave) __syncthreads();
  if (!C.active(e)) return;
  const MgDev& g = C.g;
  const LevDev &Cl = g.lev[l], &F = g.lev[l + 1];
  const int nx = F.nx, ny = F.ny;
  double* pc = C.phcor(l + 1);
  double* phif = C.phi(l + 1);
  const size_t ld = (size_t)nx + 2;
  for (int t = e.tid; t < (nx + 2) * (ny + 2); t += e.nt) pc[t] = 0.0;
  e.sync();
  prolong(C, e, Cl.nx, Cl.ny, C.phi(l), nx, ny, pc);
  int ist = 1, ifn = nx, jst = 1, jfn = ny;
  if (g.nxa == 1) ist = 2;
  if (g.nxb == 1) ifn = nx - 1;
  if (g.nyc == 1) jst = 2;
  if (g.nyd == 1) jfn = ny - 1;
  int ni = ifn - ist + 1, nj = jfn - jst + 1;
  for (int t = e.tid; t < ni * nj; t += e.nt) {
    size_t q = (size_t)(ist + t % ni) + ld * (jst + t / ni);
    phif[q] = phif[q] + pc[q];
  }
  e.sync();
  if (g.nyc == 0) {
    for (int t = e.tid; t < ni; t += e.nt) {
      int i = ist + t;
      phif[i] = phif[i + ld * (ny - 1)];
      phif[i + ld * (ny + 1)] = phif[i + ld * 2];
    }
    e.sync();
  }
  if (g.nxa == 0) {
    for (int t = e.tid; t < nj; t += e.nt) {
      int j = jst + t;
      phif[ld * j] = phif[nx - 1 + ld * j];
      phif[nx + 1 + ld * j] = phif[2 + ld * j];
    }
    e.sync();
  }
}

// kcymd2cr (src/mud2cr.f:340-499).  The control flow does not depend on the
// data, so every thread of the block runs it; each primitive is carried out
// by the executor of its level.
__device__ void kcycle(const Ctx& C, int kcur) {
  const int kcyc = 2, iprer = 2, ipost = 1;
  int kount[kMaxLev + 1];
  for (int l = 0; l <= kMaxLev; ++l) kount[l] = 0;
  for (int l = 0; l < iprer; ++l) relax(C, kcur);
  if (kcur != 1) {
    restrict_(C, kcur);
    int klevel = kcur - 1, nrel = iprer;
    while (klevel != kcur) {
      kount[klevel]++;
      for (int l = 0; l < nrel; ++l) relax(C, klevel);
      if (kount[klevel] == kcyc + 1) {
        correct(C, klevel);
        kount[klevel] = 0;
        klevel++;
        nrel = ipost;
      } else if (klevel > 1) {
        restrict_(C, klevel);
        klevel--;
        nrel = iprer;
      } else {
        for (int l = 0; l < ipost; ++l) relax(C, klevel);
        correct(C, 1);
        nrel = ipost;
        klevel = 2;
      }
    }
  }
  for (int l = 0; l < ipost; ++l) relax(C, kcur);
}

__global__ void __launch_bounds__(1024) k_mg_solve(const MgDev* __restrict__ gp, double* __restrict__ ws,
                                                   const double* __restrict__ rhs_tmp, double* __restrict__ phi_out,
                                                   const double* __restrict__ kx2, int* __restrict__ cycles,
                                                   size_t plane, int iguess) {
  extern __shared__ double lds[];
  const MgDev& g = *gp;
  const int s = blockIdx.x;
  Ctx C{g, ws + (size_t)s * g.ws_global, lds, kx2[s]};
  const int nfx = g.nfx, nfy = g.nfy, ng = g.ngrid;
  const double* rt = rhs_tmp + (size_t)s * plane;
  double* pout = phi_out + (size_t)s * plane;
  // per-solve c9 = ((ce - s9) + corr9) with ce = -kx2 J, and the coarse-level
  // coefficient planes copied into LDS
  for (int l = 1; l <= ng; ++l) {
    const LevDev& L = g.lev[l];
    const size_t n = (size_t)L.nx * L.ny;
    double* c9 = C.c9(l);
    for (size_t t = threadIdx.x; t < n; t += blockDim.x) c9[t] = ((-(C.kx2 * L.jl[t])) - L.s9[t]) + L.corr9[t];
    if (L.in_lds) {
      double* cf = lds + L.cf_off;
      for (size_t t = threadIdx.x; t < 8 * n; t += blockDim.x) {
        size_t p = t % n, q = t / n;
        cf[t] = L.coef[8 * p + q];
      }
    }
  }
  // swk2 + phi_bc: phi = initial guess (0 unless iguess), rhs = -RHS_TMP
  {
    double* phi = C.phi(ng);
    double* rhs = C.rhs(ng);
    for (int t = threadIdx.x; t < (nfx + 2) * (nfy + 2); t += blockDim.x) {
      int i = t % (nfx + 2), j = t / (nfx + 2);
      bool inner = i >= 1 && i <= nfx && j >= 1 && j <= nfy;
      phi[t] = (inner && iguess) ? pout[(i - 1) + (size_t)nfx * (j - 1)] : 0.0;
    }
    for (int t = threadIdx.x; t < nfx * nfy; t += blockDim.x) rhs[t] = -rt[t];
  }
  __syncthreads();
  // adjmd2cr: specified boundaries take phi (whole block, then a barrier)
  auto adjust = [&](int l) {
    const LevDev& L = g.lev[l];
    double* phi = C.phi(l);
    double* rhs = C.rhs(l);
    const int nx = L.nx, ny = L.ny;
    const size_t ld = (size_t)nx + 2;
    for (int t = threadIdx.x; t < nx * ny; t += blockDim.x) {
      int i = 1 + t % nx, j = 1 + t / nx;
      if ((g.nxa == 1 && i == 1) || (g.nxb == 1 && i == nx) || (g.nyc == 1 && j == 1) || (g.nyd == 1 && j == ny))
        rhs[t] = phi[(size_t)i + ld * j];
    }
    __syncthreads();
  };
  if (!iguess) {
    // trsfc2 down to every level (injection of phi and rhs)
    for (int l = ng - 1; l >= 1; --l) {
      const LevDev &F = g.lev[l + 1], &Cl = g.lev[l];
      const double* phi = C.phi(l + 1);
      const double* rhs = C.rhs(l + 1);
      double* phic = C.phi(l);
      double* rhsc = C.rhs(l);
      const int nx = F.nx, ny = F.ny, ncx = Cl.nx, ncy = Cl.ny;
      for (int t = threadIdx.x; t < (ncx + 2) * (ncy + 2); t += blockDim.x) {
        int ic = t % (ncx + 2), jc = t / (ncx + 2);
        double v = 0.0;
        if (ic >= 1 && ic <= ncx && jc >= 1 && jc <= ncy) {
          int i = (ncx < nx) ? ic + ic - 1 : ic, j = (ncy < ny) ? jc + jc - 1 : jc;
          v = phi[(size_t)i + (size_t)(nx + 2) * j];
          rhsc[(ic - 1) + (size_t)ncx * (jc - 1)] = rhs[(i - 1) + (size_t)nx * (j - 1)];
        }
        phic[t] = v;
      }
      __syncthreads();
    }
    for (int l = 1; l <= ng; ++l) adjust(l);
    // full multigrid start
    for (int k = 1; k <= ng - 1; ++k) {
      kcycle(C, k);
      const Exec e = C.ex(k + 1);
      __syncthreads();
      if (C.active(e)) prolong(C, e, g.lev[k].nx, g.lev[k].ny, C.phi(k), g.lev[k + 1].nx, g.lev[k + 1].ny, C.phi(k + 1));
    }
  } else {
    adjust(ng);
  }
  __syncthreads();
  // W cycles from the finest level with the relmax test (mud2cr1 :296-337)
  double* red = lds + g.red_l;
  double* phi = C.phi(ng);
  const size_t ld = (size_t)nfx + 2;
  const int NT = blockDim.x, TID = threadIdx.x;
  int itero = 0;
  bool done = false;
  if (!iguess)
    for (int t = TID; t < nfx * nfy; t += NT) pout[t] = 0.0;
  __syncthreads();
  for (int iter = 1; iter <= g.maxcy && !done; ++iter) {
    itero = iter;
    kcycle(C, ng);
    __syncthreads();
    if (g.tolmax > 0.0) {
      double rel = 0.0, pm = 0.0;
      for (int t = TID; t < nfx * nfy; t += NT) {
        int i = 1 + t % nfx, j = 1 + t / nfx;
        double w = phi[(size_t)i + ld * j];
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
      __syncthreads();
      if (relmax <= g.tolmax) done = true;
    }
  }
  if (!done) {
    for (int t = TID; t < nfx * nfy; t += NT) {
      int i = 1 + t % nfx, j = 1 + t / nfx;
      pout[t] = phi[(size_t)i + ld * j];
    }
  }
  if (TID == 0) cycles[s] = itero;
}

}  // namespace

// Host side of the plan: level sizes, LDS split, coefficient tables.
int mg_plan_build(MgPlan& mg, const ekm_params& p, const Dims& d, const HostMetrics& hm, int nsolve_max,
                  int block, size_t lds_budget_doubles) {
  mg.nxa = p.mudpack_bc[0];
  mg.nxb = p.mudpack_bc[1];
  mg.nyc = p.mudpack_bc[2];
  mg.nyd = p.mudpack_bc[3];
  mg.maxcy = p.mg_maxcy;
  mg.tolmax = p.mg_tolmax;
  mg.ngrid = p.mg_iex > p.mg_jey ? p.mg_iex : p.mg_jey;
  if (mg.ngrid > kMaxLev) return EKM_ERR_UNSUPPORTED;
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
  // finest level lc such that levels 1..lc (phi, rhs, c9 and the eight
  // coefficient planes each) plus the restriction/correction scratch of lc
  // and the reduction buffer fit in the LDS budget
  const size_t red = (size_t)2 * block;
  auto lev_lds = [&](int l) {
    size_t n = (size_t)mg.lev[l].nx * mg.lev[l].ny;
    return (size_t)(mg.lev[l].nx + 2) * (mg.lev[l].ny + 2) + 10 * n + 8;
  };
  mg.lc = 0;
  for (int k = 1; k <= mg.ngrid - 1; ++k) {
    size_t tot = red;
    for (int l = 1; l <= k; ++l) tot += lev_lds(l);
    tot += (size_t)mg.lev[k].nx * mg.lev[k].ny + (size_t)(mg.lev[k].nx + 2) * (mg.lev[k].ny + 2);
    if (tot <= lds_budget_doubles) mg.lc = k;
  }
  size_t go = 0, lo = 0;
  for (int k = 1; k <= mg.ngrid; ++k) {
    MgLevel& L = mg.lev[k];
    L.in_lds = k <= mg.lc;
    size_t nphi = (size_t)(L.nx + 2) * (L.ny + 2), n = (size_t)L.nx * L.ny;
    size_t& o = L.in_lds ? lo : go;
    L.phi_off = o;
    o += nphi;
    L.rhs_off = o;
    o += n;
    L.c9_off = o;
    o += n;
    L.cf_off = 0;
    if (L.in_lds) {
      L.cf_off = o;
      o += 8 * n;
    }
    o = (o + 7) & ~(size_t)7;
  }
  int nfx = d.NZ2, nfy = d.NY2;
  size_t resf_g = go;
  go += (size_t)nfx * nfy;
  size_t phcor_g = go;
  go += (size_t)(nfx + 2) * (nfy + 2);
  go = (go + 31) & ~(size_t)31;
  mg.ws_global = go;
  size_t resf_l = lo, phcor_l = 0, red_l = 0;
  if (mg.lc > 0) {
    lo += (size_t)mg.lev[mg.lc].nx * mg.lev[mg.lc].ny;
    phcor_l = lo;
    lo += (size_t)(mg.lev[mg.lc].nx + 2) * (mg.lev[mg.lc].ny + 2);
  }
  red_l = lo;
  lo += red;
  mg.ws_lds = lo;
  if (lo > lds_budget_doubles) return EKM_ERR_UNSUPPORTED;
  mg.resf_g = resf_g;
  mg.phcor_g = phcor_g;
  mg.resf_l = resf_l;
  mg.phcor_l = phcor_l;
  mg.red_l = red_l;
  // coefficient tables: c1..c8 interleaved per point ([p][8], one 64-byte
  // record), then s9, corr9 and the Jacobian of each level
  EKM_CHECK(hipMalloc(&mg.coefmem, coef_total * sizeof(double)));
  std::vector<double> host(coef_total);
  size_t off = 0;
  for (int k = 1; k <= mg.ngrid; ++k) {
    size_t n = (size_t)mg.lev[k].nx * mg.lev[k].ny;
    for (size_t p = 0; p < n; ++p)
      for (int q = 0; q < 8; ++q) host[off + 8 * p + q] = hl[k].c[q * n + p];
    mg.lev[k].coef = mg.coefmem + off;
    off += 8 * n;
    std::memcpy(&host[off], hl[k].s9.data(), n * sizeof(double));
    mg.lev[k].s9 = mg.coefmem + off;
    off += n;
    std::memcpy(&host[off], hl[k].corr9.data(), n * sizeof(double));
    mg.lev[k].corr9 = mg.coefmem + off;
    off += n;
    std::memcpy(&host[off], hl[k].jl.data(), n * sizeof(double));
    mg.lev[k].jl = mg.coefmem + off;
    off += n;
    off = (off + 7) & ~(size_t)7;
  }
  EKM_CHECK(hipMemcpy(mg.coefmem, host.data(), coef_total * sizeof(double), hipMemcpyHostToDevice));
  if (mg.ws_lds * sizeof(double) > 65536)
    EKM_CHECK(hipFuncSetAttribute((const void*)k_mg_solve, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(mg.ws_lds * sizeof(double))));
  {
    // the level descriptor lives in device memory: read with uniform
    // (scalar) loads instead of being copied per lane
    MgDev g;
    std::memset(&g, 0, sizeof(g));
    g.ngrid = mg.ngrid;
    g.nxa = mg.nxa; g.nxb = mg.nxb; g.nyc = mg.nyc; g.nyd = mg.nyd;
    g.maxcy = mg.maxcy;
    g.nfx = d.NZ2;
    g.nfy = d.NY2;
    g.tolmax = mg.tolmax;
    g.ws_global = mg.ws_global;
    g.resf_g = mg.resf_g; g.phcor_g = mg.phcor_g;
    g.resf_l = mg.resf_l; g.phcor_l = mg.phcor_l; g.red_l = mg.red_l;
    g.lc = mg.lc;
    for (int k = 1; k <= mg.ngrid; ++k) {
      const MgLevel& L = mg.lev[k];
      g.lev[k] = LevDev{L.nx, L.ny, L.in_lds, 0, L.phi_off, L.rhs_off, L.c9_off, L.cf_off,
                        L.coef, L.s9, L.corr9, L.jl};
    }
    EKM_CHECK(hipMalloc(&mg.desc, sizeof(MgDev)));
    EKM_CHECK(hipMemcpy(mg.desc, &g, sizeof(MgDev), hipMemcpyHostToDevice));
  }
  EKM_CHECK(hipMalloc(&mg.ws, (size_t)nsolve_max * mg.ws_global * sizeof(double)));
  EKM_CHECK(hipMalloc(&mg.cycles, (size_t)nsolve_max * sizeof(int)));
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
