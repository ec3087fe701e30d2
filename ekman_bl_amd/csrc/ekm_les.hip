// ekm_les.hip -- the dynamic Smagorinsky subgrid models of the velocity
// (reference: src/les_chan.F90 les_chan, LES_MODEL_TYPE 2 with
// LES_IMPLICIT = .FALSE.; compute_strain_chan :1623-1717; les_filter_chan
// :1719-1838 through FILTER_VAR, src/mpi_duct.F90:834-852) and of the scalar
// (src/les_chan_th.F90 les_chan_th, LES_MODEL_TYPE_TH 2; compute_scalar_grad
// :462-535).
//
// Every kernel is an HBM-bound FP64 pass over device-resident fields with
// consecutive threads on consecutive x (physical) or kx (Fourier) points, so
// each row is one coalesced run.  The dynamic coefficient, recomputed every
// fifth step, is built from x-filtered fields and one x reduction per (z,y)
// pencil.
#include "ekm_les.h"

namespace {

__device__ inline size_t ph(const Dims& d, int i, int k, int j) {
  return (size_t)i + (size_t)d.NX2 * ((size_t)k + (size_t)d.NZ2 * (size_t)j);
}
__device__ inline bool in(int v, int lo, int hi) { return v >= lo && v <= hi; }

#define C11(k, j, f) M.c11[mf(d, k, j, f)]
#define C12(k, j, f) M.c12[mf(d, k, j, f)]
#define C21(k, j, f) M.c21[mf(d, k, j, f)]
#define C22(k, j, f) M.c22[mf(d, k, j, f)]
#define JAC(k, j) M.jac[mi(d, k, j)]

constexpr double kAlpha = 2.44950;  // test / LES filter width ratio (les_chan.F90:37)

// the face-difference operators of the strain at (k, j) of a field with
// strides dk, dj: D2 with CJOB_12 / CJOB_22, D3 with CJOB_11 / CJOB_21
template <typename V>
__device__ inline V d2(const Dims& d, const Metrics& M, const V* f, size_t q, int k, int j, size_t dk, size_t dj);
template <>
__device__ inline double d2<double>(const Dims& d, const Metrics& M, const double* f, size_t q, int k, int j,
                                    size_t dk, size_t dj) {
  return 0.5 * C12(k + 1, j, 2) * (f[q] + f[q + dk]) - 0.5 * C12(k, j, 2) * (f[q] + f[q - dk]) +
         0.5 * C22(k, j + 1, 1) * (f[q] + f[q + dj]) - 0.5 * C22(k, j, 1) * (f[q] + f[q - dj]);
}
__device__ inline double d3(const Dims& d, const Metrics& M, const double* f, size_t q, int k, int j, size_t dk,
                            size_t dj) {
  return 0.5 * C11(k + 1, j, 2) * (f[q] + f[q + dk]) - 0.5 * C11(k, j, 2) * (f[q] + f[q - dk]) +
         0.5 * C21(k, j + 1, 1) * (f[q] + f[q + dj]) - 0.5 * C21(k, j, 1) * (f[q] + f[q - dj]);
}
// the same on the two parts of a complex field (the coefficients are real)
__device__ inline double2 d2c(const Dims& d, const Metrics& M, const double2* f, size_t q, int k, int j, size_t dk,
                              size_t dj) {
  const double* p = reinterpret_cast<const double*>(f);
  return make_double2(d2<double>(d, M, p, 2 * q, k, j, 2 * dk, 2 * dj),
                      d2<double>(d, M, p, 2 * q + 1, k, j, 2 * dk, 2 * dj));
}
__device__ inline double2 d3c(const Dims& d, const Metrics& M, const double2* f, size_t q, int k, int j, size_t dk,
                              size_t dj) {
  const double* p = reinterpret_cast<const double*>(f);
  return make_double2(d3(d, M, p, 2 * q, k, j, 2 * dk, 2 * dj), d3(d, M, p, 2 * q + 1, k, j, 2 * dk, 2 * dj));
}
// i kx f
__device__ inline double2 ikx(double kx, double2 f) { return make_double2(-kx * f.y, kx * f.x); }

// strain rate in Fourier space at the interior points, zero elsewhere
// (compute_strain_chan :1633-1680)
__global__ void k_les_strain(Dims d, Metrics M, Ranges r, const double* __restrict__ kxv, const double2* CU1,
                             const double2* CU2, const double2* CU3, double2* S0, double2* S1, double2* S2,
                             double2* S3, double2* S4, double2* S5) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)d.NXC * d.NZ2 * d.NY2;
  if (t >= n) return;
  const int m = (int)(t % d.NXC);
  const size_t rr = t / d.NXC;
  const int k = (int)(rr % d.NZ2), j = (int)(rr / d.NZ2);
  const size_t dk = d.NXC, dj = (size_t)d.NXC * d.NZ2;
  double2 s[6];
  for (int c = 0; c < 6; ++c) s[c] = make_double2(0.0, 0.0);
  if (in(j, r.js, r.je) && in(k, r.zs, r.ze)) {
    const double Jc = JAC(k, j), kx = kxv[m];
    const double2 u1 = CU1[t], u2 = CU2[t], u3 = CU3[t];
    s[0] = ikx(kx, u1);
    const double2 a = d2c(d, M, CU2, t, k, j, dk, dj), b = d3c(d, M, CU3, t, k, j, dk, dj);
    s[1] = make_double2(a.x / Jc, a.y / Jc);
    s[2] = make_double2(b.x / Jc, b.y / Jc);
    const double2 c1 = d2c(d, M, CU1, t, k, j, dk, dj), c2 = d3c(d, M, CU1, t, k, j, dk, dj);
    const double2 i2 = ikx(kx, u2), i3 = ikx(kx, u3);
    s[3] = make_double2(0.5 * (c1.x / Jc + i2.x), 0.5 * (c1.y / Jc + i2.y));
    s[4] = make_double2(0.5 * (c2.x / Jc + i3.x), 0.5 * (c2.y / Jc + i3.y));
    const double2 e = d3c(d, M, CU2, t, k, j, dk, dj), g = d2c(d, M, CU3, t, k, j, dk, dj);
    s[5] = make_double2(0.5 * (e.x + g.x) / Jc, 0.5 * (e.y + g.y) / Jc);
  }
  S0[t] = s[0];
  S1[t] = s[1];
  S2[t] = s[2];
  S3[t] = s[3];
  S4[t] = s[4];
  S5[t] = s[5];
}

// |S| at the interior points (:351-364), copied to the y ghost rows (:366-371)
__device__ inline double sabs_at(const Dims& d, const LesFields& L, int i, int k, int j) {
  const size_t q = ph(d, i, k, j);
  const double s1 = L.S[0][q], s2 = L.S[1][q], s3 = L.S[2][q], s4 = L.S[3][q], s5 = L.S[4][q], s6 = L.S[5][q];
  return sqrt(2.0 * s1 * s1 + 4.0 * s4 * s4 + 4.0 * s5 * s5 + 2.0 * s2 * s2 + 4.0 * s6 * s6 + 2.0 * s3 * s3);
}
__global__ void k_les_sabs(Dims d, Ranges r, LesFields L) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.npts) return;
  const int i = (int)(t % d.NX2);
  const size_t rr = t / d.NX2;
  const int k = (int)(rr % d.NZ2), j = (int)(rr / d.NZ2);
  double v = 0.0;
  int jj = j;
  if (j == 0) jj = 1;
  if (j == d.NY + 1) jj = d.NY;
  if (in(jj, r.js, r.je) && in(k, r.zs, r.ze)) v = sabs_at(d, L, i, k, jj);
  L.sabs[t] = v;
}

// x test filter [1/4 1/2 1/4], periodic over i < NX, the padding copied
__global__ void k_les_filter(Dims d, const double* __restrict__ a, double* __restrict__ b) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.npts) return;
  const int i = (int)(t % d.NX2);
  if (i >= d.NX) {
    b[t] = a[t];
    return;
  }
  const size_t p0 = t - i;
  const int im = i == 0 ? d.NX - 1 : i - 1, ip = i == d.NX - 1 ? 0 : i + 1;
  b[t] = ((0.0 * a[p0 + im] + 0.25 * a[p0 + im]) + 0.5 * a[t] + 0.25 * a[p0 + ip]) + 0.0 * a[p0 + ip];
}

// T1 = filter(S_ij) (alpha Delta)^2 filter(|S|) and M0 = Delta^2 |S| S_ij (:549-571)
__global__ void k_les_mprep(Dims d, const double* __restrict__ delta, const double* __restrict__ sij,
                            const double* __restrict__ sabs, const double* __restrict__ s2b, double* __restrict__ t1,
                            double* __restrict__ m0) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.npts) return;
  const size_t pl = t / d.NX2;
  const double D = delta[pl];
  t1[t] = t1[t] * ((kAlpha * D) * (kAlpha * D)) * s2b[t];
  m0[t] = (D * D) * sabs[t] * sij[t];
}

// the product of two velocity components (:585-640)
__global__ void k_les_prod(Dims d, const double* __restrict__ a, const double* __restrict__ b,
                           double* __restrict__ c) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.npts) return;
  c[t] = a[t] * b[t];
}

// per (z,y) pencil: M = T1 - filter(M0); the x sums of M (filter(u_a u_b) -
// ub_a ub_b) and of M^2, weighted by `cross`, added to the running sums
// (:571-796).  One wavefront per pencil.
__global__ void __launch_bounds__(256) k_les_numden(Dims d, const double* __restrict__ t1,
                                                    const double* __restrict__ fm0, const double* __restrict__ ft2,
                                                    const double* __restrict__ uba, const double* __restrict__ ubb,
                                                    double cross, double* __restrict__ nsum,
                                                    double* __restrict__ dsum) {
  const int lane = threadIdx.x & 63;
  const size_t pen = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (pen >= d.plane) return;
  const size_t p0 = pen * d.NX2;
  double sn = 0.0, sd = 0.0;
  for (int i = lane; i < d.NX; i += 64) {
    const size_t q = p0 + i;
    const double m = t1[q] - fm0[q];
    sn += m * (ft2[q] - uba[q] * ubb[q]);
    sd += m * m;
  }
  for (int o = 32; o > 0; o >>= 1) {
    sn += __shfl_xor(sn, o, 64);
    sd += __shfl_xor(sd, o, 64);
  }
  if (lane == 0) {
    dsum[pen] = dsum[pen] + cross * sd;
    nsum[pen] = nsum[pen] + cross * sn;
  }
}

// C_DYN = max(0, -1/2 num / den) on the interior points, zero elsewhere (:798-814)
// The plane sums are cleared here after they are read, so they start at zero
// at the next refresh (they are zero when the solver is created).
__global__ void k_les_cdyn(Dims d, Ranges r, double* __restrict__ nsum, double* __restrict__ dsum,
                           double* __restrict__ cdyn) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.plane) return;
  const int k = (int)(t % d.NZ2), j = (int)(t / d.NZ2);
  double c = 0.0;
  const double ns = nsum[t], ds = dsum[t];
  if (in(j, r.js, r.je) && in(k, r.zs, r.ze)) c = ds != 0.0 ? -0.5 * ns / ds : 0.0;
  cdyn[t] = c < 0.0 ? 0.0 : c;
  nsum[t] = 0.0;
  dsum[t] = 0.0;
}

// eddy viscosity (:825-831) and the subgrid stress with the parts that the
// implicit ADI rows carry removed (:833-960), in place of the strain
__global__ void k_les_stress(Dims d, Metrics M, Ranges r, const double* __restrict__ delta,
                             const double* __restrict__ cdyn, const double* __restrict__ U1,
                             const double* __restrict__ U2, LesFields L) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.npts) return;
  const size_t pl = t / d.NX2;
  const int k = (int)(pl % d.NZ2), j = (int)(pl / d.NZ2);
  const double D = delta[pl], C = cdyn[pl], s = L.sabs[t];
  const bool inner = in(j, r.js, r.je) && in(k, r.zs, r.ze);
  L.nut[t] = inner ? C * (D * D) * s : 0.0;
  const double f = ((-2.0 * C) * (D * D)) * s;
  if (j >= 1 && k >= 1) L.S[0][t] = f * L.S[0][t];
  // components 2 and 3 of the stress are zero (:866-880): they are neither
  // transformed nor read below, so S[1] and S[2] are left as they are
  if (inner) {
    const size_t dk = d.NX2, dj = (size_t)d.NX2 * d.NZ2;
    const double Jc = JAC(k, j);
    L.S[3][t] = f * (L.S[3][t] - (0.5 * d2<double>(d, M, U1, t, k, j, dk, dj)) / Jc);
    L.S[4][t] = f * (L.S[4][t] - (0.5 * d3(d, M, U1, t, k, j, dk, dj)) / Jc);
    L.S[5][t] = f * (L.S[5][t] - (0.5 * d3(d, M, U2, t, k, j, dk, dj)) / Jc);
  }
}

// ghost copies of the stress, y rows first (pass 0: j = NY+1 <- NY,
// j = 0 <- 1 for every k), then z rows (pass 1: k = NZ+1 <- NZ, k = 0 <- 1
// for every j) (:950-970)
__global__ void k_les_ghosts(Dims d, LesFields L, int pass) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nl = pass == 0 ? d.NZ2 : d.NY2;
  if (t >= (size_t)d.NX2 * nl * 4) return;
  const int i = (int)(t % d.NX2);
  const size_t rr = t / d.NX2;
  const int l = (int)(rr % nl), c = (int)(rr / nl);
  double* S = L.S[c == 0 ? 0 : c + 2];  // stress components 1, 4, 5, 6
  if (pass == 0) {
    S[ph(d, i, l, d.NY + 1)] = S[ph(d, i, l, d.NY)];
    S[ph(d, i, l, 0)] = S[ph(d, i, l, 1)];
  } else {
    S[ph(d, i, d.NZ + 1, l)] = S[ph(d, i, d.NZ, l)];
    S[ph(d, i, 0, l)] = S[ph(d, i, 1, l)];
  }
}

// CF -= J i kx CS + D2(CS) + D3(CS) with the stress in Fourier space (the
// components 2 and 3 are zero) (:1313-1356)
__global__ void k_les_cf(Dims d, Metrics M, Ranges r, const double* __restrict__ kxv, const double2* C1,
                         const double2* C4, const double2* C5, const double2* C6, double2* CF1, double2* CF2,
                         double2* CF3) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)d.NXC * d.NZ2 * d.NY2;
  if (t >= n) return;
  const int m = (int)(t % d.NXC);
  const size_t rr = t / d.NXC;
  const int k = (int)(rr % d.NZ2), j = (int)(rr / d.NZ2);
  if (!(in(j, r.js, r.je) && in(k, r.zs, r.ze))) return;
  const size_t dk = d.NXC, dj = (size_t)d.NXC * d.NZ2;
  const double Jc = JAC(k, j), kx = kxv[m];
  const double2 z = make_double2(0.0, 0.0);
  auto upd = [&](double2* CF, const double2* Cx, const double2 a2, const double2 a3) {
    const double2 ix = ikx(kx, Cx[t]);
    double2 f = CF[t];
    f.x = ((f.x - Jc * ix.x) - a2.x) - a3.x;
    f.y = ((f.y - Jc * ix.y) - a2.y) - a3.y;
    CF[t] = f;
  };
  upd(CF1, C1, d2c(d, M, C4, t, k, j, dk, dj), d3c(d, M, C5, t, k, j, dk, dj));
  upd(CF2, C4, z, d3c(d, M, C6, t, k, j, dk, dj));
  upd(CF3, C5, d2c(d, M, C6, t, k, j, dk, dj), z);
}

// ---------------------------------------------------------------------------
// the scalar model (src/les_chan_th.F90, model 2)
// ---------------------------------------------------------------------------
// scalar gradient in Fourier space at the velocity interior, zero elsewhere
// (compute_scalar_grad :480-496): G0 = i kx TH, G1 = D2(TH)/J, G2 = D3(TH)/J.
// G1 and G2 enter only the dynamic coefficient, so they are formed (full)
// only on the substeps that recompute it.
__global__ void k_les_thgrad(Dims d, Metrics M, Ranges r, const double* __restrict__ kxv, const double2* CTH,
                             double2* G0, double2* G1, double2* G2, int full) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)d.NXC * d.NZ2 * d.NY2;
  if (t >= n) return;
  const int m = (int)(t % d.NXC);
  const size_t rr = t / d.NXC;
  const int k = (int)(rr % d.NZ2), j = (int)(rr / d.NZ2);
  const size_t dk = d.NXC, dj = (size_t)d.NXC * d.NZ2;
  double2 g0 = make_double2(0.0, 0.0), g1 = g0, g2 = g0;
  if (in(j, r.js, r.je) && in(k, r.zs, r.ze)) {
    const double Jc = JAC(k, j);
    g0 = ikx(kxv[m], CTH[t]);
    if (full) {
      const double2 a = d2c(d, M, CTH, t, k, j, dk, dj), b = d3c(d, M, CTH, t, k, j, dk, dj);
      g1 = make_double2(a.x / Jc, a.y / Jc);
      g2 = make_double2(b.x / Jc, b.y / Jc);
    }
  }
  G0[t] = g0;
  if (full) {
    G1[t] = g1;
    G2[t] = g2;
  }
}

// the background gradient added to dTH/dy at the velocity interior (:521-531)
__global__ void k_les_thbg(Dims d, Metrics M, Ranges r, const double* __restrict__ thbar, double* __restrict__ G1) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.npts) return;
  const size_t pl = t / d.NX2;
  const int k = (int)(pl % d.NZ2), j = (int)(pl / d.NZ2);
  if (!(in(j, r.js, r.je) && in(k, r.zs, r.ze))) return;
  G1[t] = G1[t] + d2<double>(d, M, thbar, pl, k, j, 1, (size_t)d.NZ2) / JAC(k, j);
}

// KAPPA_T = C_DYN_TH Delta^2 |S| on the scalar's interior, zero elsewhere
// (:326-334), and the x flux -C_DYN_TH Delta^2 |S| dTH/dx in place of G0
// (:341-351; the y and z components the reference also forms are unused)
__global__ void k_les_kappa(Dims d, Ranges rt, const double* __restrict__ delta, const double* __restrict__ cdt,
                            const double* __restrict__ sabs, double* __restrict__ kappa, double* __restrict__ G0) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d.npts) return;
  const size_t pl = t / d.NX2;
  const int k = (int)(pl % d.NZ2), j = (int)(pl / d.NZ2);
  const double D = delta[pl], C = cdt[pl], s = sabs[t];
  kappa[t] = (in(j, rt.js, rt.je) && in(k, rt.zs, rt.ze)) ? C * (D * D) * s : 0.0;
  G0[t] = -(C * (D * D) * s * G0[t]);
}

// CFTH -= J i kx (flux)^ on the scalar's interior (:411-419)
__global__ void k_les_thcf(Dims d, Metrics M, Ranges rt, const double* __restrict__ kxv, const double2* F,
                           double2* CFTH) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)d.NXC * d.NZ2 * d.NY2;
  if (t >= n) return;
  const int m = (int)(t % d.NXC);
  const size_t rr = t / d.NXC;
  const int k = (int)(rr % d.NZ2), j = (int)(rr / d.NZ2);
  if (!(in(j, rt.js, rt.je) && in(k, rt.zs, rt.ze))) return;
  const double a = JAC(k, j) * kxv[m];
  const double2 f = F[t];
  double2 c = CFTH[t];
  c.x = c.x - (0.0 * f.x - a * f.y);
  c.y = c.y - (0.0 * f.y + a * f.x);
  CFTH[t] = c;
}

inline unsigned nblk(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

int les_chan_th_run(const FftPlan& fft, const Dims& d, const Metrics& M, const Ranges& rv, const Ranges& rt,
                    const double* kx, const double* thbar, double* const U[3], double* TH, double* CFTH,
                    const LesFields& L, const LesScratch& w, bool refresh, hipStream_t st) {
  const size_t nf = (size_t)d.NXC * d.NZ2 * d.NY2, np = d.npts;
  const long npen = (long)d.plane;
  double* G[3] = {L.S[0], L.S[1], L.S[2]};
  hipLaunchKernelGGL(k_les_thgrad, dim3(nblk(nf, 256)), dim3(256), 0, st, d, M, rv, kx, (const double2*)TH,
                     (double2*)G[0], (double2*)G[1], (double2*)G[2], refresh ? 1 : 0);
  EKM_CHECK(hipGetLastError());
  for (int c = 0; c < (refresh ? 3 : 1); ++c) {
    const int e = fft_inverse(fft, G[c], npen, st);
    if (e != EKM_OK) return e;
  }
  if (refresh) {
    hipLaunchKernelGGL(k_les_thbg, dim3(nblk(np, 256)), dim3(256), 0, st, d, M, rv, thbar, G[1]);
    EKM_CHECK(hipGetLastError());
  }
  {
    const int e = fft_inverse(fft, TH, npen, st);
    if (e != EKM_OK) return e;
  }
  if (refresh) {
    // test-filtered scalar, then the three contractions with the velocity
    // model's filtered velocities and |S| (:155-321)
    double* thb = L.S[3];
    hipLaunchKernelGGL(k_les_filter, dim3(nblk(np, 256)), dim3(256), 0, st, d, (const double*)TH, thb);
    for (int ij = 0; ij < 3; ++ij) {
      hipLaunchKernelGGL(k_les_filter, dim3(nblk(np, 256)), dim3(256), 0, st, d, (const double*)G[ij], w.t1);
      hipLaunchKernelGGL(k_les_mprep, dim3(nblk(np, 256)), dim3(256), 0, st, d, (const double*)L.delta,
                         (const double*)G[ij], (const double*)L.sabs, (const double*)w.s2b, w.t1, w.m0);
      hipLaunchKernelGGL(k_les_filter, dim3(nblk(np, 256)), dim3(256), 0, st, d, (const double*)w.m0, w.fm0);
      hipLaunchKernelGGL(k_les_prod, dim3(nblk(np, 256)), dim3(256), 0, st, d, (const double*)U[ij],
                         (const double*)TH, w.m0);
      hipLaunchKernelGGL(k_les_filter, dim3(nblk(np, 256)), dim3(256), 0, st, d, (const double*)w.m0, w.ft2);
      hipLaunchKernelGGL(k_les_numden, dim3(nblk(d.plane, 4)), dim3(256), 0, st, d, (const double*)w.t1,
                         (const double*)w.fm0, (const double*)w.ft2, (const double*)thb, (const double*)w.ub[ij],
                         1.0, w.nsum, w.dsum);
      EKM_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_les_cdyn, dim3(nblk(d.plane, 256)), dim3(256), 0, st, d, rv, w.nsum,
                       w.dsum, L.cdyn_th);
    EKM_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_les_kappa, dim3(nblk(np, 256)), dim3(256), 0, st, d, rt, (const double*)L.delta,
                     (const double*)L.cdyn_th, (const double*)L.sabs, L.kappa, G[0]);
  EKM_CHECK(hipGetLastError());
  {
    const int e = fft_forward(fft, G[0], npen, st);
    if (e != EKM_OK) return e;
  }
  hipLaunchKernelGGL(k_les_thcf, dim3(nblk(nf, 256)), dim3(256), 0, st, d, M, rt, kx, (const double2*)G[0],
                     (double2*)CFTH);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}

int les_chan_run(const FftPlan& fft, const Dims& d, const Metrics& M, const Ranges& r, const double* kx,
                 double* const U[3], double* const CF[3], const LesFields& L, const LesScratch& w, bool refresh,
                 hipStream_t st) {
  const size_t nf = (size_t)d.NXC * d.NZ2 * d.NY2, np = d.npts;
  const long npen = (long)d.plane;
  hipLaunchKernelGGL(k_les_strain, dim3(nblk(nf, 256)), dim3(256), 0, st, d, M, r, kx, (const double2*)U[0],
                     (const double2*)U[1], (const double2*)U[2], (double2*)L.S[0], (double2*)L.S[1],
                     (double2*)L.S[2], (double2*)L.S[3], (double2*)L.S[4], (double2*)L.S[5]);
  EKM_CHECK(hipGetLastError());
  for (int c = 0; c < 6; ++c) {
    const int e = fft_inverse(fft, L.S[c], npen, st);
    if (e != EKM_OK) return e;
  }
  hipLaunchKernelGGL(k_les_sabs, dim3(nblk(np, 256)), dim3(256), 0, st, d, r, L);
  EKM_CHECK(hipGetLastError());
  for (int c = 0; c < 3; ++c) {
    const int e = fft_inverse(fft, U[c], npen, st);
    if (e != EKM_OK) return e;
  }
  if (refresh) {
    // test-filtered velocity and |S|, then the six contractions
    for (int c = 0; c < 3; ++c)
      hipLaunchKernelGGL(k_les_filter, dim3(nblk(np, 256)), dim3(256), 0, st, d, (const double*)U[c], w.ub[c]);
    hipLaunchKernelGGL(k_les_filter, dim3(nblk(np, 256)), dim3(256), 0, st, d, (const double*)L.sabs, w.s2b);
    static const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
    for (int ij = 0; ij < 6; ++ij) {
      hipLaunchKernelGGL(k_les_filter, dim3(nblk(np, 256)), dim3(256), 0, st, d, (const double*)L.S[ij], w.t1);
      hipLaunchKernelGGL(k_les_mprep, dim3(nblk(np, 256)), dim3(256), 0, st, d, (const double*)L.delta,
                         (const double*)L.S[ij], (const double*)L.sabs, (const double*)w.s2b, w.t1, w.m0);
      hipLaunchKernelGGL(k_les_filter, dim3(nblk(np, 256)), dim3(256), 0, st, d, (const double*)w.m0, w.fm0);
      hipLaunchKernelGGL(k_les_prod, dim3(nblk(np, 256)), dim3(256), 0, st, d, (const double*)U[ia[ij]],
                         (const double*)U[ib[ij]], w.m0);
      hipLaunchKernelGGL(k_les_filter, dim3(nblk(np, 256)), dim3(256), 0, st, d, (const double*)w.m0, w.ft2);
      hipLaunchKernelGGL(k_les_numden, dim3(nblk(d.plane, 4)), dim3(256), 0, st, d, (const double*)w.t1,
                         (const double*)w.fm0, (const double*)w.ft2, (const double*)w.ub[ia[ij]],
                         (const double*)w.ub[ib[ij]], ij < 3 ? 1.0 : 2.0, w.nsum, w.dsum);
      EKM_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_les_cdyn, dim3(nblk(d.plane, 256)), dim3(256), 0, st, d, r, w.nsum,
                       w.dsum, L.cdyn);
    EKM_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_les_stress, dim3(nblk(np, 256)), dim3(256), 0, st, d, M, r, (const double*)L.delta,
                     (const double*)L.cdyn, (const double*)U[0], (const double*)U[1], L);
  EKM_CHECK(hipGetLastError());
  for (int pass = 0; pass < 2; ++pass) {
    const size_t ng = (size_t)d.NX2 * (pass == 0 ? d.NZ2 : d.NY2) * 4;
    hipLaunchKernelGGL(k_les_ghosts, dim3(nblk(ng, 256)), dim3(256), 0, st, d, L, pass);
    EKM_CHECK(hipGetLastError());
  }
  for (int c : {0, 3, 4, 5}) {
    const int e = fft_forward(fft, L.S[c], npen, st);
    if (e != EKM_OK) return e;
  }
  hipLaunchKernelGGL(k_les_cf, dim3(nblk(nf, 256)), dim3(256), 0, st, d, M, r, kx, (const double2*)L.S[0],
                     (const double2*)L.S[3], (const double2*)L.S[4], (const double2*)L.S[5], (double2*)CF[0],
                     (double2*)CF[1], (double2*)CF[2]);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
