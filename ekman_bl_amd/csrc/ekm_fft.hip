// ekm_fft.hip -- batched real<->complex transforms along x for gfx950.
//
// Replaces FFT_X_TO_FOURIER_OP / FFT_X_TO_PHYSICAL_OP of the reference's
// missing fft.F90 (call sites src/mpi_duct.F90:290,308; conventions fixed by
// the callers, see ekman_bl_amd.h): forward c_m = (1/NX) sum_i u_i e^{-2 pi i m i/NX}
// with modes NKX+1..NX/2 zeroed; inverse from modes 0..NKX (Im c_0 ignored),
// padding entries NX, NX+1 zeroed.
//
// A real pencil of NX doubles is read as M = NX/2 complex values z_k =
// (u_2k, u_2k+1).  One wavefront transforms one pencil (or 64/M pencils
// when M < 64): lane l holds elements l, l+64, l+128, ... in registers, the
// radix-2 decimation-in-frequency stages with partner distance >= 64 pair
// registers of the same lane and the last six stages pair lanes through
// cross-lane shuffles, so no pass goes through LDS and no block barrier is
// needed.  The bit-reversed result is put in order through LDS, where the
// even/odd split of the real transform (forward) pairs Z_k with Z_{M-k}.
// The inverse builds Z_k from the modes in registers and runs the same
// stages on the conjugate.  Loads and stores are contiguous across lanes.
#include "ekm_internal.h"

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <mutex>
#include <vector>

namespace {

constexpr int kWaves = 4;  // wavefronts per block
constexpr int kBlock = 64 * kWaves;

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// DIF stages on M = GW*R elements; element e = lane' + GW*r sits in z[r] of
// lane' (lane' = lane mod GW; GW = 64 whenever R > 1).  tw[t] = exp(-2 pi i
// t / M), t < M/2.  Partner distances h >= GW pair registers of one lane
// (compile-time indices), smaller ones pair lanes.
template <int R>
__device__ inline void dif(double2 (&z)[R], int GW, int M, int lanep, const double2* __restrict__ tw) {
#pragma unroll
  for (int hr = R / 2; hr >= 1; hr >>= 1) {
    const int h = hr * GW, step = M / (2 * h);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if ((r / hr) % 2 == 0) {
        const int e = lanep + GW * r;
        const double2 x = z[r], y = z[r + hr];
        z[r] = cadd(x, y);
        z[r + hr] = cmul(csub(x, y), tw[(e % h) * step]);
      }
    }
  }
  for (int h = GW >> 1; h >= 1; h >>= 1) {
    const int step = M / (2 * h);
    const bool top = (lanep & h) == 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double2 mine = z[r];
      double2 other;
      other.x = __shfl_xor(mine.x, h, 64);
      other.y = __shfl_xor(mine.y, h, 64);
      const int e = lanep + GW * r;
      z[r] = top ? cadd(mine, other) : cmul(csub(other, mine), tw[(e % h) * step]);
    }
  }
}

// Twiddles of the six lane stages when a wavefront holds one transform
// (GW = 64): the stage with partner distance h uses tw[(lane mod h) * M/(2h)]
// whatever the register, so all six are loaded up front, together with the
// data, instead of one dependent load per stage.
struct LaneTw {
  double2 w[6];
};
__device__ inline LaneTw lane_twiddles(int lanep, int M, const double2* __restrict__ tw) {
  LaneTw t;
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int h = 32 >> s;
    t.w[s] = tw[(lanep & (h - 1)) * (M / (2 * h))];
  }
  return t;
}

// The value of lane (lane ^ H) for H = 1, 2: a quad_perm DPP move
// ([1,0,3,2] = 0xB1, [2,3,0,1] = 0x4E) in the vector ALU, where a shuffle
// would go through the LDS permute unit; the source lane is always inside
// the quad, so the bound control never applies.
template <int H>
__device__ inline double quad_xor(double v) {
  constexpr int ctrl = H == 1 ? 0xB1 : 0x4E;
  const long long x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)x, ctrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(x >> 32), ctrl, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// dif<R> for GW = 64 with the lane-stage twiddles given
template <int R>
__device__ inline void dif64(double2 (&z)[R], int M, int lanep, const double2* __restrict__ tw, const LaneTw& lt) {
#pragma unroll
  for (int hr = R / 2; hr >= 1; hr >>= 1) {
    const int h = hr * 64, step = M / (2 * h);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if ((r / hr) % 2 == 0) {
        const int e = lanep + 64 * r;
        const double2 x = z[r], y = z[r + hr];
        z[r] = cadd(x, y);
        z[r + hr] = cmul(csub(x, y), tw[(e % h) * step]);
      }
    }
  }
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int h = 32 >> s;
    const bool top = (lanep & h) == 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double2 mine = z[r];
      double2 other;
      if (h == 2) {
        other.x = quad_xor<2>(mine.x);
        other.y = quad_xor<2>(mine.y);
      } else if (h == 1) {
        other.x = quad_xor<1>(mine.x);
        other.y = quad_xor<1>(mine.y);
      } else {
        other.x = __shfl_xor(mine.x, h, 64);
        other.y = __shfl_xor(mine.y, h, 64);
      }
      z[r] = top ? cadd(mine, other) : cmul(csub(other, mine), lt.w[s]);
    }
  }
}

// the transform of z: dif64 when a wavefront holds one transform, else the
// general form
template <int R>
__device__ inline void dif_any(double2 (&z)[R], int GW, int M, int lanep, const double2* __restrict__ tw,
                               const LaneTw& lt) {
  if (GW == 64)
    dif64<R>(z, M, lanep, tw, lt);
  else
    dif<R>(z, GW, M, lanep, tw);
}

__device__ inline int bitrev(int e, int logm) { return (int)(__brev((unsigned)e) >> (32 - logm)); }

template <int R>
__global__ void __launch_bounds__(kBlock) k_fft_fwd(double* __restrict__ v, long npencil, int nx, int logm,
                                                    const double2* __restrict__ tw,
                                                    const double2* __restrict__ tw2) {
  extern __shared__ double2 lds[];
  const int M = nx >> 1, GW = M < 64 ? M : 64, PW = 64 / GW;  // pencils per wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lanep = lane % GW, sub = lane / GW;
  const long pen = ((long)blockIdx.x * kWaves + wave) * PW + sub;
  const bool live = pen < npencil;
  const int stride = nx + 2, nkx = nx / 3;
  const double2* src = reinterpret_cast<const double2*>(v + (live ? pen : 0) * stride);
  const LaneTw lt = lane_twiddles(lanep, M, tw);
  double2 z[R];
#pragma unroll
  for (int r = 0; r < R; ++r) z[r] = live ? src[lanep + GW * r] : make_double2(0.0, 0.0);
  dif_any<R>(z, GW, M, lanep, tw, lt);
  // into natural order in this wave's LDS slice
  double2* Z = lds + (size_t)(wave * PW + sub) * M;
#pragma unroll
  for (int r = 0; r < R; ++r) Z[bitrev(lanep + GW * r, logm)] = z[r];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  // split: X_k = E_k + W^k O_k, k = 0..M (only k <= NKX kept), scaled by 1/NX
  if (!live) return;
  const double sc = 1.0 / (double)nx;
  double2* dst = reinterpret_cast<double2*>(v + pen * stride);
  double2 w2[R + 1];
#pragma unroll
  for (int i = 0; i <= R; ++i) w2[i] = tw2[min(lanep + GW * i, M)];
#pragma unroll
  for (int i = 0; i <= R; ++i) {
    const int k = lanep + GW * i;
    if (k > M) break;
    double2 x = make_double2(0.0, 0.0);
    if (k <= nkx) {
      const double2 zk = Z[k == M ? 0 : k];
      double2 y = Z[(M - k) % M];
      y.y = -y.y;  // conj(Z_{M-k})
      const double2 e = make_double2(0.5 * (zk.x + y.x), 0.5 * (zk.y + y.y));
      const double2 o = make_double2(0.5 * (zk.y - y.y), -0.5 * (zk.x - y.x));
      const double2 wo = cmul(w2[i], o);
      x = make_double2((e.x + wo.x) * sc, (e.y + wo.y) * sc);
    }
    dst[k] = x;
  }
}

template <int R>
__global__ void __launch_bounds__(kBlock) k_fft_inv(double* __restrict__ v, long npencil, int nx, int logm,
                                                    const double2* __restrict__ tw,
                                                    const double2* __restrict__ tw2) {
  extern __shared__ double2 lds[];
  const int M = nx >> 1, GW = M < 64 ? M : 64, PW = 64 / GW;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lanep = lane % GW, sub = lane / GW;
  const long pen = ((long)blockIdx.x * kWaves + wave) * PW + sub;
  const bool live = pen < npencil;
  const int stride = nx + 2, nkx = nx / 3;
  const double2* X = reinterpret_cast<const double2*>(v + (live ? pen : 0) * stride);
  // Z_k = (X_k + conj X_{M-k}) + i e^{+2 pi i k/NX} (X_k - conj X_{M-k}), k < M,
  // conjugated: the inverse transform is conj(DFT(conj Z))
  const LaneTw lt = lane_twiddles(lanep, M, tw);
  double2 z[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int k = lanep + GW * r, b = M - k;
    double2 x = (live && k <= nkx) ? X[k] : make_double2(0.0, 0.0);
    if (k == 0) x.y = 0.0;
    double2 y = (live && b <= nkx && b < M) ? X[b] : make_double2(0.0, 0.0);
    if (b == 0) y.y = 0.0;
    y.y = -y.y;
    const double2 e = cadd(x, y);
    const double2 dd = csub(x, y);
    double2 w = tw2[k];
    w.y = -w.y;  // exp(+2 pi i k / NX)
    const double2 o = cmul(dd, w);
    z[r] = make_double2(e.x - o.y, -(e.y + o.x));
  }
  dif_any<R>(z, GW, M, lanep, tw, lt);
  double2* Z = lds + (size_t)(wave * PW + sub) * M;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    double2 t = z[r];
    t.y = -t.y;
    Z[bitrev(lanep + GW * r, logm)] = t;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  if (!live) return;
  double2* dst = reinterpret_cast<double2*>(v + pen * stride);
  for (int k = lanep; k <= M; k += GW) dst[k] = (k < M) ? Z[k] : make_double2(0.0, 0.0);
}

// Fused nonlinear-flux stage of one transported field Q (RK_DUCT
// src/duct.F90:683-1161 for U1..U3, :1169-1326 for theta): per (K,J)
// pencil, one wavefront computes in registers
//   Sx  = J U1 Q                                              (:685-700)
//   Syz = -(U2b flux)_eta + nu (G12 Q_eta)_zeta
//         - (U3b flux)_zeta + nu (G12 Q_zeta)_eta             (:721-761)
// at the pencil's NX points, transforms both (forward, 1/NX, 2/3 rule) and
// accumulates CF = (CF - i kx Sx) + Syz, CR = CR + TEMP5 CF (:714, :772-777,
// :1127-1149), so the two flux fields never round-trip through HBM.  The
// arithmetic of every value is that of the separate flux, transform and
// accumulation kernels.
// LES adds the subgrid model's eddy-viscosity cross terms (:990-1120) to Syz
// before the transform: with the face averages of NU_T and GL = GMAT_12
// (U1), GMAT_12_y (U2) or GMAT_12_z (U3),
//   Sl = 1/4 [GL(K+1) nu(K+1/2) Q_eta(K+1/2) - GL(K) nu(K-1/2) Q_eta(K-1/2)
//           + GL(J+1) nu(J+1/2) Q_zeta(J+1/2) - GL(J) nu(J-1/2) Q_zeta(J-1/2)].
// The reference transforms Sl separately and adds it to CF afterwards; the
// transform is linear, so the two differ by rounding only.
#define MET(arr, k, j, f) M.arr[mf(d, k, j, f)]
template <int R, bool LES>
__global__ void __launch_bounds__(kBlock) k_flux_accum(Dims d, Metrics M, Ranges rg, const double* __restrict__ U1,
                                                       const double* __restrict__ Q, const double* __restrict__ U2b,
                                                       const double* __restrict__ U3b, double nu,
                                                       const double* __restrict__ kxv, double2* __restrict__ CF,
                                                       double2* __restrict__ CR, const double* __restrict__ temp5p, int logm,
                                                       const double2* __restrict__ tw,
                                                       const double2* __restrict__ tw2,
                                                       const double* __restrict__ nut,
                                                       const double* __restrict__ gl) {
  extern __shared__ double2 lds[];
  const int nx = d.NX, M2 = nx >> 1, GW = M2 < 64 ? M2 : 64, PW = 64 / GW;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lanep = lane % GW, sub = lane / GW;
  const int nk = rg.ze - rg.zs + 1, nj = rg.je - rg.js + 1;
  const long npen = (long)nk * nj;
  const long pen = ((long)blockIdx.x * kWaves + wave) * PW + sub;
  const bool live = pen < npen;
  const int k = rg.zs + (int)(live ? pen % nk : 0), j = rg.js + (int)(live ? pen / nk : 0);
  const size_t dj = (size_t)d.NX2 * d.NZ2, dk = d.NX2;
  const size_t q0 = (size_t)d.NX2 * ((size_t)k + (size_t)d.NZ2 * j);
  const double Jc = M.jac[mi(d, k, j)];
  const double g12kp = MET(g12, k + 1, j, 2), g12k = MET(g12, k, j, 2);
  const double g12jp = MET(g12, k, j + 1, 1), g12j = MET(g12, k, j, 1);
  double glkp = 0.0, glk = 0.0, gljp = 0.0, glj = 0.0;
  if constexpr (LES) {
    glkp = gl[mf(d, k + 1, j, 2)];
    glk = gl[mf(d, k, j, 2)];
    gljp = gl[mf(d, k, j + 1, 1)];
    glj = gl[mf(d, k, j, 1)];
  }
  const LaneTw lt = lane_twiddles(lanep, M2, tw);
  double2 zx[R], zy[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    double sx[2] = {0.0, 0.0}, sy[2] = {0.0, 0.0};
    if (live) {
      const size_t q = q0 + 2 * (size_t)(lanep + GW * r);
      const double2 qc = *reinterpret_cast<const double2*>(Q + q);
      const double2 u1 = *reinterpret_cast<const double2*>(U1 + q);
      const double2 qjp = *reinterpret_cast<const double2*>(Q + q + dj);
      const double2 qjm = *reinterpret_cast<const double2*>(Q + q - dj);
      const double2 qkp = *reinterpret_cast<const double2*>(Q + q + dk);
      const double2 qkm = *reinterpret_cast<const double2*>(Q + q - dk);
      const double2 qpp = *reinterpret_cast<const double2*>(Q + q + dk + dj);
      const double2 qpm = *reinterpret_cast<const double2*>(Q + q + dk - dj);
      const double2 qmp = *reinterpret_cast<const double2*>(Q + q - dk + dj);
      const double2 qmm = *reinterpret_cast<const double2*>(Q + q - dk - dj);
      const double2 vjp = *reinterpret_cast<const double2*>(U2b + q + dj);
      const double2 vj = *reinterpret_cast<const double2*>(U2b + q);
      const double2 wkp = *reinterpret_cast<const double2*>(U3b + q + dk);
      const double2 wk = *reinterpret_cast<const double2*>(U3b + q);
      const double QC[2] = {qc.x, qc.y}, UU[2] = {u1.x, u1.y}, QJP[2] = {qjp.x, qjp.y}, QJM[2] = {qjm.x, qjm.y};
      const double QKP[2] = {qkp.x, qkp.y}, QKM[2] = {qkm.x, qkm.y}, QPP[2] = {qpp.x, qpp.y};
      const double QPM[2] = {qpm.x, qpm.y}, QMP[2] = {qmp.x, qmp.y}, QMM[2] = {qmm.x, qmm.y};
      const double VJP[2] = {vjp.x, vjp.y}, VJ[2] = {vj.x, vj.y}, WKP[2] = {wkp.x, wkp.y}, WK[2] = {wk.x, wk.y};
      double NC[2] = {0.0, 0.0}, NKP[2] = {0.0, 0.0}, NKM[2] = {0.0, 0.0}, NJP[2] = {0.0, 0.0}, NJM[2] = {0.0, 0.0};
      if constexpr (LES) {
        const double2 nc = *reinterpret_cast<const double2*>(nut + q);
        const double2 nkp = *reinterpret_cast<const double2*>(nut + q + dk);
        const double2 nkm = *reinterpret_cast<const double2*>(nut + q - dk);
        const double2 njp = *reinterpret_cast<const double2*>(nut + q + dj);
        const double2 njm = *reinterpret_cast<const double2*>(nut + q - dj);
        NC[0] = nc.x; NC[1] = nc.y; NKP[0] = nkp.x; NKP[1] = nkp.y; NKM[0] = nkm.x; NKM[1] = nkm.y;
        NJP[0] = njp.x; NJP[1] = njp.y; NJM[0] = njm.x; NJM[1] = njm.y;
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        sx[h] = Jc * UU[h] * QC[h];
        double s1 = 0.5 * VJP[h] * (QC[h] + QJP[h]) - 0.5 * VJ[h] * (QC[h] + QJM[h]);
        double s2 = 0.5 * WKP[h] * (QC[h] + QKP[h]) - 0.5 * WK[h] * (QC[h] + QKM[h]);
        s1 = -s1 + 0.25 * nu * (g12kp * (QPP[h] + QJP[h] - QJM[h] - QPM[h]) - g12k * (QMP[h] + QJP[h] - QJM[h] - QMM[h]));
        s2 = -s2 + 0.25 * nu * (g12jp * (QPP[h] + QKP[h] - QMP[h] - QKM[h]) - g12j * (QKP[h] + QPM[h] - QKM[h] - QMM[h]));
        sy[h] = s1 + s2;
        if constexpr (LES)
          sy[h] = sy[h] + 0.25 * (glkp * 0.5 * (NKP[h] + NC[h]) * (QPP[h] + QJP[h] - QJM[h] - QPM[h]) -
                                  glk * 0.5 * (NC[h] + NKM[h]) * (QMP[h] + QJP[h] - QJM[h] - QMM[h]) +
                                  gljp * 0.5 * (NC[h] + NJP[h]) * (QPP[h] + QKP[h] - QMP[h] - QKM[h]) -
                                  glj * 0.5 * (NC[h] + NJM[h]) * (QKP[h] + QPM[h] - QKM[h] - QMM[h]));
      }
    }
    zx[r] = make_double2(sx[0], sx[1]);
    zy[r] = make_double2(sy[0], sy[1]);
  }
  dif_any<R>(zx, GW, M2, lanep, tw, lt);
  dif_any<R>(zy, GW, M2, lanep, tw, lt);
  double2* Zx = lds + (size_t)2 * (wave * PW + sub) * M2;
  double2* Zy = Zx + M2;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int b = bitrev(lanep + GW * r, logm);
    Zx[b] = zx[r];
    Zy[b] = zy[r];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  if (!live) return;
  const double sc = 1.0 / (double)nx;
  const int nkx = nx / 3;
  const size_t f0 = (size_t)d.NXC * ((size_t)k + (size_t)d.NZ2 * j);
  // the outputs m = lanep + GW i (i <= R): all global operands are loaded
  // first, then combined, so the lane pays one memory round trip
  double2 cf[R + 1], cr[R + 1], w2[R + 1];
  double kw[R + 1];
#pragma unroll
  for (int i = 0; i <= R; ++i) {
    const int m = min(lanep + GW * i, M2);
    cf[i] = CF[f0 + m];
    cr[i] = CR[f0 + m];
    w2[i] = tw2[m];
    kw[i] = kxv[m];
  }
#pragma unroll
  for (int i = 0; i <= R; ++i) {
    const int m = lanep + GW * i;
    if (m > M2) break;
    double2 X[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
    if (m <= nkx) {
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const double2* Zt = t == 0 ? Zx : Zy;
        const double2 zk = Zt[m == M2 ? 0 : m];
        double2 y = Zt[(M2 - m) % M2];
        y.y = -y.y;
        const double2 e = make_double2(0.5 * (zk.x + y.x), 0.5 * (zk.y + y.y));
        const double2 o = make_double2(0.5 * (zk.y - y.y), -0.5 * (zk.x - y.x));
        const double2 wo = cmul(w2[i], o);
        X[t] = make_double2((e.x + wo.x) * sc, (e.y + wo.y) * sc);
      }
    }
    double2 f = cf[i];
    const double w = kw[i];
    f.x = f.x - (0.0 * X[0].x - w * X[0].y);
    f.y = f.y - (0.0 * X[0].y + w * X[0].x);
    f.x = f.x + X[1].x;
    f.y = f.y + X[1].y;
    CF[f0 + m] = f;
    double2 c = cr[i];
    const double temp5 = *temp5p;
    c.x = c.x + temp5 * f.x;
    c.y = c.y + temp5 * f.y;
    CR[f0 + m] = c;
  }
}
#undef MET

template <int R>
int launch_flux_accum_r(const FftPlan& p, const Dims& d, const Metrics& M, const Ranges& rg, const double* U1,
                        const double* Q, const double* U2b, const double* U3b, double nu, const double* kx,
                        double* CF, double* CR, const double* temp5, const double* nut, const double* gl, hipStream_t st) {
  const int M2 = p.nx / 2, GW = M2 < 64 ? M2 : 64, PW = 64 / GW;
  const long npen = (long)(rg.ze - rg.zs + 1) * (rg.je - rg.js + 1);
  if (npen <= 0) return EKM_OK;
  const long per_block = (long)kWaves * PW;
  const long nb = (npen + per_block - 1) / per_block;
  const size_t lds = (size_t)2 * kWaves * PW * M2 * sizeof(double2);
  int logm = 0;
  while ((1 << logm) < M2) ++logm;
  if (nut)
    hipLaunchKernelGGL((k_flux_accum<R, true>), dim3((unsigned)nb), dim3(kBlock), lds, st, d, M, rg, U1, Q, U2b, U3b,
                       nu, kx, (double2*)CF, (double2*)CR, temp5, logm, (const double2*)p.tw, (const double2*)p.tw2,
                       nut, gl);
  else
    hipLaunchKernelGGL((k_flux_accum<R, false>), dim3((unsigned)nb), dim3(kBlock), lds, st, d, M, rg, U1, Q, U2b,
                       U3b, nu, kx, (double2*)CF, (double2*)CR, temp5, logm, (const double2*)p.tw,
                       (const double2*)p.tw2, nullptr, nullptr);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}

template <int R>
int launch_pair(bool fwd, const FftPlan& p, double* v, long npencil, hipStream_t st) {
  const int M = p.nx / 2, GW = M < 64 ? M : 64, PW = 64 / GW;
  const long per_block = (long)kWaves * PW;
  const long nb = (npencil + per_block - 1) / per_block;
  const size_t lds = (size_t)kWaves * PW * M * sizeof(double2);
  int logm = 0;
  while ((1 << logm) < M) ++logm;
  if (fwd)
    hipLaunchKernelGGL(k_fft_fwd<R>, dim3((unsigned)nb), dim3(kBlock), lds, st, v, npencil, p.nx, logm,
                       (const double2*)p.tw, (const double2*)p.tw2);
  else
    hipLaunchKernelGGL(k_fft_inv<R>, dim3((unsigned)nb), dim3(kBlock), lds, st, v, npencil, p.nx, logm,
                       (const double2*)p.tw, (const double2*)p.tw2);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}

int launch(bool fwd, const FftPlan& p, double* v, long npencil, hipStream_t st) {
  if (npencil <= 0) return EKM_OK;
  const int M = p.nx / 2;
  if (M <= 64) return launch_pair<1>(fwd, p, v, npencil, st);
  if (M == 128) return launch_pair<2>(fwd, p, v, npencil, st);
  if (M == 256) return launch_pair<4>(fwd, p, v, npencil, st);
  return launch_pair<8>(fwd, p, v, npencil, st);
}

}  // namespace

// twiddle tables of an NX-point real transform: tw = exp(-2 pi i k / M),
// k < M/2, and tw2 = exp(-2 pi i k / NX), k <= M, M = NX/2 (re, im pairs);
// tw holds max(M/2, 1) entries
void fft_twiddles(int nx, double* tw, double* tw2) {
  const int M = nx / 2;
  if (M / 2 == 0) tw[0] = tw[1] = 0.0;
  for (int k = 0; k < M / 2; ++k) {
    tw[2 * k] = std::cos(2.0 * M_PI * k / M);
    tw[2 * k + 1] = -std::sin(2.0 * M_PI * k / M);
  }
  for (int k = 0; k <= M; ++k) {
    tw2[2 * k] = std::cos(2.0 * M_PI * k / nx);
    tw2[2 * k + 1] = -std::sin(2.0 * M_PI * k / nx);
  }
}
size_t fft_tw_doubles(int nx) { return 2 * (size_t)std::max(nx / 4, 1); }
size_t fft_tw2_doubles(int nx) { return 2 * (size_t)(nx / 2 + 1); }

int fft_plan_init(FftPlan& p, int nx) {
  if (nx < 8 || nx > 1024 || (nx & (nx - 1))) return EKM_ERR_UNSUPPORTED;
  p.nx = nx;
  std::vector<double> tw(fft_tw_doubles(nx)), tw2(fft_tw2_doubles(nx));
  fft_twiddles(nx, tw.data(), tw2.data());
  EKM_CHECK(hipMalloc(&p.tw, tw.size() * sizeof(double)));
  EKM_CHECK(hipMalloc(&p.tw2, tw2.size() * sizeof(double)));
  EKM_CHECK(hipMemcpy(p.tw, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice));
  EKM_CHECK(hipMemcpy(p.tw2, tw2.data(), tw2.size() * sizeof(double), hipMemcpyHostToDevice));
  return EKM_OK;
}

void fft_plan_free(FftPlan& p) {
  if (p.tw) (void)hipFree(p.tw);
  if (p.tw2) (void)hipFree(p.tw2);
  p.tw = p.tw2 = nullptr;
}

int fft_forward(const FftPlan& p, double* v, long npencil, hipStream_t st) { return launch(true, p, v, npencil, st); }

int fft_inverse(const FftPlan& p, double* v, long npencil, hipStream_t st) { return launch(false, p, v, npencil, st); }

int launch_flux_accum(const FftPlan& p, const Dims& d, const Metrics& M, const Ranges& r, const double* U1,
                      const double* Q, const double* U2b, const double* U3b, double nu, const double* kx, double* CF,
                      double* CR, const double* temp5, hipStream_t st, const double* nut, const double* gl) {
  const int M2 = p.nx / 2;
  if (nut && !gl) return EKM_ERR_ARG;
  if (M2 <= 64) return launch_flux_accum_r<1>(p, d, M, r, U1, Q, U2b, U3b, nu, kx, CF, CR, temp5, nut, gl, st);
  if (M2 == 128) return launch_flux_accum_r<2>(p, d, M, r, U1, Q, U2b, U3b, nu, kx, CF, CR, temp5, nut, gl, st);
  if (M2 == 256) return launch_flux_accum_r<4>(p, d, M, r, U1, Q, U2b, U3b, nu, kx, CF, CR, temp5, nut, gl, st);
  return launch_flux_accum_r<8>(p, d, M, r, U1, Q, U2b, U3b, nu, kx, CF, CR, temp5, nut, gl, st);
}

// ---------------------------------------------------------------------------
// The component drop-ins FFT_X_TO_FOURIER_OP / FFT_X_TO_PHYSICAL_OP on device
// pointers: plans are built once per (device, NX) and kept for the process.
// ---------------------------------------------------------------------------

namespace {
std::mutex g_plan_mu;
std::vector<std::pair<std::pair<int, int>, FftPlan>> g_plans;

int cached_plan(int nx, const FftPlan** out) {
  int dev = 0;
  EKM_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_plan_mu);
  for (auto& e : g_plans)
    if (e.first.first == dev && e.first.second == nx) {
      *out = &e.second;
      return EKM_OK;
    }
  FftPlan p;
  const int st = fft_plan_init(p, nx);
  if (st != EKM_OK) {
    fft_plan_free(p);
    return st;
  }
  g_plans.reserve(64);
  if (g_plans.size() == g_plans.capacity()) return EKM_ERR_ARG;  // pointers into the vector stay valid
  g_plans.push_back({{dev, nx}, p});
  *out = &g_plans.back().second;
  return EKM_OK;
}

}  // namespace

extern "C" int ekm_fft_x_to_fourier_dev(double* varp, int nx, int npencil, void* stream) {
  const FftPlan* p = nullptr;
  const int e = cached_plan(nx, &p);
  if (e != EKM_OK) return e;
  const int st = fft_forward(*p, varp, npencil, (hipStream_t)stream);
  EKM_CHECK(hipStreamSynchronize((hipStream_t)stream));
  return st;
}

extern "C" int ekm_fft_x_to_physical_dev(double* varp, int nx, int npencil, void* stream) {
  const FftPlan* p = nullptr;
  const int e = cached_plan(nx, &p);
  if (e != EKM_OK) return e;
  const int st = fft_inverse(*p, varp, npencil, (hipStream_t)stream);
  EKM_CHECK(hipStreamSynchronize((hipStream_t)stream));
  return st;
}
