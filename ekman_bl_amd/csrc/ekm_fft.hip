// ekm_fft.hip -- batched real<->complex transforms along x for gfx950.
//
// Replaces FFT_X_TO_FOURIER_OP / FFT_X_TO_PHYSICAL_OP of the reference's
// missing fft.F90 (call sites src/mpi_duct.F90:290,308; conventions fixed by
// the callers, see ekman_bl_amd.h): forward c_m = (1/NX) sum_i u_i e^{-2 pi i m i/NX}
// with modes NKX+1..NX/2 zeroed; inverse from modes 0..NKX (Im c_0 ignored),
// padding entries NX, NX+1 zeroed.
//
// A real pencil of NX doubles is read as NX/2 complex values (even, odd);
// an LDS-resident radix-2 Stockham FFT of length M = NX/2 runs with one
// butterfly per thread per pass; the even/odd split is undone (forward) or
// set up (inverse) in LDS.  P = 256/(M/2) pencils share a 256-thread block,
// so global loads and stores are contiguous runs of P*(NX+2) doubles.
#include "ekm_internal.h"

#include <cmath>
#include <cstdio>

namespace {

constexpr int kBlock = 256;

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// One Stockham radix-2 pass over a pencil of M complex values held in `src`;
// result in `dst`.  j in [0, M/2).
template <bool INV>
__device__ inline void stockham_pass(const double2* src, double2* dst, int M, int ns, int j,
                                     const double2* __restrict__ tw) {
  int half = M >> 1;
  double2 v0 = src[j], v1 = src[j + half];
  int k = j & (ns - 1);
  double2 w = tw[k * (half / ns)];  // exp(-2 pi i k / (2 ns))
  if (INV) w.y = -w.y;
  v1 = cmul(v1, w);
  int d = ((j - k) << 1) + k;
  dst[d] = make_double2(v0.x + v1.x, v0.y + v1.y);
  dst[d + ns] = make_double2(v0.x - v1.x, v0.y - v1.y);
}

template <bool INV>
__device__ inline double2* run_fft(double2* a, double2* b, int M, int j, const double2* tw) {
  for (int ns = 1; ns < M; ns <<= 1) {
    stockham_pass<INV>(a, b, M, ns, j, tw);
    __syncthreads();
    double2* t = a;
    a = b;
    b = t;
  }
  return a;
}

__global__ void __launch_bounds__(kBlock) k_fft_fwd(double* __restrict__ v, long npencil, int nx,
                                                    const double2* __restrict__ tw,
                                                    const double2* __restrict__ tw2) {
  extern __shared__ double2 lds[];
  const int M = nx >> 1, T = M >> 1, P = kBlock / T;
  const int nkx = nx / 3;
  const int stride = nx + 2;
  long p0 = (long)blockIdx.x * P;
  int np = (int)((npencil - p0) < P ? (npencil - p0) : P);
  // two LDS regions of P*(M+1) values, pencil stride M+1, so either can take
  // the NX+2 output doubles of every pencil without touching the other
  const int ld = M + 1;
  double2* A = lds;
  double2* B = lds + (size_t)P * ld;
  double* Ad = reinterpret_cast<double*>(A);
  // load: the first NX doubles of each pencil are the packed (even, odd) array
  const double* src = v + p0 * stride;
  for (int t = threadIdx.x; t < np * stride; t += kBlock) {
    int p = t / stride, i = t - p * stride;
    if (i < nx) Ad[(size_t)p * 2 * ld + i] = src[t];
  }
  __syncthreads();
  int p = threadIdx.x / T, j = threadIdx.x - p * T;
  double2* Z = run_fft<false>(A + (size_t)p * ld, B + (size_t)p * ld, M, j, tw);
  // the result sits in A or B depending on the pass-count parity (same for all pencils)
  double2* Zall = (Z == A + (size_t)p * ld) ? A : B;
  double2* Out = (Zall == A) ? B : A;
  // split: X_k = E_k + W^k O_k, k = 0..M (only k <= NKX kept), scaled by 1/NX
  const double sc = 1.0 / (double)nx;
  for (int t = threadIdx.x; t < np * (M + 1); t += kBlock) {
    int q = t / (M + 1), k = t - q * (M + 1);
    double2 x = make_double2(0.0, 0.0);
    if (k <= nkx) {
      const double2* Zq = Zall + (size_t)q * ld;
      double2 z = Zq[k == M ? 0 : k];
      double2 y = Zq[(M - k) % M];
      y.y = -y.y;  // conj(Z_{M-k})
      double2 e = make_double2(0.5 * (z.x + y.x), 0.5 * (z.y + y.y));
      double2 o = make_double2(0.5 * (z.y - y.y), -0.5 * (z.x - y.x));
      double2 w = tw2[k];
      double2 wo = cmul(w, o);
      x = make_double2((e.x + wo.x) * sc, (e.y + wo.y) * sc);
    }
    Out[(size_t)q * ld + k] = x;
  }
  __syncthreads();
  double* dst = v + p0 * stride;
  const double* Od = reinterpret_cast<const double*>(Out);
  for (int t = threadIdx.x; t < np * stride; t += kBlock) dst[t] = Od[t];
}

__global__ void __launch_bounds__(kBlock) k_fft_inv(double* __restrict__ v, long npencil, int nx,
                                                    const double2* __restrict__ tw,
                                                    const double2* __restrict__ tw2) {
  extern __shared__ double2 lds[];
  const int M = nx >> 1, T = M >> 1, P = kBlock / T;
  const int nkx = nx / 3;
  const int stride = nx + 2;
  long p0 = (long)blockIdx.x * P;
  int np = (int)((npencil - p0) < P ? (npencil - p0) : P);
  double2* A = lds;                        // P*(M+1) input modes
  double2* B = lds + (size_t)P * (M + 1);  // P*M work
  double2* C = B + (size_t)P * M;          // P*M work
  double* Ad = reinterpret_cast<double*>(A);
  const double* src = v + p0 * stride;
  for (int t = threadIdx.x; t < np * stride; t += kBlock) Ad[t] = src[t];
  __syncthreads();
  // Z_k = (X_k + conj X_{M-k}) + i e^{+2 pi i k/NX} (X_k - conj X_{M-k}), k < M
  for (int t = threadIdx.x; t < np * M; t += kBlock) {
    int q = t / M, k = t - q * M;
    const double2* Xq = A + (size_t)q * (M + 1);
    int b = M - k;
    double2 x = (k <= nkx) ? Xq[k] : make_double2(0.0, 0.0);
    if (k == 0) x.y = 0.0;
    double2 y = (b <= nkx && b < M) ? Xq[b] : make_double2(0.0, 0.0);
    if (b == 0) y.y = 0.0;
    y.y = -y.y;
    double2 e = make_double2(x.x + y.x, x.y + y.y);
    double2 dd = make_double2(x.x - y.x, x.y - y.y);
    double2 w = tw2[k];
    w.y = -w.y;  // exp(+2 pi i k / NX)
    double2 o = cmul(dd, w);
    B[(size_t)q * M + k] = make_double2(e.x - o.y, e.y + o.x);
  }
  __syncthreads();
  int p = threadIdx.x / T, j = threadIdx.x - p * T;
  double2* Z = run_fft<true>(B + (size_t)p * M, C + (size_t)p * M, M, j, tw);
  double2* Zall = (Z == B + (size_t)p * M) ? B : C;
  const double* Zd = reinterpret_cast<const double*>(Zall);
  double* dst = v + p0 * stride;
  for (int t = threadIdx.x; t < np * stride; t += kBlock) {
    int q = t / stride, i = t - q * stride;
    dst[t] = (i < nx) ? Zd[(size_t)q * 2 * M + i] : 0.0;
  }
}

}  // namespace

int fft_plan_init(FftPlan& p, int nx) {
  if (nx < 8 || nx > 1024 || (nx & (nx - 1))) return EKM_ERR_UNSUPPORTED;
  p.nx = nx;
  int M = nx / 2;
  std::vector<double> tw(2 * (M / 2 > 0 ? M / 2 : 1)), tw2(2 * (M + 1));
  for (int k = 0; k < M / 2; ++k) {
    tw[2 * k] = std::cos(2.0 * M_PI * k / M);
    tw[2 * k + 1] = -std::sin(2.0 * M_PI * k / M);
  }
  for (int k = 0; k <= M; ++k) {
    tw2[2 * k] = std::cos(2.0 * M_PI * k / nx);
    tw2[2 * k + 1] = -std::sin(2.0 * M_PI * k / nx);
  }
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

int fft_forward(const FftPlan& p, double* v, long npencil, hipStream_t st) {
  int M = p.nx / 2, T = M / 2, P = kBlock / T;
  long nb = (npencil + P - 1) / P;
  size_t lds = (size_t)2 * P * (M + 1) * sizeof(double2);
  hipLaunchKernelGGL(k_fft_fwd, dim3((unsigned)nb), dim3(kBlock), lds, st, v, npencil, p.nx,
                     (const double2*)p.tw, (const double2*)p.tw2);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}

int fft_inverse(const FftPlan& p, double* v, long npencil, hipStream_t st) {
  int M = p.nx / 2, T = M / 2, P = kBlock / T;
  long nb = (npencil + P - 1) / P;
  size_t lds = (size_t)P * ((M + 1) + 2 * M) * sizeof(double2);
  hipLaunchKernelGGL(k_fft_inv, dim3((unsigned)nb), dim3(kBlock), lds, st, v, npencil, p.nx,
                     (const double2*)p.tw, (const double2*)p.tw2);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}
