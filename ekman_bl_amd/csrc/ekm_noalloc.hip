// ekm_noalloc.hip -- the allocation-free forward x-transform.
//
// ekm_fft_x_to_fourier_dev keeps one plan per (device, NX), allocated with
// hipMalloc on first use.  Callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, hosts that forbid it) use
// ekm_fft_x_to_fourier_dev_noalloc instead: the plan's twiddle tables go to
// a file-scope device array, so no call allocates.
#include <mutex>

#include "ekm_internal.h"

// Scratch of the allocation-free transform: the twiddle tables of one plan,
// NX <= kNoAllocMaxNx, 1538 doubles (12 KB).  Each device holds its own copy
// until the process ends; host code uses only the address that
// hipGetSymbolAddress returns for the current device, never the array's name.
constexpr int kNoAllocMaxNx = 1024;
constexpr int kScratchDoubles = 2 * (kNoAllocMaxNx / 4) + 2 * (kNoAllocMaxNx / 2 + 1);
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch_dev[kMaxDevices] = {};

// the scratch's device address on the current device, looked up once per device
int scratch_ptr(double** out) {
  int dev = 0;
  EKM_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return EKM_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    EKM_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = static_cast<double*>(ptr);
  }
  *out = g_scratch_dev[dev];
  return EKM_OK;
}
}  // namespace

// ekm_fft_x_to_fourier_dev without any device allocation (NX <= 1024).  Every
// call on a device shares that device's scratch: calls that overlap on one
// device race.  Each call rewrites the scratch before its transform and waits
// for the transform to end.
extern "C" int ekm_fft_x_to_fourier_dev_noalloc(double* varp, int nx, int npencil, void* stream) {
  if (nx > kNoAllocMaxNx) {
    ekm_note_error("ekm_fft_x_to_fourier_dev_noalloc: NX above 1024 does not fit the transform's scratch");
    return EKM_ERR_ARG;
  }
  if (nx < 8 || (nx & (nx - 1)) || npencil < 0) return EKM_ERR_UNSUPPORTED;
  double* sp = nullptr;
  const int e = scratch_ptr(&sp);
  if (e != EKM_OK) return e;
  const size_t n1 = fft_tw_doubles(nx), n2 = fft_tw2_doubles(nx);
  double host[kScratchDoubles];
  fft_twiddles(nx, host, host + n1);
  hipStream_t st = (hipStream_t)stream;
  EKM_CHECK(hipMemcpyAsync(sp, host, (n1 + n2) * sizeof(double), hipMemcpyHostToDevice, st));
  FftPlan p;
  p.nx = nx;
  p.tw = sp;
  p.tw2 = sp + n1;
  const int r = fft_forward(p, varp, npencil, st);
  // the host tables must outlive the copy, and the next call rewrites the scratch
  EKM_CHECK(hipStreamSynchronize(st));
  return r;
}
