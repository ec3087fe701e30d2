// ekm_wsnet.hip -- learned initial guess of the per-mode pressure solves.
//
// No counterpart in the reference: MUDPACK starts every solve from zero
// (iguess = 0, phi = 0; src/mg_solver2d.F90:157,190; the iguess branch of
// src/mud2cr.f:93,238).  Here a small convolutional encoder-decoder predicts
// the start of each real solve of the consistent projection from its
// right-hand side, the previous projection's solution of the same solve and
// kx; the multigrid then iterates to the same tolerance as from zero, so it
// stays both the fallback and what decides when a solve is done.
//
// Network (one plane per sample, activations (H, W, C) with C fastest; the
// plane's K (z) index is W, J (y) is H; H0 x W0 = (NY+2) x (NZ+2)):
//   input at level 1 (full weighting of the plane, mirrored ends), 4 channels:
//     rhs/s, prev/(s sigma), kx/kx_max, J/(H1-1), with s = max|rhs| per solve;
//   five levels 1..5 (stride-2 3x3 convolutions down, bilinear prolongation
//   up), channels 8, 16, 32, 48, 64, two 3x3 convolutions with ReLU per
//   level and direction, the skip of each encoder level concatenated ahead of
//   the prolonged decoder tensor, a 1x1 output convolution; the prediction
//   is prolonged to level 0 and scaled by s sigma.
// The 3x3 convolutions are implicit GEMMs on the bf16 matrix cores
// (v_mfma_f32_16x16x16_bf16, fp32 accumulation): the forward product, the
// input gradient and the weight gradient; activations, gradients, master
// weights and Adam moments are fp32 in HBM and rounded to bf16 (nearest
// even) when they are loaded as matrix operands.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "ekm_wsnet.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kCh[5] = {8, 16, 32, 48, 64};
constexpr int kCin0 = 4;

// fp32 -> bf16, round to nearest even (finite values)
__device__ inline unsigned short f2bf(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline bf16x4 bf4(float a, float b, float c, float d) {
  bf16x4 r;
  r[0] = (short)f2bf(a);
  r[1] = (short)f2bf(b);
  r[2] = (short)f2bf(c);
  r[3] = (short)f2bf(d);
  return r;
}
inline unsigned nblk(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ---------------------------------------------------------------------------
// 3x3 convolution, forward: rows = output pixels of the batch (16 per wave),
// columns = output channels (NT tiles of 16), K = 9 cin ordered (tap, ci) in
// chunks of 16 (lane l: pixel l%16, k = 4 (l/16) .. +3; the 4 channels of one
// tap are one float4).  Result lane layout: pixel 4 (l/16) + v, channel l%16.
// ---------------------------------------------------------------------------
struct ConvFwd {
  const float* xa;
  const float* xb;  // second input (concatenated after xa), may be null
  int ca, cb;
  int hi, wi, ho, wo, stride, n;
  const unsigned short* w;  // [cpad][kpad] bf16, k = tap*cin + ci
  int kpad, k;
  const float* bias;
  int cout, relu;
  float* y;
};

template <int NT>
__global__ void __launch_bounds__(256) k_conv3_fwd(ConvFwd a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long P = (long)a.n * a.ho * a.wo;
  const long p0 = ((long)blockIdx.x * 4 + wave) * 16;
  if (p0 >= P) return;
  const long p = p0 + (lane & 15);
  const bool pv = p < P;
  int n = 0, oy = 0, ox = 0;
  if (pv) {
    n = (int)(p / ((long)a.ho * a.wo));
    const int r = (int)(p - (long)n * a.ho * a.wo);
    oy = r / a.wo;
    ox = r - oy * a.wo;
  }
  const int kq = (lane >> 4) * 4;
  const int cin = a.ca + a.cb;
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const unsigned short* wrow = a.w + (size_t)(lane & 15) * a.kpad + kq;
  for (int k0 = 0; k0 < a.kpad; k0 += 16) {
    const int k = k0 + kq;
    bf16x4 av = bf16x4{0, 0, 0, 0};
    if (pv && k < a.k) {
      const int tap = k / cin, ci = k - tap * cin;
      const int iy = oy * a.stride + tap / 3 - 1, ix = ox * a.stride + tap % 3 - 1;
      if (iy >= 0 && iy < a.hi && ix >= 0 && ix < a.wi) {
        const long q = ((long)n * a.hi + iy) * a.wi + ix;
        const float4 v = ci < a.ca ? *reinterpret_cast<const float4*>(a.xa + q * a.ca + ci)
                                   : *reinterpret_cast<const float4*>(a.xb + q * a.cb + (ci - a.ca));
        av = bf4(v.x, v.y, v.z, v.w);
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const bf16x4 bv = *reinterpret_cast<const bf16x4*>(wrow + (size_t)16 * t * a.kpad + k0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int co = 16 * t + (lane & 15);
    if (co >= a.cout) continue;
    const float b = a.bias[co];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const long q = p0 + 4 * (lane >> 4) + v;
      if (q >= P) continue;
      float y = acc[t][v] + b;
      if (a.relu) y = fmaxf(y, 0.f);
      a.y[q * a.cout + co] = y;
    }
  }
}

// ---------------------------------------------------------------------------
// 3x3 convolution, input gradient: rows = input pixels, columns = input
// channels (NT tiles), K = 9 cout ordered (tap, co): the output gradient at
// the pixel that tap reaches from this input pixel (stride 2: only where the
// offset is even).  Written to the two input tensors' gradients (the
// concatenation split), each optionally masked by its ReLU output and
// optionally added to what is there.
// ---------------------------------------------------------------------------
struct ConvBwdD {
  const float* gy;
  int cout, ho, wo, stride, hi, wi, n;
  const unsigned short* w;  // [cbpad][kbpad] bf16, row ci, k = tap*cout + co
  int kbpad, kb;
  float* ga;
  float* gb;
  const float* ma;  // ReLU outputs masking ga (null: no mask)
  const float* mb;
  int ca, cb, acc_a, acc_b;
};

template <int NT>
__global__ void __launch_bounds__(256) k_conv3_bwdd(ConvBwdD a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long P = (long)a.n * a.hi * a.wi;
  const long p0 = ((long)blockIdx.x * 4 + wave) * 16;
  if (p0 >= P) return;
  const long p = p0 + (lane & 15);
  const bool pv = p < P;
  int n = 0, iy = 0, ix = 0;
  if (pv) {
    n = (int)(p / ((long)a.hi * a.wi));
    const int r = (int)(p - (long)n * a.hi * a.wi);
    iy = r / a.wi;
    ix = r - iy * a.wi;
  }
  const int kq = (lane >> 4) * 4;
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const unsigned short* wrow = a.w + (size_t)(lane & 15) * a.kbpad + kq;
  for (int k0 = 0; k0 < a.kbpad; k0 += 16) {
    const int k = k0 + kq;
    bf16x4 av = bf16x4{0, 0, 0, 0};
    if (pv && k < a.kb) {
      const int tap = k / a.cout, co = k - tap * a.cout;
      int oy = iy - (tap / 3 - 1), ox = ix - (tap % 3 - 1);
      bool ok = true;
      if (a.stride == 2) {
        ok = ((oy | ox) & 1) == 0;
        oy >>= 1;
        ox >>= 1;
      }
      if (ok && oy >= 0 && oy < a.ho && ox >= 0 && ox < a.wo) {
        const float4 v = *reinterpret_cast<const float4*>(a.gy + (((long)n * a.ho + oy) * a.wo + ox) * a.cout + co);
        av = bf4(v.x, v.y, v.z, v.w);
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const bf16x4 bv = *reinterpret_cast<const bf16x4*>(wrow + (size_t)16 * t * a.kbpad + k0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int ci = 16 * t + (lane & 15);
    if (ci >= a.ca + a.cb) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const long q = p0 + 4 * (lane >> 4) + v;
      if (q >= P) continue;
      float g = acc[t][v];
      if (ci < a.ca) {
        const long o = q * a.ca + ci;
        if (a.ma && !(a.ma[o] > 0.f)) g = 0.f;
        a.ga[o] = a.acc_a ? a.ga[o] + g : g;
      } else {
        const long o = q * a.cb + (ci - a.ca);
        if (a.mb && !(a.mb[o] > 0.f)) g = 0.f;
        a.gb[o] = a.acc_b ? a.gb[o] + g : g;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 3x3 convolution, weight gradient: dW[co][tap*cin+ci] = sum over output
// pixels of gy[p][co] x(p, tap, ci): rows = cout (16-tile), columns = k
// (16-tile), K = output pixels in chunks of 16; each wave reduces its range
// of pixels and adds its tile to the fp32 gradient (canonical layout
// [cout][cin][9]) with atomics.
// ---------------------------------------------------------------------------
struct ConvBwdW {
  const float* gy;
  const float* xa;
  const float* xb;
  int ca, cb, cout, ho, wo, hi, wi, stride, n;
  int k, ntk;       // 9 cin, column tiles
  long chunk;       // output pixels per workgroup
  float* gw;        // canonical [cout][cin][9]
};

__global__ void __launch_bounds__(256) k_conv3_bwdw(ConvBwdW a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tile = blockIdx.x, tr = tile / a.ntk, tc = tile - tr * a.ntk;
  const int cin = a.ca + a.cb;
  const long P = (long)a.n * a.ho * a.wo;
  const long pa = (long)blockIdx.y * a.chunk, pb = pa + a.chunk < P ? pa + a.chunk : P;
  const int co = 16 * tr + (lane & 15);          // A row of this lane
  const int kcol = 16 * tc + (lane & 15);        // B column of this lane
  const bool cok = co < a.cout, kok = kcol < a.k;
  int tap = 0, ci = 0;
  if (kok) {
    tap = kcol / cin;
    ci = kcol - tap * cin;
  }
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
  const long hw = (long)a.ho * a.wo;
  for (long q0 = pa + (long)wave * 16; q0 < pb; q0 += 64) {
    float av[4], bv[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const long q = q0 + 4 * (lane >> 4) + v;
      av[v] = 0.f;
      bv[v] = 0.f;
      if (q < pb) {
        if (cok) av[v] = a.gy[q * a.cout + co];
        if (kok) {
          const int n = (int)(q / hw);
          const int r = (int)(q - (long)n * hw), oy = r / a.wo, ox = r - oy * a.wo;
          const int iy = oy * a.stride + dy, ix = ox * a.stride + dx;
          if (iy >= 0 && iy < a.hi && ix >= 0 && ix < a.wi) {
            const long s = ((long)n * a.hi + iy) * a.wi + ix;
            bv[v] = ci < a.ca ? a.xa[s * a.ca + ci] : a.xb[s * a.cb + (ci - a.ca)];
          }
        }
      }
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(bf4(av[0], av[1], av[2], av[3]), bf4(bv[0], bv[1], bv[2], bv[3]),
                                                    acc, 0, 0, 0);
  }
  // lane holds rows 4 (l/16) + v (cout), column l%16 (k)
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int r = 16 * tr + 4 * (lane >> 4) + v;
    if (r < a.cout && kok) atomicAdd(&a.gw[((size_t)r * cin + ci) * 9 + tap], acc[v]);
  }
}

// bias gradient: db[c] += sum over the pixels of a chunk of g[p][c]
__global__ void __launch_bounds__(256) k_bias_grad(const float* __restrict__ g, int c, long P, long chunk,
                                                  float* __restrict__ db) {
  const long pa = (long)blockIdx.x * chunk, pb = pa + chunk < P ? pa + chunk : P;
  const int lanes = blockDim.x / c;
  if ((int)threadIdx.x >= lanes * c) return;
  const int ch = threadIdx.x % c, sub = threadIdx.x / c;
  float s = 0.f;
  for (long p = pa + sub; p < pb; p += lanes) s += g[p * c + ch];
  atomicAdd(&db[ch], s);
}

// 1x1 output convolution (c -> 1), forward and backward, fp32
__global__ void k_conv1_fwd(const float* __restrict__ x, int c, long P, const float* __restrict__ w,
                            const float* __restrict__ b, float* __restrict__ y) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  float s = b[0];
  for (int i = 0; i < c; ++i) s += w[i] * x[p * c + i];
  y[p] = s;
}
__global__ void __launch_bounds__(256) k_conv1_bwd(const float* __restrict__ x, int c, long P,
                                                  const float* __restrict__ w, const float* __restrict__ gy,
                                                  float* __restrict__ gx, float* __restrict__ gw,
                                                  float* __restrict__ gb) {
  __shared__ float red[256][9];
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float part[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (p < P) {
    const float g = gy[p];
    for (int i = 0; i < c; ++i) {
      const float xv = x[p * c + i];
      gx[p * c + i] = xv > 0.f ? w[i] * g : 0.f;
      part[i] = g * xv;
    }
    part[8] = g;
  }
  for (int i = 0; i < 9; ++i) red[threadIdx.x][i] = part[i];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int i = 0; i < 9; ++i) red[threadIdx.x][i] += red[threadIdx.x + s][i];
    __syncthreads();
  }
  if (threadIdx.x < (unsigned)c) atomicAdd(&gw[threadIdx.x], red[0][threadIdx.x]);
  if (threadIdx.x == 8) atomicAdd(gb, red[0][8]);
}

// bilinear prolongation (h, w) -> (2h-1, 2w-1) of c channels
__global__ void k_up_fwd(const float* __restrict__ src, int h, int w, int c, int n, float* __restrict__ dst) {
  const int hf = 2 * h - 1, wf = 2 * w - 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long N = (long)n * hf * wf * c;
  if (t >= N) return;
  const int ch = (int)(t % c);
  long r = t / c;
  const int fx = (int)(r % wf);
  r /= wf;
  const int fy = (int)(r % hf), s = (int)(r / hf);
  const int y0 = fy >> 1, x0 = fx >> 1, y1 = y0 + (fy & 1), x1 = x0 + (fx & 1);
  const float* S = src + (long)s * h * w * c;
  const float v = 0.25f * (S[((long)y0 * w + x0) * c + ch] + S[((long)y0 * w + x1) * c + ch] +
                           S[((long)y1 * w + x0) * c + ch] + S[((long)y1 * w + x1) * c + ch]);
  dst[t] = v;
}
// its adjoint, masked by the source's ReLU output
__global__ void k_up_bwd(const float* __restrict__ g, int h, int w, int c, int n, const float* __restrict__ mask,
                         float* __restrict__ gs) {
  const int hf = 2 * h - 1, wf = 2 * w - 1;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long N = (long)n * h * w * c;
  if (t >= N) return;
  const int ch = (int)(t % c);
  long r = t / c;
  const int cx = (int)(r % w);
  r /= w;
  const int cy = (int)(r % h), s = (int)(r / h);
  const float* G = g + (long)s * hf * wf * c;
  float acc = 0.f;
  for (int fy = 2 * cy - 1; fy <= 2 * cy + 1; ++fy) {
    if (fy < 0 || fy >= hf) continue;
    const float wy = fy == 2 * cy ? 1.f : 0.5f;
    float row = 0.f;
    for (int fx = 2 * cx - 1; fx <= 2 * cx + 1; ++fx) {
      if (fx < 0 || fx >= wf) continue;
      row += (fx == 2 * cx ? 1.f : 0.5f) * G[((long)fy * wf + fx) * c + ch];
    }
    acc += wy * row;
  }
  if (mask && !(mask[t] > 0.f)) acc = 0.f;
  gs[t] = acc;
}

__global__ void k_fill(float* __restrict__ a, long n, float v) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}

// value of the level-1 prediction y (h1 x w1) prolonged to level-0 point (fy, fx)
__device__ inline float up_at(const float* y, int w1, int fy, int fx) {
  const int y0 = fy >> 1, x0 = fx >> 1, y1 = y0 + (fy & 1), x1 = x0 + (fx & 1);
  return 0.25f * (y[y0 * w1 + x0] + y[y0 * w1 + x1] + y[y1 * w1 + x0] + y[y1 * w1 + x1]);
}

// per-sample sums of the loss: num = sum (up(y) - t)^2, den = sum t^2
__global__ void __launch_bounds__(256) k_loss(const float* __restrict__ y, const float* __restrict__ t0, int h1,
                                             int w1, float* __restrict__ nd) {
  __shared__ float sn[256], sd[256];
  const int s = blockIdx.x, h0 = 2 * h1 - 1, w0 = 2 * w1 - 1;
  const float* Y = y + (long)s * h1 * w1;
  const float* T = t0 + (long)s * h0 * w0;
  float a = 0.f, b = 0.f;
  for (int q = threadIdx.x; q < h0 * w0; q += blockDim.x) {
    const int fy = q / w0, fx = q - fy * w0;
    const float e = up_at(Y, w1, fy, fx) - T[q];
    a += e * e;
    b += T[q] * T[q];
  }
  sn[threadIdx.x] = a;
  sd[threadIdx.x] = b;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      sn[threadIdx.x] += sn[threadIdx.x + k];
      sd[threadIdx.x] += sd[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    nd[2 * s] = sn[0];
    nd[2 * s + 1] = sd[0];
  }
}
// gradient of the mean relative loss at level 1: gy = up^T (2 (up(y) - t) / den / B)
__global__ void k_loss_grad(const float* __restrict__ y, const float* __restrict__ t0, int h1, int w1, int nb,
                            const float* __restrict__ nd, float* __restrict__ gy) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)nb * h1 * w1) return;
  const int s = (int)(i / ((long)h1 * w1)), r = (int)(i - (long)s * h1 * w1), cy = r / w1, cx = r - cy * w1;
  const int h0 = 2 * h1 - 1, w0 = 2 * w1 - 1;
  const float den = nd[2 * s + 1];
  float acc = 0.f;
  if (den > 0.f) {
    const float* Y = y + (long)s * h1 * w1;
    const float* T = t0 + (long)s * h0 * w0;
    const float sc = 2.f / den / (float)nb;
    for (int fy = 2 * cy - 1; fy <= 2 * cy + 1; ++fy) {
      if (fy < 0 || fy >= h0) continue;
      const float wy = fy == 2 * cy ? 1.f : 0.5f;
      for (int fx = 2 * cx - 1; fx <= 2 * cx + 1; ++fx) {
        if (fx < 0 || fx >= w0) continue;
        const float wx = fx == 2 * cx ? 1.f : 0.5f;
        acc += wy * wx * sc * (up_at(Y, w1, fy, fx) - T[fy * w0 + fx]);
      }
    }
  }
  gy[i] = acc;
}

// Adam on the fp32 parameters (bias-corrected step size lr_t)
__global__ void k_adam(float* __restrict__ th, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, long n, float lr_t, float b1, float b2, float eps) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  th[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

// bf16 operand copies of every 3x3 layer: forward [cpad][kpad] (k = tap*cin +
// ci) and input-gradient [cbpad][kbpad] (row ci, k = tap*cout + co)
struct PackDesc {
  const float* w;  // canonical [cout][cin][9]
  unsigned short* wf;
  unsigned short* wb;
  int cin, cout, cpad, kpad, cbpad, kbpad;
};
__global__ void k_pack(const PackDesc* __restrict__ pd) {
  const PackDesc d = pd[blockIdx.y];
  const int nf = d.cpad * d.kpad, nb = d.cbpad * d.kbpad;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nf + nb; t += gridDim.x * blockDim.x) {
    if (t < nf) {
      const int co = t / d.kpad, k = t - co * d.kpad, tap = k / d.cin, ci = k - tap * d.cin;
      const float v = (co < d.cout && k < 9 * d.cin) ? d.w[((size_t)co * d.cin + ci) * 9 + tap] : 0.f;
      d.wf[t] = f2bf(v);
    } else {
      const int u = t - nf, ci = u / d.kbpad, k = u - ci * d.kbpad, tap = k / d.cout, co = k - tap * d.cout;
      const float v = (ci < d.cin && k < 9 * d.cout) ? d.w[((size_t)co * d.cin + ci) * 9 + tap] : 0.f;
      d.wb[u] = f2bf(v);
    }
  }
}

// per-solve scale s = max |rhs| over the plane
__global__ void __launch_bounds__(256) k_scale(const double* __restrict__ rhs, long plane, double* __restrict__ s) {
  __shared__ double red[256];
  const double* R = rhs + (long)blockIdx.x * plane;
  double m = 0.0;
  for (long q = threadIdx.x; q < plane; q += blockDim.x) m = fmax(m, fabs(R[q]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) s[blockIdx.x] = red[0];
}

__device__ inline int mirror(int i, int n) { return i < 0 ? -i : (i >= n ? 2 * n - 2 - i : i); }
// full weighting of a plane at level-1 point (cy, cx), mirrored ends
__device__ inline double fw(const double* v, int h0, int w0, int cy, int cx) {
  const double wt[3] = {0.25, 0.5, 0.25};
  double acc = 0.0;
  for (int a = 0; a < 3; ++a) {
    const int iy = mirror(2 * cy - 1 + a, h0);
    double row = 0.0;
    for (int b = 0; b < 3; ++b) row += wt[b] * v[(long)iy * w0 + mirror(2 * cx - 1 + b, w0)];
    acc += wt[a] * row;
  }
  return acc;
}
// network input at level 1 (4 channels) from the planes; the previous
// solution enters scaled by ratio (the H_BAR ratio of the two substeps)
__global__ void k_prep(const double* __restrict__ rhs, const double* __restrict__ prev,
                       const double* __restrict__ ratio, const double* __restrict__ kx2,
                       const double* __restrict__ s, int n, int h0, int w0, double sigma, double kx_max,
                       float* __restrict__ x) {
  const int h1 = (h0 + 1) / 2, w1 = (w0 + 1) / 2;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)n * h1 * w1) return;
  const int sm = (int)(i / ((long)h1 * w1)), r = (int)(i - (long)sm * h1 * w1), cy = r / w1, cx = r - cy * w1;
  const double sc = s[sm];
  float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
  if (sc > 0.0) {
    const long pl = (long)h0 * w0;
    o.x = (float)(fw(rhs + sm * pl, h0, w0, cy, cx) / sc);
    const double rt = ratio ? *ratio : 1.0;
    o.y = prev ? (float)(rt * fw(prev + sm * pl, h0, w0, cy, cx) / (sc * sigma)) : 0.f;
    o.z = (float)(sqrt(kx2[sm]) / kx_max);
    o.w = (float)cy / (float)(h1 - 1);
  }
  reinterpret_cast<float4*>(x)[i] = o;
}
// training target at level 0: what the network adds to the scaled previous
// solution, t = (phi - ratio prev) / (s sigma) (zero where s = 0)
__global__ void k_target(const double* __restrict__ phi, const double* __restrict__ prev,
                         const double* __restrict__ ratio, const double* __restrict__ s, int n, long plane,
                         double sigma, float* __restrict__ t) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * plane) return;
  const double sc = s[i / plane];
  const double base = prev ? (ratio ? *ratio : 1.0) * prev[i] : 0.0;
  t[i] = sc > 0.0 ? (float)((phi[i] - base) / (sc * sigma)) : 0.f;
}
// the guess at level 0, as doubles: ratio prev + s sigma up(y); out may
// alias prev (each point reads its own previous value first)
__global__ void k_guess_out(const float* __restrict__ y, const double* prev, const double* __restrict__ ratio,
                            const double* __restrict__ s, int n, int h0, int w0, double sigma, double* out) {
  const int h1 = (h0 + 1) / 2, w1 = (w0 + 1) / 2;
  const long plane = (long)h0 * w0;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * plane) return;
  const int sm = (int)(i / plane), q = (int)(i - sm * plane), fy = q / w0, fx = q - fy * w0;
  const double sc = s[sm];
  const double base = prev ? (ratio ? *ratio : 1.0) * prev[i] : 0.0;
  out[i] = sc > 0.0 ? base + sc * sigma * (double)up_at(y + (long)sm * h1 * w1, w1, fy, fx) : base;
}

}  // namespace

// ---------------------------------------------------------------------------
// the network object
// ---------------------------------------------------------------------------
struct WsTensor {
  int c, h, w;
  float* a = nullptr;  // [bmax][h][w][c]
  float* g = nullptr;  // gradient, when trainable
  bool relu = true;    // a ReLU output (its gradient is masked by it)
};
struct WsOp {
  int kind;  // 0 conv3, 1 up, 2 conv1
  int stride = 1, ina = -1, inb = -1, out = -1, relu = 1;
  int cin = 0, cout = 0, cpad = 0, kpad = 0, cbpad = 0, kbpad = 0;
  size_t w_off = 0, b_off = 0;
  unsigned short *wf = nullptr, *wb = nullptr;
};

struct ekm_wsnet {
  int h[6], w[6];  // level extents; level 0 is the plane (NY+2) x (NZ+2)
  int bmax = 0;
  bool train = false;
  std::vector<WsTensor> T;
  std::vector<WsOp> ops;
  size_t np = 0;
  float *theta = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  unsigned short* wpack = nullptr;
  PackDesc* pd = nullptr;
  int npack = 0;
  bool packed = false;
  long adam_t = 0;
  double sigma = 1.0, kx_max = 1.0;
  double* scale = nullptr;   // [bmax]
  float* x1 = nullptr;       // prepared input [bmax][h1][w1][4] (tensor 0's storage)
  float* lossbuf = nullptr;  // [2 bmax]
  float* gy1 = nullptr;      // loss gradient at level 1 [bmax][h1][w1]
  std::vector<float> host_nd;
};

static int tensor_new(ekm_wsnet* n, int c, int lvl, bool relu) {
  WsTensor t;
  t.c = c;
  t.h = n->h[lvl];
  t.w = n->w[lvl];
  t.relu = relu;
  n->T.push_back(t);
  return (int)n->T.size() - 1;
}

static void net_free(ekm_wsnet* n) {
  for (auto& t : n->T) {
    if (t.a) (void)hipFree(t.a);
    if (t.g) (void)hipFree(t.g);
  }
  float* f[] = {n->theta, n->grad, n->m, n->v, n->lossbuf, n->gy1};
  for (float* p : f)
    if (p) (void)hipFree(p);
  if (n->wpack) (void)hipFree(n->wpack);
  if (n->pd) (void)hipFree(n->pd);
  if (n->scale) (void)hipFree(n->scale);
}

// the layer list of the U-Net (the same as oracle/wsnet.py's NET)
static void build_ops(ekm_wsnet* n) {
  auto conv = [&](int stride, int ina, int inb, int cout, int lvl_out, bool relu) {
    WsOp o;
    o.kind = 0;
    o.stride = stride;
    o.ina = ina;
    o.inb = inb;
    o.cin = n->T[ina].c + (inb >= 0 ? n->T[inb].c : 0);
    o.cout = cout;
    o.relu = relu;
    o.out = tensor_new(n, cout, lvl_out, relu);
    n->ops.push_back(o);
    return o.out;
  };
  auto up = [&](int src, int lvl_out) {
    WsOp o;
    o.kind = 1;
    o.ina = src;
    o.out = tensor_new(n, n->T[src].c, lvl_out, false);
    n->ops.push_back(o);
    return o.out;
  };
  const int x = tensor_new(n, kCin0, 1, false);
  int e = conv(1, x, -1, kCh[0], 1, true);
  int skip[4];
  skip[0] = conv(1, e, -1, kCh[0], 1, true);
  int cur = skip[0];
  for (int l = 1; l <= 4; ++l) {
    e = conv(2, cur, -1, kCh[l], l + 1, true);
    cur = conv(1, e, -1, kCh[l], l + 1, true);
    if (l < 4) skip[l] = cur;
  }
  for (int l = 3; l >= 0; --l) {
    const int u = up(cur, l + 1);
    e = conv(1, skip[l], u, kCh[l], l + 1, true);
    cur = conv(1, e, -1, kCh[l], l + 1, true);
  }
  WsOp o;
  o.kind = 2;
  o.ina = cur;
  o.cin = kCh[0];
  o.cout = 1;
  o.relu = 0;
  o.out = tensor_new(n, 1, 1, false);
  n->ops.push_back(o);
}

static int run_forward(ekm_wsnet* n, int nb, hipStream_t st) {
  for (const WsOp& o : n->ops) {
    WsTensor& Y = n->T[o.out];
    const WsTensor& A = n->T[o.ina];
    if (o.kind == 1) {
      const long N = (long)nb * Y.h * Y.w * Y.c;
      hipLaunchKernelGGL(k_up_fwd, dim3(nblk(N, 256)), dim3(256), 0, st, A.a, A.h, A.w, A.c, nb, Y.a);
    } else if (o.kind == 2) {
      const long P = (long)nb * A.h * A.w;
      hipLaunchKernelGGL(k_conv1_fwd, dim3(nblk(P, 256)), dim3(256), 0, st, A.a, A.c, P, n->theta + o.w_off,
                         n->theta + o.b_off, Y.a);
    } else {
      ConvFwd a;
      a.xa = A.a;
      a.xb = o.inb >= 0 ? n->T[o.inb].a : nullptr;
      a.ca = A.c;
      a.cb = o.inb >= 0 ? n->T[o.inb].c : 0;
      a.hi = A.h;
      a.wi = A.w;
      a.ho = Y.h;
      a.wo = Y.w;
      a.stride = o.stride;
      a.n = nb;
      a.w = o.wf;
      a.kpad = o.kpad;
      a.k = 9 * o.cin;
      a.bias = n->theta + o.b_off;
      a.cout = o.cout;
      a.relu = o.relu;
      a.y = Y.a;
      const long P = (long)nb * Y.h * Y.w;
      const dim3 g(nblk(P, 64));
      switch (o.cpad / 16) {
        case 1: hipLaunchKernelGGL(k_conv3_fwd<1>, g, dim3(256), 0, st, a); break;
        case 2: hipLaunchKernelGGL(k_conv3_fwd<2>, g, dim3(256), 0, st, a); break;
        case 3: hipLaunchKernelGGL(k_conv3_fwd<3>, g, dim3(256), 0, st, a); break;
        default: hipLaunchKernelGGL(k_conv3_fwd<4>, g, dim3(256), 0, st, a); break;
      }
    }
    EKM_CHECK(hipGetLastError());
  }
  return EKM_OK;
}

static int ensure_packed(ekm_wsnet* n, hipStream_t st) {
  if (n->packed) return EKM_OK;
  hipLaunchKernelGGL(k_pack, dim3(64, n->npack), dim3(256), 0, st, n->pd);
  EKM_CHECK(hipGetLastError());
  n->packed = true;
  return EKM_OK;
}

// gradient of the mean loss into n->grad (zeroed first); loss sums into lossbuf
static int run_backward(ekm_wsnet* n, const float* t0, int nb, hipStream_t st) {
  const WsTensor& Yt = n->T[n->ops.back().out];
  hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, st, Yt.a, t0, Yt.h, Yt.w, n->lossbuf);
  EKM_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_fill, dim3(nblk(n->np, 256)), dim3(256), 0, st, n->grad, (long)n->np, 0.f);
  EKM_CHECK(hipGetLastError());
  {
    const long N = (long)nb * Yt.h * Yt.w;
    hipLaunchKernelGGL(k_loss_grad, dim3(nblk(N, 256)), dim3(256), 0, st, Yt.a, t0, Yt.h, Yt.w, nb, n->lossbuf,
                       n->T[n->ops.back().out].g);
    EKM_CHECK(hipGetLastError());
  }
  // gradients of tensors with two readers (the skips) are written by the
  // decoder's concatenation first and added to by the encoder's next layer;
  // ops run in reverse, so the order holds
  std::vector<int> written(n->T.size(), 0);
  for (int i = (int)n->ops.size() - 1; i >= 0; --i) {
    const WsOp& o = n->ops[i];
    WsTensor& Y = n->T[o.out];
    WsTensor& A = n->T[o.ina];
    if (o.kind == 1) {
      const long N = (long)nb * A.h * A.w * A.c;
      hipLaunchKernelGGL(k_up_bwd, dim3(nblk(N, 256)), dim3(256), 0, st, Y.g, A.h, A.w, A.c, nb,
                         A.relu ? A.a : nullptr, A.g);
      EKM_CHECK(hipGetLastError());
      written[o.ina] = 1;
      continue;
    }
    const long P = (long)nb * Y.h * Y.w;
    if (o.kind == 2) {
      hipLaunchKernelGGL(k_conv1_bwd, dim3(nblk(P, 256)), dim3(256), 0, st, A.a, A.c, P, n->theta + o.w_off, Y.g,
                         A.g, n->grad + o.w_off, n->grad + o.b_off);
      EKM_CHECK(hipGetLastError());
      written[o.ina] = 1;
      continue;
    }
    const WsTensor* B = o.inb >= 0 ? &n->T[o.inb] : nullptr;
    // weight and bias gradients
    {
      ConvBwdW a;
      a.gy = Y.g;
      a.xa = A.a;
      a.xb = B ? B->a : nullptr;
      a.ca = A.c;
      a.cb = B ? B->c : 0;
      a.cout = o.cout;
      a.ho = Y.h;
      a.wo = Y.w;
      a.hi = A.h;
      a.wi = A.w;
      a.stride = o.stride;
      a.n = nb;
      a.k = 9 * o.cin;
      a.ntk = (a.k + 15) / 16;
      a.chunk = 4096;
      a.gw = n->grad + o.w_off;
      const int ntr = (o.cout + 15) / 16;
      const dim3 g(ntr * a.ntk, nblk(P, (int)a.chunk));
      hipLaunchKernelGGL(k_conv3_bwdw, g, dim3(256), 0, st, a);
      EKM_CHECK(hipGetLastError());
      hipLaunchKernelGGL(k_bias_grad, dim3(nblk(P, 4096)), dim3(256), 0, st, Y.g, o.cout, P, 4096L,
                         n->grad + o.b_off);
      EKM_CHECK(hipGetLastError());
    }
    if (i == 0) continue;  // the input needs no gradient
    ConvBwdD a;
    a.gy = Y.g;
    a.cout = o.cout;
    a.ho = Y.h;
    a.wo = Y.w;
    a.stride = o.stride;
    a.hi = A.h;
    a.wi = A.w;
    a.n = nb;
    a.w = o.wb;
    a.kbpad = o.kbpad;
    a.kb = 9 * o.cout;
    a.ga = A.g;
    a.gb = B ? B->g : nullptr;
    a.ma = A.relu ? A.a : nullptr;
    a.mb = (B && B->relu) ? B->a : nullptr;
    a.ca = A.c;
    a.cb = B ? B->c : 0;
    a.acc_a = written[o.ina];
    a.acc_b = B ? written[o.inb] : 0;
    const long Pi = (long)nb * A.h * A.w;
    const dim3 g(nblk(Pi, 64));
    switch (o.cbpad / 16) {
      case 1: hipLaunchKernelGGL(k_conv3_bwdd<1>, g, dim3(256), 0, st, a); break;
      case 2: hipLaunchKernelGGL(k_conv3_bwdd<2>, g, dim3(256), 0, st, a); break;
      case 3: hipLaunchKernelGGL(k_conv3_bwdd<3>, g, dim3(256), 0, st, a); break;
      case 4: hipLaunchKernelGGL(k_conv3_bwdd<4>, g, dim3(256), 0, st, a); break;
      case 5: hipLaunchKernelGGL(k_conv3_bwdd<5>, g, dim3(256), 0, st, a); break;
      case 6: hipLaunchKernelGGL(k_conv3_bwdd<6>, g, dim3(256), 0, st, a); break;
      default: hipLaunchKernelGGL(k_conv3_bwdd<7>, g, dim3(256), 0, st, a); break;
    }
    EKM_CHECK(hipGetLastError());
    written[o.ina] = 1;
    if (B) written[o.inb] = 1;
  }
  return EKM_OK;
}

void wsnet_dims(const ekm_wsnet* net, int* nz2, int* ny2, int* bmax) {
  *nz2 = net->w[0];
  *ny2 = net->h[0];
  *bmax = net->bmax;
}

int wsnet_guess(ekm_wsnet* n, const double* rhs, const double* prev, const double* ratio, const double* kx2,
                int nb, double* out, hipStream_t st) {
  if (nb < 1 || nb > n->bmax) return EKM_ERR_ARG;
  if (ensure_packed(n, st) != EKM_OK) return EKM_ERR_HIP;
  const long plane = (long)n->h[0] * n->w[0];
  hipLaunchKernelGGL(k_scale, dim3(nb), dim3(256), 0, st, rhs, plane, n->scale);
  EKM_CHECK(hipGetLastError());
  const long N1 = (long)nb * n->h[1] * n->w[1];
  hipLaunchKernelGGL(k_prep, dim3(nblk(N1, 256)), dim3(256), 0, st, rhs, prev, ratio, kx2, n->scale, nb, n->h[0],
                     n->w[0], n->sigma, n->kx_max, n->T[0].a);
  EKM_CHECK(hipGetLastError());
  const int e = run_forward(n, nb, st);
  if (e != EKM_OK) return e;
  hipLaunchKernelGGL(k_guess_out, dim3(nblk((size_t)nb * plane, 256)), dim3(256), 0, st, n->T[n->ops.back().out].a,
                     prev, ratio, n->scale, nb, n->h[0], n->w[0], n->sigma, out);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int ekm_wsnet_create(int nz2, int ny2, int bmax, int trainable, ekm_wsnet** out) {
  if (!out || bmax < 1) return EKM_ERR_ARG;
  *out = nullptr;
  // five levels below the plane, odd extents down to level 4 (so that each
  // prolongation 2n-1 lands on the finer level): 32 divides the plane's n-1
  if (nz2 < 33 || ny2 < 33 || (nz2 - 1) % 32 || (ny2 - 1) % 32) return EKM_ERR_UNSUPPORTED;
  ekm_wsnet* n = new ekm_wsnet();
  n->bmax = bmax;
  n->train = trainable != 0;
  n->h[0] = ny2;
  n->w[0] = nz2;
  for (int l = 1; l < 6; ++l) {
    n->h[l] = (n->h[l - 1] + 1) / 2;
    n->w[l] = (n->w[l - 1] + 1) / 2;
  }
  build_ops(n);
  auto fail = [&](int e) {
    net_free(n);
    delete n;
    return e;
  };
  // parameters: per trainable layer W [cout][cin][taps], then b [cout]
  size_t off = 0, packn = 0;
  for (WsOp& o : n->ops) {
    if (o.kind == 1) continue;
    const int taps = o.kind == 0 ? 9 : 1;
    o.w_off = off;
    off += (size_t)o.cout * o.cin * taps;
    o.b_off = off;
    off += o.cout;
    if (o.kind == 0) {
      o.cpad = (o.cout + 15) / 16 * 16;
      o.kpad = (9 * o.cin + 15) / 16 * 16;
      o.cbpad = (o.cin + 15) / 16 * 16;
      o.kbpad = (9 * o.cout + 15) / 16 * 16;
      packn += (size_t)o.cpad * o.kpad + (size_t)o.cbpad * o.kbpad;
      ++n->npack;
    }
  }
  n->np = off;
  if (hipMalloc(&n->theta, n->np * sizeof(float)) != hipSuccess) return fail(EKM_ERR_HIP);
  if (hipMalloc(&n->wpack, packn * sizeof(unsigned short)) != hipSuccess) return fail(EKM_ERR_HIP);
  std::vector<PackDesc> pd;
  size_t po = 0;
  for (WsOp& o : n->ops) {
    if (o.kind != 0) continue;
    o.wf = n->wpack + po;
    po += (size_t)o.cpad * o.kpad;
    o.wb = n->wpack + po;
    po += (size_t)o.cbpad * o.kbpad;
    pd.push_back(PackDesc{n->theta + o.w_off, o.wf, o.wb, o.cin, o.cout, o.cpad, o.kpad, o.cbpad, o.kbpad});
  }
  if (hipMalloc(&n->pd, pd.size() * sizeof(PackDesc)) != hipSuccess) return fail(EKM_ERR_HIP);
  if (hipMemcpy(n->pd, pd.data(), pd.size() * sizeof(PackDesc), hipMemcpyHostToDevice) != hipSuccess)
    return fail(EKM_ERR_HIP);
  for (WsTensor& t : n->T) {
    const size_t e = (size_t)bmax * t.h * t.w * t.c;
    if (hipMalloc(&t.a, e * sizeof(float)) != hipSuccess) return fail(EKM_ERR_HIP);
    if (n->train && hipMalloc(&t.g, e * sizeof(float)) != hipSuccess) return fail(EKM_ERR_HIP);
  }
  if (hipMalloc(&n->scale, bmax * sizeof(double)) != hipSuccess) return fail(EKM_ERR_HIP);
  if (n->train) {
    float** f[] = {&n->grad, &n->m, &n->v};
    for (float** q : f) {
      if (hipMalloc(q, n->np * sizeof(float)) != hipSuccess) return fail(EKM_ERR_HIP);
      if (hipMemset(*q, 0, n->np * sizeof(float)) != hipSuccess) return fail(EKM_ERR_HIP);
    }
    if (hipMalloc(&n->lossbuf, 2 * (size_t)bmax * sizeof(float)) != hipSuccess) return fail(EKM_ERR_HIP);
    n->host_nd.resize(2 * (size_t)bmax);
  }
  n->x1 = n->T[0].a;
  // He initialisation, seed 0 (ekm_wsnet_set_params replaces it)
  std::vector<float> th(n->np, 0.f);
  std::mt19937_64 rng(0);
  std::normal_distribution<float> nd(0.f, 1.f);
  for (const WsOp& o : n->ops) {
    if (o.kind == 1) continue;
    const int taps = o.kind == 0 ? 9 : 1;
    const float sd = std::sqrt(2.f / (float)(o.cin * taps));
    for (size_t i = 0; i < (size_t)o.cout * o.cin * taps; ++i) th[o.w_off + i] = sd * nd(rng);
  }
  if (hipMemcpy(n->theta, th.data(), n->np * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return fail(EKM_ERR_HIP);
  *out = n;
  return EKM_OK;
}

int ekm_wsnet_destroy(ekm_wsnet* n) {
  if (!n) return EKM_OK;
  (void)hipDeviceSynchronize();
  net_free(n);
  delete n;
  return EKM_OK;
}

long ekm_wsnet_nparams(const ekm_wsnet* n) { return n ? (long)n->np : 0; }

int ekm_wsnet_set_params(ekm_wsnet* n, const float* host) {
  EKM_CHECK(hipMemcpy(n->theta, host, n->np * sizeof(float), hipMemcpyHostToDevice));
  n->packed = false;
  return EKM_OK;
}

int ekm_wsnet_get_params(ekm_wsnet* n, float* host) {
  EKM_CHECK(hipDeviceSynchronize());
  EKM_CHECK(hipMemcpy(host, n->theta, n->np * sizeof(float), hipMemcpyDeviceToHost));
  return EKM_OK;
}

int ekm_wsnet_set_norm(ekm_wsnet* n, double sigma, double kx_max) {
  if (!(sigma > 0.0) || !(kx_max > 0.0)) return EKM_ERR_ARG;
  n->sigma = sigma;
  n->kx_max = kx_max;
  return EKM_OK;
}

int ekm_wsnet_get_norm(const ekm_wsnet* n, double* sigma, double* kx_max) {
  *sigma = n->sigma;
  *kx_max = n->kx_max;
  return EKM_OK;
}

int ekm_wsnet_prepare(ekm_wsnet* n, const double* rhs, const double* prev, const double* ratio, const double* kx2,
                      int nb, float* x1, double* scale, void* stream) {
  // caller buffers only: any number of planes
  if (nb < 1) return EKM_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const long plane = (long)n->h[0] * n->w[0];
  hipLaunchKernelGGL(k_scale, dim3(nb), dim3(256), 0, st, rhs, plane, scale);
  EKM_CHECK(hipGetLastError());
  const long N1 = (long)nb * n->h[1] * n->w[1];
  hipLaunchKernelGGL(k_prep, dim3(nblk(N1, 256)), dim3(256), 0, st, rhs, prev, ratio, kx2, scale, nb, n->h[0],
                     n->w[0], n->sigma, n->kx_max, x1);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}

int ekm_wsnet_target(ekm_wsnet* n, const double* phi, const double* prev, const double* ratio, const double* scale,
                     int nb, float* t0, void* stream) {
  if (nb < 1) return EKM_ERR_ARG;
  const long plane = (long)n->h[0] * n->w[0];
  hipLaunchKernelGGL(k_target, dim3(nblk((size_t)nb * plane, 256)), dim3(256), 0, (hipStream_t)stream, phi, prev,
                     ratio, scale, nb, plane, n->sigma, t0);
  EKM_CHECK(hipGetLastError());
  return EKM_OK;
}

int ekm_wsnet_forward(ekm_wsnet* n, const float* x1, int nb, float* y1, void* stream) {
  if (nb < 1 || nb > n->bmax) return EKM_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const size_t nx = (size_t)nb * n->h[1] * n->w[1] * kCin0;
  if (x1 != n->T[0].a) EKM_CHECK(hipMemcpyAsync(n->T[0].a, x1, nx * sizeof(float), hipMemcpyDeviceToDevice, st));
  if (ensure_packed(n, st) != EKM_OK) return EKM_ERR_HIP;
  const int e = run_forward(n, nb, st);
  if (e != EKM_OK) return e;
  if (y1)
    EKM_CHECK(hipMemcpyAsync(y1, n->T[n->ops.back().out].a, (size_t)nb * n->h[1] * n->w[1] * sizeof(float),
                             hipMemcpyDeviceToDevice, st));
  return EKM_OK;
}

static int grad_pass(ekm_wsnet* n, const float* x1, const float* t0, int nb, float* loss, hipStream_t st) {
  if (!n->train) return EKM_ERR_ARG;
  {
    const int e = ekm_wsnet_forward(n, x1, nb, nullptr, st);
    if (e != EKM_OK) return e;
  }
  {
    const int e = run_backward(n, t0, nb, st);
    if (e != EKM_OK) return e;
  }
  if (loss) {
    EKM_CHECK(hipMemcpyAsync(n->host_nd.data(), n->lossbuf, 2 * (size_t)nb * sizeof(float), hipMemcpyDeviceToHost,
                             st));
    EKM_CHECK(hipStreamSynchronize(st));
    double acc = 0.0;
    for (int s = 0; s < nb; ++s)
      if (n->host_nd[2 * s + 1] > 0.f) acc += (double)n->host_nd[2 * s] / (double)n->host_nd[2 * s + 1];
    *loss = (float)(acc / nb);
  }
  return EKM_OK;
}

int ekm_wsnet_gradient(ekm_wsnet* n, const float* x1, const float* t0, int nb, float* grad_host, float* loss) {
  const int e = grad_pass(n, x1, t0, nb, loss, nullptr);
  if (e != EKM_OK) return e;
  EKM_CHECK(hipDeviceSynchronize());
  if (grad_host) EKM_CHECK(hipMemcpy(grad_host, n->grad, n->np * sizeof(float), hipMemcpyDeviceToHost));
  return EKM_OK;
}

int ekm_wsnet_train_step(ekm_wsnet* n, const float* x1, const float* t0, int nb, float lr, float* loss, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  const int e = grad_pass(n, x1, t0, nb, loss, st);
  if (e != EKM_OK) return e;
  n->adam_t += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float lr_t = lr * std::sqrt(1.f - std::pow(b2, (float)n->adam_t)) / (1.f - std::pow(b1, (float)n->adam_t));
  hipLaunchKernelGGL(k_adam, dim3(nblk(n->np, 256)), dim3(256), 0, st, n->theta, n->m, n->v, n->grad, (long)n->np,
                     lr_t, b1, b2, 1e-8f);
  EKM_CHECK(hipGetLastError());
  n->packed = false;
  return EKM_OK;
}

int ekm_wsnet_guess(ekm_wsnet* n, const double* rhs, const double* prev, const double* ratio, const double* kx2,
                    int nb, double* out,
                    void* stream) {
  return wsnet_guess(n, rhs, prev, ratio, kx2, nb, out, (hipStream_t)stream);
}

}  // extern "C"
