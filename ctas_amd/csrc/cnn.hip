// Learned front end of the rotation-axis search: a small BF16 CNN that
// predicts the centre offset, and a confidence, from a sinogram downsampled
// to 180 angles x 256 detector bins plus its copy turned by 180 degrees.
//
// Network (245,554 parameters): four 3x3 convolutions (2->32, 32->32, 32->64,
// 64->64, padding 1, leaky ReLU (slope 1/64), 2x2 max pooling after each) on
// 2x180x256, then a 11264->16 fully connected layer with leaky ReLU and a
// 16->2 output (offset in
// downsampled detector bins, confidence logit). Activations and weights are
// bfloat16; every product runs on the matrix cores (v_mfma_f32_16x16x32_bf16,
// float32 accumulation) through one implicit-GEMM kernel whose operands are
// gathered by small functors: convolution forward (W x im2col(x)), backward
// data (W^T x col(g) with the mirrored stencil), backward weights
// (g x im2col(x)^T, split over the pixel dimension and reduced in a fixed
// order) and the fully connected layers. Training is Adam on float32 master
// weights; the bfloat16 copies feed the next step.
//
// Training data are made by the engine: seeded ellipse phantoms (plus small
// dense inclusions) projected analytically on the GPU with photon noise,
// then reduced to the network's input on the GPU.
//
// Searches that use the network:
// * ctas_axis_find: the sweep of trial centres on a fixed grid (the
//   AverageGradient score of lct.cpp:872-915 at every grid offset, argmin)
//   is narrowed to a window around the network's estimate that grows until
//   its lowest score is interior; below the confidence threshold the whole
//   grid is swept. Without a network it is the plain sweep.
// * ctas_raxis_cnn: raxis algorithm 1 (lct.cpp:1133-1212) with its 8-px
//   sweep replaced by the estimate and improveMe (lct.cpp:1108-1129) run
//   from the bracket +-4 px around it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "ctas_internal.hpp"

namespace ctas {
namespace cnn {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t bf16_rne(float v) {
  uint32_t u = __float_as_uint(v);
  if ((u & 0x7f800000u) == 0x7f800000u) return static_cast<uint16_t>((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf16_f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

constexpr int kFeatA = 180, kFeatW = 256;  // feature map: angles x detector bins
// Slope of the leaky ReLU on the negative side (a power of two, so scaled
// gradients round like the unscaled ones).
constexpr float kLeak = 0.015625f;

__device__ __forceinline__ float leaky(float v) { return v > 0.0f ? v : kLeak * v; }
constexpr int kBT = 64;                    // GEMM block tile (M and N)

// ---- generic GEMM: C[m][n] = sum_k A(m, k) B(k, n) -------------------------
// Block: 256 threads = 4 waves, 64x64 C tile, each wave 32x32 as 2x2 tiles
// of gfx950's v_mfma_f32_16x16x32_bf16. K advances by 32 through LDS. Operand
// layout of that instruction: lane l holds A[l%16][8(l/16)+j] and
// B[8(l/16)+j][l%16] (j = 0..7, eight bf16 in four registers); the result
// D[4(l/16)+r][l%16] (r = 0..3). Both tiles sit in LDS with K contiguous
// ([row][k] for A, [col][k] for B, rows padded to 40 elements), so that a
// lane's eight operands are one 16-byte read; the operand functors gather
// eight consecutive k of one row or column per thread, written as one
// 16-byte store.
// gridDim.z splits K into ranges of k_chunk; each split writes its own
// partial (the store functor gets the split index).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
constexpr int kKS = 32, kKP = kKS + 8;  // K per step, padded LDS row

template <class FA, class FB, class FC>
__global__ __launch_bounds__(256) void gemm_kernel(FA fa, FB fb, FC fc, int M, int N, int K, int k_chunk) {
  __shared__ __align__(16) uint16_t As[kBT][kKP];
  __shared__ __align__(16) uint16_t Bs[kBT][kKP];
  const int tid = threadIdx.x;
  const int m0 = blockIdx.y * kBT, n0 = blockIdx.x * kBT;
  const int kb = blockIdx.z * k_chunk, ke = min(K, kb + k_chunk);
  const int wave = tid >> 6, lane = tid & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int si = tid >> 2, sk = (tid & 3) * 8;  // staging: row / column si, k offset sk
  f4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += kKS) {
    u16x8 va, vb;
    const int m = m0 + si, n = n0 + si;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = k0 + sk + j;
      va[j] = (m < M && k < ke) ? fa(m, k) : static_cast<uint16_t>(0);
      vb[j] = (n < N && k < ke) ? fb(k, n) : static_cast<uint16_t>(0);
    }
    *reinterpret_cast<u16x8 *>(&As[si][sk]) = va;
    *reinterpret_cast<u16x8 *>(&Bs[si][sk]) = vb;
    __syncthreads();
    const int r16 = lane & 15, k8 = (lane >> 4) * 8;
    bf16x8 a[2], b[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      a[t] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(&As[wm + 16 * t + r16][k8]));
      b[t] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(&Bs[wn + 16 * t + r16][k8]));
    }
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
      for (int ni = 0; ni < 2; ++ni)
        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mi], b[ni], acc[mi][ni], 0, 0, 0);
    __syncthreads();
  }
  const int col = lane & 15, row4 = (lane >> 4) * 4;
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int ni = 0; ni < 2; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * mi + row4 + r, n = n0 + wn + 16 * ni + col;
        if (m < M && n < N) fc(m, n, acc[mi][ni][r], blockIdx.z);
      }
}

template <class FA, class FB, class FC>
hipError_t gemm(FA fa, FB fb, FC fc, int M, int N, int K, hipStream_t s, int splits = 1) {
  int chunk = (K + splits - 1) / splits;
  chunk = (chunk + kKS - 1) & ~(kKS - 1);
  splits = (K + chunk - 1) / chunk;
  const dim3 grid((N + kBT - 1) / kBT, (M + kBT - 1) / kBT, splits);
  hipLaunchKernelGGL((gemm_kernel<FA, FB, FC>), grid, dim3(256), 0, s, fa, fb, fc, M, N, K, chunk);
  return hipGetLastError();
}

// ---- operand functors ---------------------------------------------------------
struct DenseA {  // A[m][k] bf16, row stride ld
  const uint16_t *p;
  int ld;
  __device__ uint16_t operator()(int m, int k) const { return p[static_cast<size_t>(m) * ld + k]; }
};
struct DenseAT {  // A(m, k) = P[k][m] bf16 (transposed), row stride ld
  const uint16_t *p;
  int ld;
  __device__ uint16_t operator()(int m, int k) const { return p[static_cast<size_t>(k) * ld + m]; }
};
struct Im2col {  // B(k, n) = x[b][ci][y+dy][x+dx], k = ci*9 + r, n = (b*H + y)*W + x
  const uint16_t *x;
  int C, H, W;
  __device__ uint16_t operator()(int k, int n) const {
    const int ci = k / 9, r = k - ci * 9;
    const int xx = n % W, t = n / W, yy = t % H, b = t / H;
    const int y2 = yy + r / 3 - 1, x2 = xx + r % 3 - 1;
    if (y2 < 0 || y2 >= H || x2 < 0 || x2 >= W) return 0;
    return x[((static_cast<size_t>(b) * C + ci) * H + y2) * W + x2];
  }
};
struct Im2colT {  // B(n, k) = im2col(k, n): the pixel dimension is the reduction
  Im2col f;
  __device__ uint16_t operator()(int n, int k) const { return f(k, n); }
};
// gradient of the pre-activation: g = d(post) * (post > 0 ? 1 : kLeak), as bf16
__device__ __forceinline__ uint16_t relu_grad(const float *d, const uint16_t *post, size_t i) {
  return bf16_rne(bf16_f(post[i]) > 0.0f ? d[i] : kLeak * d[i]);
}
struct GradA {  // A(co, n) = g[b][co][y][x], n = (b*H + y)*W + x
  const float *d;
  const uint16_t *post;
  int C, H, W;
  __device__ uint16_t operator()(int co, int n) const {
    const int hw = H * W, b = n / hw, p = n - b * hw;
    const size_t i = (static_cast<size_t>(b) * C + co) * hw + p;
    return relu_grad(d, post, i);
  }
};
struct WT {  // A(ci, k = co*9 + r) = W[co][ci*9 + r]
  const uint16_t *w;
  int Cin;
  __device__ uint16_t operator()(int ci, int k) const {
    const int co = k / 9, r = k - co * 9;
    return w[static_cast<size_t>(co) * Cin * 9 + ci * 9 + r];
  }
};
struct ColG {  // B(k = co*9 + r, n) = g[b][co][y - dy][x - dx]
  const float *d;
  const uint16_t *post;
  int C, H, W;
  __device__ uint16_t operator()(int k, int n) const {
    const int co = k / 9, r = k - co * 9;
    const int xx = n % W, t = n / W, yy = t % H, b = t / H;
    const int y2 = yy - (r / 3 - 1), x2 = xx - (r % 3 - 1);
    if (y2 < 0 || y2 >= H || x2 < 0 || x2 >= W) return 0;
    const size_t i = ((static_cast<size_t>(b) * C + co) * H + y2) * W + x2;
    return relu_grad(d, post, i);
  }
};
struct GradFlatA {  // A(o, b) = g[b][o] (fp32 gradient, ReLU-masked by post when given)
  const float *d;
  const uint16_t *post;
  int O;
  __device__ uint16_t operator()(int o, int b) const {
    const size_t i = static_cast<size_t>(b) * O + o;
    return post ? relu_grad(d, post, i) : bf16_rne(d[i]);
  }
};
struct GradFlatB {  // B(o, b) = g[b][o]
  GradFlatA f;
  __device__ uint16_t operator()(int o, int b) const { return f(o, b); }
};
struct ActB {  // B(k, n) = act[n][k] (flattened activations per sample)
  const uint16_t *p;
  int F;
  __device__ uint16_t operator()(int k, int n) const { return p[static_cast<size_t>(n) * F + k]; }
};
struct ActBT {  // B(n, k) = act[n][k]
  const uint16_t *p;
  int F;
  __device__ uint16_t operator()(int n, int k) const { return p[static_cast<size_t>(n) * F + k]; }
};

// ---- store functors -------------------------------------------------------------
struct StoreConvAct {  // y[b][co][y][x] = leaky(acc + bias[co]) as bf16
  uint16_t *y;
  const float *bias;
  int C, H, W;
  __device__ void operator()(int co, int n, float v, int) const {
    const int hw = H * W, b = n / hw, p = n - b * hw;
    y[(static_cast<size_t>(b) * C + co) * hw + p] = bf16_rne(leaky(v + bias[co]));
  }
};
struct StoreDX {  // dx[b][ci][y][x] (fp32)
  float *dx;
  int C, H, W;
  __device__ void operator()(int ci, int n, float v, int) const {
    const int hw = H * W, b = n / hw, p = n - b * hw;
    dx[(static_cast<size_t>(b) * C + ci) * hw + p] = v;
  }
};
struct StorePartial {  // ws[split][m][n]
  float *ws;
  int M, N;
  __device__ void operator()(int m, int n, float v, int z) const {
    ws[(static_cast<size_t>(z) * M + m) * N + n] = v;
  }
};
struct StoreFC {  // out[b][o] = acc + bias[o] (leaky ReLU optional), bf16 and/or fp32
  uint16_t *y16;
  float *y32;
  const float *bias;
  int O;
  bool relu;
  __device__ void operator()(int o, int b, float v, int) const {
    float r = v + bias[o];
    if (relu) r = leaky(r);
    const size_t i = static_cast<size_t>(b) * O + o;
    if (y16) y16[i] = bf16_rne(r);
    if (y32) y32[i] = r;
  }
};
struct StoreFlatDX {  // dx[b][f] fp32 from C(f, b)
  float *dx;
  int F;
  __device__ void operator()(int f, int b, float v, int) const { dx[static_cast<size_t>(b) * F + f] = v; }
};

// ---- elementwise kernels -------------------------------------------------------------
__global__ void maxpool_fwd(const uint16_t *__restrict__ x, uint16_t *__restrict__ y, int BC, int H, int W) {
  const int Ho = H / 2, Wo = W / 2;
  const size_t total = static_cast<size_t>(BC) * Ho * Wo;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < total;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int xo = static_cast<int>(i % Wo), yo = static_cast<int>((i / Wo) % Ho);
    const size_t bc = i / (static_cast<size_t>(Wo) * Ho);
    const uint16_t *p = x + (bc * H + 2 * yo) * W + 2 * xo;
    float m = bf16_f(p[0]);
    m = fmaxf(m, bf16_f(p[1]));
    m = fmaxf(m, bf16_f(p[W]));
    m = fmaxf(m, bf16_f(p[W + 1]));
    y[i] = bf16_rne(m);
  }
}

// gradient to the first maximum of each window (rows/columns beyond the
// pooled extent get zero)
__global__ void maxpool_bwd(const uint16_t *__restrict__ x, const float *__restrict__ dy, float *__restrict__ dx,
                            int BC, int H, int W) {
  const int Ho = H / 2, Wo = W / 2;
  const size_t total = static_cast<size_t>(BC) * H * W;
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < total;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int xx = static_cast<int>(i % W), yy = static_cast<int>((i / W) % H);
    const size_t bc = i / (static_cast<size_t>(W) * H);
    const int yo = yy / 2, xo = xx / 2;
    float g = 0.0f;
    if (yo < Ho && xo < Wo) {
      const uint16_t *p = x + (bc * H + 2 * yo) * W + 2 * xo;
      const float v[4] = {bf16_f(p[0]), bf16_f(p[1]), bf16_f(p[W]), bf16_f(p[W + 1])};
      int am = 0;
      for (int k = 1; k < 4; ++k)
        if (v[k] > v[am]) am = k;
      const int mine = (yy - 2 * yo) * 2 + (xx - 2 * xo);
      if (mine == am) g = dy[(bc * Ho + yo) * Wo + xo];
    }
    dx[i] = g;
  }
}

// Sum of the split-K partials (fixed order) into the weight gradient.
// Sum over the splits of a K-split product, in a fixed order: 16 outputs per
// block, each summed by 16 threads over interleaved splits (z = g, g + 16,
// ...) and their 16 partials added in order.
__global__ __launch_bounds__(256) void reduce_splits(const float *__restrict__ ws, float *__restrict__ out,
                                                     int splits, int MN) {
  const int o = threadIdx.x & 15, g = threadIdx.x >> 4;
  const int i = blockIdx.x * 16 + o;
  float s = 0.0f;
  if (i < MN)
    for (int z = g; z < splits; z += 16) s += ws[static_cast<size_t>(z) * MN + i];
  __shared__ float part[16][17];
  part[g][o] = s;
  __syncthreads();
  if (g == 0 && i < MN) {
    float t = part[0][o];
    for (int k = 1; k < 16; ++k) t += part[k][o];
    out[i] = t;
  }
}

// Epilogue of a split forward product: y[b][m][p] = bf16(leaky(sum over
// splits of ws[z][m][n] + bias[m])), n = b*HW + p.
__global__ void finish_act(const float *__restrict__ ws, int splits, int M, int N, const float *__restrict__ bias,
                           uint16_t *__restrict__ y, int HW) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * N) return;
  const int m = i / N, n = i - m * N;
  float v = 0.0f;
  for (int z = 0; z < splits; ++z) v += ws[static_cast<size_t>(z) * M * N + i];
  const int b = n / HW, p = n - b * HW;
  y[(static_cast<size_t>(b) * M + m) * HW + p] = bf16_rne(leaky(v + bias[m]));
}

// Bias gradient: sum of the slope-scaled gradient over batch and pixels (one
// block per channel, fixed-order tree reduction).
__global__ void bias_grad(const float *__restrict__ d, const uint16_t *__restrict__ post, float *__restrict__ db,
                          int B, int C, int HW) {
  const int co = blockIdx.x;
  float s = 0.0f;
  for (int b = 0; b < B; ++b)
    for (int p = threadIdx.x; p < HW; p += blockDim.x) {
      const size_t i = (static_cast<size_t>(b) * C + co) * HW + p;
      s += (!post || bf16_f(post[i]) > 0.0f) ? d[i] : kLeak * d[i];
    }
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[co] = red[0];
}

// Bias gradient of a convolution layer in two fixed-order steps that fill
// the chip: block (co, b) sums the slope-scaled gradient over the pixels of
// one channel of one sample (contiguous, coalesced) into part[co][b], then
// bias_grad_sum adds the samples in order.
__global__ __launch_bounds__(256) void bias_grad_part(const float *__restrict__ d, const uint16_t *__restrict__ post,
                                                      float *__restrict__ part, int B, int C, int HW) {
  const int co = blockIdx.x, b = blockIdx.y;
  const size_t base = (static_cast<size_t>(b) * C + co) * HW;
  float s = 0.0f;
  for (int p = threadIdx.x; p < HW; p += blockDim.x) {
    const size_t i = base + p;
    s += bf16_f(post[i]) > 0.0f ? d[i] : kLeak * d[i];
  }
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[static_cast<size_t>(co) * B + b] = red[0];
}

__global__ void bias_grad_sum(const float *__restrict__ part, float *__restrict__ db, int B, int C) {
  const int co = blockIdx.x * blockDim.x + threadIdx.x;
  if (co >= C) return;
  float s = 0.0f;
  for (int b = 0; b < B; ++b) s += part[static_cast<size_t>(co) * B + b];
  db[co] = s;
}

// Loss over targets t[b] = {offset, tolerance} (output units):
// Huber(offset error, delta 0.25) + 0.2 * BCE(confidence logit, |error| < tolerance).
// g_out = dL/d(out) / B; loss = mean over the batch.
__global__ void loss_kernel(const float *__restrict__ out, const float *__restrict__ tgt, float *__restrict__ g,
                            float *__restrict__ loss, int B) {
  __shared__ float red[256];
  float l = 0.0f;
  for (int b = threadIdx.x; b < B; b += blockDim.x) {
    const float e = out[2 * b] - tgt[2 * b];
    const float delta = 0.25f;
    const float ae = fabsf(e);
    const float hub = ae < delta ? 0.5f * e * e / delta : ae - 0.5f * delta;
    const float dh = ae < delta ? e / delta : (e > 0 ? 1.0f : -1.0f);
    const float y = ae < tgt[2 * b + 1] ? 1.0f : 0.0f;
    const float z = out[2 * b + 1];
    const float p = 1.0f / (1.0f + expf(-z));
    const float bce = fmaxf(z, 0.0f) - z * y + log1pf(expf(-fabsf(z)));
    l += hub + 0.2f * bce;
    g[2 * b] = dh / B;
    g[2 * b + 1] = 0.2f * (p - y) / B;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0] / B;
}

// Global gradient norm: per-block sums of squares, then one block sums them
// in a fixed order and writes the clipping factor min(1, clip / norm) (1
// when clip <= 0) and the norm.
constexpr int kNormBlocks = 256;

__global__ __launch_bounds__(256) void sumsq_kernel(const float *__restrict__ g, int n, float *__restrict__ part) {
  __shared__ float red[256];
  float s = 0.0f;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += kNormBlocks * 256) s += g[i] * g[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void clip_factor_kernel(const float *__restrict__ part, float clip,
                                                          float *__restrict__ out) {
  __shared__ float red[256];
  red[threadIdx.x] = threadIdx.x < kNormBlocks ? part[threadIdx.x] : 0.0f;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float norm = sqrtf(red[0]);
    out[0] = (clip > 0.0f && norm > clip) ? clip / norm : 1.0f;
    out[1] = norm;
  }
}

__global__ void adam_kernel(float *__restrict__ p, uint16_t *__restrict__ p16, const float *__restrict__ g,
                            float *__restrict__ m, float *__restrict__ v, int n, float lr, float b1, float b2,
                            float c1, float c2, float eps, const float *__restrict__ scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] * scale[0];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  p16[i] = bf16_rne(pi);
}

__global__ void to_bf16(const float *__restrict__ p, uint16_t *__restrict__ q, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) q[i] = bf16_rne(p[i]);
}

// Sinogram [N][w] (transmission) -> features [2][180][256] bf16: the mean
// of -log(max(v, 1e-6)) over each cell of angle rows x detector columns,
// standardised per sample; channel 1 is the map turned by 180 degrees
// (detector and angle order reversed), which puts the projections that
// mirror each other across the 0/180 degree seam (P(theta + pi, s) =
// P(theta, -s)) in the same rows of the two channels, shifted by twice
// the axis offset. Grid (1, 180): one block per angle cell, one thread per
// detector bin.
__global__ __launch_bounds__(256) void features_kernel(const float *__restrict__ sino, int N, int w,
                                                       float *__restrict__ tmp) {
  const int a = blockIdx.y;
  const int x = threadIdx.x;
  const int r0 = static_cast<int>(static_cast<long long>(a) * N / kFeatA);
  const int r1 = max(r0 + 1, static_cast<int>(static_cast<long long>(a + 1) * N / kFeatA));
  const int c0 = static_cast<int>(static_cast<long long>(x) * w / kFeatW);
  const int c1 = max(c0 + 1, static_cast<int>(static_cast<long long>(x + 1) * w / kFeatW));
  float s = 0.0f;
  for (int r = r0; r < r1; ++r)
    for (int c = c0; c < c1; ++c) s += -logf(fmaxf(sino[static_cast<size_t>(r) * w + c], 1e-6f));
  tmp[a * kFeatW + x] = s / static_cast<float>((r1 - r0) * (c1 - c0));
}

// Mean and standard deviation of the 180 x 256 features, then the two
// standardised bf16 channels. One block of 1024 threads (a few float4 loads
// each), partial sums reduced in a fixed order.
__global__ __launch_bounds__(1024) void standardise_kernel(const float *__restrict__ tmp, uint16_t *__restrict__ feat) {
  __shared__ float red[1024], red2[1024];
  constexpr int n = kFeatA * kFeatW;
  static_assert(n % 4 == 0, "float4 loads");
  const float4 *t4 = reinterpret_cast<const float4 *>(tmp);
  float s = 0.0f, s2 = 0.0f;
  for (int i = threadIdx.x; i < n / 4; i += 1024) {
    const float4 v = t4[i];
    s += (v.x + v.y) + (v.z + v.w);
    s2 += (v.x * v.x + v.y * v.y) + (v.z * v.z + v.w * v.w);
  }
  red[threadIdx.x] = s;
  red2[threadIdx.x] = s2;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      red[threadIdx.x] += red[threadIdx.x + o];
      red2[threadIdx.x] += red2[threadIdx.x + o];
    }
    __syncthreads();
  }
  const float mean = red[0] / n;
  const float var = fmaxf(red2[0] / n - mean * mean, 1e-12f);
  const float inv = 1.0f / sqrtf(var);
  for (int i = threadIdx.x; i < n; i += 1024) {
    const int a = i / kFeatW, x = i - a * kFeatW;
    const float v = (tmp[i] - mean) * inv;
    feat[i] = bf16_rne(v);
    feat[n + (kFeatA - 1 - a) * kFeatW + (kFeatW - 1 - x)] = bf16_rne(v);
  }
}

hipError_t launch_features(const float *d_sino, int N, int w, float *d_tmp, uint16_t *d_x, hipStream_t s) {
  hipLaunchKernelGGL(features_kernel, dim3(1, kFeatA), dim3(kFeatW), 0, s, d_sino, N, w, d_tmp);
  hipLaunchKernelGGL(standardise_kernel, dim3(1), dim3(1024), 0, s, d_tmp, d_x);
  return hipGetLastError();
}

// ---- the network -----------------------------------------------------------------
struct Layer {
  int cin, cout, H, W;  // conv: input channels, output channels, spatial size
};
constexpr Layer kConv[4] = {{2, 32, 180, 256}, {32, 32, 90, 128}, {32, 64, 45, 64}, {64, 64, 22, 32}};
constexpr int kF = 64 * 11 * 16, kH1 = 16, kOut = 2;

struct Param {
  size_t off, n;
};

}  // namespace cnn
}  // namespace ctas

struct ctas_cnn {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<ctas::cnn::Param> prm;  // w1 b1 w2 b2 w3 b3 w4 b4 wf1 bf1 wf2 bf2
  size_t nparam = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *p16 = nullptr;
  int step = 0;
  int cap = 0;  // batch capacity of the activation buffers
  uint16_t *act[4] = {}, *pool[4] = {}, *h1 = nullptr;
  float *out = nullptr, *gout = nullptr, *gh1 = nullptr, *gf = nullptr, *gbig = nullptr, *gbig2 = nullptr;
  float *ws = nullptr;
  size_t ws_cap = 0;
  float *loss = nullptr;
  float *ftmp = nullptr;
  uint16_t *xin = nullptr;  // features of one sinogram (prediction)
  float *norm = nullptr;    // gradient-norm partials, clip factor, norm
};

namespace ctas {
namespace cnn {
namespace {

int alloc(void **p, size_t bytes) {
  if (hipMalloc(p, std::max<size_t>(bytes, 256)) != hipSuccess)
    return fail(CTAS_E_NOMEM, "CNN allocation of " + std::to_string(bytes) + " bytes failed");
  return CTAS_OK;
}

void release_batch(ctas_cnn *n) {
  for (int l = 0; l < 4; ++l) {
    if (n->act[l]) (void)hipFree(n->act[l]);
    if (n->pool[l]) (void)hipFree(n->pool[l]);
    n->act[l] = n->pool[l] = nullptr;
  }
  void *ptrs[] = {n->h1, n->out, n->gout, n->gh1, n->gf, n->gbig, n->gbig2};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  n->h1 = nullptr;
  n->out = n->gout = n->gh1 = n->gf = n->gbig = n->gbig2 = nullptr;
  n->cap = 0;
}

int ensure_batch(ctas_cnn *n, int B) {
  if (B <= n->cap) return CTAS_OK;
  CTAS_HIP_TRY(hipStreamSynchronize(n->stream));
  release_batch(n);
  size_t big = 0;
  for (int l = 0; l < 4; ++l) {
    const Layer &L = kConv[l];
    const size_t a = static_cast<size_t>(B) * L.cout * L.H * L.W;
    const size_t q = static_cast<size_t>(B) * L.cout * (L.H / 2) * (L.W / 2);
    if (int rc = alloc(reinterpret_cast<void **>(&n->act[l]), a * 2)) return rc;
    if (int rc = alloc(reinterpret_cast<void **>(&n->pool[l]), q * 2)) return rc;
    big = std::max(big, a);
    big = std::max(big, static_cast<size_t>(B) * L.cin * L.H * L.W);
  }
  if (int rc = alloc(reinterpret_cast<void **>(&n->h1), static_cast<size_t>(B) * kH1 * 2)) return rc;
  if (int rc = alloc(reinterpret_cast<void **>(&n->out), static_cast<size_t>(B) * kOut * 4)) return rc;
  if (int rc = alloc(reinterpret_cast<void **>(&n->gout), static_cast<size_t>(B) * kOut * 4)) return rc;
  if (int rc = alloc(reinterpret_cast<void **>(&n->gh1), static_cast<size_t>(B) * kH1 * 4)) return rc;
  if (int rc = alloc(reinterpret_cast<void **>(&n->gf), static_cast<size_t>(B) * kF * 4)) return rc;
  if (int rc = alloc(reinterpret_cast<void **>(&n->gbig), big * 4)) return rc;
  if (int rc = alloc(reinterpret_cast<void **>(&n->gbig2), big * 4)) return rc;
  n->cap = B;
  return CTAS_OK;
}

int ensure_ws(ctas_cnn *n, size_t floats) {
  if (floats <= n->ws_cap) return CTAS_OK;
  CTAS_HIP_TRY(hipStreamSynchronize(n->stream));
  if (n->ws) (void)hipFree(n->ws);
  n->ws = nullptr;
  if (int rc = alloc(reinterpret_cast<void **>(&n->ws), floats * 4)) return rc;
  n->ws_cap = floats;
  return CTAS_OK;
}

const uint16_t *W16(const ctas_cnn *n, int i) { return n->p16 + n->prm[i].off; }
const float *P32(const ctas_cnn *n, int i) { return n->p + n->prm[i].off; }
float *G32(const ctas_cnn *n, int i) { return n->g + n->prm[i].off; }

int blocks_for(size_t n) { return static_cast<int>(std::min<size_t>((n + 255) / 256, 65535)); }

// Splits of the reduction for a product whose C tiles alone would not fill
// the chip (inference on a single sinogram); 1 when they do.
int forward_splits(int M, int N, int K) {
  const int tiles = ((M + kBT - 1) / kBT) * ((N + kBT - 1) / kBT);
  if (tiles >= 256) return 1;
  return std::max(1, std::min((256 + tiles - 1) / tiles, K / 64));
}

// A product with bias + ReLU + bf16 epilogue: direct store when it fills
// the chip, else split over K into the workspace and finished in a fixed
// order.
template <class FA, class FB>
int product_act(ctas_cnn *n, FA fa, FB fb, int M, int Ncols, int K, const float *bias, uint16_t *y, int HW) {
  hipStream_t s = n->stream;
  const int splits = forward_splits(M, Ncols, K);
  if (splits == 1) {
    CTAS_HIP_TRY(gemm(fa, fb, StoreConvAct{y, bias, M, 1, HW}, M, Ncols, K, s));
    return CTAS_OK;
  }
  if (int rc = ensure_ws(n, static_cast<size_t>(splits) * M * Ncols)) return rc;
  int chunk = (K + splits - 1) / splits;
  chunk = (chunk + kKS - 1) & ~(kKS - 1);
  const int real = (K + chunk - 1) / chunk;
  CTAS_HIP_TRY(gemm(fa, fb, StorePartial{n->ws, M, Ncols}, M, Ncols, K, s, splits));
  hipLaunchKernelGGL(finish_act, dim3((M * Ncols + 255) / 256), dim3(256), 0, s, n->ws, real, M, Ncols, bias, y, HW);
  return CTAS_OK;
}

int forward(ctas_cnn *n, const uint16_t *x, int B) {
  hipStream_t s = n->stream;
  const uint16_t *in = x;
  for (int l = 0; l < 4; ++l) {
    const Layer &L = kConv[l];
    const int HW = L.H * L.W;
    if (int rc = product_act(n, DenseA{W16(n, 2 * l), L.cin * 9}, Im2col{in, L.cin, L.H, L.W}, L.cout, B * HW,
                             L.cin * 9, P32(n, 2 * l + 1), n->act[l], HW))
      return rc;
    const size_t q = static_cast<size_t>(B) * L.cout * (L.H / 2) * (L.W / 2);
    hipLaunchKernelGGL(maxpool_fwd, dim3(blocks_for(q)), dim3(256), 0, s, n->act[l], n->pool[l], B * L.cout, L.H, L.W);
    in = n->pool[l];
  }
  if (int rc = product_act(n, DenseA{W16(n, 8), kF}, ActB{n->pool[3], kF}, kH1, B, kF, P32(n, 9), n->h1, 1))
    return rc;
  CTAS_HIP_TRY(gemm(DenseA{W16(n, 10), kH1}, ActB{n->h1, kH1}, StoreFC{nullptr, n->out, P32(n, 11), kOut, false},
                    kOut, B, kH1, s));
  return CTAS_OK;
}

// weight gradient of a GEMM whose reduction runs over K (split, reduced in a
// fixed order)
template <class FA, class FB>
int weight_grad(ctas_cnn *n, FA fa, FB fb, int M, int Ncols, int K, float *dst) {
  // enough K splits that the few output tiles of a weight gradient still
  // put several workgroups on every CU
  const int splits = std::max(1, std::min(2048, K / 1024));
  if (int rc = ensure_ws(n, static_cast<size_t>(splits) * M * Ncols)) return rc;
  int chunk = (K + splits - 1) / splits;
  chunk = (chunk + kKS - 1) & ~(kKS - 1);
  const int real = (K + chunk - 1) / chunk;
  CTAS_HIP_TRY(gemm(fa, fb, StorePartial{n->ws, M, Ncols}, M, Ncols, K, n->stream, splits));
  hipLaunchKernelGGL(reduce_splits, dim3((M * Ncols + 15) / 16), dim3(256), 0, n->stream, n->ws, dst, real,
                     M * Ncols);
  return CTAS_OK;
}

int backward(ctas_cnn *n, const uint16_t *x, int B) {
  hipStream_t s = n->stream;
  // fc2: dW[o][k] = sum_b gout[b][o] h1[b][k]; dh1[b][k] = sum_o W[o][k] gout[b][o]
  if (int rc = weight_grad(n, GradFlatA{n->gout, nullptr, kOut}, ActBT{n->h1, kH1}, kOut, kH1, B, G32(n, 10)))
    return rc;
  hipLaunchKernelGGL(bias_grad, dim3(kOut), dim3(256), 0, s, n->gout, nullptr, G32(n, 11), B, kOut, 1);
  CTAS_HIP_TRY(gemm(DenseAT{W16(n, 10), kH1}, GradFlatB{GradFlatA{n->gout, nullptr, kOut}}, StoreFlatDX{n->gh1, kH1},
                    kH1, B, kOut, s));
  // fc1 (ReLU on h1)
  if (int rc = weight_grad(n, GradFlatA{n->gh1, n->h1, kH1}, ActBT{n->pool[3], kF}, kH1, kF, B, G32(n, 8))) return rc;
  hipLaunchKernelGGL(bias_grad, dim3(kH1), dim3(256), 0, s, n->gh1, n->h1, G32(n, 9), B, kH1, 1);
  CTAS_HIP_TRY(gemm(DenseAT{W16(n, 8), kF}, GradFlatB{GradFlatA{n->gh1, n->h1, kH1}}, StoreFlatDX{n->gf, kF}, kF, B,
                    kH1, s));
  // conv layers, last to first. gpool = gradient of the layer's pooled
  // output; gact (buffer gbig) = gradient of its post-activation; the
  // gradient of its input (= the previous pooled output) goes to gbig2.
  const float *gpool = n->gf;
  float *gact = n->gbig;
  for (int l = 3; l >= 0; --l) {
    const Layer &L = kConv[l];
    const int HW = L.H * L.W;
    const size_t a = static_cast<size_t>(B) * L.cout * HW;
    hipLaunchKernelGGL(maxpool_bwd, dim3(blocks_for(a)), dim3(256), 0, s, n->act[l], gpool, gact, B * L.cout, L.H,
                       L.W);
    const uint16_t *in = l ? n->pool[l - 1] : x;
    if (int rc = weight_grad(n, GradA{gact, n->act[l], L.cout, L.H, L.W}, Im2colT{Im2col{in, L.cin, L.H, L.W}},
                             L.cout, L.cin * 9, B * HW, G32(n, 2 * l)))
      return rc;
    // the workspace's split partials were reduced by weight_grad on this stream
    if (int rc = ensure_ws(n, static_cast<size_t>(L.cout) * B)) return rc;
    hipLaunchKernelGGL(bias_grad_part, dim3(L.cout, B), dim3(256), 0, s, gact, n->act[l], n->ws, B, L.cout, HW);
    hipLaunchKernelGGL(bias_grad_sum, dim3(1), dim3(64), 0, s, n->ws, G32(n, 2 * l + 1), B, L.cout);
    if (l > 0) {
      CTAS_HIP_TRY(gemm(WT{W16(n, 2 * l), L.cin}, ColG{gact, n->act[l], L.cout, L.H, L.W},
                        StoreDX{n->gbig2, L.cin, L.H, L.W}, L.cin, B * HW, L.cout * 9, s));
      gpool = n->gbig2;
    }
  }
  return CTAS_OK;
}

}  // namespace
}  // namespace cnn
}  // namespace ctas

using namespace ctas;
using namespace ctas::cnn;

extern "C" int ctas_cnn_create(ctas_cnn **pn, uint64_t seed) {
  if (!pn) return fail(CTAS_E_ARG, "null argument");
  *pn = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(CTAS_E_DEVICE, "no HIP device is available");
  std::unique_ptr<ctas_cnn> n(new ctas_cnn);
  CTAS_HIP_TRY(hipGetDevice(&n->device));
  CTAS_HIP_TRY(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
  std::vector<std::pair<size_t, float>> shapes;  // (count, init scale)
  for (int l = 0; l < 4; ++l) {
    const size_t fan_in = static_cast<size_t>(kConv[l].cin) * 9;
    shapes.push_back({kConv[l].cout * fan_in, std::sqrt(2.0f / fan_in)});
    shapes.push_back({static_cast<size_t>(kConv[l].cout), 0.0f});
  }
  shapes.push_back({static_cast<size_t>(kH1) * kF, std::sqrt(2.0f / kF)});
  shapes.push_back({static_cast<size_t>(kH1), 0.0f});
  shapes.push_back({static_cast<size_t>(kOut) * kH1, std::sqrt(1.0f / kH1)});
  shapes.push_back({static_cast<size_t>(kOut), 0.0f});
  size_t off = 0;
  for (auto &sh : shapes) {
    n->prm.push_back({off, sh.first});
    off += (sh.first + 63) & ~static_cast<size_t>(63);
  }
  n->nparam = off;
  std::vector<float> h(off, 0.0f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  for (size_t i = 0; i < shapes.size(); ++i)
    for (size_t k = 0; k < shapes[i].first; ++k) h[n->prm[i].off + k] = shapes[i].second * nd(rng);
  if (int rc = alloc(reinterpret_cast<void **>(&n->p), off * 4)) return rc;
  if (int rc = alloc(reinterpret_cast<void **>(&n->g), off * 4)) return rc;
  if (int rc = alloc(reinterpret_cast<void **>(&n->m), off * 4)) return rc;
  if (int rc = alloc(reinterpret_cast<void **>(&n->v), off * 4)) return rc;
  if (int rc = alloc(reinterpret_cast<void **>(&n->p16), off * 2)) return rc;
  if (int rc = alloc(reinterpret_cast<void **>(&n->loss), 256)) return rc;
  if (int rc = alloc(reinterpret_cast<void **>(&n->norm), sizeof(float) * (kNormBlocks + 2))) return rc;
  if (int rc = alloc(reinterpret_cast<void **>(&n->ftmp), sizeof(float) * kFeatA * kFeatW)) return rc;
  CTAS_HIP_TRY(hipMemcpy(n->p, h.data(), off * 4, hipMemcpyHostToDevice));
  std::vector<float> z(off, 0.0f);
  CTAS_HIP_TRY(hipMemcpy(n->m, z.data(), off * 4, hipMemcpyHostToDevice));
  CTAS_HIP_TRY(hipMemcpy(n->v, z.data(), off * 4, hipMemcpyHostToDevice));
  CTAS_HIP_TRY(hipMemcpy(n->g, z.data(), off * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(to_bf16, dim3((off + 255) / 256), dim3(256), 0, n->stream, n->p, n->p16, static_cast<int>(off));
  CTAS_HIP_TRY(hipStreamSynchronize(n->stream));
  *pn = n.release();
  return CTAS_OK;
}

extern "C" int ctas_cnn_destroy(ctas_cnn *n) {
  if (!n) return CTAS_OK;
  (void)hipStreamSynchronize(n->stream);
  release_batch(n);
  void *ptrs[] = {n->p, n->g, n->m, n->v, n->p16, n->ws, n->loss, n->ftmp, n->xin, n->norm};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  (void)hipStreamDestroy(n->stream);
  delete n;
  return CTAS_OK;
}

// Parameter layout: the 12 tensors w1 b1 w2 b2 w3 b3 w4 b4 wf1 bf1 wf2 bf2
// (conv weights [cout][cin][3][3], fc weights [out][in]) packed back to back.
extern "C" int ctas_cnn_num_params(const ctas_cnn *n, int *count) {
  if (!n || !count) return fail(CTAS_E_ARG, "null argument");
  size_t c = 0;
  for (const auto &p : n->prm) c += p.n;
  *count = static_cast<int>(c);
  return CTAS_OK;
}

static int copy_params(ctas_cnn *n, float *host, const float *dev, bool to_host) {
  CTAS_HIP_TRY(hipStreamSynchronize(n->stream));
  size_t at = 0;
  for (const auto &p : n->prm) {
    if (to_host)
      CTAS_HIP_TRY(hipMemcpy(host + at, dev + p.off, p.n * 4, hipMemcpyDeviceToHost));
    else
      CTAS_HIP_TRY(hipMemcpy(const_cast<float *>(dev) + p.off, host + at, p.n * 4, hipMemcpyHostToDevice));
    at += p.n;
  }
  return CTAS_OK;
}

extern "C" int ctas_cnn_get_params(ctas_cnn *n, float *params) {
  if (!n || !params) return fail(CTAS_E_ARG, "null argument");
  return copy_params(n, params, n->p, true);
}

extern "C" int ctas_cnn_get_grads(ctas_cnn *n, float *grads) {
  if (!n || !grads) return fail(CTAS_E_ARG, "null argument");
  return copy_params(n, grads, n->g, true);
}

extern "C" int ctas_cnn_set_params(ctas_cnn *n, const float *params) {
  if (!n || !params) return fail(CTAS_E_ARG, "null argument");
  if (int rc = copy_params(n, const_cast<float *>(params), n->p, false)) return rc;
  hipLaunchKernelGGL(to_bf16, dim3((n->nparam + 255) / 256), dim3(256), 0, n->stream, n->p, n->p16,
                     static_cast<int>(n->nparam));
  CTAS_HIP_TRY(hipStreamSynchronize(n->stream));
  return CTAS_OK;
}

// The _dev calls read inputs made ready on `stream` (nullptr: the null
// stream), run on the network's own stream and return when their results
// are complete.
extern "C" int ctas_cnn_forward_dev(ctas_cnn *n, const uint16_t *d_x, int B, float *d_out, void *stream) {
  if (!n || !d_x || !d_out || B <= 0) return fail(CTAS_E_ARG, "bad arguments");
  if (int rc = stream_wait(n->stream, static_cast<hipStream_t>(stream))) return rc;
  if (int rc = ensure_batch(n, B)) return rc;
  if (int rc = forward(n, d_x, B)) return rc;
  CTAS_HIP_TRY(hipMemcpyAsync(d_out, n->out, sizeof(float) * kOut * B, hipMemcpyDeviceToDevice, n->stream));
  CTAS_HIP_TRY(hipStreamSynchronize(n->stream));
  return CTAS_OK;
}

// Loss and parameter gradients of a batch (no update); loss on the host.
extern "C" int ctas_cnn_gradients_dev(ctas_cnn *n, const uint16_t *d_x, const float *d_target, int B, float *loss,
                                      void *stream) {
  if (!n || !d_x || !d_target || B <= 0) return fail(CTAS_E_ARG, "bad arguments");
  if (int rc = stream_wait(n->stream, static_cast<hipStream_t>(stream))) return rc;
  if (int rc = ensure_batch(n, B)) return rc;
  if (int rc = forward(n, d_x, B)) return rc;
  hipLaunchKernelGGL(loss_kernel, dim3(1), dim3(256), 0, n->stream, n->out, d_target, n->gout, n->loss, B);
  if (int rc = backward(n, d_x, B)) return rc;
  CTAS_HIP_TRY(hipStreamSynchronize(n->stream));
  if (loss) CTAS_HIP_TRY(hipMemcpy(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost));
  return CTAS_OK;
}

// Gradients + one Adam step (beta 0.9 / 0.999, eps 1e-8) at rate lr, the
// gradient scaled down to global norm clip_norm when it is larger (no
// clipping for clip_norm <= 0); grad_norm (optional) receives the norm
// before clipping.
extern "C" int ctas_cnn_train_step_dev(ctas_cnn *n, const uint16_t *d_x, const float *d_target, int B, float lr,
                                       float clip_norm, float *loss, float *grad_norm, void *stream) {
  if (int rc = ctas_cnn_gradients_dev(n, d_x, d_target, B, loss, stream)) return rc;
  ++n->step;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, static_cast<float>(n->step));
  const float c2 = 1.0f - std::pow(b2, static_cast<float>(n->step));
  const int np = static_cast<int>(n->nparam);
  hipLaunchKernelGGL(sumsq_kernel, dim3(kNormBlocks), dim3(256), 0, n->stream, n->g, np, n->norm);
  hipLaunchKernelGGL(clip_factor_kernel, dim3(1), dim3(256), 0, n->stream, n->norm, clip_norm, n->norm + kNormBlocks);
  hipLaunchKernelGGL(adam_kernel, dim3((n->nparam + 255) / 256), dim3(256), 0, n->stream, n->p, n->p16, n->g, n->m,
                     n->v, np, lr, b1, b2, c1, c2, 1e-8f, n->norm + kNormBlocks);
  if (grad_norm) {
    CTAS_HIP_TRY(hipMemcpyAsync(grad_norm, n->norm + kNormBlocks + 1, sizeof(float), hipMemcpyDeviceToHost, n->stream));
  }
  CTAS_HIP_TRY(hipStreamSynchronize(n->stream));
  return CTAS_OK;
}

// Features of one sinogram [N][w] (device, transmission) into d_x
// [2][180][256]; work on `stream` (the caller's data must be ready there).
extern "C" int ctas_cnn_features_dev(const float *d_sino, int N, int w, uint16_t *d_x, void *stream) {
  if (!d_sino || !d_x || N < 1 || w < 2) return fail(CTAS_E_ARG, "bad arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  float *tmp = nullptr;
  if (hipMalloc(&tmp, sizeof(float) * kFeatA * kFeatW) != hipSuccess) return fail(CTAS_E_NOMEM, "allocation failed");
  hipError_t e = launch_features(d_sino, N, w, tmp, d_x, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(tmp);
  if (e != hipSuccess) return fail(CTAS_E_DEVICE, std::string("features: ") + hipGetErrorString(e));
  return CTAS_OK;
}

extern "C" int ctas_cnn_save(const ctas_cnn *n, const char *path) {
  if (!n || !path) return fail(CTAS_E_ARG, "null argument");
  int c = 0;
  ctas_cnn_num_params(n, &c);
  std::vector<float> h(c);
  if (int rc = ctas_cnn_get_params(const_cast<ctas_cnn *>(n), h.data())) return rc;
  std::ofstream f(path, std::ios::binary);
  if (!f) return fail(CTAS_E_ARG, std::string("cannot write ") + path);
  const char magic[8] = {'C', 'T', 'A', 'S', 'C', 'N', 'N', '1'};
  f.write(magic, 8);
  f.write(reinterpret_cast<const char *>(&c), 4);
  f.write(reinterpret_cast<const char *>(h.data()), static_cast<std::streamsize>(h.size() * 4));
  return CTAS_OK;
}

extern "C" int ctas_cnn_load(ctas_cnn *n, const char *path) {
  if (!n || !path) return fail(CTAS_E_ARG, "null argument");
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(CTAS_E_ARG, std::string("cannot read ") + path);
  char magic[8];
  int c = 0, want = 0;
  f.read(magic, 8);
  f.read(reinterpret_cast<char *>(&c), 4);
  ctas_cnn_num_params(n, &want);
  if (!f || std::memcmp(magic, "CTASCNN1", 8) != 0 || c != want)
    return fail(CTAS_E_ARG, std::string("not a CNN weight file of this network: ") + path);
  std::vector<float> h(c);
  f.read(reinterpret_cast<char *>(h.data()), static_cast<std::streamsize>(h.size() * 4));
  if (!f) return fail(CTAS_E_ARG, std::string("truncated weight file ") + path);
  return ctas_cnn_set_params(n, h.data());
}

// ---- engine-made training samples ------------------------------------------
namespace {

uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
double unit01(uint64_t h) { return (static_cast<double>(h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

constexpr int kBigEll = 20, kDots = 40;

}  // namespace

// Sample `index` of the stream `seed`: width w in {1024, 1280, ..., 4096},
// angles N in {900, 1200, 1500, 1800} over 180 degrees, centre offset c
// uniform in [-20, 20] px, photons per ray N0 log-uniform in [1e4, 1e6],
// 20 random ellipses (ctas_phantom_ellipses, semi-axes scaled by w/512)
// plus 40 dense inclusions (semi-axes 1-3 px, 0.03-0.09 per px) whose sharp
// edges give the average-gradient score its minimum at the true centre.
// ell receives (20 + 40) * 6 doubles; returns the ellipse count.
extern "C" int ctas_axis_sample(uint64_t seed, uint64_t index, int *w, int *N, double *c, double *N0, double *ell,
                                int *n_ell) {
  if (!w || !N || !c || !N0 || !ell || !n_ell) return fail(CTAS_E_ARG, "null argument");
  uint64_t st = mix64(seed ^ mix64(index + 0x51ed2701ull));
  auto next = [&]() {
    st = mix64(st);
    return unit01(st);
  };
  *w = 1024 + 256 * std::min(12, static_cast<int>(next() * 13));
  *N = 900 + 300 * std::min(3, static_cast<int>(next() * 4));
  *c = -20.0 + 40.0 * next();
  *N0 = std::pow(10.0, 4.0 + 2.0 * next());
  const uint64_t eseed = mix64(st ^ 0xe11a5e5ull);
  if (int rc = ctas_phantom_ellipses(kBigEll, *w, *w / 512.0, eseed, ell)) return rc;
  const double R = 0.5 * *w;
  for (int k = 0; k < kDots; ++k) {
    double *E = ell + 6 * (kBigEll + k);
    const double rr = 0.8 * R * std::sqrt(next());
    const double ang = 2.0 * M_PI * next();
    E[0] = rr * std::cos(ang);
    E[1] = rr * std::sin(ang);
    E[2] = 1.0 + 2.0 * next();
    E[3] = 1.0 + 2.0 * next();
    E[4] = M_PI * next();
    E[5] = 0.03 + 0.06 * next();
  }
  *n_ell = kBigEll + kDots;
  return CTAS_OK;
}

// Network input and targets of samples first .. first+B-1 of `seed`:
// d_x [B][2][180][256] bf16, d_target [B][2] = {offset in downsampled bins,
// tolerance (1 px) in bins}; offsets_px (host, optional) receives the true
// offsets and widths_px the widths. The sinograms are projected on the GPU
// (ctas_phantom_sino_dev geometry) one at a time into a scratch buffer.
extern "C" int ctas_cnn_make_batch(uint64_t seed, uint64_t first, int B, uint16_t *d_x, float *d_target,
                                   float *offsets_px, int *widths_px, void *stream) {
  if (!d_x || !d_target || B <= 0) return fail(CTAS_E_ARG, "bad arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  constexpr int kMaxEll = kBigEll + kDots;
  float *d_sino = nullptr, *d_tmp = nullptr;
  double *d_ell = nullptr;
  const size_t max_sino = static_cast<size_t>(4096) * 1800;
  int rc = CTAS_OK;
  if (hipMalloc(&d_sino, sizeof(float) * max_sino) != hipSuccess ||
      hipMalloc(&d_tmp, sizeof(float) * kFeatA * kFeatW) != hipSuccess ||
      hipMalloc(&d_ell, sizeof(double) * 6 * kMaxEll * static_cast<size_t>(B)) != hipSuccess)
    rc = fail(CTAS_E_NOMEM, "allocation failed");
  std::vector<double> ell(static_cast<size_t>(6) * kMaxEll * B);
  std::vector<float> tgt(2 * static_cast<size_t>(B));
  std::vector<int> ws(B), Ns(B), ne(B);
  std::vector<double> cs(B), n0s(B);
  for (int b = 0; b < B && !rc; ++b)
    rc = ctas_axis_sample(seed, first + b, &ws[b], &Ns[b], &cs[b], &n0s[b], ell.data() + 6 * kMaxEll * b, &ne[b]);
  if (!rc && hipMemcpyAsync(d_ell, ell.data(), sizeof(double) * ell.size(), hipMemcpyHostToDevice, s) != hipSuccess)
    rc = fail(CTAS_E_DEVICE, "copy failed");
  for (int b = 0; b < B && !rc; ++b) {
    const double bin = ws[b] / static_cast<double>(kFeatW);
    tgt[2 * b] = static_cast<float>(cs[b] / bin);
    tgt[2 * b + 1] = static_cast<float>(1.0 / bin);
    if (offsets_px) offsets_px[b] = static_cast<float>(cs[b]);
    if (widths_px) widths_px[b] = ws[b];
    const uint64_t nseed = mix64(seed * 0x9E3779B97F4A7C15ull + first + b);
    if (launch_phantom(Ns[b], ws[b], cs[b], ne[b], d_ell + 6 * kMaxEll * b, n0s[b], nseed, nullptr, d_sino, s) !=
            hipSuccess ||
        launch_features(d_sino, Ns[b], ws[b], d_tmp, d_x + static_cast<size_t>(b) * 2 * kFeatA * kFeatW, s) !=
            hipSuccess)
      rc = fail(CTAS_E_DEVICE, "sample generation launch failed");
  }
  if (!rc && hipMemcpyAsync(d_target, tgt.data(), sizeof(float) * tgt.size(), hipMemcpyHostToDevice, s) != hipSuccess)
    rc = fail(CTAS_E_DEVICE, "copy failed");
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = fail(CTAS_E_DEVICE, "sample generation failed");
  if (d_sino) (void)hipFree(d_sino);
  if (d_tmp) (void)hipFree(d_tmp);
  if (d_ell) (void)hipFree(d_ell);
  return rc;
}

// Offset (full-resolution px) and confidence of one device sinogram.
extern "C" int ctas_cnn_predict_dev(ctas_cnn *n, const float *d_sino, int N, int w, float *offset,
                                    float *confidence, void *stream) {
  if (!n || !d_sino || !offset || !confidence || N < 1 || w < 2) return fail(CTAS_E_ARG, "bad arguments");
  if (int rc = stream_wait(n->stream, static_cast<hipStream_t>(stream))) return rc;
  if (int rc = ensure_batch(n, 1)) return rc;
  if (!n->xin && alloc(reinterpret_cast<void **>(&n->xin), 2 * kFeatA * kFeatW * 2)) return CTAS_E_NOMEM;
  CTAS_HIP_TRY(launch_features(d_sino, N, w, n->ftmp, n->xin, n->stream));
  if (int rc = forward(n, n->xin, 1)) return rc;
  float o[2];
  CTAS_HIP_TRY(hipMemcpyAsync(o, n->out, sizeof(o), hipMemcpyDeviceToHost, n->stream));
  CTAS_HIP_TRY(hipStreamSynchronize(n->stream));
  *offset = o[0] * static_cast<float>(w) / kFeatW;
  *confidence = 1.0f / (1.0f + std::exp(-o[1]));
  return CTAS_OK;
}

// Same for a host sinogram.
extern "C" int ctas_cnn_predict(ctas_cnn *n, const float *sino, int N, int w, float *offset, float *confidence) {
  if (!n || !sino || !offset || !confidence || N < 1 || w < 2) return fail(CTAS_E_ARG, "bad arguments");
  float *d_s = nullptr;
  if (hipMalloc(&d_s, sizeof(float) * N * w) != hipSuccess) return fail(CTAS_E_NOMEM, "allocation failed");
  int rc = CTAS_OK;
  if (hipMemcpy(d_s, sino, sizeof(float) * N * w, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(CTAS_E_DEVICE, "copy failed");
  if (!rc) rc = ctas_cnn_predict_dev(n, d_s, N, w, offset, confidence, nullptr);
  (void)hipFree(d_s);
  return rc;
}

// Sweep of trial centres offsets -half_range + k*step (k = 0 .. K-1) on the
// device sinogram d_sino of handle h's shape (AverageGradient canon: ABS,
// RAMP); result = the offset of the lowest score (first on ties). With a
// network whose confidence reaches min_confidence, only a window of 3 grid
// points around its estimate is scored, widened by one point on the side of
// its lowest score while that is at an edge (and not the grid's end); past
// kMaxWindow points the remaining grid is scored too. info
// (optional, 4 ints): trials scored, 1 if the whole grid was swept,
// network estimate (milli-px), confidence (per mille).
constexpr int kMaxWindow = 21;

namespace {

// The sweep of ctas_axis_find on a device sinogram: with `windowed`, a
// window of 3 grid points around `start` (px), widened one point at a time
// towards a lower edge score until its minimum is interior, the whole grid
// past kMaxWindow points; otherwise the whole grid. est and conf only go to
// info.
// Argument checks and the filter of the sinogram on the handle's stream
// (after the work already queued on the null stream, which the caller's
// NullStreamJoin orders).
int axis_prepare(ctas_fbp *h, const float *d_sino, float half_range, float step, float *offset) {
  if (!h || !d_sino || !offset || !(step > 0) || !(half_range >= 0)) return fail(CTAS_E_ARG, "bad arguments");
  const int K = static_cast<int>(std::floor(2.0 * half_range / step + 0.5)) + 1;
  if (K > 4096) return fail(CTAS_E_ARG, "too many trial centres");
  if (std::ceil(half_range) >= 0.5f * h->w) return fail(CTAS_E_CENTER, "trial centres beyond half the width");
  return ctas_fbp_prepare_dev(h, d_sino, 1, h->stream);
}

// The search itself, on the filtered sinogram axis_prepare left.
int axis_search(ctas_fbp *h, float half_range, float step, bool windowed, float start, float est, float conf,
                float *offset, int *info) {
  const int w = h->w;
  const int K = static_cast<int>(std::floor(2.0 * half_range / step + 0.5)) + 1;
  const float cent = static_cast<float>(0.5 * (w - 1));
  auto grid = [&](int k) { return static_cast<float>(-static_cast<double>(half_range) + k * static_cast<double>(step)); };
  std::vector<float> sc(K, 0.0f);
  std::vector<char> have(K, 0);
  int trials = 0;
  auto eval = [&](int lo, int hi) -> int {
    std::vector<float> ax;
    std::vector<int> ks;
    for (int k = lo; k <= hi; ++k)
      if (!have[k]) {
        ax.push_back(cent + grid(k));
        ks.push_back(k);
      }
    if (ax.empty()) return CTAS_OK;
    std::vector<float> v(ax.size());
    if (int rc = ctas_axis_scores(h, ax.data(), static_cast<int>(ax.size()), v.data())) return rc;
    for (size_t i = 0; i < ks.size(); ++i) {
      sc[ks[i]] = v[i];
      have[ks[i]] = 1;
    }
    trials += static_cast<int>(ks.size());
    return CTAS_OK;
  };
  auto argmin = [&](int lo, int hi) {
    int m = lo;
    for (int k = lo + 1; k <= hi; ++k)
      if (sc[k] < sc[m]) m = k;
    return m;
  };
  int swept = 1, km = 0;
  if (windowed) {
    swept = 0;
    const int k0 = std::max(0, std::min(K - 1, static_cast<int>(std::lround((start + half_range) / step))));
    int lo = std::max(0, k0 - 1), hi = std::min(K - 1, k0 + 1);
    if (int rc = eval(lo, hi)) return rc;
    for (;;) {
      km = argmin(lo, hi);
      if (hi - lo + 1 > kMaxWindow) {
        swept = 1;
        break;
      }
      if (km == lo && lo > 0) {
        --lo;
      } else if (km == hi && hi < K - 1) {
        ++hi;
      } else {
        break;
      }
      if (int rc = eval(lo, hi)) return rc;
    }
  }
  if (swept) {
    if (int rc = eval(0, K - 1)) return rc;
    km = argmin(0, K - 1);
  }
  *offset = grid(km);
  if (info) {
    info[0] = trials;
    info[1] = swept;
    info[2] = static_cast<int>(std::lround(est * 1000.0f));
    info[3] = static_cast<int>(std::lround(conf * 1000.0f));
  }
  return CTAS_OK;
}

}  // namespace

extern "C" int ctas_axis_find(ctas_fbp *h, ctas_cnn *net, const float *d_sino, float half_range, float step,
                              float min_confidence, float *offset, int *info) {
  if (!h) return fail(CTAS_E_ARG, "bad arguments");
  // filtering on the handle's stream, the network on its own stream, both
  // after the work already queued on the null stream
  NullStreamJoin join(h->stream, nullptr);
  if (join.rc) return join.rc;
  if (int rc = axis_prepare(h, d_sino, half_range, step, offset)) return rc;
  float est = 0.0f, conf = 0.0f;
  if (net) {
    if (int rc = ctas_cnn_predict_dev(net, d_sino, h->N, h->w, &est, &conf, nullptr)) return rc;
  }
  return axis_search(h, half_range, step, net && conf >= min_confidence, est, est, conf, offset, info);
}

extern "C" int ctas_axis_find_from(ctas_fbp *h, const float *d_sino, float half_range, float step, float start,
                                   float *offset, int *info) {
  if (!h) return fail(CTAS_E_ARG, "bad arguments");
  NullStreamJoin join(h->stream, nullptr);
  if (join.rc) return join.rc;
  if (int rc = axis_prepare(h, d_sino, half_range, step, offset)) return rc;
  return axis_search(h, half_range, step, true, start, start, 1.0f, offset, info);
}

// bf16 GEMM of the network's kernel on plain row-major operands:
// C[M][N] (fp32) = A[M][K] x B[K][N] (bf16). Exposed for verification of the
// matrix-core operand layout.
extern "C" int ctas_cnn_gemm_dev(const uint16_t *d_a, const uint16_t *d_b, float *d_c, int M, int N, int K,
                                 void *stream) {
  if (!d_a || !d_b || !d_c || M <= 0 || N <= 0 || K <= 0) return fail(CTAS_E_ARG, "bad arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  CTAS_HIP_TRY(gemm(DenseA{d_a, K}, ActBT{d_b, N}, StoreDX{d_c, 1, 1, N}, M, N, K, s));  // on the caller's stream
  CTAS_HIP_TRY(hipStreamSynchronize(s));
  return CTAS_OK;
}

// CNN-fronted axis search: the network's estimate replaces the sweep of
// raxis algorithm 1 (lct.cpp:1183-1205); the score is evaluated at the
// estimate (rounded to 0.25 px) and +-4 px around it in one batch, and
// improveMe (lct.cpp:1108-1129) refines from that bracket with the same
// tolerance (0.25 px). Falls back to ctas_raxis (algorithm 1) when the
// confidence is below min_confidence, the estimate is outside maxDev, or
// the centre of the bracket is not its lowest score; the argument checks
// and their errors are those of ctas_raxis.
extern "C" int ctas_raxis_cnn(ctas_cnn *n, const float *sino, int N, int w, float anarc, float maxDev,
                              float min_confidence, float *result, int *used_sweep) {
  if (!n || !sino || !result) return fail(CTAS_E_ARG, "null argument");
  if (used_sweep) *used_sweep = 0;
  auto sweep = [&]() {
    if (used_sweep) *used_sweep = 1;
    return ctas_raxis(sino, N, w, anarc, maxDev, 1, result);
  };
  const float arc = N > 1 ? arc_rad(anarc, N) : 0.0f;
  if (N < 3 || w < 3 || std::fabs(arc) < static_cast<float>(M_PI) || maxDev < 0 || maxDev > 0.25f) return sweep();
  float off = 0.0f, conf = 0.0f;
  if (int rc = ctas_cnn_predict(n, sino, N, w, &off, &conf)) return rc;
  const float mid = static_cast<float>(0.5 * (w - 1));
  const float est_px = std::round(off * 4.0f) / 4.0f;
  if (!(conf >= min_confidence) || std::fabs(est_px) + 4.0f > maxDev * w) return sweep();
  ctas_fbp *h = nullptr;
  if (int rc = ctas_fbp_create(&h, w, N, CTAS_CONTRAST_ABS, CTAS_FILTER_RAMP, 0.0f, anarc, 0)) return rc;
  std::unique_ptr<ctas_fbp, int (*)(ctas_fbp *)> guard(h, ctas_fbp_destroy);
  std::vector<float> copy(sino, sino + static_cast<size_t>(N) * w);
  if (int rc = ctas_fbp_sino(h, copy.data())) return rc;
  const std::vector<float> offsets = {est_px - 4.0f, est_px, est_px + 4.0f};
  std::vector<float> scores(3);
  const float ax[3] = {mid + offsets[0], mid + offsets[1], mid + offsets[2]};
  if (int rc = ctas_axis_scores(h, ax, 3, scores.data())) return rc;
  if (!(scores[1] < scores[0] && scores[1] < scores[2])) {
    guard.reset();
    return sweep();
  }
  int rc = CTAS_OK;
  auto score_at = [&](float x) -> float {
    float axis = mid + x, v = 0.0f;
    if (!rc) rc = ctas_axis_scores(h, &axis, 1, &v);
    return v;
  };
  const float best = improve_me(0.25f, offsets, scores, 1, 1, score_at, &rc);
  if (rc) return rc;
  *result = mid + best;
  return CTAS_OK;
}
