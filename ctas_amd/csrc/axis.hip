// Rotation-axis search (raxis, lct.cpp:1133-1326) with the trial
// reconstructions and their sharpness scores (AverageGradient::apply,
// lct.cpp:872-915) evaluated on the GPU in batches of centres.
//
// Score of one trial slice: the one-sided Gaussian smoothing of
// lct.cpp:837-843 / 878-884 is the circular convolution with
// g[u]*g[v], g[n] = exp(-n^2 / (2 sigma^2)), evaluated separably with the
// taps above 1e-9 (a row pass through LDS, then a column pass); then the
// rows [r, w-r) with r = ceil|c| are Sobel-filtered into the matching rows
// of zeroed full-size derivative arrays, and the magnitude is read from the
// top w-2r rows of those arrays (the shape mismatch of lct.cpp:905), the
// border of that block is zeroed and the mean of its positive values is the
// score. The block sums are reduced in a fixed order (deterministic).
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "ctas_internal.hpp"

namespace ctas {
namespace {

constexpr int kMaxTaps = 64;

struct Taps {
  float g[kMaxTaps];
  int nt;
};

Taps make_taps(float sigma) {
  Taps t{};
  const float sig22 = 2 * sigma * sigma;
  t.nt = 0;
  for (int n = 0; n < kMaxTaps; ++n) {
    const float v = std::exp(-static_cast<float>(n * n) / sig22);
    t.g[n] = v;
    if (v > 1e-9f) t.nt = n + 1;
  }
  return t;
}

// i mod w for i >= -kMaxTaps: one addition unless the taps outnumber w.
__device__ __forceinline__ int wrap_index(int i, int w) {
  if (i >= 0) return i;
  i += w;
  return i >= 0 ? i : ((i % w) + w) % w;
}

// Row pass: dst[u][v] = sum_n g[n] src[u][(v - n) mod w]. Grid (w, n).
__global__ __launch_bounds__(256) void smooth_rows_kernel(const float *__restrict__ src, float *__restrict__ dst,
                                                          int w, Taps taps) {
  extern __shared__ float row[];
  const size_t base = (static_cast<size_t>(blockIdx.y) * w + blockIdx.x) * w;
  for (int v = threadIdx.x; v < w; v += blockDim.x) row[v] = src[base + v];
  __syncthreads();
  for (int v = threadIdx.x; v < w; v += blockDim.x) {
    float a = 0.0f;
    for (int n = 0; n < taps.nt; ++n) {
      a = a + taps.g[n] * row[wrap_index(v - n, w)];
    }
    dst[base + v] = a;
  }
}

// Column pass: dst[u][v] = sum_m g[m] src[(u - m) mod w][v]. Grid (w, n).
__global__ __launch_bounds__(256) void smooth_cols_kernel(const float *__restrict__ src, float *__restrict__ dst,
                                                          int w, Taps taps) {
  const int u = blockIdx.x;
  const size_t sl = static_cast<size_t>(blockIdx.y) * w * w;
  for (int v = threadIdx.x; v < w; v += blockDim.x) {
    float a = 0.0f;
    for (int m = 0; m < taps.nt; ++m) {
      a = a + taps.g[m] * src[sl + static_cast<size_t>(wrap_index(u - m, w)) * w + v];
    }
    dst[sl + static_cast<size_t>(u) * w + v] = a;
  }
}

// Sobel derivative of the cropped slice cs (h x w) at (yy, x), accumulated
// in the statement order of lct.cpp:887-902; (dy, dx, coefficient) terms.
__device__ __forceinline__ float sobel_term(const float *cs, int h, int w, int yy, int x, int a, int b, float k,
                                            float acc) {
  const int sy = yy - a, sx = x - b;
  if (sy >= 0 && sy < h && sx >= 0 && sx < w) acc = acc + k * cs[static_cast<size_t>(sy) * w + sx];
  return acc;
}

// Partial sums of the score of slice k (blockIdx.y) over the derivative rows
// y = 1 + blockIdx.x, 1 + blockIdx.x + nb, ... (nb = gridDim.x), lanes over
// the columns; partial[(k nb + b) 2 + {0, 1}] = {sum of cS, #(cS > 0)}.
__global__ __launch_bounds__(256) void sobel_score_kernel(const float *__restrict__ slices,
                                                          const float *__restrict__ centres, int w,
                                                          int smoothed, double *__restrict__ partial) {
  const int k = blockIdx.y, nb = gridDim.x;
  const float c = centres[k];
  const int r = static_cast<int>(ceilf(fabsf(c)));
  const int h = w - 2 * r;
  const float *cs = slices + static_cast<size_t>(k) * w * w + static_cast<size_t>(r) * w;
  double sm = 0.0, cnt = 0.0;
  // full-array row y holds cropped-derivative row y - r (zero above r):
  // rows y = max(1, r) .. h - 2, columns 1 .. w - 2
  for (int y = 1 + blockIdx.x; y <= h - 2; y += nb) {
    const int yy = y - r;
    if (yy < 0) continue;
    for (int x = 1 + threadIdx.x; x <= w - 2; x += blockDim.x) {
      float dH = 0.0f, dV = 0.0f;
      dH = sobel_term(cs, h, w, yy, x, -1, -1, -1.0f, dH);
      dH = sobel_term(cs, h, w, yy, x, -1, 0, -2.0f, dH);
      dH = sobel_term(cs, h, w, yy, x, -1, 1, -1.0f, dH);
      dH = sobel_term(cs, h, w, yy, x, 1, -1, 1.0f, dH);
      dH = sobel_term(cs, h, w, yy, x, 1, 0, 2.0f, dH);
      dH = sobel_term(cs, h, w, yy, x, 1, 1, 1.0f, dH);
      dV = sobel_term(cs, h, w, yy, x, -1, -1, -1.0f, dV);
      dV = sobel_term(cs, h, w, yy, x, 0, -1, -2.0f, dV);
      dV = sobel_term(cs, h, w, yy, x, 1, -1, -1.0f, dV);
      dV = sobel_term(cs, h, w, yy, x, -1, 1, 1.0f, dV);
      dV = sobel_term(cs, h, w, yy, x, 0, 1, 2.0f, dV);
      dV = sobel_term(cs, h, w, yy, x, 1, 1, 1.0f, dV);
      const float m = sqrtf(dH * dH + dV * dV);
      sm += m;
      // #(cS > 0) of lct.cpp:912: the reference's float-FFT smoothing leaves
      // round-off in every pixel, so then every derivative pixel counts.
      cnt += (smoothed || m > 0.0f) ? 1.0 : 0.0;
    }
  }
  __shared__ double ps[256], pc[256];
  ps[threadIdx.x] = sm;
  pc[threadIdx.x] = cnt;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      ps[threadIdx.x] += ps[threadIdx.x + o];
      pc[threadIdx.x] += pc[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[(static_cast<size_t>(k) * nb + blockIdx.x) * 2] = ps[0];
    partial[(static_cast<size_t>(k) * nb + blockIdx.x) * 2 + 1] = pc[0];
  }
}

// Score of slice k (blockIdx.x): the nb partials summed in a fixed order.
__global__ __launch_bounds__(256) void score_finish_kernel(const double *__restrict__ partial,
                                                           float *__restrict__ scores, int nb) {
  const int k = blockIdx.x;
  double s = 0.0, c = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    s += partial[(static_cast<size_t>(k) * nb + b) * 2];
    c += partial[(static_cast<size_t>(k) * nb + b) * 2 + 1];
  }
  __shared__ double ps[256], pc[256];
  ps[threadIdx.x] = s;
  pc[threadIdx.x] = c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      ps[threadIdx.x] += ps[threadIdx.x + o];
      pc[threadIdx.x] += pc[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float sm = static_cast<float>(ps[0]);
    scores[k] = pc[0] > 0.0 ? sm / static_cast<float>(pc[0]) : 0.0f;
  }
}

}  // namespace

hipError_t launch_ag_scores(float *d_slices, int n, int w, const float *d_centres, float sigma, float *d_tmp,
                            double *d_partial, float *d_scores, hipStream_t s) {
  if (sigma > 1.0f) {
    const Taps taps = make_taps(sigma);
    hipLaunchKernelGGL(smooth_rows_kernel, dim3(w, n), dim3(256), sizeof(float) * w, s, d_slices, d_tmp, w, taps);
    hipLaunchKernelGGL(smooth_cols_kernel, dim3(w, n), dim3(256), 0, s, d_tmp, d_slices, w, taps);
  }
  // as many row blocks per slice as fill the chip, within the partial
  // buffer's kScoreBlocks x kScoreChunk pairs
  const int nb = std::max(1, std::min(1024, kScoreBlocks * kScoreChunk / std::max(n, 1)));
  hipLaunchKernelGGL(sobel_score_kernel, dim3(nb, n), dim3(256), 0, s, d_slices, d_centres, w,
                     sigma > 1.0f ? 1 : 0, d_partial);
  hipLaunchKernelGGL(score_finish_kernel, dim3(n), dim3(256), 0, s, d_partial, d_scores, nb);
  return hipGetLastError();
}

namespace {

template <typename T>
int grow_buf(T *&ptr, size_t &cap, size_t bytes) {
  if (bytes <= cap && ptr) return CTAS_OK;
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  cap = 0;
  void *p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess)
    return fail(CTAS_E_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
  ptr = static_cast<T *>(p);
  cap = bytes;
  return CTAS_OK;
}


}  // namespace

// improveMe (lct.cpp:1108-1129): bisection towards the lower neighbour.
// The bracket around offsets[at] is halved on its wider side until both
// sides are within eps; a probe that scores no better than the current best
// becomes that side's edge, a better one the new best (the old best then
// closes the other side).
float improve_me(float eps, const std::vector<float> &offsets, const std::vector<float> &scores, int at, int dir,
                 const std::function<float(float)> &score_at, int *rc) {
  float left = offsets[at - dir];
  float right = offsets[at + dir];
  float best = offsets[at];
  float best_score = scores[at];
  while (std::fabs(left - best) > eps || std::fabs(best - right) > eps) {
    const bool left_wider = std::fabs(left - best) > std::fabs(best - right);
    float &wide = left_wider ? left : right;
    float &narrow = left_wider ? right : left;
    const float probe = (best + wide) / 2;
    const float probe_score = score_at(probe);
    if (*rc) return best;
    if (best_score <= probe_score) {
      wide = probe;
    } else {
      best_score = probe_score;
      narrow = best;
      best = probe;
    }
  }
  return best;
}

namespace {

struct HandleGuard {
  ctas_fbp *h = nullptr;
  ~HandleGuard() {
    if (h) ctas_fbp_destroy(h);
  }
};

}  // namespace
}  // namespace ctas

using namespace ctas;

extern "C" int ctas_axis_scores(ctas_fbp *h, const float *axisR, int n, float *scores) {
  if (!h || !axisR || !scores || n <= 0) return fail(CTAS_E_ARG, "bad arguments");
  if (h->kept_slices < 1) return fail(CTAS_E_STATE, "no filtered sinogram on the device: call ctas_fbp_sino first");
  const int w = h->w;
  const float sigma = static_cast<float>(step_rad(h->anarc, h->N) * static_cast<float>(w) / M_PI);
  const size_t ww = static_cast<size_t>(w) * w;
  for (int k0 = 0; k0 < n; k0 += kScoreChunk) {
    const int m = std::min(kScoreChunk, n - k0);
    std::vector<float> centres(m);
    for (int i = 0; i < m; ++i)
      centres[i] = static_cast<float>(static_cast<double>(axisR[k0 + i]) - 0.5 * (w - 1));
    if (int rc = grow_buf(h->d_out, h->out_cap, sizeof(float) * ww * m)) return rc;
    if (int rc = grow_buf(h->d_tmp, h->score_cap, sizeof(float) * ww * m)) return rc;
    size_t pcap = h->d_partial ? sizeof(double) * 2 * kScoreBlocks * kScoreChunk : 0;
    if (int rc = grow_buf(h->d_partial, pcap, sizeof(double) * 2 * kScoreBlocks * kScoreChunk)) return rc;
    if (!h->d_centres) {
      if (hipMalloc(&h->d_centres, sizeof(float) * kScoreChunk) != hipSuccess ||
          hipMalloc(&h->d_scores, sizeof(float) * kScoreChunk) != hipSuccess)
        return fail(CTAS_E_NOMEM, "allocation failed");
    }
    // all on the handle's stream: the centres' upload, the trial slices, the
    // scores and their download, then one wait (centres outlives it)
    CTAS_HIP_TRY(hipMemcpyAsync(h->d_centres, centres.data(), sizeof(float) * m, hipMemcpyHostToDevice, h->stream));
    if (int rc = ctas_fbp_repeat_dev(h, 0, h->d_out, centres.data(), m, h->stream)) return rc;
    CTAS_HIP_TRY(launch_ag_scores(h->d_out, m, w, h->d_centres, sigma, h->d_tmp, h->d_partial, h->d_scores,
                                  h->stream));
    CTAS_HIP_TRY(hipMemcpyAsync(scores + k0, h->d_scores, sizeof(float) * m, hipMemcpyDeviceToHost, h->stream));
    CTAS_HIP_TRY(hipStreamSynchronize(h->stream));
  }
  return CTAS_OK;
}

namespace ctas {
namespace {

// ---- 0/180-degree registration (raxis(proj0, proj180, maxDev)) ----------
// Column profile sum_r |p0 - 1| + |p180 - 1| (double terms added to a float
// per column, rows in order), its centre of mass and the mean of the two
// edge columns of the mirrored p180, each a sequential sum as in
// lct.cpp:1226-1251; one workgroup, w <= 16384.
__global__ __launch_bounds__(256) void pair_profile_kernel(const float *__restrict__ p0,
                                                           const float *__restrict__ p180, int rows, int w,
                                                           float *__restrict__ out) {
  extern __shared__ float prof[];
  for (int x = threadIdx.x; x < w; x += 256) {
    float acc = 0.0f;
    for (int r = 0; r < rows; ++r) {
      const size_t i = static_cast<size_t>(r) * w + x;
      acc = static_cast<float>(acc + (fabs(static_cast<double>(p0[i]) - 1.0) + fabs(static_cast<double>(p180[i]) - 1.0)));
    }
    prof[x] = acc;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double ssum = 0.0, wsum = 0.0;
  for (int x = 0; x < w; ++x) {
    ssum += prof[x];
    wsum += static_cast<double>(static_cast<float>(x) * prof[x]);
  }
  // p180 mirrored: its first column is p180's last and vice versa
  double a0 = 0.0, a1 = 0.0;
  for (int r = 0; r < rows; ++r) {
    a0 += p180[static_cast<size_t>(r) * w + (w - 1)];
    a1 += p180[static_cast<size_t>(r) * w];
  }
  out[0] = static_cast<float>(ssum == 0.0 ? 0.0 : 1.0);
  out[1] = ssum == 0.0 ? 0.0f : static_cast<float>(wsum) / static_cast<float>(ssum);
  out[2] = (static_cast<float>(a0) + static_cast<float>(a1)) / static_cast<float>(2 * rows);
}

// Registration error of trial shift x = shifts[blockIdx.x] (calcErr of
// lct.cpp:1255-1270): sum over rows and columns of (p180r(col - sh) - p0)^2,
// p180r the mirrored p180, `aver` outside the overlap, each square in float
// and summed in double; the block's partial sums are reduced in a fixed order.
__global__ __launch_bounds__(256) void pair_err_kernel(const float *__restrict__ p0, const float *__restrict__ p180,
                                                       int rows, int w, const float *__restrict__ shifts, float aver,
                                                       double *__restrict__ out) {
  __shared__ double part[256];
  const long sh = static_cast<long>(shifts[blockIdx.x]);
  const long lo = sh > 0 ? sh : 0, hi = w - 1 + (sh < 0 ? sh : 0);
  const size_t n = static_cast<size_t>(rows) * w;
  double acc = 0.0;
  for (size_t i = threadIdx.x; i < n; i += 256) {
    const long col = static_cast<long>(i % w), r = static_cast<long>(i / w);
    const long src = col - sh;
    float d = aver;
    if (col >= lo && col <= hi && src >= 0 && src < w) d = p180[r * w + (w - 1 - src)];
    d = d - p0[i];
    acc += static_cast<double>(d * d);
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (static_cast<int>(threadIdx.x) < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = part[0];
}

}  // namespace
}  // namespace ctas

using namespace ctas;

// raxis(proj0, proj180, maxDev) (lct.cpp:1217-1294): registration of the
// 0 and 180 degree projections. The profile and the error sums of every
// trial shift run on the GPU; the sweep and the bisection are host logic.
extern "C" int ctas_raxis_pair(const float *p0, const float *p180, int rows, int w, float maxDev, float *result) {
  if (!p0 || !p180 || !result) return fail(CTAS_E_ARG, "null argument");
  if (maxDev < 0 || maxDev > 0.5f)
    return fail(CTAS_E_ARG, "Invalid maximum relative shift of the rotation axis " + std::to_string(maxDev) +
                                ". Must be within [0, 0.5] range.");
  if (rows <= 0 || w <= 0) return fail(CTAS_E_ARG, "Empty input projections.");
  if (w > 16384) return fail(CTAS_E_ARG, "projection width above 16384");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(CTAS_E_DEVICE, "no HIP device is available");
  const float mid = static_cast<float>(0.5 * (w - 1));
  const size_t n = static_cast<size_t>(rows) * w;
  const int kStep = 16;
  const int n_trials = static_cast<int>(1 + 2 * (maxDev == 0.0f ? 1.0f : maxDev * static_cast<float>(w) / kStep));
  // device: p0, p180, profile results, trial shifts, their sums
  const int max_trials = std::max(n_trials, 8) + w / kStep + 4;
  struct Buf {
    void *p = nullptr;
    ~Buf() {
      if (p) (void)hipFree(p);
    }
  } buf;
  const size_t off_res = 2 * n * sizeof(float), off_sh = off_res + 256, off_acc = off_sh + 4 * ((max_trials + 63) / 64 * 64);
  if (hipMalloc(&buf.p, off_acc + sizeof(double) * max_trials) != hipSuccess)
    return fail(CTAS_E_NOMEM, "device allocation failed");
  char *base = static_cast<char *>(buf.p);
  float *d_p0 = reinterpret_cast<float *>(base), *d_p180 = d_p0 + n;
  float *d_res = reinterpret_cast<float *>(base + off_res), *d_sh = reinterpret_cast<float *>(base + off_sh);
  double *d_acc = reinterpret_cast<double *>(base + off_acc);
  CTAS_HIP_TRY(hipMemcpy(d_p0, p0, n * sizeof(float), hipMemcpyHostToDevice));
  CTAS_HIP_TRY(hipMemcpy(d_p180, p180, n * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(pair_profile_kernel, dim3(1), dim3(256), sizeof(float) * w, nullptr, d_p0, d_p180, rows, w,
                     d_res);
  CTAS_HIP_TRY(hipGetLastError());
  float res[3];
  CTAS_HIP_TRY(hipMemcpy(res, d_res, sizeof res, hipMemcpyDeviceToHost));
  if (res[0] == 0.0f) return fail(CTAS_E_AXIS, "Impossible input projection averaging to zero.");
  const float seed = res[1], mean_level = res[2];
  int dev_rc = CTAS_OK;
  // errors of trial shifts xs[0..k), one workgroup each
  auto errs = [&](const float *xs, int k, float *e) {
    if (dev_rc != CTAS_OK) return;
    std::vector<double> acc(k);
    if (hipMemcpy(d_sh, xs, sizeof(float) * k, hipMemcpyHostToDevice) != hipSuccess) {
      dev_rc = fail(CTAS_E_DEVICE, "registration: upload failed");
      return;
    }
    hipLaunchKernelGGL(pair_err_kernel, dim3(k), dim3(256), 0, nullptr, d_p0, d_p180, rows, w, d_sh, mean_level, d_acc);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpy(acc.data(), d_acc, sizeof(double) * k, hipMemcpyDeviceToHost) != hipSuccess) {
      dev_rc = fail(CTAS_E_DEVICE, "registration: error sums failed on the device");
      return;
    }
    for (int i = 0; i < k; ++i) e[i] = std::sqrt(static_cast<float>(acc[i]));
  };
  auto calc_err = [&](float x) -> float {
    float e = 0.0f;
    errs(&x, 1, &e);
    return e;
  };
  std::vector<float> offsets(n_trials), errors(n_trials);
  for (int ii = 0; ii < n_trials; ++ii) offsets[ii] = 2 * (seed - mid) + static_cast<float>(kStep * (ii - (n_trials - 1) / 2));
  errs(offsets.data(), n_trials, errors.data());
  if (dev_rc != CTAS_OK) return dev_rc;
  int idx = -1;
  float best_error = FLT_MAX;
  for (int i = 1; i < n_trials - 1; ++i)
    if (errors[i] < errors[i - 1] && errors[i] < errors[i + 1] && errors[i] < best_error) best_error = errors[idx = i];
  int rc = CTAS_OK;
  if (idx >= 0) {
    *result = mid + 0.5f * improve_me(1.0f, offsets, errors, idx, 1, calc_err, &rc);
    return dev_rc != CTAS_OK ? dev_rc : rc;
  }
  idx = errors[0] < errors[n_trials - 1] ? 0 : n_trials - 1;
  const int dir = idx == 0 ? -1 : 1;
  best_error = errors[idx];
  const int walk_limit = static_cast<int>((static_cast<float>(w / 2) - static_cast<float>(dir) * offsets[idx]) / kStep);
  if (walk_limit <= 0) return fail(CTAS_E_AXIS, "Failed to calculate rotation centre.");
  std::vector<float> walk_offsets(walk_limit + 2), walk_errors(walk_limit + 2);
  walk_offsets[0] = offsets[idx - dir];
  walk_offsets[1] = offsets[idx];
  walk_errors[0] = errors[idx - dir];
  walk_errors[1] = errors[idx];
  int walk_best = 1;
  for (int i = 0; i < walk_limit; ++i) {
    walk_offsets[i + 2] = offsets[idx] + static_cast<float>(dir * (i + 1) * kStep);
    walk_errors[i + 2] = calc_err(walk_offsets[i + 2]);
    if (dev_rc != CTAS_OK) return dev_rc;
    if (walk_errors[i + 2] < best_error) {
      best_error = walk_errors[i + 2];
      walk_best = i + 2;
    } else {
      break;
    }
  }
  if (walk_best == walk_limit + 1) return fail(CTAS_E_AXIS, "Failed to calculate rotation centre.");
  *result = mid + 0.5f * improve_me(1.0f, walk_offsets, walk_errors, walk_best, dir, calc_err, &rc);
  return dev_rc != CTAS_OK ? dev_rc : rc;
}

// raxis(sino, anarc, maxDev, algo) (lct.cpp:1133-1212). The trial
// reconstructions of the sweep run as one batch; the directional walk is
// evaluated in batches and cut at its first non-improving trial, which is
// the sequential order of AGinThread::exexafter2 (lct.cpp:1028-1040); the
// bisection is sequential.
extern "C" int ctas_raxis(const float *sino, int N, int w, float anarc, float maxDev, int algo, float *result) {
  if (!sino || !result) return fail(CTAS_E_ARG, "null argument");
  if (N < 3 || w < 3)
    return fail(CTAS_E_AXIS, "Shape of input sinogram (" + std::to_string(N) + "," + std::to_string(w) +
                                 ") is less than minimal (3,3).");
  const float arc = arc_rad(anarc, N);
  if (arc == 0.0f) return fail(CTAS_E_AXIS, "Zero sinogram arc.");
  if (std::fabs(arc) < static_cast<float>(M_PI))
    return fail(CTAS_E_AXIS, "Sinogram arc " + std::to_string(std::fabs(arc) * 180 / M_PI) + " is less than 180 degree.");
  const float kShiftLimit = 0.25f;
  if (maxDev < 0 || maxDev > kShiftLimit)
    return fail(CTAS_E_AXIS, "Invalid maximum relative shift of the rotation axis " + std::to_string(maxDev) + ".");
  const float mid = static_cast<float>(0.5 * (w - 1));
  float seed = 0.0f;
  if (algo == 0 || algo == 2) {
    const int half_turn = static_cast<int>((N - 1) * M_PI / std::fabs(arc_rad(anarc, N)));
    const int n_rows = N - half_turn;
    if (n_rows <= 0) return fail(CTAS_E_AXIS, "No opposite projections in the sinogram.");
    std::vector<float> r0(static_cast<size_t>(n_rows) * w), r180(static_cast<size_t>(n_rows) * w);
    for (int j = 0; j < n_rows; ++j)
      for (int x = 0; x < w; ++x) {
        r0[static_cast<size_t>(j) * w + x] = sino[static_cast<size_t>(j) * w + x] + 1;
        r180[static_cast<size_t>(j) * w + x] = sino[static_cast<size_t>(j + half_turn) * w + x] + 1;
      }
    float q = 0.0f;
    if (int rc = ctas_raxis_pair(r0.data(), r180.data(), n_rows, w, std::min(maxDev, 0.5f), &q)) return rc;
    if (algo == 0) {
      *result = q;
      return CTAS_OK;
    }
    seed = q;
  } else if (algo == 1) {
    seed = mid;
  } else if (algo == 3) {
    std::vector<float> profile(w, 0.0f);
    for (int j = 0; j < N; ++j)
      for (int x = 0; x < w; ++x) profile[x] = profile[x] + sino[static_cast<size_t>(j) * w + x];
    const float mx = *std::max_element(profile.begin(), profile.end());
    const float mn = *std::min_element(profile.begin(), profile.end());
    // lct.cpp:1168-1180 only ends its scans at a sample above the
    // threshold; a constant projection has none.
    if (!(mx > mn)) return fail(CTAS_E_AXIS, "Constant projection: no edge above the threshold.");
    const float threshold = static_cast<float>(0.1 * mx + 0.9 * mn);
    int a = 0, b = w - 1;
    while (!(profile[a] > threshold)) ++a;
    while (!(profile[b] > threshold)) --b;
    seed = (static_cast<float>(a) + static_cast<float>(b)) / 2.0f;
  } else {
    return fail(CTAS_E_AXIS, "Unknown algorithm for rotation axis " + std::to_string(algo) + ".");
  }
  if (std::fabs(seed - mid) / w > maxDev)
    return fail(CTAS_E_AXIS, "Rotation axis relative shift " + std::to_string(std::fabs(seed - mid) / w) +
                                 " is larger than maximum shift " + std::to_string(maxDev) + ".");
  const float box_fraction = static_cast<float>(1.0 - 2.0 * (maxDev + std::fabs(seed - mid) / w));
  if (box_fraction < 0.5f)
    return fail(CTAS_E_AXIS, "Relative box size " + std::to_string(box_fraction) + " is less than miminum window of 0.5.");

  // AverageGradient (lct.cpp:818-844): ABS contrast, RAMP, the sinogram
  // filtered once.
  HandleGuard g;
  if (int rc = ctas_fbp_create(&g.h, w, N, CTAS_CONTRAST_ABS, CTAS_FILTER_RAMP, 0.0f, anarc, 0)) return rc;
  std::vector<float> copy(sino, sino + static_cast<size_t>(N) * w);
  if (int rc = ctas_fbp_sino(g.h, copy.data())) return rc;
  int rc = CTAS_OK;
  auto score_at = [&](float x) -> float {
    float axes = mid + x, s = 0.0f;
    if (!rc) rc = ctas_axis_scores(g.h, &axes, 1, &s);
    return s;
  };

  const int kStep = 8;
  const int n_trials = static_cast<int>(1 + 2 * (maxDev == 0.0f ? 1.0f : maxDev * static_cast<float>(w) / kStep));
  const int half = (n_trials - 1) / 2;
  const float seed_offset = seed - mid;
  std::vector<float> offsets(n_trials), scores(n_trials), axes(n_trials);
  for (int k = 0; k < n_trials; ++k) {
    offsets[k] = seed_offset + static_cast<float>(kStep * (k - half));
    axes[k] = mid + offsets[k];
  }
  if ((rc = ctas_axis_scores(g.h, axes.data(), n_trials, scores.data()))) return rc;
  int best_at = 0;
  float best_score = FLT_MAX;
  for (int i = 1; i < n_trials - 1; ++i)
    if (scores[i] < scores[i - 1] && scores[i] < scores[i + 1] && scores[i] < best_score) {
      best_at = i;
      best_score = scores[i];
    }
  int dir = 1;
  if (!best_at) {
    const int idx = scores[0] < scores[n_trials - 1] ? 0 : n_trials - 1;
    dir = idx == 0 ? -1 : 1;
    const int walk_from = static_cast<int>(offsets[idx - dir * 2]);
    const int walk_limit = static_cast<int>(static_cast<float>(w) * kShiftLimit - static_cast<float>(dir * walk_from / kStep));
    if (walk_limit <= 0) return fail(CTAS_E_AXIS, "Impossible number of n_trials.");
    std::vector<float> walk_offsets, walk_scores;
    bool stop = false;
    constexpr int kWalkChunk = 8;
    for (int k0 = 0; k0 < walk_limit && !stop; k0 += kWalkChunk) {
      const int m = std::min(kWalkChunk, walk_limit - k0);
      std::vector<float> a(m), s(m);
      for (int i = 0; i < m; ++i) a[i] = mid + (static_cast<float>(walk_from) + static_cast<float>(kStep * (k0 + i) * dir));
      if ((rc = ctas_axis_scores(g.h, a.data(), m, s.data()))) return rc;
      for (int i = 0; i < m && !stop; ++i) {
        walk_offsets.push_back(static_cast<float>(walk_from) + static_cast<float>(kStep * (k0 + i) * dir));
        walk_scores.push_back(s[i]);
        if (s[i] < best_score) {
          best_score = s[i];
          best_at = k0 + i;
        } else {
          stop = true;
        }
      }
    }
    if (best_at == walk_limit - 1 || best_at == 0) {
      *result = seed;
      return CTAS_OK;
    }
    offsets = walk_offsets;
    scores = walk_scores;
  }
  const float rf = improve_me(0.25f, offsets, scores, best_at, dir, score_at, &rc);
  if (rc) return rc;
  *result = mid + rf;
  return CTAS_OK;
}
