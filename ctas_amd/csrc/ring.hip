// Ring-artefact filter of sinograms (RingFilter::apply, lct.cpp:295-327):
// subtract from every row the difference between the column mean over
// angles and its box average of half-width `box`. One thread per column;
// the column sums run over the rows in order, and the box sum and the edge
// divisions follow the reference's statement order, so the result matches a
// sequential float implementation bit for bit.
#include <hip/hip_runtime.h>

#include "ctas_internal.hpp"

namespace ctas {
namespace {

__global__ __launch_bounds__(256) void column_mean_kernel(const float *__restrict__ sino, float *__restrict__ avg,
                                                          int N, int w) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (x >= w) return;
  const float *p = sino + static_cast<size_t>(s) * N * w + x;
  float a = 0.0f;
  for (int r = 0; r < N; ++r) a = a + p[static_cast<size_t>(r) * w];
  avg[static_cast<size_t>(s) * w + x] = a / static_cast<float>(N);
}

__global__ __launch_bounds__(256) void ring_apply_kernel(float *__restrict__ sino, const float *__restrict__ avg,
                                                         int N, int w, int box) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (x >= w) return;
  const float *a = avg + static_cast<size_t>(s) * w;
  float t = 0.0f;
  for (int ii = -box; ii <= box; ++ii) {
    // dstR(w, ii) = [max(0, ii), w-1+min(0, ii)], source x - ii
    const int d0 = ii > 0 ? ii : 0, d1 = w - 1 + (ii < 0 ? ii : 0);
    if (x >= d0 && x <= d1) t = t + a[x - ii];
  }
  if (x >= box && x <= w - 1 - box) t = t / static_cast<float>(2 * box + 1);
  for (int ii = 0; ii < box; ++ii) {
    if (x == ii) t = t / static_cast<float>(box + ii + 1);
    if (x == w - 1 - ii) t = t / static_cast<float>(box + ii + 1);
  }
  t = t - a[x];
  float *p = sino + static_cast<size_t>(s) * N * w + x;
  for (int r = 0; r < N; ++r) p[static_cast<size_t>(r) * w] = p[static_cast<size_t>(r) * w] + t;
}

}  // namespace

hipError_t launch_ring(float *d_sino, float *d_avg, int S, int N, int w, int box, hipStream_t s) {
  const dim3 g((w + 255) / 256, S);
  hipLaunchKernelGGL(column_mean_kernel, g, dim3(256), 0, s, d_sino, d_avg, N, w);
  hipLaunchKernelGGL(ring_apply_kernel, g, dim3(256), 0, s, d_sino, d_avg, N, w, box);
  return hipGetLastError();
}

}  // namespace ctas
