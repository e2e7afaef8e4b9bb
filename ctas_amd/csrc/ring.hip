// Ring-artefact filter of sinograms (RingFilter::apply, lct.cpp:295-327):
// subtract from every row the difference between the column mean over
// angles and its box average of half-width `box`. One thread per column;
// the column sums run over the rows in order, and the box sum and the edge
// divisions follow the reference's statement order, so the result matches a
// sequential float implementation bit for bit.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "ctas_internal.hpp"

namespace ctas {
namespace {

// Column means of ctas_ring_filter_dev_noalloc: a file-scope device array,
// one copy per GPU, held until the process ends. Host code never passes the
// array's name to a kernel (on the host it is a stand-in, not a device
// address): each device's address comes from hipGetSymbolAddress, taken
// once per device with that device current.
constexpr long long kRingScratchFloats = 1LL << 20;  // 4 MiB: S * w <= 1,048,576
__device__ __align__(256) float scratch[kRingScratchFloats];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kMaxDevices] = {};

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

using namespace ctas;

extern "C" int ctas_ring_filter_dev_noalloc(float *d_sino, int S, int N, int w, int box, void *stream) {
  if (!d_sino || S <= 0 || N <= 0 || w <= 0 || box < 0) return fail(CTAS_E_ARG, "bad arguments");
  if (static_cast<long long>(S) * w > kRingScratchFloats)
    return fail(CTAS_E_ARG, "ring filter without allocation: " + std::to_string(S) + " slices x " +
                                std::to_string(w) + " columns exceed its " + std::to_string(kRingScratchFloats) +
                                " column means of scratch");
  if (box == 0) return CTAS_OK;
  if (int rc = check_ring_box(box, w)) return rc;
  int dev = 0;
  CTAS_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(CTAS_E_DEVICE, "device index " + std::to_string(dev) + " out of range");
  float *avg = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      void *p = nullptr;
      CTAS_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      g_scratch_ptr[dev] = static_cast<float *>(p);
    }
    avg = g_scratch_ptr[dev];
  }
  // column_mean_kernel writes every mean it reads back: nothing to clear
  const hipError_t e = launch_ring(d_sino, avg, S, N, w, box, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return fail(CTAS_E_DEVICE, std::string("ring filter: ") + hipGetErrorString(e));
  return CTAS_OK;
}
