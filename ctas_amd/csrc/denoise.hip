// Denoiser::proc (matrix.world.cpp:952-1037) without a mask: each pixel's
// mean over the disc of offsets (i, j), i^2 + j^2 <= rad^2, that stay inside
// the image, summed in the reference's offset order (i outer, j inner, both
// from -rad), times the float reciprocal of their count (invert, 1.0 / n in
// double, matrix.world.cpp:939). The pixel becomes that mean always
// (threshold 0), or when it is 0 or differs from it by more than the
// threshold (absolute for threshold > 0, relative to the mean for < 0).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "ctas_internal.hpp"

namespace ctas {
namespace {

__global__ __launch_bounds__(256) void denoise_kernel(const float *__restrict__ in, float *__restrict__ out, int H,
                                                      int W, int rad, float thr) {
#pragma clang fp contract(off)
  const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
  if (x >= W) return;
  const int r2 = rad * rad;
  float acc = 0.0f;
  int cnt = 0;
  for (int i = -rad; i <= rad; ++i) {
    const int sy = y - i;
    for (int j = -rad; j <= rad; ++j) {
      const int sx = x - j;
      if (i * i + j * j <= r2 && sy >= 0 && sy < H && sx >= 0 && sx < W) {
        acc = acc + in[static_cast<size_t>(sy) * W + sx];
        ++cnt;
      }
    }
  }
  const float wgt = cnt ? static_cast<float>(1.0 / static_cast<double>(cnt)) : 0.0f;
  const float t = acc * wgt;
  const size_t at = static_cast<size_t>(y) * W + x;
  float v = in[at];
  if (thr == 0.0f) {
    v = t;
  } else if (thr < 0.0f) {
    if (v == 0.0f || (t != 0.0f && fabsf((v - t) / t) > -thr)) v = t;
  } else {
    if (v == 0.0f || fabsf(v - t) > thr) v = t;
  }
  out[at] = v;
}

}  // namespace
}  // namespace ctas

using namespace ctas;

extern "C" {

int ctas_denoise_dev(const float *d_in, float *d_out, int rows, int cols, int rad, float threshold, void *stream) {
  if (!d_in || !d_out || rows < 0 || cols < 0) return fail(CTAS_E_ARG, "bad arguments");
  if (d_in == d_out) return fail(CTAS_E_ARG, "denoiser: input and output must be different arrays");
  rad = std::abs(rad);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!rows || !cols) return CTAS_OK;
  if (!rad) {
    CTAS_HIP_TRY(hipMemcpyAsync(d_out, d_in, sizeof(float) * rows * static_cast<size_t>(cols),
                                hipMemcpyDeviceToDevice, s));
    return CTAS_OK;
  }
  hipLaunchKernelGGL(denoise_kernel, dim3((cols + 255) / 256, rows), dim3(256), 0, s, d_in, d_out, rows, cols, rad,
                     threshold);
  CTAS_HIP_TRY(hipGetLastError());
  return CTAS_OK;
}

int ctas_denoise(float *image, int rows, int cols, int rad, float threshold) {
  if (!image || rows < 0 || cols < 0) return fail(CTAS_E_ARG, "bad arguments");
  if (!rows || !cols || !rad) return CTAS_OK;  // Denoiser::proc returns early
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(CTAS_E_DEVICE, "no HIP device is available");
  const size_t bytes = sizeof(float) * rows * static_cast<size_t>(cols);
  float *d = nullptr;
  if (hipMalloc(&d, 2 * bytes) != hipSuccess) return fail(CTAS_E_NOMEM, "device allocation failed");
  int rc = CTAS_OK;
  if (hipMemcpy(d, image, bytes, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(CTAS_E_DEVICE, "denoiser: upload failed");
  if (rc == CTAS_OK) rc = ctas_denoise_dev(d, d + bytes / sizeof(float), rows, cols, rad, threshold, nullptr);
  if (rc == CTAS_OK && (hipDeviceSynchronize() != hipSuccess ||
                        hipMemcpy(image, d + bytes / sizeof(float), bytes, hipMemcpyDeviceToHost) != hipSuccess))
    rc = fail(CTAS_E_DEVICE, "denoiser failed on the device");
  (void)hipFree(d);
  return rc;
}

}  // extern "C"
