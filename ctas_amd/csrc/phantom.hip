// Synthetic sinograms of random ellipse phantoms: analytic line integrals in
// FP64 at detector bin centres, optional photon noise and per-detector gain.
// Geometry of the backprojector: pixel (j, i) sits at x = i - hp,
// y = j - hp - c; the ray of angle theta = p*pi/N through bin k has
// coordinate s = k + 1/2 - hp - c along (cos theta, sin theta).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "ctas_internal.hpp"

namespace ctas {
namespace {

__host__ __device__ inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__host__ __device__ inline double unit(uint64_t h) { return (static_cast<double>(h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

__global__ __launch_bounds__(256) void phantom_kernel(int N, int w, double c, int n_ell, const double *__restrict__ ell,
                                                      double N0, uint64_t seed, const float *__restrict__ gain,
                                                      float *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (k >= w) return;
  const int hp = w / 2;
  const double th = p * M_PI / N;
  double st, ct;
  sincos(th, &st, &ct);
  const double s = k + 0.5 - hp - c;
  double acc = 0.0;
  for (int e = 0; e < n_ell; ++e) {
    const double *E = ell + 6 * e;
    const double s0 = E[0] * ct + E[1] * st;
    double sp, cp;
    sincos(th - E[4], &sp, &cp);
    const double r2 = E[2] * E[2] * cp * cp + E[3] * E[3] * sp * sp;
    const double ds = s - s0;
    if (ds * ds < r2) acc += 2.0 * E[5] * E[2] * E[3] * sqrt(r2 - ds * ds) / r2;
  }
  double v = acc;
  if (N0 > 0) {
    const double lam = N0 * exp(-acc);
    const uint64_t h1 = mix64(seed ^ mix64(static_cast<uint64_t>(p) * static_cast<uint64_t>(w) + k));
    const uint64_t h2 = mix64(h1);
    double cnt;
    if (lam < 64.0) {  // Poisson by multiplication of uniforms
      const double L = exp(-lam);
      double prod = unit(h1);
      uint64_t st2 = h2;
      cnt = 0;
      while (prod > L && cnt < 1000) {
        st2 = mix64(st2);
        prod *= unit(st2);
        cnt += 1;
      }
    } else {  // normal approximation of the Poisson law
      const double z = sqrt(-2.0 * log(unit(h1))) * cos(2.0 * M_PI * unit(h2));
      cnt = floor(lam + sqrt(lam) * z + 0.5);
      if (cnt < 0) cnt = 0;
    }
    v = cnt / N0;
  }
  if (gain) v *= gain[k];
  out[static_cast<size_t>(p) * w + k] = static_cast<float>(v);
}

}  // namespace

hipError_t launch_phantom(int N, int w, double c, int n_ell, const double *d_ell, double N0, uint64_t seed,
                          const float *d_gain, float *d_out, hipStream_t s) {
  hipLaunchKernelGGL(phantom_kernel, dim3((w + 255) / 256, N), dim3(256), 0, s, N, w, c, n_ell, d_ell, N0, seed,
                     d_gain, d_out);
  return hipGetLastError();
}

}  // namespace ctas

using namespace ctas;

extern "C" int ctas_phantom_ellipses(int n_ell, int w, double axis_scale, uint64_t seed, double *ell) {
  if (n_ell < 0 || w < 2 || !ell) return fail(CTAS_E_ARG, "bad arguments");
  uint64_t st = seed;
  auto next = [&]() {
    st = mix64(st);
    return unit(st);
  };
  const double R = 0.5 * w;
  for (int e = 0; e < n_ell; ++e) {
    // centre uniform in the disc of radius 0.8 R
    const double rr = 0.8 * R * std::sqrt(next());
    const double ang = 2.0 * M_PI * next();
    ell[6 * e + 0] = rr * std::cos(ang);
    ell[6 * e + 1] = rr * std::sin(ang);
    ell[6 * e + 2] = (5.0 + 195.0 * next()) * axis_scale;
    ell[6 * e + 3] = (5.0 + 195.0 * next()) * axis_scale;
    ell[6 * e + 4] = M_PI * next();
    ell[6 * e + 5] = (-0.5 + 1.5 * next()) * (2.0 / w);
  }
  return CTAS_OK;
}

extern "C" int ctas_phantom_sino_dev(int N, int w, double c, int n_ell, const double *ell, double N0, uint64_t seed,
                                     const float *d_gain, float *d_out, void *stream) {
  if (N <= 0 || w < 2 || n_ell < 0 || (n_ell && !ell) || !d_out) return fail(CTAS_E_ARG, "bad arguments");
  double *d_ell = nullptr;
  const size_t bytes = sizeof(double) * 6 * static_cast<size_t>(n_ell > 0 ? n_ell : 1);
  if (hipMalloc(&d_ell, bytes) != hipSuccess) return fail(CTAS_E_NOMEM, "allocation failed");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc = CTAS_OK;
  if (n_ell && hipMemcpy(d_ell, ell, sizeof(double) * 6 * n_ell, hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(CTAS_E_DEVICE, "copy failed");
  if (!rc && launch_phantom(N, w, c, n_ell, d_ell, N0, seed, d_gain, d_out, s) != hipSuccess)
    rc = fail(CTAS_E_DEVICE, "phantom launch failed");
  (void)hipStreamSynchronize(s);
  (void)hipFree(d_ell);
  return rc;
}
