// Radix-2 FFT filtering of one complex line held in LDS, shared by the
// sinogram filter and the tomosynthesis kernel (256 threads per workgroup).
//
// The line buf[0..Z) (Z = 2^logZ >= 2) is loaded by the caller in
// bit-reversed order. line_filter transforms it by a decimation-in-time FFT
// whose radix-2 stages are fused in pairs (four points per thread per pass),
// multiplies the spectrum, in natural order, by the real window W, and
// transforms back by a decimation-in-frequency FFT with conjugate twiddles,
// which leaves Z times the filtered line in bit-reversed order again. A real
// and even window filters the real and imaginary parts independently, so two
// real rows travel as one complex line. tw[m] = exp(-2 pi i m / Z), m < Z/2.
// Every stage ends with a workgroup barrier, so the result can be read by
// any thread on return.
//
// Point i of the line lives at buf[lds_pad(i, psh)] = buf[i + (i >> psh)]:
// one spare entry per 2^psh points spreads the strided accesses of the
// early and late stages (four points 2^(s-1) apart, consecutive groups
// 2^(s+1) apart) over the LDS banks. psh >= logZ means no padding.
#pragma once

#include <hip/hip_runtime.h>

namespace ctas {
namespace {

__device__ __forceinline__ int lds_pad(int i, int psh) { return i + (i >> psh); }
// Entries of a line of Z points with one spare per 2^psh points.
__host__ __device__ inline int padded_len(int Z, int psh) { return Z + (psh < 31 ? (Z >> psh) : 0); }
// Padding used for a line of Z points: none at Z = 8192, whose 64 KiB line
// fills a workgroup's LDS.
__host__ __device__ inline int line_psh(int Z) { return Z >= 8192 ? 31 : 4; }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

__device__ __forceinline__ void line_filter(float2 *buf, const float2 *__restrict__ tw,
                                            const float *__restrict__ W, int Z, int logZ, int tid,
                                            int psh) {
  // Forward DIT, stages s = 1..logZ (len 2^s), fused in pairs.
  for (int s = 1; s <= logZ;) {
    if (s < logZ) {
      const int h = 1 << (s - 1);
      const int sh1 = logZ - s, sh2 = logZ - s - 1;
      for (int g = tid; g < (Z >> 2); g += 256) {
        const int i = g & (h - 1);
        const int base = ((g >> (s - 1)) << (s + 1)) + i;
        float2 x0 = buf[lds_pad(base, psh)], x1 = buf[lds_pad(base + h, psh)];
        float2 x2 = buf[lds_pad(base + 2 * h, psh)], x3 = buf[lds_pad(base + 3 * h, psh)];
        const float2 w1 = tw[i << sh1];
        float2 y = cmul(x1, w1);
        x1 = csub(x0, y);
        x0 = cadd(x0, y);
        y = cmul(x3, w1);
        x3 = csub(x2, y);
        x2 = cadd(x2, y);
        const float2 w2 = tw[i << sh2];
        const float2 w3 = tw[(i + h) << sh2];
        y = cmul(x2, w2);
        x2 = csub(x0, y);
        x0 = cadd(x0, y);
        y = cmul(x3, w3);
        x3 = csub(x1, y);
        x1 = cadd(x1, y);
        buf[lds_pad(base, psh)] = x0;
        buf[lds_pad(base + h, psh)] = x1;
        buf[lds_pad(base + 2 * h, psh)] = x2;
        buf[lds_pad(base + 3 * h, psh)] = x3;
      }
      s += 2;
    } else {
      const int h = 1 << (s - 1);
      const int sh1 = logZ - s;
      for (int g = tid; g < (Z >> 1); g += 256) {
        const int i = g & (h - 1);
        const int base = ((g >> (s - 1)) << s) + i;
        const float2 x0 = buf[lds_pad(base, psh)];
        const float2 y = cmul(buf[lds_pad(base + h, psh)], tw[i << sh1]);
        buf[lds_pad(base, psh)] = cadd(x0, y);
        buf[lds_pad(base + h, psh)] = csub(x0, y);
      }
      s += 1;
    }
    __syncthreads();
  }

  // Spectrum (natural order) times the real window.
  for (int k = tid; k < Z; k += 256) {
    const float wk = W[k];
    float2 v = buf[lds_pad(k, psh)];
    v.x *= wk;
    v.y *= wk;
    buf[lds_pad(k, psh)] = v;
  }
  __syncthreads();

  // Inverse DIF, stages len = 2^s for s = logZ..1, fused in pairs; conjugate
  // twiddles; output in bit-reversed order.
  for (int s = logZ; s >= 1;) {
    if (s >= 2) {
      const int h = 1 << (s - 2);  // len 4h, then 2h
      const int sh4 = logZ - s, sh2 = logZ - s + 1;
      for (int g = tid; g < (Z >> 2); g += 256) {
        const int i = g & (h - 1);
        const int base = ((g >> (s - 2)) << s) + i;
        float2 x0 = buf[lds_pad(base, psh)], x1 = buf[lds_pad(base + h, psh)];
        float2 x2 = buf[lds_pad(base + 2 * h, psh)], x3 = buf[lds_pad(base + 3 * h, psh)];
        float2 d = csub(x0, x2);
        x0 = cadd(x0, x2);
        x2 = cmulc(d, tw[i << sh4]);
        d = csub(x1, x3);
        x1 = cadd(x1, x3);
        x3 = cmulc(d, tw[(i + h) << sh4]);
        const float2 w2 = tw[i << sh2];
        d = csub(x0, x1);
        x0 = cadd(x0, x1);
        x1 = cmulc(d, w2);
        d = csub(x2, x3);
        x2 = cadd(x2, x3);
        x3 = cmulc(d, w2);
        buf[lds_pad(base, psh)] = x0;
        buf[lds_pad(base + h, psh)] = x1;
        buf[lds_pad(base + 2 * h, psh)] = x2;
        buf[lds_pad(base + 3 * h, psh)] = x3;
      }
      s -= 2;
    } else {
      const int h = 1;  // len 2, twiddle 1
      for (int g = tid; g < (Z >> 1); g += 256) {
        const int base = g << 1;
        const float2 x0 = buf[lds_pad(base, psh)], x1 = buf[lds_pad(base + h, psh)];
        buf[lds_pad(base, psh)] = cadd(x0, x1);
        buf[lds_pad(base + h, psh)] = csub(x0, x1);
      }
      s -= 1;
    }
    __syncthreads();
  }

}

}  // namespace
}  // namespace ctas
