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


// ---- register-resident radix-8 / radix-16 pieces -----------------------------
// Small DFTs on values held in registers, natural order in and out, forward
// (exp(-2 pi i jk/n)) or inverse (INV, conjugate factors, unscaled).
constexpr float kCos16[16] = {1.0f, 0.92387953f, 0.70710678f, 0.38268343f, 0.0f, -0.38268343f, -0.70710678f,
                              -0.92387953f, -1.0f, -0.92387953f, -0.70710678f, -0.38268343f, 0.0f, 0.38268343f,
                              0.70710678f, 0.92387953f};
constexpr float kSin16[16] = {0.0f, 0.38268343f, 0.70710678f, 0.92387953f, 1.0f, 0.92387953f, 0.70710678f,
                              0.38268343f, 0.0f, -0.38268343f, -0.70710678f, -0.92387953f, -1.0f, -0.92387953f,
                              -0.70710678f, -0.38268343f};

// 4-point DFT, exp(-2 pi i/4) per step (forward) or its conjugate (INV).
template <bool INV>
__device__ __forceinline__ void dft4(float2 &x0, float2 &x1, float2 &x2, float2 &x3) {
  const float2 a0 = cadd(x0, x2), a1 = csub(x0, x2), a2 = cadd(x1, x3), a3 = csub(x1, x3);
  x0 = cadd(a0, a2);
  x2 = csub(a0, a2);
  if (INV) {
    x1 = make_float2(a1.x - a3.y, a1.y + a3.x);
    x3 = make_float2(a1.x + a3.y, a1.y - a3.x);
  } else {
    x1 = make_float2(a1.x + a3.y, a1.y - a3.x);
    x3 = make_float2(a1.x - a3.y, a1.y + a3.x);
  }
}

// The factors exp(-+2 pi i b c / 16) between the two radix-4 stages of dft16.
template <bool INV>
__device__ __forceinline__ void dft16_factors(float2 (&v)[16]) {
#pragma unroll
  for (int b = 1; b < 4; ++b)
#pragma unroll
    for (int c = 1; c < 4; ++c) {
      const float co = kCos16[b * c], si = INV ? kSin16[b * c] : -kSin16[b * c];
      v[b + 4 * c] = cmul(v[b + 4 * c], make_float2(co, si));
    }
}

// 16-point DFT in registers, natural order in and out: j = 4a + b,
// k = c + 4d; inner DFT4 over a, factor exp(-+2 pi i b c / 16), outer DFT4
// over b.
template <bool INV>
__device__ __forceinline__ void dft16(float2 (&v)[16]) {
#pragma unroll
  for (int b = 0; b < 4; ++b) dft4<INV>(v[b], v[4 + b], v[8 + b], v[12 + b]);  // v[b + 4c] = u_b[c]
  dft16_factors<INV>(v);
  float2 r[16];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    float2 y0 = v[4 * c], y1 = v[4 * c + 1], y2 = v[4 * c + 2], y3 = v[4 * c + 3];
    dft4<INV>(y0, y1, y2, y3);  // over b -> d
    r[c] = y0;
    r[c + 4] = y1;
    r[c + 8] = y2;
    r[c + 12] = y3;
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = r[k];
}

// dft16 of points whose upper half v[8..15] is zero (not read): the inner
// 4-point DFTs over (v[b], v[4 + b], 0, 0) are sums and differences.
template <bool INV>
__device__ __forceinline__ void dft16_lo8(float2 (&v)[16]) {
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const float2 x0 = v[b], x1 = v[4 + b];
    v[b] = cadd(x0, x1);
    v[8 + b] = csub(x0, x1);
    const float2 p = make_float2(x0.x + x1.y, x0.y - x1.x), m = make_float2(x0.x - x1.y, x0.y + x1.x);
    v[4 + b] = INV ? m : p;
    v[12 + b] = INV ? p : m;
  }
  dft16_factors<INV>(v);
  float2 r[16];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    float2 y0 = v[4 * c], y1 = v[4 * c + 1], y2 = v[4 * c + 2], y3 = v[4 * c + 3];
    dft4<INV>(y0, y1, y2, y3);
    r[c] = y0;
    r[c + 4] = y1;
    r[c + 8] = y2;
    r[c + 12] = y3;
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = r[k];
}

// dft16 of which only the outputs v[0..7] are formed (v[8..15] are left
// undefined): the outer 4-point DFTs give their first two outputs.
template <bool INV>
__device__ __forceinline__ void dft16_out8(float2 (&v)[16]) {
#pragma unroll
  for (int b = 0; b < 4; ++b) dft4<INV>(v[b], v[4 + b], v[8 + b], v[12 + b]);
  dft16_factors<INV>(v);
  float2 r[8];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float2 y0 = v[4 * c], y1 = v[4 * c + 1], y2 = v[4 * c + 2], y3 = v[4 * c + 3];
    const float2 a0 = cadd(y0, y2), a1 = csub(y0, y2), a2 = cadd(y1, y3), a3 = csub(y1, y3);
    r[c] = cadd(a0, a2);
    r[c + 4] = INV ? make_float2(a1.x - a3.y, a1.y + a3.x) : make_float2(a1.x + a3.y, a1.y - a3.x);
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = r[k];
}

// v[k] *= r^k, k = 1..15, with r^2, r^4, r^8 from squarings and the other
// powers as products of those (at most four factors, a few float ulps).
__device__ __forceinline__ void twiddle_powers(float2 (&v)[16], float2 r) {
  float2 p[16];
  p[1] = r;
  p[2] = cmul(r, r);
  p[4] = cmul(p[2], p[2]);
  p[8] = cmul(p[4], p[4]);
  p[3] = cmul(p[2], r);
  p[5] = cmul(p[4], r);
  p[6] = cmul(p[4], p[2]);
  p[7] = cmul(p[4], p[3]);
  p[9] = cmul(p[8], r);
  p[10] = cmul(p[8], p[2]);
  p[11] = cmul(p[8], p[3]);
  p[12] = cmul(p[8], p[4]);
  p[13] = cmul(p[8], p[5]);
  p[14] = cmul(p[8], p[6]);
  p[15] = cmul(p[8], p[7]);
#pragma unroll
  for (int k = 1; k < 16; ++k) v[k] = cmul(v[k], p[k]);
}


// 8-point DFT in registers: two 4-point DFTs of the even and odd points,
// combined with exp(-+2 pi i k / 8).
template <bool INV>
__device__ __forceinline__ void dft8(float2 (&v)[8]) {
  float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4<INV>(e0, e1, e2, e3);
  dft4<INV>(o0, o1, o2, o3);
  constexpr float h = 0.70710678f;
  const float2 w1 = make_float2(h, INV ? h : -h), w2 = make_float2(0.0f, INV ? 1.0f : -1.0f),
               w3 = make_float2(-h, INV ? h : -h);
  o1 = cmul(o1, w1);
  o2 = cmul(o2, w2);
  o3 = cmul(o3, w3);
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1);
  v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2);
  v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3);
  v[7] = csub(e3, o3);
}

// v[k] *= r^k, k = 1..7 (powers by squarings and products).
__device__ __forceinline__ void twiddle_powers8(float2 (&v)[8], float2 r) {
  const float2 p2 = cmul(r, r), p4 = cmul(p2, p2), p3 = cmul(p2, r);
  v[1] = cmul(v[1], r);
  v[2] = cmul(v[2], p2);
  v[3] = cmul(v[3], p3);
  v[4] = cmul(v[4], p4);
  v[5] = cmul(v[5], cmul(p4, r));
  v[6] = cmul(v[6], cmul(p4, p2));
  v[7] = cmul(v[7], cmul(p4, p3));
}

// 32-point DFT in registers: two 16-point DFTs of the even and odd points,
// combined with exp(-+2 pi i k / 32).
constexpr float kCos32[16] = {1.0f, 0.98078528f, 0.92387953f, 0.83146961f, 0.70710678f, 0.55557023f,
                              0.38268343f, 0.19509032f, 0.0f, -0.19509032f, -0.38268343f, -0.55557023f,
                              -0.70710678f, -0.83146961f, -0.92387953f, -0.98078528f};
constexpr float kSin32[16] = {0.0f, 0.19509032f, 0.38268343f, 0.55557023f, 0.70710678f, 0.83146961f,
                              0.92387953f, 0.98078528f, 1.0f, 0.98078528f, 0.92387953f, 0.83146961f,
                              0.70710678f, 0.55557023f, 0.38268343f, 0.19509032f};

template <bool INV>
__device__ __forceinline__ void dft32(float2 (&v)[32]) {
  float2 e[16], o[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    e[j] = v[2 * j];
    o[j] = v[2 * j + 1];
  }
  dft16<INV>(e);
  dft16<INV>(o);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const float2 t = cmul(o[k], make_float2(kCos32[k], INV ? kSin32[k] : -kSin32[k]));
    v[k] = cadd(e[k], t);
    v[k + 16] = csub(e[k], t);
  }
}

// dft32 of points whose upper half v[16..31] is zero (not read): the even
// and the odd points each have their upper half zero.
template <bool INV>
__device__ __forceinline__ void dft32_lo16(float2 (&v)[32]) {
  float2 e[16], o[16];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    e[j] = v[2 * j];
    o[j] = v[2 * j + 1];
  }
  dft16_lo8<INV>(e);
  dft16_lo8<INV>(o);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const float2 t = cmul(o[k], make_float2(kCos32[k], INV ? kSin32[k] : -kSin32[k]));
    v[k] = cadd(e[k], t);
    v[k + 16] = csub(e[k], t);
  }
}

// dft32 of which only the outputs v[0..15] are formed.
template <bool INV>
__device__ __forceinline__ void dft32_out16(float2 (&v)[32]) {
  float2 e[16], o[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    e[j] = v[2 * j];
    o[j] = v[2 * j + 1];
  }
  dft16<INV>(e);
  dft16<INV>(o);
#pragma unroll
  for (int k = 0; k < 16; ++k)
    v[k] = cadd(e[k], cmul(o[k], make_float2(kCos32[k], INV ? kSin32[k] : -kSin32[k])));
}

// exp(-2 pi i m / 8192) from the half-turn table tw[m] (m < 4096).
__device__ __forceinline__ float2 w8192(const float2 *__restrict__ tw, int m) {
  m &= 8191;
  const float2 t = tw[m & 4095];
  return (m & 4096) ? make_float2(-t.x, -t.y) : t;
}

}  // namespace
}  // namespace ctas
