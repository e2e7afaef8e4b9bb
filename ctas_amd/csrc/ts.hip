// Tomosynthesis: ts_add (lct.cpp:739-787) on the device.
//
// For one projection (rows x n) at angle a, the reference takes -log of the
// whole array (ABS contrast), filters every row in place with the window
// Filter::fill(n) by n-point FFTW transforms (and takes the running sum of
// the row for REF contrast), and then adds to result(y, x) the filtered
// sample projection(y, (long)d(x)), where
//   d(x) = (x - R) * cos(a) - P + R,  R = n/2 + center,
//   P = (plane - center) * sin(a),
// is evaluated in float and clamped to [0, n - 1]. The column map d(x) is
// the same for every row.
//
// Layout: a workgroup of 256 threads owns two rows (or four) and walks
// through a batch of projections in order. Two rows travel through the FFT
// as the real and imaginary parts of one complex line (the window is real
// and even); the filtered line ends in LDS, so the resampling reads it there
// and the only HBM traffic per projection and pixel is the raw sample in
// and, when asked for, the filtered sample out. Thread t keeps its result
// columns t + 256 j in registers across the whole batch, so the result is
// read and written once per batch; the per-pixel sums happen in projection
// order, as successive ts_add calls make them. The kernels:
//   ts2048_kernel  2048-point lines, four rows, FFT in registers (8x16x16);
//   ts4096_kernel  4096-point lines, two rows, FFT in registers (16x16x16);
//   ts8192_kernel  8192-point lines of widths up to 4096, two rows, FFT in
//                  registers (32x16x16);
//   ts_kernel      other lengths, two rows, radix-2 FFT in LDS, up to four
//                  projections in flight per workgroup.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "ctas_internal.hpp"
#include "line_fft.hpp"

namespace ctas {
namespace {

// Column index of ts_add's resampling (lct.cpp:779-782), float arithmetic
// in the reference's order, each product and sum rounded on its own.
__device__ __forceinline__ int ts_column(int x, float cs, float pc, float rc, int n) {
#pragma clang fp contract(off)
  float di = (static_cast<float>(x) - rc) * cs - pc + rc;
  di = (di < 0.0f) ? 0.0f : (di >= static_cast<float>(n)) ? static_cast<float>(n - 1) : di;
  int i = static_cast<int>(di);
  // A NaN d(x) (non-finite angle or centre) converts to an unspecified
  // integer; keep the LDS read inside the line.
  i = i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
  return i;
}

// Odd widths. FFTW's half-complex product of filter_line (lct.cpp:280-291)
// multiplies the real part of bin k by W[k] and its imaginary part by
// W[n-k]. Filter::fill's window is even except at the middle bins of an odd
// n: f = 2(n-k)/n at k = m = (n-1)/2 and 2(n-m-1)/n at k = m+1
// (lct.cpp:147-148), so W[m] != W[m+1]. The device runs the circular
// convolution of the even window, whose bin m takes their mean; the
// difference is the rank-two term
//   y[x] += (W[m] - W[m+1])/n (Re X_m cos(phi x) + Im X_m sin(phi x)),
//   phi = 2 pi m/n, X_m = sum_t x[t] exp(-i phi t).
// With u[t] = (-1)^t cos(pi t/n) = cos(phi t) and v[t] = (-1)^t sin(pi t/n)
// = -sin(phi t): y[x] += g (S_u u[x] - S_v v[x]), S_u = sum_t x[t] u[t],
// S_v = sum_t x[t] v[t]; g = Z (W[m] - W[m+1])/n for a line scaled by Z.
// `odd` holds (u[x], v[x]), x < n, for odd n >= 3 and is null otherwise.
//
// Sums S_u, S_v of the rows r0 and r1 of one projection (after -log for
// ABS): thread lt of a group of 256, reduced over its wave; lane 0 of wave
// w leaves {S_u row 0, S_u row 1, S_v row 0, S_v row 1} at stash[4 w].
__device__ __forceinline__ void odd_sums(const float *r0, const float *r1, bool has0, bool has1, int n,
                                         bool absorb, const float2 *__restrict__ odd, int lt, float *stash) {
  float su0 = 0.0f, su1 = 0.0f, sv0 = 0.0f, sv1 = 0.0f;
  for (int x = lt; x < n; x += 256) {
    float a = has0 ? r0[x] : 1.0f, c = has1 ? r1[x] : 1.0f;
    if (absorb) {
      a = -__logf(a);
      c = -__logf(c);
    }
    a = has0 ? a : 0.0f;
    c = has1 ? c : 0.0f;
    const float2 uv = odd[x];
    su0 += a * uv.x;
    su1 += c * uv.x;
    sv0 += a * uv.y;
    sv1 += c * uv.y;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    su0 += __shfl_xor(su0, o);
    su1 += __shfl_xor(su1, o);
    sv0 += __shfl_xor(sv0, o);
    sv1 += __shfl_xor(sv1, o);
  }
  if ((lt & 63) == 0) {
    float *d = stash + 4 * (lt >> 6);
    d[0] = su0;
    d[1] = su1;
    d[2] = sv0;
    d[3] = sv1;
  }
}

// The four waves' stashes summed: {S_u row 0, S_u row 1, S_v row 0, S_v row 1}.
__device__ __forceinline__ float4 odd_total(const float *stash) {
  float4 t = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    t.x += stash[4 * w];
    t.y += stash[4 * w + 1];
    t.z += stash[4 * w + 2];
    t.w += stash[4 * w + 3];
  }
  return t;
}

// The correction of point x (< n) of a row pair's line scaled by Z.
__device__ __forceinline__ float2 odd_term(const float2 *__restrict__ odd, float g, float4 t, int x) {
  const float2 uv = odd[x];
  return make_float2(g * (t.x * uv.x - t.z * uv.y), g * (t.y * uv.x - t.w * uv.y));
}

// K projections are filtered at once, one per 256-thread quarter of the
// workgroup, each in its own line of LDS; all 256 K threads then add them
// to their columns (x = tid + 256 K j, j < J) in projection order.
template <int J, int K, bool TAB>
__global__ __launch_bounds__(256 * K) void ts_kernel(const float *proj, long long proj_stride, int n_proj,
                                                     TsConsts consts, float *filtered,
                                                     float *__restrict__ result, const float *__restrict__ W,
                                                     const float2 *__restrict__ tw, int rows, int n, int Z,
                                                     int logZ, int contrast, const float2 *__restrict__ odd,
                                                     float odd_g) {
  extern __shared__ float2 lines[];
  const int tid = threadIdx.x;
  const int sub = tid >> 8, lt = tid & 255;  // quarter and thread within it
  const int psh = line_psh(Z), L = padded_len(Z, psh);  // line stride in LDS
  float2 *buf = lines + sub * L;
  // TAB: the twiddles and the window, read at every stage of every
  // projection, are staged in LDS behind the lines.
  const float2 *twp = tw;
  const float *Wp = W;
  // odd widths: 16 floats of wave sums per quarter, behind the tables
  float *ostash = reinterpret_cast<float *>(lines + K * L + (TAB ? Z : 0)) + 16 * sub;
  if constexpr (TAB) {
    float2 *tw_s = lines + K * L;
    float *W_s = reinterpret_cast<float *>(tw_s + Z / 2);
    for (int m = tid; m < Z / 2; m += 256 * K) tw_s[m] = tw[m];
    for (int m = tid; m < Z; m += 256 * K) W_s[m] = W[m];
    twp = tw_s;
    Wp = W_s;
  }
  const int t0 = 2 * blockIdx.x, t1 = t0 + 1;
  const bool has1 = t1 < rows;
  const int s0 = (Z - n) >> 1;
  const unsigned rsh = 32 - logZ;
  const size_t row0 = static_cast<size_t>(t0) * n, row1 = row0 + n;
  const bool absorb = contrast == CTAS_CONTRAST_ABS;

  // Accumulators are loaded, updated and kept without branches (columns
  // past the row and a missing second row alias valid elements and are
  // never stored): conditional updates of register arrays this long make
  // the compiler copy the whole array per update.
  const size_t row1c = has1 ? row1 : row0;
  float acc0[J], acc1[J];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int x = min(tid + 256 * K * j, n - 1);
    acc0[j] = result[row0 + x];
    acc1[j] = result[row1c + x];
  }

  for (int p0 = 0; p0 < n_proj; p0 += K) {
    const int p = p0 + sub;
    const bool live = p < n_proj;  // the last group may be short
    const float *r0 = proj + static_cast<size_t>(p) * proj_stride + row0;
    const float *r1 = r0 + n;
    // The row pair at positions s0..s0+n-1 of the zeroed line, bit-reversed.
    for (int m = lt; m < Z; m += 256) {
      const int x = m - s0;
      float a = 0.0f, c = 0.0f;
      if (live && x >= 0 && x < n) {
        a = r0[x];
        if (absorb) a = -__logf(a);
        if (has1) {
          c = r1[x];
          if (absorb) c = -__logf(c);
        }
      }
      buf[lds_pad(__brev(static_cast<unsigned>(m)) >> rsh, psh)] = make_float2(a, c);
    }
    if (odd) odd_sums(r0, r1, live, live && has1, n, absorb, odd, lt, ostash);
    __syncthreads();
    line_filter(buf, twp, Wp, Z, logZ, lt, psh);
    if (odd) {
      const float4 t = odd_total(ostash);
      for (int x = lt; x < n; x += 256) {
        float2 &e = buf[lds_pad(__brev(static_cast<unsigned>(s0 + x)) >> rsh, psh)];
        const float2 d = odd_term(odd, odd_g, t, x);
        e.x += d.x;
        e.y += d.y;
      }
      __syncthreads();
    }

    float scale = 1.0f / static_cast<float>(Z);  // exact: Z is a power of two
    if (contrast == CTAS_CONTRAST_REF) {
      // std::partial_sum of the filtered row (lct.cpp:774-775), one lane per
      // row in sequential order, written back into the line.
      if (lt == 0 || (lt == 64 && has1)) {
        float *line = reinterpret_cast<float *>(buf) + (lt == 64 ? 1 : 0);  // one float of each pair
        float run = 0.0f;
        for (int x = 0; x < n; ++x) {
          const unsigned at = 2 * lds_pad(__brev(static_cast<unsigned>(s0 + x)) >> rsh, psh);
          const float f = line[at] * scale;
          run = (x == 0) ? f : run + f;
          line[at] = run;
        }
      }
      __syncthreads();
      scale = 1.0f;
    }

    if (filtered && live) {
      float *f0 = filtered + static_cast<size_t>(p) * proj_stride + row0;
      for (int x = lt; x < n; x += 256) {
        const float2 v = buf[lds_pad(__brev(static_cast<unsigned>(s0 + x)) >> rsh, psh)];
        f0[x] = v.x * scale;
        if (has1) f0[n + x] = v.y * scale;
      }
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
      if (p0 + q >= n_proj) break;
      const float4 k = consts.k[p0 + q];  // {cos a, P, R, -}
      const float2 *lq = lines + q * L;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int x = min(tid + 256 * K * j, n - 1);
        const int i = ts_column(x, k.x, k.y, k.z, n);
        const float2 v = lq[lds_pad(__brev(static_cast<unsigned>(s0 + i)) >> rsh, psh)];
        acc0[j] = acc0[j] + v.x * scale;
        acc1[j] = acc1[j] + v.y * scale;
      }
    }
    __syncthreads();  // the next group overwrites the lines
  }

#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int x = tid + 256 * K * j;
    if (x < n) {
      result[row0 + x] = acc0[j];
      if (has1) result[row1 + x] = acc1[j];
    }
  }
}

// exp(-+2 pi i m / 2048) from the half-turn table tw[m] (m < 1024).
template <bool INV>
__device__ __forceinline__ float2 w2048(const float2 *__restrict__ tw, int m) {
  m &= 2047;
  float2 t = tw[m & 1023];
  if (m & 1024) t = make_float2(-t.x, -t.y);
  return INV ? make_float2(t.x, -t.y) : t;
}

// 2048-point lines (Z = 2048: widths 2048, and 513-1023 padded): two lines
// of two rows each (four rows) per workgroup of 256 threads, transformed in
// registers as 8 x 16 x 16. Thread t loads x[t + 256 j] (j < 8) of both
// lines, takes 8-point DFTs and multiplies by w^(t k) (w = exp(-2 pi i /
// 2048)); after an exchange through LDS, thread (line, k, t1) takes the
// 16-point DFT over t2 of the values at t = t1 + 16 t2 and multiplies by
// w^(8 t1 m2); after another exchange, thread (line, k, m2) takes the
// 16-point DFT over t1 and holds X[k + 8 m2 + 128 m1], m1 < 16, where the
// window multiplies it. Each thread's twiddles are read once from the
// FP64-computed table and kept in registers. The inverse runs the same steps backwards with
// conjugate factors and leaves thread t with the filtered points t + 256 j
// in natural order: coalesced loads, two exchanges each way, no bit
// reversal. The filtered lines then go to LDS for the resampling.
constexpr int kA = 9, kB = 17;               // padded row lengths of the exchanges
constexpr int kLineA = 256 * kA;             // [t][k], k < 8
constexpr int kLineB = 128 * kB;             // [k * 16 + m2][t1]
constexpr size_t kTs2048Lds = sizeof(float2) * (2 * kLineA + 2 * kLineB) + sizeof(float) * 32;  // + odd sums

template <int J, bool ODD>
__global__ __launch_bounds__(256) void ts2048_kernel(const float *proj, long long proj_stride, int n_proj,
                                                     TsConsts consts, float *filtered,
                                                     float *__restrict__ result, const float *__restrict__ W,
                                                     const float2 *__restrict__ tw, int rows, int n, int contrast,
                                                     const float2 *__restrict__ odd, float odd_g) {
  constexpr int Z = 2048;
  extern __shared__ float2 lds[];
  float2 *A = lds;                 // exchange 1, then the filtered lines [line][Z]
  float2 *B = lds + 2 * kLineA;    // exchange 2
  float *ostash = reinterpret_cast<float *>(B + 2 * kLineB);  // odd widths: 16 floats per line
  const int tid = threadIdx.x;
  const int r0 = 4 * blockIdx.x;   // rows r0 .. r0 + 3
  const int s0 = (Z - n) >> 1;
  const bool absorb = contrast == CTAS_CONTRAST_ABS;
  const int ln = tid >> 7, kk = (tid >> 4) & 7, lo = tid & 15;  // line, k, t1 or m2

  // The row index of each of the four rows, clamped for loads that are
  // never stored (rows past the end).
  int rr[4];
  bool rv[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    rv[q] = r0 + q < rows;
    rr[q] = rv[q] ? r0 + q : rows - 1;
  }
  float acc[4][J];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int j = 0; j < J; ++j) acc[q][j] = result[static_cast<size_t>(rr[q]) * n + min(tid + 256 * j, n - 1)];

  // window values this thread multiplies, fixed for the whole batch
  float wv[16];
#pragma unroll
  for (int m1 = 0; m1 < 16; ++m1) wv[m1] = W[kk + 8 * lo + 128 * m1];
  // this thread's twiddles, fixed for the whole batch and read from the
  // table (each rounded once): w^(t k), k < 8, and w^(8 t1 m2), m2 < 16
  float2 tw1[8], tw2[16];
#pragma unroll
  for (int k = 0; k < 8; ++k) tw1[k] = w2048<false>(tw, tid * k);
#pragma unroll
  for (int m = 0; m < 16; ++m) tw2[m] = w2048<false>(tw, 8 * lo * m);

  // The raw samples of the next projection are loaded while the current
  // one is transformed (x = t + 256 j - s0 inside the row; 0 outside).
  float2 raw[2][8];
  auto fetch = [&](int p) {
    const float *base = proj + static_cast<size_t>(p) * proj_stride;
#pragma unroll
    for (int l = 0; l < 2; ++l) {
      const float *ra = base + static_cast<size_t>(rr[2 * l]) * n;
      const float *rb = base + static_cast<size_t>(rr[2 * l + 1]) * n;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int x = tid + 256 * j - s0;
        raw[l][j] = (x >= 0 && x < n) ? make_float2(ra[x], rb[x]) : make_float2(1.0f, 1.0f);
      }
    }
  };
  fetch(0);

  for (int p = 0; p < n_proj; ++p) {
    float2 cur[2][8];
#pragma unroll
    for (int l = 0; l < 2; ++l)
#pragma unroll
      for (int j = 0; j < 8; ++j) cur[l][j] = raw[l][j];
    if (p + 1 < n_proj) fetch(p + 1);
    if (ODD) {
      const float *base = proj + static_cast<size_t>(p) * proj_stride;
#pragma unroll
      for (int l = 0; l < 2; ++l)
        odd_sums(base + static_cast<size_t>(rr[2 * l]) * n, base + static_cast<size_t>(rr[2 * l + 1]) * n,
                 rv[2 * l], rv[2 * l + 1], n, absorb, odd, tid, ostash + 16 * l);
    }
    // pass 1: thread t, both lines, x[t + 256 j]
#pragma unroll
    for (int l = 0; l < 2; ++l) {
      float2 v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int x = tid + 256 * j - s0;
        float a = 0.0f, c = 0.0f;
        if (x >= 0 && x < n) {
          a = cur[l][j].x;
          c = cur[l][j].y;
          if (absorb) {
            a = -__logf(a);
            c = -__logf(c);
          }
        }
        v[j] = make_float2(rv[2 * l] ? a : 0.0f, rv[2 * l + 1] ? c : 0.0f);
      }
      dft8<false>(v);
#pragma unroll
      for (int k = 1; k < 8; ++k) v[k] = cmul(v[k], tw1[k]);
#pragma unroll
      for (int k = 0; k < 8; ++k) A[l * kLineA + tid * kA + k] = v[k];
    }
    __syncthreads();
    // pass 2: thread (line, k, t1), the values at t = t1 + 16 t2
    {
      float2 v[16];
#pragma unroll
      for (int t2 = 0; t2 < 16; ++t2) v[t2] = A[ln * kLineA + (lo + 16 * t2) * kA + kk];
      dft16<false>(v);
#pragma unroll
      for (int m2 = 1; m2 < 16; ++m2) v[m2] = cmul(v[m2], tw2[m2]);
#pragma unroll
      for (int m2 = 0; m2 < 16; ++m2) B[ln * kLineB + (kk * 16 + m2) * kB + lo] = v[m2];
    }
    __syncthreads();
    // pass 3: thread (line, k, m2) -> X[k + 8 m2 + 128 m1]; window; inverse
    {
      float2 v[16];
#pragma unroll
      for (int t1 = 0; t1 < 16; ++t1) v[t1] = B[ln * kLineB + (kk * 16 + lo) * kB + t1];
      dft16<false>(v);
#pragma unroll
      for (int m1 = 0; m1 < 16; ++m1) {
        v[m1].x *= wv[m1];
        v[m1].y *= wv[m1];
      }
      dft16<true>(v);
#pragma unroll
      for (int t1 = 0; t1 < 16; ++t1) B[ln * kLineB + (kk * 16 + lo) * kB + t1] = v[t1];
    }
    __syncthreads();
    // inverse pass 2: thread (line, k, t1)
    {
      float2 v[16];
#pragma unroll
      for (int m2 = 0; m2 < 16; ++m2) v[m2] = B[ln * kLineB + (kk * 16 + m2) * kB + lo];
#pragma unroll
      for (int m2 = 1; m2 < 16; ++m2) v[m2] = cmulc(v[m2], tw2[m2]);
      dft16<true>(v);
#pragma unroll
      for (int t2 = 0; t2 < 16; ++t2) A[ln * kLineA + (lo + 16 * t2) * kA + kk] = v[t2];
    }
    __syncthreads();
    // inverse pass 1: thread t -> Z x the filtered points t + 256 j
    float2 y[2][8];
#pragma unroll
    for (int l = 0; l < 2; ++l) {
#pragma unroll
      for (int k = 0; k < 8; ++k) y[l][k] = A[l * kLineA + tid * kA + k];
#pragma unroll
      for (int k = 1; k < 8; ++k) y[l][k] = cmulc(y[l][k], tw1[k]);
      dft8<true>(y[l]);
    }
    if (ODD) {
#pragma unroll
      for (int l = 0; l < 2; ++l) {
        const float4 t = odd_total(ostash + 16 * l);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int x = tid + 256 * j - s0;
          if (x >= 0 && x < n) {
            const float2 d = odd_term(odd, odd_g, t, x);
            y[l][j].x += d.x;
            y[l][j].y += d.y;
          }
        }
      }
    }
    __syncthreads();  // exchange 1 is read; the lines go in its place
    constexpr float scale0 = 1.0f / static_cast<float>(Z);
#pragma unroll
    for (int l = 0; l < 2; ++l)
#pragma unroll
      for (int j = 0; j < 8; ++j) A[l * Z + tid + 256 * j] = y[l][j];
    __syncthreads();
    float scale = scale0;
    if (contrast == CTAS_CONTRAST_REF) {
      // std::partial_sum of each filtered row (lct.cpp:774-775), one lane
      // per row in sequential order, written back into its line.
      if ((tid & 63) == 0) {
        const int q = tid >> 6;  // row of the workgroup
        float *line = reinterpret_cast<float *>(A + (q >> 1) * Z) + (q & 1);
        float run = 0.0f;
        for (int x = 0; x < n; ++x) {
          const float f = line[2 * (s0 + x)] * scale0;
          run = (x == 0) ? f : run + f;
          line[2 * (s0 + x)] = run;
        }
      }
      __syncthreads();
      scale = 1.0f;
    }
    if (filtered) {
      float *fb = filtered + static_cast<size_t>(p) * proj_stride;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (!rv[q]) continue;
        const float *line = reinterpret_cast<const float *>(A + (q >> 1) * Z) + (q & 1);
        for (int x = tid; x < n; x += 256) fb[static_cast<size_t>(r0 + q) * n + x] = line[2 * (s0 + x)] * scale;
      }
    }
    const float4 kc = consts.k[p];  // {cos a, P, R, -}
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int x = min(tid + 256 * j, n - 1);
      const int i = ts_column(x, kc.x, kc.y, kc.z, n);
      const float2 a = A[s0 + i], b = A[Z + s0 + i];
      acc[0][j] = acc[0][j] + a.x * scale;
      acc[1][j] = acc[1][j] + a.y * scale;
      acc[2][j] = acc[2][j] + b.x * scale;
      acc[3][j] = acc[3][j] + b.y * scale;
    }
    __syncthreads();  // the next projection overwrites exchange 1 and the lines
  }

#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (!rv[q]) continue;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int x = tid + 256 * j;
      if (x < n) result[static_cast<size_t>(r0 + q) * n + x] = acc[q][j];
    }
  }
}

// exp(-+2 pi i m / 4096) from the half-turn table tw[m] (m < 2048).
template <bool INV>
__device__ __forceinline__ float2 w4096(const float2 *__restrict__ tw, int m) {
  m &= 4095;
  float2 t = tw[m & 2047];
  if (m & 2048) t = make_float2(-t.x, -t.y);
  return INV ? make_float2(t.x, -t.y) : t;
}

// 4096-point lines (Z = 4096: width 4096, and 1025-2047 padded): one line
// (two rows) per workgroup of 256 threads, transformed in registers as
// 16 x 16 x 16, the steps of the sinogram filter's radix-16 kernel: thread t
// holds x[t + 256 j], takes a 16-point DFT and multiplies by w^(t k); after
// an exchange thread (t1 = t % 16, k1 = t / 16) takes the DFT over t2 of the
// values at t1 + 16 t2 and multiplies by w^(16 t1 k2); after another, thread
// (k1, k2) holds X[k1 + 16 k2 + 256 k3], k3 < 16, for the window; the inverse
// runs back with conjugate factors to the filtered x[t + 256 j]. Twiddles
// (from the FP64 table) and window values stay in registers for the batch.
constexpr int kP16 = 17;  // padded row length of the exchanges
constexpr size_t kTs4096Lds = sizeof(float2) * 256 * kP16 + sizeof(float) * 16;  // + odd sums

template <int J, bool ODD>
__global__ __launch_bounds__(256) void ts4096_kernel(const float *proj, long long proj_stride, int n_proj,
                                                     TsConsts consts, float *filtered,
                                                     float *__restrict__ result, const float *__restrict__ W,
                                                     const float2 *__restrict__ tw, int rows, int n, int contrast,
                                                     const float2 *__restrict__ odd, float odd_g) {
  constexpr int Z = 4096;
  extern __shared__ float2 xch[];
  float *ostash = reinterpret_cast<float *>(xch + 256 * kP16);
  const int tid = threadIdx.x, lo4 = tid & 15, hi4 = tid >> 4;
  const int t0 = 2 * blockIdx.x;
  const bool has1 = t0 + 1 < rows;
  const int s0 = (Z - n) >> 1;
  const bool absorb = contrast == CTAS_CONTRAST_ABS;
  const size_t row0 = static_cast<size_t>(t0) * n, row1c = has1 ? row0 + n : row0;

  float acc0[J], acc1[J];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int x = min(tid + 256 * j, n - 1);
    acc0[j] = result[row0 + x];
    acc1[j] = result[row1c + x];
  }
  float2 tw1[16], tw2[16];
  float wv[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    tw1[k] = w4096<false>(tw, tid * k);
    tw2[k] = w4096<false>(tw, 16 * lo4 * k);
    wv[k] = W[hi4 + 16 * lo4 + 256 * k];
  }

  // The next projection's rows load during the current transform, except
  // at 16 columns per thread and row, where the registers would not keep
  // two waves per SIMD.
  constexpr bool kPrefetch = J <= 8;
  float2 raw[16];
  auto fetch = [&](int p) {
    const float *r0 = proj + static_cast<size_t>(p) * proj_stride + row0;
    const float *r1 = proj + static_cast<size_t>(p) * proj_stride + row1c;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int x = tid + 256 * j - s0;
      raw[j] = (x >= 0 && x < n) ? make_float2(r0[x], r1[x]) : make_float2(1.0f, 1.0f);
    }
  };
  if (kPrefetch) fetch(0);

  for (int p = 0; p < n_proj; ++p) {
    if (!kPrefetch) fetch(p);
    if (ODD)
      odd_sums(proj + static_cast<size_t>(p) * proj_stride + row0, proj + static_cast<size_t>(p) * proj_stride + row1c,
               true, has1, n, absorb, odd, tid, ostash);
    float2 v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int x = tid + 256 * j - s0;
      float a = 0.0f, c = 0.0f;
      if (x >= 0 && x < n) {
        a = raw[j].x;
        c = raw[j].y;
        if (absorb) {
          a = -__logf(a);
          c = -__logf(c);
        }
      }
      v[j] = make_float2(a, has1 ? c : 0.0f);
    }
    if (kPrefetch && p + 1 < n_proj) fetch(p + 1);

    dft16<false>(v);
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmul(v[k], tw1[k]);
#pragma unroll
    for (int k = 0; k < 16; ++k) xch[tid * kP16 + k] = v[k];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = xch[(lo4 + 16 * j) * kP16 + hi4];
    __syncthreads();
    dft16<false>(v);
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmul(v[k], tw2[k]);
#pragma unroll
    for (int k = 0; k < 16; ++k) xch[(hi4 * 16 + k) * kP16 + lo4] = v[k];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = xch[tid * kP16 + j];
    dft16<false>(v);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      v[k].x *= wv[k];
      v[k].y *= wv[k];
    }
    dft16<true>(v);
#pragma unroll
    for (int j = 0; j < 16; ++j) xch[tid * kP16 + j] = v[j];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = xch[(hi4 * 16 + k) * kP16 + lo4];
    __syncthreads();
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmulc(v[k], tw2[k]);
    dft16<true>(v);
#pragma unroll
    for (int j = 0; j < 16; ++j) xch[(lo4 + 16 * j) * kP16 + hi4] = v[j];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = xch[tid * kP16 + k];
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmulc(v[k], tw1[k]);
    dft16<true>(v);
    if (ODD) {
      const float4 t = odd_total(ostash);
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int x = tid + 256 * j - s0;
        if (x >= 0 && x < n) {
          const float2 d = odd_term(odd, odd_g, t, x);
          v[j].x += d.x;
          v[j].y += d.y;
        }
      }
    }
    __syncthreads();  // the exchange is read; the line goes in its place
#pragma unroll
    for (int j = 0; j < 16; ++j) xch[tid + 256 * j] = v[j];
    __syncthreads();

    float scale = 1.0f / static_cast<float>(Z);  // exact: Z is a power of two
    if (contrast == CTAS_CONTRAST_REF) {
      // std::partial_sum of each filtered row (lct.cpp:774-775), one lane per
      // row in sequential order, written back into the line.
      if (tid == 0 || (tid == 64 && has1)) {
        float *line = reinterpret_cast<float *>(xch) + (tid == 64 ? 1 : 0);
        float run = 0.0f;
        for (int x = 0; x < n; ++x) {
          const float f = line[2 * (s0 + x)] * scale;
          run = (x == 0) ? f : run + f;
          line[2 * (s0 + x)] = run;
        }
      }
      __syncthreads();
      scale = 1.0f;
    }
    if (filtered) {
      float *f0 = filtered + static_cast<size_t>(p) * proj_stride + row0;
      for (int x = tid; x < n; x += 256) {
        const float2 q = xch[s0 + x];
        f0[x] = q.x * scale;
        if (has1) f0[n + x] = q.y * scale;
      }
    }
    const float4 kc = consts.k[p];  // {cos a, P, R, -}
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int x = min(tid + 256 * j, n - 1);
      const float2 q = xch[s0 + ts_column(x, kc.x, kc.y, kc.z, n)];
      acc0[j] = acc0[j] + q.x * scale;
      acc1[j] = acc1[j] + q.y * scale;
    }
    __syncthreads();  // the next projection overwrites the exchange
  }

#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int x = tid + 256 * j;
    if (x < n) {
      result[row0 + x] = acc0[j];
      if (has1) result[row0 + n + x] = acc1[j];
    }
  }
}

// 8192-point lines of widths 2049-4095 (padded): one line (two rows) per
// workgroup, transformed in registers as 32 x 16 x 16 with the steps of the
// sinogram filter's 8192-point kernel (thread t holds x[t + 256 j], j < 32;
// two items per thread in each radix-16 pass; one exchange buffer of
// 512 x 17 entries, which then holds the filtered line).
constexpr int kTsPA8 = 33, kTsPB8 = 17;
constexpr size_t kTs8192Lds = sizeof(float2) * 512 * kTsPB8 + sizeof(float) * 16;  // + odd sums

template <int J, bool ODD>
__global__ __launch_bounds__(256) void ts8192_kernel(const float *proj, long long proj_stride, int n_proj,
                                                     TsConsts consts, float *filtered,
                                                     float *__restrict__ result, const float *__restrict__ W,
                                                     const float2 *__restrict__ tw, int rows, int n, int contrast,
                                                     const float2 *__restrict__ odd, float odd_g) {
  constexpr int Z = 8192;
  extern __shared__ float2 xb[];
  float *ostash = reinterpret_cast<float *>(xb + 512 * kTsPB8);
  const int tid = threadIdx.x;
  const int t0 = 2 * blockIdx.x;
  const bool has1 = t0 + 1 < rows;
  const int s0 = (Z - n) >> 1;
  const bool absorb = contrast == CTAS_CONTRAST_ABS;
  const size_t row0 = static_cast<size_t>(t0) * n, row1c = has1 ? row0 + n : row0;

  float acc0[J], acc1[J];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int x = min(tid + 256 * j, n - 1);
    acc0[j] = result[row0 + x];
    acc1[j] = result[row1c + x];
  }

  for (int p = 0; p < n_proj; ++p) {
    const float *r0 = proj + static_cast<size_t>(p) * proj_stride + row0;
    const float *r1 = proj + static_cast<size_t>(p) * proj_stride + row1c;
    if (ODD) odd_sums(r0, r1, true, has1, n, absorb, odd, tid, ostash);
    float2 v[32];
#pragma unroll
    for (int j = 0; j < 32; ++j) {
      const int x = tid + 256 * j - s0;
      float a = 0.0f, c = 0.0f;
      if (x >= 0 && x < n) {
        a = r0[x];
        c = r1[x];
        if (absorb) {
          a = -__logf(a);
          c = -__logf(c);
        }
      }
      v[j] = make_float2(a, has1 ? c : 0.0f);
    }
    dft32<false>(v);
#pragma unroll
    for (int k = 1; k < 32; ++k) v[k] = cmul(v[k], w8192(tw, tid * k));
#pragma unroll
    for (int k = 0; k < 32; ++k) xb[tid * kTsPA8 + k] = v[k];
    __syncthreads();
    float2 u[2][16];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int i = tid + 256 * s, k = i >> 4, q = i & 15;  // q = t1
#pragma unroll
      for (int t2 = 0; t2 < 16; ++t2) u[s][t2] = xb[(q + 16 * t2) * kTsPA8 + k];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int i = tid + 256 * s, k = i >> 4, q = i & 15;
      dft16<false>(u[s]);
#pragma unroll
      for (int m2 = 1; m2 < 16; ++m2) u[s][m2] = cmul(u[s][m2], w8192(tw, 32 * q * m2));
#pragma unroll
      for (int m2 = 0; m2 < 16; ++m2) xb[(k * 16 + m2) * kTsPB8 + q] = u[s][m2];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int i = tid + 256 * s, k = i >> 4, q = i & 15;  // q = m2
#pragma unroll
      for (int t1 = 0; t1 < 16; ++t1) u[s][t1] = xb[(k * 16 + q) * kTsPB8 + t1];
      dft16<false>(u[s]);
#pragma unroll
      for (int m1 = 0; m1 < 16; ++m1) {
        const float wk = W[k + 32 * q + 512 * m1];
        u[s][m1].x *= wk;
        u[s][m1].y *= wk;
      }
      dft16<true>(u[s]);
#pragma unroll
      for (int t1 = 0; t1 < 16; ++t1) xb[(k * 16 + q) * kTsPB8 + t1] = u[s][t1];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int i = tid + 256 * s, k = i >> 4, q = i & 15;  // q = t1
#pragma unroll
      for (int m2 = 0; m2 < 16; ++m2) u[s][m2] = xb[(k * 16 + m2) * kTsPB8 + q];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int i = tid + 256 * s, k = i >> 4, q = i & 15;
#pragma unroll
      for (int m2 = 1; m2 < 16; ++m2) u[s][m2] = cmulc(u[s][m2], w8192(tw, 32 * q * m2));
      dft16<true>(u[s]);
#pragma unroll
      for (int t2 = 0; t2 < 16; ++t2) xb[(q + 16 * t2) * kTsPA8 + k] = u[s][t2];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 32; ++k) v[k] = xb[tid * kTsPA8 + k];
#pragma unroll
    for (int k = 1; k < 32; ++k) v[k] = cmulc(v[k], w8192(tw, tid * k));
    dft32<true>(v);
    if (ODD) {
      const float4 t = odd_total(ostash);
#pragma unroll
      for (int j = 0; j < 32; ++j) {
        const int x = tid + 256 * j - s0;
        if (x >= 0 && x < n) {
          const float2 d = odd_term(odd, odd_g, t, x);
          v[j].x += d.x;
          v[j].y += d.y;
        }
      }
    }
    __syncthreads();  // the exchange is read; the line goes in its place
#pragma unroll
    for (int j = 0; j < 32; ++j) xb[tid + 256 * j] = v[j];
    __syncthreads();

    float scale = 1.0f / static_cast<float>(Z);  // exact: Z is a power of two
    if (contrast == CTAS_CONTRAST_REF) {
      // std::partial_sum of each filtered row (lct.cpp:774-775), one lane per
      // row in sequential order, written back into the line.
      if (tid == 0 || (tid == 64 && has1)) {
        float *line = reinterpret_cast<float *>(xb) + (tid == 64 ? 1 : 0);
        float run = 0.0f;
        for (int x = 0; x < n; ++x) {
          const float f = line[2 * (s0 + x)] * scale;
          run = (x == 0) ? f : run + f;
          line[2 * (s0 + x)] = run;
        }
      }
      __syncthreads();
      scale = 1.0f;
    }
    if (filtered) {
      float *f0 = filtered + static_cast<size_t>(p) * proj_stride + row0;
      for (int x = tid; x < n; x += 256) {
        const float2 q = xb[s0 + x];
        f0[x] = q.x * scale;
        if (has1) f0[n + x] = q.y * scale;
      }
    }
    const float4 kc = consts.k[p];  // {cos a, P, R, -}
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int x = min(tid + 256 * j, n - 1);
      const float2 q = xb[s0 + ts_column(x, kc.x, kc.y, kc.z, n)];
      acc0[j] = acc0[j] + q.x * scale;
      acc1[j] = acc1[j] + q.y * scale;
    }
    __syncthreads();  // the next projection overwrites the buffer
  }

#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int x = tid + 256 * j;
    if (x < n) {
      result[row0 + x] = acc0[j];
      if (has1) result[row0 + n + x] = acc1[j];
    }
  }
}

}  // namespace

hipError_t launch_ts(const float *proj, long long proj_stride, int n_proj, const TsConsts &consts,
                     float *filtered, float *result, const float *W, const float2 *tw, int rows, int n, int Z,
                     int logZ, int contrast, int cus, int lds_max, const float2 *odd, float odd_g,
                     hipStream_t s) {
  if (rows <= 0 || n_proj <= 0) return hipSuccess;
  if (n_proj > kTsBatch || n > kTsMaxWidth || Z > kTsMaxWidth) return hipErrorInvalidValue;
  if (Z == 2048 && kTs2048Lds <= static_cast<size_t>(lds_max)) {
    const dim3 grid((rows + 3) / 4), block(256);
    const int J = (n + 255) / 256;  // n <= 2048: J <= 8
    if (J <= 4)
      hipLaunchKernelGGL((odd ? ts2048_kernel<4, true> : ts2048_kernel<4, false>), grid, block, kTs2048Lds, s, proj, proj_stride, n_proj, consts, filtered,
                         result, W, tw, rows, n, contrast, odd, odd_g);
    else
      hipLaunchKernelGGL((odd ? ts2048_kernel<8, true> : ts2048_kernel<8, false>), grid, block, kTs2048Lds, s, proj, proj_stride, n_proj, consts, filtered,
                         result, W, tw, rows, n, contrast, odd, odd_g);
    return hipGetLastError();
  }
  if (Z == 8192 && n <= 4096 && kTs8192Lds <= static_cast<size_t>(lds_max)) {
    const dim3 grid((rows + 1) / 2), block(256);
    hipLaunchKernelGGL((odd ? ts8192_kernel<16, true> : ts8192_kernel<16, false>), grid, block, kTs8192Lds, s, proj, proj_stride, n_proj, consts, filtered,
                       result, W, tw, rows, n, contrast, odd, odd_g);
    return hipGetLastError();
  }
  if (Z == 4096 && kTs4096Lds <= static_cast<size_t>(lds_max)) {
    const dim3 grid((rows + 1) / 2), block(256);
    const int J = (n + 255) / 256;  // n <= 4096: J <= 16
    if (J <= 8)
      hipLaunchKernelGGL((odd ? ts4096_kernel<8, true> : ts4096_kernel<8, false>), grid, block, kTs4096Lds, s, proj, proj_stride, n_proj, consts, filtered,
                         result, W, tw, rows, n, contrast, odd, odd_g);
    else
      hipLaunchKernelGGL((odd ? ts4096_kernel<16, true> : ts4096_kernel<16, false>), grid, block, kTs4096Lds, s, proj, proj_stride, n_proj, consts, filtered,
                         result, W, tw, rows, n, contrast, odd, odd_g);
    return hipGetLastError();
  }
  // Projections in flight per workgroup (K): up to 4, at most the batch, as
  // long as every workgroup of the launch stays resident (160 KiB of LDS per
  // CU; a second round of workgroups would wait for the first to finish the
  // whole batch) within the workgroup limit of lds_max bytes. The twiddle
  // and window tables go to LDS too (TAB) where that keeps one line per
  // workgroup resident.
  const int psh = line_psh(Z);
  const size_t line_bytes = sizeof(float2) * padded_len(Z, psh);
  const size_t tab_bytes = (sizeof(float2) / 2 + sizeof(float)) * static_cast<size_t>(Z);
  const size_t stash_bytes = sizeof(float) * 16;  // odd-width wave sums per line
  const int pairs = (rows + 1) / 2;
  const size_t per_cu = (pairs + cus - 1) / cus;
  auto fits = [&](size_t b) { return b <= static_cast<size_t>(lds_max) && per_cu * b <= 160 * 1024; };
  const bool tab = fits(line_bytes + stash_bytes + tab_bytes);
  const size_t extra = tab ? tab_bytes : 0;
  int K = 1;
  while (K < 4 && K < n_proj && 2 * K * Z <= 8192 && fits(2 * K * (line_bytes + stash_bytes) + extra)) K *= 2;
  const size_t lds = (line_bytes + stash_bytes) * K + extra;
  if (lds > static_cast<size_t>(lds_max)) return hipErrorInvalidValue;
  const dim3 grid(pairs), block(256 * K);
  const int J = (n + 256 * K - 1) / (256 * K);
#define CTAS_TS_CASE(JJ, KK)                                                                                  \
  do {                                                                                                        \
    if (tab)                                                                                                  \
      hipLaunchKernelGGL((ts_kernel<JJ, KK, true>), grid, block, lds, s, proj, proj_stride, n_proj, consts,   \
                         filtered, result, W, tw, rows, n, Z, logZ, contrast, odd, odd_g);                    \
    else                                                                                                      \
      hipLaunchKernelGGL((ts_kernel<JJ, KK, false>), grid, block, lds, s, proj, proj_stride, n_proj, consts,  \
                         filtered, result, W, tw, rows, n, Z, logZ, contrast, odd, odd_g);                    \
  } while (0)
  // K = 4 only for Z <= 2048 (so J <= 2), K = 2 for Z <= 4096 (J <= 8).
  if (K == 4) {
    if (J <= 1) CTAS_TS_CASE(1, 4);
    else CTAS_TS_CASE(2, 4);
  } else if (K == 2) {
    if (J <= 1) CTAS_TS_CASE(1, 2);
    else if (J <= 2) CTAS_TS_CASE(2, 2);
    else if (J <= 4) CTAS_TS_CASE(4, 2);
    else CTAS_TS_CASE(8, 2);
  } else {
    if (J <= 1) CTAS_TS_CASE(1, 1);
    else if (J <= 2) CTAS_TS_CASE(2, 1);
    else if (J <= 4) CTAS_TS_CASE(4, 1);
    else if (J <= 8) CTAS_TS_CASE(8, 1);
    else if (J <= 16) CTAS_TS_CASE(16, 1);
    else CTAS_TS_CASE(32, 1);
  }
#undef CTAS_TS_CASE
  return hipGetLastError();
}

namespace {

// ts_add's per-projection constants (lct.cpp:762-765), with the reference's
// types: sin and cos of the float angle in float, P = (plane - center) *
// sin(a) in float, R = pixels * 0.5 + center in double, narrowed.
float4 ts_consts(int n, float center, float angle, int plane) {
  const float cur_sin = std::sin(angle), cur_cos = std::cos(angle);
  const float plane_cos = (static_cast<float>(plane) - center) * cur_sin;
  const float rcenter = static_cast<float>(n * 0.5 + static_cast<double>(center));
  return make_float4(cur_cos, plane_cos, rcenter, 0.0f);
}

}  // namespace
}  // namespace ctas

using namespace ctas;

struct ctas_ts {
  int device = 0;
  int rows = 0, n = 0, Z = 0, logZ = 0, contrast = 0;
  int cus = 1, lds_max = 65536;  // compute units, LDS bytes per workgroup
  float *d_W = nullptr;
  float2 *d_tw = nullptr;
  float2 *d_odd = nullptr;  // odd widths: basis of the middle-bin correction
  float odd_g = 0.0f;
};

extern "C" {

int ctas_filter_fill(int filter, float param, int n, float *W) {
  if (!W) return fail(CTAS_E_ARG, "null argument");
  std::vector<float> v;
  if (int rc = filter_fill(filter, param, n, v)) return rc;
  std::copy(v.begin(), v.end(), W);
  return CTAS_OK;
}

int ctas_ts_window(int filter, float param, int n, int *Z, float *W) {
  if (!Z) return fail(CTAS_E_ARG, "null argument");
  std::vector<float> v;
  int z = 0;
  if (int rc = ts_window(filter, param, n, z, v)) return rc;
  *Z = z;
  if (W) std::copy(v.begin(), v.end(), W);
  return CTAS_OK;
}

int ctas_ts_create(ctas_ts **ph, int rows, int pixels, int contrast, int filter, float filter_param) {
  if (!ph) return fail(CTAS_E_ARG, "null handle pointer");
  *ph = nullptr;
  if (rows < 1 || pixels < 1) return fail(CTAS_E_ARG, "tomosynthesis: empty projection");
  if (contrast < CTAS_CONTRAST_ABS || contrast > CTAS_CONTRAST_FLT)
    return fail(CTAS_E_ARG, "unknown contrast code " + std::to_string(contrast));
  if (pixels > kTsMaxWidth)
    return fail(CTAS_E_ARG, "tomosynthesis: projection width " + std::to_string(pixels) + " above " +
                                std::to_string(kTsMaxWidth));
  std::unique_ptr<ctas_ts> h(new ctas_ts);
  h->rows = rows;
  h->n = pixels;
  h->contrast = contrast;
  std::vector<float> W;
  if (int rc = ts_window(filter, filter_param, pixels, h->Z, W)) return rc;
  while ((1 << h->logZ) < h->Z) ++h->logZ;
  if (h->Z > kTsMaxWidth)
    return fail(CTAS_E_ARG, "tomosynthesis: width " + std::to_string(pixels) + " needs a line of " +
                                std::to_string(h->Z) + " points, more than the 64 KiB of LDS of one workgroup holds");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(CTAS_E_DEVICE, "no HIP device is available");
  CTAS_HIP_TRY(hipGetDevice(&h->device));
  CTAS_HIP_TRY(hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, h->device));
  CTAS_HIP_TRY(hipDeviceGetAttribute(&h->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
  h->cus = std::max(h->cus, 1);
  if (sizeof(float2) * static_cast<size_t>(padded_len(h->Z, line_psh(h->Z))) > static_cast<size_t>(h->lds_max))
    return fail(CTAS_E_ARG, "tomosynthesis: width " + std::to_string(pixels) + " needs a line of " +
                                std::to_string(h->Z) + " points, more than one workgroup's LDS holds");
  std::vector<float2> tw(std::max(h->Z / 2, 1));
  for (int m = 0; m < h->Z / 2; ++m) {
    const double a = -2.0 * M_PI * m / h->Z;
    tw[m] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
  }
  std::vector<float2> uv;
  const bool odd = ts_odd_basis(filter, filter_param, pixels, h->Z, uv, h->odd_g);
  if (hipMalloc(&h->d_W, sizeof(float) * W.size()) != hipSuccess ||
      hipMalloc(&h->d_tw, sizeof(float2) * tw.size()) != hipSuccess ||
      (odd && hipMalloc(&h->d_odd, sizeof(float2) * uv.size()) != hipSuccess)) {
    (void)hipFree(h->d_W);
    (void)hipFree(h->d_tw);
    return fail(CTAS_E_NOMEM, "device allocation failed");
  }
  CTAS_HIP_TRY(hipMemcpy(h->d_W, W.data(), sizeof(float) * W.size(), hipMemcpyHostToDevice));
  CTAS_HIP_TRY(hipMemcpy(h->d_tw, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice));
  if (odd) CTAS_HIP_TRY(hipMemcpy(h->d_odd, uv.data(), sizeof(float2) * uv.size(), hipMemcpyHostToDevice));
  *ph = h.release();
  return CTAS_OK;
}

int ctas_ts_destroy(ctas_ts *h) {
  if (!h) return CTAS_OK;
  (void)hipFree(h->d_W);
  (void)hipFree(h->d_tw);
  (void)hipFree(h->d_odd);
  delete h;
  return CTAS_OK;
}

int ctas_ts_info(const ctas_ts *h, int *rows, int *pixels, int *zidth) {
  if (!h) return fail(CTAS_E_ARG, "null handle");
  if (rows) *rows = h->rows;
  if (pixels) *pixels = h->n;
  if (zidth) *zidth = h->Z;
  return CTAS_OK;
}

int ctas_ts_add_dev(ctas_ts *h, const float *d_proj, int n_proj, const float *angles, float center, int plane,
                    float *d_filtered, float *d_result, void *stream) {
  if (!h || !d_proj || !angles || !d_result || n_proj < 0) return fail(CTAS_E_ARG, "bad arguments");
  const long long stride = static_cast<long long>(h->rows) * h->n;
  hipStream_t s = static_cast<hipStream_t>(stream);
  for (int p0 = 0; p0 < n_proj; p0 += kTsBatch) {
    const int np = std::min(kTsBatch, n_proj - p0);
    TsConsts k{};
    for (int i = 0; i < np; ++i) k.k[i] = ts_consts(h->n, center, angles[p0 + i], plane);
    const hipError_t e = launch_ts(d_proj + p0 * stride, stride, np, k, d_filtered ? d_filtered + p0 * stride : nullptr,
                                   d_result, h->d_W, h->d_tw, h->rows, h->n, h->Z, h->logZ, h->contrast, h->cus,
                                   h->lds_max, h->d_odd, h->odd_g, s);
    if (e != hipSuccess) return fail(CTAS_E_DEVICE, std::string("tomosynthesis: ") + hipGetErrorString(e));
  }
  return CTAS_OK;
}

int ctas_ts_accumulate(ctas_ts *h, const float *projections, int n_proj, const float *angles, float center,
                       int plane, float *result) {
  if (!h || !projections || !angles || !result || n_proj < 0) return fail(CTAS_E_ARG, "bad arguments");
  const size_t elems = static_cast<size_t>(h->rows) * h->n;
  // Projections go up in chunks of kTsBatch through two staging buffers, so
  // the upload of one chunk overlaps the kernel on the previous one.
  const int chunk = std::min(std::max(n_proj, 1), kTsBatch);
  float *d = nullptr;
  if (hipMalloc(&d, sizeof(float) * elems * (1 + 2 * static_cast<size_t>(chunk))) != hipSuccess)
    return fail(CTAS_E_NOMEM, "device allocation failed");
  float *d_res = d, *d_stage[2] = {d + elems, d + elems * (1 + chunk)};
  hipStream_t s = nullptr;
  hipEvent_t done[2] = {nullptr, nullptr};
  int rc = CTAS_OK;
  auto dev = [&](hipError_t e) {
    if (e != hipSuccess && rc == CTAS_OK) rc = fail(CTAS_E_DEVICE, std::string("tomosynthesis: ") + hipGetErrorString(e));
    return rc == CTAS_OK;
  };
  if (dev(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) && dev(hipEventCreate(&done[0])) &&
      dev(hipEventCreate(&done[1])) &&
      dev(hipMemcpyAsync(d_res, result, sizeof(float) * elems, hipMemcpyHostToDevice, s))) {
    for (int p0 = 0, b = 0; p0 < n_proj && rc == CTAS_OK; p0 += chunk, b ^= 1) {
      const int np = std::min(chunk, n_proj - p0);
      // the kernel that last read this staging buffer has finished
      if (!dev(hipEventSynchronize(done[b]))) break;
      if (!dev(hipMemcpyAsync(d_stage[b], projections + p0 * elems, sizeof(float) * elems * np,
                              hipMemcpyHostToDevice, s)))
        break;
      if (int r = ctas_ts_add_dev(h, d_stage[b], np, angles + p0, center, plane, nullptr, d_res, s)) {
        rc = r;
        break;
      }
      dev(hipEventRecord(done[b], s));
    }
    if (rc == CTAS_OK && dev(hipMemcpyAsync(result, d_res, sizeof(float) * elems, hipMemcpyDeviceToHost, s)))
      dev(hipStreamSynchronize(s));
  }
  if (s) (void)hipStreamSynchronize(s);
  for (hipEvent_t e : done)
    if (e) (void)hipEventDestroy(e);
  if (s) (void)hipStreamDestroy(s);
  (void)hipFree(d);
  return rc;
}

int ctas_ts_add(float *projection, int rows, int pixels, float *result, int result_rows, int result_pixels,
                int filter, float filter_param, float center, int contrast, float angle, int plane) {
  if (!projection || !result) return fail(CTAS_E_ARG, "null argument");
  if (result_pixels != pixels || result_rows != rows)
    return fail(CTAS_E_ARG, "The size of the result array does not match the requested reconstruction.");
  // The last geometry's plan and staging buffers are kept (until the
  // process ends) for the next call: the tool calls this once per
  // projection of a series. Calls are serialised.
  struct Cached {
    std::tuple<int, int, int, int, int, float> key{-1, 0, 0, 0, 0, 0.0f};
    ctas_ts *plan = nullptr;
    float *d_io = nullptr;
  };
  static Cached c;
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(CTAS_E_DEVICE, "no HIP device is available");
  const auto key = std::make_tuple(dev, rows, pixels, contrast, filter, filter_param);
  if (!c.plan || c.key != key) {
    ctas_ts_destroy(c.plan);
    c.plan = nullptr;
    if (c.d_io) (void)hipFree(c.d_io);
    c.d_io = nullptr;
    if (int rc = ctas_ts_create(&c.plan, rows, pixels, contrast, filter, filter_param)) return rc;
    if (hipMalloc(&c.d_io, 2 * sizeof(float) * rows * static_cast<size_t>(pixels)) != hipSuccess)
      return fail(CTAS_E_NOMEM, "device allocation failed");
    c.key = key;
  }
  const size_t elems = static_cast<size_t>(rows) * pixels;
  float *d_proj = c.d_io, *d_res = c.d_io + elems;
  CTAS_HIP_TRY(hipMemcpy(d_proj, projection, sizeof(float) * elems, hipMemcpyHostToDevice));
  CTAS_HIP_TRY(hipMemcpy(d_res, result, sizeof(float) * elems, hipMemcpyHostToDevice));
  // The reference filters the projection in place (the row view of lct.cpp:768).
  if (int rc = ctas_ts_add_dev(c.plan, d_proj, 1, &angle, center, plane, d_proj, d_res, nullptr)) return rc;
  CTAS_HIP_TRY(hipDeviceSynchronize());
  CTAS_HIP_TRY(hipMemcpy(projection, d_proj, sizeof(float) * elems, hipMemcpyDeviceToHost));
  CTAS_HIP_TRY(hipMemcpy(result, d_res, sizeof(float) * elems, hipMemcpyDeviceToHost));
  return CTAS_OK;
}

}  // extern "C"
