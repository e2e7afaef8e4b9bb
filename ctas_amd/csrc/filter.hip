// Sinogram preparation on the device: the deAbs clamp rule and -log, the
// zero-padded FFT convolution of each detector row with the filter window,
// and the REF prefix sum (CTrec::prepare_sino, lct.cpp:569-597).
//
// Layout: one workgroup filters two rows of one sinogram as the real and
// imaginary parts of one complex line of zidth points held in LDS. The window
// is real and even, so the two filtered rows separate exactly. The line is
// loaded in bit-reversed order, transformed by a decimation-in-time FFT whose
// radix-2 stages are fused in pairs (four points per thread per pass),
// multiplied by the window in natural order, and transformed back by a
// decimation-in-frequency FFT with conjugate twiddles, which leaves the
// result in bit-reversed order again; the centre w samples are read out of
// their bit-reversed positions. HBM traffic per row: w*4 bytes in, w*s out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>

#include "ctas_internal.hpp"
#include "line_fft.hpp"

namespace ctas {
namespace {

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float v) {
  uint32_t u = __float_as_uint(v);
  if ((u & 0x7f800000u) == 0x7f800000u) return static_cast<uint16_t>((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

template <int ST>
__device__ __forceinline__ void store_val(void *base, size_t i, float v) {
  if constexpr (ST == ST_F32) {
    static_cast<float *>(base)[i] = v;
  } else if constexpr (ST == ST_F16) {
    static_cast<_Float16 *>(base)[i] = static_cast<_Float16>(v);
  } else {
    static_cast<uint16_t *>(base)[i] = f32_to_bf16_rne(v);
  }
}

// Loads and stores through global (address space 1) pointers: the item
// table holds generic pointers, from which the compiler emits FLAT
// accesses, which also count against the LDS wait counter.
__device__ __forceinline__ float gload(const float *p, long long i) {
  return ((const __attribute__((address_space(1))) float *)(p))[i];
}
template <int ST>
__device__ __forceinline__ void gstore(void *base, size_t i, float v) {
  if constexpr (ST == ST_F32) {
    ((__attribute__((address_space(1))) float *)(base))[i] = v;
  } else if constexpr (ST == ST_F16) {
    ((__attribute__((address_space(1))) _Float16 *)(base))[i] = static_cast<_Float16>(v);
  } else {
    ((__attribute__((address_space(1))) uint16_t *)(base))[i] = f32_to_bf16_rne(v);
  }
}

// The item that owns row pair b. Batches of equal items (every slice of a
// handle) are found by one division and checked with two loads; uneven ones
// fall back to a binary search (a chain of dependent loads, each a scalar
// memory round trip at the start of every workgroup).
__device__ __forceinline__ int find_item_by_pair(const Item *items, int n, int b) {
  if (n > 1) {
    const int per = items[1].pair_base;
    if (per > 0) {
      const int g = min(b / per, n - 1);
      if (items[g].pair_base <= b && (g == n - 1 || items[g + 1].pair_base > b)) return g;
    }
  }
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].pair_base <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// deAbs (physical.world.cpp:88-100): when a sinogram's minimum is not
// positive, every sample at or below max/1e6 is raised to that value; then
// -log. Absorption data is positive as a rule, so the filter runs first
// without the clamp and notes the sinograms that hold a sample <= 0 (the
// minimum's test, sample by sample: every sample is loaded by some row
// pair, and NaN passes both tests alike); only for those the (min, max)
// partials are formed and their row pairs filtered again with the clamp.
// Two ints follow each item's kMinMaxBlocks partials: [0] the item holds a
// sample <= 0, [1] some item of the launch does (kept by the launch's
// first item).
__device__ __forceinline__ int *deabs_flags(const Item &it) {
  return reinterpret_cast<int *>(it.minmax + kMinMaxBlocks);
}

struct DeAbs {
  bool absorb;  // ABS contrast: -log of the raw samples
  bool note;    // first pass: no clamp, note samples <= 0
  bool clamp;   // second pass of a sinogram whose minimum is not positive
  float mina;   // max/1e6
};

__device__ __forceinline__ float deabs_apply(const DeAbs &d, float v) {
  if (d.clamp && v <= d.mina) v = d.mina;
  return -__logf(v);
}

// FP16 storage of PHS, REF and FLT data: the largest finite magnitude a
// thread stores, against the item's half_max once the row pair is written
// (CTAS_E_RANGE above it). -log and the filter keep ABS data below 1e3.
template <int ST>
struct RangeCheck {
  float big = 0.0f;
  __device__ __forceinline__ void see(float v) {
    if constexpr (ST == ST_F16)
      if (__builtin_isfinite(v)) big = fmaxf(big, fabsf(v));
  }
  __device__ __forceinline__ void done(const Item &it, int contrast) const {
    if constexpr (ST == ST_F16)
      if (contrast != CTAS_CONTRAST_ABS && big > it.half_max) *it.status = CTAS_E_RANGE;
  }
};

// First pass: lo is the least raw sample the thread loaded (all of them
// samples of the item's rows).
__device__ __forceinline__ void deabs_note(const Item *__restrict__ items, const Item &it, float lo) {
  if (lo <= 0.0f) {
    deabs_flags(it)[0] = 1;
    deabs_flags(items[0])[1] = 1;
  }
}

// Second pass: the clamp rule from the item's partials. Every wave reduces
// the 64 partial extrema itself (one per lane), so no LDS is needed.
__device__ __forceinline__ DeAbs deabs_rule(const Item &it, int pair, int tid) {
  const float2 m = it.minmax[tid & 63];
  float mn = m.x, mx = m.y;
  for (int o = 32; o > 0; o >>= 1) {
    mn = fminf(mn, __shfl_xor(mn, o));
    mx = fmaxf(mx, __shfl_xor(mx, o));
  }
  const DeAbs d{true, false, mn <= 0.0f, static_cast<float>(static_cast<double>(mx) / 1000000.0)};
  if (pair == 0 && tid == 0 && d.clamp && d.mina <= 0.0f) *it.status = CTAS_E_DATA;
  return d;
}

// Clears the flags and the status of every item of a launch (every
// contrast: the status also carries CTAS_E_RANGE of FP16 storage).
__global__ __launch_bounds__(256) void deabs_reset_kernel(const Item *__restrict__ items, int n_items) {
  for (int i = threadIdx.x; i < n_items; i += blockDim.x) {
    int *f = deabs_flags(items[i]);
    f[0] = 0;
    f[1] = 0;
    *items[i].status = 0;
  }
}

// Row pairs of a filter launch. The first pass takes the pair of its
// workgroup; the second walks, on a small grid, the pairs of the items
// noted by the first (skipping the others item by item) and ends at once
// when there are none.
template <bool FIX, class F>
__device__ __forceinline__ void for_pairs(const Item *__restrict__ items, int n_items, int total_pairs,
                                          int contrast, F &&f) {
  const int tid = threadIdx.x;
  if (!FIX) {
    const int b = blockIdx.x;
    const Item &it = items[find_item_by_pair(items, n_items, b)];
    const bool absorb = contrast == CTAS_CONTRAST_ABS;
    f(it, b - it.pair_base, DeAbs{absorb, absorb, false, 0.0f});
    return;
  }
  if (deabs_flags(items[0])[1] == 0) return;
  const int G = gridDim.x;
  for (int b = blockIdx.x; b < total_pairs; b += G) {
    const Item &it = items[find_item_by_pair(items, n_items, b)];
    if (deabs_flags(it)[0] == 0) {
      const int end = it.pair_base + (it.N + 1) / 2;  // first pair of the next item
      b += ((end - 1 - b) / G) * G;
      continue;
    }
    const int pair = b - it.pair_base;
    const DeAbs d = deabs_rule(it, pair, tid);
    __syncthreads();  // the previous pair's last LDS reads
    f(it, pair, d);
  }
}

// Partial (min, max) of the raw sinograms noted by the first filter pass:
// grid (kMinMaxBlocks, n_items).
__global__ __launch_bounds__(256) void minmax_kernel(const Item *__restrict__ items) {
  const Item &it = items[blockIdx.y];
  if (deabs_flags(it)[0] == 0) return;
  const float *p = it.sino;
  const size_t n = static_cast<size_t>(it.N) * it.w;
  float mn = FLT_MAX, mx = -FLT_MAX;
  const size_t stride = static_cast<size_t>(blockDim.x) * gridDim.x;
  size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if ((reinterpret_cast<uintptr_t>(p) & 15u) == 0) {
    const float4 *p4 = reinterpret_cast<const float4 *>(p);
    const size_t n4 = n >> 2;
    for (size_t k = i; k < n4; k += stride) {
      const float4 v = p4[k];
      mn = fminf(mn, fminf(fminf(v.x, v.y), fminf(v.z, v.w)));
      mx = fmaxf(mx, fmaxf(fmaxf(v.x, v.y), fmaxf(v.z, v.w)));
    }
    for (size_t k = (n4 << 2) + i; k < n; k += stride) {
      mn = fminf(mn, p[k]);
      mx = fmaxf(mx, p[k]);
    }
  } else {
    for (size_t k = i; k < n; k += stride) {
      mn = fminf(mn, p[k]);
      mx = fmaxf(mx, p[k]);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    mn = fminf(mn, __shfl_xor(mn, o));
    mx = fmaxf(mx, __shfl_xor(mx, o));
  }
  __shared__ float2 part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = make_float2(mn, mx);
  __syncthreads();
  if (threadIdx.x == 0) {
    float2 r = part[0];
    for (int k = 1; k < 4; ++k) {
      r.x = fminf(r.x, part[k].x);
      r.y = fmaxf(r.y, part[k].y);
    }
    it.minmax[blockIdx.x] = r;
  }
}

// One row pair on a line of any length in LDS (buf: the complex line).
template <int ST>
__device__ __forceinline__ void filter_pair(const Item *__restrict__ items, const Item &it, int pair, float2 *buf,
                                            int contrast, const DeAbs &d) {
  const int w = it.w, N = it.N, Z = it.Z, logZ = it.logZ;
  const int psh = line_psh(Z);
  const int s0 = (Z - w) >> 1;
  const int t0 = 2 * pair, t1 = t0 + 1;
  const bool has1 = t1 < N;
  const float *__restrict__ r0 = it.sino + static_cast<size_t>(t0) * w;
  const float2 *__restrict__ tw = it.tw;
  const int tid = threadIdx.x;
  const bool absorb = d.absorb;
  auto deabs = [&](float v) { return deabs_apply(d, v); };
  float lo = FLT_MAX;

  // Padded line z[n], n in [0, Z): row samples at s0..s0+w-1 and, from the
  // inclusive copy range of lct.cpp:580-584, the next row's first sample at
  // s0+w (0 after the last row). Stored at bit-reversed positions.
  // The points t + 256 k of each thread are loaded in chunks of 32 (two for
  // Z = 16384), all loads of a chunk issued before any is used, at indices
  // clamped into the rows (samples outside are zeroed after).
  const float *r1c = has1 ? r0 + w : r0;
  const int per = (Z + 255) >> 8;  // points per thread (the last partial below 256 points)
  for (int k0 = 0; k0 < per; k0 += 32) {
    float ra[32], rb[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      const int kk = k0 + k;
      if (kk < per && tid + 256 * kk < Z) {
        const int x = tid + 256 * kk - s0;
        const bool in = x >= 0 && x < w;
        const bool in0 = in || (x == w && t0 + 1 < N), in1 = has1 && (in || (x == w && t1 + 1 < N));
        ra[k] = gload(r0, in0 ? x : 0);
        rb[k] = gload(r1c, in1 ? x : 0);
      }
    }
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      const int kk = k0 + k;
      if (kk < per && tid + 256 * kk < Z) {
        const int n = tid + 256 * kk, x = n - s0;
        const bool in = x >= 0 && x < w;
        const bool in0 = in || (x == w && t0 + 1 < N), in1 = has1 && (in || (x == w && t1 + 1 < N));
        lo = fminf(lo, fminf(ra[k], rb[k]));
        const float a = absorb ? deabs(ra[k]) : ra[k], c = absorb ? deabs(rb[k]) : rb[k];
        buf[lds_pad(__brev(static_cast<unsigned>(n)) >> (32 - logZ), psh)] =
            make_float2(in0 ? a : 0.0f, in1 ? c : 0.0f);
      }
    }
  }
  if (d.note) deabs_note(items, it, lo);
  __syncthreads();

  line_filter(buf, tw, it.W, Z, logZ, tid, psh);

  const float invZ = 1.0f / static_cast<float>(Z);  // exact: Z is a power of two
  const size_t row0 = static_cast<size_t>(t0) * w, row1 = row0 + w;
  const unsigned rsh = 32 - logZ;
  RangeCheck<ST> range;
  if (contrast != CTAS_CONTRAST_REF) {
    for (int x = tid; x < w; x += 256) {
      const float2 v = buf[lds_pad(__brev(static_cast<unsigned>(s0 + x)) >> rsh, psh)];
      const float a = v.x * invZ, c = v.y * invZ;
      gstore<ST>(it.filt, row0 + x, a);
      range.see(a);
      if (it.filt_f32) gstore<ST_F32>(it.filt_f32, row0 + x, a);
      if (has1) {
        gstore<ST>(it.filt, row1 + x, c);
        range.see(c);
        if (it.filt_f32) gstore<ST_F32>(it.filt_f32, row1 + x, c);
      }
    }
    range.done(it, contrast);
    return;
  }
  // REF (lct.cpp:590-595): prefix sum along each filtered row, one lane per
  // row, in the sequential order of std::partial_sum.
  if (tid == 0 || (tid == 64 && has1)) {
    const bool im = tid == 64;
    const size_t rowb = im ? row1 : row0;
    float run = 0.0f;
    for (int x = 0; x < w; ++x) {
      const float2 v = buf[lds_pad(__brev(static_cast<unsigned>(s0 + x)) >> rsh, psh)];
      const float f = (im ? v.y : v.x) * invZ;
      run = (x == 0) ? f : run + f;
      store_val<ST>(it.filt, rowb + x, run);
      range.see(run);
      if (it.filt_f32) it.filt_f32[rowb + x] = run;
    }
    range.done(it, contrast);
  }
}

template <int ST, bool FIX>
__global__ __launch_bounds__(256) void filter_kernel(const Item *__restrict__ items, int n_items, int total_pairs,
                                                     int contrast) {
  // The only LDS is the complex line (Z x 8 bytes): 128 KiB at Z = 16384.
  extern __shared__ float2 buf[];
  for_pairs<FIX>(items, n_items, total_pairs, contrast, [&](const Item &it, int pair, const DeAbs &d) {
    filter_pair<ST>(items, it, pair, buf, contrast, d);
  });
}

// ---- Z = 4096: radix-16 passes in registers --------------------------------
// The 4096-point line is transformed as 16 x 16 x 16: thread t holds the 16
// points t + 256 j, takes a 16-point DFT in registers, multiplies by
// w^(t k1) (w = exp(-2 pi i / 4096)), and two more such passes follow, each
// after an exchange of the values through LDS (rows padded to 17 entries).
// The thread (k1, k2) then holds X[k1 + 16 k2 + 256 k3], k3 = 0..15. After
// the window multiply the inverse runs the same steps backwards with
// conjugate factors, so that thread t ends with the filtered points
// t + 256 j in natural order: loads and stores of the rows are coalesced and
// no index is bit-reversed.
// exp(-+2 pi i m / 4096) from the half-turn table tw[m] (m < 2048).
template <bool INV>
__device__ __forceinline__ float2 w4096(const float2 *__restrict__ tw, int m) {
  m &= 4095;
  float2 t = tw[m & 2047];
  if (m & 2048) t = make_float2(-t.x, -t.y);
  return INV ? make_float2(t.x, -t.y) : t;
}

// Raw samples of one row pair of a 4096-point line, circularly shifted
// (lct.cpp:580-584 puts the row at s0 = (Z - w)/2, the next row's first
// sample at s0 + w, zeros elsewhere; here the row is at n = 0..w-1 and that
// sample at n = w; a circular shift of the line shifts the filtered line
// alike, the window multiplying the spectrum, so the row is read back from
// n = 0..w-1). With w <= 2048 thread t holds the points t + 256 j, j < 8, of
// both rows; the upper half is zero but for n = 2048 when w = 2048 (ea, ec),
// added into thread 0's first DFT. Indices are clamped into the rows; the
// samples outside the line are zeroed when the pair is filtered.
struct Raw16 {
  float a[8], c[8], ea, ec;
};

__device__ __forceinline__ void load16(const Item &it, int pair, int tid, Raw16 &r) {
  const int w = it.w, N = it.N, t0 = 2 * pair, t1 = t0 + 1;
  const bool has1 = t1 < N;
  const float *r0 = it.sino + static_cast<size_t>(t0) * w;
  const float *r1c = has1 ? r0 + w : r0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int x = tid + 256 * j;
    const bool in = x < w;
    const bool in0 = in || (x == w && t0 + 1 < N), in1 = has1 && (in || (x == w && t1 + 1 < N));
    r.a[j] = gload(r0, in0 ? x : 0);
    r.c[j] = gload(r1c, in1 ? x : 0);
  }
  const bool e0 = w == 2048 && t0 + 1 < N, e1 = w == 2048 && has1 && t1 + 1 < N;
  r.ea = gload(r0, e0 ? w : 0);
  r.ec = gload(r1c, e1 ? w : 0);
}

// One row pair: deAbs, the 4096-point transform, window, inverse, stores.
template <int ST>
__device__ __forceinline__ void filter16_pair(const Item *__restrict__ items, const Item &it, int pair,
                                              const Raw16 &r, float2 *xch, int contrast, const DeAbs &d, int tid) {
  constexpr int Z = 4096;
  auto at = [](int row, int col) { return row * 16 + (col ^ (row & 15)); };
  const int w = it.w, N = it.N;  // w <= 2048: zidth(w) = 4096
  const int t0 = 2 * pair, t1 = t0 + 1;
  const bool has1 = t1 < N;
  const float2 *__restrict__ tw = it.tw;
  const bool absorb = d.absorb;
  auto deabs = [&](float v) { return deabs_apply(d, v); };
  if (d.note) {
    float lo = fminf(r.ea, r.ec);
#pragma unroll
    for (int j = 0; j < 8; ++j) lo = fminf(lo, fminf(r.a[j], r.c[j]));
    deabs_note(items, it, lo);
  }

  float2 v[16];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int x = tid + 256 * j;
    const bool in = x < w;
    const bool in0 = in || (x == w && t0 + 1 < N), in1 = has1 && (in || (x == w && t1 + 1 < N));
    const float a = absorb ? deabs(r.a[j]) : r.a[j], c = absorb ? deabs(r.c[j]) : r.c[j];
    v[j] = make_float2(in0 ? a : 0.0f, in1 ? c : 0.0f);
  }

  // forward: thread t -> (t1, k1) -> (k1, k2)
  const int lo4 = tid & 15, hi4 = tid >> 4;
  dft16_lo8<false>(v);
  if (w == 2048 && tid == 0) {
    // z[2048] (point j = 8 of thread 0): factor exp(-2 pi i 8 k / 16) = (-1)^k
    const bool e0 = t0 + 1 < N, e1 = has1 && t1 + 1 < N;
    const float a = e0 ? (absorb ? deabs(r.ea) : r.ea) : 0.0f;
    const float c = e1 ? (absorb ? deabs(r.ec) : r.ec) : 0.0f;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      v[k].x += (k & 1) ? -a : a;
      v[k].y += (k & 1) ? -c : c;
    }
  }
  twiddle_powers(v, w4096<false>(tw, tid));
#pragma unroll
  for (int k = 0; k < 16; ++k) xch[at(tid, k)] = v[k];
  __syncthreads();
  // thread (t1 = lo4, k1 = hi4): values over t2
#pragma unroll
  for (int j = 0; j < 16; ++j) v[j] = xch[at(lo4 + 16 * j, hi4)];
  __syncthreads();
  dft16<false>(v);
  twiddle_powers(v, w4096<false>(tw, 16 * lo4));
#pragma unroll
  for (int k = 0; k < 16; ++k) xch[at(hi4 * 16 + k, lo4)] = v[k];
  __syncthreads();
  // thread (k1 = hi4, k2 = lo4): values over t1
#pragma unroll
  for (int j = 0; j < 16; ++j) v[j] = xch[at(tid, j)];
  dft16<false>(v);

  // X[k1 + 16 k2 + 256 k3] times the window and 1/Z (a power of two: the
  // scaling is exact wherever it is applied)
  const int kbase = hi4 + 16 * lo4;
  const float invZ = 1.0f / static_cast<float>(Z);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const float wk = it.W[kbase + 256 * k] * invZ;
    v[k].x *= wk;
    v[k].y *= wk;
  }

  // inverse: the same steps backwards, conjugate factors
  dft16<true>(v);
#pragma unroll
  for (int j = 0; j < 16; ++j) xch[at(tid, j)] = v[j];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = xch[at(hi4 * 16 + k, lo4)];
  __syncthreads();
  twiddle_powers(v, w4096<true>(tw, 16 * lo4));
  dft16<true>(v);
#pragma unroll
  for (int j = 0; j < 16; ++j) xch[at(lo4 + 16 * j, hi4)] = v[j];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = xch[at(tid, k)];
  twiddle_powers(v, w4096<true>(tw, tid));
  dft16_out8<true>(v);  // the points t + 256 j, j < 8: n < 2048

  const size_t row0 = static_cast<size_t>(t0) * w, row1 = row0 + w;
  RangeCheck<ST> range;
  if (contrast != CTAS_CONTRAST_REF) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int x = tid + 256 * j;
      if (x >= w) continue;
      const float a = v[j].x, c = v[j].y;
      gstore<ST>(it.filt, row0 + x, a);
      range.see(a);
      if (it.filt_f32) gstore<ST_F32>(it.filt_f32, row0 + x, a);
      if (has1) {
        gstore<ST>(it.filt, row1 + x, c);
        range.see(c);
        if (it.filt_f32) gstore<ST_F32>(it.filt_f32, row1 + x, c);
      }
    }
    range.done(it, contrast);
    return;
  }
  // REF (lct.cpp:590-595): the line in natural order through LDS, then the
  // prefix sum along each row, one lane per row, in sequential order.
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 8; ++j) xch[tid + 256 * j] = v[j];
  __syncthreads();
  if (tid == 0 || (tid == 64 && has1)) {
    const bool im = tid == 64;
    const size_t rowb = im ? row1 : row0;
    float run = 0.0f;
    for (int x = 0; x < w; ++x) {
      const float2 q = xch[x];
      const float f = im ? q.y : q.x;
      run = (x == 0) ? f : run + f;
      store_val<ST>(it.filt, rowb + x, run);
      range.see(run);
      if (it.filt_f32) it.filt_f32[rowb + x] = run;
    }
    range.done(it, contrast);
  }
}

// The 4096-point filter: one row pair per workgroup in the first pass, its
// loads issued before any is used.
template <int ST, bool FIX>
__global__ __launch_bounds__(256) void filter16_kernel(const Item *__restrict__ items, int n_items, int total_pairs,
                                                       int contrast) {
  // Exchange buffer of 256 rows of 16 points, column c of row r stored at
  // r*16 + (c ^ (r % 16)): every access pattern of the transposes spreads
  // 16 lanes over 16 distinct 8-byte bank pairs; it is the 32 KiB of the
  // line itself.
  __shared__ float2 xch[256 * 16];
  const int tid = threadIdx.x;
  for_pairs<FIX>(items, n_items, total_pairs, contrast, [&](const Item &it, int pair, const DeAbs &d) {
    Raw16 r;
    load16(it, pair, tid, r);
    filter16_pair<ST>(items, it, pair, r, xch, contrast, d, tid);
  });
}

// ---- Z = 8192: 32 x 16 x 16 in registers --------------------------------------
// Thread t holds the 32 points t + 256 j, takes a 32-point DFT (two 16-point
// DFTs and a butterfly) and multiplies by w^(t k), w = exp(-2 pi i / 8192);
// after an exchange through LDS each thread takes two items (k, t1) and the
// 16-point DFT over t2 of the values at t = t1 + 16 t2, times w^(32 t1 m2);
// after another, items (k, m2) take the DFT over t1 and hold
// X[k + 32 m2 + 512 m1], m1 < 16, for the window. The inverse runs back with
// conjugate factors to the filtered points t + 256 j in natural order.
// Twiddles are read from the FP64-computed table. One exchange buffer of
// 512 x 17 entries (69632 B) serves both exchanges.
constexpr int kPA8 = 33, kPB8 = 17;
constexpr size_t kFilter8192Lds = sizeof(float2) * 512 * kPB8;

template <int ST>
__device__ __forceinline__ void filter8192_pair(const Item *__restrict__ items, const Item &it, int pair, float2 *xb,
                                                int contrast, const DeAbs &d) {
  constexpr int Z = 8192;
  const int w = it.w, N = it.N;  // w <= 4096: zidth(w) = 8192
  const int t0 = 2 * pair, t1r = t0 + 1;
  const bool has1 = t1r < N;
  const float *__restrict__ r0 = it.sino + static_cast<size_t>(t0) * w;
  const float *__restrict__ r1 = r0 + w;
  const float2 *__restrict__ tw = it.tw;
  const int tid = threadIdx.x;
  const bool absorb = d.absorb;
  auto deabs = [&](float v) { return deabs_apply(d, v); };

  // Padded line, circularly shifted as in filter16_kernel: the row at
  // n = 0..w-1 and the next row's first sample at n = w (lct.cpp:580-584;
  // 0 after the last row). With w <= 4096 thread t holds the points
  // t + 256 j, j < 16, and n = 4096 (w = 4096) is added into thread 0's
  // first DFT. All loads are issued before any is used, at indices clamped
  // into the rows (samples outside are zeroed afterwards).
  const float *r1c = has1 ? r1 : r0;
  float2 v[32];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int x = tid + 256 * j;
    const bool in = x < w;
    const bool in0 = in || (x == w && t0 + 1 < N), in1 = has1 && (in || (x == w && t1r + 1 < N));
    v[j] = make_float2(gload(r0, in0 ? x : 0), gload(r1c, in1 ? x : 0));
  }
  if (d.note) {
    float lo = FLT_MAX;
#pragma unroll
    for (int j = 0; j < 16; ++j) lo = fminf(lo, fminf(v[j].x, v[j].y));
    deabs_note(items, it, lo);
  }
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int x = tid + 256 * j;
    const bool in = x < w;
    const bool in0 = in || (x == w && t0 + 1 < N), in1 = has1 && (in || (x == w && t1r + 1 < N));
    const float a = absorb ? deabs(v[j].x) : v[j].x, c = absorb ? deabs(v[j].y) : v[j].y;
    v[j] = make_float2(in0 ? a : 0.0f, in1 ? c : 0.0f);
  }

  dft32_lo16<false>(v);
  if (w == 4096 && tid == 0) {
    // z[4096] (point j = 16 of thread 0): factor exp(-2 pi i 16 k / 32) = (-1)^k
    const bool e0 = t0 + 1 < N, e1 = has1 && t1r + 1 < N;
    const float a = e0 ? (absorb ? deabs(r0[w]) : r0[w]) : 0.0f;
    const float c = e1 ? (absorb ? deabs(r1[w]) : r1[w]) : 0.0f;
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      v[k].x += (k & 1) ? -a : a;
      v[k].y += (k & 1) ? -c : c;
    }
  }
#pragma unroll
  for (int k = 1; k < 32; ++k) v[k] = cmul(v[k], w8192(tw, tid * k));
#pragma unroll
  for (int k = 0; k < 32; ++k) xb[tid * kPA8 + k] = v[k];
  __syncthreads();
  float2 u[2][16];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int i = tid + 256 * s, k = i >> 4, q = i & 15;  // q = t1
#pragma unroll
    for (int t2 = 0; t2 < 16; ++t2) u[s][t2] = xb[(q + 16 * t2) * kPA8 + k];
  }
  __syncthreads();  // the first exchange is read; the second takes its place
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int i = tid + 256 * s, k = i >> 4, q = i & 15;
    dft16<false>(u[s]);
#pragma unroll
    for (int m2 = 1; m2 < 16; ++m2) u[s][m2] = cmul(u[s][m2], w8192(tw, 32 * q * m2));
#pragma unroll
    for (int m2 = 0; m2 < 16; ++m2) xb[(k * 16 + m2) * kPB8 + q] = u[s][m2];
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int i = tid + 256 * s, k = i >> 4, q = i & 15;  // q = m2
#pragma unroll
    for (int t1 = 0; t1 < 16; ++t1) u[s][t1] = xb[(k * 16 + q) * kPB8 + t1];
    dft16<false>(u[s]);
#pragma unroll
    for (int m1 = 0; m1 < 16; ++m1) {
      // the window and 1/Z (a power of two: exact wherever applied)
      const float wk = it.W[k + 32 * q + 512 * m1] * (1.0f / static_cast<float>(Z));
      u[s][m1].x *= wk;
      u[s][m1].y *= wk;
    }
    dft16<true>(u[s]);
#pragma unroll
    for (int t1 = 0; t1 < 16; ++t1) xb[(k * 16 + q) * kPB8 + t1] = u[s][t1];
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int i = tid + 256 * s, k = i >> 4, q = i & 15;  // q = t1
#pragma unroll
    for (int m2 = 0; m2 < 16; ++m2) u[s][m2] = xb[(k * 16 + m2) * kPB8 + q];
  }
  __syncthreads();  // the second exchange is read; the first takes its place
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int i = tid + 256 * s, k = i >> 4, q = i & 15;
#pragma unroll
    for (int m2 = 1; m2 < 16; ++m2) u[s][m2] = cmulc(u[s][m2], w8192(tw, 32 * q * m2));
    dft16<true>(u[s]);
#pragma unroll
    for (int t2 = 0; t2 < 16; ++t2) xb[(q + 16 * t2) * kPA8 + k] = u[s][t2];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 32; ++k) v[k] = xb[tid * kPA8 + k];
#pragma unroll
  for (int k = 1; k < 32; ++k) v[k] = cmulc(v[k], w8192(tw, tid * k));
  dft32_out16<true>(v);  // the points t + 256 j, j < 16: n < 4096

  const size_t row0 = static_cast<size_t>(t0) * w, row1 = row0 + w;
  RangeCheck<ST> range;
  if (contrast != CTAS_CONTRAST_REF) {
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int x = tid + 256 * j;
      if (x >= w) continue;
      const float a = v[j].x, c = v[j].y;
      gstore<ST>(it.filt, row0 + x, a);
      range.see(a);
      if (it.filt_f32) gstore<ST_F32>(it.filt_f32, row0 + x, a);
      if (has1) {
        gstore<ST>(it.filt, row1 + x, c);
        range.see(c);
        if (it.filt_f32) gstore<ST_F32>(it.filt_f32, row1 + x, c);
      }
    }
    range.done(it, contrast);
    return;
  }
  // REF (lct.cpp:590-595): the line in natural order through LDS, then the
  // prefix sum along each row, one lane per row, in sequential order.
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 16; ++j) xb[tid + 256 * j] = v[j];
  __syncthreads();
  if (tid == 0 || (tid == 64 && has1)) {
    const bool im = tid == 64;
    const size_t rowb = im ? row1 : row0;
    float run = 0.0f;
    for (int x = 0; x < w; ++x) {
      const float2 q = xb[x];
      const float f = im ? q.y : q.x;
      run = (x == 0) ? f : run + f;
      gstore<ST>(it.filt, rowb + x, run);
      range.see(run);
      if (it.filt_f32) it.filt_f32[rowb + x] = run;
    }
    range.done(it, contrast);
  }
}

template <int ST, bool FIX>
__global__ __launch_bounds__(256) void filter8192_kernel(const Item *__restrict__ items, int n_items, int total_pairs,
                                                         int contrast) {
  extern __shared__ float2 xb[];
  for_pairs<FIX>(items, n_items, total_pairs, contrast, [&](const Item &it, int pair, const DeAbs &d) {
    filter8192_pair<ST>(items, it, pair, xb, contrast, d);
  });
}

// FLT contrast: no filtering, only the storage conversion (and float copy).
template <int ST>
__global__ __launch_bounds__(256) void passthrough_kernel(const Item *__restrict__ items,
                                                          int n_items) {
  const int b = blockIdx.x;
  const Item &it = items[find_item_by_pair(items, n_items, b)];
  const int pair = b - it.pair_base;
  const int w = it.w, N = it.N;
  const int t0 = 2 * pair;
  const int rows = (t0 + 1 < N) ? 2 : 1;
  const size_t base = static_cast<size_t>(t0) * w;
  RangeCheck<ST> range;
  for (int x = threadIdx.x; x < rows * w; x += 256) {
    const float v = it.sino[base + x];
    store_val<ST>(it.filt, base + x, v);
    range.see(v);
    if (it.filt_f32) it.filt_f32[base + x] = v;
  }
  range.done(it, CTAS_CONTRAST_FLT);
}

// Zeros of the pads around one item's filtered sinogram (16-bit samples).
__global__ void pad_zero_kernel(const Item *__restrict__ items) {
  const Item &it = items[blockIdx.x];
  uint16_t *f = static_cast<uint16_t *>(it.filt);
  const long long n = static_cast<long long>(it.N) * it.w;
  for (int j = threadIdx.x; j < it.pad; j += blockDim.x) {
    f[j - it.pad] = 0;
    f[n + j] = 0;
  }
}

}  // namespace

hipError_t launch_pad_zero(const Item *d_items, int n_items, hipStream_t s) {
  hipLaunchKernelGGL(pad_zero_kernel, dim3(n_items), dim3(256), 0, s, d_items);
  return hipGetLastError();
}

namespace {

using FilterKernel = void (*)(const Item *, int, int, int);

template <template <int, bool> class K, bool FIX>
struct Pick {
  static FilterKernel of(int store) {
    return store == ST_F32 ? K<ST_F32, FIX>::fn : store == ST_F16 ? K<ST_F16, FIX>::fn : K<ST_BF16, FIX>::fn;
  }
};
template <int ST, bool FIX> struct K16 { static constexpr FilterKernel fn = filter16_kernel<ST, FIX>; };
template <int ST, bool FIX> struct K8192 { static constexpr FilterKernel fn = filter8192_kernel<ST, FIX>; };
template <int ST, bool FIX> struct KLds { static constexpr FilterKernel fn = filter_kernel<ST, FIX>; };

// One filter pass: the first takes one workgroup per row pair, the second as
// many workgroups as the device holds at once (one round), walking the pairs
// of the noted items. Both choose the same kernel for the same lines.
template <bool FIX>
hipError_t filter_pass(const Item *d_items, int n_items, int total_pairs, int min_logZ, int max_logZ,
                       int contrast, int store, hipStream_t s) {
  int dev = 0, lds_max = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return hipErrorInvalidDevice;
  FilterKernel k;
  size_t lds = 0;
  if (min_logZ == 12 && max_logZ == 12) {
    k = Pick<K16, FIX>::of(store);
  } else if (min_logZ == 13 && max_logZ == 13 && kFilter8192Lds <= static_cast<size_t>(lds_max)) {
    // the 8192-point register kernel's exchange buffer fits the device's
    // LDS per workgroup; otherwise the radix-2 LDS kernel filters these lines
    k = Pick<K8192, FIX>::of(store);
    lds = kFilter8192Lds;
  } else {
    // a padded line of any item fits in that of the largest
    lds = sizeof(float2) * padded_len(1 << max_logZ, line_psh(1 << max_logZ));
    if (lds > static_cast<size_t>(lds_max)) return hipErrorInvalidValue;
    k = Pick<KLds, FIX>::of(store);
  }
  int grid = total_pairs;
  if (FIX) {
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k), 256, lds) !=
            hipSuccess)
      return hipErrorInvalidDevice;
    grid = std::max(1, std::min(total_pairs, std::max(per_cu, 1) * cus));
  }
  hipLaunchKernelGGL(k, dim3(grid), dim3(256), lds, s, d_items, n_items, total_pairs, contrast);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_filter(const Item *d_items, int n_items, int total_pairs, int min_logZ, int max_logZ,
                         int contrast, int store, hipStream_t s) {
  if (total_pairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(deabs_reset_kernel, dim3(1), dim3(256), 0, s, d_items, n_items);
  if (const hipError_t e = hipGetLastError(); e != hipSuccess) return e;
  if (contrast == CTAS_CONTRAST_FLT) {
    switch (store) {
      case ST_F32: hipLaunchKernelGGL(passthrough_kernel<ST_F32>, dim3(total_pairs), dim3(256), 0, s, d_items, n_items); break;
      case ST_F16: hipLaunchKernelGGL(passthrough_kernel<ST_F16>, dim3(total_pairs), dim3(256), 0, s, d_items, n_items); break;
      default: hipLaunchKernelGGL(passthrough_kernel<ST_BF16>, dim3(total_pairs), dim3(256), 0, s, d_items, n_items); break;
    }
    return hipGetLastError();
  }
  const bool absorb = contrast == CTAS_CONTRAST_ABS;
  if (const hipError_t e = filter_pass<false>(d_items, n_items, total_pairs, min_logZ, max_logZ, contrast, store, s);
      e != hipSuccess || !absorb)
    return e;
  // sinograms with a sample <= 0: their extrema, then their rows again with the clamp
  hipLaunchKernelGGL(minmax_kernel, dim3(kMinMaxBlocks, n_items), dim3(256), 0, s, d_items);
  if (const hipError_t e = hipGetLastError(); e != hipSuccess) return e;
  return filter_pass<true>(d_items, n_items, total_pairs, min_logZ, max_logZ, contrast, store, s);
}

}  // namespace ctas
