// Pixel-driven backprojection (CTrec::reconstructOnCPU, lct.cpp:688-729, and
// the fbp kernel of ct.cl:5-44).
//
// A 256-thread workgroup owns a tile of output pixels of one slice (64x64,
// 128x64 or 128x128; each thread a 4x4, 8x4 or 8x8 lattice with stride 16).
// The angles are taken in passes of AC. For each angle of a pass the
// workgroup stages into LDS the window of the filtered row that the tile can
// touch (a fixed length L >= the tile's diagonal + margin, placed around the
// ray coordinate of the tile centre), together with the angle's constants
// {base, sin, cos, LDS offset}. Every pixel then reads one sample (CTAS mode)
// or one sample pair (LINEAR mode) per angle from LDS: a gather, not a dense
// contraction, so LDS bandwidth and VALU issue bound it, not MFMA or HBM.
//
// The parity mode (CTAS sampling, float32 storage) reproduces the reference
// arithmetic exactly: the ray coordinate
// is ((c + ((1-s)-co)*hp) + s*(j-c)) + co*i with every product and sum
// rounded separately (FP contraction is off in this file), truncated to an
// integer, and the samples are summed in angle order in float32. Given the
// same filtered sinogram and angle table, the output is bit-identical to the
// CPU restatement. The sample is addressed as the reference does (flat
// offset theta*w + idx of the slice's buffer), so even an index at w reads
// the next row's first sample. LINEAR mode interpolates between bin centres
// k + 1/2 (samples outside the detector are 0). LINEAR and the FP16/BF16
// storage modes take the angle constants from an FP64 table (theta = i*arc/N
// in double) and form the per-angle base term in FP64, rounding once: the
// reference's float table puts, e.g., cos(90 deg) at -4.4e-8, which moves
// every exact-integer ray coordinate of that angle below the integer.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "ctas_internal.hpp"

namespace ctas {
namespace {

constexpr int isqrt_ceil(int v) {
  int r = 0;
  while (r * r < v) ++r;
  return r;
}

template <int TR_, int TC_>
struct Geo {
  static constexpr int TR = TR_, TC = TC_;
  static constexpr int RT = 16 * TR, CT = 16 * TC;
  // diagonal of the tile (pixel centres), rounded up
  static constexpr int D = isqrt_ceil((RT - 1) * (RT - 1) + (CT - 1) * (CT - 1));
  static constexpr int HALF = (D + 1) / 2;
  static constexpr int L = D + 8;
};

// Waves per SIMD the register budget of each tile kind is sized for.
template <int KIND> struct KindGeo;
template <> struct KindGeo<TK_64x64> : Geo<4, 4> { static constexpr int WAVES = 4; };
template <> struct KindGeo<TK_128x64> : Geo<8, 4> { static constexpr int WAVES = 3; };
template <> struct KindGeo<TK_128x128> : Geo<8, 8> { static constexpr int WAVES = 2; };

template <int ST>
__device__ __forceinline__ float load_val(const void *base, size_t i) {
  if constexpr (ST == ST_F32) {
    return static_cast<const float *>(base)[i];
  } else if constexpr (ST == ST_F16) {
    return static_cast<float>(static_cast<const _Float16 *>(base)[i]);
  } else {
    return __uint_as_float(static_cast<uint32_t>(static_cast<const uint16_t *>(base)[i]) << 16);
  }
}

template <int MODE> struct WinT { using type = float; };
template <> struct WinT<MODE_LINEAR> { using type = float2; };

// LDS needed by one tile kind (window + per-angle constants + lo table).
template <int KIND, int AC, int MODE>
constexpr size_t lds_bytes() {
  return sizeof(typename WinT<MODE>::type) * AC * KindGeo<KIND>::L + sizeof(float4) * AC +
         sizeof(int) * AC;
}

// True when every pixel of the block rows [r0, r0+nr) x cols [c0, c0+nc)
// lies outside the reconstruction disc (radius at most hp + |c| around the
// axis point (hp, hp + c)) or the column band, so that all are masked.
__device__ __forceinline__ bool wave_outside(const Item &it, float c, int r0, int nr, int c0, int nc) {
  const int hp = it.w / 2;
  const float acent = fabsf(c);
  const float ax = static_cast<float>(hp), ay = static_cast<float>(hp) + c;
  const float dx = fmaxf(fmaxf(static_cast<float>(c0) - ax, ax - static_cast<float>(c0 + nc - 1)), 0.0f);
  const float dy = fmaxf(fmaxf(static_cast<float>(r0) - ay, ay - static_cast<float>(r0 + nr - 1)), 0.0f);
  const float rmax = static_cast<float>(hp) + acent + 2.0f;
  return (dx * dx + dy * dy > rmax * rmax) || (static_cast<float>(c0 + nc - 1) <= acent) ||
         (static_cast<float>(c0) >= static_cast<float>(it.w) - acent - 1.0f);
}

__device__ __forceinline__ int find_item_by_tile(const Item *items, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].tile_base <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// One output tile of one item.
template <int KIND, int AC, int MODE, int ST>
__device__ __forceinline__ void bp_tile(const Item &it, int tile, unsigned char *lds) {
  using G = KindGeo<KIND>;
  using WT = typename WinT<MODE>::type;
  constexpr int TR = G::TR, TC = G::TC, RT = G::RT, CT = G::CT, L = G::L;
  WT *win = reinterpret_cast<WT *>(lds);
  float4 *prm = reinterpret_cast<float4 *>(lds + sizeof(WT) * AC * L);
  int *los = reinterpret_cast<int *>(lds + sizeof(WT) * AC * L + sizeof(float4) * AC);

  const int w = it.w, N = it.N;
  const int hp = w / 2;
  const float c = it.center;
  const int ty0 = (tile / it.tiles_x) * RT;
  const int tx0 = (tile % it.tiles_x) * CT;
  const int tid = threadIdx.x;
  // wave v owns the tile rows [v 4TR, (v+1) 4TR); lane (ty4, tx) the pixels
  // (v 4TR + ty4 + 4r, tx + 16q): one LDS read instruction covers a 4 x 16
  // pixel block, and a wave whose rows lie outside the disc skips its work
  const int tx = tid & 15, ty = (tid >> 6) * (4 * TR) + ((tid >> 4) & 3);
  float *__restrict__ out = it.out;

  // Tiles that lie wholly outside the reconstruction disc (radius at most
  // hp + |c| around the axis point (hp, hp + c)) or the column band are
  // zero, as every pixel of them is masked.
  {
    const float acent = fabsf(c);
    const float ax = static_cast<float>(hp), ay = static_cast<float>(hp) + c;
    const float dx = fmaxf(fmaxf(static_cast<float>(tx0) - ax, ax - static_cast<float>(tx0 + CT - 1)), 0.0f);
    const float dy = fmaxf(fmaxf(static_cast<float>(ty0) - ay, ay - static_cast<float>(ty0 + RT - 1)), 0.0f);
    const float rmax = static_cast<float>(hp) + acent + 2.0f;
    const bool outside = (dx * dx + dy * dy > rmax * rmax) ||
                         (static_cast<float>(tx0 + CT - 1) <= acent) ||
                         (static_cast<float>(tx0) >= static_cast<float>(w) - acent - 1.0f);
    if (outside) {
#pragma unroll
      for (int r = 0; r < TR; ++r) {
        const int row = ty0 + ty + 4 * r;
#pragma unroll
        for (int q = 0; q < TC; ++q) {
          const int col = tx0 + tx + 16 * q;
          if (row < w && col < w) out[static_cast<size_t>(row) * w + col] = 0.0f;
        }
      }
      return;
    }
  }

  float jc[TR], fi[TC];
#pragma unroll
  for (int r = 0; r < TR; ++r) jc[r] = static_cast<float>(ty0 + ty + 4 * r) - c;
#pragma unroll
  for (int q = 0; q < TC; ++q) fi[q] = static_cast<float>(tx0 + tx + 16 * q);
  float acc[TR][TC];
#pragma unroll
  for (int r = 0; r < TR; ++r)
#pragma unroll
    for (int q = 0; q < TC; ++q) acc[r][q] = 0.0f;

  const float ycm = static_cast<float>(ty0) + 0.5f * (RT - 1) - c;
  const float xcm = static_cast<float>(tx0) + 0.5f * (CT - 1);
  const void *__restrict__ filt = it.filt;
  const bool wskip = wave_outside(it, c, ty0 + (tid >> 6) * (4 * TR), 4 * TR, tx0, CT);
  const long long total = static_cast<long long>(N) * w;

  for (int a0 = 0; a0 < N; a0 += AC) {
    const int na = min(AC, N - a0);
    __syncthreads();  // the previous pass is done with the LDS
    if (tid < na) {
      float s, co, base;
      if (it.sincos64) {
        // precision modes: angle constants and centre shift in FP64
        const double2 sd = it.sincos64[a0 + tid];
        double bd = static_cast<double>(c) + static_cast<double>(hp) * ((1.0 - sd.x) - sd.y);
        if (MODE == MODE_LINEAR) bd += 0.5;
        s = static_cast<float>(sd.x);
        co = static_cast<float>(sd.y);
        base = static_cast<float>(bd);
      } else {
        const float2 sc = it.sincos[a0 + tid];
        s = sc.x;
        co = sc.y;
        base = c + ((1.0f - s) - co) * static_cast<float>(hp);
        if (MODE == MODE_LINEAR) base = base + 0.5f;
      }
      const float tc = base + s * ycm + co * xcm;
      const int lo = static_cast<int>(floorf(tc)) - G::HALF - 2;
      los[tid] = lo;
      const int offb = (tid * L - lo) * static_cast<int>(sizeof(WT));
      prm[tid] = make_float4(base, s, co, __int_as_float(offb));
    }
    __syncthreads();
    // Stage the windows.
    for (int e = tid; e < na * L; e += 256) {
      const int a = e / L;
      const int o = e - a * L;
      const int k = los[a] + o;
      const int ang = a0 + a;
      if constexpr (MODE == MODE_CTAS) {
        const long long flat = static_cast<long long>(ang) * w + k;
        win[e] = (flat >= 0 && flat < total) ? load_val<ST>(filt, static_cast<size_t>(flat)) : 0.0f;
      } else {
        const size_t rb = static_cast<size_t>(ang) * w;
        const float q1 = (k >= 0 && k < w) ? load_val<ST>(filt, rb + k) : 0.0f;
        const float q0 = (k - 1 >= 0 && k - 1 < w) ? load_val<ST>(filt, rb + k - 1) : 0.0f;
        win[e] = make_float2(q0, q1 - q0);
      }
    }
    __syncthreads();
    if (wskip) continue;
    const unsigned char *wb0 = reinterpret_cast<const unsigned char *>(win);
#pragma unroll 2
    for (int a = 0; a < na; ++a) {
      const float4 p = prm[a];
      const unsigned char *wb = wb0 + __float_as_int(p.w);
      float aj[TR], ci[TC];
#pragma unroll
      for (int r = 0; r < TR; ++r) aj[r] = p.x + p.y * jc[r];
#pragma unroll
      for (int q = 0; q < TC; ++q) ci[q] = p.z * fi[q];
#pragma unroll
      for (int r = 0; r < TR; ++r) {
        // Issue the TC reads of a row before consuming any, so that the LDS
        // latency overlaps (the compiler would otherwise wait on each read).
        if constexpr (MODE == MODE_CTAS) {
          float vv[TC];
#pragma unroll
          for (int q = 0; q < TC; ++q) {
            const int k = static_cast<int>(aj[r] + ci[q]);
            vv[q] = *reinterpret_cast<const float *>(wb + (k << 2));
          }
#pragma unroll
          for (int q = 0; q < TC; ++q) acc[r][q] += vv[q];
        } else {
          float2 vv[TC];
          float ff[TC];
#pragma unroll
          for (int q = 0; q < TC; ++q) {
            const float t = aj[r] + ci[q];
            const int k = static_cast<int>(t);
            ff[q] = __builtin_amdgcn_fractf(t);
            vv[q] = *reinterpret_cast<const float2 *>(wb + (k << 3));
          }
#pragma unroll
          for (int q = 0; q < TC; ++q) {
            acc[r][q] += vv[q].x;
            acc[r][q] = fmaf(ff[q], vv[q].y, acc[r][q]);
          }
        }
      }
    }
  }

  // Masks and scaling, in the reference's float arithmetic.
  const float acent = fabsf(c);
  const float up = static_cast<float>(hp) + c, dn = static_cast<float>(hp) - c;
  const float recCof = it.recCof;
#pragma unroll
  for (int r = 0; r < TR; ++r) {
    const int row = ty0 + ty + 4 * r;
    if (row >= w) continue;
    float dy = static_cast<float>(row - hp) - c;
    dy *= dy;
#pragma unroll
    for (int q = 0; q < TC; ++q) {
      const int col = tx0 + tx + 16 * q;
      if (col >= w) continue;
      int ip2 = col - hp;
      ip2 *= ip2;
      const float d2 = dy + static_cast<float>(ip2);
      const float fr = static_cast<float>(row), fc = static_cast<float>(col);
      const bool masked = (fc <= acent) || (fc >= static_cast<float>(w) - acent - 1.0f) ||
                          (fr <= up && d2 >= up * up) || (fr >= up && d2 >= dn * dn);
      float v = acc[r][q];
      if (recCof != 0.0f) v *= recCof;
      v = v / static_cast<float>(N);
      out[static_cast<size_t>(row) * w + col] = masked ? 0.0f : v;
    }
  }
}

// One output tile of one item, FP16/BF16 storage with CTAS (floor)
// sampling in the precision path (FP64 angle constants). The LDS windows
// hold the 16-bit samples as stored. The ray coordinate of each pixel is a
// 16.16 fixed-point integer: per angle, the tile's base term (T0 - lo + the
// window slot offset, from FP64) plus a row term and a column term, each
// advanced by an integer increment, so that each sample costs one integer
// add and one shift (the byte offset is (u >> 16) << 1). Two angles are
// summed into a pixel's float accumulator with one dot2 (f16 or bf16 pair
// times (1, 1)). Accuracy: the fixed-point coordinate is within ~1e-4 px of
// the FP64 one, finer than the float arithmetic of the parity path.
template <int ST>
struct Pair16;
template <>
struct Pair16<ST_F16> {
  typedef _Float16 v2 __attribute__((ext_vector_type(2)));
  static __device__ __forceinline__ float dot(v2 a, float acc) {
    const v2 one = {static_cast<_Float16>(1.0f), static_cast<_Float16>(1.0f)};
    return __builtin_amdgcn_fdot2(a, one, acc, false);
  }
  static __device__ __forceinline__ v2 make(uint16_t lo, uint16_t hi) {
    return v2{__builtin_bit_cast(_Float16, lo), __builtin_bit_cast(_Float16, hi)};
  }
};
template <>
struct Pair16<ST_BF16> {
  typedef __bf16 v2 __attribute__((ext_vector_type(2)));
  static __device__ __forceinline__ float dot(v2 a, float acc) {
    const v2 one = {static_cast<__bf16>(1.0f), static_cast<__bf16>(1.0f)};
    return __builtin_amdgcn_fdot2_f32_bf16(a, one, acc, false);
  }
  static __device__ __forceinline__ v2 make(uint16_t lo, uint16_t hi) {
    return v2{__builtin_bit_cast(__bf16, lo), __builtin_bit_cast(__bf16, hi)};
  }
};

template <int KIND, int AC>
constexpr size_t lds_bytes_fx() {
  return sizeof(uint16_t) * AC * KindGeo<KIND>::L + sizeof(int4) * AC + sizeof(int) * AC;
}

template <int KIND, int AC, int ST>
__device__ __forceinline__ void bp_tile_fx(const Item &it, int tile, unsigned char *lds) {
  using G = KindGeo<KIND>;
  using P = Pair16<ST>;
  constexpr int TR = G::TR, TC = G::TC, RT = G::RT, CT = G::CT, L = G::L;
  static_assert(AC % 2 == 0, "angles are summed in pairs");
  uint16_t *win = reinterpret_cast<uint16_t *>(lds);
  int4 *prm = reinterpret_cast<int4 *>(lds + sizeof(uint16_t) * AC * L);
  int *los = reinterpret_cast<int *>(lds + sizeof(uint16_t) * AC * L + sizeof(int4) * AC);

  const int w = it.w, N = it.N;
  const int hp = w / 2;
  const float c = it.center;
  const int ty0 = (tile / it.tiles_x) * RT;
  const int tx0 = (tile % it.tiles_x) * CT;
  const int tid = threadIdx.x;
  // wave v owns the tile rows [v 4TR, (v+1) 4TR); lane (ty4, tx) the pixels
  // (v 4TR + ty4 + 4r, tx + 16q): one LDS read instruction covers a 4 x 16
  // pixel block, and a wave whose rows lie outside the disc skips its work
  const int tx = tid & 15, ty = (tid >> 6) * (4 * TR) + ((tid >> 4) & 3);
  float *__restrict__ out = it.out;
  {
    const float acent = fabsf(c);
    const float ax = static_cast<float>(hp), ay = static_cast<float>(hp) + c;
    const float dx = fmaxf(fmaxf(static_cast<float>(tx0) - ax, ax - static_cast<float>(tx0 + CT - 1)), 0.0f);
    const float dy = fmaxf(fmaxf(static_cast<float>(ty0) - ay, ay - static_cast<float>(ty0 + RT - 1)), 0.0f);
    const float rmax = static_cast<float>(hp) + acent + 2.0f;
    const bool outside = (dx * dx + dy * dy > rmax * rmax) ||
                         (static_cast<float>(tx0 + CT - 1) <= acent) ||
                         (static_cast<float>(tx0) >= static_cast<float>(w) - acent - 1.0f);
    if (outside) {
#pragma unroll
      for (int r = 0; r < TR; ++r) {
        const int row = ty0 + ty + 4 * r;
#pragma unroll
        for (int q = 0; q < TC; ++q) {
          const int col = tx0 + tx + 16 * q;
          if (row < w && col < w) out[static_cast<size_t>(row) * w + col] = 0.0f;
        }
      }
      return;
    }
  }
  float acc[TR][TC];
#pragma unroll
  for (int r = 0; r < TR; ++r)
#pragma unroll
    for (int q = 0; q < TC; ++q) acc[r][q] = 0.0f;
  const uint16_t *__restrict__ filt = static_cast<const uint16_t *>(it.filt);
  const long long total = static_cast<long long>(N) * w;
  const float fty = static_cast<float>(ty), ftx = static_cast<float>(tx);
  const bool wskip = wave_outside(it, c, ty0 + (tid >> 6) * (4 * TR), 4 * TR, tx0, CT);

  for (int a0 = 0; a0 < N; a0 += AC) {
    const int na = min(AC, N - a0);
    const int np = (na + 1) & ~1;  // slots used, padded to a pair
    __syncthreads();
    if (tid < np) {
      // an odd pass ends with an all-zero slot that reuses the last angle's
      // geometry
      const int ang = a0 + min(tid, na - 1);
      const double2 sd = it.sincos64[ang];
      const double cd = static_cast<double>(c);
      const double T0 = cd + hp * ((1.0 - sd.x) - sd.y) + sd.x * (ty0 - cd) + sd.y * tx0;
      const double tc = T0 + sd.x * (0.5 * (RT - 1)) + sd.y * (0.5 * (CT - 1));
      const int lo = static_cast<int>(floor(tc)) - G::HALF - 2;
      los[tid] = (tid < na) ? lo : INT_MIN;
      const int bt = static_cast<int>(llrint((T0 - lo + static_cast<double>(tid) * L) * 65536.0));
      const float Ss = static_cast<float>(sd.x * 65536.0), Cs = static_cast<float>(sd.y * 65536.0);
      prm[tid] = make_int4(bt, __float_as_int(Ss), __float_as_int(Cs), 0);
    }
    __syncthreads();
    for (int e = tid; e < np * L; e += 256) {
      const int a = e / L;
      const int o = e - a * L;
      const int lo = los[a];
      uint16_t v = 0;
      if (lo != INT_MIN) {
        const long long flat = static_cast<long long>(a0 + a) * w + lo + o;
        if (flat >= 0 && flat < total) v = filt[flat];
      }
      win[e] = v;
    }
    __syncthreads();
    if (wskip) continue;
    const unsigned char *wb = reinterpret_cast<const unsigned char *>(win);
    for (int a = 0; a < np; a += 2) {
      int A0[TR], A1[TR], C0[TC], C1[TC];
      {
        const int4 p = prm[a];
        const float Ss = __int_as_float(p.y), Cs = __int_as_float(p.z);
        const int dS = static_cast<int>(Ss * 4.0f), dC = static_cast<int>(Cs * 16.0f);
        A0[0] = p.x + static_cast<int>(Ss * fty);
#pragma unroll
        for (int r = 1; r < TR; ++r) A0[r] = A0[r - 1] + dS;
        C0[0] = static_cast<int>(Cs * ftx);
#pragma unroll
        for (int q = 1; q < TC; ++q) C0[q] = C0[q - 1] + dC;
      }
      {
        const int4 p = prm[a + 1];
        const float Ss = __int_as_float(p.y), Cs = __int_as_float(p.z);
        const int dS = static_cast<int>(Ss * 4.0f), dC = static_cast<int>(Cs * 16.0f);
        A1[0] = p.x + static_cast<int>(Ss * fty);
#pragma unroll
        for (int r = 1; r < TR; ++r) A1[r] = A1[r - 1] + dS;
        C1[0] = static_cast<int>(Cs * ftx);
#pragma unroll
        for (int q = 1; q < TC; ++q) C1[q] = C1[q - 1] + dC;
      }
#pragma unroll
      for (int r = 0; r < TR; ++r) {
        typename P::v2 v[TC];
#pragma unroll
        for (int q = 0; q < TC; ++q) {
          const unsigned u0 = static_cast<unsigned>(A0[r] + C0[q]);
          const unsigned u1 = static_cast<unsigned>(A1[r] + C1[q]);
          const uint16_t x0 = *reinterpret_cast<const uint16_t *>(wb + ((u0 >> 16) << 1));
          const uint16_t x1 = *reinterpret_cast<const uint16_t *>(wb + ((u1 >> 16) << 1));
          v[q] = P::make(x0, x1);
        }
#pragma unroll
        for (int q = 0; q < TC; ++q) acc[r][q] = P::dot(v[q], acc[r][q]);
      }
    }
  }

  const float acent = fabsf(c);
  const float up = static_cast<float>(hp) + c, dn = static_cast<float>(hp) - c;
  const float recCof = it.recCof;
#pragma unroll
  for (int r = 0; r < TR; ++r) {
    const int row = ty0 + ty + 4 * r;
    if (row >= w) continue;
    float dy = static_cast<float>(row - hp) - c;
    dy *= dy;
#pragma unroll
    for (int q = 0; q < TC; ++q) {
      const int col = tx0 + tx + 16 * q;
      if (col >= w) continue;
      int ip2 = col - hp;
      ip2 *= ip2;
      const float d2 = dy + static_cast<float>(ip2);
      const float fr = static_cast<float>(row), fc = static_cast<float>(col);
      const bool masked = (fc <= acent) || (fc >= static_cast<float>(w) - acent - 1.0f) ||
                          (fr <= up && d2 >= up * up) || (fr >= up && d2 >= dn * dn);
      float v = acc[r][q];
      if (recCof != 0.0f) v *= recCof;
      v = v / static_cast<float>(N);
      out[static_cast<size_t>(row) * w + col] = masked ? 0.0f : v;
    }
  }
}

// ---- FP16/BF16 with pair words -------------------------------------------
// LDS holds, for every window position o, the 32-bit word (v[o], v[o+1]).
// A lane owns vertically adjacent pixel pairs (rows j, j+1): their ray
// coordinates differ by s = sin(theta), so with floor sampling row j+1
// reads v[k] or v[k+1] (k = floor(t_j)) for s >= 0, v[k-1] or v[k] for
// s < 0 (the pair is then taken at k-1). One ds_read_b32 serves both pixels;
// each pixel adds its half with a dot2 against a weight word that is 1.0 in
// the half it takes. Coordinates are 16.16 fixed point formed from FP64
// per-angle constants, as in bp_tile_fx.
template <int ST> struct One16;
template <> struct One16<ST_F16> { static constexpr unsigned LO = 0x3C00u, HI = 0x3C000000u; };
template <> struct One16<ST_BF16> { static constexpr unsigned LO = 0x3F80u, HI = 0x3F800000u; };

template <int ST>
__device__ __forceinline__ float dot2w(unsigned a, unsigned b, float acc) {
  using v2 = typename Pair16<ST>::v2;
  if constexpr (ST == ST_F16)
    return __builtin_amdgcn_fdot2(__builtin_bit_cast(v2, a), __builtin_bit_cast(v2, b), acc, false);
  else
    return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(v2, a), __builtin_bit_cast(v2, b), acc, false);
}

template <int KIND, int AC>
constexpr size_t lds_bytes_px2() {
  return sizeof(uint32_t) * AC * KindGeo<KIND>::L + 2 * sizeof(int4) * AC + sizeof(int) * AC;
}

template <int KIND, int AC, int ST>
__device__ __forceinline__ void bp_tile_px2(const Item &it, int tile, unsigned char *lds) {
  using G = KindGeo<KIND>;
  constexpr int TR = G::TR, TC = G::TC, RT = G::RT, CT = G::CT, L = G::L;
  constexpr int RP = TR / 2;
  constexpr unsigned LO = One16<ST>::LO, HI = One16<ST>::HI;
  uint32_t *win = reinterpret_cast<uint32_t *>(lds);
  int4 *prm = reinterpret_cast<int4 *>(lds + sizeof(uint32_t) * AC * L);
  int *los = reinterpret_cast<int *>(lds + sizeof(uint32_t) * AC * L + 2 * sizeof(int4) * AC);

  const int w = it.w, N = it.N;
  const int hp = w / 2;
  const float c = it.center;
  const int ty0 = (tile / it.tiles_x) * RT;
  const int tx0 = (tile % it.tiles_x) * CT;
  const int tid = threadIdx.x;
  // wave v owns the tile rows [v 4TR, (v+1) 4TR); lane (ty4, tx) the pixels
  // (v 4TR + 2 ty4 + 8 rp + h, tx + 16 q), h = 0, 1; the coordinate of
  // pixel (y, x) of the tile is bt + S y + C x in 16.16 fixed point
  const int tx = tid & 15, ty = (tid >> 6) * (4 * TR) + 2 * ((tid >> 4) & 3);
  float *__restrict__ out = it.out;
  {
    const float acent = fabsf(c);
    const float ax = static_cast<float>(hp), ay = static_cast<float>(hp) + c;
    const float dx = fmaxf(fmaxf(static_cast<float>(tx0) - ax, ax - static_cast<float>(tx0 + CT - 1)), 0.0f);
    const float dy = fmaxf(fmaxf(static_cast<float>(ty0) - ay, ay - static_cast<float>(ty0 + RT - 1)), 0.0f);
    const float rmax = static_cast<float>(hp) + acent + 2.0f;
    const bool outside = (dx * dx + dy * dy > rmax * rmax) ||
                         (static_cast<float>(tx0 + CT - 1) <= acent) ||
                         (static_cast<float>(tx0) >= static_cast<float>(w) - acent - 1.0f);
    if (outside) {
#pragma unroll
      for (int rp = 0; rp < RP; ++rp)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int row = ty0 + ty + 8 * rp + h;
#pragma unroll
          for (int q = 0; q < TC; ++q) {
            const int col = tx0 + tx + 16 * q;
            if (row < w && col < w) out[static_cast<size_t>(row) * w + col] = 0.0f;
          }
        }
      return;
    }
  }
  float acc[RP][2][TC];
#pragma unroll
  for (int rp = 0; rp < RP; ++rp)
#pragma unroll
    for (int q = 0; q < TC; ++q) acc[rp][0][q] = acc[rp][1][q] = 0.0f;
  const uint16_t *__restrict__ filt = static_cast<const uint16_t *>(it.filt);
  const long long total = static_cast<long long>(N) * w;
  const bool wskip = wave_outside(it, c, ty0 + (tid >> 6) * (4 * TR), 4 * TR, tx0, CT);

  for (int a0 = 0; a0 < N; a0 += AC) {
    const int na = min(AC, N - a0);
    __syncthreads();
    if (tid < na) {
      const int ang = a0 + tid;
      const double2 sd = it.sincos64[ang];
      const double cd = static_cast<double>(c);
      const double T0 = cd + hp * ((1.0 - sd.x) - sd.y) + sd.x * (ty0 - cd) + sd.y * tx0;
      const double tc = T0 + sd.x * (0.5 * (RT - 1)) + sd.y * (0.5 * (CT - 1));
      const int lo = static_cast<int>(floor(tc)) - G::HALF - 2;
      los[tid] = lo;
      const bool neg = sd.x < 0.0;
      const int bt = static_cast<int>(llrint((T0 - lo + static_cast<double>(tid) * L) * 65536.0)) - (neg ? 65536 : 0);
      const int S = static_cast<int>(llrint(sd.x * 65536.0)), C = static_cast<int>(llrint(sd.y * 65536.0));
      const int thr = neg ? -S : 65536 - S;
      prm[2 * tid] = make_int4(bt, S, C, thr | (neg ? (1 << 30) : 0));
      prm[2 * tid + 1] = make_int4(8 * S, 16 * C, 0, 0);
    }
    __syncthreads();
    for (int e = tid; e < na * L; e += 256) {
      const int a = e / L;
      const int o = e - a * L;
      const long long flat = static_cast<long long>(a0 + a) * w + los[a] + o;
      const uint32_t v0 = (flat >= 0 && flat < total) ? filt[flat] : 0u;
      const uint32_t v1 = (flat + 1 >= 0 && flat + 1 < total) ? filt[flat + 1] : 0u;
      win[e] = v0 | (v1 << 16);
    }
    __syncthreads();
    if (wskip) continue;
    const unsigned char *wb = reinterpret_cast<const unsigned char *>(win);
    for (int a = 0; a < na; ++a) {
      const int4 p = prm[2 * a], pd = prm[2 * a + 1];
      const unsigned thr = static_cast<unsigned>(p.w & 0x3FFFFFFF);
      const unsigned wj = (p.w >> 30) ? HI : LO;
      int A[RP], C[TC];
      A[0] = __mul24(p.y, ty) + p.x;
#pragma unroll
      for (int rp = 1; rp < RP; ++rp) A[rp] = A[rp - 1] + pd.x;
      C[0] = __mul24(p.z, tx);
#pragma unroll
      for (int q = 1; q < TC; ++q) C[q] = C[q - 1] + pd.y;
#pragma unroll
      for (int rp = 0; rp < RP; ++rp) {
        unsigned u[TC], P[TC];
#pragma unroll
        for (int q = 0; q < TC; ++q) {
          u[q] = static_cast<unsigned>(A[rp] + C[q]);
          P[q] = *reinterpret_cast<const uint32_t *>(wb + ((u[q] >> 16) << 2));
        }
#pragma unroll
        for (int q = 0; q < TC; ++q) {
          const unsigned wj1 = (u[q] & 0xFFFFu) >= thr ? HI : LO;
          acc[rp][0][q] = dot2w<ST>(P[q], wj, acc[rp][0][q]);
          acc[rp][1][q] = dot2w<ST>(P[q], wj1, acc[rp][1][q]);
        }
      }
    }
  }

  const float acent = fabsf(c);
  const float up = static_cast<float>(hp) + c, dn = static_cast<float>(hp) - c;
  const float recCof = it.recCof;
#pragma unroll
  for (int rp = 0; rp < RP; ++rp)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int row = ty0 + ty + 8 * rp + h;
      if (row >= w) continue;
      float dy = static_cast<float>(row - hp) - c;
      dy *= dy;
#pragma unroll
      for (int q = 0; q < TC; ++q) {
        const int col = tx0 + tx + 16 * q;
        if (col >= w) continue;
        int ip2 = col - hp;
        ip2 *= ip2;
        const float d2 = dy + static_cast<float>(ip2);
        const float fr = static_cast<float>(row), fc = static_cast<float>(col);
        const bool masked = (fc <= acent) || (fc >= static_cast<float>(w) - acent - 1.0f) ||
                            (fr <= up && d2 >= up * up) || (fr >= up && d2 >= dn * dn);
        float v = acc[rp][h][q];
        if (recCof != 0.0f) v *= recCof;
        v = v / static_cast<float>(N);
        out[static_cast<size_t>(row) * w + col] = masked ? 0.0f : v;
      }
    }
}

// Grid-mapped launch: one workgroup per tile; blocks are remapped so that the
// round-robin dispatch over the 8 XCDs gives each XCD a contiguous range of
// tiles (neighbouring tiles share sinogram windows in that XCD's L2).
template <int KIND, int AC, int MODE, int ST>
constexpr bool use_fx() {
  return MODE == MODE_CTAS && ST != ST_F32;
}

template <int KIND, int AC, int MODE, int ST>
constexpr size_t grid_lds() {
  if constexpr (use_fx<KIND, AC, MODE, ST>()) return lds_bytes_px2<KIND, AC>();
  return lds_bytes<KIND, AC, MODE>();
}

// Waves per SIMD each variant's register budget is sized for.
template <int KIND, int AC, int MODE, int ST>
constexpr int grid_waves() {
  return KindGeo<KIND>::WAVES;
}

template <int KIND, int AC, int MODE, int ST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(grid_waves<KIND, AC, MODE, ST>(), grid_waves<KIND, AC, MODE, ST>()))) void bp_grid_kernel(const Item *__restrict__ items, int n_items) {
  __shared__ __align__(16) unsigned char lds[grid_lds<KIND, AC, MODE, ST>()];
  int b = blockIdx.x;
  const int q = gridDim.x >> 3;
  if (b < (q << 3)) b = (b & 7) * q + (b >> 3);
  const Item &it = items[find_item_by_tile(items, n_items, b)];
  if constexpr (use_fx<KIND, AC, MODE, ST>())
    bp_tile_px2<KIND, AC, ST>(it, b - it.tile_base, lds);
  else
    bp_tile<KIND, AC, MODE, ST>(it, b - it.tile_base, lds);
}

template <int A, int B>
struct Max { static constexpr size_t v = A > B ? A : B; };

constexpr int kPersistAC = 32;

template <int MODE>
constexpr size_t persist_lds() {
  return Max<Max<lds_bytes<TK_64x64, kPersistAC, MODE>(), lds_bytes<TK_128x64, kPersistAC, MODE>()>::v,
             lds_bytes<TK_128x128, kPersistAC, MODE>()>::v;
}

__global__ void reset_counter_kernel(unsigned *counter) { *counter = 0u; }

// Persistent launch: a fixed set of workgroups pulls tiles from a global
// counter until every tile of every item is done; each item carries its own
// tile kind, so one launch serves sinograms of uneven width.
template <int MODE, int ST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void bp_persistent_kernel(const Item *__restrict__ items, int n_items,
                                                            int total_tiles, unsigned *counter) {
  __shared__ __align__(16) unsigned char lds[persist_lds<MODE>()];
  __shared__ int s_job;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_job = static_cast<int>(atomicAdd(counter, 1u));
    __syncthreads();
    const int job = s_job;
    if (job >= total_tiles) break;
    const Item &it = items[find_item_by_tile(items, n_items, job)];
    const int tile = job - it.tile_base;
    switch (it.kind) {
      case TK_64x64: bp_tile<TK_64x64, kPersistAC, MODE, ST>(it, tile, lds); break;
      case TK_128x64: bp_tile<TK_128x64, kPersistAC, MODE, ST>(it, tile, lds); break;
      default: bp_tile<TK_128x128, kPersistAC, MODE, ST>(it, tile, lds); break;
    }
  }
}

template <int KIND, int AC, int MODE>
hipError_t launch_kas(const Item *d_items, int n_items, int total_tiles, int store, hipStream_t s) {
  switch (store) {
    case ST_F32: hipLaunchKernelGGL((bp_grid_kernel<KIND, AC, MODE, ST_F32>), dim3(total_tiles), dim3(256), 0, s, d_items, n_items); break;
    case ST_F16: hipLaunchKernelGGL((bp_grid_kernel<KIND, AC, MODE, ST_F16>), dim3(total_tiles), dim3(256), 0, s, d_items, n_items); break;
    default: hipLaunchKernelGGL((bp_grid_kernel<KIND, AC, MODE, ST_BF16>), dim3(total_tiles), dim3(256), 0, s, d_items, n_items); break;
  }
  return hipGetLastError();
}

template <int KIND, int MODE>
hipError_t launch_km(const Item *d_items, int n_items, int total_tiles, int ac, int store, hipStream_t s) {
  if (ac == 16) return launch_kas<KIND, 16, MODE>(d_items, n_items, total_tiles, store, s);
  return launch_kas<KIND, 32, MODE>(d_items, n_items, total_tiles, store, s);
}

template <int KIND>
hipError_t launch_k(const Item *d_items, int n_items, int total_tiles, int ac, int mode, int store, hipStream_t s) {
  if (mode == MODE_LINEAR) return launch_km<KIND, MODE_LINEAR>(d_items, n_items, total_tiles, ac, store, s);
  return launch_km<KIND, MODE_CTAS>(d_items, n_items, total_tiles, ac, store, s);
}

}  // namespace

hipError_t launch_backproject(const Item *d_items, int n_items, int total_tiles, int kind,
                              int angles_per_pass, int mode, int store, hipStream_t s) {
  if (total_tiles <= 0) return hipSuccess;
  switch (kind) {
    case TK_64x64: return launch_k<TK_64x64>(d_items, n_items, total_tiles, angles_per_pass, mode, store, s);
    case TK_128x64: return launch_k<TK_128x64>(d_items, n_items, total_tiles, angles_per_pass, mode, store, s);
    default: return launch_k<TK_128x128>(d_items, n_items, total_tiles, angles_per_pass, mode, store, s);
  }
}

hipError_t launch_backproject_persistent(const Item *d_items, int n_items, int total_tiles,
                                         int mode, int store, unsigned *d_counter,
                                         int n_workgroups, hipStream_t s) {
  if (total_tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, d_counter);
  const dim3 g(n_workgroups), b(256);
  if (mode == MODE_LINEAR) {
    switch (store) {
      case ST_F32: hipLaunchKernelGGL((bp_persistent_kernel<MODE_LINEAR, ST_F32>), g, b, 0, s, d_items, n_items, total_tiles, d_counter); break;
      case ST_F16: hipLaunchKernelGGL((bp_persistent_kernel<MODE_LINEAR, ST_F16>), g, b, 0, s, d_items, n_items, total_tiles, d_counter); break;
      default: hipLaunchKernelGGL((bp_persistent_kernel<MODE_LINEAR, ST_BF16>), g, b, 0, s, d_items, n_items, total_tiles, d_counter); break;
    }
  } else {
    switch (store) {
      case ST_F32: hipLaunchKernelGGL((bp_persistent_kernel<MODE_CTAS, ST_F32>), g, b, 0, s, d_items, n_items, total_tiles, d_counter); break;
      case ST_F16: hipLaunchKernelGGL((bp_persistent_kernel<MODE_CTAS, ST_F16>), g, b, 0, s, d_items, n_items, total_tiles, d_counter); break;
      default: hipLaunchKernelGGL((bp_persistent_kernel<MODE_CTAS, ST_BF16>), g, b, 0, s, d_items, n_items, total_tiles, d_counter); break;
    }
  }
  return hipGetLastError();
}

}  // namespace ctas
