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
#include <type_traits>

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

// floor(x) as an int in one instruction (v_cvt_flr_i32_f32; floorf and a
// conversion take two). x is a finite window coordinate.
__device__ __forceinline__ int cvt_floor(float x) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
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

// The item that owns tile b: one division for batches of equal items, a
// binary search otherwise.
__device__ __forceinline__ int find_item_by_tile(const Item *items, int n, int b) {
  if (n > 1) {
    const int per = items[1].tile_base;
    if (per > 0) {
      const int g = min(b / per, n - 1);
      if (items[g].tile_base <= b && (g == n - 1 || items[g + 1].tile_base > b)) return g;
    }
  }
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].tile_base <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Zeros of a tile that lies wholly outside the disc or the column band.
template <int KIND>
__device__ __forceinline__ bool tile_outside(const Item &it, int ty0, int tx0) {
  using G = KindGeo<KIND>;
  return wave_outside(it, it.center, ty0, G::RT, tx0, G::CT);
}

// Masks and scaling of the accumulated sums, in the reference's float
// arithmetic (lct.cpp:710-727); row(r) and col(q) give the pixel of
// accumulator [r][q].
template <int NR, int NQ, class RowF, class ColF>
__device__ __forceinline__ void store_tile(const Item &it, const float (&acc)[NR][NQ], RowF rowf, ColF colf) {
  const int w = it.w, hp = w / 2;
  const float c = it.center;
  const float acent = fabsf(c);
  const float up = static_cast<float>(hp) + c, dn = static_cast<float>(hp) - c;
  const float recCof = it.recCof;
  float *__restrict__ out = it.out;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int row = rowf(r);
    if (row >= w) continue;
    float dy = static_cast<float>(row - hp) - c;
    dy *= dy;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int col = colf(q);
      if (col < 0 || col >= w) continue;
      int ip2 = col - hp;
      ip2 *= ip2;
      const float d2 = dy + static_cast<float>(ip2);
      const float fr = static_cast<float>(row), fc = static_cast<float>(col);
      const bool masked = (fc <= acent) || (fc >= static_cast<float>(w) - acent - 1.0f) ||
                          (fr <= up && d2 >= up * up) || (fr >= up && d2 >= dn * dn);
      float v = acc[r][q];
      if (recCof != 0.0f) v *= recCof;
      v = v / static_cast<float>(it.N);
      out[static_cast<size_t>(row) * w + col] = masked ? 0.0f : v;
    }
  }
}

template <int NR, int NQ, class RowF, class ColF>
__device__ __forceinline__ void zero_tile(const Item &it, RowF rowf, ColF colf) {
  const int w = it.w;
#pragma unroll
  for (int r = 0; r < NR; ++r)
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int row = rowf(r), col = colf(q);
      if (row < w && col >= 0 && col < w) it.out[static_cast<size_t>(row) * w + col] = 0.0f;
    }
}

// ---- staging of the angle windows, software-pipelined ----------------------
// The angles are taken in passes of AC. While pass p is computed from one
// LDS buffer, the windows of pass p+1 are fetched from global memory into
// registers and written to the other buffer afterwards, so that the fetch
// latency hides behind the compute and each pass costs one barrier. Thread
// t stages angle t / TPA of the pass (it forms that angle's constants
// itself) and its window elements sub, sub + TPA, ... (sub = t % TPA).
//
// Window of an angle: L samples starting at lo = floor(ray coordinate of the
// tile centre) - HALF - 2, addressed as the reference does (flat offset
// theta*w + k of the slice; 0 outside the slice) in CTAS mode, and as
// (q0, q1 - q0) pairs of bin-centre samples (0 outside the row) in LINEAR
// mode.
// Loads through a global (address space 1) pointer. The item table holds
// generic pointers, from which the compiler would emit FLAT loads; those
// also count against the LDS wait counter, so every LDS wait of the compute
// loop would wait for the prefetch too.
template <class T>
__device__ __forceinline__ T ldg(const T *p, long long i) {
  return ((const __attribute__((address_space(1))) T *)(p))[i];
}
// 32-bit element offset from a uniform base (a slice holds fewer than 2^31
// samples): the load takes the base in scalar registers and the offset in
// one vector register.
template <class T>
__device__ __forceinline__ T ldg32(const T *p, unsigned i) {
  const unsigned byte = i * static_cast<unsigned>(sizeof(T));
  const __attribute__((address_space(1))) char *b = (const __attribute__((address_space(1))) char *)(p);
  return *(const __attribute__((address_space(1))) T *)(b + byte);
}
__device__ __forceinline__ float2 ldg(const float2 *p, long long i) {
  const float *q = reinterpret_cast<const float *>(p) + 2 * i;
  return make_float2(ldg(q, 0), ldg(q, 1));
}
__device__ __forceinline__ double2 ldg(const double2 *p, long long i) {
  const double *q = reinterpret_cast<const double *>(p) + 2 * i;
  return make_double2(ldg(q, 0), ldg(q, 1));
}

template <int ST> struct Raw { using type = uint16_t; };
template <> struct Raw<ST_F32> { using type = float; };

template <int ST>
__device__ __forceinline__ float raw_to_float(typename Raw<ST>::type r) {
  if constexpr (ST == ST_F32)
    return r;
  else if constexpr (ST == ST_F16)
    return static_cast<float>(__builtin_bit_cast(_Float16, r));
  else
    return __uint_as_float(static_cast<uint32_t>(r) << 16);
}

// The loads of a fetch are unconditional (indices clamped into the slice or
// row, the sample zeroed at store time when it was outside), so that no
// branch forces a wait on them before the compute of the current pass.
template <int KIND, int AC, int MODE, int ST>
struct StageFloat {
  using G = KindGeo<KIND>;
  using WT = typename WinT<MODE>::type;
  using R = typename Raw<ST>::type;
  static constexpr int L = G::L, TPA = 256 / AC, EPT = (L + TPA - 1) / TPA;
  static constexpr int NV = MODE == MODE_CTAS ? 1 : 2;  // raw samples per element
  // FP64 angle table except in the parity mode (CTAS sampling, float32).
  static constexpr bool F64 = !(MODE == MODE_CTAS && ST == ST_F32);
  using AT = typename std::conditional<F64, double2, float2>::type;
  static constexpr size_t WIN = sizeof(WT) * AC * L;
  static constexpr size_t HALFBUF = WIN + sizeof(float4) * AC;
  // two window/constant buffers, then two slots of AC table entries
  static constexpr size_t ANGS = 2 * HALFBUF;
  static constexpr size_t BYTES = ANGS + 2 * sizeof(AT) * AC;
  static_assert(NV * EPT <= 64, "one validity bit per raw sample");
  R v[EPT][NV];
  uint64_t ok;  // bit (NV j + i): raw sample i of element j lies inside
  float4 prm;
  AT ang, ang_next;  // table entries of the pass being fetched and of the one after

  // Branch-free: threads past the pass's last angle repeat that angle (their
  // results are not stored), so no branch makes the compiler wait on a load.
  __device__ __forceinline__ void load_angle(const Item &it, int a0, int na) {
    const int a = min(static_cast<int>(threadIdx.x) / TPA, na - 1);
    if constexpr (F64)
      ang_next = ldg(it.sincos64, a0 + a);
    else
      ang_next = ldg(it.sincos, a0 + a);
  }
  // The entry loaded with the previous fetch reaches the next one through
  // LDS, so that no load stays pending across a pass.
  __device__ __forceinline__ void put_angle(unsigned char *lds, int slot, int na) const {
    const int tid = threadIdx.x, a = tid / TPA;
    if (a < na && tid - a * TPA == 0) reinterpret_cast<AT *>(lds + ANGS)[slot * AC + a] = ang_next;
  }
  __device__ __forceinline__ void get_angle(const unsigned char *lds, int slot, int na) {
    ang = reinterpret_cast<const AT *>(lds + ANGS)[slot * AC + min(static_cast<int>(threadIdx.x) / TPA, na - 1)];
  }
  __device__ __forceinline__ void take_loaded() { ang = ang_next; }

  __device__ __forceinline__ void fetch(const Item &it, int a0, int na, float ycm, float xcm) {
    const int tid = threadIdx.x, a = min(tid / TPA, na - 1), sub = tid - (tid / TPA) * TPA;
    const int w = it.w, hp = w / 2;
    const float c = it.center;
    float s, co, base;
    if constexpr (F64) {
      // precision modes: angle constants and centre shift in FP64
      double bd = static_cast<double>(c) + static_cast<double>(hp) * ((1.0 - ang.x) - ang.y);
      if (MODE == MODE_LINEAR) bd += 0.5;
      s = static_cast<float>(ang.x);
      co = static_cast<float>(ang.y);
      base = static_cast<float>(bd);
    } else {
      s = ang.x;
      co = ang.y;
      base = c + ((1.0f - s) - co) * static_cast<float>(hp);
      if (MODE == MODE_LINEAR) base = base + 0.5f;
    }
    const float tc = base + s * ycm + co * xcm;
    const int lo = static_cast<int>(floorf(tc)) - G::HALF - 2;
    prm = make_float4(base, s, co, __int_as_float((a * L - lo) * static_cast<int>(sizeof(WT))));
    const R *__restrict__ filt = static_cast<const R *>(it.filt);
    ok = 0;
    if constexpr (MODE == MODE_CTAS) {
      const long long total = static_cast<long long>(it.N) * w;
      const long long rb = static_cast<long long>(a0 + a) * w + lo;
#pragma unroll
      for (int j = 0; j < EPT; ++j) {
        const long long flat = rb + sub + j * TPA;
        const bool in = flat >= 0 && flat < total;
        v[j][0] = ldg(filt, in ? flat : 0);
        ok |= static_cast<uint64_t>(in) << j;
      }
    } else {
      const size_t rb = static_cast<size_t>(a0 + a) * w;
#pragma unroll
      for (int j = 0; j < EPT; ++j) {
        const int k = lo + sub + j * TPA;
        const bool in1 = k >= 0 && k < w, in0 = k - 1 >= 0 && k - 1 < w;
        v[j][1] = ldg(filt, static_cast<long long>(rb) + (in1 ? k : 0));
        v[j][0] = ldg(filt, static_cast<long long>(rb) + (in0 ? k - 1 : 0));
        ok |= (static_cast<uint64_t>(in0) | (static_cast<uint64_t>(in1) << 1)) << (2 * j);
      }
    }
  }

  __device__ __forceinline__ void store(unsigned char *buf, int na) const {
    const int tid = threadIdx.x, a = tid / TPA, sub = tid - a * TPA;
    if (a >= na) return;
    WT *win = reinterpret_cast<WT *>(buf);
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int o = sub + j * TPA;
      if (o >= L) continue;
      if constexpr (MODE == MODE_CTAS) {
        win[a * L + o] = ((ok >> j) & 1u) ? raw_to_float<ST>(v[j][0]) : 0.0f;
      } else {
        const float q0 = ((ok >> (2 * j)) & 1u) ? raw_to_float<ST>(v[j][0]) : 0.0f;
        const float q1 = ((ok >> (2 * j + 1)) & 1u) ? raw_to_float<ST>(v[j][1]) : 0.0f;
        win[a * L + o] = make_float2(q0, q1 - q0);
      }
    }
    if (sub == 0) reinterpret_cast<float4 *>(buf + WIN)[a] = prm;
  }
};

// Drives the passes: fetch(p+1) / compute(p) / store(p+1) / barrier.
template <class Stage, class Compute>
__device__ __forceinline__ void run_passes(const Item &it, unsigned char *lds, int AC, float ycm, float xcm,
                                           bool wskip, Compute compute) {
  Stage st;
  const int N = it.N;
  const int P = (N + AC - 1) / AC;
  st.load_angle(it, 0, min(AC, N));
  st.take_loaded();
  st.fetch(it, 0, min(AC, N), ycm, xcm);
  st.load_angle(it, min(1, P - 1) * AC, min(AC, N - min(1, P - 1) * AC));
  st.store(lds, min(AC, N));
  st.put_angle(lds, 1, min(AC, N - min(1, P - 1) * AC));
  __syncthreads();
  for (int p = 0; p < P; ++p) {
    const int a0 = p * AC, na = min(AC, N - a0);
    // The fetches run in every iteration (past the end they repeat the last
    // pass into the buffer nobody reads), so that every load issued here is
    // consumed by this iteration's stores on every path and none stays
    // pending into the next iteration, where the compiler would wait on it.
    const int p1 = min(p + 1, P - 1), p2 = min(p + 2, P - 1);
    const int na1 = min(AC, N - p1 * AC), na2 = min(AC, N - p2 * AC);
    st.get_angle(lds, (p + 1) & 1, na1);
    st.fetch(it, p1 * AC, na1, ycm, xcm);
    st.load_angle(it, p2 * AC, na2);
    if (!wskip) compute(lds + (p & 1) * Stage::HALFBUF, na);
    st.store(lds + ((p + 1) & 1) * Stage::HALFBUF, na1);
    st.put_angle(lds, p & 1, na2);
    __syncthreads();
  }
}

// One output tile of one item, float LDS windows (FP32 storage, and LINEAR
// mode at any storage).
template <int KIND, int AC, int MODE, int ST>
__device__ __forceinline__ void bp_tile(const Item &it, int tile, unsigned char *lds) {
  using G = KindGeo<KIND>;
  using S = StageFloat<KIND, AC, MODE, ST>;
  using WT = typename S::WT;
  constexpr int TR = G::TR, TC = G::TC, RT = G::RT, CT = G::CT;
  const float c = it.center;
  const int ty0 = (tile / it.tiles_x) * RT;
  const int tx0 = (tile % it.tiles_x) * CT;
  const int tid = threadIdx.x;
  // wave v owns the tile rows [v 4TR, (v+1) 4TR); lane (ty4, tx) the pixels
  // (v 4TR + ty4 + 4r, tx + 16q): one LDS read instruction covers a 4 x 16
  // pixel block, and a wave whose rows lie outside the disc skips its work
  const int tx = tid & 15, ty = (tid >> 6) * (4 * TR) + ((tid >> 4) & 3);
  auto rowf = [&](int r) { return ty0 + ty + 4 * r; };
  auto colf = [&](int q) { return tx0 + tx + 16 * q; };
  if (tile_outside<KIND>(it, ty0, tx0)) {
    zero_tile<TR, TC>(it, rowf, colf);
    return;
  }
  float jc[TR], fi[TC];
#pragma unroll
  for (int r = 0; r < TR; ++r) jc[r] = static_cast<float>(rowf(r)) - c;
#pragma unroll
  for (int q = 0; q < TC; ++q) fi[q] = static_cast<float>(colf(q));
  float acc[TR][TC];
#pragma unroll
  for (int r = 0; r < TR; ++r)
#pragma unroll
    for (int q = 0; q < TC; ++q) acc[r][q] = 0.0f;
  const float ycm = static_cast<float>(ty0) + 0.5f * (RT - 1) - c;
  const float xcm = static_cast<float>(tx0) + 0.5f * (CT - 1);
  const bool wskip = wave_outside(it, c, ty0 + (tid >> 6) * (4 * TR), 4 * TR, tx0, CT);

  run_passes<S>(it, lds, AC, ycm, xcm, wskip, [&](const unsigned char *buf, int na) {
    const float4 *prm = reinterpret_cast<const float4 *>(buf + S::WIN);
#pragma unroll 2
    for (int a = 0; a < na; ++a) {
      const float4 p = prm[a];
      const unsigned char *wb = buf + __float_as_int(p.w);
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
          WT vv[TC];
          float ff[TC];
#pragma unroll
          for (int q = 0; q < TC; ++q) {
            // element floor(t) with weight t - floor(t): arcs past 180 degrees
            // put in-mask pixels at t < 0 when c < 0, where truncation is not
            // the floor that fract is taken from
            const float t = aj[r] + ci[q];
            const int k = cvt_floor(t);
            ff[q] = __builtin_amdgcn_fractf(t);
            vv[q] = *reinterpret_cast<const WT *>(wb + (k << 3));
          }
#pragma unroll
          for (int q = 0; q < TC; ++q) {
            acc[r][q] += vv[q].x;
            acc[r][q] = fmaf(ff[q], vv[q].y, acc[r][q]);
          }
        }
      }
    }
  });
  store_tile<TR, TC>(it, acc, rowf, colf);
}

// ---- FP16/BF16 storage, CTAS sampling, row pairs -----------------------------
// LDS holds, for every window position o, the pair (v[o], v[o+1]) as two
// floats (converted from the stored 16-bit samples when staged). A lane owns
// vertically adjacent pixel pairs (rows j, j+1): their ray coordinates
// differ by s = sin(theta), so with floor sampling row j+1 reads v[k] or
// v[k+1] (k = floor(t_j)) for s >= 0, v[k-1] or v[k] for s < 0 (the pair is
// then taken at k-1). One 8-byte LDS read serves both pixels; half as many
// LDS reads as the float path, and each sample costs an integer add, a
// compare and a select instead of float arithmetic and a conversion. The
// ray coordinate of pixel (y, x) of the tile is bt + S y + C x in 14.18
// fixed point (the word index of the sample in LDS, the window's offset
// included, and 18 fraction bits), bt, S and C rounded per angle from FP64
// constants: within (1 + y + x) 2^-19 px of the FP64 coordinate, 5e-4 px at
// the far corner of a 128 x 128 tile. (With 16.16 the rounded steps reached
// 2e-3 px there; with few angles, the samples those errors moved across a
// bin edge put slices 2x to 20x further from the FP64 oracle than the
// reference's float path, and 14.18 halves that distance.)
// The fixed-point paths index windows by floor(t); the reference reads
// sample (long)t, truncated (lct.cpp:710-713), which for a non-integer t < 0
// is floor(t) + 1. Window position k therefore holds the sample at
// trunc_pos(k). An integer t is its own truncation: at an angle whose sine
// and cosine are 0 or +-1 (multiples of 90 degrees) every coordinate of a
// tile shares the fraction of the tile's base, and when that fraction is 0
// the window holds sample k at position k (`exact`), so that a coordinate
// that is exactly a negative integer reads that sample. At other angles a
// fixed-point coordinate lands on an integer only by rounding, where the
// FP64 coordinate may lie on either side of it.
__device__ __forceinline__ int trunc_pos(int k) { return k < 0 ? k + 1 : k; }
__device__ __forceinline__ int window_pos(int k, bool exact) { return exact ? k : trunc_pos(k); }

template <int ST>
__device__ __forceinline__ float half_to_float(uint32_t h) {
  if constexpr (ST == ST_F16)
    return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(h)));
  else
    return __uint_as_float(h << 16);
}

// 14.18: fraction bits of the fixed-point coordinates of StagePair
constexpr int kFxFrac = 18;
constexpr unsigned kFxOne = 1u << kFxFrac;

template <int KIND, int AC, int ST>
struct StagePair {
  using G = KindGeo<KIND>;
  using WT = float2;
  static constexpr int L = G::L, TPA = 256 / AC, EPT = (L + TPA - 1) / TPA;
  static constexpr size_t WIN = sizeof(WT) * AC * L;
  static constexpr size_t HALFBUF = WIN + 2 * sizeof(int4) * AC;
  static constexpr size_t ANGS = 2 * HALFBUF;
  static constexpr size_t BYTES = ANGS + 2 * sizeof(double2) * AC;
  // every word index of the two windows fits the integer part
  static_assert(ANGS / sizeof(WT) <= (size_t{1} << (32 - kFxFrac)), "14.18 word index");
  uint16_t v[EPT][2];
  int4 prm0, prm1;
  double2 sd, sd_next;  // table entries of the pass being fetched and of the one after

  __device__ __forceinline__ void load_angle(const Item &it, int a0, int na) {
    sd_next = ldg(it.sincos64, a0 + min(static_cast<int>(threadIdx.x) / TPA, na - 1));
  }
  __device__ __forceinline__ void put_angle(unsigned char *lds, int slot, int na) const {
    const int tid = threadIdx.x, a = tid / TPA;
    if (a < na && tid - a * TPA == 0) reinterpret_cast<double2 *>(lds + ANGS)[slot * AC + a] = sd_next;
  }
  __device__ __forceinline__ void get_angle(const unsigned char *lds, int slot, int na) {
    sd = reinterpret_cast<const double2 *>(lds + ANGS)[slot * AC + min(static_cast<int>(threadIdx.x) / TPA, na - 1)];
  }
  __device__ __forceinline__ void take_loaded() { sd = sd_next; }

  __device__ __forceinline__ void fetch(const Item &it, int a0, int na, int ty0, int tx0) {
    const int tid = threadIdx.x, a = min(tid / TPA, na - 1), sub = tid - (tid / TPA) * TPA;
    const int w = it.w, hp = w / 2;
    const double cd = static_cast<double>(it.center);
    const double T0 = cd + hp * ((1.0 - sd.x) - sd.y) + sd.x * (ty0 - cd) + sd.y * tx0;
    const double tc = T0 + sd.x * (0.5 * (G::RT - 1)) + sd.y * (0.5 * (G::CT - 1));
    const int lo = static_cast<int>(floor(tc)) - G::HALF - 2;
    const bool neg = sd.x < 0.0;
    constexpr double one = static_cast<double>(kFxOne);
    const int b0 = static_cast<int>(llrint((T0 - lo + static_cast<double>(a) * L) * one));
    const int bt = b0 - (neg ? static_cast<int>(kFxOne) : 0);
    const int S = static_cast<int>(llrint(sd.x * one)), C = static_cast<int>(llrint(sd.y * one));
    // every coordinate of the tile an integer: sample k at position k
    const bool exact = ((b0 | S | C) & static_cast<int>(kFxOne - 1)) == 0;
    const int thr = neg ? -S : static_cast<int>(kFxOne) - S;
    prm0 = make_int4(bt, S, C, thr | (neg ? (1 << 30) : 0));
    prm1 = make_int4(8 * S, 16 * C, 0, 0);
    // reads outside the sinogram land in its zero pads
    const long long rb = static_cast<long long>(a0 + a) * w;
    const uint16_t *__restrict__ filt = static_cast<const uint16_t *>(it.filt);
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int k = lo + sub + j * TPA;
      v[j][0] = ldg(filt, rb + window_pos(k, exact));
      v[j][1] = ldg(filt, rb + window_pos(k + 1, exact));
    }
  }

  __device__ __forceinline__ void store(unsigned char *buf, int na) const {
    const int tid = threadIdx.x, a = tid / TPA, sub = tid - a * TPA;
    if (a >= na) return;
    WT *win = reinterpret_cast<WT *>(buf);
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int o = sub + j * TPA;
      if (o >= L) continue;
      win[a * L + o] = make_float2(half_to_float<ST>(v[j][0]), half_to_float<ST>(v[j][1]));
    }
    if (sub == 0) {
      // the buffer's offset from the start of LDS goes into the base
      // coordinate, so that a sample's LDS address is its word index times
      // the word size
      int4 *prm = reinterpret_cast<int4 *>(buf + WIN);
      int4 p0 = prm0;
      p0.x = static_cast<int>(static_cast<unsigned>(p0.x) +
                              (static_cast<unsigned>((buf - base) / sizeof(WT)) << kFxFrac));
      prm[2 * a] = p0;
      prm[2 * a + 1] = prm1;
    }
  }
  const unsigned char *base = nullptr;  // start of the LDS block
};

// Same pipeline for the fixed-point stages; N steps (angles or mirror slots).
template <class Stage, class Compute>
__device__ __forceinline__ void run_passes_fx(const Item &it, int N, unsigned char *lds, int AC, int ty0, int tx0,
                                              bool wskip, Compute compute) {
  Stage st;
  st.base = lds;
  const int P = (N + AC - 1) / AC;
  st.load_angle(it, 0, min(AC, N));
  st.take_loaded();
  st.fetch(it, 0, min(AC, N), ty0, tx0);
  st.load_angle(it, min(1, P - 1) * AC, min(AC, N - min(1, P - 1) * AC));
  st.store(lds, min(AC, N));
  st.put_angle(lds, 1, min(AC, N - min(1, P - 1) * AC));
  __syncthreads();
  for (int p = 0; p < P; ++p) {
    const int a0 = p * AC, na = min(AC, N - a0);
    const int p1 = min(p + 1, P - 1), p2 = min(p + 2, P - 1);
    const int na1 = min(AC, N - p1 * AC), na2 = min(AC, N - p2 * AC);
    st.get_angle(lds, (p + 1) & 1, na1);
    st.fetch(it, p1 * AC, na1, ty0, tx0);
    st.load_angle(it, p2 * AC, na2);
    if (!wskip) compute(lds + (p & 1) * Stage::HALFBUF, na);
    st.store(lds + ((p + 1) & 1) * Stage::HALFBUF, na1);
    st.put_angle(lds, p & 1, na2);
    __syncthreads();
  }
}

template <int KIND, int AC, int ST>
__device__ __forceinline__ void bp_tile_px2(const Item &it, int tile, unsigned char *lds) {
  using G = KindGeo<KIND>;
  using S = StagePair<KIND, AC, ST>;
  using W = typename S::WT;
  constexpr int TR = G::TR, TC = G::TC, RT = G::RT, CT = G::CT;
  constexpr int RP = TR / 2;
  const float c = it.center;
  const int ty0 = (tile / it.tiles_x) * RT;
  const int tx0 = (tile % it.tiles_x) * CT;
  const int tid = threadIdx.x;
  // wave v owns the tile rows [v 4TR, (v+1) 4TR); lane (ty4, tx) the pixels
  // (v 4TR + 2 ty4 + 8 rp + h, tx + 16 q), h = 0, 1
  const int tx = tid & 15, ty = (tid >> 6) * (4 * TR) + 2 * ((tid >> 4) & 3);
  auto rowf = [&](int r) { return ty0 + ty + 8 * (r >> 1) + (r & 1); };
  auto colf = [&](int q) { return tx0 + tx + 16 * q; };
  if (tile_outside<KIND>(it, ty0, tx0)) {
    zero_tile<TR, TC>(it, rowf, colf);
    return;
  }
  float acc[TR][TC];  // row 2 rp + h
#pragma unroll
  for (int r = 0; r < TR; ++r)
#pragma unroll
    for (int q = 0; q < TC; ++q) acc[r][q] = 0.0f;
  const bool wskip = wave_outside(it, c, ty0 + (tid >> 6) * (4 * TR), 4 * TR, tx0, CT);

  run_passes_fx<S>(it, it.N, lds, AC, ty0, tx0, wskip, [&](const unsigned char *buf, int na) {
    const int4 *prm = reinterpret_cast<const int4 *>(buf + S::WIN);
    const unsigned char *lds0 = lds;  // sample addresses are relative to it
    // one angle; HF (sin < 0, pair taken at k-1): row j takes the upper
    // sample. Row j+1 takes the upper sample when its fraction test holds.
    auto angle = [&](const int4 p, const int4 pd, auto hi_first) {
      constexpr bool HF = decltype(hi_first)::value;
      const unsigned thr = static_cast<unsigned>(p.w & 0x3FFFFFFF);
      // unsigned: the word index reaches 2^14 (bt near 2^32); the sums wrap
      // to the coordinate, which lies inside the window
      unsigned A[RP], C[TC];
      A[0] = static_cast<unsigned>(__mul24(p.y, ty)) + static_cast<unsigned>(p.x);
#pragma unroll
      for (int rp = 1; rp < RP; ++rp) A[rp] = A[rp - 1] + static_cast<unsigned>(pd.x);
      C[0] = static_cast<unsigned>(__mul24(p.z, tx));
#pragma unroll
      for (int q = 1; q < TC; ++q) C[q] = C[q - 1] + static_cast<unsigned>(pd.y);
#pragma unroll
      for (int rp = 0; rp < RP; ++rp) {
        unsigned u[TC];
        W P[TC];
#pragma unroll
        for (int q = 0; q < TC; ++q) {
          u[q] = A[rp] + C[q];
          P[q] = *reinterpret_cast<const W *>(lds0 + ((u[q] >> kFxFrac) * sizeof(W)));
        }
#pragma unroll
        for (int q = 0; q < TC; ++q) {
          const bool up = (u[q] & (kFxOne - 1)) >= thr;
          acc[2 * rp][q] += HF ? P[q].y : P[q].x;
          acc[2 * rp + 1][q] += up ? P[q].y : P[q].x;
        }
      }
    };
    if (it.neg_sin) {
      for (int a = 0; a < na; ++a) {
        const int4 p = prm[2 * a], pd = prm[2 * a + 1];
        if (p.w >> 30)
          angle(p, pd, std::true_type{});
        else
          angle(p, pd, std::false_type{});
      }
    } else {
#pragma unroll 2
      for (int a = 0; a < na; ++a) angle(prm[2 * a], prm[2 * a + 1], std::false_type{});
    }
  });
  store_tile<TR, TC>(it, acc, rowf, colf);
}

// ---- FP16/BF16 storage, CTAS sampling, mirrored angle pairs ----------------
// For angles a and b = pi - a (equal sines, opposite cosines), a pixel p and
// its mirror p' about the axis column hp = w/2 (column 2hp - x) have swapped
// ray coordinates: t_a(p) = t_b(p') and t_a(p') = t_b(p), exactly. So one
// coordinate and one 4-byte LDS read of the word (v_a[k], v_b[k]) give two
// updates: v_a[k] to p and v_b[k] to p'. A workgroup owns a tile of columns
// up to hp and the mirrored tile; per slot it stages two windows of such
// words, A around the tile's coordinates at angle a and B around the mirrored
// tile's, and every pixel pair reads one word of each:
//   p  += lo(A[t_a(p)])  + hi(B[t_a(p')])
//   p' += hi(A[t_a(p)])  + lo(B[t_a(p')])
// Angles without a partner (0 and 90 degrees of a 180-degree arc) are slots
// with zero high halves. The column hp is its own mirror: the first sum
// alone is its value.
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));

// acc + lo(x) + hi(x): both samples of a word, one v_dot2 with weights (1, 1)
template <int ST>
__device__ __forceinline__ float add_pair(float acc, uint32_t x) {
  if constexpr (ST == ST_F16) {
    const half2v one = {static_cast<_Float16>(1.0f), static_cast<_Float16>(1.0f)};
    return __builtin_amdgcn_fdot2(__builtin_bit_cast(half2v, x), one, acc, false);
  } else {
    uint32_t wt = 0x3F803F80u;  // bf16 (1, 1), held in a scalar register
    asm("" : "+s"(wt));
    return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2v, x), __builtin_bit_cast(bf16x2v, wt), acc,
                                           false);
  }
}

// Coordinates, 16.16 fixed point: per slot the staging threads form the
// row terms RA[y] = round(TA - loA) + SY[y] (window A, word offsets and
// buffer included), RB[y] likewise for window B, and the column terms
// CX[x], from the slot's precomputed SY[y] = round(s y 2^16) and
// CX[x] = round(c x 2^16). Pixel (y, x) then reads word RA[y] + CX[x] of
// window A and its mirror word RB[y] - CX[x] of window B, within 2^-15 px
// of the FP64 coordinate. The tables are laid out in lattice order, so that
// a lane reads its TR row terms and TC column terms as 16-byte words. Per 4
// updates (pixel and mirror, two angles): two coordinates, two 4-byte LDS
// reads, two bit-field merges and two dot2 accumulations.
// Integer constants of one slot for one tile: window starts, the 16.16
// bases of the row terms (word offsets within the slot included) and the
// sinogram offsets of the two rows (rb < 0: no partner).
// exact: bit 0 (1) every coordinate of window A is an integer, bit 1 (2)
// every coordinate of window B.
struct SlotConsts {
  int loA, loB, ba, bb, ra, rb, exact, unused;
};

template <int KIND, int AC, int ST>
struct StageMirror {
  using G = KindGeo<KIND>;
  // window of D + 6 positions from floor(centre) - ceil(D/2) - 1: the tile's
  // coordinates lie within D/2 of the centre's, so one position of margin
  // below and three above (for 64x64 tiles 2L = 192 words, six per thread)
  static constexpr int L = G::D + 6, LO = (G::D + 1) / 2 + 1;
  static constexpr int L2 = 2 * L, TPA = 256 / AC, EPT = (L2 + TPA - 1) / TPA;
  static constexpr int RT = G::RT, CT = G::CT, TR = G::TR, TC = G::TC;
  static constexpr int TAB = 2 * RT + CT;  // ints per slot: RA, RB, CX
  static constexpr int TPT = (RT + CT + TPA - 1) / TPA;  // SY and CX entries per thread
  // LINEAR (any storage, samples as float): word k of a window holds
  // (v_a[k-1], (v_a[k] - v_a[k-1]) 2^-16, v_b[k-1], (v_b[k] - v_b[k-1])
  // 2^-16), zero outside the rows; the CTAS words hold the two 16-bit
  // samples (v_a[k], v_b[k])
  static constexpr bool LIN = ST == ST_F32_LIN || ST == ST_F16_LIN || ST == ST_BF16_LIN;
  using WordT = typename std::conditional<LIN, float4, uint32_t>::type;
  static constexpr size_t WIN = sizeof(WordT) * AC * L2;
  static constexpr size_t HALFBUF = WIN + sizeof(int) * AC * TAB;
  static constexpr size_t ANGS = 2 * HALFBUF;
  static constexpr size_t BYTES = ANGS + 2 * sizeof(SlotConsts) * AC;
  static_assert(TR % 4 == 0 && TC % 4 == 0, "16-byte table reads");
  static_assert(AC <= 64, "one lane of the deriving wave per slot");
  typename std::conditional<LIN, float4, u16x2>::type v[EPT];  // the samples of word j
  int t[TPT];    // SY (rows) and CX (columns) entries of this thread
  SlotConsts k;  // constants of this thread's slot, pass being fetched
  MirrorSlot raw;  // deriving wave: the slot of its lane, pass after next
  const unsigned char *base = nullptr;  // start of the LDS block

  // Deriving wave, lane l: slot a0 + min(l, na - 1).
  __device__ __forceinline__ void load_raw(const Item &it, int a0, int na) {
    const MirrorSlot *p = it.slots + a0 + min(static_cast<int>(threadIdx.x & 63), na - 1);
    const double *d = reinterpret_cast<const double *>(p);
    const int *r = reinterpret_cast<const int *>(p) + 4;
    raw.s = ldg(d, 0);
    raw.c = ldg(d, 1);
    raw.row_a = ldg(r, 0);
    raw.row_b = ldg(r, 1);
    raw.integral = ldg(r, 2);
  }
  // Deriving wave: the constants of its lanes' slots for the tile at
  // (ty0, tx0) (the mirrored tile's columns are 2hp - x), formed in FP64,
  // into LDS slot area `par`. The other waves do no FP64 work.
  __device__ __forceinline__ void derive(const Item &it, unsigned char *lds, int par, int ty0, int tx0) const {
    const int l = threadIdx.x & 63;
    const int w = it.w, hp = w / 2;
    const double cd = static_cast<double>(it.center), s = raw.s, co = raw.c;
    const double rowterm = cd + hp * ((1.0 - s) - co) + s * (ty0 - cd);
    // LINEAR interpolates between bin centres: word floor(t + 1/2)
    const double half = LIN ? 0.5 : 0.0;
    const double TA = rowterm + co * tx0 + half;             // pixel (ty0, tx0)
    const double TB = rowterm + co * (2 * hp - tx0) + half;  // its mirror
    const double hy = s * (0.5 * (RT - 1)), hx = co * (0.5 * (CT - 1));
    SlotConsts c;
    c.loA = static_cast<int>(floor(TA + hy + hx)) - LO;
    c.loB = static_cast<int>(floor(TB + hy - hx)) - LO;
    c.ba = static_cast<int>(rint((TA - c.loA) * 65536.0));
    c.bb = static_cast<int>(rint((TB - c.loB) * 65536.0));
    c.ra = it.pad + raw.row_a * w;
    c.rb = raw.row_b >= 0 ? it.pad + raw.row_b * w : -1;
    // the row and column terms are whole pixels: the base's fraction is
    // that of every coordinate of the window
    c.exact = (!LIN && raw.integral) ? ((c.ba & 0xFFFF) == 0 ? 1 : 0) | ((c.bb & 0xFFFF) == 0 ? 2 : 0) : 0;
    c.unused = 0;
    if (l < AC) reinterpret_cast<SlotConsts *>(lds + ANGS)[par * AC + l] = c;
  }
  __device__ __forceinline__ void get(const unsigned char *lds, int par, int na) {
    k = reinterpret_cast<const SlotConsts *>(lds + ANGS)[par * AC + min(static_cast<int>(threadIdx.x) / TPA, na - 1)];
  }

  __device__ __forceinline__ void fetch(const Item &it, int a0, int na) {
    const int tid = threadIdx.x, sub = tid - (tid / TPA) * TPA;
    // the slot's row terms (lattice order of this tile kind), then its
    // column terms (lattice order)
    const int *st = it.slot_tab + (a0 + min(tid / TPA, na - 1)) * kSlotTab + lattice_offset(KIND);
#pragma unroll
    for (int j = 0; j < TPT; ++j) t[j] = ldg32(st, static_cast<unsigned>(min(sub + j * TPA, RT + CT - 1)));
    // offsets from the start of the zero pad before the sinogram; a slot
    // without a partner reads its second samples from the pad's first element
    if constexpr (LIN) {
      // samples k-1 and k of both rows, indices clamped into the row and the
      // samples outside it zeroed (LINEAR reads nothing beyond the detector);
      // a slot without a partner reads row A again for its zeroed second
      // samples. Row offsets count from the allocation's start (the 16-bit
      // sinograms' zero pad before them included).
      const int w = it.w;
      const bool hb = k.rb >= 0;
      const int rb = hb ? k.rb : k.ra;
      auto sample = [&](int i) -> float {
        if constexpr (ST == ST_F32_LIN) {
          return ldg(static_cast<const float *>(it.filt), i);
        } else {
          const uint16_t *__restrict__ b16 = static_cast<const uint16_t *>(it.filt) - it.pad;
          return half_to_float<ST == ST_F16_LIN ? ST_F16 : ST_BF16>(ldg(b16, i));
        }
      };
#pragma unroll
      for (int j = 0; j < EPT; ++j) {
        const int e = sub + j * TPA;
        const int kk = e < L ? k.loA + e : k.loB + (e - L);
        const bool in1 = kk >= 0 && kk < w, in0 = kk >= 1 && kk <= w;
        const int i1 = in1 ? kk : 0, i0 = in0 ? kk - 1 : 0;
        const float a0 = sample(k.ra + i0), a1 = sample(k.ra + i1);
        const float b0 = sample(rb + i0), b1 = sample(rb + i1);
        v[j] = make_float4(in0 ? a0 : 0.0f, in1 ? a1 : 0.0f, (hb && in0) ? b0 : 0.0f, (hb && in1) ? b1 : 0.0f);
      }
    } else {
      const uint16_t *__restrict__ base16 = static_cast<const uint16_t *>(it.filt) - it.pad;
      const int mb = ~(k.rb >> 31);
#pragma unroll
      for (int j = 0; j < EPT; ++j) {
        const int e = sub + j * TPA;  // word e of the slot: [0, L) window A, [L, 2L) window B
        const int kk = e < L ? window_pos(k.loA + e, k.exact & 1) : window_pos(k.loB + (e - L), k.exact & 2);
        v[j].x = ldg32(base16, static_cast<unsigned>(k.ra + kk));
        v[j].y = ldg32(base16, static_cast<unsigned>((k.rb + kk) & mb));
      }
    }
  }

  __device__ __forceinline__ void store(unsigned char *buf, int na) const {
    const int tid = threadIdx.x, a = tid / TPA, sub = tid - a * TPA;
    if (a >= na) return;
    WordT *win = reinterpret_cast<WordT *>(buf);
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      const int e = sub + j * TPA;
      if (e >= L2) continue;
      if constexpr (LIN) {
        // the differences scaled by 2^-16 (exact), so that the 16-bit
        // fraction of a coordinate weighs them as an integer
        const float4 q = v[j];
        win[a * L2 + e] = make_float4(q.x, (q.y - q.x) * 0x1p-16f, q.z, (q.w - q.z) * 0x1p-16f);
      } else if constexpr (ST == ST_BF16_H) {
        // BF16 -> FP32 (exact) -> FP16, rounded to nearest even
        const half2v h = {static_cast<_Float16>(__uint_as_float(static_cast<uint32_t>(v[j].x) << 16)),
                          static_cast<_Float16>(__uint_as_float(static_cast<uint32_t>(v[j].y) << 16))};
        win[a * L2 + e] = __builtin_bit_cast(uint32_t, h);
      } else {
        win[a * L2 + e] = __builtin_bit_cast(uint32_t, v[j]);
      }
    }
    // word addresses are taken from the start of the LDS block
    const int off = static_cast<int>((buf - base) / sizeof(WordT)) + a * L2;
    const int ba = k.ba + (off << 16), bb = k.bb + ((off + L) << 16);
    int *tab = reinterpret_cast<int *>(buf + WIN) + a * TAB;
#pragma unroll
    for (int j = 0; j < TPT; ++j) {
      const int e = sub + j * TPA;
      if (e < RT) {
        tab[e] = ba + t[j];
        tab[RT + e] = bb + t[j];
      } else if (e < RT + CT) {
        tab[RT + e] = t[j];
      }
    }
  }
};

// Pipeline of the mirrored passes: fetch(p+1) / compute(p) / store(p+1) /
// barrier, as run_passes_fx; one wave (in turn) also loads the slots of
// pass p+2 and derives their constants, which reach every thread through LDS.
template <class Stage, class Compute>
__device__ __forceinline__ void run_passes_mirror(const Item &it, unsigned char *lds, int AC, int ty0, int tx0,
                                                  bool wskip, Compute compute) {
  Stage st;
  st.base = lds;
  const bool first = threadIdx.x < 64;
  const int N = it.n_slots;
  const int P = (N + AC - 1) / AC;
  auto count = [&](int p) { return min(AC, N - p * AC); };
  if (first) {
    st.load_raw(it, 0, count(0));
    st.derive(it, lds, 0, ty0, tx0);
    const int p1 = min(1, P - 1);
    st.load_raw(it, p1 * AC, count(p1));
    st.derive(it, lds, 1, ty0, tx0);
  }
  __syncthreads();
  st.get(lds, 0, count(0));
  st.fetch(it, 0, count(0));
  st.store(lds, count(0));
  __syncthreads();
  for (int p = 0; p < P; ++p) {
    const int na = count(p);
    const int p1 = min(p + 1, P - 1), p2 = min(p + 2, P - 1);
    const int na1 = count(p1);
    // the four waves take turns at deriving, so that no wave waits longest
    const bool deriver = (static_cast<int>(threadIdx.x) >> 6) == (p & 3);
    st.get(lds, (p + 1) & 1, na1);
    st.fetch(it, p1 * AC, na1);
    if (deriver) st.load_raw(it, p2 * AC, count(p2));
    if (!wskip) compute(lds + (p & 1) * Stage::HALFBUF, na);
    st.store(lds + ((p + 1) & 1) * Stage::HALFBUF, na1);
    if (deriver) st.derive(it, lds, p & 1, ty0, tx0);
    __syncthreads();
  }
}

template <int KIND, int AC, int ST>
__device__ __forceinline__ void bp_tile_mirror(const Item &it, int tile, unsigned char *lds) {
  using G = KindGeo<KIND>;
  using S = StageMirror<KIND, AC, ST>;
  constexpr int TR = G::TR, TC = G::TC, RT = G::RT, CT = G::CT;
  const float c = it.center;
  const int w = it.w, hp = w / 2;
  const int ty0 = (tile / it.tiles_x) * RT;
  // tile column 0 ends at column hp; tiles to the left may start below 0
  const int tx0 = hp + 1 - CT * (tile % it.tiles_x + 1);
  const int mx0 = 2 * hp - (tx0 + CT - 1);  // first column of the mirrored tile
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = (tid >> 6) * (4 * TR) + ((tid >> 4) & 3);
  auto rowf = [&](int r) { return ty0 + ty + 4 * r; };
  auto colf = [&](int q) { return tx0 + tx + 16 * q; };
  // column hp is stored once, from the first sum
  auto colm = [&](int q) {
    const int col = tx0 + tx + 16 * q;
    return col == hp ? w : 2 * hp - col;
  };
  if (tile_outside<KIND>(it, ty0, tx0) && tile_outside<KIND>(it, ty0, mx0)) {
    zero_tile<TR, TC>(it, rowf, colf);
    zero_tile<TR, TC>(it, rowf, colm);
    return;
  }
  float acc[TR][TC], accm[TR][TC];
#pragma unroll
  for (int r = 0; r < TR; ++r)
#pragma unroll
    for (int q = 0; q < TC; ++q) acc[r][q] = accm[r][q] = 0.0f;
  const int wr0 = ty0 + (tid >> 6) * (4 * TR);
  const bool wskip = wave_outside(it, c, wr0, 4 * TR, tx0, CT) && wave_outside(it, c, wr0, 4 * TR, mx0, CT);
  // this lane's entries in the tables
  const int rbase = ((tid >> 6) * 4 + ((tid >> 4) & 3)) * TR, cbase = 2 * RT + tx * TC;

  run_passes_mirror<S>(it, lds, AC, ty0, tx0, wskip, [&](const unsigned char *buf, int na) {
    const unsigned char *lds0 = lds;  // word addresses are relative to it
    // slot a's row terms (window A, window B) and column terms of this lane
    auto terms = [&](int a, int (&A)[TR], int (&B)[TR], int (&Cq)[TC]) {
      const int *tab = reinterpret_cast<const int *>(buf + S::WIN) + a * S::TAB;
#pragma unroll
      for (int r = 0; r < TR; r += 4) {
        const int4 va = *reinterpret_cast<const int4 *>(tab + rbase + r);
        const int4 vb = *reinterpret_cast<const int4 *>(tab + RT + rbase + r);
        A[r] = va.x; A[r + 1] = va.y; A[r + 2] = va.z; A[r + 3] = va.w;
        B[r] = vb.x; B[r + 1] = vb.y; B[r + 2] = vb.z; B[r + 3] = vb.w;
      }
#pragma unroll
      for (int q = 0; q < TC; q += 4) {
        const int4 vc = *reinterpret_cast<const int4 *>(tab + cbase + q);
        Cq[q] = vc.x; Cq[q + 1] = vc.y; Cq[q + 2] = vc.z; Cq[q + 3] = vc.w;
      }
    };
    // the words of both windows for the pixel pairs of row r
    auto words = [&](const int (&A)[TR], const int (&B)[TR], const int (&Cq)[TC], int r, uint32_t (&WA)[TC],
                     uint32_t (&WB)[TC]) {
#pragma unroll
      for (int q = 0; q < TC; ++q) {
        const unsigned ua = static_cast<unsigned>(A[r] + Cq[q]);
        const unsigned ub = static_cast<unsigned>(B[r] - Cq[q]);
        WA[q] = *reinterpret_cast<const uint32_t *>(lds0 + ((ua >> 16) << 2));
        WB[q] = *reinterpret_cast<const uint32_t *>(lds0 + ((ub >> 16) << 2));
      }
    };
    if constexpr (S::LIN) {
      // LINEAR: the word of window A at t_a(p) and of window B at t_a(p')
      // carry both angles' samples and differences; with their 16-bit
      // fractions fa, fb: p += a(t_a(p)) + b(t_a(p')), p' += a(t_a(p')) +
      // b(t_a(p)), two coordinates and two 16-byte reads for four updates
#pragma unroll 2
      for (int a = 0; a < na; ++a) {
        int A[TR], B[TR], Cq[TC];
        terms(a, A, B, Cq);
#pragma unroll
        for (int r = 0; r < TR; ++r) {
          float4 WA[TC], WB[TC];
          float FA[TC], FB[TC];
#pragma unroll
          for (int q = 0; q < TC; ++q) {
            const unsigned ua = static_cast<unsigned>(A[r] + Cq[q]);
            const unsigned ub = static_cast<unsigned>(B[r] - Cq[q]);
            WA[q] = *reinterpret_cast<const float4 *>(lds0 + ((ua >> 16) << 4));
            WB[q] = *reinterpret_cast<const float4 *>(lds0 + ((ub >> 16) << 4));
            FA[q] = static_cast<float>(ua & 0xFFFFu);
            FB[q] = static_cast<float>(ub & 0xFFFFu);
          }
#pragma unroll
          for (int q = 0; q < TC; ++q) {
            acc[r][q] = fmaf(FA[q], WA[q].y, acc[r][q] + WA[q].x);
            acc[r][q] = fmaf(FB[q], WB[q].w, acc[r][q] + WB[q].z);
            accm[r][q] = fmaf(FB[q], WB[q].y, accm[r][q] + WB[q].x);
            accm[r][q] = fmaf(FA[q], WA[q].w, accm[r][q] + WA[q].z);
          }
        }
      }
    } else if constexpr (ST == ST_F16 || ST == ST_BF16_H) {
      // FP16: one packed add forms both pixels' pair sums of a slot, (lo WA
      // + hi WB, hi WA + lo WB) (the swap is the add's operand select); the
      // sums of two slots are added packed, each step rounded once to FP16,
      // and a dot2 that picks its half adds each pixel's four samples to its
      // FP32 accumulator: five instructions for eight updates
      const half2v lo = {static_cast<_Float16>(1.0f), static_cast<_Float16>(0.0f)};
      const half2v hi = {static_cast<_Float16>(0.0f), static_cast<_Float16>(1.0f)};
      auto pair_sum = [](uint32_t wa, uint32_t wb) {
        const half2v pa = __builtin_bit_cast(half2v, wa), pb = __builtin_bit_cast(half2v, wb);
        return pa + __builtin_shufflevector(pb, pb, 1, 0);
      };
      int a = 0;
      for (; a + 1 < na; a += 2) {
        int A0[TR], B0[TR], C0[TC], A1[TR], B1[TR], C1[TC];
        terms(a, A0, B0, C0);
        terms(a + 1, A1, B1, C1);
#pragma unroll
        for (int r = 0; r < TR; ++r) {
          uint32_t WA0[TC], WB0[TC], WA1[TC], WB1[TC];
          words(A0, B0, C0, r, WA0, WB0);
          words(A1, B1, C1, r, WA1, WB1);
#pragma unroll
          for (int q = 0; q < TC; ++q) {
            const half2v sum = pair_sum(WA0[q], WB0[q]) + pair_sum(WA1[q], WB1[q]);
            acc[r][q] = __builtin_amdgcn_fdot2(sum, lo, acc[r][q], false);
            accm[r][q] = __builtin_amdgcn_fdot2(sum, hi, accm[r][q], false);
          }
        }
      }
      if (a < na) {  // an odd slot count: the last slot alone
        int A[TR], B[TR], Cq[TC];
        terms(a, A, B, Cq);
#pragma unroll
        for (int r = 0; r < TR; ++r) {
          uint32_t WA[TC], WB[TC];
          words(A, B, Cq, r, WA, WB);
#pragma unroll
          for (int q = 0; q < TC; ++q) {
            const half2v sum = pair_sum(WA[q], WB[q]);
            acc[r][q] = __builtin_amdgcn_fdot2(sum, lo, acc[r][q], false);
            accm[r][q] = __builtin_amdgcn_fdot2(sum, hi, accm[r][q], false);
          }
        }
      }
    } else {
#pragma unroll 2
      for (int a = 0; a < na; ++a) {
        int A[TR], B[TR], Cq[TC];
        terms(a, A, B, Cq);
#pragma unroll
        for (int r = 0; r < TR; ++r) {
          uint32_t WA[TC], WB[TC];
          words(A, B, Cq, r, WA, WB);
#pragma unroll
          for (int q = 0; q < TC; ++q) {
            // the pixel's two samples (lo WA, hi WB) and the mirror pixel's
            // (lo WB, hi WA), merged by bit-field inserts (the asm keeps them
            // 32-bit merges) and each added by one dot2
            uint32_t X = (WA[q] & 0xFFFFu) | (WB[q] & 0xFFFF0000u);
            uint32_t Y = (WB[q] & 0xFFFFu) | (WA[q] & 0xFFFF0000u);
            asm("" : "+v"(X), "+v"(Y));
            acc[r][q] = add_pair<ST>(acc[r][q], X);
            accm[r][q] = add_pair<ST>(accm[r][q], Y);
          }
        }
      }
    }
  });
  store_tile<TR, TC>(it, acc, rowf, colf);
  store_tile<TR, TC>(it, accm, rowf, colm);
}

// Grid-mapped launch: one workgroup per tile; blocks are remapped so that the
// round-robin dispatch over the 8 XCDs gives each XCD a contiguous range of
// tiles (neighbouring tiles share sinogram windows in that XCD's L2).
template <int KIND, int AC, int MODE, int ST>
constexpr bool use_fx() {
  return MODE == MODE_CTAS && ST != ST_F32;
}

template <int KIND, int AC, int MODE, int ST>
constexpr size_t stage_bytes() {
  if constexpr (use_fx<KIND, AC, MODE, ST>())
    return StagePair<KIND, AC, ST>::BYTES;
  else
    return StageFloat<KIND, AC, MODE, ST>::BYTES;
}

// Angles per pass actually used: the double-buffered windows of a variant
// must fit the 64 KiB a kernel may allocate statically; the 8-byte windows
// of the larger tiles take 16 angles per pass instead of 32.
constexpr size_t kMaxStaticLds = 65536;
template <int KIND, int AC, int MODE, int ST>
constexpr int eff_ac() {
  return stage_bytes<KIND, AC, MODE, ST>() <= kMaxStaticLds ? AC : AC / 2;
}

// Waves per SIMD each variant's register budget is sized for.
template <int KIND, int AC, int MODE, int ST>
constexpr int grid_waves() {
  return KindGeo<KIND>::WAVES;
}

template <int KIND, int AC, int MODE, int ST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(grid_waves<KIND, AC, MODE, ST>(), grid_waves<KIND, AC, MODE, ST>()))) void bp_grid_kernel(const Item *__restrict__ items, int n_items) {
  constexpr int A = eff_ac<KIND, AC, MODE, ST>();
  __shared__ __align__(16) unsigned char lds[stage_bytes<KIND, A, MODE, ST>()];
  int b = blockIdx.x;
  const int q = gridDim.x >> 3;
  if (b < (q << 3)) b = (b & 7) * q + (b >> 3);
  const Item &it = items[find_item_by_tile(items, n_items, b)];
  if constexpr (use_fx<KIND, A, MODE, ST>())
    bp_tile_px2<KIND, A, ST>(it, b - it.tile_base, lds);
  else
    bp_tile<KIND, A, MODE, ST>(it, b - it.tile_base, lds);
}

// Mirrored launch: AC slots (2 AC angles) per pass.
template <int KIND, int AC, int ST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KindGeo<KIND>::WAVES, KindGeo<KIND>::WAVES))) void bp_mirror_kernel(const Item *__restrict__ items, int n_items) {
  __shared__ __align__(16) unsigned char lds[StageMirror<KIND, AC, ST>::BYTES];
  int b = blockIdx.x;
  const int q = gridDim.x >> 3;
  if (b < (q << 3)) b = (b & 7) * q + (b >> 3);
  const Item &it = items[find_item_by_tile(items, n_items, b)];
  bp_tile_mirror<KIND, AC, ST>(it, b - it.tile_base, lds);
}

__global__ void reset_counter_kernel(unsigned *counter) { *counter = 0u; }

// Persistent launches: a fixed set of workgroups, as many as stay resident,
// pull tiles from a global counter until every tile of every item is done.
// One launch serves items of uneven width that share a tile kind and pass
// length; each variant keeps the register budget of its grid kernel. The
// items come costliest first, so the launch ends with the small tiles.
template <class Tile>
__device__ __forceinline__ void persistent_loop(const Item *__restrict__ items, int n_items, int total_tiles,
                                                unsigned *counter, Tile tile) {
  __shared__ int s_job;
  for (;;) {
    __syncthreads();  // the previous tile is done with the LDS and s_job
    if (threadIdx.x == 0) s_job = static_cast<int>(atomicAdd(counter, 1u));
    __syncthreads();
    const int job = s_job;
    if (job >= total_tiles) break;
    const Item &it = items[find_item_by_tile(items, n_items, job)];
    tile(it, job - it.tile_base);
  }
}

template <int KIND, int AC, int MODE, int ST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(grid_waves<KIND, AC, MODE, ST>(), grid_waves<KIND, AC, MODE, ST>()))) void bp_persistent_kernel(const Item *__restrict__ items, int n_items,
                                                                 int total_tiles, unsigned *counter) {
  constexpr int A = eff_ac<KIND, AC, MODE, ST>();
  __shared__ __align__(16) unsigned char lds[stage_bytes<KIND, A, MODE, ST>()];
  persistent_loop(items, n_items, total_tiles, counter, [&](const Item &it, int tile) {
    if constexpr (use_fx<KIND, A, MODE, ST>())
      bp_tile_px2<KIND, A, ST>(it, tile, lds);
    else
      bp_tile<KIND, A, MODE, ST>(it, tile, lds);
  });
}

template <int KIND, int AC, int ST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KindGeo<KIND>::WAVES, KindGeo<KIND>::WAVES))) void bp_persistent_mirror_kernel(const Item *__restrict__ items, int n_items,
                                                                        int total_tiles, unsigned *counter) {
  __shared__ __align__(16) unsigned char lds[StageMirror<KIND, AC, ST>::BYTES];
  persistent_loop(items, n_items, total_tiles, counter,
                  [&](const Item &it, int tile) { bp_tile_mirror<KIND, AC, ST>(it, tile, lds); });
}

// Workgroups of a persistent launch: as many as the CUs hold at once, by the
// variant's waves per SIMD and its LDS.
int resident_workgroups(int cus, int waves, size_t lds_bytes) {
  const int by_lds = static_cast<int>((160 * 1024) / (lds_bytes + 64));
  return cus * std::max(1, std::min(waves, by_lds));
}

template <int KIND, int AC, int MODE, int ST>
hipError_t launch_persist(const Item *d_items, int n_items, int total_tiles, unsigned *counter, int cus, hipStream_t s) {
  constexpr int A = eff_ac<KIND, AC, MODE, ST>();
  const int g = resident_workgroups(cus, grid_waves<KIND, AC, MODE, ST>(), stage_bytes<KIND, A, MODE, ST>());
  hipLaunchKernelGGL((bp_persistent_kernel<KIND, AC, MODE, ST>), dim3(g), dim3(256), 0, s, d_items, n_items,
                     total_tiles, counter);
  return hipGetLastError();
}

template <int KIND, int AC, int MODE>
hipError_t launch_persist_st(const Item *d_items, int n_items, int total_tiles, unsigned *counter, int cus, int store,
                             hipStream_t s) {
  switch (store) {
    case ST_F32: return launch_persist<KIND, AC, MODE, ST_F32>(d_items, n_items, total_tiles, counter, cus, s);
    case ST_F16: return launch_persist<KIND, AC, MODE, ST_F16>(d_items, n_items, total_tiles, counter, cus, s);
    default: return launch_persist<KIND, AC, MODE, ST_BF16>(d_items, n_items, total_tiles, counter, cus, s);
  }
}

template <int KIND>
hipError_t launch_persist_k(const Item *d_items, int n_items, int total_tiles, int ac, int mode, int store, bool mirror,
                            unsigned *counter, int cus, hipStream_t s) {
  if (mirror && mode == MODE_LINEAR) {
    // 64x64 tiles only (the 16-byte float words of both windows at 4 slots)
    if constexpr (KIND == TK_64x64) {
      const int g = resident_workgroups(cus, KindGeo<KIND>::WAVES, StageMirror<KIND, 4, ST_F32_LIN>::BYTES);
      switch (store) {
        case ST_F32: hipLaunchKernelGGL((bp_persistent_mirror_kernel<KIND, 4, ST_F32_LIN>), dim3(g), dim3(256), 0, s, d_items, n_items, total_tiles, counter); break;
        case ST_F16: hipLaunchKernelGGL((bp_persistent_mirror_kernel<KIND, 4, ST_F16_LIN>), dim3(g), dim3(256), 0, s, d_items, n_items, total_tiles, counter); break;
        default: hipLaunchKernelGGL((bp_persistent_mirror_kernel<KIND, 4, ST_BF16_LIN>), dim3(g), dim3(256), 0, s, d_items, n_items, total_tiles, counter); break;
      }
      return hipGetLastError();
    } else {
      return hipErrorInvalidValue;
    }
  }
  if (mirror) {
    const int g = resident_workgroups(cus, KindGeo<KIND>::WAVES, StageMirror<KIND, 8, ST_F16>::BYTES);
    if (store == ST_F16)
      hipLaunchKernelGGL((bp_persistent_mirror_kernel<KIND, 8, ST_F16>), dim3(g), dim3(256), 0, s, d_items, n_items,
                         total_tiles, counter);
    else if (store == ST_BF16_H)
      hipLaunchKernelGGL((bp_persistent_mirror_kernel<KIND, 8, ST_BF16_H>), dim3(g), dim3(256), 0, s, d_items,
                         n_items, total_tiles, counter);
    else
      hipLaunchKernelGGL((bp_persistent_mirror_kernel<KIND, 8, ST_BF16>), dim3(g), dim3(256), 0, s, d_items, n_items,
                         total_tiles, counter);
    return hipGetLastError();
  }
  if (mode == MODE_LINEAR) {
    if (ac == 16) return launch_persist_st<KIND, 16, MODE_LINEAR>(d_items, n_items, total_tiles, counter, cus, store, s);
    return launch_persist_st<KIND, 32, MODE_LINEAR>(d_items, n_items, total_tiles, counter, cus, store, s);
  }
  if (ac == 16) return launch_persist_st<KIND, 16, MODE_CTAS>(d_items, n_items, total_tiles, counter, cus, store, s);
  return launch_persist_st<KIND, 32, MODE_CTAS>(d_items, n_items, total_tiles, counter, cus, store, s);
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

template <int KIND, int AC>
hipError_t launch_mirror(const Item *d_items, int n_items, int total_tiles, int store, hipStream_t s) {
  if (store == ST_F16)
    hipLaunchKernelGGL((bp_mirror_kernel<KIND, AC, ST_F16>), dim3(total_tiles), dim3(256), 0, s, d_items, n_items);
  else if (store == ST_BF16_H)
    hipLaunchKernelGGL((bp_mirror_kernel<KIND, AC, ST_BF16_H>), dim3(total_tiles), dim3(256), 0, s, d_items, n_items);
  else
    hipLaunchKernelGGL((bp_mirror_kernel<KIND, AC, ST_BF16>), dim3(total_tiles), dim3(256), 0, s, d_items, n_items);
  return hipGetLastError();
}

template <int KIND>
hipError_t launch_k(const Item *d_items, int n_items, int total_tiles, int ac, int mode, int store, bool mirror,
                    hipStream_t s) {
  // 8 slots (16 angles) per pass: the staging registers of 16 slots spill
  if (mirror) return launch_mirror<KIND, 8>(d_items, n_items, total_tiles, store, s);
  if (mode == MODE_LINEAR) return launch_km<KIND, MODE_LINEAR>(d_items, n_items, total_tiles, ac, store, s);
  return launch_km<KIND, MODE_CTAS>(d_items, n_items, total_tiles, ac, store, s);
}

}  // namespace

hipError_t launch_backproject(const Item *d_items, int n_items, int total_tiles, int kind,
                              int angles_per_pass, int mode, int store, bool mirror, hipStream_t s) {
  if (total_tiles <= 0) return hipSuccess;
  if (mirror && mode == MODE_LINEAR) {
    // 64x64 tiles: the 16-byte float words of both windows (16-bit samples
    // converted when staged) fill the static LDS of a workgroup at 4 slots
    if (kind != TK_64x64) return hipErrorInvalidValue;
    switch (store) {
      case ST_F32: hipLaunchKernelGGL((bp_mirror_kernel<TK_64x64, 4, ST_F32_LIN>), dim3(total_tiles), dim3(256), 0, s, d_items, n_items); break;
      case ST_F16: hipLaunchKernelGGL((bp_mirror_kernel<TK_64x64, 4, ST_F16_LIN>), dim3(total_tiles), dim3(256), 0, s, d_items, n_items); break;
      default: hipLaunchKernelGGL((bp_mirror_kernel<TK_64x64, 4, ST_BF16_LIN>), dim3(total_tiles), dim3(256), 0, s, d_items, n_items); break;
    }
    return hipGetLastError();
  }
  if (mirror && (mode != MODE_CTAS || store == ST_F32)) return hipErrorInvalidValue;
  switch (kind) {
    case TK_64x64: return launch_k<TK_64x64>(d_items, n_items, total_tiles, angles_per_pass, mode, store, mirror, s);
    case TK_128x64: return launch_k<TK_128x64>(d_items, n_items, total_tiles, angles_per_pass, mode, store, mirror, s);
    default: return launch_k<TK_128x128>(d_items, n_items, total_tiles, angles_per_pass, mode, store, mirror, s);
  }
}

hipError_t launch_backproject_persistent(const Item *d_items, int n_items, int total_tiles, int kind,
                                         int angles_per_pass, int mode, int store, bool mirror,
                                         unsigned *d_counter, int cus, hipStream_t s) {
  if (total_tiles <= 0) return hipSuccess;
  // mirrored: CTAS sampling of 16-bit storage, or LINEAR on 64x64 tiles
  if (mirror && (mode == MODE_CTAS ? store == ST_F32 : kind != TK_64x64)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, d_counter);
  switch (kind) {
    case TK_64x64:
      return launch_persist_k<TK_64x64>(d_items, n_items, total_tiles, angles_per_pass, mode, store, mirror, d_counter,
                                        cus, s);
    case TK_128x64:
      return launch_persist_k<TK_128x64>(d_items, n_items, total_tiles, angles_per_pass, mode, store, mirror,
                                         d_counter, cus, s);
    default:
      return launch_persist_k<TK_128x128>(d_items, n_items, total_tiles, angles_per_pass, mode, store, mirror,
                                          d_counter, cus, s);
  }
}

}  // namespace ctas
