// Internal declarations shared by the translation units of libctas_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "../../include/ctas_hip.h"

namespace ctas {

// ---- errors ---------------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
// RingFilter::apply (lct.cpp:317-320) divides temp(ii) and temp(w-1-ii) for
// every ii < box: a box of w or more indexes outside the line there, so such
// a box is rejected (CTAS_E_ARG); narrower boxes keep the reference's sums.
inline int check_ring_box(int box, int w) {
  if (box >= w)
    return fail(CTAS_E_ARG, "ring filter box " + std::to_string(box) + " is not below the sinogram width " +
                                std::to_string(w));
  return 0;
}

#define CTAS_HIP_TRY(expr)                                                              \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return ::ctas::fail(CTAS_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---- host geometry (window.cpp) -------------------------------------------
int zidth(int w);
float arc_rad(float anarc, int points);
float step_rad(float anarc, int points);
void angle_table(int n, float anarc, float *sincos);
void angle_table64(int n, float anarc, double *sincos);
// FP64 per-angle constants are used unless the mode is the parity mode
// (CTAS sampling with float32 storage).
inline bool use_fp64_angles(int flags) {
  return (flags & CTAS_INTERP_MASK) != CTAS_INTERP_CTAS || (flags & CTAS_STORE_MASK) != CTAS_STORE_FP32;
}
int filter_window(int filter, float param, int w, int contrast, std::vector<float> &W);
// Filter::fill (lct.cpp:122-209): the window of an n-point line, no gain.
int filter_fill(int filter, float param, int n, std::vector<float> &W);
// The window of ts_add's row filter as a transform of Z points (Z = n for a
// power of two n, else zidth(n)), inverse scaled by 1/Z.
int ts_window(int filter, float param, int n, int &Z, std::vector<float> &W);

// improveMe (lct.cpp:1108-1129): bisection from the bracket around vax[idx]
// down to eps; f evaluates the score of one trial offset and sets *rc on
// failure (axis.hip).
float improve_me(float eps, const std::vector<float> &offsets, const std::vector<float> &scores, int at, int dir,
                 const std::function<float(float)> &score_at, int *rc);

// ---- device descriptors -----------------------------------------------------
// One step of the mirrored backprojection (FP16/BF16 storage): angle row_a
// and angle row_b with theta_b = pi - theta_a (equal sines, opposite
// cosines), or row_b = -1 when row_a has no such partner. A pixel and its
// mirror about the axis column w/2 then have swapped ray coordinates:
// t_a(p) = t_b(p') and t_a(p') = t_b(p).
struct MirrorSlot {
  double s, c;  // FP64 sin and cos of angle row_a
  int row_a, row_b;
  int integral;  // every row and column term of every tile kind is a whole
                 // multiple of 2^16 (sin and cos in {0, +-1}): a tile's
                 // coordinates are then all integers or all non-integers
  int pad;
};
// Per slot and tile kind, the row terms round(s y 2^16) (y < rows of the
// tile) and then the column terms round(c x 2^16) (x < columns), each in
// the lattice order of the kind's lanes (lattice_row / lattice_col): the
// backprojection copies them into LDS without index arithmetic. The three
// kinds' sections lie back to back, kSlotTab ints per slot.
constexpr int kSlotTab = (64 + 64) + (128 + 64) + (128 + 128);
// Offset of a tile kind's section (TK_64x64 = 1, TK_128x64 = 2, TK_128x128 = 3).
__host__ __device__ constexpr int lattice_offset(int kind) { return kind == 1 ? 0 : kind == 2 ? 128 : 320; }
// Lane lattice of a tile with TR x TC pixels per lane (4x16-pixel blocks per
// wave instruction): position of row y and of column x in the tables.
__host__ __device__ constexpr int lattice_row(int y, int TR) {
  return ((y / (4 * TR)) * 4 + (y & 3)) * TR + (y % (4 * TR)) / 4;
}
__host__ __device__ constexpr int lattice_col(int x, int TC) { return (x & 15) * TC + x / 16; }
// Pairs first (ordered by row_a), then the angles without a partner.
// Returns the number of pairs.
int mirror_slots(const std::vector<double> &sc64, std::vector<MirrorSlot> &slots);

// One sinogram -> one slice. The backprojector walks tiles of every item;
// the filter walks row pairs of every item that needs filtering.
struct Item {
  const float *sino;      // raw sinogram [N][w] (filter input)
  void *filt;             // filtered sinogram [N][w], storage dtype
  float *filt_f32;        // optional float32 copy of the filtered rows (host API)
  float *out;             // slice [w][w]
  const float2 *sincos;   // [N] {sin, cos}: the reference's float table
  const double2 *sincos64; // [N] FP64 {sin, cos} of i*arc/N; when set, the
                           // per-angle constants are formed in FP64 and
                           // rounded once (LINEAR and FP16/BF16 modes)
  const float *W;         // [Z] window
  const float2 *tw;       // [Z/2] twiddles exp(-2 pi i m / Z)
  float2 *minmax;         // [kMinMaxStride]: kMinMaxBlocks partial (min, max) of the raw sinogram, then deAbs flags
  int *status;            // deAbs status (0 / CTAS_E_DATA)
  float center;
  float recCof;
  int w, N, Z, logZ;
  int tiles_x, tiles;     // backprojection tiles of this item
  int tile_base;          // prefix sum of tiles over the items
  int pair_base;          // prefix sum of row pairs over the items
  int kind;               // tile kind used for this item (persistent kernel)
  int neg_sin;             // some angle has sin < 0 (arcs past 180 degrees)
  int n_slots;             // mirrored launch: number of slots
  const MirrorSlot *slots; // mirrored launch: [n_slots]
  const int *slot_tab;     // mirrored launch: [n_slots][kSlotTab]
  int pad;                 // zero samples before and after filt (16-bit storage), 0 otherwise
  float half_max;          // FP16 storage: largest filtered magnitude kept finite (kHalfMax / kHalfSumMax)
};

// Zero samples kept before and after each filtered sinogram of 16-bit
// storage: every window the backprojection stages (tile coordinates within
// sqrt(2) (w/2 + 32 + |c|) of the axis, |c| < w/2, plus the window length)
// then reads inside the allocation, and its reads outside the sinogram
// return 0 without a bounds check.
inline int filt_pad(int w) { return (3 * w) / 2 + 256; }

// FP16 storage: 65504 is the largest half; the mirrored kernel adds four
// samples (two angle pairs) in half precision, so its samples must stay within
// 65504 / 4. A finite filtered sample above the limit sets CTAS_E_RANGE.
constexpr float kHalfMax = 65504.0f, kHalfSumMax = 16376.0f;

constexpr int kMinMaxBlocks = 64;
constexpr int kMinMaxStride = kMinMaxBlocks + 1;  // float2 entries per item

enum Store { ST_F32 = 0, ST_F16 = 1, ST_BF16 = 2,
             // the mirrored backprojection of BF16 storage staging its samples
             // as FP16 and summing them as the FP16 path does (see bp_store)
             ST_BF16_H = 3,
             // LINEAR sampling through mirrored angle pairs of FP32, FP16 and
             // BF16 storage (the mirrored backprojection launch only)
             ST_F32_LIN = 4, ST_F16_LIN = 5, ST_BF16_LIN = 6 };

// Storage code of the backprojection launch. BF16 samples convert to FP16
// exactly above 2^-14 in magnitude (8 significant bits into 11) and below
// it round by at most 2^-25; the mirrored kernel adds four in FP16, which
// stays finite below 16376. ABS data is far inside that: -log of a positive
// float lies in [-88.8, 103.3] and no window's spatial kernel sums to more
// than 6.5 in magnitude (PARZEN's quirk, lct.cpp:170-172), so every
// filtered sample is below 700. Other contrasts keep the BF16 path.
inline int bp_store(int st, int contrast, bool mirror) {
  return (mirror && st == ST_BF16 && contrast == CTAS_CONTRAST_ABS) ? ST_BF16_H : st;
}
enum Mode { MODE_CTAS = 0, MODE_LINEAR = 1 };

inline int store_of(int flags) {
  const int s = flags & CTAS_STORE_MASK;
  return s == CTAS_STORE_FP16 ? ST_F16 : s == CTAS_STORE_BF16 ? ST_BF16 : ST_F32;
}
inline int mode_of(int flags) { return (flags & CTAS_INTERP_MASK) == CTAS_INTERP_LINEAR ? MODE_LINEAR : MODE_CTAS; }
inline size_t store_bytes(int st) { return st == ST_F32 ? 4 : 2; }

// Tile kinds: rows x cols of output per 256-thread workgroup.
enum TileKind { TK_AUTO = 0, TK_64x64 = 1, TK_128x64 = 2, TK_128x128 = 3 };
inline int tile_rows(int kind) { return kind == TK_64x64 ? 64 : 128; }
inline int tile_cols(int kind) { return kind == TK_128x128 ? 128 : 64; }

// ---- launchers (filter.hip, backproject.hip) ------------------------------
// Zeros the pads around the filtered sinograms of the items (Item::pad).
hipError_t launch_pad_zero(const Item *d_items, int n_items, hipStream_t s);
// deAbs + padded FFT filtering + REF prefix sum for all row pairs; with ABS
// contrast it also sets every item's status (0 / CTAS_E_DATA).
// min_logZ == max_logZ == 12 (every line 4096 points) takes the radix-16 kernel.
hipError_t launch_filter(const Item *d_items, int n_items, int total_pairs, int min_logZ, int max_logZ,
                         int contrast, int store, hipStream_t s);
// Backprojection of all tiles of all items (grid-mapped).
// mirror: FP16/BF16 storage with CTAS sampling, items carry mirror slots and
// their tiles cover the columns up to w/2 together with the mirrored ones.
hipError_t launch_backproject(const Item *d_items, int n_items, int total_tiles, int kind,
                              int angles_per_pass, int mode, int store, bool mirror, hipStream_t s);
// Same, one persistent launch with a work counter over the tiles of items of
// one tile kind and pass length (widths and angle counts may differ); mirror:
// FP16/BF16 items whose angles pair up (Item::slots). cus: compute units.
hipError_t launch_backproject_persistent(const Item *d_items, int n_items, int total_tiles, int kind,
                                         int angles_per_pass, int mode, int store, bool mirror,
                                         unsigned *d_counter, int cus, hipStream_t s);
// Ring filter (RingFilter::apply) on S sinograms in place.
hipError_t launch_ring(float *d_sino, float *d_avg, int S, int N, int w, int box, hipStream_t s);

// Tomosynthesis (ts.hip): per projection of a launch {cos a, P, R, 0} of
// ts_add's column map, passed by value; projections beyond kTsBatch are
// split over launches.
constexpr int kTsBatch = 64;
constexpr int kTsMaxWidth = 8192;
struct TsConsts {
  float4 k[kTsBatch];
};
// odd (null for even widths): (u[x], v[x]) of the odd-width correction of
// the row filter, odd_g its gain (ts.hip).
hipError_t launch_ts(const float *proj, long long proj_stride, int n_proj, const TsConsts &consts,
                     float *filtered, float *result, const float *W, const float2 *tw, int rows, int n, int Z,
                     int logZ, int contrast, int cus, int lds_max, const float2 *odd, float odd_g, hipStream_t s);
// For odd n >= 3: the basis (u[x], v[x]) = (-1)^x (cos(pi x/n), sin(pi x/n)),
// x < n, and the gain Z (W[m] - W[m+1])/n (W = Filter::fill(n), m = (n-1)/2)
// of the odd-width correction of ts_add's row filter; false for other n.
bool ts_odd_basis(int filter, float param, int n, int Z, std::vector<float2> &uv, float &gain);
// Axis-score kernels (axis.hip).
// AverageGradient::apply scores of n slices [n][w][w] reconstructed at the
// given centres (n <= kScoreChunk); d_tmp holds n*w*w floats, d_partial
// kScoreBlocks*kScoreChunk*2 doubles.
hipError_t launch_ag_scores(float *d_slices, int n, int w, const float *d_centres, float sigma,
                            float *d_tmp, double *d_partial, float *d_scores, hipStream_t s);
constexpr int kScoreBlocks = 128;
constexpr int kScoreChunk = 64;  // centres per scoring batch (bounds the scratch slices)
// Phantom projector (phantom.hip).
hipError_t launch_phantom(int N, int w, double c, int n_ell, const double *d_ell, double N0,
                          uint64_t seed, const float *d_gain, float *d_out, hipStream_t s);

// Makes work queued next on `waiter` start after the work already queued on
// `producer` (nullptr: the null stream, whose work is ordered with every
// blocking stream, torch's default stream included).
int stream_wait(hipStream_t waiter, hipStream_t producer);

// A _dev call with no caller stream runs on the library's own stream as if
// it ran on the null stream: after the work already queued there, and
// before any queued there later.
struct NullStreamJoin {
  hipStream_t own;
  bool active;
  int rc;
  NullStreamJoin(hipStream_t own_stream, void *caller_stream)
      : own(own_stream), active(caller_stream == nullptr), rc(0) {
    if (active) rc = stream_wait(own, nullptr);
  }
  ~NullStreamJoin() {
    if (active) (void)stream_wait(nullptr, own);
  }
};

}  // namespace ctas

// Handle (CTrec state). Defined here so that the C ABI units share it.
struct ctas_fbp {
  int device = 0;
  int w = 0, N = 0, Z = 0, logZ = 0;
  int contrast = 0, filter = 0, flags = 0;
  float fparam = 0.f, anarc = 180.f;
  float recCof = 1.f;
  std::vector<float> W;        // window [Z]
  std::vector<float> sincos;   // angle table [2N]
  float *d_W = nullptr;
  float2 *d_sincos = nullptr;
  double2 *d_sincos64 = nullptr;
  int neg_sin = 0;             // some entry of the FP64 table has sin < 0
  ctas::MirrorSlot *d_slots = nullptr;  // [N] mirror slots of the FP64 table
  int *d_slot_tab = nullptr;             // [N][kSlotTab] their row and column terms
  int n_slots = 0, n_paired = 0;
  float2 *d_tw = nullptr;
  hipStream_t stream = nullptr;

  // workspaces (grown on demand)
  void *d_filt = nullptr;  size_t filt_cap = 0;      // bytes
  float *d_io = nullptr;   size_t io_cap = 0;        // bytes, host-API staging
  float *d_out = nullptr;  size_t out_cap = 0;       // bytes
  float2 *d_minmax = nullptr; int *d_status = nullptr; int mm_cap = 0;  // slices
  ctas::Item *d_items = nullptr; int items_cap = 0;
  float *d_centres = nullptr; int centres_cap = 0;
  double *d_partial = nullptr; float *d_tmp = nullptr; size_t score_cap = 0;
  float *d_scores = nullptr;
  int kept_slices = 0;         // filtered slices held for repeat()
  int last_S = 0;

  int tile_kind = ctas::TK_AUTO;
  int angles_per_pass = 0;

  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
};
