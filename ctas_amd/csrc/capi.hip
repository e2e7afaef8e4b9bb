// The C ABI of libctas_hip.so: CTrec-equivalent handles (lct.cpp:487-685),
// device-resident and host-API variants, hipGraph capture, uneven batches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "ctas_internal.hpp"

namespace ctas {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string &msg) { g_last_error = msg; }
int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}


namespace {

std::string fmt_float(float v) {
  char b[64];
  std::snprintf(b, sizeof b, "%g", static_cast<double>(v));
  return b;
}

int check_center(int w, float c) {
  // lct.cpp:614-616 / 651-653: abs(center) >= ish(1)/2 (integer half) throws.
  if (!(std::fabs(c) < static_cast<float>(w / 2)))
    return fail(CTAS_E_CENTER, "Rotation center " + fmt_float(c) +
                                   " is outside the sinogram width " + std::to_string(w) + ".");
  return CTAS_OK;
}

template <typename T>
int grow(T *&ptr, size_t &cap, size_t bytes) {
  if (bytes <= cap && ptr) return CTAS_OK;
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  cap = 0;
  void *p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess)
    return fail(CTAS_E_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed");
  ptr = static_cast<T *>(p);
  cap = bytes;
  return CTAS_OK;
}

int ensure_slices(ctas_fbp *h, int S) {
  if (S <= h->mm_cap && h->d_minmax) return CTAS_OK;
  if (h->d_minmax) (void)hipFree(h->d_minmax);
  if (h->d_status) (void)hipFree(h->d_status);
  h->d_minmax = nullptr;
  h->d_status = nullptr;
  h->mm_cap = 0;
  if (hipMalloc(&h->d_minmax, sizeof(float2) * kMinMaxStride * std::max(S, 1)) != hipSuccess ||
      hipMalloc(&h->d_status, sizeof(int) * std::max(S, 1)) != hipSuccess)
    return fail(CTAS_E_NOMEM, "device allocation of per-slice status failed");
  h->mm_cap = S;
  return CTAS_OK;
}

int choose_kind(const ctas_fbp *h);
int choose_ac(const ctas_fbp *h);

// Upload the item table unless it is identical to the resident one. The
// stream is drained first, since a running kernel may still read the table.
int upload_items(ctas_fbp *h, const std::vector<Item> &items, hipStream_t s, bool *changed,
                 std::vector<Item> &resident) {
  *changed = false;
  if (resident.size() == items.size() && h->d_items &&
      std::memcmp(resident.data(), items.data(), sizeof(Item) * items.size()) == 0)
    return CTAS_OK;
  *changed = true;
  // the table may still be read by work queued on the handle's stream or on s
  CTAS_HIP_TRY(hipStreamSynchronize(s));
  CTAS_HIP_TRY(hipStreamSynchronize(h->stream));
  if (static_cast<int>(items.size()) > h->items_cap) {
    if (h->d_items) (void)hipFree(h->d_items);
    h->d_items = nullptr;
    h->items_cap = 0;
    if (hipMalloc(&h->d_items, sizeof(Item) * items.size()) != hipSuccess)
      return fail(CTAS_E_NOMEM, "device allocation of the item table failed");
    h->items_cap = static_cast<int>(items.size());
  }
  CTAS_HIP_TRY(hipMemcpy(h->d_items, items.data(), sizeof(Item) * items.size(), hipMemcpyHostToDevice));
  resident = items;
  return CTAS_OK;
}

struct HandleExtra {
  std::vector<Item> resident;
};
std::mutex g_extra_mu;
std::map<const ctas_fbp *, std::unique_ptr<HandleExtra>> g_extra;

HandleExtra &extra(const ctas_fbp *h) {
  std::lock_guard<std::mutex> lk(g_extra_mu);
  auto &p = g_extra[h];
  if (!p) p.reset(new HandleExtra);
  return *p;
}
void drop_extra(const ctas_fbp *h) {
  std::lock_guard<std::mutex> lk(g_extra_mu);
  g_extra.erase(h);
}

// Mirror slots of an FP64 table of N angles into d_slots [N] and their row
// and column terms into d_slot_tab [N][kSlotTab] (allocated when null).
int build_slots(const std::vector<double> &sc64, int N, MirrorSlot *&d_slots, int *&d_slot_tab, int *n_slots,
                int *n_paired) {
  std::vector<MirrorSlot> slots;
  *n_paired = mirror_slots(sc64, slots);
  *n_slots = static_cast<int>(slots.size());
  if (!d_slots && hipMalloc(&d_slots, sizeof(MirrorSlot) * std::max(N, 1)) != hipSuccess)
    return fail(CTAS_E_NOMEM, "device allocation of the mirror slots failed");
  if (!d_slot_tab && hipMalloc(&d_slot_tab, sizeof(int) * kSlotTab * std::max(N, 1)) != hipSuccess)
    return fail(CTAS_E_NOMEM, "device allocation of the mirror slot tables failed");
  std::vector<int> tab(static_cast<size_t>(kSlotTab) * slots.size());
  for (size_t i = 0; i < slots.size(); ++i) {
    int frac = 0;  // OR of the fractional bits of every term
    for (int kind = TK_64x64; kind <= TK_128x128; ++kind) {
      const int RT = tile_rows(kind), CT = tile_cols(kind), TR = RT / 16, TC = CT / 16;
      int *t = tab.data() + i * kSlotTab + lattice_offset(kind);
      for (int y = 0; y < RT; ++y) {
        const int v = static_cast<int>(std::lrint(slots[i].s * y * 65536.0));
        t[lattice_row(y, TR)] = v;
        frac |= v & 0xFFFF;
      }
      for (int x = 0; x < CT; ++x) {
        const int v = static_cast<int>(std::lrint(slots[i].c * x * 65536.0));
        t[RT + lattice_col(x, TC)] = v;
        frac |= v & 0xFFFF;
      }
    }
    slots[i].integral = frac == 0;
  }
  CTAS_HIP_TRY(hipMemcpy(d_slots, slots.data(), sizeof(MirrorSlot) * slots.size(), hipMemcpyHostToDevice));
  CTAS_HIP_TRY(hipMemcpy(d_slot_tab, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice));
  return CTAS_OK;
}

// Mirror slots of the handle's FP64 table (one slot per angle at most).
int upload_slots(ctas_fbp *h, const std::vector<double> &sc64) {
  return build_slots(sc64, h->N, h->d_slots, h->d_slot_tab, &h->n_slots, &h->n_paired);
}

// The mirrored backprojection serves FP16/BF16 storage with CTAS sampling
// when at least half of the angles pair up as theta and pi - theta (its
// staging indexes a sinogram with 32-bit offsets).
bool mirror_applies(int flags, int N, int w, int n_paired) {
  return mode_of(flags) == MODE_CTAS && store_of(flags) != ST_F32 && 4 * n_paired >= N &&
         static_cast<long long>(N) * w + 2LL * filt_pad(w) < (1LL << 31);
}

// Filtered sinograms of 16-bit storage carry zero pads (read without
// bounds checks); float32 ones are contiguous (the host API copies them out).
int filt_pad_of(const ctas_fbp *h) { return store_of(h->flags) == ST_F32 ? 0 : filt_pad(h->w); }

// FP16/BF16 storage with CTAS sampling takes the mirrored path when at
// least half of the angles pair up as theta and pi - theta.
bool use_mirror(const ctas_fbp *h) { return mirror_applies(h->flags, h->N, h->w, h->n_paired); }

// LINEAR sampling takes the mirrored path too, on 64x64 tiles (the 16-byte
// float words of its two windows fill the static LDS of a workgroup).
bool mirror_linear_applies(int flags, int N, int w, int n_paired, int kind) {
  const int pad = store_of(flags) == ST_F32 ? 0 : filt_pad(w);
  return mode_of(flags) == MODE_LINEAR && kind == TK_64x64 && 4 * n_paired >= N &&
         static_cast<long long>(N) * w + 2LL * pad < (1LL << 31);
}
bool use_mirror(const ctas_fbp *h, int kind) {
  return use_mirror(h) || mirror_linear_applies(h->flags, h->N, h->w, h->n_paired, kind);
}

Item base_item(const ctas_fbp *h, int kind, bool mirror) {
  Item it{};
  it.sincos = h->d_sincos;
  it.sincos64 = use_fp64_angles(h->flags) ? h->d_sincos64 : nullptr;
  it.neg_sin = h->neg_sin;
  it.W = h->d_W;
  it.tw = h->d_tw;
  it.recCof = h->recCof;
  it.w = h->w;
  it.N = h->N;
  it.Z = h->Z;
  it.logZ = h->logZ;
  // mirrored: tiles cover the columns 0 .. w/2, each with its mirror image
  const int cols = mirror ? h->w / 2 + 1 : h->w;
  it.tiles_x = (cols + tile_cols(kind) - 1) / tile_cols(kind);
  it.tiles = it.tiles_x * ((h->w + tile_rows(kind) - 1) / tile_rows(kind));
  it.kind = kind;
  // the mirrored CTAS kernel adds four FP16 samples; LINEAR stages floats
  it.half_max = mirror && mode_of(h->flags) == MODE_CTAS ? kHalfSumMax : kHalfMax;
  if (mirror) {
    it.slots = h->d_slots;
    it.slot_tab = h->d_slot_tab;
    it.n_slots = h->n_slots;
  }
  return it;
}

hipStream_t pick_stream(const ctas_fbp *h, void *stream) {
  return stream ? static_cast<hipStream_t>(stream) : h->stream;
}

}  // namespace

int stream_wait(hipStream_t waiter, hipStream_t producer) {
  if (waiter == producer) return CTAS_OK;
  hipEvent_t e = nullptr;
  CTAS_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  hipError_t r = hipEventRecord(e, producer);
  if (r == hipSuccess) r = hipStreamWaitEvent(waiter, e, 0);
  (void)hipEventDestroy(e);
  if (r != hipSuccess) return fail(CTAS_E_DEVICE, std::string("stream ordering: ") + hipGetErrorString(r));
  return CTAS_OK;
}

namespace {

// Enqueue prepare (+ optional backprojection) for S slices of device data.
int enqueue_slices(ctas_fbp *h, const float *d_sino, float *d_out, float *d_f32,
                   const float *centers, int S, bool do_bp, hipStream_t s) {
  if (S <= 0) return fail(CTAS_E_ARG, "number of slices must be positive");
  if (do_bp)
    for (int i = 0; i < S; ++i)
      if (int rc = check_center(h->w, centers[i])) return rc;
  const int st = store_of(h->flags);
  const size_t slice_elems = static_cast<size_t>(h->N) * h->w;
  const int pad = filt_pad_of(h);
  const size_t stride = slice_elems + 2 * static_cast<size_t>(pad);
  if (int rc = grow(h->d_filt, h->filt_cap, stride * S * store_bytes(st))) return rc;
  if (int rc = ensure_slices(h, S)) return rc;
  const int kind = choose_kind(h);
  const bool mirror = use_mirror(h, kind);
  const int pairs = (h->N + 1) / 2;
  std::vector<Item> items(S);
  for (int i = 0; i < S; ++i) {
    Item it = base_item(h, kind, mirror);
    it.sino = d_sino + slice_elems * i;
    it.filt = static_cast<unsigned char *>(h->d_filt) + (stride * i + pad) * store_bytes(st);
    it.pad = pad;
    it.filt_f32 = d_f32 ? d_f32 + slice_elems * i : nullptr;
    it.out = d_out ? d_out + static_cast<size_t>(h->w) * h->w * i : nullptr;
    it.minmax = h->d_minmax + static_cast<size_t>(kMinMaxStride) * i;
    it.status = h->d_status + i;
    it.center = do_bp ? centers[i] : 0.0f;
    it.tile_base = it.tiles * i;
    it.pair_base = pairs * i;
    items[i] = it;
  }
  bool changed = false;
  if (int rc = upload_items(h, items, s, &changed, extra(h).resident)) return rc;
  if (pad) CTAS_HIP_TRY(launch_pad_zero(h->d_items, S, s));
  CTAS_HIP_TRY(launch_filter(h->d_items, S, pairs * S, h->logZ, h->logZ, h->contrast, st, s));
  if (do_bp)
    CTAS_HIP_TRY(launch_backproject(h->d_items, S, items[0].tiles * S, kind, choose_ac(h),
                                    mode_of(h->flags), bp_store(st, h->contrast, mirror), mirror, s));
  h->last_S = S;
  return CTAS_OK;
}

// Backproject n centres from the filtered slices already on device: all
// from slice `slice` (per_slice == false, CTrec::repeat) or centre i from
// slice i (per_slice == true).
int enqueue_repeat(ctas_fbp *h, int slice, float *d_out, const float *centers, int n, hipStream_t s,
                   bool per_slice = false) {
  if (n <= 0) return fail(CTAS_E_ARG, "number of centres must be positive");
  if (slice < 0 || slice >= h->kept_slices || (per_slice && n > h->kept_slices))
    return fail(CTAS_E_STATE, "no filtered sinogram on the device: call ctas_fbp_sino first");
  for (int i = 0; i < n; ++i)
    if (int rc = check_center(h->w, centers[i])) return rc;
  const int st = store_of(h->flags);
  const int kind = choose_kind(h);
  const bool mirror = use_mirror(h, kind);
  const size_t slice_elems = static_cast<size_t>(h->N) * h->w;
  const int pad = filt_pad_of(h);
  const size_t stride = slice_elems + 2 * static_cast<size_t>(pad);
  std::vector<Item> items(n);
  for (int i = 0; i < n; ++i) {
    Item it = base_item(h, kind, mirror);
    it.filt = static_cast<unsigned char *>(h->d_filt) + (stride * (per_slice ? i : slice) + pad) * store_bytes(st);
    it.pad = pad;
    it.out = d_out + static_cast<size_t>(h->w) * h->w * i;
    it.center = centers[i];
    it.tile_base = it.tiles * i;
    it.pair_base = 0;
    items[i] = it;
  }
  bool changed = false;
  if (int rc = upload_items(h, items, s, &changed, extra(h).resident)) return rc;
  CTAS_HIP_TRY(launch_backproject(h->d_items, n, items[0].tiles * n, kind, choose_ac(h),
                                  mode_of(h->flags), bp_store(st, h->contrast, mirror), mirror, s));
  return CTAS_OK;
}

// The status words carry deAbs (ABS contrast) and the FP16 range check (FP16
// storage of the other contrasts); other handles have nothing to report.
bool has_status(int contrast, int store) { return contrast == CTAS_CONTRAST_ABS || store == ST_F16; }

int status_error(int code) {
  if (code == CTAS_E_DATA) return fail(CTAS_E_DATA, "Bad absorption data (maximum is sub-zero).");
  return fail(CTAS_E_RANGE, "a filtered sample exceeds the FP16 range of the backprojection (" +
                                std::to_string(static_cast<int>(kHalfSumMax)) +
                                " with mirrored angle pairs in CTAS sampling, else 65504): use BF16 or FP32 storage");
}

int read_status(ctas_fbp *h, int S) {
  if (!has_status(h->contrast, store_of(h->flags))) return CTAS_OK;
  std::vector<int> st(S);
  CTAS_HIP_TRY(hipMemcpy(st.data(), h->d_status, sizeof(int) * S, hipMemcpyDeviceToHost));
  for (int i = 0; i < S; ++i)
    if (st[i] != 0) return status_error(st[i]);
  return CTAS_OK;
}

// ---- autotuner ------------------------------------------------------------
struct TuneKey {
  int wclass, nclass, mode, store, dev;
  bool operator<(const TuneKey &o) const {
    return std::tie(wclass, nclass, mode, store, dev) < std::tie(o.wclass, o.nclass, o.mode, o.store, o.dev);
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, std::pair<int, int>> g_tune;

int width_class(int w) {
  int c = 1;
  while (c < w) c <<= 1;
  return c;
}
TuneKey tune_key(const ctas_fbp *h) {
  return TuneKey{width_class(h->w), width_class(h->N), mode_of(h->flags), store_of(h->flags), h->device};
}

int choose_kind(const ctas_fbp *h) {
  if (h->tile_kind != TK_AUTO) return h->tile_kind;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(tune_key(h));
    if (it != g_tune.end()) return it->second.first;
  }
  return TK_64x64;  // the tuned choice at 512-4096 px on MI355X
}
int choose_ac(const ctas_fbp *h) {
  if (h->angles_per_pass) return h->angles_per_pass;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(tune_key(h));
    if (it != g_tune.end()) return it->second.second;
  }
  return 32;
}

// ---- shared resources of uneven batches -------------------------------------
struct GeomRes {
  float *d_W = nullptr;
  float2 *d_sincos = nullptr;
  double2 *d_sincos64 = nullptr;
  int neg_sin = 0;
  float2 *d_tw = nullptr;
  MirrorSlot *d_slots = nullptr;  // mirror slots of the FP64 table and their row/column terms
  int *d_slot_tab = nullptr;
  int n_slots = 0, n_paired = 0;
};

// 1 when some {sin, cos} entry has a negative sine.
int any_negative_sine(const std::vector<double> &sc64) {
  for (size_t i = 0; i < sc64.size(); i += 2)
    if (sc64[i] < 0.0) return 1;
  return 0;
}
std::mutex g_geo_mu;
std::map<std::tuple<int, int, int, int, float, float, int>, GeomRes> g_geo;  // (dev, w, N, filter, param, anarc, contrast)

int make_twiddles(int Z, float2 **d_tw) {
  std::vector<float2> tw(std::max(Z / 2, 1));
  for (int m = 0; m < Z / 2; ++m) {
    const double a = -2.0 * M_PI * m / Z;
    tw[m] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
  }
  if (hipMalloc(d_tw, sizeof(float2) * tw.size()) != hipSuccess)
    return fail(CTAS_E_NOMEM, "twiddle allocation failed");
  CTAS_HIP_TRY(hipMemcpy(*d_tw, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice));
  return CTAS_OK;
}

int geometry_resources(int w, int N, int filter, float param, float anarc, int contrast, GeomRes *out) {
  int dev = 0;
  CTAS_HIP_TRY(hipGetDevice(&dev));
  const auto key = std::make_tuple(dev, w, N, filter, param, anarc, contrast);
  std::lock_guard<std::mutex> lk(g_geo_mu);
  auto f = g_geo.find(key);
  if (f != g_geo.end()) {
    *out = f->second;
    return CTAS_OK;
  }
  std::vector<float> W;
  if (int rc = filter_window(filter, param, w, contrast, W)) return rc;
  std::vector<float> sc(2 * static_cast<size_t>(N));
  angle_table(N, anarc, sc.data());
  GeomRes r;
  std::vector<double> sc64(2 * static_cast<size_t>(N));
  angle_table64(N, anarc, sc64.data());
  r.neg_sin = any_negative_sine(sc64);
  if (hipMalloc(&r.d_W, sizeof(float) * W.size()) != hipSuccess ||
      hipMalloc(&r.d_sincos, sizeof(float2) * N) != hipSuccess ||
      hipMalloc(&r.d_sincos64, sizeof(double2) * N) != hipSuccess)
    return fail(CTAS_E_NOMEM, "geometry allocation failed");
  CTAS_HIP_TRY(hipMemcpy(r.d_sincos64, sc64.data(), sizeof(double) * sc64.size(), hipMemcpyHostToDevice));
  CTAS_HIP_TRY(hipMemcpy(r.d_W, W.data(), sizeof(float) * W.size(), hipMemcpyHostToDevice));
  CTAS_HIP_TRY(hipMemcpy(r.d_sincos, sc.data(), sizeof(float) * sc.size(), hipMemcpyHostToDevice));
  if (int rc = make_twiddles(zidth(w), &r.d_tw)) return rc;
  if (int rc = build_slots(sc64, N, r.d_slots, r.d_slot_tab, &r.n_slots, &r.n_paired)) return rc;
  g_geo[key] = r;
  *out = r;
  return CTAS_OK;
}

struct BatchWork {
  void *d_filt = nullptr;  size_t filt_cap = 0;
  float2 *d_mm = nullptr;  size_t mm_cap = 0;
  int *d_status = nullptr; size_t st_cap = 0;
  Item *d_items = nullptr; size_t it_cap = 0;
  Item *d_items_f = nullptr; size_t itf_cap = 0;  // the same items grouped by line length
  unsigned *d_counter = nullptr; size_t ctr_cap = 0;
};
std::mutex g_batch_mu;
std::map<int, BatchWork> g_batch;

}  // namespace
}  // namespace ctas

using namespace ctas;

extern "C" {

const char *ctas_last_error(void) { return g_last_error.c_str(); }

const char *ctas_version(void) { return "ctas-mi355x 0.1.0 (gfx950)"; }

int ctas_device_count(int *n) {
  if (!n) return fail(CTAS_E_ARG, "null argument");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *n = c;
  return CTAS_OK;
}

int ctas_zidth(int w) { return zidth(w); }
float ctas_arc_rad(float anarc, int points) { return arc_rad(anarc, points); }
float ctas_step_rad(float anarc, int points) { return step_rad(anarc, points); }

int ctas_filter_window(int filter, float param, int w, int contrast, float *W) {
  if (!W) return fail(CTAS_E_ARG, "null argument");
  std::vector<float> v;
  if (int rc = filter_window(filter, param, w, contrast, v)) return rc;
  std::memcpy(W, v.data(), sizeof(float) * v.size());
  return CTAS_OK;
}

int ctas_angles(int n, float anarc, float *sincos) {
  if (!sincos || n <= 0) return fail(CTAS_E_ARG, "bad angle table arguments");
  angle_table(n, anarc, sincos);
  return CTAS_OK;
}

int ctas_filter_parse(const char *spec, int *filter, float *param) {
  if (!spec || !filter || !param) return fail(CTAS_E_ARG, "null argument");
  std::string in(spec), name = in;
  float p = 0.0f;
  const size_t colon = in.find(':');
  if (colon != std::string::npos) {
    name = in.substr(0, colon);
    const std::string ps = in.substr(colon + 1);
    char *end = nullptr;
    p = std::strtof(ps.c_str(), &end);
    if (ps.empty() || (end && *end)) return fail(CTAS_E_ARG, "could not parse filter parameter \"" + ps + "\"");
  }
  std::string up = name;
  for (char &ch : up) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
  int f;
  // Filter::fromName (lct.cpp:71-86); "BARLETT" selects RAMP there.
  if (up == "NONE") f = CTAS_FILTER_NONE;
  else if (up.empty() || up == "RAMP") f = CTAS_FILTER_RAMP;
  else if (up == "BARLETT") f = CTAS_FILTER_RAMP;
  else if (up == "WELCH") f = CTAS_FILTER_WELCH;
  else if (up == "PARZEN") f = CTAS_FILTER_PARZEN;
  else if (up == "HANN") f = CTAS_FILTER_HANN;
  else if (up == "HAMMING") f = CTAS_FILTER_HAMMING;
  else if (up == "BLACKMAN") f = CTAS_FILTER_BLACKMAN;
  else if (up == "LANCKZOS") f = CTAS_FILTER_LANCKZOS;
  else if (up == "KAISER") f = CTAS_FILTER_KAISER;
  else if (up == "GAUSS") f = CTAS_FILTER_GAUSS;
  else return fail(CTAS_E_ARG, "The string \"" + name + "\" does not describe any known filter.");
  // Filter ctor checks (ct.h:148-151)
  if (f == CTAS_FILTER_KAISER && p < 0.0f)
    return fail(CTAS_E_ARG, "Alpha parameter of Kaiser filter must be non-negative (set " + fmt_float(p) + "). ");
  if (f == CTAS_FILTER_GAUSS && p <= 0.0f)
    return fail(CTAS_E_ARG, "Sigma parameter of Guass filter must be strictly positive (set " + fmt_float(p) + "). ");
  *filter = f;
  *param = p;
  return CTAS_OK;
}

int ctas_contrast_parse(const char *name, int *contrast) {
  if (!name || !contrast) return fail(CTAS_E_ARG, "null argument");
  std::string up(name);
  for (char &ch : up) ch = static_cast<char>(std::toupper(static_cast<unsigned char>(ch)));
  if (up == "A" || up == "ABS" || up == "ABSORPTION") *contrast = CTAS_CONTRAST_ABS;
  else if (up == "R" || up == "REF" || up == "REFRACTION") *contrast = CTAS_CONTRAST_REF;
  else if (up == "P" || up == "PHS" || up == "PHASE") *contrast = CTAS_CONTRAST_PHS;
  else if (up == "F" || up == "FLT" || up == "FILTERED") *contrast = CTAS_CONTRAST_FLT;
  else return fail(CTAS_E_ARG, std::string("The string \"") + name + "\" does not describe any known contrast.");
  return CTAS_OK;
}

int ctas_fbp_create(ctas_fbp **ph, int w, int n_angles, int contrast, int filter, float filter_param,
                    float anarc, int flags) {
  if (!ph) return fail(CTAS_E_ARG, "null handle pointer");
  *ph = nullptr;
  if (w < 2) return fail(CTAS_E_ARG, "Number of pixels in the CT reconstruction less than 2.");
  if (n_angles <= 0) return fail(CTAS_E_ARG, "Zero projections in sinogram.");
  if (contrast < CTAS_CONTRAST_ABS || contrast > CTAS_CONTRAST_FLT)
    return fail(CTAS_E_ARG, "unknown contrast code " + std::to_string(contrast));
  if ((flags & ~(CTAS_INTERP_MASK | CTAS_STORE_MASK)) != 0 || (flags & CTAS_INTERP_MASK) > CTAS_INTERP_LINEAR ||
      (flags & CTAS_STORE_MASK) > CTAS_STORE_BF16)
    return fail(CTAS_E_ARG, "unknown flags " + std::to_string(flags));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(CTAS_E_DEVICE, "no HIP device is available");
  std::unique_ptr<ctas_fbp> h(new ctas_fbp);
  CTAS_HIP_TRY(hipGetDevice(&h->device));
  h->w = w;
  h->N = n_angles;
  h->Z = zidth(w);
  h->logZ = 0;
  while ((1 << h->logZ) < h->Z) ++h->logZ;
  hipDeviceProp_t prop;
  CTAS_HIP_TRY(hipGetDeviceProperties(&prop, h->device));
  if (sizeof(float2) * static_cast<size_t>(h->Z) > prop.sharedMemPerBlock)
    return fail(CTAS_E_ARG, "width " + std::to_string(w) + " needs a padded line of " +
                                std::to_string(h->Z) + " points, more than one workgroup's LDS holds");
  h->contrast = contrast;
  h->filter = filter;
  h->fparam = filter_param;
  h->anarc = anarc;
  h->flags = flags;
  if (int rc = filter_window(filter, filter_param, w, contrast, h->W)) return rc;
  h->sincos.resize(2 * static_cast<size_t>(n_angles));
  angle_table(n_angles, anarc, h->sincos.data());
  CTAS_HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  if (hipMalloc(&h->d_W, sizeof(float) * h->Z) != hipSuccess ||
      hipMalloc(&h->d_sincos, sizeof(float2) * n_angles) != hipSuccess ||
      hipMalloc(&h->d_sincos64, sizeof(double2) * n_angles) != hipSuccess)
    return fail(CTAS_E_NOMEM, "device allocation failed");
  {
    std::vector<double> sc64(2 * static_cast<size_t>(n_angles));
    angle_table64(n_angles, anarc, sc64.data());
    h->neg_sin = any_negative_sine(sc64);
    CTAS_HIP_TRY(hipMemcpy(h->d_sincos64, sc64.data(), sizeof(double) * sc64.size(), hipMemcpyHostToDevice));
    if (int rc = upload_slots(h.get(), sc64)) return rc;
  }
  CTAS_HIP_TRY(hipMemcpy(h->d_W, h->W.data(), sizeof(float) * h->Z, hipMemcpyHostToDevice));
  CTAS_HIP_TRY(hipMemcpy(h->d_sincos, h->sincos.data(), sizeof(float) * h->sincos.size(), hipMemcpyHostToDevice));
  if (int rc = make_twiddles(h->Z, &h->d_tw)) return rc;
  *ph = h.release();
  return CTAS_OK;
}

int ctas_fbp_clone(const ctas_fbp *src, ctas_fbp **ph) {
  if (!src || !ph) return fail(CTAS_E_ARG, "null argument");
  ctas_fbp *h = nullptr;
  if (int rc = ctas_fbp_create(&h, src->w, src->N, src->contrast, src->filter, src->fparam, src->anarc, src->flags))
    return rc;
  h->recCof = src->recCof;
  h->tile_kind = src->tile_kind;
  h->angles_per_pass = src->angles_per_pass;
  if (src->sincos != h->sincos) ctas_fbp_set_angles(h, src->sincos.data());
  *ph = h;
  return CTAS_OK;
}

int ctas_fbp_destroy(ctas_fbp *h) {
  if (!h) return CTAS_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  void *ptrs[] = {h->d_W, h->d_sincos, h->d_sincos64, h->d_slots, h->d_slot_tab, h->d_tw, h->d_filt, h->d_io, h->d_out, h->d_minmax,
                  h->d_status, h->d_items, h->d_centres, h->d_partial, h->d_tmp, h->d_scores};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  drop_extra(h);
  delete h;
  return CTAS_OK;
}

int ctas_fbp_rec_shape(const ctas_fbp *h, float center, int *rows, int *cols) {
  if (!h || !rows || !cols) return fail(CTAS_E_ARG, "null argument");
  const int r = static_cast<int>(std::ceil(std::fabs(center)));
  *rows = h->w - 2 * r;
  *cols = h->w;
  return CTAS_OK;
}

int ctas_fbp_rec_crop(const ctas_fbp *h, float center, int *row_begin, int *row_end) {
  if (!h || !row_begin || !row_end) return fail(CTAS_E_ARG, "null argument");
  const int r = static_cast<int>(std::ceil(std::fabs(center)));
  *row_begin = r;
  *row_end = h->w - r;
  return CTAS_OK;
}

int ctas_fbp_info(const ctas_fbp *h, int *w, int *n, int *z, int *flags) {
  if (!h) return fail(CTAS_E_ARG, "null handle");
  if (w) *w = h->w;
  if (n) *n = h->N;
  if (z) *z = h->Z;
  if (flags) *flags = h->flags;
  return CTAS_OK;
}

int ctas_fbp_set_physics(ctas_fbp *h, float pixel_size, float lambda) {
  if (!h) return fail(CTAS_E_ARG, "null handle");
  // lct.cpp:558-566: float recCof scaled by double factors.
  float r = 1.0f;
  if (lambda > 0.0f) r = static_cast<float>(r * (lambda / (2 * M_PI)));
  if (pixel_size > 0.0f) r = static_cast<float>(r * (M_PI / pixel_size));
  h->recCof = r;
  return CTAS_OK;
}

int ctas_fbp_get_reccof(const ctas_fbp *h, float *r) {
  if (!h || !r) return fail(CTAS_E_ARG, "null argument");
  *r = h->recCof;
  return CTAS_OK;
}

int ctas_fbp_set_angles(ctas_fbp *h, const float *sincos) {
  if (!h || !sincos) return fail(CTAS_E_ARG, "null argument");
  CTAS_HIP_TRY(hipStreamSynchronize(h->stream));
  std::memcpy(h->sincos.data(), sincos, sizeof(float) * h->sincos.size());
  CTAS_HIP_TRY(hipMemcpy(h->d_sincos, h->sincos.data(), sizeof(float) * h->sincos.size(), hipMemcpyHostToDevice));
  // a caller-supplied table replaces the FP64 one too (promoted exactly)
  std::vector<double> sc64(h->sincos.begin(), h->sincos.end());
  h->neg_sin = any_negative_sine(sc64);
  CTAS_HIP_TRY(hipMemcpy(h->d_sincos64, sc64.data(), sizeof(double) * sc64.size(), hipMemcpyHostToDevice));
  return upload_slots(h, sc64);
}

int ctas_fbp_get_angles(const ctas_fbp *h, float *sincos) {
  if (!h || !sincos) return fail(CTAS_E_ARG, "null argument");
  std::memcpy(sincos, h->sincos.data(), sizeof(float) * h->sincos.size());
  return CTAS_OK;
}

int ctas_fbp_get_window(const ctas_fbp *h, float *W) {
  if (!h || !W) return fail(CTAS_E_ARG, "null argument");
  std::memcpy(W, h->W.data(), sizeof(float) * h->W.size());
  return CTAS_OK;
}

int ctas_fbp_set_tiles(ctas_fbp *h, int kind, int ac) {
  if (!h) return fail(CTAS_E_ARG, "null handle");
  if (kind < TK_AUTO || kind > TK_128x128) return fail(CTAS_E_ARG, "unknown tile kind");
  if (ac != 0 && ac != 16 && ac != 32) return fail(CTAS_E_ARG, "angles per pass must be 0, 16 or 32");
  h->tile_kind = kind;
  h->angles_per_pass = ac;
  return CTAS_OK;
}

int ctas_fbp_get_tiles(const ctas_fbp *h, int *kind, int *ac) {
  if (!h) return fail(CTAS_E_ARG, "null handle");
  if (kind) *kind = choose_kind(h);
  if (ac) *ac = choose_ac(h);
  return CTAS_OK;
}

// Host-API work: upload S slices, prepare (and backproject), download.
static int host_run(ctas_fbp *h, float *sino, float *out, const float *centers, int S, bool do_bp) {
  if (!h || !sino || S <= 0) return fail(CTAS_E_ARG, "bad arguments");
  if (do_bp && (!out || !centers)) return fail(CTAS_E_ARG, "bad arguments");
  const size_t slice_elems = static_cast<size_t>(h->N) * h->w;
  const size_t in_bytes = sizeof(float) * slice_elems * S;
  const bool f32_copy = store_of(h->flags) != ST_F32;
  if (int rc = grow(h->d_io, h->io_cap, in_bytes * (f32_copy ? 2 : 1))) return rc;
  if (do_bp)
    if (int rc = grow(h->d_out, h->out_cap, sizeof(float) * h->w * h->w * S)) return rc;
  hipStream_t s = h->stream;
  CTAS_HIP_TRY(hipStreamSynchronize(s));
  CTAS_HIP_TRY(hipMemcpy(h->d_io, sino, in_bytes, hipMemcpyHostToDevice));
  float *d_f32 = f32_copy ? h->d_io + slice_elems * S : nullptr;
  if (int rc = enqueue_slices(h, h->d_io, do_bp ? h->d_out : nullptr, d_f32, centers, S, do_bp, s)) return rc;
  CTAS_HIP_TRY(hipStreamSynchronize(s));
  if (int rc = read_status(h, S)) return rc;
  // The reference filters the caller's sinogram in place.
  if (h->contrast != CTAS_CONTRAST_FLT) {
    const void *src = f32_copy ? static_cast<const void *>(d_f32) : h->d_filt;
    CTAS_HIP_TRY(hipMemcpy(sino, src, in_bytes, hipMemcpyDeviceToHost));
  }
  if (do_bp) CTAS_HIP_TRY(hipMemcpy(out, h->d_out, sizeof(float) * h->w * h->w * S, hipMemcpyDeviceToHost));
  h->kept_slices = S;
  return CTAS_OK;
}

int ctas_fbp_prepare(ctas_fbp *h, float *sino, int S) { return host_run(h, sino, nullptr, nullptr, S, false); }

int ctas_fbp_sino(ctas_fbp *h, float *sino) { return host_run(h, sino, nullptr, nullptr, 1, false); }

int ctas_fbp_reconstruct(ctas_fbp *h, float *sino, float *out, const float *centers, int S) {
  return host_run(h, sino, out, centers, S, true);
}

int ctas_fbp_repeat(ctas_fbp *h, float *out, const float *centers, int n) {
  if (!h || !out || !centers || n <= 0) return fail(CTAS_E_ARG, "bad arguments");
  if (int rc = grow(h->d_out, h->out_cap, sizeof(float) * h->w * h->w * n)) return rc;
  if (int rc = enqueue_repeat(h, 0, h->d_out, centers, n, h->stream)) return rc;
  CTAS_HIP_TRY(hipStreamSynchronize(h->stream));
  CTAS_HIP_TRY(hipMemcpy(out, h->d_out, sizeof(float) * h->w * h->w * n, hipMemcpyDeviceToHost));
  return CTAS_OK;
}

int ctas_fbp_reconstruct_dev(ctas_fbp *h, const float *d_sino, float *d_out, const float *centers, int S,
                             void *stream) {
  if (!h || !d_sino || !d_out || !centers) return fail(CTAS_E_ARG, "bad arguments");
  NullStreamJoin join(h->stream, stream);
  if (join.rc) return join.rc;
  const int rc = enqueue_slices(h, d_sino, d_out, nullptr, centers, S, true, pick_stream(h, stream));
  if (rc == CTAS_OK) h->kept_slices = S;
  return rc;
}

int ctas_fbp_repeat_dev(ctas_fbp *h, int slice, float *d_out, const float *centers, int n, void *stream) {
  if (!h || !d_out || !centers) return fail(CTAS_E_ARG, "bad arguments");
  NullStreamJoin join(h->stream, stream);
  if (join.rc) return join.rc;
  return enqueue_repeat(h, slice, d_out, centers, n, pick_stream(h, stream));
}

int ctas_fbp_prepare_dev(ctas_fbp *h, const float *d_sino, int S, void *stream) {
  if (!h || !d_sino) return fail(CTAS_E_ARG, "bad arguments");
  NullStreamJoin join(h->stream, stream);
  if (join.rc) return join.rc;
  const int rc = enqueue_slices(h, d_sino, nullptr, nullptr, nullptr, S, false, pick_stream(h, stream));
  if (rc == CTAS_OK) h->kept_slices = S;
  return rc;
}

int ctas_fbp_backproject_dev(ctas_fbp *h, float *d_out, const float *centers, int S, void *stream) {
  if (!h || !d_out || !centers) return fail(CTAS_E_ARG, "bad arguments");
  NullStreamJoin join(h->stream, stream);
  if (join.rc) return join.rc;
  return enqueue_repeat(h, 0, d_out, centers, S, pick_stream(h, stream), true);
}

int ctas_fbp_status(ctas_fbp *h, int S, int *status) {
  if (!h || !status || S <= 0 || S > h->mm_cap) return fail(CTAS_E_ARG, "bad arguments");
  CTAS_HIP_TRY(hipStreamSynchronize(h->stream));
  if (!has_status(h->contrast, store_of(h->flags))) {
    for (int i = 0; i < S; ++i) status[i] = 0;
    return CTAS_OK;
  }
  CTAS_HIP_TRY(hipMemcpy(status, h->d_status, sizeof(int) * S, hipMemcpyDeviceToHost));
  return CTAS_OK;
}

int ctas_fbp_graph_capture(ctas_fbp *h, const float *d_sino, float *d_out, const float *centers, int S,
                           void *stream) {
  if (!h || !d_sino || !d_out || !centers) return fail(CTAS_E_ARG, "bad arguments");
  hipStream_t s = pick_stream(h, stream);
  if (h->graph_exec) {
    (void)hipGraphExecDestroy(h->graph_exec);
    h->graph_exec = nullptr;
  }
  if (h->graph) {
    (void)hipGraphDestroy(h->graph);
    h->graph = nullptr;
  }
  // Make the item table resident before capture so that no copy is recorded.
  if (!stream) CTAS_HIP_TRY(hipDeviceSynchronize());
  CTAS_HIP_TRY(hipStreamSynchronize(s));
  extra(h).resident.clear();
  {
    // first enqueue uploads the table; it also runs the work once
    if (int rc = enqueue_slices(h, d_sino, d_out, nullptr, centers, S, true, s)) return rc;
    CTAS_HIP_TRY(hipStreamSynchronize(s));
  }
  CTAS_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_slices(h, d_sino, d_out, nullptr, centers, S, true, s);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(CTAS_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  h->graph = g;
  CTAS_HIP_TRY(hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0));
  h->kept_slices = S;
  return CTAS_OK;
}

int ctas_fbp_graph_launch(ctas_fbp *h, void *stream) {
  if (!h) return fail(CTAS_E_ARG, "null handle");
  if (!h->graph_exec) return fail(CTAS_E_STATE, "no captured graph: call ctas_fbp_graph_capture first");
  NullStreamJoin join(h->stream, stream);
  if (join.rc) return join.rc;
  CTAS_HIP_TRY(hipGraphLaunch(h->graph_exec, pick_stream(h, stream)));
  return CTAS_OK;
}

int ctas_fbp_autotune(ctas_fbp *h, int S) {
  if (!h || S <= 0) return fail(CTAS_E_ARG, "bad arguments");
  const int keep_kind = h->tile_kind, keep_ac = h->angles_per_pass;
  const size_t slice_elems = static_cast<size_t>(h->N) * h->w;
  float *d_in = nullptr, *d_o = nullptr;
  if (hipMalloc(&d_in, sizeof(float) * slice_elems * S) != hipSuccess ||
      hipMalloc(&d_o, sizeof(float) * h->w * h->w * S) != hipSuccess) {
    if (d_in) (void)hipFree(d_in);
    return fail(CTAS_E_NOMEM, "autotune allocation failed");
  }
  // Every path below ends at the clean-up: buffers and events freed, the
  // handle's own tile settings restored.
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float best = 1e30f;
  int bk = TK_128x128, bac = 32;
  auto hip = [](hipError_t e, const char *what) {
    return e == hipSuccess ? CTAS_OK : fail(CTAS_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  };
  auto run = [&]() -> int {
    // Transmission-like data in (0.5, 1]: a constant is enough to time the
    // gather, whose cost does not depend on the values.
    std::vector<float> ones(slice_elems * S, 0.75f);
    if (int rc = hip(hipMemcpy(d_in, ones.data(), sizeof(float) * ones.size(), hipMemcpyHostToDevice), "hipMemcpy"))
      return rc;
    if (int rc = hip(hipEventCreate(&e0), "hipEventCreate")) return rc;
    if (int rc = hip(hipEventCreate(&e1), "hipEventCreate")) return rc;
    std::vector<float> centers(S, 0.0f);
    const int kinds[] = {TK_64x64, TK_128x64, TK_128x128};
    const int acs[] = {16, 32};
    for (int k : kinds) {
      for (int ac : acs) {
        h->tile_kind = k;
        h->angles_per_pass = ac;
        if (int rc = enqueue_slices(h, d_in, d_o, nullptr, centers.data(), S, true, h->stream)) return rc;
        // the fastest of five runs: one slow run (a clock or cache state)
        // must not decide between candidates a few percent apart
        float ms = 1e30f;
        for (int r = 0; r < 5; ++r) {
          if (int rc = hip(hipEventRecord(e0, h->stream), "hipEventRecord")) return rc;
          if (int rc = enqueue_slices(h, d_in, d_o, nullptr, centers.data(), S, true, h->stream)) return rc;
          if (int rc = hip(hipEventRecord(e1, h->stream), "hipEventRecord")) return rc;
          if (int rc = hip(hipEventSynchronize(e1), "hipEventSynchronize")) return rc;
          float one = 0.f;
          if (int rc = hip(hipEventElapsedTime(&one, e0, e1), "hipEventElapsedTime")) return rc;
          ms = std::min(ms, one);
        }
        if (ms < best) {
          best = ms;
          bk = k;
          bac = ac;
        }
      }
    }
    return CTAS_OK;
  };
  const int rc = run();
  (void)hipStreamSynchronize(h->stream);  // nothing in flight reads the buffers
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(d_in);
  (void)hipFree(d_o);
  h->tile_kind = keep_kind;
  h->angles_per_pass = keep_ac;
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[tune_key(h)] = {bk, bac};
  return CTAS_OK;
}

int ctas_fbp_batch_uneven(const ctas_batch_item *items, int n, int contrast, int filter, float filter_param,
                          float anarc, int flags, void *stream) {
  if (!items || n <= 0) return fail(CTAS_E_ARG, "bad arguments");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(CTAS_E_DEVICE, "no HIP device is available");
  int dev = 0, cus = 0;
  CTAS_HIP_TRY(hipGetDevice(&dev));
  CTAS_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int st = store_of(flags), mode = mode_of(flags);
  for (int i = 0; i < n; ++i) {
    const ctas_batch_item &bi = items[i];
    if (bi.w < 2 || bi.n_angles <= 0 || !bi.d_sino || !bi.d_out)
      return fail(CTAS_E_ARG, "bad batch item " + std::to_string(i));
    if (int rc = check_center(bi.w, bi.center)) return rc;
  }
  // Tile kind and angles per pass of each (width class, angle-count class)
  // from the autotuner: classes not tuned yet in this process are timed now
  // on a handle of the first item of the class.
  std::map<std::pair<int, int>, std::pair<int, int>> pick;
  for (int i = 0; i < n; ++i) {
    const auto cls = std::make_pair(width_class(items[i].w), width_class(items[i].n_angles));
    if (pick.count(cls)) continue;
    const TuneKey key{cls.first, cls.second, mode, st, dev};
    bool tuned;
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      tuned = g_tune.count(key) != 0;
    }
    if (!tuned) {
      ctas_fbp *t = nullptr;
      int rc = ctas_fbp_create(&t, items[i].w, items[i].n_angles, contrast, filter, filter_param, anarc, flags);
      if (rc == CTAS_OK) rc = ctas_fbp_autotune(t, 2);
      ctas_fbp_destroy(t);
      if (rc) return rc;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    pick[cls] = g_tune[key];
  }
  struct Entry {
    int src, kind, ac;
    bool mirror;
    double cost;
    GeomRes g;
  };
  std::vector<Entry> ents(n);
  for (int i = 0; i < n; ++i) {
    const ctas_batch_item &bi = items[i];
    Entry &e = ents[i];
    e.src = i;
    if (int rc = geometry_resources(bi.w, bi.n_angles, filter, filter_param, anarc, contrast, &e.g)) return rc;
    const auto &pk = pick[std::make_pair(width_class(bi.w), width_class(bi.n_angles))];
    e.kind = pk.first;
    e.ac = pk.second;
    e.mirror = mirror_applies(flags, bi.n_angles, bi.w, e.g.n_paired) ||
               mirror_linear_applies(flags, bi.n_angles, bi.w, e.g.n_paired, e.kind);
    e.cost = static_cast<double>(bi.w) * bi.w * bi.n_angles;
  }
  // One persistent launch per (tile kind, pass length, mirrored) group, its
  // items costliest first, so that each launch ends with the small tiles.
  std::stable_sort(ents.begin(), ents.end(), [](const Entry &a, const Entry &b) {
    if (a.mirror != b.mirror) return a.mirror < b.mirror;
    if (a.kind != b.kind) return a.kind < b.kind;
    if (a.ac != b.ac) return a.ac < b.ac;
    return a.cost > b.cost;
  });
  std::vector<Item> its(n);
  std::vector<size_t> filt_off(n);
  size_t filt_bytes = 0;
  std::vector<std::tuple<int, int, int, int, int, bool>> launches;  // first, items, tiles, kind, ac, mirror
  int tiles = 0, group_first = 0;
  for (int oi = 0; oi < n; ++oi) {
    const Entry &e = ents[oi];
    const ctas_batch_item &bi = items[e.src];
    if (oi > 0 && (e.kind != ents[oi - 1].kind || e.ac != ents[oi - 1].ac || e.mirror != ents[oi - 1].mirror)) {
      const Entry &p = ents[oi - 1];
      launches.emplace_back(group_first, oi - group_first, tiles, p.kind, p.ac, p.mirror);
      group_first = oi;
      tiles = 0;
    }
    Item it{};
    it.sino = bi.d_sino;
    it.out = bi.d_out;
    it.sincos = e.g.d_sincos;
    it.sincos64 = use_fp64_angles(flags) ? e.g.d_sincos64 : nullptr;
    it.neg_sin = e.g.neg_sin;
    it.W = e.g.d_W;
    it.tw = e.g.d_tw;
    it.center = bi.center;
    it.recCof = 1.0f;
    it.w = bi.w;
    it.N = bi.n_angles;
    it.Z = zidth(bi.w);
    while ((1 << it.logZ) < it.Z) ++it.logZ;
    it.kind = e.kind;
    const int cols = e.mirror ? bi.w / 2 + 1 : bi.w;
    it.tiles_x = (cols + tile_cols(e.kind) - 1) / tile_cols(e.kind);
    it.tiles = it.tiles_x * ((bi.w + tile_rows(e.kind) - 1) / tile_rows(e.kind));
    it.half_max = e.mirror && mode == MODE_CTAS ? kHalfSumMax : kHalfMax;
    if (e.mirror) {
      it.slots = e.g.d_slots;
      it.slot_tab = e.g.d_slot_tab;
      it.n_slots = e.g.n_slots;
    }
    // 16-bit filtered sinograms carry zero pads (read without bounds checks)
    it.pad = st == ST_F32 ? 0 : filt_pad(bi.w);
    it.tile_base = tiles;
    tiles += it.tiles;
    filt_off[oi] = filt_bytes + static_cast<size_t>(it.pad) * store_bytes(st);
    filt_bytes += (store_bytes(st) * (static_cast<size_t>(bi.n_angles) * bi.w + 2 * static_cast<size_t>(it.pad)) +
                   255) & ~static_cast<size_t>(255);
    its[oi] = it;
  }
  launches.emplace_back(group_first, n - group_first, tiles, ents[n - 1].kind, ents[n - 1].ac, ents[n - 1].mirror);
  std::lock_guard<std::mutex> lk(g_batch_mu);
  BatchWork &bw = g_batch[dev];
  CTAS_HIP_TRY(hipStreamSynchronize(s));
  if (int rc = grow(bw.d_filt, bw.filt_cap, filt_bytes)) return rc;
  if (int rc = grow(bw.d_mm, bw.mm_cap, sizeof(float2) * kMinMaxStride * n)) return rc;
  if (int rc = grow(bw.d_status, bw.st_cap, sizeof(int) * n)) return rc;
  if (int rc = grow(bw.d_items, bw.it_cap, sizeof(Item) * n)) return rc;
  if (int rc = grow(bw.d_items_f, bw.itf_cap, sizeof(Item) * n)) return rc;
  if (int rc = grow(bw.d_counter, bw.ctr_cap, sizeof(unsigned) * 64)) return rc;
  for (int oi = 0; oi < n; ++oi) {
    its[oi].filt = static_cast<unsigned char *>(bw.d_filt) + filt_off[oi];
    its[oi].minmax = bw.d_mm + static_cast<size_t>(kMinMaxStride) * oi;
    its[oi].status = bw.d_status + oi;
  }
  // The filter runs once per line length, each with its own kernel (the
  // register-resident FFTs for 4096 and 8192 points): the items again,
  // grouped by logZ, with their row pairs counted from each group's start.
  std::vector<Item> fits(its);
  std::stable_sort(fits.begin(), fits.end(), [](const Item &a, const Item &b) { return a.logZ < b.logZ; });
  std::vector<std::tuple<int, int, int, int>> groups;  // first item, items, pairs, logZ
  for (int i = 0; i < n;) {
    int j = i, gp = 0;
    for (; j < n && fits[j].logZ == fits[i].logZ; ++j) {
      fits[j].pair_base = gp;
      gp += (fits[j].N + 1) / 2;
    }
    groups.emplace_back(i, j - i, gp, fits[i].logZ);
    i = j;
  }
  CTAS_HIP_TRY(hipMemcpy(bw.d_items, its.data(), sizeof(Item) * n, hipMemcpyHostToDevice));
  CTAS_HIP_TRY(hipMemcpy(bw.d_items_f, fits.data(), sizeof(Item) * n, hipMemcpyHostToDevice));
  if (st != ST_F32) CTAS_HIP_TRY(launch_pad_zero(bw.d_items, n, s));
  for (const auto &g : groups)
    CTAS_HIP_TRY(launch_filter(bw.d_items_f + std::get<0>(g), std::get<1>(g), std::get<2>(g), std::get<3>(g),
                               std::get<3>(g), contrast, st, s));
  for (size_t l = 0; l < launches.size(); ++l) {
    const auto &L = launches[l];
    CTAS_HIP_TRY(launch_backproject_persistent(bw.d_items + std::get<0>(L), std::get<1>(L), std::get<2>(L),
                                               std::get<3>(L), std::get<4>(L), mode,
                                               bp_store(st, contrast, std::get<5>(L)), std::get<5>(L),
                                               bw.d_counter + l, cus, s));
  }
  CTAS_HIP_TRY(hipStreamSynchronize(s));
  if (has_status(contrast, st)) {
    std::vector<int> stv(n);
    CTAS_HIP_TRY(hipMemcpy(stv.data(), bw.d_status, sizeof(int) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i)
      if (stv[i]) return status_error(stv[i]);
  }
  return CTAS_OK;
}

int ctas_fbp_pipeline(ctas_fbp *h, long n_slices, int batch, const float *centers, int ring_box,
                      ctas_read_fn read, ctas_write_fn write, void *user, int io_threads) {
  if (!h || !centers || !read || !write || n_slices < 0 || batch < 1 || ring_box < 0 || io_threads < 1)
    return fail(CTAS_E_ARG, "bad arguments");
  if (ring_box)
    if (int rc = check_ring_box(ring_box, h->w)) return rc;
  if (n_slices == 0) return CTAS_OK;
  for (long i = 0; i < n_slices; ++i)
    if (int rc = check_center(h->w, centers[i])) return rc;
  const int B = static_cast<int>(std::min<long>(batch, n_slices));
  const long n_batches = (n_slices + B - 1) / B;
  const size_t in_elems = static_cast<size_t>(h->N) * h->w, out_elems = static_cast<size_t>(h->w) * h->w;
  struct DeviceGuard {  // the handle's device for this call, the caller's afterwards
    int prev = -1;
    ~DeviceGuard() {
      if (prev >= 0) (void)hipSetDevice(prev);
    }
  } guard;
  CTAS_HIP_TRY(hipGetDevice(&guard.prev));
  CTAS_HIP_TRY(hipSetDevice(h->device));
  CTAS_HIP_TRY(hipStreamSynchronize(h->stream));

  // Two of everything: pinned host batches, device batches, events.
  struct Res {
    float *hin[2] = {}, *hout[2] = {}, *din[2] = {}, *dout[2] = {}, *avg = nullptr;
    hipStream_t s_in = nullptr, s_comp = nullptr, s_out = nullptr;
    hipEvent_t ev_in[2] = {}, ev_comp[2] = {}, ev_out[2] = {};
    ~Res() {
      for (hipStream_t st : {s_in, s_comp, s_out})
        if (st) (void)hipStreamSynchronize(st);
      for (int b = 0; b < 2; ++b) {
        if (hin[b]) (void)hipHostFree(hin[b]);
        if (hout[b]) (void)hipHostFree(hout[b]);
        if (din[b]) (void)hipFree(din[b]);
        if (dout[b]) (void)hipFree(dout[b]);
        for (hipEvent_t e : {ev_in[b], ev_comp[b], ev_out[b]})
          if (e) (void)hipEventDestroy(e);
      }
      if (avg) (void)hipFree(avg);
      for (hipStream_t st : {s_in, s_comp, s_out})
        if (st) (void)hipStreamDestroy(st);
    }
  } r;
  for (int b = 0; b < 2; ++b) {
    if (hipHostMalloc(reinterpret_cast<void **>(&r.hin[b]), sizeof(float) * in_elems * B) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&r.hout[b]), sizeof(float) * out_elems * B) != hipSuccess ||
        hipMalloc(&r.din[b], sizeof(float) * in_elems * B) != hipSuccess ||
        hipMalloc(&r.dout[b], sizeof(float) * out_elems * B) != hipSuccess)
      return fail(CTAS_E_NOMEM, "pipeline buffer allocation failed");
    CTAS_HIP_TRY(hipEventCreateWithFlags(&r.ev_in[b], hipEventDisableTiming));
    CTAS_HIP_TRY(hipEventCreateWithFlags(&r.ev_comp[b], hipEventDisableTiming));
    CTAS_HIP_TRY(hipEventCreateWithFlags(&r.ev_out[b], hipEventDisableTiming));
  }
  if (ring_box && hipMalloc(&r.avg, sizeof(float) * h->w * B) != hipSuccess)
    return fail(CTAS_E_NOMEM, "pipeline buffer allocation failed");
  CTAS_HIP_TRY(hipStreamCreateWithFlags(&r.s_in, hipStreamNonBlocking));
  CTAS_HIP_TRY(hipStreamCreateWithFlags(&r.s_comp, hipStreamNonBlocking));
  CTAS_HIP_TRY(hipStreamCreateWithFlags(&r.s_out, hipStreamNonBlocking));

  // Progress shared by the reader, the writer and this thread: batches read
  // into host buffers, batches whose upload is queued, downloads queued,
  // batches written. A failure anywhere sets `abort`.
  std::mutex mu;
  std::condition_variable cv;
  long filled = -1, up_queued = -1, down_queued = -1, written = -1;
  bool abort = false;
  std::string err;
  int err_code = CTAS_OK;
  auto set_fail = [&](int code, const std::string &m) {
    std::lock_guard<std::mutex> lk(mu);
    if (!abort) {
      abort = true;
      err = m;
      err_code = code;
    }
    cv.notify_all();
  };
  auto count = [&](long k) { return static_cast<int>(std::min<long>(B, n_slices - k * B)); };
  // Calls f(i) for i < n: in order on this thread, or spread over
  // io_threads threads (i = t, t + T, ...); false if any call failed.
  auto for_each = [&](int n, const std::function<int(int)> &f) {
    const int T = std::min(io_threads, n);
    if (T <= 1) {
      for (int i = 0; i < n; ++i)
        if (f(i) != 0) return false;
      return true;
    }
    std::atomic<bool> ok{true};
    std::vector<std::thread> pool;
    for (int t = 0; t < T; ++t)
      pool.emplace_back([&, t] {
        for (int i = t; i < n && ok.load(); i += T)
          if (f(i) != 0) ok = false;
      });
    for (std::thread &th : pool) th.join();
    return ok.load();
  };

  std::thread reader([&] {
    (void)hipSetDevice(h->device);
    for (long k = 0; k < n_batches; ++k) {
      const int b = static_cast<int>(k & 1);
      {
        // the upload of batch k - 2 from this host buffer is queued ...
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return abort || up_queued >= k - 2; });
        if (abort) return;
      }
      // ... and finished
      if (k >= 2 && hipEventSynchronize(r.ev_in[b]) != hipSuccess) return set_fail(CTAS_E_DEVICE, "upload failed");
      std::atomic<long> bad{-1};
      if (!for_each(count(k), [&](int i) {
            const int rc = read(user, k * B + i, r.hin[b] + in_elems * i);
            if (rc != 0) bad = k * B + i;
            return rc;
          }))
        return set_fail(CTAS_E_IO, "reading sinogram " + std::to_string(bad.load()) + " failed");
      std::lock_guard<std::mutex> lk(mu);
      filled = k;
      cv.notify_all();
    }
  });
  std::thread writer([&] {
    (void)hipSetDevice(h->device);
    for (long k = 0; k < n_batches; ++k) {
      const int b = static_cast<int>(k & 1);
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return abort || down_queued >= k; });
        if (abort) return;
      }
      if (hipEventSynchronize(r.ev_out[b]) != hipSuccess) return set_fail(CTAS_E_DEVICE, "download failed");
      std::atomic<long> bad{-1};
      if (!for_each(count(k), [&](int i) {
            const int rc = write(user, k * B + i, r.hout[b] + out_elems * i);
            if (rc != 0) bad = k * B + i;
            return rc;
          }))
        return set_fail(CTAS_E_IO, "writing slice " + std::to_string(bad.load()) + " failed");
      std::lock_guard<std::mutex> lk(mu);
      written = k;
      cv.notify_all();
    }
  });

  auto run = [&]() -> int {
    for (long k = 0; k < n_batches; ++k) {
      const int b = static_cast<int>(k & 1), n = count(k);
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return abort || filled >= k; });
        if (abort) return CTAS_OK;
      }
      // upload once the compute that last read this device buffer is done
      if (k >= 2) CTAS_HIP_TRY(hipStreamWaitEvent(r.s_in, r.ev_comp[b], 0));
      CTAS_HIP_TRY(hipMemcpyAsync(r.din[b], r.hin[b], sizeof(float) * in_elems * n, hipMemcpyHostToDevice, r.s_in));
      CTAS_HIP_TRY(hipEventRecord(r.ev_in[b], r.s_in));
      {
        std::lock_guard<std::mutex> lk(mu);
        up_queued = k;
        cv.notify_all();
      }
      // the previous batch's compute is done: its deAbs status can be read
      // before this batch overwrites it
      CTAS_HIP_TRY(hipStreamSynchronize(r.s_comp));
      if (k >= 1)
        if (int rc = read_status(h, count(k - 1))) return rc;
      CTAS_HIP_TRY(hipStreamWaitEvent(r.s_comp, r.ev_in[b], 0));
      if (k >= 2) CTAS_HIP_TRY(hipStreamWaitEvent(r.s_comp, r.ev_out[b], 0));  // d_out[b] downloaded
      if (ring_box) CTAS_HIP_TRY(launch_ring(r.din[b], r.avg, n, h->N, h->w, ring_box, r.s_comp));
      if (int rc = enqueue_slices(h, r.din[b], r.dout[b], nullptr, centers + k * B, n, true, r.s_comp)) return rc;
      CTAS_HIP_TRY(hipEventRecord(r.ev_comp[b], r.s_comp));
      {
        // the writer is done with the host buffer of batch k - 2
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return abort || written >= k - 2; });
        if (abort) return CTAS_OK;
      }
      CTAS_HIP_TRY(hipStreamWaitEvent(r.s_out, r.ev_comp[b], 0));
      CTAS_HIP_TRY(hipMemcpyAsync(r.hout[b], r.dout[b], sizeof(float) * out_elems * n, hipMemcpyDeviceToHost, r.s_out));
      CTAS_HIP_TRY(hipEventRecord(r.ev_out[b], r.s_out));
      std::lock_guard<std::mutex> lk(mu);
      down_queued = k;
      cv.notify_all();
    }
    CTAS_HIP_TRY(hipStreamSynchronize(r.s_comp));
    return read_status(h, count(n_batches - 1));
  };
  const int rc = run();
  if (rc != CTAS_OK) set_fail(rc, ctas_last_error());
  reader.join();
  writer.join();
  if (abort) return fail(err_code, err);
  h->kept_slices = 0;  // the filtered sinograms of the last batch are not kept for repeat()
  return CTAS_OK;
}

int ctas_ring_filter(float *sino, int N, int w, int box) {
  if (!sino || N <= 0 || w <= 0 || box < 0) return fail(CTAS_E_ARG, "bad arguments");
  if (box == 0) return CTAS_OK;
  if (int rc = check_ring_box(box, w)) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(CTAS_E_DEVICE, "no HIP device is available");
  float *d = nullptr, *avg = nullptr;
  if (hipMalloc(&d, sizeof(float) * N * w) != hipSuccess || hipMalloc(&avg, sizeof(float) * w) != hipSuccess) {
    if (d) (void)hipFree(d);
    return fail(CTAS_E_NOMEM, "allocation failed");
  }
  int rc = CTAS_OK;
  if (hipMemcpy(d, sino, sizeof(float) * N * w, hipMemcpyHostToDevice) != hipSuccess ||
      launch_ring(d, avg, 1, N, w, box, nullptr) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(sino, d, sizeof(float) * N * w, hipMemcpyDeviceToHost) != hipSuccess)
    rc = fail(CTAS_E_DEVICE, "ring filter failed on the device");
  (void)hipFree(d);
  (void)hipFree(avg);
  return rc;
}

int ctas_ring_filter_dev(float *d_sino, int S, int N, int w, int box, void *stream) {
  if (!d_sino || S <= 0 || N <= 0 || w <= 0 || box < 0) return fail(CTAS_E_ARG, "bad arguments");
  if (box == 0) return CTAS_OK;
  if (int rc = check_ring_box(box, w)) return rc;
  // The column means need S*w floats of scratch; it is freed after the
  // stream drains.
  float *avg = nullptr;
  if (hipMalloc(&avg, sizeof(float) * S * w) != hipSuccess) return fail(CTAS_E_NOMEM, "allocation failed");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const hipError_t e = launch_ring(d_sino, avg, S, N, w, box, s);
  (void)hipStreamSynchronize(s);
  (void)hipFree(avg);
  if (e != hipSuccess) return fail(CTAS_E_DEVICE, std::string("ring filter: ") + hipGetErrorString(e));
  return CTAS_OK;
}

}  // extern "C"
