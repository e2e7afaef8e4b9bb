// C++ reconstruction API over the C ABI (see ct.h).
#include "ct.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>

namespace ctas {

void throw_error(const std::string &module, const std::string &msg) {
  // common.world.cpp:132-170: the message goes to stderr, then the throw.
  std::cerr << "ERROR! In module '" << module << "'. " << msg << std::endl;
  throw CtasErr("In module '" + module + "'. " + msg);
}

void warn(const std::string &module, const std::string &msg) {
  std::cerr << "WARNING! In module '" << module << "'. " << msg << std::endl;
}

namespace {

const std::string kRecMod = "reconstruction";

void check(int rc, const std::string &module) {
  if (rc < 0) throw_error(module, ctas_last_error());
}

std::string upper(std::string s) {
  for (char &c : s) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
  return s;
}

std::string shape_str(const Shape2 &s) { return std::to_string(s[0]) + "," + std::to_string(s[1]); }

}  // namespace

Map Crop2::apply(const Map &in) const {
  const Shape2 o = shape(in.shape());
  Map out(o);
  for (long r = 0; r < o[0]; ++r)
    std::memcpy(out.data() + static_cast<size_t>(r) * o[1], in.data() + static_cast<size_t>(r + top) * o[1],
                sizeof(float) * static_cast<size_t>(o[1]));
  return out;
}

Contrast::Contrast(const std::string &name) {
  int c = 0;
  check(ctas_contrast_parse(name.c_str(), &c), "contrast type");
  c_ = static_cast<Cntype>(c);
}

std::string Contrast::name() const {
  switch (c_) {
    case REF: return "REFRACTION";
    case ABS: return "ABSORPTION";
    case PHS: return "PHASE";
    case FLT: return "FILTERED";
  }
  return "";
}

Filter::Filter(Ftype tp, float as) : tp_(tp), as_(as) {
  if (tp_ == KAISER && as_ < 0.0f)
    throw_error("filter", "Alpha parameter of Kaiser filter must be non-negative (set " + std::to_string(as_) + "). ");
  if (tp_ == GAUSS && as_ <= 0.0f)
    throw_error("filter", "Sigma parameter of Guass filter must be strictly positive (set " + std::to_string(as_) + "). ");
}

Filter::Filter(const std::string &name, float as) : tp_(RAMP), as_(as) {
  int f = 0;
  float p = 0.0f;
  // name only; the parameter is the second argument (Filter(string, float))
  const std::string spec = upper(name) == "KAISER" || upper(name) == "GAUSS" ? name + ":" + std::to_string(as) : name;
  check(ctas_filter_parse(spec.c_str(), &f, &p), "filter");
  tp_ = static_cast<Ftype>(f);
}

std::string Filter::name() const {
  static const char *names[] = {"NONE", "RAMP", "BARLETT", "WELCH", "PARZEN", "HANN",
                                "HAMMING", "BLACKMAN", "LANCKZOS", "KAISER", "GAUSS"};
  return names[tp_];
}

Line &Filter::fill(Line &filt, int pixels) const {
  if (!pixels) pixels = static_cast<int>(filt.size());
  if (!pixels) throw_error("filter", "Zero-size filter requested.");
  filt.assign(static_cast<size_t>(pixels), 0.0f);
  check(ctas_filter_fill(tp_, as_, pixels, filt.data()), "filter");
  return filt;
}

bool _conversion(Filter *val, const std::string &in) {
  int f = 0;
  float p = 0.0f;
  if (ctas_filter_parse(in.c_str(), &f, &p) < 0) return false;
  *val = Filter(static_cast<Filter::Ftype>(f), p);
  return true;
}

CTrec::CTrec(const Shape2 &sinoshape, Contrast cn, float anarc, const Filter &ft, int flags) : ish_(sinoshape) {
  check(ctas_fbp_create(&h_, static_cast<int>(ish_[1]), static_cast<int>(ish_[0]), cn.contrast(), ft.filter(),
                        ft.param(), anarc, flags),
        kRecMod);
}

CTrec::CTrec(CTrec &other) : ish_(other.ish_) { check(ctas_fbp_clone(other.h_, &h_), kRecMod); }

CTrec::~CTrec() {
  if (h_) ctas_fbp_destroy(h_);
}

Shape2 CTrec::recShape(float center) const {
  int rows = 0, cols = 0;
  check(ctas_fbp_rec_shape(h_, center, &rows, &cols), kRecMod);
  return {rows, cols};
}

Crop2 CTrec::recCrop(float center) const {
  int b = 0, e = 0;
  check(ctas_fbp_rec_crop(h_, center, &b, &e), kRecMod);
  return Crop2{b, ish_[1] - e};
}

void CTrec::setPhysics(float pixelSize, float lambda) { check(ctas_fbp_set_physics(h_, pixelSize, lambda), kRecMod); }

void CTrec::sino(Map &sinogram) {
  if (sinogram.shape() != ish_)
    throw_error(kRecMod, "Shape of input sinogram (" + shape_str(sinogram.shape()) + ") does not match initial (" +
                             shape_str(ish_) + ").");
  check(ctas_fbp_sino(h_, sinogram.data()), kRecMod);
}

void CTrec::repeat(Map &slice, float center) {
  if (slice.shape() != Shape2{ish_[1], ish_[1]}) slice.resize(ish_[1], ish_[1]);
  check(ctas_fbp_repeat(h_, slice.data(), &center, 1), kRecMod);
}

void CTrec::reconstruct(Map &sinogram, Map &slice, float center) {
  if (sinogram.shape() != ish_)
    throw_error(kRecMod, "Wrong sinogram shape (" + shape_str(sinogram.shape()) + ") where (" + shape_str(ish_) +
                             ") is expected.");
  if (slice.shape() != Shape2{ish_[1], ish_[1]}) slice.resize(ish_[1], ish_[1]);
  check(ctas_fbp_reconstruct(h_, sinogram.data(), slice.data(), &center, 1), kRecMod);
}

void CTrec::reconstruct(std::vector<Map> &sinograms, std::vector<Map> &slices, const std::vector<float> &centers) {
  const size_t S = sinograms.size();
  if (centers.size() != S) throw_error(kRecMod, "one centre per sinogram is required");
  const size_t in = static_cast<size_t>(ish_[0]) * ish_[1], out = static_cast<size_t>(ish_[1]) * ish_[1];
  std::vector<float> sin(in * S), sl(out * S);
  for (size_t s = 0; s < S; ++s) {
    if (sinograms[s].shape() != ish_) throw_error(kRecMod, "Wrong sinogram shape (" + shape_str(sinograms[s].shape()) + ").");
    std::memcpy(sin.data() + in * s, sinograms[s].data(), sizeof(float) * in);
  }
  check(ctas_fbp_reconstruct(h_, sin.data(), sl.data(), centers.data(), static_cast<int>(S)), kRecMod);
  slices.resize(S);
  for (size_t s = 0; s < S; ++s) {
    std::memcpy(sinograms[s].data(), sin.data() + in * s, sizeof(float) * in);
    slices[s].resize(ish_[1], ish_[1]);
    std::memcpy(slices[s].data(), sl.data() + out * s, sizeof(float) * out);
  }
}

void RingFilter::apply(Map &sinogram) {
  if (!box_) return;
  if (!sinogram.rows()) throw_error("RingFilter", "Empty sinogram.");
  if (static_cast<size_t>(sinogram.columns()) != width_)
    throw_error("RingFilter", "Wrong sinogram width " + std::to_string(sinogram.columns()) + " where " +
                                  std::to_string(width_) + " is expected.");
  check(ctas_ring_filter(sinogram.data(), static_cast<int>(sinogram.rows()), static_cast<int>(sinogram.columns()),
                         static_cast<int>(box_)),
        "RingFilter");
}

void ts_add(Map &projection, Map &result, const Filter &filter, const float center, const Contrast contrast,
            const float angle, const int plane) {
  if (result.columns() != projection.columns() || result.rows() != projection.rows())
    throw_error("tomosynthesis", "The size of the result array does not match the requested reconstruction.");
  check(ctas_ts_add(projection.data(), static_cast<int>(projection.rows()), static_cast<int>(projection.columns()),
                    result.data(), static_cast<int>(result.rows()), static_cast<int>(result.columns()),
                    filter.filter(), filter.param(), center, contrast.contrast(), angle, plane),
        "tomosynthesis");
}

void Denoiser::proc(Map &iom) const {
  if (!sh_[0] || !sh_[1] || !rad_) return;
  if (iom.shape() != sh_) throw_error("denoiser", "Non matching shape of input array.");
  check(ctas_denoise(iom.data(), static_cast<int>(sh_[0]), static_cast<int>(sh_[1]), rad_, thr_), "denoiser");
}

float raxis(Map &sino, float anarc, float maxDev, int algo) {
  float r = 0.0f;
  check(ctas_raxis(sino.data(), static_cast<int>(sino.rows()), static_cast<int>(sino.columns()), anarc, maxDev, algo,
                   &r),
        "Axis shift");
  return r;
}

float raxis(Map &proj0, Map &proj180, float maxDev) {
  if (proj0.shape() != proj180.shape())
    throw_error("Axis shift", "Different shapes of input projections: (" + shape_str(proj0.shape()) + ") and (" +
                                  shape_str(proj180.shape()) + ") .");
  float r = 0.0f;
  check(ctas_raxis_pair(proj0.data(), proj180.data(), static_cast<int>(proj0.rows()),
                        static_cast<int>(proj0.columns()), maxDev, &r),
        "Axis shift");
  return r;
}

}  // namespace ctas
