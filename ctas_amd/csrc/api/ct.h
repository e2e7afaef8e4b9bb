// C++ reconstruction API of CTAS, over the C ABI of the MI355X engine.
//
// Same class and function names, argument meaning and error behaviour as
// the reference's src/common/ct.h (Filter ct.h:87-162, CTrec ct.h:235-298,
// RingFilter ct.h:215-229, raxis ct.h:327-329) and physical.world.h
// (Contrast :18-43, deAbs). Arrays are the minimal row-major float types
// below instead of blitz arrays. Errors print "ERROR! In module '<m>'. <msg>"
// to stderr and throw CtasErr, as throw_error does
// (common.world.cpp:132-170). There is no CPU reconstruction path: a
// handle that cannot reach a HIP device throws.
#pragma once

#include <array>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "ctas_hip.h"

namespace ctas {

class CtasErr : public std::runtime_error {
 public:
  explicit CtasErr(const std::string &m) : std::runtime_error(m) {}
};

[[noreturn]] void throw_error(const std::string &module, const std::string &msg);
void warn(const std::string &module, const std::string &msg);

using Shape2 = std::array<long, 2>;  // (rows, columns)

// Row-major float32 image (blitz::Array<float,2> in the reference).
class Map {
 public:
  Map() = default;
  Map(long rows, long cols) { resize(rows, cols); }
  explicit Map(const Shape2 &s) { resize(s[0], s[1]); }
  void resize(long rows, long cols) {
    rows_ = rows;
    cols_ = cols;
    data_.assign(static_cast<size_t>(rows) * cols, 0.0f);
  }
  void resize(const Shape2 &s) { resize(s[0], s[1]); }
  Shape2 shape() const { return {rows_, cols_}; }
  long rows() const { return rows_; }
  long columns() const { return cols_; }
  size_t size() const { return data_.size(); }
  float *data() { return data_.data(); }
  const float *data() const { return data_.data(); }
  float &operator()(long r, long c) { return data_[static_cast<size_t>(r) * cols_ + c]; }
  float operator()(long r, long c) const { return data_[static_cast<size_t>(r) * cols_ + c]; }

 private:
  long rows_ = 0, cols_ = 0;
  std::vector<float> data_;
};

using Line = std::vector<float>;

// Crop of rows [top, rows - bottom) (Crop<2> with Segment(r, -r)).
struct Crop2 {
  long top = 0, bottom = 0;
  Shape2 shape(const Shape2 &in) const { return {in[0] - top - bottom, in[1]}; }
  Map apply(const Map &in) const;
};

class Contrast {
 public:
  enum Cntype { ABS = CTAS_CONTRAST_ABS, REF = CTAS_CONTRAST_REF, PHS = CTAS_CONTRAST_PHS, FLT = CTAS_CONTRAST_FLT };
  Contrast(Cntype c = FLT) : c_(c) {}
  explicit Contrast(const std::string &name);
  Cntype contrast() const { return c_; }
  std::string name() const;

 private:
  Cntype c_;
};
inline bool operator==(const Contrast &a, const Contrast &b) { return a.contrast() == b.contrast(); }
inline bool operator!=(const Contrast &a, const Contrast &b) { return a.contrast() != b.contrast(); }

class Filter {
 public:
  enum Ftype {
    NONE = CTAS_FILTER_NONE, RAMP, BARLETT, WELCH, PARZEN, HANN, HAMMING, BLACKMAN, LANCKZOS, KAISER, GAUSS
  };
  Filter(Ftype tp = RAMP, float as = 0);
  explicit Filter(const std::string &name, float as = 0);
  Ftype filter() const { return tp_; }
  float param() const { return as_; }
  std::string name() const;
  // Filter::fill (lct.cpp:122-209): the window of a `pixels`-point line
  // (the line's own size when 0), without the CTrec gain.
  Line &fill(Line &filt, int pixels = 0) const;

 private:
  Ftype tp_;
  float as_;
};
inline bool operator==(const Filter &a, const Filter &b) { return a.filter() == b.filter(); }
inline bool operator!=(const Filter &a, const Filter &b) { return a.filter() != b.filter(); }
// "NAME[:FLOAT]" (lct.cpp:236-248)
bool _conversion(Filter *val, const std::string &in);

// Sampling and storage of the engine (no counterpart in the reference):
// CTAS floor sampling in float32 is the default and reproduces it.
enum EngineFlags {
  INTERP_CTAS = CTAS_INTERP_CTAS,
  INTERP_LINEAR = CTAS_INTERP_LINEAR,
  STORE_FP32 = CTAS_STORE_FP32,
  STORE_FP16 = CTAS_STORE_FP16,
  STORE_BF16 = CTAS_STORE_BF16,
};

class CTrec {
 public:
  CTrec(const Shape2 &sinoshape, Contrast cn, float anarc = 180, const Filter &ft = Filter(), int flags = 0);
  CTrec(CTrec &other);
  ~CTrec();
  CTrec &operator=(const CTrec &) = delete;

  Shape2 recShape(float center = 0.0f) const;
  Crop2 recCrop(float center = 0.0f) const;

  void setPhysics(float pixelSize = 0.0f, float lambda = 0.0f);
  void sino(Map &sinogram);
  void repeat(Map &slice, float center);
  void reconstruct(Map &sinogram, Map &slice, float center);
  // S slices in one launch: sinograms filtered in place.
  void reconstruct(std::vector<Map> &sinograms, std::vector<Map> &slices, const std::vector<float> &centers);

  // Kept for source compatibility with the reference's thread/device
  // arbitration; every reconstruction runs on the GPU.
  void doOnGPU() {}
  void allowCPU() {}

  static void reconstruct(Map &sinogram, Map &slice, Contrast cn, float arc = 180, const Filter &ft = Filter(),
                          float center = 0, float pixelSize = 0.0f, float lambda = 0.0f) {
    CTrec rec(sinogram.shape(), cn, arc, ft);
    rec.setPhysics(pixelSize, lambda);
    rec.reconstruct(sinogram, slice, center);
  }

  ctas_fbp *handle() { return h_; }

 private:
  Shape2 ish_;
  ctas_fbp *h_ = nullptr;
};

class RingFilter {
 public:
  RingFilter(size_t width, size_t ringBox = 0) : width_(width), box_(ringBox) {}
  void apply(Map &sinogram);

 private:
  size_t width_, box_;
};

// ts_add (ct.h:317, lct.cpp:739-787): projection is replaced by its
// filtered rows; result accumulates the projection resampled for the plane.
void ts_add(Map &projection, Map &result, const Filter &filter, const float center, const Contrast contrast,
            const float angle, const int plane);

// Denoiser (matrix.world.h:468-481) without the mask: proc() replaces
// pixels by their disc mean under the threshold rule, on the GPU.
class Denoiser {
 public:
  Denoiser(const Shape2 &sh, int rad, float threshold) : sh_(sh), rad_(rad < 0 ? -rad : rad), thr_(threshold) {}
  void proc(Map &iom) const;

 private:
  Shape2 sh_;
  int rad_;
  float thr_;
};

float raxis(Map &sino, float anarc, float maxDev, int algo = 0);
float raxis(Map &proj0, Map &proj180, float maxDev);

}  // namespace ctas
