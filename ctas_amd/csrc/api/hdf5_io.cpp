// HDF5 volumes as slice sources of the tools: "file:dataset[:[xyz]slices]"
// (the reference's ImagePath + HDFwrapper, external.world.cpp:104-660 and
// external.world.h:14-50). The HDF5 C library is loaded at run time
// (dlopen of libhdf5.so, $CTAS_HDF5_LIB first), so the tools build and run
// without it; a HDF5 path then fails with a message. The few types and
// constants used are those of the library's 1.8-1.14 ABI (hid_t 64 bit).
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "ct.h"
#include "image_io.h"

namespace ctas {
namespace {

using hid_t = int64_t;
using herr_t = int;
using htri_t = int;
using hsize_t = unsigned long long;
constexpr unsigned kAccRdonly = 0x0000u, kAccTrunc = 0x0002u;
constexpr hid_t kDefault = 0, kAll = 0;
constexpr int kSelectSet = 0;

struct H5 {
  void *so = nullptr;
  herr_t (*open)();
  herr_t (*eset_auto)(hid_t, void *, void *);
  htri_t (*fis_hdf5)(const char *);
  hid_t (*fopen)(const char *, unsigned, hid_t);
  hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
  herr_t (*fclose)(hid_t);
  hid_t (*dopen)(hid_t, const char *, hid_t);
  hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
  herr_t (*dclose)(hid_t);
  hid_t (*dget_space)(hid_t);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
  int (*sget_ndims)(hid_t);
  int (*sget_dims)(hid_t, hsize_t *, hsize_t *);
  hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
  herr_t (*sselect_hyperslab)(hid_t, int, const hsize_t *, const hsize_t *, const hsize_t *, const hsize_t *);
  herr_t (*sclose)(hid_t);
  hid_t *native_float = nullptr;
  std::string error;
};

// The HDF5 library is not assumed to be thread-safe: every call holds this.
std::mutex g_h5_mu;

template <typename F>
bool sym(void *so, const char *name, F &f) {
  f = reinterpret_cast<F>(dlsym(so, name));
  return f != nullptr;
}

H5 &h5() {
  static H5 lib = [] {
    H5 l{};
    std::vector<std::string> names;
    if (const char *e = std::getenv("CTAS_HDF5_LIB")) names.push_back(e);
    names.insert(names.end(), {"libhdf5.so", "libhdf5.so.103", "libhdf5_serial.so", "/opt/conda/lib/libhdf5.so"});
    for (const auto &n : names)
      if ((l.so = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
    if (!l.so) {
      l.error = "the HDF5 library (libhdf5.so) is not available; set CTAS_HDF5_LIB to its path";
      return l;
    }
    const bool ok = sym(l.so, "H5open", l.open) && sym(l.so, "H5Eset_auto2", l.eset_auto) &&
                    sym(l.so, "H5Fis_hdf5", l.fis_hdf5) && sym(l.so, "H5Fopen", l.fopen) &&
                    sym(l.so, "H5Fcreate", l.fcreate) && sym(l.so, "H5Fclose", l.fclose) &&
                    sym(l.so, "H5Dopen2", l.dopen) && sym(l.so, "H5Dcreate2", l.dcreate) &&
                    sym(l.so, "H5Dclose", l.dclose) && sym(l.so, "H5Dget_space", l.dget_space) &&
                    sym(l.so, "H5Dread", l.dread) && sym(l.so, "H5Dwrite", l.dwrite) &&
                    sym(l.so, "H5Sget_simple_extent_ndims", l.sget_ndims) &&
                    sym(l.so, "H5Sget_simple_extent_dims", l.sget_dims) &&
                    sym(l.so, "H5Screate_simple", l.screate_simple) &&
                    sym(l.so, "H5Sselect_hyperslab", l.sselect_hyperslab) && sym(l.so, "H5Sclose", l.sclose);
    l.native_float = reinterpret_cast<hid_t *>(dlsym(l.so, "H5T_NATIVE_FLOAT_g"));
    if (!ok || !l.native_float || l.open() < 0) {
      l.error = "the HDF5 library does not provide the functions it needs";
      dlclose(l.so);
      l.so = nullptr;
      return l;
    }
    l.eset_auto(kDefault, nullptr, nullptr);  // no error stack printed on stderr
    return l;
  }();
  return lib;
}

H5 &need_h5() {
  H5 &l = h5();
  if (!l.so) throw_error("HDF i/o", l.error);
  return l;
}

// "file:desc" -> file, desc (ImagePath::splitIn, external.world.h:19-22).
std::pair<std::string, std::string> split_path(const std::string &spec) {
  const size_t p = std::min(spec.find(':'), spec.size());
  return {spec.substr(0, p), spec.substr(p)};
}

std::vector<std::string> split_colons(const std::string &s) {
  std::vector<std::string> out;
  size_t b = 0;
  for (;;) {
    const size_t e = s.find(':', b);
    out.push_back(s.substr(b, e == std::string::npos ? std::string::npos : e - b));
    if (e == std::string::npos) break;
    b = e + 1;
  }
  return out;
}

// An open 3-D (or 2-D) dataset with the planes that the descriptor selects.
struct Volume {
  hid_t file = 0, dset = 0, space = 0;
  int rank = 0, dim = 0;  // dim: the axis perpendicular to the planes (0 z, 1 y, 2 x)
  hsize_t cnt[3] = {1, 1, 1};
  std::vector<long> planes;
  Shape2 face{0, 0};
  ~Volume() {
    H5 &l = h5();
    if (!l.so) return;
    if (space > 0) l.sclose(space);
    if (dset > 0) l.dclose(dset);
    if (file > 0) l.fclose(file);
  }
};

std::map<std::string, std::unique_ptr<Volume>> g_volumes;

Volume &open_volume(const std::string &spec) {
  auto f = g_volumes.find(spec);
  if (f != g_volumes.end()) return *f->second;
  H5 &l = need_h5();
  const auto [file, desc] = split_path(spec);
  std::vector<std::string> comps = split_colons(desc.empty() ? std::string() : desc.substr(1));
  if (comps.empty() || comps[0].empty()) throw_error("HDF i/o", "Invalid HDF5 path descriptor \"" + spec + "\"");
  auto v = std::make_unique<Volume>();
  // [xyz] of the slice string selects the axis perpendicular to the planes
  // (default z, the first); the rest are the slices (external.world.cpp:349-376)
  std::string slices;
  char axis = 'z';
  if (comps.size() > 1) {
    for (size_t i = 1; i < comps.size(); ++i) slices += ":" + comps[i];
    slices.erase(0, 1);
    if (!comps[1].empty()) axis = comps[1][0];
    if (!slices.empty() && std::string("xXyYzZ").find(slices[0]) != std::string::npos) slices.erase(0, 1);
  }
  v->dim = (axis == 'x' || axis == 'X') ? 2 : (axis == 'y' || axis == 'Y') ? 1 : 0;
  if (l.fis_hdf5(file.c_str()) <= 0) throw_error("HDF i/o", "No HDF file at " + file + ".");
  v->file = l.fopen(file.c_str(), kAccRdonly, kDefault);
  if (v->file <= 0 || (v->dset = l.dopen(v->file, comps[0].c_str(), kDefault)) < 0 ||
      (v->space = l.dget_space(v->dset)) < 0 || (v->rank = l.sget_ndims(v->space)) < 0)
    throw_error("HDF i/o", "Failed to open data " + spec);
  if (v->rank != 2 && v->rank != 3) throw_error("HDF i/o", "Dataset is not 2D or 3D in " + spec);
  if (l.sget_dims(v->space, v->cnt, nullptr) != v->rank) throw_error("HDF i/o", "Failed to read dataset size in " + spec);
  if (v->rank == 2) {
    v->face = {static_cast<long>(v->cnt[0]), static_cast<long>(v->cnt[1])};
    v->planes = {0};
  } else {
    long a = 0, b = 0;
    for (int d = 0, o = 0; d < 3; ++d)
      if (d != v->dim) (o++ ? b : a) = static_cast<long>(v->cnt[d]);
    // planes across x are read in the YZ plane and transposed
    v->face = v->dim == 2 ? Shape2{b, a} : Shape2{a, b};
    v->planes = slice_list(slices, static_cast<long>(v->cnt[v->dim]));
    for (long p : v->planes)
      if (p < 0 || p >= static_cast<long>(v->cnt[v->dim]))
        throw_error("HDF i/o", "Slice " + std::to_string(p) + " is outside the dataset in " + spec);
  }
  Volume &ref = *v;
  g_volumes.emplace(spec, std::move(v));
  return ref;
}

}  // namespace

// slice_str2vec (common.world.cpp:471-612): comma-separated ranges
// start[:stop[:step]] in numpy's manner, negative indices counted from the
// end, "start+n" / "start-n" for stop = start +- n; an empty string selects
// everything.
std::vector<long> slice_list(const std::string &spec, long height) {
  std::vector<long> out;
  if (spec.empty()) {
    for (long i = 0; i < std::labs(height); ++i) out.push_back(i);
    return out;
  }
  struct Range {
    long start = 0, stop = 0, step = 0;
    bool empty_start = false, empty_stop = false;
  };
  auto int_end = [](const std::string &s) {
    size_t p = s.find_first_not_of("+-");
    if (p == std::string::npos) return std::string::npos;
    return s.find_first_not_of("0123456789", p);
  };
  auto to_long = [&](const std::string &s, const std::string &whole) {
    char *e = nullptr;
    const long v = std::strtol(s.c_str(), &e, 10);
    if (s.empty() || *e) throw_error("slice string", "Range string \"" + whole + "\" must start with an integer.");
    return v;
  };
  long minsz = height;
  std::vector<Range> ranges;
  size_t b = 0;
  for (;;) {
    const size_t e = spec.find(',', b);
    const std::string rd = spec.substr(b, e == std::string::npos ? std::string::npos : e - b);
    Range r;
    std::string rem = rd.empty() ? std::string(":") : rd;
    size_t end;
    r.empty_start = rem[0] == ':';
    if (r.empty_start) {
      end = 0;
    } else {
      end = int_end(rem);
      if (end == 0) throw_error("slice string", "Range string \"" + rem + "\" must start with an integer.");
      r.start = to_long(rem.substr(0, end), rd);
    }
    minsz = std::max({minsz, -r.start, r.start + 1});
    rem = end == std::string::npos ? std::string() : rem.substr(end);
    if (rem.empty()) {
      r.stop = r.start + (r.start < 0 ? -1 : 1);
    } else {
      const char link = rem[0];
      if (std::string("+-:").find(link) == std::string::npos)
        throw_error("slice string", "Expected [+-:] in range string \"" + rd + "\".");
      rem = rem.substr(1);
      end = int_end(rem);
      std::string stop = rem.substr(0, end);
      r.empty_stop = stop.empty();
      if (r.empty_stop) {
        r.stop = minsz;
        r.step = link == '-' ? -1 : 1;
      } else {
        if (link != ':') stop = std::string(1, link) + stop;
        r.stop = to_long(stop, rd);
        if (!r.empty_start && link != ':') r.stop += r.start;
      }
      minsz = std::max({minsz, -r.stop, r.stop});
      rem = end == std::string::npos ? std::string() : rem.substr(end);
      if (!rem.empty()) {
        if (rem[0] != ':') throw_error("slice string", "Expected ':' in range string \"" + rd + "\".");
        rem = rem.substr(1);
      }
      if (!rem.empty()) {
        end = int_end(rem);
        if (end == 0) throw_error("slice string", "Can't parse integer in range string \"" + rd + "\".");
        r.step = to_long(rem.substr(0, end), rd);
        if (r.step == 0) throw_error("slice string", "Zero step in range string \"" + rd + "\".");
        if ((r.stop - r.start) * r.step < 0)
          throw_error("slice string",
                      "Range string \"" + rd + "\" describes no range due to the step sign opposite to the range.");
      }
    }
    ranges.push_back(r);
    if (e == std::string::npos) break;
    b = e + 1;
  }
  if (height > 0 && minsz > height) minsz = height;
  else if (height < 0) minsz = std::max(minsz, -height);
  for (Range r : ranges) {
    if (r.start < 0) r.start += minsz;
    if (r.stop < 0) r.stop += minsz;
    if (r.empty_start && r.empty_stop) {
      if (r.step < 0) {
        r.start = -minsz - 1;
        r.stop = -1;
      } else {
        r.start = 0;
        r.stop = minsz;
      }
    } else if (r.empty_stop) {
      r.stop = r.step < 0 ? -1 : minsz;
    } else if (r.empty_start) {
      r.start = r.step < 0 ? -minsz - 1 : 0;
    }
    if (!r.step) r.step = r.start < r.stop ? 1 : -1;
    for (long s = r.start; (r.stop - s) * r.step > 0; s += r.step) out.push_back(s);
  }
  return out;
}

bool is_hdf5_spec(const std::string &spec) {
  const auto [file, desc] = split_path(spec);
  if (desc.size() < 2) return false;
  H5 &l = h5();
  std::lock_guard<std::mutex> lk(g_h5_mu);
  return l.so && l.fis_hdf5(file.c_str()) > 0;
}

std::vector<long> hdf5_planes(const std::string &spec, Shape2 &face) {
  std::lock_guard<std::mutex> lk(g_h5_mu);
  Volume &v = open_volume(spec);
  face = v.face;
  return v.planes;
}

void hdf5_read_plane(const std::string &spec, long plane, float *dst) {
  std::lock_guard<std::mutex> lk(g_h5_mu);
  H5 &l = need_h5();
  Volume &v = open_volume(spec);
  hsize_t off[3] = {0, 0, 0}, cnt[3] = {v.cnt[0], v.cnt[1], v.cnt[2]};
  if (v.rank == 3) {
    off[v.dim] = static_cast<hsize_t>(plane);
    cnt[v.dim] = 1;
  }
  const hsize_t n = cnt[0] * cnt[1] * (v.rank == 3 ? cnt[2] : 1);
  std::vector<float> tmp;
  float *buf = dst;
  if (v.dim == 2 && v.rank == 3) {
    tmp.resize(n);
    buf = tmp.data();
  }
  const hsize_t mcnt = n;
  const hid_t fs = l.screate_simple(v.rank, v.cnt, nullptr);
  const hid_t ms = l.screate_simple(1, &mcnt, nullptr);
  const bool ok = fs >= 0 && ms >= 0 && l.sselect_hyperslab(fs, kSelectSet, off, nullptr, cnt, nullptr) >= 0 &&
                  l.dread(v.dset, *l.native_float, ms, fs, kDefault, buf) >= 0;
  if (fs >= 0) l.sclose(fs);
  if (ms >= 0) l.sclose(ms);
  if (!ok) throw_error("HDF i/o", "Failed to read slice " + std::to_string(plane) + " from " + spec + ".");
  if (v.dim == 2 && v.rank == 3) {  // the plane [z][y] of one x, transposed to [y][z]
    const long Z = static_cast<long>(cnt[0]), Y = static_cast<long>(cnt[1]);
    for (long z = 0; z < Z; ++z)
      for (long y = 0; y < Y; ++y) dst[y * Z + z] = tmp[z * Y + y];
  }
}

void hdf5_write_volume(const std::string &spec, const float *data, long d0, long d1, long d2) {
  std::lock_guard<std::mutex> lk(g_h5_mu);
  H5 &l = need_h5();
  const auto [file, desc] = split_path(spec);
  const std::vector<std::string> comps = split_colons(desc.empty() ? std::string() : desc.substr(1));
  if (comps.empty() || comps[0].empty()) throw_error("HDF i/o", "Invalid HDF5 path descriptor \"" + spec + "\"");
  g_volumes.erase(spec);
  const hsize_t dims[3] = {static_cast<hsize_t>(d0), static_cast<hsize_t>(d1), static_cast<hsize_t>(d2)};
  const hid_t f = l.fcreate(file.c_str(), kAccTrunc, kDefault, kDefault);
  const hid_t sp = f > 0 ? l.screate_simple(3, dims, nullptr) : -1;
  const hid_t ds = sp >= 0 ? l.dcreate(f, comps[0].c_str(), *l.native_float, sp, kDefault, kDefault, kDefault) : -1;
  const bool ok = ds >= 0 && l.dwrite(ds, *l.native_float, kAll, kAll, kDefault, data) >= 0;
  if (ds >= 0) l.dclose(ds);
  if (sp >= 0) l.sclose(sp);
  if (f > 0) l.fclose(f);
  if (!ok) throw_error("HDF i/o", "Failed to write " + spec + ".");
}

}  // namespace ctas

// Test and scripting helpers of the tools' I/O (not part of the engine's
// C ABI): write a float32 3-D dataset, list the planes a descriptor
// selects, read one. Return 0, or -1 with the message on stderr.
extern "C" int ctas_io_h5_write(const char *spec, const float *data, long d0, long d1, long d2) {
  try {
    ctas::hdf5_write_volume(spec, data, d0, d1, d2);
    return 0;
  } catch (const ctas::CtasErr &) {
    return -1;
  }
}

extern "C" int ctas_io_h5_planes(const char *spec, long *n_planes, long *rows, long *cols) {
  try {
    ctas::Shape2 face;
    const std::vector<long> p = ctas::hdf5_planes(spec, face);
    *n_planes = static_cast<long>(p.size());
    *rows = face[0];
    *cols = face[1];
    return 0;
  } catch (const ctas::CtasErr &) {
    return -1;
  }
}

extern "C" int ctas_io_h5_read(const char *spec, long index, float *dst) {
  try {
    ctas::Shape2 face;
    const std::vector<long> p = ctas::hdf5_planes(spec, face);
    if (index < 0 || index >= static_cast<long>(p.size())) return -1;
    ctas::hdf5_read_plane(spec, p[static_cast<size_t>(index)], dst);
    return 0;
  } catch (const ctas::CtasErr &) {
    return -1;
  }
}
