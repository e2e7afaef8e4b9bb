#include "image_io.h"

#include <fcntl.h>
#include <unistd.h>

#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>

namespace ctas {
namespace {

bool ends_with(const std::string &s, const std::string &e) {
  if (s.size() < e.size()) return false;
  std::string t = s.substr(s.size() - e.size());
  for (char &c : t) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return t == e;
}

std::vector<unsigned char> slurp(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw_error("image", "Could not open file \"" + path + "\".");
  return std::vector<unsigned char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

struct Reader {
  const std::vector<unsigned char> &b;
  bool be;
  uint32_t u16(size_t o) const {
    if (o + 2 > b.size()) throw_error("image", "truncated TIFF");
    return be ? (b[o] << 8 | b[o + 1]) : (b[o] | b[o + 1] << 8);
  }
  uint32_t u32(size_t o) const {
    if (o + 4 > b.size()) throw_error("image", "truncated TIFF");
    return be ? (uint32_t(b[o]) << 24 | uint32_t(b[o + 1]) << 16 | uint32_t(b[o + 2]) << 8 | b[o + 3])
              : (b[o] | uint32_t(b[o + 1]) << 8 | uint32_t(b[o + 2]) << 16 | uint32_t(b[o + 3]) << 24);
  }
};

Map read_tiff(const std::string &path) {
  const std::vector<unsigned char> b = slurp(path);
  if (b.size() < 8 || !((b[0] == 'I' && b[1] == 'I') || (b[0] == 'M' && b[1] == 'M')))
    throw_error("image", "\"" + path + "\" is not a TIFF file.");
  Reader r{b, b[0] == 'M'};
  if (r.u16(2) != 42) throw_error("image", "\"" + path + "\": unsupported TIFF variant.");
  const size_t ifd = r.u32(4);
  const uint32_t n = r.u16(ifd);
  uint32_t width = 0, height = 0, bps = 0, comp = 1, spp = 1, fmt = 1, rps = 0;
  std::vector<uint32_t> offs, counts;
  auto values = [&](size_t e, std::vector<uint32_t> &out) {
    const uint32_t type = r.u16(e + 2), cnt = r.u32(e + 4);
    const size_t sz = type == 3 ? 2 : 4;
    size_t at = cnt * sz <= 4 ? e + 8 : r.u32(e + 8);
    out.resize(cnt);
    for (uint32_t k = 0; k < cnt; ++k) out[k] = sz == 2 ? r.u16(at + 2 * k) : r.u32(at + 4 * k);
  };
  for (uint32_t k = 0; k < n; ++k) {
    const size_t e = ifd + 2 + 12 * k;
    const uint32_t tag = r.u16(e);
    std::vector<uint32_t> v;
    values(e, v);
    if (v.empty()) continue;
    switch (tag) {
      case 256: width = v[0]; break;
      case 257: height = v[0]; break;
      case 258: bps = v[0]; break;
      case 259: comp = v[0]; break;
      case 273: offs = v; break;
      case 277: spp = v[0]; break;
      case 278: rps = v[0]; break;
      case 279: counts = v; break;
      case 339: fmt = v[0]; break;
      default: break;
    }
  }
  if (bps != 32 || fmt != 3 || comp != 1 || spp != 1)
    throw_error("image", "\"" + path + "\": only uncompressed single-channel float32 TIFF is supported.");
  (void)rps;
  Map img(height, width);
  const size_t need = static_cast<size_t>(width) * height * 4;
  std::vector<unsigned char> raw;
  raw.reserve(need);
  for (size_t s = 0; s < offs.size() && raw.size() < need; ++s) {
    const size_t len = s < counts.size() ? counts[s] : need - raw.size();
    if (offs[s] + len > b.size()) throw_error("image", "\"" + path + "\": truncated strip.");
    raw.insert(raw.end(), b.begin() + offs[s], b.begin() + offs[s] + len);
  }
  if (raw.size() < need) throw_error("image", "\"" + path + "\": missing image data.");
  for (size_t i = 0; i < static_cast<size_t>(width) * height; ++i) {
    uint32_t u = r.be ? (uint32_t(raw[4 * i]) << 24 | uint32_t(raw[4 * i + 1]) << 16 | uint32_t(raw[4 * i + 2]) << 8 | raw[4 * i + 3])
                      : (raw[4 * i] | uint32_t(raw[4 * i + 1]) << 8 | uint32_t(raw[4 * i + 2]) << 16 | uint32_t(raw[4 * i + 3]) << 24);
    std::memcpy(img.data() + i, &u, 4);
  }
  return img;
}

// Shape (planes, rows, columns) and data offset of a float32 .npy file,
// from its header alone.
struct NpyInfo {
  long dims[3];
  size_t data;
};

NpyInfo npy_info(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw_error("image", "Could not open file \"" + path + "\".");
  unsigned char b[12] = {};
  f.read(reinterpret_cast<char *>(b), 12);
  if (f.gcount() < 10 || std::memcmp(b, "\x93NUMPY", 6) != 0) throw_error("image", "\"" + path + "\" is not a .npy file.");
  const int major = b[6];
  size_t hlen, hoff;
  if (major == 1) {
    hlen = b[8] | (b[9] << 8);
    hoff = 10;
  } else {
    hlen = b[8] | (b[9] << 8) | (size_t(b[10]) << 16) | (size_t(b[11]) << 24);
    hoff = 12;
  }
  std::string h(hlen, ' ');
  f.seekg(static_cast<std::streamoff>(hoff));
  f.read(&h[0], static_cast<std::streamsize>(hlen));
  if (static_cast<size_t>(f.gcount()) != hlen) throw_error("image", "\"" + path + "\": truncated header.");
  if (h.find("'<f4'") == std::string::npos) throw_error("image", "\"" + path + "\": only little-endian float32 .npy is supported.");
  if (h.find("'fortran_order': True") != std::string::npos) throw_error("image", "\"" + path + "\": Fortran order is not supported.");
  const size_t a = h.find('('), z = h.find(')');
  std::vector<long> dims;
  std::stringstream ss(h.substr(a + 1, z - a - 1));
  std::string tok;
  while (std::getline(ss, tok, ','))
    if (tok.find_first_not_of(" ") != std::string::npos) dims.push_back(std::stol(tok));
  if (dims.size() == 2) dims.insert(dims.begin(), 1);
  if (dims.size() != 3) throw_error("image", "\"" + path + "\": expected a 2-D or 3-D array.");
  f.seekg(0, std::ios::end);
  const size_t size = static_cast<size_t>(f.tellg());
  NpyInfo info{{dims[0], dims[1], dims[2]}, hoff + hlen};
  if (info.data + static_cast<size_t>(dims[0]) * dims[1] * dims[2] * 4 > size)
    throw_error("image", "\"" + path + "\": truncated data.");
  return info;
}

std::vector<Map> read_npy(const std::string &path) {
  const NpyInfo info = npy_info(path);
  std::vector<Map> out(info.dims[0]);
  for (long s = 0; s < info.dims[0]; ++s) {
    out[s].resize(info.dims[1], info.dims[2]);
    read_plane({path, s}, out[s].shape(), out[s].data());
  }
  return out;
}

void put16(std::vector<unsigned char> &v, uint32_t x) {
  v.push_back(x & 255);
  v.push_back((x >> 8) & 255);
}
void put32(std::vector<unsigned char> &v, uint32_t x) {
  for (int k = 0; k < 4; ++k) v.push_back((x >> (8 * k)) & 255);
}

}  // namespace

std::vector<ImagePlane> index_images(const std::vector<std::string> &paths, Shape2 &shape) {
  std::vector<ImagePlane> planes;
  shape = {0, 0};
  for (const std::string &p : paths) {
    Shape2 sh;
    long n = 1;
    if (is_hdf5_spec(p)) {
      const std::vector<long> idx = hdf5_planes(p, sh);
      if (planes.empty()) shape = sh;
      else if (sh != shape) throw_error("image", "All images must have the same shape (\"" + p + "\" differs).");
      for (long k : idx) planes.push_back({p, k, true});
      continue;
    }
    if (ends_with(p, ".npy")) {
      const NpyInfo info = npy_info(p);
      n = info.dims[0];
      sh = {info.dims[1], info.dims[2]};
    } else {
      sh = read_tiff(p).shape();
    }
    if (planes.empty()) shape = sh;
    else if (sh != shape) throw_error("image", "All images must have the same shape (\"" + p + "\" differs).");
    for (long k = 0; k < n; ++k) planes.push_back({p, k});
  }
  return planes;
}

void read_plane(const ImagePlane &p, const Shape2 &shape, float *dst) {
  const size_t elems = static_cast<size_t>(shape[0]) * shape[1];
  if (p.h5) {
    Shape2 face;
    hdf5_planes(p.path, face);
    if (face != shape) throw_error("image", "\"" + p.path + "\" changed shape.");
    hdf5_read_plane(p.path, p.plane, dst);
    return;
  }
  if (ends_with(p.path, ".npy")) {
    const NpyInfo info = npy_info(p.path);
    if (info.dims[1] != shape[0] || info.dims[2] != shape[1] || p.plane >= info.dims[0])
      throw_error("image", "\"" + p.path + "\" changed shape.");
    std::ifstream f(p.path, std::ios::binary);
    f.seekg(static_cast<std::streamoff>(info.data + elems * 4 * p.plane));
    f.read(reinterpret_cast<char *>(dst), static_cast<std::streamsize>(elems * 4));
    if (static_cast<size_t>(f.gcount()) != elems * 4) throw_error("image", "\"" + p.path + "\": truncated data.");
    return;
  }
  const Map m = read_tiff(p.path);
  if (m.shape() != shape) throw_error("image", "\"" + p.path + "\" changed shape.");
  std::memcpy(dst, m.data(), elems * 4);
}

namespace {

int sample_bytes(int bpp) { return bpp ? std::abs(bpp) / 8 : 4; }

std::string npy_descr(int bpp) {
  if (!bpp) return "<f4";
  return std::string(bpp < 0 ? "<i" : "<u") + std::to_string(std::abs(bpp) / 8);
}

template <typename T>
void limit_to(const float *data, size_t n, float mincon, float maxcon, T *out) {
  if (mincon == maxcon && n) {
    mincon = *std::min_element(data, data + n);
    maxcon = *std::max_element(data, data + n);
  }
  if (mincon == maxcon) {
    std::fill(out, out + n, T(0));
    return;
  }
  const float lo = static_cast<float>(std::numeric_limits<T>::min());
  const float type_width = static_cast<float>(std::numeric_limits<T>::max()) - lo;
  const float lims_width = maxcon - mincon;
  for (size_t i = 0; i < n; ++i) {
    const float interim = (data[i] - mincon) / lims_width;
    if (interim < 0) {
      out[i] = std::numeric_limits<T>::min();
    } else if (interim > 1) {
      out[i] = std::numeric_limits<T>::max();
    } else {
      // the float value can round up past the type's maximum (32 and 64 bits)
      const double v = static_cast<double>(lo + interim * type_width);
      out[i] = v >= static_cast<double>(std::numeric_limits<T>::max()) ? std::numeric_limits<T>::max()
                                                                        : static_cast<T>(v);
    }
  }
}

}  // namespace

std::vector<unsigned char> to_integer(const float *data, size_t n, int bpp, float mincon, float maxcon) {
  std::vector<unsigned char> out(n * static_cast<size_t>(sample_bytes(bpp)));
  void *o = out.data();
  switch (bpp) {
    case -8: limit_to(data, n, mincon, maxcon, static_cast<int8_t *>(o)); break;
    case 8: limit_to(data, n, mincon, maxcon, static_cast<uint8_t *>(o)); break;
    case -16: limit_to(data, n, mincon, maxcon, static_cast<int16_t *>(o)); break;
    case 16: limit_to(data, n, mincon, maxcon, static_cast<uint16_t *>(o)); break;
    case -32: limit_to(data, n, mincon, maxcon, static_cast<int32_t *>(o)); break;
    case 32: limit_to(data, n, mincon, maxcon, static_cast<uint32_t *>(o)); break;
    case -64: limit_to(data, n, mincon, maxcon, static_cast<int64_t *>(o)); break;
    case 64: limit_to(data, n, mincon, maxcon, static_cast<uint64_t *>(o)); break;
    default: throw_error("save image", "unsupported bits per sample " + std::to_string(bpp));
  }
  return out;
}

NpyWriter::NpyWriter(const std::string &path, long count, long rows, long cols, int bpp, float mincon,
                     float maxcon)
    : plane_bytes_(static_cast<size_t>(rows) * cols * sample_bytes(bpp)), path_(path), bpp_(bpp), mincon_(mincon),
      maxcon_(maxcon), elems_(rows * cols) {
  if (bpp) (void)to_integer(nullptr, 0, bpp, 0, 0);  // rejects an unsupported bpp
  fd_ = ::open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (fd_ < 0) throw_error("image", "Could not write file \"" + path + "\".");
  std::string h = "{'descr': '" + npy_descr(bpp) + "', 'fortran_order': False, 'shape': (" +
                  std::to_string(count) + ", " + std::to_string(rows) + ", " + std::to_string(cols) + "), }";
  while ((10 + h.size() + 1) % 64) h += ' ';
  h += '\n';
  std::string head("\x93NUMPY\x01\x00", 8);
  head += static_cast<char>(h.size() & 255);
  head += static_cast<char>(h.size() >> 8);
  head += h;
  data_ = head.size();
  if (::pwrite(fd_, head.data(), head.size(), 0) != static_cast<ssize_t>(head.size()))
    throw_error("image", "Could not write file \"" + path + "\".");
}

NpyWriter::~NpyWriter() {
  if (fd_ >= 0) ::close(fd_);
}

void NpyWriter::write(long index, const float *plane) {
  std::vector<unsigned char> conv;
  if (bpp_) conv = to_integer(plane, static_cast<size_t>(elems_), bpp_, mincon_, maxcon_);
  const char *p = bpp_ ? reinterpret_cast<const char *>(conv.data()) : reinterpret_cast<const char *>(plane);
  size_t done = 0;
  const off_t at = static_cast<off_t>(data_ + plane_bytes_ * static_cast<size_t>(index));
  while (done < plane_bytes_) {
    const ssize_t n = ::pwrite(fd_, p + done, plane_bytes_ - done, at + static_cast<off_t>(done));
    if (n <= 0) throw_error("image", "writing plane " + std::to_string(index) + " of \"" + path_ + "\" failed");
    done += static_cast<size_t>(n);
  }
}

std::vector<Map> read_images(const std::string &path) {
  if (ends_with(path, ".npy")) return read_npy(path);
  return {read_tiff(path)};
}

void write_tiff(const std::string &path, const Map &img) { write_tiff(path, img.data(), img.rows(), img.columns()); }

void write_tiff(const std::string &path, const float *data, long rows, long cols) {
  write_tiff(path, data, rows, cols, 0, 0.0f, 0.0f);
}

void write_tiff(const std::string &path, const float *data, long rows, long cols, int bpp, float mincon,
                float maxcon) {
  const uint32_t w = static_cast<uint32_t>(cols), h = static_cast<uint32_t>(rows);
  const uint32_t sb = static_cast<uint32_t>(sample_bytes(bpp));
  std::vector<unsigned char> conv;
  if (bpp) conv = to_integer(data, static_cast<size_t>(w) * h, bpp, mincon, maxcon);
  const void *pixels = bpp ? static_cast<const void *>(conv.data()) : static_cast<const void *>(data);
  std::vector<unsigned char> v;
  v.insert(v.end(), {'I', 'I'});
  put16(v, 42);
  put32(v, 8);
  const uint16_t n = 10;
  const uint32_t data_off = 8 + 2 + 12 * n + 4;
  put16(v, n);
  auto entry = [&](uint16_t tag, uint16_t type, uint32_t cnt, uint32_t val) {
    put16(v, tag);
    put16(v, type);
    put32(v, cnt);
    if (type == 3 && cnt == 1) {
      put16(v, val);
      put16(v, 0);
    } else {
      put32(v, val);
    }
  };
  entry(256, 4, 1, w);
  entry(257, 4, 1, h);
  entry(258, 3, 1, 8 * sb);
  entry(259, 3, 1, 1);
  entry(262, 3, 1, 1);
  entry(273, 4, 1, data_off);
  entry(277, 3, 1, 1);
  entry(278, 4, 1, h);
  entry(279, 4, 1, w * h * sb);
  entry(339, 3, 1, bpp == 0 ? 3 : bpp < 0 ? 2 : 1);  // IEEE float, signed or unsigned integer
  put32(v, 0);
  const size_t at = v.size();
  v.resize(at + static_cast<size_t>(w) * h * sb);
  std::memcpy(v.data() + at, pixels, static_cast<size_t>(w) * h * sb);
  std::ofstream f(path, std::ios::binary);
  if (!f) throw_error("image", "Could not write file \"" + path + "\".");
  f.write(reinterpret_cast<const char *>(v.data()), static_cast<std::streamsize>(v.size()));
}

void write_npy(const std::string &path, const std::vector<Map> &stack) {
  if (stack.empty()) throw_error("image", "nothing to write");
  const long r = stack[0].rows(), c = stack[0].columns();
  std::string h = "{'descr': '<f4', 'fortran_order': False, 'shape': (" + std::to_string(stack.size()) + ", " +
                  std::to_string(r) + ", " + std::to_string(c) + "), }";
  while ((10 + h.size() + 1) % 64) h += ' ';
  h += '\n';
  std::ofstream f(path, std::ios::binary);
  if (!f) throw_error("image", "Could not write file \"" + path + "\".");
  f.write("\x93NUMPY\x01\x00", 8);
  const unsigned char hl[2] = {static_cast<unsigned char>(h.size() & 255), static_cast<unsigned char>(h.size() >> 8)};
  f.write(reinterpret_cast<const char *>(hl), 2);
  f.write(h.data(), static_cast<std::streamsize>(h.size()));
  for (const Map &m : stack) f.write(reinterpret_cast<const char *>(m.data()), static_cast<std::streamsize>(m.size() * 4));
}

std::string mask_name(const std::string &mask, long index, long count) {
  if (count <= 1 && mask.find('@') == std::string::npos) return mask;
  int digits = 1;
  for (long c = count - 1; c >= 10; c /= 10) ++digits;
  char num[32];
  std::snprintf(num, sizeof num, "%0*ld", digits, index);
  const size_t at = mask.rfind('@');
  if (at != std::string::npos) return mask.substr(0, at) + num + mask.substr(at + 1);
  const size_t slash = mask.rfind('/');
  const size_t dot = mask.rfind('.');
  if (dot != std::string::npos && (slash == std::string::npos || dot > slash))
    return mask.substr(0, dot) + "-" + num + mask.substr(dot);
  return mask + "-" + num;
}

}  // namespace ctas

// Test and scripting helper of the tools' integer output (not part of the
// engine's C ABI): to_integer into `out` (n samples of |bpp|/8 bytes).
extern "C" int ctas_io_to_integer(const float *data, size_t n, int bpp, float mincon, float maxcon, void *out) {
  try {
    const std::vector<unsigned char> v = ctas::to_integer(data, n, bpp, mincon, maxcon);
    std::memcpy(out, v.data(), v.size());
    return 0;
  } catch (const ctas::CtasErr &) {
    return -1;
  }
}
