// Float32 slice files for the ctas-ct / ctas-ax tools: uncompressed
// single-channel TIFF (the format the reference's ct tool writes by
// default, external.world.cpp:1625-1735) and NumPy .npy (2-D or 3-D stacks).
#pragma once

#include <string>
#include <vector>

#include "ct.h"

namespace ctas {

// Reads one 2-D image, or every 2-D plane of a 3-D .npy stack.
std::vector<Map> read_images(const std::string &path);
void write_tiff(const std::string &path, const Map &img);
void write_tiff(const std::string &path, const float *data, long rows, long cols);

// One 2-D plane of an image file: a TIFF, or plane `plane` of a .npy stack.
struct ImagePlane {
  std::string path;
  long plane;
};
// The planes of the files, in order, and their common shape (headers only
// for .npy; throws when the shapes differ).
std::vector<ImagePlane> index_images(const std::vector<std::string> &paths, Shape2 &shape);
// Reads one plane of the given shape into dst.
void read_plane(const ImagePlane &p, const Shape2 &shape, float *dst);

// A .npy stack of `count` planes of rows x cols, written plane by plane at
// their positions (write() may be called concurrently for different planes).
class NpyWriter {
 public:
  NpyWriter(const std::string &path, long count, long rows, long cols);
  ~NpyWriter();
  NpyWriter(const NpyWriter &) = delete;
  NpyWriter &operator=(const NpyWriter &) = delete;
  void write(long index, const float *plane);

 private:
  int fd_ = -1;
  size_t data_ = 0, plane_bytes_;
  std::string path_;
};
void write_npy(const std::string &path, const std::vector<Map> &stack);

// The reference's output mask (external.world.cpp:1096-1112): the last '@'
// is replaced by the zero-padded slice number; without '@', "-@" is inserted
// before the extension. Single-slice outputs keep the name as is.
std::string mask_name(const std::string &mask, long index, long count);

}  // namespace ctas
