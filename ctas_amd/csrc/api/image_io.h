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

// One 2-D plane of an image file: a TIFF, plane `plane` of a .npy stack,
// or plane `plane` (along the descriptor's axis) of a HDF5 dataset given as
// "file:dataset[:[xyz]slices]" (h5).
struct ImagePlane {
  std::string path;
  long plane;
  bool h5 = false;
};
// The planes of the files, in order, and their common shape (headers only
// for .npy; throws when the shapes differ).
std::vector<ImagePlane> index_images(const std::vector<std::string> &paths, Shape2 &shape);
// Reads one plane of the given shape into dst.
void read_plane(const ImagePlane &p, const Shape2 &shape, float *dst);

// HDF5 (hdf5_io.cpp; the library is loaded at run time). slice_list:
// the reference's slice string (common.world.cpp:471-612) for an axis of
// `height` planes. is_hdf5_spec: "file:desc" naming an existing HDF5 file.
std::vector<long> slice_list(const std::string &spec, long height);
bool is_hdf5_spec(const std::string &spec);
std::vector<long> hdf5_planes(const std::string &spec, Shape2 &face);
void hdf5_read_plane(const std::string &spec, long plane, float *dst);
void hdf5_write_volume(const std::string &spec, const float *data, long d0, long d1, long d2);

// Integer output of the tools (`-m/-M/-i`, ct.cpp:112-118, 148): the
// reference's limit<T> (external.world.cpp): (v - mincon)/(maxcon - mincon)
// mapped onto the type's range, clamped below 0 and above 1; mincon ==
// maxcon takes the image's own minimum and maximum (all zero when those are
// equal). bpp: 8, 16, 32, 64 unsigned, negative for signed.
std::vector<unsigned char> to_integer(const float *data, size_t n, int bpp, float mincon, float maxcon);
// TIFF of bpp bits per sample (0: float32; an integer sample format else).
void write_tiff(const std::string &path, const float *data, long rows, long cols, int bpp, float mincon,
                float maxcon);

// A .npy stack of `count` planes of rows x cols, written plane by plane at
// their positions (write() may be called concurrently for different planes).
class NpyWriter {
 public:
  // bpp != 0: planes stored as that integer type (to_integer per plane)
  NpyWriter(const std::string &path, long count, long rows, long cols, int bpp = 0, float mincon = 0,
            float maxcon = 0);
  ~NpyWriter();
  NpyWriter(const NpyWriter &) = delete;
  NpyWriter &operator=(const NpyWriter &) = delete;
  void write(long index, const float *plane);

 private:
  int fd_ = -1;
  size_t data_ = 0, plane_bytes_;
  std::string path_;
  int bpp_ = 0;
  float mincon_ = 0, maxcon_ = 0;
  long elems_ = 0;
};
void write_npy(const std::string &path, const std::vector<Map> &stack);

// The reference's output mask (external.world.cpp:1096-1112): the last '@'
// is replaced by the zero-padded slice number; without '@', "-@" is inserted
// before the extension. Single-slice outputs keep the name as is.
std::string mask_name(const std::string &mask, long index, long count);

}  // namespace ctas
