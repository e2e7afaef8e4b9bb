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
void write_npy(const std::string &path, const std::vector<Map> &stack);

// The reference's output mask (external.world.cpp:1096-1112): the last '@'
// is replaced by the zero-padded slice number; without '@', "-@" is inserted
// before the extension. Single-slice outputs keep the name as is.
std::string mask_name(const std::string &mask, long index, long count);

}  // namespace ctas
