// ctas-ts: tomosynthesis reconstruction of one plane, the drop-in for the
// reference's `ctas ts` (src/executables/ts.cpp:41-222): same arguments,
// options, list-file format and messages. The projections are streamed to
// the GPU and added to the plane there (ctas_ts_accumulate) instead of one
// CPU ts_add call per projection.
#include <getopt.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <utility>
#include <vector>

#include "ct.h"
#include "image_io.h"

using namespace ctas;

namespace {

void usage(const char *cmd) {
  std::cout << "Usage: " << cmd << " [OPTIONS] <data file> [result]\n"
            << "Tomosynthesis reconstruction.\n"
            << "The program reads file which describes the images containing TS projections and\n"
            << "performs TS reconstruction of the requested plane. The data file holds lines\n"
            << "    <float> <string>\n"
            << "the angle in degrees at which the projection was recorded and the image file of\n"
            << "the projection, relative to the data file's directory.\n"
            << "  -C, --contrast NAME     ABS|REF|PHS|FLT (default ABS)\n"
            << "  -y, --plane INT         plane of the reconstruction, relative to the rotation centre\n"
            << "  -c, --center FLOAT      rotation centre (default 0)\n"
            << "  -v, --verbose\n";
}

// upgrade() of the reference (the prefix goes before the file name).
std::string upgrade(const std::string &path, const std::string &prefix) {
  const size_t slash = path.rfind('/');
  if (slash == std::string::npos) return prefix + path;
  return path.substr(0, slash + 1) + prefix + path.substr(slash + 1);
}

// proj_info (ts.cpp:131-192): "angle name" pairs, angles in radians
// (angle * M_PI / 180 in double, kept as float), names relative to the data
// file's directory unless absolute.
std::vector<std::pair<float, std::string>> read_list(const std::string &data_name) {
  FILE *in = std::fopen(data_name.c_str(), "r");
  if (!in) throw_error("scan input", "Could not open input data file '" + data_name + "' for reading.");
  const std::string fdir(data_name, 0, data_name.rfind('/') + 1);
  std::vector<std::pair<float, std::string>> list;
  float nangle = 0;
  char name[1024];
  while (!std::feof(in) && std::fscanf(in, "%f %1023s\n", &nangle, name) == 2) {
    const std::string n(name);
    list.emplace_back(static_cast<float>(nangle * M_PI / 180), n.size() && n[0] == '/' ? n : fdir + n);
  }
  if (!std::feof(in))
    warn("scan input", "Certain part of the input file \"" + data_name +
                           "\" could not be scanned. succeeded in " + std::to_string(list.size()) + " strings.");
  std::fclose(in);
  if (list.empty()) throw_error("scan input", "Input file \"" + data_name + "\" does not provide any input info.");
  return list;
}

}  // namespace

int main(int argc, char **argv) {
  std::string contrast = "ABS";
  float center = 0;
  int plane = 0;
  bool have_plane = false, verbose = false;
  static option opts[] = {{"contrast", required_argument, nullptr, 'C'}, {"plane", required_argument, nullptr, 'y'},
                          {"center", required_argument, nullptr, 'c'},   {"int", required_argument, nullptr, 'i'},
                          {"verbose", no_argument, nullptr, 'v'},        {"help", no_argument, nullptr, 'h'},
                          {nullptr, 0, nullptr, 0}};
  int ch;
  while ((ch = getopt_long(argc, argv, "C:y:c:i:vh", opts, nullptr)) != -1) {
    switch (ch) {
      case 'C': contrast = optarg; break;
      case 'y': plane = std::atoi(optarg); have_plane = true; break;
      case 'c': center = std::strtof(optarg, nullptr); break;
      case 'i':
        std::cerr << "Integer output is not supported: the plane is written as float32.\n";
        return 1;
      case 'v': verbose = true; break;
      default: usage(argv[0]); return ch == 'h' ? 0 : 1;
    }
  }
  if (optind >= argc) {
    usage(argv[0]);
    return 0;
  }
  const std::string cmd = "ts";
  try {
    if (!have_plane)
      throw_error(cmd, "The reconstruction plane must be explicitly set. Use option -y, --plane for this.");
    const std::string data_name = argv[optind];
    const std::string out =
        optind + 1 < argc ? argv[optind + 1] : upgrade(data_name, "ts-") + "-" + std::to_string(plane) + ".tif";
    const auto list = read_list(data_name);
    std::vector<Map> first = read_images(list[0].second);
    const Shape2 shape = first.at(0).shape();
    if (shape[1] / 2 <= std::abs(plane))
      throw_error(cmd, "The reconstruction plane is outside the image: Image width: " + std::to_string(shape[1]) +
                           ", reconstruction plane: " + std::to_string(plane) + ".");
    const size_t elems = static_cast<size_t>(shape[0]) * shape[1];
    std::vector<float> projections(elems * list.size()), angles(list.size());
    for (size_t k = 0; k < list.size(); ++k) {
      std::vector<Map> img = k ? read_images(list[k].second) : std::move(first);
      if (img.at(0).shape() != shape)
        throw_error(cmd, "Image \"" + list[k].second + "\" does not match the shape of the first projection.");
      std::copy(img[0].data(), img[0].data() + elems, projections.begin() + k * elems);
      angles[k] = list[k].first;
    }
    const Filter ft;  // the reference's ts uses the default filter, RAMP
    ctas_ts *h = nullptr;
    if (ctas_ts_create(&h, static_cast<int>(shape[0]), static_cast<int>(shape[1]), Contrast(contrast).contrast(),
                       ft.filter(), ft.param()) != CTAS_OK)
      throw_error(cmd, ctas_last_error());
    Map result(shape);
    const int rc = ctas_ts_accumulate(h, projections.data(), static_cast<int>(list.size()), angles.data(), center,
                                      plane, result.data());
    ctas_ts_destroy(h);
    if (rc != CTAS_OK) throw_error(cmd, ctas_last_error());
    if (verbose) std::cerr << "reconstruction: " << list.size() << " projections\n";
    write_tiff(out, result);
  } catch (const CtasErr &) {
    return 1;
  }
  return 0;
}
