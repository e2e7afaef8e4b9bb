// ctas-ax: rotation-axis estimate, the drop-in for the reference's `ctas ax`
// (src/executables/axis.cpp:43-208). One sinogram: prints the centre offset
// found by each of the four algorithms, one per line ("Failed to calculate
// COR." on stderr for a failing one). Two images: 0/180 deg projections.
#include <getopt.h>

#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>

#include "ct.h"
#include "image_io.h"

using namespace ctas;

int main(int argc, char **argv) {
  float arc = 180, maxShift = 0.25f;
  int denoiseRad = 0;
  float denoiseThr = 0;
  std::string outmask;
  static option opts[] = {{"contrast", required_argument, nullptr, 'k'}, {"arcan", required_argument, nullptr, 'a'},
                          {"noise-rad", required_argument, nullptr, 'n'}, {"noise-thr", required_argument, nullptr, 'N'},
                          {"output", required_argument, nullptr, 'o'},  {"max", required_argument, nullptr, 'm'},
                          {"verbose", no_argument, nullptr, 'v'},       {"help", no_argument, nullptr, 'h'},
                          {nullptr, 0, nullptr, 0}};
  int ch;
  while ((ch = getopt_long(argc, argv, "k:a:n:N:o:m:vh", opts, nullptr)) != -1) {
    switch (ch) {
      case 'k': break;  // parsed but unused by the reference (axis.cpp:77, 146)
      case 'a': arc = std::strtof(optarg, nullptr); break;
      case 'n': denoiseRad = std::atoi(optarg); break;
      case 'N': denoiseThr = std::strtof(optarg, nullptr); break;
      case 'o': outmask = optarg; break;
      case 'm': maxShift = std::strtof(optarg, nullptr); break;
      case 'v': break;
      default:
        std::cout << "Usage: " << argv[0] << " [-a ARC] [-m MAXSHIFT] [-o MASK] [-n RAD] [-N THR] sinogram | proj0 proj180\n";
        return ch == 'h' ? 0 : 1;
    }
  }
  if (optind >= argc) {
    std::cout << "Usage: " << argv[0] << " [-a ARC] [-m MAXSHIFT] [-o MASK] [-n RAD] [-N THR] sinogram | proj0 proj180\n";
    return 0;
  }
  if (arc <= 0.0f) { std::cerr << "CT arc (given by -a): must be strictly positive.\n"; return 1; }
  if (maxShift <= 0.0f) { std::cerr << "Impossible maximum shift " << maxShift << ": must be strictly positive.\n"; return 1; }
  const int nimg = argc - optind;
  try {
    Map img0 = read_images(argv[optind]).at(0);
    const Shape2 ish = img0.shape();
    if (ish[1] < 2 || ish[0] < 2) throw_error(argv[0], "Size of input image is less than 2x2.");
    float cent = 0.0f;
    auto rotax = [&](auto &&fn) {
      cent = 0.0f;
      try {
        cent = fn() - 0.5f * (ish[1] - 1);
        std::cout << cent << "\n" << std::flush;
      } catch (const CtasErr &) {
        std::cerr << "Failed to calculate COR.\n" << std::flush;
      }
    };
    // axis.cpp:139-142, 176-180: the images go through the denoiser first
    const Denoiser dnz(ish, denoiseRad, denoiseThr);
    if (nimg == 1) {
      if (denoiseRad) dnz.proc(img0);
      const float ang = arc / (std::fabs(arc) <= 1.0f ? 1.0f : static_cast<float>(ish[0] - 1));
      CTrec *rec = nullptr;
      Map recr;
      if (!outmask.empty()) {
        rec = new CTrec(ish, Contrast(Contrast::ABS), arc);
        Map copy = img0;
        rec->sino(copy);
      }
      for (int algo = 0; algo < 4; ++algo) {
        rotax([&] { return raxis(img0, ang, maxShift, algo); });
        if (rec) {
          rec->repeat(recr, cent);
          const std::string name = mask_name(outmask, algo, 4);
          const size_t dot = name.rfind('.');
          write_tiff((dot == std::string::npos ? name : name.substr(0, dot)) + "_" + std::to_string(cent) + ".tif", recr);
        }
      }
      delete rec;
    } else if (nimg == 2) {
      Map p180 = read_images(argv[optind + 1]).at(0);
      if (p180.shape() != ish) throw_error(argv[0], "Different shapes of input projections.");
      if (denoiseRad) {
        dnz.proc(img0);
        dnz.proc(p180);
      }
      rotax([&] { return raxis(img0, p180, maxShift); });
    } else {
      throw_error(argv[0], "Can accept only one (sinogram) or two (opposit projections) images.");
    }
  } catch (const CtasErr &) {
    return 1;
  }
  return 0;
}
