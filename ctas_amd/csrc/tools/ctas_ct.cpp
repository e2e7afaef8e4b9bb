// ctas-ct: CT reconstruction of sinogram(s), the drop-in for the reference's
// `ctas ct` (src/executables/ct.cpp:43-244): same options and defaults; the
// slices go to the GPU in batches instead of one per CPU thread.
#include <getopt.h>

#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "ct.h"
#include "image_io.h"

using namespace ctas;

namespace {

void usage(const char *cmd) {
  std::cout << "Usage: " << cmd << " [OPTIONS] sinogram [sinogram ...]\n"
            << "CT reconstruction of sinogram(s).\n"
            << "  -o, --output MASK       output mask (default reconstructed-<sinogram>)\n"
            << "  -k, --contrast NAME     ABS|REF|PHS|FLT (default ABS)\n"
            << "  -c, --center FLOAT      rotation centre offset (default 0)\n"
            << "  -a, --arcan FLOAT       arc (> 1 deg) or step (< 1 deg) (default 180)\n"
            << "  -f, --filter NAME[:P]   filter window (default RAMP)\n"
            << "  -r, --resolution FLOAT  pixel size (micron)\n"
            << "  -w, --wavelength FLOAT  wavelength (Angstrom)\n"
            << "  -R, --ring INT          half size of the ring filter (default 0)\n"
            << "      --cpu               accepted for compatibility (the GPU is always used)\n"
            << "      --interp ctas|linear  detector sampling (default ctas, the reference's)\n"
            << "      --store fp32|fp16|bf16  filtered-sinogram storage (default fp32)\n"
            << "      --batch INT         slices per GPU launch (default 16)\n"
            << "  -v, --verbose\n";
}

std::string upgrade(const std::string &path, const std::string &prefix) {
  const size_t slash = path.rfind('/');
  if (slash == std::string::npos) return prefix + path;
  return path.substr(0, slash + 1) + prefix + path.substr(slash + 1);
}

}  // namespace

int main(int argc, char **argv) {
  std::string out, contrast = "ABS", filter = "RAMP", interp = "ctas", store = "fp32";
  float center = 0, arc = 180, dd = 0, lambda = 0;
  int ring = 0, batch = 16;
  bool verbose = false, have_out = false;
  static option opts[] = {{"output", required_argument, nullptr, 'o'},   {"contrast", required_argument, nullptr, 'k'},
                          {"center", required_argument, nullptr, 'c'},   {"arcan", required_argument, nullptr, 'a'},
                          {"filter", required_argument, nullptr, 'f'},   {"resolution", required_argument, nullptr, 'r'},
                          {"wavelength", required_argument, nullptr, 'w'}, {"ring", required_argument, nullptr, 'R'},
                          {"min", required_argument, nullptr, 'm'},      {"max", required_argument, nullptr, 'M'},
                          {"int", required_argument, nullptr, 'i'},      {"cpu", no_argument, nullptr, 1},
                          {"interp", required_argument, nullptr, 2},     {"store", required_argument, nullptr, 3},
                          {"batch", required_argument, nullptr, 4},      {"verbose", no_argument, nullptr, 'v'},
                          {"help", no_argument, nullptr, 'h'},           {nullptr, 0, nullptr, 0}};
  int ch;
  while ((ch = getopt_long(argc, argv, "o:k:c:a:f:r:w:R:m:M:i:vh", opts, nullptr)) != -1) {
    switch (ch) {
      case 'o': out = optarg; have_out = true; break;
      case 'k': contrast = optarg; break;
      case 'c': center = std::strtof(optarg, nullptr); break;
      case 'a': arc = std::strtof(optarg, nullptr); break;
      case 'f': filter = optarg; break;
      case 'r':
        dd = std::strtof(optarg, nullptr);
        if (dd <= 0.0f) { std::cerr << "Negative pixel size (given by -r).\n"; return 1; }
        dd /= 1.0E6f;
        break;
      case 'w':
        lambda = std::strtof(optarg, nullptr);
        if (lambda <= 0.0f) { std::cerr << "Negative wawelwngth (given by -w).\n"; return 1; }
        lambda /= 1.0E10f;
        break;
      case 'R': ring = std::atoi(optarg); break;
      case 'm': case 'M': case 'i':
        std::cerr << "Integer output is not supported: slices are written as float32.\n";
        return 1;
      case 1: break;
      case 2: interp = optarg; break;
      case 3: store = optarg; break;
      case 4: batch = std::max(1, std::atoi(optarg)); break;
      case 'v': verbose = true; break;
      default: usage(argv[0]); return ch == 'h' ? 0 : 1;
    }
  }
  if (optind >= argc) {
    usage(argv[0]);
    return 0;
  }
  if (arc <= 0.0f) { std::cerr << "CT arc (given by -a) must be strictly positive.\n"; return 1; }
  try {
    std::vector<Map> sinos;
    for (int k = optind; k < argc; ++k)
      for (Map &m : read_images(argv[k])) sinos.push_back(std::move(m));
    if (!have_out) out = upgrade(argv[optind], "reconstructed-");
    const Shape2 ish = sinos.at(0).shape();
    for (const Map &m : sinos)
      if (m.shape() != ish) throw_error("ct", "All sinograms must have the same shape.");
    int flags = 0;
    if (interp == "linear") flags |= INTERP_LINEAR;
    else if (interp != "ctas") throw_error("ct", "Unknown interpolation \"" + interp + "\".");
    if (store == "fp16") flags |= STORE_FP16;
    else if (store == "bf16") flags |= STORE_BF16;
    else if (store != "fp32") throw_error("ct", "Unknown storage \"" + store + "\".");
    Filter ft;
    if (!_conversion(&ft, filter)) throw_error("ct", "Bad filter \"" + filter + "\".");
    CTrec rec(ish, Contrast(contrast), arc, ft, flags);
    rec.setPhysics(dd, lambda);
    RingFilter rf(static_cast<size_t>(ish[1]), static_cast<size_t>(ring));
    const long S = static_cast<long>(sinos.size());
    const bool npy = out.size() > 4 && out.substr(out.size() - 4) == ".npy";
    std::vector<Map> all;
    for (long s0 = 0; s0 < S; s0 += batch) {
      const long n = std::min<long>(batch, S - s0);
      std::vector<Map> part(sinos.begin() + s0, sinos.begin() + s0 + n), slices;
      for (Map &m : part) rf.apply(m);
      rec.reconstruct(part, slices, std::vector<float>(n, center));
      for (long k = 0; k < n; ++k) {
        if (npy) all.push_back(std::move(slices[k]));
        else write_tiff(mask_name(out, s0 + k, S), slices[k]);
      }
      if (verbose) std::cerr << "CT'ing: " << (s0 + n) << "/" << S << "\n";
    }
    if (npy) write_npy(out, all);
  } catch (const CtasErr &) {
    return 1;
  }
  return 0;
}
