// ctas-ct: CT reconstruction of sinogram(s), the drop-in for the reference's
// `ctas ct` (src/executables/ct.cpp:43-244): same options and defaults
// (inputs: TIFF, .npy stacks, HDF5 "file:dataset[:[xyz]slices]"); the
// slices go to the GPU in batches instead of one per CPU thread, with the
// reading of the next batch and the writing of the previous one overlapping
// the device work (ctas_fbp_pipeline).
#include <getopt.h>

#include <atomic>
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <memory>
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
            << "  -m, --min FLOAT         pixel value corresponding to black\n"
            << "  -M, --max FLOAT         pixel value corresponding to white\n"
            << "  -i, --int BPP           bits per pixel of integer output (8, 16, 32, 64; negative: signed;\n"
            << "                          default 0, float32, unless -m or -M is given: then 8)\n"
            << "      --cpu               accepted for compatibility (the GPU is always used)\n"
            << "      --interp ctas|linear  detector sampling (default ctas, the reference's)\n"
            << "      --store fp32|fp16|bf16  filtered-sinogram storage (default fp32)\n"
            << "      --batch INT         slices per GPU launch (default 16)\n"
            << "      --io-threads INT    threads reading / writing the slices of a batch (default 8)\n"
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
  int ring = 0, batch = 16, io_threads = 8, bpp = 0;
  float mincon = 0, maxcon = 0;
  bool verbose = false, have_out = false, have_limits = false, have_bpp = false;
  static option opts[] = {{"output", required_argument, nullptr, 'o'},   {"contrast", required_argument, nullptr, 'k'},
                          {"center", required_argument, nullptr, 'c'},   {"arcan", required_argument, nullptr, 'a'},
                          {"filter", required_argument, nullptr, 'f'},   {"resolution", required_argument, nullptr, 'r'},
                          {"wavelength", required_argument, nullptr, 'w'}, {"ring", required_argument, nullptr, 'R'},
                          {"min", required_argument, nullptr, 'm'},      {"max", required_argument, nullptr, 'M'},
                          {"int", required_argument, nullptr, 'i'},      {"cpu", no_argument, nullptr, 1},
                          {"interp", required_argument, nullptr, 2},     {"store", required_argument, nullptr, 3},
                          {"batch", required_argument, nullptr, 4},      {"io-threads", required_argument, nullptr, 5},
                          {"verbose", no_argument, nullptr, 'v'},
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
      case 'm': mincon = std::strtof(optarg, nullptr); have_limits = true; break;
      case 'M': maxcon = std::strtof(optarg, nullptr); have_limits = true; break;
      case 'i': bpp = std::atoi(optarg); have_bpp = true; break;
      case 1: break;
      case 2: interp = optarg; break;
      case 3: store = optarg; break;
      case 4: batch = std::max(1, std::atoi(optarg)); break;
      case 5: io_threads = std::max(1, std::atoi(optarg)); break;
      case 'v': verbose = true; break;
      default: usage(argv[0]); return ch == 'h' ? 0 : 1;
    }
  }
  if (optind >= argc) {
    usage(argv[0]);
    return 0;
  }
  if (arc <= 0.0f) { std::cerr << "CT arc (given by -a) must be strictly positive.\n"; return 1; }
  // ct.cpp:148: limits without -i mean 8-bit output
  if (have_limits && !have_bpp) bpp = 8;
  if (bpp != 0 && bpp != 8 && bpp != -8 && bpp != 16 && bpp != -16 && bpp != 32 && bpp != -32 && bpp != 64 &&
      bpp != -64) {
    std::cerr << "Unsupported bits per sample " << bpp << ".\n";
    return 1;
  }
  const auto t_start = std::chrono::steady_clock::now();
  auto secs = [&] { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count(); };
  try {
    const std::vector<std::string> paths(argv + optind, argv + argc);
    Shape2 ish;
    const std::vector<ImagePlane> planes = index_images(paths, ish);
    if (!have_out) out = upgrade(argv[optind], "reconstructed-");
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
    const long S = static_cast<long>(planes.size());
    const bool npy = out.size() > 4 && out.substr(out.size() - 4) == ".npy";
    const Shape2 osh{ish[1], ish[1]};  // full slices, as the reference's tool writes them
    // Sinograms are read, reconstructed and written in overlapping batches
    // (reader and writer threads around the device work).
    struct Ctx {
      const std::vector<ImagePlane> &planes;
      Shape2 ish, osh;
      std::string out;
      long S;
      std::unique_ptr<NpyWriter> npy;
      bool verbose;
      long batch;
      int bpp;
      float mincon, maxcon;
      std::atomic<long> done{0};
    } ctx{planes, ish, osh, out, S,
          npy ? std::make_unique<NpyWriter>(out, S, osh[0], osh[1], bpp, mincon, maxcon) : nullptr, verbose, batch,
          bpp, mincon, maxcon};
    auto read = [](void *u, long i, float *sino) -> int {
      Ctx &c = *static_cast<Ctx *>(u);
      try {
        read_plane(c.planes[i], c.ish, sino);
        return 0;
      } catch (const CtasErr &) {
        return 1;
      }
    };
    auto write = [](void *u, long i, const float *slice) -> int {
      Ctx &c = *static_cast<Ctx *>(u);
      try {
        if (c.npy) c.npy->write(i, slice);
        else write_tiff(mask_name(c.out, i, c.S), slice, c.osh[0], c.osh[1], c.bpp, c.mincon, c.maxcon);
        const long done = ++c.done;
        if (c.verbose && (done % c.batch == 0 || done == c.S)) std::cerr << "CT'ing: " << done << "/" << c.S << "\n";
        return 0;
      } catch (const CtasErr &) {
        return 1;
      }
    };
    const std::vector<float> centers(S, center);
    const double t_setup = secs();
    if (ctas_fbp_pipeline(rec.handle(), S, batch, centers.data(), ring, read, write, &ctx, io_threads) != CTAS_OK)
      throw_error("ct", ctas_last_error());
    if (verbose)
      std::cerr << "setup " << t_setup << " s (device and handle), reconstruction " << (secs() - t_setup) << " s\n";
  } catch (const CtasErr &) {
    return 1;
  }
  return 0;
}
