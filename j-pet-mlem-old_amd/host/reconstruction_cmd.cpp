// 2d_strip_reconstruction — the reference's command-line tool
// (src/2d/strip/reconstruction_cmd.cpp:105-301) on the MI355X library.
//
// Same options (options.cpp:13-124, common/options.cpp:5-23), the same
// <events>.cfg auto-loading, event formats, iteration blocks (-i) x iterations
// (-I), and the same output files.  Every reconstruction runs the HIP kernels
// through include/strip_gpu_reconstruction.hpp:
//   --gpu / -G   the reference's --gpu driver protocol: <out>_sensitivity
//                first, early iterations of block 0, full-matrix .txt;
//   otherwise    the reference CPU path's protocol: sensitivity PNG, one
//                output per block, sparse "x y value" .txt, --png-max.
// The arithmetic is the reference CPU path's (GaussianKernel, no
// sensitivity, guarded 1/D) unless --legacy-gpu-kernel selects the old CUDA
// kernel's (AnalyticKernel, 5-point sensitivity in the denominator).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "cmdline.hpp"
#include "io.hpp"
#include "strip_types.hpp"
// after PET2D::PixelMap / Pixel
#include "strip_gpu_reconstruction.hpp"

using Scanner = PET2D::Strip::Scanner<float, short>;
using Response = PET2D::Strip::Response<float>;
using Image = PET2D::PixelMap<PET2D::Pixel<short>, float>;

namespace {

// util::progress (src/util/progress.h:17-154): -v estimate line on stderr,
// -vv per-iteration benchmark table on stdout.
class Progress {
 public:
  Progress(int verbosity, int total) : verbosity_(verbosity), total_(total) {
    if (verbosity_ >= 2) std::cout << "# it     time (ms)" << std::endl;
  }
  void operator()(int completed, bool finished) {
    if (!verbosity_) return;
    if (verbosity_ >= 2) {
      if (!finished) {
        start_ = clock::now();
        return;
      }
      const double ms =
          std::chrono::duration<double, std::milli>(clock::now() - start_)
              .count();
      total_ms_ += ms;
      std::cout << std::setiosflags(std::ios::fixed) << std::setprecision(2)
                << std::setw(4) << (completed + 1) << ' ' << std::setw(10)
                << ms << std::endl;
      if (completed == total_ - 1)
        std::cout << "# av " << std::setw(10) << total_ms_ / total_
                  << "   total " << total_ms_ << std::endl;
      std::cout << std::resetiosflags(std::ios::fixed);
      return;
    }
    if (finished) {
      std::cerr << ' ' << (int)std::round(100.0 * (completed + 1) / total_)
                << "% " << (completed + 1) << "/" << total_
                << ((completed + 1 == total_) ? "\n" : "\r");
    }
  }

 private:
  using clock = std::chrono::high_resolution_clock;
  int verbosity_, total_;
  clock::time_point start_;
  double total_ms_ = 0;
};

void add_options(strip_cli::Parser& cl) {
  using strip_cli::Kind;
  cl.add("config", 'c', Kind::Path, "", "load config file");
  cl.add("radius", 'r', Kind::Doubles, "0.5", "R distance between scintillator");
  cl.add("length", 'l', Kind::Double, "1", "scintillator length");
  cl.add("s-pixel", 'p', Kind::Double, "0.01", "pixel size");
  cl.add("n-pixels", 'n', Kind::Int, "", "number of pixels");
  cl.add("n-z-pixels", 0, Kind::Int, "", "number of z pixels");
  cl.add("n-y-pixels", 0, Kind::Int, "", "number of y pixels");
  cl.add("s-z", 0, Kind::Double, "0.015", "TOF sigma along z axis");
  cl.add("s-dl", 0, Kind::Double, "0.06", "TOF sigma delta-l");
  cl.add("n-emissions", 'e', Kind::Int, "0", "number of emissions", true);
  cl.add("verbose", 'v', Kind::Flag, "",
         "print progress information (-v) or benchmark (-vv)");
  cl.add("n-threads", 'T', Kind::Int, "", "number of host threads");
  cl.add("3d-response", 0, Kind::Flag, "",
         "input response contains additional LOR information");
  cl.add("dl-is-time", 0, Kind::Flag, "",
         "delta is time, convert using speed of light");
  cl.add("speed-of-light", 0, Kind::Double, "299792458", "speed of light");
  cl.add("scale-length", 0, Kind::Double, "1", "scale input length values");
  cl.add("scale-time", 0, Kind::Double, "1", "scale input time values");
  cl.add("blocks", 'i', Kind::Int, "0", "number of iteration blocks");
  cl.add("iterations", 'I', Kind::Int, "1", "number of iterations (per block)");
  cl.add("output", 'o', Kind::Path, "", "output files prefix (png)", true);
  cl.add("png-max", 0, Kind::Double, "0", "maximum value mapped to 255 in PNG");
  cl.add("gpu", 'G', Kind::Flag, "", "use the --gpu driver output protocol");
  cl.add("cuda-device", 'D', Kind::Int, "0", "GPU device");
  cl.add("cuda-blocks", 'B', Kind::Int, "64", "blocks (bbox kernel)");
  cl.add("cuda-threads", 'W', Kind::Int, "512",
         "threads per block (bbox kernel, capped at 256)");
  cl.add("rho", 0, Kind::Path, "", "start rho (eg. existing iteration)");
  cl.add("rho-is-int", 0, Kind::Flag, "",
         "start rho uses integer representation (not float)");
  cl.add("legacy-gpu-kernel", 0, Kind::Flag, "",
         "arithmetic of the old CUDA kernel (AnalyticKernel, 5-point "
         "sensitivity in the denominator, unguarded 1/D)");
}

// calculate_scanner_options (options.cpp:106-124)
void complete_scanner_options(strip_cli::Parser& cl) {
  if (cl.exist("n-pixels")) {
    cl.set("n-z-pixels", cl.text("n-pixels"));
    cl.set("n-y-pixels", cl.text("n-pixels"));
    return;
  }
  const auto radius = cl.get_doubles("radius");
  if (radius.empty()) throw std::runtime_error("at least one --radius has to be specified");
  if (radius.size() > 1)
    std::cerr << "warning: more than one radius specified with --radius"
              << std::endl;
  const double s_pixel = cl.get_double("s-pixel");
  if (!cl.exist("n-z-pixels"))
    cl.set("n-z-pixels",
           std::to_string((int)std::ceil(cl.get_double("length") / s_pixel)));
  if (!cl.exist("n-y-pixels"))
    cl.set("n-y-pixels", std::to_string((int)std::ceil(2 * radius[0] / s_pixel)));
}

int run(int argc, char** argv) {
  strip_cli::Parser cl;
  add_options(cl);
  cl.parse(argc, argv);
  if (cl.rest().empty()) {
    if (argc == 1) {
      std::cerr << cl.usage(argv[0]);
      return 0;
    }
    throw std::runtime_error(
        "at least one responses input file expected, consult --help");
  }
  // <events>.cfg accompanying the first input that has one
  // (cmdline_hooks.cpp:110-123); command-line values take precedence
  for (const std::string& fn : cl.rest()) {
    const std::string cfg = strip_cli::path::wo_ext(fn) + ".cfg";
    std::ifstream probe(cfg);
    if (probe) {
      probe.close();
      cl.load_file(cfg);
    }
  }
  complete_scanner_options(cl);

  const Scanner scanner((float)cl.get_doubles("radius")[0],
                        (float)cl.get_double("length"),
                        (short)cl.get_int("n-y-pixels"),
                        (short)cl.get_int("n-z-pixels"),
                        (float)cl.get_double("s-pixel"),
                        (float)cl.get_double("s-pixel"),
                        (float)cl.get_double("s-z"),
                        (float)cl.get_double("s-dl"));
  const int verbose = cl.count("verbose");
  const int n_z = scanner.n_z_pixels, n_y = scanner.n_y_pixels;
  if (verbose)
    std::cout << "# image: " << n_y << "x" << n_z << std::endl;

  Image rho(scanner.n_z_pixels, scanner.n_y_pixels, 1.0f);
  if (cl.exist("rho"))
    strip_io::read_image(cl.text("rho"), cl.exist("rho-is-int"), rho.data,
                         (size_t)rho.size);

  // events
  std::vector<float> events;
  const bool is_3d = cl.exist("3d-response");
  for (const std::string& fn : cl.rest()) {
    const size_t before = events.size();
    if (strip_cli::path::ext(fn) == ".txt") {
      strip_io::append_text_events(fn, is_3d, events);
    } else {
      if (is_3d) throw std::runtime_error("3D input must have .txt extension");
      strip_io::append_binary_events(fn, events);
    }
    if (verbose)
      std::cerr << "# read " << events.size() / 3 << " responsess from " << fn
                << std::endl;
    // unit scaling of this file's responses (reconstruction_cmd.cpp:194-210)
    const bool dl_is_time = cl.exist("dl-is-time");
    if (cl.exist("scale-length")) {
      const double scale = cl.get_double("scale-length");
      for (size_t i = before; i < events.size(); i += 3) {
        events[i] *= scale;
        events[i + 1] *= scale;
        if (!dl_is_time) events[i + 2] *= scale;
      }
    }
    if (dl_is_time) {
      const double scale = cl.get_double("scale-time");
      const double c = cl.get_double("speed-of-light");
      for (size_t i = before; i < events.size(); i += 3)
        events[i + 2] = events[i + 2] * c * scale;
    }
  }
  const int n_events = (int)(events.size() / 3);

  const int n_blocks = (int)cl.get_int("blocks");
  const int n_iterations_in_block = (int)cl.get_int("iterations");
  const int n_iterations = n_blocks * n_iterations_in_block;
  const std::string output_name = cl.text("output");
  const std::string base = strip_cli::path::wo_ext(output_name);
  const std::string ext = strip_cli::path::ext(output_name);
  const bool output_txt = ext == ".txt";
  const double png_max = cl.get_double("png-max");
  const bool gpu_protocol = cl.exist("gpu");
  strip_mlem_mode mode =
      cl.exist("legacy-gpu-kernel")
          ? PET2D::Strip::GPU::Reconstruction::legacy_mode()
          : PET2D::Strip::GPU::Reconstruction::cpu_mode();

  Progress progress(verbose, n_iterations);
  auto* responses = reinterpret_cast<Response*>(events.data());

  if (gpu_protocol) {
    PET2D::Strip::GPU::Reconstruction::run<float>(
        scanner, responses, n_events, rho, n_blocks, n_iterations_in_block,
        [&](int iteration, const Image& output) {
          if (base.empty()) return;
          const std::string fn =
              iteration >= 0 ? strip_cli::path::add_index(base, iteration,
                                                          n_iterations)
                             : base + "_sensitivity";
          strip_io::write_png(fn + ".png", output.data, output.width,
                              output.height);
          if (output_txt) {
            strip_io::write_matrix_text(fn + ".txt", output.data, output.width,
                                        output.height);
          } else if (ext != ".png") {
            strip_io::write_raw(fn + ext, output.data, (size_t)output.size);
          }
          strip_io::write_nrrd(fn + ".nrrd", strip_cli::path::wo_path(fn) + ext,
                               output.width, output.height, output_txt);
        },
        [&](int completed, bool finished) { progress(completed, finished); },
        (int)cl.get_int("cuda-device"), (int)cl.get_int("cuda-blocks"),
        (int)cl.get_int("cuda-threads"),
        [&](const char* device_name) {
          if (verbose) std::cerr << "# device: " << device_name << std::endl;
        },
        mode);
    return 0;
  }

  // the CPU path's protocol (reconstruction_cmd.cpp:264-292)
  strip_mlem_geometry geometry{scanner.radius, scanner.scintillator_length,
                               n_z, n_y, scanner.pixel_width,
                               scanner.pixel_height, scanner.sigma_z,
                               scanner.sigma_dl, 0.0, 0.0};
  strip_mlem_ctx* ctx = nullptr;
  auto check = [&](int rc) {
    if (rc != STRIP_MLEM_OK) {
      std::string msg = strip_mlem_last_error();
      if (ctx) strip_mlem_destroy(ctx);
      throw std::runtime_error(msg);
    }
  };
  check(strip_mlem_create(&geometry, &mode, (int)cl.get_int("cuda-device"),
                          &ctx));
  if (verbose) {
    char name[256];
    strip_mlem_device_name(ctx, name, sizeof(name));
    std::cerr << "# device: " << name << std::endl;
  }
  if (!base.empty()) {
    std::vector<float> sens((size_t)n_z * n_y);
    check(strip_mlem_get_sensitivity(ctx, 0, sens.data()));
    strip_io::write_png(base + "_sensitivity.png", sens.data(), n_z, n_y);
  }
  check(strip_mlem_set_rho(ctx, rho.data));
  check(strip_mlem_set_events(ctx, events.data(), n_events));
  for (int block = 0; block < n_blocks; ++block) {
    for (int i = 0; i < n_iterations_in_block; ++i) {
      const int iteration = block * n_iterations_in_block + i;
      progress(iteration, false);
      check(strip_mlem_iterate(ctx, 1));
      progress(iteration, true);
    }
    check(strip_mlem_get_rho(ctx, rho.data));
    if (base.empty()) continue;
    const std::string fn = strip_cli::path::add_index(
        base, (block + 1) * n_iterations_in_block, n_iterations);
    strip_io::write_png(fn + ".png", rho.data, n_z, n_y, png_max);
    if (output_txt) {
      strip_io::write_tuples_text(fn + ".txt", rho.data, n_z, n_y);
    } else if (ext != ".png") {
      strip_io::write_raw(fn + ext, rho.data, (size_t)rho.size);
      strip_io::write_nrrd(fn + ".nrrd", strip_cli::path::wo_path(fn) + ext,
                           n_z, n_y, false);
    }
  }
  strip_mlem_destroy(ctx);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& ex) {
    std::cerr << "error: " << ex.what() << std::endl;
  }
  return 1;
}
