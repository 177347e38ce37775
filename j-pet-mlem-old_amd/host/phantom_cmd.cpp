// 2d_strip_phantom — the reference's phantom simulation tool
// (src/2d/strip/phantom_cmd.cpp:136-444) on the MI355X event generator
// (include/strip_phantom.h).
//
// Same options (options.cpp:13-34,78-102), phantom description files
// ("ellipse x y a b angle intensity", "rectangle x y a b intensity",
// "point x y intensity", phantom.h:137-181) and outputs:
//   <output>                 detected responses with error (binary float
//                            triples; ".txt": "z_u z_d dl" lines)
//   <base>_wo_error<ext>     (--full) the same responses without error
//   <base>.cfg               the scanner options, for 2d_strip_reconstruction
//   <base>_emitted|_detected|_naive(.nrrd, .png)
//                            int32 images of emission points, detected
//                            emission points and points reconstructed from
//                            the responses with error (pixel_grid.h:59-69)
//   <base>_events<ext>       (--full) emission point and direction of each
//                            detected event ("z y dz dy"; binary: 4 floats)
//   <base>_full_response<ext>
//                            (--full) the exact responses
//   <base>_tan_bins(.nrrd)   (--tan-bins N) int32 counts of the detected
//                            events by the point and tangent reconstructed
//                            from their responses with error, N tangent
//                            planes of the image grid
//   <base>_tan_kernel(.nrrd) (--tan-bins N) float map of the analytic kernel
//                            at --kernel-point (default 0,0) for an event in
//                            each bin, times the bin's volume (and divided by
//                            the kernel point's sensitivity with --detected)
// -e counts emissions, or detected events with --detected.  Additions:
// --seed (the generator's stream is Philox4x64-10 keyed by it) and
// --device.  One error draw per detected event serves both the written
// response and the _naive / _tan_bins images (the reference draws the
// written response's error a second time, phantom_cmd.cpp:300-331; the
// distributions are the same).
#include <cmath>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "cmdline.hpp"
#include "io.hpp"
#include "strip_phantom.h"

namespace {

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
  cl.add("verbose", 'v', Kind::Flag, "", "print progress information");
  cl.add("n-threads", 'T', Kind::Int, "", "number of host threads (unused)");
  cl.add("scale", 0, Kind::Double, "1", "Scale phantom with given constant");
  cl.add("additive", 0, Kind::Flag, "",
         "phantom regions are additive, not disjunctive");
  cl.add("tan-bins", 0, Kind::Int, "",
         "Bin reconstructed events using given number of tangent bins", true);
  cl.add("kernel-point", 0, Kind::Doubles, "", "Point to evaluate kernel in",
         true);
  cl.add("full", 0, Kind::Flag, "", "Emit additional full information", true);
  cl.add("no-responses", 0, Kind::Flag, "", "Do not emit responses", true);
  cl.add("point", 0, Kind::Doubles, "", "Add point phantom x,y[,intensity]");
  cl.add("output", 'o', Kind::Path, "phantom.bin", "responses file", true);
  cl.add("detected", 0, Kind::Flag, "", "collects detected emissions");
  cl.add("image-origin", 0, Kind::Doubles, "",
         "Image origin x,y (default 0,0)", true);
  cl.add("seed", 0, Kind::Int, "0", "random stream seed");
  cl.add("device", 'D', Kind::Int, "0", "GPU device");
}

// phantom.h:137-181; positions and sizes times --scale
void read_phantom(const std::string& fn, double scale,
                  std::vector<strip_phantom_region>& out) {
  std::ifstream in(fn);
  if (!in) throw std::runtime_error("cannot open phantom description: " + fn);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream iss(line);
    std::string type;
    if (!(iss >> type)) continue;
    strip_phantom_region r{};
    if (type == "ellipse") {
      double x, y, a, b, angle, v;
      if (!(iss >> x >> y >> a >> b >> angle >> v)) break;
      r = {STRIP_REGION_ELLIPSE, x * scale, y * scale, a * scale, b * scale,
           angle, v};
    } else if (type == "rectangle") {
      double x, y, a, b, v;
      if (!(iss >> x >> y >> a >> b >> v)) break;
      r = {STRIP_REGION_RECTANGLE, x * scale, y * scale, a * scale, b * scale,
           0, v};
    } else if (type == "point") {
      double x, y, v;
      if (!(iss >> x >> y >> v)) break;
      r = {STRIP_REGION_POINT, x, y, 0, 0, 0, v};
    } else {
      throw std::runtime_error("unknown phantom type: " + type);
    }
    out.push_back(r);
  }
}

// Raw map plus its NRRD header; the header names the raw file as the
// reference's writer gets it, "./" + the output base with its directories
// (phantom_cmd.cpp:374-386,388-391).
// depth 0: a 2D image; depth >= 1: a 3D voxel map of `depth` planes.
template <typename T>
void write_map(const std::string& raw, const char* type, const T* data,
               int n_z, int n_y, int depth) {
  const size_t n = (size_t)n_z * n_y * (depth > 0 ? depth : 1);
  {
    std::ofstream out(raw, std::ios::binary);
    out.write(reinterpret_cast<const char*>(data), sizeof(T) * n);
  }
  std::ofstream nrrd(raw + ".nrrd");
  nrrd << "NRRD0001" << std::endl
       << "type: " << type << std::endl
       << "endian: little" << std::endl
       << "dimension: " << (depth > 0 ? 3 : 2) << std::endl
       << "sizes: " << n_z << " " << n_y;
  if (depth > 0) nrrd << " " << depth;
  nrrd << std::endl
       << "data file: ./" << raw << std::endl
       << "encoding: raw" << std::endl;
}

void write_int_image(const std::string& base, const std::string& name,
                     const std::vector<int>& img, int n_z, int n_y) {
  const std::string raw = base + name;
  write_map(raw, "int", img.data(), n_z, n_y, 0);
  std::vector<float> f(img.begin(), img.end());
  strip_io::write_png(raw + ".png", f.data(), n_z, n_y);
}

}  // namespace

int main(int argc, char** argv) {
  try {
    strip_cli::Parser cl;
    add_options(cl);
    cl.parse(argc, argv);
    if (cl.rest().empty() && !cl.exist("point")) {
      std::cerr << cl.usage("2d_strip_phantom");
      throw std::runtime_error(
          "at least one input phantom description expected, consult --help");
    }
    const int tan_bins = cl.exist("tan-bins") ? (int)cl.get_int("tan-bins") : 0;
    double kernel_z = 0, kernel_y = 0;
    if (cl.exist("kernel-point")) {
      const std::vector<double> k = cl.get_doubles("kernel-point");
      if (k.size() != 2)
        throw std::runtime_error(
            "--kernel-point must supply exactly 2 values for x, y");
      kernel_z = k[0];
      kernel_y = k[1];
    }
    // calculate_scanner_options (options.cpp:104-124)
    const std::vector<double> radius = cl.get_doubles("radius");
    if (radius.empty()) throw std::runtime_error("at least one --radius has to be specified");
    const double length = cl.get_double("length");
    const double s_pixel = cl.get_double("s-pixel");
    int n_z, n_y;
    if (cl.exist("n-pixels")) {
      n_z = n_y = (int)cl.get_int("n-pixels");
    } else {
      n_z = cl.exist("n-z-pixels") ? (int)cl.get_int("n-z-pixels")
                                   : (int)std::ceil(length / s_pixel);
      n_y = cl.exist("n-y-pixels") ? (int)cl.get_int("n-y-pixels")
                                   : (int)std::ceil(2 * radius[0] / s_pixel);
    }
    const int verbose = cl.count("verbose");
    const int64_t n = cl.get_int("n-emissions");
    const double scale = cl.get_double("scale");

    std::vector<strip_phantom_region> regions;
    for (const std::string& fn : cl.rest()) read_phantom(fn, scale, regions);
    if (cl.exist("point")) {
      const std::vector<double> p = cl.get_doubles("point");
      if (p.size() < 2 || p.size() > 3)
        throw std::runtime_error("--point must specify x,y[,intensity]");
      regions.push_back({STRIP_REGION_POINT, p[0], p[1], 0, 0, 0,
                         p.size() == 3 ? p[2] : 1.0});
    }
    if (regions.empty()) throw std::runtime_error("must specify at least one region");

    strip_phantom_params params{};
    params.radius = radius[0];
    params.scintillator_length = length;
    params.sigma_z = cl.get_double("s-z");
    params.sigma_dl = cl.get_double("s-dl");
    params.seed = (uint64_t)cl.get_int("seed");
    params.additive = cl.exist("additive") ? 1 : 0;
    params.only_detected = cl.exist("detected") ? 1 : 0;
    params.n_z = n_z;
    params.n_y = n_y;
    params.pixel = s_pixel;
    if (cl.exist("image-origin")) {
      const std::vector<double> o = cl.get_doubles("image-origin");
      if (o.size() != 2)
        throw std::runtime_error("--image-origin must specify two values for x,y");
      params.image_origin_z = o[0];
      params.image_origin_y = o[1];
    }
    const bool full = cl.exist("full");
    if (verbose)
      std::cerr << "    image size: " << n_z << " x " << n_y << std::endl
                << "    pixel size: " << s_pixel << " x " << s_pixel
                << std::endl;

    const size_t n1 = (size_t)std::max<int64_t>(n, 1);
    std::vector<float> resp(3 * n1);
    std::vector<float> exact(full ? 3 * n1 : 0), events(full ? 4 * n1 : 0);
    const size_t np = (size_t)n_z * n_y;
    std::vector<int> emitted(np), detected(np), naive(np);
    std::vector<int> tan_map(tan_bins > 0 ? np * tan_bins : 0);
    strip_phantom_outputs outs{};
    outs.responses = resp.data();
    outs.exact = full ? exact.data() : nullptr;
    outs.events = full ? events.data() : nullptr;
    outs.image_emitted = emitted.data();
    outs.image_detected = detected.data();
    outs.image_detected_w_error = naive.data();
    outs.tan_bins = tan_bins > 0 ? tan_bins : 0;
    outs.image_tan_bins = tan_bins > 0 ? tan_map.data() : nullptr;
    const int device = (int)cl.get_int("device");
    int rc = strip_phantom_generate_full(device, &params, regions.data(),
                                         (int)regions.size(), n, &outs);
    if (rc) throw std::runtime_error(strip_phantom_last_error());
    const int64_t n_det = outs.n_detected, n_em = outs.n_emitted;
    if (verbose)
      std::cerr << " emitted: " << n_em << " events" << std::endl
                << "detected: " << n_det << " events" << std::endl;

    const std::string output = cl.text("output");
    const std::string base = strip_cli::path::wo_ext(output);
    const std::string ext = strip_cli::path::ext(output);
    if (!base.empty() && !cl.exist("no-responses")) {
      // k floats per detected event: "a b c[ d]" lines or raw floats
      auto write = [&](const std::string& fn, const std::vector<float>& v,
                       int k) {
        if (ext == ".txt") {
          std::ofstream out(fn);
          for (int64_t i = 0; i < n_det; ++i) {
            for (int c = 0; c < k; ++c)
              out << (c ? " " : "") << v[(size_t)k * i + c];
            out << "\n";
          }
        } else {
          strip_io::write_raw(fn, v.data(), (size_t)k * n_det);
        }
      };
      write(output, resp, 3);
      if (full) {
        write(base + "_wo_error" + ext, exact, 3);
        write(base + "_events" + ext, events, 4);
        write(base + "_full_response" + ext, exact, 3);
      }
    }
    if (!base.empty()) {
      // <base>.cfg: the scanner options for the reconstruction tool
      std::ofstream cfg(base + ".cfg");
      cfg << "--radius=" << radius[0] << "\n"
          << "--length=" << length << "\n"
          << "--s-pixel=" << s_pixel << "\n"
          << "--n-z-pixels=" << n_z << "\n"
          << "--n-y-pixels=" << n_y << "\n"
          << "--s-z=" << params.sigma_z << "\n"
          << "--s-dl=" << params.sigma_dl << "\n";
      write_int_image(base, "_detected", detected, n_z, n_y);
      write_int_image(base, "_emitted", emitted, n_z, n_y);
      write_int_image(base, "_naive", naive, n_z, n_y);
      if (tan_bins > 0) {
        write_map(base + "_tan_bins", "int", tan_map.data(), n_z, n_y,
                  tan_bins);
        if (verbose) std::cerr << "generating kernel map..." << std::endl;
        std::vector<float> kmap(np * tan_bins);
        rc = strip_phantom_tan_kernel_map(device, &params, tan_bins, kernel_z,
                                          kernel_y, kmap.data());
        if (rc) throw std::runtime_error(strip_phantom_last_error());
        write_map(base + "_tan_kernel", "float", kmap.data(), n_z, n_y,
                  tan_bins);
      }
    }
    return 0;
  } catch (const std::exception& ex) {
    std::cerr << "error: " << ex.what() << std::endl;
  } catch (const char* ex) {
    std::cerr << "error: " << ex << std::endl;
  }
  return 1;
}
