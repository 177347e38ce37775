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
// -e counts emissions, or detected events with --detected.  Additions:
// --seed (the generator's stream is Philox4x64-10 keyed by it) and
// --device.  Not available: --tan-bins / --kernel-point maps and the
// --full _events / _full_response files.
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
  cl.add("tan-bins", 0, Kind::Int, "", "tangent bins (not available)", true);
  cl.add("kernel-point", 0, Kind::Doubles, "", "kernel point (not available)",
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

void write_int_image(const std::string& base, const std::string& name,
                     const std::vector<int>& img, int n_z, int n_y) {
  const std::string raw = base + name;
  {
    std::ofstream out(raw, std::ios::binary);
    out.write(reinterpret_cast<const char*>(img.data()),
              sizeof(int) * img.size());
  }
  {
    std::ofstream nrrd(raw + ".nrrd");
    nrrd << "NRRD0001" << std::endl
         << "type: int" << std::endl
         << "endian: little" << std::endl
         << "dimension: 2" << std::endl
         << "sizes: " << n_z << " " << n_y << std::endl
         << "data file: ./" << strip_cli::path::wo_path(raw) << std::endl
         << "encoding: raw" << std::endl;
  }
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
    if (cl.exist("tan-bins") || cl.exist("kernel-point"))
      throw std::runtime_error("--tan-bins / --kernel-point are not available");
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

    std::vector<float> resp(3 * (size_t)std::max<int64_t>(n, 1));
    std::vector<float> exact(full ? resp.size() : 0);
    const size_t np = (size_t)n_z * n_y;
    std::vector<int> emitted(np), detected(np), naive(np);
    int64_t n_det = 0, n_em = 0;
    const int rc = strip_phantom_generate(
        (int)cl.get_int("device"), &params, regions.data(),
        (int)regions.size(), n, resp.data(), full ? exact.data() : nullptr,
        &n_det, &n_em, emitted.data(), detected.data(), naive.data());
    if (rc) throw std::runtime_error(strip_phantom_last_error());
    if (verbose)
      std::cerr << " emitted: " << n_em << " events" << std::endl
                << "detected: " << n_det << " events" << std::endl;

    const std::string output = cl.text("output");
    const std::string base = strip_cli::path::wo_ext(output);
    const std::string ext = strip_cli::path::ext(output);
    if (!base.empty() && !cl.exist("no-responses")) {
      auto write = [&](const std::string& fn, const std::vector<float>& v) {
        if (ext == ".txt") {
          std::ofstream out(fn);
          for (int64_t i = 0; i < n_det; ++i)
            out << v[3 * i] << ' ' << v[3 * i + 1] << ' ' << v[3 * i + 2]
                << "\n";
        } else {
          strip_io::write_raw(fn, v.data(), 3 * (size_t)n_det);
        }
      };
      write(output, resp);
      if (full) write(base + "_wo_error" + ext, exact);
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
    }
    return 0;
  } catch (const std::exception& ex) {
    std::cerr << "error: " << ex.what() << std::endl;
  } catch (const char* ex) {
    std::cerr << "error: " << ex << std::endl;
  }
  return 1;
}
