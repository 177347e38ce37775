// Harness around the REFERENCE's own CPU strip reconstruction.
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile (target `ref`)
// directly from the reference's headers where they lie under
// /root/reference/src; the output goes to oracle/_ref/ (git-ignored).  No
// reference source is copied into this repository: this file only
// instantiates the reference templates and exposes them through extern "C"
// entry points so tests can compare the oracle restatement with them.
//
// Instantiated reference code:
//   PET2D::Strip::Reconstruction<F, GaussianKernel<F>>   reconstruction.h:32-344
//   PET2D::Strip::Scanner<F, short>                      scanner.h:26-291
//   PET2D::Strip::GaussianKernel<F>, AnalyticKernel<F>   gaussian_kernel.h, analytic_kernel.h
//   util::text_parser (fast .txt event loading)          util/text_parser.h:29-175
//   PET3D::Hybrid::Reconstruction<Scanner, GaussianKernel<float>>
//                                                        3d/hybrid/reconstruction.h:32-549
//   PET2D::Barrel::GeometrySOA (filled from caller arrays) 2d/barrel/geometry_soa.h:31-287
//   PET2D::Barrel::Geometry read from its binary file and flattened by
//   GeometrySOA(geometry, 1)                             2d/barrel/geometry.h:56-105,
//                                                        lor_geometry.h:58-84,
//                                                        geometry_soa.h:187-206
//   PET2D::PixelMap text/binary streams, util::nrrd_writer,
//   util::obstream / ibstream (image outputs and --rho inputs)
//                                                        pixel_map.h:113-157, nrrd_writer.h, bstream.h
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "2d/strip/reconstruction.h"
#include "2d/strip/gaussian_kernel.h"
#include "2d/strip/analytic_kernel.h"
#include "2d/strip/scanner.h"
#include "2d/geometry/pixel_map.h"
#include "common/types.h"
#include "util/bstream.h"
#include "util/nrrd_writer.h"
#include "2d/barrel/square_detector.h"
#include "2d/barrel/generic_scanner.h"
#include "3d/hybrid/scanner.h"
#include "3d/hybrid/reconstruction.h"
#include "2d/barrel/geometry.h"
#include "2d/barrel/geometry_soa.h"

#include "strip_oracle.h"  // oracle_geometry only (plain struct)

namespace {

struct NoProgress {
  void operator()(size_t, bool = false) {}
};

template <typename F>
PET2D::Strip::Scanner<F, short> make_scanner(const oracle_geometry* g) {
  // Same argument order as PET2D_STRIP_SCANNER_CL (options.h:42-50)
  return PET2D::Strip::Scanner<F, short>(
      (F)g->radius, (F)g->scintillator_length, (short)g->n_y_pixels,
      (short)g->n_z_pixels, (F)g->pixel_height, (F)g->pixel_width,
      (F)g->sigma_z, (F)g->sigma_dl, (F)g->center_y, (F)g->center_z);
}

template <typename F, typename E>
int run_mlem(const oracle_geometry* g, const E* events, int64_t n, F* rho,
             int n_iterations, int n_threads) {
#if _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  using Kernel = PET2D::Strip::GaussianKernel<F>;
  PET2D::Strip::Reconstruction<F, Kernel> rec(make_scanner<F>(g));
  rec.responses.reserve(n);
  for (int64_t e = 0; e < n; ++e)
    rec.responses.emplace_back((F)events[3 * e], (F)events[3 * e + 1],
                               (F)events[3 * e + 2]);
  const int64_t np = (int64_t)g->n_z_pixels * g->n_y_pixels;
  for (int64_t i = 0; i < np; ++i) rec.rho[i] = rho[i];
  NoProgress progress;
  rec(progress, n_iterations, 0);
  for (int64_t i = 0; i < np; ++i) rho[i] = rec.rho[i];
  return 0;
}

}  // namespace

extern "C" {

int ref_mlem_f32(const oracle_geometry* g, const float* events, int64_t n,
                 float* rho, int n_iterations, int n_threads) {
  return run_mlem<float>(g, events, n, rho, n_iterations, n_threads);
}

int ref_mlem_f64(const oracle_geometry* g, const double* events, int64_t n,
                 double* rho, int n_iterations, int n_threads) {
  return run_mlem<double>(g, events, n, rho, n_iterations, n_threads);
}

// Reconstruction ctor's sensitivity map (reconstruction.h:67-73)
int ref_sensitivity_map_f32(const oracle_geometry* g, float* out) {
  using Kernel = PET2D::Strip::GaussianKernel<float>;
  PET2D::Strip::Reconstruction<float, Kernel> rec(make_scanner<float>(g));
  const int64_t np = (int64_t)g->n_z_pixels * g->n_y_pixels;
  for (int64_t i = 0; i < np; ++i) out[i] = rec.sensitivity[i];
  return 0;
}

// Legacy --gpu sensitivity (cuda/reconstruction.cu:115-126 via
// Scanner::pixel_sensitivity, scanner.h:174-185), evaluated on the host.
int ref_pixel_sensitivity_map_f32(const oracle_geometry* g, float* out) {
  auto s = make_scanner<float>(g);
  for (int y = 0; y < g->n_y_pixels; ++y)
    for (int x = 0; x < g->n_z_pixels; ++x)
      out[(int64_t)y * g->n_z_pixels + x] =
          s.pixel_sensitivity(PET2D::Pixel<short>(x, y));
  return 0;
}

float ref_gaussian_kernel_f32(float sz, float sdl, float y, float tan,
                              float sec, float R, float dz, float dy) {
  PET2D::Strip::GaussianKernel<float> k(sz, sdl);
  return k(y, tan, sec, R, PET2D::Vector<float>(dz, dy));
}

double ref_gaussian_kernel_f64(double sz, double sdl, double y, double tan,
                               double sec, double R, double dz, double dy) {
  PET2D::Strip::GaussianKernel<double> k(sz, sdl);
  return k(y, tan, sec, R, PET2D::Vector<double>(dz, dy));
}

float ref_analytic_kernel_f32(float sz, float sdl, float y, float tan,
                              float sec, float R, float dz, float dy) {
  PET2D::Strip::AnalyticKernel<float> k(sz, sdl);
  return k(y, tan, sec, R, PET2D::Vector<float>(dz, dy));
}

double ref_analytic_kernel_f64(double sz, double sdl, double y, double tan,
                               double sec, double R, double dz, double dy) {
  PET2D::Strip::AnalyticKernel<double> k(sz, sdl);
  return k(y, tan, sec, R, PET2D::Vector<double>(dz, dy));
}

// Response::calculate_tan_y_z + GaussianKernel::ellipse_bb (float)
void ref_event_params_f32(const oracle_geometry* g, float z_u, float z_d,
                          float dl, float* out) {
  auto s = make_scanner<float>(g);
  PET2D::Strip::Response<float> r(z_u, z_d, dl);
  float tan, y, z;
  r.calculate_tan_y_z(s.radius, tan, y, z);
  PET2D::Strip::GaussianKernel<float> k(s.sigma_z, s.sigma_dl);
  float sec, A, B, C, bb_y, bb_z;
  k.ellipse_bb(tan, sec, A, B, C, bb_y, bb_z);
  float v[9] = { tan, y, z, sec, A, B, C, bb_y, bb_z };
  std::memcpy(out, v, sizeof(v));
}

void ref_pixel_at_f32(const oracle_geometry* g, float z, float y, int* px,
                      int* py) {
  auto s = make_scanner<float>(g);
  auto p = s.pixel_at(PET2D::Point<float>(z, y));
  *px = p.x;
  *py = p.y;
}

float ref_sensitivity_f32(const oracle_geometry* g, float z, float y) {
  auto s = make_scanner<float>(g);
  return s.sensitivity(PET2D::Point<float>(z, y));
}

// Fast text event loader (reconstruction.h:170-195, text_parser.h:140-175):
// returns the number of events; writes up to `cap` triples.
int64_t ref_load_txt_events(const char* fn, int is_3d, float* out,
                            int64_t cap) {
  oracle_geometry g = { 1, 1, 1, 1, 1, 1, 1, 1, 0, 0 };
  using Kernel = PET2D::Strip::GaussianKernel<float>;
  PET2D::Strip::Reconstruction<float, Kernel> rec(make_scanner<float>(&g));
  try {
    rec.fast_load_txt_events(fn, is_3d != 0);
  } catch (...) {
    return -1;
  }
  int64_t n = (int64_t)rec.responses.size();
  for (int64_t e = 0; e < n && e < cap; ++e) {
    out[3 * e] = rec.responses[e].z_u;
    out[3 * e + 1] = rec.responses[e].z_d;
    out[3 * e + 2] = rec.responses[e].dl;
  }
  return n;
}

// The image files 2d_strip_reconstruction writes, through the reference's
// own writers, into directory `dir`:
//   matrix.txt  PixelMap text stream, the --gpu protocol's .txt
//               (pixel_map.h:134-147, reconstruction_cmd.cpp:242-244)
//   tuples.txt  Reconstruction::output_tuples with precision 12, fixed: the
//               CPU protocol's .txt (reconstruction.h:203-213,
//               reconstruction_cmd.cpp:281-285)
//   image.bin   raw floats (pixel_map.h:154-157)
//   raw.nrrd    NRRD header of image.bin (nrrd_writer.h:19-28)
//   ascii.nrrd  NRRD header of matrix.txt, textual
int ref_write_images(const oracle_geometry* g, const float* img,
                     const char* dir) {
  using Image = PET2D::PixelMap<PET2D::Pixel<short>, float>;
  const short w = (short)g->n_z_pixels, h = (short)g->n_y_pixels;
  Image map(w, h);
  for (int i = 0; i < map.size; ++i) map[i] = img[i];
  const std::string d(dir);
  {
    std::ofstream txt(d + "/matrix.txt");
    txt << map;
  }
  {
    using Kernel = PET2D::Strip::GaussianKernel<float>;
    PET2D::Strip::Reconstruction<float, Kernel> rec(make_scanner<float>(g));
    for (int i = 0; i < map.size; ++i) rec.rho[i] = img[i];
    std::ofstream txt(d + "/tuples.txt");
    txt.precision(12);
    txt << std::fixed;
    rec.output_tuples(txt);
  }
  {
    util::obstream bin(d + "/image.bin");
    bin << map;
  }
  {
    util::nrrd_writer nrrd(d + "/raw.nrrd", "image.bin");
    nrrd << map;
  }
  {
    util::nrrd_writer nrrd(d + "/ascii.nrrd", "matrix.txt", true);
    nrrd << map;
  }
  return 0;
}

// A --rho start image read as reconstruction_cmd.cpp:145-165 reads it into
// the Reconstruction's rho (all ones, reconstruction.h:59): text through the
// PixelMap text stream, binary floats through ibstream, or int32 counts
// plus 1000 with --rho-is-int.
int ref_read_start_image(const oracle_geometry* g, const char* fn, int is_txt,
                         int is_int, float* out) {
  using Image = PET2D::PixelMap<PET2D::Pixel<short>, float>;
  const short w = (short)g->n_z_pixels, h = (short)g->n_y_pixels;
  Image rho(w, h, 1.0f);
  if (is_txt) {
    std::ifstream txt(fn);
    txt >> rho;
  } else {
    util::ibstream bin(fn);
    if (is_int) {
      PET2D::PixelMap<PET2D::Pixel<short>, Hit> int_rho(w, h);
      bin >> int_rho;
      for (int i = 0; i < int_rho.size; ++i) rho[i] = int_rho[i] + 1000;
    } else {
      bin >> rho;
    }
  }
  for (int i = 0; i < rho.size; ++i) out[i] = rho[i];
  return 0;
}

// 3D hybrid reconstruction (reconstruction_cmd.cpp:459-667 without the file
// handling) over a GeometrySOA filled from the caller's arrays:
//   seg: 4 floats per LOR (start x, y = the second detector's centre, end
//   x, y = the first's), widths: 1 per LOR, pixels: 2 int16 per pixel info,
//   positions: 1 per info, weights: n_infos x n_planes_half, info_begin /
//   info_end: per LOR (LOR index first (first + 1) / 2 + second).
// weights_mode 1 calls calculate_weight (no system matrix); sens_mode 0
// calculate_sensitivity, 1 set_sensitivity_to_one, 2 copies sens_in (depth
// planes); then normalize_geometry_weights, the events (first, second,
// z_up, z_dn, dl per event) through translate_to_frame and the grid filter,
// and n_iterations of operator().  rho (n_columns x n_rows x n_planes) is
// the start image and receives the result; sens_out (may be null) the
// sensitivity, weights_out (may be null) the normalised weights.
int ref_hybrid_mlem(int n_detectors, const float* seg, const float* widths,
                    int64_t n_infos, const int16_t* pixels,
                    const float* positions, const float* weights,
                    int n_planes_half, const int64_t* info_begin,
                    const int64_t* info_end, int n_columns, int n_rows,
                    float pixel_size, int n_planes, float z_left,
                    float sigma_z, float sigma_dl, float length,
                    int weights_mode, int sens_mode, int use_3d_sensitivity,
                    const float* sens_in, int sens_depth,
                    const float* events, int64_t n_events, float* rho,
                    int n_iterations, int n_threads, int64_t* n_used,
                    float* sens_out, float* weights_out) {
#if _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  using Detector = PET2D::Barrel::SquareDetector<float>;
  using Scanner2D = PET2D::Barrel::GenericScanner<Detector, short>;
  using Scanner = PET3D::Hybrid::Scanner<Scanner2D>;
  using Kernel = PET2D::Strip::GaussianKernel<float>;
  using Rec = PET3D::Hybrid::Reconstruction<Scanner, Kernel>;
  Scanner2D barrel;
  Scanner scanner(barrel, length);
  scanner.set_sigmas(sigma_z, sigma_dl);
  PET2D::PixelGrid<float, short> pixel_grid((short)n_columns, (short)n_rows,
                                            pixel_size);
  Rec::Grid grid(pixel_grid, z_left, (short)n_planes);
  Rec::Geometry geo((short)n_detectors, (size_t)n_infos,
                    (size_t)n_planes_half);
  for (size_t l = 0; l < geo.n_lors; ++l) {
    geo.lor_line_segments[l] = PET2D::LineSegment<float>(
        PET2D::Point<float>(seg[4 * l], seg[4 * l + 1]),
        PET2D::Point<float>(seg[4 * l + 2], seg[4 * l + 3]));
    geo.lor_widths[l] = widths[l];
    geo.lor_pixel_info_begin[l] = (size_t)info_begin[l];
    geo.lor_pixel_info_end[l] = (size_t)info_end[l];
  }
  for (int64_t i = 0; i < n_infos; ++i) {
    geo.pixels[i] = PET2D::Pixel<short>(pixels[2 * i], pixels[2 * i + 1]);
    geo.pixel_positions[i] = positions[i];
  }
  for (int64_t i = 0; i < n_infos * n_planes_half; ++i)
    geo.pixel_weights[i] = weights[i];
  Rec rec(scanner, grid, geo, use_3d_sensitivity != 0);
  if (weights_mode == 1) rec.calculate_weight();
  if (sens_mode == 2) {
    const int64_t n = (int64_t)rec.sensitivity.size;
    for (int64_t i = 0; i < n; ++i) rec.sensitivity[i] = sens_in[i];
    (void)sens_depth;
  } else if (sens_mode == 1) {
    rec.set_sensitivity_to_one();
  } else {
    rec.calculate_sensitivity();
  }
  rec.normalize_geometry_weights();
  std::stringstream in;
  for (int64_t e = 0; e < n_events; ++e) {
    // Response read as text: "first second z_up z_dn dl"
    // (3d/hybrid/scanner.h Response(std::istream&))
    in << (int)events[5 * e] << ' ' << (int)events[5 * e + 1] << ' ';
    in.precision(9);
    in << events[5 * e + 2] << ' ' << events[5 * e + 3] << ' '
       << events[5 * e + 4] << '\n';
  }
  rec << in;
  const int64_t nv = (int64_t)grid.n_voxels;
  for (int64_t i = 0; i < nv; ++i) rec.rho[i] = rho[i];
  int used = 0;
  for (int it = 0; it < n_iterations; ++it) used = rec();
  for (int64_t i = 0; i < nv; ++i) rho[i] = rec.rho[i];
  if (n_used) *n_used = n_iterations ? used : rec.n_events();
  if (sens_out)
    for (int i = 0; i < rec.sensitivity.size; ++i)
      sens_out[i] = rec.sensitivity[i];
  if (weights_out)
    for (int64_t i = 0; i < n_infos * n_planes_half; ++i)
      weights_out[i] = geo.pixel_weights[i];
  return 0;
}

// A PET2D::Barrel::Geometry<float, short> binary file (the reference's
// barrel geometry tool writes it) read by the reference's own constructor
// and flattened by GeometrySOA(geometry, 1).  Call first with null arrays:
// *n_detectors, *n_infos and the grid (n_columns, n_rows, pixel_size,
// lower-left x, y) are set; then with arrays of those sizes (n_lors x 4
// segments, n_lors widths, n_infos x 2 pixels, positions, weights, n_lors
// begin / end).  Returns -1 where the file cannot be read.
int ref_read_geometry(const char* path, int* n_detectors, int64_t* n_infos,
                      int* n_columns, int* n_rows, float* grid4,
                      float* seg, float* widths, int16_t* pixels,
                      float* positions, float* weights, int64_t* begin,
                      int64_t* end) {
  try {
    util::ibstream in(path);
    if (!in) return -1;
    PET2D::Barrel::Geometry<float, short> geometry(in);
    PET2D::Barrel::GeometrySOA<float, short> soa(geometry, 1);
    *n_detectors = soa.n_detectors;
    *n_infos = (int64_t)soa.n_pixel_infos;
    *n_columns = geometry.grid.n_columns;
    *n_rows = geometry.grid.n_rows;
    grid4[0] = geometry.grid.pixel_size;
    grid4[1] = geometry.grid.lower_left.x;
    grid4[2] = geometry.grid.lower_left.y;
    grid4[3] = 0.0f;
    if (!seg) return 0;
    for (size_t l = 0; l < soa.n_lors; ++l) {
      seg[4 * l] = soa.lor_line_segments[l].start.x;
      seg[4 * l + 1] = soa.lor_line_segments[l].start.y;
      seg[4 * l + 2] = soa.lor_line_segments[l].end.x;
      seg[4 * l + 3] = soa.lor_line_segments[l].end.y;
      widths[l] = soa.lor_widths[l];
      begin[l] = (int64_t)soa.lor_pixel_info_begin[l];
      end[l] = (int64_t)soa.lor_pixel_info_end[l];
    }
    for (size_t i = 0; i < soa.n_pixel_infos; ++i) {
      pixels[2 * i] = soa.pixels[i].x;
      pixels[2 * i + 1] = soa.pixels[i].y;
      positions[i] = soa.pixel_positions[i];
      weights[i] = soa.pixel_weights[i];
    }
    return 0;
  } catch (...) {
    return -1;
  }
}

}  // extern "C"
