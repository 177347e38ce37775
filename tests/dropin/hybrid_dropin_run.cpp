// Run harness of the hybrid drop-in (test infrastructure): the reference's
// own command-line preparation and event reader (src/3d/hybrid/
// reconstruction_cmd.cpp:459-560 without the file handling: calculate_weight,
// calculate_sensitivity, normalize_geometry_weights, then the responses read
// through Reconstruction::operator<<, i.e. translate_to_frame and the grid
// filter, reconstruction.h:101-212) followed by the `--gpu` call of
// reconstruction_cmd.cpp:596-640 with include/hybrid_gpu_reconstruction.hpp
// in place of cuda/reconstruction.h.  Built by tests/dropin/Makefile from
// the reference's headers where the reference tree is present.
#include <cstdint>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "hybrid_gpu_reconstruction.hpp"

namespace {

namespace GPU = PET3D::Hybrid::GPU::Reconstruction;
using Rec = GPU::ReconstructionBase;
using Output = GPU::Output;

void copy_text(const std::string& s, char* out, int size) {
  if (!out || size <= 0) return;
  const size_t n = std::min(s.size(), (size_t)size - 1);
  std::memcpy(out, s.data(), n);
  out[n] = '\0';
}

}  // namespace

extern "C" {

// Geometry arrays as the oracle's harness takes them (4 floats per LOR
// segment, widths, 2 int16 per pixel info, positions, weights, per-LOR info
// begin / end); events as (first, second, z_up, z_dn, dl) per response.
// rho: start image in, final image out ((z n_rows + y) n_columns + x).  The
// run's outputs (iteration number and image) are kept up to max_outputs.
int dropin_hybrid_run(int n_detectors, const float* seg, const float* widths,
                      int64_t n_infos, const int16_t* pixels,
                      const float* positions, const float* weights,
                      const int64_t* info_begin, const int64_t* info_end,
                      int n_columns, int n_rows, float pixel_size,
                      int n_planes, float z_left, float sigma_z,
                      float sigma_dl, float length, const float* events,
                      int64_t n_events, int n_blocks,
                      int n_iterations_in_block, int start_iteration,
                      int semantics, float* rho, int max_outputs,
                      int* output_iterations, float* outputs, int* n_outputs,
                      int64_t* n_frame_events, char* device_name,
                      int name_size, char* error, int error_size) {
  try {
    GPU::Scanner2D barrel;
    GPU::Scanner scanner(barrel, length);
    scanner.set_sigmas(sigma_z, sigma_dl);
    PET2D::PixelGrid<float, short> pixel_grid((short)n_columns, (short)n_rows,
                                              pixel_size);
    Rec::Grid grid(pixel_grid, z_left, (short)n_planes);
    Rec::Geometry geometry_soa((short)n_detectors, (size_t)n_infos, 1);
    for (size_t l = 0; l < geometry_soa.n_lors; ++l) {
      geometry_soa.lor_line_segments[l] = PET2D::LineSegment<float>(
          PET2D::Point<float>(seg[4 * l], seg[4 * l + 1]),
          PET2D::Point<float>(seg[4 * l + 2], seg[4 * l + 3]));
      geometry_soa.lor_widths[l] = widths[l];
      geometry_soa.lor_pixel_info_begin[l] = (size_t)info_begin[l];
      geometry_soa.lor_pixel_info_end[l] = (size_t)info_end[l];
    }
    for (int64_t i = 0; i < n_infos; ++i) {
      geometry_soa.pixels[i] =
          PET2D::Pixel<short>(pixels[2 * i], pixels[2 * i + 1]);
      geometry_soa.pixel_positions[i] = positions[i];
      geometry_soa.pixel_weights[i] = weights[i];
    }
    Rec reconstruction(scanner, grid, geometry_soa, false);
    reconstruction.calculate_weight();
    reconstruction.calculate_sensitivity();
    reconstruction.normalize_geometry_weights();
    std::stringstream in;
    in.precision(9);
    for (int64_t e = 0; e < n_events; ++e)
      in << (int)events[5 * e] << ' ' << (int)events[5 * e + 1] << ' '
         << events[5 * e + 2] << ' ' << events[5 * e + 3] << ' '
         << events[5 * e + 4] << '\n';
    reconstruction << in;
    if (n_frame_events) *n_frame_events = reconstruction.n_events();
    const int64_t nv = (int64_t)grid.n_voxels;
    for (int64_t i = 0; i < nv; ++i) reconstruction.rho[i] = rho[i];
    int kept = 0;
    GPU::run(
        geometry_soa,
        reconstruction.sensitivity,
        reconstruction.events().data(),
        reconstruction.n_events(),
        reconstruction.scanner.sigma_z(),
        reconstruction.scanner.sigma_dl(),
        reconstruction.grid,
        scanner.length,
        n_blocks,
        n_iterations_in_block,
        reconstruction.rho,
        start_iteration,
        [&](int iteration, const Output& output) {
          if (kept >= max_outputs) return;
          output_iterations[kept] = iteration;
          for (int64_t i = 0; i < nv; ++i) outputs[kept * nv + i] = output[i];
          ++kept;
        },
        [&](int, bool) {},
        0, 0, 0,
        [&](const char* name) { copy_text(name, device_name, name_size); },
        semantics);
    if (n_outputs) *n_outputs = kept;
    for (int64_t i = 0; i < nv; ++i) rho[i] = reconstruction.rho[i];
    return 0;
  } catch (const std::exception& ex) {
    copy_text(ex.what(), error, error_size);
    return -1;
  }
}

}  // extern "C"
