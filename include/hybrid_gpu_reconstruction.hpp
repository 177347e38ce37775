// hybrid_gpu_reconstruction.hpp — C++ drop-in for the reference's
//   src/3d/hybrid/cuda/reconstruction.h
// i.e. PET3D::Hybrid::GPU::Reconstruction::run(...) (reconstruction.h:
// 36-55, cuda/reconstruction.cu:36-187), implemented over the C ABI of
// hybrid_mlem.h (libstrip_mlem.so, HIP kernels for gfx950).
//
// Include it where src/3d/hybrid/reconstruction_cmd.cpp includes
// "cuda/reconstruction.h", with the reference's src/ on the include path.
// The call in reconstruction_cmd.cpp:596-640 then compiles unchanged: the
// GeometrySOA with its normalised weights, the sensitivity map, the frame
// events of the CPU reader, the voxel grid, the image and the callbacks are
// taken as they are.
//
// Differences from the CUDA original, all on purpose:
//  - the update follows the reference CPU path by default (a voxel that one
//    event visits twice gets the weight cached last on both visits,
//    reconstruction.h:347-369,417-429); pass HYBRID_MLEM_LEGACY_GPU as the
//    last argument for the CUDA kernel's own-weight update;
//  - errors throw std::runtime_error (the original aborted on CUDA errors);
//  - n_blocks / n_threads_per_block are not used: the kernels size their
//    own launches;
//  - single-plane weights and a 2D sensitivity map only (the command line's
//    default); the others are refused.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "common/types.h"
#include "util/delegate.h"

#include "2d/barrel/generic_scanner.h"
#include "2d/barrel/geometry_soa.h"
#include "2d/barrel/square_detector.h"
#include "2d/strip/gaussian_kernel.h"
#include "3d/geometry/point.h"
#include "3d/geometry/voxel.h"
#include "3d/geometry/voxel_grid.h"
#include "3d/geometry/voxel_map.h"
#include "3d/hybrid/reconstruction.h"
#include "3d/hybrid/scanner.h"

#include "hybrid_mlem.h"

namespace PET3D {
namespace Hybrid {
namespace GPU {
namespace Reconstruction {

/// \cond PRIVATE
using Voxel = PET3D::Voxel<S>;
using Detector = PET2D::Barrel::SquareDetector<F>;
using Scanner2D = PET2D::Barrel::GenericScanner<Detector, S>;
using Scanner = PET3D::Hybrid::Scanner<Scanner2D>;
using Kernel = PET2D::Strip::GaussianKernel<F>;
using ReconstructionBase = PET3D::Hybrid::Reconstruction<Scanner, Kernel>;
using Event = ReconstructionBase::FrameEvent;
using Geometry = PET2D::Barrel::GeometrySOA<F, S>;
using Output = VoxelMap<Voxel, F>;
using Grid = PET3D::VoxelGrid<F, S>;
/// \endcond

namespace detail {

inline void check(int status) {
  if (status != STRIP_MLEM_OK)
    throw std::runtime_error(std::string("hybrid MLEM: ") +
                             strip_mlem_last_error());
}

struct CtxDeleter {
  void operator()(hybrid_mlem_ctx* c) const { hybrid_mlem_destroy(c); }
};

}  // namespace detail

/// The reconstruction of cuda/reconstruction.cu:36-187 on the MI355X.
inline void run(const Geometry& geometry,
                const Output& sensitivity,
                const Event* events,
                int n_events,
                F sigma_z,
                F sigma_dl,
                const Grid& grid,
                const F barrel_length,
                int n_iteration_blocks,
                int n_iterations_in_block,
                const Output& rho_output,
                const int start_iteration,
                util::delegate<void(int iteration, const Output& output)>
                    output,
                util::delegate<void(int completed, bool finished)> progress,
                int device,
                int n_blocks,
                int n_threads_per_block,
                util::delegate<void(const char* device_name)> info,
                int semantics = HYBRID_MLEM_CPU) {
  (void)n_blocks;
  (void)n_threads_per_block;
  const auto& pg = grid.pixel_grid;
  if (geometry.n_planes_half != 1 || sensitivity.depth != 1)
    throw std::runtime_error(
        "hybrid MLEM: only single-plane weights with a 2D sensitivity map");
  if (pg.lower_left.x != -pg.pixel_size * pg.n_columns / 2 ||
      pg.lower_left.y != -pg.pixel_size * pg.n_rows / 2)
    throw std::runtime_error("hybrid MLEM: the pixel grid must be centred");

  // geometry arrays in the C ABI's plain layout
  const size_t n_lors = geometry.n_lors;
  const size_t n_infos = geometry.n_pixel_infos;
  std::vector<float> segments(4 * n_lors), widths(n_lors);
  std::vector<int64_t> begin(n_lors), end(n_lors);
  for (size_t l = 0; l < n_lors; ++l) {
    const auto& s = geometry.lor_line_segments[l];
    segments[4 * l] = s.start.x;
    segments[4 * l + 1] = s.start.y;
    segments[4 * l + 2] = s.end.x;
    segments[4 * l + 3] = s.end.y;
    widths[l] = geometry.lor_widths[l];
    begin[l] = static_cast<int64_t>(geometry.lor_pixel_info_begin[l]);
    end[l] = static_cast<int64_t>(geometry.lor_pixel_info_end[l]);
  }
  std::vector<int16_t> pixels(2 * n_infos);
  for (size_t i = 0; i < n_infos; ++i) {
    pixels[2 * i] = geometry.pixels[i].x;
    pixels[2 * i + 1] = geometry.pixels[i].y;
  }
  hybrid_mlem_geometry g;
  g.n_detectors = geometry.n_detectors;
  g.n_pixel_infos = static_cast<int64_t>(n_infos);
  g.n_planes_half = 1;
  g.lor_line_segments = segments.data();
  g.lor_widths = widths.data();
  g.pixels = pixels.data();
  g.pixel_positions = geometry.pixel_positions;
  g.pixel_weights = geometry.pixel_weights;
  g.lor_pixel_info_begin = begin.data();
  g.lor_pixel_info_end = end.data();
  g.n_columns = pg.n_columns;
  g.n_rows = pg.n_rows;
  g.n_planes = grid.n_planes;
  g.pixel_size = pg.pixel_size;
  g.z_left = grid.z_left;
  g.sigma_z = sigma_z;
  g.sigma_dl = sigma_dl;
  g.length = barrel_length;

  hybrid_mlem_ctx* raw = nullptr;
  detail::check(hybrid_mlem_create(&g, semantics, device, &raw));
  std::unique_ptr<hybrid_mlem_ctx, detail::CtxDeleter> ctx(raw);
  // the CLI normalised the weights and computed the sensitivity already
  detail::check(hybrid_mlem_prepare_weights(
      ctx.get(), 0, HYBRID_MLEM_SENSITIVITY_NORMALISED, sensitivity.data));

  std::vector<hybrid_mlem_frame_event> frames(n_events > 0 ? n_events : 0);
  for (int i = 0; i < n_events; ++i) {
    const Event& e = events[i];
    hybrid_mlem_frame_event& f = frames[i];
    f.first = e.lor.first;
    f.second = e.lor.second;
    f.up = e.up;
    f.right = e.right;
    f.tan = e.tan;
    f.sec = e.sec;
    f.info_begin = static_cast<int64_t>(e.pixel_info_begin);
    f.info_end = static_cast<int64_t>(e.pixel_info_end);
    f.plane_begin = e.plane_begin;
    f.plane_end = e.plane_end;
  }
  detail::check(
      hybrid_mlem_set_frame_events(ctx.get(), frames.data(), n_events));

  char name[256];
  detail::check(hybrid_mlem_device_name(ctx.get(), name, sizeof(name)));
  info(name);

  // the image is read from and written back into rho_output's buffer, as
  // the original's output_rho does (cuda/reconstruction.cu:89-90,178)
  float* image = const_cast<float*>(rho_output.data);
  detail::check(hybrid_mlem_set_rho(ctx.get(), image));

  for (int block = 0; block < n_iteration_blocks; ++block) {
    for (int i = 0; i < n_iterations_in_block; ++i) {
      const int iteration = block * n_iterations_in_block + i;
      if (iteration < start_iteration) continue;
      progress(iteration, false);
      detail::check(hybrid_mlem_iterate(ctx.get(), 1));
      progress(iteration, true);
      // the first block also outputs iterations 1-5, 10, 15, 20, 30, 50,
      // 100 (cuda/reconstruction.cu:165-172)
      if (!block && i < n_iterations_in_block - 1 &&
          (i < 5 || i == 9 || i == 14 || i == 19 || i == 29 || i == 49 ||
           i == 99)) {
        detail::check(hybrid_mlem_get_rho(ctx.get(), image));
        output(i + 1, rho_output);
      }
    }
    const int iteration = (block + 1) * n_iterations_in_block;
    if (iteration <= start_iteration) continue;
    detail::check(hybrid_mlem_get_rho(ctx.get(), image));
    output(iteration, rho_output);
  }
  progress(n_iteration_blocks * n_iterations_in_block, false);
}

}  // namespace Reconstruction
}  // namespace GPU
}  // namespace Hybrid
}  // namespace PET3D
