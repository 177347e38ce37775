// strip_gpu_reconstruction.hpp — C++ drop-in for the reference's
//   src/2d/strip/cuda/reconstruction.h
// i.e. PET2D::Strip::GPU::Reconstruction::run<F>(...) (reconstruction.h:31-44,
// cuda/reconstruction.cu:31-113), implemented over the C ABI of
// strip_mlem.h (libstrip_mlem.so, HIP kernels for gfx950).
//
// Include it where the reference includes "cuda/reconstruction.h", after the
// headers that declare PET2D::Pixel and PET2D::PixelMap (the reference's
// src/2d/geometry/pixel_map.h, or this project's host/strip_types.hpp).  The
// call in src/2d/strip/reconstruction_cmd.cpp:225-261 then compiles
// unchanged: the scanner, the Response<F> array, the PixelMap image and the
// three callbacks are taken as they are.
//
// Differences from the CUDA original, all on purpose:
//  - the arithmetic follows the reference CPU path by default (GaussianKernel,
//    no sensitivity, guarded 1/D; reconstruction.h:220-298); pass
//    legacy_mode() as the last argument for the old CUDA semantics
//    (AnalyticKernel, 5-point sensitivity in the denominator, unguarded);
//  - errors throw std::runtime_error (the original printed file:line and
//    called abort(), src/util/cuda/debug.h:17-27); the reference CLI's
//    CMDLINE_CATCH prints it and exits with status 1;
//  - n_blocks / n_threads_per_block (-B/-W) keep their meaning for the bbox
//    kernel; the row-band kernels tile themselves (strip_mlem_set_band_tiling).
#pragma once

#include <stdexcept>
#include <string>
#include <type_traits>

#include "strip_mlem.h"

namespace PET2D {
namespace Strip {
namespace GPU {
namespace Reconstruction {

/// Image type passed to the output callback (cuda/reconstruction.h:27-29).
using Output = PET2D::PixelMap<PET2D::Pixel<short>, float>;

/// Reference CPU semantics (the default).
inline strip_mlem_mode cpu_mode() { return strip_mlem_mode{0, 0, 0, 1}; }
/// Semantics of the legacy CUDA warp kernel.
inline strip_mlem_mode legacy_mode() { return strip_mlem_mode{1, 1, 1, 0}; }

namespace detail {

template <typename OutputFn, typename ProgressFn, typename NameFn>
struct Callbacks {
  OutputFn* output;
  ProgressFn* progress;
  NameFn* device_name;

  static void on_output(void* user, int iteration, const float* image,
                        int n_z, int n_y) {
    auto* self = static_cast<Callbacks*>(user);
    // a non-owning view, valid during the call only (as in the original,
    // whose sensitivity image lived on the stack, cuda/reconstruction.cu:65)
    Output view(static_cast<short>(n_z), static_cast<short>(n_y),
                const_cast<float*>(image));
    (*self->output)(iteration, view);
  }
  static void on_progress(void* user, int completed, int finished) {
    auto* self = static_cast<Callbacks*>(user);
    (*self->progress)(completed, finished != 0);
  }
  static void on_device_name(void* user, const char* name) {
    auto* self = static_cast<Callbacks*>(user);
    (*self->device_name)(name);
  }
};

}  // namespace detail

/// GPU reconstruction entry point (cuda/reconstruction.h:31-44).
///  scanner        Scanner<F,S> (radius, scintillator_length, n_*_pixels,
///                 pixel_width/height, sigma_z/sigma_dl, center_y/z)
///  responses      n_responses Response<F> = {z_u, z_d, dl}
///  input_output   start image in, image after the last block out
///  output         called with (-1, sensitivity) first, then with the image
///                 after early iterations of block 0 and after every block
///  progress       (iteration, false) before / (iteration, true) after each
///                 iteration, then (total, false)
template <typename F, typename ScannerT, typename ResponseT,
          typename OutputFn, typename ProgressFn, typename NameFn>
void run(ScannerT& scanner, ResponseT* responses, int n_responses,
         Output& input_output, int n_iteration_blocks,
         int n_iterations_in_block, OutputFn output, ProgressFn progress,
         int device, int n_blocks, int n_threads_per_block,
         NameFn device_name, strip_mlem_mode mode = cpu_mode()) {
  static_assert(std::is_same<F, float>::value,
                "only the float instantiation exists (cuda/reconstruction.cu:128-141)");
  static_assert(sizeof(ResponseT) == 3 * sizeof(float),
                "Response<float> must be three packed floats (z_u, z_d, dl)");
  strip_mlem_geometry geometry;
  geometry.radius = scanner.radius;
  geometry.scintillator_length = scanner.scintillator_length;
  geometry.n_z_pixels = scanner.n_z_pixels;
  geometry.n_y_pixels = scanner.n_y_pixels;
  geometry.pixel_width = scanner.pixel_width;
  geometry.pixel_height = scanner.pixel_height;
  geometry.sigma_z = scanner.sigma_z;
  geometry.sigma_dl = scanner.sigma_dl;
  geometry.center_y = scanner.center_y;
  geometry.center_z = scanner.center_z;
  detail::Callbacks<OutputFn, ProgressFn, NameFn> callbacks{
      &output, &progress, &device_name};
  using CB = detail::Callbacks<OutputFn, ProgressFn, NameFn>;
  const int rc = strip_mlem_run(
      &geometry, &mode, reinterpret_cast<const float*>(responses),
      n_responses, input_output.data, n_iteration_blocks,
      n_iterations_in_block, &CB::on_output, &CB::on_progress, &callbacks,
      device, n_blocks, n_threads_per_block, &CB::on_device_name);
  if (rc != STRIP_MLEM_OK)
    throw std::runtime_error(std::string("GPU reconstruction failed: ") +
                             strip_mlem_last_error());
}

}  // namespace Reconstruction
}  // namespace GPU
}  // namespace Strip
}  // namespace PET2D
