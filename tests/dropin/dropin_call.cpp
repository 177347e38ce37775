// Compile check of the drop-in header against the reference's own headers:
// the --gpu call below has the shape of src/2d/strip/reconstruction_cmd.cpp
// (scanner, Response<float> vector, PixelMap image, output / progress /
// device-name lambdas, -D/-B/-W integers) and must compile unchanged with
// include/strip_gpu_reconstruction.hpp in place of cuda/reconstruction.h.
#include <iostream>
#include <vector>

#include "2d/geometry/pixel.h"
#include "2d/geometry/pixel_map.h"
#include "2d/strip/response.h"
#include "2d/strip/scanner.h"
#include "strip_gpu_reconstruction.hpp"

using F = float;
using Scanner = PET2D::Strip::Scanner<F, short>;
using Image = PET2D::PixelMap<PET2D::Pixel<short>, F>;

void reconstruct_on_gpu(Scanner& scanner,
                        std::vector<PET2D::Strip::Response<F>>& responses,
                        Image& rho, int n_blocks, int n_iterations_in_block,
                        int device, int cuda_blocks, int cuda_threads,
                        bool verbose, std::vector<float>& last_output) {
  PET2D::Strip::GPU::Reconstruction::run<F>(
      scanner, responses.data(), responses.size(), rho, n_blocks,
      n_iterations_in_block,
      [&](int iteration,
          const PET2D::Strip::GPU::Reconstruction::Output& output) {
        if (iteration < -1) return;
        last_output.assign(output.begin(), output.end());
      },
      [&](int completed, bool finished) {
        if (verbose) std::cerr << completed << (finished ? " done\n" : "\n");
      },
      device, cuda_blocks, cuda_threads,
      [=](const char* device_name) {
        if (verbose) std::cerr << "# device: " << device_name << std::endl;
      });
}
