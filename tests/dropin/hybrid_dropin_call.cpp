// Compile check of the hybrid drop-in header against the reference's own
// headers: the --gpu call below has the shape of
// src/3d/hybrid/reconstruction_cmd.cpp:596-640 (GeometrySOA, the
// Reconstruction's sensitivity, frame events, voxel grid and image, the
// output / progress / device-name lambdas, -D/-B/-W integers) and must
// compile unchanged with include/hybrid_gpu_reconstruction.hpp in place of
// cuda/reconstruction.h.
#include <iostream>
#include <vector>

#include "hybrid_gpu_reconstruction.hpp"

using Rec = PET3D::Hybrid::GPU::Reconstruction::ReconstructionBase;
using Output = PET3D::Hybrid::GPU::Reconstruction::Output;

void reconstruct_on_gpu(Rec::Scanner& scanner, Rec& reconstruction,
                        Rec::Geometry& geometry_soa, int n_blocks,
                        int n_iterations_in_block, int start_iteration,
                        int device, int cuda_blocks, int cuda_threads,
                        bool verbose, std::vector<float>& last_output) {
  PET3D::Hybrid::GPU::Reconstruction::run(
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
