// Device side of the event generator (strip_phantom.hip), shared with the
// context's in-place generation (strip_mlem.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "strip_device.hpp"
#include "strip_phantom.h"

namespace strip {

constexpr int64_t kPhantomBatch = (int64_t)1 << 22;  // emissions per launch

struct DevRegion {
  int kind;
  float cx, cy, a, b;
  float sin_t, cos_t;  // rotation of the ellipse
  float qa, qb, qc;    // ellipse test a dx^2 + 2 c dx dy + b dy^2 <= 1
  float cdf;           // cumulative weight / total
};

struct PhantomKernelArgs {
  const DevRegion* regions;
  int n_regions;
  int additive;
  uint64_t seed;
  float radius, half_length, sigma_z, sigma_dl;
  int img_nz, img_ny;
  float img_pixel, img_z0, img_y0;
  // tangent bins of the points reconstructed with error (0: none): plane =
  // floor((tan - tan_left) / tan_voxel)
  int tan_bins;
  float tan_left, tan_voxel;
};

struct PhantomDevice {
  DevRegion* regions = nullptr;
  PhantomKernelArgs args{};
  int64_t next = 0;     // next emission index
  int64_t emitted = 0;  // emissions used so far
  // batch buffers
  uint8_t* flags = nullptr;
  float* resp = nullptr;
  float* exact = nullptr;
  int* pix_origin = nullptr;
  int* pix_err = nullptr;
  int* pix_tan = nullptr;
  float* ev4 = nullptr;      // emission point and direction per emission
  float* sel_ev4 = nullptr;
  float* sel_resp = nullptr;
  float* sel_exact = nullptr;
  int64_t* sel_index = nullptr;
  int64_t* index = nullptr;
  int64_t* count = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
};

// FP32 Scanner and kernel constants of a geometry (strip_mlem.hip).
DeviceGeometry make_device_geometry(const strip_mlem_geometry& geo);

// Tangent axis of the --tan-bins grid: first plane's lower edge and width.
void phantom_tan_axis(double radius, double length, int tan_bins,
                      float* left, float* voxel);

// Regions to the device; the scanner values given here (not those in p) are
// the ones used, so a context can impose its own.
int phantom_prepare(PhantomDevice& P, const strip_phantom_params& p,
                    const strip_phantom_region* regions, int n_regions,
                    double radius, double length, double sigma_z,
                    double sigma_dl, std::string* err);
void phantom_release(PhantomDevice& P);

// Emissions from P.next on, until `need` detected responses are produced or
// max_emissions emissions are used; responses (AoS triples, emission order)
// go to d_resp (and d_exact without error, d_events the emission point and
// direction, 4 floats each); images (device, may be null) count the
// emissions used, img_tan (tan_bins planes of n_y x n_z) the detected ones by
// the point and tangent reconstructed from the response with error.
int phantom_detect(PhantomDevice& P, int64_t need, int64_t max_emissions,
                   float* d_resp, float* d_exact, int* img_em, int* img_det,
                   int* img_err, int64_t* produced, hipStream_t stream,
                   std::string* err, float* d_events = nullptr,
                   int* img_tan = nullptr);

}  // namespace strip
