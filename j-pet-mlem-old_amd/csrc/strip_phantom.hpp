// Device side of the event generator (strip_phantom.hip), shared with the
// context's in-place generation (strip_mlem.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

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
  float* sel_resp = nullptr;
  float* sel_exact = nullptr;
  int64_t* sel_index = nullptr;
  int64_t* index = nullptr;
  int64_t* count = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
};

// Regions to the device; the scanner values given here (not those in p) are
// the ones used, so a context can impose its own.
int phantom_prepare(PhantomDevice& P, const strip_phantom_params& p,
                    const strip_phantom_region* regions, int n_regions,
                    double radius, double length, double sigma_z,
                    double sigma_dl, std::string* err);
void phantom_release(PhantomDevice& P);

// Emissions from P.next on, until `need` detected responses are produced or
// max_emissions emissions are used; responses (AoS triples, emission order)
// go to d_resp (and d_exact without error); images (device, may be null)
// count the emissions used.
int phantom_detect(PhantomDevice& P, int64_t need, int64_t max_emissions,
                   float* d_resp, float* d_exact, int* img_em, int* img_det,
                   int* img_err, int64_t* produced, hipStream_t stream,
                   std::string* err);

}  // namespace strip
