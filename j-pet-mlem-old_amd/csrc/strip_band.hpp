// Row-band algorithm of the strip MLEM iteration (host interface).
//
// Each event is split into (event, band) pairs, a band being H consecutive
// image rows over the full image width.  Pairs are radix-sorted once per
// event upload by (band, bbox x0).  One iteration then runs
//   band_forward   partial D per pair; the rho band lives in LDS
//                  column-major (word = column*H + row), one image row per
//                  lane, so lanes never share an LDS bank;
//   band_inverse   D_e = sum of the event's partials, inv_e = 1/D_e (guarded);
//   band_backward  each wavefront sweeps x0-sorted pairs of one band through
//                  a private sliding window of W columns x H rows in LDS,
//                  accumulating K*inv_e with plain read-modify-write (no LDS
//                  atomics: lanes own distinct rows, waves own their
//                  window); columns leaving the window are added to the
//                  column-major back-projection image with coalesced
//                  atomics;
//   band_finalize  rho'(p) = rho(p) * bp(p) in column-major layout.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "strip_device.hpp"

namespace strip {

struct BandPlan {
  bool ready = false;
  int H = 0;          // rows per band (16, 32 or 64); lanes per row = 64/H
  int W = 0;          // backward window width in columns (power of two)
  int n_bands = 0;
  int max_bands = 0;  // max bands any event touches (partial slots/event)
  int64_t n_pairs = 0;
  int64_t n_events = 0;
  int* pair_event = nullptr;  // pair -> event, sorted by (band, x0)
  int4* fwd_tasks = nullptr;  // (band, begin, end, 0) per workgroup
  int n_fwd_tasks = 0;
  int4* bwd_tasks = nullptr;  // (band, begin, end, 0) per wavefront
  int n_bwd_tasks = 0;
  float* partial = nullptr;   // n_events * max_bands
  float* inv = nullptr;       // n_events
  float* rhoT[2] = {nullptr, nullptr};  // column-major images (n_z * n_y)
  int curT = 0;
  float* bpT = nullptr;
  float* sensT = nullptr;
  size_t fwd_lds = 0, bwd_lds = 0;
  int64_t pair_cap = 0, event_cap = 0;
};

// H for this grid, or 0 when the band algorithm cannot hold a band in LDS.
int band_rows_for(const DeviceGeometry& g);

// Builds the pair lists for the resident events.  Returns 0 or an error
// message in *err.
int band_prepare(BandPlan& plan, const DeviceGeometry& g, const EventSoA& ev,
                 int64_t n_events, hipStream_t stream, std::string* err);

void band_release(BandPlan& plan);

// Row-major <-> column-major image copies (sens may be null).
int band_load_images(BandPlan& plan, const DeviceGeometry& g,
                     const float* rho, const float* sens, hipStream_t stream,
                     std::string* err);
int band_store_image(BandPlan& plan, const DeviceGeometry& g, float* rho,
                     hipStream_t stream, std::string* err);

// One iteration.  t_fwd (optional) is recorded after the forward kernel,
// t_bwd after the back-projection kernel.
int band_iteration(BandPlan& plan, const DeviceGeometry& g,
                   const EventSoA& ev, int sensitivity_mode, int guard_zero,
                   hipStream_t stream, hipEvent_t t_fwd, hipEvent_t t_bwd,
                   std::string* err);

// Per-event forward projections of the current column-major image.
int band_forward_only(BandPlan& plan, const DeviceGeometry& g,
                      const EventSoA& ev, float* denominators,
                      hipStream_t stream, std::string* err);

// Kernel-evaluation count of the span walk (must equal the exact count).
int band_count(BandPlan& plan, const DeviceGeometry& g, const EventSoA& ev,
               unsigned long long* d_total, hipStream_t stream,
               std::string* err);

}  // namespace strip
