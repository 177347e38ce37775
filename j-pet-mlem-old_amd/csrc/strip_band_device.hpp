// Device helpers shared by the row-band kernels (strip_band.hip) and the
// adjoint kernels (strip_adjoint.hip): exact row runs of the 3-sigma ellipse,
// the span-table word, DPP wavefront sums and the batched pair gather.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "strip_band.hpp"
#include "strip_device.hpp"

namespace strip {
namespace {

constexpr int kFwdWaves = 8;  // wavefronts per forward workgroup
constexpr int kSpanAlign = 64;  // span-table slots start on such rows

// Pointers that a kernel reads from a descriptor in memory are generic, and
// accesses through them compile to flat instructions, which count against
// the LDS wait counter as well: every LDS wait would then also wait for the
// outstanding global loads.  The hot accesses go through global-space
// pointers instead (with native vector types: a copy of a HIP vector struct
// through such a pointer falls back to flat loads).
template <typename T>
using global_t = __attribute__((address_space(1))) T;
typedef float gfloat4 __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ T gload(const T* p) {
  return *(const global_t<T>*)p;
}
template <typename T>
__device__ __forceinline__ void gstore(T* p, T v) {
  *(global_t<T>*)p = v;
}
// relaxed, device scope, no return value
__device__ __forceinline__ void gatomic_add(float* p, float v) {
  __hip_atomic_fetch_add((global_t<float>*)p, v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
}

#define BAND_TRY(expr)                                              \
  do {                                                              \
    hipError_t _e = (expr);                                         \
    if (_e != hipSuccess) {                                         \
      if (err) *err = std::string(#expr) + ": " + hipGetErrorString(_e); \
      return -1;                                                    \
    }                                                               \
  } while (0)

// Exact per-row terms of the 3-sigma test (gaussian_kernel.h:131-137):
// Q(dz) = ((A dy) dy + (B dy) dz) + (C dz) dz, evaluated like the reference.
__device__ inline void row_terms_exact(float A, float B, float dy,
                                       float& rowA, float& bdy) {
#pragma clang fp contract(off)
  rowA = (A * dy) * dy;
  bdy = B * dy;
}

__device__ inline float row_q_exact(const DeviceGeometry& g, float cz,
                                    float C, float rowA, float bdy, int col) {
#pragma clang fp contract(off)
  float dz = pixel_center_z(g, col) - cz;
  return (rowA + bdy * dz) + (C * dz * dz);
}

// The pixels of one image row inside an event's support: the row's columns
// in the clamped bbox [x0, x1) whose centre passes Q <= 9.  The interval is
// found from the real roots of Q(dz) = 9; where a root lies within 2e-3
// pixel of a pixel centre (or the span is narrower than a pixel) the end
// pixels are re-tested with the exact FP32 form, so the set equals the
// reference's (reconstruction.h:255-264) pixel for pixel.
struct RowSpan {
  int lo;      // first column
  int n;       // number of columns (0 = empty)
  float rowA;  // (A dy) dy
  float bdy;   // B dy
  float pre;   // c / sqrt(a . S^-1 . a) of this row
  float tdy;   // tan dy: the row's Q is least at dz = tan dy
  float qd;    // 4 id sec^2 dy^2: the row's least Q
};

// The kernel's exponent along a row in vertex form,
//   Q(dz) = C (dz - tan dy)^2 + 4 id sec^2 dy^2,
// the sum of squares b . S^-1 . b of gaussian_kernel.h:51-79 (C = 2 iz).
// Both terms are non-negative, so FP32 keeps its relative precision where
// the expanded form A dy^2 + B dy dz + C dz^2 cancels (thin, steep
// ellipses).  tan = -B / (2C) and sec^2 = 1 + tan^2 from the event's B.
__device__ inline void row_vertex(const DeviceGeometry& g, float B, float dy,
                                  float& tdy, float& qd) {
  const float t = -B * g.inv_2c;
  const float ad = 4.0f * g.inv_pow_sigma_dl * __builtin_fmaf(t, t, 1.0f);
  tdy = t * dy;
  qd = ad * dy * dy;
}

__device__ inline RowSpan row_span(const DeviceGeometry& g, float cz,
                                   float cy, float A, float B, float P,
                                   float Q0, int x0, int x1, int y0, int y1,
                                   int iy) {
  const float C = g.C, inv_2c = g.inv_2c, inv_w = g.inv_pixel_width;
  RowSpan s;
  s.lo = 0;
  s.n = 0;
  s.rowA = 0.f;
  s.bdy = 0.f;
  s.pre = 0.f;
  s.tdy = 0.f;
  s.qd = 0.f;
  if (iy < y0 || iy >= y1) return s;
  const float py = pixel_center_y(g, iy);
  const float dy = py - cy;
  row_terms_exact(A, B, dy, s.rowA, s.bdy);
  row_vertex(g, B, dy, s.tdy, s.qd);
  const float yp = dy + cy;
  s.pre = __builtin_amdgcn_rsqf(__builtin_fmaf(P, yp * yp, Q0));
  const float disc = s.bdy * s.bdy - 4.0f * C * (s.rowA - 9.0f);
  if (disc < 0.0f) return s;
  const float half = __builtin_sqrtf(disc) * inv_2c;
  const float mid = cz - s.bdy * inv_2c;
  float fl = (mid - half - g.tl_z_half_w) * inv_w;
  float fh = (mid + half - g.tl_z_half_w) * inv_w;
  fl = fminf(fmaxf(fl, (float)x0 - 2.0f), (float)x1 + 1.0f);
  fh = fminf(fmaxf(fh, (float)x0 - 2.0f), (float)x1 + 1.0f);
  int lo = (int)ceilf(fl);
  int hi = (int)floorf(fh);
  const float eps = 2e-3f;
  const bool check = (fl - floorf(fl) < eps) || (ceilf(fl) - fl < eps) ||
                     (fh - floorf(fh) < eps) || (ceilf(fh) - fh < eps) ||
                     (half * inv_w < 1.0f);
  lo = max(lo, x0);
  hi = min(hi, x1 - 1);
  if (check) {
    if (lo - 1 >= x0 && row_q_exact(g, cz, C, s.rowA, s.bdy, lo - 1) <= 9.0f)
      lo -= 1;
    else if (lo <= hi && row_q_exact(g, cz, C, s.rowA, s.bdy, lo) > 9.0f)
      lo += 1;
    if (hi + 1 <= x1 - 1 &&
        row_q_exact(g, cz, C, s.rowA, s.bdy, hi + 1) <= 9.0f)
      hi += 1;
    else if (hi >= lo && row_q_exact(g, cz, C, s.rowA, s.bdy, hi) > 9.0f)
      hi -= 1;
  }
  s.lo = lo;
  s.n = max(hi - lo + 1, 0);
  return s;
}

// The bands of an event's bbox rows [y0, y1) (none for an empty box).  A
// forward pair list aligns its bands to the event: band k holds rows
// y0 + k H .. y0 + k H + H - 1, so only the last band is partial.  The
// back-projection's bands are the image's, rows b H .. b H + H - 1.  Either
// way band k belongs to the tasks of image band y0 / H + k, the band of its
// first row.
__device__ inline int band_count(int y0, int y1, bool has_columns, int H,
                                 bool aligned) {
  if (!has_columns || y1 <= y0) return 0;
  return aligned ? (y1 - y0 + H - 1) / H : (y1 - 1) / H - y0 / H + 1;
}

// First row of the event's band in image band b (H a power of two).
__device__ inline int band_first_row(int y0, int b, int H, bool aligned) {
  return b * H + (aligned ? (y0 & (H - 1)) : 0);
}

__device__ inline void record_box(const BandRecord& rec, int& x0, int& x1,
                                  int& y0, int& y1) {
  x0 = rec.bx & 0xffff;
  x1 = rec.bx >> 16;
  y0 = rec.by & 0xffff;
  y1 = rec.by >> 16;
}

// Row iy's run from a span-table word, with the row's exact terms computed as
// in row_span.
__device__ inline RowSpan stored_span(const DeviceGeometry& g, uint32_t word,
                                      float cy, float A, float B, float P,
                                      float Q0, int x0, int iy) {
  RowSpan s;
  s.n = (int)(word >> 8);
  s.lo = x0 + (int)(word & 0xff);
  const float py = pixel_center_y(g, iy);
  const float dy = py - cy;
  row_terms_exact(A, B, dy, s.rowA, s.bdy);
  row_vertex(g, B, dy, s.tdy, s.qd);
  const float yp = dy + cy;
  s.pre = __builtin_amdgcn_rsqf(__builtin_fmaf(P, yp * yp, Q0));
  return s;
}

// Span-table word of row iy of record e.  A record's slot holds the rows
// from its bbox top rounded down to a multiple of kSpanAlign (rows outside
// the bbox hold 0, an empty run), so every band of the record (H divides
// kSpanAlign) reads within the slot and a 16-row band is one aligned 32-byte
// piece.  The index is clamped into the slot so the load is unconditional
// and always in bounds (keeping the load free of conditions lets it stay in
// flight across pairs).
__device__ inline uint32_t fetch_span(const uint16_t* __restrict__ spans,
                                      int stride, int e, uint32_t by, int iy) {
  const int y0a = (int)(by & 0xffff) & ~(kSpanAlign - 1);
  const int k = min(max(iy - y0a, 0), stride - 1);
  return gload(spans + (int64_t)e * stride + k);
}

template <int CTRL>
__device__ inline float dpp_move(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL,
                                         0xF, 0xF, false));
}

__device__ inline float lane_value(float v, int lane) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

// Sum over the wavefront with DPP inside each 16-lane row (quad swaps, then
// row rotations by 4 and 8), then the four row totals read as scalars.
// Returns the total (wave-uniform).
__device__ inline float wave_total(float v) {
  v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x124>(v);  // row_ror:4
  v += dpp_move<0x128>(v);  // row_ror:8
  return (lane_value(v, 0) + lane_value(v, 16)) +
         (lane_value(v, 32) + lane_value(v, 48));
}

// Records of up to 64 consecutive pairs, one pair per lane, gathered with a
// handful of vector loads so the random event reads of a whole batch are in
// flight together; each pair's fields are then handed to the wave with
// v_readlane (wave-uniform scalars).
struct PairBatch {
  int e;
  float cz, cy, A, B, P, Q0, inv;
  uint32_t bx, by;
};

template <bool WITH_INV>
__device__ inline PairBatch gather_batch(const BandRecord* __restrict__ rec,
                                         const int* __restrict__ pair_event,
                                         const float* __restrict__ inv,
                                         int first, int end, int lane) {
  PairBatch pb;
  const int p = first + lane;
  pb.e = (p < end) ? gload(pair_event + p) : 0;
  // the 32-byte record as two 16-byte loads: (cz, cy, A, B), (P, Q0, bx, by)
  const gfloat4* r = reinterpret_cast<const gfloat4*>(rec + pb.e);
  const gfloat4 lo = gload(r), hi = gload(r + 1);
  pb.cz = lo.x;
  pb.cy = lo.y;
  pb.A = lo.z;
  pb.B = lo.w;
  pb.P = hi.x;
  pb.Q0 = hi.y;
  // (the elements are copied out first: a bit cast applied directly to a
  // vector element expression reads element 0)
  const float bx = hi.z, by = hi.w;
  pb.bx = __builtin_bit_cast(uint32_t, bx);
  pb.by = __builtin_bit_cast(uint32_t, by);
  pb.inv = WITH_INV ? gload(inv + pb.e) : 0.0f;
  return pb;
}

__device__ inline int lane_int(int v, int j) {
  return __builtin_amdgcn_readlane(v, j);
}
__device__ inline float lane_float(float v, int j) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), j));
}

struct PairView {
  int e;
  float cz, cy, A, B, P, Q0, inv;
  int x0, x1, y0, y1;
};

__device__ inline PairView pick(const PairBatch& pb, int j) {
  PairView v;
  v.e = lane_int(pb.e, j);
  v.cz = lane_float(pb.cz, j);
  v.cy = lane_float(pb.cy, j);
  v.A = lane_float(pb.A, j);
  v.B = lane_float(pb.B, j);
  v.P = lane_float(pb.P, j);
  v.Q0 = lane_float(pb.Q0, j);
  v.inv = lane_float(pb.inv, j);
  const uint32_t bx = (uint32_t)lane_int((int)pb.bx, j);
  const uint32_t by = (uint32_t)lane_int((int)pb.by, j);
  v.x0 = bx & 0xffff;
  v.x1 = bx >> 16;
  v.y0 = by & 0xffff;
  v.y1 = by >> 16;
  return v;
}

// The fields of batch lane src (which may differ between lanes) for this
// lane, through the LDS crossbar (ds_bpermute; no LDS memory is touched).
template <bool WITH_INV>
__device__ inline PairView pick_lane(const PairBatch& pb, int src) {
  const int a = src << 2;
  auto mv = [a](float x) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_ds_bpermute(a, __builtin_bit_cast(int, x)));
  };
  PairView v;
  v.e = __builtin_amdgcn_ds_bpermute(a, pb.e);
  v.cz = mv(pb.cz);
  v.cy = mv(pb.cy);
  v.A = mv(pb.A);
  v.B = mv(pb.B);
  v.P = mv(pb.P);
  v.Q0 = mv(pb.Q0);
  v.inv = WITH_INV ? mv(pb.inv) : 0.0f;
  const uint32_t bx = (uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)pb.bx);
  const uint32_t by = (uint32_t)__builtin_amdgcn_ds_bpermute(a, (int)pb.by);
  v.x0 = bx & 0xffff;
  v.x1 = bx >> 16;
  v.y0 = by & 0xffff;
  v.y1 = by >> 16;
  return v;
}

// Pair j + group of a batch for this lane: wave-uniform scalars when a step
// holds one pair (G = 1), else the lane group's own pair.
template <int G, bool WITH_INV>
__device__ inline PairView step_pair(const PairBatch& pb, int j, int group) {
  if constexpr (G == 1)
    return pick(pb, j);
  else
    return pick_lane<WITH_INV>(pb, j + group);
}

// Sum over each group of N consecutive lanes (N = 16, 32 or 64), returned
// in every lane of the group.
template <int N>
__device__ inline float group_total(float v) {
  v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x124>(v);  // row_ror:4
  v += dpp_move<0x128>(v);  // row_ror:8
  if constexpr (N >= 32) v += __shfl_xor(v, 16);
  if constexpr (N >= 64) v += __shfl_xor(v, 32);
  return v;
}
}  // namespace
}  // namespace strip
