// Device-side math of the strip TOF MLEM path (gfx950, wave64).
//
// Everything here is FP32.  Functions marked "exact" are evaluated in the
// reference's operation order with contraction disabled, so that they give
// the same bits as a plain IEEE FP32 evaluation of the reference formulas;
// they decide integer quantities (bounding boxes, the 3-sigma support test),
// which must not depend on rounding choices.  The per-pixel kernel value uses
// the fast hardware exp2/rsqrt instead.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace strip {

// Scanner + GaussianKernel constants (src/2d/strip/scanner.h:39-67,
// gaussian_kernel.h:27-33), computed on the host in FP32.
struct DeviceGeometry {
  float radius, half_length, pixel_width, pixel_height;
  float tl_y, tl_z, tl_y_half_h, tl_z_half_w;
  float inv_pow_sigma_z, inv_pow_sigma_dl, norm_c;
  int n_z, n_y;
  // derived constants of the span walk (not part of the reference state)
  float inv_pixel_width;  // 1 / pixel_width
  float C;                // 2 / sigma_z^2, the dz^2 coefficient of Q
  float inv_2c;           // 1 / (2 C)
};

// Per-event parameters, precomputed once per event upload.
//   a = (z, y, tan, sec)   ellipse centre and angle (response.h:22-37)
//   b = (A, B, C, 0)       3-sigma ellipse quadratic form (gaussian_kernel.h:121-123)
//   box = (x0 | x1<<16, y0 | y1<<16), clamped bbox, x1/y1 exclusive
//         (reconstruction.h:225-256)
//   c = (P', Q0', 0, 0)    row prefactor c/sqrt(a.S^-1.a) = rsqrt(P' yp^2 + Q0')
//                          (gaussian_kernel.h:58-74 with a.S^-1.a expanded:
//                          2 iz s^4 (yp^2 + R^2) + 4 id s^2 t^2 yp^2)
struct EventSoA {
  float4* a;
  float4* b;
  float4* c;
  uint2* box;
};

// -(1/2) * log2(e): exp(-Q/2) == exp2(Q * kExp2Scale)
constexpr float kExp2Scale = -0.72134752044448170368f;

// float -> short as Pixel(S,S) converts it (truncation toward zero,
// pixel.h:15); saturating instead of undefined out of range.
__device__ __host__ inline int to_short_trunc(float v) {
  if (!(v > -32768.0f)) return -32768;
  if (!(v < 32767.0f)) return 32767;
  return (int)(short)(int)v;
}

__device__ __host__ inline int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Response::calculate_tan_y_z (response.h:22-37), ellipse_bb
// (gaussian_kernel.h:108-128) and the bbox of bb_update
// (reconstruction.h:225-246).  Exact.
__device__ inline void setup_event_exact(const DeviceGeometry& g, float z_u,
                                         float z_d, float dl, float4& a,
                                         float4& b, float4& c, uint2& box) {
#pragma clang fp contract(off)
  const float iz = g.inv_pow_sigma_z, id = g.inv_pow_sigma_dl;
  float tan = (z_u - z_d) / (2.0f * g.radius);
  float y = -0.5f * (dl / sqrtf(1.0f + tan * tan));
  float z = 0.5f * (z_u + z_d + (2.0f * y * tan));
  float tan_sq = tan * tan;
  float sec = sqrtf(1.0f + tan * tan);
  float A = 4.0f * sec * sec * id + 2.0f * tan_sq * iz;
  float B = -4.0f * tan * iz;
  float C = 2.0f * iz;
  float B_2 = (B / 2.0f) * (B / 2.0f);
  float bb_y = 3.0f / sqrtf(A - (B_2 / C));
  float bb_z = 3.0f / sqrtf(C - (B_2 / A));
  int cx = to_short_trunc((z - g.tl_z) / g.pixel_width);
  int cy = to_short_trunc((g.tl_y - y) / g.pixel_height);
  int hw = (int)(bb_z / g.pixel_width + 0.5f);
  int hh = (int)(bb_y / g.pixel_height + 0.5f);
  int x0 = (short)(cx - hw), y0 = (short)(cy - hh);
  int x1 = (short)(cx + hw), y1 = (short)(cy + hh);
  x0 = clampi(x0, 0, g.n_z - 1);
  x1 = clampi(x1, 0, g.n_z - 1);
  y0 = clampi(y0, 0, g.n_y - 1);
  y1 = clampi(y1, 0, g.n_y - 1);
  a = make_float4(z, y, tan, sec);
  b = make_float4(A, B, C, 0.0f);
  {
    // row prefactor coefficients, normalised by c^2 (not decision-making,
    // contraction harmless)
    float s2 = 1.0f + tan * tan;
    float s4 = s2 * s2;
    float inv_c2 = 1.0f / (g.norm_c * g.norm_c);
    float P = (2.0f * iz * s4 + 4.0f * id * s2 * tan * tan) * inv_c2;
    float Q0 = (2.0f * iz * s4 * g.radius * g.radius) * inv_c2;
    c = make_float4(P, Q0, 0.0f, 0.0f);
  }
  box = make_uint2((uint32_t)x0 | ((uint32_t)x1 << 16),
                   (uint32_t)y0 | ((uint32_t)y1 << 16));
}

// Scanner::pixel_center (scanner.h:136-139).  Exact.
__device__ inline float pixel_center_z(const DeviceGeometry& g, int ix) {
#pragma clang fp contract(off)
  return g.tl_z_half_w + (float)ix * g.pixel_width;
}
__device__ inline float pixel_center_y(const DeviceGeometry& g, int iy) {
#pragma clang fp contract(off)
  return g.tl_y_half_h - (float)iy * g.pixel_height;
}

// GaussianKernel::in_ellipse quadratic form (gaussian_kernel.h:131-137).
// Exact; membership is Q <= 9.
__device__ inline float ellipse_q_exact(float A, float B, float C, float dy,
                                        float dz) {
#pragma clang fp contract(off)
  return (A * dy * dy) + (B * dy * dz) + (C * dz * dz);
}

// Row prefactor of GaussianKernel::operator() (gaussian_kernel.h:58-74):
//   c / sqrt(a . Sigma^-1 . a), a = (-(yp-R) s^2, -(yp+R) s^2, -2 yp s t),
// which depends only on the pixel row (yp = dy + y) and the event angle.
__device__ inline float gauss_row_prefactor(const DeviceGeometry& g, float yp,
                                            float tan, float sec) {
  float sec_sq = sec * sec;
  float ap = -(yp - g.radius) * sec_sq;
  float aq = -(yp + g.radius) * sec_sq;
  float ar = -2.0f * yp * sec * tan;
  float a_ic_a = g.inv_pow_sigma_z * (ap * ap + aq * aq) +
                 g.inv_pow_sigma_dl * (ar * ar);
  return g.norm_c * __builtin_amdgcn_rsqf(a_ic_a);
}

// exp(-Q/2) with the hardware exp2 (v_exp_f32).  Q <= 9 keeps the argument
// in [-6.5, 0], far from denormals.
__device__ inline float gauss_exp(float q) {
  return __builtin_amdgcn_exp2f(q * kExp2Scale);
}

// AnalyticKernel::operator() (analytic_kernel.h:28-62), accurate libm-grade
// FP32 math.  Used only by the legacy --gpu semantics mode.
__device__ inline float analytic_kernel(const DeviceGeometry& g, float y,
                                        float tan, float sec, float dz,
                                        float dy) {
  const float iz = g.inv_pow_sigma_z, id = g.inv_pow_sigma_dl;
  float sec_sq = sec * sec;
  float yp = dy + y;
  float ap = -(yp - g.radius) * sec_sq;
  float aq = -(yp + g.radius) * sec_sq;
  float ar = -2.0f * yp * sec * tan;
  float op = ap * tan, oq = aq * tan;
  float orr = -yp * sec * (1.0f + 2.0f * tan * tan);
  float b_pq = dz - dy * tan;
  float br = -2.0f * dy * sec;
  float a_ic_a = ap * iz * ap + aq * iz * aq + ar * id * ar;
  float b_ic_a = b_pq * iz * ap + b_pq * iz * aq + br * id * ar;
  float b_ic_b = b_pq * iz * b_pq + b_pq * iz * b_pq + br * id * br;
  float o_ic_b = op * iz * b_pq + oq * iz * b_pq + orr * id * br;
  float norm = a_ic_a + 2.0f * o_ic_b;
  float element_before_exp = g.norm_c / sqrtf(norm);
  float exp_arg = -0.5f * (b_ic_b - b_ic_a * b_ic_a / norm);
  return element_before_exp * expf(exp_arg);
}

// Scanner::sensitivity (scanner.h:147-171), accurate FP32.
__device__ inline float scanner_sensitivity(const DeviceGeometry& g, float z,
                                            float y) {
#pragma clang fp contract(off)
  float L_plus = g.half_length + z;
  float L_minus = g.half_length - z;
  float R_plus = g.radius + y;
  float R_minus = g.radius - y;
  if (R_plus <= 0.0f) R_plus = 1.0f;
  if (R_minus <= 0.0f) R_minus = 1.0f;
  float u = L_minus / R_minus, v = L_plus / R_plus;
  float p = -L_plus / R_minus, q = -L_minus / R_plus;
  float mn = (v < u) ? v : u;
  float mx = (p < q) ? q : p;
  return 0.31830988618379067154f * (atanf(mn) - atanf(mx));
}

// Sum over the 64 lanes of a wavefront; every lane receives the total.
__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1)
    v += __shfl_xor(v, offset, 64);
  return v;
}

__device__ inline int wave_sum_int(int v) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1)
    v += __shfl_xor(v, offset, 64);
  return v;
}

}  // namespace strip
