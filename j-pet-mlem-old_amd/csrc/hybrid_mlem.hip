// MI355X (gfx950) 3D hybrid list-mode MLEM: kernels + C ABI.
//
// One iteration (src/3d/hybrid/reconstruction.h:276-447) over the frame
// events e, each a LOR with a box of planes [pb, pe) and of pixel infos
// [ib, ie) along the LOR (translate_to_frame, reconstruction.h:101-147):
//   w_e(i, z) = K_e(i, z) pw_i rho(pixel_i, z)
//               K_e = GaussianKernel::normalized (gaussian_kernel.h:93-105),
//               skipped where <= 0
//   D_e       = sum_{i, z} w_e(i, z) s(pixel_i)
//   rho'(v)   = sum_e sum_{(i, z) -> v} w_e(i, z) / D_e      (D_e > 0)
// One wavefront per event; its lanes walk the box's (info, plane) pairs
// plane-fastest.  On the device the image is stored plane-fastest
// (index pixel n_planes + z), so the lanes that share an info touch
// consecutive addresses in both the rho gather and the atomic update.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "hybrid_mlem.h"
#include "strip_device.hpp"

namespace strip {
int fail(int code, const std::string& message);
}

// the C ABI's records, as the binding lays them out
static_assert(sizeof(hybrid_mlem_response) == 16, "response layout");
static_assert(sizeof(hybrid_mlem_frame_event) == 48 &&
                  offsetof(hybrid_mlem_frame_event, info_begin) == 24 &&
                  offsetof(hybrid_mlem_frame_event, plane_begin) == 40,
              "frame event layout");

namespace hybrid {
namespace {

struct Grid {
  int n_columns, n_rows, n_planes, n_pixels;
  float pixel_size, z_left;
  float llc_x, llc_y;  // centre of pixel (0, 0) (pixel_grid.h:25-32)
};

// GaussianKernel(sigma_z, sigma_dl) (gaussian_kernel.h:27-33) and the
// scintillator length of GaussianKernel::normalized.
struct Kernel {
  float iz, id, norm_c, length;
};

// Event arrays (frame events), one entry per event.
//   ev_i = (lor, info begin, info end, plane begin | plane end << 16)
//   ev_f = (up, right, tan, sec)
//   ev_dup = info index of the range whose pixel repeats the pixel of the
//            range's last info when that info lies past the LOR's own
//            infos, else -1
struct Events {
  int4* ev_i = nullptr;
  float4* ev_f = nullptr;
  int* ev_dup = nullptr;
  int64_t n = 0, capacity = 0;
};

__device__ inline int pixel_x(int p) { return (int)(short)(p & 0xffff); }
__device__ inline int pixel_y(int p) { return p >> 16; }

__device__ inline bool in_grid(const Grid& g, int x, int y) {
  return x >= 0 && x < g.n_columns && y >= 0 && y < g.n_rows;
}

// PixelGrid::center_at (pixel_grid.h:54-57).  Exact.
__device__ inline void pixel_centre(const Grid& g, int x, int y, float& cx,
                                    float& cy) {
#pragma clang fp contract(off)
  cx = g.llc_x + (float)x * g.pixel_size;
  cy = g.llc_y + (float)y * g.pixel_size;
}

// GaussianKernel::sensitivity (gaussian_kernel.h:140-150), the frame
// sensitivity at p = (z, up) of a LOR of half length R.
__device__ inline float frame_sensitivity(float px, float py, float R,
                                          float L) {
#pragma clang fp contract(off)
  float L2 = L / 2.0f;
  float u = (L2 - px) / (R - py), v = (L2 + px) / (R + py);
  float p = -(L2 + px) / (R - py), q = (-L2 + px) / (R + py);
  float mn = (v < u) ? v : u;
  float mx = (p < q) ? q : p;
  return (atanf(mn) - atanf(mx)) / 3.14159265358979323846f;
}

// GaussianKernel::operator() (gaussian_kernel.h:51-79).
__device__ inline float gauss(const Kernel& k, float y, float tan, float sec,
                              float R, float dz, float dy) {
  float sec_sq = sec * sec;
  float yp = dy + y;
  float ap = -(yp - R) * sec_sq;
  float aq = -(yp + R) * sec_sq;
  float ar = -2.0f * yp * sec * tan;
  float b_pq = dz - dy * tan;
  float br = -2.0f * dy * sec;
  float a_ic_a = ap * k.iz * ap + aq * k.iz * aq + ar * k.id * ar;
  float b_ic_b = b_pq * k.iz * b_pq + b_pq * k.iz * b_pq + br * k.id * br;
  return (k.norm_c / sqrtf(a_ic_a)) * expf(-0.5f * b_ic_b);
}

// kernel2d of one (info, plane) of an event (reconstruction.h:333-341):
// normalized(Point2D(right, up), tan, sec, R, L, Point2D(z, up_pixel)).
__device__ inline float kernel2d(const Grid& g, const Kernel& k,
                                 const float4& sa, float R, const float4& ef,
                                 int x, int y, int plane) {
  float cx, cy;
  pixel_centre(g, x, y, cx, cy);
  float up_pixel, z;
  {
#pragma clang fp contract(off)
    // LineSegment::projection_relative_middle (line_segment.h:82-84)
    up_pixel = (cx - sa.x) * sa.z + (cy - sa.y) * sa.w;
    // VoxelGrid::center_z_at (voxel_grid.h:40-42)
    z = ((float)plane + 0.5f) * g.pixel_size + g.z_left;
  }
  float un = gauss(k, ef.x, ef.z, ef.w, R, z - ef.y, up_pixel - ef.x);
  return un / frame_sensitivity(z, up_pixel, R, k.length);
}

// ---------------------------------------------------------------------------
// Geometry preparation
// ---------------------------------------------------------------------------

// LineSegment(start, end) (line_segment.h:27-34): a = (mid x, mid y,
// direction x, y), b = (normal x, y, distance from origin, length / 2).
__global__ __launch_bounds__(256) void segments_kernel(
    const float* __restrict__ s4, int64_t n_lors, float4* __restrict__ a,
    float4* __restrict__ b) {
#pragma clang fp contract(off)
  int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n_lors) return;
  float sx = s4[4 * l], sy = s4[4 * l + 1], ex = s4[4 * l + 2],
        ey = s4[4 * l + 3];
  float vx = ex - sx, vy = ey - sy;
  float len = sqrtf(vx * vx + vy * vy);
  float dx = vx / len, dy = vy / len;
  float nx = -dy, ny = dx;
  a[l] = make_float4((sx + ex) / 2.0f, (sy + ey) / 2.0f, dx, dy);
  b[l] = make_float4(nx, ny, sx * nx + sy * ny, len / 2.0f);
}

// calculate_weight (reconstruction.h:218-238): one wavefront per LOR.
__global__ __launch_bounds__(256) void calculate_weight_kernel(
    Grid g, const float4* __restrict__ seg_b, const float* __restrict__ widths,
    const int2* __restrict__ ranges, const int* __restrict__ pix,
    int64_t n_lors, float* __restrict__ w) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t l = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
       l < n_lors; l += n_waves) {
    const float4 sb = seg_b[l];
    const float sigma_w = 0.3f * widths[l];
    const double norm_w = 1.0 / ((double)sigma_w * sqrt(2.0 * M_PI));
    const float inv_sigma2_w = 1.0f / (2.0f * sigma_w * sigma_w);
    const int2 r = ranges[l];
    for (int i = r.x + lane; i < r.y; i += 64) {
      const int p = pix[i];
      float cx, cy;
      pixel_centre(g, pixel_x(p), pixel_y(p), cx, cy);
      // LineSegment::distance_from (line_segment.h:65-67)
      float d = (cx * sb.x + cy * sb.y) - sb.z;
      w[i] = (float)(norm_w * (double)expf(-d * d * inv_sigma2_w));
    }
  }
}

// Per-pixel info counts of the infos inside the grid.
__global__ __launch_bounds__(256) void pixel_count_kernel(
    Grid g, const int* __restrict__ pix, int64_t n, int* __restrict__ count,
    unsigned* __restrict__ keys) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = pix[i];
  const int x = pixel_x(p), y = pixel_y(p);
  unsigned key = 0xffffffffu;
  if (in_grid(g, x, y)) {
    key = (unsigned)(y * g.n_columns + x);
    atomicAdd(&count[key], 1);
  }
  keys[i] = key;
}

// calculate_sensitivity (reconstruction.h:240-251): each pixel sums its
// weights in info order (the infos of the pixel, sorted stably by pixel).
__global__ __launch_bounds__(256) void sensitivity_sum_kernel(
    const int* __restrict__ offsets, const int* __restrict__ order,
    const float* __restrict__ w, int n_pixels, float* __restrict__ sens) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pixels) return;
  float s = 0.0f;
  for (int k = offsets[p]; k < offsets[p + 1]; ++k) s += w[order[k]];
  sens[p] = s;
}

// normalize_geometry_weights (reconstruction.h:253-271).
__global__ __launch_bounds__(256) void normalize_kernel(
    Grid g, const int* __restrict__ pix, const float* __restrict__ sens,
    int64_t n, float* __restrict__ w) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int p = pix[i];
  const int x = pixel_x(p), y = pixel_y(p);
  if (in_grid(g, x, y)) w[i] = w[i] / sens[y * g.n_columns + x];
}

// ---------------------------------------------------------------------------
// Events
// ---------------------------------------------------------------------------

// translate_to_frame (reconstruction.h:101-147) and
// bb_intersects_grid_with_positive_weight (reconstruction.h:151-162).
// Exact (contraction off), as the box decides which infos and planes an
// event visits.
__global__ __launch_bounds__(256) void frame_events_kernel(
    Grid g, Kernel k, const hybrid_mlem_response* __restrict__ responses,
    int64_t n, int64_t n_lors, const float4* __restrict__ seg_b,
    const int2* __restrict__ ranges, const float* __restrict__ pos,
    const int* __restrict__ pix, const float* __restrict__ w, int n_infos,
    int4* __restrict__ ev_i, float4* __restrict__ ev_f,
    int* __restrict__ ev_dup, char* __restrict__ keep,
    int* __restrict__ bad) {
#pragma clang fp contract(off)
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const hybrid_mlem_response r = responses[i];
  const int first = r.first, second = r.second;
  // LOR::index (lor.h:36-38), first and second as read (lor.h:29-33)
  const int64_t lor = (int64_t)(first * (first + 1) / 2 + second);
  if (first < 0 || second < 0 || lor >= n_lors) {
    keep[i] = 0;
    atomicOr(bad, 1);
    return;
  }
  const float R = seg_b[lor].w;
  // Response::calculate_tan_y_z (2d/strip/response.h:22-37)
  const float tan = (r.z_up - r.z_dn) / (2.0f * R);
  const float up = -0.5f * (r.dl / sqrtf(1.0f + tan * tan));
  const float right = 0.5f * (r.z_up + r.z_dn + (2.0f * up * tan));
  // GaussianKernel::ellipse_bb (gaussian_kernel.h:108-128)
  const float sec = sqrtf(1.0f + tan * tan);
  const float A = 4.0f * sec * sec * k.id + 2.0f * (tan * tan) * k.iz;
  const float B = -4.0f * tan * k.iz;
  const float C = 2.0f * k.iz;
  const float B_2 = (B / 2.0f) * (B / 2.0f);
  const float half_up = 3.0f / sqrtf(A - (B_2 / C));
  const float half_right = 3.0f / sqrtf(C - (B_2 / A));
  // plane(z) = S((z - z_left) / pixel_size) (reconstruction.h:149)
  const int pl =
      strip::to_short_trunc(((right - half_right) - g.z_left) / g.pixel_size);
  const int pr =
      strip::to_short_trunc(((right + half_right) - g.z_left) / g.pixel_size);
  const int pb = pl > 0 ? pl : 0;
  int pe = (int)(short)(pr + 1);
  pe = pe < g.n_planes ? pe : g.n_planes;
  const float t_up = ((up + half_up) + R) / (2.0f * R);
  const float t_dn = ((up - half_up) + R) / (2.0f * R);
  const int2 be = ranges[lor];
  // upper_bound(t_up) + 1 and lower_bound(t_dn) over the LOR's positions
  int lo = be.x, hi = be.y;
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if (t_up < pos[mid])
      hi = mid;
    else
      lo = mid + 1;
  }
  // the info after the LOR's last one is past the arrays for the last LOR
  // with infos; the range stops at the arrays' end there
  const int ie = lo + 1 < n_infos ? lo + 1 : n_infos;
  lo = be.x;
  hi = be.y;
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if (pos[mid] < t_dn)
      lo = mid + 1;
    else
      hi = mid;
  }
  const int ib = lo;
  bool kept = !(pe <= 0 || pb >= g.n_planes);
  if (kept) {
    kept = false;
    for (int j = ib; j < ie; ++j) {
      const int p = pix[j];
      if (in_grid(g, pixel_x(p), pixel_y(p)) && w[j] > 0.0f) {
        kept = true;
        break;
      }
    }
  }
  int dup = -1;
  const int last = ie - 1;
  if (kept && last >= be.y) {
    const int px = pix[last];
    for (int j = ib; j < last; ++j)
      if (pix[j] == px) {
        dup = j;
        break;
      }
  }
  keep[i] = kept ? 1 : 0;
  ev_i[i] = make_int4((int)lor, ib, ie, pb | (pe << 16));
  ev_f[i] = make_float4(up, right, tan, sec);
  ev_dup[i] = dup;
}

// Frame events given by the caller: checked against the geometry, the
// info end clipped to the arrays, the repeated-visit index found as in
// frame_events_kernel.
__global__ __launch_bounds__(256) void frame_import_kernel(
    Grid g, const hybrid_mlem_frame_event* __restrict__ in, int64_t n,
    int64_t n_lors, const int2* __restrict__ ranges,
    const int* __restrict__ pix, int n_infos, int4* __restrict__ ev_i,
    float4* __restrict__ ev_f, int* __restrict__ ev_dup,
    int* __restrict__ bad) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const hybrid_mlem_frame_event f = in[i];
  const int first = f.first, second = f.second;
  const int64_t lor = (int64_t)(first * (first + 1) / 2 + second);
  const int64_t ie = f.info_end <= n_infos ? f.info_end : n_infos;
  if (first < 0 || second < 0 || lor >= n_lors || f.info_begin < 0 ||
      f.info_begin > ie || f.plane_begin < 0 || f.plane_end > g.n_planes) {
    atomicOr(bad, 1);
    ev_i[i] = make_int4(0, 0, 0, 0);
    ev_f[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    ev_dup[i] = -1;
    return;
  }
  const int2 be = ranges[lor];
  const int ib = (int)f.info_begin, iend = (int)ie;
  int dup = -1;
  const int last = iend - 1;
  if (last >= ib && last >= be.y) {
    const int px = pix[last];
    for (int j = ib; j < last; ++j)
      if (pix[j] == px) {
        dup = j;
        break;
      }
  }
  ev_i[i] = make_int4((int)lor, ib, iend,
                      (int)f.plane_begin | ((int)f.plane_end << 16));
  ev_f[i] = make_float4(f.up, f.right, f.tan, f.sec);
  ev_dup[i] = dup;
}

__global__ __launch_bounds__(256) void visits_kernel(
    const int4* __restrict__ ev_i, int64_t n,
    unsigned long long* __restrict__ total) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = 0;
  if (i < n) {
    const int4 e = ev_i[i];
    const int n_pl = (e.w >> 16) - (e.w & 0xffff);
    v = (unsigned long long)(e.z - e.y) * (unsigned long long)(n_pl > 0 ? n_pl : 0);
  }
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(total, v);
}

// ---------------------------------------------------------------------------
// Iteration
// ---------------------------------------------------------------------------

// One wavefront per event (grid-stride).  Pass 1 forms D_e (wavefront sum),
// pass 2 recomputes each weight and adds weight / D_e with float atomics
// into the plane-fastest output image.  LAST_WRITE: the CPU path's kernel
// cache semantics for the repeated pixel of the range's extra info (the
// earlier visit adds the extra info's weight).
template <bool LAST_WRITE>
__global__ __launch_bounds__(256) void event_kernel(
    Grid g, Kernel k, const float4* __restrict__ seg_a,
    const float4* __restrict__ seg_b, const int* __restrict__ pix,
    const float* __restrict__ w, const int4* __restrict__ ev_i,
    const float4* __restrict__ ev_f, const int* __restrict__ ev_dup,
    int64_t n_events, const float* __restrict__ rho,
    const float* __restrict__ sens, float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t e = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
       e < n_events; e += n_waves) {
    const int4 ei = ev_i[e];
    const float4 ef = ev_f[e];
    const int pb = ei.w & 0xffff, pe = ei.w >> 16;
    const int n_pl = pe - pb;
    const int n_info = ei.z - ei.y;
    if (n_pl <= 0 || n_info <= 0) continue;
    const float4 sa = seg_a[ei.x];
    const float R = seg_b[ei.x].w;
    const int total = n_info * n_pl;
    // lane's first pair (q info offset, r plane offset) and the stride of 64
    const int sq = 64 / n_pl, sr = 64 - sq * n_pl;
    float den = 0.0f;
    {
      int q = lane / n_pl, r = lane - (lane / n_pl) * n_pl;
      for (int idx = lane; idx < total; idx += 64) {
        const int i = ei.y + q;
        const int p = pix[i];
        const int x = pixel_x(p), y = pixel_y(p);
        if (in_grid(g, x, y)) {
          const float kv = kernel2d(g, k, sa, R, ef, x, y, pb + r);
          if (!(kv <= 0.0f)) {
            const int pixel = y * g.n_columns + x;
            const float wt =
                kv * w[i] * rho[(int64_t)pixel * g.n_planes + pb + r];
            den += wt * sens[pixel];
          }
        }
        q += sq;
        r += sr;
        if (r >= n_pl) {
          r -= n_pl;
          ++q;
        }
      }
    }
    den = strip::wave_sum(den);
    if (!(den > 0.0f)) continue;
    const float inv = 1.0f / den;
    const int dup = LAST_WRITE ? ev_dup[e] : -1;
    int q = lane / n_pl, r = lane - (lane / n_pl) * n_pl;
    for (int idx = lane; idx < total; idx += 64) {
      const int i = ei.y + q;
      const int p = pix[i];
      const int x = pixel_x(p), y = pixel_y(p);
      if (in_grid(g, x, y)) {
        const float kv = kernel2d(g, k, sa, R, ef, x, y, pb + r);
        if (!(kv <= 0.0f)) {
          const int pixel = y * g.n_columns + x;
          const int64_t v = (int64_t)pixel * g.n_planes + pb + r;
          const float pw = (LAST_WRITE && i == dup) ? w[ei.z - 1] : w[i];
          const float wt = kv * pw * rho[v];
          atomicAdd(&out[v], wt * inv);
        }
      }
      q += sq;
      r += sr;
      if (r >= n_pl) {
        r -= n_pl;
        ++q;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// LOR-space iteration
// ---------------------------------------------------------------------------
//
// The frame sensitivity and the pixel weight of a visit depend on the info
// (its pixel and LOR) and the plane, not on the event, so
//   G(i, z) = pw_i / S(z, up_i; R_lor, L)
// is tabulated once per geometry (n_infos x n_planes, plane-fastest), with
// 0 where kernel2d would be skipped (S negative).  An event's kernel splits
// into a per-info factor and a Gaussian in z:
//   K_e(i, z) = A_ei exp(-iz b^2),  b = (z - right) - dy tan,
//   A_ei = c / sqrt(a.S^-1.a) exp(-id (2 dy sec)^2 / 2),  dy = up_i - up
// (gaussian_kernel.h:51-79 with b_ic_b = 2 iz b^2 + id (2 dy sec)^2).
//   forward:  D_e = sum K_e G rho s + the extra info's visits
//   backward: bp(i, z) = sum_e K_e(i, z) / D_e, accumulated per LOR in LDS
//   gather:   rho'(p, z) = rho(p, z) sum_{i of pixel p} G(i, z) bp(i, z)
//                          + extra(p, z)
// The extra info past a LOR's own infos (the range's last, when it lies
// beyond the LOR) is evaluated in full and added to extra(p, z) with
// atomics; so is a repeated visit's update under the kernel-cache
// semantics.

// LDS words of a wavefront's area: an event's two factors per info
constexpr int kWaveWords = 512;
// dynamic LDS of the back-projection: 160 KiB less room for its static
// queue slot
constexpr size_t kLdsBytes = 160 * 1024 - 256;

// A_ei and dy tan of one info of an event.
__device__ inline void info_factors(const Kernel& k, const float4& ef,
                                    float R, float up_i, float& A,
                                    float& t) {
  const float y = ef.x, tan = ef.z, sec = ef.w;
  const float dy = up_i - y;
  const float sec_sq = sec * sec;
  const float yp = dy + y;
  const float ap = -(yp - R) * sec_sq;
  const float aq = -(yp + R) * sec_sq;
  const float ar = -2.0f * yp * sec * tan;
  const float a_ic_a = ap * k.iz * ap + aq * k.iz * aq + ar * k.id * ar;
  const float br = -2.0f * dy * sec;
  A = (k.norm_c / sqrtf(a_ic_a)) * expf(-0.5f * (br * k.id * br));
  t = dy * tan;
}

__device__ inline float plane_centre(const Grid& g, int plane) {
#pragma clang fp contract(off)
  return ((float)plane + 0.5f) * g.pixel_size + g.z_left;
}

// info -> LOR, and per-info pixel index (-1 outside the grid) and up, the
// projection of the pixel centre on its LOR relative to the middle.
__global__ __launch_bounds__(256) void info_lor_kernel(
    Grid g, const float4* __restrict__ seg_a, const int2* __restrict__ ranges,
    const int* __restrict__ pix, int64_t n_lors, int* __restrict__ info_lor,
    int* __restrict__ ipix, float* __restrict__ upi) {
  const int lane = threadIdx.x & 63;
  const int64_t n_waves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t l = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
       l < n_lors; l += n_waves) {
    const int2 r = ranges[l];
    const float4 sa = seg_a[l];
    for (int i = r.x + lane; i < r.y; i += 64) {
      const int p = pix[i];
      const int x = pixel_x(p), y = pixel_y(p);
      float cx, cy;
      pixel_centre(g, x, y, cx, cy);
      float up;
      {
#pragma clang fp contract(off)
        up = (cx - sa.x) * sa.z + (cy - sa.y) * sa.w;
      }
      info_lor[i] = (int)l;
      ipix[i] = in_grid(g, x, y) ? y * g.n_columns + x : -1;
      upi[i] = up;
    }
  }
}

// G(i, z) for every info of a LOR (0 for the others).
__global__ __launch_bounds__(256) void frame_table_kernel(
    Grid g, Kernel k, const float4* __restrict__ seg_b,
    const int* __restrict__ info_lor, const int* __restrict__ ipix,
    const float* __restrict__ upi, const float* __restrict__ w, int n_infos,
    float* __restrict__ G) {
  // a grid-stride loop: the table can hold more cells than one launch has
  // work-items (a dispatch counts them in 32 bits)
  const int64_t n_cells = (int64_t)n_infos * g.n_planes;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
       idx < n_cells; idx += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / g.n_planes);
    const int z = (int)(idx - (int64_t)i * g.n_planes);
    const int l = info_lor[i];
    float v = 0.0f;
    if (l >= 0 && ipix[i] >= 0) {
      const float S =
          frame_sensitivity(plane_centre(g, z), upi[i], seg_b[l].w, k.length);
      // kernel2d = un / S is skipped where <= 0 (S negative or -0)
      v = signbit(S) ? 0.0f : w[i] / S;
    }
    G[idx] = v;
  }
}

// Orders this wavefront's LDS accesses before and after (a wavefront's
// LDS instructions complete in issue order; the fences keep the compiler
// from moving them across).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The walk of an event's rows x planes with each lane's plane fixed: for a
// chunk of w <= 64 planes from plane chunk0 of the box, lane l takes plane
// chunk0 + l mod w and rows l / w, l / w + rpi, ... (rpi = 64 / w rows per
// step); lanes from rpi w on stay idle.
struct PlaneWalk {
  int q0, r, rpi;
  bool active;
  __device__ PlaneWalk(int lane, int n_pl, int chunk0) {
    const int w = n_pl - chunk0 < 64 ? n_pl - chunk0 : 64;
    // l / w for 0 <= l <= 64 as the integer part of (l + 1/2) / w, which
    // lies at least 1 / (2 w) >= 1 / 128 from an integer: far more than the
    // error of the hardware reciprocal and of the product
    const float rcp = __builtin_amdgcn_rcpf((float)w);
    rpi = (int)(64.5f * rcp);
    q0 = (int)(((float)lane + 0.5f) * rcp);
    r = chunk0 + lane - __mul24(q0, w);
    active = q0 < rpi;
  }
};

// Event parameters of 64 events, one per lane, read back for one event at a
// time with readlane (wave-uniform values in scalar registers).
struct EventBatch {
  int4 i;
  float4 f;
  float inv;
  int dup;
  __device__ void load(const int4* ev_i, const float4* ev_f,
                       const float* inv_in, const int* ev_dup, int e,
                       bool valid) {
    if (valid) {
      i = ev_i[e];
      f = ev_f[e];
      inv = inv_in ? inv_in[e] : 0.0f;
      dup = ev_dup[e];
    } else {
      i = make_int4(0, 0, 0, 0);
      f = make_float4(0.f, 0.f, 0.f, 0.f);
      inv = 0.0f;
      dup = -1;
    }
  }
};

__device__ inline int lane_int(int v, int j) {
  return __builtin_amdgcn_readlane(v, j);
}
__device__ inline float lane_float(float v, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

// The factors (A / D_e, dy tan) of a back-projection unit: event lane j,
// rows r0 .. min(r0 + 64, hi_j) - 1 of its box; lane t holds row r0 + t.
__device__ inline float2 unit_factors(const float2* __restrict__ fac, int a0,
                                      int hi, int64_t off, int j, int r0,
                                      int lane) {
  const int ja0 = lane_int(a0, j), jhi = lane_int(hi, j);
  const int64_t joff =
      (int64_t)(((uint64_t)(uint32_t)lane_int((int)(off >> 32), j) << 32) |
                (uint32_t)lane_int((int)off, j));
  return lane < jhi - r0 ? fac[joff + (r0 - ja0) + lane]
                         : make_float2(0.f, 0.f);
}

// Inclusive prefix sums over the 64 lanes.
__device__ inline int wave_scan_int(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  return v;
}
__device__ inline float wave_scan_float(float v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const float u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  return v;
}

// The end of an event's forward step: D_e = the lanes' partial sums plus
// the extra info's terms (the info one past the LOR's range, with the full
// kernel), then 1 / D_e, the back-projection's factors F = A / D_e (0 for
// the visit the extra info repeats under the kernel-cache semantics) and
// dy tan, and the extra info's update added to the image-space sums (twice
// for that repeated visit).  wAT holds the event's (A, dy tan) per info.
template <bool LAST_WRITE>
__device__ inline void finish_event(
    const Grid& g, const Kernel& k, const float4& sa, float R,
    const float4& ef, const int4& ei, int e, int dup, int lor_end, int pb,
    int pe, int n_info, int np, int lane, float den,
    const int* __restrict__ pix, const float* __restrict__ w,
    const float* __restrict__ rho, const float* __restrict__ sens,
    const int64_t* __restrict__ fac_off, const float2* wAT,
    float* __restrict__ inv, float2* __restrict__ fac,
    float* __restrict__ extra) {
  int xx = 0, yy = 0, xpixel = -1;
  if (ei.z > lor_end && pe > pb) {
    const int pxy = pix[ei.z - 1];
    xx = pixel_x(pxy);
    yy = pixel_y(pxy);
    if (in_grid(g, xx, yy)) xpixel = yy * g.n_columns + xx;
  }
  if (xpixel >= 0) {
    for (int z = pb + lane; z < pe; z += 64) {
      const float kv = kernel2d(g, k, sa, R, ef, xx, yy, z);
      if (!(kv <= 0.0f))
        den += kv * w[ei.z - 1] * rho[xpixel * np + z] * sens[xpixel];
    }
  }
  den = strip::wave_sum(den);
  const float iv = den > 0.0f ? 1.0f / den : 0.0f;
  if (lane == 0) inv[e] = iv;
  if (iv > 0.0f) {
    float2* f = fac + fac_off[e];
    for (int t = lane; t < n_info; t += 64) {
      const float2 at = wAT[t];
      f[t] = make_float2(
          (LAST_WRITE && ei.y + t == dup) ? 0.0f : at.x * iv, at.y);
    }
    if (xpixel >= 0) {
      const float mult = (LAST_WRITE && dup >= 0) ? 2.0f : 1.0f;
      for (int z = pb + lane; z < pe; z += 64) {
        const float kv = kernel2d(g, k, sa, R, ef, xx, yy, z);
        if (!(kv <= 0.0f)) {
          const int v = xpixel * np + z;
          atomicAdd(&extra[v], mult * (kv * w[ei.z - 1] * rho[v] * iv));
        }
      }
    }
  }
}

// One LOR-space iteration's event work.  A persistent grid, one block per
// CU, takes LORs from a queue (LORs with events, most events first).  For
// each LOR the block's LDS slab (the LOR's infos x all planes, laid out as
// its rows of bp) first holds P = G rho s of every (info, plane); where the
// LOR's infos x planes exceed slab_budget floats (with its infos' up), the
// slab holds a chunk of the planes at a time and each step below runs chunk
// by chunk, the forward pass summing each event's D_e over the chunks in
// memory before a last pass forms 1 / D_e and the factors.
//   Forward: the waves take the LOR's events in turn (parameters read 64 at
//   a time, one per lane); per event the per-info factors (A, dy tan) go
//   to the wave's LDS area, D_e = sum A_ei exp(-iz b^2) P over the box
//   (each lane on one plane, rows in steps; plus the extra info in full),
//   then 1 / D_e and the back-projection's factors F = A_ei / D_e and
//   dy tan are written to global memory, and the extra info's update to the
//   image-space extra sums (twice, for the visit it repeats, under the
//   kernel-cache semantics, which also leaves that visit's F at 0).
//   Backward: the slab is cleared and split into one rectangle per wave
//   (n_waves / 2 contiguous row ranges with about equal (row, plane) pair
//   counts over the events' boxes, each cut in two at the boxes' mean
//   plane); every wave reads the events 64 at a time, picks by ballot
//   those whose box meets its rectangle, and adds F exp(-iz b^2) over its
//   part of each box, up to 64 rows per unit, loading the next unit's
//   factors while it sums the current one.  Each slab word has one writer,
//   so the sums are plain LDS read-modify-writes (LDS float atomics run at
//   about one lane per clock), in event order per word.
// The slab is finally written to bp.

template <bool LAST_WRITE>
__global__ __launch_bounds__(1024) void lor_iteration_kernel(
    Grid g, Kernel k, float kexp, const float4* __restrict__ seg_a,
    const float4* __restrict__ seg_b, const int2* __restrict__ ranges,
    const int* __restrict__ pix, const int* __restrict__ ipix,
    const float* __restrict__ upi, const float* __restrict__ w,
    const float* __restrict__ sens, const float* __restrict__ G,
    const float* __restrict__ rho, const int4* __restrict__ ev_i,
    const float4* __restrict__ ev_f, const int* __restrict__ ev_dup,
    const int64_t* __restrict__ fac_off, const int* __restrict__ lor_events,
    const int* __restrict__ lor_order, int n_active, int* __restrict__ queue,
    float* __restrict__ inv, float2* __restrict__ fac, float* __restrict__ bp,
    float* __restrict__ extra, int slab_budget) {
  extern __shared__ __align__(16) float lds[];
  __shared__ int s_next;
  __shared__ unsigned long long s_zw[2];  // box pairs, pairs x (pb + pe)
  const int np = g.n_planes;
  const int n_waves = blockDim.x >> 6;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (;;) {
    if (threadIdx.x == 0) s_next = atomicAdd(queue, 1);
    __syncthreads();
    const int qi = s_next;
    __syncthreads();  // every thread has read s_next
    if (qi >= n_active) break;
    const int lor = lor_order[qi];
    const int e0 = lor_events[lor], e1 = lor_events[lor + 1];
    const int2 lr = ranges[lor];
    const int n_lor = lr.y - lr.x;
    // planes per slab chunk: the slab holds the LOR's infos x pc planes (all
    // of them where they fit), then the infos' up, then the waves' areas
    int pc = n_lor > 0 ? (slab_budget - n_lor - 2) / n_lor : np;
    pc = pc < np ? pc : np;
    const int n_chunks = (np + pc - 1) / pc;
    float* slab = lds;
    float* s_up = slab + (size_t)n_lor * pc;
    // the waves' areas of kWaveWords floats (8-byte aligned): an event's
    // factors (A, dy tan) per info
    float* areas = lds + (((size_t)n_lor * pc + n_lor + 2) & ~(size_t)1);
    float* warea = areas + (size_t)wave * kWaveWords;
    float2* wAT = (float2*)warea;
    const float4 sa = seg_a[lor];
    const float R = seg_b[lor].w;
    for (int t = threadIdx.x; t < n_lor; t += blockDim.x)
      s_up[t] = upi[lr.x + t];
    // the rows' pixel and sensitivity, staged for each chunk in the areas
    // (the forward pass then takes them over)
    int* s_pix = (int*)areas;
    float* s_sens = (float*)(s_pix + n_lor);
    for (int ch = 0; ch < n_chunks; ++ch) {
      const int z0 = ch * pc;
      const int cw = pc < np - z0 ? pc : np - z0;  // the chunk's planes
      __syncthreads();  // the last chunk's P and areas are no longer read
      for (int t = threadIdx.x; t < n_lor; t += blockDim.x) {
        const int p = ipix[lr.x + t];
        s_pix[t] = p;
        s_sens[t] = p >= 0 ? sens[p] : 0.0f;
      }
      __syncthreads();
      // P = G rho s of the chunk's planes (row pitch cw), eight cells per
      // thread in flight
      {
        const int n_cells = n_lor * cw;
        const float* Grows = G + (int64_t)lr.x * np + z0;
        const float inv_cw = 1.0f / (float)cw;
        for (int c0 = threadIdx.x; c0 < n_cells; c0 += 8 * blockDim.x) {
          float gv[8], rv[8], sv[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int c = c0 + u * blockDim.x;
            gv[u] = rv[u] = sv[u] = 0.0f;
            if (c < n_cells) {
              // the row from a float estimate (within one of it; the slab
              // has fewer than 2^16 cells)
              int row = (int)((float)c * inv_cw);
              if (row * cw > c) --row;
              if ((row + 1) * cw <= c) ++row;
              const int p = s_pix[row];
              if (p >= 0) {
                const int zz = c - row * cw;
                gv[u] = Grows[row * np + zz];
                rv[u] = rho[p * np + z0 + zz];
                sv[u] = s_sens[row];
              }
            }
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int c = c0 + u * blockDim.x;
            if (c < n_cells) slab[c] = gv[u] * rv[u] * sv[u];
          }
        }
      }
      __syncthreads();

      // forward: wave w takes events e0 + w, e0 + w + n_waves, ... (the
      // same events in every chunk)
      for (int base = e0 + wave; base < e1; base += 64 * n_waves) {
        EventBatch bt;
        const int mine = base + lane * n_waves;
        bt.load(ev_i, ev_f, nullptr, ev_dup, mine, mine < e1);
        const int cnt = min(64, (e1 - base + n_waves - 1) / n_waves);
        for (int j = 0; j < cnt; ++j) {
          const int e = base + j * n_waves;
          const int4 ei = make_int4(lane_int(bt.i.x, j), lane_int(bt.i.y, j),
                                    lane_int(bt.i.z, j), lane_int(bt.i.w, j));
          const float4 ef = make_float4(lane_float(bt.f.x, j),
                                        lane_float(bt.f.y, j),
                                        lane_float(bt.f.z, j),
                                        lane_float(bt.f.w, j));
          const int dup = LAST_WRITE ? lane_int(bt.dup, j) : -1;
          const int pb = ei.w & 0xffff, pe = ei.w >> 16;
          const int ie_l = ei.z < lr.y ? ei.z : lr.y;
          const int a0 = ei.y - lr.x;
          const int n_info = ie_l - ei.y;
          // the box's planes in this chunk
          const int qb = pb > z0 ? pb : z0;
          const int qe = pe < z0 + cw ? pe : z0 + cw;
          if (n_chunks > 1 && qe <= qb) {
            // D_e is summed over the chunks in inv[e], by this wave's lane 0
            if (ch == 0 && lane == 0) inv[e] = 0.0f;
            continue;
          }
          for (int t = lane; t < n_info; t += 64) {
            float A, T;
            info_factors(k, ef, R, s_up[a0 + t], A, T);
            wAT[t] = make_float2(A, T);
          }
          wave_lds_sync();
          float den = 0.0f;
          const int n_q = qe - qb;
          for (int c0 = 0; c0 < n_q; c0 += 64) {
            const PlaneWalk pw(lane, n_q, c0);
            if (!pw.active) continue;
            const int z = qb + pw.r;
            const float c = plane_centre(g, z) - ef.y;
            const float* prow = slab + __mul24(a0 + pw.q0, cw) + (z - z0);
            const int stride = __mul24(pw.rpi, cw);
            for (int q = pw.q0; q < n_info; q += pw.rpi) {
              const float2 at = wAT[q];
              const float d = c - at.y;
              const float K = __builtin_amdgcn_exp2f(kexp * d * d);
              den += at.x * K * *prow;
              prow += stride;
            }
          }
          if (n_chunks > 1) {
            den = strip::wave_sum(den);
            if (lane == 0) inv[e] = (ch == 0 ? 0.0f : inv[e]) + den;
          } else {
            finish_event<LAST_WRITE>(g, k, sa, R, ef, ei, e, dup, lr.y, pb, pe,
                                     n_info, np, lane, den, pix, w, rho,
                                     sens, fac_off, wAT, inv, fac, extra);
          }
          wave_lds_sync();  // this event's factors are read before the next's
        }
      }
    }
    if (n_chunks > 1) {
      // D_e summed over the chunks: 1 / D_e, the factors and the extra
      // info's update, each event by the wave that summed it
      for (int base = e0 + wave; base < e1; base += 64 * n_waves) {
        EventBatch bt;
        const int mine = base + lane * n_waves;
        bt.load(ev_i, ev_f, nullptr, ev_dup, mine, mine < e1);
        const int cnt = min(64, (e1 - base + n_waves - 1) / n_waves);
        for (int j = 0; j < cnt; ++j) {
          const int e = base + j * n_waves;
          const int4 ei = make_int4(lane_int(bt.i.x, j), lane_int(bt.i.y, j),
                                    lane_int(bt.i.z, j), lane_int(bt.i.w, j));
          const float4 ef = make_float4(lane_float(bt.f.x, j),
                                        lane_float(bt.f.y, j),
                                        lane_float(bt.f.z, j),
                                        lane_float(bt.f.w, j));
          const int dup = LAST_WRITE ? lane_int(bt.dup, j) : -1;
          const int pb = ei.w & 0xffff, pe = ei.w >> 16;
          const int ie_l = ei.z < lr.y ? ei.z : lr.y;
          const int a0 = ei.y - lr.x;
          const int n_info = ie_l - ei.y;
          for (int t = lane; t < n_info; t += 64) {
            float A, T;
            info_factors(k, ef, R, s_up[a0 + t], A, T);
            wAT[t] = make_float2(A, T);
          }
          wave_lds_sync();
          // the chunks' sum enters through lane 0, the extra info's terms
          // through every lane
          const float acc = inv[e];
          finish_event<LAST_WRITE>(g, k, sa, R, ef, ei, e, dup, lr.y, pb, pe,
                                   n_info, np, lane, lane == 0 ? acc : 0.0f,
                                   pix, w, rho, sens, fac_off, wAT, inv, fac,
                                   extra);
          wave_lds_sync();
        }
      }
    }
    // every 1 / D_e and factor is written and P is no longer read; other
    // waves of this block read them back from global memory, which the
    // barrier's workgroup-scope release and acquire order (one CU, one L1)
    __syncthreads();
    // back-projection over rectangles of the slab, one per wave: the rows in
    // n_waves / 2 contiguous ranges of about equal (row, plane) pair counts
    // of the event boxes, each range split in two at the boxes' mean plane
    // (weighted by their pairs; the chunk's middle plane where the slab
    // holds a chunk of the planes).  The rows' weights (n_pl per covered
    // row) go in as differences, then their running sum C(r), where the
    // infos' up was (no longer read)
    int* s_dw = (int*)s_up;
    float* s_C = (float*)s_dw;
    for (int t = threadIdx.x; t <= n_lor; t += blockDim.x) s_dw[t] = 0;
    if (threadIdx.x == 0) s_zw[0] = s_zw[1] = 0;
    __syncthreads();
    for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
      const int4 ei = ev_i[e];
      const int pb = ei.w & 0xffff, pe = ei.w >> 16;
      const int n_pl = pe - pb;
      const int ie_l = ei.z < lr.y ? ei.z : lr.y;
      const int a0 = ei.y - lr.x, a1 = ie_l - lr.x;
      if (n_pl > 0 && a1 > a0 && inv[e] > 0.0f) {
        atomicAdd(&s_dw[a0], n_pl);
        atomicAdd(&s_dw[a1], -n_pl);
        const unsigned long long pairs =
            (unsigned long long)(a1 - a0) * (unsigned long long)n_pl;
        atomicAdd(&s_zw[0], pairs);
        atomicAdd(&s_zw[1], pairs * (unsigned long long)(pb + pe));
      }
    }
    __syncthreads();
    if (wave == 0) {
      int carry_v = 0;
      float carry_c = 0.0f;
      for (int c0 = 0; c0 < n_lor; c0 += 64) {
        const int r = c0 + lane;
        const int v = wave_scan_int(r < n_lor ? s_dw[r] : 0) + carry_v;
        carry_v = lane_int(v, 63);
        const float c = wave_scan_float((float)v) + carry_c;
        carry_c = lane_float(c, 63);
        if (r < n_lor) s_C[r] = c;
      }
    }
    __syncthreads();
    // row r goes to range floor(n_ranges (C(r - 1) + C(r)) / 2 /
    // C(n_lor - 1)); range k takes rows [#{owner < k}, #{owner <= k}), a
    // partition of the rows into contiguous ranges whatever the rounding.
    // Wave w takes range w / 2 and the planes below (w even) or from (w odd)
    // the split plane.
    const int n_ranges = n_waves >> 1;
    const int range = wave >> 1;
    int r_start = 0, r_end = 0;
    {
      const float W = n_lor > 0 ? s_C[n_lor - 1] : 0.0f;
      const float scale = W > 0.0f ? (float)n_ranges / W : 0.0f;
      for (int c0 = 0; c0 < n_lor; c0 += 64) {
        const int r = c0 + lane;
        int o = n_ranges;
        if (r < n_lor) {
          const float mid = 0.5f * ((r > 0 ? s_C[r - 1] : 0.0f) + s_C[r]);
          o = (int)(mid * scale);
          o = o < n_ranges - 1 ? o : n_ranges - 1;
        }
        r_start += __popcll(__ballot(r < n_lor && o < range));
        r_end += __popcll(__ballot(r < n_lor && o <= range));
      }
    }
    int z_split = np >> 1;
    if (s_zw[0] > 0) {
      // the pair-weighted mean of the boxes' plane midpoints (pb + pe) / 2
      z_split = (int)((double)s_zw[1] / (2.0 * (double)s_zw[0]) + 0.5);
      z_split = z_split < 1 ? 1 : (z_split > np - 1 ? np - 1 : z_split);
    }
    for (int ch = 0; ch < n_chunks; ++ch) {
      const int z0 = ch * pc;
      const int cw = pc < np - z0 ? pc : np - z0;  // the chunk's planes
      const int zs = n_chunks > 1 ? z0 + (cw >> 1) : z_split;
      const int z_lo = (wave & 1) ? zs : z0;
      const int z_hi = (wave & 1) ? z0 + cw : zs;
      for (int t = threadIdx.x; t < n_lor * cw; t += blockDim.x)
        slab[t] = 0.0f;
      __syncthreads();
      for (int base = e0; base < e1; base += 64) {
        const int e = base + lane;
        int4 ei = make_int4(0, 0, 0, 0);
        float iv = 0.0f, right = 0.0f;
        int64_t off = 0;
        if (e < e1) {
          ei = ev_i[e];
          iv = inv[e];
          right = ev_f[e].y;
          off = fac_off[e];
        }
        const int ie_l = ei.z < lr.y ? ei.z : lr.y;
        const int a0 = ei.y - lr.x, a1 = ie_l - lr.x;
        // this wave's rows and planes of the event's box
        const int lo = a0 > r_start ? a0 : r_start;
        const int hi = a1 < r_end ? a1 : r_end;
        const int pb = (ei.w & 0xffff) > z_lo ? (ei.w & 0xffff) : z_lo;
        const int pe = (ei.w >> 16) < z_hi ? (ei.w >> 16) : z_hi;
        const bool mine = e < e1 && iv > 0.0f && pe > pb && hi > lo;
        uint64_t mask = __ballot(mine);
        // the wave's units (event j, up to 64 rows from r0) in order; the
        // factors of the next unit are loaded while the current one is
        // summed
        int nj = -1, nr0 = 0;
        float2 nf = make_float2(0.f, 0.f);
        if (mask) {
          nj = __builtin_ctzll(mask);
          mask &= mask - 1;
          nr0 = lane_int(lo, nj);
        }
        if (nj >= 0) nf = unit_factors(fac, a0, hi, off, nj, nr0, lane);
        while (nj >= 0) {
          const int j = nj, r0 = nr0;
          const float2 f = nf;
          const int jhi = lane_int(hi, j);
          const int jpb = lane_int(pb, j), jpe = lane_int(pe, j);
          const float jright = lane_float(right, j);
          if (r0 + 64 < jhi) {
            nr0 = r0 + 64;
          } else if (mask) {
            nj = __builtin_ctzll(mask);
            mask &= mask - 1;
            nr0 = lane_int(lo, nj);
          } else {
            nj = -1;
          }
          if (nj >= 0) nf = unit_factors(fac, a0, hi, off, nj, nr0, lane);
          const int rows = (r0 + 64 < jhi ? r0 + 64 : jhi) - r0;
          const int n_pl = jpe - jpb;
          // the unit's row factors in the wave's area
          if (lane < rows) wAT[lane] = f;
          wave_lds_sync();
          for (int c0 = 0; c0 < n_pl; c0 += 64) {
            const PlaneWalk pw(lane, n_pl, c0);
            if (!pw.active) continue;
            const int z = jpb + pw.r;
            const float c = plane_centre(g, z) - jright;
            float* srow = slab + __mul24(r0 + pw.q0, cw) + (z - z0);
            const int stride = __mul24(pw.rpi, cw);
            for (int q = pw.q0; q < rows; q += pw.rpi) {
              const float2 ft = wAT[q];
              const float d = c - ft.y;
              const float K = __builtin_amdgcn_exp2f(kexp * d * d);
              *srow += ft.x * K;
              srow += stride;
            }
          }
          wave_lds_sync();  // the area is read before the next unit's
        }
      }
      __syncthreads();
      // the chunk's columns of the LOR's rows of bp
      float* rows_out = bp + (int64_t)lr.x * np + z0;
      if (cw == np) {
        for (int t = threadIdx.x; t < n_lor * np; t += blockDim.x)
          rows_out[t] = slab[t];
      } else {
        const float inv_cw = 1.0f / (float)cw;
        for (int t = threadIdx.x; t < n_lor * cw; t += blockDim.x) {
          int row = (int)((float)t * inv_cw);
          if (row * cw > t) --row;
          if ((row + 1) * cw <= t) ++row;
          rows_out[(int64_t)row * np + (t - row * cw)] = slab[t];
        }
      }
      __syncthreads();  // the slab is read before the next chunk or LOR
    }
  }
}

// LORs with events, most events first: (count, LOR) sorted descending.
__global__ __launch_bounds__(256) void lor_count_keys_kernel(
    const int* __restrict__ lor_events, int64_t n_lors,
    unsigned* __restrict__ keys, int* __restrict__ idx,
    int* __restrict__ n_active) {
  int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n_lors) return;
  const int c = lor_events[l + 1] - lor_events[l];
  keys[l] = (unsigned)c;
  idx[l] = (int)l;
  if (c > 0) atomicAdd(n_active, 1);
}

// rho'(p, z) = rho(p, z) sum_i G(i, z) bp(i, z) + extra(p, z) over the
// infos of pixel p (in info order): one wavefront per pixel, lanes over
// planes.
__global__ __launch_bounds__(256) void lor_gather_kernel(
    const int* __restrict__ offsets, const int* __restrict__ order,
    const float* __restrict__ G, const float* __restrict__ bp,
    const float* __restrict__ rho, const float* __restrict__ extra,
    int n_pixels, int n_planes, float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int n_waves = gridDim.x * (blockDim.x >> 6);
  const int64_t np = n_planes;
  for (int p = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
       p < n_pixels; p += n_waves) {
    const int k0 = offsets[p], k1 = offsets[p + 1];
    for (int z0 = 0; z0 < n_planes; z0 += 128) {
      // two plane chunks per pass; the pixel's info indices are read 64 at
      // a time, one per lane, and four infos are summed per step
      const int za = z0 + lane, zb = z0 + 64 + lane;
      const bool in_a = za < n_planes, in_b = zb < n_planes;
      float acc_a = 0.0f, acc_b = 0.0f;
      for (int kb = k0; kb < k1; kb += 64) {
        const int mine = kb + lane < k1 ? order[kb + lane] : 0;
        const int cnt = k1 - kb < 64 ? k1 - kb : 64;
        int j = 0;
        for (; j + 4 <= cnt; j += 4) {
          float ga[4], ba[4], gb[4], bb[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int64_t row = (int64_t)lane_int(mine, j + u) * np;
            ga[u] = in_a ? G[row + za] : 0.0f;
            ba[u] = in_a ? bp[row + za] : 0.0f;
            gb[u] = in_b ? G[row + zb] : 0.0f;
            bb[u] = in_b ? bp[row + zb] : 0.0f;
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            acc_a += ga[u] * ba[u];
            acc_b += gb[u] * bb[u];
          }
        }
        for (; j < cnt; ++j) {
          const int64_t row = (int64_t)lane_int(mine, j) * np;
          if (in_a) acc_a += G[row + za] * bp[row + za];
          if (in_b) acc_b += G[row + zb] * bp[row + zb];
        }
      }
      if (in_a) {
        const int64_t v = (int64_t)p * np + za;
        out[v] = rho[v] * acc_a + extra[v];
      }
      if (in_b) {
        const int64_t v = (int64_t)p * np + zb;
        out[v] = rho[v] * acc_b + extra[v];
      }
    }
  }
}

// Plan: LOR keys of the events, per-LOR event offsets, the events whose
// range has an extra info, and the largest per-event info count.
__global__ __launch_bounds__(256) void event_keys_kernel(
    const int4* __restrict__ ev_i, int64_t n, unsigned* __restrict__ keys,
    int* __restrict__ idx) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  keys[e] = (unsigned)ev_i[e].x;
  idx[e] = (int)e;
}

__global__ __launch_bounds__(256) void permute_events_kernel(
    const int* __restrict__ order, int64_t n, const int4* __restrict__ a_i,
    const float4* __restrict__ a_f, const int* __restrict__ a_dup,
    const int2* __restrict__ ranges, int4* __restrict__ b_i,
    float4* __restrict__ b_f, int* __restrict__ b_dup,
    int* __restrict__ lor_count, int* __restrict__ max_infos,
    int64_t* __restrict__ n_info) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int src = order[e];
  const int4 ei = a_i[src];
  b_i[e] = ei;
  b_f[e] = a_f[src];
  b_dup[e] = a_dup[src];
  atomicAdd(&lor_count[ei.x], 1);
  const int2 lr = ranges[ei.x];
  const int ie_l = ei.z < lr.y ? ei.z : lr.y;
  atomicMax(max_infos, ie_l - ei.y);
  n_info[e] = ie_l - ei.y;
}

__global__ __launch_bounds__(256) void iota_kernel(int* __restrict__ p,
                                                   int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (int)i;
}

// Reference voxel order (z n_rows + y) n_columns + x <-> plane-fastest
// pixel n_planes + z.
__global__ __launch_bounds__(256) void to_device_order_kernel(
    const float* __restrict__ in, int n_pixels, int n_planes,
    float* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n_pixels * n_planes) return;
  const int64_t z = i / n_pixels, p = i - z * n_pixels;
  out[p * n_planes + z] = in[i];
}

__global__ __launch_bounds__(256) void to_reference_order_kernel(
    const float* __restrict__ in, int n_pixels, int n_planes,
    float* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n_pixels * n_planes) return;
  const int64_t z = i / n_pixels, p = i - z * n_pixels;
  out[i] = in[p * n_planes + z];
}

__global__ __launch_bounds__(256) void fill_kernel(float* __restrict__ p,
                                                   int64_t n, float v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

inline unsigned blocks_for(int64_t n, int per = 256) {
  return (unsigned)((n + per - 1) / per);
}

}  // namespace
}  // namespace hybrid

using namespace hybrid;

#define HY_TRY(expr)                                                      \
  do {                                                                    \
    hipError_t _err = (expr);                                             \
    if (_err != hipSuccess)                                               \
      return strip::fail(STRIP_MLEM_ERR_HIP, std::string(#expr) + ": " +  \
                                                 hipGetErrorString(_err)); \
  } while (0)

struct hybrid_mlem_ctx {
  int device = 0;
  int semantics = HYBRID_MLEM_CPU;
  hipStream_t stream = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  Grid grid{};
  Kernel kernel{};
  int64_t n_lors = 0;
  int n_infos = 0;
  bool prepared = false;
  // geometry
  float* segments = nullptr;  // n_lors x 4 as given
  float* widths = nullptr;
  float4* seg_a = nullptr;
  float4* seg_b = nullptr;
  int2* ranges = nullptr;
  int* pix = nullptr;
  float* pos = nullptr;
  float* w_given = nullptr;
  float* w = nullptr;
  float* sens = nullptr;
  // image (plane-fastest) and its update
  float* rho = nullptr;
  float* out = nullptr;
  float* stage = nullptr;  // reference-order staging
  Events ev;
  float last_ms = 0.0f;
  // per-iteration timing events of the last hybrid_mlem_iterate call:
  // start, before the gather, end
  std::vector<hipEvent_t> t_start, t_mid, t_end;
  double timed_event_ms = 0.0, timed_gather_ms = 0.0;
  int timed_iterations = 0;
  // per-info LOR, pixel index and up; infos sorted by pixel (CSR)
  int* info_lor = nullptr;
  int* ipix = nullptr;
  float* upi = nullptr;
  int* csr_offsets = nullptr;
  int* csr_order = nullptr;
  int max_lor_infos = 0;
  // LOR-space tables (n_infos x n_planes) and the image-space extra sums
  float* G = nullptr;
  float* bp = nullptr;
  float* extra = nullptr;
  bool tables = false;
  int lor_waves = 0;  // waves per block of the LOR-space event kernel
  size_t lor_lds = 0;
  int lor_slab_budget = 0;  // LDS floats for a LOR's slab and infos' up
  // algorithm: 0 automatic, 1 wavefront per event with image atomics,
  // 2 LOR space
  int algorithm_request = 0;
  int algorithm = 1;
  // plan of the LOR-space iteration: events sorted by LOR
  bool plan_valid = false;
  int4* pl_i = nullptr;
  float4* pl_f = nullptr;
  int* pl_dup = nullptr;
  float* inv = nullptr;
  int* lor_events = nullptr;
  int64_t* fac_off = nullptr;  // per event: its first (event, info) factor
  float2* fac = nullptr;       // per (event, info): A / D_e, dy tan
  int64_t fac_capacity = 0;
  int* lor_order = nullptr;  // LORs with events, most events first
  int* queue = nullptr;      // the back-projection's LOR counter
  int n_active = 0;
  int n_cus = 0;
  int max_event_infos = 0;
  int64_t pl_capacity = 0;

  ~hybrid_mlem_ctx() {
    if (stream) (void)hipStreamSynchronize(stream);
    void* ptrs[] = {segments, widths, seg_a, seg_b, ranges, pix, pos,
                    w_given, w, sens, rho, out, stage, ev.ev_i, ev.ev_f,
                    ev.ev_dup, info_lor, ipix, upi, csr_offsets, csr_order,
                    G, bp, extra, pl_i, pl_f, pl_dup, inv, lor_events,
                    lor_order, queue, fac_off, fac};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    for (auto* v : {&t_start, &t_mid, &t_end})
      for (hipEvent_t e : *v) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
  int64_t n_voxels() const { return (int64_t)grid.n_pixels * grid.n_planes; }
};

namespace {

template <typename T>
int device_alloc(T** p, size_t count) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return strip::fail(STRIP_MLEM_ERR_ALLOC,
                       std::string("hipMalloc of ") +
                           std::to_string(sizeof(T) * count) +
                           " bytes failed: " + hipGetErrorString(e));
  }
  return STRIP_MLEM_OK;
}

#define HY_CHECK(call)            \
  do {                            \
    int _rc = (call);             \
    if (_rc != STRIP_MLEM_OK) return _rc; \
  } while (0)

int grow_events(hybrid_mlem_ctx* c, int64_t need) {
  if (need <= c->ev.capacity) return STRIP_MLEM_OK;
  int64_t cap = std::max<int64_t>(need, c->ev.capacity * 3 / 2);
  int4* a = nullptr;
  float4* b = nullptr;
  int* d = nullptr;
  HY_CHECK(device_alloc(&a, cap));
  if (device_alloc(&b, cap) != STRIP_MLEM_OK ||
      device_alloc(&d, cap) != STRIP_MLEM_OK) {
    (void)hipFree(a);
    if (b) (void)hipFree(b);
    return STRIP_MLEM_ERR_ALLOC;
  }
  if (c->ev.n) {
    HY_TRY(hipMemcpyAsync(a, c->ev.ev_i, sizeof(int4) * c->ev.n,
                          hipMemcpyDeviceToDevice, c->stream));
    HY_TRY(hipMemcpyAsync(b, c->ev.ev_f, sizeof(float4) * c->ev.n,
                          hipMemcpyDeviceToDevice, c->stream));
    HY_TRY(hipMemcpyAsync(d, c->ev.ev_dup, sizeof(int) * c->ev.n,
                          hipMemcpyDeviceToDevice, c->stream));
    HY_TRY(hipStreamSynchronize(c->stream));
  }
  if (c->ev.ev_i) (void)hipFree(c->ev.ev_i);
  if (c->ev.ev_f) (void)hipFree(c->ev.ev_f);
  if (c->ev.ev_dup) (void)hipFree(c->ev.ev_dup);
  c->ev.ev_i = a;
  c->ev.ev_f = b;
  c->ev.ev_dup = d;
  c->ev.capacity = cap;
  return STRIP_MLEM_OK;
}

// Infos inside the grid sorted stably by pixel (radix sort of (pixel,
// info) pairs) with per-pixel offsets: the order of calculate_sensitivity's
// per-pixel sums (reconstruction.h:240-251) and the gather's index.
int build_pixel_index(hybrid_mlem_ctx* c) {
  const Grid& g = c->grid;
  const int n = c->n_infos;
  int rc = STRIP_MLEM_OK;
  if ((rc = device_alloc(&c->csr_offsets, (size_t)g.n_pixels + 1)) ||
      (rc = device_alloc(&c->csr_order, n)))
    return rc;
  if (n == 0) {
    HY_TRY(hipMemsetAsync(c->csr_offsets, 0,
                          sizeof(int) * ((size_t)g.n_pixels + 1), c->stream));
    return STRIP_MLEM_OK;
  }
  int* count = nullptr;
  unsigned *keys = nullptr, *keys_sorted = nullptr;
  int* iota = nullptr;
  char* tmp = nullptr;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(c->stream);
    void* ps[] = {count, keys, keys_sorted, iota, tmp};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  };
  size_t sort_bytes = 0, scan_bytes = 0;
  if ((rc = device_alloc(&count, (size_t)g.n_pixels + 1)) ||
      (rc = device_alloc(&keys, n)) || (rc = device_alloc(&keys_sorted, n)) ||
      (rc = device_alloc(&iota, n))) {
    cleanup();
    return rc;
  }
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(
      nullptr, sort_bytes, keys, keys_sorted, iota, c->csr_order, n, 0, 32,
      c->stream);
  if (e == hipSuccess)
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, count,
                                         c->csr_offsets, g.n_pixels + 1,
                                         c->stream);
  if (e == hipSuccess &&
      (rc = device_alloc(&tmp, std::max(sort_bytes, scan_bytes)))) {
    cleanup();
    return rc;
  }
  if (e == hipSuccess)
    e = hipMemsetAsync(count, 0, sizeof(int) * ((size_t)g.n_pixels + 1),
                       c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(iota_kernel, dim3(blocks_for(n)), dim3(256), 0,
                       c->stream, iota, (int64_t)n);
    hipLaunchKernelGGL(pixel_count_kernel, dim3(blocks_for(n)), dim3(256), 0,
                       c->stream, g, c->pix, (int64_t)n, count, keys);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, keys, keys_sorted,
                                           iota, c->csr_order, n, 0, 32,
                                           c->stream);
  if (e == hipSuccess)
    e = hipcub::DeviceScan::ExclusiveSum(tmp, scan_bytes, count,
                                         c->csr_offsets, g.n_pixels + 1,
                                         c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  cleanup();
  if (e != hipSuccess)
    return strip::fail(STRIP_MLEM_ERR_HIP,
                       std::string("pixel index: ") + hipGetErrorString(e));
  return STRIP_MLEM_OK;
}

// LOR-space tables: G (recomputed with every weight preparation) and bp
// (rows of LORs without events stay zero), the extra accumulator, and the
// iteration kernel's block: the largest LOR's slab (infos x planes) and
// its infos' up with a 2 KiB area per wave (an event's factors), within
// 160 KiB of LDS, 16 waves (or 8); else all of LDS with 16 waves, the slabs
// then holding chunks of the planes (at least one plane of the largest
// LOR).  Where that does not fit, or the tables do not fit in device
// memory, the context stays on the wavefront-per-event algorithm.
int prepare_tables(hybrid_mlem_ctx* c) {
  const Grid& g = c->grid;
  const size_t slab =
      sizeof(float) * ((size_t)c->max_lor_infos * ((size_t)g.n_planes + 1) + 2);
  const size_t per_wave = sizeof(float) * kWaveWords;
  // every plane of the largest LOR with 16 (or 8) waves where that fits,
  // else all of LDS with 16 waves and the slab in chunks of planes
  int waves = 16;
  if (slab + waves * per_wave > kLdsBytes) waves = 8;
  if (slab + waves * per_wave <= kLdsBytes) {
    c->lor_lds = slab + waves * per_wave;
  } else {
    waves = 16;
    c->lor_lds = kLdsBytes;
  }
  const size_t budget = c->lor_lds / sizeof(float) - (size_t)waves * kWaveWords;
  // at least one plane of the largest LOR; the factor area also stages the
  // rows' pixel and sensitivity
  if (budget < 2 * (size_t)c->max_lor_infos + 2 ||
      2 * sizeof(float) * (size_t)c->max_lor_infos > waves * per_wave) {
    c->tables = false;
    return STRIP_MLEM_OK;
  }
  c->lor_waves = waves;
  c->lor_slab_budget = (int)budget;
  const int64_t cells = (int64_t)c->n_infos * g.n_planes;
  if (!c->G) {
    const bool ok =
        hipFuncSetAttribute((const void*)lor_iteration_kernel<true>,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kLdsBytes) == hipSuccess &&
        hipFuncSetAttribute((const void*)lor_iteration_kernel<false>,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kLdsBytes) == hipSuccess &&
        device_alloc(&c->G, cells) == STRIP_MLEM_OK &&
        device_alloc(&c->bp, cells) == STRIP_MLEM_OK &&
        device_alloc(&c->extra, c->n_voxels()) == STRIP_MLEM_OK;
    if (!ok) {
      (void)hipGetLastError();
      void* ps[] = {c->G, c->bp, c->extra};
      for (void* q : ps)
        if (q) (void)hipFree(q);
      c->G = c->bp = c->extra = nullptr;
      c->tables = false;
      return STRIP_MLEM_OK;
    }
    HY_TRY(hipMemsetAsync(c->bp, 0, sizeof(float) * cells, c->stream));
  }
  if (cells)
    hipLaunchKernelGGL(frame_table_kernel,
                       dim3((unsigned)std::min<int64_t>(blocks_for(cells),
                                                        (int64_t)1 << 20)),
                       dim3(256), 0, c->stream, g, c->kernel, c->seg_b,
                       c->info_lor, c->ipix, c->upi, c->w, c->n_infos, c->G);
  HY_TRY(hipGetLastError());
  c->tables = true;
  return STRIP_MLEM_OK;
}

// Events sorted by LOR (a stable radix sort), per-LOR offsets and the
// largest per-event info count within a LOR.  Rows of bp whose LOR loses
// all its events are cleared.
int build_plan(hybrid_mlem_ctx* c) {
  const int64_t n = c->ev.n;
  int rc = STRIP_MLEM_OK;
  if (n > c->pl_capacity) {
    void* ps[] = {c->pl_i, c->pl_f, c->pl_dup, c->inv, c->fac_off};
    for (void* p : ps)
      if (p) (void)hipFree(p);
    c->pl_i = nullptr;
    c->pl_f = nullptr;
    c->pl_dup = nullptr;
    c->inv = nullptr;
    c->fac_off = nullptr;
    c->pl_capacity = 0;
    if ((rc = device_alloc(&c->pl_i, n)) || (rc = device_alloc(&c->pl_f, n)) ||
        (rc = device_alloc(&c->pl_dup, n)) || (rc = device_alloc(&c->inv, n)) ||
        (rc = device_alloc(&c->fac_off, n + 1)))
      return rc;
    c->pl_capacity = n;
  }
  if (!c->lor_events &&
      (rc = device_alloc(&c->lor_events, (size_t)c->n_lors + 1)))
    return rc;
  unsigned *keys = nullptr, *keys_sorted = nullptr;
  int *idx = nullptr, *order = nullptr, *count = nullptr, *stats = nullptr;
  int64_t* n_info = nullptr;
  char* tmp = nullptr;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(c->stream);
    void* ps[] = {keys, keys_sorted, idx, order, count, stats, n_info, tmp};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  };
  const int64_t nn = n > 0 ? n : 1;
  if ((rc = device_alloc(&keys, nn)) || (rc = device_alloc(&keys_sorted, nn)) ||
      (rc = device_alloc(&idx, nn)) || (rc = device_alloc(&order, nn)) ||
      (rc = device_alloc(&count, (size_t)c->n_lors + 1)) ||
      (rc = device_alloc(&stats, 4)) || (rc = device_alloc(&n_info, nn + 1))) {
    cleanup();
    return rc;
  }
  size_t b1 = 0, b2 = 0, b3 = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(
      nullptr, b1, keys, keys_sorted, idx, order, (int)nn, 0, 32, c->stream);
  if (e == hipSuccess)
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, b2, count, c->lor_events,
                                         (int)(c->n_lors + 1), c->stream);
  if (e == hipSuccess)
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, b3, n_info, c->fac_off,
                                         (int)(nn + 1), c->stream);
  const size_t tb = std::max(b1, std::max(b2, b3));
  if (e == hipSuccess && (rc = device_alloc(&tmp, tb))) {
    cleanup();
    return rc;
  }
  if (e == hipSuccess)
    e = hipMemsetAsync(count, 0, sizeof(int) * ((size_t)c->n_lors + 1),
                       c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(stats, 0, sizeof(int) * 4, c->stream);
  if (e == hipSuccess && n > 0) {
    hipLaunchKernelGGL(event_keys_kernel, dim3(blocks_for(n)), dim3(256), 0,
                       c->stream, c->ev.ev_i, n, keys, idx);
    e = hipGetLastError();
    size_t t1 = tb, t2 = tb;
    if (e == hipSuccess)
      e = hipcub::DeviceRadixSort::SortPairs(tmp, t1, keys, keys_sorted, idx,
                                             order, (int)n, 0, 32, c->stream);
    if (e == hipSuccess)
      e = hipMemsetAsync(n_info + n, 0, sizeof(int64_t), c->stream);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(permute_events_kernel, dim3(blocks_for(n)),
                         dim3(256), 0, c->stream, order, n, c->ev.ev_i,
                         c->ev.ev_f, c->ev.ev_dup, c->ranges, c->pl_i,
                         c->pl_f, c->pl_dup, count, stats, n_info);
      e = hipGetLastError();
    }
    if (e == hipSuccess)
      e = hipcub::DeviceScan::ExclusiveSum(tmp, t2, count, c->lor_events,
                                           (int)(c->n_lors + 1), c->stream);
    size_t t3 = tb;
    if (e == hipSuccess)
      e = hipcub::DeviceScan::ExclusiveSum(tmp, t3, n_info, c->fac_off,
                                           (int)(n + 1), c->stream);
  } else if (e == hipSuccess) {
    e = hipMemsetAsync(c->lor_events, 0,
                       sizeof(int) * ((size_t)c->n_lors + 1), c->stream);
  }
  // the LOR queue: LORs with events, most events first
  if (!c->lor_order && (rc = device_alloc(&c->lor_order, (size_t)c->n_lors))) {
    cleanup();
    return rc;
  }
  if (!c->queue && (rc = device_alloc(&c->queue, 1))) {
    cleanup();
    return rc;
  }
  if (e == hipSuccess && c->n_lors) {
    unsigned* ck = nullptr;
    unsigned* ck_sorted = nullptr;
    int* li = nullptr;
    char* t2 = nullptr;
    size_t tb2 = 0;
    if ((rc = device_alloc(&ck, (size_t)c->n_lors)) ||
        (rc = device_alloc(&ck_sorted, (size_t)c->n_lors)) ||
        (rc = device_alloc(&li, (size_t)c->n_lors))) {
      if (ck) (void)hipFree(ck);
      if (ck_sorted) (void)hipFree(ck_sorted);
      cleanup();
      return rc;
    }
    hipLaunchKernelGGL(lor_count_keys_kernel, dim3(blocks_for(c->n_lors)),
                       dim3(256), 0, c->stream, c->lor_events, c->n_lors, ck,
                       li, stats + 1);
    e = hipGetLastError();
    if (e == hipSuccess)
      e = hipcub::DeviceRadixSort::SortPairsDescending(
          nullptr, tb2, ck, ck_sorted, li, c->lor_order, (int)c->n_lors, 0,
          32, c->stream);
    if (e == hipSuccess && (rc = device_alloc(&t2, tb2)) == STRIP_MLEM_OK)
      e = hipcub::DeviceRadixSort::SortPairsDescending(
          t2, tb2, ck, ck_sorted, li, c->lor_order, (int)c->n_lors, 0, 32,
          c->stream);
    (void)hipStreamSynchronize(c->stream);
    void* ps[] = {ck, ck_sorted, li, t2};
    for (void* q : ps)
      if (q) (void)hipFree(q);
    if (rc) {
      cleanup();
      return rc;
    }
  }
  // a LOR whose events are gone keeps no stale rows
  if (e == hipSuccess && c->bp)
    e = hipMemsetAsync(c->bp, 0,
                       sizeof(float) * (size_t)c->n_infos * c->grid.n_planes,
                       c->stream);
  int host_stats[4] = {0, 0, 0, 0};
  if (e == hipSuccess)
    e = hipMemcpyAsync(host_stats, stats, sizeof(int) * 4,
                       hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  cleanup();
  if (e != hipSuccess)
    return strip::fail(STRIP_MLEM_ERR_HIP,
                       std::string("event plan: ") + hipGetErrorString(e));
  int64_t total = 0;
  if (n > 0) {
    HY_TRY(hipMemcpy(&total, c->fac_off + n, sizeof(int64_t),
                     hipMemcpyDeviceToHost));
  }
  if (total > c->fac_capacity) {
    if (c->fac) (void)hipFree(c->fac);
    c->fac = nullptr;
    c->fac_capacity = 0;
    HY_CHECK(device_alloc(&c->fac, total));
    c->fac_capacity = total;
  }
  c->max_event_infos = host_stats[0];
  c->n_active = n > 0 ? host_stats[1] : 0;
  c->plan_valid = true;
  return STRIP_MLEM_OK;
}

}  // namespace

extern "C" {

int hybrid_mlem_create(const hybrid_mlem_geometry* G, int semantics,
                       int device, hybrid_mlem_ctx** out) {
  if (!out) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (!G) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "geometry is NULL");
  if (semantics != HYBRID_MLEM_CPU && semantics != HYBRID_MLEM_LEGACY_GPU)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "unknown semantics");
  if (G->n_planes_half != 1)
    return strip::fail(STRIP_MLEM_ERR_UNSUPPORTED,
                       "only single-plane pixel weights (n_planes_half = 1) "
                       "are supported");
  if (G->n_detectors <= 0 || G->n_detectors > 32767 ||
      G->n_pixel_infos < 0 || G->n_pixel_infos >= (int64_t)1 << 31)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT,
                       "n_detectors or n_pixel_infos out of range");
  if (G->n_columns <= 0 || G->n_rows <= 0 || G->n_planes <= 0 ||
      G->n_columns > 32767 || G->n_rows > 32767 || G->n_planes > 32767 ||
      !(G->pixel_size > 0) || !(G->sigma_z > 0) || !(G->sigma_dl > 0) ||
      !(G->length > 0))
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "bad grid or kernel sizes");
  if (!G->lor_line_segments || !G->lor_widths || !G->lor_pixel_info_begin ||
      !G->lor_pixel_info_end ||
      (G->n_pixel_infos &&
       (!G->pixels || !G->pixel_positions || !G->pixel_weights)))
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "geometry array is NULL");
  const int64_t n_lors = (int64_t)G->n_detectors * (G->n_detectors + 1) / 2;
  const int n_infos = (int)G->n_pixel_infos;
  const int64_t n_pixels = (int64_t)G->n_columns * G->n_rows;
  if (n_pixels * G->n_planes >= (int64_t)1 << 31)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT,
                       "more than 2^31 voxels (the reference's int index)");
  // host-side checks of the layout the kernels rely on
  std::vector<int2> ranges(n_lors);
  std::vector<int> stamp(n_pixels, -1);
  for (int64_t l = 0; l < n_lors; ++l) {
    const int64_t b = G->lor_pixel_info_begin[l], e = G->lor_pixel_info_end[l];
    if (b < 0 || e < b || e > n_infos)
      return strip::fail(STRIP_MLEM_ERR_ARGUMENT,
                         "LOR " + std::to_string(l) +
                             ": pixel info range outside [0, n_pixel_infos]");
    for (int64_t i = b; i < e; ++i) {
      const int x = G->pixels[2 * i], y = G->pixels[2 * i + 1];
      if (x < 0 || x >= G->n_columns || y < 0 || y >= G->n_rows) continue;
      int& s = stamp[(int64_t)y * G->n_columns + x];
      if (s == (int)l)
        return strip::fail(STRIP_MLEM_ERR_ARGUMENT,
                           "LOR " + std::to_string(l) +
                               ": a pixel appears twice among its infos");
      s = (int)l;
    }
    ranges[l] = make_int2((int)b, (int)e);
  }
  std::vector<int> packed(n_infos);
  for (int i = 0; i < n_infos; ++i)
    packed[i] = (int)(uint16_t)G->pixels[2 * i] |
                ((int)G->pixels[2 * i + 1] << 16);

  hipError_t he = hipSetDevice(device);
  if (he != hipSuccess)
    return strip::fail(STRIP_MLEM_ERR_HIP,
                       std::string("hipSetDevice: ") + hipGetErrorString(he));
  hybrid_mlem_ctx* c = new hybrid_mlem_ctx();
  c->device = device;
  c->semantics = semantics;
  c->n_lors = n_lors;
  c->n_infos = n_infos;
  auto bail = [&](int rc) {
    delete c;
    return rc;
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) !=
          hipSuccess ||
      hipEventCreate(&c->t0) != hipSuccess ||
      hipEventCreate(&c->t1) != hipSuccess)
    return bail(strip::fail(STRIP_MLEM_ERR_HIP, "stream/event creation"));
  Grid& g = c->grid;
  g.n_columns = G->n_columns;
  g.n_rows = G->n_rows;
  g.n_planes = G->n_planes;
  g.n_pixels = (int)n_pixels;
  g.pixel_size = (float)G->pixel_size;
  g.z_left = (float)G->z_left;
  {
    // PixelGrid(n_columns, n_rows, pixel_size): lower left corner
    // (-pixel_size n_columns / 2, -pixel_size n_rows / 2), centre of
    // pixel (0, 0) half a pixel further (pixel_grid.h:25-40)
    volatile float ps = g.pixel_size;
    volatile float llx = -ps * (float)(short)g.n_columns / 2.0f;
    volatile float lly = -ps * (float)(short)g.n_rows / 2.0f;
    volatile float half = ps / 2.0f;
    g.llc_x = llx + half;
    g.llc_y = lly + half;
  }
  Kernel& k = c->kernel;
  {
    volatile float sz = (float)G->sigma_z, sdl = (float)G->sigma_dl;
    volatile float iz = 1.0f / (sz * sz);
    volatile float id = 1.0f / (sdl * sdl);
    volatile float prod = iz * iz * id;
    k.iz = iz;
    k.id = id;
    k.norm_c = (float)(1.0 / (2.0 * M_PI * M_PI) * (double)sqrtf(prod));
    k.length = (float)G->length;
  }
  int rc = STRIP_MLEM_OK;
  const int64_t nv = c->n_voxels();
  if ((rc = device_alloc(&c->segments, n_lors * 4)) ||
      (rc = device_alloc(&c->widths, n_lors)) ||
      (rc = device_alloc(&c->seg_a, n_lors)) ||
      (rc = device_alloc(&c->seg_b, n_lors)) ||
      (rc = device_alloc(&c->ranges, n_lors)) ||
      (rc = device_alloc(&c->pix, n_infos)) ||
      (rc = device_alloc(&c->pos, n_infos)) ||
      (rc = device_alloc(&c->w_given, n_infos)) ||
      (rc = device_alloc(&c->w, n_infos)) ||
      (rc = device_alloc(&c->sens, n_pixels)) ||
      (rc = device_alloc(&c->rho, nv)) || (rc = device_alloc(&c->out, nv)) ||
      (rc = device_alloc(&c->stage, nv)))
    return bail(rc);
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice,
                                  c->stream)
                 : hipSuccess;
  };
  if (up(c->segments, G->lor_line_segments, sizeof(float) * 4 * n_lors) ||
      up(c->widths, G->lor_widths, sizeof(float) * n_lors) ||
      up(c->ranges, ranges.data(), sizeof(int2) * n_lors) ||
      up(c->pix, packed.data(), sizeof(int) * n_infos) ||
      up(c->pos, G->pixel_positions, sizeof(float) * n_infos) ||
      up(c->w_given, G->pixel_weights, sizeof(float) * n_infos))
    return bail(strip::fail(STRIP_MLEM_ERR_HIP, "geometry upload"));
  hipLaunchKernelGGL(segments_kernel, dim3(blocks_for(n_lors)), dim3(256), 0,
                     c->stream, c->segments, n_lors, c->seg_a, c->seg_b);
  hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(nv)), dim3(256), 0,
                     c->stream, c->rho, nv, 1.0f);
  hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(n_pixels)), dim3(256), 0,
                     c->stream, c->sens, n_pixels, 0.0f);
  if (n_infos)
    he = hipMemcpyAsync(c->w, c->w_given, sizeof(float) * n_infos,
                        hipMemcpyDeviceToDevice, c->stream);
  if (he != hipSuccess || hipGetLastError() != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return bail(strip::fail(STRIP_MLEM_ERR_HIP, "geometry setup"));
  for (int64_t l = 0; l < n_lors; ++l)
    c->max_lor_infos = std::max(c->max_lor_infos, ranges[l].y - ranges[l].x);
  if ((rc = device_alloc(&c->info_lor, n_infos)) ||
      (rc = device_alloc(&c->ipix, n_infos)) ||
      (rc = device_alloc(&c->upi, n_infos)))
    return bail(rc);
  if (n_infos) {
    // infos outside every LOR: LOR -1, pixel -1, up 0
    if (hipMemsetAsync(c->info_lor, 0xff, sizeof(int) * n_infos, c->stream) ||
        hipMemsetAsync(c->ipix, 0xff, sizeof(int) * n_infos, c->stream) ||
        hipMemsetAsync(c->upi, 0, sizeof(float) * n_infos, c->stream))
      return bail(strip::fail(STRIP_MLEM_ERR_HIP, "per-info arrays"));
    const int64_t waves = std::min<int64_t>(n_lors, 65536);
    hipLaunchKernelGGL(info_lor_kernel, dim3(blocks_for(waves, 4)), dim3(256),
                       0, c->stream, g, c->seg_a, c->ranges, c->pix, n_lors,
                       c->info_lor, c->ipix, c->upi);
    if (hipGetLastError() != hipSuccess)
      return bail(strip::fail(STRIP_MLEM_ERR_HIP, "per-info arrays"));
  }
  if ((rc = build_pixel_index(c))) return bail(rc);
  if (hipDeviceGetAttribute(&c->n_cus, hipDeviceAttributeMultiprocessorCount,
                            device) != hipSuccess ||
      c->n_cus <= 0)
    c->n_cus = 1;
  *out = c;
  return STRIP_MLEM_OK;
}

void hybrid_mlem_destroy(hybrid_mlem_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  delete ctx;
}

int hybrid_mlem_prepare_weights(hybrid_mlem_ctx* c, int calculate_weight,
                                int sensitivity_mode,
                                const float* sensitivity) {
  if (!c) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  if (sensitivity_mode < 0 || sensitivity_mode > 3 ||
      (sensitivity_mode >= 2 && !sensitivity))
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "bad sensitivity mode");
  HY_TRY(hipSetDevice(c->device));
  const Grid& g = c->grid;
  const int n = c->n_infos;
  if (n)
    HY_TRY(hipMemcpyAsync(c->w, c->w_given, sizeof(float) * n,
                          hipMemcpyDeviceToDevice, c->stream));
  if (calculate_weight) {
    const int64_t waves = std::min<int64_t>(c->n_lors, 65536);
    hipLaunchKernelGGL(calculate_weight_kernel, dim3(blocks_for(waves, 4)),
                       dim3(256), 0, c->stream, g, c->seg_b, c->widths,
                       c->ranges, c->pix, c->n_lors, c->w);
    HY_TRY(hipGetLastError());
  }
  if (sensitivity_mode >= 2) {
    HY_TRY(hipMemcpyAsync(c->sens, sensitivity, sizeof(float) * g.n_pixels,
                          hipMemcpyHostToDevice, c->stream));
  } else if (sensitivity_mode == 1) {
    hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(g.n_pixels)), dim3(256),
                       0, c->stream, c->sens, (int64_t)g.n_pixels, 1.0f);
  } else {
    // each pixel sums its weights in info order
    hipLaunchKernelGGL(sensitivity_sum_kernel, dim3(blocks_for(g.n_pixels)),
                       dim3(256), 0, c->stream, c->csr_offsets, c->csr_order,
                       c->w, g.n_pixels, c->sens);
    HY_TRY(hipGetLastError());
  }
  if (n && sensitivity_mode != HYBRID_MLEM_SENSITIVITY_NORMALISED)
    hipLaunchKernelGGL(normalize_kernel, dim3(blocks_for(n)), dim3(256), 0,
                       c->stream, g, c->pix, c->sens, (int64_t)n, c->w);
  HY_TRY(hipGetLastError());
  HY_CHECK(prepare_tables(c));
  HY_TRY(hipStreamSynchronize(c->stream));
  c->prepared = true;
  return STRIP_MLEM_OK;
}

int hybrid_mlem_add_responses(hybrid_mlem_ctx* c,
                              const hybrid_mlem_response* responses,
                              int64_t n) {
  if (!c) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  if (n < 0 || (n && !responses))
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "bad responses");
  if (!c->prepared)
    return strip::fail(STRIP_MLEM_ERR_STATE,
                       "hybrid_mlem_prepare_weights must come first");
  if (n == 0) return STRIP_MLEM_OK;
  HY_TRY(hipSetDevice(c->device));
  hybrid_mlem_response* d_resp = nullptr;
  int4* ti = nullptr;
  float4* tf = nullptr;
  int* td = nullptr;
  char* keep = nullptr;
  int* flags = nullptr;  // bad LOR flag, then the selected count
  void* tmp = nullptr;
  int rc = STRIP_MLEM_OK;
  auto cleanup = [&]() {
    void* ps[] = {d_resp, ti, tf, td, keep, flags, tmp};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  };
  if ((rc = device_alloc(&d_resp, n)) || (rc = device_alloc(&ti, n)) ||
      (rc = device_alloc(&tf, n)) || (rc = device_alloc(&td, n)) ||
      (rc = device_alloc(&keep, n)) || (rc = device_alloc(&flags, 4))) {
    cleanup();
    return rc;
  }
  hipError_t e = hipMemcpyAsync(d_resp, responses,
                                sizeof(hybrid_mlem_response) * n,
                                hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = hipMemsetAsync(flags, 0, sizeof(int) * 4, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(frame_events_kernel, dim3(blocks_for(n)), dim3(256), 0,
                       c->stream, c->grid, c->kernel, d_resp, n, c->n_lors,
                       c->seg_b, c->ranges, c->pos, c->pix, c->w, c->n_infos,
                       ti, tf, td, keep, flags);
    e = hipGetLastError();
  }
  int host_flags[4] = {0, 0, 0, 0};
  if (e == hipSuccess)
    e = hipMemcpyAsync(host_flags, flags, sizeof(int) * 4,
                       hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    cleanup();
    return strip::fail(STRIP_MLEM_ERR_HIP,
                       std::string("frame events: ") + hipGetErrorString(e));
  }
  if (host_flags[0]) {
    cleanup();
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT,
                       "a response's LOR index is outside the geometry");
  }
  // order-preserving compaction of the kept events
  if ((rc = grow_events(c, c->ev.n + n))) {
    cleanup();
    return rc;
  }
  size_t tb = 0, b2 = 0, b3 = 0;
  e = hipcub::DeviceSelect::Flagged(nullptr, tb, ti, keep,
                                    c->ev.ev_i + c->ev.n, flags + 1, n,
                                    c->stream);
  if (e == hipSuccess)
    e = hipcub::DeviceSelect::Flagged(nullptr, b2, tf, keep,
                                      c->ev.ev_f + c->ev.n, flags + 1, n,
                                      c->stream);
  if (e == hipSuccess)
    e = hipcub::DeviceSelect::Flagged(nullptr, b3, td, keep,
                                      c->ev.ev_dup + c->ev.n, flags + 1, n,
                                      c->stream);
  tb = std::max(tb, std::max(b2, b3));
  if (e == hipSuccess && (rc = device_alloc((char**)&tmp, tb))) {
    cleanup();
    return rc;
  }
  size_t t1 = tb, t2 = tb, t3 = tb;
  if (e == hipSuccess)
    e = hipcub::DeviceSelect::Flagged(tmp, t1, ti, keep, c->ev.ev_i + c->ev.n,
                                      flags + 1, n, c->stream);
  if (e == hipSuccess)
    e = hipcub::DeviceSelect::Flagged(tmp, t2, tf, keep, c->ev.ev_f + c->ev.n,
                                      flags + 1, n, c->stream);
  if (e == hipSuccess)
    e = hipcub::DeviceSelect::Flagged(tmp, t3, td, keep,
                                      c->ev.ev_dup + c->ev.n, flags + 1, n,
                                      c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(host_flags, flags, sizeof(int) * 4,
                       hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  cleanup();
  if (e != hipSuccess)
    return strip::fail(STRIP_MLEM_ERR_HIP,
                       std::string("event compaction: ") +
                           hipGetErrorString(e));
  c->ev.n += host_flags[1];
  c->plan_valid = false;
  return STRIP_MLEM_OK;
}

int hybrid_mlem_clear_events(hybrid_mlem_ctx* c) {
  if (!c) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  c->ev.n = 0;
  c->plan_valid = false;
  return STRIP_MLEM_OK;
}

int64_t hybrid_mlem_n_events(const hybrid_mlem_ctx* c) {
  return c ? c->ev.n : -1;
}

int hybrid_mlem_set_frame_events(hybrid_mlem_ctx* c,
                                 const hybrid_mlem_frame_event* events,
                                 int64_t n) {
  if (!c) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  if (n < 0 || (n && !events))
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "bad frame events");
  HY_TRY(hipSetDevice(c->device));
  c->ev.n = 0;
  c->plan_valid = false;
  if (n == 0) return STRIP_MLEM_OK;
  HY_CHECK(grow_events(c, n));
  hybrid_mlem_frame_event* d = nullptr;
  int* bad = nullptr;
  int rc = STRIP_MLEM_OK;
  if ((rc = device_alloc(&d, n)) || (rc = device_alloc(&bad, 1))) {
    if (d) (void)hipFree(d);
    return rc;
  }
  int host_bad = 0;
  hipError_t e = hipMemcpyAsync(d, events, sizeof(hybrid_mlem_frame_event) * n,
                                hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(bad, 0, sizeof(int), c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(frame_import_kernel, dim3(blocks_for(n)), dim3(256), 0,
                       c->stream, c->grid, d, n, c->n_lors, c->ranges, c->pix,
                       c->n_infos, c->ev.ev_i, c->ev.ev_f, c->ev.ev_dup, bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(&host_bad, bad, sizeof(int), hipMemcpyDeviceToHost,
                       c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  (void)hipFree(bad);
  if (e != hipSuccess)
    return strip::fail(STRIP_MLEM_ERR_HIP,
                       std::string("frame events: ") + hipGetErrorString(e));
  if (host_bad)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT,
                       "a frame event's LOR, info or plane range is outside "
                       "the geometry");
  c->ev.n = n;
  return STRIP_MLEM_OK;
}

int hybrid_mlem_device_name(hybrid_mlem_ctx* c, char* name, int size) {
  if (!c || !name || size <= 0)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "bad arguments");
  hipDeviceProp_t prop;
  HY_TRY(hipGetDeviceProperties(&prop, c->device));
  std::snprintf(name, (size_t)size, "%s (%s)", prop.name, prop.gcnArchName);
  return STRIP_MLEM_OK;
}

int hybrid_mlem_get_frame_events(hybrid_mlem_ctx* c, int32_t* ints,
                                 float* floats) {
  if (!c) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  HY_TRY(hipSetDevice(c->device));
  if (c->ev.n && ints)
    HY_TRY(hipMemcpyAsync(ints, c->ev.ev_i, sizeof(int4) * c->ev.n,
                          hipMemcpyDeviceToHost, c->stream));
  if (c->ev.n && floats)
    HY_TRY(hipMemcpyAsync(floats, c->ev.ev_f, sizeof(float4) * c->ev.n,
                          hipMemcpyDeviceToHost, c->stream));
  HY_TRY(hipStreamSynchronize(c->stream));
  return STRIP_MLEM_OK;
}

int hybrid_mlem_set_rho(hybrid_mlem_ctx* c, const float* rho) {
  if (!c || !rho) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "NULL argument");
  HY_TRY(hipSetDevice(c->device));
  const int64_t nv = c->n_voxels();
  HY_TRY(hipMemcpyAsync(c->stage, rho, sizeof(float) * nv,
                        hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(to_device_order_kernel, dim3(blocks_for(nv)), dim3(256),
                     0, c->stream, c->stage, c->grid.n_pixels,
                     c->grid.n_planes, c->rho);
  HY_TRY(hipGetLastError());
  HY_TRY(hipStreamSynchronize(c->stream));
  return STRIP_MLEM_OK;
}

int hybrid_mlem_get_rho(hybrid_mlem_ctx* c, float* rho) {
  if (!c || !rho) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "NULL argument");
  HY_TRY(hipSetDevice(c->device));
  const int64_t nv = c->n_voxels();
  hipLaunchKernelGGL(to_reference_order_kernel, dim3(blocks_for(nv)),
                     dim3(256), 0, c->stream, c->rho, c->grid.n_pixels,
                     c->grid.n_planes, c->stage);
  HY_TRY(hipGetLastError());
  HY_TRY(hipMemcpyAsync(rho, c->stage, sizeof(float) * nv,
                        hipMemcpyDeviceToHost, c->stream));
  HY_TRY(hipStreamSynchronize(c->stream));
  return STRIP_MLEM_OK;
}

int hybrid_mlem_get_sensitivity(hybrid_mlem_ctx* c, float* sensitivity) {
  if (!c || !sensitivity)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "NULL argument");
  HY_TRY(hipSetDevice(c->device));
  HY_TRY(hipMemcpyAsync(sensitivity, c->sens,
                        sizeof(float) * c->grid.n_pixels,
                        hipMemcpyDeviceToHost, c->stream));
  HY_TRY(hipStreamSynchronize(c->stream));
  return STRIP_MLEM_OK;
}

int hybrid_mlem_get_weights(hybrid_mlem_ctx* c, float* weights) {
  if (!c || !weights)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "NULL argument");
  HY_TRY(hipSetDevice(c->device));
  if (c->n_infos)
    HY_TRY(hipMemcpyAsync(weights, c->w, sizeof(float) * c->n_infos,
                          hipMemcpyDeviceToHost, c->stream));
  HY_TRY(hipStreamSynchronize(c->stream));
  return STRIP_MLEM_OK;
}

int hybrid_mlem_set_algorithm(hybrid_mlem_ctx* c, int algorithm) {
  if (!c) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  if (algorithm < 0 || algorithm > 2)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "unknown algorithm");
  c->algorithm_request = algorithm;
  return STRIP_MLEM_OK;
}

int hybrid_mlem_get_algorithm(hybrid_mlem_ctx* c) {
  if (!c) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  return c->algorithm;
}

namespace {

// The algorithm the next iteration uses: the LOR-space one where its tables
// exist and every event's infos fit the per-event LDS factors, else one
// wavefront per event.
int choose_algorithm(hybrid_mlem_ctx* c) {
  if (c->algorithm_request == 1) {
    c->algorithm = 1;
    return STRIP_MLEM_OK;
  }
  if (!c->tables) {
    if (c->algorithm_request == 2)
      return strip::fail(STRIP_MLEM_ERR_UNSUPPORTED,
                         "LOR-space tables unavailable (memory or a LOR "
                         "with too many infos for the LDS slab)");
    c->algorithm = 1;
    return STRIP_MLEM_OK;
  }
  if (!c->plan_valid) HY_CHECK(build_plan(c));
  if (2 * c->max_event_infos > kWaveWords) {
    if (c->algorithm_request == 2)
      return strip::fail(STRIP_MLEM_ERR_UNSUPPORTED,
                         "an event spans more than " +
                             std::to_string(kWaveWords / 2) + " infos");
    c->algorithm = 1;
    return STRIP_MLEM_OK;
  }
  c->algorithm = 2;
  return STRIP_MLEM_OK;
}

int iterate_lor_space(hybrid_mlem_ctx* c, hipEvent_t mid) {
  const Grid& g = c->grid;
  const int64_t nv = c->n_voxels();
  const float kexp = -c->kernel.iz * 1.44269504088896340736f;
  hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(nv)), dim3(256), 0,
                     c->stream, c->extra, nv, 0.0f);
  if (c->ev.n && c->n_active) {
    HY_TRY(hipMemsetAsync(c->queue, 0, sizeof(int), c->stream));
    const dim3 grid((unsigned)std::max(1, std::min(c->n_cus, c->n_active))),
        block(64 * c->lor_waves);
    if (c->semantics == HYBRID_MLEM_CPU)
      hipLaunchKernelGGL(lor_iteration_kernel<true>, grid, block, c->lor_lds,
                         c->stream, g, c->kernel, kexp, c->seg_a, c->seg_b,
                         c->ranges, c->pix, c->ipix, c->upi, c->w, c->sens,
                         c->G, c->rho, c->pl_i, c->pl_f, c->pl_dup,
                         c->fac_off, c->lor_events, c->lor_order, c->n_active,
                         c->queue, c->inv, c->fac, c->bp, c->extra,
                         c->lor_slab_budget);
    else
      hipLaunchKernelGGL(lor_iteration_kernel<false>, grid, block,
                         c->lor_lds, c->stream, g, c->kernel, kexp, c->seg_a,
                         c->seg_b, c->ranges, c->pix, c->ipix, c->upi, c->w,
                         c->sens, c->G, c->rho, c->pl_i, c->pl_f, c->pl_dup,
                         c->fac_off, c->lor_events, c->lor_order, c->n_active,
                         c->queue, c->inv, c->fac, c->bp, c->extra,
                         c->lor_slab_budget);
  }
  if (mid) HY_TRY(hipEventRecord(mid, c->stream));
  const unsigned gblocks =
      (unsigned)std::min<int64_t>((g.n_pixels + 3) / 4, 65536);
  hipLaunchKernelGGL(lor_gather_kernel, dim3(gblocks), dim3(256), 0,
                     c->stream, c->csr_offsets, c->csr_order, c->G, c->bp,
                     c->rho, c->extra, g.n_pixels, g.n_planes, c->out);
  HY_TRY(hipGetLastError());
  return STRIP_MLEM_OK;
}

}  // namespace

int hybrid_mlem_iterate(hybrid_mlem_ctx* c, int n_iterations) {
  if (!c) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  if (n_iterations < 0)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "negative iteration count");
  if (!c->prepared)
    return strip::fail(STRIP_MLEM_ERR_STATE,
                       "hybrid_mlem_prepare_weights must come first");
  HY_TRY(hipSetDevice(c->device));
  HY_CHECK(choose_algorithm(c));
  const int64_t nv = c->n_voxels();
  const int64_t n = c->ev.n;
  const unsigned blocks =
      (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 3) / 4, 16384));
  for (auto* v : {&c->t_start, &c->t_mid, &c->t_end})
    while ((int)v->size() < n_iterations) {
      hipEvent_t e;
      HY_TRY(hipEventCreate(&e));
      v->push_back(e);
    }
  for (int it = 0; it < n_iterations; ++it) {
    if (c->algorithm == 1)
      hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(nv)), dim3(256), 0,
                         c->stream, c->out, nv, 0.0f);
    HY_TRY(hipEventRecord(c->t_start[it], c->stream));
    HY_TRY(hipEventRecord(c->t0, c->stream));
    if (c->algorithm == 2) {
      HY_CHECK(iterate_lor_space(c, c->t_mid[it]));
    } else if (n) {
      if (c->semantics == HYBRID_MLEM_CPU)
        hipLaunchKernelGGL(event_kernel<true>, dim3(blocks), dim3(256), 0,
                           c->stream, c->grid, c->kernel, c->seg_a, c->seg_b,
                           c->pix, c->w, c->ev.ev_i, c->ev.ev_f, c->ev.ev_dup,
                           n, c->rho, c->sens, c->out);
      else
        hipLaunchKernelGGL(event_kernel<false>, dim3(blocks), dim3(256), 0,
                           c->stream, c->grid, c->kernel, c->seg_a, c->seg_b,
                           c->pix, c->w, c->ev.ev_i, c->ev.ev_f, c->ev.ev_dup,
                           n, c->rho, c->sens, c->out);
    }
    HY_TRY(hipGetLastError());
    if (c->algorithm != 2) HY_TRY(hipEventRecord(c->t_mid[it], c->stream));
    HY_TRY(hipEventRecord(c->t1, c->stream));
    HY_TRY(hipEventRecord(c->t_end[it], c->stream));
    std::swap(c->rho, c->out);
  }
  HY_TRY(hipStreamSynchronize(c->stream));
  if (n_iterations > 0) HY_TRY(hipEventElapsedTime(&c->last_ms, c->t0, c->t1));
  c->timed_event_ms = c->timed_gather_ms = 0.0;
  for (int it = 0; it < n_iterations; ++it) {
    float a = 0.0f, b = 0.0f;
    HY_TRY(hipEventElapsedTime(&a, c->t_start[it], c->t_mid[it]));
    HY_TRY(hipEventElapsedTime(&b, c->t_mid[it], c->t_end[it]));
    c->timed_event_ms += a;
    c->timed_gather_ms += b;
  }
  c->timed_iterations = n_iterations;
  return STRIP_MLEM_OK;
}

int hybrid_mlem_get_timing(hybrid_mlem_ctx* c, double* event_kernels_ms,
                           double* gather_ms, int* n_iterations) {
  if (!c) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  if (event_kernels_ms) *event_kernels_ms = c->timed_event_ms;
  if (gather_ms) *gather_ms = c->timed_gather_ms;
  if (n_iterations) *n_iterations = c->timed_iterations;
  return STRIP_MLEM_OK;
}

int hybrid_mlem_last_iteration_ms(hybrid_mlem_ctx* c, float* ms) {
  if (!c || !ms) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "NULL argument");
  *ms = c->last_ms;
  return STRIP_MLEM_OK;
}

int hybrid_mlem_voxel_visits(hybrid_mlem_ctx* c, int64_t* visits) {
  if (!c || !visits)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "NULL argument");
  HY_TRY(hipSetDevice(c->device));
  unsigned long long* d = nullptr;
  HY_CHECK(device_alloc(&d, 1));
  unsigned long long h = 0;
  hipError_t e = hipMemsetAsync(d, 0, sizeof(unsigned long long), c->stream);
  if (e == hipSuccess && c->ev.n) {
    hipLaunchKernelGGL(visits_kernel, dim3(blocks_for(c->ev.n)), dim3(256), 0,
                       c->stream, c->ev.ev_i, c->ev.n, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess)
    return strip::fail(STRIP_MLEM_ERR_HIP, hipGetErrorString(e));
  *visits = (int64_t)h;
  return STRIP_MLEM_OK;
}

}  // extern "C"
