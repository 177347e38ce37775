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
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "hybrid_mlem.h"
#include "strip_device.hpp"

namespace strip {
int fail(int code, const std::string& message);
}

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

  ~hybrid_mlem_ctx() {
    if (stream) (void)hipStreamSynchronize(stream);
    void* ptrs[] = {segments, widths, seg_a, seg_b, ranges, pix, pos,
                    w_given, w, sens, rho, out, stage, ev.ev_i, ev.ev_f,
                    ev.ev_dup};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
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

// calculate_sensitivity on the device: the infos inside the grid sorted
// stably by pixel (radix sort of (pixel, info) pairs), per-pixel offsets
// from a count and a scan, then one thread per pixel sums its weights in
// info order, the order of the reference's loop.
int sensitivity_from_weights(hybrid_mlem_ctx* c) {
  const Grid& g = c->grid;
  const int n = c->n_infos;
  if (n == 0) {
    hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(g.n_pixels)), dim3(256),
                       0, c->stream, c->sens, (int64_t)g.n_pixels, 0.0f);
    HY_TRY(hipGetLastError());
    return STRIP_MLEM_OK;
  }
  int *count = nullptr, *offsets = nullptr;
  unsigned *keys = nullptr, *keys_sorted = nullptr;
  int *order = nullptr, *iota = nullptr;
  char* tmp = nullptr;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(c->stream);
    void* ps[] = {count, offsets, keys, keys_sorted, order, iota, tmp};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  };
  int rc = STRIP_MLEM_OK;
  size_t sort_bytes = 0, scan_bytes = 0;
  hipError_t e = hipSuccess;
  std::vector<int> seq(n);
  for (int i = 0; i < n; ++i) seq[i] = i;
  if ((rc = device_alloc(&count, (size_t)g.n_pixels + 1)) ||
      (rc = device_alloc(&offsets, (size_t)g.n_pixels + 1)) ||
      (rc = device_alloc(&keys, n)) || (rc = device_alloc(&keys_sorted, n)) ||
      (rc = device_alloc(&order, n)) || (rc = device_alloc(&iota, n))) {
    cleanup();
    return rc;
  }
  e = hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys,
                                         keys_sorted, iota, order, n, 0, 32,
                                         c->stream);
  if (e == hipSuccess)
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, count, offsets,
                                         g.n_pixels + 1, c->stream);
  if (e == hipSuccess &&
      (rc = device_alloc(&tmp, std::max(sort_bytes, scan_bytes)))) {
    cleanup();
    return rc;
  }
  if (e == hipSuccess)
    e = hipMemsetAsync(count, 0, sizeof(int) * ((size_t)g.n_pixels + 1),
                       c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(iota, seq.data(), sizeof(int) * n,
                       hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(pixel_count_kernel, dim3(blocks_for(n)), dim3(256), 0,
                       c->stream, g, c->pix, (int64_t)n, count, keys);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, keys, keys_sorted,
                                           iota, order, n, 0, 32, c->stream);
  if (e == hipSuccess)
    e = hipcub::DeviceScan::ExclusiveSum(tmp, scan_bytes, count, offsets,
                                         g.n_pixels + 1, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(sensitivity_sum_kernel, dim3(blocks_for(g.n_pixels)),
                       dim3(256), 0, c->stream, offsets, order, c->w,
                       g.n_pixels, c->sens);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  cleanup();
  if (e != hipSuccess)
    return strip::fail(STRIP_MLEM_ERR_HIP,
                       std::string("sensitivity: ") + hipGetErrorString(e));
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
  if (sensitivity_mode < 0 || sensitivity_mode > 2 ||
      (sensitivity_mode == 2 && !sensitivity))
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
  if (sensitivity_mode == 2) {
    HY_TRY(hipMemcpyAsync(c->sens, sensitivity, sizeof(float) * g.n_pixels,
                          hipMemcpyHostToDevice, c->stream));
  } else if (sensitivity_mode == 1) {
    hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(g.n_pixels)), dim3(256),
                       0, c->stream, c->sens, (int64_t)g.n_pixels, 1.0f);
  } else {
    // stable sort of the infos by pixel; each pixel then sums its weights
    // in info order
    HY_CHECK(sensitivity_from_weights(c));
  }
  if (n)
    hipLaunchKernelGGL(normalize_kernel, dim3(blocks_for(n)), dim3(256), 0,
                       c->stream, g, c->pix, c->sens, (int64_t)n, c->w);
  HY_TRY(hipGetLastError());
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
  return STRIP_MLEM_OK;
}

int hybrid_mlem_clear_events(hybrid_mlem_ctx* c) {
  if (!c) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  c->ev.n = 0;
  return STRIP_MLEM_OK;
}

int64_t hybrid_mlem_n_events(const hybrid_mlem_ctx* c) {
  return c ? c->ev.n : -1;
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

int hybrid_mlem_iterate(hybrid_mlem_ctx* c, int n_iterations) {
  if (!c) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  if (n_iterations < 0)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "negative iteration count");
  if (!c->prepared)
    return strip::fail(STRIP_MLEM_ERR_STATE,
                       "hybrid_mlem_prepare_weights must come first");
  HY_TRY(hipSetDevice(c->device));
  const int64_t nv = c->n_voxels();
  const int64_t n = c->ev.n;
  const unsigned blocks =
      (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 3) / 4, 16384));
  for (int it = 0; it < n_iterations; ++it) {
    hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(nv)), dim3(256), 0,
                       c->stream, c->out, nv, 0.0f);
    HY_TRY(hipEventRecord(c->t0, c->stream));
    if (n) {
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
    HY_TRY(hipEventRecord(c->t1, c->stream));
    std::swap(c->rho, c->out);
  }
  HY_TRY(hipStreamSynchronize(c->stream));
  if (n_iterations > 0) HY_TRY(hipEventElapsedTime(&c->last_ms, c->t0, c->t1));
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
