// Seeded event generation on the GPU (semantics and random stream described
// in include/strip_phantom.h).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "strip_device.hpp"
#include "strip_phantom.h"
#include "strip_phantom.hpp"

namespace strip {
namespace {

thread_local std::string g_phantom_error;

int phantom_fail(int code, const std::string& msg) {
  g_phantom_error = msg;
  return code;
}

#define PH_TRY(expr)                                                  \
  do {                                                                \
    hipError_t _e = (expr);                                           \
    if (_e != hipSuccess) {                                           \
      if (err) *err = std::string(#expr) + ": " + hipGetErrorString(_e); \
      return -1;                                                      \
    }                                                                 \
  } while (0)

// ---------------------------------------------------------------------------
// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): ten rounds of two 64 x 64 -> 128-bit products,
// key bumped by the Weyl constants between rounds.
// ---------------------------------------------------------------------------
struct Block4 {
  uint64_t v[4];
};

__host__ __device__ inline void mulhilo64(uint64_t a, uint64_t b, uint64_t& hi,
                                          uint64_t& lo) {
#ifdef __HIP_DEVICE_COMPILE__
  hi = __umul64hi(a, b);
  lo = a * b;
#else
  const unsigned __int128 p = (unsigned __int128)a * b;
  hi = (uint64_t)(p >> 64);
  lo = (uint64_t)p;
#endif
}

__host__ __device__ inline Block4 philox4x64(uint64_t c0, uint64_t c1,
                                             uint64_t c2, uint64_t c3,
                                             uint64_t k0, uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  const uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
  uint64_t x0 = c0, x1 = c1, x2 = c2, x3 = c3;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t hi0, lo0, hi1, lo1;
    mulhilo64(M0, x0, hi0, lo0);
    mulhilo64(M1, x2, hi1, lo1);
    const uint64_t y0 = hi1 ^ x1 ^ k0, y2 = hi0 ^ x3 ^ k1;
    x0 = y0;
    x1 = lo1;
    x2 = y2;
    x3 = lo0;
    k0 += W0;
    k1 += W1;
  }
  Block4 b;
  b.v[0] = x0;
  b.v[1] = x1;
  b.v[2] = x2;
  b.v[3] = x3;
  return b;
}

// uniform in [0, 1) from the top 24 bits
__device__ inline float unit(uint64_t w) {
  return (float)(w >> 40) * (1.0f / 16777216.0f);
}

// purposes of the counter's third word
constexpr uint64_t kEmission = 0;  // region, point, direction (per attempt)
constexpr uint64_t kSmear = 1;     // Gaussian errors

__device__ inline bool region_contains(const DevRegion& r, float px,
                                       float py) {
  const float dx = px - r.cx, dy = py - r.cy;
  if (r.kind == STRIP_REGION_ELLIPSE)  // ellipse.h:29-33
    return r.qa * dx * dx + 2.0f * r.qc * dx * dy + r.qb * dy * dy <= 1.0f;
  if (r.kind == STRIP_REGION_RECTANGLE)  // rectangle.h:22-25
    return fabsf(dx) <= r.a && fabsf(dy) <= r.b;
  return px == r.cx && py == r.cy;  // phantom.h:47
}

__global__ __launch_bounds__(256) void emit_kernel(
    PhantomKernelArgs a, int64_t first, int64_t count,
    uint8_t* __restrict__ flags, float* __restrict__ resp,
    float* __restrict__ exact, int* __restrict__ pix_origin,
    int* __restrict__ pix_err, float* __restrict__ ev4,
    int* __restrict__ pix_tan) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const uint64_t index = (uint64_t)(first + j);
  float px = 0.f, py = 0.f, phi_u = 0.f;
  // region and point; earlier regions win in a non-additive phantom
  for (uint64_t attempt = 0;; ++attempt) {
    const Block4 b = philox4x64(index, attempt, kEmission, 0, a.seed, 0);
    const float u = unit(b.v[0]);
    int k = 0;
    while (k < a.n_regions - 1 && u > a.regions[k].cdf) ++k;
    const DevRegion& r = a.regions[k];
    if (r.kind == STRIP_REGION_ELLIPSE) {
      const float ang = 6.28318530717958647692f * unit(b.v[1]);
      const float rad = sqrtf(unit(b.v[2]));
      const float x = r.a * rad * cosf(ang), y = r.b * rad * sinf(ang);
      px = r.cos_t * x - r.sin_t * y + r.cx;
      py = r.sin_t * x + r.cos_t * y + r.cy;
    } else if (r.kind == STRIP_REGION_RECTANGLE) {
      px = r.cx + (2.0f * unit(b.v[1]) - 1.0f) * r.a;
      py = r.cy + (2.0f * unit(b.v[2]) - 1.0f) * r.b;
    } else {
      px = r.cx;
      py = r.cy;
    }
    phi_u = unit(b.v[3]);
    bool taken = false;
    if (!a.additive)
      for (int i = 0; i < k && !taken; ++i)
        taken = region_contains(a.regions[i], px, py);
    // every thread ends: an attempt cap far beyond any phantom's need
    if (!taken || attempt >= 1u << 20) break;
  }
  const float phi = 1.57079632679489661923f * (2.0f * phi_u - 1.0f);
  const float sin_phi = sinf(phi), cos_phi = cosf(phi);
  const float t = sin_phi / cos_phi;
  if (ev4) {
    // Event(origin, phi): direction (sin phi, cos phi) (event.h:36-37)
    ev4[4 * j] = px;
    ev4[4 * j + 1] = py;
    ev4[4 * j + 2] = sin_phi;
    ev4[4 * j + 3] = cos_phi;
  }
  // to_projection_space_tan (scanner.h:107-113)
  const float z_u = px + (a.radius - py) * t;
  const float z_d = px - (a.radius + py) * t;
  const float dl = -2.0f * py * sqrtf(t * t + 1.0f);
  const bool detected =
      fabsf(z_u) < a.half_length && fabsf(z_d) < a.half_length;
  flags[j] = detected ? 1 : 0;
  if (pix_origin) {
    const int ix = (int)floorf((px - a.img_z0) / a.img_pixel);
    const int iy = (int)floorf((py - a.img_y0) / a.img_pixel);
    pix_origin[j] = (ix >= 0 && ix < a.img_nz && iy >= 0 && iy < a.img_ny)
                        ? iy * a.img_nz + ix
                        : -1;
  }
  // Gaussian errors: Box-Muller on the smear block
  const Block4 s = philox4x64(index, 0, kSmear, 0, a.seed, 0);
  const float r0 = sqrtf(-2.0f * logf(1.0f - unit(s.v[0])));
  const float r1 = sqrtf(-2.0f * logf(1.0f - unit(s.v[2])));
  const float th0 = 6.28318530717958647692f * unit(s.v[1]);
  const float th1 = 6.28318530717958647692f * unit(s.v[3]);
  const float e_u = a.sigma_z * r0 * cosf(th0);
  const float e_d = a.sigma_z * r0 * sinf(th0);
  const float e_l = a.sigma_dl * r1 * cosf(th1);
  const float w_u = z_u + e_u, w_d = z_d + e_d, w_l = dl + e_l;
  resp[3 * j] = w_u;
  resp[3 * j + 1] = w_d;
  resp[3 * j + 2] = w_l;
  if (exact) {
    exact[3 * j] = z_u;
    exact[3 * j + 1] = z_d;
    exact[3 * j + 2] = dl;
  }
  if (pix_err) {
    // the point reconstructed from the response with error
    // (calculate_tan_y_z, response.h:22-37)
    const float tt = (w_u - w_d) / (2.0f * a.radius);
    const float yy = -0.5f * (w_l / sqrtf(1.0f + tt * tt));
    const float zz = 0.5f * (w_u + w_d + 2.0f * yy * tt);
    const int ix = (int)floorf((zz - a.img_z0) / a.img_pixel);
    const int iy = (int)floorf((yy - a.img_y0) / a.img_pixel);
    const bool in_image =
        detected && ix >= 0 && ix < a.img_nz && iy >= 0 && iy < a.img_ny;
    pix_err[j] = in_image ? iy * a.img_nz + ix : -1;
    if (pix_tan) {
      // its tangent bin (VariableVoxelSizeVoxelGrid::voxel_at,
      // voxel_grid.h:109-118; VoxelMap index (plane n_y + row) n_z + column)
      const int it = (int)floorf((tt - a.tan_left) / a.tan_voxel);
      pix_tan[j] = (in_image && it >= 0 && it < a.tan_bins)
                       ? (it * a.img_ny + iy) * a.img_nz + ix
                       : -1;
    }
  }
}

// Emissions j < used: count emitted points; detected ones also in the two
// detected images.
__global__ __launch_bounds__(256) void image_kernel(
    int64_t used, const uint8_t* __restrict__ flags,
    const int* __restrict__ pix_origin, const int* __restrict__ pix_err,
    const int* __restrict__ pix_tan, int* __restrict__ img_em,
    int* __restrict__ img_det, int* __restrict__ img_err,
    int* __restrict__ img_tan) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= used) return;
  const int p = pix_origin[j];
  if (p >= 0 && img_em) atomicAdd(&img_em[p], 1);
  if (!flags[j]) return;
  if (p >= 0 && img_det) atomicAdd(&img_det[p], 1);
  const int q = pix_err[j];
  if (q >= 0 && img_err) atomicAdd(&img_err[q], 1);
  if (img_tan) {
    const int v = pix_tan[j];
    if (v >= 0) atomicAdd(&img_tan[v], 1);
  }
}

// The --tan-bins kernel map (phantom_cmd.cpp:405-431): for every voxel
// (column x, row y, tangent plane t) of the tangent-binned grid, the analytic
// kernel of an event at the voxel centre with the plane's tangent, evaluated
// at the kernel point, times the bin's phase-space volume
// 4 R pixel^2 dtan sqrt(1 + tan^2), times inv_sens.
__global__ __launch_bounds__(256) void tan_kernel_map_kernel(
    DeviceGeometry g, int n_z, int n_y, int n_t, float pixel, float z0,
    float y0, float tan_left, float tan_voxel, float kz, float ky,
    float inv_sens, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)n_z * n_y * n_t;
  if (i >= n) return;
  const int x = (int)(i % n_z);
  const int y = (int)((i / n_z) % n_y);
  const int t = (int)(i / ((int64_t)n_z * n_y));
  const float tan = ((float)t + 0.5f) * tan_voxel + tan_left;
  const float sec = sqrtf(1.0f + tan * tan);
  // PixelGrid::center_at (pixel_grid.h:54-57)
  const float half = pixel / 2.0f;
  const float cz = (z0 + half) + (float)x * pixel;
  const float cy = (y0 + half) + (float)y * pixel;
  const float k = analytic_kernel(g, cy, tan, sec, kz - cz, ky - cy);
  const float bin_volume = pixel * pixel * tan_voxel;
  const float volume = 4.0f * g.radius * bin_volume * sqrtf(1.0f + tan * tan);
  out[i] = k * volume * inv_sens;
}

// Scanner::sensitivity at one point (for the kernel map's normalisation).
__global__ void sensitivity_at_kernel(DeviceGeometry g, float z, float y,
                                      float* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *out = scanner_sensitivity(g, z, y);
}

__global__ __launch_bounds__(256) void iota_kernel(int64_t first,
                                                   int64_t count,
                                                   int64_t* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < count) out[j] = first + j;
}

__global__ void random_blocks_kernel(uint64_t k0, uint64_t k1, uint64_t first,
                                     uint64_t c1, uint64_t c2, int n,
                                     uint64_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Block4 b = philox4x64(first + (uint64_t)i, c1, c2, 0, k0, k1);
  for (int w = 0; w < 4; ++w) out[4 * i + w] = b.v[w];
}

struct Triple {
  float v[3];
};

struct Quad {
  float v[4];
};

}  // namespace

int phantom_prepare(PhantomDevice& P, const strip_phantom_params& p,
                    const strip_phantom_region* regions, int n_regions,
                    double radius, double length, double sigma_z,
                    double sigma_dl, std::string* err) {
  if (n_regions < 1 || !regions) {
    if (err) *err = "must specify at least one region";
    return -1;
  }
  std::vector<DevRegion> host(n_regions);
  std::vector<double> weight(n_regions);
  double total = 0;
  for (int i = 0; i < n_regions; ++i) {
    const strip_phantom_region& r = regions[i];
    DevRegion& d = host[i];
    d.kind = r.kind;
    d.cx = (float)r.x;
    d.cy = (float)r.y;
    d.a = (float)r.a;
    d.b = (float)r.b;
    const double t = r.angle_deg * M_PI / 180.0;
    d.sin_t = (float)std::sin(t);
    d.cos_t = (float)std::cos(t);
    // ellipse.h:29-33 quadratic form of the rotated ellipse
    const double ia2 = 1.0 / (r.a * r.a), ib2 = 1.0 / (r.b * r.b);
    const double s = std::sin(t), c = std::cos(t);
    d.qa = (float)(c * c * ia2 + s * s * ib2);
    d.qb = (float)(s * s * ia2 + c * c * ib2);
    d.qc = (float)(s * c * (ia2 - ib2));
    if (r.kind == STRIP_REGION_ELLIPSE)
      weight[i] = r.intensity * M_PI * r.a * r.b;
    else if (r.kind == STRIP_REGION_RECTANGLE)
      weight[i] = r.intensity * 4.0 * r.a * r.b;
    else if (r.kind == STRIP_REGION_POINT)
      weight[i] = r.intensity;
    else {
      if (err) *err = "unknown region kind";
      return -1;
    }
    if (r.kind != STRIP_REGION_POINT && !(r.a > 0 && r.b > 0)) {
      if (err) *err = "region sizes must be > 0";
      return -1;
    }
    total += weight[i];
  }
  if (!(total > 0)) {
    if (err) *err = "phantom weights must add up to > 0";
    return -1;
  }
  double acc = 0;
  for (int i = 0; i < n_regions; ++i) {
    acc += weight[i];
    host[i].cdf = (float)(acc / total);
  }
  host[n_regions - 1].cdf = 1.0f;
  PH_TRY(hipMalloc(&P.regions, sizeof(DevRegion) * n_regions));
  PH_TRY(hipMemcpy(P.regions, host.data(), sizeof(DevRegion) * n_regions,
                   hipMemcpyHostToDevice));
  P.args.regions = P.regions;
  P.args.n_regions = n_regions;
  P.args.additive = p.additive ? 1 : 0;
  P.args.seed = p.seed;
  P.args.radius = (float)radius;
  P.args.half_length = (float)(length / 2);
  P.args.sigma_z = (float)sigma_z;
  P.args.sigma_dl = (float)sigma_dl;
  P.args.img_nz = p.n_z;
  P.args.img_ny = p.n_y;
  P.args.img_pixel = p.pixel > 0 ? (float)p.pixel : 1.0f;
  P.args.img_z0 = (float)(-p.pixel * p.n_z / 2 + p.image_origin_z);
  P.args.img_y0 = (float)(-p.pixel * p.n_y / 2 + p.image_origin_y);
  P.args.tan_bins = 0;
  P.args.tan_left = 0.0f;
  P.args.tan_voxel = 1.0f;
  return 0;
}

// Tangent axis of the --tan-bins grid (phantom_cmd.cpp:246-250), in FP32:
// max_tan = 1.1 L / (2 R), planes of max_tan / bins * 2 from -max_tan.
void phantom_tan_axis(double radius, double length, int tan_bins,
                      float* left, float* voxel) {
  const float max_tan =
      (float)1.1 * (float)length / (2.0f * (float)radius);
  *left = -max_tan;
  *voxel = max_tan / (float)tan_bins * 2.0f;
}

void phantom_release(PhantomDevice& P) {
  if (P.regions) (void)hipFree(P.regions);
  for (void* q : {(void*)P.flags, (void*)P.resp, (void*)P.exact,
                  (void*)P.pix_origin, (void*)P.pix_err, (void*)P.pix_tan,
                  (void*)P.ev4, (void*)P.sel_ev4, (void*)P.sel_resp,
                  (void*)P.sel_exact, (void*)P.sel_index, (void*)P.index,
                  (void*)P.count, P.tmp})
    if (q) (void)hipFree(q);
  P = PhantomDevice{};
}

int phantom_detect(PhantomDevice& P, int64_t need, int64_t max_emissions,
                   float* d_resp, float* d_exact, int* img_em, int* img_det,
                   int* img_err, int64_t* produced, hipStream_t stream,
                   std::string* err, float* d_events, int* img_tan) {
  if (img_tan && P.args.tan_bins < 1) {
    if (err) *err = "tangent-bin image without tangent bins";
    return -1;
  }
  const bool images = img_em || img_det || img_err || img_tan;
  if (!P.flags) {
    const int64_t B = kPhantomBatch;
    PH_TRY(hipMalloc(&P.flags, B));
    PH_TRY(hipMalloc(&P.resp, sizeof(float) * 3 * B));
    PH_TRY(hipMalloc(&P.exact, sizeof(float) * 3 * B));
    PH_TRY(hipMalloc(&P.pix_origin, sizeof(int) * B));
    PH_TRY(hipMalloc(&P.pix_err, sizeof(int) * B));
    PH_TRY(hipMalloc(&P.sel_resp, sizeof(float) * 3 * B));
    PH_TRY(hipMalloc(&P.sel_exact, sizeof(float) * 3 * B));
    PH_TRY(hipMalloc(&P.sel_index, sizeof(int64_t) * B));
    PH_TRY(hipMalloc(&P.index, sizeof(int64_t) * B));
    PH_TRY(hipMalloc(&P.count, sizeof(int64_t)));
    size_t b1 = 0, b2 = 0;
    PH_TRY(hipcub::DeviceSelect::Flagged(
        nullptr, b1, reinterpret_cast<Triple*>(P.resp), P.flags,
        reinterpret_cast<Triple*>(P.sel_resp), P.count, (int)B, stream));
    PH_TRY(hipcub::DeviceSelect::Flagged(nullptr, b2, P.index, P.flags,
                                         P.sel_index, P.count, (int)B,
                                         stream));
    size_t b3 = 0;
    PH_TRY(hipcub::DeviceSelect::Flagged(
        nullptr, b3, reinterpret_cast<Quad*>(P.resp), P.flags,
        reinterpret_cast<Quad*>(P.sel_resp), P.count, (int)B, stream));
    P.tmp_bytes = std::max(std::max(b1, b2), b3);
    PH_TRY(hipMalloc(&P.tmp, P.tmp_bytes));
  }
  if (img_tan && !P.pix_tan)
    PH_TRY(hipMalloc(&P.pix_tan, sizeof(int) * kPhantomBatch));
  if (d_events && !P.ev4) {
    PH_TRY(hipMalloc(&P.ev4, sizeof(float) * 4 * kPhantomBatch));
    PH_TRY(hipMalloc(&P.sel_ev4, sizeof(float) * 4 * kPhantomBatch));
  }
  int64_t done = 0;
  int64_t emitted = 0;
  while (done < need && emitted < max_emissions) {
    const int64_t count =
        std::min<int64_t>(kPhantomBatch, max_emissions - emitted);
    const int64_t first = P.next;
    const unsigned blocks = (unsigned)((count + 255) / 256);
    hipLaunchKernelGGL(emit_kernel, dim3(blocks), dim3(256), 0, stream,
                       P.args, first, count, P.flags, P.resp,
                       d_exact ? P.exact : nullptr,
                       images ? P.pix_origin : nullptr,
                       images ? P.pix_err : nullptr,
                       d_events ? P.ev4 : nullptr,
                       img_tan ? P.pix_tan : nullptr);
    PH_TRY(hipGetLastError());
    // emission numbers of the batch, compacted with the responses
    hipLaunchKernelGGL(iota_kernel, dim3(blocks), dim3(256), 0, stream,
                       first, count, P.index);
    PH_TRY(hipGetLastError());
    size_t bytes = P.tmp_bytes;
    PH_TRY(hipcub::DeviceSelect::Flagged(
        P.tmp, bytes, reinterpret_cast<Triple*>(P.resp), P.flags,
        reinterpret_cast<Triple*>(P.sel_resp), P.count, (int)count, stream));
    int64_t n_sel = 0;
    PH_TRY(hipMemcpyAsync(&n_sel, P.count, sizeof(int64_t),
                          hipMemcpyDeviceToHost, stream));
    PH_TRY(hipStreamSynchronize(stream));
    const int64_t take = std::min(n_sel, need - done);
    // emissions used: all of the batch, or up to the last taken one
    int64_t used = count;
    if (take < n_sel) {
      bytes = P.tmp_bytes;
      PH_TRY(hipcub::DeviceSelect::Flagged(P.tmp, bytes, P.index, P.flags,
                                           P.sel_index, P.count, (int)count,
                                           stream));
      int64_t last = 0;
      PH_TRY(hipMemcpyAsync(&last, P.sel_index + (take - 1), sizeof(int64_t),
                            hipMemcpyDeviceToHost, stream));
      PH_TRY(hipStreamSynchronize(stream));
      used = last - first + 1;
    }
    if (take > 0) {
      PH_TRY(hipMemcpyAsync(d_resp + 3 * done, P.sel_resp,
                            sizeof(float) * 3 * take, hipMemcpyDeviceToDevice,
                            stream));
      if (d_exact) {
        bytes = P.tmp_bytes;
        PH_TRY(hipcub::DeviceSelect::Flagged(
            P.tmp, bytes, reinterpret_cast<Triple*>(P.exact), P.flags,
            reinterpret_cast<Triple*>(P.sel_exact), P.count, (int)count,
            stream));
        PH_TRY(hipMemcpyAsync(d_exact + 3 * done, P.sel_exact,
                              sizeof(float) * 3 * take,
                              hipMemcpyDeviceToDevice, stream));
      }
      if (d_events) {
        bytes = P.tmp_bytes;
        PH_TRY(hipcub::DeviceSelect::Flagged(
            P.tmp, bytes, reinterpret_cast<Quad*>(P.ev4), P.flags,
            reinterpret_cast<Quad*>(P.sel_ev4), P.count, (int)count,
            stream));
        PH_TRY(hipMemcpyAsync(d_events + 4 * done, P.sel_ev4,
                              sizeof(float) * 4 * take,
                              hipMemcpyDeviceToDevice, stream));
      }
    }
    if (images) {
      hipLaunchKernelGGL(image_kernel, dim3((unsigned)((used + 255) / 256)),
                         dim3(256), 0, stream, used, P.flags, P.pix_origin,
                         P.pix_err, P.pix_tan, img_em, img_det, img_err,
                         img_tan);
      PH_TRY(hipGetLastError());
    }
    PH_TRY(hipStreamSynchronize(stream));
    done += take;
    emitted += used;
    P.next += used;
    P.emitted += used;
  }
  if (produced) *produced = done;
  return 0;
}

}  // namespace strip

extern "C" {

const char* strip_phantom_last_error(void) {
  return strip::g_phantom_error.c_str();
}

int strip_phantom_random_blocks(int device, uint64_t key0, uint64_t key1,
                                uint64_t first, uint64_t c1, uint64_t c2,
                                int n_blocks, uint64_t* out) {
  using namespace strip;
  g_phantom_error.clear();
  if (n_blocks < 0 || (n_blocks > 0 && !out))
    return phantom_fail(STRIP_MLEM_ERR_ARGUMENT, "bad output");
  if (n_blocks == 0) return STRIP_MLEM_OK;
  uint64_t* d = nullptr;
  if (hipSetDevice(device) != hipSuccess ||
      hipMalloc(&d, sizeof(uint64_t) * 4 * n_blocks) != hipSuccess)
    return phantom_fail(STRIP_MLEM_ERR_HIP, "device allocation failed");
  hipLaunchKernelGGL(random_blocks_kernel, dim3((n_blocks + 255) / 256),
                     dim3(256), 0, 0, key0, key1, first, c1, c2, n_blocks, d);
  hipError_t e = hipMemcpy(out, d, sizeof(uint64_t) * 4 * n_blocks,
                           hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return phantom_fail(STRIP_MLEM_ERR_HIP, hipGetErrorString(e));
  return STRIP_MLEM_OK;
}

int strip_phantom_generate_full(int device, const strip_phantom_params* params,
                                const strip_phantom_region* regions,
                                int n_regions, int64_t n,
                                strip_phantom_outputs* out) {
  using namespace strip;
  g_phantom_error.clear();
  if (!params || !out || n < 0 || (n > 0 && !out->responses))
    return phantom_fail(STRIP_MLEM_ERR_ARGUMENT, "bad arguments");
  int* image_tan = out->tan_bins > 0 ? out->image_tan_bins : nullptr;
  if (out->tan_bins < 0 || (out->tan_bins > 0 && !out->image_tan_bins))
    return phantom_fail(STRIP_MLEM_ERR_ARGUMENT,
                        "tan_bins > 0 needs image_tan_bins");
  const bool images = out->image_emitted || out->image_detected ||
                      out->image_detected_w_error || image_tan;
  if (images && (params->n_z <= 0 || params->n_y <= 0 || !(params->pixel > 0)))
    return phantom_fail(STRIP_MLEM_ERR_ARGUMENT,
                        "images need n_z, n_y and pixel > 0");
  if (hipSetDevice(device) != hipSuccess)
    return phantom_fail(STRIP_MLEM_ERR_HIP, "hipSetDevice failed");
  PhantomDevice P;
  std::string err;
  hipStream_t stream = nullptr;
  float* d_resp = nullptr;
  float* d_exact = nullptr;
  float* d_events = nullptr;
  int* d_img = nullptr;
  const int64_t np = images ? (int64_t)params->n_z * params->n_y : 0;
  const int64_t nt = image_tan ? np * out->tan_bins : 0;
  const int64_t n1 = std::max<int64_t>(n, 1);
  int rc = STRIP_MLEM_OK;
  int64_t produced = 0;
  do {
    if (phantom_prepare(P, *params, regions, n_regions, params->radius,
                        params->scintillator_length, params->sigma_z,
                        params->sigma_dl, &err)) {
      rc = STRIP_MLEM_ERR_ARGUMENT;
      break;
    }
    if (image_tan) {
      P.args.tan_bins = out->tan_bins;
      phantom_tan_axis(params->radius, params->scintillator_length,
                       out->tan_bins, &P.args.tan_left, &P.args.tan_voxel);
    }
    if (hipStreamCreate(&stream) != hipSuccess ||
        hipMalloc(&d_resp, sizeof(float) * 3 * n1) != hipSuccess ||
        (out->exact &&
         hipMalloc(&d_exact, sizeof(float) * 3 * n1) != hipSuccess) ||
        (out->events &&
         hipMalloc(&d_events, sizeof(float) * 4 * n1) != hipSuccess) ||
        (images &&
         (hipMalloc(&d_img, sizeof(int) * (3 * np + nt)) != hipSuccess ||
          hipMemsetAsync(d_img, 0, sizeof(int) * (3 * np + nt), stream) !=
              hipSuccess))) {
      err = "device allocation failed";
      rc = STRIP_MLEM_ERR_ALLOC;
      break;
    }
    const int64_t max_em = params->only_detected ? INT64_MAX : n;
    if (phantom_detect(P, n, max_em, d_resp, d_exact,
                       out->image_emitted ? d_img : nullptr,
                       out->image_detected ? d_img + np : nullptr,
                       out->image_detected_w_error ? d_img + 2 * np : nullptr,
                       &produced, stream, &err, d_events,
                       image_tan ? d_img + 3 * np : nullptr)) {
      rc = STRIP_MLEM_ERR_HIP;
      break;
    }
    if (produced > 0 &&
        (hipMemcpy(out->responses, d_resp, sizeof(float) * 3 * produced,
                   hipMemcpyDeviceToHost) != hipSuccess ||
         (out->exact &&
          hipMemcpy(out->exact, d_exact, sizeof(float) * 3 * produced,
                    hipMemcpyDeviceToHost) != hipSuccess) ||
         (out->events &&
          hipMemcpy(out->events, d_events, sizeof(float) * 4 * produced,
                    hipMemcpyDeviceToHost) != hipSuccess))) {
      err = "copy to host failed";
      rc = STRIP_MLEM_ERR_HIP;
      break;
    }
    int* outs[3] = {out->image_emitted, out->image_detected,
                    out->image_detected_w_error};
    for (int k = 0; k < 3 && images; ++k)
      if (outs[k] && hipMemcpy(outs[k], d_img + k * np, sizeof(int) * np,
                               hipMemcpyDeviceToHost) != hipSuccess) {
        err = "copy to host failed";
        rc = STRIP_MLEM_ERR_HIP;
      }
    if (image_tan && hipMemcpy(image_tan, d_img + 3 * np, sizeof(int) * nt,
                               hipMemcpyDeviceToHost) != hipSuccess) {
      err = "copy to host failed";
      rc = STRIP_MLEM_ERR_HIP;
    }
  } while (false);
  out->n_detected = produced;
  out->n_emitted = P.emitted;
  phantom_release(P);
  if (d_resp) (void)hipFree(d_resp);
  if (d_exact) (void)hipFree(d_exact);
  if (d_events) (void)hipFree(d_events);
  if (d_img) (void)hipFree(d_img);
  if (stream) (void)hipStreamDestroy(stream);
  if (rc) return phantom_fail(rc, err);
  return STRIP_MLEM_OK;
}

int strip_phantom_generate(int device, const strip_phantom_params* params,
                           const strip_phantom_region* regions, int n_regions,
                           int64_t n, float* responses, float* exact,
                           int64_t* n_detected, int64_t* n_emitted,
                           int* image_emitted, int* image_detected,
                           int* image_detected_w_error) {
  strip_phantom_outputs out{};
  out.responses = responses;
  out.exact = exact;
  out.image_emitted = image_emitted;
  out.image_detected = image_detected;
  out.image_detected_w_error = image_detected_w_error;
  const int rc = strip_phantom_generate_full(device, params, regions,
                                             n_regions, n, &out);
  if (n_detected) *n_detected = out.n_detected;
  if (n_emitted) *n_emitted = out.n_emitted;
  return rc;
}

int strip_phantom_tan_kernel_map(int device, const strip_phantom_params* params,
                                 int tan_bins, double kernel_z,
                                 double kernel_y, float* out) {
  using namespace strip;
  g_phantom_error.clear();
  if (!params || !out || tan_bins < 1 || params->n_z <= 0 ||
      params->n_y <= 0 || !(params->pixel > 0))
    return phantom_fail(STRIP_MLEM_ERR_ARGUMENT, "bad arguments");
  if (hipSetDevice(device) != hipSuccess)
    return phantom_fail(STRIP_MLEM_ERR_HIP, "hipSetDevice failed");
  strip_mlem_geometry geo{};
  geo.radius = params->radius;
  geo.scintillator_length = params->scintillator_length;
  geo.n_z_pixels = params->n_z;
  geo.n_y_pixels = params->n_y;
  geo.pixel_width = params->pixel;
  geo.pixel_height = params->pixel;
  geo.sigma_z = params->sigma_z;
  geo.sigma_dl = params->sigma_dl;
  const DeviceGeometry g = make_device_geometry(geo);
  float left = 0.f, voxel = 1.f;
  phantom_tan_axis(params->radius, params->scintillator_length, tan_bins,
                   &left, &voxel);
  const int64_t n = (int64_t)params->n_z * params->n_y * tan_bins;
  float* d = nullptr;
  if (hipMalloc(&d, sizeof(float) * (n + 1)) != hipSuccess)
    return phantom_fail(STRIP_MLEM_ERR_ALLOC, "device allocation failed");
  // only detected emissions: normalised by the kernel point's sensitivity
  float inv_sens = 1.0f;
  hipError_t e = hipSuccess;
  if (params->only_detected) {
    hipLaunchKernelGGL(sensitivity_at_kernel, dim3(1), dim3(64), 0, 0, g,
                       (float)kernel_z, (float)kernel_y, d + n);
    float s = 0.f;
    e = hipMemcpy(&s, d + n, sizeof(float), hipMemcpyDeviceToHost);
    inv_sens = 1.0f / s;
  }
  if (e == hipSuccess) {
    const float pixel = (float)params->pixel;
    const float z0 = (float)(-params->pixel * params->n_z / 2 +
                             params->image_origin_z);
    const float y0 = (float)(-params->pixel * params->n_y / 2 +
                             params->image_origin_y);
    hipLaunchKernelGGL(tan_kernel_map_kernel,
                       dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, g,
                       params->n_z, params->n_y, tan_bins, pixel, z0, y0,
                       left, voxel, (float)kernel_z, (float)kernel_y, inv_sens,
                       d);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpy(out, d, sizeof(float) * n, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess)
    return phantom_fail(STRIP_MLEM_ERR_HIP, hipGetErrorString(e));
  return STRIP_MLEM_OK;
}

}  // extern "C"
