// MI355X (gfx950) J-PET 2D strip list-mode TOF MLEM: kernels + C ABI.
//
// One MLEM iteration (src/2d/strip/reconstruction.h:117-152) is
//   D_e   = sum_{p in E_e} K_e(p) rho(p)                      (forward)
//   bp(p) = sum_{e : p in E_e} K_e(p) * inv_e,  inv_e = D_e>0 ? 1/D_e : 0
//   rho'(p) = rho(p) * bp(p)                                  (finalize)
// which equals the reference's  rho'(p) = sum_e (K_e(p) rho(p)) inv_e
// (reconstruction.h:277-296) with rho(p) factored out of the event sum.
// E_e is the set of pixels of the event's clamped bounding box whose centre
// passes the 3-sigma ellipse test (reconstruction.h:255-264).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "strip_band.hpp"
#include "strip_device.hpp"
#include "strip_phantom.hpp"
#include "strip_mlem.h"

using namespace strip;

namespace strip {

thread_local std::string g_last_error;

int fail(int code, const std::string& message) {
  g_last_error = message;
  return code;
}

}  // namespace strip

#define HIP_TRY(expr)                                                     \
  do {                                                                    \
    hipError_t _err = (expr);                                             \
    if (_err != hipSuccess)                                               \
      return strip::fail(STRIP_MLEM_ERR_HIP, std::string(#expr) + ": " +  \
                                                 hipGetErrorString(_err)); \
  } while (0)

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Per-event precompute from AoS (z_u, z_d, dl) triples.
__global__ __launch_bounds__(256) void setup_events_kernel(
    DeviceGeometry g, const float* __restrict__ responses, int64_t n,
    float4* __restrict__ ev_a, float4* __restrict__ ev_b,
    float4* __restrict__ ev_c, uint2* __restrict__ ev_box) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 a, b, c;
  uint2 box;
  setup_event_exact(g, responses[3 * i], responses[3 * i + 1],
                    responses[3 * i + 2], a, b, c, box);
  ev_a[i] = a;
  ev_b[i] = b;
  ev_c[i] = c;
  ev_box[i] = box;
}

// Sensitivity map: kind 0 = Scanner::sensitivity(pixel centre)
// (reconstruction.h:67-73), kind 1 = Scanner::pixel_sensitivity 5-point
// rule (scanner.h:174-185, cuda/reconstruction.cu:115-126).
__global__ __launch_bounds__(256) void sensitivity_kernel(DeviceGeometry g,
                                                          int kind,
                                                          float* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n_z * g.n_y) return;
  int ix = i % g.n_z, iy = i / g.n_z;
  float z = pixel_center_z(g, ix), y = pixel_center_y(g, iy);
  float v;
  if (kind == 0) {
    v = scanner_sensitivity(g, z, y);
  } else {
#pragma clang fp contract(off)
    float hw = g.pixel_width / 2.0f, hh = g.pixel_height / 2.0f;
    v = scanner_sensitivity(g, z, y) / 3.0f +
        scanner_sensitivity(g, z + hw, y + hh) / 6.0f +
        scanner_sensitivity(g, z - hw, y - hh) / 6.0f +
        scanner_sensitivity(g, z + (-hw), y + hh) / 6.0f +
        scanner_sensitivity(g, z - (-hw), y - hh) / 6.0f;
  }
  out[i] = v;
}

struct ModeFlags {
  int kernel;       // 0 gaussian, 1 analytic
  int sensitivity;  // 0 none, 1 denominator, 2 numerator
  int guard_zero;
};

// Bounding-box wavefront kernel ("bbox" algorithm): one wave64 per event,
// grid-stride over events.  Lanes walk the clamped bbox linearised row-major
// like the reference CPU loop (reconstruction.h:255-288).  Pass 1 forms D_e
// with a wavefront reduction; pass 2 recomputes K and adds K*inv_e into the
// back-projection image with L2 float atomics.
//  FAST = true: Gaussian kernel, no sensitivity, guarded inverse (CPU
//  parity mode) with the row-hoisted prefactor and hardware exp2.
//  FAST = false: generic semantics (legacy --gpu mode and variants).
template <bool FAST, bool BACKPROJECT>
__global__ __launch_bounds__(256) void bbox_event_kernel(
    DeviceGeometry g, ModeFlags m, const float4* __restrict__ ev_a,
    const float4* __restrict__ ev_b, const uint2* __restrict__ ev_box,
    int64_t n_events, const float* __restrict__ rho,
    const float* __restrict__ sens, float* __restrict__ bp,
    float* __restrict__ denominators) {
  const int lane = threadIdx.x & 63;
  const int64_t waves_per_block = blockDim.x >> 6;
  const int64_t n_waves = (int64_t)gridDim.x * waves_per_block;
  const bool legacy_split = !FAST && m.kernel == 1;
  for (int64_t e = (int64_t)blockIdx.x * waves_per_block + (threadIdx.x >> 6);
       e < n_events; e += n_waves) {
    const float4 a = ev_a[e];
    const float4 b = ev_b[e];
    const uint2 box = ev_box[e];
    const int x0 = box.x & 0xffff, x1 = box.x >> 16;
    const int y0 = box.y & 0xffff, y1 = box.y >> 16;
    const int width = x1 - x0, height = y1 - y0;
    if (width <= 0 || height <= 0) {
      if (!BACKPROJECT && lane == 0) denominators[e] = 0.0f;
      continue;
    }
    const int n = width * height;
    const float z = a.x, y = a.y, tan = a.z, sec = a.w;
    const float inv_width_f = 1.0f / (float)width;

    float acc = 0.0f;
    for (int idx = lane; idx < n; idx += 64) {
      int r, c;
      if (legacy_split) {
        // warp_granularity.cuh:162-176 (float reciprocal row split)
        r = (int)(short)(int)((float)idx * inv_width_f);
        c = idx - width * r;
      } else {
        r = idx / width;
        c = idx - r * width;
      }
      const int ix = x0 + c, iy = y0 + r;
      // the float split can step one column outside the bbox; keep every
      // access inside the image
      if (!FAST && (ix < 0 || ix >= g.n_z || iy >= g.n_y)) continue;
      const float dz = pixel_center_z(g, ix) - z;
      const float dy = pixel_center_y(g, iy) - y;
      const float q = ellipse_q_exact(b.x, b.y, b.z, dy, dz);
      if (q <= 9.0f) {
        const int p = iy * g.n_z + ix;
        float k;
        if (FAST || m.kernel == 0)
          k = gauss_row_prefactor(g, dy + y, tan, sec) * gauss_exp(q);
        else
          k = analytic_kernel(g, y, tan, sec, dz, dy);
        float kr = k * rho[p];
        if (!FAST && m.sensitivity == 1) kr *= sens[p];
        acc += kr;
      }
    }
    const float d = wave_sum(acc);
    if (!BACKPROJECT) {
      if (lane == 0) denominators[e] = d;
      continue;
    }
    float inv;
    if (FAST || m.guard_zero)
      inv = (d > 0.0f) ? 1.0f / d : 0.0f;
    else
      inv = 1.0f / d;

    for (int idx = lane; idx < n; idx += 64) {
      int r, c;
      if (legacy_split) {
        r = (int)(short)(int)((float)idx * inv_width_f);
        c = idx - width * r;
      } else {
        r = idx / width;
        c = idx - r * width;
      }
      const int ix = x0 + c, iy = y0 + r;
      // the float split can step one column outside the bbox; keep every
      // access inside the image
      if (!FAST && (ix < 0 || ix >= g.n_z || iy >= g.n_y)) continue;
      const float dz = pixel_center_z(g, ix) - z;
      const float dy = pixel_center_y(g, iy) - y;
      const float q = ellipse_q_exact(b.x, b.y, b.z, dy, dz);
      if (q <= 9.0f) {
        const int p = iy * g.n_z + ix;
        float k;
        if (FAST || m.kernel == 0)
          k = gauss_row_prefactor(g, dy + y, tan, sec) * gauss_exp(q);
        else
          k = analytic_kernel(g, y, tan, sec, dz, dy);
        unsafeAtomicAdd(&bp[p], k * inv);
      }
    }
  }
}

// rho'(p) = rho(p) * bp(p) [/ s(p) in numerator mode]; clears bp for the
// next iteration (a fill kernel rather than hipMemsetAsync, so iterations can
// be captured in a HIP graph).
__global__ __launch_bounds__(256) void finalize_kernel(
    int n_pixels, const float* __restrict__ rho, float* __restrict__ bp,
    const float* __restrict__ sens, int divide_by_sens,
    float* __restrict__ rho_next) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_pixels;
       i += gridDim.x * blockDim.x) {
    float v = rho[i] * bp[i];
    if (divide_by_sens) v = v / sens[i];
    rho_next[i] = v;
    bp[i] = 0.0f;
  }
}

__global__ __launch_bounds__(256) void add_kernel(int64_t n, float value,
                                                  float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] += value;
}

__global__ __launch_bounds__(256) void fill_kernel(float* __restrict__ out,
                                                   int64_t n, float value) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = value;
}

// Exact kernel-evaluation count sum_e P_e: one wave per event, same support
// test as the iteration kernels.
__global__ __launch_bounds__(256) void count_support_kernel(
    DeviceGeometry g, const float4* __restrict__ ev_a,
    const float4* __restrict__ ev_b, const uint2* __restrict__ ev_box,
    int64_t n_events, unsigned long long* __restrict__ total) {
  const int lane = threadIdx.x & 63;
  const int64_t waves_per_block = blockDim.x >> 6;
  const int64_t n_waves = (int64_t)gridDim.x * waves_per_block;
  unsigned long long mine = 0;
  for (int64_t e = (int64_t)blockIdx.x * waves_per_block + (threadIdx.x >> 6);
       e < n_events; e += n_waves) {
    const float4 a = ev_a[e];
    const float4 b = ev_b[e];
    const uint2 box = ev_box[e];
    const int x0 = box.x & 0xffff, x1 = box.x >> 16;
    const int y0 = box.y & 0xffff, y1 = box.y >> 16;
    const int width = x1 - x0, height = y1 - y0;
    if (width <= 0 || height <= 0) continue;
    const int n = width * height;
    int cnt = 0;
    for (int idx = lane; idx < n; idx += 64) {
      int r = idx / width, c = idx - r * width;
      const float dz = pixel_center_z(g, x0 + c) - a.x;
      const float dy = pixel_center_y(g, y0 + r) - a.y;
      cnt += ellipse_q_exact(b.x, b.y, b.z, dy, dz) <= 9.0f;
    }
    mine += (unsigned long long)wave_sum_int(cnt);
  }
  if (lane == 0 && mine) atomicAdd(total, mine);
}


// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

namespace strip {

// Scanner<float,short> constants (scanner.h:39-67) and GaussianKernel
// constants (gaussian_kernel.h:27-33) in FP32, as the reference computes
// them on the host.
DeviceGeometry make_device_geometry(const strip_mlem_geometry& geo) {
#pragma clang fp contract(off)
  DeviceGeometry g;
  const float R = (float)geo.radius, L = (float)geo.scintillator_length;
  const float w = (float)geo.pixel_width, h = (float)geo.pixel_height;
  const float cy = (float)geo.center_y, cz = (float)geo.center_z;
  const short ny = (short)geo.n_y_pixels, nz = (short)geo.n_z_pixels;
  const float size_y = (float)ny * h;  // scanner.h:61
  const float size_z = (float)nz * w;  // scanner.h:62
  g.radius = R;
  g.half_length = L / 2.0f;
  g.pixel_width = w;
  g.pixel_height = h;
  g.tl_y = cy + size_y / 2.0f;
  g.tl_z = cz - size_z / 2.0f;
  g.tl_y_half_h = g.tl_y - h / 2.0f;
  g.tl_z_half_w = g.tl_z + w / 2.0f;
  const float sz = (float)geo.sigma_z, sdl = (float)geo.sigma_dl;
  g.inv_pow_sigma_z = 1.0f / (sz * sz);
  g.inv_pow_sigma_dl = 1.0f / (sdl * sdl);
  // 1/(2 pi^2) folded in double (M_PI is a double), sqrt taken in float
  g.norm_c = (float)(1.0 / (2.0 * M_PI * M_PI) *
                     (double)std::sqrt(g.inv_pow_sigma_z * g.inv_pow_sigma_z *
                                       g.inv_pow_sigma_dl));
  g.n_z = nz;
  g.n_y = ny;
  g.inv_pixel_width = 1.0f / w;
  g.C = 2.0f * g.inv_pow_sigma_z;  // as ellipse_bb computes C
  g.inv_2c = 0.5f / g.C;
  return g;
}

}  // namespace strip

struct strip_mlem_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  strip_mlem_geometry geometry{};
  strip_mlem_mode mode{};
  DeviceGeometry g{};
  int n_pixels = 0;
  int64_t n_events = 0, cap_events = 0;
  EventSoA ev{nullptr, nullptr, nullptr, nullptr};
  float* rho[2] = {nullptr, nullptr};
  int cur = 0;
  float* bp = nullptr;
  float* sens = nullptr;  // map used by the mode (only when it needs one)
  float* sens_out = nullptr;  // scratch for strip_mlem_get_sensitivity
  float* staging = nullptr;   // AoS staging for host uploads
  int64_t staging_cap = 0;
  float* denominators = nullptr;
  int64_t denominators_cap = 0;
  unsigned long long* counter = nullptr;
  int n_blocks = 0, threads = 256;
  int algorithm = 0;  // requested: 0 = automatic, 1 = bbox, 2 = row band
  int cu_count = 0;
  strip::BandPlan band;
  bool band_dirty = true;
  // differentiation: work buffers, per-iteration image checkpoints
  // (column-major) and row-major input/output staging
  strip::AdjointWork adjoint;
  // ordered subsets: plans over contiguous event ranges sharing images
  int n_subsets = 1;
  std::vector<strip::BandPlan> subsets;
  bool subsets_dirty = true;
  float* osemT = nullptr;  // 3 column-major images: rho (2), back-projection
  int osem_cur = 0;
  // events per row-band plan of plain MLEM (more are split over plans)
  int64_t plan_events = int64_t(1) << 25;
  // streamed uploads: pinned host and device staging, double-buffered
  hipStream_t copy_stream = nullptr;
  float* pinned[2] = {nullptr, nullptr};
  float* dstage[2] = {nullptr, nullptr};
  int64_t stage_cap = 0;
  hipEvent_t copied[2] = {nullptr, nullptr};
  hipEvent_t consumed[2] = {nullptr, nullptr};
  float* ckpt = nullptr;
  int64_t ckpt_images = 0;
  float* io = nullptr;  // 2 row-major images
  char name[256] = {0};
  // timing
  bool timing = false;
  std::vector<hipEvent_t> event_pool;
  size_t event_used = 0;
  double forward_ms = 0.0, backward_ms = 0.0, iteration_ms = 0.0;
  int64_t timed_iterations = 0;
};

namespace {

#define CTX_CHECK(ctx)                                               \
  do {                                                               \
    if (!(ctx)) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL"); \
    HIP_TRY(hipSetDevice((ctx)->device));                            \
  } while (0)

int check_geometry(const strip_mlem_geometry* g) {
  if (!g) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "geometry is NULL");
  if (g->n_z_pixels <= 0 || g->n_y_pixels <= 0 || g->n_z_pixels > 32767 ||
      g->n_y_pixels > 32767)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT,
                       "pixel counts must be in [1, 32767]");
  if (!(g->pixel_width > 0) || !(g->pixel_height > 0))
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "pixel size must be > 0");
  if (!(g->sigma_z > 0) || !(g->sigma_dl > 0))
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "sigmas must be > 0");
  if (!(g->radius > 0))
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "radius must be > 0");
  return STRIP_MLEM_OK;
}

int check_mode(const strip_mlem_mode* m) {
  if (m->kernel < 0 || m->kernel > 1 || m->sensitivity < 0 ||
      m->sensitivity > 2 || m->sensitivity_map < 0 || m->sensitivity_map > 1)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "invalid mode");
  return STRIP_MLEM_OK;
}

bool is_fast_mode(const strip_mlem_mode& m) {
  return m.kernel == 0 && m.sensitivity == 0 && m.guard_zero != 0;
}

int grid_for_events(const strip_mlem_ctx* ctx, int64_t n_events) {
  const int waves_per_block = ctx->threads / 64;
  int64_t needed = (n_events + waves_per_block - 1) / waves_per_block;
  int64_t cap = ctx->n_blocks > 0 ? ctx->n_blocks
                                  : (int64_t)std::max(ctx->cu_count, 1) * 16;
  return (int)std::max<int64_t>(1, std::min(needed, cap));
}

hipEvent_t next_event(strip_mlem_ctx* ctx) {
  if (ctx->event_used == ctx->event_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ctx->event_pool.push_back(e);
  }
  return ctx->event_pool[ctx->event_used++];
}

// Sums the recorded (start, after forward, after back-projection, end)
// quadruples.
int collect_timing(strip_mlem_ctx* ctx) {
  if (ctx->event_used == 0) return STRIP_MLEM_OK;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (size_t i = 0; i + 4 <= ctx->event_used; i += 4) {
    float f = 0.f, b = 0.f, t = 0.f;
    HIP_TRY(hipEventElapsedTime(&f, ctx->event_pool[i], ctx->event_pool[i + 1]));
    HIP_TRY(hipEventElapsedTime(&b, ctx->event_pool[i + 1], ctx->event_pool[i + 2]));
    HIP_TRY(hipEventElapsedTime(&t, ctx->event_pool[i], ctx->event_pool[i + 3]));
    ctx->forward_ms += f;
    ctx->backward_ms += b;
    ctx->iteration_ms += t;
    ctx->timed_iterations += 1;
  }
  ctx->event_used = 0;
  return STRIP_MLEM_OK;
}

int ensure_mode_sensitivity(strip_mlem_ctx* ctx) {
  if (ctx->mode.sensitivity == 0 || ctx->sens) return STRIP_MLEM_OK;
  HIP_TRY(hipMalloc(&ctx->sens, sizeof(float) * ctx->n_pixels));
  const int blocks = (ctx->n_pixels + 255) / 256;
  hipLaunchKernelGGL(sensitivity_kernel, dim3(blocks), dim3(256), 0,
                     ctx->stream, ctx->g, ctx->mode.sensitivity_map,
                     ctx->sens);
  HIP_TRY(hipGetLastError());
  return STRIP_MLEM_OK;
}

// Event arrays for at least n events; the first `keep` resident events are
// preserved when the arrays grow.
int reserve_events(strip_mlem_ctx* ctx, int64_t n, int64_t keep = 0) {
  if (n <= ctx->cap_events) return STRIP_MLEM_OK;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  EventSoA next{nullptr, nullptr, nullptr, nullptr};
  if (hipMalloc(&next.a, sizeof(float4) * n) != hipSuccess ||
      hipMalloc(&next.b, sizeof(float4) * n) != hipSuccess ||
      hipMalloc(&next.c, sizeof(float4) * n) != hipSuccess ||
      hipMalloc(&next.box, sizeof(uint2) * n) != hipSuccess) {
    (void)hipFree(next.a);
    (void)hipFree(next.b);
    (void)hipFree(next.c);
    (void)hipFree(next.box);
    return strip::fail(STRIP_MLEM_ERR_ALLOC, "cannot allocate event arrays");
  }
  if (keep > 0) {
    HIP_TRY(hipMemcpyAsync(next.a, ctx->ev.a, sizeof(float4) * keep,
                           hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(next.b, ctx->ev.b, sizeof(float4) * keep,
                           hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(next.c, ctx->ev.c, sizeof(float4) * keep,
                           hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(next.box, ctx->ev.box, sizeof(uint2) * keep,
                           hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  (void)hipFree(ctx->ev.a);
  (void)hipFree(ctx->ev.b);
  (void)hipFree(ctx->ev.c);
  (void)hipFree(ctx->ev.box);
  ctx->ev = next;
  ctx->cap_events = n;
  return STRIP_MLEM_OK;
}

int run_setup(strip_mlem_ctx* ctx, const float* d_responses, int64_t offset,
              int64_t n) {
  if (n <= 0) return STRIP_MLEM_OK;
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(setup_events_kernel, dim3((unsigned)blocks), dim3(256),
                     0, ctx->stream, ctx->g, d_responses, n,
                     ctx->ev.a + offset, ctx->ev.b + offset,
                     ctx->ev.c + offset, ctx->ev.box + offset);
  HIP_TRY(hipGetLastError());
  return STRIP_MLEM_OK;
}

// Host AoS events -> resident events [offset, offset + n): chunks of up to
// kStreamChunk events are copied into pinned memory, sent by a copy stream
// and set up on the compute stream, two chunks in flight.
constexpr int64_t kStreamChunk = (int64_t)1 << 23;

int upload_events(strip_mlem_ctx* ctx, const float* responses, int64_t offset,
                  int64_t n) {
  if (n <= 0) return STRIP_MLEM_OK;
  const int64_t chunk = std::min(n, kStreamChunk);
  if (!ctx->copy_stream) {
    HIP_TRY(hipStreamCreate(&ctx->copy_stream));
    for (int i = 0; i < 2; ++i) {
      HIP_TRY(hipEventCreateWithFlags(&ctx->copied[i], hipEventDisableTiming));
      HIP_TRY(hipEventCreateWithFlags(&ctx->consumed[i],
                                      hipEventDisableTiming));
    }
  }
  if (chunk > ctx->stage_cap) {
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->copy_stream));
    for (int i = 0; i < 2; ++i) {
      if (ctx->pinned[i]) HIP_TRY(hipHostFree(ctx->pinned[i]));
      if (ctx->dstage[i]) HIP_TRY(hipFree(ctx->dstage[i]));
      ctx->pinned[i] = nullptr;
      ctx->dstage[i] = nullptr;
    }
    ctx->stage_cap = 0;
    for (int i = 0; i < 2; ++i)
      if (hipHostMalloc(&ctx->pinned[i], sizeof(float) * 3 * chunk) !=
              hipSuccess ||
          hipMalloc(&ctx->dstage[i], sizeof(float) * 3 * chunk) != hipSuccess)
        return strip::fail(STRIP_MLEM_ERR_ALLOC, "cannot allocate staging");
    ctx->stage_cap = chunk;
  }
  int64_t k = 0;
  for (int64_t off = 0; off < n; off += chunk, ++k) {
    const int64_t m = std::min(chunk, n - off);
    const int i = (int)(k & 1);
    // buffers i are free once chunk k-2 has been set up
    if (k >= 2) HIP_TRY(hipEventSynchronize(ctx->consumed[i]));
    std::memcpy(ctx->pinned[i], responses + 3 * off, sizeof(float) * 3 * m);
    HIP_TRY(hipMemcpyAsync(ctx->dstage[i], ctx->pinned[i],
                           sizeof(float) * 3 * m, hipMemcpyHostToDevice,
                           ctx->copy_stream));
    HIP_TRY(hipEventRecord(ctx->copied[i], ctx->copy_stream));
    HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->copied[i], 0));
    int rc = run_setup(ctx, ctx->dstage[i], offset + off, m);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(ctx->consumed[i], ctx->stream));
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return STRIP_MLEM_OK;
}

ModeFlags mode_flags(const strip_mlem_mode& m) {
  return ModeFlags{m.kernel, m.sensitivity, m.guard_zero};
}

// One event pass: forward + back-projection of every resident event into bp.
int launch_event_pass(strip_mlem_ctx* ctx, const float* rho) {
  if (ctx->n_events == 0) return STRIP_MLEM_OK;
  const int grid = grid_for_events(ctx, ctx->n_events);
  if (is_fast_mode(ctx->mode)) {
    hipLaunchKernelGGL((bbox_event_kernel<true, true>), dim3(grid),
                       dim3(ctx->threads), 0, ctx->stream, ctx->g,
                       mode_flags(ctx->mode), ctx->ev.a, ctx->ev.b,
                       ctx->ev.box, ctx->n_events, rho, ctx->sens, ctx->bp,
                       nullptr);
  } else {
    hipLaunchKernelGGL((bbox_event_kernel<false, true>), dim3(grid),
                       dim3(ctx->threads), 0, ctx->stream, ctx->g,
                       mode_flags(ctx->mode), ctx->ev.a, ctx->ev.b,
                       ctx->ev.box, ctx->n_events, rho, ctx->sens, ctx->bp,
                       nullptr);
  }
  HIP_TRY(hipGetLastError());
  return STRIP_MLEM_OK;
}

bool band_supported(const strip_mlem_ctx* ctx) {
  return ctx->mode.kernel == 0 && ctx->mode.sensitivity != 1 &&
         strip::band_supported_grid(ctx->g) &&
         strip::band_window_fits(ctx->g);
}

int effective_algorithm(const strip_mlem_ctx* ctx) {
  if (ctx->algorithm == 1) return 1;
  if (ctx->algorithm == 2) return 2;
  return band_supported(ctx) ? 2 : 1;
}

// Number of row-band plans the events are split into: the ordered subsets,
// or, for plain MLEM with the CPU semantics, ceil(n / plan_events) (the plans,
// contiguous event ranges like the subsets', then accumulate one update
// together, keeping every plan's 32-bit pair indices in range).
int plan_parts(const strip_mlem_ctx* ctx) {
  if (ctx->n_subsets > 1) return ctx->n_subsets;
  if (ctx->mode.sensitivity != 0 || ctx->n_events <= ctx->plan_events)
    return 1;
  return (int)((ctx->n_events + ctx->plan_events - 1) / ctx->plan_events);
}

int record_marks(strip_mlem_ctx* ctx, hipEvent_t* marks) {
  for (int i = 0; i < 4; ++i) {
    marks[i] = next_event(ctx);
    if (!marks[i])
      return strip::fail(STRIP_MLEM_ERR_HIP, "hipEventCreate failed");
  }
  return STRIP_MLEM_OK;
}

int launch_iteration(strip_mlem_ctx* ctx) {
  const float* rho = ctx->rho[ctx->cur];
  float* rho_next = ctx->rho[ctx->cur ^ 1];
  hipEvent_t marks[4] = {nullptr, nullptr, nullptr, nullptr};
  if (ctx->timing) {
    int rc = record_marks(ctx, marks);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(marks[0], ctx->stream));
  }
  int rc = launch_event_pass(ctx, rho);
  if (rc) return rc;
  if (ctx->timing) {
    HIP_TRY(hipEventRecord(marks[1], ctx->stream));
    HIP_TRY(hipEventRecord(marks[2], ctx->stream));
  }
  const int blocks = std::max(1, std::min((ctx->n_pixels + 255) / 256, 4096));
  hipLaunchKernelGGL(finalize_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                     ctx->n_pixels, rho, ctx->bp, ctx->sens,
                     ctx->mode.sensitivity == 2 ? 1 : 0, rho_next);
  HIP_TRY(hipGetLastError());
  if (ctx->timing) HIP_TRY(hipEventRecord(marks[3], ctx->stream));
  ctx->cur ^= 1;
  return STRIP_MLEM_OK;
}

// Row-band plan for the resident events (rebuilt after set_events).
int ensure_band(strip_mlem_ctx* ctx) {
  if (!ctx->band_dirty && ctx->band.ready) return STRIP_MLEM_OK;
  std::string err;
  if (strip::band_prepare(ctx->band, ctx->g, ctx->ev, ctx->n_events,
                          ctx->stream, &err))
    return strip::fail(STRIP_MLEM_ERR_HIP, "band plan: " + err);
  ctx->band_dirty = false;
  return STRIP_MLEM_OK;
}

int band_iterations(strip_mlem_ctx* ctx, int n_iterations) {
  int rc = ensure_band(ctx);
  if (rc) return rc;
  std::string err;
  if (strip::band_load_images(ctx->band, ctx->g, ctx->rho[ctx->cur],
                              ctx->mode.sensitivity == 2 ? ctx->sens : nullptr,
                              ctx->stream, &err))
    return strip::fail(STRIP_MLEM_ERR_HIP, err);
  for (int i = 0; i < n_iterations; ++i) {
    hipEvent_t marks[4] = {nullptr, nullptr, nullptr, nullptr};
    if (ctx->timing) {
      if ((rc = record_marks(ctx, marks))) return rc;
      HIP_TRY(hipEventRecord(marks[0], ctx->stream));
    }
    if (strip::band_iteration(ctx->band, ctx->g, ctx->ev,
                              ctx->mode.sensitivity, ctx->mode.guard_zero,
                              ctx->stream, marks[1], marks[2], &err))
      return strip::fail(STRIP_MLEM_ERR_HIP, err);
    if (ctx->timing) HIP_TRY(hipEventRecord(marks[3], ctx->stream));
  }
  if (strip::band_store_image(ctx->band, ctx->g, ctx->rho[ctx->cur],
                              ctx->stream, &err))
    return strip::fail(STRIP_MLEM_ERR_HIP, err);
  return STRIP_MLEM_OK;
}

// Differentiation needs the CPU-path semantics on the row-band plan.
int fill(strip_mlem_ctx* ctx, float* ptr, int64_t n, float value);
int ensure_subsets(strip_mlem_ctx* ctx);

int adjoint_ready(strip_mlem_ctx* ctx, bool full_plan = true) {
  if (ctx->mode.kernel != 0 || ctx->mode.sensitivity != 0 ||
      ctx->mode.guard_zero != 1)
    return strip::fail(STRIP_MLEM_ERR_UNSUPPORTED,
                       "gradients are defined for the CPU-path semantics "
                       "(Gaussian kernel, no sensitivity, guarded 1/D)");
  if (effective_algorithm(ctx) != 2)
    return strip::fail(STRIP_MLEM_ERR_UNSUPPORTED,
                       "gradients need the row-band algorithm");
  if (!full_plan) return STRIP_MLEM_OK;
  int rc = ensure_band(ctx);
  if (rc) return rc;
  std::string err;
  if (strip::band_adjoint_prepare(ctx->band, ctx->g, ctx->adjoint, &err))
    return strip::fail(STRIP_MLEM_ERR_ALLOC, err);
  return STRIP_MLEM_OK;
}

// At least n_images column-major checkpoint images and the row-major staging.
int ensure_images(strip_mlem_ctx* ctx, int64_t n_images) {
  if (!ctx->io) {
    if (hipMalloc(&ctx->io, sizeof(float) * 2 * ctx->n_pixels) != hipSuccess) {
      ctx->io = nullptr;
      return strip::fail(STRIP_MLEM_ERR_ALLOC, "cannot allocate image staging");
    }
  }
  if (n_images > ctx->ckpt_images) {
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->ckpt) HIP_TRY(hipFree(ctx->ckpt));
    ctx->ckpt = nullptr;
    ctx->ckpt_images = 0;
    if (hipMalloc(&ctx->ckpt, sizeof(float) * n_images * ctx->n_pixels) !=
        hipSuccess) {
      ctx->ckpt = nullptr;
      return strip::fail(STRIP_MLEM_ERR_ALLOC, "cannot allocate checkpoints");
    }
    ctx->ckpt_images = n_images;
  }
  return STRIP_MLEM_OK;
}

int loglik_gradient_device(strip_mlem_ctx* ctx, const float* d_rho,
                           double* value, float* d_grad, double* grad_sigma) {
  int rc = adjoint_ready(ctx, plan_parts(ctx) == 1);
  if (rc) return rc;
  if ((rc = ensure_images(ctx, 2))) return rc;
  std::string err;
  float* rhoT = ctx->ckpt;
  float* gradT = ctx->ckpt + ctx->n_pixels;
  double gl[2] = {0.0, 0.0};
  if (strip::band_to_column_major(ctx->g, d_rho ? d_rho : ctx->rho[ctx->cur],
                                  rhoT, ctx->stream, &err))
    return strip::fail(STRIP_MLEM_ERR_HIP, err);
  if (plan_parts(ctx) > 1) {
    // the sum over the subsets' plans (each holds a part of the events)
    if ((rc = ensure_subsets(ctx))) return rc;
    if (d_grad && (rc = fill(ctx, gradT, ctx->n_pixels, 0.0f))) return rc;
    double sum_log = 0.0, total = 0.0;
    for (auto& p : ctx->subsets) {
      double part = 0.0;
      if (strip::band_adjoint_prepare(p, ctx->g, ctx->adjoint, &err) ||
          strip::band_loglik_terms(p, ctx->g, ctx->adjoint, rhoT,
                                   d_grad ? gradT : nullptr, &part, gl,
                                   ctx->stream, &err))
        return strip::fail(STRIP_MLEM_ERR_HIP, err);
      sum_log += part;
    }
    if (strip::band_image_total(ctx->g, ctx->adjoint, rhoT, &total,
                                ctx->stream, &err))
      return strip::fail(STRIP_MLEM_ERR_HIP, err);
    if (value) *value = sum_log - total;
    if (d_grad) {
      if (strip::band_to_row_major(ctx->g, gradT, d_grad, ctx->stream, &err))
        return strip::fail(STRIP_MLEM_ERR_HIP, err);
      hipLaunchKernelGGL(add_kernel, dim3(256), dim3(256), 0, ctx->stream,
                         (int64_t)ctx->n_pixels, -1.0f, d_grad);
      HIP_TRY(hipGetLastError());
    }
  } else if (strip::band_loglik_gradient(ctx->band, ctx->g, ctx->adjoint, rhoT,
                                         d_grad ? gradT : nullptr, value, gl,
                                         ctx->stream, &err) ||
             (d_grad && strip::band_to_row_major(ctx->g, gradT, d_grad,
                                                 ctx->stream, &err))) {
    return strip::fail(STRIP_MLEM_ERR_HIP, err);
  }
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (grad_sigma) {
    grad_sigma[0] = gl[0] / ctx->geometry.sigma_z;
    grad_sigma[1] = gl[1] / ctx->geometry.sigma_dl;
  }
  return STRIP_MLEM_OK;
}

int unrolled_gradient_device(strip_mlem_ctx* ctx, const float* d_rho0,
                             int n_iterations, const float* d_target,
                             double* loss, float* d_grad, double* grad_sigma) {
  if (n_iterations < 0 || !d_target)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT,
                       "n_iterations >= 0 and a target image are required");
  int rc = adjoint_ready(ctx);
  if (rc) return rc;
  // checkpoints rho_0 .. rho_N, the target, the adjoint image, then the
  // back-projections B_0 .. B_{N-1} of the forward iterations
  if ((rc = ensure_images(ctx, 2 * (int64_t)n_iterations + 3))) return rc;
  std::string err;
  strip::BandPlan& plan = ctx->band;
  const int64_t np = ctx->n_pixels;
  auto ck = [&](int64_t k) { return ctx->ckpt + k * np; };
  float* targetT = ck(n_iterations + 1);
  float* gbarT = ck(n_iterations + 2);
  auto bk = [&](int64_t k) { return ck(n_iterations + 3 + k); };
  if (strip::band_to_column_major(ctx->g, d_rho0 ? d_rho0 : ctx->rho[ctx->cur],
                                  ck(0), ctx->stream, &err) ||
      strip::band_to_column_major(ctx->g, d_target, targetT, ctx->stream,
                                  &err))
    return strip::fail(STRIP_MLEM_ERR_HIP, err);
  for (int k = 0; k < n_iterations; ++k) {
    HIP_TRY(hipMemcpyAsync(plan.rhoT[plan.curT], ck(k), sizeof(float) * np,
                           hipMemcpyDeviceToDevice, ctx->stream));
    if (strip::band_iteration(plan, ctx->g, ctx->ev, 0, 1, ctx->stream,
                              nullptr, nullptr, &err, 1.0, bk(k)))
      return strip::fail(STRIP_MLEM_ERR_HIP, err);
    HIP_TRY(hipMemcpyAsync(ck(k + 1), plan.rhoT[plan.curT], sizeof(float) * np,
                           hipMemcpyDeviceToDevice, ctx->stream));
  }
  double gl[2] = {0.0, 0.0};
  if (strip::band_mse_seed(ctx->g, ctx->adjoint, ck(n_iterations), targetT,
                           gbarT, loss, ctx->stream, &err))
    return strip::fail(STRIP_MLEM_ERR_HIP, err);
  for (int k = n_iterations - 1; k >= 0; --k)
    if (strip::band_mlem_adjoint_step(plan, ctx->g, ctx->adjoint, ck(k),
                                      gbarT, gl, ctx->stream, &err, bk(k)))
      return strip::fail(STRIP_MLEM_ERR_HIP, err);
  // the context keeps rho_N as its image, as after n_iterations iterations
  if ((d_grad &&
       strip::band_to_row_major(ctx->g, gbarT, d_grad, ctx->stream, &err)) ||
      strip::band_to_row_major(ctx->g, ck(n_iterations), ctx->rho[ctx->cur],
                               ctx->stream, &err))
    return strip::fail(STRIP_MLEM_ERR_HIP, err);
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  if (grad_sigma) {
    grad_sigma[0] = gl[0] / ctx->geometry.sigma_z;
    grad_sigma[1] = gl[1] / ctx->geometry.sigma_dl;
  }
  return STRIP_MLEM_OK;
}

// The plans the iterations run: the subset (or split) plans when there are
// several (they share one tuning), else the whole-list plan.
int active_plans(strip_mlem_ctx* ctx,
                 std::vector<const strip::BandPlan*>& plans) {
  plans.clear();
  if (plan_parts(ctx) > 1) {
    int rc = ensure_subsets(ctx);
    if (rc) return rc;
    for (const strip::BandPlan& p : ctx->subsets) plans.push_back(&p);
  } else {
    int rc = ensure_band(ctx);
    if (rc) return rc;
    plans.push_back(&ctx->band);
  }
  return STRIP_MLEM_OK;
}

// Plans of the ordered subsets: contiguous event ranges, shared images.
int ensure_subsets(strip_mlem_ctx* ctx) {
  const int S = plan_parts(ctx);
  if (!ctx->subsets_dirty && (int)ctx->subsets.size() == S) {
    bool ready = true;
    for (auto& p : ctx->subsets) ready = ready && p.ready;
    if (ready) return STRIP_MLEM_OK;
  }
  for (auto& p : ctx->subsets) strip::band_release(p);
  ctx->subsets.clear();
  const int64_t np = ctx->n_pixels;
  if (!ctx->osemT) {
    if (hipMalloc(&ctx->osemT, sizeof(float) * 3 * np) != hipSuccess) {
      ctx->osemT = nullptr;
      return strip::fail(STRIP_MLEM_ERR_ALLOC, "cannot allocate images");
    }
    int rc = fill(ctx, ctx->osemT + 2 * np, np, 0.0f);
    if (rc) return rc;
  }
  ctx->subsets.resize(S);
  std::string err;
  for (int s = 0; s < S; ++s) {
    const int64_t lo = ctx->n_events * s / S;
    const int64_t hi = ctx->n_events * (s + 1) / S;
    strip::BandPlan& p = ctx->subsets[s];
    p.tuning = ctx->band.tuning;
    p.rhoT[0] = ctx->osemT;
    p.rhoT[1] = ctx->osemT + np;
    p.bpT = ctx->osemT + 2 * np;
    p.owns_images = false;
    EventSoA view{ctx->ev.a + lo, ctx->ev.b + lo, ctx->ev.c + lo,
                  ctx->ev.box + lo};
    if (strip::band_prepare(p, ctx->g, view, hi - lo, ctx->stream, &err))
      return strip::fail(STRIP_MLEM_ERR_HIP, "subset plan: " + err);
  }
  ctx->subsets_dirty = false;
  return STRIP_MLEM_OK;
}

// n full OSEM iterations: every subset s updates the image with
//   rho' = rho * (sum_{e in s} K_e / D_e) * (N / N_s),
// the count ratio N / N_s applied in FP64.
int osem_iterations(strip_mlem_ctx* ctx, int n_iterations) {
  int rc = ensure_subsets(ctx);
  if (rc) return rc;
  std::string err;
  const int64_t np = ctx->n_pixels;
  if (strip::band_to_column_major(ctx->g, ctx->rho[ctx->cur], ctx->osemT,
                                  ctx->stream, &err))
    return strip::fail(STRIP_MLEM_ERR_HIP, err);
  ctx->osem_cur = 0;
  for (int it = 0; it < n_iterations; ++it) {
    hipEvent_t marks[4] = {nullptr, nullptr, nullptr, nullptr};
    if (ctx->timing) {
      if ((rc = record_marks(ctx, marks))) return rc;
      HIP_TRY(hipEventRecord(marks[0], ctx->stream));
    }
    for (auto& p : ctx->subsets) {
      if (p.n_events == 0) continue;
      p.curT = ctx->osem_cur;
      const double ratio = (double)ctx->n_events / (double)p.n_events;
      if (strip::band_iteration(p, ctx->g, ctx->ev, 0, ctx->mode.guard_zero,
                                ctx->stream, nullptr, nullptr, &err, ratio))
        return strip::fail(STRIP_MLEM_ERR_HIP, err);
      ctx->osem_cur = p.curT;
    }
    if (ctx->timing) {
      // a whole pass over the subsets is reported as its forward part
      HIP_TRY(hipEventRecord(marks[1], ctx->stream));
      HIP_TRY(hipEventRecord(marks[2], ctx->stream));
      HIP_TRY(hipEventRecord(marks[3], ctx->stream));
    }
  }
  if (strip::band_to_row_major(ctx->g, ctx->osemT + ctx->osem_cur * np,
                               ctx->rho[ctx->cur], ctx->stream, &err))
    return strip::fail(STRIP_MLEM_ERR_HIP, err);
  return STRIP_MLEM_OK;
}

// n plain MLEM iterations over several plans (more events than one plan
// holds): every plan adds its back-projection into the shared accumulator,
// then one update rho' = rho * bp.
int split_iterations(strip_mlem_ctx* ctx, int n_iterations) {
  int rc = ensure_subsets(ctx);
  if (rc) return rc;
  std::string err;
  const int64_t np = ctx->n_pixels;
  if (strip::band_to_column_major(ctx->g, ctx->rho[ctx->cur], ctx->osemT,
                                  ctx->stream, &err))
    return strip::fail(STRIP_MLEM_ERR_HIP, err);
  ctx->osem_cur = 0;
  for (int it = 0; it < n_iterations; ++it) {
    hipEvent_t marks[4] = {nullptr, nullptr, nullptr, nullptr};
    if (ctx->timing) {
      if ((rc = record_marks(ctx, marks))) return rc;
      HIP_TRY(hipEventRecord(marks[0], ctx->stream));
    }
    for (auto& p : ctx->subsets) {
      p.curT = ctx->osem_cur;
      if (p.n_events > 0 &&
          strip::band_iteration(p, ctx->g, ctx->ev, 0, ctx->mode.guard_zero,
                                ctx->stream, nullptr, nullptr, &err, 1.0,
                                nullptr, false))
        return strip::fail(STRIP_MLEM_ERR_HIP, err);
    }
    if (ctx->timing) {
      // the pass over the plans is reported as its forward part
      HIP_TRY(hipEventRecord(marks[1], ctx->stream));
      HIP_TRY(hipEventRecord(marks[2], ctx->stream));
    }
    strip::BandPlan& p0 = ctx->subsets[0];
    if (strip::band_finalize(p0, ctx->g, 0, ctx->stream, &err))
      return strip::fail(STRIP_MLEM_ERR_HIP, err);
    ctx->osem_cur = p0.curT;
    if (ctx->timing) HIP_TRY(hipEventRecord(marks[3], ctx->stream));
  }
  if (strip::band_to_row_major(ctx->g, ctx->osemT + ctx->osem_cur * np,
                               ctx->rho[ctx->cur], ctx->stream, &err))
    return strip::fail(STRIP_MLEM_ERR_HIP, err);
  return STRIP_MLEM_OK;
}

int fill(strip_mlem_ctx* ctx, float* ptr, int64_t n, float value) {
  int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                     ptr, n, value);
  HIP_TRY(hipGetLastError());
  return STRIP_MLEM_OK;
}

}  // namespace

extern "C" {

const char* strip_mlem_last_error(void) { return strip::g_last_error.c_str(); }

int strip_mlem_create(const strip_mlem_geometry* geometry,
                      const strip_mlem_mode* mode, int device,
                      strip_mlem_ctx** out) {
  strip::g_last_error.clear();
  if (!out) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "out is NULL");
  *out = nullptr;
  int rc = check_geometry(geometry);
  if (rc) return rc;
  strip_mlem_mode m = {0, 0, 0, 1};
  if (mode) {
    rc = check_mode(mode);
    if (rc) return rc;
    m = *mode;
  }
  int n_devices = 0;
  HIP_TRY(hipGetDeviceCount(&n_devices));
  if (device < 0 || device >= n_devices)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT,
                       "device index out of range (" +
                           std::to_string(n_devices) + " devices)");
  HIP_TRY(hipSetDevice(device));
  strip_mlem_ctx* ctx = new strip_mlem_ctx;
  ctx->device = device;
  ctx->geometry = *geometry;
  ctx->mode = m;
  ctx->g = strip::make_device_geometry(*geometry);
  ctx->n_pixels = geometry->n_z_pixels * geometry->n_y_pixels;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    ctx->cu_count = prop.multiProcessorCount;
    std::snprintf(ctx->name, sizeof(ctx->name), "%s (%s)", prop.name,
                  prop.gcnArchName);
  }
  auto cleanup = [&](int code) {
    strip_mlem_destroy(ctx);
    return code;
  };
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) !=
      hipSuccess)
    return cleanup(strip::fail(STRIP_MLEM_ERR_HIP, "hipStreamCreate failed"));
  if (hipMalloc(&ctx->rho[0], sizeof(float) * ctx->n_pixels) != hipSuccess ||
      hipMalloc(&ctx->rho[1], sizeof(float) * ctx->n_pixels) != hipSuccess ||
      hipMalloc(&ctx->bp, sizeof(float) * ctx->n_pixels) != hipSuccess ||
      hipMalloc(&ctx->counter, sizeof(unsigned long long)) != hipSuccess)
    return cleanup(strip::fail(STRIP_MLEM_ERR_ALLOC, "cannot allocate images"));
  if ((rc = fill(ctx, ctx->rho[0], ctx->n_pixels, 1.0f)) ||
      (rc = fill(ctx, ctx->bp, ctx->n_pixels, 0.0f)) ||
      (rc = ensure_mode_sensitivity(ctx)))
    return cleanup(rc);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess)
    return cleanup(strip::fail(STRIP_MLEM_ERR_HIP, "initialisation failed"));
  *out = ctx;
  return STRIP_MLEM_OK;
}

void strip_mlem_destroy(strip_mlem_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
  (void)hipFree(ctx->ev.a);
  (void)hipFree(ctx->ev.b);
  (void)hipFree(ctx->ev.c);
  (void)hipFree(ctx->ev.box);
  strip::band_release(ctx->band);
  for (auto& p : ctx->subsets) strip::band_release(p);
  (void)hipFree(ctx->osemT);
  for (int i = 0; i < 2; ++i) {
    if (ctx->pinned[i]) (void)hipHostFree(ctx->pinned[i]);
    (void)hipFree(ctx->dstage[i]);
    if (ctx->copied[i]) (void)hipEventDestroy(ctx->copied[i]);
    if (ctx->consumed[i]) (void)hipEventDestroy(ctx->consumed[i]);
  }
  if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
  strip::adjoint_release(ctx->adjoint);
  (void)hipFree(ctx->ckpt);
  (void)hipFree(ctx->io);
  (void)hipFree(ctx->rho[0]);
  (void)hipFree(ctx->rho[1]);
  (void)hipFree(ctx->bp);
  (void)hipFree(ctx->sens);
  (void)hipFree(ctx->sens_out);
  (void)hipFree(ctx->staging);
  (void)hipFree(ctx->denominators);
  (void)hipFree(ctx->counter);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int strip_mlem_set_events(strip_mlem_ctx* ctx, const float* responses,
                          int64_t n) {
  CTX_CHECK(ctx);
  if (n < 0 || (n > 0 && !responses))
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "bad event array");
  int rc = reserve_events(ctx, n);
  if (rc) return rc;
  ctx->n_events = 0;
  ctx->band_dirty = true;
  ctx->subsets_dirty = true;
  if ((rc = upload_events(ctx, responses, 0, n))) return rc;
  ctx->n_events = n;
  return STRIP_MLEM_OK;
}

int strip_mlem_reserve_events(strip_mlem_ctx* ctx, int64_t n_events) {
  CTX_CHECK(ctx);
  if (n_events < 0)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "n_events < 0");
  return reserve_events(ctx, n_events, ctx->n_events);
}

int strip_mlem_append_events(strip_mlem_ctx* ctx, const float* responses,
                             int64_t n) {
  CTX_CHECK(ctx);
  if (n < 0 || (n > 0 && !responses))
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "bad event array");
  const int64_t have = ctx->n_events;
  if (have + n > ctx->cap_events) {
    const int64_t grow = std::max(have + n, have + have / 2);
    int rc = reserve_events(ctx, grow, have);
    if (rc) return rc;
  }
  ctx->band_dirty = true;
  ctx->subsets_dirty = true;
  int rc = upload_events(ctx, responses, have, n);
  if (rc) return rc;
  ctx->n_events = have + n;
  return STRIP_MLEM_OK;
}

int strip_mlem_set_events_device(strip_mlem_ctx* ctx,
                                 const float* device_responses, int64_t n) {
  CTX_CHECK(ctx);
  if (n < 0 || (n > 0 && !device_responses))
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "bad event array");
  int rc = reserve_events(ctx, n);
  if (rc) return rc;
  if ((rc = run_setup(ctx, device_responses, 0, n))) return rc;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->n_events = n;
  ctx->band_dirty = true;
  ctx->subsets_dirty = true;
  return STRIP_MLEM_OK;
}

int64_t strip_mlem_n_events(const strip_mlem_ctx* ctx) {
  return ctx ? ctx->n_events : -1;
}

int strip_mlem_set_rho(strip_mlem_ctx* ctx, const float* rho) {
  CTX_CHECK(ctx);
  if (!rho) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "rho is NULL");
  HIP_TRY(hipMemcpyAsync(ctx->rho[ctx->cur], rho,
                         sizeof(float) * ctx->n_pixels, hipMemcpyHostToDevice,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return STRIP_MLEM_OK;
}

int strip_mlem_get_rho(strip_mlem_ctx* ctx, float* rho) {
  CTX_CHECK(ctx);
  if (!rho) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "rho is NULL");
  HIP_TRY(hipMemcpyAsync(rho, ctx->rho[ctx->cur],
                         sizeof(float) * ctx->n_pixels, hipMemcpyDeviceToHost,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return STRIP_MLEM_OK;
}

int strip_mlem_set_rho_device(strip_mlem_ctx* ctx, const float* device_rho) {
  CTX_CHECK(ctx);
  if (!device_rho) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "rho is NULL");
  HIP_TRY(hipMemcpyAsync(ctx->rho[ctx->cur], device_rho,
                         sizeof(float) * ctx->n_pixels,
                         hipMemcpyDeviceToDevice, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return STRIP_MLEM_OK;
}

int strip_mlem_get_rho_device(strip_mlem_ctx* ctx, float* device_rho) {
  CTX_CHECK(ctx);
  if (!device_rho) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "rho is NULL");
  HIP_TRY(hipMemcpyAsync(device_rho, ctx->rho[ctx->cur],
                         sizeof(float) * ctx->n_pixels,
                         hipMemcpyDeviceToDevice, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return STRIP_MLEM_OK;
}

int strip_mlem_loglik_gradient_device(strip_mlem_ctx* ctx, const float* rho,
                                     double* value, float* grad_rho,
                                     double* grad_sigma) {
  CTX_CHECK(ctx);
  return loglik_gradient_device(ctx, rho, value, grad_rho, grad_sigma);
}

int strip_mlem_loglik_gradient(strip_mlem_ctx* ctx, const float* rho,
                               double* value, float* grad_rho,
                               double* grad_sigma) {
  CTX_CHECK(ctx);
  int rc = ensure_images(ctx, 2);
  if (rc) return rc;
  const size_t bytes = sizeof(float) * ctx->n_pixels;
  if (rho)
    HIP_TRY(hipMemcpyAsync(ctx->io, rho, bytes, hipMemcpyHostToDevice,
                           ctx->stream));
  if ((rc = loglik_gradient_device(ctx, rho ? ctx->io : nullptr, value,
                                   grad_rho ? ctx->io + ctx->n_pixels : nullptr,
                                   grad_sigma)))
    return rc;
  if (grad_rho) {
    HIP_TRY(hipMemcpyAsync(grad_rho, ctx->io + ctx->n_pixels, bytes,
                           hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  return STRIP_MLEM_OK;
}

int strip_mlem_unrolled_gradient_device(strip_mlem_ctx* ctx, const float* rho0,
                                        int n_iterations, const float* target,
                                        double* loss, float* grad_rho0,
                                        double* grad_sigma) {
  CTX_CHECK(ctx);
  return unrolled_gradient_device(ctx, rho0, n_iterations, target, loss,
                                  grad_rho0, grad_sigma);
}

int strip_mlem_unrolled_gradient(strip_mlem_ctx* ctx, const float* rho0,
                                 int n_iterations, const float* target,
                                 double* loss, float* grad_rho0,
                                 double* grad_sigma) {
  CTX_CHECK(ctx);
  if (!target) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "target is NULL");
  int rc = ensure_images(ctx, 2);
  if (rc) return rc;
  const size_t bytes = sizeof(float) * ctx->n_pixels;
  float* d_rho0 = ctx->io;
  float* d_target = ctx->io + ctx->n_pixels;
  if (rho0)
    HIP_TRY(hipMemcpyAsync(d_rho0, rho0, bytes, hipMemcpyHostToDevice,
                           ctx->stream));
  HIP_TRY(hipMemcpyAsync(d_target, target, bytes, hipMemcpyHostToDevice,
                         ctx->stream));
  // the gradient reuses the first staging image once rho0 has been read
  if ((rc = unrolled_gradient_device(ctx, rho0 ? d_rho0 : nullptr,
                                     n_iterations, d_target, loss,
                                     grad_rho0 ? d_rho0 : nullptr,
                                     grad_sigma)))
    return rc;
  if (grad_rho0) {
    HIP_TRY(hipMemcpyAsync(grad_rho0, d_rho0, bytes, hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  return STRIP_MLEM_OK;
}

int strip_mlem_get_sensitivity(strip_mlem_ctx* ctx, int kind,
                               float* sensitivity) {
  CTX_CHECK(ctx);
  if (!sensitivity || kind < 0 || kind > 1)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "bad sensitivity request");
  if (!ctx->sens_out)
    HIP_TRY(hipMalloc(&ctx->sens_out, sizeof(float) * ctx->n_pixels));
  const int blocks = (ctx->n_pixels + 255) / 256;
  hipLaunchKernelGGL(sensitivity_kernel, dim3(blocks), dim3(256), 0,
                     ctx->stream, ctx->g, kind, ctx->sens_out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(sensitivity, ctx->sens_out,
                         sizeof(float) * ctx->n_pixels, hipMemcpyDeviceToHost,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return STRIP_MLEM_OK;
}

int strip_mlem_iterate_async(strip_mlem_ctx* ctx, int n_iterations) {
  CTX_CHECK(ctx);
  if (n_iterations < 0)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "negative iteration count");
  if (n_iterations == 0) return STRIP_MLEM_OK;
  if (ctx->n_subsets > 1) return osem_iterations(ctx, n_iterations);
  if (effective_algorithm(ctx) == 2)
    return plan_parts(ctx) > 1 ? split_iterations(ctx, n_iterations)
                               : band_iterations(ctx, n_iterations);
  for (int i = 0; i < n_iterations; ++i) {
    int rc = launch_iteration(ctx);
    if (rc) return rc;
  }
  return STRIP_MLEM_OK;
}

int strip_mlem_generate_events(strip_mlem_ctx* ctx,
                               const strip_phantom_params* params,
                               const strip_phantom_region* regions,
                               int n_regions, int64_t n_detected) {
  CTX_CHECK(ctx);
  if (!params || n_detected < 0)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "bad arguments");
  const int64_t have = ctx->n_events;
  int rc = reserve_events(ctx, have + n_detected, have);
  if (rc) return rc;
  strip::PhantomDevice P;
  std::string err;
  const strip_mlem_geometry& G = ctx->geometry;
  if (strip::phantom_prepare(P, *params, regions, n_regions, G.radius,
                             G.scintillator_length, G.sigma_z, G.sigma_dl,
                             &err)) {
    strip::phantom_release(P);
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, err);
  }
  const int64_t chunk = std::min<int64_t>(std::max<int64_t>(n_detected, 1),
                                          strip::kPhantomBatch);
  float* stage = nullptr;
  if (hipMalloc(&stage, sizeof(float) * 3 * chunk) != hipSuccess) {
    strip::phantom_release(P);
    return strip::fail(STRIP_MLEM_ERR_ALLOC, "cannot allocate staging");
  }
  ctx->band_dirty = true;
  ctx->subsets_dirty = true;
  int64_t done = 0;
  while (done < n_detected && !rc) {
    int64_t got = 0;
    const int64_t want = std::min(chunk, n_detected - done);
    if (strip::phantom_detect(P, want, INT64_MAX, stage, nullptr, nullptr,
                              nullptr, nullptr, &got, ctx->stream, &err)) {
      rc = strip::fail(STRIP_MLEM_ERR_HIP, err);
      break;
    }
    if ((rc = run_setup(ctx, stage, have + done, got))) break;
    done += got;
    ctx->n_events = have + done;
  }
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(stage);
  strip::phantom_release(P);
  return rc;
}

int strip_mlem_set_subsets(strip_mlem_ctx* ctx, int n_subsets) {
  CTX_CHECK(ctx);
  if (n_subsets < 1 || n_subsets > 4096)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "n_subsets must be 1..4096");
  if (n_subsets > 1 &&
      (ctx->mode.kernel != 0 || ctx->mode.sensitivity != 0 ||
       effective_algorithm(ctx) != 2))
    return strip::fail(STRIP_MLEM_ERR_UNSUPPORTED,
                       "ordered subsets need the CPU-path semantics on the "
                       "row-band algorithm");
  if (n_subsets != ctx->n_subsets) {
    ctx->n_subsets = n_subsets;
    ctx->subsets_dirty = true;
  }
  return STRIP_MLEM_OK;
}

int strip_mlem_set_plan_events(strip_mlem_ctx* ctx, int64_t n) {
  CTX_CHECK(ctx);
  if (n < 1 || n > (int64_t(1) << 30))
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "plan events must be 1..2^30");
  if (n != ctx->plan_events) {
    ctx->plan_events = n;
    ctx->subsets_dirty = true;
  }
  return STRIP_MLEM_OK;
}

int strip_mlem_image_total(strip_mlem_ctx* ctx, double* total) {
  CTX_CHECK(ctx);
  if (!total) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "total is NULL");
  std::string err;
  // a row-major image sums like a column-major one of the same size
  if (strip::band_image_total(ctx->g, ctx->adjoint, ctx->rho[ctx->cur], total,
                              ctx->stream, &err))
    return strip::fail(STRIP_MLEM_ERR_HIP, err);
  return STRIP_MLEM_OK;
}

int strip_mlem_synchronize(strip_mlem_ctx* ctx) {
  CTX_CHECK(ctx);
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return STRIP_MLEM_OK;
}

int strip_mlem_iterate(strip_mlem_ctx* ctx, int n_iterations) {
  int rc = strip_mlem_iterate_async(ctx, n_iterations);
  if (rc) return rc;
  return strip_mlem_synchronize(ctx);
}

int strip_mlem_set_launch(strip_mlem_ctx* ctx, int n_blocks,
                          int threads_per_block) {
  if (!ctx) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  if (n_blocks < 0 || threads_per_block < 0 || threads_per_block % 64 ||
      threads_per_block > 256)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT,
                       "threads per block must be a multiple of 64, <= 256");
  ctx->n_blocks = n_blocks;
  ctx->threads = threads_per_block ? threads_per_block : 256;
  return STRIP_MLEM_OK;
}

int strip_mlem_set_algorithm(strip_mlem_ctx* ctx, int algorithm) {
  if (!ctx) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  if (algorithm < 0 || algorithm > 2)
    return strip::fail(STRIP_MLEM_ERR_UNSUPPORTED, "unknown algorithm");
  if (algorithm == 2 && !band_supported(ctx))
    return strip::fail(STRIP_MLEM_ERR_UNSUPPORTED,
                       "the row-band algorithm needs the Gaussian kernel, no "
                       "sensitivity in the denominator, n_z and n_y <= 32767 "
                       "and a 16-row window of the widest bbox within 160 KiB "
                       "of LDS");
  if (algorithm == 1 && ctx->n_subsets > 1)
    return strip::fail(STRIP_MLEM_ERR_STATE,
                       "ordered subsets run on the row-band algorithm: set "
                       "one subset before choosing the bbox kernel");
  ctx->algorithm = algorithm;
  return STRIP_MLEM_OK;
}

int strip_mlem_get_algorithm(const strip_mlem_ctx* ctx) {
  return ctx ? effective_algorithm(ctx) : -1;
}

int strip_mlem_set_band_tiling(strip_mlem_ctx* ctx, int forward_rows,
                               int backward_rows, int bucket_columns,
                               int forward_pairs, int backward_pairs) {
  if (!ctx) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  if ((forward_rows != 16 && forward_rows != 32 && forward_rows != 64) ||
      (backward_rows != 16 && backward_rows != 32 && backward_rows != 64) ||
      (bucket_columns != 8 && bucket_columns != 16 && bucket_columns != 32 &&
       bucket_columns != 64) ||
      forward_pairs < 1 || backward_pairs < 1)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT,
                       "band rows 16/32/64, bucket 8/16/32/64, pairs >= 1");
  ctx->band.tuning.fwd_rows = forward_rows;
  ctx->band.tuning.bwd_rows = backward_rows;
  ctx->band.tuning.bucket = bucket_columns;
  ctx->band.tuning.fwd_pairs = forward_pairs;
  ctx->band.tuning.bwd_pairs = backward_pairs;
  ctx->band_dirty = true;
  ctx->subsets_dirty = true;
  return STRIP_MLEM_OK;
}

int strip_mlem_set_band_groups(strip_mlem_ctx* ctx, int forward_group,
                               int backward_group) {
  if (!ctx) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  auto ok = [](int v) { return v == 0 || v == 1 || v == 2 || v == 4; };
  if (!ok(forward_group) || !ok(backward_group))
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT,
                       "groups must be 1, 2 or 4 (0: automatic)");
  const bool automatic = backward_group == 0;
  const strip::BandTuning defaults;
  if (forward_group == 0) forward_group = defaults.fwd_group;
  if (automatic) backward_group = defaults.bwd_group;
  if (ctx->band.tuning.fwd_group != forward_group ||
      ctx->band.tuning.bwd_group != backward_group ||
      ctx->band.tuning.auto_groups != automatic) {
    ctx->band.tuning.fwd_group = forward_group;
    ctx->band.tuning.bwd_group = backward_group;
    ctx->band.tuning.auto_groups = automatic;
    ctx->band_dirty = true;
    ctx->subsets_dirty = true;
  }
  return STRIP_MLEM_OK;
}

int strip_mlem_get_band_groups(strip_mlem_ctx* ctx, int* forward_group,
                               int* backward_group) {
  CTX_CHECK(ctx);
  if (effective_algorithm(ctx) != 2)
    return strip::fail(STRIP_MLEM_ERR_STATE, "row-band algorithm not in use");
  std::vector<const strip::BandPlan*> plans;
  int rc = active_plans(ctx, plans);
  if (rc) return rc;
  if (forward_group) *forward_group = plans[0]->fwd_group;
  if (backward_group) *backward_group = plans[0]->bwd_group;
  return STRIP_MLEM_OK;
}

int strip_mlem_set_backprojection_waves(strip_mlem_ctx* ctx, int waves) {
  if (!ctx) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  if (waves < 0 || waves > 4)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "waves must be 0..4");
  if (ctx->band.tuning.bwd_waves != waves) {
    ctx->band.tuning.bwd_waves = waves;
    ctx->band_dirty = true;
    ctx->subsets_dirty = true;
  }
  return STRIP_MLEM_OK;
}

int strip_mlem_set_backprojection_occupancy(strip_mlem_ctx* ctx,
                                            int cu_waves) {
  if (!ctx) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  if (cu_waves < 4 || cu_waves > 32)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "cu_waves must be 4..32");
  if (ctx->band.tuning.bwd_cu_waves != cu_waves) {
    ctx->band.tuning.bwd_cu_waves = cu_waves;
    ctx->band_dirty = true;
    ctx->subsets_dirty = true;
  }
  return STRIP_MLEM_OK;
}

int strip_mlem_set_exp_mode(strip_mlem_ctx* ctx, int mode) {
  if (!ctx) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  if (mode != 0 && mode != 1)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "exp mode must be 0 or 1");
  ctx->band.tuning.recurrence = mode;
  for (auto& p : ctx->subsets) p.tuning.recurrence = mode;
  return STRIP_MLEM_OK;
}

int strip_mlem_set_span_table(strip_mlem_ctx* ctx, int enable) {
  if (!ctx) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  if (enable != 0 && enable != 1)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "enable must be 0 or 1");
  if (ctx->band.tuning.spans != enable) {
    ctx->band.tuning.spans = enable;
    ctx->band_dirty = true;
    ctx->subsets_dirty = true;
  }
  return STRIP_MLEM_OK;
}

int strip_mlem_get_span_table(strip_mlem_ctx* ctx, int* in_use) {
  CTX_CHECK(ctx);
  if (!in_use) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "NULL output");
  *in_use = 0;
  if (effective_algorithm(ctx) != 2) return STRIP_MLEM_OK;
  std::vector<const strip::BandPlan*> plans;
  int rc = active_plans(ctx, plans);
  if (rc) return rc;
  bool any = false, all = true;
  for (const strip::BandPlan* p : plans)
    if (p->n_events > 0) {
      any = true;
      all = all && p->spans != nullptr;
    }
  *in_use = any && all ? 1 : 0;
  return STRIP_MLEM_OK;
}

int strip_mlem_get_band_tiling(strip_mlem_ctx* ctx, int* forward_rows,
                               int* backward_rows, int* window_columns,
                               int64_t* forward_pairs,
                               int64_t* backward_pairs) {
  CTX_CHECK(ctx);
  if (effective_algorithm(ctx) != 2)
    return strip::fail(STRIP_MLEM_ERR_STATE, "row-band algorithm not in use");
  std::vector<const strip::BandPlan*> plans;
  int rc = active_plans(ctx, plans);
  if (rc) return rc;
  if (forward_rows) *forward_rows = plans[0]->fwd.H;
  if (backward_rows) *backward_rows = plans[0]->bwd.H;
  int wt = 0;
  int64_t nf = 0, nb = 0;
  for (const strip::BandPlan* p : plans) {
    wt = std::max(wt, p->bwd.Wt);
    nf += p->fwd.n_pairs;
    nb += p->bwd.n_pairs;
  }
  if (window_columns) *window_columns = wt;
  if (forward_pairs) *forward_pairs = nf;
  if (backward_pairs) *backward_pairs = nb;
  return STRIP_MLEM_OK;
}

int strip_mlem_kernel_evaluations(strip_mlem_ctx* ctx, int64_t* count) {
  CTX_CHECK(ctx);
  if (!count) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "count is NULL");
  HIP_TRY(hipMemsetAsync(ctx->counter, 0, sizeof(unsigned long long),
                         ctx->stream));
  if (ctx->n_events > 0 && effective_algorithm(ctx) == 2) {
    std::string err;
    if (plan_parts(ctx) > 1) {
      int rc = ensure_subsets(ctx);
      if (rc) return rc;
      for (auto& p : ctx->subsets)
        if (p.n_events > 0 && strip::band_count(p, ctx->g, ctx->ev,
                                                ctx->counter, ctx->stream,
                                                &err))
          return strip::fail(STRIP_MLEM_ERR_HIP, err);
    } else {
      int rc = ensure_band(ctx);
      if (rc) return rc;
      if (strip::band_count(ctx->band, ctx->g, ctx->ev, ctx->counter,
                            ctx->stream, &err))
        return strip::fail(STRIP_MLEM_ERR_HIP, err);
    }
  } else if (ctx->n_events > 0) {
    const int grid = grid_for_events(ctx, ctx->n_events);
    hipLaunchKernelGGL(count_support_kernel, dim3(grid), dim3(ctx->threads), 0,
                       ctx->stream, ctx->g, ctx->ev.a, ctx->ev.b, ctx->ev.box,
                       ctx->n_events, ctx->counter);
    HIP_TRY(hipGetLastError());
  }
  unsigned long long v = 0;
  HIP_TRY(hipMemcpyAsync(&v, ctx->counter, sizeof(v), hipMemcpyDeviceToHost,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  *count = (int64_t)v;
  return STRIP_MLEM_OK;
}

int strip_mlem_forward(strip_mlem_ctx* ctx, float* denominators) {
  CTX_CHECK(ctx);
  if (!denominators && ctx->n_events)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "denominators is NULL");
  if (ctx->n_events == 0) return STRIP_MLEM_OK;
  if (ctx->denominators_cap < ctx->n_events) {
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->denominators) HIP_TRY(hipFree(ctx->denominators));
    ctx->denominators = nullptr;
    ctx->denominators_cap = 0;
    HIP_TRY(hipMalloc(&ctx->denominators, sizeof(float) * ctx->n_events));
    ctx->denominators_cap = ctx->n_events;
  }
  if (effective_algorithm(ctx) == 2) {
    std::string err;
    if (plan_parts(ctx) > 1) {
      // every plan projects its event range into its part of the output
      int rc = ensure_subsets(ctx);
      if (rc) return rc;
      ctx->subsets[0].curT = 0;
      if (strip::band_load_images(ctx->subsets[0], ctx->g,
                                  ctx->rho[ctx->cur], nullptr, ctx->stream,
                                  &err))
        return strip::fail(STRIP_MLEM_ERR_HIP, err);
      int64_t lo = 0;
      for (auto& p : ctx->subsets) {
        p.curT = 0;
        if (p.n_events > 0 &&
            strip::band_forward_only(p, ctx->g, ctx->ev,
                                     ctx->denominators + lo, ctx->stream,
                                     &err))
          return strip::fail(STRIP_MLEM_ERR_HIP, err);
        lo += p.n_events;
      }
    } else {
      int rc = ensure_band(ctx);
      if (rc) return rc;
      if (strip::band_load_images(ctx->band, ctx->g, ctx->rho[ctx->cur],
                                  nullptr, ctx->stream, &err) ||
          strip::band_forward_only(ctx->band, ctx->g, ctx->ev,
                                   ctx->denominators, ctx->stream, &err))
        return strip::fail(STRIP_MLEM_ERR_HIP, err);
    }
    HIP_TRY(hipMemcpyAsync(denominators, ctx->denominators,
                           sizeof(float) * ctx->n_events,
                           hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return STRIP_MLEM_OK;
  }
  const int grid = grid_for_events(ctx, ctx->n_events);
  const float* rho = ctx->rho[ctx->cur];
  if (is_fast_mode(ctx->mode))
    hipLaunchKernelGGL((bbox_event_kernel<true, false>), dim3(grid),
                       dim3(ctx->threads), 0, ctx->stream, ctx->g,
                       mode_flags(ctx->mode), ctx->ev.a, ctx->ev.b,
                       ctx->ev.box, ctx->n_events, rho, ctx->sens, ctx->bp,
                       ctx->denominators);
  else
    hipLaunchKernelGGL((bbox_event_kernel<false, false>), dim3(grid),
                       dim3(ctx->threads), 0, ctx->stream, ctx->g,
                       mode_flags(ctx->mode), ctx->ev.a, ctx->ev.b,
                       ctx->ev.box, ctx->n_events, rho, ctx->sens, ctx->bp,
                       ctx->denominators);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(denominators, ctx->denominators,
                         sizeof(float) * ctx->n_events, hipMemcpyDeviceToHost,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return STRIP_MLEM_OK;
}

int strip_mlem_set_timing(strip_mlem_ctx* ctx, int enable) {
  if (!ctx) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "ctx is NULL");
  int rc = collect_timing(ctx);
  if (rc) return rc;
  ctx->timing = enable != 0;
  return STRIP_MLEM_OK;
}

int strip_mlem_get_timing(strip_mlem_ctx* ctx, double* kernel_ms,
                          double* iteration_ms, int64_t* n_iterations) {
  double f = 0, b = 0;
  int rc = strip_mlem_get_timing_detail(ctx, &f, &b, iteration_ms,
                                        n_iterations);
  if (rc) return rc;
  if (kernel_ms) *kernel_ms = f + b;
  return STRIP_MLEM_OK;
}

int strip_mlem_get_timing_detail(strip_mlem_ctx* ctx, double* forward_ms,
                                 double* backward_ms, double* iteration_ms,
                                 int64_t* n_iterations) {
  CTX_CHECK(ctx);
  int rc = collect_timing(ctx);
  if (rc) return rc;
  if (forward_ms) *forward_ms = ctx->forward_ms;
  if (backward_ms) *backward_ms = ctx->backward_ms;
  if (iteration_ms) *iteration_ms = ctx->iteration_ms;
  if (n_iterations) *n_iterations = ctx->timed_iterations;
  ctx->forward_ms = ctx->backward_ms = ctx->iteration_ms = 0.0;
  ctx->timed_iterations = 0;
  return STRIP_MLEM_OK;
}

int strip_mlem_device_name(strip_mlem_ctx* ctx, char* buffer, int size) {
  if (!ctx || !buffer || size <= 0)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "bad arguments");
  std::snprintf(buffer, (size_t)size, "%s", ctx->name);
  return STRIP_MLEM_OK;
}

int strip_mlem_run(const strip_mlem_geometry* geometry,
                   const strip_mlem_mode* mode, const float* responses,
                   int n_responses, float* rho_inout, int n_iteration_blocks,
                   int n_iterations_in_block, strip_mlem_output_fn output,
                   strip_mlem_progress_fn progress, void* user, int device,
                   int n_blocks, int n_threads_per_block,
                   strip_mlem_device_name_fn device_name) {
  if (!rho_inout)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "rho_inout is NULL");
  if (n_iteration_blocks < 0 || n_iterations_in_block < 0)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "negative iteration counts");
  strip_mlem_ctx* ctx = nullptr;
  int rc = strip_mlem_create(geometry, mode, device, &ctx);
  if (rc) return rc;
  struct Guard {
    strip_mlem_ctx* c;
    ~Guard() { strip_mlem_destroy(c); }
  } guard{ctx};
  if (device_name) device_name(user, ctx->name);
  if (n_threads_per_block > 256) n_threads_per_block = 256;
  n_threads_per_block -= n_threads_per_block % 64;
  if ((rc = strip_mlem_set_launch(ctx, n_blocks > 0 ? n_blocks : 0,
                                  n_threads_per_block > 0 ? n_threads_per_block
                                                          : 0)))
    return rc;
  const int nz = geometry->n_z_pixels, ny = geometry->n_y_pixels;
  // cuda/reconstruction.cu:63-70: the 5-point sensitivity goes out first
  std::vector<float> host_image((size_t)nz * ny);
  if ((rc = strip_mlem_get_sensitivity(ctx, 1, host_image.data()))) return rc;
  if (output) output(user, -1, host_image.data(), nz, ny);
  if ((rc = strip_mlem_set_rho(ctx, rho_inout))) return rc;
  if ((rc = strip_mlem_set_events(ctx, responses, n_responses))) return rc;
  for (int block = 0; block < n_iteration_blocks; ++block) {
    for (int i = 0; i < n_iterations_in_block; ++i) {
      const int iteration = block * n_iterations_in_block + i;
      if (progress) progress(user, iteration, 0);
      if ((rc = strip_mlem_iterate(ctx, 1))) return rc;
      if (progress) progress(user, iteration, 1);
      // cuda/reconstruction.cu:98-104: early outputs of block 0
      if (!block && i < n_iterations_in_block - 1 &&
          (i < 5 || i == 9 || i == 14 || i == 19 || i == 29 || i == 49 ||
           i == 99)) {
        if ((rc = strip_mlem_get_rho(ctx, rho_inout))) return rc;
        if (output) output(user, i + 1, rho_inout, nz, ny);
      }
    }
    if ((rc = strip_mlem_get_rho(ctx, rho_inout))) return rc;
    if (output)
      output(user, (block + 1) * n_iterations_in_block, rho_inout, nz, ny);
  }
  if (progress) progress(user, n_iteration_blocks * n_iterations_in_block, 0);
  return STRIP_MLEM_OK;
}

}  // extern "C"

// ===========================================================================
// Batched reconstructions: many independent event lists and grids iterated
// together.  Each member keeps its own row-band plan (records, span table,
// pair lists); the batch merges the members' task lists, tagging every task
// with its member's index, and runs each pass of an iteration as ONE launch
// over all members (persistent workgroups walking the merged task list).
// Iterations are recorded once into a HIP graph and replayed.
// ===========================================================================

struct strip_mlem_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<strip_mlem_ctx*> items;
  strip::BandTuning tuning;
  int grid_factor = 4;  // persistent workgroups per CU
  bool use_graph = true;
  // merged state, valid for the members' plan generations below
  std::vector<int64_t> generations;
  strip::BandInstance* desc[2] = {nullptr, nullptr};  // image parity 0 / 1
  int4* fwd_tasks = nullptr;
  int4* bwd_tasks = nullptr;
  int* queue = nullptr;  // task counters of the forward / back-projection
  int n_fwd = 0, n_bwd = 0;
  int H_fwd = 0, H_bwd = 0;
  int G_fwd = 1, G_bwd = 1;  // pairs per wavefront step
  int fwd_Wt = 0, bwd_waves = 1;  // widest member forward window
  int bwd_wave_words = 0;  // LDS words per back-projection wavefront
  size_t fwd_lds = 0;
  bool spans = true;
  int64_t max_events = 0, max_pixels = 0;
  int cu_count = 0;
  int cur = 0;  // image parity of every member's plan
  bool images_loaded = false;
  hipGraphExec_t graph = nullptr;  // two iterations, parity 0 then 1
  // timing of the last iterate call
  hipEvent_t t0 = nullptr, t1 = nullptr;
  double last_ms = 0.0;
};

namespace {

#define BATCH_CHECK(b)                                                 \
  do {                                                                 \
    if (!(b)) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "batch is NULL"); \
    HIP_TRY(hipSetDevice((b)->device));                                \
  } while (0)

void batch_drop_merge(strip_mlem_batch* b) {
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  b->graph = nullptr;
  for (auto& d : b->desc) {
    if (d) (void)hipFree(d);
    d = nullptr;
  }
  if (b->fwd_tasks) (void)hipFree(b->fwd_tasks);
  if (b->bwd_tasks) (void)hipFree(b->bwd_tasks);
  if (b->queue) (void)hipFree(b->queue);
  b->fwd_tasks = b->bwd_tasks = nullptr;
  b->queue = nullptr;
  b->n_fwd = b->n_bwd = 0;
  b->generations.clear();
}

bool batch_current(const strip_mlem_batch* b) {
  if (b->generations.size() != b->items.size()) return false;
  for (size_t i = 0; i < b->items.size(); ++i)
    if (b->items[i]->band_dirty || !b->items[i]->band.ready ||
        b->items[i]->band.generation != b->generations[i])
      return false;
  return true;
}

// Member plans (same tiling everywhere) and the merged task lists.
int batch_merge(strip_mlem_batch* b) {
  if (batch_current(b)) return STRIP_MLEM_OK;
  batch_drop_merge(b);
  const int n = (int)b->items.size();
  if (n > 32767)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "at most 32767 members");
  std::vector<strip::BandInstance> d0(n), d1(n);
  std::vector<int4> fwd, bwd;
  b->fwd_Wt = 1;
  b->bwd_wave_words = 0;
  b->spans = true;
  b->max_events = b->max_pixels = 0;
  for (int i = 0; i < n; ++i) {
    strip_mlem_ctx* c = b->items[i];
    if (c->band.tuning.fwd_rows != b->tuning.fwd_rows ||
        c->band.tuning.bwd_rows != b->tuning.bwd_rows ||
        c->band.tuning.bucket != b->tuning.bucket ||
        c->band.tuning.fwd_pairs != b->tuning.fwd_pairs ||
        c->band.tuning.bwd_pairs != b->tuning.bwd_pairs ||
        c->band.tuning.fwd_group != b->tuning.fwd_group ||
        c->band.tuning.bwd_group != b->tuning.bwd_group ||
        c->band.tuning.auto_groups != b->tuning.auto_groups ||
        c->band.tuning.bwd_cu_waves != b->tuning.bwd_cu_waves ||
        c->band.tuning.spans != b->tuning.spans) {
      c->band.tuning = b->tuning;
      c->band_dirty = true;
    }
    int rc = ensure_band(c);
    if (rc) return rc;
    const strip::BandPlan& P = c->band;
    // the pairs per wavefront step only shape the launches: the batch takes
    // the smallest of its members' forward G and back-projection limit, and
    // the largest back-projection LDS budget (the widest member's windows)
    if (i == 0) {
      b->H_fwd = P.fwd.H;
      b->H_bwd = P.bwd.H;
      b->G_fwd = P.fwd_group;
      b->G_bwd = P.bwd_group;
    } else if (P.fwd.H != b->H_fwd || P.bwd.H != b->H_bwd) {
      return strip::fail(STRIP_MLEM_ERR_UNSUPPORTED,
                         "members ended with different band heights");
    }
    b->G_fwd = std::min(b->G_fwd, P.fwd_group);
    b->G_bwd = std::min(b->G_bwd, P.bwd_group);
    b->fwd_Wt = std::max(b->fwd_Wt, P.fwd.Wt);
    b->bwd_wave_words = std::max(b->bwd_wave_words, P.bwd_wave_words);
    b->spans = b->spans && P.spans != nullptr;
    b->max_events = std::max<int64_t>(b->max_events, P.n_events);
    b->max_pixels = std::max<int64_t>(b->max_pixels, c->n_pixels);
    d0[i] = strip::band_descriptor(P, c->g, 0);
    d1[i] = strip::band_descriptor(P, c->g, 1);
    for (int4 t : P.fwd.tasks_host) {
      t.x |= i << 16;
      fwd.push_back(t);
    }
    for (int4 t : P.bwd.tasks_host) {
      t.x |= i << 16;
      bwd.push_back(t);
    }
    b->generations.push_back(P.generation);
  }
  // forward windows hold 2 H rows (bands aligned to their events)
  b->fwd_lds = sizeof(float) * (size_t)b->fwd_Wt * 2 * b->H_fwd;
  const size_t window = sizeof(float) * (size_t)b->bwd_wave_words;
  if (b->fwd_lds > 160 * 1024 || window > 160 * 1024)
    return strip::fail(STRIP_MLEM_ERR_UNSUPPORTED,
                       "a member's task window does not fit in LDS");
  b->bwd_waves = (int)std::max<size_t>(1, std::min<size_t>(4, 160 * 1024 / window));
  for (int c = 0; c < 2; ++c) {
    HIP_TRY(hipMalloc(&b->desc[c], sizeof(strip::BandInstance) * std::max(n, 1)));
    HIP_TRY(hipMemcpy(b->desc[c], c ? d1.data() : d0.data(),
                      sizeof(strip::BandInstance) * n, hipMemcpyHostToDevice));
  }
  // the queues hand out the largest tasks first, so the launches end on
  // small ones
  auto by_size = [](const int4& a, const int4& c) {
    return (int64_t)(a.z - a.y) * (a.w >> 16) >
           (int64_t)(c.z - c.y) * (c.w >> 16);
  };
  std::stable_sort(fwd.begin(), fwd.end(), by_size);
  std::stable_sort(bwd.begin(), bwd.end(), by_size);
  HIP_TRY(hipMalloc(&b->queue, 2 * sizeof(int)));
  b->n_fwd = (int)fwd.size();
  b->n_bwd = (int)bwd.size();
  if (b->n_fwd) {
    HIP_TRY(hipMalloc(&b->fwd_tasks, sizeof(int4) * fwd.size()));
    HIP_TRY(hipMemcpy(b->fwd_tasks, fwd.data(), sizeof(int4) * fwd.size(),
                      hipMemcpyHostToDevice));
  }
  if (b->n_bwd) {
    HIP_TRY(hipMalloc(&b->bwd_tasks, sizeof(int4) * bwd.size()));
    HIP_TRY(hipMemcpy(b->bwd_tasks, bwd.data(), sizeof(int4) * bwd.size(),
                      hipMemcpyHostToDevice));
  }
  b->images_loaded = false;
  return STRIP_MLEM_OK;
}

// Every member's image into its plan's parity-0 column-major buffer.
int batch_load_images(strip_mlem_batch* b) {
  if (b->images_loaded) return STRIP_MLEM_OK;
  std::string err;
  for (strip_mlem_ctx* c : b->items) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->band.curT = 0;
    if (strip::band_load_images(c->band, c->g, c->rho[c->cur], nullptr,
                                b->stream, &err))
      return strip::fail(STRIP_MLEM_ERR_HIP, err);
  }
  b->cur = 0;
  b->images_loaded = true;
  return STRIP_MLEM_OK;
}

int batch_store_images(strip_mlem_batch* b) {
  if (!b->images_loaded) return STRIP_MLEM_OK;
  std::string err;
  for (strip_mlem_ctx* c : b->items) {
    c->band.curT = b->cur;
    if (strip::band_store_image(c->band, c->g, c->rho[c->cur], b->stream,
                                &err))
      return strip::fail(STRIP_MLEM_ERR_HIP, err);
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->images_loaded = false;
  return STRIP_MLEM_OK;
}

// One iteration of every member with image parity c (four launches).
int batch_launch_iteration(strip_mlem_batch* b, int c) {
  std::string err;
  const int n = (int)b->items.size();
  const int persistent = std::max(1, b->cu_count * b->grid_factor);
  const bool rec = b->tuning.recurrence != 0;
  if (strip::band_zero_ints(b->queue, 2, b->stream, &err) ||
      strip::band_launch_forward(b->desc[c], b->fwd_tasks, b->n_fwd, b->H_fwd,
                                 b->G_fwd, b->spans, rec, b->fwd_lds,
                                 std::min(b->n_fwd, persistent), b->queue,
                                 b->stream, &err) ||
      strip::band_launch_batch_update(b->desc[c], n, b->max_events,
                                      b->max_pixels, b->H_fwd, true,
                                      b->stream, &err) ||
      strip::band_launch_backward(
          b->desc[c], b->bwd_tasks, b->n_bwd, b->H_bwd, b->G_bwd, b->spans,
          rec, b->bwd_wave_words, b->bwd_waves,
          std::min((b->n_bwd + b->bwd_waves - 1) / b->bwd_waves, persistent),
          b->queue + 1, b->stream, &err) ||
      strip::band_launch_batch_update(b->desc[c], n, b->max_events,
                                      b->max_pixels, b->H_fwd, false,
                                      b->stream, &err))
    return strip::fail(STRIP_MLEM_ERR_HIP, err);
  return STRIP_MLEM_OK;
}

int batch_iterate(strip_mlem_batch* b, int n_iterations) {
  int rc = batch_merge(b);
  if (rc) return rc;
  if ((rc = batch_load_images(b))) return rc;
  if (b->items.empty() || n_iterations <= 0) return batch_store_images(b);
  HIP_TRY(hipEventRecord(b->t0, b->stream));
  int done = 0;
  if (b->use_graph) {
    // eager iterations until the parity is 0 (and the LDS limits are set),
    // then the two-iteration graph
    if (!b->graph) {
      if ((rc = batch_launch_iteration(b, b->cur))) return rc;
      b->cur ^= 1;
      ++done;
      if (b->cur != 0 && done < n_iterations) {
        if ((rc = batch_launch_iteration(b, b->cur))) return rc;
        b->cur ^= 1;
        ++done;
      }
      if (b->cur == 0 && n_iterations - done >= 2) {
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int r0 = batch_launch_iteration(b, 0);
        int r1 = r0 ? r0 : batch_launch_iteration(b, 1);
        hipError_t ec = hipStreamEndCapture(b->stream, &g);
        if (r1) {
          if (g) (void)hipGraphDestroy(g);
          return r1;
        }
        if (ec != hipSuccess) return strip::fail(STRIP_MLEM_ERR_HIP,
                                                 "graph capture failed");
        hipError_t ei = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
          b->graph = nullptr;
          return strip::fail(STRIP_MLEM_ERR_HIP, "graph instantiation failed");
        }
      }
    } else if (b->cur != 0) {
      if ((rc = batch_launch_iteration(b, b->cur))) return rc;
      b->cur ^= 1;
      ++done;
    }
    while (b->graph && n_iterations - done >= 2) {
      HIP_TRY(hipGraphLaunch(b->graph, b->stream));
      done += 2;
    }
  }
  for (; done < n_iterations; ++done) {
    if ((rc = batch_launch_iteration(b, b->cur))) return rc;
    b->cur ^= 1;
  }
  HIP_TRY(hipEventRecord(b->t1, b->stream));
  HIP_TRY(hipEventSynchronize(b->t1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, b->t0, b->t1));
  b->last_ms = ms;
  return batch_store_images(b);
}

}  // namespace

extern "C" {

int strip_mlem_batch_create(int device, strip_mlem_batch** out) {
  strip::g_last_error.clear();
  if (!out) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "out is NULL");
  *out = nullptr;
  HIP_TRY(hipSetDevice(device));
  auto* b = new strip_mlem_batch();
  b->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
      hipStreamCreate(&b->stream) != hipSuccess ||
      hipEventCreate(&b->t0) != hipSuccess ||
      hipEventCreate(&b->t1) != hipSuccess) {
    delete b;
    return strip::fail(STRIP_MLEM_ERR_HIP, "cannot create batch stream");
  }
  b->cu_count = prop.multiProcessorCount;
  *out = b;
  return STRIP_MLEM_OK;
}

void strip_mlem_batch_destroy(strip_mlem_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_merge(b);
  for (strip_mlem_ctx* c : b->items) strip_mlem_destroy(c);
  if (b->t0) (void)hipEventDestroy(b->t0);
  if (b->t1) (void)hipEventDestroy(b->t1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int strip_mlem_batch_add(strip_mlem_batch* b,
                         const strip_mlem_geometry* geometry,
                         const float* responses, int64_t n_responses,
                         int* index) {
  BATCH_CHECK(b);
  if ((int)b->items.size() >= 32767)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "at most 32767 members");
  strip_mlem_mode mode{0, 0, 0, 1};
  strip_mlem_ctx* c = nullptr;
  int rc = strip_mlem_create(geometry, &mode, b->device, &c);
  if (rc) return rc;
  if (!strip::band_supported_grid(c->g) || !strip::band_window_fits(c->g)) {
    strip_mlem_destroy(c);
    return strip::fail(STRIP_MLEM_ERR_UNSUPPORTED,
                       "grid too large or bbox too wide for a batch");
  }
  c->algorithm = 2;
  c->band.tuning = b->tuning;
  if ((rc = batch_store_images(b)) ||
      (rc = strip_mlem_set_events(c, responses, n_responses))) {
    strip_mlem_destroy(c);
    return rc;
  }
  b->items.push_back(c);
  if (index) *index = (int)b->items.size() - 1;
  return STRIP_MLEM_OK;
}

int strip_mlem_batch_size(const strip_mlem_batch* b) {
  return b ? (int)b->items.size() : -1;
}

int strip_mlem_batch_context(strip_mlem_batch* b, int index,
                             strip_mlem_ctx** ctx) {
  BATCH_CHECK(b);
  if (!ctx || index < 0 || index >= (int)b->items.size())
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "bad member index");
  int rc = batch_store_images(b);
  if (rc) return rc;
  *ctx = b->items[index];
  return STRIP_MLEM_OK;
}

int strip_mlem_batch_set_tiling(strip_mlem_batch* b, int forward_rows,
                                int backward_rows, int bucket_columns,
                                int forward_pairs, int backward_pairs,
                                int grid_factor) {
  BATCH_CHECK(b);
  if ((forward_rows != 16 && forward_rows != 32 && forward_rows != 64) ||
      (backward_rows != 16 && backward_rows != 32 && backward_rows != 64) ||
      (bucket_columns != 8 && bucket_columns != 16 && bucket_columns != 32 &&
       bucket_columns != 64) ||
      forward_pairs < 1 || backward_pairs < 1 || grid_factor < 1 ||
      grid_factor > 64)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT,
                       "band rows 16/32/64, bucket 8/16/32/64, pairs >= 1, "
                       "grid factor 1..64");
  int rc = batch_store_images(b);
  if (rc) return rc;
  strip::BandTuning& t = b->tuning;
  if (t.fwd_rows != forward_rows || t.bwd_rows != backward_rows ||
      t.bucket != bucket_columns || t.fwd_pairs != forward_pairs ||
      t.bwd_pairs != backward_pairs)
    b->generations.clear();  // the members' plans are rebuilt and re-merged
  t.fwd_rows = forward_rows;
  t.bwd_rows = backward_rows;
  t.bucket = bucket_columns;
  t.fwd_pairs = forward_pairs;
  t.bwd_pairs = backward_pairs;
  if (b->grid_factor != grid_factor && b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  b->grid_factor = grid_factor;
  return STRIP_MLEM_OK;
}

int strip_mlem_batch_set_groups(strip_mlem_batch* b, int forward_group,
                                int backward_group) {
  BATCH_CHECK(b);
  auto ok = [](int v) { return v == 0 || v == 1 || v == 2 || v == 4; };
  if (!ok(forward_group) || !ok(backward_group))
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT,
                       "groups must be 1, 2 or 4 (0: automatic)");
  const bool automatic = backward_group == 0;
  const strip::BandTuning defaults;
  if (forward_group == 0) forward_group = defaults.fwd_group;
  if (automatic) backward_group = defaults.bwd_group;
  int rc = batch_store_images(b);
  if (rc) return rc;
  if (b->tuning.fwd_group != forward_group ||
      b->tuning.bwd_group != backward_group ||
      b->tuning.auto_groups != automatic)
    b->generations.clear();  // the members' plans are rebuilt and re-merged
  b->tuning.fwd_group = forward_group;
  b->tuning.bwd_group = backward_group;
  b->tuning.auto_groups = automatic;
  return STRIP_MLEM_OK;
}

int strip_mlem_batch_get_groups(const strip_mlem_batch* b, int* forward_group,
                                int* backward_group) {
  if (!b) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "batch is NULL");
  if (forward_group) *forward_group = b->tuning.fwd_group;
  if (backward_group) *backward_group = b->tuning.bwd_group;
  return STRIP_MLEM_OK;
}

int strip_mlem_batch_set_backprojection_occupancy(strip_mlem_batch* b,
                                                  int cu_waves) {
  BATCH_CHECK(b);
  if (cu_waves < 4 || cu_waves > 32)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "cu_waves must be 4..32");
  int rc = batch_store_images(b);
  if (rc) return rc;
  if (b->tuning.bwd_cu_waves != cu_waves)
    b->generations.clear();  // the members' plans are rebuilt and re-merged
  b->tuning.bwd_cu_waves = cu_waves;
  return STRIP_MLEM_OK;
}

int strip_mlem_batch_get_backprojection_occupancy(const strip_mlem_batch* b,
                                                  int* cu_waves) {
  if (!b || !cu_waves) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "NULL");
  *cu_waves = b->tuning.bwd_cu_waves;
  return STRIP_MLEM_OK;
}

int strip_mlem_batch_get_tiling(const strip_mlem_batch* b, int* tiling) {
  if (!b || !tiling) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "NULL");
  tiling[0] = b->tuning.fwd_rows;
  tiling[1] = b->tuning.bwd_rows;
  tiling[2] = b->tuning.bucket;
  tiling[3] = b->tuning.fwd_pairs;
  tiling[4] = b->tuning.bwd_pairs;
  tiling[5] = b->grid_factor;
  return STRIP_MLEM_OK;
}

int strip_mlem_batch_set_graph(strip_mlem_batch* b, int enable) {
  BATCH_CHECK(b);
  b->use_graph = enable != 0;
  return STRIP_MLEM_OK;
}

int strip_mlem_batch_iterate(strip_mlem_batch* b, int n_iterations) {
  BATCH_CHECK(b);
  if (n_iterations < 0)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "n_iterations < 0");
  return batch_iterate(b, n_iterations);
}

int strip_mlem_batch_last_ms(const strip_mlem_batch* b, double* ms) {
  if (!b || !ms) return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "NULL");
  *ms = b->last_ms;
  return STRIP_MLEM_OK;
}

int strip_mlem_batch_set_rho(strip_mlem_batch* b, int index, const float* rho) {
  strip_mlem_ctx* c = nullptr;
  int rc = strip_mlem_batch_context(b, index, &c);
  return rc ? rc : strip_mlem_set_rho(c, rho);
}

int strip_mlem_batch_get_rho(strip_mlem_batch* b, int index, float* rho) {
  strip_mlem_ctx* c = nullptr;
  int rc = strip_mlem_batch_context(b, index, &c);
  return rc ? rc : strip_mlem_get_rho(c, rho);
}

// Coordinate search over the tiling knobs: every candidate is timed on one
// batch iteration (images restored afterwards), the fastest is kept.
int strip_mlem_batch_autotune(strip_mlem_batch* b, double* best_ms) {
  BATCH_CHECK(b);
  if (b->items.empty()) return STRIP_MLEM_OK;
  int rc = batch_store_images(b);
  if (rc) return rc;
  // keep the members' images: tuning iterations must not advance them
  std::vector<std::vector<float>> saved(b->items.size());
  for (size_t i = 0; i < b->items.size(); ++i) {
    saved[i].resize(b->items[i]->n_pixels);
    if ((rc = strip_mlem_get_rho(b->items[i], saved[i].data()))) return rc;
  }
  // coordinates: the six tiling knobs, one index into the pairs-per-step
  // choices (0 = automatic), the back-projection wavefronts per CU that
  // automatic groups keep resident
  static const int kGroups[][2] = {{0, 0}, {1, 0}, {4, 0}, {1, 1}, {2, 1},
                                   {2, 2}, {2, 4}, {4, 2}, {4, 4}};
  constexpr int kNGroups = sizeof(kGroups) / sizeof(kGroups[0]);
  constexpr int kKnobs = 8;
  int cur[kKnobs];
  strip_mlem_batch_get_tiling(b, cur);
  cur[6] = 0;
  cur[7] = b->tuning.bwd_cu_waves;
  for (int g = 1; g < kNGroups; ++g) {
    const int bwd = b->tuning.auto_groups ? 0 : b->tuning.bwd_group;
    if (kGroups[g][0] == b->tuning.fwd_group && kGroups[g][1] == bwd)
      cur[6] = g;
  }
  // a candidate's time is the faster of two iterations after the one that
  // builds its plans and sets its LDS limits
  auto time_it = [&](const int* t, double* ms) -> int {
    int r = strip_mlem_batch_set_tiling(b, t[0], t[1], t[2], t[3], t[4], t[5]);
    if (!r) r = strip_mlem_batch_set_groups(b, kGroups[t[6]][0], kGroups[t[6]][1]);
    if (!r) r = strip_mlem_batch_set_backprojection_occupancy(b, t[7]);
    if (r) return r;
    const bool graph = b->use_graph;
    b->use_graph = false;
    r = batch_iterate(b, 1);
    double fastest = 0.0;
    for (int rep = 0; rep < 2 && !r; ++rep) {
      r = batch_iterate(b, 1);
      if (!r && (rep == 0 || b->last_ms < fastest)) fastest = b->last_ms;
    }
    b->use_graph = graph;
    *ms = fastest;
    return r;
  };
  double best = 0.0;
  if ((rc = time_it(cur, &best))) return rc;
  const std::vector<std::vector<int>> choices = {
      {64, 32, 16}, {16, 32, 64}, {8, 16, 32}, {1024, 2048, 4096},
      {256, 512, 1024}, {2, 4, 8}, {0, 1, 2, 3, 4, 5, 6, 7, 8},
      {8, 12, 16, 20}};
  // the knobs interact (the bucket width sets the windows that bound the
  // pairs per step), so sweeps repeat until one changes nothing
  for (int pass = 0; pass < 3; ++pass) {
    bool changed = false;
    for (int k = 0; k < kKnobs; ++k) {
      // the occupancy only matters with automatic back-projection groups
      if (k == 7 && kGroups[cur[6]][1] != 0) continue;
      for (int v : choices[k]) {
        if (v == cur[k]) continue;
        int cand[kKnobs];
        std::copy(cur, cur + kKnobs, cand);
        cand[k] = v;
        double ms = 0.0;
        if (time_it(cand, &ms)) continue;  // a shape that does not fit
        if (ms < best) {
          best = ms;
          std::copy(cand, cand + kKnobs, cur);
          changed = true;
        }
      }
    }
    if (!changed) break;
  }
  // the chosen tiling, its plans built and merged before returning
  if ((rc = strip_mlem_batch_set_tiling(b, cur[0], cur[1], cur[2], cur[3],
                                        cur[4], cur[5])) ||
      (rc = strip_mlem_batch_set_groups(b, kGroups[cur[6]][0],
                                        kGroups[cur[6]][1])) ||
      (rc = strip_mlem_batch_set_backprojection_occupancy(b, cur[7])) ||
      (rc = batch_merge(b)))
    return rc;
  for (size_t i = 0; i < b->items.size(); ++i)
    if ((rc = strip_mlem_set_rho(b->items[i], saved[i].data()))) return rc;
  if (best_ms) *best_ms = best;
  strip::g_last_error.clear();
  return STRIP_MLEM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Forward projection that allocates nothing
// ---------------------------------------------------------------------------

namespace {

// Scratch of strip_mlem_forward_noalloc: the per-event parameters of up to
// STRIP_MLEM_NOALLOC_MAX_EVENTS events, as four arrays (a, b, c: float4 each,
// clamped bbox: uint2; 56 bytes per event, 14 MiB).  Every GPU the library
// runs on holds its own copy until the process ends; host code reaches it
// only through the per-device address looked up below.
constexpr int64_t kNoallocEvents = STRIP_MLEM_NOALLOC_MAX_EVENTS;
static __device__ __align__(256) unsigned char
    scratch[kNoallocEvents * (3 * sizeof(float4) + sizeof(uint2))];

constexpr int kNoallocDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[kNoallocDevices] = {};

// Device address of the current device's scratch array, looked up once per
// device (with that device current) and kept.
int noalloc_scratch(int device, unsigned char** out) {
  if (device < 0 || device >= kNoallocDevices)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return strip::fail(STRIP_MLEM_ERR_HIP,
                         std::string("hipGetSymbolAddress(scratch): ") +
                             hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[device];
  return STRIP_MLEM_OK;
}

}  // namespace

extern "C" {

int strip_mlem_forward_noalloc(const strip_mlem_geometry* geometry,
                               const strip_mlem_mode* mode,
                               const float* device_responses, int64_t n_events,
                               const float* device_rho,
                               float* device_denominators, void* stream) {
  strip::g_last_error.clear();
  int rc = check_geometry(geometry);
  if (rc) return rc;
  strip_mlem_mode m = {0, 0, 0, 1};
  if (mode) {
    if ((rc = check_mode(mode))) return rc;
    m = *mode;
  }
  if (m.sensitivity != 0)
    return strip::fail(STRIP_MLEM_ERR_UNSUPPORTED,
                       "the allocation-free forward projection has no "
                       "sensitivity map: use a context for sensitivity modes");
  if (n_events < 0 || n_events > kNoallocEvents)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT,
                       "n_events must be 0.." +
                           std::to_string(kNoallocEvents) +
                           " (the allocation-free scratch holds that many)");
  if (n_events == 0) return STRIP_MLEM_OK;
  if (!device_responses || !device_rho || !device_denominators)
    return strip::fail(STRIP_MLEM_ERR_ARGUMENT, "NULL device pointer");
  int device = 0;
  HIP_TRY(hipGetDevice(&device));
  unsigned char* base = nullptr;
  if ((rc = noalloc_scratch(device, &base))) return rc;
  float4* a = reinterpret_cast<float4*>(base);
  float4* b = a + kNoallocEvents;
  float4* c = b + kNoallocEvents;
  uint2* box = reinterpret_cast<uint2*>(c + kNoallocEvents);
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                device));
  const strip::DeviceGeometry g = strip::make_device_geometry(*geometry);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(setup_events_kernel,
                     dim3((unsigned)((n_events + 255) / 256)), dim3(256), 0, s,
                     g, device_responses, n_events, a, b, c, box);
  HIP_TRY(hipGetLastError());
  // one wavefront per event, four per workgroup, at most 16 workgroups/CU
  const int grid = (int)std::max<int64_t>(
      1, std::min<int64_t>((n_events + 3) / 4,
                           (int64_t)std::max(cus, 1) * 16));
  if (is_fast_mode(m))
    hipLaunchKernelGGL((bbox_event_kernel<true, false>), dim3(grid),
                       dim3(256), 0, s, g, mode_flags(m), a, b, box, n_events,
                       device_rho, nullptr, nullptr, device_denominators);
  else
    hipLaunchKernelGGL((bbox_event_kernel<false, false>), dim3(grid),
                       dim3(256), 0, s, g, mode_flags(m), a, b, box, n_events,
                       device_rho, nullptr, nullptr, device_denominators);
  HIP_TRY(hipGetLastError());
  return STRIP_MLEM_OK;
}

}  // extern "C"
