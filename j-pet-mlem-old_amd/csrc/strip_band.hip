// Row-band algorithm of the strip MLEM iteration (kernels + host driver).
// The algorithm and its data layout are described in strip_band.hpp.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "strip_band.hpp"
#include "strip_device.hpp"

namespace strip {
namespace {

constexpr int kFwdThreads = 1024;      // one workgroup holds a rho band
constexpr int kBwdWaves = 4;           // wavefronts per backward workgroup
constexpr int kFwdPairsPerTask = 8192; // pairs per forward workgroup
constexpr int kBwdPairsPerTask = 1024; // pairs per backward wavefront sweep

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
};

__device__ inline RowSpan row_span(const DeviceGeometry& g, float inv_w,
                                   float inv_2c, float cz, float cy, float A,
                                   float B, float C, float P, float Q0,
                                   int x0, int x1, int y0, int y1, int iy) {
  RowSpan s;
  s.lo = 0;
  s.n = 0;
  s.rowA = 0.f;
  s.bdy = 0.f;
  s.pre = 0.f;
  if (iy < y0 || iy >= y1) return s;
  const float py = pixel_center_y(g, iy);
  const float dy = py - cy;
  row_terms_exact(A, B, dy, s.rowA, s.bdy);
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

struct EventView {
  float cz, cy, A, B, C, P, Q0;
  int x0, x1, y0, y1;
};

__device__ inline EventView load_event(const EventSoA& ev, int e) {
  const float4 a = ev.a[e];
  const float4 b = ev.b[e];
  const float4 c = ev.c[e];
  const uint2 box = ev.box[e];
  EventView v;
  v.cz = a.x;
  v.cy = a.y;
  v.A = b.x;
  v.B = b.y;
  v.C = b.z;
  v.P = c.x;
  v.Q0 = c.y;
  v.x0 = box.x & 0xffff;
  v.x1 = box.x >> 16;
  v.y0 = box.y & 0xffff;
  v.y1 = box.y >> 16;
  return v;
}

// ---------------------------------------------------------------------------
// pair list construction
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void count_pairs_kernel(
    const uint2* __restrict__ box, int64_t n, int H, int* __restrict__ counts) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint2 b = box[e];
  const int x0 = b.x & 0xffff, x1 = b.x >> 16;
  const int y0 = b.y & 0xffff, y1 = b.y >> 16;
  int c = 0;
  if (x1 > x0 && y1 > y0) c = (y1 - 1) / H - y0 / H + 1;
  counts[e] = c;
}

__global__ __launch_bounds__(256) void emit_pairs_kernel(
    const uint2* __restrict__ box, int64_t n, int H,
    const int* __restrict__ counts, const int* __restrict__ offsets,
    uint32_t* __restrict__ keys, int* __restrict__ values) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int c = counts[e];
  if (!c) return;
  const uint2 b = box[e];
  const uint32_t x0 = b.x & 0xffff;
  const int y0 = b.y & 0xffff;
  const int base = offsets[e];
  for (int k = 0; k < c; ++k) {
    keys[base + k] = ((uint32_t)(y0 / H + k) << 16) | x0;
    values[base + k] = (int)e;
  }
}

__global__ __launch_bounds__(256) void box_width_kernel(
    const uint2* __restrict__ box, int64_t n, int* __restrict__ width) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint2 b = box[e];
  width[e] = max((int)(b.x >> 16) - (int)(b.x & 0xffff), 0);
}

__global__ __launch_bounds__(256) void band_bounds_kernel(
    const uint32_t* __restrict__ keys, int64_t n, int* __restrict__ first) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = keys[i] >> 16;
  if (i == 0 || (int)(keys[i - 1] >> 16) != b) first[b] = (int)i;
}

// ---------------------------------------------------------------------------
// image layout kernels
// ---------------------------------------------------------------------------

// out[x * n_y + y] = in[y * n_z + x] * (mul ? mul[y * n_z + x] : 1)
__global__ __launch_bounds__(256) void to_column_major_kernel(
    const float* __restrict__ in, const float* __restrict__ mul, int n_z,
    int n_y, float* __restrict__ out) {
  __shared__ float tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int x = bx + tx, y = by + j;
    float v = 0.f;
    if (x < n_z && y < n_y) {
      v = in[(int64_t)y * n_z + x];
      if (mul) v *= mul[(int64_t)y * n_z + x];
    }
    tile[j][tx] = v;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int x = bx + j, y = by + tx;
    if (x < n_z && y < n_y) out[(int64_t)x * n_y + y] = tile[tx][j];
  }
}

__global__ __launch_bounds__(256) void to_row_major_kernel(
    const float* __restrict__ in, int n_z, int n_y, float* __restrict__ out) {
  __shared__ float tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;  // bx over y, by over x
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int j = ty; j < 32; j += 8) {
    const int y = bx + tx, x = by + j;
    float v = 0.f;
    if (x < n_z && y < n_y) v = in[(int64_t)x * n_y + y];
    tile[j][tx] = v;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int y = bx + j, x = by + tx;
    if (x < n_z && y < n_y) out[(int64_t)y * n_z + x] = tile[tx][j];
  }
}

// rhoT'(p) = rhoT(p) * bpT(p) [/ sT(p)]; bpT = 0
__global__ __launch_bounds__(256) void band_finalize_kernel(
    int64_t n, const float* __restrict__ rho, float* __restrict__ bp,
    const float* __restrict__ sens, int divide, float* __restrict__ next) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    float v = rho[i] * bp[i];
    if (divide) v = v / sens[i];
    next[i] = v;
    bp[i] = 0.0f;
  }
}

// ---------------------------------------------------------------------------
// iteration kernels
// ---------------------------------------------------------------------------

// Forward projection: workgroup = one rho band (column-major in LDS), waves
// take the task's pairs round-robin; lane = q*H + r handles row r of the band
// and every (64/H)-th column of the row's span starting at offset q.
// MODE 0: partial D per pair; MODE 1: count the span pixels instead.
template <int H, int MODE>
__global__ __launch_bounds__(kFwdThreads) void band_forward_kernel(
    DeviceGeometry g, EventSoA ev, const int* __restrict__ pair_event,
    const int4* __restrict__ tasks, const float* __restrict__ rhoT,
    float* __restrict__ partial, int max_bands,
    unsigned long long* __restrict__ count) {
  constexpr int LR = 64 / H;
  extern __shared__ float band[];  // n_z * H, word = col * H + row
  const int4 task = tasks[blockIdx.x];
  const int b = task.x;
  const int y_base = b * H;
  if (MODE == 0) {
    const int n_words = g.n_z * H;
    for (int i = threadIdx.x; i < n_words; i += blockDim.x) {
      const int col = i / H, r = i - col * H;
      const int iy = y_base + r;
      band[i] = (iy < g.n_y) ? rhoT[(int64_t)col * g.n_y + iy] : 0.0f;
    }
    __syncthreads();
  }
  const int lane = threadIdx.x & 63;
  const int r = lane & (H - 1);
  const int q = lane / H;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int n_waves = blockDim.x >> 6;
  const float inv_w = 1.0f / g.pixel_width;
  const float step = (float)LR * g.pixel_width;
  unsigned long long counted = 0;
  for (int p = task.y + wave; p < task.z; p += n_waves) {
    const int e = __builtin_amdgcn_readfirstlane(pair_event[p]);
    const EventView v = load_event(ev, e);
    const float inv_2c = 0.5f / v.C;
    const RowSpan s = row_span(g, inv_w, inv_2c, v.cz, v.cy, v.A, v.B, v.C,
                               v.P, v.Q0, v.x0, v.x1, v.y0, v.y1, y_base + r);
    const int nq = (s.n > q) ? (s.n - q + LR - 1) / LR : 0;
    if (MODE == 1) {
      counted += (unsigned long long)wave_sum_int(q == 0 ? s.n : 0);
      continue;
    }
    const int col0 = s.lo + q;
    float dz = __builtin_fmaf((float)col0, g.pixel_width,
                              g.tl_z_half_w - v.cz);
    const float qa = s.rowA * kExp2Scale;
    const float qb = s.bdy * kExp2Scale;
    const float qc = v.C * kExp2Scale;
    const float* src = band + col0 * H + r;
    float acc = 0.0f;
    for (int k = 0; k < nq; ++k) {
      const float qq = __builtin_fmaf(__builtin_fmaf(qc, dz, qb), dz, qa);
      acc = __builtin_fmaf(__builtin_amdgcn_exp2f(qq), src[k * LR * H], acc);
      dz += step;
    }
    const float d = wave_sum(acc * s.pre);
    if (lane == 0) partial[(int64_t)e * max_bands + (b - v.y0 / H)] = d;
  }
  if (MODE == 1 && lane == 0 && counted) atomicAdd(count, counted);
}

// D_e = sum of the event's band partials; inv_e = 1/D_e (guarded or not).
__global__ __launch_bounds__(256) void band_inverse_kernel(
    const uint2* __restrict__ box, int64_t n, int H, int max_bands,
    const float* __restrict__ partial, int guard, float* __restrict__ inv,
    float* __restrict__ denominators) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint2 b = box[e];
  const int x0 = b.x & 0xffff, x1 = b.x >> 16;
  const int y0 = b.y & 0xffff, y1 = b.y >> 16;
  float d = 0.0f;
  if (x1 > x0 && y1 > y0) {
    const int nb = (y1 - 1) / H - y0 / H + 1;
    for (int k = 0; k < nb; ++k) d += partial[e * max_bands + k];
  }
  if (denominators) denominators[e] = d;
  if (inv) inv[e] = guard ? ((d > 0.0f) ? 1.0f / d : 0.0f) : 1.0f / d;
}

// Adds window columns [c0, c1) to the column-major back-projection image and
// clears them.  lane = q*H + r handles row r of columns c0+q, c0+q+LR, ...
template <int H>
__device__ inline void flush_columns(float* win, int W, int c0, int c1,
                                     int y_base, const DeviceGeometry& g,
                                     float* __restrict__ bpT, int lane) {
  constexpr int LR = 64 / H;
  const int r = lane & (H - 1), q = lane / H;
  const int iy = y_base + r;
  for (int c = c0 + q; c < c1; c += LR) {
    float* w = win + (c & (W - 1)) * H + r;
    const float v = *w;
    *w = 0.0f;
    if (v != 0.0f && iy < g.n_y && c < g.n_z)
      unsafeAtomicAdd(&bpT[(int64_t)c * g.n_y + iy], v);
  }
}

// Back-projection: each wavefront sweeps one task (pairs of one band sorted by
// x0) through its private LDS window of W columns; plain read-modify-write.
template <int H>
__global__ __launch_bounds__(64 * kBwdWaves) void band_backward_kernel(
    DeviceGeometry g, EventSoA ev, const int* __restrict__ pair_event,
    const int4* __restrict__ tasks, int n_tasks, const float* __restrict__ inv,
    int W, float* __restrict__ bpT) {
  constexpr int LR = 64 / H;
  extern __shared__ float windows[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int t = blockIdx.x * kBwdWaves + wave;
  if (t >= n_tasks) return;
  float* win = windows + wave * W * H;
  for (int i = lane; i < W * H; i += 64) win[i] = 0.0f;
  const int4 task = tasks[t];
  const int b = task.x;
  const int y_base = b * H;
  const int r = lane & (H - 1);
  const int q = lane / H;
  const float inv_w = 1.0f / g.pixel_width;
  const float step = (float)LR * g.pixel_width;
  int F = -1;  // window holds columns [F, F + W)
  for (int p = task.y; p < task.z; ++p) {
    const int e = __builtin_amdgcn_readfirstlane(pair_event[p]);
    const EventView v = load_event(ev, e);
    if (F < 0) F = v.x0;
    if (v.x1 > F + W) {
      const int nf = v.x1 - W;
      flush_columns<H>(win, W, F, nf, y_base, g, bpT, lane);
      F = nf;
    }
    const float s_inv = inv[e];
    if (s_inv == 0.0f) continue;
    const float inv_2c = 0.5f / v.C;
    const RowSpan s = row_span(g, inv_w, inv_2c, v.cz, v.cy, v.A, v.B, v.C,
                               v.P, v.Q0, v.x0, v.x1, v.y0, v.y1, y_base + r);
    const int nq = (s.n > q) ? (s.n - q + LR - 1) / LR : 0;
    const int col0 = s.lo + q;
    float dz = __builtin_fmaf((float)col0, g.pixel_width,
                              g.tl_z_half_w - v.cz);
    const float qa = s.rowA * kExp2Scale;
    const float qb = s.bdy * kExp2Scale;
    const float qc = v.C * kExp2Scale;
    const float scale = s.pre * s_inv;
    int col = col0;
    int k = 0;
    // two columns per step: both reads are issued before the writes (the two
    // words differ, and no other lane or wave touches them)
    for (; k + 1 < nq; k += 2) {
      const float q0 = __builtin_fmaf(__builtin_fmaf(qc, dz, qb), dz, qa);
      const float dz1 = dz + step;
      const float q1 = __builtin_fmaf(__builtin_fmaf(qc, dz1, qb), dz1, qa);
      float* w0 = win + (col & (W - 1)) * H + r;
      float* w1 = win + ((col + LR) & (W - 1)) * H + r;
      const float a0 = *w0, a1 = *w1;
      *w0 = __builtin_fmaf(__builtin_amdgcn_exp2f(q0), scale, a0);
      *w1 = __builtin_fmaf(__builtin_amdgcn_exp2f(q1), scale, a1);
      dz += 2.0f * step;
      col += 2 * LR;
    }
    if (k < nq) {
      const float q0 = __builtin_fmaf(__builtin_fmaf(qc, dz, qb), dz, qa);
      float* w0 = win + (col & (W - 1)) * H + r;
      *w0 = __builtin_fmaf(__builtin_amdgcn_exp2f(q0), scale, *w0);
    }
  }
  if (F >= 0) flush_columns<H>(win, W, F, F + W, y_base, g, bpT, lane);
}

template <typename Kern>
int set_lds(Kern kernel, size_t bytes, std::string* err) {
  BAND_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes));
  return 0;
}

template <int H>
int launch_forward(BandPlan& plan, const DeviceGeometry& g, const EventSoA& ev,
                   int mode, unsigned long long* count, hipStream_t stream,
                   std::string* err) {
  if (plan.n_fwd_tasks == 0) return 0;
  if (mode == 0) {
    auto kern = band_forward_kernel<H, 0>;
    if (set_lds(kern, plan.fwd_lds, err)) return -1;
    hipLaunchKernelGGL(kern, dim3(plan.n_fwd_tasks), dim3(kFwdThreads),
                       plan.fwd_lds, stream, g, ev, plan.pair_event,
                       plan.fwd_tasks, plan.rhoT[plan.curT], plan.partial,
                       plan.max_bands, count);
  } else {
    auto kern = band_forward_kernel<H, 1>;
    hipLaunchKernelGGL(kern, dim3(plan.n_fwd_tasks), dim3(kFwdThreads), 0,
                       stream, g, ev, plan.pair_event, plan.fwd_tasks,
                       plan.rhoT[plan.curT], plan.partial, plan.max_bands,
                       count);
  }
  BAND_TRY(hipGetLastError());
  return 0;
}

template <int H>
int launch_backward(BandPlan& plan, const DeviceGeometry& g,
                    const EventSoA& ev, hipStream_t stream, std::string* err) {
  if (plan.n_bwd_tasks == 0) return 0;
  auto kern = band_backward_kernel<H>;
  if (set_lds(kern, plan.bwd_lds, err)) return -1;
  const int blocks = (plan.n_bwd_tasks + kBwdWaves - 1) / kBwdWaves;
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * kBwdWaves), plan.bwd_lds,
                     stream, g, ev, plan.pair_event, plan.bwd_tasks,
                     plan.n_bwd_tasks, plan.inv, plan.W, plan.bpT);
  BAND_TRY(hipGetLastError());
  return 0;
}

int dispatch_forward(BandPlan& plan, const DeviceGeometry& g,
                     const EventSoA& ev, int mode, unsigned long long* count,
                     hipStream_t stream, std::string* err) {
  switch (plan.H) {
    case 16: return launch_forward<16>(plan, g, ev, mode, count, stream, err);
    case 32: return launch_forward<32>(plan, g, ev, mode, count, stream, err);
    case 64: return launch_forward<64>(plan, g, ev, mode, count, stream, err);
  }
  if (err) *err = "unsupported band height";
  return -1;
}

int dispatch_backward(BandPlan& plan, const DeviceGeometry& g,
                      const EventSoA& ev, hipStream_t stream,
                      std::string* err) {
  switch (plan.H) {
    case 16: return launch_backward<16>(plan, g, ev, stream, err);
    case 32: return launch_backward<32>(plan, g, ev, stream, err);
    case 64: return launch_backward<64>(plan, g, ev, stream, err);
  }
  if (err) *err = "unsupported band height";
  return -1;
}

template <typename T>
void free_ptr(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace

int band_rows_for(const DeviceGeometry& g) {
  const size_t limit = 128 * 1024;
  for (int H : {64, 32, 16})
    if ((size_t)g.n_z * H * sizeof(float) <= limit) return H;
  return 0;
}

void band_release(BandPlan& plan) {
  free_ptr(plan.pair_event);
  free_ptr(plan.fwd_tasks);
  free_ptr(plan.bwd_tasks);
  free_ptr(plan.partial);
  free_ptr(plan.inv);
  free_ptr(plan.rhoT[0]);
  free_ptr(plan.rhoT[1]);
  free_ptr(plan.bpT);
  free_ptr(plan.sensT);
  plan = BandPlan{};
}

int band_prepare(BandPlan& plan, const DeviceGeometry& g, const EventSoA& ev,
                 int64_t n_events, hipStream_t stream, std::string* err) {
  const int H = band_rows_for(g);
  if (!H) {
    if (err) *err = "image too wide for the row-band algorithm";
    return -1;
  }
  if (n_events > 0x7fffffff) {
    if (err) *err = "too many events for one band plan";
    return -1;
  }
  const int64_t n_pixels = (int64_t)g.n_z * g.n_y;
  // keep the images; rebuild the lists
  free_ptr(plan.pair_event);
  free_ptr(plan.fwd_tasks);
  free_ptr(plan.bwd_tasks);
  free_ptr(plan.partial);
  free_ptr(plan.inv);
  plan.ready = false;
  plan.H = H;
  plan.n_bands = (g.n_y + H - 1) / H;
  plan.n_events = n_events;
  if (!plan.rhoT[0]) {
    BAND_TRY(hipMalloc(&plan.rhoT[0], sizeof(float) * n_pixels));
    BAND_TRY(hipMalloc(&plan.rhoT[1], sizeof(float) * n_pixels));
    BAND_TRY(hipMalloc(&plan.bpT, sizeof(float) * n_pixels));
    BAND_TRY(hipMemsetAsync(plan.bpT, 0, sizeof(float) * n_pixels, stream));
  }
  const int64_t n = n_events;
  int* counts = nullptr;
  int* offsets = nullptr;
  BAND_TRY(hipMalloc(&counts, sizeof(int) * std::max<int64_t>(n, 1)));
  BAND_TRY(hipMalloc(&offsets, sizeof(int) * std::max<int64_t>(n, 1)));
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (n > 0) {
    hipLaunchKernelGGL(count_pairs_kernel, dim3(blocks), dim3(256), 0, stream,
                       ev.box, n, H, counts);
    BAND_TRY(hipGetLastError());
  }
  // scan and maximum
  int* d_max = nullptr;
  BAND_TRY(hipMalloc(&d_max, sizeof(int)));
  size_t tmp_bytes = 0, t1 = 0, t2 = 0;
  if (n > 0) {
    BAND_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, t1, counts, offsets,
                                               (int)n, stream));
    BAND_TRY(hipcub::DeviceReduce::Max(nullptr, t2, counts, d_max, (int)n,
                                       stream));
  }
  tmp_bytes = std::max(t1, t2);
  void* tmp = nullptr;
  if (tmp_bytes) BAND_TRY(hipMalloc(&tmp, tmp_bytes));
  int last_count = 0, last_offset = 0, max_count = 0;
  if (n > 0) {
    BAND_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, t1, counts, offsets, (int)n,
                                               stream));
    BAND_TRY(hipcub::DeviceReduce::Max(tmp, t2, counts, d_max, (int)n,
                                       stream));
    BAND_TRY(hipMemcpyAsync(&last_count, counts + n - 1, sizeof(int),
                            hipMemcpyDeviceToHost, stream));
    BAND_TRY(hipMemcpyAsync(&last_offset, offsets + n - 1, sizeof(int),
                            hipMemcpyDeviceToHost, stream));
    BAND_TRY(hipMemcpyAsync(&max_count, d_max, sizeof(int),
                            hipMemcpyDeviceToHost, stream));
    BAND_TRY(hipStreamSynchronize(stream));
  }
  if (tmp) BAND_TRY(hipFree(tmp));
  tmp = nullptr;
  BAND_TRY(hipFree(d_max));
  const int64_t n_pairs = (int64_t)last_offset + last_count;
  plan.n_pairs = n_pairs;
  plan.max_bands = std::max(max_count, 1);
  // window width: power of two >= widest bbox (x1 - x0)
  int max_width = 1;
  if (n > 0) {
    hipLaunchKernelGGL(box_width_kernel, dim3(blocks), dim3(256), 0, stream,
                       ev.box, n, counts);
    BAND_TRY(hipGetLastError());
    int* d_w = nullptr;
    BAND_TRY(hipMalloc(&d_w, sizeof(int)));
    size_t wb = 0;
    BAND_TRY(hipcub::DeviceReduce::Max(nullptr, wb, counts, d_w, (int)n,
                                       stream));
    BAND_TRY(hipMalloc(&tmp, wb));
    BAND_TRY(hipcub::DeviceReduce::Max(tmp, wb, counts, d_w, (int)n, stream));
    BAND_TRY(hipMemcpyAsync(&max_width, d_w, sizeof(int),
                            hipMemcpyDeviceToHost, stream));
    BAND_TRY(hipStreamSynchronize(stream));
    BAND_TRY(hipFree(tmp));
    tmp = nullptr;
    BAND_TRY(hipFree(d_w));
    // the width pass overwrote the per-event pair counts: recompute them
    hipLaunchKernelGGL(count_pairs_kernel, dim3(blocks), dim3(256), 0, stream,
                       ev.box, n, H, counts);
    BAND_TRY(hipGetLastError());
  }
  uint32_t* keys = nullptr;
  uint32_t* keys_alt = nullptr;
  int* vals_alt = nullptr;
  if (n_pairs > 0) {
    BAND_TRY(hipMalloc(&keys, sizeof(uint32_t) * n_pairs));
    BAND_TRY(hipMalloc(&keys_alt, sizeof(uint32_t) * n_pairs));
    BAND_TRY(hipMalloc(&plan.pair_event, sizeof(int) * n_pairs));
    BAND_TRY(hipMalloc(&vals_alt, sizeof(int) * n_pairs));
    hipLaunchKernelGGL(emit_pairs_kernel, dim3(blocks), dim3(256), 0, stream,
                       ev.box, n, H, counts, offsets, keys, plan.pair_event);
    BAND_TRY(hipGetLastError());
    int end_bit = 16;
    while ((1 << (end_bit - 16)) < plan.n_bands + 1) ++end_bit;
    hipcub::DoubleBuffer<uint32_t> kb(keys, keys_alt);
    hipcub::DoubleBuffer<int> vb(plan.pair_event, vals_alt);
    size_t sort_bytes = 0;
    BAND_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, kb, vb,
                                                (int)n_pairs, 0, end_bit,
                                                stream));
    BAND_TRY(hipMalloc(&tmp, sort_bytes));
    BAND_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, kb, vb,
                                                (int)n_pairs, 0, end_bit,
                                                stream));
    BAND_TRY(hipStreamSynchronize(stream));
    BAND_TRY(hipFree(tmp));
    tmp = nullptr;
    if (kb.Current() != keys) std::swap(keys, keys_alt);
    if (vb.Current() != plan.pair_event) std::swap(plan.pair_event, vals_alt);
    BAND_TRY(hipFree(keys_alt));
    BAND_TRY(hipFree(vals_alt));
  }
  // band boundaries
  std::vector<int> first(plan.n_bands + 1, -1);
  if (n_pairs > 0) {
    int* d_first = nullptr;
    BAND_TRY(hipMalloc(&d_first, sizeof(int) * (plan.n_bands + 1)));
    BAND_TRY(hipMemsetAsync(d_first, 0xff, sizeof(int) * (plan.n_bands + 1),
                            stream));
    const unsigned pb = (unsigned)((n_pairs + 255) / 256);
    hipLaunchKernelGGL(band_bounds_kernel, dim3(pb), dim3(256), 0, stream,
                       keys, n_pairs, d_first);
    BAND_TRY(hipGetLastError());
    BAND_TRY(hipMemcpyAsync(first.data(), d_first,
                            sizeof(int) * (plan.n_bands + 1),
                            hipMemcpyDeviceToHost, stream));
    BAND_TRY(hipStreamSynchronize(stream));
    BAND_TRY(hipFree(d_first));
  }
  if (keys) BAND_TRY(hipFree(keys));
  BAND_TRY(hipFree(counts));
  BAND_TRY(hipFree(offsets));
  // band ranges
  std::vector<int> begin(plan.n_bands), end(plan.n_bands);
  {
    int next = (int)n_pairs;
    for (int b = plan.n_bands - 1; b >= 0; --b) {
      if (first[b] >= 0) {
        begin[b] = first[b];
        end[b] = next;
        next = first[b];
      } else {
        begin[b] = end[b] = next;
      }
    }
  }
  std::vector<int4> fwd, bwd;
  for (int b = 0; b < plan.n_bands; ++b) {
    for (int p = begin[b]; p < end[b]; p += kFwdPairsPerTask)
      fwd.push_back(make_int4(b, p, std::min(end[b], p + kFwdPairsPerTask), 0));
    for (int p = begin[b]; p < end[b]; p += kBwdPairsPerTask)
      bwd.push_back(make_int4(b, p, std::min(end[b], p + kBwdPairsPerTask), 0));
  }
  plan.n_fwd_tasks = (int)fwd.size();
  plan.n_bwd_tasks = (int)bwd.size();
  if (!fwd.empty()) {
    BAND_TRY(hipMalloc(&plan.fwd_tasks, sizeof(int4) * fwd.size()));
    BAND_TRY(hipMemcpyAsync(plan.fwd_tasks, fwd.data(),
                            sizeof(int4) * fwd.size(), hipMemcpyHostToDevice,
                            stream));
  }
  if (!bwd.empty()) {
    BAND_TRY(hipMalloc(&plan.bwd_tasks, sizeof(int4) * bwd.size()));
    BAND_TRY(hipMemcpyAsync(plan.bwd_tasks, bwd.data(),
                            sizeof(int4) * bwd.size(), hipMemcpyHostToDevice,
                            stream));
  }
  int W = 32;
  while (W < max_width) W *= 2;
  plan.W = W;
  plan.fwd_lds = (size_t)g.n_z * H * sizeof(float);
  plan.bwd_lds = (size_t)kBwdWaves * W * H * sizeof(float);
  if (plan.bwd_lds > 160 * 1024) {
    if (err) *err = "back-projection window does not fit in LDS";
    return -1;
  }
  if (n > 0) {
    BAND_TRY(hipMalloc(&plan.partial, sizeof(float) * n * plan.max_bands));
    BAND_TRY(hipMalloc(&plan.inv, sizeof(float) * n));
  }
  BAND_TRY(hipStreamSynchronize(stream));
  plan.ready = true;
  return 0;
}

int band_load_images(BandPlan& plan, const DeviceGeometry& g,
                     const float* rho, const float* sens, hipStream_t stream,
                     std::string* err) {
  const int64_t n_pixels = (int64_t)g.n_z * g.n_y;
  dim3 grid((g.n_z + 31) / 32, (g.n_y + 31) / 32);
  // denominator-sensitivity mode folds s(p) into the image the forward pass
  // reads; the stored image stays rho (see band_iteration)
  (void)n_pixels;
  hipLaunchKernelGGL(to_column_major_kernel, grid, dim3(256), 0, stream, rho,
                     nullptr, g.n_z, g.n_y, plan.rhoT[plan.curT]);
  BAND_TRY(hipGetLastError());
  if (sens) {
    if (!plan.sensT)
      BAND_TRY(hipMalloc(&plan.sensT, sizeof(float) * n_pixels));
    hipLaunchKernelGGL(to_column_major_kernel, grid, dim3(256), 0, stream,
                       sens, nullptr, g.n_z, g.n_y, plan.sensT);
    BAND_TRY(hipGetLastError());
  }
  return 0;
}

int band_store_image(BandPlan& plan, const DeviceGeometry& g, float* rho,
                     hipStream_t stream, std::string* err) {
  dim3 grid((g.n_y + 31) / 32, (g.n_z + 31) / 32);
  hipLaunchKernelGGL(to_row_major_kernel, grid, dim3(256), 0, stream,
                     plan.rhoT[plan.curT], g.n_z, g.n_y, rho);
  BAND_TRY(hipGetLastError());
  return 0;
}

int band_iteration(BandPlan& plan, const DeviceGeometry& g,
                   const EventSoA& ev, int sensitivity_mode, int guard_zero,
                   hipStream_t stream, hipEvent_t t_fwd, hipEvent_t t_bwd,
                   std::string* err) {
  if (!plan.ready) {
    if (err) *err = "band plan not prepared";
    return -1;
  }
  const int64_t n_pixels = (int64_t)g.n_z * g.n_y;
  float* rho = plan.rhoT[plan.curT];
  float* next = plan.rhoT[plan.curT ^ 1];
  if (sensitivity_mode == 1) {
    if (err) *err = "the band algorithm does not fold sensitivity into D";
    return -1;
  }
  if (dispatch_forward(plan, g, ev, 0, nullptr, stream, err)) return -1;
  if (t_fwd) BAND_TRY(hipEventRecord(t_fwd, stream));
  if (plan.n_events > 0) {
    const unsigned blocks = (unsigned)((plan.n_events + 255) / 256);
    hipLaunchKernelGGL(band_inverse_kernel, dim3(blocks), dim3(256), 0, stream,
                       ev.box, plan.n_events, plan.H, plan.max_bands,
                       plan.partial, guard_zero, plan.inv, nullptr);
    BAND_TRY(hipGetLastError());
  }
  if (dispatch_backward(plan, g, ev, stream, err)) return -1;
  if (t_bwd) BAND_TRY(hipEventRecord(t_bwd, stream));
  const int blocks = (int)std::min<int64_t>((n_pixels + 255) / 256, 4096);
  hipLaunchKernelGGL(band_finalize_kernel, dim3(blocks), dim3(256), 0, stream,
                     n_pixels, rho, plan.bpT, plan.sensT,
                     sensitivity_mode == 2 ? 1 : 0, next);
  BAND_TRY(hipGetLastError());
  plan.curT ^= 1;
  return 0;
}

int band_forward_only(BandPlan& plan, const DeviceGeometry& g,
                      const EventSoA& ev, float* denominators,
                      hipStream_t stream, std::string* err) {
  if (dispatch_forward(plan, g, ev, 0, nullptr, stream, err)) return -1;
  if (plan.n_events > 0) {
    const unsigned blocks = (unsigned)((plan.n_events + 255) / 256);
    hipLaunchKernelGGL(band_inverse_kernel, dim3(blocks), dim3(256), 0, stream,
                       ev.box, plan.n_events, plan.H, plan.max_bands,
                       plan.partial, 1, nullptr, denominators);
    BAND_TRY(hipGetLastError());
  }
  return 0;
}

int band_count(BandPlan& plan, const DeviceGeometry& g, const EventSoA& ev,
               unsigned long long* d_total, hipStream_t stream,
               std::string* err) {
  return dispatch_forward(plan, g, ev, 1, d_total, stream, err);
}

}  // namespace strip
