// Differentiation of the strip MLEM path on the row-band plan: gradients of
// the Poisson log-likelihood and reverse steps through unrolled MLEM
// iterations, with respect to the image and to (sigma_z, sigma_dl).
//
// Kernel in log form (gaussian_kernel.h:51-79 with the inverse covariance
// diag(iz, iz, id), iz = 1/sigma_z^2, id = 1/sigma_dl^2):
//   ln K = ln c - ln sqrt(a.a) - Q/2,  c = iz sqrt(id) / (2 pi^2),
//   a.a = iz (ap^2 + aq^2) + id ar^2,  Q = Qz + Qd,
//   Qz = iz 2 (dz - dy t)^2,  Qd = id (2 dy sec)^2 = (A - B^2 / (4C)) dy^2.
// With w = iz (ap^2 + aq^2) / a.a = Q0' (1 + yp^2 / R^2) pre^2 (pre = the row
// prefactor c / sqrt(a.a) = rsqrt(P' yp^2 + Q0')):
//   d ln K / d ln sigma_z  = -2 + w + Qz = -2 + Q + (w - Qd),
//   d ln K / d ln sigma_dl = Qd - w.
// w and Qd are constant along an image row, so a pass needs per row the sums
// sum K x and sum K x Q only.  The support (bbox, Q <= 9) of each event stays
// that of the plan: the derivative is the almost-everywhere derivative of the
// pixel sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "strip_band.hpp"
#include "strip_band_device.hpp"
#include "strip_device.hpp"

namespace strip {
namespace {

constexpr int kMoments = 3;  // per image: sum K x, sum K x Q, sum K x (w - Qd)

// Per (event, band) moments of NIMG column-major images x0T (and x1T):
// moments[(e * max_bands + band) * 3 * NIMG + 3 i + {0, 1, 2}] =
//   sum K x_i, sum K x_i Q, sum K x_i (w - Qd)   over the band's pixels.
// Same tiling as band_forward_kernel: one workgroup per task, the windows of
// the NIMG images (2H rows each, the pairs' bands being aligned to their
// events) staged in LDS, one row of a pair's band per lane.
template <int H, int NIMG, bool SPANS>
__global__ __launch_bounds__(64 * kFwdWaves) void band_moments_kernel(
    DeviceGeometry g, const BandRecord* __restrict__ rec,
    const int* __restrict__ pair_event, const int4* __restrict__ tasks,
    const uint16_t* __restrict__ spans, int span_stride,
    const float* __restrict__ x0T, const float* __restrict__ x1T, int Wt,
    float* __restrict__ moments, int max_bands) {
  extern __shared__ float win[];
  const int4 task = tasks[blockIdx.x];
  const int b = task.x & 0xffff;
  const int c0 = task.w & 0xffff;
  const int y_base = b * H;
  constexpr int WH = 2 * H;  // window rows
  // the columns the task's pairs reach (Wt: the window stride)
  const int n_words = min(task.w >> 16, Wt) * WH;
  for (int i = threadIdx.x; i < n_words; i += blockDim.x) {
    const int col = c0 + i / WH, iy = y_base + (i & (WH - 1));
    const bool in = col < g.n_z && iy < g.n_y;
    const int64_t at = (int64_t)col * g.n_y + iy;
    win[i] = in ? x0T[at] : 0.0f;
    if (NIMG == 2) win[n_words + i] = in ? x1T[at] : 0.0f;
  }
  __syncthreads();
  constexpr int LR = 64 / H;
  const int lane = threadIdx.x & 63;
  const int r = lane & (H - 1);
  const int q = lane / H;
  const float step = (float)LR * g.pixel_width;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int total = task.z - task.y;
  const int share = (total + kFwdWaves - 1) / kFwdWaves;
  const int begin = task.y + wave * share;
  const int end = min(task.z, begin + share);
  const float qc = g.C * kExp2Scale;
  const float two_qc_step = 2.0f * qc * step;
  const float d2q = two_qc_step * step;  // second difference of the exponent
  const float G = __builtin_amdgcn_exp2f(d2q);
  const float inv_scale = 1.0f / kExp2Scale;
  const float inv_r2 = 1.0f / (g.radius * g.radius);
  PairBatch next = gather_batch<false>(rec, pair_event, nullptr, begin, end,
                                       lane);
  for (int base = begin; base < end; base += 64) {
    const PairBatch cur = next;
    if (base + 64 < end)
      next = gather_batch<false>(rec, pair_event, nullptr, base + 64, end, lane);
    const int n_here = min(64, end - base);
    uint32_t sp = 0;
    if (SPANS) {
      const uint32_t by0 = (uint32_t)lane_int((int)cur.by, 0);
      sp = fetch_span(spans, span_stride, lane_int(cur.e, 0), by0,
                      y_base + (int)(by0 & (H - 1)) + r);
    }
    for (int j = 0; j < n_here; ++j) {
      const PairView v = pick(cur, j);
      // this lane's row of the pair's band and its row in the window
      const int wrow = (v.y0 & (H - 1)) + r;
      const int iy = y_base + wrow;
      RowSpan s;
      if (SPANS) {
        const uint32_t word = (iy >= v.y0 && iy < v.y1) ? sp : 0u;
        __builtin_amdgcn_sched_barrier(0);
        const int j1 = min(j + 1, n_here - 1);
        const uint32_t by1 = (uint32_t)lane_int((int)cur.by, j1);
        sp = fetch_span(spans, span_stride, lane_int(cur.e, j1), by1,
                        y_base + (int)(by1 & (H - 1)) + r);
        s = stored_span(g, word, v.cy, v.A, v.B, v.P, v.Q0, v.x0, iy);
      } else {
        s = row_span(g, v.cz, v.cy, v.A, v.B, v.P, v.Q0, v.x0, v.x1, v.y0,
                     v.y1, iy);
      }
      const int nq = (s.n > q) ? (s.n - q + LR - 1) / LR : 0;
      const int col0 = s.lo + q;
      // offset from the row's vertex (the pixel centre first, then the
      // difference, as the reference rounds them), and the exponent's
      // coefficients
      float dz = (__builtin_fmaf((float)col0, g.pixel_width, g.tl_z_half_w) -
                  v.cz) -
                 s.tdy;
      const float qa = s.qd * kExp2Scale;
      const float qb = 0.0f;
      const float db = __builtin_fmaf(qc * step, step, qb * step);
      const float* src0 = win + (col0 - c0) * WH + wrow;
      const float* src1 = src0 + n_words;
      float a0 = 0.f, aq0 = 0.f, a1 = 0.f, aq1 = 0.f;
      int k = 0;
      // e = exp2(q) by the recurrence of the forward pass, q by its first and
      // second differences; both restarted exactly every 8 columns
      for (; k + 8 <= nq; k += 8) {
        float qq = __builtin_fmaf(__builtin_fmaf(qc, dz, qb), dz, qa);
        float dq = __builtin_fmaf(two_qc_step, dz, db);
        float e = __builtin_amdgcn_exp2f(qq);
        float rr = __builtin_amdgcn_exp2f(dq);
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) {
          const float t0 = e * src0[(k + kk) * LR * WH];
          a0 += t0;
          aq0 = __builtin_fmaf(t0, qq, aq0);
          if (NIMG == 2) {
            const float t1 = e * src1[(k + kk) * LR * WH];
            a1 += t1;
            aq1 = __builtin_fmaf(t1, qq, aq1);
          }
          e *= rr;
          rr *= G;
          qq += dq;
          dq += d2q;
        }
        dz += 8.0f * step;
      }
      for (; k < nq; ++k) {
        const float qq = __builtin_fmaf(__builtin_fmaf(qc, dz, qb), dz, qa);
        const float e = __builtin_amdgcn_exp2f(qq);
        const float t0 = e * src0[k * LR * WH];
        a0 += t0;
        aq0 = __builtin_fmaf(t0, qq, aq0);
        if (NIMG == 2) {
          const float t1 = e * src1[k * LR * WH];
          a1 += t1;
          aq1 = __builtin_fmaf(t1, qq, aq1);
        }
        dz += step;
      }
      // row constants of the sigma derivatives
      const float py = pixel_center_y(g, iy);
      const float dy = py - v.cy;
      const float yp = dy + v.cy;
      const float w = v.Q0 * __builtin_fmaf(yp * yp, inv_r2, 1.0f) * s.pre *
                      s.pre;
      const float wq = w - s.qd;
      const float R0 = s.pre * a0;
      const float m00 = wave_total(R0);
      const float m01 = wave_total(s.pre * aq0 * inv_scale);
      const float m02 = wave_total(R0 * wq);
      float m10 = 0.f, m11 = 0.f, m12 = 0.f;
      if (NIMG == 2) {
        const float R1 = s.pre * a1;
        m10 = wave_total(R1);
        m11 = wave_total(s.pre * aq1 * inv_scale);
        m12 = wave_total(R1 * wq);
      }
      if (lane == 0) {
        float* out = moments + ((int64_t)v.e * max_bands + (b - v.y0 / H)) *
                                   (kMoments * NIMG);
        out[0] = m00;
        out[1] = m01;
        out[2] = m02;
        if (NIMG == 2) {
          out[3] = m10;
          out[4] = m11;
          out[5] = m12;
        }
      }
    }
  }
}

// Per event: sums the band moments and forms
//   NIMG 1 (log-likelihood): w1 = 1/D, value = log D,
//     gz = sum K rho (Q + w - Qd) / D - 2, gd = sum K rho (Qd - w) / D;
//   NIMG 2 (MLEM reverse step, x0 = rho_k, x1 = gbar * rho_k):
//     w1 = 1/D, w2 = G/D^2 with D = sum K x0, G = sum K x1,
//     gz = (Z1 - Z0 G/D) / D, gd = (Y1 - Y0 G/D) / D
//     (the -2 terms of d ln K / d ln sigma_z cancel here).
// Events with D <= 0 contribute nothing (the reference's guard).
template <int NIMG>
__global__ __launch_bounds__(256) void adjoint_event_kernel(
    const BandRecord* __restrict__ rec, int64_t n, int H, int max_bands,
    const float* __restrict__ moments, float* __restrict__ w1,
    float* __restrict__ w2, float* __restrict__ value, float* __restrict__ gz,
    float* __restrict__ gd) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int x0, x1, y0, y1;
  record_box(rec[e], x0, x1, y0, y1);
  float m[kMoments * 2] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const int nb = band_count(y0, y1, x1 > x0, H, true);
  for (int k = 0; k < nb; ++k) {
    const float* src = moments + (e * max_bands + k) * (kMoments * NIMG);
#pragma unroll
    for (int i = 0; i < kMoments * NIMG; ++i) m[i] += src[i];
  }
  const float D = m[0];
  float v1 = 0.f, v2 = 0.f, v = 0.f, vz = 0.f, vd = 0.f;
  if (D > 0.0f) {
    const float inv = 1.0f / D;
    v1 = inv;
    if (NIMG == 1) {
      v = logf(D);
      vz = (m[1] + m[2]) * inv - 2.0f;
      vd = -m[2] * inv;
    } else {
      const float Gr = m[3] * inv;  // G / D
      v2 = Gr * inv;
      vz = ((m[4] + m[5]) - (m[1] + m[2]) * Gr) * inv;
      vd = (-m[5] + m[2] * Gr) * inv;
    }
  }
  w1[e] = v1;
  if (w2) w2[e] = v2;
  if (value) value[e] = v;
  gz[e] = vz;
  gd[e] = vd;
}

// out = a * b (b may be null: out = a); n elements
__global__ __launch_bounds__(256) void mul_kernel(int64_t n,
                                                  const float* __restrict__ a,
                                                  const float* __restrict__ b,
                                                  float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = b ? a[i] * b[i] : a[i];
}

// gbar = gbar * bp1 - bp2
__global__ __launch_bounds__(256) void reverse_combine_kernel(
    int64_t n, const float* __restrict__ bp1, const float* __restrict__ bp2,
    float* __restrict__ gbar) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    gbar[i] = __builtin_fmaf(gbar[i], bp1[i], -bp2[i]);
}

// out = bp - 1
__global__ __launch_bounds__(256) void minus_one_kernel(
    int64_t n, const float* __restrict__ bp, float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = bp[i] - 1.0f;
}

// gbar = 2 (rho - target) / n_pixels, sq = (rho - target)^2
__global__ __launch_bounds__(256) void mse_seed_kernel(
    int64_t n, const float* __restrict__ rho, const float* __restrict__ target,
    float scale, float* __restrict__ gbar, float* __restrict__ sq) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float d = rho[i] - target[i];
    gbar[i] = scale * d;
    sq[i] = d * d;
  }
}

// Deterministic FP64 sum of n floats: kSumBlocks block partials (each block
// a fixed grid-stride share, summed in LDS in a fixed order), then one block
// adds the partials in index order into *out.  No atomics, so repeated calls
// give the same bits.
constexpr int kSumBlocks = 512;

__global__ __launch_bounds__(256) void sum_partials_kernel(
    int64_t n, const float* __restrict__ in, double* __restrict__ partials) {
  __shared__ double part[256];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    acc += (double)in[i];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = part[0];
}

__global__ __launch_bounds__(256) void sum_final_kernel(
    const double* __restrict__ partials, int n_partials,
    double* __restrict__ out) {
  __shared__ double part[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n_partials; i += blockDim.x) acc += partials[i];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = part[0];
}

__global__ __launch_bounds__(256) void zero_kernel(int64_t n,
                                                   float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = 0.0f;
}

unsigned grid_for(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096));
}

template <typename T>
void free_dev(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

int ensure_work(AdjointWork& wk, const BandPlan& plan, int64_t n_pixels,
                std::string* err) {
  const int64_t n = std::max<int64_t>(plan.n_events, 1);
  const int64_t m = n * plan.fwd.max_bands * kMoments * 2;
  if (m > wk.moments_cap) {
    free_dev(wk.moments);
    BAND_TRY(hipMalloc(&wk.moments, sizeof(float) * m));
    wk.moments_cap = m;
  }
  if (n > wk.ev_cap) {
    free_dev(wk.ev_buf);
    BAND_TRY(hipMalloc(&wk.ev_buf, sizeof(float) * 5 * n));
    wk.ev_cap = n;
  }
  if (n_pixels > wk.img_cap) {
    free_dev(wk.img);
    BAND_TRY(hipMalloc(&wk.img, sizeof(float) * 3 * n_pixels));
    wk.img_cap = n_pixels;
  }
  if (!wk.sums)
    BAND_TRY(hipMalloc(&wk.sums, sizeof(double) * (kSumSlots + kSumBlocks)));
  return 0;
}

// Device FP64 sum of n floats into wk.sums[slot] (deterministic).
int reduce_sum(AdjointWork& wk, const float* in, int64_t n, int slot,
               hipStream_t stream, std::string* err) {
  if (!wk.sums || slot < 0 || slot >= kSumSlots) {
    if (err) *err = "adjoint reduction buffer not prepared";
    return -1;
  }
  double* partials = wk.sums + kSumSlots;
  hipLaunchKernelGGL(sum_partials_kernel, dim3(kSumBlocks), dim3(256), 0,
                     stream, n, in, partials);
  hipLaunchKernelGGL(sum_final_kernel, dim3(1), dim3(256), 0, stream,
                     (const double*)partials, kSumBlocks, wk.sums + slot);
  BAND_TRY(hipGetLastError());
  return 0;
}

template <int NIMG>
int launch_moments(BandPlan& plan, const DeviceGeometry& g, AdjointWork& wk,
                   const float* x0T, const float* x1T, hipStream_t stream,
                   std::string* err) {
  const PairList& L = plan.fwd;
  if (L.n_tasks == 0) return 0;
  const size_t lds = (size_t)L.Wt * 2 * L.H * sizeof(float) * NIMG;
  if (lds > 160 * 1024) {
    if (err) *err = "moment windows do not fit in LDS";
    return -1;
  }
  auto pick_kernel = [&](auto h_tag) {
    constexpr int H = decltype(h_tag)::value;
    return plan.spans ? band_moments_kernel<H, NIMG, true>
                      : band_moments_kernel<H, NIMG, false>;
  };
  auto kern = L.H == 16   ? pick_kernel(std::integral_constant<int, 16>{})
              : L.H == 32 ? pick_kernel(std::integral_constant<int, 32>{})
                          : pick_kernel(std::integral_constant<int, 64>{});
  BAND_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
  hipLaunchKernelGGL(kern, dim3(L.n_tasks), dim3(64 * kFwdWaves), lds, stream,
                     g, plan.records, L.pair_event, L.tasks, plan.spans,
                     plan.span_stride, x0T, x1T, L.Wt, wk.moments,
                     L.max_bands);
  BAND_TRY(hipGetLastError());
  return 0;
}

}  // namespace

void adjoint_release(AdjointWork& wk) {
  free_dev(wk.moments);
  free_dev(wk.ev_buf);
  free_dev(wk.img);
  free_dev(wk.sums);
  wk = AdjointWork{};
}

int band_loglik_terms(BandPlan& plan, const DeviceGeometry& g,
                      AdjointWork& wk, const float* rhoT, float* bp_accum,
                      double* sum_log, double grad_ln_sigma[2],
                      hipStream_t stream, std::string* err) {
  if (!plan.ready) {
    if (err) *err = "band plan not prepared";
    return -1;
  }
  const int64_t n_pixels = (int64_t)g.n_z * g.n_y;
  if (ensure_work(wk, plan, n_pixels, err)) return -1;
  const int64_t n = plan.n_events;
  *sum_log = 0.0;
  if (n == 0) return 0;
  float* w1 = wk.ev_buf;
  float* val = wk.ev_buf + 2 * wk.ev_cap;
  float* gz = wk.ev_buf + 3 * wk.ev_cap;
  float* gd = wk.ev_buf + 4 * wk.ev_cap;
  if (launch_moments<1>(plan, g, wk, rhoT, nullptr, stream, err)) return -1;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(adjoint_event_kernel<1>, dim3(blocks), dim3(256), 0,
                     stream, plan.records, n, plan.fwd.H, plan.fwd.max_bands,
                     wk.moments, w1, nullptr, val, gz, gd);
  BAND_TRY(hipGetLastError());
  if (reduce_sum(wk, val, n, 0, stream, err) ||
      reduce_sum(wk, gz, n, 1, stream, err) ||
      reduce_sum(wk, gd, n, 2, stream, err))
    return -1;
  if (bp_accum && band_backproject(plan, g, w1, bp_accum, stream, err))
    return -1;
  double sums[3] = {0.0, 0.0, 0.0};
  BAND_TRY(hipMemcpyAsync(sums, wk.sums, sizeof(double) * 3,
                          hipMemcpyDeviceToHost, stream));
  BAND_TRY(hipStreamSynchronize(stream));
  *sum_log = sums[0];
  grad_ln_sigma[0] += sums[1];
  grad_ln_sigma[1] += sums[2];
  return 0;
}

int band_image_total(const DeviceGeometry& g, AdjointWork& wk,
                     const float* imgT, double* total, hipStream_t stream,
                     std::string* err) {
  const int64_t n = (int64_t)g.n_z * g.n_y;
  if (!wk.sums)
    BAND_TRY(hipMalloc(&wk.sums, sizeof(double) * (kSumSlots + kSumBlocks)));
  if (reduce_sum(wk, imgT, n, 4, stream, err)) return -1;
  double host = 0.0;
  BAND_TRY(hipMemcpyAsync(&host, wk.sums + 4, sizeof(double),
                          hipMemcpyDeviceToHost, stream));
  BAND_TRY(hipStreamSynchronize(stream));
  *total = host;
  return 0;
}

int band_loglik_gradient(BandPlan& plan, const DeviceGeometry& g,
                         AdjointWork& wk, const float* rhoT, float* gradT,
                         double* value, double grad_ln_sigma[2],
                         hipStream_t stream, std::string* err) {
  const int64_t n_pixels = (int64_t)g.n_z * g.n_y;
  if (ensure_work(wk, plan, n_pixels, err)) return -1;
  if (gradT) {
    hipLaunchKernelGGL(zero_kernel, dim3(grid_for(n_pixels)), dim3(256), 0,
                       stream, n_pixels, gradT);
    BAND_TRY(hipGetLastError());
  }
  double sum_log = 0.0, total = 0.0;
  grad_ln_sigma[0] = grad_ln_sigma[1] = 0.0;
  if (band_loglik_terms(plan, g, wk, rhoT, gradT, &sum_log, grad_ln_sigma,
                        stream, err) ||
      band_image_total(g, wk, rhoT, &total, stream, err))
    return -1;
  if (gradT) {
    hipLaunchKernelGGL(minus_one_kernel, dim3(grid_for(n_pixels)), dim3(256),
                       0, stream, n_pixels, gradT, gradT);
    BAND_TRY(hipGetLastError());
  }
  BAND_TRY(hipStreamSynchronize(stream));
  if (value) *value = sum_log - total;
  return 0;
}

int band_mse_seed(const DeviceGeometry& g, AdjointWork& wk, const float* rhoT,
                  const float* targetT, float* gbarT, double* loss,
                  hipStream_t stream, std::string* err) {
  const int64_t n_pixels = (int64_t)g.n_z * g.n_y;
  if (n_pixels > wk.img_cap || !wk.sums) {
    if (err) *err = "adjoint work buffers not prepared";
    return -1;
  }
  float* sq = wk.img;
  hipLaunchKernelGGL(mse_seed_kernel, dim3(grid_for(n_pixels)), dim3(256), 0,
                     stream, n_pixels, rhoT, targetT, 2.0f / (float)n_pixels,
                     gbarT, sq);
  BAND_TRY(hipGetLastError());
  if (reduce_sum(wk, sq, n_pixels, 0, stream, err)) return -1;
  double total = 0.0;
  BAND_TRY(hipMemcpyAsync(&total, wk.sums, sizeof(double),
                          hipMemcpyDeviceToHost, stream));
  BAND_TRY(hipStreamSynchronize(stream));
  if (loss) *loss = total / (double)n_pixels;
  return 0;
}

int band_adjoint_prepare(BandPlan& plan, const DeviceGeometry& g,
                         AdjointWork& wk, std::string* err) {
  return ensure_work(wk, plan, (int64_t)g.n_z * g.n_y, err);
}

int band_mlem_adjoint_step(BandPlan& plan, const DeviceGeometry& g,
                           AdjointWork& wk, const float* rhoT_k, float* gbarT,
                           double grad_ln_sigma[2], hipStream_t stream,
                           std::string* err, const float* bpT_k) {
  if (!plan.ready) {
    if (err) *err = "band plan not prepared";
    return -1;
  }
  const int64_t n_pixels = (int64_t)g.n_z * g.n_y;
  if (ensure_work(wk, plan, n_pixels, err)) return -1;
  const int64_t n = plan.n_events;
  float* hT = wk.img;
  float* bp1 = wk.img + wk.img_cap;
  float* bp2 = wk.img + 2 * wk.img_cap;
  float* w1 = wk.ev_buf;
  float* w2 = wk.ev_buf + wk.ev_cap;
  float* gz = wk.ev_buf + 3 * wk.ev_cap;
  float* gd = wk.ev_buf + 4 * wk.ev_cap;
  const unsigned pg = grid_for(n_pixels);
  hipLaunchKernelGGL(mul_kernel, dim3(pg), dim3(256), 0, stream, n_pixels,
                     gbarT, rhoT_k, hT);
  BAND_TRY(hipGetLastError());
  if (!bpT_k)
    hipLaunchKernelGGL(zero_kernel, dim3(pg), dim3(256), 0, stream, n_pixels,
                       bp1);
  hipLaunchKernelGGL(zero_kernel, dim3(pg), dim3(256), 0, stream, n_pixels,
                     bp2);
  BAND_TRY(hipGetLastError());
  double sums[2] = {0.0, 0.0};
  if (n > 0) {
    if (launch_moments<2>(plan, g, wk, rhoT_k, hT, stream, err)) return -1;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(adjoint_event_kernel<2>, dim3(blocks), dim3(256), 0,
                       stream, plan.records, n, plan.fwd.H,
                       plan.fwd.max_bands, wk.moments, w1, w2, nullptr, gz,
                       gd);
    BAND_TRY(hipGetLastError());
    if (reduce_sum(wk, gz, n, 0, stream, err) ||
        reduce_sum(wk, gd, n, 1, stream, err))
      return -1;
    if ((!bpT_k && band_backproject(plan, g, w1, bp1, stream, err)) ||
        band_backproject(plan, g, w2, bp2, stream, err))
      return -1;
    BAND_TRY(hipMemcpyAsync(sums, wk.sums, sizeof(double) * 2,
                            hipMemcpyDeviceToHost, stream));
  }
  hipLaunchKernelGGL(reverse_combine_kernel, dim3(pg), dim3(256), 0, stream,
                     n_pixels, bpT_k ? bpT_k : bp1, bp2, gbarT);
  BAND_TRY(hipGetLastError());
  BAND_TRY(hipStreamSynchronize(stream));
  grad_ln_sigma[0] += sums[0];
  grad_ln_sigma[1] += sums[1];
  return 0;
}

}  // namespace strip
