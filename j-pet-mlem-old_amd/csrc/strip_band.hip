// Row-band algorithm of the strip MLEM iteration (kernels + host driver).
// The algorithm and its data layout are described in strip_band.hpp.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "strip_band.hpp"
#include "strip_band_device.hpp"
#include "strip_device.hpp"

namespace strip {
namespace {

// ---------------------------------------------------------------------------
// pair list construction
// ---------------------------------------------------------------------------

// total += sum of counts, in 64 bits (the pair offsets are 32-bit)
__global__ __launch_bounds__(256) void sum_counts_kernel(
    const int* __restrict__ counts, int64_t n,
    unsigned long long* __restrict__ total) {
  unsigned long long mine = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    mine += (unsigned long long)counts[i];
  for (int o = 32; o > 0; o >>= 1) mine += __shfl_xor(mine, o);
  if ((threadIdx.x & 63) == 0 && mine) atomicAdd(total, mine);
}

__global__ __launch_bounds__(256) void count_pairs_kernel(
    const BandRecord* __restrict__ rec, int64_t n, int H, int aligned,
    int* __restrict__ counts) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int x0, x1, y0, y1;
  record_box(rec[e], x0, x1, y0, y1);
  counts[e] = band_count(y0, y1, x1 > x0, H, aligned != 0);
}

// bounds[e * max_bands + k] = lo | hi << 16: the columns [lo, hi) reached
// by the runs of band k of record e (the span table's runs, or the bbox
// columns without one; an empty slice gives [x0, x0 + 1)).  One wavefront
// per record (records in grid strides: a launch counts its work-items in 32
// bits), lane = row offset from the start of its first band, so a band's
// rows are H consecutive lanes (H divides 64).
__global__ __launch_bounds__(256) void band_bounds_kernel(
    const BandRecord* __restrict__ rec, const uint16_t* __restrict__ spans,
    int stride, int64_t n, int H, int aligned, int max_bands,
    int* __restrict__ bounds) {
  const int lane = threadIdx.x & 63;
  for (int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); e < n;
       e += (int64_t)gridDim.x * 4) {
  int x0, x1, y0, y1;
  record_box(rec[e], x0, x1, y0, y1);
  if (x1 <= x0 || y1 <= y0) continue;  // uniform over the wavefront
  const int row0 = band_first_row(y0, y0 / H, H, aligned != 0);
  const int rows = band_count(y0, y1, true, H, aligned != 0) * H;
  for (int base = 0; base < rows; base += 64) {
    const int i = base + lane;
    const int iy = row0 + i;
    int lo = 0x7fff, hi = 0;
    if (i < rows && iy >= y0 && iy < y1) {
      if (spans) {
        const uint32_t w = spans[e * stride + (iy - (y0 & ~(kSpanAlign - 1)))];
        if (w >> 8) {
          lo = x0 + (int)(w & 0xff);
          hi = lo + (int)(w >> 8);
        }
      } else {
        lo = x0;
        hi = x1;
      }
    }
    for (int o = 1; o < H; o <<= 1) {
      lo = min(lo, __shfl_xor(lo, o));
      hi = max(hi, __shfl_xor(hi, o));
    }
    if ((lane & (H - 1)) == 0 && i < rows) {
      if (hi <= lo) {
        lo = x0;
        hi = x0 + 1;
      }
      bounds[e * max_bands + i / H] = lo | (hi << 16);
    }
  }
  }
}

// keys of the event's pairs: band << 16 | (first column the band's runs
// reach, rounded down to the 2^shift-column bucket) | a width code in the low
// shift bits, so a bucket's pairs sort by the width of their run slices and
// the pairs a wavefront step takes together have similar trip counts
__global__ __launch_bounds__(256) void emit_pairs_kernel(
    const BandRecord* __restrict__ rec, const int* __restrict__ bounds,
    int max_bands, int64_t n, int H, int shift,
    const int* __restrict__ counts, const int* __restrict__ offsets,
    uint32_t* __restrict__ keys, int* __restrict__ values) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int c = counts[e];
  if (!c) return;
  const int b0 = (int)(rec[e].by & 0xffff) / H;
  const int base = offsets[e];
  const uint32_t code_max = (1u << shift) - 1;
  for (int k = 0; k < c; ++k) {
    const uint32_t bd = (uint32_t)bounds[e * max_bands + k];
    const uint32_t lo = bd & 0xffff, width = (bd >> 16) - lo;
    const uint32_t code = min(width >> (7 - shift), code_max);
    keys[base + k] = ((uint32_t)(b0 + k) << 16) | (lo & ~code_max) | code;
    values[base + k] = (int)e;
  }
}

// extent[e] = clamped bbox width (rows = 0) or its rows rounded out to
// multiples of kSpanAlign (rows = 1: the record's span-table slot)
__global__ __launch_bounds__(256) void box_extent_kernel(
    const uint2* __restrict__ box, int64_t n, int rows,
    int* __restrict__ extent) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint2 b = box[e];
  const uint32_t v = rows ? b.y : b.x;
  const int lo = (int)(v & 0xffff), hi = (int)(v >> 16);
  if (!rows || hi <= lo)
    extent[e] = max(hi - lo, 0);
  else
    extent[e] = ((hi + kSpanAlign - 1) & ~(kSpanAlign - 1)) -
                (lo & ~(kSpanAlign - 1));
}

// spatial order of the events: key = (bbox y0, bbox x0)
__global__ __launch_bounds__(256) void order_keys_kernel(
    const uint2* __restrict__ box, int64_t n, uint32_t* __restrict__ keys,
    int* __restrict__ values) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint2 b = box[e];
  keys[e] = ((b.y & 0xffff) << 16) | (b.x & 0xffff);
  values[e] = (int)e;
}

// rec[i] = packed parameters of event order[i]
__global__ __launch_bounds__(256) void pack_records_kernel(
    EventSoA ev, const int* __restrict__ order, int64_t n,
    BandRecord* __restrict__ rec) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int e = order[i];
  const float4 a = ev.a[e];
  const float4 b = ev.b[e];
  const float4 c = ev.c[e];
  const uint2 box = ev.box[e];
  BandRecord r;
  r.cz = a.x;
  r.cy = a.y;
  r.A = b.x;
  r.B = b.y;
  r.P = c.x;
  r.Q0 = c.y;
  r.bx = box.x;
  r.by = box.y;
  rec[i] = r;
}

// out[order[i]] = in[i]
__global__ __launch_bounds__(256) void unpermute_kernel(
    const float* __restrict__ in, const int* __restrict__ order, int64_t n,
    float* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[order[i]] = in[i];
}

// Span table: one wavefront per record, lanes over the rows of its slot;
// spans[i * stride + k] = (lo - x0) | n << 8 of row y0a + k, y0a = y0
// rounded down to a multiple of kSpanAlign (0 outside [y0, y1); stride >=
// every slot, lo - x0 and n below 256).
__global__ __launch_bounds__(256) void span_table_kernel(
    DeviceGeometry g, const BandRecord* __restrict__ rec, int64_t n,
    int stride, uint16_t* __restrict__ spans) {
  const int lane = threadIdx.x & 63;
  // records in grid strides (a launch counts its work-items in 32 bits)
  for (int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); i < n;
       i += (int64_t)gridDim.x * 4) {
  const BandRecord r = rec[i];
  int x0, x1, y0, y1;
  record_box(r, x0, x1, y0, y1);
  uint16_t* out = spans + i * stride;
  const int y0a = y0 & ~(kSpanAlign - 1);
  for (int k = lane; k < stride; k += 64) {
    const int iy = y0a + k;
    uint16_t word = 0;
    if (iy >= y0 && iy < y1) {
      const RowSpan s =
          row_span(g, r.cz, r.cy, r.A, r.B, r.P, r.Q0, x0, x1, y0, y1, iy);
      if (s.n > 0) word = (uint16_t)((s.lo - x0) | (s.n << 8));
    }
    out[k] = word;
  }
  }
}

// first[k] = index of the first sorted pair whose (band, x0 bucket) key is k
__global__ __launch_bounds__(256) void group_bounds_kernel(
    const uint32_t* __restrict__ keys, int64_t n, int shift,
    int* __restrict__ first) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i] >> shift;
  if (i == 0 || (keys[i - 1] >> shift) != k) first[k] = (int)i;
}

// task.w |= (columns its pairs reach, counted from its first column) << 16,
// at most Wt: one wavefront per task
__global__ __launch_bounds__(256) void task_extent_kernel(
    const BandRecord* __restrict__ rec, const int* __restrict__ bounds,
    int max_bands, int H, const int* __restrict__ pair_event,
    int4* __restrict__ tasks, int n_tasks, int Wt) {
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= n_tasks) return;  // uniform over the wavefront
  const int lane = threadIdx.x & 63;
  const int4 task = tasks[t];
  int reach = 1;
  for (int p = task.y + lane; p < task.z; p += 64) {
    const int e = pair_event[p];
    const int k = (task.x & 0xffff) - (int)(rec[e].by & 0xffff) / H;
    const int hi = (int)((uint32_t)bounds[(int64_t)e * max_bands + k] >> 16);
    reach = max(reach, hi - task.w);
  }
  for (int o = 32; o > 0; o >>= 1) reach = max(reach, __shfl_xor(reach, o));
  if (lane == 0) tasks[t].w = task.w | (min(reach, Wt) << 16);
}

__global__ __launch_bounds__(256) void zero_ints_kernel(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
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
// (scale != 1: the update is multiplied by scale in FP64, the count ratio of
// an ordered-subset update)
__global__ __launch_bounds__(256) void band_finalize_kernel(
    int64_t n, const float* __restrict__ rho, float* __restrict__ bp,
    const float* __restrict__ sens, int divide, double scale,
    float* __restrict__ next, float* __restrict__ bp_out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (bp_out) bp_out[i] = bp[i];
    float v = rho[i] * bp[i];
    if (divide) v = v / sens[i];
    if (scale != 1.0) v = (float)((double)v * scale);
    next[i] = v;
    bp[i] = 0.0f;
  }
}

// ---------------------------------------------------------------------------
// iteration kernels
// ---------------------------------------------------------------------------

// D_e = sum of the event's band partials; inv_e = 1/D_e (guarded or not).
__global__ __launch_bounds__(256) void band_inverse_kernel(
    const BandRecord* __restrict__ rec, int64_t n, int H, int max_bands,
    const float* __restrict__ partial, int guard, float* __restrict__ inv,
    float* __restrict__ denominators) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int x0, x1, y0, y1;
  record_box(rec[e], x0, x1, y0, y1);
  float d = 0.0f;
  const int nb = band_count(y0, y1, x1 > x0, H, true);
  for (int k = 0; k < nb; ++k) d += partial[e * max_bands + k];
  if (denominators) denominators[e] = d;
  if (inv) inv[e] = guard ? ((d > 0.0f) ? 1.0f / d : 0.0f) : 1.0f / d;
}

// Forward projection.  Workgroup = one task (pairs of one band and one x0
// bucket); the pairs' bands are aligned to their events (a pair's H rows
// start at the event's bbox top row, plus a multiple of H, inside image band
// b), so the rho window holds 2H rows from b H, Wt columns from c0, in LDS
// column-major (word = (col - c0) * 2H + row; lanes of a column touch
// distinct rows, hence distinct banks).  Wave w takes a contiguous share of
// the task's pairs, G pairs per step: lane = g*(H*LR) + q*H + r covers row r
// of the band of pair g of the step and every LR-th column of the row's span
// from offset q (G = 64 / (H*LR); with G > 1 the lane groups read their
// pairs' fields through the LDS crossbar, so one instruction stream serves G
// pairs).
// MODE 0: partial D per pair; MODE 1: count the span pixels instead.
template <int H, int LR, int MODE, bool REC, bool SPANS>
__global__ __launch_bounds__(64 * kFwdWaves) void band_forward_kernel(
    const BandInstance* __restrict__ desc, const int4* __restrict__ tasks,
    int n_tasks, int* __restrict__ queue,
    unsigned long long* __restrict__ count) {
  constexpr int G = 64 / (H * LR);
  static_assert(G * H * LR == 64, "lane layout");
  extern __shared__ float win[];
  __shared__ int queued;
  const int lane = threadIdx.x & 63;
  const int r = lane & (H - 1);
  const int q = (lane / H) % LR;
  const int grp = lane / (H * LR);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  unsigned long long counted = 0;
  // Resident workgroups walk the task list: from a queue counter (tasks
  // handed out as workgroups finish) or in grid strides.
  for (int t = blockIdx.x;;) {
  if (queue) {
    __syncthreads();  // the previous task's window and index are done with
    if (threadIdx.x == 0) queued = atomicAdd(queue, 1);
    __syncthreads();
    t = queued;
  }
  if (t >= n_tasks) break;
  const int4 task = tasks[t];
  t += gridDim.x;
  const int b = task.x & 0xffff;
  const BandInstance& d = desc[task.x >> 16];
  const DeviceGeometry g = d.g;
  const BandRecord* __restrict__ rec = d.rec;
  const int* __restrict__ pair_event = d.fwd_pairs;
  const uint16_t* __restrict__ spans = d.spans;
  const int span_stride = d.span_stride;
  float* __restrict__ partial = d.partial;
  const int max_bands = d.fwd_max_bands;
  constexpr int WH = 2 * H;  // window rows
  const int c0 = task.w & 0xffff;
  const int y_base = b * H;
  if (MODE == 0) {
    const float* __restrict__ rhoT = d.rhoT;
    // the columns the task's pairs reach
    const int n_words = (task.w >> 16) * WH;
    __syncthreads();  // the previous task's window is no longer read
    for (int i = threadIdx.x; i < n_words; i += blockDim.x) {
      const int col = c0 + i / WH, row = y_base + (i & (WH - 1));
      win[i] = (col < g.n_z && row < g.n_y)
                   ? gload(rhoT + (int64_t)col * g.n_y + row)
                   : 0.0f;
    }
    __syncthreads();
  }
  const float step = (float)LR * g.pixel_width;
  const int total = task.z - task.y;
  // shares of whole steps
  const int share = ((total + kFwdWaves * G - 1) / (kFwdWaves * G)) * G;
  const int begin = task.y + wave * share;
  const int end = min(task.z, begin + share);
  const float qc = g.C * kExp2Scale;
  const float two_qc_step = 2.0f * qc * step;
  PairBatch next = gather_batch<false>(rec, pair_event, nullptr, begin, end,
                                       lane);
  for (int base = begin; base < end; base += 64) {
    const PairBatch cur = next;
    if (base + 64 < end)
      next = gather_batch<false>(rec, pair_event, nullptr, base + 64, end, lane);
    const int n_here = min(64, end - base);
    // the next step's pair and span word are in flight while one step is
    // processed
    // two register sets for a step's pair and span word, used alternately:
    // a step reads one set and fills the other with the next step's
    PairView va = step_pair<G, false>(cur, 0, grp), vb = va;
    uint32_t spa = 0, spb = 0;
    if (SPANS)
      spa = fetch_span(spans, span_stride, va.e, (uint32_t)va.y0,
                       y_base + (va.y0 & (H - 1)) + r);
    auto run_step = [&](int j, const PairView& v, uint32_t sp, PairView& nv,
                        uint32_t& sp_next) {
      const bool valid = G == 1 || j + grp < n_here;
      // this lane's row of the pair's band and its row in the window
      const int wrow = (v.y0 & (H - 1)) + r;
      const int iy = y_base + wrow;
      RowSpan s;
      if (SPANS) {
        // rows past the bbox (the last band's) hold an empty run
        const uint32_t word = (valid && iy < v.y1) ? sp : 0u;
        // the word is consumed before the next load is issued, so that load
        // stays in flight through this step
        __builtin_amdgcn_sched_barrier(0);
        const int jn = G == 1 ? min(j + 1, n_here - 1) : min(j + G, 64 - G);
        nv = step_pair<G, false>(cur, jn, grp);
        sp_next = fetch_span(spans, span_stride, nv.e, (uint32_t)nv.y0,
                             y_base + (nv.y0 & (H - 1)) + r);
        s = stored_span(g, word, v.cy, v.A, v.B, v.P, v.Q0, v.x0, iy);
      } else {
        if (G > 1) nv = step_pair<G, false>(cur, min(j + G, 64 - G), grp);
        s = row_span(g, v.cz, v.cy, v.A, v.B, v.P, v.Q0, v.x0, v.x1, v.y0,
                     v.y1, iy);
        if (!valid) s.n = 0;
      }
      if (G == 1 && !SPANS) nv = pick(cur, min(j + 1, n_here - 1));
      const int nq = (s.n > q) ? (s.n - q + LR - 1) / LR : 0;
      if (MODE == 1) {
        if (G == 1)
          counted += (unsigned long long)wave_sum_int(q == 0 ? s.n : 0);
        else
          counted += (unsigned long long)(q == 0 ? s.n : 0);
        return;
      }
      const int col0 = s.lo + q;
      // offset from the row's vertex (the pixel centre first, then the
      // difference, as the reference rounds them), and the exponent's
      // coefficients
      float dz = (__builtin_fmaf((float)col0, g.pixel_width, g.tl_z_half_w) -
                  v.cz) -
                 s.tdy;
      const float qa = s.qd * kExp2Scale;
      const float qb = 0.0f;
      const float* src = win + (col0 - c0) * WH + wrow;
      float acc = 0.0f;
      int k = 0;
      if (REC) {
        // exp2 of a quadratic along the row: e[j+1] = e[j] r[j] with
        // r[j] = r[0] G2^j, so a block's sum is Horner's scheme
        //   sum_j rho[j] e[j] = e[0] (rho[0] + r[0] (rho[1] + ... r[6] rho[7]))
        // with r[j] taken downward from r[6] (r[j-1] = r[j] / G2): two
        // operations per column; e[0] and r[6] from exp2 in every block of 8
        // (4 in a tail)
        const float db = __builtin_fmaf(qc * step, step, qb * step);
        const float Ginv = __builtin_amdgcn_exp2f(-two_qc_step * step);
        const float db6 = __builtin_fmaf(6.0f * two_qc_step, step, db);
        const float db2 = __builtin_fmaf(2.0f * two_qc_step, step, db);
        constexpr int S = LR * WH;
        for (; k + 8 <= nq; k += 8) {
          const float e0 = __builtin_amdgcn_exp2f(
              __builtin_fmaf(__builtin_fmaf(qc, dz, qb), dz, qa));
          float rr =
              __builtin_amdgcn_exp2f(__builtin_fmaf(two_qc_step, dz, db6));
          const float* p = src + k * S;
          float t = p[7 * S];
#pragma unroll
          for (int i = 6; i >= 0; --i) {
            t = __builtin_fmaf(rr, t, p[i * S]);
            if (i) rr *= Ginv;
          }
          acc = __builtin_fmaf(e0, t, acc);
          dz += 8.0f * step;
        }
        if (k + 4 <= nq) {
          const float e0 = __builtin_amdgcn_exp2f(
              __builtin_fmaf(__builtin_fmaf(qc, dz, qb), dz, qa));
          float rr =
              __builtin_amdgcn_exp2f(__builtin_fmaf(two_qc_step, dz, db2));
          const float* p = src + k * S;
          float t = p[3 * S];
#pragma unroll
          for (int i = 2; i >= 0; --i) {
            t = __builtin_fmaf(rr, t, p[i * S]);
            if (i) rr *= Ginv;
          }
          acc = __builtin_fmaf(e0, t, acc);
          dz += 4.0f * step;
          k += 4;
        }
      }
      for (; k < nq; ++k) {
        const float qq = __builtin_fmaf(__builtin_fmaf(qc, dz, qb), dz, qa);
        acc = __builtin_fmaf(__builtin_amdgcn_exp2f(qq), src[k * LR * WH],
                             acc);
        dz += step;
      }
      if constexpr (G == 1) {
        const float dsum = wave_total(acc * s.pre);
        if (lane == 0)
          gstore(partial + (int64_t)v.e * max_bands + (b - v.y0 / H), dsum);
      } else {
        const float dsum = group_total<H * LR>(acc * s.pre);
        if ((lane & (H * LR - 1)) == 0 && valid)
          gstore(partial + (int64_t)v.e * max_bands + (b - v.y0 / H), dsum);
      }
    };
    for (int j = 0; j < n_here; j += 2 * G) {
      run_step(j, va, spa, vb, spb);
      if (j + G >= n_here) break;
      run_step(j + G, vb, spb, va, spa);
    }
  }
  }
  if (MODE == 1) {
    if (G > 1)
      for (int o = 32; o > 0; o >>= 1) counted += __shfl_xor(counted, o);
    if (lane == 0 && counted) atomicAdd(count, counted);
  }
}

// Back-projection of one task by one wavefront, G pairs per step: G
// private windows of Wt columns x H rows in LDS at wave_win (word = (col -
// c0) * H + row, window k at k * Wt * H), one per lane group: every lane
// writes only its own (row, column) words of its group's window, and no
// other wave sees them, so plain read-modify-write suffices.  lane =
// g*(H*LR) + q*H + r as in the forward projection.  The windows are then
// summed and added to the column-major image.
template <int H, int LR, bool REC, bool SPANS>
__device__ __forceinline__ void backward_task(const BandInstance& d,
                                              const int4 task,
                                              float* const wave_win,
                                              const int lane) {
  constexpr int G = 64 / (H * LR);
  static_assert(G * H * LR == 64, "lane layout");
  const int r = lane & (H - 1);
  const int q = (lane / H) % LR;
  const int grp = lane / (H * LR);
  const int b = task.x & 0xffff;
  const DeviceGeometry g = d.g;
  const BandRecord* __restrict__ rec = d.rec;
  const int* __restrict__ pair_event = d.bwd_pairs;
  const uint16_t* __restrict__ spans = d.spans;
  const int span_stride = d.span_stride;
  const float* __restrict__ inv = d.inv;
  float* __restrict__ bpT = d.bpT;
  const int Wt = task.w >> 16;  // the columns the task's pairs reach
  const int window_words = Wt * H;
  float* const win = wave_win + grp * window_words;
  // each lane group clears its own window
  for (int i = lane & (H * LR - 1); i < Wt * H; i += H * LR) win[i] = 0.0f;
  const int c0 = task.w & 0xffff;
  const int y_base = b * H;
  const int iy = y_base + r;
  constexpr int S = LR * H;  // word stride between a lane's columns
  const float step = (float)LR * g.pixel_width;
  const float qc = g.C * kExp2Scale;
  const float two_qc_step = 2.0f * qc * step;
  const float G2 = __builtin_amdgcn_exp2f(two_qc_step * step);
  PairBatch next = gather_batch<true>(rec, pair_event, inv, task.y, task.z,
                                      lane);
  for (int base = task.y; base < task.z; base += 64) {
    const PairBatch cur = next;
    if (base + 64 < task.z)
      next = gather_batch<true>(rec, pair_event, inv, base + 64, task.z, lane);
    const int n_here = min(64, task.z - base);
    // two register sets for a step's pair and span word, used alternately
    PairView va = step_pair<G, true>(cur, 0, grp), vb = va;
    uint32_t spa = 0, spb = 0;
    if (SPANS) spa = fetch_span(spans, span_stride, va.e, (uint32_t)va.y0, iy);
    auto run_step = [&](int j, const PairView& v, uint32_t sp, PairView& nv,
                        uint32_t& sp_next) {
      const bool valid = G == 1 || j + grp < n_here;
      uint32_t word = 0;
      if (SPANS) {
        // an event with inv = 0 adds nothing: empty run; the word is
        // consumed before the next load is issued
        word = (valid && v.inv != 0.0f) ? sp : 0u;
        __builtin_amdgcn_sched_barrier(0);
        const int jn = G == 1 ? min(j + 1, n_here - 1) : min(j + G, 64 - G);
        nv = step_pair<G, true>(cur, jn, grp);
        sp_next = fetch_span(spans, span_stride, nv.e, (uint32_t)nv.y0, iy);
      } else {
        nv = step_pair<G, true>(
            cur, G == 1 ? min(j + 1, n_here - 1) : min(j + G, 64 - G), grp);
        if (G == 1 && v.inv == 0.0f) return;
      }
      RowSpan s =
          SPANS ? stored_span(g, word, v.cy, v.A, v.B, v.P, v.Q0, v.x0, iy)
                : row_span(g, v.cz, v.cy, v.A, v.B, v.P, v.Q0, v.x0, v.x1,
                           v.y0, v.y1, iy);
      if (!SPANS && G > 1 && (!valid || v.inv == 0.0f)) s.n = 0;
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
      const float scale = s.pre * v.inv;
      float* w = win + (col0 - c0) * H + r;
      int k = 0;
      if (REC) {
        // e[k+1] = e[k] r[k], r[k+1] = r[k] G2, restarted every 8 columns
        // (4 for a tail): the block's words are all read before any is
        // written, so eight reads are in flight together
        const float db = __builtin_fmaf(qc * step, step, qb * step);
        for (; k + 8 <= nq; k += 8) {
          float a[8];
#pragma unroll
          for (int i = 0; i < 8; ++i) a[i] = w[i * S];
          float e = __builtin_amdgcn_exp2f(
              __builtin_fmaf(__builtin_fmaf(qc, dz, qb), dz, qa));
          float rr =
              __builtin_amdgcn_exp2f(__builtin_fmaf(two_qc_step, dz, db));
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            w[i * S] = __builtin_fmaf(e, scale, a[i]);
            e *= rr;
            rr *= G2;
          }
          dz += 8.0f * step;
          w += 8 * S;
        }
        for (; k + 4 <= nq; k += 4) {
          const float a0 = w[0], a1 = w[S], a2 = w[2 * S], a3 = w[3 * S];
          const float e0 = __builtin_amdgcn_exp2f(
              __builtin_fmaf(__builtin_fmaf(qc, dz, qb), dz, qa));
          const float r0 =
              __builtin_amdgcn_exp2f(__builtin_fmaf(two_qc_step, dz, db));
          const float e1 = e0 * r0, r1 = r0 * G2;
          const float e2 = e1 * r1, r2 = r1 * G2;
          const float e3 = e2 * r2;
          w[0] = __builtin_fmaf(e0, scale, a0);
          w[S] = __builtin_fmaf(e1, scale, a1);
          w[2 * S] = __builtin_fmaf(e2, scale, a2);
          w[3 * S] = __builtin_fmaf(e3, scale, a3);
          dz += 4.0f * step;
          w += 4 * S;
        }
      }
      for (; k + 4 <= nq; k += 4) {
        const float a0 = w[0], a1 = w[S], a2 = w[2 * S], a3 = w[3 * S];
        const float dz1 = dz + step, dz2 = dz1 + step, dz3 = dz2 + step;
        const float q0 = __builtin_fmaf(__builtin_fmaf(qc, dz, qb), dz, qa);
        const float q1 = __builtin_fmaf(__builtin_fmaf(qc, dz1, qb), dz1, qa);
        const float q2 = __builtin_fmaf(__builtin_fmaf(qc, dz2, qb), dz2, qa);
        const float q3 = __builtin_fmaf(__builtin_fmaf(qc, dz3, qb), dz3, qa);
        w[0] = __builtin_fmaf(__builtin_amdgcn_exp2f(q0), scale, a0);
        w[S] = __builtin_fmaf(__builtin_amdgcn_exp2f(q1), scale, a1);
        w[2 * S] = __builtin_fmaf(__builtin_amdgcn_exp2f(q2), scale, a2);
        w[3 * S] = __builtin_fmaf(__builtin_amdgcn_exp2f(q3), scale, a3);
        dz = dz3 + step;
        w += 4 * S;
      }
      for (; k < nq; ++k) {
        const float qq = __builtin_fmaf(__builtin_fmaf(qc, dz, qb), dz, qa);
        *w = __builtin_fmaf(__builtin_amdgcn_exp2f(qq), scale, *w);
        dz += step;
        w += S;
      }
    };
    for (int j = 0; j < n_here; j += 2 * G) {
      run_step(j, va, spa, vb, spb);
      if (j + G >= n_here) break;
      run_step(j + G, vb, spb, va, spa);
    }
  }
  // flush: lane = q'*H + r adds row r of columns c0+q', c0+q'+64/H, ...,
  // summed over the wave's G windows
  constexpr int LC = 64 / H;
  const int qc0 = lane / H;
  for (int c = qc0; c < Wt; c += LC) {
    float val = wave_win[c * H + r];
#pragma unroll
    for (int k = 1; k < G; ++k) val += wave_win[k * window_words + c * H + r];
    const int col = c0 + c;
    if (val != 0.0f && iy < g.n_y && col < g.n_z)
      gatomic_add(bpT + (int64_t)col * g.n_y + iy, val);
  }
}

// Back-projection.  Each wavefront takes one task at a time and owns
// wave_words words of LDS; a task runs with the most pairs per step (up to
// max_group, halving) whose G windows of Wt x H words fit there, so narrow
// tasks share the per-pair work among more lane groups than wide ones.
template <int H, bool REC, bool SPANS>
__global__ __launch_bounds__(256) void band_backward_kernel(
    const BandInstance* __restrict__ desc, const int4* __restrict__ tasks,
    int n_tasks, int* __restrict__ queue, int wave_words, int max_group) {
  extern __shared__ float windows[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int n_waves = blockDim.x >> 6;
  float* const wave_win = windows + wave * wave_words;
  // each wavefront walks the task list: from a queue counter or in grid
  // strides of the launch's wavefronts
  for (int t = blockIdx.x * n_waves + wave;;) {
    if (queue) {
      int got = 0;
      if (lane == 0) got = atomicAdd(queue, 1);
      t = __builtin_amdgcn_readfirstlane(got);
    }
    if (t >= n_tasks) break;
    const int4 task = tasks[t];
    t += gridDim.x * n_waves;
    const BandInstance& d = desc[task.x >> 16];
    const int window = (task.w >> 16) * H;
    if constexpr (64 / H >= 4) {
      if (max_group >= 4 && 4 * window <= wave_words) {
        backward_task<H, 64 / (H * 4), REC, SPANS>(d, task, wave_win, lane);
        continue;
      }
    }
    if constexpr (64 / H >= 2) {
      if (max_group >= 2 && 2 * window <= wave_words) {
        backward_task<H, 64 / (H * 2), REC, SPANS>(d, task, wave_win, lane);
        continue;
      }
    }
    backward_task<H, 64 / H, REC, SPANS>(d, task, wave_win, lane);
  }
}

// Raises a kernel's dynamic LDS limit only when a launch needs more than
// was granted before, so launches recorded into a graph make no attribute
// calls once the sizes have been seen.
template <typename Kern>
int set_lds(Kern kernel, size_t bytes, std::string* err) {
  static std::mutex lock;
  static std::vector<std::pair<const void*, size_t>> granted;
  const void* key = reinterpret_cast<const void*>(kernel);
  std::lock_guard<std::mutex> guard(lock);
  size_t* have = nullptr;
  for (auto& kv : granted)
    if (kv.first == key) have = &kv.second;
  if (have && bytes <= *have) return 0;
  BAND_TRY(hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes));
  if (have)
    *have = bytes;
  else
    granted.emplace_back(key, bytes);
  return 0;
}

// Per instance (blockIdx.y): D_e = sum of the record's band partials,
// inv_e = D>0 ? 1/D : 0 (the reference's guard).
__global__ __launch_bounds__(256) void batch_inverse_kernel(
    const BandInstance* __restrict__ desc, int H) {
  const BandInstance& d = desc[blockIdx.y];
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d.n_events) return;
  int x0, x1, y0, y1;
  record_box(d.rec[e], x0, x1, y0, y1);
  float sum = 0.0f;
  const int nb = band_count(y0, y1, x1 > x0, H, true);
  for (int k = 0; k < nb; ++k) sum += d.partial[e * d.fwd_max_bands + k];
  d.inv[e] = (sum > 0.0f) ? 1.0f / sum : 0.0f;
}

// Per instance (blockIdx.y): rho_next = rho * bp, bp = 0.
__global__ __launch_bounds__(256) void batch_finalize_kernel(
    const BandInstance* __restrict__ desc) {
  const BandInstance& d = desc[blockIdx.y];
  const int64_t n = (int64_t)d.g.n_z * d.g.n_y;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    d.rho_next[i] = d.rhoT[i] * d.bpT[i];
    d.bpT[i] = 0.0f;
  }
}

template <int H, int LR>
int launch_forward_t(const BandInstance* desc, const int4* tasks, int n_tasks,
                     bool sp, bool recurrence, int mode, size_t lds, int grid,
                     int* queue, unsigned long long* count,
                     hipStream_t stream, std::string* err) {
  if (mode == 0) {
    auto kern =
        recurrence ? (sp ? band_forward_kernel<H, LR, 0, true, true>
                         : band_forward_kernel<H, LR, 0, true, false>)
                   : (sp ? band_forward_kernel<H, LR, 0, false, true>
                         : band_forward_kernel<H, LR, 0, false, false>);
    if (set_lds(kern, lds, err)) return -1;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * kFwdWaves), lds, stream,
                       desc, tasks, n_tasks, queue, count);
  } else {
    auto kern = sp ? band_forward_kernel<H, LR, 1, false, true>
                   : band_forward_kernel<H, LR, 1, false, false>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * kFwdWaves), 0, stream,
                       desc, tasks, n_tasks, queue, count);
  }
  BAND_TRY(hipGetLastError());
  return 0;
}

template <int H>
int launch_backward_t(const BandInstance* desc, const int4* tasks,
                      int n_tasks, bool sp, bool recurrence, int wave_words,
                      int max_group, int waves, int grid, int* queue,
                      hipStream_t stream, std::string* err) {
  auto kern = recurrence ? (sp ? band_backward_kernel<H, true, true>
                               : band_backward_kernel<H, true, false>)
                         : (sp ? band_backward_kernel<H, false, true>
                               : band_backward_kernel<H, false, false>);
  const size_t lds = sizeof(float) * (size_t)wave_words * waves;
  if (set_lds(kern, lds, err)) return -1;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * waves), lds, stream, desc,
                     tasks, n_tasks, queue, wave_words, max_group);
  BAND_TRY(hipGetLastError());
  return 0;
}

// (band rows, pairs per step) of the forward projection
int launch_forward_any(int H, int G, const BandInstance* desc,
                       const int4* tasks, int n_tasks, bool sp,
                       bool recurrence, int mode, size_t lds, int grid,
                       int* queue, unsigned long long* count,
                       hipStream_t stream, std::string* err) {
  if (n_tasks == 0) return 0;
#define STRIP_FWD(HH, GG)                                                  \
  if (H == HH && G == GG)                                                  \
    return launch_forward_t<HH, 64 / (HH * GG)>(desc, tasks, n_tasks, sp,  \
                                                recurrence, mode, lds, grid, \
                                                queue, count, stream, err);
  STRIP_FWD(64, 1)
  STRIP_FWD(32, 1)
  STRIP_FWD(32, 2)
  STRIP_FWD(16, 1)
  STRIP_FWD(16, 2)
  STRIP_FWD(16, 4)
#undef STRIP_FWD
  if (err) *err = "unsupported forward band rows / pairs per step";
  return -1;
}

// (band rows) of the back-projection; max_group = most pairs per step
int launch_backward_any(int H, int max_group, const BandInstance* desc,
                        const int4* tasks, int n_tasks, bool sp,
                        bool recurrence, int wave_words, int waves,
                        int grid, int* queue, hipStream_t stream,
                        std::string* err) {
  if (n_tasks == 0) return 0;
  if (max_group < 1 || max_group * H > 64) {
    if (err) *err = "unsupported back-projection pairs per step";
    return -1;
  }
#define STRIP_BWD(HH)                                                        \
  if (H == HH)                                                               \
    return launch_backward_t<HH>(desc, tasks, n_tasks, sp, recurrence,       \
                                 wave_words, max_group, waves, grid, queue,  \
                                 stream, err);
  STRIP_BWD(16)
  STRIP_BWD(32)
  STRIP_BWD(64)
#undef STRIP_BWD
  if (err) *err = "unsupported back-projection band rows";
  return -1;
}

int upload_slot(BandPlan& plan, int slot, hipStream_t stream,
                std::string* err) {
  BAND_TRY(hipMemcpyAsync(plan.desc + slot, &plan.desc_host[slot],
                          sizeof(BandInstance), hipMemcpyHostToDevice,
                          stream));
  BAND_TRY(hipStreamSynchronize(stream));
  return 0;
}

int dispatch_forward(BandPlan& plan, const DeviceGeometry& g, int mode,
                     unsigned long long* count, hipStream_t stream,
                     std::string* err) {
  const PairList& L = plan.fwd;
  const bool sp = plan.spans != nullptr;
  const bool rec = plan.tuning.recurrence != 0;
  const BandInstance* desc = plan.desc + plan.curT;
  return launch_forward_any(L.H, plan.fwd_group, desc, L.tasks, L.n_tasks, sp,
                            rec, mode, plan.fwd_lds, L.n_tasks, nullptr,
                            count, stream, err);
}

int dispatch_backward(BandPlan& plan, const DeviceGeometry& g,
                      const float* weights, float* outT, hipStream_t stream,
                      std::string* err) {
  const PairList& L = plan.bwd;
  int slot = plan.curT;
  if (weights != plan.inv || outT != plan.bpT) {
    slot = -1;
    for (int k = 2; k < kDescSlots; ++k)
      if (plan.desc_host[k].inv == weights && plan.desc_host[k].bpT == outT &&
          plan.desc_host[k].rec == plan.records)
        slot = k;
    if (slot < 0) {
      slot = 2 + (plan.n_custom++ % (kDescSlots - 2));
      plan.desc_host[slot] = band_descriptor(plan, g, plan.curT);
      plan.desc_host[slot].inv = const_cast<float*>(weights);
      plan.desc_host[slot].bpT = outT;
      if (upload_slot(plan, slot, stream, err)) return -1;
    }
  }
  const bool sp = plan.spans != nullptr;
  const bool rec = plan.tuning.recurrence != 0;
  const int grid = (L.n_tasks + plan.bwd_waves - 1) / plan.bwd_waves;
  return launch_backward_any(L.H, plan.bwd_group, plan.desc + slot, L.tasks,
                             L.n_tasks, sp, rec, plan.bwd_wave_words,
                             plan.bwd_waves, grid, nullptr, stream, err);
}

template <typename T>
void free_ptr(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void free_list(PairList& L) {
  if (L.owns_pairs) {
    free_ptr(L.pair_event);
    free_ptr(L.bounds);
  }
  L.pair_event = nullptr;
  free_ptr(L.tasks);
  L = PairList{};
}

constexpr size_t kLdsLimit = 160 * 1024;

}  // namespace

// LDS bytes per back-projection wavefront that keep cu_waves wavefronts
// resident per CU in workgroups of four (each workgroup's share a whole
// number of KiB).
size_t backprojection_budget(int cu_waves) {
  const size_t groups = (size_t)std::max(1, (cu_waves + 3) / 4);
  return ((kLdsLimit / groups) & ~(size_t)1023) / 4;
}

namespace {

// Sorted (event, band) pairs of band height H for the resident events, and
// the index of the first pair of every (band, x0 >> shift) group; aligned:
// the bands start at each event's bbox top (forward pair lists).
int build_pairs(PairList& L, int H, bool aligned, int shift,
                const DeviceGeometry& g,
                const BandRecord* rec, const uint16_t* spans, int stride,
                int64_t n, hipStream_t stream, std::vector<int>& first,
                std::string* err) {
  L.H = H;
  L.n_bands = (g.n_y + H - 1) / H;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  int* counts = nullptr;
  int* offsets = nullptr;
  int* d_red = nullptr;
  void* tmp = nullptr;
  BAND_TRY(hipMalloc(&counts, sizeof(int) * std::max<int64_t>(n, 1)));
  BAND_TRY(hipMalloc(&offsets, sizeof(int) * std::max<int64_t>(n, 1)));
  BAND_TRY(hipMalloc(&d_red, sizeof(int)));
  int last_count = 0, last_offset = 0, max_count = 0;
  if (n > 0) {
    hipLaunchKernelGGL(count_pairs_kernel, dim3(blocks), dim3(256), 0, stream,
                       rec, n, H, aligned ? 1 : 0, counts);
    BAND_TRY(hipGetLastError());
    // the pair offsets below are 32-bit: refuse before they could wrap
    unsigned long long* d_total = nullptr;
    unsigned long long total = 0;
    BAND_TRY(hipMalloc(&d_total, sizeof(unsigned long long)));
    BAND_TRY(hipMemsetAsync(d_total, 0, sizeof(unsigned long long), stream));
    hipLaunchKernelGGL(sum_counts_kernel,
                       dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)),
                       dim3(256), 0, stream, counts, n, d_total);
    BAND_TRY(hipGetLastError());
    BAND_TRY(hipMemcpyAsync(&total, d_total, sizeof(total),
                            hipMemcpyDeviceToHost, stream));
    BAND_TRY(hipStreamSynchronize(stream));
    BAND_TRY(hipFree(d_total));
    if (total > 0x7fffffffull) {
      (void)hipFree(counts);
      (void)hipFree(offsets);
      (void)hipFree(d_red);
      if (err) *err = "more than 2^31 (event, band) pairs in one plan";
      return -1;
    }
    size_t t1 = 0, t2 = 0;
    BAND_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, t1, counts, offsets,
                                               (int)n, stream));
    BAND_TRY(hipcub::DeviceReduce::Max(nullptr, t2, counts, d_red, (int)n,
                                       stream));
    BAND_TRY(hipMalloc(&tmp, std::max(t1, t2)));
    BAND_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, t1, counts, offsets,
                                               (int)n, stream));
    BAND_TRY(hipcub::DeviceReduce::Max(tmp, t2, counts, d_red, (int)n,
                                       stream));
    BAND_TRY(hipMemcpyAsync(&last_count, counts + n - 1, sizeof(int),
                            hipMemcpyDeviceToHost, stream));
    BAND_TRY(hipMemcpyAsync(&last_offset, offsets + n - 1, sizeof(int),
                            hipMemcpyDeviceToHost, stream));
    BAND_TRY(hipMemcpyAsync(&max_count, d_red, sizeof(int),
                            hipMemcpyDeviceToHost, stream));
    BAND_TRY(hipStreamSynchronize(stream));
    BAND_TRY(hipFree(tmp));
    tmp = nullptr;
  }
  BAND_TRY(hipFree(d_red));
  const int64_t n_pairs = (int64_t)last_offset + last_count;
  if (n_pairs > 0x7fffffff) {
    if (err) *err = "too many (event, band) pairs";
    return -1;
  }
  L.n_pairs = n_pairs;
  L.max_bands = std::max(max_count, 1);
  const int n_groups = L.n_bands << (16 - shift);
  first.assign(n_groups + 1, -1);
  if (n_pairs > 0) {
    uint32_t* keys = nullptr;
    uint32_t* keys_alt = nullptr;
    int* vals_alt = nullptr;
    BAND_TRY(hipMalloc(&keys, sizeof(uint32_t) * n_pairs));
    BAND_TRY(hipMalloc(&keys_alt, sizeof(uint32_t) * n_pairs));
    BAND_TRY(hipMalloc(&L.pair_event, sizeof(int) * n_pairs));
    BAND_TRY(hipMalloc(&vals_alt, sizeof(int) * n_pairs));
    BAND_TRY(hipMalloc(&L.bounds, sizeof(int) * n * L.max_bands));
    hipLaunchKernelGGL(band_bounds_kernel,
                       dim3((unsigned)std::min<int64_t>((n + 3) / 4, 1 << 20)),
                       dim3(256), 0, stream, rec, spans, stride, n, H,
                       aligned ? 1 : 0, L.max_bands, L.bounds);
    BAND_TRY(hipGetLastError());
    hipLaunchKernelGGL(emit_pairs_kernel, dim3(blocks), dim3(256), 0, stream,
                       rec, L.bounds, L.max_bands, n, H, shift, counts,
                       offsets, keys, L.pair_event);
    BAND_TRY(hipGetLastError());
    int end_bit = 16;
    while ((1 << (end_bit - 16)) < L.n_bands + 1) ++end_bit;
    hipcub::DoubleBuffer<uint32_t> kb(keys, keys_alt);
    hipcub::DoubleBuffer<int> vb(L.pair_event, vals_alt);
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
    if (vb.Current() != L.pair_event) std::swap(L.pair_event, vals_alt);
    BAND_TRY(hipFree(keys_alt));
    BAND_TRY(hipFree(vals_alt));
    int* d_first = nullptr;
    BAND_TRY(hipMalloc(&d_first, sizeof(int) * (n_groups + 1)));
    BAND_TRY(hipMemsetAsync(d_first, 0xff, sizeof(int) * (n_groups + 1),
                            stream));
    const unsigned pb = (unsigned)((n_pairs + 255) / 256);
    hipLaunchKernelGGL(group_bounds_kernel, dim3(pb), dim3(256), 0, stream,
                       keys, n_pairs, shift, d_first);
    BAND_TRY(hipGetLastError());
    BAND_TRY(hipMemcpyAsync(first.data(), d_first,
                            sizeof(int) * (n_groups + 1),
                            hipMemcpyDeviceToHost, stream));
    BAND_TRY(hipStreamSynchronize(stream));
    BAND_TRY(hipFree(d_first));
    BAND_TRY(hipFree(keys));
  }
  BAND_TRY(hipFree(counts));
  BAND_TRY(hipFree(offsets));
  return 0;
}

// Tasks of at most `chunk` consecutive pairs inside one group, each with the
// columns its pairs reach; L.Wt = the widest task.
int build_tasks(PairList& L, int shift, const std::vector<int>& first,
                int chunk, const BandRecord* rec, int Wt, hipStream_t stream,
                std::string* err) {
  const int n_groups = (int)first.size() - 1;
  std::vector<int4> tasks;
  std::vector<int> end(n_groups, 0);
  int next = (int)L.n_pairs;
  for (int k = n_groups - 1; k >= 0; --k) {
    end[k] = next;
    if (first[k] >= 0) next = first[k];
  }
  chunk = std::max(chunk, 1);
  for (int k = 0; k < n_groups; ++k) {
    if (first[k] < 0) continue;
    const int band = k >> (16 - shift);
    const int c0 = (k & ((1 << (16 - shift)) - 1)) << shift;
    for (int p = first[k]; p < end[k]; p += chunk)
      tasks.push_back(make_int4(band, p, std::min(end[k], p + chunk), c0));
  }
  L.n_tasks = (int)tasks.size();
  if (!tasks.empty()) {
    BAND_TRY(hipMalloc(&L.tasks, sizeof(int4) * tasks.size()));
    BAND_TRY(hipMemcpyAsync(L.tasks, tasks.data(), sizeof(int4) * tasks.size(),
                            hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(task_extent_kernel, dim3((L.n_tasks + 3) / 4),
                       dim3(256), 0, stream, rec, L.bounds, L.max_bands,
                       L.H, L.pair_event, L.tasks, L.n_tasks, Wt);
    BAND_TRY(hipGetLastError());
    BAND_TRY(hipMemcpyAsync(tasks.data(), L.tasks, sizeof(int4) * tasks.size(),
                            hipMemcpyDeviceToHost, stream));
    BAND_TRY(hipStreamSynchronize(stream));
  }
  L.Wt = 1;
  for (const int4& t : tasks) L.Wt = std::max(L.Wt, t.w >> 16);
  L.tasks_host = tasks;
  return 0;
}

}  // namespace

bool band_supported_grid(const DeviceGeometry& g) {
  // group keys hold the band index above 16 bits of column
  return g.n_z <= 32767 && g.n_y <= 32767;
}

// Upper bound of the clamped bbox width in columns for any event.  The bbox
// half width is int(bb_z / w + 0.5) with bb_z = 3 / sqrt(C - B^2 / (4A))
// (gaussian_kernel.h:108-128); with u = tan^2, C - B^2 / (4A) =
// 2 iz - 4 u iz^2 / (4 (1 + u) id + 2 u iz) falls monotonically in u towards
// 8 iz id / (4 id + 2 iz), so that limit bounds bb_z for every angle (one
// more column of margin covers the FP32 evaluation).
int band_max_box_width(const DeviceGeometry& g) {
  const double iz = g.inv_pow_sigma_z, id = g.inv_pow_sigma_dl;
  const double k = 8.0 * iz * id / (4.0 * id + 2.0 * iz);
  if (!(k > 0.0)) return g.n_z;
  const double hw = std::floor(3.0 / std::sqrt(k) / g.pixel_width + 0.5) + 1.0;
  return (int)std::min<double>(2.0 * hw, (double)g.n_z);
}

bool band_window_fits(const DeviceGeometry& g) {
  // the narrowest band (16 rows) and the widest bucket (64 columns)
  return (size_t)(64 + band_max_box_width(g)) * 16 * sizeof(float) <=
         kLdsLimit;
}

void band_release(BandPlan& plan) {
  free_list(plan.bwd);
  free_list(plan.fwd);
  free_ptr(plan.partial);
  free_ptr(plan.inv);
  free_ptr(plan.records);
  free_ptr(plan.order);
  free_ptr(plan.spans);
  if (plan.owns_images) {
    free_ptr(plan.rhoT[0]);
    free_ptr(plan.rhoT[1]);
    free_ptr(plan.bpT);
  }
  free_ptr(plan.sensT);
  free_ptr(plan.desc);
  BandTuning keep = plan.tuning;
  plan = BandPlan{};
  plan.tuning = keep;
}

namespace {

int band_prepare_once(BandPlan& plan, const DeviceGeometry& g,
                      const EventSoA& ev, int64_t n_events, hipStream_t stream,
                      bool allow_spans, bool* built_spans, std::string* err) {
  *built_spans = false;
  if (n_events > 0x7fffffff) {
    if (err) *err = "too many events for one band plan";
    return -1;
  }
  const int64_t n_pixels = (int64_t)g.n_z * g.n_y;
  free_list(plan.bwd);
  free_list(plan.fwd);
  free_ptr(plan.partial);
  free_ptr(plan.inv);
  free_ptr(plan.records);
  free_ptr(plan.order);
  free_ptr(plan.spans);
  plan.ready = false;
  plan.n_events = n_events;
  if (!plan.rhoT[0]) {
    BAND_TRY(hipMalloc(&plan.rhoT[0], sizeof(float) * n_pixels));
    BAND_TRY(hipMalloc(&plan.rhoT[1], sizeof(float) * n_pixels));
    BAND_TRY(hipMalloc(&plan.bpT, sizeof(float) * n_pixels));
    BAND_TRY(hipMemsetAsync(plan.bpT, 0, sizeof(float) * n_pixels, stream));
  }
  const int64_t n = n_events;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  // widest and tallest clamped bbox
  int max_extent[2] = {1, 1};
  if (n > 0) {
    int* extent = nullptr;
    int* d_red = nullptr;
    void* tmp = nullptr;
    BAND_TRY(hipMalloc(&extent, sizeof(int) * n));
    BAND_TRY(hipMalloc(&d_red, sizeof(int)));
    size_t wb = 0;
    BAND_TRY(hipcub::DeviceReduce::Max(nullptr, wb, extent, d_red, (int)n,
                                       stream));
    BAND_TRY(hipMalloc(&tmp, wb));
    for (int rows = 0; rows < 2; ++rows) {
      hipLaunchKernelGGL(box_extent_kernel, dim3(blocks), dim3(256), 0,
                         stream, ev.box, n, rows, extent);
      BAND_TRY(hipGetLastError());
      BAND_TRY(hipcub::DeviceReduce::Max(tmp, wb, extent, d_red, (int)n,
                                         stream));
      BAND_TRY(hipMemcpyAsync(&max_extent[rows], d_red, sizeof(int),
                              hipMemcpyDeviceToHost, stream));
      BAND_TRY(hipStreamSynchronize(stream));
      max_extent[rows] = std::max(max_extent[rows], 1);
    }
    BAND_TRY(hipFree(tmp));
    BAND_TRY(hipFree(d_red));
    BAND_TRY(hipFree(extent));
  }
  const int max_width = max_extent[0];
  plan.max_width = max_width;
  // spatial renumbering and packed records
  if (n > 0) {
    uint32_t* keys = nullptr;
    uint32_t* keys_alt = nullptr;
    int* vals_alt = nullptr;
    void* tmp = nullptr;
    BAND_TRY(hipMalloc(&keys, sizeof(uint32_t) * n));
    BAND_TRY(hipMalloc(&keys_alt, sizeof(uint32_t) * n));
    BAND_TRY(hipMalloc(&plan.order, sizeof(int) * n));
    BAND_TRY(hipMalloc(&vals_alt, sizeof(int) * n));
    BAND_TRY(hipMalloc(&plan.records, sizeof(BandRecord) * n));
    hipLaunchKernelGGL(order_keys_kernel, dim3(blocks), dim3(256), 0, stream,
                       ev.box, n, keys, plan.order);
    BAND_TRY(hipGetLastError());
    hipcub::DoubleBuffer<uint32_t> kb(keys, keys_alt);
    hipcub::DoubleBuffer<int> vb(plan.order, vals_alt);
    size_t sort_bytes = 0;
    BAND_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, kb, vb,
                                                (int)n, 0, 32, stream));
    BAND_TRY(hipMalloc(&tmp, sort_bytes));
    BAND_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, kb, vb,
                                                (int)n, 0, 32, stream));
    BAND_TRY(hipStreamSynchronize(stream));
    BAND_TRY(hipFree(tmp));
    if (vb.Current() != plan.order) std::swap(plan.order, vals_alt);
    BAND_TRY(hipFree(keys));
    BAND_TRY(hipFree(keys_alt));
    BAND_TRY(hipFree(vals_alt));
    hipLaunchKernelGGL(pack_records_kernel, dim3(blocks), dim3(256), 0, stream,
                       ev, plan.order, n, plan.records);
    BAND_TRY(hipGetLastError());
  }
  // span table, when its byte runs fit, it takes at most half the free
  // device memory and what the rest of the plan needs still fits beside it:
  // pair lists with their sort buffers (16 B per pair at 16-row bands, a
  // record reaching at most rows / 16 + 2 bands, in both passes) and the
  // per-(event, band) partials
  plan.span_stride = max_extent[1];
  if (n > 0 && allow_spans && plan.tuning.spans && max_width <= 255) {
    const size_t bytes = sizeof(uint16_t) * (size_t)n * plan.span_stride;
    const size_t bands = (size_t)max_extent[1] / 16 + 2;
    const size_t rest = (size_t)n * bands * (2 * 16 + sizeof(float)) +
                        (size_t)n * 16;
    size_t free_bytes = 0, total_bytes = 0;
    BAND_TRY(hipMemGetInfo(&free_bytes, &total_bytes));
    if (bytes <= free_bytes / 2 && bytes + rest <= free_bytes) {
      *built_spans = true;
      BAND_TRY(hipMalloc(&plan.spans, bytes));
      const unsigned span_blocks =
          (unsigned)std::min<int64_t>((n + 3) / 4, 1 << 20);
      hipLaunchKernelGGL(span_table_kernel, dim3(span_blocks), dim3(256), 0,
                         stream, g, plan.records, n, plan.span_stride,
                         plan.spans);
      BAND_TRY(hipGetLastError());
    }
  }
  const BandTuning& tu = plan.tuning;
  const int bucket = (tu.bucket == 8 || tu.bucket == 16 || tu.bucket == 64)
                         ? tu.bucket
                         : 32;
  const int shift = bucket == 8 ? 3 : bucket == 16 ? 4 : (bucket == 64 ? 6 : 5);
  // a task reaches at most its bucket plus the widest bbox (the bound used
  // to pick band heights); its actual window is set by its pairs' runs
  plan.Wt = bucket + max_width;
  // band heights as asked, halved (down to 16 rows) where a task window of
  // the widest bbox would not fit in LDS (the forward window holds 2 Hf
  // rows: its pairs' bands start anywhere in the task's image band)
  int Hf = tu.fwd_rows == 16 ? 16 : (tu.fwd_rows == 32 ? 32 : 64);
  int Hb = tu.bwd_rows == 64 ? 64 : (tu.bwd_rows == 16 ? 16 : 32);
  while (Hf > 16 && (size_t)plan.Wt * 2 * Hf * sizeof(float) > kLdsLimit)
    Hf /= 2;
  while (Hb > 16 && (size_t)plan.Wt * Hb * sizeof(float) > kLdsLimit) Hb /= 2;
  if ((size_t)plan.Wt * 2 * Hf * sizeof(float) > kLdsLimit) {
    if (err) *err = "task window does not fit in LDS";
    return -1;
  }
  const uint16_t* sp = plan.spans;
  const int ss = plan.span_stride;
  std::vector<int> first;
  if (build_pairs(plan.fwd, Hf, true, shift, g, plan.records, sp, ss, n,
                  stream, first, err))
    return -1;
  if (build_tasks(plan.fwd, shift, first, tu.fwd_pairs, plan.records,
                  plan.Wt, stream, err))
    return -1;
  if (build_pairs(plan.bwd, Hb, false, shift, g, plan.records, sp, ss, n,
                  stream, first, err))
    return -1;
  if (build_tasks(plan.bwd, shift, first, tu.bwd_pairs, plan.records,
                  plan.Wt, stream, err))
    return -1;
  // the per-(record, band) column bounds are needed only to cut the tasks
  if (plan.bwd.owns_pairs) free_ptr(plan.bwd.bounds);
  plan.bwd.bounds = nullptr;
  free_ptr(plan.fwd.bounds);
  // pairs per wavefront step: as asked where the lane layout allows it
  // (G * H <= 64)
  plan.fwd_group = std::max(1, std::min(tu.fwd_group, 64 / Hf));
  plan.bwd_group = std::max(1, std::min(tu.bwd_group, 64 / Hb));
  if (plan.fwd_group == 3) plan.fwd_group = 2;
  if (plan.bwd_group == 3) plan.bwd_group = 2;
  plan.fwd_lds = (size_t)plan.fwd.Wt * 2 * Hf * sizeof(float);
  // back-projection: every wavefront owns a fixed LDS budget and runs each
  // task with the most pairs per step (up to bwd_group, halving) whose
  // windows fit in it.  Automatic groups: the budget keeps bwd_cu_waves
  // wavefronts per CU resident (more pairs per step share the per-pair work
  // but need more LDS, and below ~12 resident wavefronts the LDS latency
  // costs more than the shared work saves); where that share cannot give
  // the widest tasks two pairs per step, it is one widest window instead,
  // so more wavefronts stay resident (the 0.5 mm grid's 111-column windows:
  // 20 per CU, 3% faster than 16).  Explicit groups: bwd_group windows of
  // the widest task, so every task takes bwd_group pairs per step.
  const size_t window_bytes = (size_t)plan.bwd.Wt * Hb * sizeof(float);
  size_t budget;
  if (tu.auto_groups) {
    budget = backprojection_budget(tu.bwd_cu_waves);
    if (2 * window_bytes > budget) budget = window_bytes;
  } else {
    while (plan.bwd_group > 1 && window_bytes * plan.bwd_group > kLdsLimit)
      plan.bwd_group /= 2;
    budget = window_bytes * plan.bwd_group;
  }
  plan.bwd_wave_words = (int)(budget / sizeof(float));
  plan.bwd_waves = (int)std::min<size_t>(4, kLdsLimit / budget);
  if (tu.bwd_waves > 0) plan.bwd_waves = std::min(plan.bwd_waves, tu.bwd_waves);
  if (plan.bwd_waves < 1) {
    if (err) *err = "back-projection window does not fit in LDS";
    return -1;
  }
  plan.bwd_lds = budget * plan.bwd_waves;
  if (n > 0) {
    BAND_TRY(hipMalloc(&plan.partial,
                       sizeof(float) * n * plan.fwd.max_bands));
    BAND_TRY(hipMalloc(&plan.inv, sizeof(float) * n));
  }
  // descriptors of the two iteration images; cached custom back-projection
  // descriptors refer to the old records and are dropped
  if (!plan.desc)
    BAND_TRY(hipMalloc(&plan.desc, sizeof(BandInstance) * kDescSlots));
  for (int k = 0; k < kDescSlots; ++k) plan.desc_host[k] = BandInstance{};
  plan.n_custom = 0;
  static std::atomic<int64_t> generations{0};
  plan.generation = ++generations;
  for (int c = 0; c < 2; ++c) plan.desc_host[c] = band_descriptor(plan, g, c);
  BAND_TRY(hipMemcpyAsync(plan.desc, plan.desc_host,
                          sizeof(BandInstance) * kDescSlots,
                          hipMemcpyHostToDevice, stream));
  BAND_TRY(hipStreamSynchronize(stream));
  plan.ready = true;
  return 0;
}

}  // namespace

int band_prepare(BandPlan& plan, const DeviceGeometry& g, const EventSoA& ev,
                 int64_t n_events, hipStream_t stream, std::string* err) {
  std::string first;
  bool spans = false;
  if (!band_prepare_once(plan, g, ev, n_events, stream, true, &spans, &first))
    return 0;
  // a plan that ran out of memory beside its span table is built again
  // without it (the kernels then solve every row's run themselves)
  if (spans && first.find("out of memory") != std::string::npos) {
    (void)hipGetLastError();
    free_ptr(plan.spans);
    if (!band_prepare_once(plan, g, ev, n_events, stream, false, &spans, err))
      return 0;
    return -1;
  }
  if (err) *err = first;
  return -1;
}

int band_zero_ints(int* p, int n, hipStream_t stream, std::string* err) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 255) / 256), dim3(256), 0,
                     stream, p, n);
  BAND_TRY(hipGetLastError());
  return 0;
}

BandInstance band_descriptor(const BandPlan& plan, const DeviceGeometry& g,
                             int c) {
  BandInstance d{};
  d.g = g;
  d.rec = plan.records;
  d.fwd_pairs = plan.fwd.pair_event;
  d.bwd_pairs = plan.bwd.pair_event;
  d.spans = plan.spans;
  d.rhoT = plan.rhoT[c];
  d.rho_next = plan.rhoT[c ^ 1];
  d.partial = plan.partial;
  d.inv = plan.inv;
  d.bpT = plan.bpT;
  d.sensT = plan.sensT;
  d.n_events = plan.n_events;
  d.span_stride = plan.span_stride;
  d.Wt = plan.Wt;
  d.fwd_max_bands = plan.fwd.max_bands;
  return d;
}

int band_launch_forward(const BandInstance* desc, const int4* tasks,
                        int n_tasks, int H, int G, bool spans, bool recurrence,
                        size_t lds, int grid, int* queue, hipStream_t stream,
                        std::string* err) {
  return launch_forward_any(H, G, desc, tasks, n_tasks, spans, recurrence, 0,
                            lds, grid, queue, nullptr, stream, err);
}

int band_launch_batch_update(const BandInstance* desc, int n_instances,
                             int64_t max_events, int64_t max_pixels, int H,
                             bool inverse, hipStream_t stream,
                             std::string* err) {
  if (n_instances <= 0) return 0;
  if (inverse) {
    if (max_events <= 0) return 0;
    dim3 grid((unsigned)((max_events + 255) / 256), (unsigned)n_instances);
    hipLaunchKernelGGL(batch_inverse_kernel, grid, dim3(256), 0, stream, desc,
                       H);
  } else {
    const int64_t blocks = std::min<int64_t>((max_pixels + 255) / 256, 1024);
    dim3 grid((unsigned)std::max<int64_t>(blocks, 1), (unsigned)n_instances);
    hipLaunchKernelGGL(batch_finalize_kernel, grid, dim3(256), 0, stream,
                       desc);
  }
  BAND_TRY(hipGetLastError());
  return 0;
}

int band_launch_backward(const BandInstance* desc, const int4* tasks,
                         int n_tasks, int H, int max_group, bool spans,
                         bool recurrence, int wave_words, int waves,
                         int grid, int* queue, hipStream_t stream,
                         std::string* err) {
  return launch_backward_any(H, max_group, desc, tasks, n_tasks, spans,
                             recurrence, wave_words, waves, grid, queue,
                             stream, err);
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

int band_finalize(BandPlan& plan, const DeviceGeometry& g,
                  int sensitivity_mode, hipStream_t stream, std::string* err,
                  double scale, float* bp_out) {
  const int64_t n_pixels = (int64_t)g.n_z * g.n_y;
  const int blocks = (int)std::min<int64_t>((n_pixels + 255) / 256, 4096);
  hipLaunchKernelGGL(band_finalize_kernel, dim3(blocks), dim3(256), 0, stream,
                     n_pixels, plan.rhoT[plan.curT], plan.bpT, plan.sensT,
                     sensitivity_mode == 2 ? 1 : 0, scale,
                     plan.rhoT[plan.curT ^ 1], bp_out);
  BAND_TRY(hipGetLastError());
  plan.curT ^= 1;
  return 0;
}

int band_iteration(BandPlan& plan, const DeviceGeometry& g,
                   const EventSoA& ev, int sensitivity_mode, int guard_zero,
                   hipStream_t stream, hipEvent_t t_fwd, hipEvent_t t_bwd,
                   std::string* err, double scale, float* bp_out,
                   bool finalize) {
  if (!plan.ready) {
    if (err) *err = "band plan not prepared";
    return -1;
  }
  if (sensitivity_mode == 1) {
    if (err) *err = "the band algorithm does not fold sensitivity into D";
    return -1;
  }
  if (dispatch_forward(plan, g, 0, nullptr, stream, err)) return -1;
  if (t_fwd) BAND_TRY(hipEventRecord(t_fwd, stream));
  if (plan.n_events > 0) {
    const unsigned blocks = (unsigned)((plan.n_events + 255) / 256);
    hipLaunchKernelGGL(band_inverse_kernel, dim3(blocks), dim3(256), 0, stream,
                       plan.records, plan.n_events, plan.fwd.H,
                       plan.fwd.max_bands, plan.partial, guard_zero, plan.inv,
                       nullptr);
    BAND_TRY(hipGetLastError());
  }
  if (dispatch_backward(plan, g, plan.inv, plan.bpT, stream, err)) return -1;
  if (t_bwd) BAND_TRY(hipEventRecord(t_bwd, stream));
  if (!finalize) return 0;
  return band_finalize(plan, g, sensitivity_mode, stream, err, scale, bp_out);
}

int band_forward_only(BandPlan& plan, const DeviceGeometry& g,
                      const EventSoA& ev, float* denominators,
                      hipStream_t stream, std::string* err) {
  if (dispatch_forward(plan, g, 0, nullptr, stream, err)) return -1;
  if (plan.n_events > 0) {
    const unsigned blocks = (unsigned)((plan.n_events + 255) / 256);
    // D in record order into plan.inv (free between iterations), then back
    // to the caller's event order
    hipLaunchKernelGGL(band_inverse_kernel, dim3(blocks), dim3(256), 0, stream,
                       plan.records, plan.n_events, plan.fwd.H,
                       plan.fwd.max_bands, plan.partial, 1, nullptr,
                       plan.inv);
    BAND_TRY(hipGetLastError());
    hipLaunchKernelGGL(unpermute_kernel, dim3(blocks), dim3(256), 0, stream,
                       plan.inv, plan.order, plan.n_events, denominators);
    BAND_TRY(hipGetLastError());
  }
  return 0;
}

int band_to_column_major(const DeviceGeometry& g, const float* in, float* outT,
                         hipStream_t stream, std::string* err) {
  dim3 grid((g.n_z + 31) / 32, (g.n_y + 31) / 32);
  hipLaunchKernelGGL(to_column_major_kernel, grid, dim3(256), 0, stream, in,
                     nullptr, g.n_z, g.n_y, outT);
  BAND_TRY(hipGetLastError());
  return 0;
}

int band_to_row_major(const DeviceGeometry& g, const float* inT, float* out,
                      hipStream_t stream, std::string* err) {
  dim3 grid((g.n_y + 31) / 32, (g.n_z + 31) / 32);
  hipLaunchKernelGGL(to_row_major_kernel, grid, dim3(256), 0, stream, inT,
                     g.n_z, g.n_y, out);
  BAND_TRY(hipGetLastError());
  return 0;
}

int band_backproject(BandPlan& plan, const DeviceGeometry& g,
                     const float* weights, float* outT, hipStream_t stream,
                     std::string* err) {
  if (!plan.ready) {
    if (err) *err = "band plan not prepared";
    return -1;
  }
  return dispatch_backward(plan, g, weights, outT, stream, err);
}

int band_count(BandPlan& plan, const DeviceGeometry& g, const EventSoA& ev,
               unsigned long long* d_total, hipStream_t stream,
               std::string* err) {
  return dispatch_forward(plan, g, 1, d_total, stream, err);
}

}  // namespace strip
