// Row-band algorithm of the strip MLEM iteration (host interface).
//
// Each event is split into (event, band) pairs, a band being H consecutive
// image rows (H = 16, 32 or 64).  Pairs are radix-sorted once per event
// upload by (band, first column the band's row runs reach) and cut into
// tasks that never leave one (band, column bucket) group; every task's LDS
// window spans only the columns its pairs reach (at most bucket + widest
// run slice).  One iteration then runs
//   band_forward   a workgroup stages the task's rho window (columns x H
//                  rows, column-major: word = column*H + row) in LDS; a
//                  wavefront step takes G pairs, one per group of 64/G lanes,
//                  with 64/(G*H) lanes per image row, so lanes of a group
//                  never share an LDS bank; partial D per pair;
//   band_inverse   D_e = sum of the event's partials, inv_e = 1/D_e (guarded);
//   band_backward  each wavefront owns G private windows in LDS (one per lane
//                  group) and accumulates K * inv_e of its task's pairs with
//                  plain read-modify-write (lanes own distinct words, groups
//                  and waves own their windows: no LDS atomics), then adds the
//                  summed windows to the column-major back-projection image
//                  with coalesced atomics;
//   band_finalize  rho'(p) = rho(p) * bp(p) [/ s(p)], column-major.
// The in-ellipse pixels of an event in one image row form one run of
// columns; the runs of every bbox row of every event are found once per
// upload (span table) and read by both per-iteration kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "strip_device.hpp"

namespace strip {

struct BandTuning {
  int fwd_rows = 32;     // forward band height: 64, 32 or 16
  int bwd_rows = 16;     // back-projection band height: 64, 32 or 16
  int bwd_waves = 0;     // wavefronts per back-projection workgroup (0 = as
                         // many as fit, at most 4)
  int bucket = 8;        // column bucket width: 8, 16, 32 or 64
  int fwd_pairs = 2048;  // pairs per forward workgroup task
  int bwd_pairs = 512;   // pairs per back-projection wavefront task
  // pairs a wavefront processes together (1, 2 or 4, at most 64 / rows):
  // lane groups of 64 / G lanes take one pair each, with 64 / (G * rows)
  // lanes per image row
  int fwd_group = 2;
  int bwd_group = 4;  // the most; a task takes fewer when its windows are wide
  // true until a caller sets the groups: each back-projection wavefront then
  // owns the LDS that keeps bwd_cu_waves wavefronts per CU resident (or one
  // widest window, where that share would leave the widest tasks one pair),
  // and a task takes the most pairs per step (up to bwd_group) whose windows
  // fit there; with set groups every task takes bwd_group
  bool auto_groups = true;
  int bwd_cu_waves = 16;
  // kernel values along a row: 0 = one exp2 per pixel; 1 = the exact
  // two-term recurrence e[k+1] = e[k] r[k], r[k+1] = r[k] G of exp2 of a
  // quadratic, restarted from exp2 every 8 pixels (4 in a tail); the forward
  // projection sums each block in Horner form with r[k] taken downward
  int recurrence = 1;
  // 1 = build the span table when it fits, 0 = solve every row's run in the
  // iteration kernels
  int spans = 1;
};

// Per-event parameters of the row-band kernels, packed in one 32-byte record
// so a pair costs one cache-line access: ellipse centre (cz, cy), quadratic
// form coefficients A, B (C is a constant of the geometry), row prefactor
// coefficients P', Q0' and the clamped bbox (x0 | x1<<16, y0 | y1<<16).
struct alignas(32) BandRecord {
  float cz, cy, A, B, P, Q0;
  uint32_t bx, by;
};

// (event, band) pairs of one band height, sorted by (band, x0), and the
// tasks cut from them.
struct PairList {
  int H = 0;                  // rows per band
  int n_bands = 0;
  int max_bands = 0;          // most bands one event touches
  int64_t n_pairs = 0;
  int* pair_event = nullptr;  // pair -> event
  // (band | instance << 16, begin, end, first column | columns << 16)
  int4* tasks = nullptr;
  int n_tasks = 0;
  int Wt = 1;                 // widest task window (columns)
  // while the plan is built: columns [lo, hi) each (record, band) reaches,
  // lo | hi << 16 at [record * max_bands + band offset]
  int* bounds = nullptr;
  bool owns_pairs = true;
  std::vector<int4> tasks_host;  // host copy (instance 0)
};

// What the band kernels need of one reconstruction, read per task from a
// device array so that one launch can serve many reconstructions (a task's
// x field carries its instance index above 16 bits of band index).
struct BandInstance {
  DeviceGeometry g;
  const BandRecord* rec;
  const int* fwd_pairs;     // pair -> record, forward band height
  const int* bwd_pairs;     // pair -> record, back-projection band height
  const uint16_t* spans;    // span table or null
  const float* rhoT;        // image the forward pass reads (column-major)
  float* rho_next;          // image the update writes
  float* partial;           // partial D per (record, forward band)
  float* inv;               // back-projection weight per record
  float* bpT;               // back-projection accumulator (column-major)
  const float* sensT;       // sensitivity (numerator mode) or null
  int64_t n_events;
  int span_stride;
  int Wt;                   // task window width (columns)
  int fwd_max_bands;
  int pad;
};

// descriptor slots of a plan: 0/1 = iteration with image rhoT[0]/rhoT[1],
// the rest cache back-projections with caller-given weights and output
constexpr int kDescSlots = 6;

struct BandPlan {
  bool ready = false;
  BandTuning tuning;
  PairList fwd, bwd;
  int Wt = 0;         // bound on the task windows: bucket + widest bbox
  int bwd_waves = 4;  // wavefronts per back-projection workgroup
  int fwd_group = 1;  // pairs per wavefront step in use
  int bwd_group = 1;  // the most pairs per back-projection step
  int bwd_wave_words = 0;  // LDS words per back-projection wavefront
  int max_width = 0;  // widest clamped bbox (x1 - x0)
  int64_t n_events = 0;
  // events renumbered in (bbox y0, x0) order so neighbouring pairs of a task
  // share cache lines; order[i] = original index of record i
  BandRecord* records = nullptr;
  int* order = nullptr;
  // exact support of every bbox row of every record, computed once per
  // upload: spans[i * span_stride + (iy - y0a)] = (lo - x0) | n << 8 (first
  // column and length of row iy's in-ellipse run, 0 outside the bbox), y0a =
  // y0 rounded down to a multiple of 64 so the rows of a band are one aligned
  // piece; null when a bbox is wider than 255 columns or the table does not
  // fit, then the kernels solve each row's span themselves
  uint16_t* spans = nullptr;
  int span_stride = 0;  // largest slot: bbox rows rounded out to 64
  float* partial = nullptr;   // n_events * fwd.max_bands
  float* inv = nullptr;       // n_events
  float* rhoT[2] = {nullptr, nullptr};  // column-major images (n_z * n_y)
  bool owns_images = true;  // false: rhoT and bpT are shared, set by the owner
  int curT = 0;
  float* bpT = nullptr;
  float* sensT = nullptr;
  size_t fwd_lds = 0, bwd_lds = 0;
  int64_t generation = 0;  // changes at every band_prepare
  BandInstance* desc = nullptr;  // device, kDescSlots
  BandInstance desc_host[kDescSlots] = {};
  int n_custom = 0;              // custom slots in use (round robin)
};

// Descriptor of a prepared plan reading image rhoT[c] (weights plan.inv, into
// plan.bpT).
BandInstance band_descriptor(const BandPlan& plan, const DeviceGeometry& g,
                             int c);

// Launches of the band kernels over a task list of one or many instances.
// grid = workgroups (tasks are taken in grid strides); lds = dynamic LDS
// bytes of the forward projection.
// queue (may be null): a device counter, zero at launch, from which the
// resident workgroups (forward) / wavefronts (back-projection) take tasks.
// H = band rows, G = pairs per forward wavefront step.  The back-projection
// gives every wavefront wave_words words of LDS (at least one widest task
// window, Wt x H) in workgroups of `waves`; a task takes the most pairs per
// step, up to max_group and halving, whose windows fit in them.
int band_launch_forward(const BandInstance* desc, const int4* tasks,
                        int n_tasks, int H, int G, bool spans, bool recurrence,
                        size_t lds, int grid, int* queue, hipStream_t stream,
                        std::string* err);
int band_launch_backward(const BandInstance* desc, const int4* tasks,
                         int n_tasks, int H, int max_group, bool spans,
                         bool recurrence, int wave_words, int waves,
                         int grid, int* queue, hipStream_t stream,
                         std::string* err);

// LDS bytes per back-projection wavefront that keep cu_waves wavefronts
// resident per CU in workgroups of four.
size_t backprojection_budget(int cu_waves);

// Zeroes n ints on the stream with a kernel (graph-recordable).
int band_zero_ints(int* p, int n, hipStream_t stream, std::string* err);

// Per-instance updates of a batch (blockIdx.y = instance): inverse = true
// forms inv = D>0 ? 1/D : 0 from the partials, inverse = false writes
// rho_next = rhoT * bpT and clears bpT.
int band_launch_batch_update(const BandInstance* desc, int n_instances,
                             int64_t max_events, int64_t max_pixels, int H,
                             bool inverse, hipStream_t stream,
                             std::string* err);

// True when the row-band algorithm can run on this grid.
bool band_supported_grid(const DeviceGeometry& g);
// Upper bound of any event's clamped bbox width (columns), from sigma and the
// pixel width alone.
int band_max_box_width(const DeviceGeometry& g);
// A 16-row task window of the widest possible bbox fits in LDS.
bool band_window_fits(const DeviceGeometry& g);

// Builds the pair lists and tasks for the resident events.
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
// scale multiplies the update (in FP64): 1 for MLEM, the ratio of all events
// to the plan's events for an ordered-subset update.  bp_out (column-major,
// may be null) receives the iteration's back-projection sum_e K_e / D_e.
// finalize = false stops after the back-projection (added into plan.bpT), so
// several plans sharing the images can accumulate one MLEM update before
// band_finalize.
int band_iteration(BandPlan& plan, const DeviceGeometry& g,
                   const EventSoA& ev, int sensitivity_mode, int guard_zero,
                   hipStream_t stream, hipEvent_t t_fwd, hipEvent_t t_bwd,
                   std::string* err, double scale = 1.0,
                   float* bp_out = nullptr, bool finalize = true);

// rhoT[curT ^ 1] = rhoT[curT] * bpT [/ sensT] (x scale in FP64), bpT = 0
// (bp_out, if given, keeps bpT), then curT flips.
int band_finalize(BandPlan& plan, const DeviceGeometry& g,
                  int sensitivity_mode, hipStream_t stream, std::string* err,
                  double scale = 1.0, float* bp_out = nullptr);

// Per-event forward projections of the current column-major image.
int band_forward_only(BandPlan& plan, const DeviceGeometry& g,
                      const EventSoA& ev, float* denominators,
                      hipStream_t stream, std::string* err);

// Image layout changes: row-major (y * n_z + x) <-> column-major
// (x * n_y + y).
int band_to_column_major(const DeviceGeometry& g, const float* in, float* outT,
                         hipStream_t stream, std::string* err);
int band_to_row_major(const DeviceGeometry& g, const float* inT, float* out,
                      hipStream_t stream, std::string* err);

// outT (column-major, not cleared) += sum_e K_e * weights[e] over the
// resident events (weights in record order).
int band_backproject(BandPlan& plan, const DeviceGeometry& g,
                     const float* weights, float* outT, hipStream_t stream,
                     std::string* err);

// ---------------------------------------------------------------------------
// differentiation (strip_adjoint.hip)
// ---------------------------------------------------------------------------

// Device work buffers of the gradient passes.
struct AdjointWork {
  float* moments = nullptr;  // 6 floats per (event, forward band)
  int64_t moments_cap = 0;
  float* ev_buf = nullptr;   // 5 per-event arrays: w1, w2, value, gz, gd
  int64_t ev_cap = 0;
  float* img = nullptr;      // 3 column-major images: h, bp1, bp2
  int64_t img_cap = 0;
  // FP64 reduction results (kSumSlots) followed by the block partials
  double* sums = nullptr;
};
constexpr int kSumSlots = 8;

void adjoint_release(AdjointWork& wk);

// The plan's part of the log-likelihood of column-major image rhoT:
// *sum_log = sum_{e: D_e > 0} log D_e, grad_ln_sigma += its d / d ln sigma,
// and (bp_accum not null) bp_accum += sum_e K_e / D_e.
int band_loglik_terms(BandPlan& plan, const DeviceGeometry& g,
                      AdjointWork& wk, const float* rhoT, float* bp_accum,
                      double* sum_log, double grad_ln_sigma[2],
                      hipStream_t stream, std::string* err);

// Sum of a column-major image in FP64 (*total) and in FP32 order-free form.
int band_image_total(const DeviceGeometry& g, AdjointWork& wk,
                     const float* imgT, double* total, hipStream_t stream,
                     std::string* err);

// Poisson log-likelihood L = sum_{e: D_e > 0} log D_e - sum_p rho(p) of the
// column-major image rhoT, with d L / d rho into gradT (column-major, may be
// null) and d L / d ln sigma_z, d L / d ln sigma_dl.  The support of every
// event is held fixed (the almost-everywhere derivative).
int band_loglik_gradient(BandPlan& plan, const DeviceGeometry& g,
                         AdjointWork& wk, const float* rhoT, float* gradT,
                         double* value, double grad_ln_sigma[2],
                         hipStream_t stream, std::string* err);

// One reverse step through rho_{k+1} = rho_k * sum_e K_e / D_e(rho_k):
// gbarT (column-major d loss / d rho_{k+1}) is replaced by d loss / d rho_k
// and the step's d loss / d ln sigma is added to grad_ln_sigma.  bpT_k (may
// be null) is the forward iteration's back-projection sum_e K_e / D_e kept
// from the forward pass; without it the step computes it again.
int band_mlem_adjoint_step(BandPlan& plan, const DeviceGeometry& g,
                           AdjointWork& wk, const float* rhoT_k, float* gbarT,
                           double grad_ln_sigma[2], hipStream_t stream,
                           std::string* err, const float* bpT_k = nullptr);

// Allocates the work buffers for the plan's event count and grid.
int band_adjoint_prepare(BandPlan& plan, const DeviceGeometry& g,
                         AdjointWork& wk, std::string* err);

// Seed of the reverse pass of loss = mean_p (rho(p) - target(p))^2:
// gbarT = 2 (rhoT - targetT) / n_pixels; *loss receives the loss.
int band_mse_seed(const DeviceGeometry& g, AdjointWork& wk, const float* rhoT,
                  const float* targetT, float* gbarT, double* loss,
                  hipStream_t stream, std::string* err);

// Kernel-evaluation count of the span walk (must equal the exact count).
int band_count(BandPlan& plan, const DeviceGeometry& g, const EventSoA& ev,
               unsigned long long* d_total, hipStream_t stream,
               std::string* err);

}  // namespace strip
