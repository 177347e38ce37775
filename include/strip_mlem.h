/*
 * strip_mlem.h — C ABI of the MI355X (gfx950) J-PET 2D strip list-mode TOF
 * MLEM reconstruction.
 *
 * This is the drop-in boundary for the reference's `--gpu` path:
 *   template <typename F> void PET2D::Strip::GPU::Reconstruction::run(
 *       Scanner<F,S>&, Response<F>*, int, Output&, int, int,
 *       delegate<void(int, const Output&)>, delegate<void(int, bool)>,
 *       int device, int n_blocks, int n_threads_per_block,
 *       delegate<void(const char*)>)
 *   (src/2d/strip/cuda/reconstruction.h:31-44, .cu:31-113)
 * Plain pointers and sizes only; no C++ or torch types cross it.  A C++
 * wrapper that reproduces the templated `run<F>` call on top of this ABI is
 * in strip_gpu_reconstruction.hpp.
 *
 * Error convention: every function returns STRIP_MLEM_OK (0) or a negative
 * status; strip_mlem_last_error() gives a message for the calling thread.
 * (The reference aborts on any CUDA error, src/util/cuda/debug.h:17-55; the
 * C++ wrapper restores that behaviour by throwing.)
 *
 * Threading: a context is single-threaded (one host thread at a time); all
 * device work is issued on the context's own HIP stream.  Several contexts
 * may coexist in one process (the reference's global textures made its run()
 * non-reentrant, cuda/reconstruction.cu:12-13; this ABI is not).
 */
#ifndef STRIP_MLEM_H
#define STRIP_MLEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STRIP_MLEM_OK 0
#define STRIP_MLEM_ERR_ARGUMENT (-1)
#define STRIP_MLEM_ERR_HIP (-2)
#define STRIP_MLEM_ERR_STATE (-3)
#define STRIP_MLEM_ERR_ALLOC (-4)
#define STRIP_MLEM_ERR_UNSUPPORTED (-5)

/* Scanner geometry, the arguments of Scanner<F,S>(R, L, n_y, n_z, h, w,
 * sigma_z, sigma_dl, center_y, center_z) (src/2d/strip/scanner.h:39-67).
 * All lengths in the same unit (the CLI uses millimetres). */
typedef struct strip_mlem_geometry {
  double radius;              /* R, half distance between the strips */
  double scintillator_length; /* L */
  int n_z_pixels;             /* image width  (z) */
  int n_y_pixels;             /* image height (y) */
  double pixel_width;         /* along z */
  double pixel_height;        /* along y */
  double sigma_z;             /* TOF sigma along z (--s-z) */
  double sigma_dl;            /* TOF sigma of delta-l (--s-dl) */
  double center_y;            /* 0 in the CLI */
  double center_z;            /* 0 in the CLI */
} strip_mlem_geometry;

/* Reconstruction semantics.
 * STRIP_MLEM_MODE_CPU reproduces the reference CPU path (the graded parity
 * mode: GaussianKernel, no sensitivity, zero-denominator guard,
 * src/2d/strip/reconstruction.h:220-298).  STRIP_MLEM_MODE_LEGACY_GPU
 * reproduces the old CUDA warp kernel (AnalyticKernel, 5-point pixel
 * sensitivity in the denominator, no guard,
 * cuda/reconstruction/warp_granularity.cuh:88-152). */
typedef struct strip_mlem_mode {
  int kernel;          /* 0 = Gaussian, 1 = Analytic */
  int sensitivity;     /* 0 = none, 1 = multiplies the denominator,
                          2 = divides the back-projected numerator */
  int sensitivity_map; /* 0 = sensitivity(pixel centre), 1 = 5-point */
  int guard_zero;      /* 1 = inv = D>0 ? 1/D : 0 ;  0 = inv = 1/D */
} strip_mlem_mode;

#define STRIP_MLEM_KERNEL_GAUSSIAN 0
#define STRIP_MLEM_KERNEL_ANALYTIC 1

typedef struct strip_mlem_ctx strip_mlem_ctx;

/* Creates a context on HIP device `device` (the reference's -D/--cuda-device,
 * src/common/options.cpp:8).  mode may be NULL (= CPU parity mode).  The
 * image starts at rho = 1 (reconstruction.h:59). */
int strip_mlem_create(const strip_mlem_geometry* geometry,
                      const strip_mlem_mode* mode, int device,
                      strip_mlem_ctx** out);
void strip_mlem_destroy(strip_mlem_ctx* ctx);

/* Replaces the event list.  responses: host AoS triples (z_u, z_d, dl), the
 * layout of Response<float> (src/2d/strip/response.h:13-20) and of the
 * binary event file.  The events are copied to HBM and their per-event
 * ellipse parameters precomputed once (the reference recomputes them every
 * iteration, reconstruction.h:132-134).  (Replaces ResponsesSOA,
 * cuda/gpu_events_soa.h:15-44.) */
int strip_mlem_set_events(strip_mlem_ctx* ctx, const float* responses,
                          int64_t n_responses);
/* Same, from AoS triples already resident in device memory (not retained). */
int strip_mlem_set_events_device(strip_mlem_ctx* ctx,
                                 const float* device_responses,
                                 int64_t n_responses);
int64_t strip_mlem_n_events(const strip_mlem_ctx* ctx);

/* Streaming: strip_mlem_reserve_events sizes the device event arrays for
 * n_events (resident events kept); strip_mlem_append_events adds host AoS
 * events after the resident ones.  Uploads go in chunks of up to 2^23
 * events through two pinned buffers on a copy stream, overlapped with the
 * per-event setup of the previous chunk. */
int strip_mlem_reserve_events(strip_mlem_ctx* ctx, int64_t n_events);
int strip_mlem_append_events(strip_mlem_ctx* ctx, const float* responses,
                             int64_t n_responses);

/* Ordered subsets (no counterpart in the reference, whose update uses every
 * event): n_subsets > 1 splits the resident events by index into that many
 * contiguous ranges, and each strip_mlem_iterate iteration then applies one
 * update per subset s:  rho' = rho * (sum_{e in s} K_e / D_e) * (N / N_s),
 * the count ratio N / N_s applied in FP64.  1 = plain MLEM (default).  CPU-path
 * semantics on the row-band algorithm only. */
int strip_mlem_set_subsets(strip_mlem_ctx* ctx, int n_subsets);

/* Events per row-band plan of plain MLEM (default 2^25): a context holding
 * more splits them into contiguous ranges with a plan each, which accumulate
 * one MLEM update together (the plan's pair indices are 32-bit; a smaller
 * value also bounds the per-plan memory).  No counterpart in the reference,
 * whose CPU loop has no plan. */
int strip_mlem_set_plan_events(strip_mlem_ctx* ctx, int64_t n);

/* Sum of the current image's pixels, accumulated in FP64 on the device (for
 * MLEM it equals the number of events with D_e > 0 after an iteration,
 * reconstruction.h:290-296). */
int strip_mlem_image_total(strip_mlem_ctx* ctx, double* total);

/* Start image / current image, n_y*n_z floats row-major (index y*n_z+z,
 * src/2d/geometry/pixel.h:29).  (Replaces output_rho.copy_to_device /
 * copy_from_device, cuda/reconstruction.cu:73-74,107.) */
int strip_mlem_set_rho(strip_mlem_ctx* ctx, const float* rho);
int strip_mlem_get_rho(strip_mlem_ctx* ctx, float* rho);
/* Device-resident variants (pointers to n_y*n_z floats in device memory;
 * the copies are ordered on the context stream). */
int strip_mlem_set_rho_device(strip_mlem_ctx* ctx, const float* device_rho);
int strip_mlem_get_rho_device(strip_mlem_ctx* ctx, float* device_rho);

/* Sensitivity map used by the mode (centre or 5-point), n_y*n_z floats.
 * (Replaces fill_with_sensitivity, cuda/reconstruction.cu:115-126, and the
 * CPU map of reconstruction.h:67-73.) */
int strip_mlem_get_sensitivity(strip_mlem_ctx* ctx, int kind,
                               float* sensitivity);

/* Runs n MLEM iterations on the resident events (one iteration of
 * Reconstruction::operator(), reconstruction.h:117-152, per step) and
 * waits for them to finish. */
int strip_mlem_iterate(strip_mlem_ctx* ctx, int n_iterations);
/* Same, without waiting (work is queued on the context stream). */
int strip_mlem_iterate_async(strip_mlem_ctx* ctx, int n_iterations);
int strip_mlem_synchronize(strip_mlem_ctx* ctx);

/* Launch shape, replacing -B/--cuda-blocks and -W/--cuda-threads
 * (src/common/options.cpp:9-11).  0 keeps the automatic choice.
 * threads_per_block must be a multiple of 64 (one wavefront per event). */
int strip_mlem_set_launch(strip_mlem_ctx* ctx, int n_blocks,
                          int threads_per_block);
/* Selects the hot-path implementation: 0 = automatic (row band when the
 * mode and grid allow it), 1 = bbox wavefront kernel (every mode), 2 =
 * row-band LDS kernels (Gaussian kernel, sensitivity none or numerator).
 * Returns STRIP_MLEM_ERR_UNSUPPORTED when the requested one cannot run the
 * context's mode or grid.  strip_mlem_get_algorithm returns the one in use. */
int strip_mlem_set_algorithm(strip_mlem_ctx* ctx, int algorithm);
int strip_mlem_get_algorithm(const strip_mlem_ctx* ctx);

/* Tile shapes of the row-band algorithm (the knobs an autotuner turns in
 * place of -B/-W): forward_rows / backward_rows = image rows per band of the
 * forward-projection and back-projection kernels (16, 32 or 64),
 * bucket_columns = width of the column buckets bounding a task's window (8, 16,
 * 32 or 64), forward_pairs / backward_pairs = (event, band) pairs per forward
 * workgroup / per back-projection wavefront.  Takes effect at the next
 * iteration. */
int strip_mlem_set_band_tiling(strip_mlem_ctx* ctx, int forward_rows,
                               int backward_rows, int bucket_columns,
                               int forward_pairs, int backward_pairs);
/* (Event, band) pairs one wavefront of the forward / back-projection kernel
 * processes together: 1, 2 or 4 (at most 64 / band rows).  Lane groups of
 * 64 / G lanes then take one pair each, so one instruction stream serves G
 * pairs.  The back-projection keeps one LDS window per group: with a set
 * backward_group every task takes it (halved until one wavefront's windows
 * fit).  0 = automatic, the default: forward 2; back-projection: every
 * wavefront owns the LDS that keeps strip_mlem_set_backprojection_occupancy
 * wavefronts per CU resident (or one widest task window, where that share
 * would leave the widest tasks one pair per step), each task taking the
 * most pairs per step, up to 4 and halving, whose windows fit there (wide
 * windows fewer, narrow ones more).  strip_mlem_get_band_groups returns the forward G and the
 * back-projection's most. */
int strip_mlem_set_band_groups(strip_mlem_ctx* ctx, int forward_group,
                               int backward_group);
int strip_mlem_get_band_groups(strip_mlem_ctx* ctx, int* forward_group,
                               int* backward_group);
/* How the row-band kernels evaluate exp(-Q/2) along an image row: 0 = one
 * hardware exp2 per pixel; 1 = the recurrence e[k+1] = e[k] r[k],
 * r[k+1] = r[k] G (exact for exp2 of a quadratic in the column), restarted
 * from exp2 every 8 pixels (4 in a tail); the forward projection sums each
 * block in Horner form, e[0] (rho[0] + r[0] (rho[1] + ...)).  Both stay
 * within the FP32 parity tolerance of the oracle. */
int strip_mlem_set_exp_mode(strip_mlem_ctx* ctx, int mode);
/* Wavefronts per back-projection workgroup (each owns an LDS window);
 * 0 = as many as fit, at most 4. */
int strip_mlem_set_backprojection_waves(strip_mlem_ctx* ctx, int waves);
/* Back-projection wavefronts per CU (4..32, default 16) whose LDS windows the
 * automatic groups keep resident; no counterpart in the reference (it
 * replaces part of -B/-W). */
int strip_mlem_set_backprojection_occupancy(strip_mlem_ctx* ctx,
                                            int cu_waves);
/* Row runs of the row-band algorithm: 1 (default) = the in-ellipse column run
 * of every bbox row of every event is solved once per upload and stored
 * (2 bytes per row; used when every clamped bbox is at most 255 columns wide
 * and the table takes at most half the free device memory), 0 = each
 * iteration solves the runs again.  Same pixels either way.  Takes effect at
 * the next upload-time plan (next iteration).  strip_mlem_get_span_table
 * reports whether the plan in use has the table (1) or not (0). */
int strip_mlem_set_span_table(strip_mlem_ctx* ctx, int enable);
int strip_mlem_get_span_table(strip_mlem_ctx* ctx, int* in_use);

/* The plan in use: band rows of both kernels, the widest back-projection
 * task window (columns) and the number of (event, band) pairs of both
 * kernels. */
int strip_mlem_get_band_tiling(strip_mlem_ctx* ctx, int* forward_rows,
                               int* backward_rows, int* window_columns,
                               int64_t* forward_pairs,
                               int64_t* backward_pairs);

/* Exact number of kernel evaluations of one iteration, sum_e P_e, where P_e
 * counts the pixels of event e's clamped bounding box that pass the 3-sigma
 * ellipse test (gaussian_kernel.h:131-137).  Computed on the device. */
int strip_mlem_kernel_evaluations(strip_mlem_ctx* ctx, int64_t* count);

/* Per-event forward projections D_e = sum_p K_e(p) rho(p) [s(p)] of the
 * current image (host array of n_events floats). */
int strip_mlem_forward(strip_mlem_ctx* ctx, float* denominators);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): strip_mlem_forward's projections D_e of device-resident
 * responses (n_events AoS float triples z_u, z_d, dl) through a
 * device-resident row-major n_y x n_z image into device_denominators, on
 * `stream` (a hipStream_t; NULL = the default stream) of the current device,
 * asynchronously.  It calls no hipMalloc / hipFree: the per-event parameters
 * live in a fixed device array of STRIP_MLEM_NOALLOC_MAX_EVENTS events
 * (14 MiB) that every GPU keeps until the process ends, so n_events above it
 * is refused, and calls that overlap on one GPU race on that array.  Modes
 * without a sensitivity map only (sensitivity 0); NULL mode = the CPU
 * semantics. */
#define STRIP_MLEM_NOALLOC_MAX_EVENTS (1 << 18)
int strip_mlem_forward_noalloc(const strip_mlem_geometry* geometry,
                               const strip_mlem_mode* mode,
                               const float* device_responses, int64_t n_events,
                               const float* device_rho,
                               float* device_denominators, void* stream);

/* ---- differentiation (no counterpart in the reference; the adjoint form of
 * the same kernels).  Defined for the CPU-path semantics (Gaussian kernel, no
 * sensitivity, guarded 1/D) on the row-band algorithm; other modes return
 * STRIP_MLEM_ERR_UNSUPPORTED.  The support of every event (clamped bbox and
 * Q <= 9 at the context's sigmas) is held fixed, so the sigma derivatives are
 * the almost-everywhere derivatives of the pixel sums.  Images are row-major
 * n_y x n_z (PixelMap order). ---- */

/* Poisson list-mode log-likelihood of image rho (NULL = the context image)
 *   L = sum_{e: D_e > 0} log D_e - sum_p rho(p)
 * (the objective whose EM step is the reference's update,
 * reconstruction.h:110-154), its gradient d L / d rho (may be NULL) and
 * d L / d sigma_z, d L / d sigma_dl in grad_sigma[0..1] (may be NULL).
 * Host pointers; the _device variant takes device pointers. */
int strip_mlem_loglik_gradient(strip_mlem_ctx* ctx, const float* rho,
                               double* value, float* grad_rho,
                               double* grad_sigma);
int strip_mlem_loglik_gradient_device(strip_mlem_ctx* ctx, const float* rho,
                                      double* value, float* grad_rho,
                                      double* grad_sigma);

/* Unrolled MLEM: n_iterations reference updates from rho0 (NULL = the context
 * image), every iterate kept as a checkpoint; loss = mean_p (rho_N(p) -
 * target(p))^2; its gradient with respect to rho0 (may be NULL) and to
 * (sigma_z, sigma_dl) (may be NULL), by reverse steps through the iterations.
 * The context image becomes rho_N.  Host pointers; the _device variant takes
 * device pointers. */
int strip_mlem_unrolled_gradient(strip_mlem_ctx* ctx, const float* rho0,
                                 int n_iterations, const float* target,
                                 double* loss, float* grad_rho0,
                                 double* grad_sigma);
int strip_mlem_unrolled_gradient_device(strip_mlem_ctx* ctx, const float* rho0,
                                        int n_iterations, const float* target,
                                        double* loss, float* grad_rho0,
                                        double* grad_sigma);

/* ---- batched reconstructions (no counterpart in the reference, which runs
 * one reconstruction per process).  Many independent event lists, each with
 * its own geometry and grid, iterated together: every pass of an iteration is
 * one launch over all members (persistent workgroups over the merged task
 * list) and the iterations are recorded once into a HIP graph and replayed.
 * Members use the reference CPU semantics and start at rho = 1. ---- */
typedef struct strip_mlem_batch strip_mlem_batch;

int strip_mlem_batch_create(int device, strip_mlem_batch** out);
void strip_mlem_batch_destroy(strip_mlem_batch* batch);
/* Adds a reconstruction (geometry + host AoS events); *index = its member
 * number (may be NULL). */
int strip_mlem_batch_add(strip_mlem_batch* batch,
                         const strip_mlem_geometry* geometry,
                         const float* responses, int64_t n_responses,
                         int* index);
int strip_mlem_batch_size(const strip_mlem_batch* batch);
/* The member's own context (owned by the batch): per-member calls such as
 * strip_mlem_iterate run it alone, the plain single-reconstruction form. */
int strip_mlem_batch_context(strip_mlem_batch* batch, int index,
                             strip_mlem_ctx** ctx);
/* n_iterations MLEM iterations of every member. */
int strip_mlem_batch_iterate(strip_mlem_batch* batch, int n_iterations);
int strip_mlem_batch_set_rho(strip_mlem_batch* batch, int index,
                             const float* rho);
int strip_mlem_batch_get_rho(strip_mlem_batch* batch, int index, float* rho);
/* Shared tiling of the members' plans (as strip_mlem_set_band_tiling) and
 * persistent workgroups per compute unit; tiling[6] reads them back. */
int strip_mlem_batch_set_tiling(strip_mlem_batch* batch, int forward_rows,
                                int backward_rows, int bucket_columns,
                                int forward_pairs, int backward_pairs,
                                int grid_factor);
int strip_mlem_batch_get_tiling(const strip_mlem_batch* batch, int* tiling);
/* Shared pairs per wavefront step of the members' plans (as
 * strip_mlem_set_band_groups, 0 = automatic; the batch launches with the
 * smallest of its members' forward G and back-projection limit). */
int strip_mlem_batch_set_groups(strip_mlem_batch* batch, int forward_group,
                                int backward_group);
int strip_mlem_batch_get_groups(const strip_mlem_batch* batch,
                                int* forward_group, int* backward_group);
/* Shared back-projection occupancy of the members' plans (as
 * strip_mlem_set_backprojection_occupancy; the batch gives every wavefront
 * the largest member's LDS budget). */
int strip_mlem_batch_set_backprojection_occupancy(strip_mlem_batch* batch,
                                                  int cu_waves);
int strip_mlem_batch_get_backprojection_occupancy(
    const strip_mlem_batch* batch, int* cu_waves);
/* 1 (default) = replay graph-recorded iterations, 0 = launch each. */
int strip_mlem_batch_set_graph(strip_mlem_batch* batch, int enable);
/* Device time (ms) of the last strip_mlem_batch_iterate call. */
int strip_mlem_batch_last_ms(const strip_mlem_batch* batch, double* ms);
/* Autotuner: coordinate search over the tiling knobs, the pairs per step and
 * the back-projection occupancy, each candidate timed on batch iterations
 * (sweeps repeat until one changes nothing, at most three); keeps the
 * fastest and restores the members' images.
 * *best_ms (may be NULL) = time of one iteration with the chosen tiling. */
int strip_mlem_batch_autotune(strip_mlem_batch* batch, double* best_ms);

/* Timing of the hot kernel(s) with HIP events on the context stream.
 * After strip_mlem_set_timing(ctx, 1), every iteration records the start and
 * end of its event-loop kernel(s); strip_mlem_get_timing returns the summed
 * milliseconds of those kernels, of whole iterations, and the iteration
 * count, then resets them. */
int strip_mlem_set_timing(strip_mlem_ctx* ctx, int enable);
int strip_mlem_get_timing(strip_mlem_ctx* ctx, double* kernel_ms,
                          double* iteration_ms, int64_t* n_iterations);
/* Same, with the forward-projection and back-projection kernels apart (the
 * bbox algorithm runs both in one kernel: all of it is reported as
 * forward). */
int strip_mlem_get_timing_detail(strip_mlem_ctx* ctx, double* forward_ms,
                                 double* backward_ms, double* iteration_ms,
                                 int64_t* n_iterations);

/* Name of the device (the reference's device_name callback,
 * cuda/reconstruction.cu:54-57). */
int strip_mlem_device_name(strip_mlem_ctx* ctx, char* buffer, int size);

/* Message for the last failing call on this thread ("" if none). */
const char* strip_mlem_last_error(void);

/* ---- One-call driver with the reference's output protocol ---------------
 * Mirrors GPU::Reconstruction::run (cuda/reconstruction.cu:31-113):
 *  - output(user, -1, sensitivity) once, with the 5-point pixel sensitivity;
 *  - per iteration: progress(user, iteration, 0) ... progress(user, it, 1);
 *  - in block 0, output(user, i+1, rho) after iterations i < 5 and
 *    i = 9, 14, 19, 29, 49, 99 unless i is the block's last iteration;
 *  - output(user, (block+1)*n_iterations_in_block, rho) after each block;
 *  - a final progress(user, n_blocks*n_iterations_in_block, 0).
 * rho_inout holds the start image and receives the final one.  The image
 * pointer passed to `output` is only valid during the callback.  The mode
 * decides the arithmetic (NULL = CPU parity mode). */
typedef void (*strip_mlem_output_fn)(void* user, int iteration,
                                     const float* image, int n_z, int n_y);
typedef void (*strip_mlem_progress_fn)(void* user, int completed,
                                       int finished);
typedef void (*strip_mlem_device_name_fn)(void* user, const char* name);

int strip_mlem_run(const strip_mlem_geometry* geometry,
                   const strip_mlem_mode* mode, const float* responses,
                   int n_responses, float* rho_inout, int n_iteration_blocks,
                   int n_iterations_in_block, strip_mlem_output_fn output,
                   strip_mlem_progress_fn progress, void* user, int device,
                   int n_blocks, int n_threads_per_block,
                   strip_mlem_device_name_fn device_name);

#ifdef __cplusplus
}
#endif
#endif
