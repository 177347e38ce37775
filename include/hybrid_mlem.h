/*
 * hybrid_mlem.h — C ABI of the MI355X (gfx950) 3D hybrid list-mode MLEM
 * reconstruction: a barrel's 2D LOR geometry (pixels along each LOR) crossed
 * with the strip's TOF Gaussian kernel along z.
 *
 * This is the drop-in boundary for the reference's `--gpu` path of
 * `3d_hybrid_reconstruction`:
 *   void PET3D::Hybrid::GPU::Reconstruction::run(
 *       const GeometrySOA<F,S>&, const VoxelMap& sensitivity,
 *       const FrameEvent* events, int n_events, F sigma_z, F sigma_dl,
 *       const VoxelGrid<F,S>&, F barrel_length, int n_iteration_blocks,
 *       int n_iterations_in_block, const VoxelMap& rho_output,
 *       int start_iteration, delegate output, delegate progress,
 *       int device, int n_blocks, int n_threads_per_block, delegate info)
 *   (src/3d/hybrid/cuda/reconstruction.h:36-55, .cu:36-187)
 * and for the CPU steps the command line runs before it
 * (src/3d/hybrid/reconstruction_cmd.cpp:459-560): calculate_weight,
 * calculate_sensitivity / set_sensitivity_to_one / a sensitivity file,
 * normalize_geometry_weights, and the event reader's translate_to_frame +
 * bb_intersects_grid_with_positive_weight (reconstruction.h:101-212), which
 * run on the device here.  Plain pointers and sizes only.
 *
 * Error convention as strip_mlem.h: STRIP_MLEM_OK (0) or a negative status,
 * message from strip_mlem_last_error().
 *
 * Supported: single-plane pixel weights (n_planes_half = 1) with a 2D
 * sensitivity map, the command line's default.  The multi-plane and 3D
 * sensitivity branches (reconstruction.h:347-364) index plane
 * |iz - n_planes_half| = n_planes_half at iz = 0, one past their arrays;
 * they are refused (STRIP_MLEM_ERR_UNSUPPORTED).
 */
#ifndef HYBRID_MLEM_H
#define HYBRID_MLEM_H

#include <stddef.h>
#include <stdint.h>

#include "strip_mlem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* GeometrySOA (src/2d/barrel/geometry_soa.h:30-58) + VoxelGrid
 * (src/3d/geometry/voxel_grid.h:22-26, pixel grid centred on the origin,
 * src/2d/geometry/pixel_grid.h:33-38) + the kernel's sigmas and the
 * scintillator length.  Lengths in one unit (the command line uses metres).
 * The arrays are copied to the device by hybrid_mlem_create. */
typedef struct hybrid_mlem_geometry {
  int32_t n_detectors;   /* n_lors = n_detectors (n_detectors + 1) / 2 */
  int64_t n_pixel_infos;
  int32_t n_planes_half; /* must be 1 */
  const float* lor_line_segments; /* n_lors x 4: start x, y, end x, y */
  const float* lor_widths;        /* n_lors */
  const int16_t* pixels;          /* n_pixel_infos x 2: x, y */
  const float* pixel_positions;   /* n_pixel_infos, ascending per LOR */
  const float* pixel_weights;     /* n_pixel_infos */
  const int64_t* lor_pixel_info_begin; /* n_lors */
  const int64_t* lor_pixel_info_end;   /* n_lors */
  int32_t n_columns, n_rows, n_planes;
  double pixel_size;
  double z_left;
  double sigma_z, sigma_dl;
  double length;          /* scintillator (barrel) length */
} hybrid_mlem_geometry;

/* One list-mode response as the reference's binary event file stores it
 * (PET3D::Hybrid::Scanner::Response, src/3d/hybrid/scanner.h:71-87:
 * LOR (first, second) as shorts, then z_up, z_dn, dl). */
typedef struct hybrid_mlem_response {
  int16_t first, second;
  float z_up, z_dn, dl;
} hybrid_mlem_response;

/* Update semantics of a voxel that one event visits twice.  An event's
 * pixel-info range ends one info past the first position above the box
 * (reconstruction.h:131-138); when that extra info belongs to the next LOR
 * its pixel can repeat one of the range's.  HYBRID_MLEM_CPU: the CPU
 * path's per-thread kernel cache keeps the last weight written for the
 * voxel and both visits add it (reconstruction.h:347-369,417-429).
 * HYBRID_MLEM_LEGACY_GPU: the CUDA kernel recomputes each visit's own
 * weight (cuda/reconstruction/warp_granularity.cuh:136-168). */
#define HYBRID_MLEM_CPU 0
#define HYBRID_MLEM_LEGACY_GPU 1

typedef struct hybrid_mlem_ctx hybrid_mlem_ctx;

/* Copies the geometry to device `device` and checks it: LOR ranges inside
 * [0, n_pixel_infos], pixels of one LOR distinct (the layout both of the
 * reference's geometry builders produce).  The image starts at rho = 1
 * (reconstruction.h:78-81); the sensitivity map at zero until
 * hybrid_mlem_prepare_weights. */
int hybrid_mlem_create(const hybrid_mlem_geometry* geometry, int semantics,
                       int device, hybrid_mlem_ctx** out);
void hybrid_mlem_destroy(hybrid_mlem_ctx* ctx);

/* Weights and sensitivity as the command line prepares them
 * (reconstruction_cmd.cpp:499-519):
 *   calculate_weight 1: weight = N(distance of the pixel centre from the
 *     LOR's line; sigma = 0.3 LOR width) (reconstruction.h:218-238);
 *     0: the weights given to hybrid_mlem_create;
 *   sensitivity_mode 0: per-pixel sum of the weights
 *     (calculate_sensitivity, reconstruction.h:240-251); 1: ones
 *     (set_sensitivity_to_one); 2: `sensitivity` (n_rows x n_columns
 *     floats, row-major, the --sensitivity file);
 * then every weight is divided by its pixel's sensitivity
 * (normalize_geometry_weights, reconstruction.h:253-271).  Sums run in info
 * order per pixel, as the reference's loop does.
 *   sensitivity_mode 3 (HYBRID_MLEM_SENSITIVITY_NORMALISED): `sensitivity`
 *     as given and the weights given to hybrid_mlem_create taken as already
 *     normalised (no division): the state in which the reference's CLI
 *     hands geometry and sensitivity to its GPU run
 *     (reconstruction_cmd.cpp:596-600). */
#define HYBRID_MLEM_SENSITIVITY_NORMALISED 3
int hybrid_mlem_prepare_weights(hybrid_mlem_ctx* ctx, int calculate_weight,
                                int sensitivity_mode,
                                const float* sensitivity);

/* Reads responses (host array): translate_to_frame and the box filter
 * (reconstruction.h:101-162,190-212) on the device; kept events are
 * appended after those already held, in input order.  Must follow
 * hybrid_mlem_prepare_weights (the filter reads the normalised weights).
 * A response whose LOR index is outside the geometry is an argument
 * error (the reference reads past its arrays). */
int hybrid_mlem_add_responses(hybrid_mlem_ctx* ctx,
                              const hybrid_mlem_response* responses,
                              int64_t n_responses);
int hybrid_mlem_clear_events(hybrid_mlem_ctx* ctx);
int64_t hybrid_mlem_n_events(const hybrid_mlem_ctx* ctx);

/* A frame event as the reference's event reader leaves it (FrameEvent,
 * reconstruction.h:53-63): LOR, up, right, tan, sec, pixel-info range and
 * plane range. */
typedef struct hybrid_mlem_frame_event {
  int16_t first, second;
  float up, right, tan, sec;
  int64_t info_begin, info_end;
  int16_t plane_begin, plane_end;
} hybrid_mlem_frame_event;

/* Replaces the events held by frame events computed elsewhere (the
 * reference's GPU run receives them from its CPU reader,
 * cuda/reconstruction.cu:36-57); no filter is applied.  An info end one
 * past the arrays (the reference's upper_bound + 1 on the last LOR with
 * infos) stops at the arrays' end; a LOR, info or plane range outside the
 * geometry is an argument error. */
int hybrid_mlem_set_frame_events(hybrid_mlem_ctx* ctx,
                                 const hybrid_mlem_frame_event* events,
                                 int64_t n_events);

/* Name of the context's device (the run's info callback,
 * cuda/reconstruction.cu:69-71). */
int hybrid_mlem_device_name(hybrid_mlem_ctx* ctx, char* name, int size);

/* Frame events (FrameEvent, reconstruction.h:53-63) of the events held:
 * per event lor index, info begin, info end, plane begin | plane end << 16
 * (int32 x 4) and up, right, tan, sec (float x 4).  Either pointer may be
 * NULL. */
int hybrid_mlem_get_frame_events(hybrid_mlem_ctx* ctx, int32_t* ints,
                                 float* floats);

/* Image in the reference's voxel order, index (z n_rows + y) n_columns + x
 * (voxel_grid.h:54-58), n_planes n_rows n_columns floats.  (Replaces
 * output_rho.copy_to_device / copy_from_device, cuda/reconstruction.cu:
 * 89-90,178.) */
int hybrid_mlem_set_rho(hybrid_mlem_ctx* ctx, const float* rho);
int hybrid_mlem_get_rho(hybrid_mlem_ctx* ctx, float* rho);
/* Sensitivity map (n_rows x n_columns) and normalised weights
 * (n_pixel_infos) in use. */
int hybrid_mlem_get_sensitivity(hybrid_mlem_ctx* ctx, float* sensitivity);
int hybrid_mlem_get_weights(hybrid_mlem_ctx* ctx, float* weights);

/* n_iterations MLEM updates (Reconstruction::operator(), reconstruction.h:
 * 276-447; the CUDA loop body, cuda/reconstruction.cu:110-176). */
int hybrid_mlem_iterate(hybrid_mlem_ctx* ctx, int n_iterations);

/* How an iteration runs (no counterpart in the reference, whose CUDA
 * kernel is one warp per event with global atomics,
 * cuda/reconstruction/warp_granularity.cuh:28-173):
 *   1: one wavefront per event: D_e from the full kernel per visit, the
 *      update added to the image with float atomics;
 *   2: LOR space: the frame sensitivity and pixel weight tabulated per
 *      (info, plane) once (n_pixel_infos x n_planes floats, with as many
 *      again for the back-projection's LOR-space sums), one workgroup per
 *      LOR taking its events, their updates summed in LDS over the LOR's
 *      infos x planes, then gathered per voxel over the pixel's infos;
 *   0: automatic (default): 2 where its tables fit in device memory, a
 *      plane of the largest LOR's infos fits in LDS (a LOR whose infos x
 *      planes do not is summed in chunks of planes) and no event spans
 *      more than 256 infos, else 1.
 * hybrid_mlem_get_algorithm returns the algorithm of the last iteration. */
int hybrid_mlem_set_algorithm(hybrid_mlem_ctx* ctx, int algorithm);
int hybrid_mlem_get_algorithm(hybrid_mlem_ctx* ctx);

/* Device time of the last iteration of the last hybrid_mlem_iterate, in
 * ms (HIP events on the context stream around the iteration's kernels). */
int hybrid_mlem_last_iteration_ms(hybrid_mlem_ctx* ctx, float* ms);
/* Sums over the iterations of the last hybrid_mlem_iterate: the event
 * kernels (algorithm 2: the LOR-space iteration kernel and the clearing of
 * the extra sums; algorithm 1: the event kernel) and the gather (algorithm
 * 2; 0 for algorithm 1), in ms, and the iteration count. */
int hybrid_mlem_get_timing(hybrid_mlem_ctx* ctx, double* event_kernels_ms,
                           double* gather_ms, int* n_iterations);

/* Voxel visits (info x plane pairs) of one iteration over the events held:
 * the sum over events of (info end - info begin) (plane end - plane
 * begin). */
int hybrid_mlem_voxel_visits(hybrid_mlem_ctx* ctx, int64_t* visits);

#ifdef __cplusplus
}
#endif
#endif
