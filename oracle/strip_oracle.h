/*
 * CPU oracle for the J-PET 2D strip list-mode TOF MLEM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is called by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * never by the product path.  It restates, in plain C, the reference's CPU
 * (OpenMP) reconstruction and the helpers it uses; every function cites the
 * reference file:line it follows (paths relative to the reference root).
 *
 * Each function exists in an FP32 (`_f32`) and an FP64 (`_f64`) flavour.  The
 * FP32 flavour reproduces the reference's float arithmetic, including its
 * double-precision constant folding where the reference mixes `M_PI` (double)
 * into float expressions.
 */
#ifndef STRIP_ORACLE_H
#define STRIP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scanner parameters as the CLI hands them to `Scanner(...)`
 * (src/2d/strip/options.h:42-50, scanner.h:39-67).  Values are converted to
 * the flavour's float type exactly as the reference's constructor arguments
 * are. */
typedef struct {
  double radius;              /* R: half distance between strips */
  double scintillator_length; /* L */
  int n_z_pixels;
  int n_y_pixels;
  double pixel_width;  /* along z */
  double pixel_height; /* along y */
  double sigma_z;
  double sigma_dl;
  double center_y;
  double center_z;
} oracle_geometry;

/* Reconstruction semantics.  {0,0,0,1} is the reference CPU path
 * (reconstruction.h:220-298); {1,1,1,0} is the legacy CUDA warp path
 * (cuda/reconstruction/warp_granularity.cuh:42-155). */
typedef struct {
  int kernel;          /* 0 = GaussianKernel, 1 = AnalyticKernel */
  int sensitivity;     /* 0 = none, 1 = in denominator, 2 = divides numerator */
  int sensitivity_map; /* 0 = sensitivity(centre), 1 = 5-point pixel_sensitivity */
  int guard_zero;      /* 1 = inv = D>0 ? 1/D : 0 ; 0 = inv = 1/D */
} oracle_mode;

/* ---- per-event geometry (response.h:22-37, gaussian_kernel.h:108-128) ---- */
void oracle_event_params_f32(const oracle_geometry* g, float z_u, float z_d,
                             float dl, float* out /* tan,y,z,sec,A,B,C,bb_y,bb_z */,
                             int* bbox /* x0,y0,x1,y1 (clamped, x1/y1 exclusive) */);
void oracle_event_params_f64(const oracle_geometry* g, double z_u, double z_d,
                             double dl, double* out, int* bbox);

/* ---- kernels (gaussian_kernel.h:51-79, analytic_kernel.h:28-62) ---- */
float oracle_gaussian_kernel_f32(float sigma_z, float sigma_dl, float y,
                                 float tan, float sec, float R, float dz,
                                 float dy);
double oracle_gaussian_kernel_f64(double sigma_z, double sigma_dl, double y,
                                  double tan, double sec, double R, double dz,
                                  double dy);
float oracle_analytic_kernel_f32(float sigma_z, float sigma_dl, float y,
                                 float tan, float sec, float R, float dz,
                                 float dy);
double oracle_analytic_kernel_f64(double sigma_z, double sigma_dl, double y,
                                  double tan, double sec, double R, double dz,
                                  double dy);
/* bb_y/bb_z helpers (gaussian_kernel.h:159-160) */
double oracle_bb_y_f64(double A, double C, double B_2);
double oracle_bb_z_f64(double A, double C, double B_2);

/* ---- scanner (scanner.h:136-185) ---- */
void oracle_pixel_at_f32(const oracle_geometry* g, float z, float y, int* px,
                         int* py);
void oracle_pixel_at_f64(const oracle_geometry* g, double z, double y, int* px,
                         int* py);
void oracle_pixel_center_f32(const oracle_geometry* g, int px, int py,
                             float* z, float* y);
void oracle_pixel_center_f64(const oracle_geometry* g, int px, int py,
                             double* z, double* y);
float oracle_sensitivity_f32(const oracle_geometry* g, float z, float y);
double oracle_sensitivity_f64(const oracle_geometry* g, double z, double y);
/* whole map, row-major y*n_z+z; kind 0 = centre, 1 = 5-point */
void oracle_sensitivity_map_f32(const oracle_geometry* g, int kind, float* out);
void oracle_sensitivity_map_f64(const oracle_geometry* g, int kind,
                                double* out);

/* ---- projection-space conversions (scanner.h:107-127, event.h:25-42) ---- */
void oracle_to_projection_f64(double R, double y, double z, double tan,
                              double* z_u, double* z_d, double* dl);

/* ---- MLEM (reconstruction.h:110-154 + 220-298) ----
 * events: AoS (z_u, z_d, dl) triples.  rho: n_y*n_z row-major, updated in
 * place after n_iterations.  n_threads <= 0 uses OpenMP's default; 1 is a
 * deterministic serial run.  Returns 0 on success. */
int oracle_mlem_f32(const oracle_geometry* g, const oracle_mode* m,
                    const float* events, int64_t n_events, float* rho,
                    int n_iterations, int n_threads);
int oracle_mlem_f64(const oracle_geometry* g, const oracle_mode* m,
                    const double* events, int64_t n_events, double* rho,
                    int n_iterations, int n_threads);

/* Forward projection only: D_e = sum_p K_e(p) rho(p) [* s(p)] per event. */
int oracle_forward_f32(const oracle_geometry* g, const oracle_mode* m,
                       const float* events, int64_t n_events, const float* rho,
                       float* denominators, int n_threads);
/* Ordered subsets: per iteration and subset s (events [N s/S, N (s+1)/S)),
 * the reference update over the subset times N / N_s (in double). */
int oracle_osem_f32(const oracle_geometry* g, const oracle_mode* m,
                    const float* events, int64_t n_events, float* rho,
                    int n_iterations, int n_subsets, int n_threads);
int oracle_osem_f64(const oracle_geometry* g, const oracle_mode* m,
                    const double* events, int64_t n_events, double* rho,
                    int n_iterations, int n_subsets, int n_threads);

int oracle_forward_f64(const oracle_geometry* g, const oracle_mode* m,
                       const double* events, int64_t n_events,
                       const double* rho, double* denominators, int n_threads);

/* Per-event support statistics in FP32 semantics: bbox pixel count B_e and
 * in-ellipse pixel count P_e (the kernel-evaluation count). */
int oracle_support_counts_f32(const oracle_geometry* g, const float* events,
                              int64_t n_events, int64_t* bbox_pixels,
                              int64_t* ellipse_pixels, int n_threads);

/* Poisson list-mode log-likelihood  sum_e log(D_e) - sum_p s(p) rho(p)
 * in FP64 (the adjoint's objective). */
double oracle_loglik_f64(const oracle_geometry* g, const oracle_mode* m,
                         const double* events, int64_t n_events,
                         const double* rho, const double* sens_weights,
                         int n_threads);

/* ---- differentiation checker (FP64) ----
 * The kernel is evaluated at sigma = {sigma_z, sigma_dl} while every event
 * keeps the support of the geometry's own sigmas (the pixel sums are smooth in
 * sigma at fixed support; their derivative is the almost-everywhere derivative
 * of the reference's objective).
 *
 * L = sum_{e: D_e > 0} log D_e - sum_p rho(p) with its gradients d L/d rho
 * (n_y x n_z row-major, may be NULL) and d L / d{sigma_z, sigma_dl}. */
int oracle_loglik_grad_f64(const oracle_geometry* g, const double* events,
                           int64_t n_events, const double* rho,
                           const double* sigma, double* value,
                           double* grad_rho, double* grad_sigma);

/* Unrolled MLEM loss: rho_{k+1} = rho_k * sum_e K_e / D_e (guarded) for
 * n_iterations steps from rho0, loss = mean_p (rho_N - target)^2.  When jvp is
 * not NULL and a tangent (d_rho0 and/or d_sigma) is given, *jvp receives the
 * directional derivative of the loss along it (forward-mode propagation).
 * rho_out (may be NULL) receives rho_N. */
int oracle_mlem_loss_f64(const oracle_geometry* g, const double* events,
                         int64_t n_events, const double* rho0,
                         int n_iterations, const double* target,
                         const double* sigma, const double* d_rho0,
                         const double* d_sigma, double* loss, double* jvp,
                         double* rho_out);

/* ---- 3D hybrid reconstruction (src/3d/hybrid/reconstruction.h:32-549) ----
 * Geometry in the reference's GeometrySOA layout (geometry_soa.h:262-287):
 * per LOR (index first (first + 1) / 2 + second, lor.h:36-38) the segment
 * from the second detector's centre to the first's (4 floats), its width,
 * and the range [info_begin, info_end) of its pixel infos (pixel x, y as
 * int16 pairs, scaled position along the segment, sorted).  The voxel grid
 * is the pixel grid centred on the origin (pixel_grid.h:33-40) times
 * n_planes planes from z_left (voxel_grid.h:22-59). */
typedef struct {
  int n_detectors;
  int64_t n_infos;
  int n_planes_half;
  const float* seg;
  const float* widths;
  const int16_t* pixels;
  const float* positions;
  const int64_t* info_begin;
  const int64_t* info_end;
  int n_columns, n_rows;
  double pixel_size;
  int n_planes;
  double z_left;
  double sigma_z, sigma_dl;
  double length;
} oracle_hybrid;

/* The CLI's preparation (reconstruction_cmd.cpp:459-560) and n_iterations of
 * Reconstruction::operator() (reconstruction.h:274-445): weights (n_infos x
 * n_planes_half) as given (weights_mode 0) or from calculate_weight (1,
 * plane 0); sensitivity from calculate_sensitivity (sens_mode 0), ones (1)
 * or sens_in (2, depth planes); normalize_geometry_weights; events (first,
 * second, z_up, z_dn, dl per event) through translate_to_frame and
 * bb_intersects_grid_with_positive_weight.  rho: n_planes x n_rows x
 * n_columns voxels, in and out.  n_used: events kept; sens_out (depth
 * planes) and weights_out may be NULL.  legacy = 0: the CPU path's update
 * (each visit adds the voxel's cached weight, the last one written);
 * legacy = 1: the CUDA kernel's (each visit adds its own weight,
 * cuda/reconstruction/warp_granularity.cuh:136-168). */
int oracle_hybrid_mlem_f32(const oracle_hybrid* h, const float* weights,
                           int weights_mode, int sens_mode,
                           int use_3d_sensitivity, const float* sens_in,
                           const float* events, int64_t n_events, float* rho,
                           int n_iterations, int n_threads, int64_t* n_used,
                           float* sens_out, float* weights_out, int legacy);
int oracle_hybrid_mlem_f64(const oracle_hybrid* h, const float* weights,
                           int weights_mode, int sens_mode,
                           int use_3d_sensitivity, const float* sens_in,
                           const float* events, int64_t n_events, double* rho,
                           int n_iterations, int n_threads, int64_t* n_used,
                           double* sens_out, double* weights_out, int legacy);

/* translate_to_frame and the box filter of every event (FP32), against
 * normalised weights: per event lor, info begin, info end (the reference's
 * upper_bound + 1, not clipped), plane begin, plane end (ints, 5 per
 * event), up, right, tan, sec (4 floats) and the filter's verdict. */
int oracle_hybrid_frames(const oracle_hybrid* h, const float* weights,
                         const float* events, int64_t n_events,
                         int64_t* ints, float* floats, int8_t* keep);

#ifdef __cplusplus
}
#endif
#endif
