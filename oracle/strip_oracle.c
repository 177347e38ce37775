/*
 * CPU oracle for the J-PET 2D strip list-mode TOF MLEM path.
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; the product never links it.
 *
 * Pinned by (see tests/test_oracle.py):
 *   - the reference's own known-answer tests: GaussianKernel values
 *     (src/2d/strip/gaussian_kernel_test.cpp:11-51), Scanner::sensitivity
 *     (kernel_test.cpp:31-49), bb_y/bb_z (kernel_test.cpp:81-112), pixel_at /
 *     pixel_center (scanner_test.cpp:15-66), projection conversions
 *     (response_test.cpp:11-55);
 *   - the reference itself, compiled from its headers into oracle/_ref
 *     (ref_harness.cpp) where /root/reference is present;
 *   - closed forms and invariants: Q identity, mass conservation,
 *     single-event closed form, linearity, brute-force support enumeration.
 */
#define _GNU_SOURCE
#include "strip_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef M_1_PI
#define M_1_PI 0.31830988618379067154
#endif

#define REAL float
#define SFX(x) x##_f32
#define SQRT sqrtf
#define EXP expf
#define ATAN atanf
#include "strip_oracle_impl.inc"
#undef REAL
#undef SFX
#undef SQRT
#undef EXP
#undef ATAN

#define REAL double
#define SFX(x) x##_f64
#define SQRT sqrt
#define EXP exp
#define ATAN atan
#include "strip_oracle_impl.inc"
#undef REAL
#undef SFX
#undef SQRT
#undef EXP
#undef ATAN

/* GaussianKernel::bb_y / bb_z (gaussian_kernel.h:159-160) */
double oracle_bb_y_f64(double A, double C, double B_2) {
  return 3 / sqrt(A - (B_2 / C));
}
double oracle_bb_z_f64(double A, double C, double B_2) {
  return 3 / sqrt(C - (B_2 / A));
}

/* Scanner::to_projection_space_tan (scanner.h:107-113) */
void oracle_to_projection_f64(double R, double y, double z, double tan,
                              double* z_u, double* z_d, double* dl) {
  *z_u = z + (R - y) * tan;
  *z_d = z - (R + y) * tan;
  *dl = -2 * y * sqrt(tan * tan + 1);
}

/* Support statistics: B_e = bbox pixels visited by the loop at
 * reconstruction.h:255-256, P_e = those passing in_ellipse
 * (gaussian_kernel.h:131-137), i.e. kernel evaluations per event-update. */
int oracle_support_counts_f32(const oracle_geometry* g, const float* events,
                              int64_t n_events, int64_t* bbox_pixels,
                              int64_t* ellipse_pixels, int n_threads) {
  scanner_t_f32 s;
  make_scanner_f32(g, &s);
  int nt = set_threads_f32(n_threads);
  (void)nt;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256) num_threads(nt)
#endif
  for (int64_t e = 0; e < n_events; ++e) {
    event_t_f32 ev;
    setup_event_f32(&s, events[3 * e], events[3 * e + 1], events[3 * e + 2],
                    &ev);
    int64_t nb = 0, ne = 0;
    for (int iy = ev.y0; iy < ev.y1; ++iy)
      for (int ix = ev.x0; ix < ev.x1; ++ix) {
        float pz, py;
        pixel_center_s_f32(&s, ix, iy, &pz, &py);
        ++nb;
        ne += in_ellipse_f32(ev.A, ev.B, ev.C, py - ev.y, pz - ev.z);
      }
    if (bbox_pixels) bbox_pixels[e] = nb;
    if (ellipse_pixels) ellipse_pixels[e] = ne;
  }
  return 0;
}

/* Poisson list-mode log-likelihood of the image, FP64:
 *   L(rho) = sum_{e: D_e > 0} log D_e - sum_p w(p) rho(p)
 * with D_e the forward projection above and w the sensitivity weights
 * (all ones when sens_weights is NULL). */
double oracle_loglik_f64(const oracle_geometry* g, const oracle_mode* m,
                         const double* events, int64_t n_events,
                         const double* rho, const double* sens_weights,
                         int n_threads) {
  double* d = (double*)malloc(sizeof(double) * (n_events > 0 ? n_events : 1));
  if (!d) return NAN;
  oracle_forward_f64(g, m, events, n_events, rho, d, n_threads);
  double acc = 0;
  for (int64_t e = 0; e < n_events; ++e)
    if (d[e] > 0) acc += log(d[e]);
  free(d);
  int64_t np = (int64_t)g->n_z_pixels * g->n_y_pixels;
  for (int64_t i = 0; i < np; ++i)
    acc -= (sens_weights ? sens_weights[i] : 1.0) * rho[i];
  return acc;
}
