/*
 * CPU oracle for the J-PET 2D strip list-mode TOF MLEM path and the 3D
 * hybrid reconstruction built on its kernel.
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
 *     single-event closed form, linearity, brute-force support enumeration;
 *   - the legacy --gpu composition (warp_granularity.cuh:53-152, not
 *     buildable here): its single-event closed form with the analytic kernel
 *     and the 5-point sensitivity in the denominator, the NaN of its
 *     unguarded 1/D, and the pixels its float-reciprocal row split visits;
 *   - the FP64 differentiation checker: central finite differences
 *     (tests/test_oracle_diff.py);
 *   - the 3D hybrid restatement (hybrid_oracle_impl.inc): the reference's
 *     own PET3D::Hybrid::Reconstruction built into oracle/_ref, on the same
 *     geometry arrays and events (weights, sensitivity, kept events, images
 *     including a repeated voxel visit); calculate_weight against an FP64
 *     closed form; the event-count identity of one update
 *     (tests/test_hybrid_oracle.py).  Its legacy (CUDA kernel) update is
 *     pinned by equality with the cached one where no visit repeats and by
 *     the closed-form ratio 2 w_x / (w_j + w_x) on the repeated voxel.
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
#define HY_FRAMES_ONCE 1
#include "strip_oracle_impl.inc"
#include "hybrid_oracle_impl.inc"
#undef HY_FRAMES_ONCE
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
#include "hybrid_oracle_impl.inc"
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

/* ======================================================================
 * Differentiation checker (FP64).
 *
 * The Gaussian kernel (gaussian_kernel.h:51-79) written with the inverse
 * covariance diag(iz, iz, id), iz = 1/sigma_z^2, id = 1/sigma_dl^2:
 *   K = c / sqrt(a.a) * exp(-(Qz + Qd) / 2),   c = iz sqrt(id) / (2 pi^2)
 *   a.a = iz (ap^2 + aq^2) + id ar^2,  Qz = iz * 2 (dz - dy t)^2,
 *   Qd = id * (2 dy sec)^2
 * so, with w = iz (ap^2 + aq^2) / a.a,
 *   d ln K / d ln sigma_z  = -2 + w + Qz,
 *   d ln K / d ln sigma_dl = -w + Qd.
 * The kernel is evaluated at sigma = (sigma_z, sigma_dl) given by the caller
 * while every event keeps the support (clamped bbox, Q <= 9) of the
 * geometry's own sigmas: the objectives below are then smooth in sigma and
 * their derivative is the almost-everywhere derivative of the pixel sums.
 * ====================================================================== */

typedef struct {
  double iz, id, c;
} diff_sigma_t;

static void diff_sigma(const double* sigma, diff_sigma_t* k) {
  k->iz = 1 / (sigma[0] * sigma[0]);
  k->id = 1 / (sigma[1] * sigma[1]);
  k->c = 1 / (2 * M_PI * M_PI) * sqrt(k->iz * k->iz * k->id);
}

/* K at the caller's sigma and its two log-derivatives. */
static double diff_kernel(const diff_sigma_t* k, const event_t_f64* ev,
                          double R, double dz, double dy, double* gz,
                          double* gd) {
  double sec_sq = ev->sec * ev->sec;
  double yp = dy + ev->y;
  double ap = -(yp - R) * sec_sq, aq = -(yp + R) * sec_sq;
  double ar = -2 * yp * ev->sec * ev->tan;
  double b_pq = dz - dy * ev->tan, br = -2 * dy * ev->sec;
  double alpha = ap * ap + aq * aq, beta = ar * ar;
  double aa = k->iz * alpha + k->id * beta;
  double qz = k->iz * 2 * b_pq * b_pq, qd = k->id * br * br;
  double w = k->iz * alpha / aa;
  if (gz) *gz = -2 + w + qz;
  if (gd) *gd = -w + qd;
  return k->c / sqrt(aa) * exp(-0.5 * (qz + qd));
}

/* Visits the support pixels of one event: calls fn(idx, K, gz, gd, user). */
typedef void (*diff_visit_fn)(int64_t idx, double K, double gz, double gd,
                              void* user);

static void diff_event(const scanner_t_f64* s, const diff_sigma_t* k,
                       const double* e3, diff_visit_fn fn, void* user) {
  event_t_f64 ev;
  setup_event_f64(s, e3[0], e3[1], e3[2], &ev);
  for (int iy = ev.y0; iy < ev.y1; ++iy)
    for (int ix = ev.x0; ix < ev.x1; ++ix) {
      double pz, py;
      pixel_center_s_f64(s, ix, iy, &pz, &py);
      double dy = py - ev.y, dz = pz - ev.z;
      if (!in_ellipse_f64(ev.A, ev.B, ev.C, dy, dz)) continue;
      double gz, gd;
      double K = diff_kernel(k, &ev, s->radius, dz, dy, &gz, &gd);
      fn((int64_t)iy * s->n_z + ix, K, gz, gd, user);
    }
}

/* per-event forward sums of up to two images with the log-derivative
 * weights */
typedef struct {
  const double* x0;
  const double* x1;
  double t0, z0, y0, t1, z1, y1;
} diff_fwd_t;

static void diff_fwd_visit(int64_t idx, double K, double gz, double gd,
                           void* user) {
  diff_fwd_t* f = (diff_fwd_t*)user;
  double k0 = K * f->x0[idx];
  f->t0 += k0;
  f->z0 += k0 * gz;
  f->y0 += k0 * gd;
  if (f->x1) {
    double k1 = K * f->x1[idx];
    f->t1 += k1;
    f->z1 += k1 * gz;
    f->y1 += k1 * gd;
  }
}

typedef struct {
  double* out;
  double wt;
  const double* x;   /* optional per-pixel factor of the tangent term */
  double wz, wd;     /* tangent weights of the log-derivatives */
  double* out2;
  double w2;
} diff_bwd_t;

/* out[idx] += K * wt ;  out2[idx] += K * (wz gz + wd gd) * w2 */
static void diff_bwd_visit(int64_t idx, double K, double gz, double gd,
                           void* user) {
  diff_bwd_t* b = (diff_bwd_t*)user;
  b->out[idx] += K * b->wt;
  if (b->out2) b->out2[idx] += K * (b->wz * gz + b->wd * gd) * b->w2;
}

/* Events are split over the OpenMP threads (OMP_NUM_THREADS); each thread
 * keeps its own sums and gradient image, added in thread order at the end. */
int oracle_loglik_grad_f64(const oracle_geometry* g, const double* events,
                           int64_t n_events, const double* rho,
                           const double* sigma, double* value,
                           double* grad_rho, double* grad_sigma) {
  scanner_t_f64 s;
  make_scanner_f64(g, &s);
  diff_sigma_t k;
  diff_sigma(sigma, &k);
  int64_t np = (int64_t)s.n_z * s.n_y;
  int nt = omp_get_max_threads();
  double* part = (double*)calloc((size_t)3 * nt, sizeof(double));
  double* imgs = grad_rho ? (double*)calloc((size_t)np * nt, sizeof(double))
                          : NULL;
  if (!part || (grad_rho && !imgs)) {
    free(part);
    free(imgs);
    return -1;
  }
#pragma omp parallel num_threads(nt)
  {
    int t = omp_get_thread_num();
    double* gimg = imgs ? imgs + np * t : NULL;
    double L = 0, gz = 0, gd = 0;
#pragma omp for schedule(dynamic, 64)
    for (int64_t e = 0; e < n_events; ++e) {
      diff_fwd_t f = {rho, NULL, 0, 0, 0, 0, 0, 0};
      diff_event(&s, &k, events + 3 * e, diff_fwd_visit, &f);
      if (!(f.t0 > 0)) continue;
      L += log(f.t0);
      gz += f.z0 / f.t0;
      gd += f.y0 / f.t0;
      if (gimg) {
        diff_bwd_t b = {gimg, 1 / f.t0, NULL, 0, 0, NULL, 0};
        diff_event(&s, &k, events + 3 * e, diff_bwd_visit, &b);
      }
    }
    part[3 * t] = L;
    part[3 * t + 1] = gz;
    part[3 * t + 2] = gd;
  }
  double L = 0, gz = 0, gd = 0;
  for (int t = 0; t < nt; ++t) {
    L += part[3 * t];
    gz += part[3 * t + 1];
    gd += part[3 * t + 2];
  }
  for (int64_t i = 0; i < np; ++i) L -= rho[i];
  if (grad_rho) {
    for (int64_t i = 0; i < np; ++i) {
      double acc = 0;
      for (int t = 0; t < nt; ++t) acc += imgs[np * t + i];
      grad_rho[i] = acc - 1;
    }
  }
  free(part);
  free(imgs);
  if (value) *value = L;
  if (grad_sigma) {
    grad_sigma[0] = gz / sigma[0];
    grad_sigma[1] = gd / sigma[1];
  }
  return 0;
}

/* Forward-mode (tangent) propagation through the iterations:
 *   D_e = sum K rho,  D'_e = sum (K' rho + K rho')
 *   rho_{k+1} = rho_k * sum_e K / D_e
 *   rho'_{k+1} = rho'_k * sum_e K / D_e
 *              + rho_k * sum_e (K' / D_e - K D'_e / D_e^2)
 * with K' = K (dz gz + dd gd), (dz, dd) = d_sigma / sigma. */
int oracle_mlem_loss_f64(const oracle_geometry* g, const double* events,
                         int64_t n_events, const double* rho0,
                         int n_iterations, const double* target,
                         const double* sigma, const double* d_rho0,
                         const double* d_sigma, double* loss, double* jvp,
                         double* rho_out) {
  scanner_t_f64 s;
  make_scanner_f64(g, &s);
  diff_sigma_t k;
  diff_sigma(sigma, &k);
  int64_t np = (int64_t)s.n_z * s.n_y;
  int tangent = (d_rho0 || d_sigma) && jvp;
  double wz = d_sigma ? d_sigma[0] / sigma[0] : 0;
  double wd = d_sigma ? d_sigma[1] / sigma[1] : 0;
  double* rho = (double*)malloc(sizeof(double) * np);
  double* drho = (double*)calloc(np, sizeof(double));
  double* b1 = (double*)malloc(sizeof(double) * np);  /* sum_e K / D */
  double* b2 = (double*)malloc(sizeof(double) * np);  /* sum_e (K'/D - K D'/D^2) */
  if (!rho || !drho || !b1 || !b2) {
    free(rho);
    free(drho);
    free(b1);
    free(b2);
    return -1;
  }
  memcpy(rho, rho0, sizeof(double) * np);
  if (d_rho0) memcpy(drho, d_rho0, sizeof(double) * np);
  int nt = omp_get_max_threads();
  /* per-thread b1 and b2 images, added in thread order */
  double* priv = (double*)malloc(sizeof(double) * (size_t)np * 2 * nt);
  if (!priv) {
    free(rho);
    free(drho);
    free(b1);
    free(b2);
    return -1;
  }
  for (int it = 0; it < n_iterations; ++it) {
    memset(priv, 0, sizeof(double) * (size_t)np * 2 * nt);
#pragma omp parallel num_threads(nt)
    {
      int t = omp_get_thread_num();
      double* p1 = priv + (size_t)np * 2 * t;
      double* p2 = p1 + np;
#pragma omp for schedule(dynamic, 64)
      for (int64_t e = 0; e < n_events; ++e) {
        diff_fwd_t f = {rho, tangent ? drho : NULL, 0, 0, 0, 0, 0, 0};
        diff_event(&s, &k, events + 3 * e, diff_fwd_visit, &f);
        if (!(f.t0 > 0)) continue;
        double D = f.t0;
        /* D' = sum K' rho + sum K rho' = wz z0 + wd y0 + t1 */
        double dD = wz * f.z0 + wd * f.y0 + f.t1;
        diff_bwd_t b = {p1, 1 / D, NULL, wz, wd, tangent ? p2 : NULL, 1 / D};
        diff_event(&s, &k, events + 3 * e, diff_bwd_visit, &b);
        if (tangent) {
          /* the K D'/D^2 part of b2 */
          diff_bwd_t c = {p2, -dD / (D * D), NULL, 0, 0, NULL, 0};
          diff_event(&s, &k, events + 3 * e, diff_bwd_visit, &c);
        }
      }
    }
    for (int64_t i = 0; i < np; ++i) {
      double a1 = 0, a2 = 0;
      for (int t = 0; t < nt; ++t) {
        a1 += priv[(size_t)np * 2 * t + i];
        a2 += priv[(size_t)np * 2 * t + np + i];
      }
      b1[i] = a1;
      b2[i] = a2;
    }
    for (int64_t i = 0; i < np; ++i) {
      double r = rho[i];
      rho[i] = r * b1[i];
      if (tangent) drho[i] = drho[i] * b1[i] + r * b2[i];
    }
  }
  free(priv);
  double l = 0, j = 0;
  for (int64_t i = 0; i < np; ++i) {
    double d = rho[i] - target[i];
    l += d * d;
    j += 2 * d * drho[i];
  }
  if (loss) *loss = l / np;
  if (jvp) *jvp = tangent ? j / np : 0;
  if (rho_out) memcpy(rho_out, rho, sizeof(double) * np);
  free(rho);
  free(drho);
  free(b1);
  free(b2);
  return 0;
}
