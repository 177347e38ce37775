/*
 * strip_phantom.h — seeded event generation for the strip tools on the GPU.
 *
 * Replaces the Monte-Carlo of 2d_strip_phantom
 * (src/2d/strip/phantom_cmd.cpp:136-444, src/common/phantom_monte_carlo.h:
 * 33-109) with the same distributional semantics:
 *   - region chosen with probability weight / total weight, weight =
 *     intensity x area (ellipse, rectangle) or intensity (point)
 *     (src/2d/geometry/phantom.h:44-95,120-133,183-191);
 *   - point uniform in the region: ellipse angle 2 pi U, radius sqrt(U)
 *     (ellipse.h:83-92), rectangle centre + U(-1,1) x half sizes
 *     (rectangle.h:42-47), point regions emit at the point;
 *   - non-additive phantoms draw the point again while it lies in an earlier
 *     region (phantom.h:193-204);
 *   - direction phi = pi/2 U(-1,1), tan = sin(phi)/cos(phi)
 *     (phantom.h:211-215, event.h:25-28);
 *   - exact detection when |z_u|, |z_d| < L/2 on the exact projection
 *     (scanner.h:107-113,196-214), then N(0, sigma_z) added to z_u and z_d and
 *     N(0, sigma_dl) to dl (scanner.h:257-267).
 * The random numbers are Philox4x64-10 (Salmon et al., SC'11), counter =
 * (emission index, attempt, purpose, 0), key = (seed, 0): every emission's
 * outcome depends only on (seed, index), so results do not depend on the
 * launch shape.  The reference's libstdc++ distributions re-seeded through
 * srand48 cannot be reproduced (nor are they by the reference itself), so the
 * stream is this library's own.
 */
#ifndef STRIP_PHANTOM_H
#define STRIP_PHANTOM_H

#include <stdint.h>

#include "strip_mlem.h"

#ifdef __cplusplus
extern "C" {
#endif

#define STRIP_REGION_ELLIPSE 0
#define STRIP_REGION_RECTANGLE 1
#define STRIP_REGION_POINT 2

/* One phantom line: "ellipse x y a b angle intensity" (angle in degrees),
 * "rectangle x y a b intensity" (a, b half sizes), "point x y intensity";
 * x is the z coordinate of the strip scanner.  The CLI's --scale applies to
 * positions and sizes before they reach this structure. */
typedef struct strip_phantom_region {
  int kind;
  double x, y, a, b, angle_deg, intensity;
} strip_phantom_region;

typedef struct strip_phantom_params {
  double radius;              /* R */
  double scintillator_length; /* L */
  double sigma_z;
  double sigma_dl;
  uint64_t seed;
  int additive;      /* 1 = regions add up, 0 = earlier regions win */
  int only_detected; /* 1 = n counts detected events (--detected),
                        0 = n counts emissions */
  /* histogram grid of the emitted / detected images (n_z = 0: none):
   * pixel = floor((p - origin) / pixel) (pixel_grid.h:59-69), origin =
   * (-pixel n_z / 2, -pixel n_y / 2) + image_origin */
  int n_z, n_y;
  double pixel;
  double image_origin_z, image_origin_y;
} strip_phantom_params;

/* Generates n emissions (or, with only_detected, emissions until n are
 * detected) on device `device`.  responses (host, room for 3 n floats)
 * receives the detected responses with error in emission order, exact (may be
 * NULL) the same responses without error.  *n_detected and *n_emitted (may be
 * NULL) receive the counts.  The images (may be NULL, n_y x n_z ints, row
 * index = y pixel) count emission points, detected emission points and the
 * points reconstructed from the responses with error. */
int strip_phantom_generate(int device, const strip_phantom_params* params,
                           const strip_phantom_region* regions, int n_regions,
                           int64_t n, float* responses, float* exact,
                           int64_t* n_detected, int64_t* n_emitted,
                           int* image_emitted, int* image_detected,
                           int* image_detected_w_error);

/* Everything 2d_strip_phantom can write (phantom_cmd.cpp:210-307,330-444):
 * the arrays the caller does not want are NULL. */
typedef struct strip_phantom_outputs {
  float* responses; /* 3 n: detected responses with error (z_u, z_d, dl) */
  float* exact;     /* 3 n: the same without error (= the full response,
                       scanner.h:252-255) */
  float* events;    /* 4 n: emission point (z, y) and direction
                       (sin phi, cos phi) of each detected event
                       (Event(origin, phi), event.h:36-37) */
  int* image_emitted;          /* n_y x n_z, emission points */
  int* image_detected;         /* n_y x n_z, detected emission points */
  int* image_detected_w_error; /* n_y x n_z, points from the responses with
                                  error */
  int tan_bins;                /* > 0: image_tan_bins is filled */
  int* image_tan_bins;         /* tan_bins x n_y x n_z: detected events by the
                                  point and tangent reconstructed from the
                                  response with error, tangent planes of
                                  width 2 max_tan / tan_bins from -max_tan,
                                  max_tan = 1.1 L / (2 R)
                                  (phantom_cmd.cpp:246-252,350-359) */
  int64_t n_detected;          /* out */
  int64_t n_emitted;           /* out */
} strip_phantom_outputs;

/* strip_phantom_generate with every output of the phantom tool
 * (--full dumps and --tan-bins counts).  Replaces the Monte-Carlo loop of
 * phantom_cmd.cpp:300-371. */
int strip_phantom_generate_full(int device, const strip_phantom_params* params,
                                const strip_phantom_region* regions,
                                int n_regions, int64_t n,
                                strip_phantom_outputs* out);

/* The --tan-bins kernel map (phantom_cmd.cpp:393-431): for each voxel
 * (column, row, tangent plane) of the tangent-binned image grid of params
 * (n_z, n_y, pixel, image origin), the AnalyticKernel of an event at the
 * voxel centre with the plane's central tangent, evaluated at the kernel
 * point (kernel_z, kernel_y), times 4 R pixel^2 dtan sqrt(1 + tan^2), and
 * divided by Scanner::sensitivity at the kernel point when
 * params->only_detected.  out: tan_bins x n_y x n_z floats. */
int strip_phantom_tan_kernel_map(int device, const strip_phantom_params* params,
                                 int tan_bins, double kernel_z,
                                 double kernel_y, float* out);

/* Appends n_detected generated events directly to the context's resident
 * events (no host copy; the MI355X path for very large event lists).  The
 * context's own R, L, sigma_z and sigma_dl are used; params supplies seed,
 * additivity (only_detected and the image grid are ignored). */
int strip_mlem_generate_events(strip_mlem_ctx* ctx,
                               const strip_phantom_params* params,
                               const strip_phantom_region* regions,
                               int n_regions, int64_t n_detected);

/* The raw stream: n_blocks Philox4x64-10 blocks for counters
 * (first + i, c1, c2, 0), key (key0, key1), 4 uint64 each, into out (host). */
int strip_phantom_random_blocks(int device, uint64_t key0, uint64_t key1,
                                uint64_t first, uint64_t c1, uint64_t c2,
                                int n_blocks, uint64_t* out);

const char* strip_phantom_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
