/*
 * strip_events.h — event-list ingestion of the strip tools (host only).
 *
 * Replaces the readers of PET2D::Strip::Reconstruction:
 *   operator<< on util::ibstream  (binary float32 triples,
 *                                  src/2d/strip/reconstruction.h:156-167)
 *   fast_load_txt_events          (text lines "z_u z_d dl" or, with is_3d,
 *                                  "i j z_u z_d dl"; reconstruction.h:170-195
 *                                  with util::text_parser, text_parser.h:29-175)
 * and the unit scaling of reconstruction_cmd.cpp:194-210.  The text reader
 * accumulates digits exactly like the reference's parser, so both produce
 * the same float values (that parser is not correctly rounded).
 */
#ifndef STRIP_EVENTS_H
#define STRIP_EVENTS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reads every event of `path` (".txt" = text, anything else = binary) and
 * appends them to a buffer owned by the library: *events receives n*3 floats
 * (z_u, z_d, dl), release it with strip_events_free.  Returns 0, or -1 with
 * strip_mlem_last_error()-style text in strip_events_last_error(). */
int strip_events_read(const char* path, int is_3d, float** events,
                      int64_t* n_events);
void strip_events_free(float* events);

/* In-place unit conversion (reconstruction_cmd.cpp:194-210): lengths times
 * scale_length (dl too unless dl_is_time); then, if dl_is_time,
 * dl = dl * speed_of_light * scale_time. */
void strip_events_scale(float* events, int64_t n_events, int apply_length,
                        double scale_length, int dl_is_time,
                        double speed_of_light, double scale_time);

const char* strip_events_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
