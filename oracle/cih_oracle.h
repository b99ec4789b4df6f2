/*
 * cih_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * The CPU oracle: a plain-C restatement of canvas_ity's rasterise-and-
 * composite pipeline (canvas_ity.hpp), used by the test suite, smoke() and
 * bench.py's CPU baseline as the checker of the HIP library.  The product
 * never links it.  Every function in cih_oracle.c cites the reference lines it
 * follows.  How it is pinned (reference build, known-answer vectors, closed
 * forms) is recorded in tests/test_oracle.py.
 */
#ifndef CIH_ORACLE_H
#define CIH_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cih_oracle cih_oracle;

cih_oracle *cih_oracle_create(int width, int height);
void cih_oracle_destroy(cih_oracle *canvas);
/* Replays a command stream (include/cih_stream.h); 0 or -1 if malformed. */
int cih_oracle_replay(cih_oracle *canvas, const void *stream, size_t bytes);
void cih_oracle_get_image_data(cih_oracle *canvas, unsigned char *image, int width,
                               int height, int stride, int x, int y);
void cih_oracle_put_image_data(cih_oracle *canvas, const unsigned char *image, int width,
                               int height, int stride, int x, int y);
/* is_point_in_path(x, y) on the canvas' current path: 1 inside or on it. */
int cih_oracle_is_point_in_path(cih_oracle *canvas, float x, float y);
/* measure_text of `text` after set_font(font, bytes, size) on a fresh canvas. */
float cih_oracle_measure_text(const unsigned char *font, size_t bytes, float size,
                              const char *text);
/* The fp32 linear premultiplied RGBA bitmap. */
void cih_oracle_read_linear(cih_oracle *canvas, float *rgba);
/* Stage seams, laid out as the reference build's (oracle/ref_shim.cpp):
 * lines after path_to_lines (and stroke_lines), lines after dash_lines, and
 * the sorted, merged runs of the current lines.  Up to `capacity` points
 * (x, y) and `subpath_capacity` (count, closed) pairs are written; the
 * return value counts all points (runs), *subpaths all subpaths. */
long cih_oracle_path_lines(cih_oracle *canvas, int stroking, float *points, long capacity,
                           int *subpath_info, long subpath_capacity, long *subpaths);
long cih_oracle_dash_lines(cih_oracle *canvas, float *points, long capacity, int *subpath_info,
                           long subpath_capacity, long *subpaths);
long cih_oracle_lines_to_runs(cih_oracle *canvas, float offset_x, float offset_y, int padding,
                              float *runs, long capacity);

#ifdef __cplusplus
}
#endif

#endif /* CIH_ORACLE_H */
