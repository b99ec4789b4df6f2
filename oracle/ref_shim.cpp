// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Compiles the unmodified reference header (canvas_ity.hpp, read where it lies
// under the reference checkout; nothing is copied) and exposes it through a
// C ABI so the checker can replay command streams into the real reference,
// read its private fp32 bitmap (hpp:1171) and call its private stage
// functions (hpp:1175-1188) as differential-test seams.  Built by
// oracle/Makefile into oracle/_ref/libcih_ref.so; never linked by the product.

// Standard headers first, so that the access override below only reaches the
// reference's own class.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <numeric>
#include <vector>

#define private public
#define CANVAS_ITY_IMPLEMENTATION
#include "canvas_ity.hpp"
#undef private

#include "cih_replay.hpp"

using canvas_ity::canvas;

typedef cih::replayer<canvas, canvas_ity::brush_type, canvas_ity::repetition_style,
                      canvas_ity::composite_operation, canvas_ity::cap_style,
                      canvas_ity::join_style>
    ref_replayer;

extern "C" {

void *cih_ref_create(int width, int height)
{
    if (width < 1 || height < 1 || width > 32768 || height > 32768)
        return 0;
    return new canvas(width, height);
}

void cih_ref_destroy(void *handle) { delete static_cast<canvas *>(handle); }

int cih_ref_width(void *handle) { return static_cast<canvas *>(handle)->size_x; }
int cih_ref_height(void *handle) { return static_cast<canvas *>(handle)->size_y; }

// Replays a command stream (its header's size is ignored: the canvas exists).
int cih_ref_replay(void *handle, const void *stream, size_t bytes)
{
    ref_replayer replay(*static_cast<canvas *>(handle));
    return replay.run(stream, bytes);
}

void cih_ref_get_image_data(void *handle, unsigned char *image, int width,
                            int height, int stride, int x, int y)
{
    static_cast<canvas *>(handle)->get_image_data(image, width, height, stride, x, y);
}

// Copies the private linear premultiplied fp32 bitmap (hpp:1171), RGBA order.
void cih_ref_read_linear(void *handle, float *out)
{
    canvas *c = static_cast<canvas *>(handle);
    std::memcpy(out, c->bitmap,
                sizeof(float) * 4 * (size_t)c->size_x * (size_t)c->size_y);
}

// Stage seam: flattens the current path (hpp:1495) and, when stroking, expands
// it (hpp:2070).  Writes up to `capacity` points (x,y pairs) and up to
// `subpath_capacity` (count, closed) pairs; returns the number of points and
// stores the number of subpaths in *subpaths.
long cih_ref_path_lines(void *handle, int stroking, float *points, long capacity,
                        int *subpath_info, long subpath_capacity, long *subpaths)
{
    canvas *c = static_cast<canvas *>(handle);
    c->path_to_lines(stroking != 0);
    if (stroking)
        c->stroke_lines();
    long n = (long)c->lines.points.size();
    for (long k = 0; k < n && k < capacity; ++k) {
        points[2 * k + 0] = c->lines.points[(size_t)k].x;
        points[2 * k + 1] = c->lines.points[(size_t)k].y;
    }
    long s = (long)c->lines.subpaths.size();
    for (long k = 0; k < s && k < subpath_capacity; ++k) {
        subpath_info[2 * k + 0] = (int)c->lines.subpaths[(size_t)k].count;
        subpath_info[2 * k + 1] = c->lines.subpaths[(size_t)k].closed ? 1 : 0;
    }
    *subpaths = s;
    return n;
}

// Stage seam: flattens for stroking (hpp:1495) and applies the dash pattern
// (hpp:1858) without stroke expansion.
long cih_ref_dash_lines(void *handle, float *points, long capacity, int *subpath_info,
                        long subpath_capacity, long *subpaths)
{
    canvas *c = static_cast<canvas *>(handle);
    c->path_to_lines(true);
    c->dash_lines();
    long n = (long)c->lines.points.size();
    for (long k = 0; k < n && k < capacity; ++k) {
        points[2 * k + 0] = c->lines.points[(size_t)k].x;
        points[2 * k + 1] = c->lines.points[(size_t)k].y;
    }
    long s = (long)c->lines.subpaths.size();
    for (long k = 0; k < s && k < subpath_capacity; ++k) {
        subpath_info[2 * k + 0] = (int)c->lines.subpaths[(size_t)k].count;
        subpath_info[2 * k + 1] = c->lines.subpaths[(size_t)k].closed ? 1 : 0;
    }
    *subpaths = s;
    return n;
}

// Stage seam: scan-converts the current polylines (hpp:2193) with the given
// offset and padding; writes up to `capacity` runs as (x, y, delta) triples
// (x and y stored as floats) and returns the number of runs.
long cih_ref_lines_to_runs(void *handle, float offset_x, float offset_y, int padding,
                           float *runs, long capacity)
{
    canvas *c = static_cast<canvas *>(handle);
    c->lines_to_runs(canvas_ity::xy(offset_x, offset_y), padding);
    long n = (long)c->runs.size();
    for (long k = 0; k < n && k < capacity; ++k) {
        runs[3 * k + 0] = (float)c->runs[(size_t)k].x;
        runs[3 * k + 1] = (float)c->runs[(size_t)k].y;
        runs[3 * k + 2] = c->runs[(size_t)k].delta;
    }
    return n;
}

// The current path in device space as the reference stores it (hpp:1162):
// up to `capacity` points and `subpath_capacity` (count, closed) pairs;
// returns the number of points, *subpaths the number of subpaths.
long cih_ref_path(void *handle, float *points, long capacity, int *subpath_info,
                  long subpath_capacity, long *subpaths)
{
    canvas *c = static_cast<canvas *>(handle);
    long n = (long)c->path.points.size();
    for (long k = 0; k < n && k < capacity; ++k) {
        points[2 * k + 0] = c->path.points[(size_t)k].x;
        points[2 * k + 1] = c->path.points[(size_t)k].y;
    }
    long s = (long)c->path.subpaths.size();
    for (long k = 0; k < s && k < subpath_capacity; ++k) {
        subpath_info[2 * k + 0] = (int)c->path.subpaths[(size_t)k].count;
        subpath_info[2 * k + 1] = c->path.subpaths[(size_t)k].closed ? 1 : 0;
    }
    *subpaths = s;
    return n;
}

// The reference's set_font (hpp:3205-3274) on a fresh canvas: whether it
// accepts the font, and measure_text (hpp:3297-3311) after it.
int cih_ref_set_font(const unsigned char *font, int bytes, float size, const char *text,
                     float *width)
{
    canvas c(0, 0);
    bool ok = c.set_font(font, bytes, size);
    *width = c.measure_text(text);
    return ok ? 1 : 0;
}

// The reference's is_point_in_path (hpp:3101-3132) on the current path.
int cih_ref_is_point_in_path(void *handle, float x, float y)
{
    return static_cast<canvas *>(handle)->is_point_in_path(x, y) ? 1 : 0;
}

} // extern "C"
