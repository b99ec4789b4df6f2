/*
 * cih_abi.h -- C ABI of libcanvas_ity_hip.so, the MI355X rasteriser behind
 * canvas_ity_hip.hpp.
 *
 * The reference (canvas_ity.hpp) is a single C++ class with no FFI; its
 * pixel pipeline is the private chain path_to_lines -> dash_lines ->
 * stroke_lines -> lines_to_runs -> render_shadow -> render_main
 * (canvas_ity.hpp:1195-1212) plus draw_image and get_image_data.  This ABI is
 * the seam where canvas_ity_hip.hpp hands that chain to the GPU: the header
 * keeps the public API and the host-side state machine (transforms, styles,
 * path building), records each fill/stroke/draw_image as a draw packet, and
 * the library runs the pipeline on the device when pixels are observed.
 *
 * Conventions: every function returns int status (CIH_OK == 0, negative on
 * error) and never throws; cih_last_error() gives a thread-local message.
 * Handles are opaque.  Inputs are copied before a call returns (as the
 * reference copies, e.g. canvas_ity.hpp:2761-2765, 2852-2860).  One canvas
 * must not be used from two threads at once; distinct canvases may.
 */
#ifndef CIH_ABI_H
#define CIH_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define CIH_OK 0
#define CIH_ERROR_INVALID (-1)       /* bad argument or handle               */
#define CIH_ERROR_OUT_OF_MEMORY (-2) /* host or device allocation failed     */
#define CIH_ERROR_DEVICE (-3)        /* HIP runtime error or no device       */
#define CIH_ERROR_MALFORMED (-4)     /* malformed command stream or packet   */
#define CIH_ERROR_BUDGET (-5)        /* request exceeds the memory budget    */
#define CIH_ERROR_TIMEOUT (-6)       /* nothing completed within the timeout */

typedef struct cih_device cih_device;
typedef struct cih_canvas cih_canvas;

#if defined(__GNUC__)
#define CIH_API __attribute__((visibility("default")))
#else
#define CIH_API
#endif

/* ---- draw packets: the device-level command buffer ------------------------
 * One packet per fill/stroke/fill_rectangle/stroke_rectangle/clear_rectangle/
 * draw_image call that reaches render_main (canvas_ity.hpp:2551).  Geometry is
 * in device space, exactly as the reference stores paths (hpp:2878-2950).
 * A packet is the header below followed by, in order:
 *   float    points[2 * points]        bezier path (x0,y0 then 3 points per
 *                                      cubic) or polylines, per `kind`
 *   uint32_t subpath[2 * subpaths]     (point count, closed) per subpath
 *   float    dash[dashes]              dash pattern (already doubled if odd)
 *   float    colors[4 * colors]        brush colours (hpp:162-165 layout)
 *   float    stops[stops]              gradient stop offsets, ascending
 * and padded to a multiple of 16 bytes; `bytes` is the padded total.
 */
enum cih_draw_kind {
    CIH_DRAW_FILL_PATH = 1,   /* fill():   path_to_lines(false) (hpp:3044)   */
    CIH_DRAW_STROKE_PATH = 2, /* stroke(): path_to_lines(true)+stroke_lines  */
    CIH_DRAW_FILL_LINES = 3,  /* fill_rectangle / draw_image polylines       */
    CIH_DRAW_STROKE_LINES = 4, /* stroke_rectangle polylines + stroke_lines  */
    CIH_DRAW_CLIP = 5,         /* clip(): the path's fill coverage times the
                                  mask `clip_mask` becomes mask `image`
                                  (hpp:3057-3099); paints nothing             */
    CIH_DRAW_PUT_IMAGE = 6     /* put_image_data recorded in draw order (batch
                                  and serving canvases): the texels of device
                                  image `image` (image_width x image_height,
                                  brush CIH_BRUSH_IMAGE) replace the pixels at
                                  the integer position (start_x, start_y),
                                  whatever the clip, alpha or operator
                                  (hpp:3383-3408); no geometry, no shadow     */
};

enum cih_brush_kind {
    CIH_BRUSH_COLOR = 0,   /* paint_brush::color   (hpp:162)                */
    CIH_BRUSH_LINEAR = 1,  /* paint_brush::linear                           */
    CIH_BRUSH_RADIAL = 2,  /* paint_brush::radial                           */
    CIH_BRUSH_PATTERN = 3, /* paint_brush::pattern, wraps (hpp:2303-2305)   */
    CIH_BRUSH_IMAGE = 4    /* draw_image's image_brush, clamps (hpp:2306)   */
};

#define CIH_DRAW_FLAG_EVENODD 1u /* extension: even-odd fill rule           */
/* Path packets (fill / stroke) whose cubics carry their own start points:
 * 4 points per cubic (start, control 1, control 2, end), subpaths of 4k
 * points whose polyline begins at the first cubic's start.  Glyph outlines
 * use it: the reference starts each glyph curve at the last point its
 * tessellation emitted (canvas_ity.hpp:1660-1665, 1674-1686), which can
 * differ by rounding from the previous curve's end point. */
#define CIH_DRAW_FLAG_SEGMENTS 2u

typedef struct cih_draw_packet {
    uint32_t bytes;       /* padded size of header + arrays                   */
    uint32_t kind;        /* cih_draw_kind                                    */
    uint32_t op;          /* composite_operation value (hpp:146-149)          */
    uint32_t flags;       /* CIH_DRAW_FLAG_*                                  */
    float global_alpha;
    float line_width, miter_limit, dash_offset;
    int32_t line_cap, line_join;
    float forward[6];       /* affine a,b,c,d,e,f at the call (hpp:1153)      */
    float inverse[6];       /* its inverse (stroke expansion, dashes)         */
    float paint_inverse[6]; /* inverse used for paint (draw_image adjusts it) */
    float shadow_color[4];  /* premultiplied linear; alpha 0 = no shadow pass */
    float shadow_blur, shadow_offset_x, shadow_offset_y;
    int32_t brush;          /* cih_brush_kind                                 */
    float start_x, start_y, end_x, end_y, start_radius, end_radius;
    int32_t image;          /* image handle for pattern/image brushes         */
    int32_t image_width, image_height, repetition;
    uint32_t points, subpaths, dashes, colors, stops;
    uint32_t clip_mask;     /* the clip mask the draw is composited under:
                               0 = the whole canvas (the constructor's mask,
                               hpp:2636-2643), k = the mask made by the clip
                               packet with image = k (ids reused once freed) */
} cih_draw_packet;

/* ---- device ---------------------------------------------------------------- */
/* Opens HIP device `ordinal`.  No reference counterpart: the reference is
 * CPU-only and holds no device state (.github/README.md:280-282). */
CIH_API int cih_device_open(int ordinal, cih_device **device);
CIH_API int cih_device_close(cih_device *device);
/* Waits for all work queued on the device by this library. */
CIH_API int cih_device_synchronize(cih_device *device);
/* Thread-local text of the last error (never NULL). */
CIH_API const char *cih_last_error(void);

/* ---- canvas ------------------------------------------------------------- */
/* Replaces canvas::canvas(width, height) (canvas_ity.hpp:194, 2607-2644):
 * a width x height fp32 linear premultiplied RGBA bitmap in HBM, cleared to
 * transparent black, default (full) clip. */
CIH_API int cih_canvas_create(cih_device *device, int width, int height, cih_canvas **canvas);
/* Replaces canvas::~canvas() (canvas_ity.hpp:211, 2646-2655). */
CIH_API int cih_canvas_destroy(cih_canvas *canvas);
CIH_API int cih_canvas_size(cih_canvas *canvas, int *width, int *height);
/* The device a canvas lives on. */
CIH_API int cih_canvas_device(cih_canvas *canvas, cih_device **device);

/* Uploads RGBA8 texels for pattern/image brushes and returns a handle.  It
 * replaces the texel copy and linearisation of set_pattern
 * (canvas_ity.hpp:592, 2839-2864) and draw_image (hpp:3326-3328); identical
 * texel contents share one device image. */
CIH_API int cih_image_create(cih_device *device, const unsigned char *texels, int width,
                     int height, int stride, int32_t *image);

/* Appends draw packets (see above) to the canvas' pending list.  Replaces the
 * pixel side of render_main (canvas_ity.hpp:2551-2605) and render_shadow
 * (hpp:2395-2539): drawing is deferred until pixels are observed. */
CIH_API int cih_canvas_submit(cih_canvas *canvas, const void *packets, size_t bytes);
/* Runs every pending draw on the device (asynchronously). */
CIH_API int cih_canvas_flush(cih_canvas *canvas);

/* Replaces canvas::get_image_data (canvas_ity.hpp:1092, 3348-3381): flushes,
 * converts to sRGB8 with the 4x4 ordered dither and copies the rectangle
 * (x, y, width, height) into image (row stride `stride` bytes); pixels outside
 * the canvas read as transparent black.  A NULL image does nothing. */
CIH_API int cih_canvas_get_image_data(cih_canvas *canvas, unsigned char *image, int width,
                              int height, int stride, int x, int y);
/* Replaces canvas::put_image_data (canvas_ity.hpp:1123, 3383-3408). */
CIH_API int cih_canvas_put_image_data(cih_canvas *canvas, const unsigned char *image,
                              int width, int height, int stride, int x, int y);
/* Test seam: flushes and copies the fp32 linear premultiplied RGBA bitmap
 * (the reference's private `bitmap`, canvas_ity.hpp:1171) to rgba. */
CIH_API int cih_canvas_read_linear(cih_canvas *canvas, float *rgba);

/* Runs the canvas' last flushed draw list again from HBM (optionally after
 * clearing the canvas, optionally followed by the sRGB8 conversion into a
 * device buffer), asynchronously.  For benchmarks: the inputs stay resident
 * and nothing crosses PCIe.  No reference counterpart. */
CIH_API int cih_canvas_execute(cih_canvas *canvas, int clear, int readback);

/* Work counts of the canvas' last draw list, from one synchronous run of it
 * (the canvas is cleared first): stats[0] pixels composited (draw and shadow
 * layers), stats[1] coverage runs, stats[2] polygon/scratch points,
 * stats[3] shadow image floats.  Benchmarks use them for algorithmic bytes. */
CIH_API int cih_canvas_stats(cih_canvas *canvas, uint64_t *stats);

/* Kernel timing: while enabled, every frame records HIP events around each
 * of its kernels on the device's stream; cih_device_profile_read waits for the
 * device and writes {"frames": n, "<kernel>": total_ms, ...} as JSON, then
 * restarts the count.  No reference counterpart. */
CIH_API int cih_device_profile(cih_device *device, int enable);
CIH_API int cih_device_profile_read(cih_device *device, char *json, size_t capacity);

/* Test seams for stage-level differential tests.  They read the last frame
 * run on the canvas' device, so call them right after a flush whose draws fit
 * one frame, before any other canvas on the device is flushed and while the
 * canvas that ran it is alive (its draw records are read).
 * cih_trace_polygons: the closed polygons draw `draw` produced (flattening,
 * dashing and stroking output, device space) -- the reference's `lines`
 * after path_to_lines/stroke_lines (canvas_ity.hpp:1495, 2070).
 * cih_trace_runs: that draw's signed coverage runs before coalescing (pass 0
 * main, pass 1 shadow) -- the reference's `runs` of lines_to_runs
 * (canvas_ity.hpp:2193) before its sort and merge.  Arrays are filled up to
 * their capacity; the full counts are returned. */
CIH_API int cih_trace_polygons(cih_canvas *canvas, int draw, float *points,
                               size_t point_capacity, uint32_t *subpaths,
                               size_t subpath_capacity, size_t *npoints, size_t *nsubpaths);
/* cih_trace_stage: stage 0 flattening output (path_to_lines, hpp:1495; a
 * polyline draw's own lines), 1 dashing output (dash_lines, hpp:1858; empty
 * unless the draw is a stroke with a dash pattern), 2 the final polygons (as
 * cih_trace_polygons).  A stage a draw does not have is empty. */
CIH_API int cih_trace_stage(cih_canvas *canvas, int draw, int stage, float *points,
                            size_t point_capacity, uint32_t *subpaths, size_t subpath_capacity,
                            size_t *npoints, size_t *nsubpaths);
CIH_API int cih_trace_runs(cih_canvas *canvas, int draw, int pass, int32_t *xs, int32_t *ys,
                           float *deltas, size_t capacity, size_t *nruns);

/* Replays a command stream (cih_stream.h) through canvas_ity_hip.hpp's
 * canvas API into this canvas.  The stream header's size must match. */
CIH_API int cih_canvas_replay(cih_canvas *canvas, const void *stream, size_t bytes);

/* is_point_in_path (canvas_ity.hpp:3101-3132): whether (x, y) lies inside
 * the path by the non-zero rule or on its boundary, for a path in device
 * space given as the packets carry it (a start point and 3 points per cubic
 * per subpath; subpaths as (count, closed) pairs).  The path is flattened for
 * filling (hpp:1495-1524) and crossings are counted with the reference's
 * float arithmetic.  A host-side query: it needs no device. */
CIH_API int cih_path_contains(const float *points, size_t npoints, const uint32_t *subpaths,
                              size_t nsubpaths, float x, float y, int *inside);

/* ---- text: TrueType layout, on the host ------------------------------------
 * Host-side functions (no device) behind canvas_ity_hip.hpp's set_font,
 * measure_text, fill_text and stroke_text; `font` is a whole TrueType file
 * and `scale` the font size over its units per em (hpp:3271-3273).
 * cih_font_check replaces set_font's table checks (canvas_ity.hpp:955,
 * 3205-3274): CIH_OK with the head table's units per em, or
 * CIH_ERROR_MALFORMED.  cih_text_measure replaces measure_text (hpp:1025,
 * 3297-3311).  cih_text_outline replaces text_to_lines and add_glyph
 * (hpp:1533-1846): the glyph outlines of `text` placed at (x, y) under the
 * affine `forward` with text_align `align`, text_baseline `baseline` and
 * `maximum_width`, as the geometry of a CIH_DRAW_FLAG_SEGMENTS packet in
 * device space, one closed subpath per glyph contour; `stroking` and
 * `line_width` choose the flattening the segment start points follow
 * (hpp:1497-1500).  If an array is too small it returns CIH_ERROR_BUDGET
 * with *npoints and *nsubpaths set to the sizes needed. */
CIH_API int cih_font_check(const unsigned char *font, size_t bytes, int *units_per_em);
CIH_API int cih_text_measure(const unsigned char *font, size_t font_bytes, float scale,
                             const char *text, float *width);
CIH_API int cih_text_outline(const unsigned char *font, size_t font_bytes, float scale,
                             const char *text, float x, float y, float maximum_width,
                             int align, int baseline, const float forward[6], int stroking,
                             float line_width, float *points, size_t point_capacity,
                             uint32_t *subpaths, size_t subpath_capacity, size_t *npoints,
                             size_t *nsubpaths);

/* ---- batches: many independent canvases in shared frames ------------------
 * A batch replays every command stream through canvas_ity_hip.hpp (host
 * threads in parallel), keeps all canvases, draw records and packets resident
 * in HBM, and renders them with frames that each hold many canvases' draws.
 * The first execute sizes the arenas and captures the launch sequence in a
 * HIP graph; later executes replay the graph.  Every execute starts from
 * cleared canvases (as fresh canvas objects would) and ends with the sRGB8
 * conversion of every canvas into one packed device buffer; `to_host` copies
 * it to pinned host memory and waits.  No reference counterpart: the
 * reference renders one canvas per object (canvas_ity.hpp:194). */
typedef struct cih_batch cih_batch;
/* A batch holds at most 2^30 canvas pixels (CIH_ERROR_BUDGET beyond) and any
 * number of draws.  It cuts its draws, in order, into frames: a frame over
 * several canvases holds at most 2048 draws and 2M (2 * 2^20) canvas pixels,
 * one over a single canvas at most 8192 draws (as the plain path's frames).
 * A server batch admits requests up to 2048 draws and 2M pixels; a single
 * request above them runs as a batch of its own. */
CIH_API int cih_batch_create(cih_device *device, int count, const void *const *streams,
                             const size_t *bytes, cih_batch **batch);
CIH_API int cih_batch_execute(cih_batch *batch, int to_host);
CIH_API int cih_batch_size(cih_batch *batch, int index, int *width, int *height);
/* Copies canvas `index`'s width*height*4 sRGB8 bytes of the last execute. */
CIH_API int cih_batch_image(cih_batch *batch, int index, unsigned char *image);
/* Copies canvas `index`'s fp32 linear premultiplied bitmap (test seam). */
CIH_API int cih_batch_linear(cih_batch *batch, int index, float *rgba);
/* stats: canvases, draws, frames, pixels. */
CIH_API int cih_batch_stats(cih_batch *batch, uint64_t *stats);
/* info of frame `index` (0 .. frames - 1): first draw, draws, canvases, and
 * the pixels of those canvases (test seam; no reference counterpart). */
CIH_API int cih_batch_frame(cih_batch *batch, int index, uint64_t *info);
/* Fine-stage launch shape of a batch (no reference counterpart): config
 * 2 * w + p, w = 0, 1, 2 for work items of 256, 128 or 64 pixels of a canvas
 * row, p = 1 for one persistent grid of waves draining per-XCD item queues
 * instead of one workgroup per item.  Every configuration gives bit-identical
 * pixels; config 0 is the default.  cih_batch_tune times all of them on the
 * batch (`reps` executes each after one untimed), keeps the fastest, writes
 * each one's milliseconds per execute to ms[CIH_FINE_CONFIGS] (if not NULL)
 * and its index to *best (if not NULL). */
#define CIH_FINE_CONFIGS 6
CIH_API int cih_batch_configure(cih_batch *batch, int config);
CIH_API int cih_batch_tune(cih_batch *batch, int reps, double *ms, int *best);
CIH_API int cih_batch_destroy(cih_batch *batch);

/* ---- serving: continuous batching ----------------------------------------
 * A server takes command streams as they arrive (submit copies them), replays
 * them on worker threads, admits prepared requests in arrival order while
 * their canvases fit a pool of `hbm_budget` bytes of HBM, and runs admitted
 * requests as batches on two device streams, so requests that arrive while
 * others run are admitted as soon as memory and a stream free up.
 * cih_server_wait returns the next finished request (CIH_ERROR_TIMEOUT if
 * none within timeout_ms); cih_server_result copies its sRGB8 pixels
 * (width*height*4 bytes), reports submit-to-finish latency and releases it.
 * stats: requests served, batches, most requests in flight, re-runs.
 * No reference counterpart. */
typedef struct cih_server cih_server;
CIH_API int cih_server_start(cih_device *device, uint64_t hbm_budget, int max_batch,
                             int workers, cih_server **server);
CIH_API int cih_server_submit(cih_server *server, const void *stream, size_t bytes,
                              uint64_t *request);
CIH_API int cih_server_wait(cih_server *server, uint64_t *request, int timeout_ms);
CIH_API int cih_server_result(cih_server *server, uint64_t request, unsigned char *image,
                              int *width, int *height, double *latency_ms);
CIH_API int cih_server_stats(cih_server *server, uint64_t *stats);
CIH_API int cih_server_stop(cih_server *server);

/* ---- synthetic workloads ----------------------------------------------------
 * Generates the seeded command stream of a benchmark configuration:
 * 1 plain-512, 2 large-2048, 3 effects-1024, 4 one batch canvas,
 * 5 one serving request.  width/height/shapes of 0 take the configuration's
 * defaults.  *stream is malloc'd; release it with cih_free. */
#define CIH_SCENE_EVENODD 1   /* config 2: 25% of fills use even-odd       */
#define CIH_SCENE_NO_IMAGES 2 /* configs 3/5: no draw_image calls          */
#define CIH_SCENE_SHADOWS 4   /* config 5: force the shadowed variant      */
CIH_API int cih_scene_generate(int config, uint64_t seed, int width, int height, int shapes,
                       int flags, void **stream, size_t *bytes);
/* Per-canvas / per-request seed: seed ^ index * 0x9E3779B97F4A7C15. */
CIH_API uint64_t cih_scene_seed(uint64_t seed, uint64_t index);
CIH_API void cih_free(void *pointer);

#ifdef __cplusplus
}
#endif

#endif /* CIH_ABI_H */
