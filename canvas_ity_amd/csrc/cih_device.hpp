// cih_device.hpp -- data shared by the kernels (cih_kernels.hip) and the
// host runtime (cih_runtime.cpp).  Plain structs, no HIP types.
#ifndef CIH_DEVICE_HPP
#define CIH_DEVICE_HPP

#include <stdint.h>

namespace cih {

// One draw as the kernels see it, decoded from a cih_draw_packet by the host.
// Offsets are in bytes from the start of the frame's payload buffer, which
// the host keeps below 4 GiB (CIH_ERROR_BUDGET beyond): 32-bit offsets keep
// k_fine's per-draw registers down.
struct draw_desc {
    int32_t kind, op, flags, canvas;
    float global_alpha, line_width, miter_limit, dash_offset;
    int32_t line_cap, line_join;
    float forward[6], inverse[6], paint_inverse[6];
    float shadow_color[4];
    float shadow_offset_x, shadow_offset_y;
    int32_t shadow; // 1 when render_shadow runs (hpp:2398-2401)
    int32_t radius, border;
    float weight_1, weight_2;
    int32_t brush;
    float start_x, start_y, end_x, end_y, start_radius, end_radius;
    int32_t image, image_width, image_height, repetition; // clip: image = the new mask id
    uint32_t points_at, npoints, subpaths_at, nsubpaths;
    uint32_t dash_at, ndash, colors_at, ncolors, stops_at, nstops;
    int32_t clip_mask, pad_; // mask the draw is composited under (0: whole canvas)
};

// Results the kernels leave for the later stages of the same frame.
struct draw_state {
    // final polygons: device addresses of xy points and (count, closed) pairs
    uint64_t poly_points, poly_counts;
    uint32_t poly_npoints, poly_nsub;
    // intermediate polylines (test seams): flattening and dashing output
    uint64_t flat_points, flat_counts, dash_points, dash_counts;
    uint32_t flat_npoints, flat_nsub, dash_npoints, dash_nsub;
    // main scan conversion: rows [row0, row0 + rows); run offsets per row
    int32_t row0, rows;
    uint64_t rowoff; // word-arena index of rows + 1 offsets
    uint64_t runs;   // run-arena index of the first run
    int32_t minx, maxx;
    // shadow scan conversion (extended canvas coordinates)
    int32_t srow0, srows;
    uint64_t srowoff, sruns;
    int32_t slast; // last row with a shadow step (its last pixel is never painted)
    int32_t clast; // clip draws: last row with a coalesced run (-1: none)
    int32_t pad_;
    // shadow image: [left, left + width) x [top, top + height) of the
    // extended canvas, stored row-major at `image` (and `image2` ping-pong)
    int32_t left, top, width, height;
    uint64_t image, image2;
};

struct image_desc {
    uint64_t texels; // device address of float4 premultiplied linear texels
    int32_t width, height;
};

struct canvas_desc {
    uint64_t bitmap; // device address of float4 pixels
    int32_t width, height;
    int32_t first_draw, ndraws;
    uint64_t masks; // device address of the canvas' mask_desc table (0: none)
};

// A clip mask (the reference's `mask` run list, hpp:1169): per canvas row,
// up to width + 2 pieces (x, delta) in x order, and the running visibility
// sum after each piece as steps (x, sum) for lookups.  Row y's entries start
// at y * (width + 2); counts[y] of them are used.
struct mask_desc {
    uint64_t pieces; // device address of {int32 x, float delta}
    uint64_t steps;  // device address of {int32 x, float sum}
    uint64_t counts; // device address of uint32 per row
    int32_t last_row; // last row with a piece (-1: the mask is empty)
    int32_t pad_;
};

struct run {
    int32_t x;
    float delta;
};

// A coverage step: the row's running sum of run deltas after pixel x.
struct step {
    int32_t x;
    float sum;
};

// Arena counters (device, 64-bit): requested amounts keep counting past the
// capacity so the host learns how much a frame needed.
enum counter {
    COUNTER_POINTS = 0,
    COUNTER_WORDS = 1,
    COUNTER_RUNS = 2,
    COUNTER_SHADOW = 3,
    COUNTER_OVERFLOW = 4,
    COUNTER_QUEUE = 5, // coalesce work list: (entries << 40) + rows
    COUNTER_PIXEL_DRAWS = 6, // pixels composited by k_fine (main and shadow layers)
    // draws whose geometry (k_geometry) and passes whose scan conversion
    // (k_runs) completed: after an overflow they tell the host which share
    // of the frame its arena requests cover
    COUNTER_GEOMETRY_DONE = 7,
    COUNTER_FINE_QUEUE = 8, // 8 cursors of the persistent fine kernel's item ranges
    COUNTER_RUNS_DONE = 16,
    COUNTER_COUNT = 17
};

struct frame {
    const unsigned char *payload;
    const draw_desc *draws;
    draw_state *state;
    const canvas_desc *canvases;
    const image_desc *images;
    int32_t ndraws, ncanvases;
    // arenas
    float *points; // xy pairs
    uint64_t points_cap;
    uint32_t *words;
    uint64_t words_cap;
    run *runs;
    step *steps; // same indexing and capacity as runs
    uint64_t runs_cap;
    float *shadow;
    uint64_t shadow_cap;
    unsigned long long *counters;
    // coalesce work list: one entry per (draw, pass) with rows, in the order
    // k_runs appended them: (first row in the list << 20) | (2 * draw + pass)
    unsigned long long *queue;
    // fine-stage work list: (canvas, row, segment) items, per canvas prefix
    const uint32_t *items;
    uint32_t nitems;
    int32_t segment;         // pixels per fine work item: 64, 128 or 256
    int32_t fine_persistent; // k_fine as a fixed grid draining item queues
};

} // namespace cih

#endif // CIH_DEVICE_HPP
