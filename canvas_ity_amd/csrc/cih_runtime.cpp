// cih_runtime.cpp -- host side of the C ABI (cih_abi.h): devices, canvases,
// images, packet decoding, frame scheduling and read-back.
//
// A flush turns a canvas' pending draw packets into one or more frames.  A
// frame uploads the packets once, decodes them into draw_desc records and
// launches the kernel chain of cih_kernels.hip on the device's stream.
// Variable-size intermediates (polylines, runs, shadow images) come from
// device arenas by atomic bump allocation; if a frame runs out, the kernels
// flag it before any pixel is written, the host grows the arenas to the
// requested size and runs the frame again.
#include <hip/hip_runtime.h>

#include "cih_abi.h"
#include "cih_device.hpp"
#include "cih_geometry.hpp"
#include "cih_stream.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <system_error>
#include <thread>
#include <string>
#include <unordered_map>
#include <vector>

namespace cih {
hipError_t launch_readback_batch(const void *canvases, const void *pixel_prefix, int ncanvases,
                                 uint64_t total, void *out, hipStream_t stream);
hipError_t launch_clear_batch(const void *canvases, const void *pixel_prefix, int ncanvases,
                              uint64_t total, hipStream_t stream);
const int kFrameKernels = 8; // reset, geometry, runs, coalesce, shadow raster, blur rows/cols, fine

// What one frame over several canvases (of a batch or a server batch) may
// hold: the largest multi-canvas frames that have run cleanly on the MI355X
// (1,913 draws over 10 canvases; 1.9M pixels over 29 canvases), rounded up.
// One frame of 25,187 draws over 128 canvases (14.4M pixels) has faulted the
// device, and the cause has not been confirmed on the device.  A batch cuts its draws into frames within
// these bounds; the server admits requests into a batch up to them and runs a
// request above them as a batch of its own (a single-canvas frame, like the
// plain path).
const uint64_t kBatchPixelLimit = (uint64_t)2 << 20;
const size_t kBatchDrawLimit = 2048;
// Draws per frame over one canvas (the plain path's frames).
const size_t kFrameDrawLimit = 8192;
// Canvas pixels one batch may hold: 16 bytes of fp32 bitmap and 4 of sRGB8
// each, 20 GiB of HBM at the bound.
const uint64_t kBatchTotalPixelLimit = (uint64_t)1 << 30;

// Bytes of draw packets one canvas program, batch or server batch may hold:
// draw records address the payload with 32-bit offsets.
const uint64_t kPayloadLimit = 0xffffffffull;

// Distinct images one device holds (the image table is allocated once).
const size_t kMaxImages = (size_t)1 << 16;

// Clip mask ids one canvas may use at once (its mask table is allocated once;
// the header reuses ids, so this bounds the depth of nested clips), and the
// device memory its masks may take.
const uint32_t kMaxMasks = 256;
const uint64_t kMaskBytesLimit = (uint64_t)8 << 30;
extern const char *const kFrameKernelNames[kFrameKernels];
hipError_t launch_frame(const frame &F, int shadow_rows, int shadow_cols, bool any_shadow,
                        const int32_t *clips, int nclips, const int32_t *puts, int nputs,
                        bool masked, hipEvent_t *events, hipStream_t stream);
hipError_t launch_clear(void *pixels, size_t n, hipStream_t stream);
hipError_t launch_prepare_image(const void *rgba8, void *texels, size_t n, const float *lut,
                                hipStream_t stream);
hipError_t launch_put_image(const void *rgba8, int w, int h, void *bitmap, int W, int H, int x,
                            int y, const float *lut, hipStream_t stream);
hipError_t launch_readback(const void *bitmap, int W, int H, void *out, int w, int h, int x,
                           int y, hipStream_t stream);
} // namespace cih

using namespace cih;

namespace {

thread_local std::string g_error = "no error";

int fail(int status, const std::string &message) {
    g_error = message;
    return status;
}

} // namespace

// Shared with the other host translation units of the library.
int cih_internal_fail(int status, const char *message) { return fail(status, message); }

namespace {

int hip_fail(hipError_t e, const char *what) {
    return fail(CIH_ERROR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return hip_fail(e_, #expr);                                                       \
    } while (0)

// A device allocation that only grows.
struct dbuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int reserve(size_t need) {
        if (need <= bytes)
            return CIH_OK;
        size_t want = std::max(need, bytes + bytes / 2);
        if (ptr)
            hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&ptr, want);
        if (e != hipSuccess) {
            ptr = nullptr;
            (void)hipGetLastError();
            e = hipMalloc(&ptr, need);
            if (e != hipSuccess) {
                ptr = nullptr;
                (void)hipGetLastError();
                return fail(CIH_ERROR_OUT_OF_MEMORY, "device allocation of " +
                                                         std::to_string(need) + " bytes failed");
            }
            want = need;
        }
        bytes = want;
        return CIH_OK;
    }
    void release() {
        if (ptr)
            hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

// Pinned host staging that only grows.
struct hbuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int reserve(size_t need) {
        if (need <= bytes)
            return CIH_OK;
        size_t want = std::max(need, bytes + bytes / 2);
        if (ptr)
            hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
        if (hipHostMalloc(&ptr, want, hipHostMallocDefault) != hipSuccess) {
            ptr = nullptr;
            (void)hipGetLastError();
            return fail(CIH_ERROR_OUT_OF_MEMORY, "pinned allocation failed");
        }
        bytes = want;
        return CIH_OK;
    }
    void release() {
        if (ptr)
            hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

float linearized(float v) { // hpp:1273-1275, evaluated on the host like the reference
    return v < 0.04045f ? v / 12.92f : powf((v + 0.055f) / 1.055f, 2.4f);
}

struct image_entry {
    uint64_t hash;
    int width, height;
    std::vector<unsigned char> rgba; // dense copy, for exact de-duplication
    dbuf texels;
};

} // namespace

// A stream with its own arenas and frame buffers: frames on different lanes
// can be in flight together.
struct exec_lane {
    hipStream_t stream = nullptr;
    dbuf points, words, runs, steps, shadow, counters; // arenas and their counters
    dbuf state;                                  // per-draw results and the work list
    hbuf host_counters;
    // the lane's last frame, for the trace seams: its draw count, draw
    // records and packet payload (device memory)
    size_t last_ndraws = 0;
    const draw_desc *last_draws = nullptr;
    const unsigned char *last_payload = nullptr;
    void release() {
        dbuf *all[] = {&points, &words, &runs, &steps, &shadow, &counters, &state};
        for (dbuf *b : all)
            b->release();
        host_counters.release();
    }
};

struct cih_device {
    int ordinal = 0;
    hipStream_t stream = nullptr;
    std::recursive_mutex mutex;
    exec_lane main; // the lane of canvases and batches (stream == `stream`)
    hbuf staging;
    // images
    std::vector<image_entry *> images;
    std::unordered_multimap<uint64_t, int32_t> image_index;
    dbuf image_table, lut;
    dbuf upload;
    // read-back
    dbuf readback;
    hbuf readback_host;
    size_t frames = 0, retries = 0;
    // kernel timing with HIP events on the device's stream
    bool profiling = false;
    bool capturing = false; // a HIP graph capture is recording on the stream
    std::vector<hipEvent_t> event_pool;
    size_t events_used = 0;
};

// Fine-stage item widths: every frame carries the (canvas, row, segment) item
// prefix for each, so a batch can switch between them (the autotuner does).
const int kFineVariants = 3;
const int kFineSegments[kFineVariants] = {256, 128, 64};

// How a frame's fine stage runs: the segment width of a work item (index into
// kFineSegments) and whether k_fine is a persistent grid draining queues.
struct fine_config {
    int variant = 0;
    bool persistent = false;
};

// What a frame needs from the arenas, and its fine-stage work items.
struct frame_shape {
    bool any_shadow = false;
    int max_rows = 0, max_cols = 0;
    uint64_t shadow_bound = 0, in_points = 0;
    size_t ndraws = 0;
    // per-canvas prefix of (row, segment) items, one array of ncanvases + 1
    // entries per item width of kFineSegments, in that order
    std::vector<uint32_t> items;
    std::vector<int32_t> clips;  // (draw, canvas height) of each clip draw, in order
    std::vector<int32_t> puts;   // (draw, texel count) of each put, in order
    bool masked = false;         // some draw is composited under a clip mask
    size_t nshadowed = 0;        // draws with a shadow pass
};

// A canvas' decoded draw list, resident in HBM so it can be run again, with
// each frame's canvas record and fine-stage work items (uploaded once).
struct program {
    dbuf payload, draws;
    dbuf canvases, items; // one canvas_desc per frame; the frames' item prefixes
    std::vector<draw_desc> descs; // host copy
    struct span {
        size_t first, count, items_at;
        frame_shape shape;
    };
    std::vector<span> frames;
    size_t bytes = 0;
};

struct cih_canvas {
    cih_device *device;
    int width, height;
    void *pixels = nullptr; // float4 bitmap in HBM (owned by `bitmap` or by a batch)
    dbuf bitmap;
    // Batch and serving canvases record put_image_data as a packet in draw
    // order (their pixels are drawn later, from cleared bitmaps); other
    // canvases apply it at once.
    bool deferred = false;
    std::vector<unsigned char> pending;
    program program;
    // clip masks: a table of kMaxMasks mask_desc in HBM (entry 0 unused: the
    // whole canvas) and each id's storage, allocated when a clip defines it
    dbuf mask_table;
    std::vector<dbuf> mask_store;
    std::vector<bool> mask_defined;
    uint64_t mask_bytes = 0;
};

namespace {

int set_device(cih_device *d) {
    HIP_TRY(hipSetDevice(d->ordinal));
    return CIH_OK;
}

// Host <-> device copies of the library go on the device's stream and are
// waited for there.  The stream is non-blocking, so a copy on the NULL stream
// (plain hipMemcpy) is not ordered before the kernels the stream runs next,
// and a pageable host-to-device hipMemcpy may return before its transfer has
// landed; kernels could then read half-written draw, canvas or mask tables.
hipError_t copy_sync(cih_device *d, void *dst, const void *src, size_t bytes,
                     hipMemcpyKind kind) {
    if (!bytes)
        return hipSuccess;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, d->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(d->stream);
    return e;
}

// Decodes one packet; returns the number of bytes consumed or 0 if malformed.
size_t decode_packet(const unsigned char *base, size_t at, size_t end, draw_desc &out,
                     const cih_device *dev) {
    if (end - at < sizeof(cih_draw_packet))
        return 0;
    cih_draw_packet p;
    std::memcpy(&p, base + at, sizeof(p));
    size_t body = sizeof(p) + 8 * (size_t)p.points + 8 * (size_t)p.subpaths +
                  4 * (size_t)p.dashes + 16 * (size_t)p.colors + 4 * (size_t)p.stops;
    if (p.bytes < body || p.bytes > end - at || (p.bytes & 15))
        return 0;
    if (p.kind < CIH_DRAW_FILL_PATH || p.kind > CIH_DRAW_PUT_IMAGE)
        return 0;
    // a put: an image and an integer position, nothing else
    if (p.kind == CIH_DRAW_PUT_IMAGE &&
        (p.points || p.subpaths || p.dashes || p.colors || p.stops || p.flags ||
         p.brush != CIH_BRUSH_IMAGE || p.clip_mask || !(fabsf(p.start_x) <= 16777216.0f) ||
         !(fabsf(p.start_y) <= 16777216.0f) || p.start_x != floorf(p.start_x) ||
         p.start_y != floorf(p.start_y)))
        return 0;
    if (p.clip_mask >= kMaxMasks ||
        (p.kind == CIH_DRAW_CLIP && (p.image < 1 || p.image >= (int32_t)kMaxMasks)))
        return 0;
    // subpath counts must add up to the points; bezier subpaths are 1 + 3k,
    // or 4k with explicit segment starts (fill and stroke paths only)
    bool bezier = p.kind == CIH_DRAW_FILL_PATH || p.kind == CIH_DRAW_STROKE_PATH ||
                  p.kind == CIH_DRAW_CLIP;
    bool segments = (p.flags & CIH_DRAW_FLAG_SEGMENTS) != 0;
    if ((p.flags & ~(CIH_DRAW_FLAG_EVENODD | CIH_DRAW_FLAG_SEGMENTS)) ||
        (segments && (!bezier || p.kind == CIH_DRAW_CLIP)))
        return 0;
    const unsigned char *sub = base + at + sizeof(p) + 8 * (size_t)p.points;
    uint64_t total = 0;
    for (uint32_t s = 0; s < p.subpaths; ++s) {
        uint32_t c;
        std::memcpy(&c, sub + 8 * (size_t)s, 4);
        if (c == 0 || (bezier && (segments ? c % 4 != 0 : (c - 1) % 3 != 0)))
            return 0;
        total += c;
    }
    if (total != p.points)
        return 0;
    if ((p.brush == CIH_BRUSH_LINEAR || p.brush == CIH_BRUSH_RADIAL) && p.stops != p.colors)
        return 0;
    if (p.kind == CIH_DRAW_CLIP) {
        if (p.brush != CIH_BRUSH_COLOR)
            return 0;
    } else if (p.brush == CIH_BRUSH_PATTERN || p.brush == CIH_BRUSH_IMAGE) {
        if (p.image < 0 || (size_t)p.image >= dev->images.size() ||
            dev->images[(size_t)p.image]->width != p.image_width ||
            dev->images[(size_t)p.image]->height != p.image_height)
            return 0;
    } else if (p.brush < 0 || p.brush > CIH_BRUSH_IMAGE)
        return 0;
    std::memset(&out, 0, sizeof(out));
    out.kind = (int32_t)p.kind;
    out.op = (int32_t)p.op;
    out.flags = (int32_t)p.flags;
    out.global_alpha = p.global_alpha;
    out.line_width = p.line_width;
    out.miter_limit = p.miter_limit;
    out.dash_offset = p.dash_offset;
    out.line_cap = p.line_cap;
    out.line_join = p.line_join;
    std::memcpy(out.forward, p.forward, sizeof(out.forward));
    std::memcpy(out.inverse, p.inverse, sizeof(out.inverse));
    std::memcpy(out.paint_inverse, p.paint_inverse, sizeof(out.paint_inverse));
    std::memcpy(out.shadow_color, p.shadow_color, sizeof(out.shadow_color));
    out.shadow_offset_x = p.shadow_offset_x;
    out.shadow_offset_y = p.shadow_offset_y;
    out.shadow = p.shadow_color[3] != 0.0f && p.kind != CIH_DRAW_CLIP &&
                         p.kind != CIH_DRAW_PUT_IMAGE
                     ? 1
                     : 0;
    out.clip_mask = (int32_t)p.clip_mask;
    // Extended-box blur constants (hpp:2402-2405, 2453-2459), in float as
    // the reference computes them.
    float sigma_squared = 0.25f * p.shadow_blur * p.shadow_blur;
    size_t radius = static_cast<size_t>(0.5f * sqrtf(4.0f * sigma_squared + 1.0f) - 0.5f);
    float alpha = static_cast<float>(2 * radius + 1) *
                  (static_cast<float>(radius * (radius + 1)) - sigma_squared) /
                  (2.0f * sigma_squared - static_cast<float>(6 * (radius + 1) * (radius + 1)));
    float divisor = 2.0f * (alpha + static_cast<float>(radius)) + 1.0f;
    out.radius = (int32_t)radius;
    out.border = 3 * ((int32_t)radius + 1);
    out.weight_1 = alpha / divisor;
    out.weight_2 = (1.0f - alpha) / divisor;
    if (out.shadow && (radius > 4096 || !(p.shadow_blur < 1.0e4f)))
        return 0;
    out.brush = p.brush;
    out.start_x = p.start_x, out.start_y = p.start_y;
    out.end_x = p.end_x, out.end_y = p.end_y;
    out.start_radius = p.start_radius, out.end_radius = p.end_radius;
    out.image = p.image;
    out.image_width = p.image_width;
    out.image_height = p.image_height;
    out.repetition = p.repetition;
    uint32_t a = (uint32_t)(at + sizeof(p));
    out.points_at = a;
    out.npoints = p.points;
    a += 8 * p.points;
    out.subpaths_at = a;
    out.nsubpaths = p.subpaths;
    a += 8 * p.subpaths;
    out.dash_at = a;
    out.ndash = p.dashes;
    a += 4 * p.dashes;
    out.colors_at = a;
    out.ncolors = p.colors;
    a += 16 * p.colors;
    out.stops_at = a;
    out.nstops = p.stops;
    return p.bytes;
}

// A lane's state buffer for frames of up to n draws: the draw_state records,
// then the coalesce work list (at most two entries per draw).
static_assert(sizeof(draw_state) % 8 == 0, "the work list after the records is 8-byte aligned");
size_t state_bytes(size_t n) { return (sizeof(draw_state) + 16) * std::max<size_t>(n, 1); }

// Runs and steps share indices: a fallback allocation of the exact size can
// leave one smaller than the other, so the smaller bounds both.
uint64_t run_capacity(const exec_lane &L) {
    return std::min(L.runs.bytes / sizeof(run), L.steps.bytes / sizeof(step));
}

int ensure_arenas(exec_lane &L, uint64_t pts, uint64_t words, uint64_t runs, uint64_t shadow) {
    int s;
    if ((s = L.points.reserve(8 * pts)) || (s = L.words.reserve(4 * words)) ||
        (s = L.runs.reserve(sizeof(run) * runs)) || (s = L.steps.reserve(sizeof(step) * runs)) ||
        (s = L.shadow.reserve(4 * shadow)))
        return s;
    return CIH_OK;
}

// Hands out kFrameKernels + 1 events for one profiled frame.
int profile_events(cih_device *d, hipEvent_t **events) {
    size_t need = d->events_used + kFrameKernels + 1;
    while (d->event_pool.size() < need) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        d->event_pool.push_back(e);
    }
    *events = &d->event_pool[d->events_used];
    d->events_used = need;
    return CIH_OK;
}



frame_shape shape_of(const draw_desc *host, const canvas_desc *canvases, size_t ncanvases) {
    frame_shape f;
    size_t n1 = ncanvases + 1;
    f.items.assign(kFineVariants * n1, 0);
    for (size_t c = 0; c < ncanvases; ++c) {
        const canvas_desc &cd = canvases[c];
        for (int v = 0; v < kFineVariants; ++v) {
            int seg = kFineSegments[v];
            uint64_t segs = (uint64_t)((cd.width + seg - 1) / seg);
            f.items[v * n1 + c + 1] = f.items[v * n1 + c] + (uint32_t)(segs * (uint64_t)cd.height);
        }
        for (int k = cd.first_draw; k < cd.first_draw + cd.ndraws; ++k) {
            const draw_desc &dd = host[k];
            f.in_points += dd.npoints;
            f.ndraws++;
            if (dd.kind == CIH_DRAW_CLIP) {
                f.clips.push_back(k);
                f.clips.push_back(cd.height);
            }
            if (dd.kind == CIH_DRAW_PUT_IMAGE) {
                f.puts.push_back(k);
                f.puts.push_back(dd.image_width * dd.image_height);
            }
            f.masked = f.masked || (dd.clip_mask != 0 && dd.kind != CIH_DRAW_CLIP);
            if (dd.shadow) {
                f.any_shadow = true;
                f.nshadowed++;
                int b = dd.border;
                f.max_rows = std::max(f.max_rows, cd.height + 2 * b + 1);
                f.max_cols = std::max(f.max_cols, cd.width + 2 * b + 1);
                f.shadow_bound += 2ull * (uint64_t)(cd.width + 2 * b + 1) *
                                  (uint64_t)(cd.height + 2 * b + 1);
            }
        }
    }
    return f;
}

// Launches the kernels of one frame whose canvas list and work items are
// Arena capacities for the next attempt of a frame that overflowed.  Once an
// arena is full, draws that start later return without asking, so the
// counters hold only the requests of the draws that got through (and those in
// flight).  Each overflowed arena is grown to its request scaled by the
// share of the frame still missing: geometry's when k_geometry did not
// complete every draw, else scan conversion's (every main pass, and the
// shadow pass of each shadowed draw when the frame has shadows) with a 1.5x
// margin; at least twice the request, at most 16 times.  caps[] holds points,
// words, runs and shadow capacities and is updated in place.
void grow_after_overflow(const unsigned long long *hc, size_t ndraws, size_t passes,
                         uint64_t caps[4]) {
    static const int which[4] = {COUNTER_POINTS, COUNTER_WORDS, COUNTER_RUNS, COUNTER_SHADOW};
    double geometry = hc[COUNTER_GEOMETRY_DONE]
                          ? (double)ndraws / (double)hc[COUNTER_GEOMETRY_DONE]
                          : 16.0;
    double runs = hc[COUNTER_RUNS_DONE] ? (double)passes / (double)hc[COUNTER_RUNS_DONE] : 16.0;
    double scale = hc[COUNTER_GEOMETRY_DONE] < ndraws ? geometry : runs;
    scale = std::min(16.0, std::max(2.0, 1.5 * scale)); // draws differ: a margin
    for (int k = 0; k < 4; ++k) {
        uint64_t asked = hc[which[k]];
        if (asked >= caps[k])
            caps[k] = std::max<uint64_t>(caps[k], (uint64_t)((double)asked * scale));
    }
}

// Launches the kernels of one frame whose canvas list and work items are
// already in HBM.  With `check`, waits for it and, if an arena ran out, grows
// the arenas to what the frame asked for and runs it again (kernels write no
// pixel once an arena has overflowed).  Without it nothing synchronises or
// allocates (arenas must already fit: graph capture relies on that).
int run_kernels(cih_device *d, exec_lane &L, const unsigned char *payload,
                const draw_desc *draws, size_t n,
                const canvas_desc *canvases_dev, size_t ncanvases, const uint32_t *items_dev,
                const frame_shape &shape, bool check, fine_config fine = fine_config()) {
    int s;
    if (shape.masked)
        fine = fine_config(); // k_fine's masked variant runs 256-pixel items only
    if (check) {
        if ((s = L.state.reserve(state_bytes(n))) ||
            (s = L.counters.reserve(8 * COUNTER_COUNT)) ||
            (s = L.host_counters.reserve(8 * COUNTER_COUNT)))
            return s;
        uint64_t want_points = std::max<uint64_t>(1 << 22, 64 * shape.in_points);
        uint64_t want_words = std::max<uint64_t>(1 << 22, 8 * shape.in_points + 64 * n);
        uint64_t want_runs = std::max<uint64_t>(1 << 22, 16 * shape.in_points);
        if ((s = ensure_arenas(L, std::max<uint64_t>(want_points, L.points.bytes / 8),
                               std::max<uint64_t>(want_words, L.words.bytes / 4),
                               std::max<uint64_t>(want_runs, run_capacity(L)),
                               std::max<uint64_t>(shape.shadow_bound, L.shadow.bytes / 4))))
            return s;
    }
    L.last_ndraws = n;
    L.last_draws = draws;
    L.last_payload = payload;
    for (int attempt = 0; attempt < 8; ++attempt) {
        frame F;
        std::memset(&F, 0, sizeof(F));
        F.payload = payload;
        F.draws = draws;
        F.state = (draw_state *)L.state.ptr;
        F.queue = (unsigned long long *)((char *)L.state.ptr + sizeof(draw_state) * n);
        F.canvases = canvases_dev;
        F.images = (const image_desc *)d->image_table.ptr;
        F.ndraws = (int32_t)n;
        F.ncanvases = (int32_t)ncanvases;
        F.points = (float *)L.points.ptr;
        F.points_cap = L.points.bytes / 8;
        F.words = (uint32_t *)L.words.ptr;
        F.words_cap = L.words.bytes / 4;
        F.runs = (run *)L.runs.ptr;
        F.steps = (step *)L.steps.ptr;
        F.runs_cap = run_capacity(L);
        F.shadow = (float *)L.shadow.ptr;
        F.shadow_cap = L.shadow.bytes / 4;
        F.counters = (unsigned long long *)L.counters.ptr;
        F.items = items_dev + (size_t)fine.variant * (ncanvases + 1);
        F.nitems = shape.items[(size_t)fine.variant * (ncanvases + 1) + ncanvases];
        F.segment = kFineSegments[fine.variant];
        F.fine_persistent = fine.persistent ? 1 : 0;
        hipEvent_t *events = nullptr;
        if (d->profiling && !d->capturing) {
            if ((s = profile_events(d, &events)))
                return s;
        }
        HIP_TRY(launch_frame(F, shape.max_rows, shape.max_cols, shape.any_shadow,
                             shape.clips.data(), (int)(shape.clips.size() / 2),
                             shape.puts.data(), (int)(shape.puts.size() / 2), shape.masked,
                             events, L.stream));
        d->frames++;
        if (!check)
            return CIH_OK;
        HIP_TRY(hipMemcpyAsync(L.host_counters.ptr, L.counters.ptr, 8 * COUNTER_COUNT,
                               hipMemcpyDeviceToHost, L.stream));
        HIP_TRY(hipStreamSynchronize(L.stream));
        const unsigned long long *hc = (const unsigned long long *)L.host_counters.ptr;
        if (!hc[COUNTER_OVERFLOW])
            return CIH_OK;
        d->retries++;
        uint64_t caps[4] = {F.points_cap, F.words_cap, F.runs_cap, F.shadow_cap};
        grow_after_overflow(hc, n, n + (shape.any_shadow ? shape.nshadowed : 0), caps);
        if ((s = ensure_arenas(L, caps[0], caps[1], caps[2], caps[3])))
            return s;
    }
    return fail(CIH_ERROR_OUT_OF_MEMORY, "frame did not fit after repeated arena growth");
}


// Device storage for clip mask `id` of a canvas (kept once allocated: the id
// is reused by later clips, which overwrite the mask when they run).
int ensure_mask(cih_canvas *cv, uint32_t id) {
    int s;
    if (!cv->mask_table.ptr) {
        if ((s = cv->mask_table.reserve(sizeof(mask_desc) * kMaxMasks)))
            return s;
        std::vector<mask_desc> zero(kMaxMasks);
        std::memset(zero.data(), 0, sizeof(mask_desc) * kMaxMasks);
        HIP_TRY(copy_sync(cv->device, cv->mask_table.ptr, zero.data(),
                          sizeof(mask_desc) * kMaxMasks, hipMemcpyHostToDevice));
        cv->mask_store.resize(kMaxMasks);
        cv->mask_defined.assign(kMaxMasks, false);
    }
    if (cv->mask_store[id].ptr)
        return CIH_OK;
    uint64_t cap = (uint64_t)cv->width + 2, rows = (uint64_t)cv->height;
    uint64_t bytes = 16 * rows * cap + 4 * rows;
    if (cv->mask_bytes + bytes > kMaskBytesLimit)
        return fail(CIH_ERROR_BUDGET, "the canvas' clip masks exceed " +
                                          std::to_string(kMaskBytesLimit) + " bytes");
    if ((s = cv->mask_store[id].reserve(bytes)))
        return s;
    cv->mask_bytes += bytes;
    char *base = (char *)cv->mask_store[id].ptr;
    mask_desc m;
    std::memset(&m, 0, sizeof(m));
    m.pieces = (uint64_t)(uintptr_t)base;
    m.steps = (uint64_t)(uintptr_t)(base + 8 * rows * cap);
    m.counts = (uint64_t)(uintptr_t)(base + 16 * rows * cap);
    m.last_row = -1;
    // only the new entry: the device has written the others' last rows
    HIP_TRY(copy_sync(cv->device, (mask_desc *)cv->mask_table.ptr + id, &m, sizeof(m),
                      hipMemcpyHostToDevice));
    return CIH_OK;
}

void release_masks(cih_canvas *cv) {
    cv->mask_table.release();
    for (dbuf &m : cv->mask_store)
        m.release();
    cv->mask_store.clear();
    cv->mask_defined.clear();
    cv->mask_bytes = 0;
}

// The clip-mask bookkeeping of one decoded draw, in submission order: a clip
// gets storage for the id it defines; any draw may use mask 0 or a defined id.
int note_masks(cih_canvas *cv, const draw_desc &desc) {
    int s;
    if (desc.kind == CIH_DRAW_CLIP && (s = ensure_mask(cv, (uint32_t)desc.image)))
        return s;
    if (desc.clip_mask && (cv->mask_defined.empty() || !cv->mask_defined[desc.clip_mask]))
        return fail(CIH_ERROR_MALFORMED, "a draw uses a clip mask no clip has defined");
    if (desc.kind == CIH_DRAW_CLIP)
        cv->mask_defined[(uint32_t)desc.image] = true;
    return CIH_OK;
}

// Whether a frame must end before draw `dd` of canvas `canvas`: a clip that
// redefines a mask id the frame already uses on that canvas (a frame builds
// all its masks before compositing, so each id must name one mask within
// it), or a put on a canvas that already has painting draws in the frame (a
// frame applies its puts before compositing).  `used` (kCutSlots entries)
// and `current` track the frame; they reset at a new canvas.
const size_t kCutSlots = kMaxMasks + 1; // mask ids, then "the canvas has painted"
bool frame_cut(const draw_desc &dd, int canvas, std::vector<unsigned char> &used, int &current) {
    if (canvas != current) {
        std::fill(used.begin(), used.end(), (unsigned char)0);
        current = canvas;
    }
    if (dd.kind == CIH_DRAW_PUT_IMAGE)
        return used[kMaxMasks] != 0;
    if (dd.kind == CIH_DRAW_CLIP && used[(uint32_t)dd.image])
        return true;
    used[(uint32_t)dd.clip_mask] = 1;
    if (dd.kind == CIH_DRAW_CLIP)
        used[(uint32_t)dd.image] = 1;
    else
        used[kMaxMasks] = 1;
    return false;
}

// Decodes pending packets into the canvas' resident program (payload and
// draw records in HBM) split into frames of bounded size.
int prepare_program(cih_canvas *cv) {
    cih_device *d = cv->device;
    int s;
    std::vector<unsigned char> bytes;
    bytes.swap(cv->pending);
    program &P = cv->program;
    P.descs.clear();
    P.frames.clear();
    if (bytes.size() > kPayloadLimit)
        return fail(CIH_ERROR_BUDGET, "a canvas may hold at most " +
                                          std::to_string(kPayloadLimit) +
                                          " bytes of draw packets between flushes");
    size_t at = 0;
    while (at < bytes.size()) {
        draw_desc desc;
        size_t used = decode_packet(bytes.data(), at, bytes.size(), desc, d);
        if (!used)
            return fail(CIH_ERROR_MALFORMED, "malformed draw packet at byte " +
                                                 std::to_string(at));
        desc.canvas = 0;
        if ((s = note_masks(cv, desc)))
            return s;
        P.descs.push_back(desc);
        at += used;
    }
    if ((s = P.payload.reserve(std::max<size_t>(bytes.size(), 16))) ||
        (s = P.draws.reserve(sizeof(draw_desc) * std::max<size_t>(P.descs.size(), 1))))
        return s;
    HIP_TRY(copy_sync(d, P.payload.ptr, bytes.data(), bytes.size(), hipMemcpyHostToDevice));
    HIP_TRY(copy_sync(d, P.draws.ptr, P.descs.data(), sizeof(draw_desc) * P.descs.size(),
                      hipMemcpyHostToDevice));
    // Frames of bounded size: shadow images are the largest intermediates.
    // A frame also ends where frame_cut says (mask id reuse, puts after
    // painting draws).
    const uint64_t shadow_budget = (uint64_t)1 << 30; // floats
    std::vector<unsigned char> used(kCutSlots);
    size_t first = 0;
    while (first < P.descs.size()) {
        size_t last = first;
        uint64_t shadow = 0;
        int current = -1;
        while (last < P.descs.size() && last - first < kFrameDrawLimit) {
            if (frame_cut(P.descs[last], 0, used, current) && last > first)
                break;
            uint64_t need = 0;
            if (P.descs[last].shadow) {
                int b = P.descs[last].border;
                need = 2ull * (uint64_t)(cv->width + 2 * b + 1) * (uint64_t)(cv->height + 2 * b + 1);
            }
            if (last > first && shadow + need > shadow_budget)
                break;
            shadow += need;
            ++last;
        }
        program::span sp;
        sp.first = first;
        sp.count = last - first;
        sp.items_at = 0;
        P.frames.push_back(sp);
        first = last;
    }
    // each frame's canvas record and work items, resident with the program
    std::vector<canvas_desc> records;
    std::vector<uint32_t> items;
    for (program::span &sp : P.frames) {
        canvas_desc cd;
        cd.bitmap = (uint64_t)(uintptr_t)cv->pixels;
        cd.width = cv->width;
        cd.height = cv->height;
        cd.first_draw = 0;
        cd.ndraws = (int32_t)sp.count;
        cd.masks = (uint64_t)(uintptr_t)cv->mask_table.ptr;
        sp.shape = shape_of(P.descs.data() + sp.first, &cd, 1);
        sp.items_at = items.size();
        items.insert(items.end(), sp.shape.items.begin(), sp.shape.items.end());
        records.push_back(cd);
    }
    if ((s = P.canvases.reserve(sizeof(canvas_desc) * std::max<size_t>(records.size(), 1))) ||
        (s = P.items.reserve(4 * std::max<size_t>(items.size(), 1))))
        return s;
    HIP_TRY(copy_sync(d, P.canvases.ptr, records.data(), sizeof(canvas_desc) * records.size(),
                      hipMemcpyHostToDevice));
    HIP_TRY(copy_sync(d, P.items.ptr, items.data(), 4 * items.size(), hipMemcpyHostToDevice));
    P.bytes = bytes.size();
    return CIH_OK;
}

// Launches frame k of the canvas' program.
int run_frame(cih_canvas *cv, size_t k, bool check) {
    program &P = cv->program;
    const program::span &sp = P.frames[k];
    return run_kernels(cv->device, cv->device->main, (const unsigned char *)P.payload.ptr,
                       (const draw_desc *)P.draws.ptr + sp.first, sp.count,
                       (const canvas_desc *)P.canvases.ptr + k, 1,
                       (const uint32_t *)P.items.ptr + sp.items_at, sp.shape, check);
}

// Runs the canvas' program (all its frames) in order.
int run_program(cih_canvas *cv, bool check) {
    int s;
    for (size_t k = 0; k < cv->program.frames.size(); ++k)
        if ((s = run_frame(cv, k, check)))
            return s;
    return CIH_OK;
}

// Decodes and runs all pending packets of one canvas.
int flush_canvas(cih_canvas *cv) {
    if (cv->pending.empty())
        return CIH_OK;
    int s;
    if ((s = prepare_program(cv)))
        return s;
    return run_program(cv, true);
}

uint64_t fnv1a(const unsigned char *p, size_t n, uint64_t h) {
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

} // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

const char *cih_last_error(void) { return g_error.c_str(); }

int cih_device_open(int ordinal, cih_device **device) {
    if (!device)
        return fail(CIH_ERROR_INVALID, "device pointer is NULL");
    *device = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return fail(CIH_ERROR_DEVICE, "no HIP device available");
    }
    if (ordinal < 0 || ordinal >= count)
        return fail(CIH_ERROR_INVALID, "device ordinal out of range");
    cih_device *d = new (std::nothrow) cih_device;
    if (!d)
        return fail(CIH_ERROR_OUT_OF_MEMORY, "out of host memory");
    d->ordinal = ordinal;
    int s;
    if ((s = set_device(d)) != CIH_OK) {
        delete d;
        return s;
    }
    if ((e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking)) != hipSuccess) {
        delete d;
        return hip_fail(e, "hipStreamCreate");
    }
    d->main.stream = d->stream;
    float table[256];
    for (int v = 0; v < 256; ++v)
        table[v] = linearized(v / 255.0f);
    // The image table never moves: frames in flight on other streams (the
    // serving lanes) hold its address while replay threads add images.
    if ((s = d->lut.reserve(sizeof(table))) ||
        (s = d->image_table.reserve(sizeof(image_desc) * kMaxImages)) ||
        (e = copy_sync(d, d->lut.ptr, table, sizeof(table), hipMemcpyHostToDevice)) !=
            hipSuccess) {
        d->lut.release();
        d->image_table.release();
        hipStreamDestroy(d->stream);
        delete d;
        return s ? s : hip_fail(e, "lut upload");
    }
    *device = d;
    return CIH_OK;
}

int cih_device_close(cih_device *d) {
    if (!d)
        return fail(CIH_ERROR_INVALID, "device is NULL");
    set_device(d);
    hipStreamSynchronize(d->stream);
    for (image_entry *img : d->images) {
        img->texels.release();
        delete img;
    }
    d->main.release();
    dbuf *all[] = {&d->image_table, &d->lut, &d->upload, &d->readback};
    for (hipEvent_t e : d->event_pool)
        hipEventDestroy(e);
    for (dbuf *b : all)
        b->release();
    d->staging.release();
    d->readback_host.release();
    hipStreamDestroy(d->stream);
    delete d;
    return CIH_OK;
}

int cih_device_synchronize(cih_device *d) {
    if (!d)
        return fail(CIH_ERROR_INVALID, "device is NULL");
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    HIP_TRY(hipStreamSynchronize(d->stream));
    return CIH_OK;
}

int cih_canvas_create(cih_device *d, int width, int height, cih_canvas **canvas) {
    if (!d || !canvas)
        return fail(CIH_ERROR_INVALID, "NULL device or canvas pointer");
    *canvas = nullptr;
    if (width < 1 || height < 1 || width > 32768 || height > 32768)
        return fail(CIH_ERROR_INVALID, "canvas size must be 1..32768 per side");
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    int s;
    if ((s = set_device(d)))
        return s;
    cih_canvas *c = new (std::nothrow) cih_canvas;
    if (!c)
        return fail(CIH_ERROR_OUT_OF_MEMORY, "out of host memory");
    c->device = d;
    c->width = width;
    c->height = height;
    size_t n = (size_t)width * (size_t)height;
    if ((s = c->bitmap.reserve(16 * n))) {
        delete c;
        return s;
    }
    c->pixels = c->bitmap.ptr;
    hipError_t e = launch_clear(c->pixels, n, d->stream);
    if (e != hipSuccess) {
        c->bitmap.release();
        delete c;
        return hip_fail(e, "clear");
    }
    *canvas = c;
    return CIH_OK;
}

int cih_canvas_destroy(cih_canvas *c) {
    if (!c)
        return fail(CIH_ERROR_INVALID, "canvas is NULL");
    cih_device *d = c->device;
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    set_device(d);
    hipStreamSynchronize(d->stream);
    c->bitmap.release();
    c->program.payload.release();
    c->program.draws.release();
    c->program.canvases.release();
    c->program.items.release();
    release_masks(c);
    delete c;
    return CIH_OK;
}

int cih_canvas_size(cih_canvas *c, int *width, int *height) {
    if (!c)
        return fail(CIH_ERROR_INVALID, "canvas is NULL");
    if (width)
        *width = c->width;
    if (height)
        *height = c->height;
    return CIH_OK;
}

int cih_canvas_device(cih_canvas *c, cih_device **device) {
    if (!c || !device)
        return fail(CIH_ERROR_INVALID, "NULL argument");
    *device = c->device;
    return CIH_OK;
}

int cih_image_create(cih_device *d, const unsigned char *texels, int width, int height,
                     int stride, int32_t *image) {
    if (!d || !texels || !image || width <= 0 || height <= 0)
        return fail(CIH_ERROR_INVALID, "invalid image arguments");
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    int s;
    if ((s = set_device(d)))
        return s;
    size_t n = (size_t)width * (size_t)height;
    std::vector<unsigned char> dense(4 * n);
    for (int y = 0; y < height; ++y)
        std::memcpy(&dense[4 * (size_t)y * width], texels + (ptrdiff_t)y * stride, 4 * (size_t)width);
    uint64_t h = fnv1a(dense.data(), dense.size(), 0xcbf29ce484222325ull ^ (uint64_t)width ^
                                                       ((uint64_t)height << 32));
    auto range = d->image_index.equal_range(h);
    for (auto it = range.first; it != range.second; ++it) {
        image_entry *e = d->images[(size_t)it->second];
        if (e->width == width && e->height == height && e->rgba == dense) {
            *image = it->second;
            return CIH_OK;
        }
    }
    if (d->images.size() >= kMaxImages)
        return fail(CIH_ERROR_BUDGET, "a device holds at most " + std::to_string(kMaxImages) +
                                          " distinct images");
    image_entry *e = new (std::nothrow) image_entry;
    if (!e)
        return fail(CIH_ERROR_OUT_OF_MEMORY, "out of host memory");
    e->hash = h;
    e->width = width;
    e->height = height;
    if ((s = e->texels.reserve(16 * n)) || (s = d->upload.reserve(4 * n))) {
        delete e;
        return s;
    }
    hipError_t err;
    if ((err = hipMemcpyAsync(d->upload.ptr, dense.data(), 4 * n, hipMemcpyHostToDevice,
                              d->stream)) != hipSuccess ||
        (err = launch_prepare_image(d->upload.ptr, e->texels.ptr, n, (const float *)d->lut.ptr,
                                    d->stream)) != hipSuccess ||
        (err = hipStreamSynchronize(d->stream)) != hipSuccess) {
        e->texels.release();
        delete e;
        return hip_fail(err, "image upload");
    }
    // only the new entry is written: frames in flight read the others
    image_desc entry;
    entry.texels = (uint64_t)(uintptr_t)e->texels.ptr;
    entry.width = width;
    entry.height = height;
    int32_t id = (int32_t)d->images.size();
    if ((err = hipMemcpyAsync((image_desc *)d->image_table.ptr + id, &entry, sizeof(entry),
                              hipMemcpyHostToDevice, d->stream)) != hipSuccess ||
        (err = hipStreamSynchronize(d->stream)) != hipSuccess) {
        e->texels.release();
        delete e;
        return hip_fail(err, "image table update");
    }
    e->rgba.swap(dense);
    d->images.push_back(e);
    d->image_index.emplace(h, id);
    *image = id;
    return CIH_OK;
}

int cih_canvas_submit(cih_canvas *c, const void *packets, size_t bytes) {
    if (!c || (!packets && bytes))
        return fail(CIH_ERROR_INVALID, "invalid submit arguments");
    if (bytes & 15)
        return fail(CIH_ERROR_MALFORMED, "packet bytes must be a multiple of 16");
    const unsigned char *p = (const unsigned char *)packets;
    try {
        c->pending.insert(c->pending.end(), p, p + bytes);
    } catch (const std::bad_alloc &) {
        return fail(CIH_ERROR_OUT_OF_MEMORY, "out of host memory");
    }
    return CIH_OK;
}

int cih_canvas_flush(cih_canvas *c) {
    if (!c)
        return fail(CIH_ERROR_INVALID, "canvas is NULL");
    std::lock_guard<std::recursive_mutex> lock(c->device->mutex);
    int s;
    if ((s = set_device(c->device)))
        return s;
    try {
        return flush_canvas(c);
    } catch (const std::bad_alloc &) {
        return fail(CIH_ERROR_OUT_OF_MEMORY, "out of host memory");
    }
}

int cih_canvas_get_image_data(cih_canvas *c, unsigned char *image, int width, int height,
                              int stride, int x, int y) {
    if (!c)
        return fail(CIH_ERROR_INVALID, "canvas is NULL");
    if (!image)
        return CIH_OK; // hpp:3356-3357
    cih_device *d = c->device;
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    int s;
    if ((s = cih_canvas_flush(c)))
        return s;
    if (width <= 0 || height <= 0)
        return CIH_OK;
    size_t n = (size_t)width * (size_t)height;
    if ((s = d->readback.reserve(4 * n)) || (s = d->readback_host.reserve(4 * n)))
        return s;
    HIP_TRY(launch_readback(c->pixels, c->width, c->height, d->readback.ptr, width, height,
                            x, y, d->stream));
    HIP_TRY(hipMemcpyAsync(d->readback_host.ptr, d->readback.ptr, 4 * n, hipMemcpyDeviceToHost,
                           d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    const unsigned char *src = (const unsigned char *)d->readback_host.ptr;
    for (int row = 0; row < height; ++row)
        std::memcpy(image + (ptrdiff_t)row * stride, src + 4 * (size_t)row * width,
                    4 * (size_t)width);
    return CIH_OK;
}

int cih_canvas_put_image_data(cih_canvas *c, const unsigned char *image, int width, int height,
                              int stride, int x, int y) {
    if (!c)
        return fail(CIH_ERROR_INVALID, "canvas is NULL");
    if (!image || width <= 0 || height <= 0)
        return CIH_OK; // hpp:3392-3393
    cih_device *d = c->device;
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    int s;
    if (c->deferred) {
        // the texels become a resident device image; the packet puts them
        // where painter's order reaches it
        int32_t id = -1;
        if ((s = cih_image_create(d, image, width, height, stride, &id)))
            return s;
        cih_draw_packet p;
        std::memset(&p, 0, sizeof(p));
        p.bytes = (uint32_t)sizeof(p);
        p.kind = CIH_DRAW_PUT_IMAGE;
        p.op = 14; // source_over: no whole-canvas effect
        p.global_alpha = 1.0f;
        p.brush = CIH_BRUSH_IMAGE;
        p.image = id;
        p.image_width = width;
        p.image_height = height;
        // positions past 2^20 put nothing on a canvas of at most 32768^2
        const int far = 1 << 20;
        p.start_x = (float)std::min(std::max(x, -far), far);
        p.start_y = (float)std::min(std::max(y, -far), far);
        try {
            const unsigned char *b = (const unsigned char *)&p;
            c->pending.insert(c->pending.end(), b, b + sizeof(p));
        } catch (const std::bad_alloc &) {
            return fail(CIH_ERROR_OUT_OF_MEMORY, "out of host memory");
        }
        return CIH_OK;
    }
    if ((s = cih_canvas_flush(c)))
        return s;
    size_t n = (size_t)width * (size_t)height;
    if ((s = d->upload.reserve(4 * n)) || (s = d->staging.reserve(4 * n)))
        return s;
    unsigned char *dst = (unsigned char *)d->staging.ptr;
    for (int row = 0; row < height; ++row)
        std::memcpy(dst + 4 * (size_t)row * width, image + (ptrdiff_t)row * stride,
                    4 * (size_t)width);
    HIP_TRY(hipMemcpyAsync(d->upload.ptr, d->staging.ptr, 4 * n, hipMemcpyHostToDevice,
                           d->stream));
    HIP_TRY(launch_put_image(d->upload.ptr, width, height, c->pixels, c->width, c->height,
                             x, y, (const float *)d->lut.ptr, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return CIH_OK;
}

int cih_canvas_read_linear(cih_canvas *c, float *rgba) {
    if (!c || !rgba)
        return fail(CIH_ERROR_INVALID, "NULL argument");
    cih_device *d = c->device;
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    int s;
    if ((s = cih_canvas_flush(c)))
        return s;
    HIP_TRY(hipMemcpyAsync(rgba, c->pixels, 16 * (size_t)c->width * c->height,
                           hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return CIH_OK;
}


int cih_canvas_execute(cih_canvas *c, int clear, int readback) {
    if (!c)
        return fail(CIH_ERROR_INVALID, "canvas is NULL");
    cih_device *d = c->device;
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    int s;
    if ((s = set_device(d)) || (s = cih_canvas_flush(c)))
        return s;
    size_t n = (size_t)c->width * (size_t)c->height;
    if (clear)
        HIP_TRY(launch_clear(c->pixels, n, d->stream));
    if ((s = run_program(c, false)))
        return s;
    if (readback) {
        if ((s = d->readback.reserve(4 * n)))
            return s;
        HIP_TRY(launch_readback(c->pixels, c->width, c->height, d->readback.ptr, c->width,
                                c->height, 0, 0, d->stream));
    }
    return CIH_OK;
}

int cih_canvas_stats(cih_canvas *c, uint64_t *stats) {
    if (!c || !stats)
        return fail(CIH_ERROR_INVALID, "NULL argument");
    cih_device *d = c->device;
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    int s;
    if ((s = set_device(d)) || (s = cih_canvas_flush(c)))
        return s;
    for (int k = 0; k < 4; ++k)
        stats[k] = 0;
    HIP_TRY(launch_clear(c->pixels, (size_t)c->width * c->height, d->stream));
    for (size_t k = 0; k < c->program.frames.size(); ++k) {
        if ((s = run_frame(c, k, true)))
            return s;
        const unsigned long long *hc = (const unsigned long long *)d->main.host_counters.ptr;
        stats[0] += hc[COUNTER_PIXEL_DRAWS];
        stats[1] += hc[COUNTER_RUNS];
        stats[2] += hc[COUNTER_POINTS];
        stats[3] += hc[COUNTER_SHADOW];
    }
    return CIH_OK;
}

int cih_device_profile(cih_device *d, int enable) {
    if (!d)
        return fail(CIH_ERROR_INVALID, "device is NULL");
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    d->profiling = enable != 0;
    d->events_used = 0;
    return CIH_OK;
}

int cih_device_profile_read(cih_device *d, char *json, size_t capacity) {
    if (!d || !json || !capacity)
        return fail(CIH_ERROR_INVALID, "invalid arguments");
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    HIP_TRY(hipStreamSynchronize(d->stream));
    double total[kFrameKernels] = {0};
    size_t frames = d->events_used / (kFrameKernels + 1);
    for (size_t f = 0; f < frames; ++f)
        for (int k = 0; k < kFrameKernels; ++k) {
            float ms = 0.0f;
            hipEvent_t *e = &d->event_pool[f * (kFrameKernels + 1)];
            HIP_TRY(hipEventElapsedTime(&ms, e[k], e[k + 1]));
            total[k] += ms;
        }
    std::string out = "{\"frames\": " + std::to_string(frames);
    for (int k = 0; k < kFrameKernels; ++k) {
        char buf[128];
        std::snprintf(buf, sizeof(buf), ", \"%s\": %.6f", kFrameKernelNames[k], total[k]);
        out += buf;
    }
    out += "}";
    std::snprintf(json, capacity, "%s", out.c_str());
    d->events_used = 0;
    return CIH_OK;
}

int cih_trace_polygons(cih_canvas *c, int draw, float *points, size_t point_capacity,
                       uint32_t *subpaths, size_t subpath_capacity, size_t *npoints,
                       size_t *nsubpaths) {
    return cih_trace_stage(c, draw, 2, points, point_capacity, subpaths, subpath_capacity,
                           npoints, nsubpaths);
}

int cih_trace_stage(cih_canvas *c, int draw, int stage, float *points, size_t point_capacity,
                    uint32_t *subpaths, size_t subpath_capacity, size_t *npoints,
                    size_t *nsubpaths) {
    if (!c || draw < 0 || !npoints || !nsubpaths)
        return fail(CIH_ERROR_INVALID, "invalid trace arguments");
    cih_device *d = c->device;
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    if ((size_t)draw >= d->main.last_ndraws)
        return fail(CIH_ERROR_INVALID, "no such draw in the last frame");
    if (stage < 0 || stage > 2)
        return fail(CIH_ERROR_INVALID, "stage must be 0 (flatten), 1 (dash) or 2 (polygons)");
    draw_state S;
    draw_desc D;
    HIP_TRY(copy_sync(d, &S, (draw_state *)d->main.state.ptr + draw, sizeof(S),
                      hipMemcpyDeviceToHost));
    HIP_TRY(copy_sync(d, &D, d->main.last_draws + draw, sizeof(D), hipMemcpyDeviceToHost));
    // k_geometry records a stage only where it runs: flattening for paths and
    // clips (a polyline draw's own lines are its flatten stage), dashing for
    // dashed strokes; the other stages of a draw are empty
    bool path = D.kind == CIH_DRAW_FILL_PATH || D.kind == CIH_DRAW_STROKE_PATH ||
                D.kind == CIH_DRAW_CLIP;
    bool lines = D.kind == CIH_DRAW_FILL_LINES || D.kind == CIH_DRAW_STROKE_LINES;
    bool stroke = D.kind == CIH_DRAW_STROKE_PATH || D.kind == CIH_DRAW_STROKE_LINES;
    uint64_t pts = 0, cnt = 0;
    size_t np = 0, ns = 0;
    if (stage == 0 && path) {
        pts = S.flat_points, cnt = S.flat_counts, np = S.flat_npoints, ns = S.flat_nsub;
    } else if (stage == 0 && lines) {
        pts = (uint64_t)(uintptr_t)(d->main.last_payload + D.points_at);
        cnt = (uint64_t)(uintptr_t)(d->main.last_payload + D.subpaths_at);
        np = D.npoints, ns = D.nsubpaths;
    } else if (stage == 1 && stroke && D.ndash) {
        pts = S.dash_points, cnt = S.dash_counts, np = S.dash_npoints, ns = S.dash_nsub;
    } else if (stage == 2 && (path || lines)) {
        pts = S.poly_points, cnt = S.poly_counts, np = S.poly_npoints, ns = S.poly_nsub;
    }
    *npoints = np;
    *nsubpaths = ns;
    if (points && point_capacity && np)
        HIP_TRY(copy_sync(d, points, (const void *)(uintptr_t)pts,
                          8 * std::min(point_capacity, np), hipMemcpyDeviceToHost));
    if (subpaths && subpath_capacity && ns)
        HIP_TRY(copy_sync(d, subpaths, (const void *)(uintptr_t)cnt,
                          8 * std::min(subpath_capacity, ns), hipMemcpyDeviceToHost));
    return CIH_OK;
}

int cih_trace_runs(cih_canvas *c, int draw, int pass, int32_t *xs, int32_t *ys, float *deltas,
                   size_t capacity, size_t *nruns) {
    if (!c || draw < 0 || !nruns)
        return fail(CIH_ERROR_INVALID, "invalid trace arguments");
    cih_device *d = c->device;
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    if ((size_t)draw >= d->main.last_ndraws)
        return fail(CIH_ERROR_INVALID, "no such draw in the last frame");
    draw_state S;
    HIP_TRY(copy_sync(d, &S, (draw_state *)d->main.state.ptr + draw, sizeof(S),
                      hipMemcpyDeviceToHost));
    int row0 = pass ? S.srow0 : S.row0, rows = pass ? S.srows : S.rows;
    uint64_t rowoff = pass ? S.srowoff : S.rowoff, base = pass ? S.sruns : S.runs;
    *nruns = 0;
    if (rows <= 0)
        return CIH_OK;
    std::vector<uint32_t> off((size_t)rows + 1);
    HIP_TRY(copy_sync(d, off.data(), (uint32_t *)d->main.words.ptr + rowoff, 4 * off.size(),
                      hipMemcpyDeviceToHost));
    size_t total = off.back();
    *nruns = total;
    std::vector<run> r(total);
    if (total)
        HIP_TRY(copy_sync(d, r.data(), (run *)d->main.runs.ptr + base, sizeof(run) * total,
                          hipMemcpyDeviceToHost));
    size_t k = 0;
    for (int row = 0; row < rows; ++row)
        for (uint32_t i = off[(size_t)row]; i < off[(size_t)row + 1] && k < capacity; ++i, ++k) {
            if (xs)
                xs[k] = r[i].x;
            if (ys)
                ys[k] = row0 + row;
            if (deltas)
                deltas[k] = r[i].delta;
        }
    return CIH_OK;
}

} // extern "C"

int cih_path_contains(const float *points, size_t npoints, const uint32_t *subpaths,
                      size_t nsubpaths, float x, float y, int *inside) {
    if (!inside || (npoints && !points) || (nsubpaths && !subpaths))
        return fail(CIH_ERROR_INVALID, "invalid path arguments");
    *inside = 0;
    uint64_t total = 0;
    for (size_t s = 0; s < nsubpaths; ++s) {
        uint32_t c = subpaths[2 * s];
        if (c == 0 || (c - 1) % 3 != 0)
            return fail(CIH_ERROR_MALFORMED, "a subpath holds a start point and 3 per cubic");
        total += c;
    }
    if (total != npoints)
        return fail(CIH_ERROR_MALFORMED, "subpath counts do not add up to the points");
    try {
        // path_to_lines(false): start point, then each cubic's flattening
        // with no angular test (fills emit end points only)
        const cih::f2 *in = (const cih::f2 *)points;
        std::vector<cih::f2> line;
        std::vector<size_t> ends;
        struct sink {
            std::vector<cih::f2> *v;
            void point(cih::f2 p) { v->push_back(p); }
        } out = {&line};
        size_t index = 0;
        for (size_t s = 0; s < nsubpaths; ++s) {
            size_t ending = index + subpaths[2 * s];
            cih::f2 p1 = in[index++];
            line.push_back(p1);
            for (; index < ending; index += 3) {
                cih::add_bezier(out, p1, in[index], in[index + 1], in[index + 2], -1.0f);
                p1 = in[index + 2];
            }
            ends.push_back(line.size());
        }
        // crossings of the horizontal through (x, y), each subpath implicitly
        // closed; a point on an edge counts as inside
        int winding = 0;
        size_t beginning = 0;
        for (size_t s = 0; s < ends.size(); ++s) {
            size_t ending = ends[s];
            for (size_t i = beginning; i < ending; ++i) {
                cih::f2 from = line[i], to = line[i + 1 < ending ? i + 1 : beginning];
                if ((from.y < y && y <= to.y) || (to.y < y && y <= from.y)) {
                    float side = cih::dot(cih::perp(cih::sub(to, from)),
                                          cih::sub(cih::mk(x, y), from));
                    if (side == 0.0f) {
                        *inside = 1;
                        return CIH_OK;
                    }
                    winding += side > 0.0f ? 1 : -1;
                } else if (from.y == y && y == to.y &&
                           ((from.x <= x && x <= to.x) || (to.x <= x && x <= from.x))) {
                    *inside = 1;
                    return CIH_OK;
                }
            }
            beginning = ending;
        }
        *inside = winding != 0;
    } catch (const std::bad_alloc &) {
        return fail(CIH_ERROR_OUT_OF_MEMORY, "out of host memory");
    }
    return CIH_OK;
}

// ---------------------------------------------------------------------------
// Batches: many independent canvases rendered by shared frames.

struct cih_batch {
    cih_device *device = nullptr;
    std::vector<cih_canvas *> views;   // canvases over the shared bitmap storage
    std::vector<uint64_t> pixel_prefix; // host, count + 1
    dbuf bitmaps, all_canvases, prefix_dev, srgb;
    hbuf srgb_host;
    dbuf payload, draws;
    std::vector<draw_desc> descs;
    struct frame_rec {
        size_t first_draw, ndraws, canvases_at, ncanvases, items_at;
        uint64_t pixels; // of the canvases the frame draws on
        frame_shape shape;
    };
    std::vector<frame_rec> frames;
    dbuf frame_canvases, frame_items;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void *> graph_buffers; // the lane buffers the graph's kernels use
    bool ran = false;
    bool use_graph = true;
    fine_config fine; // fine-stage launch shape (cih_batch_configure / cih_batch_tune)
};

namespace {

// Replays streams [first, last) into their views (host work, any thread).
void replay_range(cih_batch *b, const void *const *streams, const size_t *bytes, size_t first,
                  size_t last, int *status) {
    for (size_t i = first; i < last; ++i) {
        int st = cih_canvas_replay(b->views[i], streams[i], bytes[i]);
        if (st != CIH_OK) {
            *status = st;
            return;
        }
    }
}

// The device buffers a captured frame bakes into its kernel arguments.  Any
// later growth of the lane (another canvas or batch on the device) moves some
// of them, and the graph must then be captured again.
std::vector<void *> lane_buffers(const exec_lane &L) {
    return {L.points.ptr, L.words.ptr, L.runs.ptr, L.steps.ptr, L.shadow.ptr, L.counters.ptr,
            L.state.ptr};
}

void drop_graph(cih_batch *b) {
    if (b->exec)
        hipGraphExecDestroy(b->exec);
    if (b->graph)
        hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_buffers.clear();
}

void destroy_batch(cih_batch *b) {
    drop_graph(b);
    for (cih_canvas *v : b->views) {
        release_masks(v);
        delete v;
    }
    dbuf *all[] = {&b->bitmaps, &b->all_canvases, &b->prefix_dev, &b->srgb, &b->payload,
                   &b->draws, &b->frame_canvases, &b->frame_items};
    for (dbuf *x : all)
        x->release();
    b->srgb_host.release();
    delete b;
}

// Records the whole batch (clear, frames, sRGB8 conversion) on the stream.
int enqueue_batch(cih_batch *b, bool check) {
    cih_device *d = b->device;
    int s;
    size_t n = b->views.size();
    uint64_t total = b->pixel_prefix.back();
    HIP_TRY(launch_clear_batch(b->all_canvases.ptr, b->prefix_dev.ptr, (int)n, total, d->stream));
    for (const cih_batch::frame_rec &f : b->frames) {
        if ((s = run_kernels(d, d->main, (const unsigned char *)b->payload.ptr,
                             (const draw_desc *)b->draws.ptr + f.first_draw, f.ndraws,
                             (const canvas_desc *)b->frame_canvases.ptr + f.canvases_at,
                             f.ncanvases, (const uint32_t *)b->frame_items.ptr + f.items_at,
                             f.shape, check, b->fine)))
            return s;
    }
    HIP_TRY(launch_readback_batch(b->all_canvases.ptr, b->prefix_dev.ptr, (int)n, total,
                                  b->srgb.ptr, d->stream));
    return CIH_OK;
}

} // namespace

extern "C" {

int cih_batch_create(cih_device *d, int count, const void *const *streams, const size_t *bytes,
                     cih_batch **out) {
    if (!d || count <= 0 || !streams || !bytes || !out)
        return fail(CIH_ERROR_INVALID, "invalid batch arguments");
    *out = nullptr;
    std::unique_lock<std::recursive_mutex> lock(d->mutex);
    int s;
    if ((s = set_device(d)))
        return s;
    cih_batch *b = new (std::nothrow) cih_batch;
    if (!b)
        return fail(CIH_ERROR_OUT_OF_MEMORY, "out of host memory");
    b->device = d;
    try {
        b->pixel_prefix.assign((size_t)count + 1, 0);
        for (int i = 0; i < count; ++i) {
            cih_reader r;
            if (cih_reader_init(&r, streams[i], bytes[i]) != 0) {
                destroy_batch(b);
                return fail(CIH_ERROR_MALFORMED, "malformed stream header in batch");
            }
            cih_canvas *v = new cih_canvas;
            v->device = d;
            v->deferred = true;
            v->width = r.width;
            v->height = r.height;
            b->views.push_back(v);
            b->pixel_prefix[(size_t)i + 1] =
                b->pixel_prefix[(size_t)i] + (uint64_t)r.width * (uint64_t)r.height;
        }
        uint64_t total = b->pixel_prefix.back();
        if (total > kBatchTotalPixelLimit) {
            destroy_batch(b);
            return fail(CIH_ERROR_BUDGET, "a batch may hold at most " +
                                              std::to_string(kBatchTotalPixelLimit) +
                                              " canvas pixels (it has " +
                                              std::to_string(total) + ")");
        }
        if ((s = b->bitmaps.reserve(16 * total)) || (s = b->srgb.reserve(4 * total)) ||
            (s = b->prefix_dev.reserve(8 * b->pixel_prefix.size())) ||
            (s = b->all_canvases.reserve(sizeof(canvas_desc) * (size_t)count))) {
            destroy_batch(b);
            return s;
        }
        std::vector<canvas_desc> table((size_t)count);
        for (int i = 0; i < count; ++i) {
            cih_canvas *v = b->views[(size_t)i];
            v->pixels = (char *)b->bitmaps.ptr + 16 * b->pixel_prefix[(size_t)i];
            table[(size_t)i].bitmap = (uint64_t)(uintptr_t)v->pixels;
            table[(size_t)i].width = v->width;
            table[(size_t)i].height = v->height;
            table[(size_t)i].first_draw = 0;
            table[(size_t)i].ndraws = 0;
        }
        // host state machine: replay every stream into packets, in parallel.
        // The replay threads add images to the device (which takes the device
        // lock), so this thread lets go of it until they are done.
        lock.unlock();
        unsigned workers = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        workers = std::min<unsigned>(workers, (unsigned)count);
        std::vector<int> status(workers, CIH_OK);
        std::vector<std::thread> pool;
        pool.reserve(workers);
        size_t per = ((size_t)count + workers - 1) / workers;
        for (unsigned w = 0; w < workers; ++w) {
            size_t first = w * per, last = std::min((size_t)count, first + per);
            try {
                pool.emplace_back(replay_range, b, streams, bytes, first, last, &status[w]);
            } catch (const std::system_error &) { // no thread: replay the range here
                replay_range(b, streams, bytes, first, last, &status[w]);
            }
        }
        for (std::thread &t : pool)
            t.join();
        lock.lock();
        for (int st : status)
            if (st != CIH_OK) {
                destroy_batch(b);
                return fail(st, "a batch stream failed to replay");
            }
        // decode all packets into one resident draw list
        std::vector<unsigned char> payload;
        std::vector<size_t> first_draw((size_t)count + 1, 0);
        for (int i = 0; i < count; ++i) {
            cih_canvas *v = b->views[(size_t)i];
            size_t base = payload.size();
            if (base + v->pending.size() > kPayloadLimit) {
                destroy_batch(b);
                return fail(CIH_ERROR_BUDGET, "a batch may hold at most " +
                                                  std::to_string(kPayloadLimit) +
                                                  " bytes of draw packets");
            }
            payload.insert(payload.end(), v->pending.begin(), v->pending.end());
            size_t at = base;
            while (at < payload.size()) {
                draw_desc desc;
                size_t used = decode_packet(payload.data(), at, payload.size(), desc, d);
                if (!used) {
                    destroy_batch(b);
                    return fail(CIH_ERROR_MALFORMED, "malformed draw packet in batch");
                }
                if ((s = note_masks(v, desc))) {
                    destroy_batch(b);
                    return s;
                }
                desc.canvas = i;
                b->descs.push_back(desc);
                at += used;
            }
            std::vector<unsigned char>().swap(v->pending);
            first_draw[(size_t)i + 1] = b->descs.size();
        }
        // each canvas' clip masks (allocated while decoding)
        for (int i = 0; i < count; ++i)
            table[(size_t)i].masks = (uint64_t)(uintptr_t)b->views[(size_t)i]->mask_table.ptr;
        // frames: consecutive draw ranges of bounded size; canvas lists per
        // frame.  A frame over one canvas takes up to kFrameDrawLimit draws; one
        // that spans several keeps to kBatchDrawLimit draws and kBatchPixelLimit
        // canvas pixels (a first canvas above that runs in a frame of its own).
        const size_t max_draws = kFrameDrawLimit;
        const uint64_t shadow_budget = (uint64_t)1 << 30;
        std::vector<canvas_desc> frame_canvases;
        std::vector<uint32_t> frame_items;
        std::vector<unsigned char> used(kCutSlots);
        size_t draw = 0;
        while (draw < b->descs.size()) {
            size_t last = draw;
            uint64_t shadow = 0, pixels = 0;
            int current = -1, first_canvas = b->descs[draw].canvas, last_canvas = -1;
            while (last < b->descs.size() && last - draw < max_draws) {
                const draw_desc &dd = b->descs[last];
                if (dd.canvas != first_canvas && last - draw >= kBatchDrawLimit)
                    break;
                bool new_canvas = dd.canvas != last_canvas;
                uint64_t px = new_canvas ? b->pixel_prefix[(size_t)dd.canvas + 1] -
                                               b->pixel_prefix[(size_t)dd.canvas]
                                         : 0;
                if (last > draw && pixels + px > kBatchPixelLimit)
                    break;
                if (frame_cut(dd, dd.canvas, used, current) && last > draw)
                    break;
                uint64_t need = 0;
                if (dd.shadow) {
                    const cih_canvas *v = b->views[(size_t)dd.canvas];
                    int bo = dd.border;
                    need = 2ull * (uint64_t)(v->width + 2 * bo + 1) * (uint64_t)(v->height + 2 * bo + 1);
                }
                if (last > draw && shadow + need > shadow_budget)
                    break;
                shadow += need;
                pixels += px;
                if (new_canvas)
                    last_canvas = dd.canvas;
                ++last;
            }
            cih_batch::frame_rec f;
            f.first_draw = draw;
            f.ndraws = last - draw;
            f.canvases_at = frame_canvases.size();
            f.pixels = pixels;
            int c0 = b->descs[draw].canvas, c1 = b->descs[last - 1].canvas;
            for (int c = c0; c <= c1; ++c) {
                size_t lo = std::max(first_draw[(size_t)c], draw);
                size_t hi = std::min(first_draw[(size_t)c + 1], last);
                if (hi <= lo)
                    continue;
                canvas_desc cd = table[(size_t)c];
                cd.first_draw = (int32_t)(lo - draw);
                cd.ndraws = (int32_t)(hi - lo);
                for (size_t k = lo; k < hi; ++k) // draw_desc.canvas: index in the frame
                    b->descs[k].canvas = (int32_t)(frame_canvases.size() - f.canvases_at);
                frame_canvases.push_back(cd);
            }
            f.ncanvases = frame_canvases.size() - f.canvases_at;
            f.shape = shape_of(b->descs.data() + draw, frame_canvases.data() + f.canvases_at,
                               f.ncanvases);
            f.items_at = frame_items.size();
            frame_items.insert(frame_items.end(), f.shape.items.begin(), f.shape.items.end());
            b->frames.push_back(f);
            draw = last;
        }
        if ((s = b->payload.reserve(std::max<size_t>(payload.size(), 16))) ||
            (s = b->draws.reserve(sizeof(draw_desc) * std::max<size_t>(b->descs.size(), 1))) ||
            (s = b->frame_canvases.reserve(sizeof(canvas_desc) *
                                           std::max<size_t>(frame_canvases.size(), 1))) ||
            (s = b->frame_items.reserve(4 * std::max<size_t>(frame_items.size(), 1)))) {
            destroy_batch(b);
            return s;
        }
        // uploaded on the stream the batch's kernels run on, and waited for
        hipError_t e;
        if ((e = copy_sync(d, b->payload.ptr, payload.data(), payload.size(),
                           hipMemcpyHostToDevice)) != hipSuccess ||
            (e = copy_sync(d, b->draws.ptr, b->descs.data(), sizeof(draw_desc) * b->descs.size(),
                           hipMemcpyHostToDevice)) != hipSuccess ||
            (e = copy_sync(d, b->frame_canvases.ptr, frame_canvases.data(),
                           sizeof(canvas_desc) * frame_canvases.size(),
                           hipMemcpyHostToDevice)) != hipSuccess ||
            (e = copy_sync(d, b->frame_items.ptr, frame_items.data(), 4 * frame_items.size(),
                           hipMemcpyHostToDevice)) != hipSuccess ||
            (e = copy_sync(d, b->all_canvases.ptr, table.data(),
                           sizeof(canvas_desc) * table.size(), hipMemcpyHostToDevice)) !=
                hipSuccess ||
            (e = copy_sync(d, b->prefix_dev.ptr, b->pixel_prefix.data(),
                           8 * b->pixel_prefix.size(), hipMemcpyHostToDevice)) != hipSuccess) {
            destroy_batch(b);
            return hip_fail(e, "batch upload");
        }
    } catch (const std::bad_alloc &) {
        destroy_batch(b);
        return fail(CIH_ERROR_OUT_OF_MEMORY, "out of host memory");
    } catch (const std::system_error &) {
        destroy_batch(b);
        return fail(CIH_ERROR_DEVICE, "system error while creating the batch");
    }
    *out = b;
    return CIH_OK;
}

int cih_batch_execute(cih_batch *b, int to_host) {
    if (!b)
        return fail(CIH_ERROR_INVALID, "batch is NULL");
    cih_device *d = b->device;
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    int s;
    if ((s = set_device(d)))
        return s;
    if (b->exec && b->graph_buffers != lane_buffers(d->main)) {
        // the lane grew since the capture: run checked and capture again
        HIP_TRY(hipStreamSynchronize(d->stream));
        drop_graph(b);
        b->ran = false;
    }
    if (!b->ran) {
        // first run: checked frames size the arenas; then capture the graph
        if ((s = enqueue_batch(b, true)))
            return s;
        b->ran = true;
        if (b->use_graph && !d->profiling) {
            HIP_TRY(hipStreamSynchronize(d->stream));
            d->capturing = true;
            hipError_t e = hipStreamBeginCapture(d->stream, hipStreamCaptureModeRelaxed);
            if (e == hipSuccess) {
                s = enqueue_batch(b, false);
                e = hipStreamEndCapture(d->stream, &b->graph);
            }
            d->capturing = false;
            if (s)
                return s;
            if (e == hipSuccess)
                e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                b->use_graph = false; // fall back to stream launches of the same kernels
                drop_graph(b);
            } else {
                b->graph_buffers = lane_buffers(d->main);
            }
        }
    } else if (b->exec && !d->profiling) {
        HIP_TRY(hipGraphLaunch(b->exec, d->stream));
    } else if ((s = enqueue_batch(b, false))) {
        return s;
    }
    if (to_host) {
        uint64_t total = b->pixel_prefix.back();
        if ((s = b->srgb_host.reserve(4 * total)))
            return s;
        HIP_TRY(hipMemcpyAsync(b->srgb_host.ptr, b->srgb.ptr, 4 * total, hipMemcpyDeviceToHost,
                               d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
    }
    return CIH_OK;
}

namespace {

fine_config fine_of(int config) {
    fine_config f;
    f.variant = config / 2;
    f.persistent = (config & 1) != 0;
    return f;
}

} // namespace

int cih_batch_configure(cih_batch *b, int config) {
    if (!b || config < 0 || config >= CIH_FINE_CONFIGS)
        return fail(CIH_ERROR_INVALID, "invalid fine configuration");
    cih_device *d = b->device;
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    int s;
    if ((s = set_device(d)))
        return s;
    fine_config f = fine_of(config);
    if (f.variant == b->fine.variant && f.persistent == b->fine.persistent)
        return CIH_OK;
    HIP_TRY(hipStreamSynchronize(d->stream));
    drop_graph(b); // the graph bakes the fine stage's launch shape
    b->fine = f;
    b->ran = false; // the next execute runs checked and captures again
    return CIH_OK;
}

int cih_batch_tune(cih_batch *b, int reps, double *ms, int *best) {
    if (!b || reps < 1)
        return fail(CIH_ERROR_INVALID, "invalid tuning arguments");
    cih_device *d = b->device;
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    int s;
    if ((s = set_device(d)))
        return s;
    if (!b->ran && (s = cih_batch_execute(b, 0))) // sizes the arenas
        return s;
    HIP_TRY(hipStreamSynchronize(d->stream));
    hipEvent_t start, stop;
    HIP_TRY(hipEventCreate(&start));
    if (hipEventCreate(&stop) != hipSuccess) {
        hipEventDestroy(start);
        return fail(CIH_ERROR_DEVICE, "hipEventCreate failed");
    }
    double times[CIH_FINE_CONFIGS];
    int chosen = 0;
    fine_config keep = b->fine;
    s = CIH_OK;
    for (int c = 0; c < CIH_FINE_CONFIGS && s == CIH_OK; ++c) {
        drop_graph(b);
        b->fine = fine_of(c);
        // one untimed execute, then `reps` timed ones (stream launches; the
        // arenas already fit every configuration)
        hipError_t e = hipSuccess;
        if ((s = enqueue_batch(b, false)))
            break;
        if ((e = hipEventRecord(start, d->stream)) != hipSuccess) {
            s = hip_fail(e, "hipEventRecord");
            break;
        }
        for (int r = 0; r < reps && s == CIH_OK; ++r)
            s = enqueue_batch(b, false);
        if (s)
            break;
        float t = 0.0f;
        if ((e = hipEventRecord(stop, d->stream)) != hipSuccess ||
            (e = hipEventSynchronize(stop)) != hipSuccess ||
            (e = hipEventElapsedTime(&t, start, stop)) != hipSuccess) {
            s = hip_fail(e, "batch tuning timer");
            break;
        }
        times[c] = (double)t / reps;
        if (times[c] < times[chosen])
            chosen = c;
    }
    hipEventDestroy(start);
    hipEventDestroy(stop);
    drop_graph(b);
    b->ran = false; // the next execute runs checked and captures again
    if (s) {
        b->fine = keep;
        return s;
    }
    b->fine = fine_of(chosen);
    if (ms)
        for (int c = 0; c < CIH_FINE_CONFIGS; ++c)
            ms[c] = times[c];
    if (best)
        *best = chosen;
    return CIH_OK;
}

int cih_batch_size(cih_batch *b, int index, int *width, int *height) {
    if (!b || index < 0 || (size_t)index >= b->views.size())
        return fail(CIH_ERROR_INVALID, "invalid batch index");
    if (width)
        *width = b->views[(size_t)index]->width;
    if (height)
        *height = b->views[(size_t)index]->height;
    return CIH_OK;
}

int cih_batch_image(cih_batch *b, int index, unsigned char *image) {
    if (!b || !image || index < 0 || (size_t)index >= b->views.size())
        return fail(CIH_ERROR_INVALID, "invalid batch image arguments");
    if (!b->srgb_host.ptr)
        return fail(CIH_ERROR_INVALID, "execute the batch with to_host first");
    std::memcpy(image, (const unsigned char *)b->srgb_host.ptr + 4 * b->pixel_prefix[(size_t)index],
                4 * (size_t)(b->pixel_prefix[(size_t)index + 1] - b->pixel_prefix[(size_t)index]));
    return CIH_OK;
}

int cih_batch_linear(cih_batch *b, int index, float *rgba) {
    if (!b || !rgba || index < 0 || (size_t)index >= b->views.size())
        return fail(CIH_ERROR_INVALID, "invalid batch arguments");
    cih_device *d = b->device;
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    const cih_canvas *v = b->views[(size_t)index];
    HIP_TRY(hipMemcpyAsync(rgba, v->pixels, 16 * (size_t)v->width * v->height,
                           hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
    return CIH_OK;
}

int cih_batch_stats(cih_batch *b, uint64_t *stats) {
    if (!b || !stats)
        return fail(CIH_ERROR_INVALID, "NULL argument");
    stats[0] = b->views.size();
    stats[1] = b->descs.size();
    stats[2] = b->frames.size();
    stats[3] = b->pixel_prefix.back();
    return CIH_OK;
}

int cih_batch_frame(cih_batch *b, int index, uint64_t *info) {
    if (!b || !info)
        return fail(CIH_ERROR_INVALID, "NULL argument");
    if (index < 0 || (size_t)index >= b->frames.size())
        return fail(CIH_ERROR_INVALID, "frame index out of range");
    const cih_batch::frame_rec &f = b->frames[(size_t)index];
    info[0] = f.first_draw;
    info[1] = f.ndraws;
    info[2] = f.ncanvases;
    info[3] = f.pixels;
    return CIH_OK;
}

int cih_batch_destroy(cih_batch *b) {
    if (!b)
        return fail(CIH_ERROR_INVALID, "batch is NULL");
    cih_device *d = b->device;
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    set_device(d);
    hipStreamSynchronize(d->stream);
    destroy_batch(b);
    return CIH_OK;
}

} // extern "C"

// ---------------------------------------------------------------------------
// Serving: continuous batching of requests over two execution lanes.
//
// submit() copies a request's command stream and queues it; worker threads
// replay it through canvas_ity_hip.hpp into draw packets (host state machine)
// and decode them; the scheduler admits prepared requests in arrival order
// while their canvases fit the HBM pool, launches them as one batch on a free
// lane, and keeps admitting into the other lane while the first runs.  A
// finished batch's sRGB8 results are copied out and its pool space freed, so
// requests that arrived meanwhile get in as soon as memory and a lane free up.

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over [0, size) with coalescing of freed ranges.
struct range_allocator {
    std::map<uint64_t, uint64_t> free_ranges; // offset -> size
    void reset(uint64_t size) {
        free_ranges.clear();
        if (size)
            free_ranges[0] = size;
    }
    bool alloc(uint64_t n, uint64_t &at) {
        for (auto it = free_ranges.begin(); it != free_ranges.end(); ++it)
            if (it->second >= n) {
                at = it->first;
                uint64_t rest = it->second - n;
                free_ranges.erase(it);
                if (rest)
                    free_ranges[at + n] = rest;
                return true;
            }
        return false;
    }
    void release(uint64_t at, uint64_t n) {
        auto next = free_ranges.lower_bound(at);
        if (next != free_ranges.end() && at + n == next->first) {
            n += next->second;
            free_ranges.erase(next);
        }
        auto it = free_ranges.lower_bound(at);
        if (it != free_ranges.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == at) {
                prev->second += n;
                return;
            }
        }
        free_ranges[at] = n;
    }
};

struct server_request {
    uint64_t id = 0;
    std::vector<unsigned char> stream, packets, result;
    std::vector<draw_desc> descs;
    int width = 0, height = 0, status = CIH_OK;
    uint64_t pool_at = 0, bytes = 0;
    double t_submit = 0, t_done = 0;
    cih_canvas *masks = nullptr; // the request's clip masks, while it has any
    void drop_masks() {
        if (masks) {
            release_masks(masks);
            delete masks;
            masks = nullptr;
        }
    }
};

struct server_batch {
    std::vector<server_request *> requests;
    std::vector<uint64_t> pixel_prefix;
    std::vector<draw_desc> descs;
    struct frame_rec {
        size_t first_draw, ndraws, canvases_at, ncanvases, items_at;
        frame_shape shape;
    };
    std::vector<frame_rec> frames;
    dbuf payload, draws, frame_canvases, frame_items, all_canvases, prefix, srgb;
    hbuf staging, srgb_host, counters_host;
    hipEvent_t done = nullptr;
    void release() {
        dbuf *all[] = {&payload, &draws, &frame_canvases, &frame_items, &all_canvases, &prefix,
                       &srgb};
        for (dbuf *b : all)
            b->release();
        staging.release();
        srgb_host.release();
        counters_host.release();
        if (done)
            hipEventDestroy(done);
        done = nullptr;
    }
};

} // namespace

struct cih_server {
    cih_device *device = nullptr;
    uint64_t budget = 0;
    int max_batch = 0;
    dbuf pool;
    range_allocator allocator;
    exec_lane lanes[2];
    server_batch *inflight[2] = {nullptr, nullptr};
    std::mutex mu;
    std::condition_variable work, done;
    std::deque<server_request *> arrived, prepared;
    std::deque<uint64_t> completed;
    std::unordered_map<uint64_t, server_request *> requests;
    uint64_t next_id = 1;
    bool stopping = false;
    std::vector<std::thread> workers;
    std::thread scheduler;
    uint64_t batches = 0, served = 0, retries = 0, max_inflight_requests = 0;
    int error = CIH_OK;
};

namespace {

void server_worker(cih_server *sv) {
    cih_device *d = sv->device;
    for (;;) {
        server_request *r;
        {
            std::unique_lock<std::mutex> lock(sv->mu);
            sv->work.wait(lock, [sv] { return sv->stopping || !sv->arrived.empty(); });
            if (sv->arrived.empty())
                return;
            r = sv->arrived.front();
            sv->arrived.pop_front();
        }
        cih_canvas *view = new (std::nothrow) cih_canvas;
        int st = view ? CIH_OK : CIH_ERROR_OUT_OF_MEMORY;
        if (view) {
            view->device = d;
            view->deferred = true;
            view->width = r->width;
            view->height = r->height;
            st = cih_canvas_replay(view, r->stream.data(), r->stream.size());
        }
        if (st == CIH_OK && view->pending.size() > kPayloadLimit)
            st = CIH_ERROR_BUDGET; // draw records address packets with 32-bit offsets
        if (st == CIH_OK) {
            std::lock_guard<std::recursive_mutex> dlock(d->mutex); // images table
            r->packets.swap(view->pending);
            size_t at = 0;
            while (at < r->packets.size()) {
                draw_desc desc;
                size_t used = decode_packet(r->packets.data(), at, r->packets.size(), desc, d);
                if (!used) {
                    st = CIH_ERROR_MALFORMED;
                    break;
                }
                if ((st = note_masks(view, desc)) != CIH_OK)
                    break;
                r->descs.push_back(desc);
                at += used;
            }
        }
        if (view && view->mask_table.ptr && st == CIH_OK) {
            r->masks = view; // kept for the request's frames
        } else if (view) {
            release_masks(view);
            delete view;
        }
        std::vector<unsigned char>().swap(r->stream);
        std::lock_guard<std::mutex> lock(sv->mu);
        r->status = st;
        if (st != CIH_OK) {
            r->t_done = now_ms();
            sv->completed.push_back(r->id);
            sv->done.notify_all();
        } else {
            sv->prepared.push_back(r);
            sv->work.notify_all();
        }
    }
}

// Builds and launches a batch on lane k (caller holds sv->mu).
int launch_server_batch(cih_server *sv, int k, server_batch *b) {
    cih_device *d = sv->device;
    exec_lane &L = sv->lanes[k];
    int s;
    size_t n = b->requests.size();
    b->pixel_prefix.assign(n + 1, 0);
    std::vector<canvas_desc> table(n);
    std::vector<unsigned char> payload;
    std::vector<size_t> first_draw(n + 1, 0);
    for (size_t i = 0; i < n; ++i) {
        server_request *r = b->requests[i];
        b->pixel_prefix[i + 1] = b->pixel_prefix[i] + (uint64_t)r->width * r->height;
        table[i].bitmap = (uint64_t)(uintptr_t)((char *)sv->pool.ptr + r->pool_at);
        table[i].width = r->width;
        table[i].height = r->height;
        table[i].masks = r->masks ? (uint64_t)(uintptr_t)r->masks->mask_table.ptr : 0;
        table[i].first_draw = 0;
        table[i].ndraws = 0;
        uint32_t base = (uint32_t)payload.size();
        payload.insert(payload.end(), r->packets.begin(), r->packets.end());
        for (draw_desc dd : r->descs) { // packet offsets move with the concatenation
            dd.points_at += base, dd.subpaths_at += base, dd.dash_at += base;
            dd.colors_at += base, dd.stops_at += base;
            dd.canvas = (int32_t)i;
            b->descs.push_back(dd);
        }
        first_draw[i + 1] = b->descs.size();
    }
    std::vector<canvas_desc> frame_canvases;
    std::vector<uint32_t> frame_items;
    size_t draw = 0;
    const size_t max_draws = kFrameDrawLimit;
    const uint64_t shadow_budget = (uint64_t)1 << 30; // shadow floats per frame
    std::vector<unsigned char> used(kCutSlots);
    while (draw < b->descs.size()) {
        size_t last = draw;
        uint64_t shadow = 0;
        int current = -1;
        while (last < b->descs.size() && last - draw < max_draws) {
            const draw_desc &dd = b->descs[last];
            if (frame_cut(dd, dd.canvas, used, current) && last > draw)
                break;
            uint64_t need = 0;
            if (dd.shadow) {
                const server_request *r = b->requests[(size_t)dd.canvas];
                int bo = dd.border;
                need = 2ull * (uint64_t)(r->width + 2 * bo + 1) * (uint64_t)(r->height + 2 * bo + 1);
            }
            if (last > draw && shadow + need > shadow_budget)
                break;
            shadow += need;
            ++last;
        }
        server_batch::frame_rec f;
        f.first_draw = draw;
        f.ndraws = last - draw;
        f.canvases_at = frame_canvases.size();
        int c0 = b->descs[draw].canvas, c1 = b->descs[last - 1].canvas;
        for (int c = c0; c <= c1; ++c) {
            size_t lo = std::max(first_draw[(size_t)c], draw);
            size_t hi = std::min(first_draw[(size_t)c + 1], last);
            if (hi <= lo)
                continue;
            canvas_desc cd = table[(size_t)c];
            cd.first_draw = (int32_t)(lo - draw);
            cd.ndraws = (int32_t)(hi - lo);
            for (size_t q = lo; q < hi; ++q)
                b->descs[q].canvas = (int32_t)(frame_canvases.size() - f.canvases_at);
            frame_canvases.push_back(cd);
        }
        f.ncanvases = frame_canvases.size() - f.canvases_at;
        f.shape = shape_of(b->descs.data() + draw, frame_canvases.data() + f.canvases_at,
                           f.ncanvases);
        f.items_at = frame_items.size();
        frame_items.insert(frame_items.end(), f.shape.items.begin(), f.shape.items.end());
        b->frames.push_back(f);
        draw = last;
    }
    uint64_t total = b->pixel_prefix.back();
    size_t sz_payload = std::max<size_t>(payload.size(), 16);
    size_t sz_draws = sizeof(draw_desc) * std::max<size_t>(b->descs.size(), 1);
    size_t sz_fc = sizeof(canvas_desc) * std::max<size_t>(frame_canvases.size(), 1);
    size_t sz_fi = 4 * std::max<size_t>(frame_items.size(), 1);
    size_t sz_tab = sizeof(canvas_desc) * n, sz_pre = 8 * (n + 1);
    if ((s = b->payload.reserve(sz_payload)) || (s = b->draws.reserve(sz_draws)) ||
        (s = b->frame_canvases.reserve(sz_fc)) || (s = b->frame_items.reserve(sz_fi)) ||
        (s = b->all_canvases.reserve(sz_tab)) || (s = b->prefix.reserve(sz_pre)) ||
        (s = b->srgb.reserve(4 * total)) || (s = b->srgb_host.reserve(4 * total)) ||
        (s = b->counters_host.reserve(8 * COUNTER_COUNT *
                                      std::max<size_t>(b->frames.size(), 1))) ||
        (s = b->staging.reserve(sz_payload + sz_draws + sz_fc + sz_fi + sz_tab + sz_pre)))
        return s;
    unsigned char *st = (unsigned char *)b->staging.ptr;
    size_t o = 0;
    auto put = [&](dbuf &dst, const void *src, size_t bytes) -> hipError_t {
        if (bytes)
            std::memcpy(st + o, src, bytes);
        hipError_t e = bytes ? hipMemcpyAsync(dst.ptr, st + o, bytes, hipMemcpyHostToDevice,
                                              L.stream)
                             : hipSuccess;
        o += bytes;
        return e;
    };
    HIP_TRY(put(b->payload, payload.data(), payload.size()));
    HIP_TRY(put(b->draws, b->descs.data(), sizeof(draw_desc) * b->descs.size()));
    HIP_TRY(put(b->frame_canvases, frame_canvases.data(), sizeof(canvas_desc) * frame_canvases.size()));
    HIP_TRY(put(b->frame_items, frame_items.data(), 4 * frame_items.size()));
    HIP_TRY(put(b->all_canvases, table.data(), sz_tab));
    HIP_TRY(put(b->prefix, b->pixel_prefix.data(), sz_pre));
    // size the lane's arenas for the largest frame (overflow is still checked)
    uint64_t in_points = 0, shadow = 0;
    size_t ndraws = 0;
    for (const server_batch::frame_rec &f : b->frames) {
        in_points = std::max(in_points, f.shape.in_points);
        shadow = std::max(shadow, f.shape.shadow_bound);
        ndraws = std::max(ndraws, f.ndraws);
    }
    if ((s = L.state.reserve(state_bytes(ndraws))) ||
        (s = L.counters.reserve(8 * COUNTER_COUNT)) ||
        (s = ensure_arenas(L, std::max<uint64_t>(64 * in_points, 1 << 22),
                           std::max<uint64_t>(8 * in_points + 64 * ndraws, 1 << 22),
                           std::max<uint64_t>(16 * in_points, 1 << 22), shadow)))
        return s;
    HIP_TRY(launch_clear_batch(b->all_canvases.ptr, b->prefix.ptr, (int)n, total, L.stream));
    for (size_t f = 0; f < b->frames.size(); ++f) {
        const server_batch::frame_rec &fr = b->frames[f];
        if ((s = run_kernels(d, L, (const unsigned char *)b->payload.ptr,
                             (const draw_desc *)b->draws.ptr + fr.first_draw, fr.ndraws,
                             (const canvas_desc *)b->frame_canvases.ptr + fr.canvases_at,
                             fr.ncanvases, (const uint32_t *)b->frame_items.ptr + fr.items_at,
                             fr.shape, false)))
            return s;
        HIP_TRY(hipMemcpyAsync((char *)b->counters_host.ptr + 8 * COUNTER_COUNT * f,
                               L.counters.ptr, 8 * COUNTER_COUNT,
                               hipMemcpyDeviceToHost, L.stream));
    }
    HIP_TRY(launch_readback_batch(b->all_canvases.ptr, b->prefix.ptr, (int)n, total, b->srgb.ptr,
                                  L.stream));
    HIP_TRY(hipMemcpyAsync(b->srgb_host.ptr, b->srgb.ptr, 4 * total, hipMemcpyDeviceToHost,
                           L.stream));
    HIP_TRY(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(b->done, L.stream));
    return CIH_OK;
}

// A finished batch: re-run if an arena overflowed, else deliver the results.
int finish_server_batch(cih_server *sv, int k) {
    server_batch *b = sv->inflight[k];
    exec_lane &L = sv->lanes[k];
    bool overflow = false;
    for (size_t f = 0; f < b->frames.size(); ++f) {
        const unsigned long long *c =
            (const unsigned long long *)((const char *)b->counters_host.ptr +
                                         8 * COUNTER_COUNT * f);
        if (c[COUNTER_OVERFLOW]) {
            overflow = true;
            // grow only the arenas this frame ran out of (grow_after_overflow)
            const server_batch::frame_rec &fr = b->frames[f];
            uint64_t caps[4] = {L.points.bytes / 8, L.words.bytes / 4, run_capacity(L),
                                L.shadow.bytes / 4};
            grow_after_overflow(c, fr.ndraws,
                                fr.ndraws + (fr.shape.any_shadow ? fr.shape.nshadowed : 0), caps);
            int s = ensure_arenas(L, caps[0], caps[1], caps[2], caps[3]);
            if (s)
                return s;
        }
    }
    if (overflow) {
        sv->retries++;
        std::vector<server_request *> requests = b->requests;
        b->release();
        *b = server_batch();
        b->requests = requests;
        return launch_server_batch(sv, k, b);
    }
    double t = now_ms();
    for (size_t i = 0; i < b->requests.size(); ++i) {
        server_request *r = b->requests[i];
        size_t bytes = 4 * (size_t)(b->pixel_prefix[i + 1] - b->pixel_prefix[i]);
        r->result.assign((const unsigned char *)b->srgb_host.ptr + 4 * b->pixel_prefix[i],
                         (const unsigned char *)b->srgb_host.ptr + 4 * b->pixel_prefix[i] + bytes);
        sv->allocator.release(r->pool_at, r->bytes);
        r->t_done = t;
        r->drop_masks();
        std::vector<unsigned char>().swap(r->packets);
        std::vector<draw_desc>().swap(r->descs);
        sv->completed.push_back(r->id);
        sv->served++;
    }
    b->release();
    delete b;
    sv->inflight[k] = nullptr;
    sv->done.notify_all();
    return CIH_OK;
}

// A batch that failed to launch or to finish: its requests complete with the
// error and give back their pool space.  Copies and kernels it already queued
// on the lane are waited for before its buffers are freed.
void abort_server_batch(cih_server *sv, int k, server_batch *b, int status) {
    sv->error = status;
    hipStreamSynchronize(sv->lanes[k].stream);
    (void)hipGetLastError();
    for (server_request *r : b->requests) {
        sv->allocator.release(r->pool_at, r->bytes);
        r->drop_masks();
        r->status = status;
        r->t_done = now_ms();
        sv->completed.push_back(r->id);
    }
    b->release();
    delete b;
    sv->inflight[k] = nullptr;
    sv->done.notify_all();
}

void server_scheduler(cih_server *sv) {
    cih_device *d = sv->device;
    hipSetDevice(d->ordinal);
    std::unique_lock<std::mutex> lock(sv->mu);
    for (;;) {
        bool busy = false;
        for (int k = 0; k < 2; ++k) {
            server_batch *b = sv->inflight[k];
            if (!b)
                continue;
            busy = true;
            hipError_t q = hipEventQuery(b->done);
            if (q == hipErrorNotReady)
                continue;
            int s = q == hipSuccess ? finish_server_batch(sv, k) : CIH_ERROR_DEVICE;
            if (s != CIH_OK)
                abort_server_batch(sv, k, b, s);
        }
        for (int k = 0; k < 2 && !sv->prepared.empty(); ++k) {
            if (sv->inflight[k])
                continue;
            server_batch *b = new server_batch;
            size_t draws = 0;
            uint64_t pixels = 0, bytes = 0;
            while (!sv->prepared.empty() && (int)b->requests.size() < sv->max_batch) {
                server_request *r = sv->prepared.front();
                uint64_t rp = (uint64_t)r->width * (uint64_t)r->height;
                if (!b->requests.empty() &&
                    (draws + r->descs.size() > kBatchDrawLimit || pixels + rp > kBatchPixelLimit ||
                     bytes + r->packets.size() > kPayloadLimit))
                    break;
                if (!sv->allocator.alloc(r->bytes, r->pool_at))
                    break; // wait for memory: admission by HBM footprint
                sv->prepared.pop_front();
                b->requests.push_back(r);
                draws += r->descs.size();
                pixels += rp;
                bytes += r->packets.size();
            }
            if (b->requests.empty()) {
                delete b;
                break;
            }
            sv->inflight[k] = b;
            sv->batches++;
            uint64_t inflight = 0;
            for (int j = 0; j < 2; ++j)
                if (sv->inflight[j])
                    inflight += sv->inflight[j]->requests.size();
            sv->max_inflight_requests = std::max(sv->max_inflight_requests, inflight);
            int s = launch_server_batch(sv, k, b);
            if (s != CIH_OK)
                abort_server_batch(sv, k, b, s);
            busy = true;
        }
        if (sv->stopping && sv->arrived.empty() && sv->prepared.empty() && !sv->inflight[0] &&
            !sv->inflight[1])
            return;
        if (busy)
            sv->work.wait_for(lock, std::chrono::microseconds(50));
        else
            sv->work.wait_for(lock, std::chrono::milliseconds(5));
    }
}

} // namespace

extern "C" {

int cih_server_start(cih_device *d, uint64_t hbm_budget, int max_batch, int workers,
                     cih_server **out) {
    if (!d || !out || max_batch <= 0 || hbm_budget < (1u << 20))
        return fail(CIH_ERROR_INVALID, "invalid server arguments");
    *out = nullptr;
    std::lock_guard<std::recursive_mutex> dlock(d->mutex);
    int s;
    if ((s = set_device(d)))
        return s;
    cih_server *sv = new (std::nothrow) cih_server;
    if (!sv)
        return fail(CIH_ERROR_OUT_OF_MEMORY, "out of host memory");
    sv->device = d;
    sv->budget = hbm_budget;
    sv->max_batch = max_batch;
    if ((s = sv->pool.reserve(hbm_budget))) {
        delete sv;
        return s;
    }
    sv->allocator.reset(hbm_budget);
    for (int k = 0; k < 2; ++k) {
        hipError_t e = hipStreamCreateWithFlags(&sv->lanes[k].stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            sv->pool.release();
            delete sv;
            return hip_fail(e, "server stream");
        }
    }
    if (workers <= 0)
        workers = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    try {
        for (int w = 0; w < workers; ++w)
            sv->workers.emplace_back(server_worker, sv);
        sv->scheduler = std::thread(server_scheduler, sv);
    } catch (const std::system_error &) {
        {
            std::lock_guard<std::mutex> lock(sv->mu);
            sv->stopping = true;
        }
        sv->work.notify_all();
        for (std::thread &t : sv->workers)
            t.join();
        return fail(CIH_ERROR_DEVICE, "could not start server threads");
    }
    *out = sv;
    return CIH_OK;
}

int cih_server_submit(cih_server *sv, const void *stream, size_t bytes, uint64_t *request) {
    if (!sv || !stream || !request)
        return fail(CIH_ERROR_INVALID, "invalid submit arguments");
    cih_reader reader;
    if (cih_reader_init(&reader, stream, bytes) != 0)
        return fail(CIH_ERROR_MALFORMED, "malformed command stream header");
    uint64_t need = 16ull * (uint64_t)reader.width * (uint64_t)reader.height;
    if (need > sv->budget)
        return fail(CIH_ERROR_BUDGET, "request does not fit the server's HBM budget");
    server_request *r = new (std::nothrow) server_request;
    if (!r)
        return fail(CIH_ERROR_OUT_OF_MEMORY, "out of host memory");
    try {
        r->stream.assign((const unsigned char *)stream, (const unsigned char *)stream + bytes);
    } catch (const std::bad_alloc &) {
        delete r;
        return fail(CIH_ERROR_OUT_OF_MEMORY, "out of host memory");
    }
    r->width = reader.width;
    r->height = reader.height;
    r->bytes = (need + 255) & ~(uint64_t)255;
    r->t_submit = now_ms();
    std::lock_guard<std::mutex> lock(sv->mu);
    if (sv->stopping) {
        delete r;
        return fail(CIH_ERROR_INVALID, "server is stopping");
    }
    r->id = sv->next_id++;
    sv->requests[r->id] = r;
    sv->arrived.push_back(r);
    *request = r->id;
    sv->work.notify_all();
    return CIH_OK;
}

int cih_server_wait(cih_server *sv, uint64_t *request, int timeout_ms) {
    if (!sv || !request)
        return fail(CIH_ERROR_INVALID, "invalid wait arguments");
    std::unique_lock<std::mutex> lock(sv->mu);
    if (!sv->done.wait_for(lock, std::chrono::milliseconds(timeout_ms < 0 ? 0 : timeout_ms),
                           [sv] { return !sv->completed.empty(); }))
        return fail(CIH_ERROR_TIMEOUT, "no request completed in time");
    *request = sv->completed.front();
    sv->completed.pop_front();
    return CIH_OK;
}

int cih_server_result(cih_server *sv, uint64_t request, unsigned char *image, int *width,
                      int *height, double *latency_ms) {
    if (!sv)
        return fail(CIH_ERROR_INVALID, "server is NULL");
    server_request *r;
    {
        std::lock_guard<std::mutex> lock(sv->mu);
        auto it = sv->requests.find(request);
        if (it == sv->requests.end() || (it->second->t_done == 0))
            return fail(CIH_ERROR_INVALID, "unknown or unfinished request");
        r = it->second;
        sv->requests.erase(it);
    }
    int status = r->status;
    if (width)
        *width = r->width;
    if (height)
        *height = r->height;
    if (latency_ms)
        *latency_ms = r->t_done - r->t_submit;
    if (image && status == CIH_OK)
        std::memcpy(image, r->result.data(), r->result.size());
    r->drop_masks();
    delete r;
    return status == CIH_OK ? CIH_OK : fail(status, "request failed");
}

int cih_server_stats(cih_server *sv, uint64_t *stats) {
    if (!sv || !stats)
        return fail(CIH_ERROR_INVALID, "NULL argument");
    std::lock_guard<std::mutex> lock(sv->mu);
    stats[0] = sv->served;
    stats[1] = sv->batches;
    stats[2] = sv->max_inflight_requests;
    stats[3] = sv->retries;
    return CIH_OK;
}

int cih_server_stop(cih_server *sv) {
    if (!sv)
        return fail(CIH_ERROR_INVALID, "server is NULL");
    {
        std::lock_guard<std::mutex> lock(sv->mu);
        sv->stopping = true;
    }
    sv->work.notify_all();
    for (std::thread &t : sv->workers)
        t.join();
    if (sv->scheduler.joinable())
        sv->scheduler.join();
    for (int k = 0; k < 2; ++k) {
        hipStreamSynchronize(sv->lanes[k].stream);
        sv->lanes[k].release();
        hipStreamDestroy(sv->lanes[k].stream);
    }
    for (auto &kv : sv->requests) {
        kv.second->drop_masks();
        delete kv.second;
    }
    sv->pool.release();
    delete sv;
    return CIH_OK;
}

} // extern "C"
