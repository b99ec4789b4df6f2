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
#include "cih_stream.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <system_error>
#include <thread>
#include <string>
#include <unordered_map>
#include <vector>

namespace cih {
int fine_segment();
hipError_t launch_readback_batch(const void *canvases, const void *pixel_prefix, int ncanvases,
                                 uint64_t total, void *out, hipStream_t stream);
hipError_t launch_clear_batch(const void *canvases, const void *pixel_prefix, int ncanvases,
                              uint64_t total, hipStream_t stream);
const int kFrameKernels = 7; // reset, geometry, runs, shadow raster, blur rows/cols, fine
extern const char *const kFrameKernelNames[kFrameKernels];
hipError_t launch_frame(const frame &F, int shadow_rows, int shadow_cols, bool any_shadow,
                        hipEvent_t *events, hipStream_t stream);
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

struct cih_device {
    int ordinal = 0;
    hipStream_t stream = nullptr;
    std::recursive_mutex mutex;
    // arenas
    dbuf points, words, runs, shadow, counters;
    // frame inputs
    dbuf state, canvases, items;
    hbuf staging, host_counters;
    // images
    std::vector<image_entry *> images;
    std::unordered_multimap<uint64_t, int32_t> image_index;
    dbuf image_table, lut;
    dbuf upload;
    // read-back
    dbuf readback;
    hbuf readback_host;
    size_t frames = 0, retries = 0;
    size_t last_ndraws = 0; // draws of the last frame (for the trace seams)
    // kernel timing with HIP events on the device's stream
    bool profiling = false;
    bool capturing = false; // a HIP graph capture is recording on the stream
    std::vector<hipEvent_t> event_pool;
    size_t events_used = 0;
};

// A canvas' decoded draw list, resident in HBM so it can be run again.
struct program {
    dbuf payload, draws;
    std::vector<draw_desc> descs; // host copy
    struct span {
        size_t first, count;
    };
    std::vector<span> frames;
    size_t bytes = 0;
};

struct cih_canvas {
    cih_device *device;
    int width, height;
    void *pixels = nullptr; // float4 bitmap in HBM (owned by `bitmap` or by a batch)
    dbuf bitmap;
    std::vector<unsigned char> pending;
    program program;
};

namespace {

int set_device(cih_device *d) {
    HIP_TRY(hipSetDevice(d->ordinal));
    return CIH_OK;
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
    if (p.kind < CIH_DRAW_FILL_PATH || p.kind > CIH_DRAW_STROKE_LINES)
        return 0;
    // subpath counts must add up to the points; bezier subpaths are 1 + 3k
    const unsigned char *sub = base + at + sizeof(p) + 8 * (size_t)p.points;
    uint64_t total = 0;
    for (uint32_t s = 0; s < p.subpaths; ++s) {
        uint32_t c;
        std::memcpy(&c, sub + 8 * (size_t)s, 4);
        if (c == 0 || ((p.kind == CIH_DRAW_FILL_PATH || p.kind == CIH_DRAW_STROKE_PATH) &&
                       (c - 1) % 3 != 0))
            return 0;
        total += c;
    }
    if (total != p.points)
        return 0;
    if ((p.brush == CIH_BRUSH_LINEAR || p.brush == CIH_BRUSH_RADIAL) && p.stops != p.colors)
        return 0;
    if (p.brush == CIH_BRUSH_PATTERN || p.brush == CIH_BRUSH_IMAGE) {
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
    out.shadow = p.shadow_color[3] != 0.0f ? 1 : 0;
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

int ensure_arenas(cih_device *d, uint64_t pts, uint64_t words, uint64_t runs, uint64_t shadow) {
    int s;
    if ((s = d->points.reserve(8 * pts)) || (s = d->words.reserve(4 * words)) ||
        (s = d->runs.reserve(sizeof(run) * runs)) || (s = d->shadow.reserve(4 * shadow)))
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

// A frame: draws (already in HBM) of one or more canvases, run together.
struct frame_input {
    const unsigned char *payload; // device
    const draw_desc *draws;       // device, n records
    const draw_desc *host;        // host copy of the same records
    size_t n;
    std::vector<canvas_desc> canvases; // first_draw relative to `draws`
};

// What a frame needs from the arenas, and its fine-stage work items.
struct frame_shape {
    bool any_shadow = false;
    int max_rows = 0, max_cols = 0;
    uint64_t shadow_bound = 0, in_points = 0;
    size_t ndraws = 0;
    std::vector<uint32_t> items; // per-canvas prefix of (row, segment) items
};

frame_shape shape_of(const draw_desc *host, const canvas_desc *canvases, size_t ncanvases) {
    frame_shape f;
    int seg = fine_segment();
    f.items.assign(ncanvases + 1, 0);
    for (size_t c = 0; c < ncanvases; ++c) {
        const canvas_desc &cd = canvases[c];
        uint64_t segs = (uint64_t)((cd.width + seg - 1) / seg);
        f.items[c + 1] = f.items[c] + (uint32_t)(segs * (uint64_t)cd.height);
        for (int k = cd.first_draw; k < cd.first_draw + cd.ndraws; ++k) {
            const draw_desc &dd = host[k];
            f.in_points += dd.npoints;
            f.ndraws++;
            if (dd.shadow) {
                f.any_shadow = true;
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
// already in HBM.  With `check`, waits for it and, if an arena ran out, grows
// the arenas to what the frame asked for and runs it again (kernels write no
// pixel once an arena has overflowed).  Without it nothing synchronises or
// allocates (arenas must already fit: graph capture relies on that).
int run_kernels(cih_device *d, const unsigned char *payload, const draw_desc *draws, size_t n,
                const canvas_desc *canvases_dev, size_t ncanvases, const uint32_t *items_dev,
                const frame_shape &shape, bool check) {
    int s;
    if (check) {
        if ((s = d->state.reserve(sizeof(draw_state) * std::max<size_t>(n, 1))) ||
            (s = d->counters.reserve(64)) || (s = d->host_counters.reserve(64)))
            return s;
        uint64_t want_points = std::max<uint64_t>(1 << 22, 64 * shape.in_points);
        uint64_t want_words = std::max<uint64_t>(1 << 22, 8 * shape.in_points + 64 * n);
        uint64_t want_runs = std::max<uint64_t>(1 << 22, 16 * shape.in_points);
        if ((s = ensure_arenas(d, std::max<uint64_t>(want_points, d->points.bytes / 8),
                               std::max<uint64_t>(want_words, d->words.bytes / 4),
                               std::max<uint64_t>(want_runs, d->runs.bytes / sizeof(run)),
                               std::max<uint64_t>(shape.shadow_bound, d->shadow.bytes / 4))))
            return s;
    }
    d->last_ndraws = n;
    for (int attempt = 0; attempt < 8; ++attempt) {
        frame F;
        std::memset(&F, 0, sizeof(F));
        F.payload = payload;
        F.draws = draws;
        F.state = (draw_state *)d->state.ptr;
        F.canvases = canvases_dev;
        F.images = (const image_desc *)d->image_table.ptr;
        F.ndraws = (int32_t)n;
        F.ncanvases = (int32_t)ncanvases;
        F.points = (float *)d->points.ptr;
        F.points_cap = d->points.bytes / 8;
        F.words = (uint32_t *)d->words.ptr;
        F.words_cap = d->words.bytes / 4;
        F.runs = (run *)d->runs.ptr;
        F.runs_cap = d->runs.bytes / sizeof(run);
        F.shadow = (float *)d->shadow.ptr;
        F.shadow_cap = d->shadow.bytes / 4;
        F.counters = (unsigned long long *)d->counters.ptr;
        F.items = items_dev;
        F.nitems = shape.items.back();
        F.segment = fine_segment();
        hipEvent_t *events = nullptr;
        if (d->profiling && !d->capturing) {
            if ((s = profile_events(d, &events)))
                return s;
        }
        HIP_TRY(launch_frame(F, shape.max_rows, shape.max_cols, shape.any_shadow, events,
                             d->stream));
        d->frames++;
        if (!check)
            return CIH_OK;
        HIP_TRY(hipMemcpyAsync(d->host_counters.ptr, d->counters.ptr, 8 * COUNTER_COUNT,
                               hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
        const unsigned long long *hc = (const unsigned long long *)d->host_counters.ptr;
        if (!hc[COUNTER_OVERFLOW])
            return CIH_OK;
        d->retries++;
        uint64_t np = std::max<uint64_t>(hc[COUNTER_POINTS], F.points_cap);
        uint64_t nw = std::max<uint64_t>(hc[COUNTER_WORDS], F.words_cap);
        uint64_t nr = std::max<uint64_t>(hc[COUNTER_RUNS], F.runs_cap);
        uint64_t ns = std::max<uint64_t>(hc[COUNTER_SHADOW], F.shadow_cap);
        if ((s = ensure_arenas(d, hc[COUNTER_POINTS] >= F.points_cap ? 2 * np : np,
                               hc[COUNTER_WORDS] >= F.words_cap ? 2 * nw : nw,
                               hc[COUNTER_RUNS] >= F.runs_cap ? 2 * nr : nr,
                               hc[COUNTER_SHADOW] >= F.shadow_cap ? 2 * ns : ns)))
            return s;
    }
    return fail(CIH_ERROR_OUT_OF_MEMORY, "frame did not fit after repeated arena growth");
}

// Launches one frame of a single canvas (uploads its small canvas list).
int launch(cih_device *d, const frame_input &in, bool check) {
    int s;
    frame_shape shape = shape_of(in.host, in.canvases.data(), in.canvases.size());
    if ((s = d->canvases.reserve(sizeof(canvas_desc) * in.canvases.size())) ||
        (s = d->items.reserve(4 * shape.items.size())))
        return s;
    HIP_TRY(hipMemcpyAsync(d->canvases.ptr, in.canvases.data(),
                           sizeof(canvas_desc) * in.canvases.size(), hipMemcpyHostToDevice,
                           d->stream));
    HIP_TRY(hipMemcpyAsync(d->items.ptr, shape.items.data(), 4 * shape.items.size(),
                           hipMemcpyHostToDevice, d->stream));
    return run_kernels(d, in.payload, in.draws, in.n, (const canvas_desc *)d->canvases.ptr,
                       in.canvases.size(), (const uint32_t *)d->items.ptr, shape, check);
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
    size_t at = 0;
    while (at < bytes.size()) {
        draw_desc desc;
        size_t used = decode_packet(bytes.data(), at, bytes.size(), desc, d);
        if (!used)
            return fail(CIH_ERROR_MALFORMED, "malformed draw packet at byte " +
                                                 std::to_string(at));
        desc.canvas = 0;
        P.descs.push_back(desc);
        at += used;
    }
    if ((s = P.payload.reserve(std::max<size_t>(bytes.size(), 16))) ||
        (s = P.draws.reserve(sizeof(draw_desc) * std::max<size_t>(P.descs.size(), 1))) ||
        (s = d->staging.reserve(bytes.size() + sizeof(draw_desc) * P.descs.size())))
        return s;
    // one pinned staging copy, two transfers
    unsigned char *stage = (unsigned char *)d->staging.ptr;
    std::memcpy(stage, bytes.data(), bytes.size());
    std::memcpy(stage + bytes.size(), P.descs.data(), sizeof(draw_desc) * P.descs.size());
    HIP_TRY(hipMemcpyAsync(P.payload.ptr, stage, bytes.size(), hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipMemcpyAsync(P.draws.ptr, stage + bytes.size(), sizeof(draw_desc) * P.descs.size(),
                           hipMemcpyHostToDevice, d->stream));
    // Frames of bounded size: shadow images are the largest intermediates.
    const uint64_t shadow_budget = (uint64_t)1 << 30; // floats
    size_t first = 0;
    while (first < P.descs.size()) {
        size_t last = first;
        uint64_t shadow = 0;
        while (last < P.descs.size() && last - first < 8192) {
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
        program::span sp = {first, last - first};
        P.frames.push_back(sp);
        first = last;
    }
    P.bytes = bytes.size();
    return CIH_OK;
}

// Runs the canvas' program (all its frames) in order.
int run_program(cih_canvas *cv, bool check) {
    program &P = cv->program;
    int s;
    for (const program::span &sp : P.frames) {
        frame_input in;
        in.payload = (const unsigned char *)P.payload.ptr;
        in.draws = (const draw_desc *)P.draws.ptr + sp.first;
        in.host = P.descs.data() + sp.first;
        in.n = sp.count;
        canvas_desc cd;
        cd.bitmap = (uint64_t)(uintptr_t)cv->pixels;
        cd.width = cv->width;
        cd.height = cv->height;
        cd.first_draw = 0;
        cd.ndraws = (int32_t)sp.count;
        in.canvases.push_back(cd);
        if ((s = launch(cv->device, in, check)))
            return s;
    }
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
    float table[256];
    for (int v = 0; v < 256; ++v)
        table[v] = linearized(v / 255.0f);
    if ((s = d->lut.reserve(sizeof(table))) ||
        (e = hipMemcpy(d->lut.ptr, table, sizeof(table), hipMemcpyHostToDevice)) != hipSuccess) {
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
    dbuf *all[] = {&d->points, &d->words, &d->runs, &d->shadow, &d->counters, &d->state,
                   &d->canvases, &d->items, &d->image_table, &d->lut, &d->upload, &d->readback};
    for (hipEvent_t e : d->event_pool)
        hipEventDestroy(e);
    for (dbuf *b : all)
        b->release();
    d->staging.release();
    d->host_counters.release();
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
    e->rgba.swap(dense);
    int32_t id = (int32_t)d->images.size();
    d->images.push_back(e);
    d->image_index.emplace(h, id);
    std::vector<image_desc> table(d->images.size());
    for (size_t k = 0; k < table.size(); ++k) {
        table[k].texels = (uint64_t)(uintptr_t)d->images[k]->texels.ptr;
        table[k].width = d->images[k]->width;
        table[k].height = d->images[k]->height;
    }
    if ((s = d->image_table.reserve(sizeof(image_desc) * table.size())))
        return s;
    HIP_TRY(hipMemcpy(d->image_table.ptr, table.data(), sizeof(image_desc) * table.size(),
                      hipMemcpyHostToDevice));
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
    for (const program::span &sp : c->program.frames) {
        frame_input in;
        in.payload = (const unsigned char *)c->program.payload.ptr;
        in.draws = (const draw_desc *)c->program.draws.ptr + sp.first;
        in.host = c->program.descs.data() + sp.first;
        in.n = sp.count;
        canvas_desc cd;
        cd.bitmap = (uint64_t)(uintptr_t)c->pixels;
        cd.width = c->width;
        cd.height = c->height;
        cd.first_draw = 0;
        cd.ndraws = (int32_t)sp.count;
        in.canvases.push_back(cd);
        if ((s = launch(d, in, true)))
            return s;
        const unsigned long long *hc = (const unsigned long long *)d->host_counters.ptr;
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
    if (!c || draw < 0 || !npoints || !nsubpaths)
        return fail(CIH_ERROR_INVALID, "invalid trace arguments");
    cih_device *d = c->device;
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    if ((size_t)draw >= d->last_ndraws)
        return fail(CIH_ERROR_INVALID, "no such draw in the last frame");
    draw_state S;
    HIP_TRY(hipMemcpy(&S, (draw_state *)d->state.ptr + draw, sizeof(S), hipMemcpyDeviceToHost));
    *npoints = S.poly_npoints;
    *nsubpaths = S.poly_nsub;
    if (points && point_capacity)
        HIP_TRY(hipMemcpy(points, (const void *)(uintptr_t)S.poly_points,
                          8 * std::min<size_t>(point_capacity, S.poly_npoints),
                          hipMemcpyDeviceToHost));
    if (subpaths && subpath_capacity)
        HIP_TRY(hipMemcpy(subpaths, (const void *)(uintptr_t)S.poly_counts,
                          8 * std::min<size_t>(subpath_capacity, S.poly_nsub),
                          hipMemcpyDeviceToHost));
    return CIH_OK;
}

int cih_trace_runs(cih_canvas *c, int draw, int pass, int32_t *xs, int32_t *ys, float *deltas,
                   size_t capacity, size_t *nruns) {
    if (!c || draw < 0 || !nruns)
        return fail(CIH_ERROR_INVALID, "invalid trace arguments");
    cih_device *d = c->device;
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
    if ((size_t)draw >= d->last_ndraws)
        return fail(CIH_ERROR_INVALID, "no such draw in the last frame");
    draw_state S;
    HIP_TRY(hipMemcpy(&S, (draw_state *)d->state.ptr + draw, sizeof(S), hipMemcpyDeviceToHost));
    int row0 = pass ? S.srow0 : S.row0, rows = pass ? S.srows : S.rows;
    uint64_t rowoff = pass ? S.srowoff : S.rowoff, base = pass ? S.sruns : S.runs;
    *nruns = 0;
    if (rows <= 0)
        return CIH_OK;
    std::vector<uint32_t> off((size_t)rows + 1);
    HIP_TRY(hipMemcpy(off.data(), (uint32_t *)d->words.ptr + rowoff, 4 * off.size(),
                      hipMemcpyDeviceToHost));
    size_t total = off.back();
    *nruns = total;
    std::vector<run> r(total);
    if (total)
        HIP_TRY(hipMemcpy(r.data(), (run *)d->runs.ptr + base, sizeof(run) * total,
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
        frame_shape shape;
    };
    std::vector<frame_rec> frames;
    dbuf frame_canvases, frame_items;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool ran = false;
    bool use_graph = true;
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

void destroy_batch(cih_batch *b) {
    if (b->exec)
        hipGraphExecDestroy(b->exec);
    if (b->graph)
        hipGraphDestroy(b->graph);
    for (cih_canvas *v : b->views)
        delete v;
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
        if ((s = run_kernels(d, (const unsigned char *)b->payload.ptr,
                             (const draw_desc *)b->draws.ptr + f.first_draw, f.ndraws,
                             (const canvas_desc *)b->frame_canvases.ptr + f.canvases_at,
                             f.ncanvases, (const uint32_t *)b->frame_items.ptr + f.items_at,
                             f.shape, check)))
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
    std::lock_guard<std::recursive_mutex> lock(d->mutex);
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
            v->width = r.width;
            v->height = r.height;
            b->views.push_back(v);
            b->pixel_prefix[(size_t)i + 1] =
                b->pixel_prefix[(size_t)i] + (uint64_t)r.width * (uint64_t)r.height;
        }
        uint64_t total = b->pixel_prefix.back();
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
        // host state machine: replay every stream into packets, in parallel
        unsigned workers = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        workers = std::min<unsigned>(workers, (unsigned)count);
        std::vector<int> status(workers, CIH_OK);
        std::vector<std::thread> pool;
        size_t per = ((size_t)count + workers - 1) / workers;
        for (unsigned w = 0; w < workers; ++w) {
            size_t first = w * per, last = std::min((size_t)count, first + per);
            pool.emplace_back(replay_range, b, streams, bytes, first, last, &status[w]);
        }
        for (std::thread &t : pool)
            t.join();
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
            payload.insert(payload.end(), v->pending.begin(), v->pending.end());
            size_t at = base;
            while (at < payload.size()) {
                draw_desc desc;
                size_t used = decode_packet(payload.data(), at, payload.size(), desc, d);
                if (!used) {
                    destroy_batch(b);
                    return fail(CIH_ERROR_MALFORMED, "malformed draw packet in batch");
                }
                desc.canvas = i;
                b->descs.push_back(desc);
                at += used;
            }
            std::vector<unsigned char>().swap(v->pending);
            first_draw[(size_t)i + 1] = b->descs.size();
        }
        // frames: consecutive draw ranges of bounded size; canvas lists per frame
        const size_t max_draws = 32768;
        const uint64_t shadow_budget = (uint64_t)1 << 30;
        std::vector<canvas_desc> frame_canvases;
        std::vector<uint32_t> frame_items;
        size_t draw = 0;
        while (draw < b->descs.size()) {
            size_t last = draw;
            uint64_t shadow = 0;
            while (last < b->descs.size() && last - draw < max_draws) {
                const draw_desc &dd = b->descs[last];
                uint64_t need = 0;
                if (dd.shadow) {
                    const cih_canvas *v = b->views[(size_t)dd.canvas];
                    int bo = dd.border;
                    need = 2ull * (uint64_t)(v->width + 2 * bo + 1) * (uint64_t)(v->height + 2 * bo + 1);
                }
                if (last > draw && shadow + need > shadow_budget)
                    break;
                shadow += need;
                ++last;
            }
            cih_batch::frame_rec f;
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
        hipError_t e;
        if ((e = hipMemcpy(b->payload.ptr, payload.data(), payload.size(),
                           hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(b->draws.ptr, b->descs.data(), sizeof(draw_desc) * b->descs.size(),
                           hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(b->frame_canvases.ptr, frame_canvases.data(),
                           sizeof(canvas_desc) * frame_canvases.size(),
                           hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(b->frame_items.ptr, frame_items.data(), 4 * frame_items.size(),
                           hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(b->all_canvases.ptr, table.data(), sizeof(canvas_desc) * table.size(),
                           hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(b->prefix_dev.ptr, b->pixel_prefix.data(),
                           8 * b->pixel_prefix.size(), hipMemcpyHostToDevice)) != hipSuccess) {
            destroy_batch(b);
            return hip_fail(e, "batch upload");
        }
    } catch (const std::bad_alloc &) {
        destroy_batch(b);
        return fail(CIH_ERROR_OUT_OF_MEMORY, "out of host memory");
    } catch (const std::system_error &) {
        destroy_batch(b);
        return fail(CIH_ERROR_DEVICE, "could not start replay threads");
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
                b->exec = nullptr;
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
