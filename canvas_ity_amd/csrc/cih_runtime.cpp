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

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

namespace cih {
int fine_segment();
hipError_t launch_frame(const frame &F, int shadow_rows, int shadow_cols, bool any_shadow,
                        hipStream_t stream);
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
    dbuf payload, draws, state, canvases, items;
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
};

struct cih_canvas {
    cih_device *device;
    int width, height;
    dbuf bitmap;
    std::vector<unsigned char> pending;
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

// One canvas' slice of a frame.
struct frame_canvas {
    cih_canvas *canvas;
    size_t first, count; // draws in the frame's desc list
};

// Runs one frame (already decoded draws of one or more canvases, payload
// uploaded) to completion, growing arenas and retrying on overflow.
int run_frame(cih_device *d, std::vector<draw_desc> &descs,
              const std::vector<frame_canvas> &slices, bool sync) {
    int s;
    size_t n = descs.size();
    std::vector<canvas_desc> canv(slices.size());
    std::vector<uint32_t> items(slices.size() + 1, 0);
    int seg = fine_segment();
    int max_rows = 0, max_cols = 0;
    bool any_shadow = false;
    uint64_t shadow_bound = 0;
    uint64_t in_points = 0;
    for (size_t c = 0; c < slices.size(); ++c) {
        cih_canvas *cv = slices[c].canvas;
        canv[c].bitmap = (uint64_t)(uintptr_t)cv->bitmap.ptr;
        canv[c].width = cv->width;
        canv[c].height = cv->height;
        canv[c].first_draw = (int32_t)slices[c].first;
        canv[c].ndraws = (int32_t)slices[c].count;
        uint64_t segs = (uint64_t)((cv->width + seg - 1) / seg);
        items[c + 1] = items[c] + (uint32_t)(segs * (uint64_t)cv->height);
        for (size_t k = slices[c].first; k < slices[c].first + slices[c].count; ++k) {
            descs[k].canvas = (int32_t)c;
            in_points += descs[k].npoints;
            if (descs[k].shadow) {
                any_shadow = true;
                int b = descs[k].border;
                max_rows = std::max(max_rows, cv->height + 2 * b + 1);
                max_cols = std::max(max_cols, cv->width + 2 * b + 1);
                shadow_bound += 2ull * (uint64_t)(cv->width + 2 * b + 1) *
                                (uint64_t)(cv->height + 2 * b + 1);
            }
        }
    }
    if ((s = d->draws.reserve(sizeof(draw_desc) * std::max<size_t>(n, 1))) ||
        (s = d->state.reserve(sizeof(draw_state) * std::max<size_t>(n, 1))) ||
        (s = d->canvases.reserve(sizeof(canvas_desc) * canv.size())) ||
        (s = d->items.reserve(4 * items.size())) || (s = d->counters.reserve(64)) ||
        (s = d->host_counters.reserve(64)))
        return s;
    // first-guess arena sizes; overflow grows them to what the frame asked for
    uint64_t want_points = std::max<uint64_t>(1 << 22, 64 * in_points);
    uint64_t want_words = std::max<uint64_t>(1 << 22, 8 * in_points + 64 * n);
    uint64_t want_runs = std::max<uint64_t>(1 << 22, 16 * in_points);
    if ((s = ensure_arenas(d, std::max<uint64_t>(want_points, d->points.bytes / 8),
                           std::max<uint64_t>(want_words, d->words.bytes / 4),
                           std::max<uint64_t>(want_runs, d->runs.bytes / sizeof(run)),
                           std::max<uint64_t>(shadow_bound, d->shadow.bytes / 4))))
        return s;
    HIP_TRY(hipMemcpyAsync(d->draws.ptr, descs.data(), sizeof(draw_desc) * n,
                           hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipMemcpyAsync(d->canvases.ptr, canv.data(), sizeof(canvas_desc) * canv.size(),
                           hipMemcpyHostToDevice, d->stream));
    HIP_TRY(hipMemcpyAsync(d->items.ptr, items.data(), 4 * items.size(), hipMemcpyHostToDevice,
                           d->stream));
    d->last_ndraws = n;
    for (int attempt = 0; attempt < 8; ++attempt) {
        frame F;
        std::memset(&F, 0, sizeof(F));
        F.payload = (const unsigned char *)d->payload.ptr;
        F.draws = (const draw_desc *)d->draws.ptr;
        F.state = (draw_state *)d->state.ptr;
        F.canvases = (const canvas_desc *)d->canvases.ptr;
        F.images = (const image_desc *)d->image_table.ptr;
        F.ndraws = (int32_t)n;
        F.ncanvases = (int32_t)canv.size();
        F.points = (float *)d->points.ptr;
        F.points_cap = d->points.bytes / 8;
        F.words = (uint32_t *)d->words.ptr;
        F.words_cap = d->words.bytes / 4;
        F.runs = (run *)d->runs.ptr;
        F.runs_cap = d->runs.bytes / sizeof(run);
        F.shadow = (float *)d->shadow.ptr;
        F.shadow_cap = d->shadow.bytes / 4;
        F.counters = (unsigned long long *)d->counters.ptr;
        F.items = (const uint32_t *)d->items.ptr;
        F.nitems = items.back();
        F.segment = seg;
        HIP_TRY(launch_frame(F, max_rows, max_cols, any_shadow, d->stream));
        d->frames++;
        HIP_TRY(hipMemcpyAsync(d->host_counters.ptr, d->counters.ptr,
                               8 * COUNTER_COUNT, hipMemcpyDeviceToHost, d->stream));
        HIP_TRY(hipStreamSynchronize(d->stream));
        const unsigned long long *hc = (const unsigned long long *)d->host_counters.ptr;
        if (!hc[COUNTER_OVERFLOW])
            return CIH_OK;
        d->retries++;
        uint64_t np = std::max<uint64_t>(hc[COUNTER_POINTS], F.points_cap);
        uint64_t nw = std::max<uint64_t>(hc[COUNTER_WORDS], F.words_cap);
        uint64_t nr = std::max<uint64_t>(hc[COUNTER_RUNS], F.runs_cap);
        uint64_t ns = std::max<uint64_t>(hc[COUNTER_SHADOW], F.shadow_cap);
        // whichever arena ran out asked for more than it has; grow all that did
        if ((s = ensure_arenas(d, hc[COUNTER_POINTS] >= F.points_cap ? 2 * np : np,
                               hc[COUNTER_WORDS] >= F.words_cap ? 2 * nw : nw,
                               hc[COUNTER_RUNS] >= F.runs_cap ? 2 * nr : nr,
                               hc[COUNTER_SHADOW] >= F.shadow_cap ? 2 * ns : ns)))
            return s;
    }
    return fail(CIH_ERROR_OUT_OF_MEMORY, "frame did not fit after repeated arena growth");
}

// Decodes and runs all pending packets of one canvas.
int flush_canvas(cih_canvas *cv) {
    cih_device *d = cv->device;
    if (cv->pending.empty())
        return CIH_OK;
    int s;
    std::vector<unsigned char> bytes;
    bytes.swap(cv->pending);
    std::vector<draw_desc> descs;
    size_t at = 0;
    while (at < bytes.size()) {
        draw_desc desc;
        size_t used = decode_packet(bytes.data(), at, bytes.size(), desc, d);
        if (!used)
            return fail(CIH_ERROR_MALFORMED, "malformed draw packet at byte " +
                                                 std::to_string(at));
        descs.push_back(desc);
        at += used;
    }
    if ((s = d->payload.reserve(bytes.size())) || (s = d->staging.reserve(bytes.size())))
        return s;
    std::memcpy(d->staging.ptr, bytes.data(), bytes.size());
    HIP_TRY(hipMemcpyAsync(d->payload.ptr, d->staging.ptr, bytes.size(), hipMemcpyHostToDevice,
                           d->stream));
    // Frames of bounded size: shadow images are the largest intermediates.
    const uint64_t shadow_budget = (uint64_t)1 << 30; // floats
    size_t first = 0;
    while (first < descs.size()) {
        size_t last = first;
        uint64_t shadow = 0;
        while (last < descs.size() && last - first < 8192) {
            uint64_t need = 0;
            if (descs[last].shadow) {
                int b = descs[last].border;
                need = 2ull * (uint64_t)(cv->width + 2 * b + 1) * (uint64_t)(cv->height + 2 * b + 1);
            }
            if (last > first && shadow + need > shadow_budget)
                break;
            shadow += need;
            ++last;
        }
        std::vector<draw_desc> chunk(descs.begin() + (ptrdiff_t)first,
                                     descs.begin() + (ptrdiff_t)last);
        std::vector<frame_canvas> slices(1);
        slices[0].canvas = cv;
        slices[0].first = 0;
        slices[0].count = chunk.size();
        if ((s = run_frame(d, chunk, slices, true)))
            return s;
        first = last;
    }
    return CIH_OK;
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
    dbuf *all[] = {&d->points, &d->words, &d->runs, &d->shadow, &d->counters, &d->payload,
                   &d->draws, &d->state, &d->canvases, &d->items, &d->image_table, &d->lut,
                   &d->upload, &d->readback};
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
    hipError_t e = launch_clear(c->bitmap.ptr, n, d->stream);
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
    HIP_TRY(launch_readback(c->bitmap.ptr, c->width, c->height, d->readback.ptr, width, height,
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
    HIP_TRY(launch_put_image(d->upload.ptr, width, height, c->bitmap.ptr, c->width, c->height,
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
    HIP_TRY(hipMemcpyAsync(rgba, c->bitmap.ptr, 16 * (size_t)c->width * c->height,
                           hipMemcpyDeviceToHost, d->stream));
    HIP_TRY(hipStreamSynchronize(d->stream));
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
