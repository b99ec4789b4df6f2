// canvas_ity_hip.hpp -- the canvas_ity canvas API on an MI355X.
//
// Same public methods, fields and numeric enum values as canvas_ity::canvas
// (canvas_ity.hpp:146-1148), in namespace canvas_ity_hip.  Define
// CANVAS_ITY_HIP_AS_CANVAS_ITY before including to also get the names under
// namespace canvas_ity (a drop-in for code that includes canvas_ity.hpp).
//
// Split of work:
//   host (this header)  transforms, style setters, path building and the
//                       save/restore stack -- the reference's state machine,
//                       with its float arithmetic kept in the same order so
//                       device-space paths are bit-identical (hpp:2657-3042);
//   device (libcanvas_ity_hip.so through cih_abi.h)  flattening, dashing,
//                       stroking, scan conversion, shadows, shading,
//                       compositing, image resampling and read-back.
// Each fill/stroke/*_rectangle/draw_image call becomes a draw packet; packets
// are executed on the GPU when pixels are observed (get_image_data,
// put_image_data) or when the pending list grows large.  Results match the
// reference's immediate mode because nothing else observes pixels.
//
// clip() makes device-side masks that later draws are composited under;
// is_point_in_path is a host-side query answered by the library without the
// device.  Text: the library lays glyph outlines out on the host
// (cih_text_outline) and they are drawn as device-space paths like any fill
// or stroke.  Extension: set_fill_rule.
#ifndef CANVAS_ITY_HIP_HPP
#define CANVAS_ITY_HIP_HPP

#include "cih_abi.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

namespace canvas_ity_hip {

enum composite_operation {
    source_in = 1, source_copy, source_out, destination_in,
    destination_atop = 7, lighter = 10, destination_over, destination_out,
    source_atop, source_over, exclusive_or };
enum cap_style { butt, square, circle };
enum join_style { miter, bevel, rounded };
enum brush_type { fill_style, stroke_style };
enum repetition_style { repeat, repeat_x, repeat_y, no_repeat };
enum align_style { leftward, rightward, center, start = 0, ending };
enum baseline_style { alphabetic, top, middle, bottom, hanging, ideographic = 3 };
enum fill_rule { nonzero, evenodd };

namespace detail {

struct xy { float x, y; };
struct rgba { float r, g, b, a; };
struct affine { float a, b, c, d, e, f; };

inline xy make_xy(float x, float y) { xy v = {x, y}; return v; }
inline xy add(xy l, xy r) { return make_xy(l.x + r.x, l.y + r.y); }
inline xy sub(xy l, xy r) { return make_xy(l.x - r.x, l.y - r.y); }
inline xy mul(float s, xy v) { return make_xy(v.x * s, v.y * s); }
// Matrix-vector product in the reference's association (hpp:1238-1240).
inline xy apply(affine const &m, xy v) {
    return make_xy(m.a * v.x + m.c * v.y + m.e, m.b * v.x + m.d * v.y + m.f);
}
inline float dot(xy l, xy r) { return l.x * r.x + l.y * r.y; }
inline float length(xy v) { return sqrtf(dot(v, v)); }
inline float direction(xy v) { return atan2f(v.y, v.x); }
inline xy normalized(xy v) { return mul(1.0f / std::max(1.0e-6f, length(v)), v); }
inline xy perpendicular(xy v) { return make_xy(-v.y, v.x); }
inline xy lerp(xy from, xy to, float ratio) { return add(from, mul(ratio, sub(to, from))); }

// sRGB <-> linear and alpha helpers (hpp:1273-1297).
inline float linearized(float v) {
    return v < 0.04045f ? v / 12.92f : powf((v + 0.055f) / 1.055f, 2.4f);
}
inline rgba clamped(rgba c) {
    rgba o = {std::min(std::max(c.r, 0.0f), 1.0f), std::min(std::max(c.g, 0.0f), 1.0f),
              std::min(std::max(c.b, 0.0f), 1.0f), std::min(std::max(c.a, 0.0f), 1.0f)};
    return o;
}
inline rgba linearized(rgba c) {
    rgba o = {linearized(c.r), linearized(c.g), linearized(c.b), c.a};
    return o;
}
inline rgba premultiplied(rgba c) {
    rgba o = {c.r * c.a, c.g * c.a, c.b * c.a, c.a};
    return o;
}

struct brush {
    int type; // cih_brush_kind (color, linear, radial, pattern)
    std::vector<rgba> colors;
    std::vector<float> stops;
    xy start, end;
    float start_radius, end_radius;
    int width, height;
    int repetition;
    int32_t image; // device image for patterns
    rgba front;    // first texel of a pattern (colors.front() in hpp:2272)
    brush() : type(CIH_BRUSH_COLOR), start(make_xy(0, 0)), end(make_xy(0, 0)),
              start_radius(0), end_radius(0), width(0), height(0), repetition(0),
              image(-1) { front.r = front.g = front.b = front.a = 0.0f; }
    bool empty() const { return colors.empty(); }
};

struct subpath { size_t count; bool closed; };

// One process-wide device, opened on first use.  Without a device the
// failure is reported once on stderr and canvases draw nothing (read back as
// transparent black), like a canvas nothing was drawn on; the reference never
// aborts.  Define CANVAS_ITY_HIP_ABORT_WITHOUT_DEVICE to abort instead.
inline cih_device *shared_device() {
    static std::mutex guard;
    static cih_device *device = 0;
    static bool reported = false;
    std::lock_guard<std::mutex> lock(guard);
    if (!device && !reported && cih_device_open(0, &device) != CIH_OK) {
        std::fprintf(stderr, "canvas_ity_hip: no HIP device, canvases will draw nothing: %s\n",
                     cih_last_error());
#ifdef CANVAS_ITY_HIP_ABORT_WITHOUT_DEVICE
        std::abort();
#endif
        reported = true;
        device = 0;
    }
    return device;
}

} // namespace detail

class canvas {
public:
    // ---- lifecycle (hpp:194-211, 2607-2655) ----
    canvas(int width, int height)
        : global_composite_operation(source_over), shadow_offset_x(0.0f),
          shadow_offset_y(0.0f), line_cap(butt), line_join(miter),
          line_dash_offset(0.0f), text_align(start), text_baseline(alphabetic),
          size_x(width), size_y(height), global_alpha(1.0f), shadow_blur(0.0f),
          line_width(1.0f), miter_limit(10.0f), rule(nonzero), clip_id(0), face_units(0),
          face_scale(0.0f), device(0), handle(0), owned(true), saves(0) {
        reset_state();
        if (width <= 0 || height <= 0)
            return; // state-only instance (save stack entries)
        device = detail::shared_device();
        if (device && cih_canvas_create(device, width, height, &handle) != CIH_OK) {
            std::fprintf(stderr, "canvas_ity_hip: canvas %dx%d will draw nothing: %s\n", width,
                         height, cih_last_error());
#ifdef CANVAS_ITY_HIP_ABORT_WITHOUT_DEVICE
            std::abort();
#endif
            handle = 0;
        }
    }
    // Extension: a canvas API over an existing device canvas, which stays
    // owned by the caller (used to replay command streams).
    explicit canvas(cih_canvas *borrowed)
        : global_composite_operation(source_over), shadow_offset_x(0.0f),
          shadow_offset_y(0.0f), line_cap(butt), line_join(miter),
          line_dash_offset(0.0f), text_align(start), text_baseline(alphabetic),
          size_x(0), size_y(0), global_alpha(1.0f), shadow_blur(0.0f),
          line_width(1.0f), miter_limit(10.0f), rule(nonzero), clip_id(0), face_units(0),
          face_scale(0.0f), device(0), handle(borrowed), owned(false), saves(0) {
        reset_state();
        if (handle) {
            cih_canvas_size(handle, &size_x, &size_y);
            cih_canvas_device(handle, &device);
        }
    }
    ~canvas() {
        submit();
        if (handle && owned)
            cih_canvas_destroy(handle);
        while (canvas *head = saves) {
            saves = head->saves;
            head->saves = 0;
            delete head;
        }
    }

    // ---- transforms (hpp:2657-2711) ----
    void scale(float x, float y) { transform(x, 0.0f, 0.0f, y, 0.0f, 0.0f); }
    void rotate(float angle) {
        float cosine = cosf(angle);
        float sine = sinf(angle);
        transform(cosine, sine, -sine, cosine, 0.0f, 0.0f);
    }
    void translate(float x, float y) { transform(1.0f, 0.0f, 0.0f, 1.0f, x, y); }
    void transform(float a, float b, float c, float d, float e, float f) {
        set_transform(forward.a * a + forward.c * b, forward.b * a + forward.d * b,
                      forward.a * c + forward.c * d, forward.b * c + forward.d * d,
                      forward.a * e + forward.c * f + forward.e,
                      forward.b * e + forward.d * f + forward.f);
    }
    void set_transform(float a, float b, float c, float d, float e, float f) {
        float determinant = a * d - b * c;
        float scaling = determinant != 0.0f ? 1.0f / determinant : 0.0f;
        detail::affine new_forward = {a, b, c, d, e, f};
        detail::affine new_inverse = {scaling * d, scaling * -b, scaling * -c, scaling * a,
                                      scaling * (c * f - d * e), scaling * (b * e - a * f)};
        forward = new_forward;
        inverse = new_inverse;
    }

    // ---- compositing and shadow state (hpp:309-390, 2713-2736) ----
    void set_global_alpha(float alpha) {
        if (0.0f <= alpha && alpha <= 1.0f)
            global_alpha = alpha;
    }
    composite_operation global_composite_operation;
    void set_shadow_color(float red, float green, float blue, float alpha) {
        detail::rgba c = {red, green, blue, alpha};
        shadow_color = detail::premultiplied(detail::linearized(detail::clamped(c)));
    }
    float shadow_offset_x;
    float shadow_offset_y;
    void set_shadow_blur(float level) {
        if (0.0f <= level)
            shadow_blur = level;
    }

    // ---- line style (hpp:392-470, 2738-2766) ----
    void set_line_width(float width) {
        if (0.0f < width)
            line_width = width;
    }
    cap_style line_cap;
    join_style line_join;
    void set_miter_limit(float limit) {
        if (0.0f < limit)
            miter_limit = limit;
    }
    float line_dash_offset;
    void set_line_dash(float const *segments, int count) {
        for (int index = 0; index < count; ++index)
            if (segments && segments[index] < 0.0f)
                return;
        line_dash.clear();
        if (!segments)
            return;
        for (int index = 0; index < count; ++index)
            line_dash.push_back(segments[index]);
        if (count & 1)
            for (int index = 0; index < count; ++index)
                line_dash.push_back(segments[index]);
    }

    // ---- fill and stroke styles (hpp:478-600, 2768-2864) ----
    void set_color(brush_type type, float red, float green, float blue, float alpha) {
        detail::brush &b = type == fill_style ? fill_brush : stroke_brush;
        b.type = CIH_BRUSH_COLOR;
        b.colors.clear();
        detail::rgba c = {red, green, blue, alpha};
        b.colors.push_back(detail::premultiplied(detail::linearized(detail::clamped(c))));
    }
    void set_linear_gradient(brush_type type, float start_x, float start_y, float end_x,
                             float end_y) {
        detail::brush &b = type == fill_style ? fill_brush : stroke_brush;
        b.type = CIH_BRUSH_LINEAR;
        b.colors.clear();
        b.stops.clear();
        b.start = detail::make_xy(start_x, start_y);
        b.end = detail::make_xy(end_x, end_y);
    }
    void set_radial_gradient(brush_type type, float start_x, float start_y,
                             float start_radius, float end_x, float end_y,
                             float end_radius) {
        if (start_radius < 0.0f || end_radius < 0.0f)
            return;
        detail::brush &b = type == fill_style ? fill_brush : stroke_brush;
        b.type = CIH_BRUSH_RADIAL;
        b.colors.clear();
        b.stops.clear();
        b.start = detail::make_xy(start_x, start_y);
        b.end = detail::make_xy(end_x, end_y);
        b.start_radius = start_radius;
        b.end_radius = end_radius;
    }
    void add_color_stop(brush_type type, float offset, float red, float green, float blue,
                        float alpha) {
        detail::brush &b = type == fill_style ? fill_brush : stroke_brush;
        if ((b.type != CIH_BRUSH_LINEAR && b.type != CIH_BRUSH_RADIAL) || offset < 0.0f ||
            1.0f < offset)
            return;
        // Insert after any equal offsets (std::upper_bound, hpp:2831-2836).
        size_t index = (size_t)(std::upper_bound(b.stops.begin(), b.stops.end(), offset) -
                                b.stops.begin());
        detail::rgba c = {red, green, blue, alpha};
        b.colors.insert(b.colors.begin() + (ptrdiff_t)index,
                        detail::linearized(detail::clamped(c)));
        b.stops.insert(b.stops.begin() + (ptrdiff_t)index, offset);
    }
    void set_pattern(brush_type type, unsigned char const *image, int width, int height,
                     int stride, repetition_style repetition) {
        if (!image || width <= 0 || height <= 0)
            return;
        detail::brush &b = type == fill_style ? fill_brush : stroke_brush;
        load_pattern(b, image, width, height, stride, repetition);
    }

    // ---- path building (hpp:607-780, 2866-3042) ----
    void begin_path() {
        points.clear();
        subpaths.clear();
    }
    void move_to(float x, float y) {
        if (!subpaths.empty() && subpaths.back().count == 1) {
            points.back() = detail::apply(forward, detail::make_xy(x, y));
            return;
        }
        detail::subpath s = {1, false};
        points.push_back(detail::apply(forward, detail::make_xy(x, y)));
        subpaths.push_back(s);
    }
    void close_path() {
        if (subpaths.empty())
            return;
        detail::xy first = points[points.size() - subpaths.back().count];
        detail::affine saved_forward = forward;
        detail::affine identity = {1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f};
        forward = identity;
        line_to(first.x, first.y);
        subpaths.back().closed = true;
        move_to(first.x, first.y);
        forward = saved_forward;
    }
    void line_to(float x, float y) {
        if (subpaths.empty()) {
            move_to(x, y);
            return;
        }
        detail::xy point_1 = points.back();
        detail::xy point_2 = detail::apply(forward, detail::make_xy(x, y));
        detail::xy d = detail::sub(point_2, point_1);
        if (detail::dot(d, d) == 0.0f)
            return;
        points.push_back(point_1);
        points.push_back(point_2);
        points.push_back(point_2);
        subpaths.back().count += 3;
    }
    void quadratic_curve_to(float control_x, float control_y, float x, float y) {
        if (subpaths.empty())
            move_to(control_x, control_y);
        detail::xy point_1 = points.back();
        detail::xy control = detail::apply(forward, detail::make_xy(control_x, control_y));
        detail::xy point_2 = detail::apply(forward, detail::make_xy(x, y));
        points.push_back(detail::lerp(point_1, control, 2.0f / 3.0f));
        points.push_back(detail::lerp(point_2, control, 2.0f / 3.0f));
        points.push_back(point_2);
        subpaths.back().count += 3;
    }
    void bezier_curve_to(float control_1_x, float control_1_y, float control_2_x,
                         float control_2_y, float x, float y) {
        if (subpaths.empty())
            move_to(control_1_x, control_1_y);
        points.push_back(detail::apply(forward, detail::make_xy(control_1_x, control_1_y)));
        points.push_back(detail::apply(forward, detail::make_xy(control_2_x, control_2_y)));
        points.push_back(detail::apply(forward, detail::make_xy(x, y)));
        subpaths.back().count += 3;
    }
    void arc_to(float vertex_x, float vertex_y, float x, float y, float radius) {
        using namespace detail;
        if (radius < 0.0f || forward.a * forward.d - forward.b * forward.c == 0.0f)
            return;
        if (subpaths.empty())
            move_to(vertex_x, vertex_y);
        xy point_1 = apply(inverse, points.back());
        xy vertex = make_xy(vertex_x, vertex_y);
        xy point_2 = make_xy(x, y);
        xy edge_1 = normalized(sub(point_1, vertex));
        xy edge_2 = normalized(sub(point_2, vertex));
        float sine = fabsf(dot(perpendicular(edge_1), edge_2));
        if (sine < 1.0e-4f) {
            line_to(vertex_x, vertex_y);
            return;
        }
        xy offset = mul(radius / sine, add(edge_1, edge_2));
        xy center = add(vertex, offset);
        float angle_1 = direction(sub(mul(dot(offset, edge_1), edge_1), offset));
        float angle_2 = direction(sub(mul(dot(offset, edge_2), edge_2), offset));
        bool reverse = static_cast<int>(floorf((angle_2 - angle_1) / 3.14159265f)) & 1;
        arc(center.x, center.y, radius, angle_1, angle_2, reverse);
    }
    void arc(float x, float y, float radius, float start_angle, float end_angle,
             bool counter_clockwise = false) {
        using namespace detail;
        if (radius < 0.0f)
            return;
        static float const tau = 6.28318531f;
        float winding = counter_clockwise ? -1.0f : 1.0f;
        float from = fmodf(start_angle, tau);
        float span = fmodf(end_angle, tau) - from;
        if ((end_angle - start_angle) * winding >= tau)
            span = tau * winding;
        else if (span * winding < 0.0f)
            span += tau * winding;
        xy centered_1 = mul(radius, make_xy(cosf(from), sinf(from)));
        line_to(x + centered_1.x, y + centered_1.y);
        if (span == 0.0f)
            return;
        int steps = static_cast<int>(std::max(1.0f, roundf(16.0f / tau * span * winding)));
        float segment = span / static_cast<float>(steps);
        float alpha = 4.0f / 3.0f * tanf(0.25f * segment);
        for (int step = 0; step < steps; ++step) {
            float angle = from + static_cast<float>(step + 1) * segment;
            xy centered_2 = mul(radius, make_xy(cosf(angle), sinf(angle)));
            xy point_1 = add(make_xy(x, y), centered_1);
            xy point_2 = add(make_xy(x, y), centered_2);
            xy control_1 = add(point_1, mul(alpha, perpendicular(centered_1)));
            xy control_2 = sub(point_2, mul(alpha, perpendicular(centered_2)));
            bezier_curve_to(control_1.x, control_1.y, control_2.x, control_2.y, point_2.x,
                            point_2.y);
            centered_1 = centered_2;
        }
    }
    void rectangle(float x, float y, float width, float height) {
        move_to(x, y);
        line_to(x + width, y);
        line_to(x + width, y + height);
        line_to(x, y + height);
        close_path();
    }

    // ---- drawing (hpp:793-911, 3044-3203) ----
    void fill() { emit_path(CIH_DRAW_FILL_PATH, fill_brush, points, path_info(), 0); }
    void stroke() { emit_path(CIH_DRAW_STROKE_PATH, stroke_brush, points, path_info(), 0); }
    void set_fill_rule(int new_rule) { rule = new_rule == evenodd ? evenodd : nonzero; }
    // hpp:828, 3057-3099: the current path's fill coverage, multiplied into
    // the current mask, becomes the new mask.  The device computes it (the
    // path is rasterised like a fill); the host names it with the lowest id
    // that no saved state still refers to, so ids are reused.
    void clip() {
        uint32_t id = clip_id;
        for (canvas *s = saves; s; s = s->saves)
            id = std::max(id, s->clip_id);
        ++id;
        std::vector<uint32_t> info(2 * subpaths.size());
        for (size_t k = 0; k < subpaths.size(); ++k) {
            info[2 * k] = (uint32_t)subpaths[k].count;
            info[2 * k + 1] = subpaths[k].closed ? 1u : 0u;
        }
        detail::brush none;
        encode(CIH_DRAW_CLIP, none, CIH_BRUSH_COLOR, points.empty() ? 0 : &points[0],
               points.size(), info.empty() ? 0 : &info[0], subpaths.size(), inverse, forward,
               (int32_t)id);
        clip_id = id;
    }
    // hpp:843, 3101-3132: the current path's fill flattening and crossing
    // test, evaluated on the host by the library (no device work).
    bool is_point_in_path(float x, float y) {
        std::vector<uint32_t> info(2 * subpaths.size());
        for (size_t k = 0; k < subpaths.size(); ++k) {
            info[2 * k] = (uint32_t)subpaths[k].count;
            info[2 * k + 1] = subpaths[k].closed ? 1u : 0u;
        }
        int inside = 0;
        if (cih_path_contains(points.empty() ? 0 : &points[0].x, points.size(),
                              info.empty() ? 0 : &info[0], subpaths.size(), x, y,
                              &inside) != 0)
            return false;
        return inside != 0;
    }

    void clear_rectangle(float x, float y, float width, float height) {
        composite_operation saved_operation = global_composite_operation;
        float saved_global_alpha = global_alpha;
        float saved_alpha = shadow_color.a;
        int saved_type = fill_brush.type;
        global_composite_operation = destination_out;
        global_alpha = 1.0f;
        shadow_color.a = 0.0f;
        fill_brush.type = CIH_BRUSH_COLOR; // paints colors.front() (hpp:2271-2272)
        fill_rectangle(x, y, width, height);
        fill_brush.type = saved_type;
        shadow_color.a = saved_alpha;
        global_alpha = saved_global_alpha;
        global_composite_operation = saved_operation;
    }
    void fill_rectangle(float x, float y, float width, float height) {
        using namespace detail;
        if (width == 0.0f || height == 0.0f)
            return;
        xy quad[4] = {apply(forward, make_xy(x, y)), apply(forward, make_xy(x + width, y)),
                      apply(forward, make_xy(x + width, y + height)),
                      apply(forward, make_xy(x, y + height))};
        emit_lines(CIH_DRAW_FILL_LINES, fill_brush, quad, 4, true, inverse, fill_brush.type,
                   forward);
    }
    void stroke_rectangle(float x, float y, float width, float height) {
        using namespace detail;
        if (width == 0.0f && height == 0.0f)
            return;
        if (width == 0.0f || height == 0.0f) {
            xy line[2] = {apply(forward, make_xy(x, y)),
                          apply(forward, make_xy(x + width, y + height))};
            emit_lines(CIH_DRAW_STROKE_LINES, stroke_brush, line, 2, false, inverse,
                       stroke_brush.type, forward);
        } else {
            xy loop[5] = {apply(forward, make_xy(x, y)), apply(forward, make_xy(x + width, y)),
                          apply(forward, make_xy(x + width, y + height)),
                          apply(forward, make_xy(x, y + height)),
                          apply(forward, make_xy(x, y))};
            emit_lines(CIH_DRAW_STROKE_LINES, stroke_brush, loop, 5, true, inverse,
                       stroke_brush.type, forward);
        }
    }

    // ---- images (hpp:1057-1130, 3313-3408) ----
    void draw_image(unsigned char const *image, int width, int height, int stride, float x,
                    float y, float to_width, float to_height) {
        using namespace detail;
        if (!image || width <= 0 || height <= 0 || to_width == 0.0f || to_height == 0.0f)
            return;
        load_pattern(image_brush, image, width, height, stride, repeat);
        xy quad[4] = {apply(forward, make_xy(x, y)), apply(forward, make_xy(x + to_width, y)),
                      apply(forward, make_xy(x + to_width, y + to_height)),
                      apply(forward, make_xy(x, y + to_height))};
        affine saved_forward = forward;
        affine saved_inverse = inverse;
        translate(x + std::min(0.0f, to_width), y + std::min(0.0f, to_height));
        scale(fabsf(to_width) / static_cast<float>(width),
              fabsf(to_height) / static_cast<float>(height));
        emit_lines(CIH_DRAW_FILL_LINES, image_brush, quad, 4, true, inverse, CIH_BRUSH_IMAGE,
                   forward);
        forward = saved_forward;
        inverse = saved_inverse;
    }
    void get_image_data(unsigned char *image, int width, int height, int stride, int x,
                        int y) {
        if (!image)
            return;
        if (!handle) { // no device canvas: transparent black, as an untouched canvas
            if (width > 0 && height > 0)
                for (int row = 0; row < height; ++row)
                    std::memset(image + (std::ptrdiff_t)row * stride, 0, 4 * (size_t)width);
            return;
        }
        submit();
        cih_canvas_get_image_data(handle, image, width, height, stride, x, y);
    }
    void put_image_data(unsigned char const *image, int width, int height, int stride, int x,
                        int y) {
        if (!image || !handle)
            return;
        submit();
        cih_canvas_put_image_data(handle, image, width, height, stride, x, y);
    }

    // ---- state stack (hpp:1141-1148, 3410-3465) ----
    void save() {
        canvas *state = new canvas(0, 0);
        copy_state(*state, *this);
        state->saves = saves;
        saves = state;
    }
    void restore() {
        if (!saves)
            return;
        canvas *state = saves;
        copy_state(*this, *state);
        saves = state->saves;
        state->saves = 0;
        delete state;
    }

    // ---- text (hpp:924-1029, 3205-3311) ----
    align_style text_align;
    baseline_style text_baseline;
    // The font is copied (whole TrueType file); a null font or zero bytes
    // keeps the current font and only sets the size.  Returns false, with no
    // font left, if the file lacks what text needs.
    bool set_font(unsigned char const *font, int bytes, float size) {
        if (font && bytes) {
            face.reset();
            int units = 0;
            if (bytes < 6 || cih_font_check(font, (size_t)bytes, &units) != CIH_OK)
                return false;
            face = std::make_shared<std::vector<unsigned char> const>(font, font + bytes);
            face_units = units;
        }
        if (!face)
            return false;
        face_scale = size / static_cast<float>(face_units);
        return true;
    }
    void fill_text(char const *text, float x, float y, float maximum_width = 1.0e30f) {
        text_outline(text, x, y, maximum_width, false);
        int saved_rule = rule; // text fills by the non-zero rule
        rule = nonzero;
        emit_path(CIH_DRAW_FILL_PATH, fill_brush, glyph_points, glyph_info,
                  CIH_DRAW_FLAG_SEGMENTS);
        rule = saved_rule;
    }
    void stroke_text(char const *text, float x, float y, float maximum_width = 1.0e30f) {
        text_outline(text, x, y, maximum_width, true);
        emit_path(CIH_DRAW_STROKE_PATH, stroke_brush, glyph_points, glyph_info,
                  CIH_DRAW_FLAG_SEGMENTS);
    }
    float measure_text(char const *text) {
        float width = 0.0f;
        if (!face || !text ||
            cih_text_measure(&(*face)[0], face->size(), face_scale, text, &width) != CIH_OK)
            return 0.0f;
        return width;
    }

    // Extension: the device canvas behind this object (NULL if none).
    cih_canvas *device_canvas() { return handle; }
    // Extension: sends pending draws to the device without reading back.
    void submit() {
        if (handle && !pending.empty()) {
            cih_canvas_submit(handle, &pending[0], pending.size());
            pending.clear();
        }
    }

private:
    int size_x, size_y;
    detail::affine forward, inverse;
    float global_alpha;
    detail::rgba shadow_color;
    float shadow_blur;
    float line_width, miter_limit;
    std::vector<float> line_dash;
    detail::brush fill_brush, stroke_brush, image_brush;
    std::vector<detail::xy> points;
    std::vector<detail::subpath> subpaths;
    int rule;
    uint32_t clip_id; // the current clip mask (0: the whole canvas)
    std::shared_ptr<std::vector<unsigned char> const> face; // TrueType file
    int face_units;   // its units per em
    float face_scale; // font size / units per em
    std::vector<detail::xy> glyph_points; // the last text's outline segments
    std::vector<uint32_t> glyph_info;
    cih_device *device;
    cih_canvas *handle;
    bool owned;
    canvas *saves;
    std::vector<unsigned char> pending;

    canvas(canvas const &);
    canvas &operator=(canvas const &);

    void reset_state() {
        detail::affine identity = {1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f};
        forward = identity;
        inverse = identity;
        shadow_color.r = shadow_color.g = shadow_color.b = shadow_color.a = 0.0f;
        set_color(fill_style, 0.0f, 0.0f, 0.0f, 1.0f);
        set_color(stroke_style, 0.0f, 0.0f, 0.0f, 1.0f);
    }

    static void copy_state(canvas &to, canvas const &from) {
        to.clip_id = from.clip_id;
        to.global_composite_operation = from.global_composite_operation;
        to.shadow_offset_x = from.shadow_offset_x;
        to.shadow_offset_y = from.shadow_offset_y;
        to.line_cap = from.line_cap;
        to.line_join = from.line_join;
        to.line_dash_offset = from.line_dash_offset;
        to.text_align = from.text_align;
        to.text_baseline = from.text_baseline;
        to.forward = from.forward;
        to.inverse = from.inverse;
        to.global_alpha = from.global_alpha;
        to.shadow_color = from.shadow_color;
        to.shadow_blur = from.shadow_blur;
        to.line_width = from.line_width;
        to.miter_limit = from.miter_limit;
        to.line_dash = from.line_dash;
        to.fill_brush = from.fill_brush;
        to.stroke_brush = from.stroke_brush;
        to.rule = from.rule;
        to.face = from.face;
        to.face_units = from.face_units;
        to.face_scale = from.face_scale;
    }

    void load_pattern(detail::brush &b, unsigned char const *image, int width, int height,
                      int stride, int repetition) {
        b.type = CIH_BRUSH_PATTERN;
        b.colors.clear();
        b.stops.clear();
        b.width = width;
        b.height = height;
        b.repetition = repetition;
        detail::rgba first = {image[0] / 255.0f, image[1] / 255.0f, image[2] / 255.0f,
                              image[3] / 255.0f};
        b.front = detail::premultiplied(detail::linearized(first));
        b.colors.push_back(b.front); // colors.front() is the first texel
        b.image = -1;
        if (device)
            cih_image_create(device, image, width, height, stride, &b.image);
    }

    bool singular() const { return forward.a * forward.d - forward.b * forward.c == 0.0f; }

    std::vector<uint32_t> path_info() const {
        std::vector<uint32_t> info(2 * subpaths.size());
        for (size_t k = 0; k < subpaths.size(); ++k) {
            info[2 * k] = (uint32_t)subpaths[k].count;
            info[2 * k + 1] = subpaths[k].closed ? 1u : 0u;
        }
        return info;
    }

    void emit_path(uint32_t kind, detail::brush const &b, std::vector<detail::xy> const &pts,
                   std::vector<uint32_t> const &info, uint32_t flags) {
        if (singular())
            return; // render_main draws nothing (hpp:2554) and stroke_lines bails (hpp:2072)
        encode(kind, b, b.type, pts.empty() ? 0 : &pts[0], pts.size(),
               info.empty() ? 0 : &info[0], info.size() / 2, inverse, forward, 0, flags);
    }

    // text_to_lines (hpp:1793-1846): the outline of `text` as segment-start
    // cubics in glyph_points / glyph_info; empty without a font, text or
    // positive maximum width (the draw still composites, as render_main does).
    void text_outline(char const *text, float x, float y, float maximum_width, bool stroking) {
        glyph_points.clear();
        glyph_info.clear();
        if (!face || !text || maximum_width <= 0.0f)
            return;
        const float fw[6] = {forward.a, forward.b, forward.c, forward.d, forward.e, forward.f};
        size_t np = 0, ns = 0;
        int status = CIH_ERROR_BUDGET;
        for (int attempt = 0; attempt < 2 && status == CIH_ERROR_BUDGET; ++attempt) {
            glyph_points.resize(std::max(np, glyph_points.capacity()));
            glyph_info.resize(std::max(2 * ns, glyph_info.capacity()));
            status = cih_text_outline(
                &(*face)[0], face->size(), face_scale, text, x, y, maximum_width,
                (int)text_align, (int)text_baseline, fw, stroking ? 1 : 0, line_width,
                glyph_points.empty() ? 0 : &glyph_points[0].x, glyph_points.size(),
                glyph_info.empty() ? 0 : &glyph_info[0], glyph_info.size() / 2, &np, &ns);
        }
        if (status != CIH_OK)
            np = ns = 0;
        glyph_points.resize(np);
        glyph_info.resize(2 * ns);
    }

    void emit_lines(uint32_t kind, detail::brush const &b, detail::xy const *pts, size_t n,
                    bool closed, detail::affine const &paint_inverse, int brush_kind,
                    detail::affine const &check) {
        if (check.a * check.d - check.b * check.c == 0.0f)
            return;
        uint32_t info[2] = {(uint32_t)n, closed ? 1u : 0u};
        encode(kind, b, brush_kind, pts, n, info, 1, paint_inverse, forward);
    }

    void encode(uint32_t kind, detail::brush const &b, int brush_kind, detail::xy const *pts,
                size_t npts, uint32_t const *info, size_t nsub,
                detail::affine const &paint_inverse, detail::affine const &fwd,
                int32_t new_mask = 0, uint32_t flags = 0) {
        if (!handle)
            return;
        cih_draw_packet p;
        std::memset(&p, 0, sizeof(p));
        p.kind = kind;
        p.clip_mask = clip_id;
        p.op = (uint32_t)global_composite_operation;
        p.flags = flags | (rule == evenodd && kind == CIH_DRAW_FILL_PATH ? CIH_DRAW_FLAG_EVENODD
                                                                          : 0u);
        p.global_alpha = global_alpha;
        p.line_width = line_width;
        p.miter_limit = miter_limit;
        p.dash_offset = line_dash_offset;
        p.line_cap = line_cap;
        p.line_join = line_join;
        const float fw[6] = {fwd.a, fwd.b, fwd.c, fwd.d, fwd.e, fwd.f};
        const float iv[6] = {inverse.a, inverse.b, inverse.c, inverse.d, inverse.e, inverse.f};
        const float pv[6] = {paint_inverse.a, paint_inverse.b, paint_inverse.c,
                             paint_inverse.d, paint_inverse.e, paint_inverse.f};
        std::memcpy(p.forward, fw, sizeof(fw));
        std::memcpy(p.inverse, iv, sizeof(iv));
        std::memcpy(p.paint_inverse, pv, sizeof(pv));
        // render_shadow's early-out (hpp:2398-2401) is decided here.
        bool shadow = kind != CIH_DRAW_CLIP &&
                      !(shadow_color.a == 0.0f ||
                        (shadow_blur == 0.0f && shadow_offset_x == 0.0f &&
                         shadow_offset_y == 0.0f));
        p.shadow_color[0] = shadow_color.r;
        p.shadow_color[1] = shadow_color.g;
        p.shadow_color[2] = shadow_color.b;
        p.shadow_color[3] = shadow ? shadow_color.a : 0.0f;
        p.shadow_blur = shadow_blur;
        p.shadow_offset_x = shadow_offset_x;
        p.shadow_offset_y = shadow_offset_y;
        // Brush.  A colour-typed paint of a gradient or pattern brush
        // (clear_rectangle) paints colors.front() as stored (hpp:2271-2272).
        std::vector<float> colors, stops;
        p.brush = brush_kind;
        if (brush_kind == CIH_BRUSH_COLOR) {
            if (!b.empty()) {
                detail::rgba c = b.colors.front();
                colors.push_back(c.r), colors.push_back(c.g);
                colors.push_back(c.b), colors.push_back(c.a);
            }
        } else if (brush_kind == CIH_BRUSH_LINEAR || brush_kind == CIH_BRUSH_RADIAL) {
            for (size_t k = 0; k < b.colors.size(); ++k) {
                colors.push_back(b.colors[k].r), colors.push_back(b.colors[k].g);
                colors.push_back(b.colors[k].b), colors.push_back(b.colors[k].a);
            }
            stops = b.stops;
            p.start_x = b.start.x, p.start_y = b.start.y;
            p.end_x = b.end.x, p.end_y = b.end.y;
            p.start_radius = b.start_radius, p.end_radius = b.end_radius;
        } else {
            p.image = b.image;
            p.image_width = b.width;
            p.image_height = b.height;
            p.repetition = b.repetition;
        }
        if (kind == CIH_DRAW_CLIP) {
            colors.clear();
            p.image = new_mask;
        }
        p.points = (uint32_t)npts;
        p.subpaths = (uint32_t)nsub;
        bool strokes = kind == CIH_DRAW_STROKE_PATH || kind == CIH_DRAW_STROKE_LINES;
        p.dashes = strokes ? (uint32_t)line_dash.size() : 0u;
        p.colors = (uint32_t)(colors.size() / 4);
        p.stops = (uint32_t)stops.size();
        size_t body = sizeof(p) + 8 * npts + 8 * nsub + 4 * (size_t)p.dashes +
                      4 * colors.size() + 4 * stops.size();
        size_t padded = (body + 15) & ~(size_t)15;
        p.bytes = (uint32_t)padded;
        size_t at = pending.size();
        pending.resize(at + padded, 0);
        unsigned char *out = &pending[at];
        std::memcpy(out, &p, sizeof(p));
        out += sizeof(p);
        if (npts)
            std::memcpy(out, pts, 8 * npts), out += 8 * npts;
        if (nsub)
            std::memcpy(out, info, 8 * nsub), out += 8 * nsub;
        if (p.dashes)
            std::memcpy(out, &line_dash[0], 4 * (size_t)p.dashes), out += 4 * (size_t)p.dashes;
        if (!colors.empty())
            std::memcpy(out, &colors[0], 4 * colors.size()), out += 4 * colors.size();
        if (!stops.empty())
            std::memcpy(out, &stops[0], 4 * stops.size());
        if (pending.size() > ((size_t)256 << 20))
            submit();
    }
};

} // namespace canvas_ity_hip

#ifdef CANVAS_ITY_HIP_AS_CANVAS_ITY
namespace canvas_ity = canvas_ity_hip;
#endif

#endif // CANVAS_ITY_HIP_HPP
