/*
 * cih_oracle.c -- TEST INFRASTRUCTURE ONLY: the CPU oracle.
 *
 * A plain-C99 restatement of canvas_ity's rasterise-and-composite pipeline,
 * used as the checker of the HIP library (tests/, smoke(), bench.py's CPU
 * baseline).  It follows the reference's algorithm step by step, with its float
 * operations in the reference's order and association, because flattening,
 * dashing and stroking decisions are threshold tests on float arithmetic.
 * Each function cites the canvas_ity.hpp lines ("hpp:") it restates.
 *
 * Pinning (tests/test_oracle.py): bit-identical fp32 bitmaps against the
 * reference header compiled from source (oracle/_ref) on seeded scenes of
 * every configuration; the reference's exact get_image_data known-answer hash
 * (test.cpp:1800-1830); closed forms (sRGB transfer, dither, coverage of
 * aligned and fractional rectangles, polygon areas, blur weights).  Where the
 * reference build is unavailable the closed-form and KAT checks still run.
 *
 * Text (hpp:1299-1314, 1526-1846, 3205-3311) is restated too: glyph outlines
 * go straight into the polylines, as in the reference.  Font reads are
 * bounds-checked (out-of-range bytes read as 0), which only matters for
 * malformed fonts, where the reference's reads are undefined.
 *
 * is_point_in_path (hpp:3101-3132) is restated for the test suite's own
 * query-dependent tests.
 */
#include "cih_oracle.h"
#include "cih_stream.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float x, y; } xy;
typedef struct { float r, g, b, a; } rgba;
typedef struct { float a, b, c, d, e, f; } affine;
typedef struct { size_t count; int closed; } subpath;
typedef struct { unsigned short x, y; float delta; } pixel_run; /* hpp:174 */

typedef struct { xy *v; size_t n, cap; } xyvec;
typedef struct { subpath *v; size_t n, cap; } subvec;
typedef struct { pixel_run *v; size_t n, cap; } runvec;
typedef struct { float *v; size_t n, cap; } fvec;
typedef struct { rgba *v; size_t n, cap; } rgbavec;

static void *grow(void *p, size_t *cap, size_t need, size_t elem)
{
    size_t c = *cap ? *cap : 16;
    while (c < need)
        c *= 2;
    p = realloc(p, c * elem);
    if (!p)
        abort();
    *cap = c;
    return p;
}

#define RESERVE(vec, need)                                                        \
    do {                                                                          \
        if ((need) > (vec).cap)                                                   \
            (vec).v = grow((vec).v, &(vec).cap, (need), sizeof(*(vec).v));        \
    } while (0)
#define PUSH(vec, value)                                                          \
    do {                                                                          \
        RESERVE(vec, (vec).n + 1);                                                \
        (vec).v[(vec).n++] = (value);                                             \
    } while (0)

typedef struct { xyvec points; subvec subpaths; } line_path; /* hpp:169-173 */

typedef struct { /* paint_brush, hpp:162-165 */
    int type; /* 0 color, 1 linear, 2 radial, 3 pattern */
    rgbavec colors;
    fvec stops;
    xy start, end;
    float start_radius, end_radius;
    int width, height, repetition;
} brush;

struct cih_oracle { /* canvas state, hpp:1150-1172 */
    int op; /* global_composite_operation */
    float shadow_offset_x, shadow_offset_y;
    int line_cap, line_join;
    float line_dash_offset;
    int size_x, size_y;
    affine forward, inverse;
    float global_alpha;
    rgba shadow_color;
    float shadow_blur;
    fvec shadow;
    float line_width, miter_limit;
    fvec line_dash;
    brush fill_brush, stroke_brush, image_brush;
    line_path path, lines, scratch;
    runvec runs, mask;
    rgba *bitmap;
    struct cih_oracle *saves;
    int evenodd; /* extension: even-odd fill rule */
    int text_align, text_baseline;
    struct { /* font_face, hpp:166-168: offset table + kept tables */
        unsigned char *data;
        size_t n;
        int cmap, glyf, head, hhea, hmtx, loca, maxp, os_2;
        float scale;
    } face;
};

enum { OP_SOURCE_OVER = 14, OP_DESTINATION_OUT = 12 };
enum { CAP_BUTT, CAP_SQUARE, CAP_CIRCLE };
enum { JOIN_MITER, JOIN_BEVEL, JOIN_ROUNDED };

/* ---- vector and colour math (hpp:1223-1297) ---------------------------- */

static xy mkxy(float x, float y) { xy v; v.x = x; v.y = y; return v; }
static xy add(xy l, xy r) { return mkxy(l.x + r.x, l.y + r.y); }
static xy sub(xy l, xy r) { return mkxy(l.x - r.x, l.y - r.y); }
static xy scale(float s, xy v) { return mkxy(v.x * s, v.y * s); }
static xy apply(const affine *m, xy v)
{
    return mkxy(m->a * v.x + m->c * v.y + m->e, m->b * v.x + m->d * v.y + m->f);
}
static float dot(xy l, xy r) { return l.x * r.x + l.y * r.y; }
static float length(xy v) { return sqrtf(dot(v, v)); }
static float direction(xy v) { return atan2f(v.y, v.x); }
static float max_std(float a, float b) { return (a < b) ? b : a; } /* std::max */
static float min_std(float a, float b) { return (b < a) ? b : a; } /* std::min */
static xy normalized(xy v) { return scale(1.0f / max_std(1.0e-6f, length(v)), v); }
static xy perpendicular(xy v) { return mkxy(-v.y, v.x); }
static xy lerp(xy from, xy to, float ratio) { return add(from, scale(ratio, sub(to, from))); }

static rgba mkrgba(float r, float g, float b, float a)
{
    rgba c; c.r = r; c.g = g; c.b = b; c.a = a; return c;
}
static rgba rgba_add(rgba l, rgba r) { return mkrgba(l.r + r.r, l.g + r.g, l.b + r.b, l.a + r.a); }
static rgba rgba_sub(rgba l, rgba r) { return mkrgba(l.r - r.r, l.g - r.g, l.b - r.b, l.a - r.a); }
static rgba rgba_scale(float s, rgba c) { return mkrgba(c.r * s, c.g * s, c.b * s, c.a * s); }
static float linearized_1(float v)
{
    return v < 0.04045f ? v / 12.92f : powf((v + 0.055f) / 1.055f, 2.4f);
}
static rgba linearized(rgba c)
{
    return mkrgba(linearized_1(c.r), linearized_1(c.g), linearized_1(c.b), c.a);
}
static rgba premultiplied(rgba c) { return mkrgba(c.r * c.a, c.g * c.a, c.b * c.a, c.a); }
static float delinearized_1(float v)
{
    return v < 0.0031308f ? 12.92f * v : 1.055f * powf(v, 1.0f / 2.4f) - 0.055f;
}
static rgba delinearized(rgba c)
{
    return mkrgba(delinearized_1(c.r), delinearized_1(c.g), delinearized_1(c.b), c.a);
}
static rgba unpremultiplied(rgba c)
{
    const float threshold = 1.0f / 8160.0f;
    return c.a < threshold ? mkrgba(0.0f, 0.0f, 0.0f, 0.0f)
                           : mkrgba(1.0f / c.a * c.r, 1.0f / c.a * c.g, 1.0f / c.a * c.b, c.a);
}
static float clamp01(float v) { return min_std(max_std(v, 0.0f), 1.0f); }
static rgba clamped(rgba c) { return mkrgba(clamp01(c.r), clamp01(c.g), clamp01(c.b), clamp01(c.a)); }

/* ---- flattening (hpp:1331-1524) ------------------------------------------ */

static void add_tessellation(cih_oracle *c, xy point_1, xy control_1, xy control_2, xy point_2,
                             float angular, int limit) /* hpp:1331-1387 */
{
    const float tolerance = 0.125f;
    float flatness = tolerance * tolerance;
    xy edge_1 = sub(control_1, point_1);
    xy edge_2 = sub(control_2, control_1);
    xy edge_3 = sub(point_2, control_2);
    xy segment = sub(point_2, point_1);
    float squared_1 = dot(edge_1, edge_1);
    float squared_2 = dot(edge_2, edge_2);
    float squared_3 = dot(edge_3, edge_3);
    const float epsilon = 1.0e-4f;
    float length_squared = max_std(epsilon, dot(segment, segment));
    float projection_1 = dot(edge_1, segment) / length_squared;
    float projection_2 = dot(edge_3, segment) / length_squared;
    float clamped_1 = min_std(max_std(projection_1, 0.0f), 1.0f);
    float clamped_2 = min_std(max_std(projection_2, 0.0f), 1.0f);
    xy to_line_1 = sub(add(point_1, scale(clamped_1, segment)), control_1);
    xy to_line_2 = sub(sub(point_2, scale(clamped_2, segment)), control_2);
    float cosine = 1.0f;
    if (angular > -1.0f) {
        if (squared_1 * squared_3 != 0.0f)
            cosine = dot(edge_1, edge_3) / sqrtf(squared_1 * squared_3);
        else if (squared_1 * squared_2 != 0.0f)
            cosine = dot(edge_1, edge_2) / sqrtf(squared_1 * squared_2);
        else if (squared_2 * squared_3 != 0.0f)
            cosine = dot(edge_2, edge_3) / sqrtf(squared_2 * squared_3);
    }
    if ((dot(to_line_1, to_line_1) <= flatness && dot(to_line_2, to_line_2) <= flatness &&
         cosine >= angular) || !limit) {
        if (angular > -1.0f && squared_1 != 0.0f)
            PUSH(c->lines.points, control_1);
        if (angular > -1.0f && squared_2 != 0.0f)
            PUSH(c->lines.points, control_2);
        if (angular == -1.0f || squared_3 != 0.0f)
            PUSH(c->lines.points, point_2);
        return;
    }
    {
        xy left_1 = lerp(point_1, control_1, 0.5f);
        xy middle = lerp(control_1, control_2, 0.5f);
        xy right_2 = lerp(control_2, point_2, 0.5f);
        xy left_2 = lerp(left_1, middle, 0.5f);
        xy right_1 = lerp(middle, right_2, 0.5f);
        xy split = lerp(left_2, right_1, 0.5f);
        add_tessellation(c, point_1, left_1, left_2, split, angular, limit - 1);
        add_tessellation(c, split, right_1, right_2, point_2, angular, limit - 1);
    }
}

static void add_bezier(cih_oracle *c, xy point_1, xy control_1, xy control_2, xy point_2,
                       float angular) /* hpp:1398-1487 */
{
    xy edge_1 = sub(control_1, point_1);
    xy edge_2 = sub(control_2, control_1);
    xy edge_3 = sub(point_2, control_2);
    float at[7] = {0.0f, 1.0f};
    int cuts = 2, axis, index;
    const float epsilon = 1.0e-4f;
    xy extrema_a, extrema_b, extrema_c, split_point_1;
    float determinant_1, determinant_2, determinant_3, curve_a, curve_b;
    if (dot(edge_1, edge_1) == 0.0f && dot(edge_3, edge_3) == 0.0f) {
        PUSH(c->lines.points, point_2);
        return;
    }
    extrema_a = add(scale(-9.0f, edge_2), scale(3.0f, sub(point_2, point_1)));
    extrema_b = sub(scale(6.0f, add(point_1, control_2)), scale(12.0f, control_1));
    extrema_c = scale(3.0f, edge_1);
    for (axis = 0; axis < 2; ++axis) { /* x then y, hpp:1420-1449 */
        float a = axis ? extrema_a.y : extrema_a.x;
        float b = axis ? extrema_b.y : extrema_b.x;
        float k = axis ? extrema_c.y : extrema_c.x;
        if (fabsf(a) > epsilon) {
            float discriminant = b * b - 4.0f * a * k;
            if (discriminant >= 0.0f) {
                float sign = b > 0.0f ? 1.0f : -1.0f;
                float term = -b - sign * sqrtf(discriminant);
                float extremum_1 = term / (2.0f * a);
                at[cuts++] = extremum_1;
                at[cuts++] = k / (a * extremum_1);
            }
        } else if (fabsf(b) > epsilon)
            at[cuts++] = -k / b;
    }
    determinant_1 = dot(perpendicular(edge_1), edge_2);
    determinant_2 = dot(perpendicular(edge_1), edge_3);
    determinant_3 = dot(perpendicular(edge_2), edge_3);
    curve_a = determinant_1 - determinant_2 + determinant_3;
    curve_b = -2.0f * determinant_1 + determinant_2;
    if (fabsf(curve_a) > epsilon && fabsf(curve_b) > epsilon)
        at[cuts++] = -0.5f * curve_b / curve_a;
    for (index = 1; index < cuts; ++index) {
        float value = at[index];
        int sorted = index - 1;
        for (; 0 <= sorted && value < at[sorted]; --sorted)
            at[sorted + 1] = at[sorted];
        at[sorted + 1] = value;
    }
    split_point_1 = point_1;
    for (index = 0; index < cuts - 1; ++index) {
        float t, ratio;
        xy partial_1, partial_2, partial_3, partial_4, partial_5, partial_6;
        xy split_point_2, split_control_1, split_control_2;
        if (!(0.0f <= at[index] && at[index + 1] <= 1.0f && at[index] != at[index + 1]))
            continue;
        t = at[index + 1];
        ratio = at[index] / t;
        partial_1 = lerp(point_1, control_1, t);
        partial_2 = lerp(control_1, control_2, t);
        partial_3 = lerp(control_2, point_2, t);
        partial_4 = lerp(partial_1, partial_2, t);
        partial_5 = lerp(partial_2, partial_3, t);
        partial_6 = lerp(partial_1, partial_4, ratio);
        split_point_2 = lerp(partial_4, partial_5, t);
        split_control_2 = lerp(partial_4, split_point_2, ratio);
        split_control_1 = lerp(partial_6, split_control_2, ratio);
        add_tessellation(c, split_point_1, split_control_1, split_control_2, split_point_2,
                         angular, 20);
        split_point_1 = split_point_2;
    }
}

static void path_to_lines(cih_oracle *c, int stroking) /* hpp:1495-1524 */
{
    const float tolerance = 0.125f;
    float ratio = tolerance / max_std(0.5f * c->line_width, tolerance);
    float angular = stroking ? (ratio - 2.0f) * ratio * 2.0f + 1.0f : -1.0f;
    size_t index = 0, ending = 0, s;
    c->lines.points.n = 0;
    c->lines.subpaths.n = 0;
    for (s = 0; s < c->path.subpaths.n; ++s) {
        size_t first = c->lines.points.n;
        xy point_1 = c->path.points.v[index++];
        subpath entry;
        ending += c->path.subpaths.v[s].count;
        PUSH(c->lines.points, point_1);
        for (; index < ending; index += 3) {
            xy control_1 = c->path.points.v[index + 0];
            xy control_2 = c->path.points.v[index + 1];
            xy point_2 = c->path.points.v[index + 2];
            add_bezier(c, point_1, control_1, control_2, point_2, angular);
            point_1 = point_2;
        }
        entry.count = c->lines.points.n - first;
        entry.closed = c->path.subpaths.v[s].closed;
        PUSH(c->lines.subpaths, entry);
    }
}

/* ---- dashing and stroking (hpp:1858-2100) -------------------------------- */

static void swap_paths(line_path *a, line_path *b)
{
    line_path t = *a;
    *a = *b;
    *b = t;
}

static void rotate_points(xy *first, xy *middle, xy *last) /* std::rotate */
{
    size_t n = (size_t)(last - first), k = (size_t)(middle - first), i;
    xy *tmp;
    if (!k || k == n)
        return;
    tmp = (xy *)malloc(n * sizeof(xy));
    if (!tmp)
        abort();
    for (i = 0; i < n; ++i)
        tmp[i] = first[(i + k) % n];
    memcpy(first, tmp, n * sizeof(xy));
    free(tmp);
}

static void dash_lines(cih_oracle *c) /* hpp:1858-1934 */
{
    float total = 0.0f, offset;
    size_t start = 0, ending = 0, s, k;
    size_t ndash = c->line_dash.n;
    const float *dash = c->line_dash.v;
    if (!ndash)
        return;
    swap_paths(&c->lines, &c->scratch);
    c->lines.points.n = 0;
    c->lines.subpaths.n = 0;
    for (k = 0; k < ndash; ++k)
        total = total + dash[k];
    offset = fmodf(c->line_dash_offset, total);
    if (offset < 0.0f)
        offset += total;
    while (offset >= dash[start]) {
        offset -= dash[start];
        start = start + 1 < ndash ? start + 1 : 0;
    }
    for (s = 0; s < c->scratch.subpaths.n; ++s) {
        size_t index = ending;
        size_t first = c->lines.points.n;
        size_t segment = start;
        int emit = (int)(~start & 1);
        size_t merge_point = c->lines.points.n;
        size_t merge_subpath = c->lines.subpaths.n;
        int merge_emit = emit;
        float next = dash[start] - offset;
        ending += c->scratch.subpaths.v[s].count;
        for (; index + 1 < ending; ++index) {
            xy from = c->scratch.points.v[index];
            xy to = c->scratch.points.v[index + 1];
            float line;
            if (emit)
                PUSH(c->lines.points, from);
            line = length(sub(apply(&c->inverse, to), apply(&c->inverse, from)));
            while (next < line) {
                PUSH(c->lines.points, lerp(from, to, next / line));
                if (emit) {
                    subpath entry;
                    entry.count = c->lines.points.n - first;
                    entry.closed = 0;
                    PUSH(c->lines.subpaths, entry);
                    first = c->lines.points.n;
                }
                segment = segment + 1 < ndash ? segment + 1 : 0;
                emit = !emit;
                next += dash[segment];
            }
            next -= line;
        }
        if (emit) {
            subpath entry;
            PUSH(c->lines.points, c->scratch.points.v[index]);
            entry.count = c->lines.points.n - first;
            entry.closed = 0;
            PUSH(c->lines.subpaths, entry);
            if (c->scratch.subpaths.v[s].closed && merge_emit) {
                if (c->lines.subpaths.n == merge_subpath + 1)
                    c->lines.subpaths.v[c->lines.subpaths.n - 1].closed = 1;
                else {
                    size_t count = c->lines.subpaths.v[c->lines.subpaths.n - 1].count;
                    rotate_points(c->lines.points.v + merge_point,
                                  c->lines.points.v + c->lines.points.n - count,
                                  c->lines.points.v + c->lines.points.n);
                    c->lines.subpaths.v[merge_subpath].count += count;
                    c->lines.subpaths.n--;
                }
            }
        }
    }
}

static void push_forward(cih_oracle *c, xy p) { PUSH(c->lines.points, apply(&c->forward, p)); }

static void add_half_stroke(cih_oracle *c, size_t beginning, size_t ending,
                            int closed) /* hpp:1949-2058 */
{
    float half = c->line_width * 0.5f;
    float ratio = c->miter_limit * c->miter_limit * half * half;
    xy in_direction = mkxy(0.0f, 0.0f);
    float in_length = 0.0f;
    xy point = apply(&c->inverse, c->scratch.points.v[beginning]);
    size_t finish = beginning, index = beginning;
    const float epsilon = 1.0e-4f;
    xy ahead, side;
    do {
        xy next = apply(&c->inverse, c->scratch.points.v[index]);
        xy out_direction = normalized(sub(next, point));
        float out_length = length(sub(next, point));
        if (in_length != 0.0f && out_length >= epsilon) {
            xy side_in, side_out, offset, t;
            float turn;
            int tight;
            if (closed && finish == beginning)
                finish = index;
            side_in = add(point, scale(half, perpendicular(in_direction)));
            side_out = add(point, scale(half, perpendicular(out_direction)));
            turn = dot(perpendicular(in_direction), out_direction);
            if (fabsf(turn) < epsilon)
                turn = 0.0f;
            offset = turn == 0.0f ? mkxy(0.0f, 0.0f)
                                  : scale(half / turn, sub(out_direction, in_direction));
            tight = (dot(offset, in_direction) < -in_length &&
                     dot(offset, out_direction) > out_length);
            if (turn > 0.0f && tight) {
                t = side_in, side_in = side_out, side_out = t;
                t = in_direction, in_direction = out_direction, out_direction = t;
                push_forward(c, side_out);
                push_forward(c, point);
                push_forward(c, side_in);
            }
            if ((turn > 0.0f && !tight) ||
                (turn != 0.0f && c->line_join == JOIN_MITER && dot(offset, offset) <= ratio))
                push_forward(c, add(point, offset));
            else if (c->line_join == JOIN_ROUNDED) {
                float cosine = dot(in_direction, out_direction);
                float angle = acosf(min_std(max_std(cosine, -1.0f), 1.0f));
                float alpha = 4.0f / 3.0f * tanf(0.25f * angle);
                push_forward(c, side_in);
                add_bezier(c, apply(&c->forward, side_in),
                           apply(&c->forward, add(side_in, scale(alpha * half, in_direction))),
                           apply(&c->forward, sub(side_out, scale(alpha * half, out_direction))),
                           apply(&c->forward, side_out), -1.0f);
            } else {
                push_forward(c, side_in);
                push_forward(c, side_out);
            }
            if (turn > 0.0f && tight) {
                push_forward(c, side_out);
                push_forward(c, point);
                push_forward(c, side_in);
                t = in_direction, in_direction = out_direction, out_direction = t;
            }
        }
        if (out_length >= epsilon) {
            in_direction = out_direction;
            in_length = out_length;
            point = next;
        }
        index = (index == ending ? beginning : ending > beginning ? index + 1 : index - 1);
    } while (index != finish);
    if (closed || in_length == 0.0f)
        return;
    ahead = scale(half, in_direction);
    side = perpendicular(ahead);
    if (c->line_cap == CAP_BUTT) {
        push_forward(c, add(point, side));
        push_forward(c, sub(point, side));
    } else if (c->line_cap == CAP_SQUARE) {
        push_forward(c, add(add(point, ahead), side));
        push_forward(c, sub(add(point, ahead), side));
    } else if (c->line_cap == CAP_CIRCLE) {
        const float alpha = 0.55228475f;
        push_forward(c, add(point, side));
        add_bezier(c, apply(&c->forward, add(point, side)),
                   apply(&c->forward, add(add(point, side), scale(alpha, ahead))),
                   apply(&c->forward, add(add(point, ahead), scale(alpha, side))),
                   apply(&c->forward, add(point, ahead)), -1.0f);
        add_bezier(c, apply(&c->forward, add(point, ahead)),
                   apply(&c->forward, sub(add(point, ahead), scale(alpha, side))),
                   apply(&c->forward, add(sub(point, side), scale(alpha, ahead))),
                   apply(&c->forward, sub(point, side)), -1.0f);
    }
}

static void stroke_lines(cih_oracle *c) /* hpp:2070-2100 */
{
    size_t ending = 0, s;
    if (c->forward.a * c->forward.d - c->forward.b * c->forward.c == 0.0f)
        return;
    dash_lines(c);
    swap_paths(&c->lines, &c->scratch);
    c->lines.points.n = 0;
    c->lines.subpaths.n = 0;
    for (s = 0; s < c->scratch.subpaths.n; ++s) {
        size_t beginning = ending, first;
        int closed = c->scratch.subpaths.v[s].closed;
        subpath entry;
        ending += c->scratch.subpaths.v[s].count;
        if (ending - beginning < 2)
            continue;
        first = c->lines.points.n;
        add_half_stroke(c, beginning, ending - 1, closed);
        if (closed) {
            entry.count = c->lines.points.n - first;
            entry.closed = 1;
            PUSH(c->lines.subpaths, entry);
            first = c->lines.points.n;
        }
        add_half_stroke(c, ending - 1, beginning, closed);
        entry.count = c->lines.points.n - first;
        entry.closed = 1;
        PUSH(c->lines.subpaths, entry);
    }
}

/* ---- scan conversion (hpp:2109-2253) ------------------------------------- */

static void push_run(cih_oracle *c, float x, float y, float delta)
{
    pixel_run r;
    r.x = (unsigned short)x;
    r.y = (unsigned short)y;
    r.delta = delta;
    PUSH(c->runs, r);
}

static void add_runs(cih_oracle *c, xy from, xy to) /* hpp:2109-2170 */
{
    const float epsilon = 2.0e-5f;
    float sign, y_step;
    xy now, pixel, corner, slope, next_x, next_y;
    if (fabsf(to.y - from.y) < epsilon)
        return;
    sign = to.y > from.y ? 1.0f : -1.0f;
    if (from.x > to.x) {
        xy t = from;
        from = to;
        to = t;
    }
    now = from;
    pixel = mkxy(floorf(now.x), floorf(now.y));
    corner = add(pixel, mkxy(1.0f, to.y > from.y ? 1.0f : 0.0f));
    slope = mkxy((to.x - from.x) / (to.y - from.y), (to.y - from.y) / (to.x - from.x));
    next_x = (to.x - from.x < epsilon) ? to
                                       : mkxy(corner.x, now.y + (corner.x - now.x) * slope.y);
    next_y = mkxy(now.x + (corner.y - now.y) * slope.x, corner.y);
    if ((from.y < to.y && to.y < next_y.y) || (from.y > to.y && to.y > next_y.y))
        next_y = to;
    y_step = to.y > from.y ? 1.0f : -1.0f;
    do {
        float carry = 0.0f, strip, mid, area;
        while (next_x.x < next_y.x) {
            strip = min_std(max_std((next_x.y - now.y) * y_step, 0.0f), 1.0f);
            mid = (next_x.x + now.x) * 0.5f;
            area = (mid - pixel.x) * strip;
            push_run(c, pixel.x, pixel.y, (carry + strip - area) * sign);
            carry = area;
            now = next_x;
            next_x.x += 1.0f;
            next_x.y = (next_x.x - from.x) * slope.y + from.y;
            pixel.x += 1.0f;
        }
        strip = min_std(max_std((next_y.y - now.y) * y_step, 0.0f), 1.0f);
        mid = (next_y.x + now.x) * 0.5f;
        area = (mid - pixel.x) * strip;
        push_run(c, pixel.x, pixel.y, (carry + strip - area) * sign);
        push_run(c, pixel.x + 1.0f, pixel.y, area * sign);
        now = next_y;
        next_y.y += y_step;
        next_y.x = (next_y.y - from.y) * slope.x + from.x;
        pixel.y += y_step;
        if ((from.y < to.y && to.y < next_y.y) || (from.y > to.y && to.y > next_y.y))
            next_y = to;
    } while (now.y != to.y);
}

static int run_less(pixel_run l, pixel_run r) /* operator<, hpp:2172-2181 */
{
    return l.y < r.y ? 1 : l.y > r.y ? 0 : l.x < r.x ? 1 : l.x > r.x ? 0
                                                             : fabsf(l.delta) < fabsf(r.delta);
}

static int run_compare(const void *a, const void *b)
{
    pixel_run l = *(const pixel_run *)a, r = *(const pixel_run *)b;
    if (run_less(l, r))
        return -1;
    if (run_less(r, l))
        return 1;
    return l.delta < r.delta ? -1 : l.delta > r.delta ? 1 : 0;
}

static void lines_to_runs(cih_oracle *c, xy offset, int padding) /* hpp:2193-2253 */
{
    float width = (float)(c->size_x + padding);
    float height = (float)(c->size_y + padding);
    size_t ending = 0, s, to, from;
    xyvec *pts = &c->scratch.points;
    c->runs.n = 0;
    for (s = 0; s < c->lines.subpaths.n; ++s) {
        size_t beginning = ending, index, last;
        int edge;
        ending += c->lines.subpaths.v[s].count;
        pts->n = 0;
        for (index = beginning; index < ending; ++index)
            PUSH(*pts, add(offset, c->lines.points.v[index]));
        for (edge = 0; edge < 4; ++edge) { /* Sutherland-Hodgman, hpp:2208-2229 */
            xy normal = mkxy(edge == 0 ? 1.0f : edge == 2 ? -1.0f : 0.0f,
                             edge == 1 ? 1.0f : edge == 3 ? -1.0f : 0.0f);
            float place = edge == 2 ? width : edge == 3 ? height : 0.0f;
            size_t first = pts->n;
            for (index = 0; index < first; ++index) {
                xy a = pts->v[(index ? index : first) - 1];
                xy b = pts->v[index];
                float from_side = dot(a, normal) + place;
                float to_side = dot(b, normal) + place;
                if (from_side * to_side < 0.0f)
                    PUSH(*pts, lerp(a, b, from_side / (from_side - to_side)));
                if (to_side >= 0.0f)
                    PUSH(*pts, b);
            }
            memmove(pts->v, pts->v + first, (pts->n - first) * sizeof(xy));
            pts->n -= first;
        }
        last = pts->n;
        for (index = 0; index < last; ++index) {
            xy a = pts->v[(index ? index : last) - 1];
            xy b = pts->v[index];
            add_runs(c, mkxy(min_std(max_std(a.x, 0.0f), width), min_std(max_std(a.y, 0.0f), height)),
                     mkxy(min_std(max_std(b.x, 0.0f), width), min_std(max_std(b.y, 0.0f), height)));
        }
    }
    if (!c->runs.n)
        return;
    qsort(c->runs.v, c->runs.n, sizeof(pixel_run), run_compare);
    to = 0;
    for (from = 1; from < c->runs.n; ++from)
        if (c->runs.v[from].x == c->runs.v[to].x && c->runs.v[from].y == c->runs.v[to].y)
            c->runs.v[to].delta += c->runs.v[from].delta;
        else if (c->runs.v[from].delta != 0.0f)
            c->runs.v[++to] = c->runs.v[from];
    c->runs.n = to + 1;
}

/* Coverage of a running sum: non-zero (hpp:2570) or, as an extension, even-odd. */
static float coverage_of(const cih_oracle *c, float sum)
{
    if (c->evenodd) {
        float a = fmodf(fabsf(sum), 2.0f);
        return min_std(a > 1.0f ? 2.0f - a : a, 1.0f);
    }
    return min_std(fabsf(sum), 1.0f);
}

/* ---- paint (hpp:2265-2377) ----------------------------------------------- */

static rgba paint_pixel(const cih_oracle *c, xy point, const brush *b, int image)
{
    float offset;
    xy relative, line;
    float gradient, span;
    size_t index, n;
    if (!b->colors.n)
        return mkrgba(0.0f, 0.0f, 0.0f, 0.0f);
    if (b->type == 0)
        return b->colors.v[0];
    point = apply(&c->inverse, point);
    if (b->type == 3) {
        float width = (float)b->width, height = (float)b->height;
        float scale_x, scale_y, reciprocal_x, reciprocal_y, total_weight = 0.0f;
        int left, top, right, bottom, py, px;
        rgba total = mkrgba(0.0f, 0.0f, 0.0f, 0.0f);
        if (((b->repetition & 2) && (point.x < 0.0f || width <= point.x)) ||
            ((b->repetition & 1) && (point.y < 0.0f || height <= point.y)))
            return mkrgba(0.0f, 0.0f, 0.0f, 0.0f);
        scale_x = fabsf(c->inverse.a) + fabsf(c->inverse.c);
        scale_y = fabsf(c->inverse.b) + fabsf(c->inverse.d);
        scale_x = max_std(1.0f, min_std(scale_x, width * 0.25f));
        scale_y = max_std(1.0f, min_std(scale_y, height * 0.25f));
        reciprocal_x = 1.0f / scale_x;
        reciprocal_y = 1.0f / scale_y;
        point = sub(point, mkxy(0.5f, 0.5f));
        left = (int)ceilf(point.x - scale_x * 2.0f);
        top = (int)ceilf(point.y - scale_y * 2.0f);
        right = (int)ceilf(point.x + scale_x * 2.0f);
        bottom = (int)ceilf(point.y + scale_y * 2.0f);
        for (py = top; py < bottom; ++py) {
            float y = fabsf(reciprocal_y * ((float)py - point.y));
            float weight_y = y < 1.0f ? (1.5f * y - 2.5f) * y * y + 1.0f
                                      : ((-0.5f * y + 2.5f) * y - 4.0f) * y + 2.0f;
            int wrapped_y = py % b->height;
            if (wrapped_y < 0)
                wrapped_y += b->height;
            if (image)
                wrapped_y = py < 0 ? 0 : py > b->height - 1 ? b->height - 1 : py;
            for (px = left; px < right; ++px) {
                float x = fabsf(reciprocal_x * ((float)px - point.x));
                float weight_x = x < 1.0f ? (1.5f * x - 2.5f) * x * x + 1.0f
                                          : ((-0.5f * x + 2.5f) * x - 4.0f) * x + 2.0f;
                float weight;
                int wrapped_x = px % b->width;
                if (wrapped_x < 0)
                    wrapped_x += b->width;
                if (image)
                    wrapped_x = px < 0 ? 0 : px > b->width - 1 ? b->width - 1 : px;
                weight = weight_x * weight_y;
                total = rgba_add(total, rgba_scale(weight,
                                 b->colors.v[(size_t)(wrapped_y * b->width + wrapped_x)]));
                total_weight += weight;
            }
        }
        return rgba_scale(1.0f / total_weight, total);
    }
    relative = sub(point, b->start);
    line = sub(b->end, b->start);
    gradient = dot(relative, line);
    span = dot(line, line);
    if (b->type == 1) {
        if (span == 0.0f)
            return mkrgba(0.0f, 0.0f, 0.0f, 0.0f);
        offset = gradient / span;
    } else {
        float initial = b->start_radius;
        float change = b->end_radius - initial;
        float a = span - change * change;
        float bb = -2.0f * (gradient + initial * change);
        float cc = dot(relative, relative) - initial * initial;
        float discriminant = bb * bb - 4.0f * a * cc;
        float root, reciprocal, offset_1, offset_2, radius_1, radius_2;
        if (discriminant < 0.0f || (span == 0.0f && change == 0.0f))
            return mkrgba(0.0f, 0.0f, 0.0f, 0.0f);
        root = sqrtf(discriminant);
        reciprocal = 1.0f / (2.0f * a);
        offset_1 = (-bb - root) * reciprocal;
        offset_2 = (-bb + root) * reciprocal;
        radius_1 = initial + change * offset_1;
        radius_2 = initial + change * offset_2;
        if (radius_2 >= 0.0f)
            offset = offset_2;
        else if (radius_1 >= 0.0f)
            offset = offset_1;
        else
            return mkrgba(0.0f, 0.0f, 0.0f, 0.0f);
    }
    n = b->stops.n; /* std::upper_bound */
    for (index = 0; index < n && !(offset < b->stops.v[index]); ++index)
        ;
    if (index == 0)
        return premultiplied(b->colors.v[0]);
    if (index == n)
        return premultiplied(b->colors.v[n - 1]);
    {
        float mix = (offset - b->stops.v[index - 1]) /
                    (b->stops.v[index] - b->stops.v[index - 1]);
        rgba delta = rgba_sub(b->colors.v[index], b->colors.v[index - 1]);
        return premultiplied(rgba_add(b->colors.v[index - 1], rgba_scale(mix, delta)));
    }
}

/* ---- render (hpp:2395-2605) ---------------------------------------------- */

static rgba composite(rgba back, rgba fore, int operation, float visibility)
{
    float mix_fore = operation & 1 ? back.a : 0.0f;
    float mix_back = operation & 4 ? fore.a : 0.0f;
    rgba blend;
    if (operation & 2)
        mix_fore = 1.0f - mix_fore;
    if (operation & 8)
        mix_back = 1.0f - mix_back;
    blend = rgba_add(rgba_scale(mix_fore, fore), rgba_scale(mix_back, back));
    blend.a = min_std(blend.a, 1.0f);
    return rgba_add(rgba_scale(visibility, blend), rgba_scale(1.0f - visibility, back));
}

static void render_shadow(cih_oracle *c, const brush *b, int image) /* hpp:2395-2539 */
{
    float sigma_squared, alpha, divisor, weight_1, weight_2;
    size_t radius, width, height, working, index, x, y;
    int border, left, right, top, bottom, operation, cx, cy;
    float sum;
    xy offset;
    const float threshold = 1.0f / 8160.0f;
    if (c->shadow_color.a == 0.0f ||
        (c->shadow_blur == 0.0f && c->shadow_offset_x == 0.0f && c->shadow_offset_y == 0.0f))
        return;
    sigma_squared = 0.25f * c->shadow_blur * c->shadow_blur;
    radius = (size_t)(0.5f * sqrtf(4.0f * sigma_squared + 1.0f) - 0.5f);
    border = 3 * ((int)radius + 1);
    offset = mkxy((float)border + c->shadow_offset_x, (float)border + c->shadow_offset_y);
    lines_to_runs(c, offset, 2 * border);
    left = c->size_x + 2 * border;
    right = 0;
    top = c->size_y + 2 * border;
    bottom = 0;
    for (index = 0; index < c->runs.n; ++index) {
        int rx = c->runs.v[index].x, ry = c->runs.v[index].y;
        left = rx < left ? rx : left;
        right = rx > right ? rx : right;
        top = ry < top ? ry : top;
        bottom = ry > bottom ? ry : bottom;
    }
    left = left - border > 0 ? left - border : 0;
    right = (right + border < c->size_x + 2 * border ? right + border : c->size_x + 2 * border) + 1;
    top = top - border > 0 ? top - border : 0;
    bottom = bottom + border < c->size_y + 2 * border ? bottom + border : c->size_y + 2 * border;
    width = (size_t)(right - left > 0 ? right - left : 0);
    height = (size_t)(bottom - top > 0 ? bottom - top : 0);
    working = width * height;
    c->shadow.n = 0;
    RESERVE(c->shadow, working + (width > height ? width : height) + 1);
    c->shadow.n = working + (width > height ? width : height);
    memset(c->shadow.v, 0, c->shadow.n * sizeof(float));
    { /* alpha raster, hpp:2427-2452 */
        int px = -1, py = -1;
        sum = 0.0f;
        for (index = 0; index < c->runs.n; ++index) {
            pixel_run next = c->runs.v[index];
            float cov = coverage_of(c, sum);
            int to = next.y == py ? next.x : px + 1;
            if (cov >= threshold)
                for (; px < to; ++px)
                    c->shadow.v[(size_t)(py - top) * width + (size_t)(px - left)] =
                        cov * paint_pixel(c, sub(mkxy((float)px + 0.5f, (float)py + 0.5f), offset),
                                          b, image).a;
            if (next.y != py)
                sum = 0.0f;
            px = next.x;
            py = next.y;
            sum += next.delta;
        }
    }
    alpha = (float)(2 * radius + 1) * ((float)(radius * (radius + 1)) - sigma_squared) /
            (2.0f * sigma_squared - (float)(6 * (radius + 1) * (radius + 1)));
    divisor = 2.0f * (alpha + (float)radius) + 1.0f;
    weight_1 = alpha / divisor;
    weight_2 = (1.0f - alpha) / divisor;
    for (y = 0; y < height; ++y) { /* rows, hpp:2460-2481 */
        int pass;
        for (pass = 0; pass < 3; ++pass) {
            float *w = c->shadow.v + working, *row = c->shadow.v + y * width, running;
            for (x = 0; x < width; ++x)
                w[x] = row[x];
            running = weight_1 * w[radius + 1];
            for (x = 0; x <= radius; ++x)
                running += (weight_1 + weight_2) * w[x];
            row[0] = running;
            for (x = 1; x < width; ++x) {
                if (x >= radius + 1)
                    running -= weight_2 * w[x - radius - 1];
                if (x >= radius + 2)
                    running -= weight_1 * w[x - radius - 2];
                if (x + radius < width)
                    running += weight_2 * w[x + radius];
                if (x + radius + 1 < width)
                    running += weight_1 * w[x + radius + 1];
                row[x] = running;
            }
        }
    }
    for (x = 0; x < width; ++x) { /* columns, hpp:2482-2503 */
        int pass;
        for (pass = 0; pass < 3; ++pass) {
            float *w = c->shadow.v + working, running;
            for (y = 0; y < height; ++y)
                w[y] = c->shadow.v[y * width + x];
            running = weight_1 * w[radius + 1];
            for (y = 0; y <= radius; ++y)
                running += (weight_1 + weight_2) * w[y];
            c->shadow.v[x] = running;
            for (y = 1; y < height; ++y) {
                if (y >= radius + 1)
                    running -= weight_2 * w[y - radius - 1];
                if (y >= radius + 2)
                    running -= weight_1 * w[y - radius - 2];
                if (y + radius < height)
                    running += weight_2 * w[y + radius];
                if (y + radius + 1 < height)
                    running += weight_1 * w[y + radius + 1];
                c->shadow.v[y * width + x] = running;
            }
        }
    }
    operation = c->op; /* composite under the mask, hpp:2504-2538 */
    cx = -1;
    cy = -1;
    sum = 0.0f;
    for (index = 0; index < c->mask.n; ++index) {
        pixel_run next = c->mask.v[index];
        float visibility = min_std(fabsf(sum), 1.0f);
        int to = next.y == cy ? next.x : cx + 1;
        if (right - border < to)
            to = right - border;
        if (visibility >= threshold && top <= cy + border && cy + border < bottom)
            for (; cx < to; ++cx) {
                rgba *back = &c->bitmap[cy * c->size_x + cx];
                rgba fore = rgba_scale(c->global_alpha *
                                           c->shadow.v[(size_t)(cy + border - top) * width +
                                                       (size_t)(cx + border - left)],
                                       c->shadow_color);
                *back = composite(*back, fore, operation, visibility);
            }
        if (next.y != cy)
            sum = 0.0f;
        cx = next.x > left - border ? next.x : left - border;
        cy = next.y;
        sum += next.delta;
    }
}

static void render_main(cih_oracle *c, const brush *b, int image) /* hpp:2551-2605 */
{
    int operation, x = -1, y = -1;
    float path_sum = 0.0f, clip_sum = 0.0f;
    size_t path_index = 0, clip_index = 0;
    const float threshold = 1.0f / 8160.0f;
    if (c->forward.a * c->forward.d - c->forward.b * c->forward.c == 0.0f)
        return;
    render_shadow(c, b, image);
    lines_to_runs(c, mkxy(0.0f, 0.0f), 0);
    operation = c->op;
    while (clip_index < c->mask.n) {
        int which = path_index < c->runs.n &&
                    run_less(c->runs.v[path_index], c->mask.v[clip_index]);
        pixel_run next = which ? c->runs.v[path_index] : c->mask.v[clip_index];
        float cov = coverage_of(c, path_sum);
        float visibility = min_std(fabsf(clip_sum), 1.0f);
        int to = next.y == y ? next.x : x + 1;
        if ((cov >= threshold || (~operation & 8)) && visibility >= threshold)
            for (; x < to; ++x) {
                rgba *back = &c->bitmap[y * c->size_x + x];
                rgba fore = rgba_scale(cov * c->global_alpha,
                                       paint_pixel(c, mkxy((float)x + 0.5f, (float)y + 0.5f),
                                                   b, image));
                *back = composite(*back, fore, operation, visibility);
            }
        x = next.x;
        if (next.y != y) {
            y = next.y;
            path_sum = 0.0f;
            clip_sum = 0.0f;
        }
        if (which)
            path_sum += c->runs.v[path_index++].delta;
        else
            clip_sum += c->mask.v[clip_index++].delta;
    }
}

/* ---- public API (hpp:2607-3465) ------------------------------------------ */

static void brush_free(brush *b)
{
    free(b->colors.v);
    free(b->stops.v);
    memset(b, 0, sizeof(*b));
}

static void brush_copy(brush *to, const brush *from)
{
    brush_free(to);
    *to = *from;
    to->colors.v = NULL, to->colors.n = to->colors.cap = 0;
    to->stops.v = NULL, to->stops.n = to->stops.cap = 0;
    RESERVE(to->colors, from->colors.n + 1);
    RESERVE(to->stops, from->stops.n + 1);
    if (from->colors.n)
        memcpy(to->colors.v, from->colors.v, from->colors.n * sizeof(rgba));
    if (from->stops.n)
        memcpy(to->stops.v, from->stops.v, from->stops.n * sizeof(float));
    to->colors.n = from->colors.n;
    to->stops.n = from->stops.n;
}

static void set_transform(cih_oracle *c, float a, float b, float cc, float d, float e,
                          float f) /* hpp:2695-2711 */
{
    float determinant = a * d - b * cc;
    float scaling = determinant != 0.0f ? 1.0f / determinant : 0.0f;
    affine fw, iv;
    fw.a = a, fw.b = b, fw.c = cc, fw.d = d, fw.e = e, fw.f = f;
    iv.a = scaling * d, iv.b = scaling * -b, iv.c = scaling * -cc, iv.d = scaling * a;
    iv.e = scaling * (cc * f - d * e), iv.f = scaling * (b * e - a * f);
    c->forward = fw;
    c->inverse = iv;
}

static void transform(cih_oracle *c, float a, float b, float cc, float d, float e,
                      float f) /* hpp:2679-2693 */
{
    affine m = c->forward;
    set_transform(c, m.a * a + m.c * b, m.b * a + m.d * b, m.a * cc + m.c * d,
                  m.b * cc + m.d * d, m.a * e + m.c * f + m.e, m.b * e + m.d * f + m.f);
}

static void set_color(cih_oracle *c, int type, float r, float g, float b, float a) /* hpp:2768-2779 */
{
    brush *br = type == 0 ? &c->fill_brush : &c->stroke_brush;
    br->type = 0;
    br->colors.n = 0;
    PUSH(br->colors, premultiplied(linearized(clamped(mkrgba(r, g, b, a)))));
}

static void set_pattern(brush *br, const unsigned char *image, int width,
                        int height, int stride, int repetition) /* hpp:2839-2864 */
{
    int x, y;
    if (!image || width <= 0 || height <= 0)
        return;
    br->type = 3;
    br->colors.n = 0;
    for (y = 0; y < height; ++y)
        for (x = 0; x < width; ++x) {
            int index = y * stride + x * 4;
            rgba color = mkrgba(image[index + 0] / 255.0f, image[index + 1] / 255.0f,
                                image[index + 2] / 255.0f, image[index + 3] / 255.0f);
            PUSH(br->colors, premultiplied(linearized(color)));
        }
    br->width = width;
    br->height = height;
    br->repetition = repetition;
}

static void move_to(cih_oracle *c, float x, float y) /* hpp:2872-2883 */
{
    subpath s;
    if (c->path.subpaths.n && c->path.subpaths.v[c->path.subpaths.n - 1].count == 1) {
        c->path.points.v[c->path.points.n - 1] = apply(&c->forward, mkxy(x, y));
        return;
    }
    s.count = 1;
    s.closed = 0;
    PUSH(c->path.points, apply(&c->forward, mkxy(x, y)));
    PUSH(c->path.subpaths, s);
}

static void line_to(cih_oracle *c, float x, float y) /* hpp:2900-2916 */
{
    xy point_1, point_2;
    if (!c->path.subpaths.n) {
        move_to(c, x, y);
        return;
    }
    point_1 = c->path.points.v[c->path.points.n - 1];
    point_2 = apply(&c->forward, mkxy(x, y));
    if (dot(sub(point_2, point_1), sub(point_2, point_1)) == 0.0f)
        return;
    PUSH(c->path.points, point_1);
    PUSH(c->path.points, point_2);
    PUSH(c->path.points, point_2);
    c->path.subpaths.v[c->path.subpaths.n - 1].count += 3;
}

static void close_path(cih_oracle *c) /* hpp:2885-2898 */
{
    xy first;
    affine saved = c->forward, identity = {1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f};
    if (!c->path.subpaths.n)
        return;
    first = c->path.points.v[c->path.points.n - c->path.subpaths.v[c->path.subpaths.n - 1].count];
    c->forward = identity;
    line_to(c, first.x, first.y);
    c->path.subpaths.v[c->path.subpaths.n - 1].closed = 1;
    move_to(c, first.x, first.y);
    c->forward = saved;
}

static void bezier_curve_to(cih_oracle *c, float c1x, float c1y, float c2x, float c2y, float x,
                            float y) /* hpp:2935-2951 */
{
    if (!c->path.subpaths.n)
        move_to(c, c1x, c1y);
    PUSH(c->path.points, apply(&c->forward, mkxy(c1x, c1y)));
    PUSH(c->path.points, apply(&c->forward, mkxy(c2x, c2y)));
    PUSH(c->path.points, apply(&c->forward, mkxy(x, y)));
    c->path.subpaths.v[c->path.subpaths.n - 1].count += 3;
}

static void quadratic_curve_to(cih_oracle *c, float cx, float cy, float x,
                               float y) /* hpp:2918-2933 */
{
    xy point_1, control, point_2;
    if (!c->path.subpaths.n)
        move_to(c, cx, cy);
    point_1 = c->path.points.v[c->path.points.n - 1];
    control = apply(&c->forward, mkxy(cx, cy));
    point_2 = apply(&c->forward, mkxy(x, y));
    PUSH(c->path.points, lerp(point_1, control, 2.0f / 3.0f));
    PUSH(c->path.points, lerp(point_2, control, 2.0f / 3.0f));
    PUSH(c->path.points, point_2);
    c->path.subpaths.v[c->path.subpaths.n - 1].count += 3;
}

static void arc(cih_oracle *c, float x, float y, float radius, float start_angle,
                float end_angle, int counter_clockwise) /* hpp:2990-3029 */
{
    const float tau = 6.28318531f;
    float winding, from, span, segment, alpha;
    int steps, step;
    xy centered_1;
    if (radius < 0.0f)
        return;
    winding = counter_clockwise ? -1.0f : 1.0f;
    from = fmodf(start_angle, tau);
    span = fmodf(end_angle, tau) - from;
    if ((end_angle - start_angle) * winding >= tau)
        span = tau * winding;
    else if (span * winding < 0.0f)
        span += tau * winding;
    centered_1 = scale(radius, mkxy(cosf(from), sinf(from)));
    line_to(c, x + centered_1.x, y + centered_1.y);
    if (span == 0.0f)
        return;
    steps = (int)max_std(1.0f, roundf(16.0f / tau * span * winding));
    segment = span / (float)steps;
    alpha = 4.0f / 3.0f * tanf(0.25f * segment);
    for (step = 0; step < steps; ++step) {
        float angle = from + (float)(step + 1) * segment;
        xy centered_2 = scale(radius, mkxy(cosf(angle), sinf(angle)));
        xy point_1 = add(mkxy(x, y), centered_1);
        xy point_2 = add(mkxy(x, y), centered_2);
        xy control_1 = add(point_1, scale(alpha, perpendicular(centered_1)));
        xy control_2 = sub(point_2, scale(alpha, perpendicular(centered_2)));
        bezier_curve_to(c, control_1.x, control_1.y, control_2.x, control_2.y, point_2.x,
                        point_2.y);
        centered_1 = centered_2;
    }
}

static void arc_to(cih_oracle *c, float vx, float vy, float x, float y,
                   float radius) /* hpp:2953-2988 */
{
    xy point_1, vertex, point_2, edge_1, edge_2, offset, center;
    float sine, angle_1, angle_2;
    int reverse;
    if (radius < 0.0f || c->forward.a * c->forward.d - c->forward.b * c->forward.c == 0.0f)
        return;
    if (!c->path.subpaths.n)
        move_to(c, vx, vy);
    point_1 = apply(&c->inverse, c->path.points.v[c->path.points.n - 1]);
    vertex = mkxy(vx, vy);
    point_2 = mkxy(x, y);
    edge_1 = normalized(sub(point_1, vertex));
    edge_2 = normalized(sub(point_2, vertex));
    sine = fabsf(dot(perpendicular(edge_1), edge_2));
    if (sine < 1.0e-4f) {
        line_to(c, vx, vy);
        return;
    }
    offset = scale(radius / sine, add(edge_1, edge_2));
    center = add(vertex, offset);
    angle_1 = direction(sub(scale(dot(offset, edge_1), edge_1), offset));
    angle_2 = direction(sub(scale(dot(offset, edge_2), edge_2), offset));
    reverse = (int)floorf((angle_2 - angle_1) / 3.14159265f) & 1;
    arc(c, center.x, center.y, radius, angle_1, angle_2, reverse);
}

static void fill(cih_oracle *c) /* hpp:3044-3048 */
{
    path_to_lines(c, 0);
    render_main(c, &c->fill_brush, 0);
}

static void stroke(cih_oracle *c) /* hpp:3050-3055 */
{
    int rule = c->evenodd;
    path_to_lines(c, 1);
    stroke_lines(c);
    c->evenodd = 0;
    render_main(c, &c->stroke_brush, 0);
    c->evenodd = rule;
}

static void clip(cih_oracle *c) /* hpp:3057-3099 */
{
    size_t part, index_1, index_2;
    int y = -1;
    float last = 0.0f, sum_1 = 0.0f, sum_2 = 0.0f;
    path_to_lines(c, 0);
    lines_to_runs(c, mkxy(0.0f, 0.0f), 0);
    part = c->runs.n;
    RESERVE(c->runs, c->runs.n + c->mask.n);
    if (c->mask.n)
        memcpy(c->runs.v + c->runs.n, c->mask.v, c->mask.n * sizeof(pixel_run));
    c->runs.n += c->mask.n;
    c->mask.n = 0;
    index_1 = 0;
    index_2 = part;
    while (index_1 < part && index_2 < c->runs.n) {
        int which = run_less(c->runs.v[index_1], c->runs.v[index_2]);
        pixel_run next = which ? c->runs.v[index_1] : c->runs.v[index_2];
        float visibility;
        if (next.y != y) {
            y = next.y;
            last = 0.0f;
            sum_1 = 0.0f;
            sum_2 = 0.0f;
        }
        if (which)
            sum_1 += c->runs.v[index_1++].delta;
        else
            sum_2 += c->runs.v[index_2++].delta;
        visibility = min_std(fabsf(sum_1), 1.0f) * min_std(fabsf(sum_2), 1.0f);
        if (visibility == last)
            continue;
        if (c->mask.n && c->mask.v[c->mask.n - 1].x == next.x &&
            c->mask.v[c->mask.n - 1].y == next.y)
            c->mask.v[c->mask.n - 1].delta += visibility - last;
        else {
            pixel_run piece;
            piece.x = next.x;
            piece.y = next.y;
            piece.delta = visibility - last;
            PUSH(c->mask, piece);
        }
        last = visibility;
    }
}

/* ---- text (hpp:1299-1314, 1526-1846, 3205-3311) --------------------------- */

static int font_u8(const cih_oracle *c, long at) /* hpp:1300-1301 */
{
    return at >= 0 && (size_t)at < c->face.n ? c->face.data[at] : 0;
}
static int font_s8(const cih_oracle *c, long at) { return (int)(signed char)font_u8(c, at); }
static int font_u16(const cih_oracle *c, long at) /* hpp:1305-1307 */
{
    return font_u8(c, at) << 8 | font_u8(c, at + 1);
}
static int font_s16(const cih_oracle *c, long at) { return (int)(short)font_u16(c, at); }
static int font_s32(const cih_oracle *c, long at) /* hpp:1311-1314 */
{
    return (int)((unsigned)font_u8(c, at) << 24 | (unsigned)font_u8(c, at + 1) << 16 |
                 (unsigned)font_u8(c, at + 2) << 8 | (unsigned)font_u8(c, at + 3));
}

static void face_clear(cih_oracle *c)
{
    free(c->face.data);
    c->face.data = NULL;
    c->face.n = 0;
    c->face.cmap = c->face.glyf = c->face.head = c->face.hhea = 0;
    c->face.hmtx = c->face.loca = c->face.maxp = c->face.os_2 = 0;
}

static void face_append(cih_oracle *c, const unsigned char *bytes, size_t n)
{
    c->face.data = (unsigned char *)realloc(c->face.data, c->face.n + n + 1);
    if (!c->face.data)
        abort();
    memcpy(c->face.data + c->face.n, bytes, n);
    c->face.n += n;
}

/* set_font (hpp:3205-3274): keeps the offset table, then appends the eight
 * tables text uses, recording where each landed. */
static int set_font(cih_oracle *c, const unsigned char *font, int bytes, float size)
{
    if (font && bytes) {
        int version, tables, index;
        face_clear(c);
        if (bytes < 6)
            return 0;
        version = (int)((unsigned)font[0] << 24 | (unsigned)font[1] << 16 |
                        (unsigned)font[2] << 8 | (unsigned)font[3]);
        tables = font[4] << 8 | font[5];
        if ((version != 0x00010000 && version != 0x74727565) || bytes < tables * 16 + 12)
            return 0;
        face_append(c, font, (size_t)(tables * 16 + 12));
        for (index = 0; index < tables; ++index) {
            int tag = font_s32(c, index * 16 + 12);
            long offset = font_s32(c, index * 16 + 20);
            long span = font_s32(c, index * 16 + 24);
            int place = (int)c->face.n;
            if (offset < 0 || span < 0 || bytes < offset + span) {
                face_clear(c);
                return 0;
            }
            if (tag == 0x636d6170) c->face.cmap = place;
            else if (tag == 0x676c7966) c->face.glyf = place;
            else if (tag == 0x68656164) c->face.head = place;
            else if (tag == 0x68686561) c->face.hhea = place;
            else if (tag == 0x686d7478) c->face.hmtx = place;
            else if (tag == 0x6c6f6361) c->face.loca = place;
            else if (tag == 0x6d617870) c->face.maxp = place;
            else if (tag == 0x4f532f32) c->face.os_2 = place;
            else continue;
            face_append(c, font + offset, (size_t)span);
        }
        if (!c->face.cmap || !c->face.glyf || !c->face.head || !c->face.hhea ||
            !c->face.hmtx || !c->face.loca || !c->face.maxp || !c->face.os_2) {
            face_clear(c);
            return 0;
        }
    }
    if (!c->face.n)
        return 0;
    c->face.scale = size / (float)font_u16(c, c->face.head + 18);
    return 1;
}

/* character_to_glyph (hpp:1709-1784). */
static int character_to_glyph(const cih_oracle *c, const unsigned char *text, long *index)
{
    unsigned lead = text[*index];
    int bytes = (lead & 0x80) == 0x00 ? 1 : (lead & 0xe0) == 0xc0 ? 2
              : (lead & 0xf0) == 0xe0 ? 3 : (lead & 0xf8) == 0xf0 ? 4 : 0;
    static const int masks[] = {0x0, 0x7f, 0x1f, 0x0f, 0x07};
    int codepoint = bytes ? (int)(lead & (unsigned)masks[bytes]) : 0xfffd;
    int tables, table;
    long format_12 = 0, format_4 = 0, format_0 = 0;
    ++*index;
    while (--bytes > 0) {
        if ((text[*index] & 0xc0) == 0x80)
            codepoint = codepoint << 6 | (text[(*index)++] & 0x3f);
        else {
            codepoint = 0xfffd;
            break;
        }
    }
    if (codepoint == '\t' || codepoint == '\v' || codepoint == '\f' || codepoint == '\r' ||
        codepoint == '\n')
        codepoint = ' ';
    tables = font_u16(c, c->face.cmap + 2);
    for (table = 0; table < tables; ++table) {
        int platform = font_u16(c, c->face.cmap + table * 8 + 4);
        int encoding = font_u16(c, c->face.cmap + table * 8 + 6);
        long offset = font_s32(c, c->face.cmap + table * 8 + 8);
        int format = font_u16(c, c->face.cmap + offset);
        if (platform == 3 && encoding == 10 && format == 12)
            format_12 = c->face.cmap + offset;
        else if (platform == 3 && encoding == 1 && format == 4)
            format_4 = c->face.cmap + offset;
        else if (format == 0)
            format_0 = c->face.cmap + offset;
    }
    if (format_12) {
        long groups = font_s32(c, format_12 + 12), group;
        for (group = 0; group < groups && (size_t)(format_12 + 16 + group * 12) < c->face.n;
             ++group) {
            int start = font_s32(c, format_12 + 16 + group * 12);
            int end = font_s32(c, format_12 + 20 + group * 12);
            int glyph = font_s32(c, format_12 + 24 + group * 12);
            if (start <= codepoint && codepoint <= end)
                return codepoint - start + glyph;
        }
    } else if (format_4) {
        int segments = font_u16(c, format_4 + 6), segment;
        long end_array = format_4 + 14;
        long start_array = end_array + 2 + segments;
        long delta_array = start_array + segments;
        long range_array = delta_array + segments;
        for (segment = 0; segment < segments; segment += 2) {
            int start = font_u16(c, start_array + segment);
            int end = font_u16(c, end_array + segment);
            int delta = font_s16(c, delta_array + segment);
            int range = font_u16(c, range_array + segment);
            if (start <= codepoint && codepoint <= end)
                return range ? font_u16(c, range_array + segment +
                                               (long)(codepoint - start) * 2 + range)
                             : (codepoint + delta) & 0xffff;
        }
    } else if (format_0 && 0 <= codepoint && codepoint < 256)
        return font_u8(c, format_0 + 6 + codepoint);
    return 0;
}

static float measure_text(const cih_oracle *c, const char *text) /* hpp:3297-3311 */
{
    int hmetrics, width = 0;
    long index = 0;
    if (!c->face.n || !text)
        return 0.0f;
    hmetrics = font_u16(c, c->face.hhea + 34);
    while (text[index]) {
        int glyph = character_to_glyph(c, (const unsigned char *)text, &index);
        int entry = glyph < hmetrics - 1 ? glyph : hmetrics - 1; /* std::min */
        width += font_u16(c, c->face.hmtx + (long)entry * 4);
    }
    return (float)width * c->face.scale;
}

/* add_glyph (hpp:1533-1696): the glyph's contours, through c->forward, as
 * closed polylines appended to c->lines. */
static void add_glyph(cih_oracle *c, int glyph, float angular, int depth)
{
    int loc_format = font_u16(c, c->face.head + 50);
    long offset = c->face.glyf + (loc_format ? (long)font_s32(c, c->face.loca + (long)glyph * 4)
                                             : (long)font_u16(c, c->face.loca + (long)glyph * 2) * 2);
    long next = c->face.glyf + (loc_format ? (long)font_s32(c, c->face.loca + (long)glyph * 4 + 4)
                                           : (long)font_u16(c, c->face.loca + (long)glyph * 2 + 2) * 2);
    int contours, hmetrics, left_side_bearing, x_min, points, instructions;
    long flags_array, flags_size = 0, x_size = 0, x_array, y_array, index;
    int x, y, flags = 0, repeated = 0, contour;
    if (offset == next)
        return;
    contours = font_s16(c, offset);
    if (contours < 0) { /* composite: transformed components */
        if (depth >= 16)
            return;
        offset += 10;
        for (;;) {
            int fl = font_u16(c, offset);
            int component = font_u16(c, offset + 2);
            float e, f, a, b, cc, d;
            affine saved_forward, saved_inverse;
            if (!(fl & 2))
                return;
            e = (float)(fl & 1 ? font_s16(c, offset + 4) : font_s8(c, offset + 4));
            f = (float)(fl & 1 ? font_s16(c, offset + 6) : font_s8(c, offset + 5));
            offset += fl & 1 ? 8 : 6;
            a = fl & 200 ? (float)font_s16(c, offset) / 16384.0f : 1.0f;
            b = fl & 128 ? (float)font_s16(c, offset + 2) / 16384.0f : 0.0f;
            cc = fl & 128 ? (float)font_s16(c, offset + 4) / 16384.0f : 0.0f;
            d = fl & 8 ? a
              : fl & 64 ? (float)font_s16(c, offset + 2) / 16384.0f
              : fl & 128 ? (float)font_s16(c, offset + 6) / 16384.0f
              : 1.0f;
            offset += fl & 8 ? 2 : fl & 64 ? 4 : fl & 128 ? 8 : 0;
            saved_forward = c->forward;
            saved_inverse = c->inverse;
            transform(c, a, b, cc, d, e, f);
            add_glyph(c, component, angular, depth + 1);
            c->forward = saved_forward;
            c->inverse = saved_inverse;
            if (!(fl & 32) || (size_t)offset >= c->face.n)
                return;
        }
    }
    hmetrics = font_u16(c, c->face.hhea + 34);
    left_side_bearing = glyph < hmetrics
        ? font_s16(c, c->face.hmtx + (long)glyph * 4 + 2)
        : font_s16(c, c->face.hmtx + (long)hmetrics * 2 + (long)glyph * 2);
    x_min = font_s16(c, offset + 2);
    points = font_u16(c, offset + 8 + (long)contours * 2) + 1;
    instructions = font_u16(c, offset + 10 + (long)contours * 2);
    flags_array = offset + 12 + (long)contours * 2 + instructions;
    for (index = 0; index < points;) {
        int fl, rep;
        if ((size_t)(flags_array + flags_size) >= c->face.n)
            return;
        fl = font_u8(c, flags_array + flags_size++);
        rep = fl & 8 ? font_u8(c, flags_array + flags_size++) + 1 : 1;
        x_size += rep * (fl & 2 ? 1 : fl & 16 ? 0 : 2);
        index += rep;
    }
    x_array = flags_array + flags_size;
    y_array = x_array + x_size;
    x = left_side_bearing - x_min;
    y = 0;
    index = 0;
    for (contour = 0; contour < contours; ++contour) {
        long beginning = index;
        long ending = font_u16(c, offset + 10 + (long)contour * 2);
        xy begin_point = mkxy(0.0f, 0.0f), end_point = mkxy(0.0f, 0.0f);
        int begin_on = 0, end_on = 0;
        size_t first = c->lines.points.n;
        subpath entry;
        for (; index <= ending && index < points; ++index) {
            xy point;
            int on_curve;
            if (repeated)
                --repeated;
            else {
                flags = font_u8(c, flags_array++);
                if (flags & 8)
                    repeated = font_u8(c, flags_array++);
            }
            if (flags & 2)
                x += font_u8(c, x_array) * (flags & 16 ? 1 : -1);
            else if (!(flags & 16))
                x += font_s16(c, x_array);
            if (flags & 4)
                y += font_u8(c, y_array) * (flags & 32 ? 1 : -1);
            else if (!(flags & 32))
                y += font_s16(c, y_array);
            x_array += flags & 2 ? 1 : flags & 16 ? 0 : 2;
            y_array += flags & 4 ? 1 : flags & 32 ? 0 : 2;
            point = apply(&c->forward, mkxy((float)x, (float)y));
            on_curve = flags & 1;
            if (index == beginning) {
                begin_point = point;
                begin_on = on_curve;
                if (on_curve)
                    PUSH(c->lines.points, point);
            } else {
                xy point_2 = on_curve ? point : lerp(end_point, point, 0.5f);
                if (c->lines.points.n == first || (end_on && on_curve))
                    PUSH(c->lines.points, point_2);
                else if (!end_on || on_curve) {
                    xy point_1 = c->lines.points.v[c->lines.points.n - 1];
                    xy control_1 = lerp(point_1, end_point, 2.0f / 3.0f);
                    xy control_2 = lerp(point_2, end_point, 2.0f / 3.0f);
                    add_bezier(c, point_1, control_1, control_2, point_2, angular);
                }
            }
            end_point = point;
            end_on = on_curve;
        }
        if (c->lines.points.n == first)
            continue; /* one off-curve point: no outline */
        if (begin_on ^ end_on) {
            xy point_1 = c->lines.points.v[c->lines.points.n - 1];
            xy point_2 = c->lines.points.v[first];
            xy control = end_on ? begin_point : end_point;
            xy control_1 = lerp(point_1, control, 2.0f / 3.0f);
            xy control_2 = lerp(point_2, control, 2.0f / 3.0f);
            add_bezier(c, point_1, control_1, control_2, point_2, angular);
        } else if (!begin_on && !end_on) {
            xy point_1 = c->lines.points.v[c->lines.points.n - 1];
            xy split = lerp(begin_point, end_point, 0.5f);
            xy point_2 = c->lines.points.v[first];
            xy left_1 = lerp(point_1, end_point, 2.0f / 3.0f);
            xy left_2 = lerp(split, end_point, 2.0f / 3.0f);
            xy right_1 = lerp(split, begin_point, 2.0f / 3.0f);
            xy right_2 = lerp(point_2, begin_point, 2.0f / 3.0f);
            add_bezier(c, point_1, left_1, left_2, split, angular);
            add_bezier(c, split, right_1, right_2, point_2, angular);
        }
        PUSH(c->lines.points, c->lines.points.v[first]);
        entry.count = c->lines.points.n - first;
        entry.closed = 1;
        PUSH(c->lines.subpaths, entry);
    }
}

/* text_to_lines (hpp:1793-1846). */
static void text_to_lines(cih_oracle *c, const char *text, xy position, float maximum_width,
                          int stroking)
{
    const float tolerance = 0.125f;
    float ratio = tolerance / max_std(0.5f * c->line_width, tolerance);
    float angular = stroking ? (ratio - 2.0f) * ratio * 2.0f + 1.0f : -1.0f;
    float width, reduction, units_per_em, ascender, descender, normalize;
    xy scaling;
    affine saved_forward, saved_inverse;
    int hmetrics, place = 0;
    long index = 0;
    c->lines.points.n = 0;
    c->lines.subpaths.n = 0;
    if (!c->face.n || !text || maximum_width <= 0.0f)
        return;
    width = maximum_width == 1.0e30f && c->text_align == 0 ? 0.0f : measure_text(c, text);
    reduction = maximum_width / max_std(maximum_width, width);
    if (c->text_align == 1) /* rightward */
        position.x -= width * reduction;
    else if (c->text_align == 2) /* center */
        position.x -= 0.5f * width * reduction;
    scaling = scale(c->face.scale, mkxy(reduction, 1.0f));
    units_per_em = (float)font_u16(c, c->face.head + 18);
    ascender = (float)font_s16(c, c->face.os_2 + 68);
    descender = (float)font_s16(c, c->face.os_2 + 70);
    normalize = c->face.scale * units_per_em / (ascender - descender);
    if (c->text_baseline == 1) /* top */
        position.y += ascender * normalize;
    else if (c->text_baseline == 2) /* middle */
        position.y += (ascender + descender) * 0.5f * normalize;
    else if (c->text_baseline == 3) /* bottom, ideographic */
        position.y += descender * normalize;
    else if (c->text_baseline == 4) /* hanging */
        position.y += 0.6f * c->face.scale * units_per_em;
    saved_forward = c->forward;
    saved_inverse = c->inverse;
    hmetrics = font_u16(c, c->face.hhea + 34);
    while (text[index]) {
        int glyph = character_to_glyph(c, (const unsigned char *)text, &index);
        int entry = glyph < hmetrics - 1 ? glyph : hmetrics - 1;
        c->forward = saved_forward;
        transform(c, scaling.x, 0.0f, 0.0f, -scaling.y,
                  position.x + (float)place * scaling.x, position.y);
        add_glyph(c, glyph, angular, 0);
        place += font_u16(c, c->face.hmtx + (long)entry * 4);
    }
    c->forward = saved_forward;
    c->inverse = saved_inverse;
}

static void fill_text(cih_oracle *c, const char *text, float x, float y,
                      float maximum_width) /* hpp:3276-3284 */
{
    int rule = c->evenodd;
    text_to_lines(c, text, mkxy(x, y), maximum_width, 0);
    c->evenodd = 0;
    render_main(c, &c->fill_brush, 0);
    c->evenodd = rule;
}

static void stroke_text(cih_oracle *c, const char *text, float x, float y,
                        float maximum_width) /* hpp:3286-3295 */
{
    int rule = c->evenodd;
    text_to_lines(c, text, mkxy(x, y), maximum_width, 1);
    stroke_lines(c);
    c->evenodd = 0;
    render_main(c, &c->stroke_brush, 0);
    c->evenodd = rule;
}

static void fill_rectangle(cih_oracle *c, float x, float y, float w, float h) /* hpp:3155-3172 */
{
    subpath s;
    if (w == 0.0f || h == 0.0f)
        return;
    c->lines.points.n = 0;
    c->lines.subpaths.n = 0;
    PUSH(c->lines.points, apply(&c->forward, mkxy(x, y)));
    PUSH(c->lines.points, apply(&c->forward, mkxy(x + w, y)));
    PUSH(c->lines.points, apply(&c->forward, mkxy(x + w, y + h)));
    PUSH(c->lines.points, apply(&c->forward, mkxy(x, y + h)));
    s.count = 4;
    s.closed = 1;
    PUSH(c->lines.subpaths, s);
    {
        int rule = c->evenodd;
        c->evenodd = 0;
        render_main(c, &c->fill_brush, 0);
        c->evenodd = rule;
    }
}

static void clear_rectangle(cih_oracle *c, float x, float y, float w, float h) /* hpp:3134-3153 */
{
    int saved_operation = c->op, saved_type = c->fill_brush.type;
    float saved_global_alpha = c->global_alpha, saved_alpha = c->shadow_color.a;
    c->op = OP_DESTINATION_OUT;
    c->global_alpha = 1.0f;
    c->shadow_color.a = 0.0f;
    c->fill_brush.type = 0;
    fill_rectangle(c, x, y, w, h);
    c->fill_brush.type = saved_type;
    c->shadow_color.a = saved_alpha;
    c->global_alpha = saved_global_alpha;
    c->op = saved_operation;
}

static void stroke_rectangle(cih_oracle *c, float x, float y, float w, float h) /* hpp:3174-3203 */
{
    subpath s;
    int rule = c->evenodd;
    if (w == 0.0f && h == 0.0f)
        return;
    c->lines.points.n = 0;
    c->lines.subpaths.n = 0;
    if (w == 0.0f || h == 0.0f) {
        PUSH(c->lines.points, apply(&c->forward, mkxy(x, y)));
        PUSH(c->lines.points, apply(&c->forward, mkxy(x + w, y + h)));
        s.count = 2;
        s.closed = 0;
    } else {
        PUSH(c->lines.points, apply(&c->forward, mkxy(x, y)));
        PUSH(c->lines.points, apply(&c->forward, mkxy(x + w, y)));
        PUSH(c->lines.points, apply(&c->forward, mkxy(x + w, y + h)));
        PUSH(c->lines.points, apply(&c->forward, mkxy(x, y + h)));
        PUSH(c->lines.points, apply(&c->forward, mkxy(x, y)));
        s.count = 5;
        s.closed = 1;
    }
    PUSH(c->lines.subpaths, s);
    stroke_lines(c);
    c->evenodd = 0;
    render_main(c, &c->stroke_brush, 0);
    c->evenodd = rule;
}

static void draw_image(cih_oracle *c, const unsigned char *image, int width, int height,
                       int stride, float x, float y, float to_width,
                       float to_height) /* hpp:3313-3346 */
{
    subpath s;
    affine saved_forward, saved_inverse;
    int rule = c->evenodd;
    if (!image || width <= 0 || height <= 0 || to_width == 0.0f || to_height == 0.0f)
        return;
    set_pattern(&c->image_brush, image, width, height, stride, 0);
    c->lines.points.n = 0;
    c->lines.subpaths.n = 0;
    PUSH(c->lines.points, apply(&c->forward, mkxy(x, y)));
    PUSH(c->lines.points, apply(&c->forward, mkxy(x + to_width, y)));
    PUSH(c->lines.points, apply(&c->forward, mkxy(x + to_width, y + to_height)));
    PUSH(c->lines.points, apply(&c->forward, mkxy(x, y + to_height)));
    s.count = 4;
    s.closed = 1;
    PUSH(c->lines.subpaths, s);
    saved_forward = c->forward;
    saved_inverse = c->inverse;
    transform(c, 1.0f, 0.0f, 0.0f, 1.0f, x + min_std(0.0f, to_width), y + min_std(0.0f, to_height));
    transform(c, fabsf(to_width) / (float)width, 0.0f, 0.0f, fabsf(to_height) / (float)height,
              0.0f, 0.0f);
    c->evenodd = 0;
    render_main(c, &c->image_brush, 1);
    c->evenodd = rule;
    c->forward = saved_forward;
    c->inverse = saved_inverse;
}

static void init_state(cih_oracle *c, int width, int height) /* hpp:2607-2644 */
{
    affine identity = {1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f};
    memset(c, 0, sizeof(*c));
    c->op = OP_SOURCE_OVER;
    c->line_cap = CAP_BUTT;
    c->line_join = JOIN_MITER;
    c->size_x = width;
    c->size_y = height;
    c->global_alpha = 1.0f;
    c->line_width = 1.0f;
    c->miter_limit = 10.0f;
    c->forward = identity;
    c->inverse = identity;
    set_color(c, 0, 0.0f, 0.0f, 0.0f, 1.0f);
    set_color(c, 1, 0.0f, 0.0f, 0.0f, 1.0f);
}

cih_oracle *cih_oracle_create(int width, int height)
{
    cih_oracle *c;
    int y;
    if (width < 1 || height < 1 || width > 32768 || height > 32768)
        return NULL;
    c = (cih_oracle *)malloc(sizeof(cih_oracle));
    if (!c)
        return NULL;
    init_state(c, width, height);
    c->bitmap = (rgba *)calloc((size_t)width * (size_t)height, sizeof(rgba));
    if (!c->bitmap) {
        free(c);
        return NULL;
    }
    for (y = 0; y < height; ++y) { /* full clip mask, hpp:2636-2643 */
        pixel_run a, b;
        a.x = 0, a.y = (unsigned short)y, a.delta = 1.0f;
        b.x = (unsigned short)width, b.y = (unsigned short)y, b.delta = -1.0f;
        PUSH(c->mask, a);
        PUSH(c->mask, b);
    }
    return c;
}

static void free_state(cih_oracle *c)
{
    free(c->shadow.v);
    free(c->line_dash.v);
    brush_free(&c->fill_brush);
    brush_free(&c->stroke_brush);
    brush_free(&c->image_brush);
    free(c->path.points.v);
    free(c->path.subpaths.v);
    free(c->lines.points.v);
    free(c->lines.subpaths.v);
    free(c->scratch.points.v);
    free(c->scratch.subpaths.v);
    free(c->runs.v);
    free(c->mask.v);
    free(c->bitmap);
    free(c->face.data);
}

void cih_oracle_destroy(cih_oracle *c)
{
    if (!c)
        return;
    while (c->saves) {
        cih_oracle *head = c->saves;
        c->saves = head->saves;
        free_state(head);
        free(head);
    }
    free_state(c);
    free(c);
}

static void copy_saved(cih_oracle *to, const cih_oracle *from) /* hpp:3410-3465 */
{
    to->op = from->op;
    to->shadow_offset_x = from->shadow_offset_x;
    to->shadow_offset_y = from->shadow_offset_y;
    to->line_cap = from->line_cap;
    to->line_join = from->line_join;
    to->line_dash_offset = from->line_dash_offset;
    to->forward = from->forward;
    to->inverse = from->inverse;
    to->global_alpha = from->global_alpha;
    to->shadow_color = from->shadow_color;
    to->shadow_blur = from->shadow_blur;
    to->line_width = from->line_width;
    to->miter_limit = from->miter_limit;
    to->line_dash.n = 0;
    RESERVE(to->line_dash, from->line_dash.n + 1);
    if (from->line_dash.n)
        memcpy(to->line_dash.v, from->line_dash.v, from->line_dash.n * sizeof(float));
    to->line_dash.n = from->line_dash.n;
    brush_copy(&to->fill_brush, &from->fill_brush);
    brush_copy(&to->stroke_brush, &from->stroke_brush);
    to->mask.n = 0;
    RESERVE(to->mask, from->mask.n + 1);
    if (from->mask.n)
        memcpy(to->mask.v, from->mask.v, from->mask.n * sizeof(pixel_run));
    to->mask.n = from->mask.n;
    to->evenodd = from->evenodd;
    to->text_align = from->text_align;
    to->text_baseline = from->text_baseline;
    free(to->face.data);
    to->face = from->face;
    to->face.data = NULL;
    if (from->face.n) {
        to->face.data = (unsigned char *)malloc(from->face.n);
        if (!to->face.data)
            abort();
        memcpy(to->face.data, from->face.data, from->face.n);
    }
}

static void save(cih_oracle *c)
{
    cih_oracle *state = (cih_oracle *)malloc(sizeof(cih_oracle));
    if (!state)
        abort();
    init_state(state, 0, 0);
    copy_saved(state, c);
    state->saves = c->saves;
    c->saves = state;
}

static void restore(cih_oracle *c)
{
    cih_oracle *state = c->saves;
    if (!state)
        return;
    copy_saved(c, state);
    c->saves = state->saves;
    free_state(state);
    free(state);
}

void cih_oracle_get_image_data(cih_oracle *c, unsigned char *image, int width, int height,
                               int stride, int x, int y) /* hpp:3348-3381 */
{
    static const float bayer[4][4] = {{0.03125f, 0.53125f, 0.15625f, 0.65625f},
                                      {0.78125f, 0.28125f, 0.90625f, 0.40625f},
                                      {0.21875f, 0.71875f, 0.09375f, 0.59375f},
                                      {0.96875f, 0.46875f, 0.84375f, 0.34375f}};
    int iy, ix;
    if (!image)
        return;
    for (iy = 0; iy < height; ++iy)
        for (ix = 0; ix < width; ++ix) {
            int index = iy * stride + ix * 4;
            int cx = x + ix, cy = y + iy;
            rgba color = mkrgba(0.0f, 0.0f, 0.0f, 0.0f);
            float t;
            if (0 <= cx && cx < c->size_x && 0 <= cy && cy < c->size_y)
                color = c->bitmap[cy * c->size_x + cx];
            t = bayer[cy & 3][cx & 3];
            color = rgba_add(mkrgba(t, t, t, t),
                             rgba_scale(255.0f, delinearized(clamped(unpremultiplied(color)))));
            image[index + 0] = (unsigned char)color.r;
            image[index + 1] = (unsigned char)color.g;
            image[index + 2] = (unsigned char)color.b;
            image[index + 3] = (unsigned char)color.a;
        }
}

void cih_oracle_put_image_data(cih_oracle *c, const unsigned char *image, int width, int height,
                               int stride, int x, int y) /* hpp:3383-3408 */
{
    int iy, ix;
    if (!image)
        return;
    for (iy = 0; iy < height; ++iy)
        for (ix = 0; ix < width; ++ix) {
            int index = iy * stride + ix * 4;
            int cx = x + ix, cy = y + iy;
            rgba color;
            if (cx < 0 || c->size_x <= cx || cy < 0 || c->size_y <= cy)
                continue;
            color = mkrgba(image[index + 0] / 255.0f, image[index + 1] / 255.0f,
                           image[index + 2] / 255.0f, image[index + 3] / 255.0f);
            c->bitmap[cy * c->size_x + cx] = premultiplied(linearized(color));
        }
}

int cih_oracle_is_point_in_path(cih_oracle *c, float x, float y) /* hpp:3101-3132 */
{
    size_t index, subpath = 0, beginning = 0, ending = 0;
    int winding = 0;
    path_to_lines(c, 0);
    for (index = 0; index < c->lines.points.n; ++index) {
        xy from, to;
        while (index >= ending) {
            beginning = ending;
            ending += c->lines.subpaths.v[subpath++].count;
        }
        from = c->lines.points.v[index];
        to = c->lines.points.v[index + 1 < ending ? index + 1 : beginning];
        if ((from.y < y && y <= to.y) || (to.y < y && y <= from.y)) {
            float side = dot(perpendicular(sub(to, from)), sub(mkxy(x, y), from));
            if (side == 0.0f)
                return 1;
            winding += side > 0.0f ? 1 : -1;
        } else if (from.y == y && y == to.y &&
                   ((from.x <= x && x <= to.x) || (to.x <= x && x <= from.x)))
            return 1;
    }
    return winding != 0;
}

float cih_oracle_measure_text(const unsigned char *font, size_t bytes, float size,
                              const char *text)
{
    cih_oracle state;
    float width;
    init_state(&state, 0, 0);
    set_font(&state, font, (int)bytes, size);
    width = measure_text(&state, text);
    free_state(&state);
    return width;
}

void cih_oracle_read_linear(cih_oracle *c, float *out)
{
    memcpy(out, c->bitmap, sizeof(rgba) * (size_t)c->size_x * (size_t)c->size_y);
}

/* ---- command stream replay (include/cih_stream.h) ------------------------ */

typedef struct { const unsigned char *texels; int width, height; } stream_image;

static float fa(const uint32_t *a, int i) { return cih_w2f(a[i]); }
static int ia(const uint32_t *a, int i) { return (int)(int32_t)a[i]; }

int cih_oracle_replay(cih_oracle *c, const void *stream, size_t bytes)
{
    cih_reader reader;
    cih_record r;
    int status;
    stream_image *images;
    if (cih_reader_init(&reader, stream, bytes) != 0)
        return -1;
    images = (stream_image *)calloc(CIH_STREAM_MAX_IMAGES, sizeof(stream_image));
    if (!images)
        return -1;
    while ((status = cih_reader_next(&reader, &r)) == 1) {
        const uint32_t *a = r.args;
        switch (r.op) {
        case CIH_OP_SCALE: transform(c, fa(a, 0), 0.0f, 0.0f, fa(a, 1), 0.0f, 0.0f); break;
        case CIH_OP_ROTATE: {
            float cosine = cosf(fa(a, 0)), sine = sinf(fa(a, 0));
            transform(c, cosine, sine, -sine, cosine, 0.0f, 0.0f);
            break;
        }
        case CIH_OP_TRANSLATE: transform(c, 1.0f, 0.0f, 0.0f, 1.0f, fa(a, 0), fa(a, 1)); break;
        case CIH_OP_TRANSFORM:
            transform(c, fa(a, 0), fa(a, 1), fa(a, 2), fa(a, 3), fa(a, 4), fa(a, 5));
            break;
        case CIH_OP_SET_TRANSFORM:
            set_transform(c, fa(a, 0), fa(a, 1), fa(a, 2), fa(a, 3), fa(a, 4), fa(a, 5));
            break;
        case CIH_OP_SET_GLOBAL_ALPHA:
            if (0.0f <= fa(a, 0) && fa(a, 0) <= 1.0f)
                c->global_alpha = fa(a, 0);
            break;
        case CIH_OP_SET_COMPOSITE_OP: c->op = ia(a, 0); break;
        case CIH_OP_SET_SHADOW_COLOR:
            c->shadow_color =
                premultiplied(linearized(clamped(mkrgba(fa(a, 0), fa(a, 1), fa(a, 2), fa(a, 3)))));
            break;
        case CIH_OP_SET_SHADOW_BLUR:
            if (0.0f <= fa(a, 0))
                c->shadow_blur = fa(a, 0);
            break;
        case CIH_OP_SET_SHADOW_OFFSET:
            c->shadow_offset_x = fa(a, 0);
            c->shadow_offset_y = fa(a, 1);
            break;
        case CIH_OP_SET_LINE_WIDTH:
            if (0.0f < fa(a, 0))
                c->line_width = fa(a, 0);
            break;
        case CIH_OP_SET_LINE_CAP: c->line_cap = ia(a, 0); break;
        case CIH_OP_SET_LINE_JOIN: c->line_join = ia(a, 0); break;
        case CIH_OP_SET_MITER_LIMIT:
            if (0.0f < fa(a, 0))
                c->miter_limit = fa(a, 0);
            break;
        case CIH_OP_SET_LINE_DASH: { /* hpp:2752-2766 */
            int count = ia(a, 0), k, bad = 0;
            for (k = 0; k < count; ++k)
                if (fa(a, 1 + k) < 0.0f)
                    bad = 1;
            if (bad)
                break;
            c->line_dash.n = 0;
            if (count < 0)
                break;
            for (k = 0; k < count; ++k)
                PUSH(c->line_dash, fa(a, 1 + k));
            if (count & 1)
                for (k = 0; k < count; ++k)
                    PUSH(c->line_dash, fa(a, 1 + k));
            break;
        }
        case CIH_OP_SET_LINE_DASH_OFFSET: c->line_dash_offset = fa(a, 0); break;
        case CIH_OP_SET_COLOR: set_color(c, ia(a, 0), fa(a, 1), fa(a, 2), fa(a, 3), fa(a, 4)); break;
        case CIH_OP_SET_LINEAR_GRADIENT: { /* hpp:2781-2794 */
            brush *b = ia(a, 0) == 0 ? &c->fill_brush : &c->stroke_brush;
            b->type = 1;
            b->colors.n = 0;
            b->stops.n = 0;
            b->start = mkxy(fa(a, 1), fa(a, 2));
            b->end = mkxy(fa(a, 3), fa(a, 4));
            break;
        }
        case CIH_OP_SET_RADIAL_GRADIENT: { /* hpp:2796-2815 */
            brush *b = ia(a, 0) == 0 ? &c->fill_brush : &c->stroke_brush;
            if (fa(a, 3) < 0.0f || fa(a, 6) < 0.0f)
                break;
            b->type = 2;
            b->colors.n = 0;
            b->stops.n = 0;
            b->start = mkxy(fa(a, 1), fa(a, 2));
            b->end = mkxy(fa(a, 4), fa(a, 5));
            b->start_radius = fa(a, 3);
            b->end_radius = fa(a, 6);
            break;
        }
        case CIH_OP_ADD_COLOR_STOP: { /* hpp:2817-2837 */
            brush *b = ia(a, 0) == 0 ? &c->fill_brush : &c->stroke_brush;
            float offset = fa(a, 1);
            size_t index = 0;
            rgba color;
            if ((b->type != 1 && b->type != 2) || offset < 0.0f || 1.0f < offset)
                break;
            while (index < b->stops.n && !(offset < b->stops.v[index]))
                ++index;
            color = linearized(clamped(mkrgba(fa(a, 2), fa(a, 3), fa(a, 4), fa(a, 5))));
            RESERVE(b->colors, b->colors.n + 1);
            RESERVE(b->stops, b->stops.n + 1);
            memmove(b->colors.v + index + 1, b->colors.v + index,
                    (b->colors.n - index) * sizeof(rgba));
            memmove(b->stops.v + index + 1, b->stops.v + index,
                    (b->stops.n - index) * sizeof(float));
            b->colors.v[index] = color;
            b->stops.v[index] = offset;
            b->colors.n++;
            b->stops.n++;
            break;
        }
        case CIH_OP_SET_PATTERN: {
            int id = ia(a, 1);
            if (id < 0 || id >= CIH_STREAM_MAX_IMAGES || !images[id].texels)
                goto malformed;
            set_pattern(ia(a, 0) == 0 ? &c->fill_brush : &c->stroke_brush, images[id].texels,
                        images[id].width, images[id].height, images[id].width * 4, ia(a, 2));
            break;
        }
        case CIH_OP_BEGIN_PATH:
            c->path.points.n = 0;
            c->path.subpaths.n = 0;
            break;
        case CIH_OP_MOVE_TO: move_to(c, fa(a, 0), fa(a, 1)); break;
        case CIH_OP_CLOSE_PATH: close_path(c); break;
        case CIH_OP_LINE_TO: line_to(c, fa(a, 0), fa(a, 1)); break;
        case CIH_OP_QUADRATIC_CURVE_TO:
            quadratic_curve_to(c, fa(a, 0), fa(a, 1), fa(a, 2), fa(a, 3));
            break;
        case CIH_OP_BEZIER_CURVE_TO:
            bezier_curve_to(c, fa(a, 0), fa(a, 1), fa(a, 2), fa(a, 3), fa(a, 4), fa(a, 5));
            break;
        case CIH_OP_ARC_TO: arc_to(c, fa(a, 0), fa(a, 1), fa(a, 2), fa(a, 3), fa(a, 4)); break;
        case CIH_OP_ARC: arc(c, fa(a, 0), fa(a, 1), fa(a, 2), fa(a, 3), fa(a, 4), ia(a, 5) != 0); break;
        case CIH_OP_RECTANGLE: { /* hpp:3031-3042 */
            float x = fa(a, 0), y = fa(a, 1), w = fa(a, 2), h = fa(a, 3);
            move_to(c, x, y);
            line_to(c, x + w, y);
            line_to(c, x + w, y + h);
            line_to(c, x, y + h);
            close_path(c);
            break;
        }
        case CIH_OP_FILL: fill(c); break;
        case CIH_OP_STROKE: stroke(c); break;
        case CIH_OP_CLIP: clip(c); break;
        case CIH_OP_CLEAR_RECTANGLE: clear_rectangle(c, fa(a, 0), fa(a, 1), fa(a, 2), fa(a, 3)); break;
        case CIH_OP_FILL_RECTANGLE: fill_rectangle(c, fa(a, 0), fa(a, 1), fa(a, 2), fa(a, 3)); break;
        case CIH_OP_STROKE_RECTANGLE:
            stroke_rectangle(c, fa(a, 0), fa(a, 1), fa(a, 2), fa(a, 3));
            break;
        case CIH_OP_DRAW_IMAGE: {
            int id = ia(a, 0);
            if (id < 0 || id >= CIH_STREAM_MAX_IMAGES || !images[id].texels)
                goto malformed;
            draw_image(c, images[id].texels, images[id].width, images[id].height,
                       images[id].width * 4, fa(a, 1), fa(a, 2), fa(a, 3), fa(a, 4));
            break;
        }
        case CIH_OP_SAVE: save(c); break;
        case CIH_OP_RESTORE: restore(c); break;
        case CIH_OP_DEFINE_IMAGE: {
            int id = ia(a, 0);
            images[id].texels = (const unsigned char *)(a + 3);
            images[id].width = ia(a, 1);
            images[id].height = ia(a, 2);
            break;
        }
        case CIH_OP_PUT_IMAGE_DATA: {
            int id = ia(a, 0);
            if (id < 0 || id >= CIH_STREAM_MAX_IMAGES || !images[id].texels)
                goto malformed;
            cih_oracle_put_image_data(c, images[id].texels, images[id].width, images[id].height,
                                      images[id].width * 4, ia(a, 1), ia(a, 2));
            break;
        }
        case CIH_OP_SET_FILL_RULE: c->evenodd = ia(a, 0) == 1; break;
        case CIH_OP_SET_FONT: {
            int n = ia(a, 0);
            set_font(c, n ? (const unsigned char *)(a + 2) : NULL, n, fa(a, 1));
            break;
        }
        case CIH_OP_FILL_TEXT:
        case CIH_OP_STROKE_TEXT: {
            size_t n = (size_t)ia(a, 0);
            char *text = (char *)malloc(n + 1);
            if (!text)
                goto malformed;
            memcpy(text, a + 4, n);
            text[n] = 0;
            if (r.op == CIH_OP_FILL_TEXT)
                fill_text(c, text, fa(a, 1), fa(a, 2), fa(a, 3));
            else
                stroke_text(c, text, fa(a, 1), fa(a, 2), fa(a, 3));
            free(text);
            break;
        }
        case CIH_OP_SET_TEXT_ALIGN: c->text_align = ia(a, 0); break;
        case CIH_OP_SET_TEXT_BASELINE: c->text_baseline = ia(a, 0); break;
        default: goto malformed;
        }
    }
    free(images);
    return status < 0 ? -1 : 0;
malformed:
    free(images);
    return -1;
}

/* ---- stage seams ---------------------------------------------------------
 * The intermediate data of this restatement's stages, in the layout of the
 * reference build's seams (oracle/ref_shim.cpp), so tests can compare the
 * GPU's polygons and runs with them where the reference build is absent. */

static long copy_lines(const cih_oracle *c, float *points, long capacity, int *subpath_info,
                       long subpath_capacity, long *subpaths)
{
    long n = (long)c->lines.points.n, s = (long)c->lines.subpaths.n, k;
    for (k = 0; k < n && k < capacity; ++k) {
        points[2 * k + 0] = c->lines.points.v[k].x;
        points[2 * k + 1] = c->lines.points.v[k].y;
    }
    for (k = 0; k < s && k < subpath_capacity; ++k) {
        subpath_info[2 * k + 0] = (int)c->lines.subpaths.v[k].count;
        subpath_info[2 * k + 1] = c->lines.subpaths.v[k].closed ? 1 : 0;
    }
    *subpaths = s;
    return n;
}

/* path_to_lines (hpp:1495) of the current path and, when stroking,
 * stroke_lines (hpp:2070), which dashes first. */
long cih_oracle_path_lines(cih_oracle *c, int stroking, float *points, long capacity,
                           int *subpath_info, long subpath_capacity, long *subpaths)
{
    path_to_lines(c, stroking != 0);
    if (stroking)
        stroke_lines(c);
    return copy_lines(c, points, capacity, subpath_info, subpath_capacity, subpaths);
}

/* path_to_lines(true) (hpp:1495) and dash_lines (hpp:1858), unstroked. */
long cih_oracle_dash_lines(cih_oracle *c, float *points, long capacity, int *subpath_info,
                           long subpath_capacity, long *subpaths)
{
    path_to_lines(c, 1);
    dash_lines(c);
    return copy_lines(c, points, capacity, subpath_info, subpath_capacity, subpaths);
}

/* lines_to_runs (hpp:2193) of the current lines: (x, y, delta) triples. */
long cih_oracle_lines_to_runs(cih_oracle *c, float offset_x, float offset_y, int padding,
                              float *runs, long capacity)
{
    long n, k;
    lines_to_runs(c, mkxy(offset_x, offset_y), padding);
    n = (long)c->runs.n;
    for (k = 0; k < n && k < capacity; ++k) {
        runs[3 * k + 0] = (float)c->runs.v[k].x;
        runs[3 * k + 1] = (float)c->runs.v[k].y;
        runs[3 * k + 2] = c->runs.v[k].delta;
    }
    return n;
}
