// cih_geometry.hpp -- per-thread geometry stages of the rasteriser.
//
// Flattening (canvas_ity.hpp:1331-1524), dashing (hpp:1858-1934), stroke
// expansion (hpp:1949-2100), polygon clipping (hpp:2205-2238) and signed-area
// scan conversion (hpp:2109-2170), written as pure functions over "sinks" so
// one body serves a counting pass and an emitting pass.  The kernels in
// cih_device.hip run one thread per cubic, per subpath or per edge.
//
// Every flattening, dashing and stroking decision is a threshold test on float
// arithmetic, so each expression keeps the reference's operand order and
// association, no multiply-add is contracted (the library is built with
// -ffp-contract=off), and min/max keep std::min/std::max's NaN behaviour.
#ifndef CIH_GEOMETRY_HPP
#define CIH_GEOMETRY_HPP

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define CIH_HD __host__ __device__ inline
#else
#define CIH_HD inline
#endif

namespace cih {

struct f2 {
    float x, y;
};

CIH_HD f2 mk(float x, float y) {
    f2 v;
    v.x = x;
    v.y = y;
    return v;
}
CIH_HD f2 add(f2 a, f2 b) { return mk(a.x + b.x, a.y + b.y); }
CIH_HD f2 sub(f2 a, f2 b) { return mk(a.x - b.x, a.y - b.y); }
CIH_HD f2 mul(float s, f2 v) { return mk(v.x * s, v.y * s); }
CIH_HD float dot(f2 a, f2 b) { return a.x * b.x + a.y * b.y; }
CIH_HD float len(f2 v) { return sqrtf(dot(v, v)); }
CIH_HD float smax(float a, float b) { return (a < b) ? b : a; } // std::max
CIH_HD float smin(float a, float b) { return (b < a) ? b : a; } // std::min
CIH_HD f2 normalized(f2 v) { return mul(1.0f / smax(1.0e-6f, len(v)), v); }
CIH_HD f2 perp(f2 v) { return mk(-v.y, v.x); }
CIH_HD f2 lerp(f2 from, f2 to, float ratio) { return add(from, mul(ratio, sub(to, from))); }
// acosf/tanf of round joins (hpp:1994-1997), evaluated as glibc's libm
// evaluates them, so the join's cubic matches the reference's bit for bit on
// host and device alike (a correctly rounded acos or tan differs from glibc's
// on about 0.25% and 0.07% of these inputs, by an ulp, which moves the arc).
// Both follow the float versions of fdlibm's e_acos.c and k_tan.c, which are
// glibc's acosf and, for |x| <= pi/4, tanf:
//   Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
//   Developed at SunPro, a Sun Microsystems, Inc. business.
//   Permission to use, copy, modify, and distribute this software is freely
//   granted, provided that this notice is preserved.
// tests/test_libm_parity.py compares them with the host libm's acosf over all
// of [-1, 1] and tanf over all of [0, pi/4] (in strides, or every input).
CIH_HD uint32_t float_bits(float v) {
    uint32_t u;
    __builtin_memcpy(&u, &v, sizeof(u));
    return u;
}
CIH_HD float bits_float(uint32_t u) {
    float v;
    __builtin_memcpy(&v, &u, sizeof(v));
    return v;
}
// R(z) = p(z)/q(z), the rational approximation of (asin(s) - s)/s at z = s*s.
CIH_HD float acos_ratio(float z) {
    const float p0 = 0.166666672f, p1 = -0.325565815f, p2 = 0.201212525f,
                p3 = -0.0400555357f, p4 = 0.000791535014f, p5 = 3.47933092e-05f;
    const float q1 = -2.40339494f, q2 = 2.02094579f, q3 = -0.68828398f,
                q4 = 0.077038154f;
    float p = z * (p0 + z * (p1 + z * (p2 + z * (p3 + z * (p4 + z * p5)))));
    float q = 1.0f + z * (q1 + z * (q2 + z * (q3 + z * q4)));
    return p / q;
}
CIH_HD float join_acos(float x) {
    const float pi = 3.1415925f, pio2_hi = 1.57079625f, pio2_lo = 7.54978942e-08f;
    uint32_t hx = float_bits(x), ix = hx & 0x7fffffffu;
    if (ix == 0x3f800000u) // |x| = 1
        return (hx >> 31) ? pi + 2.0f * pio2_lo : 0.0f;
    if (ix > 0x3f800000u) // |x| > 1 or NaN
        return (x - x) / (x - x);
    if (ix < 0x3f000000u) { // |x| < 1/2: pi/2 - (x + x R(x^2))
        if (ix <= 0x32800000u)
            return pio2_hi + pio2_lo;
        float r = acos_ratio(x * x);
        return pio2_hi - (x - (pio2_lo - x * r));
    }
    if (hx >> 31) { // x <= -1/2: pi - 2 asin(sqrt((1 + x)/2))
        float z = (1.0f + x) * 0.5f;
        float s = sqrtf(z);
        float w = acos_ratio(z) * s - pio2_lo;
        return pi - 2.0f * (s + w);
    }
    // x >= 1/2: 2 asin(sqrt((1 - x)/2)), the root split into a head with a
    // short mantissa and a correction
    float z = (1.0f - x) * 0.5f;
    float s = sqrtf(z);
    float df = bits_float(float_bits(s) & 0xfffff000u);
    float c = (z - df * df) / (s + df);
    float w = acos_ratio(z) * s + c;
    return 2.0f * (df + w);
}
// tan(x) for 0 <= x <= float(pi)/4 (the join's quarter angle): an odd
// polynomial in x, and for x >= 0.6744 the same on pi/4 - x through
// tan(pi/4 - y) = (1 - tan y)/(1 + tan y).  At float(pi)/4, which lies above
// pi/4, glibc's tanf reduces the argument and returns exactly 1.
CIH_HD float join_tan(float x) {
    const float pio4 = 0.785398126f, pio4lo = 3.77489471e-08f;
    const float t0 = 0.333333343f, t1 = 0.13333334f, t2 = 0.0539682545f,
                t3 = 0.0218694881f, t4 = 0.00886323955f, t5 = 0.00359207904f,
                t6 = 0.00145620946f, t7 = 0.000588041265f, t8 = 0.00024646314f,
                t9 = 7.81794442e-05f, t10 = 7.14072521e-05f, t11 = -1.85586377e-05f,
                t12 = 2.59073058e-05f;
    uint32_t ix = float_bits(x) & 0x7fffffffu;
    if (ix == 0x3f490fdbu)
        return 1.0f;
    if (ix < 0x39000000u) // |x| < 2^-13
        return x;
    bool big = ix >= 0x3f2ca140u;
    float y = 0.0f;
    if (big) {
        float z = pio4 - x;
        float w = pio4lo - y;
        x = z + w;
        y = 0.0f;
    }
    float z = x * x;
    float w = z * z;
    float r = t1 + w * (t3 + w * (t5 + w * (t7 + w * (t9 + w * t11))));
    float v = z * (t2 + w * (t4 + w * (t6 + w * (t8 + w * (t10 + w * t12)))));
    float s = z * x;
    r = y + z * (s * (r + v) + y);
    r += t0 * s;
    w = x + r;
    if (big)
        return 1.0f - 2.0f * (x - (w * w / (w + 1.0f) - r));
    return w;
}

// Affine a,b,c,d,e,f applied as (a*x + c*y) + e, (b*x + d*y) + f.
CIH_HD f2 apply(const float *m, f2 v) {
    return mk(m[0] * v.x + m[2] * v.y + m[4], m[1] * v.x + m[3] * v.y + m[5]);
}

// ---------------------------------------------------------------------------
// Flattening.  A point sink needs `void point(f2)`.

// Adaptive subdivision of one cubic (hpp:1331-1387), iterative and
// depth-first so leaves come out in the recursion's order.  Flatness 0.125 px,
// at most 20 levels; `angular` > -1 adds the turn-cosine test and makes leaves
// emit their control points (stroking).
template <class Sink>
CIH_HD void tessellate(Sink &sink, f2 p1, f2 c1, f2 c2, f2 p2, float angular) {
    struct node {
        f2 p1, c1, c2, p2;
        int limit;
    };
    node stack[22];
    int top = 0;
    stack[top].p1 = p1, stack[top].c1 = c1, stack[top].c2 = c2, stack[top].p2 = p2;
    stack[top++].limit = 20;
    const float flatness = 0.125f * 0.125f;
    while (top) {
        node n = stack[--top];
        f2 edge_1 = sub(n.c1, n.p1);
        f2 edge_2 = sub(n.c2, n.c1);
        f2 edge_3 = sub(n.p2, n.c2);
        f2 segment = sub(n.p2, n.p1);
        float squared_1 = dot(edge_1, edge_1);
        float squared_2 = dot(edge_2, edge_2);
        float squared_3 = dot(edge_3, edge_3);
        float length_squared = smax(1.0e-4f, dot(segment, segment));
        float projection_1 = dot(edge_1, segment) / length_squared;
        float projection_2 = dot(edge_3, segment) / length_squared;
        float clamped_1 = smin(smax(projection_1, 0.0f), 1.0f);
        float clamped_2 = smin(smax(projection_2, 0.0f), 1.0f);
        f2 to_line_1 = sub(add(n.p1, mul(clamped_1, segment)), n.c1);
        f2 to_line_2 = sub(sub(n.p2, mul(clamped_2, segment)), n.c2);
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
             cosine >= angular) ||
            !n.limit) {
            if (angular > -1.0f && squared_1 != 0.0f)
                sink.point(n.c1);
            if (angular > -1.0f && squared_2 != 0.0f)
                sink.point(n.c2);
            if (angular == -1.0f || squared_3 != 0.0f)
                sink.point(n.p2);
            continue;
        }
        f2 left_1 = lerp(n.p1, n.c1, 0.5f);
        f2 middle = lerp(n.c1, n.c2, 0.5f);
        f2 right_2 = lerp(n.c2, n.p2, 0.5f);
        f2 left_2 = lerp(left_1, middle, 0.5f);
        f2 right_1 = lerp(middle, right_2, 0.5f);
        f2 split = lerp(left_2, right_1, 0.5f);
        // right half below the left half: the left is finished first
        stack[top].p1 = split, stack[top].c1 = right_1, stack[top].c2 = right_2;
        stack[top].p2 = n.p2, stack[top++].limit = n.limit - 1;
        stack[top].p1 = n.p1, stack[top].c1 = left_1, stack[top].c2 = left_2;
        stack[top].p2 = split, stack[top++].limit = n.limit - 1;
    }
}

// Cuts a cubic at its x/y extrema and curvature extremum, then tessellates
// each piece (hpp:1398-1487).
template <class Sink>
CIH_HD void add_bezier(Sink &sink, f2 point_1, f2 control_1, f2 control_2, f2 point_2,
                       float angular) {
    f2 edge_1 = sub(control_1, point_1);
    f2 edge_2 = sub(control_2, control_1);
    f2 edge_3 = sub(point_2, control_2);
    if (dot(edge_1, edge_1) == 0.0f && dot(edge_3, edge_3) == 0.0f) {
        sink.point(point_2);
        return;
    }
    float at[7] = {0.0f, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    int cuts = 2;
    f2 extrema_a = add(mul(-9.0f, edge_2), mul(3.0f, sub(point_2, point_1)));
    f2 extrema_b = sub(mul(6.0f, add(point_1, control_2)), mul(12.0f, control_1));
    f2 extrema_c = mul(3.0f, edge_1);
    const float epsilon = 1.0e-4f;
    for (int axis = 0; axis < 2; ++axis) {
        float a = axis ? extrema_a.y : extrema_a.x;
        float b = axis ? extrema_b.y : extrema_b.x;
        float c = axis ? extrema_c.y : extrema_c.x;
        if (fabsf(a) > epsilon) {
            float discriminant = b * b - 4.0f * a * c;
            if (discriminant >= 0.0f) {
                float sign = b > 0.0f ? 1.0f : -1.0f;
                float term = -b - sign * sqrtf(discriminant);
                float extremum_1 = term / (2.0f * a);
                at[cuts++] = extremum_1;
                at[cuts++] = c / (a * extremum_1);
            }
        } else if (fabsf(b) > epsilon)
            at[cuts++] = -c / b;
    }
    float determinant_1 = dot(perp(edge_1), edge_2);
    float determinant_2 = dot(perp(edge_1), edge_3);
    float determinant_3 = dot(perp(edge_2), edge_3);
    float curve_a = determinant_1 - determinant_2 + determinant_3;
    float curve_b = -2.0f * determinant_1 + determinant_2;
    if (fabsf(curve_a) > epsilon && fabsf(curve_b) > epsilon)
        at[cuts++] = -0.5f * curve_b / curve_a;
    for (int index = 1; index < cuts; ++index) { // insertion sort, as std order
        float value = at[index];
        int sorted = index - 1;
        for (; 0 <= sorted && value < at[sorted]; --sorted)
            at[sorted + 1] = at[sorted];
        at[sorted + 1] = value;
    }
    f2 split_point_1 = point_1;
    for (int index = 0; index < cuts - 1; ++index) {
        if (!(0.0f <= at[index] && at[index + 1] <= 1.0f && at[index] != at[index + 1]))
            continue;
        float t = at[index + 1];
        float ratio = at[index] / t;
        f2 partial_1 = lerp(point_1, control_1, t);
        f2 partial_2 = lerp(control_1, control_2, t);
        f2 partial_3 = lerp(control_2, point_2, t);
        f2 partial_4 = lerp(partial_1, partial_2, t);
        f2 partial_5 = lerp(partial_2, partial_3, t);
        f2 partial_6 = lerp(partial_1, partial_4, ratio);
        f2 split_point_2 = lerp(partial_4, partial_5, t);
        f2 split_control_2 = lerp(partial_4, split_point_2, ratio);
        f2 split_control_1 = lerp(partial_6, split_control_2, ratio);
        tessellate(sink, split_point_1, split_control_1, split_control_2, split_point_2,
                   angular);
        split_point_1 = split_point_2;
    }
}

// Turn-cosine limit for stroking (hpp:1497-1500); -1 (no limit) for fills.
CIH_HD float angular_limit(bool stroking, float line_width) {
    const float tolerance = 0.125f;
    float ratio = tolerance / smax(0.5f * line_width, tolerance);
    return stroking ? (ratio - 2.0f) * ratio * 2.0f + 1.0f : -1.0f;
}

// ---------------------------------------------------------------------------
// Line-path sinks need `void point(f2)` and `void subpath(uint32_t count,
// bool closed)` plus the running counts `points` and `subpaths`.  Dashing
// also needs `void close_last()` and `void merge_last(uint32_t merge_point,
// uint32_t merge_subpath)` (the std::rotate merge of hpp:1916-1931).

struct dash_setup {
    int start;
    float offset;
};

// Dash phase shared by all subpaths (hpp:1866-1876).  The phase walk is
// bounded so a degenerate pattern cannot spin a GPU thread forever.
CIH_HD dash_setup dash_prepare(const float *dash, int count, float dash_offset) {
    float total = 0.0f;
    for (int k = 0; k < count; ++k)
        total = total + dash[k];
    float offset = fmodf(dash_offset, total);
    if (offset < 0.0f)
        offset += total;
    int start = 0;
    for (int guard = 0; offset >= dash[start] && guard < (1 << 20); ++guard) {
        offset -= dash[start];
        start = start + 1 < count ? start + 1 : 0;
    }
    dash_setup s;
    s.start = start;
    s.offset = offset;
    return s;
}

// Dashes one subpath of `count` points (hpp:1877-1933).
template <class Sink>
CIH_HD void dash_subpath(Sink &out, const f2 *points, uint32_t count, bool closed,
                         const float *dash, int ndash, dash_setup setup,
                         const float *inverse) {
    uint32_t index = 0;
    uint32_t first = out.points;
    int segment = setup.start;
    bool emit = (~setup.start & 1) != 0;
    uint32_t merge_point = out.points;
    uint32_t merge_subpath = out.subpaths;
    bool merge_emit = emit;
    float next = dash[setup.start] - setup.offset;
    for (; index + 1 < count; ++index) {
        f2 from = points[index];
        f2 to = points[index + 1];
        if (emit)
            out.point(from);
        float line = len(sub(apply(inverse, to), apply(inverse, from)));
        for (int guard = 0; next < line && guard < (1 << 24); ++guard) {
            out.point(lerp(from, to, next / line));
            if (emit) {
                out.subpath(out.points - first, false);
                first = out.points;
            }
            segment = segment + 1 < ndash ? segment + 1 : 0;
            emit = !emit;
            next += dash[segment];
        }
        next -= line;
    }
    if (emit) {
        out.point(points[index]);
        out.subpath(out.points - first, false);
        if (closed && merge_emit) {
            if (out.subpaths == merge_subpath + 1)
                out.close_last();
            else
                out.merge_last(merge_point, merge_subpath);
        }
    }
}

// One side of a stroke (hpp:1949-2058) is a walk over points[beginning..ending]
// in either direction, in user space (through `inverse`), that keeps a point
// only when it lies at least 1e-4 from the last kept one, and emits a join at
// each kept point between its incoming and outgoing segments (and, for open
// subpaths, a cap at the end) as device-space points (through `forward`).
// The walk is split in two: stroke_walk decides which points are kept and
// reports each join as the indices (previous kept point, joint, next point)
// and the cap as (previous kept point, end point); stroke_join and
// stroke_cap emit the points of one join or cap from those indices alone.
// half_stroke chains them in walk order; the GPU runs the joins in parallel.

// Reports rec.join(prev, at, next) for every join and rec.cap(prev, at) for
// the end cap, in the order the reference emits them.
template <class Rec>
CIH_HD void stroke_walk(Rec &rec, const f2 *points, uint32_t beginning, uint32_t ending,
                        bool closed, const float *inverse) {
    const float epsilon = 1.0e-4f;
    f2 point = apply(inverse, points[beginning]);
    float in_length = 0.0f;
    uint32_t prev = beginning, at = beginning;
    uint32_t finish = beginning;
    uint32_t index = beginning;
    do {
        f2 next = apply(inverse, points[index]);
        float out_length = len(sub(next, point));
        if (in_length != 0.0f && out_length >= epsilon) {
            if (closed && finish == beginning)
                finish = index;
            rec.join(prev, at, index);
        }
        if (out_length >= epsilon) {
            in_length = out_length;
            prev = at;
            at = index;
            point = next;
        }
        index = (index == ending ? beginning : ending > beginning ? index + 1 : index - 1);
    } while (index != finish);
    if (!closed && in_length != 0.0f)
        rec.cap(prev, at);
}

// The points of the join at points[at] between the segments from
// points[prev] and to points[next] (hpp:1966-2024).
template <class Sink>
CIH_HD void stroke_join(Sink &out, const f2 *points, uint32_t prev, uint32_t at, uint32_t nexti,
                        float line_width, float miter_limit, int join, const float *forward,
                        const float *inverse) {
    float half = line_width * 0.5f;
    float ratio = miter_limit * miter_limit * half * half;
    const float epsilon = 1.0e-4f;
    f2 before = apply(inverse, points[prev]);
    f2 point = apply(inverse, points[at]);
    f2 next = apply(inverse, points[nexti]);
    f2 in_direction = normalized(sub(point, before));
    float in_length = len(sub(point, before));
    f2 out_direction = normalized(sub(next, point));
    float out_length = len(sub(next, point));
    f2 side_in = add(point, mul(half, perp(in_direction)));
    f2 side_out = add(point, mul(half, perp(out_direction)));
    float turn = dot(perp(in_direction), out_direction);
    if (fabsf(turn) < epsilon)
        turn = 0.0f;
    f2 offset = turn == 0.0f ? mk(0.0f, 0.0f) : mul(half / turn, sub(out_direction, in_direction));
    bool tight = (dot(offset, in_direction) < -in_length && dot(offset, out_direction) > out_length);
    if (turn > 0.0f && tight) {
        f2 t = side_in;
        side_in = side_out;
        side_out = t;
        t = in_direction;
        in_direction = out_direction;
        out_direction = t;
        out.point(apply(forward, side_out));
        out.point(apply(forward, point));
        out.point(apply(forward, side_in));
    }
    if ((turn > 0.0f && !tight) || (turn != 0.0f && join == 0 && dot(offset, offset) <= ratio))
        out.point(apply(forward, add(point, offset)));
    else if (join == 2) {
        float cosine = dot(in_direction, out_direction);
        float angle = join_acos(smin(smax(cosine, -1.0f), 1.0f));
        float alpha = 4.0f / 3.0f * join_tan(0.25f * angle);
        out.point(apply(forward, side_in));
        add_bezier(out, apply(forward, side_in),
                   apply(forward, add(side_in, mul(alpha * half, in_direction))),
                   apply(forward, sub(side_out, mul(alpha * half, out_direction))),
                   apply(forward, side_out), -1.0f);
    } else {
        out.point(apply(forward, side_in));
        out.point(apply(forward, side_out));
    }
    if (turn > 0.0f && tight) {
        out.point(apply(forward, side_out));
        out.point(apply(forward, point));
        out.point(apply(forward, side_in));
    }
}

// The end cap at points[at], reached from points[prev] (hpp:2029-2057).
template <class Sink>
CIH_HD void stroke_cap(Sink &out, const f2 *points, uint32_t prev, uint32_t at,
                       float line_width, int cap, const float *forward, const float *inverse) {
    float half = line_width * 0.5f;
    f2 point = apply(inverse, points[at]);
    f2 in_direction = normalized(sub(point, apply(inverse, points[prev])));
    f2 ahead = mul(half, in_direction);
    f2 side = perp(ahead);
    if (cap == 0) {
        out.point(apply(forward, add(point, side)));
        out.point(apply(forward, sub(point, side)));
    } else if (cap == 1) {
        out.point(apply(forward, add(add(point, ahead), side)));
        out.point(apply(forward, sub(add(point, ahead), side)));
    } else if (cap == 2) {
        const float alpha = 0.55228475f; // 4/3*tan(pi/8)
        out.point(apply(forward, add(point, side)));
        add_bezier(out, apply(forward, add(point, side)),
                   apply(forward, add(add(point, side), mul(alpha, ahead))),
                   apply(forward, add(add(point, ahead), mul(alpha, side))),
                   apply(forward, add(point, ahead)), -1.0f);
        add_bezier(out, apply(forward, add(point, ahead)),
                   apply(forward, sub(add(point, ahead), mul(alpha, side))),
                   apply(forward, add(sub(point, side), mul(alpha, ahead))),
                   apply(forward, sub(point, side)), -1.0f);
    }
}

template <class Sink>
struct stroke_chain {
    Sink &out;
    const f2 *points;
    float line_width, miter_limit;
    int join, cap;
    const float *forward, *inverse;
    CIH_HD void join_at(uint32_t prev, uint32_t at, uint32_t next) {
        stroke_join(out, points, prev, at, next, line_width, miter_limit, join, forward, inverse);
    }
    CIH_HD void cap_at(uint32_t prev, uint32_t at) {
        stroke_cap(out, points, prev, at, line_width, cap, forward, inverse);
    }
};

template <class Sink>
struct stroke_chain_rec {
    stroke_chain<Sink> *c;
    CIH_HD void join(uint32_t prev, uint32_t at, uint32_t next) { c->join_at(prev, at, next); }
    CIH_HD void cap(uint32_t prev, uint32_t at) { c->cap_at(prev, at); }
};

// One side of a stroke, emitted in walk order.
template <class Sink>
CIH_HD void half_stroke(Sink &out, const f2 *points, uint32_t beginning, uint32_t ending,
                        bool closed, float line_width, float miter_limit, int join, int cap,
                        const float *forward, const float *inverse) {
    stroke_chain<Sink> chain = {out, points, line_width, miter_limit, join, cap, forward, inverse};
    stroke_chain_rec<Sink> rec = {&chain};
    stroke_walk(rec, points, beginning, ending, closed, inverse);
}

// Stroke outline of one subpath (hpp:2081-2098): one closed loop for an open
// subpath, an outer and an inner loop for a closed one.
template <class Sink>
CIH_HD void stroke_subpath(Sink &out, const f2 *points, uint32_t count, bool closed,
                           float line_width, float miter_limit, int join, int cap,
                           const float *forward, const float *inverse) {
    if (count < 2)
        return;
    uint32_t first = out.points;
    half_stroke(out, points, 0, count - 1, closed, line_width, miter_limit, join, cap,
                forward, inverse);
    if (closed) {
        out.subpath(out.points - first, true);
        first = out.points;
    }
    half_stroke(out, points, count - 1, 0, closed, line_width, miter_limit, join, cap,
                forward, inverse);
    out.subpath(out.points - first, true);
}

// ---------------------------------------------------------------------------
// Clipping: one Sutherland-Hodgman stage for the half-plane of `edge` (0: x>=0,
// 1: y>=0, 2: x<=width, 3: y<=height), for the polygon edge from -> to
// (hpp:2210-2225).  Emits up to two points.
template <class Sink>
CIH_HD void clip_edge(Sink &out, f2 from, f2 to, int edge, float width, float height) {
    f2 normal = mk(edge == 0 ? 1.0f : edge == 2 ? -1.0f : 0.0f,
                   edge == 1 ? 1.0f : edge == 3 ? -1.0f : 0.0f);
    float place = edge == 2 ? width : edge == 3 ? height : 0.0f;
    float from_side = dot(from, normal) + place;
    float to_side = dot(to, normal) + place;
    if (from_side * to_side < 0.0f)
        out.point(lerp(from, to, from_side / (from_side - to_side)));
    if (to_side >= 0.0f)
        out.point(to);
}

CIH_HD bool clip_inside(f2 p, float width, float height) {
    // the four stage tests of clip_edge for a single point
    return p.x * 1.0f + p.y * 0.0f + 0.0f >= 0.0f && p.x * 0.0f + p.y * 1.0f + 0.0f >= 0.0f &&
           p.x * -1.0f + p.y * 0.0f + width >= 0.0f && p.x * 0.0f + p.y * -1.0f + height >= 0.0f;
}

CIH_HD f2 clamp_point(f2 p, float width, float height) {
    return mk(smin(smax(p.x, 0.0f), width), smin(smax(p.y, 0.0f), height));
}

// ---------------------------------------------------------------------------
// Scan conversion of one edge into signed coverage changes (hpp:2109-2170).
// A run sink needs `void run(int x, int y, float delta)`; runs arrive row by
// row (y monotone) and left to right within a row.  Loops are bounded by the
// clamped coordinates; the guards only stop malformed input.
// add_runs row by row.  The DDA's state at the start of each pixel row has a
// closed form, so every (edge, row) pair can be emitted on its own with the
// same float operations as the sequential walk:
// * row k ends at next_y_k: y = corner.y + k * y_step (integer steps are
//   exact in float), x from the initial formula for k = 0 and from
//   (y - from.y) * slope.x + from.x after, replaced by `to` where the walk
//   clamps it; row k starts at next_y_{k-1} (or `from`);
// * next_x advances by exactly 1 from corner.x, so the x steps taken through
//   row k are the integers g >= 0 with corner.x + g < the largest next_y.x so
//   far (at most one when next_x starts at `to`);
// * the carry resets at each row, and each piece needs only its own and the
//   previous next_x.
struct edge_rows {
    bool live; // the edge makes runs (finite, |dy| >= 2e-5)
    bool flat_x;
    float sign, y_step;
    f2 from, to, slope, corner, pixel, next_x0, next_y0;
    int rows;
};

CIH_HD edge_rows edge_rows_setup(f2 from, f2 to) {
    const float epsilon = 2.0e-5f;
    edge_rows E;
    E.live = false;
    E.rows = 0;
    bool finite = fabsf(from.x) <= 3.0e38f && fabsf(from.y) <= 3.0e38f &&
                  fabsf(to.x) <= 3.0e38f && fabsf(to.y) <= 3.0e38f;
    if (!finite || fabsf(to.y - from.y) < epsilon)
        return E;
    E.live = true;
    E.sign = to.y > from.y ? 1.0f : -1.0f;
    if (from.x > to.x) {
        f2 t = from;
        from = to;
        to = t;
    }
    E.from = from;
    E.to = to;
    E.pixel = mk(floorf(from.x), floorf(from.y));
    E.corner = add(E.pixel, mk(1.0f, to.y > from.y ? 1.0f : 0.0f));
    E.slope = mk((to.x - from.x) / (to.y - from.y), (to.y - from.y) / (to.x - from.x));
    E.flat_x = to.x - from.x < epsilon;
    E.next_x0 = E.flat_x ? to : mk(E.corner.x, from.y + (E.corner.x - from.x) * E.slope.y);
    E.next_y0 = mk(from.x + (E.corner.y - from.y) * E.slope.x, E.corner.y);
    E.y_step = to.y > from.y ? 1.0f : -1.0f;
    // the walk ends after the first row whose end reaches to.y
    float ahead = E.y_step > 0.0f ? ceilf(to.y) - E.corner.y : E.corner.y - floorf(to.y);
    E.rows = (ahead > 0.0f ? (int)ahead : 0) + 1;
    return E;
}

CIH_HD f2 edge_row_end(const edge_rows &E, int k) {
    f2 n = E.next_y0;
    if (k > 0) {
        n.y = E.corner.y + (float)k * E.y_step;
        n.x = (n.y - E.from.y) * E.slope.x + E.from.x;
    }
    if ((E.from.y < E.to.y && E.to.y < n.y) || (E.from.y > E.to.y && E.to.y > n.y))
        n = E.to;
    return n;
}

// x steps taken once next_y.x has reached v.  When next_x starts at `to`
// (|dx| < 2e-5), rounding can still put next_y.x past to.x: the walk then
// takes one step, after which next_x.x = to.x + 1 is out of reach.
CIH_HD int edge_steps(const edge_rows &E, float v) {
    if (E.flat_x)
        return E.to.x < v ? 1 : 0;
    if (!(v > E.corner.x))
        return 0;
    return (int)ceilf(v) - (int)E.corner.x;
}

CIH_HD f2 edge_next_x(const edge_rows &E, int g) {
    if (g == 0)
        return E.next_x0;
    float x = E.corner.x + (float)g;
    return mk(x, (x - E.from.x) * E.slope.y + E.from.y);
}

// Runs of row k of the edge (0 <= k < rows): the x steps, then two pieces.
struct edge_row {
    f2 start, end;
    int first, steps; // x steps [first, first + steps)
};

CIH_HD edge_row edge_row_at(const edge_rows &E, int k) {
    edge_row r;
    r.end = edge_row_end(E, k);
    if (k == 0) {
        r.start = E.from;
        r.first = 0;
        r.steps = edge_steps(E, r.end.x);
    } else {
        r.start = edge_row_end(E, k - 1);
        r.first = edge_steps(E, r.start.x);
        r.steps = edge_steps(E, r.end.x > r.start.x ? r.end.x : r.start.x) - r.first;
    }
    return r;
}

CIH_HD int edge_row_runs(const edge_rows &E, int k) { return edge_row_at(E, k).steps + 2; }

template <class Sink>
CIH_HD void add_runs_row(Sink &sink, const edge_rows &E, int k) {
    edge_row r = edge_row_at(E, k);
    f2 now = r.start;
    float px = E.pixel.x + (float)r.first;
    float py = E.pixel.y + (float)k * E.y_step;
    float carry = 0.0f;
    for (int i = 0; i < r.steps; ++i) {
        f2 nx = edge_next_x(E, r.first + i);
        float strip = smin(smax((nx.y - now.y) * E.y_step, 0.0f), 1.0f);
        float mid = (nx.x + now.x) * 0.5f;
        float area = (mid - px) * strip;
        sink.run((int)px, (int)py, (carry + strip - area) * E.sign);
        carry = area;
        now = nx;
        px += 1.0f;
    }
    float strip = smin(smax((r.end.y - now.y) * E.y_step, 0.0f), 1.0f);
    float mid = (r.end.x + now.x) * 0.5f;
    float area = (mid - px) * strip;
    sink.run((int)px, (int)py, (carry + strip - area) * E.sign);
    sink.run((int)(px + 1.0f), (int)py, area * E.sign);
}

template <class Sink>
CIH_HD void add_runs(Sink &sink, f2 from, f2 to) {
    const float epsilon = 2.0e-5f;
    if (fabsf(to.y - from.y) < epsilon)
        return;
    float sign = to.y > from.y ? 1.0f : -1.0f;
    if (from.x > to.x) {
        f2 t = from;
        from = to;
        to = t;
    }
    f2 now = from;
    f2 pixel = mk(floorf(now.x), floorf(now.y));
    f2 corner = add(pixel, mk(1.0f, to.y > from.y ? 1.0f : 0.0f));
    f2 slope = mk((to.x - from.x) / (to.y - from.y), (to.y - from.y) / (to.x - from.x));
    f2 next_x = (to.x - from.x < epsilon) ? to
                                          : mk(corner.x, now.y + (corner.x - now.x) * slope.y);
    f2 next_y = mk(now.x + (corner.y - now.y) * slope.x, corner.y);
    if ((from.y < to.y && to.y < next_y.y) || (from.y > to.y && to.y > next_y.y))
        next_y = to;
    float y_step = to.y > from.y ? 1.0f : -1.0f;
    int rows = 0;
    do {
        float carry = 0.0f;
        int steps = 0;
        while (next_x.x < next_y.x && steps++ < (1 << 17)) {
            float strip = smin(smax((next_x.y - now.y) * y_step, 0.0f), 1.0f);
            float mid = (next_x.x + now.x) * 0.5f;
            float area = (mid - pixel.x) * strip;
            sink.run((int)pixel.x, (int)pixel.y, (carry + strip - area) * sign);
            carry = area;
            now = next_x;
            next_x.x += 1.0f;
            next_x.y = (next_x.x - from.x) * slope.y + from.y;
            pixel.x += 1.0f;
        }
        float strip = smin(smax((next_y.y - now.y) * y_step, 0.0f), 1.0f);
        float mid = (next_y.x + now.x) * 0.5f;
        float area = (mid - pixel.x) * strip;
        sink.run((int)pixel.x, (int)pixel.y, (carry + strip - area) * sign);
        sink.run((int)(pixel.x + 1.0f), (int)pixel.y, area * sign);
        now = next_y;
        next_y.y += y_step;
        next_y.x = (next_y.y - from.y) * slope.x + from.x;
        pixel.y += y_step;
        if ((from.y < to.y && to.y < next_y.y) || (from.y > to.y && to.y > next_y.y))
            next_y = to;
    } while (now.y != to.y && ++rows < (1 << 17));
}

} // namespace cih

#endif // CIH_GEOMETRY_HPP
