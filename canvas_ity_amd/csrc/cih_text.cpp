// cih_text.cpp -- TrueType text layout on the host: font table lookup,
// UTF-8 to glyph mapping, advance widths and glyph outlines for set_font,
// measure_text, fill_text and stroke_text (canvas_ity.hpp:1526-1846,
// 3205-3311).
//
// The reference tessellates each glyph straight into its polyline list.  Here
// a glyph becomes device-space cubic segments that the device flattens,
// dashes and strokes like any path, sent in a CIH_DRAW_FLAG_SEGMENTS packet:
// every segment carries its own start point.  The reference starts a glyph
// curve at the last point its tessellation emitted before it (hpp:1660-1665,
// 1674-1676), and that point can differ by rounding from the previous curve's
// end point, so each curve's emitted end point is found here by running the
// same add_bezier with a sink that keeps only the last point.  Straight
// segments are degenerate cubics (p, p, q, q), which flatten to the single
// point q exactly as the reference's direct push does.
//
// Reads are bounds-checked: out-of-range bytes read as zero.  For well-formed
// fonts this changes nothing; for malformed ones the reference's reads are
// undefined.  Composite glyphs nest at most kMaxComponentDepth deep.
#include "cih_abi.h"
#include "cih_geometry.hpp"

#include <climits>
#include <cstring>
#include <new>
#include <vector>

int cih_internal_fail(int status, const char *message);

namespace {

using cih::f2;

const int kMaxComponentDepth = 16;

// Big-endian reads over the font file (hpp:1300-1314), offsets into the file.
struct font_file {
    const unsigned char *data;
    long size;
    long cmap, glyf, head, hhea, hmtx, loca, maxp, os_2;
    int u8(long at) const { return at >= 0 && at < size ? data[at] : 0; }
    int s8(long at) const { return (int)(signed char)u8(at); }
    int u16(long at) const { return u8(at) << 8 | u8(at + 1); }
    int s16(long at) const { return (int)(short)u16(at); }
    int s32(long at) const {
        return (int)((uint32_t)u8(at) << 24 | (uint32_t)u8(at + 1) << 16 |
                     (uint32_t)u8(at + 2) << 8 | (uint32_t)u8(at + 3));
    }
};

// The table directory checks of set_font (hpp:3210-3270): a TrueType
// version tag, a complete offset table, every table inside the file, and the
// eight tables text needs present (the last entry of a tag wins).
bool open_font(const unsigned char *font, size_t bytes, font_file &f) {
    if (!font || bytes < 6 || bytes > (size_t)INT_MAX)
        return false;
    f.data = font;
    f.size = (long)bytes;
    uint32_t version = (uint32_t)f.s32(0);
    long tables = f.u16(4);
    if ((version != 0x00010000u && version != 0x74727565u) || f.size < tables * 16 + 12)
        return false;
    f.cmap = f.glyf = f.head = f.hhea = f.hmtx = f.loca = f.maxp = f.os_2 = -1;
    for (long index = 0; index < tables; ++index) {
        uint32_t tag = (uint32_t)f.s32(index * 16 + 12);
        long offset = f.s32(index * 16 + 20);
        long span = f.s32(index * 16 + 24);
        if (offset < 0 || span < 0 || f.size < offset + span)
            return false;
        switch (tag) {
        case 0x636d6170u: f.cmap = offset; break; // cmap
        case 0x676c7966u: f.glyf = offset; break; // glyf
        case 0x68656164u: f.head = offset; break; // head
        case 0x68686561u: f.hhea = offset; break; // hhea
        case 0x686d7478u: f.hmtx = offset; break; // hmtx
        case 0x6c6f6361u: f.loca = offset; break; // loca
        case 0x6d617870u: f.maxp = offset; break; // maxp
        case 0x4f532f32u: f.os_2 = offset; break; // OS/2
        default: break;
        }
    }
    return f.cmap >= 0 && f.glyf >= 0 && f.head >= 0 && f.hhea >= 0 && f.hmtx >= 0 &&
           f.loca >= 0 && f.maxp >= 0 && f.os_2 >= 0;
}

// Next codepoint of a NUL-terminated UTF-8 string, mapped to a glyph index
// (hpp:1709-1784).  Invalid sequences decode as U+FFFD and resume at the
// offending byte; tab, vertical tab, form feed, CR and LF map to space.  The
// cmap lookup prefers a (3,10) format-12 subtable, then (3,1) format 4, then
// any format 0; a missing glyph is glyph 0.
int next_glyph(const font_file &f, const unsigned char *text, long &index) {
    unsigned lead = text[index];
    int bytes = (lead & 0x80) == 0x00 ? 1 : (lead & 0xe0) == 0xc0 ? 2
              : (lead & 0xf0) == 0xe0 ? 3 : (lead & 0xf8) == 0xf0 ? 4 : 0;
    static const int masks[] = {0x0, 0x7f, 0x1f, 0x0f, 0x07};
    int codepoint = bytes ? (int)(lead & masks[bytes]) : 0xfffd;
    ++index;
    while (--bytes > 0) {
        if ((text[index] & 0xc0) != 0x80) {
            codepoint = 0xfffd;
            break;
        }
        codepoint = codepoint << 6 | (text[index++] & 0x3f);
    }
    if (codepoint == '\t' || codepoint == '\v' || codepoint == '\f' || codepoint == '\r' ||
        codepoint == '\n')
        codepoint = ' ';
    long format_12 = 0, format_4 = 0, format_0 = 0;
    int tables = f.u16(f.cmap + 2);
    for (int table = 0; table < tables; ++table) {
        int platform = f.u16(f.cmap + table * 8 + 4);
        int encoding = f.u16(f.cmap + table * 8 + 6);
        long offset = f.s32(f.cmap + table * 8 + 8);
        int format = f.u16(f.cmap + offset);
        if (platform == 3 && encoding == 10 && format == 12)
            format_12 = f.cmap + offset;
        else if (platform == 3 && encoding == 1 && format == 4)
            format_4 = f.cmap + offset;
        else if (format == 0)
            format_0 = f.cmap + offset;
    }
    if (format_12) {
        long groups = f.s32(format_12 + 12);
        for (long group = 0; group < groups; ++group) {
            long at = format_12 + 16 + group * 12;
            if (at >= f.size)
                break;
            int start = f.s32(at), end = f.s32(at + 4), glyph = f.s32(at + 8);
            if (start <= codepoint && codepoint <= end)
                return codepoint - start + glyph;
        }
    } else if (format_4) {
        // seg_x2 counts bytes: the four arrays hold 2-byte entries
        int seg_x2 = f.u16(format_4 + 6);
        long ends = format_4 + 14;
        long starts = ends + 2 + seg_x2;
        long deltas = starts + seg_x2;
        long ranges = deltas + seg_x2;
        for (int segment = 0; segment < seg_x2; segment += 2) {
            int start = f.u16(starts + segment);
            int end = f.u16(ends + segment);
            int delta = f.s16(deltas + segment);
            int range = f.u16(ranges + segment);
            if (start <= codepoint && codepoint <= end)
                return range ? f.u16(ranges + segment + (long)(codepoint - start) * 2 + range)
                             : (codepoint + delta) & 0xffff;
        }
    } else if (format_0 && 0 <= codepoint && codepoint < 256)
        return f.u8(format_0 + 6 + codepoint);
    return 0;
}

// Advance width of a glyph: glyphs past the last horizontal metric repeat it.
int advance(const font_file &f, int glyph) {
    int hmetrics = f.u16(f.hhea + 34);
    int entry = glyph < hmetrics - 1 ? glyph : hmetrics - 1;
    return f.u16(f.hmtx + (long)entry * 4);
}

// measure_text (hpp:3297-3311): the sum of advance widths, in font units.
long text_units(const font_file &f, const unsigned char *text) {
    long width = 0;
    for (long index = 0; text[index];)
        width += advance(f, next_glyph(f, text, index));
    return width;
}

// m * [a c e; b d f], in the association of canvas::transform (hpp:2683-2693).
void compose(float m[6], float a, float b, float c, float d, float e, float f) {
    float r[6] = {m[0] * a + m[2] * b, m[1] * a + m[3] * b, m[0] * c + m[2] * d,
                  m[1] * c + m[3] * d, m[0] * e + m[2] * f + m[4],
                  m[1] * e + m[3] * f + m[5]};
    std::memcpy(m, r, sizeof(r));
}

struct last_point {
    f2 p;
    bool any;
    void point(f2 q) { p = q, any = true; }
};

// Outline collector: one subpath per glyph contour, cubic segments with
// explicit start points.
struct outline {
    const font_file &f;
    float angular;
    std::vector<f2> points;
    std::vector<uint32_t> subpaths;
    // the contour being built
    size_t contour_at;
    bool started;
    f2 start, last; // first emitted point and last emitted point

    outline(const font_file &font, float angle) : f(font), angular(angle), contour_at(0),
                                                  started(false) {}

    void segment(f2 p1, f2 c1, f2 c2, f2 p2) {
        points.push_back(p1), points.push_back(c1);
        points.push_back(c2), points.push_back(p2);
    }
    // A point the reference appends as is.
    void push(f2 p) {
        if (!started) {
            started = true;
            start = last = p;
            return;
        }
        segment(last, last, p, p);
        last = p;
    }
    // A curve the reference tessellates with add_bezier.
    void curve(f2 p1, f2 c1, f2 c2, f2 p2) {
        segment(p1, c1, c2, p2);
        last_point end = {last, false};
        cih::add_bezier(end, p1, c1, c2, p2, angular);
        if (end.any)
            last = end.p;
    }

    // add_glyph (hpp:1533-1696) under the glyph transform m.
    void glyph(int index, const float m[6], int depth) {
        long loc_format = f.u16(f.head + 50);
        long offset = f.glyf + (loc_format ? (long)f.s32(f.loca + (long)index * 4)
                                           : (long)f.u16(f.loca + (long)index * 2) * 2);
        long next = f.glyf + (loc_format ? (long)f.s32(f.loca + (long)index * 4 + 4)
                                         : (long)f.u16(f.loca + (long)index * 2 + 2) * 2);
        if (offset == next)
            return;
        int contours = f.s16(offset);
        if (contours < 0) {
            component(offset + 10, m, depth);
            return;
        }
        simple(index, offset, contours, m);
    }

    // Composite glyph: components placed by x/y offsets with an optional
    // scale, x/y scales or 2x2 matrix in 2.14 fixed point; the flag test for
    // the first matrix entry is the reference's (flags & 200).
    void component(long offset, const float m[6], int depth) {
        if (depth >= kMaxComponentDepth)
            return;
        for (;;) {
            int flags = f.u16(offset);
            int index = f.u16(offset + 2);
            if (!(flags & 2))
                return; // matching points are not supported
            float e = (float)(flags & 1 ? f.s16(offset + 4) : f.s8(offset + 4));
            float g = (float)(flags & 1 ? f.s16(offset + 6) : f.s8(offset + 5));
            offset += flags & 1 ? 8 : 6;
            float a = flags & 200 ? (float)f.s16(offset) / 16384.0f : 1.0f;
            float b = flags & 128 ? (float)f.s16(offset + 2) / 16384.0f : 0.0f;
            float c = flags & 128 ? (float)f.s16(offset + 4) / 16384.0f : 0.0f;
            float d = flags & 8     ? a
                      : flags & 64  ? (float)f.s16(offset + 2) / 16384.0f
                      : flags & 128 ? (float)f.s16(offset + 6) / 16384.0f
                                    : 1.0f;
            offset += flags & 8 ? 2 : flags & 64 ? 4 : flags & 128 ? 8 : 0;
            float placed[6];
            std::memcpy(placed, m, sizeof(placed));
            compose(placed, a, b, c, d, e, g);
            glyph(index, placed, depth + 1);
            if (!(flags & 32))
                return;
            if (offset >= f.size)
                return;
        }
    }

    void simple(int index, long offset, int contours, const float m[6]) {
        int hmetrics = f.u16(f.hhea + 34);
        int bearing = index < hmetrics ? f.s16(f.hmtx + (long)index * 4 + 2)
                                       : f.s16(f.hmtx + (long)hmetrics * 2 + (long)index * 2);
        int x_min = f.s16(offset + 2);
        long count = (long)f.u16(offset + 8 + (long)contours * 2) + 1;
        long instructions = f.u16(offset + 10 + (long)contours * 2);
        long flag_at = offset + 12 + (long)contours * 2 + instructions;
        // sizes of the flag and x arrays
        long flag_bytes = 0, x_bytes = 0;
        for (long k = 0; k < count;) {
            if (flag_at + flag_bytes >= f.size)
                return;
            int flags = f.u8(flag_at + flag_bytes++);
            long repeat = flags & 8 ? f.u8(flag_at + flag_bytes++) + 1 : 1;
            x_bytes += repeat * (flags & 2 ? 1 : flags & 16 ? 0 : 2);
            k += repeat;
        }
        long x_at = flag_at + flag_bytes, y_at = x_at + x_bytes;
        int x = bearing - x_min, y = 0;
        int flags = 0, repeat = 0;
        long k = 0;
        for (int contour = 0; contour < contours; ++contour) {
            long ending = f.u16(offset + 10 + (long)contour * 2);
            long beginning = k;
            f2 begin_point = cih::mk(0.0f, 0.0f), end_point = cih::mk(0.0f, 0.0f);
            bool begin_on = false, end_on = false;
            contour_at = points.size();
            started = false;
            for (; k <= ending && k < count; ++k) {
                if (repeat)
                    --repeat;
                else {
                    flags = f.u8(flag_at++);
                    if (flags & 8)
                        repeat = f.u8(flag_at++);
                }
                if (flags & 2)
                    x += f.u8(x_at) * (flags & 16 ? 1 : -1);
                else if (!(flags & 16))
                    x += f.s16(x_at);
                if (flags & 4)
                    y += f.u8(y_at) * (flags & 32 ? 1 : -1);
                else if (!(flags & 32))
                    y += f.s16(y_at);
                x_at += flags & 2 ? 1 : flags & 16 ? 0 : 2;
                y_at += flags & 4 ? 1 : flags & 32 ? 0 : 2;
                f2 point = cih::apply(m, cih::mk((float)x, (float)y));
                bool on = (flags & 1) != 0;
                if (k == beginning) {
                    begin_point = point;
                    begin_on = on;
                    if (on)
                        push(point);
                } else {
                    f2 point_2 = on ? point : cih::lerp(end_point, point, 0.5f);
                    if (!started || (end_on && on))
                        push(point_2);
                    else if (!end_on || on)
                        curve(last, cih::lerp(last, end_point, 2.0f / 3.0f),
                              cih::lerp(point_2, end_point, 2.0f / 3.0f), point_2);
                }
                end_point = point;
                end_on = on;
            }
            if (!started) {
                // a contour of one off-curve point (or none) has no outline
                points.resize(contour_at);
                continue;
            }
            if (begin_on != end_on) {
                f2 control = end_on ? begin_point : end_point;
                curve(last, cih::lerp(last, control, 2.0f / 3.0f),
                      cih::lerp(start, control, 2.0f / 3.0f), start);
            } else if (!begin_on && !end_on) {
                f2 split = cih::lerp(begin_point, end_point, 0.5f);
                curve(last, cih::lerp(last, end_point, 2.0f / 3.0f),
                      cih::lerp(split, end_point, 2.0f / 3.0f), split);
                curve(split, cih::lerp(split, begin_point, 2.0f / 3.0f),
                      cih::lerp(start, begin_point, 2.0f / 3.0f), start);
            }
            push(start);
            subpaths.push_back((uint32_t)(points.size() - contour_at));
            subpaths.push_back(1u);
        }
    }
};

} // namespace

extern "C" {

CIH_API int cih_font_check(const unsigned char *font, size_t bytes, int *units_per_em) {
    font_file f;
    if (!units_per_em)
        return cih_internal_fail(CIH_ERROR_INVALID, "units_per_em is NULL");
    if (!open_font(font, bytes, f))
        return cih_internal_fail(CIH_ERROR_MALFORMED, "not a usable TrueType font");
    *units_per_em = f.u16(f.head + 18);
    return CIH_OK;
}

CIH_API int cih_text_measure(const unsigned char *font, size_t font_bytes, float scale,
                             const char *text, float *width) {
    font_file f;
    if (!width || !text)
        return cih_internal_fail(CIH_ERROR_INVALID, "invalid measure arguments");
    if (!open_font(font, font_bytes, f))
        return cih_internal_fail(CIH_ERROR_MALFORMED, "not a usable TrueType font");
    *width = (float)text_units(f, (const unsigned char *)text) * scale;
    return CIH_OK;
}

CIH_API int cih_text_outline(const unsigned char *font, size_t font_bytes, float scale,
                             const char *text, float x, float y, float maximum_width,
                             int align, int baseline, const float forward[6], int stroking,
                             float line_width, float *points, size_t point_capacity,
                             uint32_t *subpaths, size_t subpath_capacity, size_t *npoints,
                             size_t *nsubpaths) {
    font_file f;
    if (!text || !forward || !npoints || !nsubpaths || (point_capacity && !points) ||
        (subpath_capacity && !subpaths))
        return cih_internal_fail(CIH_ERROR_INVALID, "invalid outline arguments");
    *npoints = *nsubpaths = 0;
    if (!open_font(font, font_bytes, f))
        return cih_internal_fail(CIH_ERROR_MALFORMED, "not a usable TrueType font");
    if (maximum_width <= 0.0f)
        return CIH_OK;
    try {
        const unsigned char *utf8 = (const unsigned char *)text;
        // text_to_lines (hpp:1793-1846): alignment and width reduction
        float width = maximum_width == 1.0e30f && align == 0
                          ? 0.0f
                          : (float)text_units(f, utf8) * scale;
        float reduction = maximum_width / cih::smax(maximum_width, width);
        if (align == 1)
            x -= width * reduction;
        else if (align == 2)
            x -= 0.5f * width * reduction;
        float scale_x = scale * reduction, scale_y = scale * 1.0f;
        float units_per_em = (float)f.u16(f.head + 18);
        float ascender = (float)f.s16(f.os_2 + 68);
        float descender = (float)f.s16(f.os_2 + 70);
        float normalize = scale * units_per_em / (ascender - descender);
        if (baseline == 1)
            y += ascender * normalize;
        else if (baseline == 2)
            y += (ascender + descender) * 0.5f * normalize;
        else if (baseline == 3)
            y += descender * normalize;
        else if (baseline == 4)
            y += 0.6f * scale * units_per_em;
        outline out(f, cih::angular_limit(stroking != 0, line_width));
        long place = 0;
        for (long index = 0; utf8[index];) {
            int glyph = next_glyph(f, utf8, index);
            float m[6];
            std::memcpy(m, forward, sizeof(m));
            compose(m, scale_x, 0.0f, 0.0f, -scale_y, x + (float)place * scale_x, y);
            out.glyph(glyph, m, 0);
            place += advance(f, glyph);
        }
        *npoints = out.points.size();
        *nsubpaths = out.subpaths.size() / 2;
        if (*npoints > point_capacity || *nsubpaths > subpath_capacity)
            return CIH_ERROR_BUDGET; // sizes reported: call again with room
        if (*npoints)
            std::memcpy(points, &out.points[0], sizeof(f2) * *npoints);
        if (*nsubpaths)
            std::memcpy(subpaths, &out.subpaths[0], sizeof(uint32_t) * 2 * *nsubpaths);
    } catch (const std::bad_alloc &) {
        return cih_internal_fail(CIH_ERROR_OUT_OF_MEMORY, "out of host memory");
    }
    return CIH_OK;
}

} // extern "C"
