"""Scenes with closed-form answers, shared by the oracle and GPU tests.

Each builder returns a command stream; each check compares a rendering (the
fp32 linear premultiplied bitmap and/or the sRGB8 read-back) with what the
mathematics fixes independently of canvas_ity's code: pixel/rectangle overlap
areas, the shoelace area of a polygon, supersampled coverage, the
Porter-Duff operator table of the W3C 2D canvas specification, the IEC
61966-2-1 sRGB transfer, linear interpolation of gradient stops, mass
conservation of a normalised blur, identity resampling, stroke areas.
"""
import json
import math
import os

import numpy as np

from canvas_ity_amd.stream import COMPOSITE, Stream

BAYER = (np.array([[0, 8, 2, 10], [12, 4, 14, 6], [3, 11, 1, 9], [15, 7, 13, 5]]) + 0.5) / 16


def srgb_to_linear(v):
    v = np.asarray(v, dtype=np.float64)
    return np.where(v < 0.04045, v / 12.92, ((v + 0.055) / 1.055) ** 2.4)


def linear_to_srgb(v):
    v = np.asarray(v, dtype=np.float64)
    return np.where(v < 0.0031308, 12.92 * v, 1.055 * np.power(np.maximum(v, 0), 1 / 2.4) - 0.055)


# ---- coverage ------------------------------------------------------------------

def fractional_rect(size=32, rect=(3.25, 5.6, 17.3, 9.15)):
    s = Stream(size, size)
    s.set_color(0, 1, 1, 1, 1)
    s.fill_rectangle(*rect)
    return s.tobytes()


def expected_rect_alpha(size, rect):
    x0, y0, w, h = rect
    xs = np.arange(size, dtype=np.float64)
    ox = np.clip(np.minimum(xs + 1, x0 + w) - np.maximum(xs, x0), 0, 1)
    oy = np.clip(np.minimum(xs + 1, y0 + h) - np.maximum(xs, y0), 0, 1)
    return oy[:, None] * ox[None, :]


STAR = [(20.5, 3.25), (25.1, 16.9), (39.0, 17.3), (27.7, 25.6), (32.2, 39.1),
        (20.4, 30.5), (8.9, 39.0), (13.2, 25.5), (2.1, 17.0), (15.9, 16.7)]


def polygon(points=STAR, size=44):
    s = Stream(size, size)
    s.set_color(0, 1, 1, 1, 1)
    s.begin_path()
    s.move_to(*points[0])
    for p in points[1:]:
        s.line_to(*p)
    s.close_path()
    s.fill()
    return s.tobytes()


def shoelace(points):
    """Polygon area, from the float32 vertices the canvas actually receives."""
    points = [(float(np.float32(x)), float(np.float32(y))) for x, y in points]
    a = 0.0
    for (x0, y0), (x1, y1) in zip(points, points[1:] + points[:1]):
        a += x0 * y1 - x1 * y0
    return abs(a) / 2


def point_in_polygon(px, py, points):
    inside = np.zeros(px.shape, dtype=bool)
    m = len(points)
    for i in range(m):
        x0, y0 = points[i]
        x1, y1 = points[(i + 1) % m]
        cond = (y0 > py) != (y1 > py)
        xint = x0 + (py - y0) * (x1 - x0) / (y1 - y0 + 1e-300)
        inside ^= cond & (px < xint)
    return inside


def supersampled_coverage(points, size, n=64):
    offs = (np.arange(n) + 0.5) / n
    cov = np.zeros((size, size))
    for y in range(size):
        py = (y + offs)[:, None] * np.ones((1, n))
        for x in range(size):
            px = np.ones((n, 1)) * (x + offs)[None, :]
            cov[y, x] = point_in_polygon(px, py, points).mean()
    return cov


TRIANGLE = [(2.3, 1.7), (13.9, 4.2), (5.1, 12.6)]


def triangle(size=16):
    return polygon(TRIANGLE, size)


# ---- compositing ----------------------------------------------------------------

# Porter-Duff factors (Fa for the source, Fb for the destination) of the
# W3C 2D context compositing operators, in terms of source and destination
# alpha.  "clears" marks the operators that also affect pixels the source
# does not cover.
PORTER_DUFF = {
    "source_atop": (lambda sa, da: da, lambda sa, da: 1 - sa, False),
    "source_in": (lambda sa, da: da, lambda sa, da: 0.0, True),
    "source_out": (lambda sa, da: 1 - da, lambda sa, da: 0.0, True),
    "source_over": (lambda sa, da: 1.0, lambda sa, da: 1 - sa, False),
    "destination_atop": (lambda sa, da: 1 - da, lambda sa, da: sa, True),
    "destination_in": (lambda sa, da: 0.0, lambda sa, da: sa, True),
    "destination_out": (lambda sa, da: 0.0, lambda sa, da: 1 - sa, False),
    "destination_over": (lambda sa, da: 1 - da, lambda sa, da: 1.0, False),
    "lighter": (lambda sa, da: 1.0, lambda sa, da: 1.0, False),
    "source_copy": (lambda sa, da: 1.0, lambda sa, da: 0.0, True),
    "exclusive_or": (lambda sa, da: 1 - da, lambda sa, da: 1 - sa, False),
}

BACK = (0.8, 0.3, 0.1, 0.6)
FORE = (0.2, 0.5, 0.9, 0.7)


def composite_scene(op, size=16):
    s = Stream(size, size)
    s.set_color(0, *BACK)
    s.fill_rectangle(0, 0, size, size)
    s.set_composite_op(COMPOSITE[op])
    s.set_color(0, *FORE)
    s.fill_rectangle(0, 0, size // 2, size)
    return s.tobytes()


def premul_linear(c):
    lin = srgb_to_linear(np.asarray(c[:3]))
    return np.concatenate([lin * c[3], [c[3]]])


def expected_composite(op, size=16):
    fa, fb, clears = PORTER_DUFF[op]
    back = premul_linear(BACK)
    fore = premul_linear(FORE)
    inside = fa(fore[3], back[3]) * fore + fb(fore[3], back[3]) * back
    inside[3] = min(inside[3], 1.0)
    if clears:
        outside = fb(0.0, back[3]) * back
    else:
        outside = back
    out = np.zeros((size, size, 4))
    out[:, : size // 2] = inside
    out[:, size // 2:] = outside
    return out


# ---- colour ----------------------------------------------------------------------

SOLID = (0.71, 0.33, 0.05, 0.83)


def solid(size=8, color=SOLID):
    s = Stream(size, size)
    s.set_color(0, *color)
    s.fill_rectangle(0, 0, size, size)
    return s.tobytes()


def expected_solid_srgb8(size=8, color=SOLID):
    lin = premul_linear(color)
    a = lin[3]
    rgb = linear_to_srgb(np.clip(lin[:3] / a, 0, 1))
    vals = np.concatenate([rgb, [a]]) * 255.0
    ys, xs = np.mgrid[0:size, 0:size]
    return vals[None, None, :] + BAYER[ys & 3, xs & 3][..., None]


GRADIENT = ((1.0, 0.2, 0.1, 1.0), (0.1, 0.3, 1.0, 0.4))


def linear_gradient(size=32):
    s = Stream(size, 4)
    s.set_linear_gradient(0, 0, 0, size, 0)
    s.add_color_stop(0, 0.0, *GRADIENT[0])
    s.add_color_stop(0, 1.0, *GRADIENT[1])
    s.fill_rectangle(0, 0, size, 4)
    return s.tobytes()


def expected_linear_gradient(size=32):
    t = (np.arange(size) + 0.5) / size
    a = np.array([*srgb_to_linear(GRADIENT[0][:3]), GRADIENT[0][3]])
    b = np.array([*srgb_to_linear(GRADIENT[1][:3]), GRADIENT[1][3]])
    c = a[None, :] + t[:, None] * (b - a)[None, :]
    c[:, :3] *= c[:, 3:4]
    return np.broadcast_to(c[None], (4, size, 4))


# ---- shadows -----------------------------------------------------------------------

SHADOW_RECT = (6.0, 6.0, 14.0, 10.0)
SHADOW_COLOR = (0.2, 0.4, 0.6, 0.75)


SHADOW_OFFSET = (80.0, 76.0)


def shadow_scene(blur, size=192, offset=SHADOW_OFFSET):
    s = Stream(size, size)
    s.set_shadow_blur(blur)
    s.set_shadow_offset(*offset)
    s.set_shadow_color(*SHADOW_COLOR)
    s.set_color(0, 1, 1, 1, 1)
    s.fill_rectangle(*SHADOW_RECT)
    return s.tobytes()


def shadow_mass_expected():
    return SHADOW_RECT[2] * SHADOW_RECT[3] * SHADOW_COLOR[3]


# ---- images ------------------------------------------------------------------------

def noise_image(w=24, h=20, seed=5):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)


def identity_image(img):
    h, w = img.shape[:2]
    s = Stream(w, h)
    s.define_image(0, img)
    s.draw_image(0, 0, 0, w, h)
    return s.tobytes()


def expected_image_linear(img):
    c = img.astype(np.float64) / 255.0
    out = np.empty(c.shape)
    out[..., :3] = srgb_to_linear(c[..., :3]) * c[..., 3:4]
    out[..., 3] = c[..., 3]
    return out


# ---- strokes -------------------------------------------------------------------------

def straight_stroke(cap, width=6.0, x0=10.0, x1=50.0, y=20.0, size=64):
    s = Stream(size, 40)
    s.set_line_width(width)
    s.set_line_cap(cap)
    s.set_color(1, 1, 1, 1, 1)
    s.begin_path()
    s.move_to(x0, y)
    s.line_to(x1, y)
    s.stroke()
    return s.tobytes()


def stroke_area(cap, width=6.0, x0=10.0, x1=50.0):
    length = x1 - x0
    return {0: length * width, 1: (length + width) * width,
            2: length * width + math.pi * (width / 2) ** 2}[cap]


# ---- even-odd (extension) ------------------------------------------------------------

def nested_squares(evenodd, size=32):
    s = Stream(size, size)
    s.set_color(0, 1, 1, 1, 1)
    s.set_fill_rule(1 if evenodd else 0)
    s.begin_path()
    s.rectangle(4, 4, 24, 24)
    s.rectangle(10, 10, 12, 12)
    s.fill()
    return s.tobytes()


# Fractional windings.  Every coordinate below is a dyadic rational, so each
# pixel's overlap fraction a with a rectangle is exact in float and the
# coverage sums are exact too.  Even-odd takes the triangle wave of the
# winding sum s (period 2): s for s <= 1, 2 - s for 1 < s <= 2.  Omitting the
# fold (min(fmod(|s|, 2), 1)) or folding with period 1 changes every pixel
# whose s lies strictly between 1 and 2, and both scenes have such pixels.

DOUBLE_RECT = (3.25, 5.5, 17.375, 9.125)


def double_rect(evenodd, size=32, rect=DOUBLE_RECT):
    """The same fractional rectangle twice in one path: s = 2a per pixel."""
    s = Stream(size, size)
    s.set_color(0, 1, 1, 1, 1)
    s.set_fill_rule(1 if evenodd else 0)
    s.begin_path()
    s.rectangle(*rect)
    s.rectangle(*rect)
    s.fill()
    return s.tobytes()


RING_OUTER = (2.0, 2.0, 28.0, 28.0)
RING_INNER = (8.25, 9.5, 12.375, 10.625)


def ring(evenodd, size=32):
    """A pixel-aligned square with a fractional square inside it, both in the
    same orientation: s = 1 + a_inner per pixel of the outer square."""
    s = Stream(size, size)
    s.set_color(0, 1, 1, 1, 1)
    s.set_fill_rule(1 if evenodd else 0)
    s.begin_path()
    s.rectangle(*RING_OUTER)
    s.rectangle(*RING_INNER)
    s.fill()
    return s.tobytes()


def even_odd(s):
    """Even-odd coverage of a winding sum s in [0, 2]."""
    s = np.abs(np.asarray(s, dtype=np.float64))
    return np.where(s > 1.0, 2.0 - s, s)


def expected_double_rect(evenodd, size=32, rect=DOUBLE_RECT):
    s = 2.0 * expected_rect_alpha(size, rect)
    return even_odd(s) if evenodd else np.minimum(s, 1.0)


def expected_ring(evenodd, size=32):
    outer = expected_rect_alpha(size, RING_OUTER)
    s = outer + expected_rect_alpha(size, RING_INNER)
    return even_odd(s) if evenodd else np.minimum(s, 1.0)


def winding_sums(size=32):
    """The winding sums of both scenes, for checking that they reach the
    interval (1, 2) where the fold matters."""
    return (2.0 * expected_rect_alpha(size, DOUBLE_RECT),
            expected_rect_alpha(size, RING_OUTER) + expected_rect_alpha(size, RING_INNER))


# ---- the reference's exact get_image_data known-answer test --------------------------

def kat_stream():
    """The drawing of test.cpp:1802-1811: 100 3x3 squares with colours from
    the index bits, on a 256x256 canvas."""
    s = Stream(256, 256)
    for index in range(100):
        s.set_color(0, float(index // 2 % 2), float(index // 4 % 2), float(index // 8 % 2),
                    float(index // 16 % 2))
        s.fill_rectangle(3.0 * (index % 10), 3.0 * (index // 10), 3.0, 3.0)
    return s.tobytes()


KAT_HASH = 0xF53F9792  # test.cpp:1822


def kat_hash(get_image_data):
    """Reads back 35x35 pixels at (-10,-10) into a 4939-byte LCG-filled buffer
    at offset 2 with stride 141 and hashes every byte (test.cpp:1812-1821)."""
    data = np.zeros(4939, dtype=np.uint8)
    data[0] = 150
    for index in range(1, 4939):
        data[index] = (int(data[index - 1]) * 137 + 53) & 255
    get_image_data(data[2:], 35, 35, 141, -10, -10)
    h = 0
    for byte in data.tolist():
        h = ((((h & 0x1FFFF) << 15) | (h >> 17)) ^ byte) & 0xFFFFFFFF
    return h


# ---- clip masks (hpp:3057-3099) ----------------------------------------------

def clip_scenes():
    """Streams exercising clip(): nested clips under save/restore, every
    whole-canvas operator under a mask, shadows, gradients and images drawn
    through a mask, and curved and self-intersecting clip paths."""
    out = {}
    # nested clips with save/restore and a curved clip path
    s = Stream(96, 80)
    s.set_color(0, 0.9, 0.9, 0.9, 1.0)
    s.fill_rectangle(0, 0, 96, 80)
    s.save()
    s.begin_path(); s.arc(48, 40, 30, 0.0, 6.2831853); s.clip()
    s.set_color(0, 0.8, 0.2, 0.1, 0.9)
    s.fill_rectangle(5, 5, 70, 50)
    s.save()
    s.begin_path(); s.rectangle(30.5, 10.25, 50, 35.5); s.clip()
    s.set_linear_gradient(0, 10, 10, 80, 60)
    s.add_color_stop(0, 0.0, 0.1, 0.2, 0.9, 1.0)
    s.add_color_stop(0, 1.0, 0.9, 0.8, 0.1, 0.6)
    s.fill_rectangle(0, 0, 96, 80)
    s.restore()
    s.set_color(0, 0.1, 0.7, 0.2, 0.7)
    s.begin_path(); s.move_to(10, 70); s.line_to(90, 5); s.line_to(60, 78); s.close_path()
    s.fill()
    s.restore()
    s.set_color(1, 0.0, 0.0, 0.0, 1.0)
    s.set_line_width(3.0)
    s.stroke_rectangle(8, 8, 40, 30)
    out["nested"] = s.tobytes()
    # whole-canvas operators under a mask
    for op in ("source_copy", "source_in", "destination_in", "destination_atop", "source_out"):
        s = Stream(64, 48)
        s.set_color(0, 0.2, 0.3, 0.9, 0.8)
        s.fill_rectangle(4, 4, 50, 36)
        s.begin_path(); s.move_to(10, 2); s.bezier_curve_to(70, 10, -10, 40, 55, 46)
        s.line_to(30, 20); s.close_path(); s.clip()
        s.set_composite_op(COMPOSITE[op])
        s.set_color(0, 0.9, 0.5, 0.1, 0.6)
        s.begin_path(); s.arc(32, 24, 15.5, 0.3, 5.9); s.fill()
        out["op_" + op] = s.tobytes()
    # a shadowed draw and an image through a mask
    s = Stream(80, 64)
    s.begin_path(); s.rectangle(12.5, 8.5, 50, 44); s.clip()
    s.set_shadow_color(0.1, 0.1, 0.3, 0.8)
    s.set_shadow_blur(6.0)
    s.set_shadow_offset(5.0, 4.0)
    s.set_color(0, 0.9, 0.8, 0.2, 1.0)
    s.fill_rectangle(20, 15, 30, 20)
    s.set_shadow_color(0.0, 0.0, 0.0, 0.0)
    s.define_image(1, noise_image(24, 20, seed=9))
    s.draw_image(1, 30, 28, 48, 40)
    out["shadow_image"] = s.tobytes()
    # a self-intersecting clip path (non-zero winding 2 inside) and a clip
    # that leaves nothing
    s = Stream(64, 64)
    s.begin_path()
    s.move_to(32, 2); s.line_to(50, 60); s.line_to(2, 22); s.line_to(62, 22)
    s.line_to(14, 60); s.close_path(); s.clip()
    s.set_color(0, 0.3, 0.8, 0.6, 1.0)
    s.fill_rectangle(0, 0, 64, 64)
    s.save()
    s.begin_path(); s.rectangle(100, 100, 10, 10); s.clip()
    s.set_color(0, 1.0, 0.0, 0.0, 1.0)
    s.fill_rectangle(0, 0, 64, 64)
    s.restore()
    out["star_empty"] = s.tobytes()
    return out


def clipped_fill(clips, size=32):
    """Clips to each (x, y, w, h) rectangle in turn, then fills the whole
    canvas opaque white: a pixel's alpha is the product of its overlaps with
    the clip rectangles (visibility = min(|coverage|, 1) multiplied per clip,
    and source_over onto transparent black leaves alpha = visibility)."""
    s = Stream(size, size)
    for rect in clips:
        s.begin_path()
        s.rectangle(*rect)
        s.clip()
    s.set_color(0, 1.0, 1.0, 1.0, 1.0)
    s.fill_rectangle(0, 0, size, size)
    return s.tobytes()


def expected_clipped_alpha(clips, size=32):
    alpha = np.ones((size, size))
    for rect in clips:
        alpha = alpha * expected_rect_alpha(size, rect)
    return alpha


def reference_clip_tests(size=256):
    """The call sequences of the reference's own clip-dependent tests
    (test.cpp:440-467 global_composite_operation, 551-577
    shadow_blur_composite, 1396-1418 clip, 1420-1447 clip_winding, 1855-1873
    save_restore) at the test suite's 256x256, as command streams."""
    w = h = float(size)
    out = {}
    s = Stream(size, size)
    ops = [("source_in", 1), ("source_copy", 2), ("source_out", 3), ("destination_in", 4),
           ("destination_atop", 7), ("lighter", 10), ("destination_over", 11),
           ("destination_out", 12), ("source_atop", 13), ("source_over", 14),
           ("exclusive_or", 15)]
    bw, bh = 0.25 * w, 0.25 * h
    for _, op in ops:
        col, row = float(op % 4), float(op // 4)
        s.save()
        s.begin_path(); s.rectangle(col * bw, row * bh, bw, bh); s.clip()
        s.set_color(0, 0.0, 0.0, 1.0, 1.0)
        s.fill_rectangle((col + 0.4) * bw, (row + 0.4) * bh, 0.4 * bw, 0.4 * bh)
        s.set_composite_op(op)
        s.set_color(0, 1.0, 0.0, 0.0, 1.0)
        s.fill_rectangle((col + 0.2) * bw, (row + 0.2) * bh, 0.4 * bw, 0.4 * bh)
        s.restore()
    out["global_composite_operation"] = s.tobytes()

    s = Stream(size, size)
    radius = min(w, h) * 0.5
    s.arc(0.5 * w, 0.5 * h, radius, 0.0, 6.28318531); s.clip()
    s.set_color(0, 0.0, 0.0, 1.0, 1.0)
    s.set_shadow_color(0.0, 0.0, 0.0, 1.0)
    s.fill_rectangle(0.4 * w, 0.0, 0.2 * w, h)
    s.set_composite_op(7)  # destination_atop
    s.set_color(1, 1.0, 0.0, 0.0, 1.0)
    s.set_line_dash([16.0, 4.0])
    s.set_line_width(15.0)
    s.set_shadow_offset(5.0, 5.0)
    s.set_shadow_blur(6.0)
    s.set_shadow_color(0.0, 0.0, 0.0, 1.0)
    s.begin_path(); s.arc(0.45 * w, 0.85 * h, radius * 0.5, 0.0, 6.28318531); s.close_path()
    s.stroke()
    s.set_composite_op(14)  # source_over
    s.begin_path(); s.arc(0.75 * w, 0.25 * h, radius, 0.0, 6.28318531); s.close_path()
    s.stroke()
    out["shadow_blur_composite"] = s.tobytes()

    s = Stream(size, size)
    s.set_line_width(8.0)
    for step in range(8):
        angle = np.float32(np.float32(step) / np.float32(8.0)) * np.float32(6.28318531)
        s.set_color(1, 0.0, float(step & 1), 0.0, 1.0)
        s.begin_path()
        s.arc(0.5 * w + 0.8 * radius * float(np.cos(angle, dtype=np.float32)),
              0.5 * h + 0.8 * radius * float(np.sin(angle, dtype=np.float32)),
              radius, 0.0, 6.28318531)
        s.close_path(); s.stroke(); s.clip()
    s.begin_path(); s.clip()
    s.set_color(0, 1.0, 0.0, 1.0, 1.0)
    s.fill_rectangle(0.0, 0.0, w, h)
    out["clip"] = s.tobytes()

    s = Stream(size, size)
    for x, y in ((0.125, 0.125), (0.625, 0.125), (0.625, 0.625), (0.125, 0.625)):
        (s.move_to if (x, y) == (0.125, 0.125) else s.line_to)(x * w, y * h)
    s.move_to(0.25 * w, 0.25 * h); s.line_to(0.75 * w, 0.25 * h)
    s.line_to(0.75 * w, 0.75 * h); s.line_to(0.25 * w, 0.75 * h)
    s.move_to(0.375 * w, 0.375 * h); s.line_to(0.375 * w, 0.875 * h)
    s.line_to(0.875 * w, 0.875 * h); s.line_to(0.875 * w, 0.375 * h)
    s.set_color(0, 1.0, 0.0, 0.0, 1.0)
    s.fill(); s.clip()
    s.set_line_width(4.0); s.stroke()
    s.set_line_width(6.0)
    s.begin_path()
    for step in range(32):
        s.move_to(step / 16.0 * w, 0.0); s.line_to(step / 16.0 * w - w, h)
    s.stroke()
    out["clip_winding"] = s.tobytes()

    s = Stream(size, size)
    s.rectangle(w * 0.25, h * 0.25, w * 0.25, h * 0.25)
    s.set_color(1, 0.0, 0.0, 1.0, 1.0)
    s.set_line_width(8.0)
    s.save(); s.clip()
    s.begin_path(); s.rectangle(w * 0.25, h * 0.25, w * 0.5, h * 0.5)
    s.set_color(1, 1.0, 0.0, 0.0, 1.0)
    s.set_line_width(1.0)
    s.restore(); s.restore()
    s.stroke()
    s.save(); s.save()
    out["save_restore"] = s.tobytes()
    return out


# ---- the reference test suite's text tests and image hash -----------------

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def font(name):
    """A TrueType fixture from the reference's tests (tests/golden/fonts)."""
    with open(os.path.join(GOLDEN, "fonts", "font_%s.ttf" % name), "rb") as f:
        return f.read()


def reference_hashes():
    """{test name: {"hash", "width", "height"}} from the reference's table."""
    with open(os.path.join(GOLDEN, "reference_hashes.json")) as f:
        return json.load(f)


_HASH_ROLLS = {}


def _hash_rolls(count):
    """The rotation of each (pixel, channel) step of the hash's xorshift state
    (test.cpp:2370, 2397-2402); it does not depend on the image."""
    if count not in _HASH_ROLLS:
        state, rolls = 0xFFFFFFFF, np.empty(count, dtype=np.uint32)
        for k in range(count):
            state ^= (state & 0x7FFFF) << 13
            state ^= state >> 17
            state ^= (state & 0x7FFFFFF) << 5
            state &= 0xFFFFFFFF
            rolls[k] = state >> 27
        _HASH_ROLLS[count] = rolls
    return _HASH_ROLLS[count]


def image_hash(rgba8):
    """The reference suite's 32-bit locality-sensitive image hash
    (test.cpp:2365-2406): per channel, edge flags against the right and lower
    neighbours (wrapping; colour channels weighted by alpha), each rotated by
    a pseudorandom amount and XORed together."""
    img = np.asarray(rgba8, dtype=np.int64)
    h, w = img.shape[:2]
    alpha = img[:, :, 3:4]
    weighted = np.concatenate([img[:, :, :3] * alpha, img[:, :, 3:4]], axis=2)
    threshold = np.array([8 * 255] * 3 + [8], dtype=np.int64)
    down = np.roll(weighted, -1, axis=0)
    right = np.roll(weighted, -1, axis=1)

    def flags(d, hi, lo):
        return (np.where(d > threshold * 16, hi, 0) | np.where(d > threshold, lo, 0))
    edges = (flags(weighted - down, 128, 64) | flags(down - weighted, 32, 16) |
             flags(weighted - right, 8, 4) | flags(right - weighted, 2, 1))
    edges = edges.reshape(-1).astype(np.uint64)
    rolls = _hash_rolls(edges.size).astype(np.uint64)
    rotated = ((edges << rolls) | (edges >> ((32 - rolls) % 32) * (rolls != 0))) & 0xFFFFFFFF
    return int(np.bitwise_xor.reduce(rotated))


def hash_distance(a, b):
    return bin((a ^ b) & 0xFFFFFFFF).count("1")


def _f(v):
    return np.float32(v)


def reference_text_tests(size=256, measure=None):
    """The call sequences of the reference's text tests (test.cpp:1604-1741
    text_align, text_baseline, font, fill_text, stroke_text, measure_text and
    1875-1907 example_button) as command streams, with float32 arithmetic
    for every computed coordinate.  `measure(font, size, text)` supplies
    measure_text's values, which the reference test adds up to place glyphs."""
    w = h = _f(size)
    fa, fb, fc = font("a"), font("b"), font("c")
    out = {}
    line = "CE\xc3\x8d\xf4\x8f\xbf\xbd\xf0I".encode("latin-1")

    def guides(s):
        s.set_color(1, 1.0, 0.0, 0.0, 0.5)
        s.set_line_width(1.0)
        s.move_to(0.0, _f(0.5) * h); s.line_to(w, _f(0.5) * h)
        s.move_to(_f(0.5) * w, 0.0); s.line_to(_f(0.5) * w, h)
        s.stroke()

    s = Stream(size, size)
    s.set_font(fa, _f(0.2) * h)
    s.rotate(0.2)
    s.set_color(0, 0.0, 0.0, 0.0, 1.0)
    for index, align in enumerate((0, 2, 1, 0, 1)):  # leftward center rightward start ending
        base = (_f(0.1) + _f(0.2) * _f(index)) * h
        s.set_text_align(align)
        s.fill_text("HIty", _f(0.5) * w, base)
    guides(s)
    out["text_align"] = s.tobytes()

    s = Stream(size, size)
    s.set_font(fa, _f(0.2) * h)
    s.rotate(0.2)
    s.set_color(0, 0.0, 0.0, 0.0, 1.0)
    for index, baseline in enumerate((0, 1, 2, 3, 4, 3)):  # ... hanging ideographic
        left = (_f(0.1) + _f(0.15) * _f(index)) * w
        s.set_text_baseline(baseline)
        s.fill_text("Iy", left, _f(0.5) * h)
    guides(s)
    out["text_baseline"] = s.tobytes()

    s = Stream(size, size)
    s.set_color(1, 1.0, 0.0, 0.0, 1.0)
    s.stroke_text("D", _f(0.8) * w, _f(0.95) * h)
    s.set_color(0, 1.0, 0.0, 0.0, 1.0)
    s.fill_text("D", _f(0.9) * w, _f(0.95) * h)
    s.set_font(None, _f(0.1) * h)
    s.set_color(0, 0.0, 0.0, 0.0, 1.0)
    s.set_font(fa, _f(0.2) * h)
    s.fill_text(line, 0.0, _f(0.20) * h)
    s.set_font(None, _f(0.1) * h)
    s.fill_text(line, _f(0.65) * w, _f(0.20) * h)
    s.set_font(fb, _f(0.2) * h)
    s.fill_text(line, 0.0, _f(0.45) * h)
    s.set_font(fc, _f(0.2) * h)
    s.fill_text(line, 0.0, _f(0.70) * h)
    s.set_color(0, 1.0, 0.0, 0.0, 1.0)
    for name, x in (("d", 0.1), ("e", 0.2), ("f", 0.3), ("g", 0.4)):
        s.set_font(font(name), _f(0.2) * h)
        s.fill_text("D", _f(x) * w, _f(0.95) * h)
    out["font"] = s.tobytes()

    s = Stream(size, size)
    s.set_linear_gradient(0, _f(0.4) * w, 0.0, _f(0.6) * w, 0.0)
    for stop in ((0.00, 0.0, 0.00, 1.0, 1.0), (0.45, 0.0, 0.25, 0.5, 1.0),
                 (0.50, 1.0, 0.00, 0.0, 1.0), (0.55, 0.0, 0.25, 0.5, 1.0),
                 (1.00, 0.0, 0.50, 0.0, 1.0)):
        s.add_color_stop(0, *stop)
    s.set_font(fa, _f(0.3) * h)
    s.rotate(0.2)
    s.set_shadow_offset(2.0, 2.0)
    s.set_shadow_blur(4.0)
    s.set_shadow_color(0.0, 0.0, 0.0, 0.75)
    s.move_to(0.0, _f(0.2) * h)
    s.fill_text("Canvas", _f(0.1) * w, _f(0.2) * h)
    s.line_to(w, _f(0.2) * h)
    s.fill_text("Ity\n*", _f(0.2) * w, _f(0.5) * h, w)
    s.move_to(0.0, _f(0.5) * h)
    s.fill_text("*Canvas\fIty*", _f(0.2) * w, _f(0.8) * h, _f(0.7) * w)
    s.set_color(0, 1.0, 0.0, 0.0, 1.0)
    s.fill_text("****", _f(0.1) * w, _f(0.35) * h, 0.0)
    s.line_to(w, _f(0.5) * h)
    s.set_shadow_color(0.0, 0.0, 0.0, 0.0)
    s.set_color(1, 1.0, 0.0, 0.0, 1.0)
    s.set_line_width(2.0)
    s.stroke()
    out["fill_text"] = s.tobytes()

    s = Stream(size, size)
    s.set_color(1, 0.0, 0.0, 0.0, 1.0)
    s.set_font(fa, _f(0.3) * h)
    s.rotate(0.2)
    s.set_line_width(2.0)
    s.set_line_dash([8.0, 2.0])
    s.move_to(0.0, _f(0.2) * h)
    s.stroke_text("Canvas", _f(0.1) * w, _f(0.2) * h)
    s.line_to(w, _f(0.2) * h)
    s.stroke_text("Ity\n*", _f(0.2) * w, _f(0.5) * h, w)
    s.move_to(0.0, _f(0.5) * h)
    s.stroke_text("*Canvas\fIty*", _f(0.2) * w, _f(0.8) * h, _f(0.7) * w)
    s.set_color(1, 1.0, 0.0, 0.0, 1.0)
    s.stroke_text("****", _f(0.1) * w, _f(0.35) * h, 0.0)
    s.line_to(w, _f(0.5) * h)
    s.set_line_dash(None)
    s.stroke()
    out["stroke_text"] = s.tobytes()

    if measure is not None:
        s = Stream(size, size)
        s.set_color(0, 0.0, 0.0, 0.0, 1.0)
        place = _f(0.1) * w
        place = _f(place + _f(0.0))  # measure_text("C") before any font: 0
        size_a = _f(0.3) * h
        s.set_font(fa, size_a)
        s.rotate(0.5)
        s.scale(1.15, 1.0)
        for text, y in (("C", 0.2), ("a", 0.25), ("nv", 0.2), ("a", 0.15), ("s", 0.2)):
            s.fill_text(text, place, _f(y) * h)
            place = _f(place + _f(measure(fa, size_a, text)))
        out["measure_text"] = s.tobytes()

    s = Stream(size, size)
    def roundf(v):  # halves away from zero; the values here are positive
        return _f(math.floor(float(v) + 0.5))
    left, right = roundf(_f(0.25) * w), roundf(_f(0.75) * w)
    top, bottom = roundf(_f(0.375) * h), roundf(_f(0.625) * h)
    mid_x, mid_y = (left + right) * _f(0.5), (top + bottom) * _f(0.5)
    half = _f(0.5)
    s.set_shadow_offset(3.0, 3.0)
    s.set_shadow_blur(3.0)
    s.set_shadow_color(0.0, 0.0, 0.0, 0.5)
    s.set_linear_gradient(0, 0.0, top, 0.0, bottom)
    s.add_color_stop(0, 0.0, 0.3, 0.3, 0.3, 1.0)
    s.add_color_stop(0, 1.0, 0.2, 0.2, 0.2, 1.0)
    s.move_to(left + half, mid_y)
    s.arc_to(left + half, top + half, mid_x, top + half, 4.0)
    s.arc_to(right - half, top + half, right - half, mid_y, 4.0)
    s.arc_to(right - half, bottom - half, mid_x, bottom - half, 4.0)
    s.arc_to(left + half, bottom - half, left + half, mid_y, 4.0)
    s.close_path()
    s.fill()
    s.set_shadow_color(0.0, 0.0, 0.0, 0.0)
    s.set_font(fa, _f(0.075) * h)
    s.set_text_align(2)     # center
    s.set_text_baseline(2)  # middle
    s.set_color(0, 0.8, 0.8, 0.8, 1.0)
    s.fill_text("* Cats", _f(0.5) * w, _f(0.5) * h)
    s.set_color(0, 0.4, 0.4, 0.4, 1.0)
    s.fill_rectangle(left + _f(4.0), top + _f(1.0), right - left - _f(8.0), _f(1.0))
    s.set_color(1, 0.1, 0.1, 0.1, 1.0)
    s.stroke()
    out["example_button"] = s.tobytes()
    return out


TEXT_CHARS = ["C", "D", "E", "F", "G", "H", "I", "a", "n", "s", "t", "v", "y", "*", " ", "\n",
              "\t", "Í", "\U0010fffd", "?", "Q"]


def text_scene(seed, size=160):
    """A seeded text scene: fill_text and stroke_text of random strings over
    the reference's three valid test fonts, with random transforms, alignment,
    baselines, maximum widths, dashes, shadows, gradients and composite
    operations.  Characters the fonts lack draw glyph 0."""
    rng = np.random.default_rng(seed)
    u = lambda lo, hi: float(np.float32(rng.uniform(lo, hi)))
    s = Stream(size, size)
    fonts = [font("a"), font("b"), font("c")]
    ops = sorted(COMPOSITE.values())
    for k in range(int(rng.integers(3, 7))):
        s.save()
        s.translate(u(0, size * 0.5), u(size * 0.2, size * 0.8))
        s.rotate(u(-0.8, 0.8))
        if rng.integers(0, 3) == 0:
            s.scale(u(0.6, 1.5), u(0.6, 1.5))
        s.set_font(fonts[int(rng.integers(0, 3))], u(8, size * 0.3))
        s.set_text_align(int(rng.integers(0, 3)))
        s.set_text_baseline(int(rng.integers(0, 5)))
        if rng.integers(0, 4) == 0:
            s.set_linear_gradient(int(rng.integers(0, 2)), 0.0, 0.0, u(20, size), u(0, 40))
            for t in sorted(rng.uniform(0, 1, size=int(rng.integers(2, 5)))):
                s.add_color_stop(0, float(t), u(0, 1), u(0, 1), u(0, 1), u(0.5, 1))
                s.add_color_stop(1, float(t), u(0, 1), u(0, 1), u(0, 1), u(0.5, 1))
        else:
            s.set_color(0, u(0, 1), u(0, 1), u(0, 1), u(0.4, 1))
            s.set_color(1, u(0, 1), u(0, 1), u(0, 1), u(0.4, 1))
        if rng.integers(0, 4) == 0:
            s.set_shadow_color(u(0, 1), u(0, 1), u(0, 1), u(0.3, 0.9))
            s.set_shadow_offset(u(-6, 6), u(-6, 6))
            s.set_shadow_blur(u(0, 6))
        if rng.integers(0, 5) == 0:
            s.set_composite_op(ops[int(rng.integers(0, len(ops)))])
        text = "".join(TEXT_CHARS[int(i)] for i in rng.integers(0, len(TEXT_CHARS),
                                                                  size=int(rng.integers(1, 9))))
        maximum = [1.0e30, u(5, size), 0.0][int(rng.choice(3, p=[0.6, 0.3, 0.1]))]
        if rng.integers(0, 2):
            s.fill_text(text, 0.0, 0.0, maximum)
        else:
            s.set_line_width(u(0.5, 5))
            s.set_line_join(int(rng.integers(0, 3)))
            s.set_line_cap(int(rng.integers(0, 3)))
            if rng.integers(0, 3) == 0:
                s.set_line_dash([u(1, 12) for _ in range(int(rng.integers(1, 4)))])
                s.set_line_dash_offset(u(0, 10))
            s.stroke_text(text, 0.0, 0.0, maximum)
        s.restore()
    return s.tobytes()


def put_scene(seed, size=48):
    """put_image_data mixed with draws: at the start, between draws,
    overlapping puts, under a clip (puts ignore it), off the canvas, and
    before a whole-canvas operator -- hpp:3383-3408 replaces pixels in
    painter's order whatever the clip, alpha or operator."""
    rng = np.random.default_rng(seed)
    s = Stream(size, size)
    s.define_image(0, rng.integers(0, 256, size=(10, 12, 4), dtype=np.uint8))
    s.define_image(1, rng.integers(0, 256, size=(7, 5, 4), dtype=np.uint8))
    s.put_image_data(0, -3, 4)
    s.set_color(0, 1.0, 0.5, 0.0, 0.7)
    s.fill_rectangle(5.0, 5.0, 20.0, 30.0)
    s.put_image_data(1, 10, 10)
    s.put_image_data(0, 12, 8)
    s.begin_path()
    s.rectangle(0.0, 0.0, 30.0, 30.0)
    s.clip()
    s.set_color(0, 0.0, 0.2, 1.0, 0.5)
    s.fill_rectangle(8.0, 2.0, 30.0, 20.0)
    s.put_image_data(1, size - 3, size - 2)
    s.put_image_data(1, 1 << 22, 0)
    s.set_composite_op(COMPOSITE["source_copy"])
    s.fill_rectangle(20.0, 20.0, 10.0, 10.0)
    return s.tobytes()


def single_draw(stream, k):
    """Draw k (1-based) of a command stream on its own: (the stream with every
    state and path record before that draw and the draw itself, the same
    without the draw).  Paths start with begin_path, so dropping the earlier
    draws leaves the state and path draw k sees unchanged."""
    from canvas_ity_amd.stream import DRAW_OPS, records
    words = np.frombuffer(stream, dtype=np.uint32)
    keep = [words[:4]]
    seen = 0
    for code, at, n in records(stream):
        if code in DRAW_OPS:
            seen += 1
            if seen == k:
                before = np.concatenate(keep).tobytes()
                return np.concatenate(keep + [words[at:at + 1 + n]]).tobytes(), before
            continue
        keep.append(words[at:at + 1 + n])
    raise IndexError("the stream has %d draws" % seen)


def round_join_strokes():
    """Strokes whose round joins go through acosf/tanf (canvas_ity.hpp:
    1994-1997): draw 157 of batch-mix canvas 40 (seed 0xC4), whose join a
    correctly rounded acos moves by a few ulps, and seeded five-segment paths
    with round joins and each cap, open and closed."""
    import canvas_ity_amd as cia
    out = {"mix40_157": single_draw(cia.scene(4, cia.scene_seed(0xC4, 40)), 157)}
    for seed in range(6):
        rng = np.random.default_rng(0x501 + seed)
        s = Stream(96, 80)
        s.set_line_width(float(rng.uniform(2, 14))).set_line_join(2).set_line_cap(seed % 3)
        s.begin_path().move_to(*rng.uniform(-5, 100, 2))
        for _ in range(5):
            if seed % 2:
                s.line_to(*rng.uniform(-5, 100, 2))
            else:
                s.bezier_curve_to(*rng.uniform(-5, 100, 6))
        if seed % 3 == 1:
            s.close_path()
        before = s.tobytes()
        s.set_color(1, 0.2, 0.5, 0.9, 0.8)  # stroke_style
        s.stroke()
        out["seeded_%d" % seed] = (s.tobytes(), before)
    return out


# (configuration, seed, width, height, shapes, flags): two seeded scenes per
# configuration at small sizes and two of configuration 2 with even-odd fills
# (flags 1), rendered bit for bit as the oracle renders them
BIT_EXACT_SCENES = [(1, 0xa00641a9f2, 124, 119, 26, 0), (2, 0x39a72dabd6, 134, 52, 17, 0),
                    (3, 0xdfa132d749, 150, 48, 23, 0), (4, 0xcc0cbdf921, 139, 57, 22, 0),
                    (5, 0x4d93887ad1, 86, 70, 31, 0), (1, 0x71f08551fe, 76, 86, 24, 0),
                    (2, 0x8db200a195, 105, 127, 33, 0), (3, 0x9f472352b3, 136, 75, 39, 0),
                    (4, 0x371e788978, 142, 60, 35, 0), (5, 0xb3fc88f1e, 116, 83, 9, 0),
                    (2, 0x20e9fa1023, 140, 95, 27, 1), (2, 0xed9f3cd366, 116, 68, 26, 1)]
