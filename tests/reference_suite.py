"""The reference's test suite (test/test.cpp:326-2163) as command streams.

Each function below restates one test of the reference's suite -- the same
canvas calls with the same arguments, every computed coordinate in float32 as
the C++ computes it -- against `Canvas`, a recorder with the canvas API that
writes a command stream (include/cih_stream.h).  The suite's own pass criterion
is an image hash within Hamming distance 5 of its golden table
(test.cpp:2186-2261, 2614-2618; tests/golden/reference_hashes.json), so each
restatement is checked against the reference's golden vector through the
oracle before it is used to test the device.

Three tests branch on what the canvas answers (is_point_in_path twice,
get_image_data once).  A stream cannot branch, so the recorder answers those
queries from an oracle canvas that has replayed the stream so far; the device
then draws what the oracle decided.  The text and clip-dependent tests live in
tests/scenes.py (reference_text_tests, reference_clip_tests).

TEST INFRASTRUCTURE ONLY.
"""
import ctypes
import ctypes.util

import numpy as np

import oracle
import scenes
from canvas_ity_amd.stream import Stream

F = np.float32
FILL, STROKE = 0, 1
BUTT, SQUARE, CIRCLE = 0, 1, 2
MITER, BEVEL, ROUNDED = 0, 1, 2
REPEAT, REPEAT_X, REPEAT_Y, NO_REPEAT = 0, 1, 2, 3
SOURCE_ATOP, DESTINATION_OVER, DESTINATION_ATOP = 13, 11, 7
LIGHTER, SOURCE_OVER = 10, 14

_m = ctypes.CDLL(ctypes.util.find_library("m"))
for _name, _n in (("cosf", 1), ("sinf", 1), ("sqrtf", 1), ("floorf", 1), ("fabsf", 1),
                  ("fmodf", 2), ("hypotf", 2)):
    getattr(_m, _name).restype = ctypes.c_float
    getattr(_m, _name).argtypes = [ctypes.c_float] * _n


def _libm(name):
    fn = getattr(_m, name)
    return lambda *args: F(fn(*[float(a) for a in args]))


cosf, sinf, sqrtf, floorf = _libm("cosf"), _libm("sinf"), _libm("sqrtf"), _libm("floorf")
fabsf, fmodf, hypotf = _libm("fabsf"), _libm("fmodf"), _libm("hypotf")


def frange(start, stop, step=1.0, inclusive=False):
    """for (float v = start; v < stop (or <=); v += step) in float32."""
    v, stop, step = F(start), F(stop), F(step)
    while (v <= stop) if inclusive else (v < stop):
        yield v
        v = F(v + step)


def packed(data, width, height, stride, offset=0):
    """The width x height RGBA8 image at data[offset:] with row stride `stride`."""
    data = np.frombuffer(bytes(data), dtype=np.uint8)
    rows = [data[offset + r * stride:offset + r * stride + 4 * width] for r in range(height)]
    return np.stack(rows).reshape(height, width, 4)


class Canvas:
    """Records canvas_ity::canvas calls (canvas_ity.hpp:194-1148) as a stream.
    Field assignments become the stream's field records."""

    def __init__(self, width, height, probe=False):
        self.stream = Stream(width, height)
        self._offset = [0.0, 0.0]
        self._images = {}
        self._probe = oracle.Canvas(width, height, "port") if probe else None
        self._replayed = 4

    def __setattr__(self, name, value):
        s = self.__dict__.get("stream")
        if name == "global_composite_operation":
            s.set_composite_op(value)
        elif name in ("shadow_offset_x", "shadow_offset_y"):
            self._offset[name == "shadow_offset_y"] = value
            s.set_shadow_offset(*self._offset)
        elif name == "line_cap":
            s.set_line_cap(value)
        elif name == "line_join":
            s.set_line_join(value)
        elif name == "line_dash_offset":
            s.set_line_dash_offset(value)
        else:
            object.__setattr__(self, name, value)

    def __getattr__(self, name):
        return getattr(self.__dict__["stream"], name)

    def set_line_dash(self, segments, count):
        self.stream.set_line_dash(None if segments is None else list(segments)[:count])

    def arc(self, x, y, radius, start, end, counter=False):
        self.stream.arc(x, y, radius, start, end, bool(counter))

    def _image(self, image):
        key = (image.shape, image.tobytes())
        if key not in self._images:
            self._images[key] = len(self._images)
            self.stream.define_image(self._images[key], image)
        return self._images[key]

    def set_pattern(self, kind, data, width, height, stride, repetition):
        if data is None:  # a null image is ignored (hpp:2847-2848)
            return
        self.stream.set_pattern(kind, self._image(packed(data, width, height, stride)),
                                repetition)

    def draw_image(self, data, width, height, stride, x, y, to_width, to_height):
        if data is None:  # ignored (hpp:3323-3325)
            return
        self.stream.draw_image(self._image(packed(data, width, height, stride)), x, y,
                               to_width, to_height)

    def put_image_data(self, data, width, height, stride, x, y, offset=0):
        if data is None:
            return
        self.stream.put_image_data(self._image(packed(data, width, height, stride, offset)),
                                   x, y)

    # ---- queries, answered by the oracle on the stream so far ----
    def _catch_up(self):
        words = self.stream.words
        header = [words[0], words[1], words[2], words[3]]
        self._probe.replay(np.asarray(header + words[self._replayed:],
                                      dtype=np.uint32).tobytes())
        self._replayed = len(words)

    def is_point_in_path(self, x, y):
        self._catch_up()
        lib = self._probe.fns["lib"]
        return lib.cih_oracle_is_point_in_path(self._probe.handle, float(x), float(y)) != 0

    def get_image_data_into(self, buffer, offset, width, height, stride, x, y):
        self._catch_up()
        out = (ctypes.c_ubyte * len(buffer)).from_buffer(buffer)
        self._probe.fns["get_image_data"](self._probe.handle,
                                          ctypes.addressof(out) + offset, width, height,
                                          stride, x, y)

    def tobytes(self):
        return self.stream.tobytes()


# ---- transforms (test.cpp:326-420) -------------------------------------------

def scale_uniform(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_line_dash([1.0], 1)
    c.line_cap = CIRCLE
    size = F(8.0)
    while size < min(w, h):
        c.scale(2.0, 2.0)
        c.stroke_rectangle(0.0, 0.0, 8.0, 8.0)
        size = F(size * F(2.0))


def scale_non_uniform(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_line_dash([4.0], 1)
    c.scale(4.0, 0.5)
    c.stroke_rectangle(w * F(0.125) / F(4.0), h * F(0.125) / F(0.5),
                       w * F(0.75) / F(4.0), h * F(0.75) / F(0.5))


def rotate(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    for _ in range(64):
        c.rotate(F(3.14159265) / F(2.0) / F(64.0))
        c.stroke_rectangle(0.0, 0.0, w, h)


def translate(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    for step in frange(0.0, 32.0):
        c.translate((F(0.5) - step / F(32.0)) * w * F(0.2), h / F(32.0))
        c.begin_path()
        c.arc(0.0, 0.0, w * F(0.125), 0.0, 6.28318531)
        c.close_path()
        c.stroke()


def transform(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    for _ in range(8):
        c.transform(1.0, 0.0, 0.1, 1.0, w * F(-0.05), 0.0)
        c.stroke_rectangle(w * F(0.25), h * F(0.25), w * F(0.5), h * F(0.5))


def transform_fill(c, w, h):
    checker = bytes(((i >> 5 & 1) ^ (i >> 9 & 1) ^ ((i & 3) == 3)) * 255 for i in range(1024))
    c.set_pattern(FILL, checker, 16, 16, 64, REPEAT)
    c.begin_path()
    c.rectangle(w * F(0.2), h * F(0.2), w * F(0.6), h * F(0.6))
    c.transform(1.0, 0.5, -0.5, 1.0, 0.0, 0.0)
    c.fill()


def transform_stroke(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_line_width(8.0)
    c.set_line_dash([22.0, 8.0, 10.0, 8.0], 4)
    c.begin_path()
    c.arc(w * F(0.5), h * F(0.5), min(w, h) * F(0.4), 0.0, 6.28318531)
    c.close_path()
    c.transform(1.0, 1.0, 0.0, 2.0, 0.0, 0.0)
    c.stroke()


def set_transform(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    for _ in range(8):
        c.set_transform(1.0, 0.0, 0.1, 1.0, w * F(-0.05), 0.0)
    c.stroke_rectangle(w * F(0.25), h * F(0.25), w * F(0.5), h * F(0.5))


# ---- compositing and shadows (test.cpp:422-577) -------------------------------

def global_alpha(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_line_width(3.0)
    for y in frange(0.0, 6.0):
        for x in frange(0.0, 6.0):
            c.set_color(FILL, x / F(5.0), 1.0, y / F(5.0), x / F(5.0))
            c.set_global_alpha(y / F(4.0) - F(0.25))
            c.begin_path()
            c.rectangle((x + F(0.1)) / F(6.0) * w, (y + F(0.1)) / F(6.0) * h,
                        F(0.8) / F(6.0) * w, F(0.8) / F(6.0) * h)
            c.fill()
            c.stroke()


def shadow_color(c, w, h):
    c.shadow_offset_x = 5.0
    c.shadow_offset_y = 5.0
    c.set_shadow_blur(1.0)
    for row in frange(0.0, 5.0):
        y = (row + F(0.25)) * F(0.2) * h
        c.set_color(FILL, 0.0, 0.0, 0.0, F(0.25) * row)
        for color, x in (((1.0, -1.0, 0.0, 0.25), 0.05), ((0.0, 1.0, 0.0, 0.5), 0.30),
                         ((0.0, 0.0, 2.0, 0.75), 0.55), ((1.0, 1.0, 1.0, 100.0), 0.80)):
            c.set_shadow_color(*color)
            c.fill_rectangle(F(x) * w, y, F(0.15) * w, F(0.1) * h)


def shadow_offset(c, w, h):
    c.set_shadow_blur(2.0)
    c.set_color(FILL, 1.0, 1.0, 1.0, 1.0)
    c.set_shadow_color(0.0, 0.0, 0.0, 1.0)
    for y in frange(0.0, 5.0):
        for x in frange(0.0, 5.0):
            c.shadow_offset_x = (x - F(2.0)) * F(4.0)
            c.shadow_offset_y = (y - F(2.0)) * F(4.0)
            c.fill_rectangle((x + F(0.25)) * F(0.2) * w, (y + F(0.25)) * F(0.2) * h,
                             F(0.1) * w, F(0.1) * h)


def shadow_offset_offscreen(c, w, h):
    c.shadow_offset_x = w
    c.set_color(FILL, 1.0, 0.0, 0.0, 1.0)
    c.set_shadow_color(0.0, 0.0, 0.0, 0.5)
    c.fill_rectangle(w * F(-0.6875), h * F(0.0625), w * F(0.375), h * F(0.375))
    c.begin_path()
    c.arc(w * F(0.5), h * F(0.75), min(w, h) * F(0.2), 0.0, 6.28318531)
    c.close_path()
    c.fill()


def shadow_blur(c, w, h):
    c.set_color(FILL, 1.0, 1.0, 1.0, 1.0)
    c.set_shadow_color(0.0, 0.0, 0.0, 1.0)
    c.shadow_offset_x = 5.0
    c.shadow_offset_y = 5.0
    for x in frange(0.0, 5.0):
        y = F(4.0)
        while y >= F(0.0):
            c.set_shadow_blur((y * F(5.0) + x) * F(0.5) - F(0.5))
            c.fill_rectangle((x + F(0.25)) * F(0.2) * w, (y + F(0.25)) * F(0.2) * h,
                             F(0.1) * w, F(0.1) * h)
            y = F(y - F(1.0))


def shadow_blur_offscreen(c, w, h):
    c.set_color(FILL, 0.0, 0.0, 0.0, 1.0)
    c.set_shadow_color(0.0, 0.0, 0.0, 1.0)
    c.set_shadow_blur(5.0)
    c.fill_rectangle(0.0, h * F(2.0), w, h)
    c.fill_rectangle(0.0, h * F(-2.0), w, h)
    c.fill_rectangle(w + F(1.0), 0.0, w, h)
    c.fill_rectangle(-w - F(1.0), 0.0, w, h)


# ---- line styles (test.cpp:579-868) ---------------------------------------------

def line_width(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_line_width(4.0)
    for step in frange(0.0, 16.0):
        left = (step + F(0.25)) / F(16.0) * w
        right = (step + F(0.75)) / F(16.0) * w
        c.begin_path()
        c.move_to(left, 0.0)
        c.bezier_curve_to(left, F(0.5) * h, right, F(0.5) * h, right, h)
        c.set_line_width(F(0.5) * (step - F(1.0)))
        c.stroke()
    c.set_color(FILL, 1.0, 1.0, 1.0, 1.0)
    c.global_composite_operation = SOURCE_ATOP
    c.fill_rectangle(0.0, F(0.5) * h, w, F(0.5) * h)
    c.global_composite_operation = DESTINATION_OVER
    c.fill_rectangle(0.0, F(0.25) * h, w, F(0.25) * h)
    c.set_color(FILL, 0.0, 0.0, 0.0, 1.0)
    c.fill_rectangle(0.0, F(0.5) * h, w, F(0.25) * h)


def line_width_angular(c, w, h):
    for step in frange(0.0, 5.0):
        grey = (step + F(1.0)) / F(5.0)
        c.set_color(STROKE, grey, grey, grey, 1.0)
        c.begin_path()
        c.move_to(F(0.1) * w, F(0.1) * h)
        c.bezier_curve_to(F(1.2) * w, F(1.0) * h, F(1.2) * w, F(-0.0) * h,
                          F(0.1) * w, F(0.9) * h)
        c.set_line_width(F(30.0) - F(7.0) * step)
        c.stroke()


def line_cap(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_line_width(24.0)
    for index, cap in enumerate((BUTT, SQUARE, CIRCLE)):
        right = F(index + 1) / F(3.0) * w - F(20.0)
        c.begin_path()
        c.move_to(right, F(0.125) * h)
        c.bezier_curve_to(right, F(0.125) * h + F(100.0), right - F(100.0), F(0.875) * h,
                          right, F(0.875) * h)
        c.line_cap = cap
        c.stroke()


def line_cap_offscreen(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_line_width(36.0)
    for index, cap in enumerate((BUTT, SQUARE, CIRCLE)):
        x = (F(index) + F(0.5)) / F(3.0) * w
        y = (F(index) + F(0.5)) / F(3.0) * h
        c.begin_path()
        c.move_to(x, -19.0)
        c.line_to(x, -9.0)
        c.move_to(x, h + F(17.0))
        c.line_to(x, h + F(27.0))
        c.move_to(-27.0, y)
        c.line_to(-17.0, y)
        c.move_to(w + F(9.0), y)
        c.line_to(w + F(19.0), y)
        c.line_cap = cap
        c.stroke()


def line_join(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_line_width(16.0)
    for index, join in enumerate((MITER, BEVEL, ROUNDED)):
        left = (F(index) + F(0.25)) / F(3.0) * w
        right = (F(index) + F(0.75)) / F(3.0) * w
        c.begin_path()
        c.move_to(left, F(0.2) * h)
        for x, y in ((left, 0.1), (left, 0.2), (right, 0.2), (left, 0.2), (left, 0.3),
                     (right, 0.3), (right, 0.4), (right, 0.5), (left, 0.4), (left, 0.5),
                     (right, 0.6)):
            c.line_to(x, F(y) * h)
        c.bezier_curve_to(right, h, left, F(0.4) * h, left, F(0.7) * h)
        c.bezier_curve_to(left, F(0.8) * h, right, F(0.8) * h, right, F(0.9) * h)
        c.line_join = join
        c.stroke()


def line_join_offscreen(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_line_width(36.0)
    for index, join in enumerate((MITER, BEVEL, ROUNDED)):
        x = (F(index) + F(0.5)) / F(3.0) * w
        y = (F(index) + F(0.5)) / F(3.0) * h
        c.begin_path()
        c.move_to(x - F(10.0), -55.0)
        c.line_to(x - F(10.0), -5.0)
        c.line_to(x + F(10.0), -55.0)
        c.move_to(x - F(10.0), h + F(130.0))
        c.line_to(x + F(10.0), h + F(80.0))
        c.line_to(x + F(10.0), h + F(130.0))
        c.move_to(-130.0, y - F(10.0))
        c.line_to(-80.0, y - F(10.0))
        c.line_to(-130.0, y + F(10.0))
        c.move_to(h + F(55.0), y - F(10.0))
        c.line_to(h + F(5.0), y + F(10.0))
        c.line_to(h + F(55.0), y + F(10.0))
        c.line_join = join
        c.stroke()


def miter_limit(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    for line in frange(0.0, 4.0):
        c.set_line_width(F(1.5) * line + F(1.0))
        c.set_miter_limit(20.0)
        for limit in frange(0.0, 8.0):
            left = (limit + F(0.2)) / F(8.0) * w
            middle = (limit + F(0.5)) / F(8.0) * w
            right = (limit + F(0.7)) / F(8.0) * w
            top = (line + F(0.3)) / F(4.0) * h
            bottom = (line + F(0.7)) / F(4.0) * h
            c.begin_path()
            c.move_to(left, bottom)
            c.line_to(left, top)
            c.line_to(right, bottom)
            c.line_to(middle, top)
            c.set_miter_limit(F(1.5) * limit)
            c.stroke()


def line_dash_offset(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_line_width(6.0)
    c.set_line_dash([20.0, 8.0, 8.0, 8.0], 4)
    for step in frange(0.0, 16.0):
        left = (step + F(0.125)) / F(16.0) * w
        right = (step + F(0.875)) / F(16.0) * w
        c.begin_path()
        c.move_to(left, 0.0)
        c.line_to(right, F(0.125) * h)
        c.line_to(left, F(0.375) * h)
        c.line_to(right, F(0.625) * h)
        c.line_to(left, F(0.875) * h)
        c.line_to(right, h)
        c.line_dash_offset = (step / F(16.0) - F(0.5)) * F(44.0)
        c.stroke()


def line_dash(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_line_width(6.0)
    c.set_line_dash([10.0], 1)
    c.stroke()
    c.move_to(0.0, 0.0)
    c.stroke()
    for x, dash in ((0.25, None), (0.375, [20.0, -8.0]), (0.5, [20.0, 8.0, 8.0, 8.0]),
                    (0.625, [0.0, 8.0, 2.0, 8.0]), (0.75, "null")):
        if dash == "null":
            c.set_line_dash(None, 0)
        elif dash is not None:
            c.set_line_dash(dash, len(dash))
        c.begin_path()
        c.move_to(w * F(x), 0.0)
        c.line_to(w * F(x), h)
        c.stroke()


def line_dash_closed(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_line_width(32.0)
    c.set_line_dash([96.0, 32.0], 2)
    c.line_dash_offset = -80.0
    c.stroke_rectangle(F(0.25) * w, F(0.25) * h, F(0.5) * w, F(0.5) * h)
    c.set_line_dash([96.0, 32.0, 1024.0, 16.0], 4)
    c.line_dash_offset = 128.0
    c.stroke_rectangle(F(0.09375) * w, F(0.09375) * h, F(0.8125) * w, F(0.8125) * h)


def line_dash_overlap(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_color(FILL, 1.0, 0.0, 0.0, 1.0)
    c.line_cap = CIRCLE
    c.set_line_width(16.0)
    c.set_line_dash([14.0, 12.0], 2)
    for index in range(4):
        flip = F(-1.0 if index == 3 else 1.0)
        top_y = F(0.25 if index & 1 else 0.1) * h
        bottom_y = F(0.9 if index & 1 else 0.75) * h
        mid_x = F(0.75 if index & 2 else 0.25) * w
        top_width = F(0.25 if index & 1 else 0.55) * flip * w
        bottom_width = F(0.55 if index & 1 else 0.25) * flip * w
        c.move_to(mid_x, top_y)
        c.bezier_curve_to(mid_x - top_width, top_y, mid_x + bottom_width, bottom_y,
                          mid_x, bottom_y)
        c.bezier_curve_to(mid_x - bottom_width, bottom_y, mid_x + top_width, top_y,
                          mid_x, top_y)
        c.close_path()
    c.fill()
    c.stroke()


def line_dash_offscreen(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_line_width(6.0)
    c.set_line_dash([0.0, w * F(20.5) * F(3.14159265) - h * F(0.5) + F(1.0), h - F(2.0), 0.0],
                    4)
    for step in frange(-2.0, 2.0, inclusive=True):
        c.begin_path()
        c.arc(w * F(-20.0), h * F(0.5), w * (F(20.5) - step * F(0.1)), 3.14159265,
              1.5707963268)
        c.line_dash_offset = w * step * F(0.1) * F(3.14159265)
        c.stroke()


# ---- fill and stroke styles (test.cpp:870-994) ------------------------------------

def color(c, w, h):
    radius = min(w, h) * F(0.4)
    c.set_color(FILL, 2.0, -1.0, 0.0, 0.5)
    c.set_color(STROKE, 0.0, 0.0, 1.0, 1.5)
    c.set_line_width(16.0)
    c.arc(F(0.5) * w, F(0.5) * h, radius, 0.0, 6.28318531)
    c.close_path()
    c.fill()
    c.stroke()


def linear_gradient(c, w, h):
    radius = min(w, h) * F(0.4)
    c.set_linear_gradient(FILL, F(0.3) * w, F(0.3) * h, F(0.7) * w, F(0.7) * h)
    c.add_color_stop(FILL, 0.0, 0.0, 1.0, 0.0, 0.5)
    c.add_color_stop(FILL, 1.0, 1.0, 0.0, 1.0, 100.0)
    c.set_linear_gradient(STROKE, F(0.3) * w, F(0.7) * h, F(0.7) * w, F(0.3) * h)
    c.add_color_stop(STROKE, 0.0, 0.0, 0.0, 1.0, 0.5)
    c.add_color_stop(STROKE, 1.0, 1.0, 1.0, 0.0, 1.0)
    c.set_line_width(16.0)
    c.arc(F(0.5) * w, F(0.5) * h, radius, 0.0, 6.28318531)
    c.close_path()
    c.fill()
    c.stroke()
    c.set_linear_gradient(STROKE, F(0.5) * w, F(0.5) * h, F(0.5) * w, F(0.5) * h)
    c.add_color_stop(STROKE, 0.0, 1.0, 0.0, 0.0, 0.5)
    c.add_color_stop(STROKE, 1.0, 0.0, 1.0, 1.0, 1.0)
    c.stroke()


def radial_gradient(c, w, h):
    radius = min(w, h) * F(0.4)
    c.set_radial_gradient(FILL, 0.0, 0.0, radius, w, h, F(0.5) * radius)
    c.add_color_stop(FILL, 0.0, 0.0, 1.0, 0.0, 0.5)
    c.add_color_stop(FILL, 1.0, 1.0, 0.0, 1.0, 1.0)
    c.set_radial_gradient(STROKE, 0.0, h, radius, w, 0.0, F(0.5) * radius)
    c.add_color_stop(STROKE, 0.0, 0.0, 0.0, 1.0, 0.5)
    c.add_color_stop(STROKE, 1.0, 1.0, 1.0, 0.0, 1.0)
    c.set_line_width(16.0)
    c.arc(F(0.5) * w, F(0.5) * h, radius, 0.0, 6.28318531)
    c.close_path()
    c.fill()
    c.stroke()
    c.set_radial_gradient(STROKE, F(0.5) * w, F(0.4) * h, 10.0, F(0.5) * w, F(0.6) * h, 0.0)
    c.set_radial_gradient(STROKE, 0.0, F(0.5) * h, -10.0, w, F(0.5) * h, 10.0)
    c.add_color_stop(STROKE, 0.0, 1.0, 0.0, 0.0, 0.5)
    c.add_color_stop(STROKE, 1.0, 0.0, 1.0, 1.0, 1.0)
    c.stroke()
    c.set_radial_gradient(FILL, F(0.5) * w, F(0.5) * h, 0.0, F(0.5) * w, F(0.5) * h, radius)
    c.add_color_stop(FILL, 0.15, 0.0, 0.0, 0.0, 1.0)
    c.add_color_stop(FILL, 0.20, 0.0, 0.0, 0.0, 0.0)
    c.fill()


def color_stop(c, w, h):
    c.add_color_stop(FILL, 0.5, 1.0, 0.0, 1.0, 1.0)
    c.set_linear_gradient(FILL, F(0.1) * w, 0.0, F(0.9) * w, 0.0)
    c.fill_rectangle(0.0, 0.0, w, F(0.1) * h)
    for stop in ((-1.0, 0.0, 1.0, 0.0, 1.0), (2.0, 1.0, 0.0, 0.0, 1.0),
                 (0.3, -1.0, 0.0, 2.0, 2.0), (0.3, 1.0, 1.0, 1.0, 1.0),
                 (0.3, 0.0, 0.0, 0.0, 1.0), (0.0, 0.0, 0.0, 0.8, 1.0),
                 (1.0, 1.0, 1.0, 1.0, 1.0), (0.7, 0.9, 0.9, 0.9, 1.0),
                 (0.6, 0.1, 0.1, 0.1, 1.0)):
        c.add_color_stop(FILL, *stop)
    c.fill_rectangle(0.0, F(0.1) * h, w, F(0.4) * h)
    c.fill_rectangle(0.0, F(0.5) * h, w, F(0.4) * h)


def pattern(c, w, h):
    checker = bytes((((i >> 2 & 1) ^ (i >> 5 & 1)) | ((i & 3) == 3)) * 255 for i in range(256))
    c.arc(F(0.5) * w, F(0.5) * h, 32.0, 0.0, 6.28318531)
    c.close_path()
    c.set_line_width(20.0)
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_pattern(STROKE, None, 8, 8, 32, REPEAT)
    c.stroke()
    c.set_line_width(16.0)
    c.set_pattern(STROKE, checker, 8, 8, 32, REPEAT)
    c.stroke()
    scale = F(8.0)
    while scale >= F(1.0):
        c.set_transform(1.0, 0.0, 0.0, 1.0, 0.0, 0.0)
        c.scale(scale, scale)
        size_x = F(0.5) * w / scale
        size_y = F(0.5) * h / scale
        c.set_pattern(FILL, checker, 8, 8, 32, NO_REPEAT)
        c.fill_rectangle(0.0, 0.0, size_x, size_y)
        c.set_pattern(FILL, checker, 8, 8, 32, REPEAT_X)
        c.fill_rectangle(size_x, 0.0, size_x, size_y)
        c.set_pattern(FILL, checker, 8, 8, 32, REPEAT_Y)
        c.fill_rectangle(0.0, size_y, size_x, size_y)
        c.set_pattern(FILL, checker, 8, 8, 32, REPEAT)
        c.fill_rectangle(size_x, size_y, size_x, size_y)
        scale = F(scale / F(2.0))


# ---- path building (test.cpp:996-1221) ---------------------------------------------

def begin_path(c, w, h):
    c.move_to(0.0, 0.0)
    c.line_to(w, h)
    c.stroke()
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.begin_path()
    c.begin_path()
    c.move_to(w, 0.0)
    c.line_to(0.0, h)
    c.stroke()
    c.begin_path()
    c.line_to(F(0.5) * w, h)
    c.stroke()


def _styles(c, width):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_color(FILL, 1.0, 0.0, 0.0, 1.0)
    c.set_line_width(width)


def move_to(c, w, h):
    _styles(c, 8.0)
    c.move_to(F(0.6) * w, h)
    c.move_to(F(0.4) * w, F(0.1) * h)
    c.line_to(F(0.2) * w, F(0.5) * h)
    c.line_to(F(0.4) * w, F(0.9) * h)
    c.move_to(F(0.6) * w, F(0.2) * h)
    c.line_to(F(0.8) * w, F(0.4) * h)
    c.move_to(F(0.8) * w, F(0.6) * h)
    c.line_to(F(0.6) * w, F(0.8) * h)
    c.move_to(F(0.7) * w, F(0.5) * h)
    c.line_to(F(0.7) * w, F(0.5) * h)
    c.fill()
    c.stroke()


def close_path(c, w, h):
    _styles(c, 8.0)
    c.close_path()
    c.line_to(F(0.5) * w, F(0.5) * h)
    c.line_to(F(0.2) * w, F(0.8) * h)
    c.line_to(F(0.2) * w, F(0.2) * h)
    c.close_path()
    c.line_to(F(0.5) * w, F(0.2) * h)
    c.line_to(F(0.8) * w, F(0.2) * h)
    c.close_path()
    c.close_path()
    c.move_to(F(0.5) * w, F(0.8) * h)
    c.line_to(F(0.8) * w, F(0.8) * h)
    c.line_to(F(0.8) * w, F(0.5) * h)
    c.close_path()
    c.fill()
    c.stroke()


def line_to(c, w, h):
    _styles(c, 16.0)
    for x, y in ((0.1, 0.2), (0.1, 0.2), (0.2, 0.5), (0.2, 0.5), (0.3, 0.8), (0.4, 0.2),
                 (0.4, 0.2), (0.6, 0.8), (0.6, 0.8)):
        c.line_to(F(x) * w, F(y) * h)
    c.move_to(F(0.7) * w, F(0.4) * h)
    for x, y in ((0.9, 0.4), (0.9, 0.6), (0.7, 0.6), (0.7, 0.4)):
        c.line_to(F(x) * w, F(y) * h)
    c.fill()
    c.stroke()


def quadratic_curve_to(c, w, h):
    _styles(c, 8.0)

    def q(a, b, x, y):
        c.quadratic_curve_to(F(a) * w, F(b) * h, F(x) * w, F(y) * h)
    q(0.1, 0.2, 0.1, 0.2)
    q(0.2, 0.5, 0.2, 0.5)
    q(0.3, 0.8, 0.4, 0.2)
    q(0.6, 0.8, 0.7, 0.2)
    c.move_to(F(0.7) * w, F(0.6) * h)
    q(0.9, 0.6, 0.9, 0.8)
    q(0.9, 0.9, 0.7, 0.9)
    c.close_path()
    c.move_to(F(0.1) * w, F(0.9) * h)
    q(0.5, 0.5, 0.1, 0.9)
    c.fill()
    c.stroke()


def bezier_curve_to(c, w, h):
    _styles(c, 8.0)

    def b(*v):
        c.bezier_curve_to(*[F(v[k]) * (w if k % 2 == 0 else h) for k in range(6)])

    def m(x, y):
        c.move_to(F(x) * w, F(y) * h)
    b(0.9, 0.9, 0.6, 0.6, 0.6, 0.9)
    m(0.1, 0.1)
    b(0.9, 0.9, 0.9, 0.1, 0.1, 0.9)
    m(0.4, 0.1)
    b(0.1, 0.3, 0.7, 0.3, 0.4, 0.1)
    m(0.9, 0.1)
    b(0.6, 0.2, 0.9, 0.1, 0.6, 0.2)
    m(0.7, 0.3)
    b(0.9, 0.3, 0.9, 0.4, 0.8, 0.5)
    b(0.7, 0.6, 0.7, 0.7, 0.9, 0.7)
    c.fill()
    c.stroke()


def arc_to(c, w, h):
    radius = min(w, h) * F(0.5)
    _styles(c, 8.0)

    def a(vx, vy, x, y, r):
        c.arc_to(F(vx) * w, F(vy) * h, F(x) * w, F(y) * h, r)
    a(0.3, 0.3, 0.5, 0.5, 16.0)
    c.move_to(F(0.4) * w, F(0.4) * h)
    a(0.7, 0.1, 0.7, 0.4, 0.0)
    a(0.9, 0.5, 0.7, 0.7, F(0.125) * radius)
    a(0.5, 0.9, 0.3, 0.8, F(0.25) * radius)
    a(0.1, 0.7, 0.4, 0.4, F(0.375) * radius)
    c.close_path()
    c.move_to(F(0.1) * w, F(0.6) * h)
    c.transform(1.0, -1.0, 0.0, 1.0, 0.0, F(0.2) * h)
    a(0.1, 0.9, 0.5, 0.9, F(0.3) * radius)
    c.set_transform(1.0, 0.0, 0.0, 1.0, 0.0, 0.0)
    c.close_path()
    c.move_to(F(0.2) * w, F(0.1) * h)
    a(0.1, 0.1, 0.1, 0.7, F(0.6) * radius)
    a(0.2, 0.4, 0.2, 0.4, F(0.5) * radius)
    a(0.4, 0.2, 0.2, 0.4, F(0.5) * radius)
    a(0.5, 0.5, 0.9, 0.1, -1.0)
    c.move_to(F(0.6) * w, F(0.9) * h)
    c.set_transform(0.0, 0.0, 0.0, 1.0, F(0.9) * w, 0.0)
    a(0.9, 0.9, 0.9, 0.6, F(0.3) * radius)
    c.set_transform(1.0, 0.0, 0.0, 1.0, 0.0, 0.0)
    a(0.9, 0.6, 0.9, 0.6, 0.0)
    c.fill()
    c.stroke()


def arc(c, w, h):
    _styles(c, 8.0)
    for i in range(4):
        for j in range(3):
            x = (F(j) + F(0.5)) * w / F(3.0)
            y = (F(i) + F(0.5)) * h / F(4.0)
            radius = min(w, h) * F(0.1)
            start = (F(3.14159265) + F(1.0e-6)) * F(i % 2)
            end = (F(3.14159265) + F(1.0e-6)) * (F(1.0) + F(0.5) * F(j))
            counter = i // 2 != 0
            c.begin_path()
            c.arc(x, y, -radius, start, end, counter)
            c.arc(x, y, radius, start, end, counter)
            c.close_path()
            c.fill()
            c.stroke()


def rectangle(c, w, h):
    _styles(c, 8.0)
    c.move_to(F(0.3) * w, F(0.3) * h)
    c.line_to(F(0.7) * w, F(0.3) * h)
    c.line_to(F(0.7) * w, F(0.7) * h)
    c.line_to(F(0.3) * w, F(0.7) * h)
    c.close_path()
    c.move_to(0.0, 0.0)
    for y in frange(-1.0, 1.0, inclusive=True):
        for x in frange(-1.0, 1.0, inclusive=True):
            c.rectangle((F(0.5) + F(0.1) * x) * w, (F(0.5) + F(0.1) * y) * h,
                        x * F(0.3) * w, y * F(0.3) * h)
    c.line_to(w, h)
    c.fill()
    c.stroke()


# ---- fill and stroke (test.cpp:1223-1394) --------------------------------------------

def _star_polygon(c, w, h, radius):
    c.begin_path()
    for step in frange(0.0, 128.0):
        angle = step * (F(59.0) / F(128.0) * F(6.28318531))
        c.line_to(cosf(angle) * radius + w / F(2.0), sinf(angle) * radius + h / F(2.0))
    c.close_path()


def fill(c, w, h):
    radius = min(w, h) * F(0.45)
    c.set_color(FILL, 0.0, 0.0, 0.0, 1.0)
    c.fill()
    _star_polygon(c, w, h, radius)
    c.fill()
    c.set_color(FILL, 1.0, 0.0, 0.0, 1.0)
    c.scale(0.0, 1.0)
    c.fill()


def fill_rounding(c, w, h):
    c.set_color(FILL, 0.0, 0.0, 0.0, 1.0)
    c.begin_path()
    c.move_to(4.00000191, 4.00000763)
    c.line_to(3.99999809, 192.0)
    c.line_to(28.0000019, 192.0)
    c.close_path()
    c.move_to(-10390.0664, 52.3311195)
    c.line_to(-10389.9941, 47.6248589)
    c.line_to(-10395.9941, 47.5328255)
    c.line_to(-10396.0664, 52.2478294)
    c.close_path()
    c.move_to(110.0, 256.0)
    c.line_to(124.086205, 255.998276)
    c.line_to(123.203453, 0.0)
    c.close_path()
    c.fill()


def fill_converging(c, w, h):
    radius = min(w, h) * F(0.48)
    c.set_color(FILL, 0.0, 0.0, 0.0, 1.0)
    for step in frange(0.0, 256.0):
        angle_1 = (step + F(0.0)) / F(256.0) * F(6.28318531)
        angle_2 = (step + F(0.5)) / F(256.0) * F(6.28318531)
        c.move_to(w / F(2.0) + F(0.5), h / F(2.0) + F(0.5))
        c.line_to(cosf(angle_1) * radius + w / F(2.0) + F(0.5),
                  sinf(angle_1) * radius + h / F(2.0) + F(0.5))
        c.line_to(cosf(angle_2) * radius + w / F(2.0) + F(0.5),
                  sinf(angle_2) * radius + h / F(2.0) + F(0.5))
        c.close_path()
    c.fill()


def fill_zone_plate(c, w, h):
    radius = floorf(min(w, h) * F(0.48) / F(4.0)) * F(4.0)
    c.set_color(FILL, 0.0, 0.0, 0.0, 1.0)
    for step in frange(0.0, radius, 2.0):
        inner = sqrtf((step + F(0.0)) / radius) * radius
        outer = sqrtf((step + F(1.0)) / radius) * radius
        c.move_to(w / F(2.0) + inner, h / F(2.0))
        c.arc(w / F(2.0), h / F(2.0), inner, 0.0, 6.28318531)
        c.close_path()
        c.move_to(w / F(2.0) + outer, h / F(2.0))
        c.arc(w / F(2.0), h / F(2.0), outer, 6.28318531, 0.0, True)
        c.close_path()
    c.fill()


def stroke(c, w, h):
    radius = min(w, h) * F(0.45)
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.stroke()
    _star_polygon(c, w, h, radius)
    c.stroke()
    c.set_color(STROKE, 1.0, 0.0, 0.0, 1.0)
    c.scale(0.0, 1.0)
    c.stroke()


def stroke_wide(c, w, h):
    c.scale(w / F(256.0), h / F(256.0))
    c.line_join = ROUNDED
    c.move_to(24.0, 104.0)
    c.bezier_curve_to(112.0, 24.0, 16.0, 24.0, 104.0, 104.0)
    c.move_to(152.0, 104.0)
    c.bezier_curve_to(232.8, 24.0, 151.2, 24.0, 232.0, 104.0)
    c.move_to(24.0, 232.0)
    c.bezier_curve_to(104.0, 152.0, 24.0, 152.0, 104.0, 232.0)
    c.move_to(188.0, 232.0)
    c.bezier_curve_to(196.0, 184.0, 188.0, 184.0, 196.0, 192.0)
    c.set_line_width(40.0)
    c.stroke()
    c.set_color(STROKE, 1.0, 0.0, 0.0, 1.0)
    c.set_line_width(1.0)
    c.stroke()


def stroke_inner_join(c, w, h):
    for index, join in enumerate((MITER, BEVEL, ROUNDED)):
        center = (F(index) + F(0.5)) / F(3.0) * w
        c.begin_path()
        c.move_to(center - F(0.05) * w, F(0.275) * h)
        c.line_to(center, F(0.225) * h)
        c.line_to(center + F(0.025) * w, F(0.25) * h)
        c.move_to(center - F(0.05) * w, F(0.775) * h)
        c.bezier_curve_to(center, F(0.725) * h, center, F(0.725) * h,
                          center + F(0.025) * w, F(0.75) * h)
        c.line_join = join
        c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
        c.set_line_width(F(0.3) * w)
        c.stroke()
        c.set_color(STROKE, 1.0, 0.0, 0.0, 1.0)
        c.set_line_width(1.0)
        c.stroke()


def stroke_spiral(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_line_width(2.0)
    c.begin_path()
    outside = min(w, h) * F(0.48)
    for step in frange(0.0, 2048.0, inclusive=True):
        parameter = (step - F(1024.0)) / F(1024.0)
        angle = fabsf(parameter) * F(12.0) * F(6.28318531)
        radius = parameter * outside
        c.line_to(cosf(angle) * radius + w * F(0.5), sinf(angle) * radius + h * F(0.5))
    c.stroke()


def stroke_long(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    for step in frange(0.0, 29.0, inclusive=True):
        c.move_to(F(0.4) * w, F(-23.0) * h)
        c.line_to(w * step / F(29.0), h)
        c.move_to(F(-23.0) * w, F(0.4) * h)
        c.line_to(w, h * step / F(29.0))
    c.stroke()


# ---- queries and rectangles (test.cpp:1449-1602) ---------------------------------------

def _bit_reversed_probes(c, w, h, dx, dy):
    for index in range(256):
        bits = index
        bits = (bits << 1 & 0xaa) | (bits >> 1 & 0x55)
        bits = (bits << 2 & 0xcc) | (bits >> 2 & 0x33)
        bits = (bits << 4 & 0xf0) | (bits >> 4 & 0x0f)
        x = F(bits) / F(256.0) * w
        y = F(index) / F(256.0) * h
        c.rotate(0.5)
        inside = c.is_point_in_path(x + dx, y - dy)
        c.set_transform(1.0, 0.0, 0.0, 1.0, 0.0, 0.0)
        c.set_color(STROKE, 1.0 - float(inside), float(inside), 0.0, 1.0)
        c.stroke_rectangle(x - F(1.5), y - F(1.5), 3.0, 3.0)


def is_point_in_path(c, w, h):
    c.set_color(FILL, 0.0, 0.0, 1.0, 1.0)
    c.set_color(STROKE, 1.0, 1.0, 1.0, 1.0)
    if c.is_point_in_path(0.0, 0.0):
        c.fill_rectangle(0.0, 0.0, 16.0, 16.0)
    c.scale(w / F(256.0), h / F(256.0))
    c.begin_path()
    c.move_to(65.0, 16.0)
    c.line_to(113.0, 24.0)
    c.bezier_curve_to(113.0, 24.0, 93.0, 126.0, 119.0, 160.0)
    c.bezier_curve_to(133.0, 180.0, 170.0, 196.0, 186.0, 177.0)
    c.bezier_curve_to(198.0, 162.0, 182.0, 130.0, 166.0, 118.0)
    c.bezier_curve_to(123.0, 80.0, 84.0, 124.0, 84.0, 124.0)
    c.line_to(35.0, 124.0)
    c.line_to(18.0, 56.0)
    c.line_to(202.0, 56.0)
    c.line_to(202.0, 90.0)
    c.bezier_curve_to(202.0, 90.0, 240.0, 168.0, 209.0, 202.0)
    c.bezier_curve_to(175.0, 240.0, 65.0, 187.0, 65.0, 187.0)
    c.close_path()
    c.translate(40.0, 160.0)
    c.move_to(110.0, 0.0)
    c.line_to(0.0, 0.0)
    c.line_to(0.0, 0.0)
    c.bezier_curve_to(0.0, 90.0, 110.0, 90.0, 110.0, 40.0)
    c.close_path()
    c.fill()
    c.stroke()
    _bit_reversed_probes(c, w, h, F(0.0), F(0.0))


def is_point_in_path_offscreen(c, w, h):
    c.set_color(FILL, 0.0, 0.0, 1.0, 1.0)
    c.set_color(STROKE, 1.0, 1.0, 1.0, 1.0)
    c.scale(w / F(256.0), h / F(256.0))
    c.begin_path()
    c.move_to(321.0, -240.0)
    c.line_to(369.0, -232.0)
    c.bezier_curve_to(369.0, -232.0, 349.0, -130.0, 375.0, -96.0)
    c.bezier_curve_to(389.0, -76.0, 426.0, -60.0, 442.0, -79.0)
    c.bezier_curve_to(454.0, -94.0, 438.0, -126.0, 422.0, -138.0)
    c.bezier_curve_to(379.0, -176.0, 340.0, -132.0, 340.0, -132.0)
    c.line_to(291.0, -132.0)
    c.line_to(274.0, -200.0)
    c.line_to(458.0, -200.0)
    c.line_to(458.0, -166.0)
    c.bezier_curve_to(458.0, -166.0, 496.0, -88.0, 465.0, -54.0)
    c.bezier_curve_to(431.0, -16.0, 321.0, -69.0, 321.0, -69.0)
    c.close_path()
    c.translate(40.0, 160.0)
    c.move_to(366.0, -256.0)
    c.line_to(256.0, -256.0)
    c.line_to(256.0, -256.0)
    c.bezier_curve_to(256.0, -166.0, 366.0, -166.0, 366.0, -216.0)
    c.close_path()
    c.fill()
    c.stroke()
    _bit_reversed_probes(c, w, h, w, h)


def _rectangle_grid(c, w, h, call):
    for y in frange(-1.0, 1.0, inclusive=True):
        for x in frange(-1.0, 1.0, inclusive=True):
            call((F(0.5) + F(0.05) * x) * w, (F(0.5) + F(0.05) * y) * h,
                 x * F(0.2) * w, y * F(0.2) * h)


def clear_rectangle(c, w, h):
    c.set_color(STROKE, 1.0, 1.0, 1.0, 1.0)
    c.set_color(FILL, 0.4, 0.05, 0.2, 1.0)
    c.move_to(0.0, 0.0)
    c.line_to(w, 0.0)
    c.line_to(w, h)
    c.line_to(0.0, h)
    c.fill()
    c.rotate(0.2)
    c.begin_path()
    c.move_to(F(0.2) * w, F(0.2) * h)
    c.line_to(F(0.8) * w, F(0.2) * h)
    c.line_to(F(0.8) * w, F(0.8) * h)
    c.shadow_offset_x = 5.0
    c.set_shadow_color(0.0, 0.0, 0.0, 1.0)
    c.global_composite_operation = DESTINATION_ATOP
    c.set_global_alpha(0.5)
    _rectangle_grid(c, w, h, c.clear_rectangle)
    c.set_global_alpha(1.0)
    c.global_composite_operation = SOURCE_OVER
    c.set_shadow_color(0.0, 0.0, 0.0, 0.0)
    c.line_to(F(0.2) * w, F(0.8) * h)
    c.close_path()
    c.stroke()


def fill_rectangle(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.set_color(FILL, 0.4, 0.05, 0.2, 1.0)
    c.rotate(0.2)
    c.move_to(F(0.2) * w, F(0.2) * h)
    c.line_to(F(0.8) * w, F(0.2) * h)
    c.line_to(F(0.8) * w, F(0.8) * h)
    _rectangle_grid(c, w, h, c.fill_rectangle)
    c.line_to(F(0.2) * w, F(0.8) * h)
    c.close_path()
    c.stroke()


def stroke_rectangle(c, w, h):
    c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
    c.rotate(0.2)
    c.begin_path()
    c.move_to(F(0.2) * w, F(0.2) * h)
    c.line_to(F(0.8) * w, F(0.2) * h)
    c.line_to(F(0.8) * w, F(0.8) * h)
    _rectangle_grid(c, w, h, c.stroke_rectangle)
    c.line_to(F(0.2) * w, F(0.8) * h)
    c.close_path()
    c.stroke()


# ---- images (test.cpp:1743-1853) ---------------------------------------------------

def draw_image(c, w, h):
    checker = bytes((((i >> 2 & 1) ^ (i >> 6 & 1)) | ((i & 3) == 3)) * 255 for i in range(1024))
    c.draw_image(checker, 16, 16, 64, 0.0, 0.0, w * F(0.75), h * F(0.75))
    for row in frange(0.0, 4.0):
        for column in frange(0.0, 4.0):
            c.draw_image(checker, 16, 16, 64, column * F(17.25), row * F(17.25), 16.0, 16.0)
    c.draw_image(checker, 16, 16, 64, 128.0, 0.0, 32.0, 8.0)
    c.draw_image(checker, 16, 16, 64, 128.0, 48.0, 32.0, -32.0)
    c.draw_image(checker, 16, 16, 64, 200.0, 16.0, -32.0, 32.0)
    c.draw_image(checker, 16, 16, 64, 128.0, 64.0, 32.0, 0.0)
    c.draw_image(None, 16, 16, 64, 200.0, 64.0, 32.0, 32.0)
    pixel = bytes([0, 255, 0, 255])
    c.draw_image(pixel, 1, 1, 4, w * F(0.875), h * F(0.25), 1.0, 1.0)
    c.draw_image(pixel, 1, 1, 4, w * F(0.875), h * F(0.5), 16.0, 16.0)
    c.rotate(0.2)
    c.global_composite_operation = LIGHTER
    c.set_global_alpha(1.0)
    c.draw_image(checker, 16, 16, 64, F(0.25) * w, F(0.25) * h, F(0.5) * w, F(0.5) * h)


def draw_image_matted(c, w, h):
    c.set_color(FILL, 0.0, 1.0, 0.0, 0.0)
    c.fill_rectangle(0.0, 0.0, w, h)
    checker = bytes([0, 0, 255, 255, 255, 0, 0, 0, 0, 0, 255, 255,
                     255, 0, 0, 0, 0, 0, 255, 255, 255, 0, 0, 0,
                     0, 0, 255, 255, 255, 0, 0, 0, 0, 0, 255, 255])
    y, size_y = F(0.5), F(3.0)
    step_y = 0
    while step_y < 20 and y < h:
        x, size_x = F(0.5), F(3.0)
        step_x = 0
        while step_x < 20 and x < w:
            c.draw_image(checker, 3, 3, 12, x, y, size_x, size_y)
            x = F(x + F(size_x + F(5.0)))
            size_x = F(size_x * F(1.5))
            step_x += 1
        y = F(y + F(size_y + F(5.0)))
        size_y = F(size_y * F(1.5))
        step_y += 1


def get_image_data(c, w, h):
    for index in range(100):
        c.set_color(FILL, float(index // 2 % 2), float(index // 4 % 2), float(index // 8 % 2),
                    float(index // 16 % 2))
        c.fill_rectangle(F(3.0) * F(index % 10), F(3.0) * F(index // 10), 3.0, 3.0)
    data = bytearray(4939)
    data[0] = 150
    for index in range(1, 4939):
        data[index] = (data[index - 1] * 137 + 53) & 255
    c.get_image_data_into(data, 2, 35, 35, 141, -10, -10)
    hashed = 0
    for value in data:
        hashed = ((((hashed & 0x1ffff) << 15) | (hashed >> 17)) ^ value) & 0xFFFFFFFF
    expected = 0xf53f9792
    c.set_color(FILL, float(hashed != expected), float(hashed == expected), 0.0, 1.0)
    c.fill_rectangle(30.0, 0.0, w, 30.0)
    c.set_linear_gradient(FILL, 0.0, 0.0, w, 0.0)
    c.add_color_stop(FILL, 0.0, 1.0, 0.0, 0.0, 1.0)
    c.add_color_stop(FILL, 1.0, 0.0, 1.0, 0.0, 1.0)
    c.fill_rectangle(0.0, F(0.45) * h, w, F(0.1) * h)


def put_image_data(c, w, h):
    checker = bytearray(2052)
    for index in range(2048):
        checker[index + 2] = ((((index >> 2 & 1) ^ (index >> 7 & 1)) | ((index & 3) == 3)) &
                              (index >> 10 & 1)) * 255
    checker[0] = checker[1] = checker[2050] = checker[2051] = 157
    c.set_color(FILL, 0.4, 0.05, 0.2, 1.0)
    c.fill_rectangle(0.0, 0.0, F(0.25) * w, F(0.25) * h)
    c.set_global_alpha(0.5)
    c.global_composite_operation = LIGHTER
    c.rotate(0.2)
    for y in range(-10, int(h), 29):
        for x in range(-10, int(w), 29):
            c.put_image_data(checker, 16, 16, 128, x, y, offset=6)
    c.put_image_data(None, 32, 32, 128, 0, 0)


# ---- examples (test.cpp:1909-2163) --------------------------------------------------

def example_smiley(c, w, h):
    center_x, center_y = F(0.5) * w, F(0.5) * h
    radius = min(w, h) * F(0.4)
    c.set_radial_gradient(FILL, center_x, center_y, 0.0, center_x, center_y, radius)
    c.add_color_stop(FILL, 0.0, 1.0, 0.9, 0.2, 1.0)
    c.add_color_stop(FILL, 0.95, 0.95, 0.65, 0.15, 1.0)
    c.add_color_stop(FILL, 1.0, 0.9, 0.55, 0.0, 1.0)
    c.arc(center_x, center_y, radius, 0.0, 6.28318531)
    c.fill()
    c.set_linear_gradient(FILL, center_x, center_y - F(0.95) * radius, center_x, center_y)
    c.add_color_stop(FILL, 0.0, 1.0, 1.0, 1.0, 0.5)
    c.add_color_stop(FILL, 1.0, 1.0, 1.0, 1.0, 0.0)
    c.begin_path()
    c.arc(center_x, center_y - F(0.15) * radius, F(0.8) * radius, 0.0, 6.28318531)
    c.fill()
    c.set_color(STROKE, 0.0, 0.0, 0.0, 0.95)
    c.set_line_width(F(0.2) * radius)
    c.line_cap = CIRCLE
    c.begin_path()
    c.move_to(center_x - F(0.2) * radius, center_y - F(0.5) * radius)
    c.line_to(center_x - F(0.2) * radius, center_y - F(0.2) * radius)
    c.move_to(center_x + F(0.2) * radius, center_y - F(0.5) * radius)
    c.line_to(center_x + F(0.2) * radius, center_y - F(0.2) * radius)
    c.stroke()
    c.set_color(FILL, 0.0, 0.0, 0.0, 0.95)
    c.begin_path()
    c.move_to(center_x - F(0.6) * radius, center_y + F(0.1) * radius)
    c.bezier_curve_to(center_x - F(0.3) * radius, center_y + F(0.8) * radius,
                      center_x + F(0.3) * radius, center_y + F(0.8) * radius,
                      center_x + F(0.6) * radius, center_y + F(0.1) * radius)
    c.bezier_curve_to(center_x + F(0.3) * radius, center_y + F(0.3) * radius,
                      center_x - F(0.3) * radius, center_y + F(0.3) * radius,
                      center_x - F(0.6) * radius, center_y + F(0.1) * radius)
    c.fill()


KNOT = [[0.0, -1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0],
        [-1.0, -1.0, -1.0, 0.0, 1.0, 0.0, 1.0, 1.0],
        [2.0, 1.0, 2.0, -2.0, -1.0, -2.0, -1.0, -1.0],
        [-2.0, -1.0, -2.0, 2.0, 1.0, 2.0, 1.0, 1.0],
        [-2.0, -1.0, -2.0, -3.0, 0.0, -3.0, 0.0, -1.0],
        [2.0, 1.0, 2.0, 3.0, 0.0, 3.0, 0.0, 1.0]]


def example_knot(c, w, h):
    c.translate(w * F(0.5), h * F(0.5))
    c.scale(w * F(0.17), h * F(0.17))
    c.rotate(F(-15.0) * F(3.14159265) / F(180.0))
    for row in KNOT:
        p = [F(v) for v in row]
        c.begin_path()
        c.move_to(F(1.01) * p[0] - F(0.01) * p[2], F(1.01) * p[1] - F(0.01) * p[3])
        c.line_to(p[0], p[1])
        c.bezier_curve_to(p[2], p[3], p[4], p[5], p[6], p[7])
        c.line_to(F(-0.01) * p[4] + F(1.01) * p[6], F(-0.01) * p[5] + F(1.01) * p[7])
        c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
        c.set_line_width(0.75)
        c.line_cap = BUTT
        c.stroke()
        c.set_radial_gradient(STROKE, 0.0, 0.0, 0.0, 0.0, 0.0, 3.0)
        c.add_color_stop(STROKE, 0.0, 0.8, 1.0, 0.6, 1.0)
        c.add_color_stop(STROKE, 1.0, 0.1, 0.5, 0.1, 1.0)
        c.set_line_width(0.5)
        c.line_cap = CIRCLE
        c.stroke()


def example_icon(c, w, h):
    c.set_shadow_color(0.0, 0.0, 0.0, 1.0)
    c.shadow_offset_x = w / F(64.0)
    c.shadow_offset_y = h / F(64.0)
    c.set_shadow_blur(min(w, h) / F(32.0))
    c.scale(w / F(32.0), h / F(32.0))
    c.set_color(FILL, 0.4, 0.05, 0.2, 1.0)
    c.move_to(15.5, 1.0)
    c.arc_to(30.0, 1.0, 30.0, 15.5, 6.0)
    c.arc_to(30.0, 30.0, 15.5, 30.0, 6.0)
    c.arc_to(1.0, 30.0, 1.0, 15.5, 6.0)
    c.arc_to(1.0, 1.0, 15.5, 1.0, 6.0)
    c.fill()
    c.set_color(STROKE, 0.5, 0.5, 0.5, 1.0)
    c.begin_path()
    c.move_to(11.0, 16.0)
    c.line_to(27.0, 16.0)
    c.move_to(2.0, 23.0)
    c.line_to(29.0, 23.0)
    c.stroke()
    c.set_color(STROKE, 0.75, 0.75, 0.75, 1.0)
    c.begin_path()
    c.arc(25.0, 22.0, 0.5, 0.0, 6.28318531)
    for a, b in (((19.0, 6.0), (18.5, 8.0)), ((20.0, 6.0), (20.0, 8.0)),
                 ((21.0, 6.0), (21.5, 8.0)), ((17.0, 14.0), (16.0, 18.0)),
                 ((20.0, 14.0), (20.0, 18.0)), ((23.0, 14.0), (24.0, 18.0)),
                 ((18.0, 9.0), (22.0, 9.0)), ((18.0, 13.0), (22.0, 13.0))):
        c.move_to(*a)
        c.line_to(*b)
    c.rectangle(16.0, 8.0, 8.0, 6.0)
    c.stroke()
    c.set_color(STROKE, 1.0, 1.0, 1.0, 1.0)
    c.begin_path()
    c.arc(19.0, 12.0, 9.0, 0.0, 6.28318531)
    for a, b in (((12.3, 17.3), (3.3, 26.3)), ((13.0, 18.0), (4.0, 27.0)),
                 ((13.7, 18.7), (4.7, 27.7))):
        c.move_to(*a)
        c.line_to(*b)
    c.stroke()


def example_illusion(c, w, h):
    c.set_color(FILL, 0.0, 0.4, 1.0, 1.0)
    c.fill_rectangle(0.0, 0.0, w, h)
    c.set_color(FILL, 0.8, 0.8, 0.0, 1.0)
    c.set_line_width(0.4)
    for spot in frange(0.0, 240.0):
        angle = fmodf(spot * F(0.61803398875), F(1.0)) * F(6.28318531)
        radius = spot / F(240.0) * F(0.5) * hypotf(w, h)
        size = min(w, h) * sqrtf(spot) / F(240.0)
        c.set_transform(1.0, 0.0, 0.0, 1.0, 0.0, 0.0)
        c.translate(F(0.5) * w + radius * cosf(angle), F(0.5) * h + radius * sinf(angle))
        c.rotate(angle - F(1.3))
        c.scale(F(0.8) * size, F(0.6) * size)
        c.rotate(1.3)
        c.begin_path()
        c.arc(0.0, 0.0, 1.0, 0.0, 6.28318531)
        c.fill()
        c.begin_path()
        c.arc(0.0, 0.0, 1.0, 0.0, 3.14159265)
        c.set_color(STROKE, 1.0, 1.0, 1.0, 1.0)
        c.stroke()
        c.begin_path()
        c.arc(0.0, 0.0, 1.0, 3.14159265, 6.28318531)
        c.set_color(STROKE, 0.0, 0.0, 0.0, 1.0)
        c.stroke()


def example_star(c, w, h):
    c.scale(w / F(256.0), h / F(256.0))
    c.move_to(128.0, 28.0)
    for x, y in ((157.0, 87.0), (223.0, 97.0), (175.0, 143.0), (186.0, 208.0),
                 (128.0, 178.0), (69.0, 208.0), (80.0, 143.0), (32.0, 97.0), (98.0, 87.0)):
        c.line_to(x, y)
    c.close_path()
    c.set_shadow_blur(8.0)
    c.shadow_offset_y = 4.0
    c.set_shadow_color(0.0, 0.0, 0.0, 0.5)
    c.set_color(FILL, 1.0, 0.9, 0.2, 1.0)
    c.fill()
    c.line_join = ROUNDED
    c.set_line_width(12.0)
    c.set_color(STROKE, 0.9, 0.0, 0.5, 1.0)
    c.stroke()
    c.set_line_dash([21.0, 9.0, 1.0, 9.0, 7.0, 9.0, 1.0, 9.0], 8)
    c.line_dash_offset = 10.0
    c.line_cap = CIRCLE
    c.set_line_width(6.0)
    c.set_color(STROKE, 0.95, 0.65, 0.15, 1.0)
    c.stroke()
    c.set_shadow_color(0.0, 0.0, 0.0, 0.0)
    c.set_linear_gradient(FILL, 64.0, 0.0, 192.0, 256.0)
    c.add_color_stop(FILL, 0.30, 1.0, 1.0, 1.0, 0.0)
    c.add_color_stop(FILL, 0.35, 1.0, 1.0, 1.0, 0.8)
    c.add_color_stop(FILL, 0.45, 1.0, 1.0, 1.0, 0.8)
    c.add_color_stop(FILL, 0.50, 1.0, 1.0, 1.0, 0.0)
    c.global_composite_operation = SOURCE_ATOP
    c.fill_rectangle(0.0, 0.0, 256.0, 256.0)


NEON = [  # (move_to point, then bezier_curve_to control/end triples) per stroke
    ((45.5, 96.2), [(45.5, 96.2, 31.3, 106.2, 31.5, 113.1),
                    (31.7, 119.5, 50.6, 104.8, 50.6, 93.9),
                    (50.6, 91.1, 46.6, 89.1, 43.3, 89.4),
                    (27.5, 90.6, 8.5, 108.2, 8.8, 121.8),
                    (9.1, 133.1, 21.3, 136.6, 29.8, 136.3),
                    (52.4, 135.5, 62.3, 115.6, 62.3, 115.6)]),
    ((81.0, 120.2), [(81.0, 120.2, 60.2, 123.0, 59.7, 130.8),
                     (59.2, 140.6, 73.8, 136.4, 78.3, 125.3)]),
    ((80.7, 130.5), [(79.5, 132.4, 80.9, 135.0, 83.4, 135.0),
                     (95.8, 135.6, 99.3, 122.5, 111.4, 121.6),
                     (112.8, 121.5, 114.0, 123.0, 114.0, 124.3),
                     (113.9, 126.1, 106.7, 133.9, 106.7, 133.9)]),
    ((118.5, 122.9), [(118.5, 122.9, 122.1, 118.8, 126.1, 122.0),
                      (131.4, 126.4, 118.7, 131.6, 124.3, 134.7),
                      (130.0, 137.8, 150.0, 116.5, 156.0, 120.2),
                      (160.2, 122.8, 149.0, 133.5, 155.6, 133.6),
                      (162.0, 133.4, 173.8, 118.3, 168.0, 117.8)]),
    ((173.1, 123.2), [(177.8, 124.8, 182.8, 123.2, 187.0, 119.7)]),
    ((206.1, 118.6), [(206.1, 118.6, 185.3, 121.3, 185.1, 129.1),
                      (185.0, 138.7, 199.9, 135.4, 203.6, 123.6)]),
    ((205.6, 129.9), [(204.4, 131.8, 205.8, 134.4, 208.3, 134.4),
                      (220.3, 134.4, 246.6, 117.1, 246.6, 117.1)]),
    ((247.0, 122.4), [(245.9, 128.5, 243.9, 139.7, 231.2, 131.5)]),
]


def example_neon(c, w, h):
    c.scale(w / F(256.0), h / F(256.0))
    c.set_color(FILL, 0.0, 0.0625, 0.125, 1.0)
    c.fill_rectangle(0.0, 0.0, 256.0, 256.0)
    for start, curves in NEON:
        c.move_to(*start)
        for curve in curves:
            c.bezier_curve_to(*curve)
    c.line_cap = CIRCLE
    c.set_shadow_color(1.0, 0.5, 0.0, 1.0)
    c.set_shadow_blur(20.0)
    c.set_line_width(4.0)
    c.set_color(STROKE, 1.0, 0.5, 0.0, 1.0)
    c.stroke()
    c.set_shadow_blur(5.0)
    c.set_line_width(3.0)
    c.set_color(STROKE, 1.0, 0.625, 0.0, 1.0)
    c.stroke()


RESTATED = [scale_uniform, scale_non_uniform, rotate, translate, transform, transform_fill,
            transform_stroke, set_transform, global_alpha, shadow_color, shadow_offset,
            shadow_offset_offscreen, shadow_blur, shadow_blur_offscreen, line_width,
            line_width_angular, line_cap, line_cap_offscreen, line_join, line_join_offscreen,
            miter_limit, line_dash_offset, line_dash, line_dash_closed, line_dash_overlap,
            line_dash_offscreen, color, linear_gradient, radial_gradient, color_stop, pattern,
            begin_path, move_to, close_path, line_to, quadratic_curve_to, bezier_curve_to,
            arc_to, arc, rectangle, fill, fill_rounding, fill_converging, fill_zone_plate,
            stroke, stroke_wide, stroke_inner_join, stroke_spiral, stroke_long,
            is_point_in_path, is_point_in_path_offscreen, clear_rectangle, fill_rectangle,
            stroke_rectangle, draw_image, draw_image_matted, get_image_data, put_image_data,
            example_smiley, example_knot, example_icon, example_illusion, example_star,
            example_neon]
QUERIES = {"is_point_in_path", "is_point_in_path_offscreen", "get_image_data"}


def stream(name, size=256):
    """The command stream of one reference test at size x size."""
    for fn in RESTATED:
        if fn.__name__ == name:
            c = Canvas(size, size, probe=name in QUERIES)
            fn(c, F(size), F(size))
            return c.tobytes()
    if name in scenes.reference_clip_tests(size):
        return scenes.reference_clip_tests(size)[name]
    return scenes.reference_text_tests(size, measure=oracle.measure_text)[name]


def names():
    """Every test of the reference's table, in its order."""
    return list(scenes.reference_hashes())
