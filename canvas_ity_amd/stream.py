"""Writer for the canvas command stream (include/cih_stream.h).

A Stream records canvas API calls by name, mirroring canvas_ity::canvas
(canvas_ity.hpp:194-1148), and serialises them to the wire format that
cih_canvas_replay and the serving loop accept.
"""
import struct

import numpy as np

MAGIC, VERSION = 0x53484943, 1

OPS = dict(scale=1, rotate=2, translate=3, transform=4, set_transform=5,
           set_global_alpha=6, set_composite_op=7, set_shadow_color=8,
           set_shadow_blur=9, set_shadow_offset=10, set_line_width=11, set_line_cap=12,
           set_line_join=13, set_miter_limit=14, set_line_dash=15,
           set_line_dash_offset=16, set_color=17, set_linear_gradient=18,
           set_radial_gradient=19, add_color_stop=20, set_pattern=21, begin_path=22,
           move_to=23, close_path=24, line_to=25, quadratic_curve_to=26,
           bezier_curve_to=27, arc_to=28, arc=29, rectangle=30, fill=31, stroke=32,
           clip=33, clear_rectangle=34, fill_rectangle=35, stroke_rectangle=36,
           draw_image=37, save=38, restore=39, define_image=40, put_image_data=41,
           set_fill_rule=42, set_font=43, fill_text=44, stroke_text=45, set_text_align=46,
           set_text_baseline=47)

# argument kinds: f float, i int
KINDS = dict(scale="ff", rotate="f", translate="ff", transform="ffffff",
             set_transform="ffffff", set_global_alpha="f", set_composite_op="i",
             set_shadow_color="ffff", set_shadow_blur="f", set_shadow_offset="ff",
             set_line_width="f", set_line_cap="i", set_line_join="i", set_miter_limit="f",
             set_line_dash_offset="f", set_color="iffff", set_linear_gradient="iffff",
             set_radial_gradient="iffffff", add_color_stop="ifffff", set_pattern="iii",
             begin_path="", move_to="ff", close_path="", line_to="ff",
             quadratic_curve_to="ffff", bezier_curve_to="ffffff", arc_to="fffff",
             arc="fffffi", rectangle="ffff", fill="", stroke="", clip="",
             clear_rectangle="ffff", fill_rectangle="ffff", stroke_rectangle="ffff",
             draw_image="iffff", save="", restore="", put_image_data="iii",
             set_fill_rule="i", set_text_align="i", set_text_baseline="i")

# text_align / text_baseline values (canvas_ity.hpp:150-156)
ALIGN = dict(leftward=0, rightward=1, center=2, start=0, ending=1)
BASELINE = dict(alphabetic=0, top=1, middle=2, bottom=3, hanging=4, ideographic=3)
COMPOSITE = dict(source_in=1, source_copy=2, source_out=3, destination_in=4,
                 destination_atop=7, lighter=10, destination_over=11, destination_out=12,
                 source_atop=13, source_over=14, exclusive_or=15)
CAP = dict(butt=0, square=1, circle=2)
JOIN = dict(miter=0, bevel=1, rounded=2)
FILL_STYLE, STROKE_STYLE = 0, 1


class Stream:
    def __init__(self, width, height):
        self.width, self.height = width, height
        self.words = [MAGIC, VERSION, width, height]

    def _emit(self, name, args):
        kinds = KINDS[name]
        if len(args) != len(kinds):
            raise TypeError("%s takes %d arguments" % (name, len(kinds)))
        self.words.append(OPS[name])
        for kind, value in zip(kinds, args):
            if kind == "f":
                self.words.append(struct.unpack("<I", struct.pack("<f", float(value)))[0])
            else:
                self.words.append(int(value) & 0xFFFFFFFF)
        return self

    def __getattr__(self, name):
        if name in KINDS:
            return lambda *args: self._emit(name, args)
        raise AttributeError(name)

    def arc(self, x, y, r, start, end, ccw=False):
        return self._emit("arc", (x, y, r, start, end, 1 if ccw else 0))

    def set_line_dash(self, segments):
        self.words.append(OPS["set_line_dash"])
        if segments is None:
            self.words.append(0xFFFFFFFF)
            return self
        self.words.append(len(segments))
        for v in segments:
            self.words.append(struct.unpack("<I", struct.pack("<f", float(v)))[0])
        return self

    def define_image(self, image_id, rgba):
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        h, w = rgba.shape[0], rgba.shape[1]
        self.words += [OPS["define_image"], image_id, w, h]
        self.words += np.frombuffer(rgba.tobytes(), dtype="<u4").tolist()
        return self

    def _blob(self, data):
        data = bytes(data)
        padded = data + b"\0" * (-len(data) % 4)
        self.words += np.frombuffer(padded, dtype="<u4").tolist()

    def set_font(self, font, size):
        """set_font(font, bytes, size); font None (or empty) only sets the size."""
        font = bytes(font) if font else b""
        self.words += [OPS["set_font"], len(font),
                       struct.unpack("<I", struct.pack("<f", float(size)))[0]]
        self._blob(font)
        return self

    def _text(self, name, text, x, y, maximum_width):
        data = text.encode("utf-8") if isinstance(text, str) else bytes(text)
        self.words += [OPS[name], len(data)]
        for v in (x, y, maximum_width):
            self.words.append(struct.unpack("<I", struct.pack("<f", float(v)))[0])
        self._blob(data)
        return self

    def fill_text(self, text, x, y, maximum_width=1.0e30):
        return self._text("fill_text", text, x, y, maximum_width)

    def stroke_text(self, text, x, y, maximum_width=1.0e30):
        return self._text("stroke_text", text, x, y, maximum_width)

    def tobytes(self):
        return np.asarray(self.words, dtype=np.uint32).tobytes()


ARITY = {1: 2, 2: 1, 3: 2, 4: 6, 5: 6, 6: 1, 7: 1, 8: 4, 9: 1, 10: 2, 11: 1, 12: 1, 13: 1, 14: 1,
         16: 1, 17: 5, 18: 5, 19: 7, 20: 6, 21: 3, 22: 0, 23: 2, 24: 0, 25: 2, 26: 4, 27: 6,
         28: 5, 29: 6, 30: 4, 31: 0, 32: 0, 33: 0, 34: 4, 35: 4, 36: 4, 37: 5, 38: 0, 39: 0,
         41: 3, 42: 1, 46: 1, 47: 1}
# fill, stroke, clear/fill/stroke_rectangle, draw_image, fill_text, stroke_text
DRAW_OPS = (31, 32, 34, 35, 36, 37, 44, 45)


def records(stream):
    """[(opcode, word offset of the opcode, argument count)] of a stream."""
    words = np.frombuffer(stream, dtype=np.uint32)
    at, out = 4, []
    while at < len(words):
        code = int(words[at])
        if code == 15:
            n = 1 + max(int(np.int32(words[at + 1])), 0)
        elif code == 40:
            n = 3 + int(words[at + 2]) * int(words[at + 3])
        elif code in (43, 44, 45):
            n = (2 if code == 43 else 4) + (int(words[at + 1]) + 3) // 4
        else:
            n = ARITY[code]
        out.append((code, at, n))
        at += 1 + n
    return out


def count_draws(stream):
    return sum(1 for code, _, _ in records(stream) if code in DRAW_OPS)


def truncate_draws(stream, draws):
    """The stream up to and including its `draws`-th drawing call."""
    seen = 0
    for code, at, n in records(stream):
        if code in DRAW_OPS:
            seen += 1
            if seen == draws:
                return bytes(stream[:4 * (at + 1 + n)])
    return bytes(stream)
