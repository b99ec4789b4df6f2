"""TEST INFRASTRUCTURE ONLY -- the checker side of the parity tests.

Two CPU implementations of the canvas pipeline, loaded with ctypes:

* ``port``      liboracle.so, the C restatement in cih_oracle.c (each
                function cites the canvas_ity.hpp lines it follows);
* ``reference`` _ref/libcih_ref.so, the reference header itself compiled
                from its source by oracle/Makefile (absent where the
                reference checkout and its prebuilt file are both missing).

Both replay the same command streams (include/cih_stream.h) as the HIP
library.  Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline
use this package; the product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PORT = os.path.join(HERE, "liboracle.so")
REFERENCE = os.path.join(HERE, "_ref", "libcih_ref.so")

_P = ctypes.c_void_p
_libs = {}


def build():
    """Builds liboracle.so (and _ref/ when the reference checkout exists)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def _load(kind):
    if kind in _libs:
        return _libs[kind]
    path, prefix = (PORT, "cih_oracle") if kind == "port" else (REFERENCE, "cih_ref")
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    lib = ctypes.CDLL(path)
    fns = {}
    for name, res, args in (
            ("create", _P, [ctypes.c_int, ctypes.c_int]),
            ("destroy", None, [_P]),
            ("replay", ctypes.c_int, [_P, _P, ctypes.c_size_t]),
            ("get_image_data", None, [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int]),
            ("read_linear", None, [_P, _P])):
        fn = getattr(lib, prefix + "_" + name)
        fn.restype = res
        fn.argtypes = args
        fns[name] = fn
    if kind == "reference":
        fn = lib.cih_ref_path_lines
        fn.restype = ctypes.c_long
        fn.argtypes = [_P, ctypes.c_int, _P, ctypes.c_long, _P, ctypes.c_long,
                       ctypes.POINTER(ctypes.c_long)]
        fn = lib.cih_ref_dash_lines
        fn.restype = ctypes.c_long
        fn.argtypes = [_P, _P, ctypes.c_long, _P, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
        fn = lib.cih_ref_lines_to_runs
        fn.restype = ctypes.c_long
        fn.argtypes = [_P, ctypes.c_float, ctypes.c_float, ctypes.c_int, _P, ctypes.c_long]
        fn = lib.cih_ref_path
        fn.restype = ctypes.c_long
        fn.argtypes = [_P, _P, ctypes.c_long, _P, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
        fn = lib.cih_ref_is_point_in_path
        fn.restype = ctypes.c_int
        fn.argtypes = [_P, ctypes.c_float, ctypes.c_float]
        fn = lib.cih_ref_set_font
        fn.restype = ctypes.c_int
        fn.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_float, ctypes.c_char_p,
                       ctypes.POINTER(ctypes.c_float)]
    else:
        fn = lib.cih_oracle_is_point_in_path
        fn.restype = ctypes.c_int
        fn.argtypes = [_P, ctypes.c_float, ctypes.c_float]
        for name in ("path_lines", "dash_lines", "lines_to_runs"):
            fn = getattr(lib, "cih_oracle_" + name)
            fn.restype = ctypes.c_long
        lib.cih_oracle_path_lines.argtypes = [_P, ctypes.c_int, _P, ctypes.c_long, _P,
                                              ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
        lib.cih_oracle_dash_lines.argtypes = [_P, _P, ctypes.c_long, _P, ctypes.c_long,
                                              ctypes.POINTER(ctypes.c_long)]
        lib.cih_oracle_lines_to_runs.argtypes = [_P, ctypes.c_float, ctypes.c_float,
                                                 ctypes.c_int, _P, ctypes.c_long]
        fn = lib.cih_oracle_measure_text
        fn.restype = ctypes.c_float
        fn.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_float, ctypes.c_char_p]
    fns["seam"] = "cih_ref_" if kind == "reference" else "cih_oracle_"
    fns["lib"] = lib
    _libs[kind] = fns
    return fns


def measure_text(font, size, text, kind="port"):
    """measure_text of `text` (str or UTF-8 bytes) after set_font(font, size)
    on a fresh canvas (0.0 if the font is rejected)."""
    data = text.encode("utf-8") if isinstance(text, str) else bytes(text)
    fns = _load(kind)
    if kind == "port":
        return fns["lib"].cih_oracle_measure_text(font, len(font), size, data)
    width = ctypes.c_float()
    fns["lib"].cih_ref_set_font(font, len(font), size, data, ctypes.byref(width))
    return width.value


def font_accepted(font, kind="reference"):
    """Whether set_font accepts a font (the reference build's answer)."""
    width = ctypes.c_float()
    return _load(kind)["lib"].cih_ref_set_font(font, len(font), 16.0, b"",
                                               ctypes.byref(width)) == 1


def available(kind):
    try:
        _load(kind)
        return True
    except (FileNotFoundError, OSError):
        return False


class Canvas:
    """A CPU canvas of the given kind ('port' or 'reference')."""

    def __init__(self, width, height, kind="port"):
        self.fns = _load(kind)
        self.width, self.height = width, height
        self.handle = self.fns["create"](width, height)
        if not self.handle:
            raise ValueError("invalid canvas size")

    def replay(self, stream):
        buf = ctypes.create_string_buffer(bytes(stream), len(stream))
        if self.fns["replay"](self.handle, buf, len(stream)) != 0:
            raise ValueError("malformed command stream")

    def get_image_data(self, width=None, height=None, stride=None, x=0, y=0, out=None):
        width = self.width if width is None else width
        height = self.height if height is None else height
        stride = 4 * width if stride is None else stride
        if out is None:
            out = np.zeros((height, stride), dtype=np.uint8)
        self.fns["get_image_data"](self.handle, out.ctypes.data_as(_P), width, height,
                                   stride, x, y)
        return out

    def rgba8(self):
        return self.get_image_data().reshape(self.height, self.width, 4)

    def linear(self):
        out = np.zeros((self.height, self.width, 4), dtype=np.float32)
        self.fns["read_linear"](self.handle, out.ctypes.data_as(_P))
        return out

    def _seam(self, name):
        return getattr(self.fns["lib"], self.fns["seam"] + name)

    def path_lines(self, stroking):
        """Stage seam: lines after path_to_lines (+ stroke_lines)."""
        cap = 1 << 22
        pts = np.zeros((cap, 2), dtype=np.float32)
        sub = np.zeros((cap // 2, 2), dtype=np.int32)
        ns = ctypes.c_long()
        n = self._seam("path_lines")(self.handle, int(stroking), pts.ctypes.data_as(_P), cap,
                                     sub.ctypes.data_as(_P), cap // 2, ctypes.byref(ns))
        return pts[:n].copy(), sub[:ns.value].copy()

    def dash_lines(self):
        """Stage seam: lines after path_to_lines(true) and dash_lines."""
        cap = 1 << 22
        pts = np.zeros((cap, 2), dtype=np.float32)
        sub = np.zeros((cap // 2, 2), dtype=np.int32)
        ns = ctypes.c_long()
        n = self._seam("dash_lines")(self.handle, pts.ctypes.data_as(_P), cap,
                                     sub.ctypes.data_as(_P), cap // 2, ctypes.byref(ns))
        return pts[:n].copy(), sub[:ns.value].copy()

    def lines_to_runs(self, offset_x=0.0, offset_y=0.0, padding=0):
        """Stage seam: sorted, coalesced runs of the current lines."""
        cap = 1 << 23
        runs = np.zeros((cap, 3), dtype=np.float32)
        n = self._seam("lines_to_runs")(self.handle, offset_x, offset_y, padding,
                                        runs.ctypes.data_as(_P), cap)
        return runs[:n].copy()

    def close(self):
        if self.handle:
            self.fns["destroy"](self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render(stream, kind="port"):
    """Replays a command stream on a fresh CPU canvas; returns (sRGB8, linear)."""
    words = np.frombuffer(stream[:16], dtype=np.uint32)
    canvas = Canvas(int(words[2]), int(words[3]), kind)
    try:
        canvas.replay(stream)
        return canvas.rgba8(), canvas.linear()
    finally:
        canvas.close()
