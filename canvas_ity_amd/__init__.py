"""canvas_ity_amd -- Python bindings to libcanvas_ity_hip.so (the C ABI in
include/cih_abi.h) for tests, benchmarks and the serving front end.

The product is the C++ header include/canvas_ity_hip.hpp plus the HIP
library; this module only drives the library through its C ABI with ctypes.
Nothing here computes pixels: every call goes to the GPU, and loading fails
loudly when the library has not been built (python -m canvas_ity_amd.build).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBRARY = os.path.join(HERE, "libcanvas_ity_hip.so")

OK = 0
ERRORS = {-1: "invalid argument", -2: "out of memory", -3: "device error",
          -4: "malformed input", -5: "over memory budget", -6: "timeout"}
ERROR_BUDGET = -5

# Fine-stage configurations of a batch (cih_batch_configure): item width in
# pixels of a canvas row, and one workgroup per item or a persistent grid.
FINE_CONFIGS = 6
FINE_CONFIG_NAMES = ["256 px, grid", "256 px, persistent", "128 px, grid",
                     "128 px, persistent", "64 px, grid", "64 px, persistent"]

# composite_operation values (canvas_ity.hpp:146-149)
OPS = {"source_in": 1, "source_copy": 2, "source_out": 3, "destination_in": 4,
       "destination_atop": 7, "lighter": 10, "destination_over": 11,
       "destination_out": 12, "source_atop": 13, "source_over": 14, "exclusive_or": 15}

# Every function declared in include/cih_abi.h, with its ctypes signature.
_c = ctypes
_P = _c.c_void_p
SIGNATURES = {
    "cih_device_open": (_c.c_int, [_c.c_int, _c.POINTER(_P)]),
    "cih_device_close": (_c.c_int, [_P]),
    "cih_device_synchronize": (_c.c_int, [_P]),
    "cih_last_error": (_c.c_char_p, []),
    "cih_canvas_create": (_c.c_int, [_P, _c.c_int, _c.c_int, _c.POINTER(_P)]),
    "cih_canvas_destroy": (_c.c_int, [_P]),
    "cih_canvas_size": (_c.c_int, [_P, _c.POINTER(_c.c_int), _c.POINTER(_c.c_int)]),
    "cih_canvas_device": (_c.c_int, [_P, _c.POINTER(_P)]),
    "cih_image_create": (_c.c_int, [_P, _P, _c.c_int, _c.c_int, _c.c_int,
                                    _c.POINTER(_c.c_int32)]),
    "cih_canvas_submit": (_c.c_int, [_P, _P, _c.c_size_t]),
    "cih_canvas_flush": (_c.c_int, [_P]),
    "cih_canvas_get_image_data": (_c.c_int, [_P, _P, _c.c_int, _c.c_int, _c.c_int,
                                             _c.c_int, _c.c_int]),
    "cih_canvas_put_image_data": (_c.c_int, [_P, _P, _c.c_int, _c.c_int, _c.c_int,
                                             _c.c_int, _c.c_int]),
    "cih_canvas_read_linear": (_c.c_int, [_P, _P]),
    "cih_canvas_replay": (_c.c_int, [_P, _P, _c.c_size_t]),
    "cih_canvas_execute": (_c.c_int, [_P, _c.c_int, _c.c_int]),
    "cih_canvas_stats": (_c.c_int, [_P, _c.POINTER(_c.c_uint64)]),
    "cih_device_profile": (_c.c_int, [_P, _c.c_int]),
    "cih_device_profile_read": (_c.c_int, [_P, _c.c_char_p, _c.c_size_t]),
    "cih_trace_polygons": (_c.c_int, [_P, _c.c_int, _P, _c.c_size_t, _P, _c.c_size_t,
                                      _c.POINTER(_c.c_size_t), _c.POINTER(_c.c_size_t)]),
    "cih_trace_stage": (_c.c_int, [_P, _c.c_int, _c.c_int, _P, _c.c_size_t, _P, _c.c_size_t,
                                   _c.POINTER(_c.c_size_t), _c.POINTER(_c.c_size_t)]),
    "cih_trace_runs": (_c.c_int, [_P, _c.c_int, _c.c_int, _P, _P, _P, _c.c_size_t,
                                  _c.POINTER(_c.c_size_t)]),
    "cih_batch_create": (_c.c_int, [_P, _c.c_int, _c.POINTER(_P), _c.POINTER(_c.c_size_t),
                                    _c.POINTER(_P)]),
    "cih_batch_execute": (_c.c_int, [_P, _c.c_int]),
    "cih_batch_size": (_c.c_int, [_P, _c.c_int, _c.POINTER(_c.c_int), _c.POINTER(_c.c_int)]),
    "cih_batch_image": (_c.c_int, [_P, _c.c_int, _P]),
    "cih_batch_linear": (_c.c_int, [_P, _c.c_int, _P]),
    "cih_batch_stats": (_c.c_int, [_P, _c.POINTER(_c.c_uint64)]),
    "cih_batch_frame": (_c.c_int, [_P, _c.c_int, _c.POINTER(_c.c_uint64)]),
    "cih_batch_configure": (_c.c_int, [_P, _c.c_int]),
    "cih_batch_tune": (_c.c_int, [_P, _c.c_int, _c.POINTER(_c.c_double),
                                  _c.POINTER(_c.c_int)]),
    "cih_batch_destroy": (_c.c_int, [_P]),
    "cih_server_start": (_c.c_int, [_P, _c.c_uint64, _c.c_int, _c.c_int, _c.POINTER(_P)]),
    "cih_server_submit": (_c.c_int, [_P, _P, _c.c_size_t, _c.POINTER(_c.c_uint64)]),
    "cih_server_wait": (_c.c_int, [_P, _c.POINTER(_c.c_uint64), _c.c_int]),
    "cih_server_result": (_c.c_int, [_P, _c.c_uint64, _P, _c.POINTER(_c.c_int),
                                     _c.POINTER(_c.c_int), _c.POINTER(_c.c_double)]),
    "cih_server_stats": (_c.c_int, [_P, _c.POINTER(_c.c_uint64)]),
    "cih_server_stop": (_c.c_int, [_P]),
    "cih_scene_generate": (_c.c_int, [_c.c_int, _c.c_uint64, _c.c_int, _c.c_int, _c.c_int,
                                      _c.c_int, _c.POINTER(_P), _c.POINTER(_c.c_size_t)]),
    "cih_scene_seed": (_c.c_uint64, [_c.c_uint64, _c.c_uint64]),
    "cih_free": (None, [_P]),
    "cih_path_contains": (_c.c_int, [_P, _c.c_size_t, _P, _c.c_size_t, _c.c_float, _c.c_float,
                                     _c.POINTER(_c.c_int)]),
    "cih_font_check": (_c.c_int, [_P, _c.c_size_t, _c.POINTER(_c.c_int)]),
    "cih_text_measure": (_c.c_int, [_P, _c.c_size_t, _c.c_float, _c.c_char_p,
                                    _c.POINTER(_c.c_float)]),
    "cih_text_outline": (_c.c_int, [_P, _c.c_size_t, _c.c_float, _c.c_char_p, _c.c_float,
                                    _c.c_float, _c.c_float, _c.c_int, _c.c_int, _P, _c.c_int,
                                    _c.c_float, _P, _c.c_size_t, _P, _c.c_size_t,
                                    _c.POINTER(_c.c_size_t), _c.POINTER(_c.c_size_t)]),
}

_lib = None


def library():
    """Loads libcanvas_ity_hip.so; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY):
            raise RuntimeError("libcanvas_ity_hip.so is not built: run "
                               "python -m canvas_ity_amd.build (needs hipcc)")
        lib = ctypes.CDLL(LIBRARY)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class CihError(RuntimeError):
    pass


def _check(status, what):
    if status != OK:
        msg = library().cih_last_error().decode(errors="replace")
        raise CihError("%s failed (%s): %s" % (what, ERRORS.get(status, status), msg))


def scene(config, seed, width=0, height=0, shapes=0, flags=0):
    """The seeded command stream of a benchmark configuration (bytes)."""
    lib = library()
    out = _P()
    size = _c.c_size_t()
    _check(lib.cih_scene_generate(config, seed, width, height, shapes, flags,
                                  _c.byref(out), _c.byref(size)), "cih_scene_generate")
    try:
        return _c.string_at(out, size.value)
    finally:
        lib.cih_free(out)


def path_contains(points, subpaths, x, y):
    """is_point_in_path for a device-space path: points (n, 2) float32 (a start
    point and 3 per cubic per subpath), subpaths (m, 2) uint32 (count, closed).
    Runs on the host (cih_path_contains); no device is needed."""
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 2)
    sub = np.ascontiguousarray(subpaths, dtype=np.uint32).reshape(-1, 2)
    inside = _c.c_int()
    _check(library().cih_path_contains(pts.ctypes.data_as(_P), pts.shape[0],
                                       sub.ctypes.data_as(_P), sub.shape[0], float(x), float(y),
                                       _c.byref(inside)), "cih_path_contains")
    return bool(inside.value)


def _text_bytes(text):
    return text.encode("utf-8") if isinstance(text, str) else bytes(text)


def font_units(font):
    """Units per em of a TrueType file (bytes), or None if set_font would
    reject it (cih_font_check)."""
    units = _c.c_int()
    if library().cih_font_check(font, len(font), _c.byref(units)) != OK:
        return None
    return units.value


def measure_text(font, size, text):
    """measure_text of `text` (str or UTF-8 bytes) in a font (bytes) at `size`,
    with the reference's scale size / units per em (cih_text_measure)."""
    units = font_units(font)
    if units is None:
        return 0.0
    scale = np.float32(np.float32(size) / np.float32(units))
    width = _c.c_float()
    _check(library().cih_text_measure(font, len(font), float(scale), _text_bytes(text),
                                      _c.byref(width)), "cih_text_measure")
    return width.value


def text_outline(font, size, text, x, y, maximum_width=1.0e30, align=0, baseline=0,
                 forward=(1.0, 0.0, 0.0, 1.0, 0.0, 0.0), stroking=False, line_width=1.0):
    """Glyph outline segments of fill_text/stroke_text (cih_text_outline):
    (points (4k, 2) float32 -- start, control, control, end per cubic --,
    subpaths (m, 2) uint32 (count, closed))."""
    units = font_units(font)
    if units is None:
        raise CihError("not a usable TrueType font")
    scale = np.float32(np.float32(size) / np.float32(units))
    fw = np.asarray(forward, dtype=np.float32)
    np_, ns = _c.c_size_t(), _c.c_size_t()
    lib = library()
    args = (font, len(font), float(scale), _text_bytes(text), float(x), float(y),
            float(maximum_width), int(align), int(baseline), fw.ctypes.data_as(_P),
            1 if stroking else 0, float(line_width))
    status = lib.cih_text_outline(*args, None, 0, None, 0, _c.byref(np_), _c.byref(ns))
    if status not in (OK, ERROR_BUDGET):
        _check(status, "cih_text_outline")
    pts = np.zeros((max(np_.value, 1), 2), dtype=np.float32)
    sub = np.zeros((max(ns.value, 1), 2), dtype=np.uint32)
    _check(lib.cih_text_outline(*args, pts.ctypes.data_as(_P), pts.shape[0],
                                sub.ctypes.data_as(_P), sub.shape[0], _c.byref(np_),
                                _c.byref(ns)), "cih_text_outline")
    return pts[:np_.value].copy(), sub[:ns.value].copy()


def scene_seed(seed, index):
    return library().cih_scene_seed(seed, index)


SCENE_EVENODD, SCENE_NO_IMAGES, SCENE_SHADOWS = 1, 2, 4


def stream_size(stream):
    words = np.frombuffer(stream[:16], dtype=np.uint32)
    return int(words[2]), int(words[3])


class Device:
    def __init__(self, ordinal=0):
        self.handle = _P()
        _check(library().cih_device_open(ordinal, _c.byref(self.handle)), "cih_device_open")

    def synchronize(self):
        _check(library().cih_device_synchronize(self.handle), "cih_device_synchronize")

    def profile(self, enable=True):
        _check(library().cih_device_profile(self.handle, int(enable)), "cih_device_profile")

    def profile_read(self):
        import json
        buf = _c.create_string_buffer(4096)
        _check(library().cih_device_profile_read(self.handle, buf, 4096),
               "cih_device_profile_read")
        return json.loads(buf.value.decode())

    def close(self):
        if self.handle:
            library().cih_device_close(self.handle)
            self.handle = _P()


_default = None


def default_device():
    global _default
    if _default is None:
        _default = Device(0)
    return _default


class Canvas:
    """A device canvas driven through the C ABI."""

    def __init__(self, width, height, device=None):
        self.device = device or default_device()
        self.width, self.height = width, height
        self.handle = _P()
        _check(library().cih_canvas_create(self.device.handle, width, height,
                                           _c.byref(self.handle)), "cih_canvas_create")

    def replay(self, stream):
        buf = _c.create_string_buffer(bytes(stream), len(stream))
        _check(library().cih_canvas_replay(self.handle, buf, len(stream)), "cih_canvas_replay")

    def flush(self):
        _check(library().cih_canvas_flush(self.handle), "cih_canvas_flush")

    def stats(self):
        """Work counts of the last draw list (runs it once, synchronously)."""
        out = (_c.c_uint64 * 4)()
        _check(library().cih_canvas_stats(self.handle, out), "cih_canvas_stats")
        return {"pixel_draws": out[0], "runs": out[1], "points": out[2],
                "shadow_floats": out[3]}

    def execute(self, clear=True, readback=True):
        """Runs the last flushed draw list again from HBM (asynchronous)."""
        _check(library().cih_canvas_execute(self.handle, int(clear), int(readback)),
               "cih_canvas_execute")

    def get_image_data(self, width=None, height=None, stride=None, x=0, y=0, out=None):
        width = self.width if width is None else width
        height = self.height if height is None else height
        stride = 4 * width if stride is None else stride
        if out is None:
            out = np.zeros((height, stride), dtype=np.uint8)
        _check(library().cih_canvas_get_image_data(
            self.handle, out.ctypes.data_as(_P), width, height, stride, x, y),
            "cih_canvas_get_image_data")
        return out

    def rgba8(self):
        return self.get_image_data().reshape(self.height, self.width, 4)

    def linear(self):
        out = np.zeros((self.height, self.width, 4), dtype=np.float32)
        _check(library().cih_canvas_read_linear(self.handle, out.ctypes.data_as(_P)),
               "cih_canvas_read_linear")
        return out

    def put_image_data(self, image, x=0, y=0):
        image = np.ascontiguousarray(image, dtype=np.uint8)
        h, w = image.shape[0], image.shape[1]
        _check(library().cih_canvas_put_image_data(self.handle, image.ctypes.data_as(_P),
                                                   w, h, 4 * w, x, y),
               "cih_canvas_put_image_data")

    def trace_polygons(self, draw):
        """(points [n,2], subpaths [m,2]) that draw `draw` of the last frame made."""
        lib = library()
        n, m = _c.c_size_t(), _c.c_size_t()
        _check(lib.cih_trace_polygons(self.handle, draw, None, 0, None, 0, _c.byref(n),
                                      _c.byref(m)), "cih_trace_polygons")
        pts = np.zeros((n.value, 2), dtype=np.float32)
        sub = np.zeros((m.value, 2), dtype=np.uint32)
        _check(lib.cih_trace_polygons(self.handle, draw, pts.ctypes.data_as(_P), n.value,
                                      sub.ctypes.data_as(_P), m.value, _c.byref(n),
                                      _c.byref(m)), "cih_trace_polygons")
        return pts, sub

    def trace_stage(self, draw, stage):
        """(points, subpaths) of stage 0 flatten, 1 dash, 2 polygons (last frame)."""
        lib = library()
        n, m = _c.c_size_t(), _c.c_size_t()
        _check(lib.cih_trace_stage(self.handle, draw, stage, None, 0, None, 0, _c.byref(n),
                                   _c.byref(m)), "cih_trace_stage")
        pts = np.zeros((n.value, 2), dtype=np.float32)
        sub = np.zeros((m.value, 2), dtype=np.uint32)
        _check(lib.cih_trace_stage(self.handle, draw, stage, pts.ctypes.data_as(_P), n.value,
                                   sub.ctypes.data_as(_P), m.value, _c.byref(n), _c.byref(m)),
               "cih_trace_stage")
        return pts, sub

    def trace_runs(self, draw, shadow=False):
        """(x, y, delta) arrays of the draw's runs in the last frame, unsorted."""
        lib = library()
        n = _c.c_size_t()
        _check(lib.cih_trace_runs(self.handle, draw, int(shadow), None, None, None, 0,
                                  _c.byref(n)), "cih_trace_runs")
        xs = np.zeros(n.value, dtype=np.int32)
        ys = np.zeros(n.value, dtype=np.int32)
        ds = np.zeros(n.value, dtype=np.float32)
        _check(lib.cih_trace_runs(self.handle, draw, int(shadow), xs.ctypes.data_as(_P),
                                  ys.ctypes.data_as(_P), ds.ctypes.data_as(_P), n.value,
                                  _c.byref(n)), "cih_trace_runs")
        return xs, ys, ds

    def close(self):
        if self.handle:
            library().cih_canvas_destroy(self.handle)
            self.handle = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many canvases rendered together (cih_batch_* in include/cih_abi.h)."""

    def __init__(self, streams, device=None):
        self.device = device or default_device()
        self._keep = [_c.create_string_buffer(bytes(s), len(s)) for s in streams]
        n = len(streams)
        ptrs = (_P * n)(*[_c.cast(b, _P) for b in self._keep])
        sizes = (_c.c_size_t * n)(*[len(s) for s in streams])
        self.handle = _P()
        _check(library().cih_batch_create(self.device.handle, n, ptrs, sizes,
                                          _c.byref(self.handle)), "cih_batch_create")
        self.count = n
        self.sizes = []
        for i in range(n):
            w, h = _c.c_int(), _c.c_int()
            library().cih_batch_size(self.handle, i, _c.byref(w), _c.byref(h))
            self.sizes.append((w.value, h.value))

    def execute(self, to_host=True):
        _check(library().cih_batch_execute(self.handle, int(to_host)), "cih_batch_execute")

    def rgba8(self, index):
        w, h = self.sizes[index]
        out = np.zeros((h, w, 4), dtype=np.uint8)
        _check(library().cih_batch_image(self.handle, index, out.ctypes.data_as(_P)),
               "cih_batch_image")
        return out

    def linear(self, index):
        w, h = self.sizes[index]
        out = np.zeros((h, w, 4), dtype=np.float32)
        _check(library().cih_batch_linear(self.handle, index, out.ctypes.data_as(_P)),
               "cih_batch_linear")
        return out

    def stats(self):
        out = (_c.c_uint64 * 4)()
        _check(library().cih_batch_stats(self.handle, out), "cih_batch_stats")
        return {"canvases": out[0], "draws": out[1], "frames": out[2], "pixels": out[3]}

    def frames(self):
        """Each frame's first draw, draws, canvases and canvas pixels."""
        out = []
        info = (_c.c_uint64 * 4)()
        for k in range(self.stats()["frames"]):
            _check(library().cih_batch_frame(self.handle, k, info), "cih_batch_frame")
            out.append({"first_draw": info[0], "draws": info[1], "canvases": info[2],
                        "pixels": info[3]})
        return out

    def configure(self, config):
        """Fine-stage launch shape, 0..FINE_CONFIGS-1 (FINE_CONFIG_NAMES)."""
        _check(library().cih_batch_configure(self.handle, int(config)), "cih_batch_configure")

    def tune(self, reps=3):
        """Times every fine-stage configuration on this batch and keeps the
        fastest; returns (best, milliseconds per execute of each)."""
        ms = (_c.c_double * FINE_CONFIGS)()
        best = _c.c_int()
        _check(library().cih_batch_tune(self.handle, int(reps), ms, _c.byref(best)),
               "cih_batch_tune")
        return best.value, list(ms)

    def close(self):
        if self.handle:
            library().cih_batch_destroy(self.handle)
            self.handle = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop (cih_server_* in include/cih_abi.h)."""

    def __init__(self, hbm_budget=64 << 30, max_batch=64, workers=0, device=None):
        self.device = device or default_device()
        self.handle = _P()
        _check(library().cih_server_start(self.device.handle, hbm_budget, max_batch, workers,
                                          _c.byref(self.handle)), "cih_server_start")
        self.sizes = {}

    def submit(self, stream):
        rid = _c.c_uint64()
        buf = _c.create_string_buffer(bytes(stream), len(stream))
        _check(library().cih_server_submit(self.handle, buf, len(stream), _c.byref(rid)),
               "cih_server_submit")
        self.sizes[rid.value] = stream_size(stream)
        return rid.value

    def wait(self, timeout_ms=1000):
        """Next finished request id, or None on timeout."""
        rid = _c.c_uint64()
        status = library().cih_server_wait(self.handle, _c.byref(rid), timeout_ms)
        if status == -6:
            return None
        _check(status, "cih_server_wait")
        return rid.value

    def result(self, rid):
        """(sRGB8 image, latency in ms) of a finished request."""
        w, h = self.sizes.pop(rid)
        out = np.zeros((h, w, 4), dtype=np.uint8)
        lat = _c.c_double()
        _check(library().cih_server_result(self.handle, rid, out.ctypes.data_as(_P), None, None,
                                           _c.byref(lat)), "cih_server_result")
        return out, lat.value

    def stats(self):
        out = (_c.c_uint64 * 4)()
        _check(library().cih_server_stats(self.handle, out), "cih_server_stats")
        return {"served": out[0], "batches": out[1], "max_in_flight": out[2],
                "reruns": out[3]}

    def stop(self):
        if self.handle:
            library().cih_server_stop(self.handle)
            self.handle = _P()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass


def render(stream, device=None):
    """Replays a command stream on a fresh canvas; returns (sRGB8, linear)."""
    w, h = stream_size(stream)
    canvas = Canvas(w, h, device)
    try:
        canvas.replay(stream)
        return canvas.rgba8(), canvas.linear()
    finally:
        canvas.close()
