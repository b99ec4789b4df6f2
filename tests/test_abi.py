"""The C-ABI library builds, loads and exports every declared entry point;
the command stream format and the seeded generator behave (no GPU needed)."""
import os
import re

import numpy as np
import pytest

import canvas_ity_amd as cia
from canvas_ity_amd.stream import Stream

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"CIH_API\s+[\w\s\*]*?\b(cih_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = cia.library()
    names = declared("cih_abi.h")
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
        assert name in cia.SIGNATURES, name


def test_no_device_is_reported_cleanly():
    # In a container without a GPU, opening a device fails with a status code
    # and a message instead of crashing; on a GPU machine it succeeds.
    import ctypes
    handle = ctypes.c_void_p()
    status = cia.library().cih_device_open(0, ctypes.byref(handle))
    if status == 0:
        cia.library().cih_device_close(handle)
    else:
        assert status == -3
        assert b"device" in cia.library().cih_last_error()


def test_scene_generator_is_deterministic():
    a = cia.scene(1, 0xC1)
    b = cia.scene(1, 0xC1)
    c = cia.scene(1, 0xC2)
    assert a == b and a != c
    assert cia.stream_size(a) == (512, 512)
    assert cia.stream_size(cia.scene(2, 1)) == (2048, 2048)
    assert cia.stream_size(cia.scene(3, 1)) == (1024, 1024)


@pytest.mark.parametrize("index", range(8))
def test_batch_and_serve_sizes(index):
    w, h = cia.stream_size(cia.scene(4, cia.scene_seed(0xC4, index)))
    assert 64 <= w <= 1024 and 64 <= h <= 1024
    w, h = cia.stream_size(cia.scene(5, cia.scene_seed(0xC5, index)))
    assert 128 <= w <= 2048 and 128 <= h <= 2048


def count_ops(stream, op):
    from canvas_ity_amd.stream import OPS
    words = np.frombuffer(stream, dtype=np.uint32)
    arity = {1: 2, 2: 1, 3: 2, 4: 6, 5: 6, 6: 1, 7: 1, 8: 4, 9: 1, 10: 2, 11: 1, 12: 1, 13: 1,
             14: 1, 16: 1, 17: 5, 18: 5, 19: 7, 20: 6, 21: 3, 22: 0, 23: 2, 24: 0, 25: 2,
             26: 4, 27: 6, 28: 5, 29: 6, 30: 4, 31: 0, 32: 0, 33: 0, 34: 4, 35: 4, 36: 4,
             37: 5, 38: 0, 39: 0, 41: 3, 42: 1}
    at, n = 4, 0
    while at < len(words):
        code = int(words[at])
        if code == 15:
            size = 1 + max(int(np.int32(words[at + 1])), 0)
        elif code == 40:
            size = 3 + int(words[at + 2]) * int(words[at + 3])
        else:
            size = arity[code]
        n += code == OPS[op]
        at += 1 + size
    assert at == len(words)
    return n


def test_configuration_draw_counts():
    s1 = cia.scene(1, 0xC1)
    assert count_ops(s1, "fill") == 200 and count_ops(s1, "stroke") == 50
    s2 = cia.scene(2, 0xC2)
    assert count_ops(s2, "fill") + count_ops(s2, "stroke") == 5000
    s3 = cia.scene(3, 0xC3)
    assert count_ops(s3, "draw_image") == 100
    assert count_ops(s3, "fill") + count_ops(s3, "stroke") == 500


def test_stream_writer_round_trip():
    s = Stream(16, 8).set_color(0, 1, 0.5, 0.25, 1).fill_rectangle(1, 2, 3, 4)
    s.set_line_dash([1.5, 2.5]).set_line_dash(None)
    words = np.frombuffer(s.tobytes(), dtype=np.uint32)
    assert words[0] == 0x53484943 and words[2] == 16 and words[3] == 8
    assert words[4] == 17 and np.frombuffer(words[6:10].tobytes(), np.float32)[1] == 0.5
    assert count_ops(s.tobytes(), "fill_rectangle") == 1


def test_header_without_a_device_draws_nothing(tmp_path):
    # A C++ caller of the drop-in header on a machine without a GPU: the
    # reference never aborts, so the header reports the missing device once
    # and its canvases read back as transparent black.
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    probe = cia.library()
    import ctypes
    handle = ctypes.c_void_p()
    if probe.cih_device_open(0, ctypes.byref(handle)) == 0:
        probe.cih_device_close(handle)
        pytest.skip("a GPU is present")
    src = tmp_path / "caller.cpp"
    src.write_text(r'''
#include "canvas_ity_hip.hpp"
#include <cstdio>
#include <vector>
int main() {
    canvas_ity_hip::canvas c(16, 8);
    c.set_color(canvas_ity_hip::fill_style, 1.0f, 0.5f, 0.0f, 1.0f);
    c.fill_rectangle(2.0f, 2.0f, 10.0f, 4.0f);
    std::vector<unsigned char> rgba(16 * 8 * 4, 0xAB);
    c.get_image_data(&rgba[0], 16, 8, 16 * 4, 0, 0);
    for (unsigned char v : rgba)
        if (v)
            return 3;
    std::printf("zeros\n");
    return 0;
}
''')
    exe = tmp_path / "caller"
    libdir = os.path.join(ROOT, "canvas_ity_amd")
    subprocess.run(["g++", "-std=c++11", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe), "-L" + libdir, "-lcanvas_ity_hip",
                    "-Wl,-rpath," + libdir], check=True)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert p.stdout.strip() == "zeros"
    assert "no HIP device" in p.stderr


def test_tuning_cache_keys_and_round_trip(tmp_path):
    # the fine-stage autotuner's cache: one entry per batch size class, a
    # miss tunes the batch, a hit applies the cached configuration
    from canvas_ity_amd import tuning
    from canvas_ity_amd.stream import count_draws
    streams = [cia.scene(4, cia.scene_seed(0xC4, i)) for i in range(6)]
    key = tuning.size_class([cia.stream_size(s) for s in streams],
                            [count_draws(s) for s in streams])
    assert key.startswith("w") and "_n8" in key

    class FakeBatch:
        def __init__(self):
            self.applied = None
            self.tuned = 0

        def tune(self, reps):
            self.tuned += 1
            return 3, [1.0, 0.9, 0.8, 0.5, 0.7, 0.6]

        def configure(self, config):
            self.applied = config

    path = str(tmp_path / "sub" / "tune.json")
    b = FakeBatch()
    assert tuning.configure(b, streams, path=path) == (3, key, True)
    assert tuning.load(path)[key] == {"config": 3, "ms": [1.0, 0.9, 0.8, 0.5, 0.7, 0.6]}
    b2 = FakeBatch()
    assert tuning.configure(b2, streams, path=path) == (3, key, False)
    assert b2.applied == 3 and b2.tuned == 0
