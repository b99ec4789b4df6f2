"""is_point_in_path (canvas_ity.hpp:3101-3132) on the host, against the
reference compiled from source and against closed forms.

The library answers the query without a device (cih_path_contains), so these
run in the CPU suite.  Each seeded path is built by replaying the same
command stream into the reference; its device-space path (the reference's
private `path`, read through the oracle's seam) is handed to the library, and
every query point -- random points, the path's own vertices and edge
midpoints, where the on-boundary rules decide -- must get the reference's
answer.
"""
import ctypes
import math

import numpy as np
import pytest

import canvas_ity_amd as cia
import oracle
from canvas_ity_amd.stream import Stream

needs_reference = pytest.mark.skipif(not oracle.available("reference"),
                                     reason="the reference build (oracle/_ref) is missing")


def random_path(seed, size=96):
    rng = np.random.default_rng(seed)
    u = lambda lo, hi: float(rng.uniform(lo, hi))
    s = Stream(size, size)
    s.translate(u(-8, 8), u(-8, 8))
    s.rotate(u(-0.6, 0.6))
    s.scale(u(0.7, 1.3), u(0.7, 1.3))
    s.begin_path()
    for sub in range(int(rng.integers(1, 4))):
        kind = int(rng.integers(0, 4))
        if kind == 3:
            s.rectangle(u(0, size / 2), u(0, size / 2), u(4, size / 2), u(4, size / 2))
            continue
        if kind == 2:
            s.arc(u(20, size - 20), u(20, size - 20), u(4, 30), u(0, 6.3), u(0, 6.3),
                  bool(rng.integers(0, 2)))
        else:
            s.move_to(u(0, size), u(0, size))
        for _ in range(int(rng.integers(2, 8))):
            op = int(rng.integers(0, 3))
            if op == 0:
                s.line_to(u(0, size), u(0, size))
            elif op == 1:
                s.quadratic_curve_to(u(0, size), u(0, size), u(0, size), u(0, size))
            else:
                s.bezier_curve_to(u(0, size), u(0, size), u(0, size), u(0, size),
                                  u(0, size), u(0, size))
        if rng.integers(0, 3):
            s.close_path()
    return s.tobytes()


def reference_path(lib, handle):
    cap = 1 << 16
    pts = np.zeros((cap, 2), dtype=np.float32)
    info = np.zeros((cap, 2), dtype=np.int32)
    nsub = ctypes.c_long()
    n = lib.cih_ref_path(handle, pts.ctypes.data_as(ctypes.c_void_p), cap,
                         info.ctypes.data_as(ctypes.c_void_p), cap, ctypes.byref(nsub))
    return pts[:n].copy(), info[:nsub.value].astype(np.uint32)


@needs_reference
@pytest.mark.parametrize("seed", range(8))
def test_matches_reference(seed):
    ref = oracle._load("reference")
    lib = ref["lib"]
    stream = random_path(seed)
    w, h = cia.stream_size(stream)
    handle = ref["create"](w, h)
    buf = ctypes.create_string_buffer(stream, len(stream))
    assert ref["replay"](handle, buf, len(stream)) == 0
    try:
        pts, sub = reference_path(lib, handle)
        assert len(pts) > 0
        rng = np.random.default_rng(1000 + seed)
        queries = [tuple(q) for q in rng.uniform(-10, 110, size=(400, 2)).astype(np.float32)]
        queries += [tuple(p) for p in pts]  # vertices: boundary cases
        queries += [tuple((pts[i] + pts[i + 1]) * np.float32(0.5)) for i in range(len(pts) - 1)]
        queries += [(float(p[0]), float(p[1]) + 0.5) for p in pts[:50]]
        mismatches = []
        for x, y in queries:
            want = lib.cih_ref_is_point_in_path(handle, float(x), float(y)) != 0
            got = cia.path_contains(pts, sub, x, y)
            if got != want:
                mismatches.append((x, y, want, got))
        assert not mismatches, mismatches[:5]
    finally:
        ref["destroy"](handle)


def rect_path(x0, y0, x1, y1):
    # move_to + three line_to + close_path: degenerate cubics (p1, p2, p2)
    corners = [(x0, y0), (x1, y0), (x1, y1), (x0, y1), (x0, y0)]
    pts = [corners[0]]
    for a, b in zip(corners, corners[1:]):
        pts += [a, b, b]
    return np.array(pts, dtype=np.float32), np.array([[len(pts), 1]], dtype=np.uint32)


def test_rectangle_closed_form():
    pts, sub = rect_path(10.0, 10.0, 50.0, 30.0)
    assert cia.path_contains(pts, sub, 20.0, 20.0)
    assert not cia.path_contains(pts, sub, 60.0, 20.0)
    assert not cia.path_contains(pts, sub, 20.0, 5.0)
    # the boundary counts as inside: a vertical edge and a horizontal edge
    assert cia.path_contains(pts, sub, 10.0, 20.0)
    assert cia.path_contains(pts, sub, 30.0, 10.0)


def test_nonzero_rule_and_orientation():
    outer, so = rect_path(0.0, 0.0, 40.0, 40.0)
    inner, si = rect_path(10.0, 10.0, 30.0, 30.0)
    both = np.concatenate([outer, inner])
    sub = np.array([[len(outer), 1], [len(inner), 1]], dtype=np.uint32)
    # same orientation: winding 2 in the hole -> inside (non-zero)
    assert cia.path_contains(both, sub, 20.0, 20.0)
    # reversed inner loop: winding 0 in the hole -> outside
    rev = np.concatenate([outer, inner[::-1]])
    assert not cia.path_contains(rev, sub, 20.0, 20.0)
    assert cia.path_contains(rev, sub, 5.0, 20.0)


def test_circle_closed_form():
    # a full circle of 8 cubic arcs, radius 20 at (50, 50)
    k = 4.0 / 3.0 * math.tan(math.pi / 16)
    pts = [(70.0, 50.0)]
    for i in range(8):
        a0, a1 = i * math.pi / 4, (i + 1) * math.pi / 4
        p0 = (50 + 20 * math.cos(a0), 50 + 20 * math.sin(a0))
        p1 = (50 + 20 * math.cos(a1), 50 + 20 * math.sin(a1))
        c0 = (p0[0] - k * 20 * math.sin(a0), p0[1] + k * 20 * math.cos(a0))
        c1 = (p1[0] + k * 20 * math.sin(a1), p1[1] - k * 20 * math.cos(a1))
        pts += [c0, c1, p1]
    pts = np.array(pts, dtype=np.float32)
    sub = np.array([[len(pts), 1]], dtype=np.uint32)
    rng = np.random.default_rng(7)
    for x, y in rng.uniform(25, 75, size=(300, 2)):
        r = math.hypot(x - 50, y - 50)
        if abs(r - 20) > 0.2:  # away from the flattening tolerance band
            assert cia.path_contains(pts, sub, x, y) == (r < 20), (x, y)


def test_malformed_path_rejected():
    pts = np.zeros((3, 2), dtype=np.float32)
    with pytest.raises(cia.CihError):
        cia.path_contains(pts, np.array([[3, 0]], dtype=np.uint32), 0.0, 0.0)
    with pytest.raises(cia.CihError):
        cia.path_contains(pts, np.array([[4, 0]], dtype=np.uint32), 0.0, 0.0)
