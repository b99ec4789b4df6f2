"""Pins the CPU oracle (oracle/cih_oracle.c) before it is trusted as checker.

* the reference's exact known-answer test for get_image_data
  (test.cpp:1800-1830, hash 0xf53f9792);
* bit-identical fp32 bitmaps and sRGB8 against the reference header compiled
  from source (oracle/_ref), on seeded scenes of all five configurations and
  on feature scenes (skipped only where that build is absent);
* closed forms (tests/scenes.py) that hold independently of canvas_ity's
  code: rectangle overlap areas, polygon areas, supersampled coverage, the
  W3C Porter-Duff table, the IEC sRGB transfer with the ordered dither,
  gradient interpolation, blur mass conservation, identity resampling and
  stroke areas.  A planted mistake in any of those stages fails one of them.
"""
import math

import numpy as np
import pytest

import canvas_ity_amd as cia
import oracle
import scenes
from canvas_ity_amd.stream import COMPOSITE

REFERENCE = oracle.available("reference")


def render(stream, kind="port"):
    return oracle.render(stream, kind)


def test_kat_get_image_data():
    c = oracle.Canvas(256, 256, "port")
    c.replay(scenes.kat_stream())
    h = scenes.kat_hash(lambda out, w, hh, stride, x, y: c.get_image_data(w, hh, stride, x, y,
                                                                          out=out))
    assert h == scenes.KAT_HASH


@pytest.mark.skipif(not REFERENCE, reason="reference build (oracle/_ref) not present")
def test_kat_reference_build():
    c = oracle.Canvas(256, 256, "reference")
    c.replay(scenes.kat_stream())
    h = scenes.kat_hash(lambda out, w, hh, stride, x, y: c.get_image_data(w, hh, stride, x, y,
                                                                          out=out))
    assert h == scenes.KAT_HASH


CONFIG_SCENES = [(1, 0xC1, 0, 0, 0, 0), (2, 0xC2, 384, 320, 250, 0), (2, 0x2C2, 256, 256, 120, 1),
                 (3, 0xC3, 320, 288, 60, 0), (4, 0xC4, 0, 0, 0, 0), (4, 0x1C4, 0, 0, 0, 0),
                 (5, 0xC5, 256, 200, 80, 0), (5, 0x1C5, 200, 256, 60, 4)]


@pytest.mark.skipif(not REFERENCE, reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("cfg,seed,w,h,n,flags", CONFIG_SCENES)
def test_oracle_bit_identical_to_reference(cfg, seed, w, h, n, flags):
    stream = cia.scene(cfg, seed, w, h, n, flags)
    ref8, refl = render(stream, "reference")
    got8, gotl = render(stream, "port")
    if flags & cia.SCENE_EVENODD:
        pytest.skip("even-odd is an extension the reference does not have")
    assert np.array_equal(refl.view(np.uint32), gotl.view(np.uint32))
    assert np.array_equal(ref8, got8)


def seam_stream(seed):
    """A seeded five-segment bezier path, closed on two seeds of three, with
    a seeded line width, join, cap and (odd seeds) dash pattern."""
    from canvas_ity_amd.stream import Stream
    rng = np.random.default_rng(seed)
    s = Stream(160, 160)
    s.set_line_width(float(rng.uniform(1, 12))).set_line_join(seed % 3).set_line_cap(seed % 3)
    if seed % 2:
        s.set_line_dash([float(v) for v in rng.uniform(2, 30, 3)]).set_line_dash_offset(3.3)
    s.begin_path().move_to(*rng.uniform(-10, 170, 2))
    for _ in range(5):
        s.bezier_curve_to(*rng.uniform(-10, 170, 6))
    if seed % 3:
        s.close_path()
    return s.tobytes()


def seams(stream, kind):
    """(name, array) of every stage seam of a stream's path."""
    out = []
    for stroking in (False, True):
        c = oracle.Canvas(160, 160, kind)
        c.replay(stream)
        pts, subs = c.path_lines(stroking)
        out += [("lines", pts), ("subpaths", subs), ("runs", c.lines_to_runs()),
                ("runs", c.lines_to_runs(0.5, -2.0, 1))]
    c = oracle.Canvas(160, 160, kind)
    c.replay(stream)
    pts, subs = c.dash_lines()
    return out + [("lines", pts), ("subpaths", subs)]


@pytest.mark.skipif(not REFERENCE, reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("seed", range(12))
def test_stage_seams_match_the_reference(seed):
    # the oracle's polygons (filled, stroked, dashed) bit for bit, and its
    # runs: the same pixels in the same order, and deltas within 1e-6.  The
    # reference merges runs after std::sort, which leaves the order of runs it
    # ranks equal (same pixel and |delta|) unspecified, so a merged delta may
    # differ in its last bits; GPU tests compare with these seams where the
    # reference build is absent
    stream = seam_stream(seed)
    for k, ((name, ref), (_, got)) in enumerate(zip(seams(stream, "reference"),
                                                    seams(stream, "port"))):
        assert ref.shape == got.shape and ref.dtype == got.dtype, (k, name)
        if name != "runs":
            assert np.array_equal(ref.view(np.uint32), got.view(np.uint32)), (k, name)
            continue
        assert np.array_equal(ref[:, :2], got[:, :2]), k
        assert np.abs(ref[:, 2] - got[:, 2]).max(initial=0.0) <= 1e-6, k
        assert np.count_nonzero(ref[:, 2] != got[:, 2]) <= max(2, len(ref) // 200), k


@pytest.mark.parametrize("x0", [4.0, 4.5, 4.25])
def test_stage_seams_closed_form(x0):
    # a rectangle from (x0, 5) to (6, 8): one closed subpath of its corners,
    # and runs whose running sum along each covered row is minus the covered
    # fraction of each pixel (the path is clockwise on screen), 0 from x = 6
    from canvas_ity_amd.stream import Stream
    s = Stream(16, 16)
    s.begin_path().move_to(x0, 5).line_to(6, 5).line_to(6, 8).line_to(x0, 8).close_path()
    c = oracle.Canvas(16, 16, "port")
    c.replay(s.tobytes())
    pts, subs = c.path_lines(False)
    assert subs[0].tolist() == [5, 1]
    assert pts[:5].tolist() == [[x0, 5], [6, 5], [6, 8], [x0, 8], [x0, 5]]
    runs = c.lines_to_runs()
    assert set(runs[:, 1].astype(int).tolist()) == {5, 6, 7}
    for y in (5, 6, 7):
        row = runs[runs[:, 1] == y]
        sums = np.zeros(16)
        for x, _, d in row:
            sums[int(x):] += d
        expect = np.clip(np.minimum(np.arange(16) + 1.0, 6.0) - np.maximum(np.arange(16), x0),
                         0.0, 1.0)
        assert np.array_equal(-sums, expect), (y, sums.tolist())


def test_fractional_rectangle_coverage():
    rect = (3.25, 5.6, 17.3, 9.15)
    _, lin = render(scenes.fractional_rect(32, rect))
    exp = scenes.expected_rect_alpha(32, rect)
    assert np.abs(lin[..., 3] - exp).max() < 2e-6
    assert np.abs(lin[..., 0] - exp).max() < 2e-6  # white: every channel = coverage


def test_polygon_area():
    _, lin = render(scenes.polygon())
    assert abs(lin[..., 3].astype(np.float64).sum() - scenes.shoelace(scenes.STAR)) < 5e-4


def test_supersampled_triangle():
    _, lin = render(scenes.triangle(16))
    exp = scenes.supersampled_coverage(scenes.TRIANGLE, 16, n=64)
    assert np.abs(lin[..., 3] - exp).max() < 0.03
    assert abs(lin[..., 3].astype(np.float64).sum() - scenes.shoelace(scenes.TRIANGLE)) < 5e-4


def test_orientation_and_integer_translation():
    pts = scenes.STAR
    _, a = render(scenes.polygon(pts))
    _, b = render(scenes.polygon(pts[::-1]))
    assert np.abs(a - b).max() < 1e-6
    moved = [(x + 3, y + 2) for x, y in pts]
    _, c = render(scenes.polygon(moved, 47))
    assert np.abs(c[2:46, 3:47] - a).max() < 1e-5


@pytest.mark.parametrize("op", sorted(scenes.PORTER_DUFF))
def test_porter_duff_table(op):
    _, lin = render(scenes.composite_scene(op))
    assert np.abs(lin - scenes.expected_composite(op)).max() < 2e-6


def test_srgb_readback_and_dither():
    rgba8, _ = render(scenes.solid())
    exp = scenes.expected_solid_srgb8()
    # truncation of 255*v + dither; float32 vs float64 may differ by one LSB
    # only where the value is within rounding of an integer
    near = np.abs(exp - np.round(exp)) < 1e-3
    got = rgba8.astype(np.float64)
    assert np.all((got == np.floor(exp)) | near)


def test_linear_gradient():
    _, lin = render(scenes.linear_gradient())
    assert np.abs(lin - scenes.expected_linear_gradient()).max() < 2e-6


@pytest.mark.parametrize("blur", [0.0, 3.0, 11.5, 32.0])
def test_shadow_mass_and_symmetry(blur):
    _, lin = render(scenes.shadow_scene(blur))
    # the shape covers [6, 20) x [6, 16); its shadow, offset by (80, 76) and
    # spread by at most 3 (r + 1) <= 48 px, does not overlap it
    alpha = lin[..., 3].astype(np.float64)
    shape = scenes.SHADOW_RECT[2] * scenes.SHADOW_RECT[3]
    assert abs(alpha[:21, :21].sum() - shape) < 1e-4
    shadow = alpha.copy()
    shadow[:21, :21] = 0.0
    assert abs(shadow.sum() - scenes.shadow_mass_expected()) < 2e-3
    ys, xs = np.nonzero(shadow > 1e-4)
    assert abs((xs.min() + xs.max()) / 2 + 0.5 - (13 + scenes.SHADOW_OFFSET[0])) <= 0.5 + 1e-9
    assert abs((ys.min() + ys.max()) / 2 + 0.5 - (11 + scenes.SHADOW_OFFSET[1])) <= 0.5 + 1e-9


def test_identity_resampling():
    img = scenes.noise_image()
    _, lin = render(scenes.identity_image(img))
    assert np.abs(lin - scenes.expected_image_linear(img)).max() < 2e-6


@pytest.mark.parametrize("cap", [0, 1, 2])
def test_stroke_area(cap):
    _, lin = render(scenes.straight_stroke(cap))
    area = lin[..., 3].astype(np.float64).sum()
    tol = 1e-3 if cap < 2 else 0.02 * scenes.stroke_area(cap)
    assert abs(area - scenes.stroke_area(cap)) < tol


def test_even_odd_extension():
    _, nz = render(scenes.nested_squares(False))
    _, eo = render(scenes.nested_squares(True))
    assert np.all(nz[12:20, 12:20, 3] == 1.0)
    assert np.all(eo[12:20, 12:20, 3] == 0.0)
    assert np.all(eo[5:9, 5:27, 3] == 1.0)


@pytest.mark.parametrize("evenodd", [False, True])
@pytest.mark.parametrize("name", ["double_rect", "ring"])
def test_even_odd_fractional_windings(name, evenodd):
    """The oracle's coverage rule against closed forms at windings in (1, 2):
    two copies of one fractional rectangle (s = 2a) and a square with a
    fractional inner square (s = 1 + a)."""
    build = getattr(scenes, name)
    expected = getattr(scenes, "expected_" + name)(evenodd)
    _, lin = render(build(evenodd))
    assert np.abs(lin[..., 3].astype(np.float64) - expected).max() <= 1e-6


def test_even_odd_scenes_reach_the_fold():
    """Both fractional-winding scenes have pixels with 1 < s < 2, where an
    even-odd rule without the triangle fold, or folded with period 1, gives
    a different coverage than the closed form."""
    for s in scenes.winding_sums():
        fold = (s > 1.0) & (s < 2.0)
        assert fold.sum() >= 20
        right = scenes.even_odd(s)[fold]
        no_fold = np.minimum(np.fmod(s, 2.0), 1.0)[fold]
        period_one = np.fmod(s, 1.0)[fold]
        assert np.abs(right - no_fold).min() > 0.01
        assert np.abs(right - period_one).max() > 0.01


def test_malformed_stream_rejected():
    stream = bytearray(scenes.solid())
    stream[16:20] = (99).to_bytes(4, "little")
    c = oracle.Canvas(8, 8, "port")
    with pytest.raises(ValueError):
        c.replay(bytes(stream))


# ---- clip masks (hpp:3057-3099) ------------------------------------------------

CLIP_RECTS = [(3.25, 5.6, 17.3, 9.15), (8.0, 2.5, 12.75, 20.0)]


@pytest.mark.skipif(not oracle.available("reference"), reason="reference build absent")
@pytest.mark.parametrize("name", sorted(scenes.clip_scenes()))
def test_clip_oracle_matches_reference(name):
    stream = scenes.clip_scenes()[name]
    r8, rl = render(stream, "reference")
    p8, pl = render(stream, "port")
    assert np.array_equal(rl.view(np.uint32), pl.view(np.uint32))
    assert np.array_equal(r8, p8)


@pytest.mark.parametrize("n", [1, 2])
def test_clip_closed_form(n):
    clips = CLIP_RECTS[:n]
    _, lin = render(scenes.clipped_fill(clips))
    exp = scenes.expected_clipped_alpha(clips)
    assert np.abs(lin[..., 3] - exp).max() < 4e-6
    assert np.all(lin[exp == 0.0] == 0.0)  # nothing outside the mask


@pytest.mark.skipif(not oracle.available("reference"), reason="reference build absent")
@pytest.mark.parametrize("name", sorted(scenes.reference_clip_tests()))
def test_reference_clip_tests_oracle(name):
    stream = scenes.reference_clip_tests()[name]
    r8, rl = render(stream, "reference")
    p8, pl = render(stream, "port")
    assert np.array_equal(rl.view(np.uint32), pl.view(np.uint32))
    assert np.array_equal(r8, p8)

