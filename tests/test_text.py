"""Text (canvas_ity.hpp:1299-1314, 1526-1846, 3205-3311) on the CPU.

The checks that need no device:

* the font fixtures (tests/golden/fonts, decoded from the reference's test
  suite by tests/golden/make_fixtures.py) are accepted or rejected alike by
  the library's cih_font_check, the oracle and the reference build;
* measure_text agrees across the library (cih_text_measure), the oracle and
  the reference on strings that exercise every cmap path, invalid UTF-8 and
  the whitespace mapping, and meets its closed forms (additivity, whitespace
  as space, invalid bytes as U+FFFD);
* the oracle's renders of the reference suite's text tests are bit-identical
  to the reference build's and their image hashes equal the reference's own
  golden table (test.cpp:2186-2261; pass within Hamming distance 5);
* the outline segments the library lays out (cih_text_outline) are well
  formed: closed subpaths of 4 points per cubic, straight segments as
  (p, p, q, q), maximum_width compresses x about the anchor.

Device renders of the same scenes are compared with the oracle in the GPU
suite (test_gpu_parity.py) and in the CPU emulation (test_cpu_emulation.py).
"""
import numpy as np
import pytest

import canvas_ity_amd as cia
import oracle
import scenes

needs_reference = pytest.mark.skipif(not oracle.available("reference"),
                                     reason="the reference build (oracle/_ref) is missing")

FONTS = "abcdefg"
STRINGS = [b"", b"C", b"D", b"DDD", b"Canvas Ity", b"HIty", b"* Cats", b"\t\v\f\r\n ",
           "Í".encode(), "\U0010fffd".encode(), b"\xf0I", b"\x80\xbf", b"\xc3", b"\xe2\x82",
           b"\xf4\x8f\xbf", "Z€é".encode(), b"nvasty*", b"QQQ?", bytes(range(1, 128))]


def test_font_fixtures_accepted_alike():
    for name in FONTS:
        data = scenes.font(name)
        ours = cia.font_units(data) is not None
        assert ours == (name in "abc"), name
        assert (oracle.measure_text(data, 20.0, b"D") > 0) == ours, name
        if oracle.available("reference"):
            assert oracle.font_accepted(data) == ours, name


def test_truncated_and_damaged_fonts_rejected():
    good = scenes.font("a")
    for cut in (0, 5, 11, 12, 100, len(good) - 1):
        data = good[:cut]
        assert cia.font_units(data) is None, cut
        if oracle.available("reference") and cut:
            assert not oracle.font_accepted(data), cut
    bad_magic = b"\x00\x02" + good[2:]
    assert cia.font_units(bad_magic) is None


@pytest.mark.parametrize("name", "abc")
def test_measure_text_agrees(name):
    data = scenes.font(name)
    for size in (10.0, 51.2, 77.7):
        for text in STRINGS:
            ours = cia.measure_text(data, size, text)
            port = oracle.measure_text(data, size, text)
            assert np.float32(ours) == np.float32(port), (name, size, text)
            if oracle.available("reference"):
                ref = oracle.measure_text(data, size, text, "reference")
                assert np.float32(ours) == np.float32(ref), (name, size, text)


@pytest.mark.parametrize("name", "abc")
def test_measure_text_closed_forms(name):
    data = scenes.font(name)
    units = cia.font_units(data)
    size = np.float32(64.0)
    scale = np.float32(size / np.float32(units))
    one = cia.measure_text(data, size, b"D")
    assert one > 0
    # advances in font units are integers: width / scale recovers them exactly
    assert float(np.float32(one) / scale) == round(float(np.float32(one) / scale))
    assert cia.measure_text(data, size, b"DDDD") == pytest.approx(4 * one, rel=1e-6)
    # tab, vertical tab, form feed, CR and LF measure as spaces
    assert cia.measure_text(data, size, b"C\nD\tE") == cia.measure_text(data, size, b"C D E")
    # an invalid byte is U+FFFD and the next byte starts a new character
    assert cia.measure_text(data, size, b"\xf0I") == cia.measure_text(
        data, size, "�I".encode())
    assert cia.measure_text(data, size, b"") == 0.0


REFERENCE_TESTS = ["text_align", "text_baseline", "font", "fill_text", "stroke_text",
                   "measure_text", "example_button"]


def reference_text_tests():
    return scenes.reference_text_tests(256, measure=oracle.measure_text)


@pytest.mark.parametrize("name", REFERENCE_TESTS)
def test_reference_text_tests_match_golden_hashes(name):
    # the oracle's render of the reference's own test hashes (test.cpp:2365)
    # within the suite's Hamming distance of the reference's golden table
    stream = reference_text_tests()[name]
    img, _ = oracle.render(stream, "port")
    golden = scenes.reference_hashes()[name]["hash"]
    assert scenes.hash_distance(scenes.image_hash(img), golden) <= 5


@needs_reference
@pytest.mark.parametrize("name", REFERENCE_TESTS)
def test_reference_text_tests_oracle_matches_reference(name):
    stream = reference_text_tests()[name]
    r8, rl = oracle.render(stream, "reference")
    p8, pl = oracle.render(stream, "port")
    assert np.array_equal(rl.view(np.uint32), pl.view(np.uint32))
    assert np.array_equal(r8, p8)


@needs_reference
@pytest.mark.parametrize("seed", range(10))
def test_text_scene_oracle_matches_reference(seed):
    # Glyph edges on a common grid give coverage runs that tie on (y, x,
    # |delta|) with opposite signs; the reference sorts runs with std::sort
    # (hpp:2243), which leaves such ties in an order of the library's making,
    # and the oracle puts the negative one first.  The sums then differ in the
    # last place: the linear buffers agree to 1e-6 and the sRGB8 images match.
    stream = scenes.text_scene(seed)
    r8, rl = oracle.render(stream, "reference")
    p8, pl = oracle.render(stream, "port")
    assert np.abs(rl - pl).max() <= 1e-6
    assert np.array_equal(r8, p8)


def test_image_hash_detects_shift_and_tolerates_noise():
    img, _ = oracle.render(reference_text_tests()["text_align"], "port")
    h = scenes.image_hash(img)
    assert scenes.hash_distance(h, scenes.image_hash(np.roll(img, 1, axis=1))) > 5
    noisy = img.astype(np.int16)
    noisy[100:110, 100:110, :3] += 1
    assert scenes.hash_distance(h, scenes.image_hash(np.clip(noisy, 0, 255).astype(np.uint8))) <= 5


@pytest.mark.parametrize("stroking", [False, True])
def test_outline_well_formed(stroking):
    data = scenes.font("a")
    pts, sub = cia.text_outline(data, 40.0, "Canvas Ity* \U0010fffd", 10.0, 50.0,
                                stroking=stroking, line_width=3.0)
    assert len(sub) > 0 and np.all(sub[:, 1] == 1)
    assert np.all(sub[:, 0] % 4 == 0) and sub[:, 0].sum() == len(pts)
    q = pts.reshape(-1, 4, 2)
    at = 0
    for count, _ in sub:
        seg = q[at // 4:(at + count) // 4]
        # each contour ends where it began (the closing segment)
        assert np.array_equal(seg[-1, 3], seg[0, 0])
        # straight segments are (p, p, q, q)
        straight = np.all(seg[:, 0] == seg[:, 1], axis=1) & np.all(seg[:, 2] == seg[:, 3], axis=1)
        assert straight.any()
        at += count


def test_outline_maximum_width_and_alignment():
    data = scenes.font("a")
    natural = cia.measure_text(data, 40.0, "Canvas")
    full, _ = cia.text_outline(data, 40.0, "Canvas", 20.0, 50.0)
    half, _ = cia.text_outline(data, 40.0, "Canvas", 20.0, 50.0, maximum_width=natural / 2)
    # the width reduction scales x about the anchor and leaves y alone
    assert np.allclose((half[:, 0] - 20.0), (full[:, 0] - 20.0) * 0.5, atol=1e-3)
    assert np.allclose(half[:, 1], full[:, 1], atol=1e-4)
    assert len(cia.text_outline(data, 40.0, "Canvas", 20.0, 50.0, maximum_width=0.0)[0]) == 0
    right, _ = cia.text_outline(data, 40.0, "Canvas", 20.0, 50.0, align=1)
    center, _ = cia.text_outline(data, 40.0, "Canvas", 20.0, 50.0, align=2)
    assert np.allclose(right[:, 0], full[:, 0] - natural, atol=1e-3)
    assert np.allclose(center[:, 0], full[:, 0] - natural / 2, atol=1e-3)
    top, _ = cia.text_outline(data, 40.0, "Canvas", 20.0, 50.0, baseline=1)
    assert np.allclose(top[:, 0], full[:, 0]) and np.all(top[:, 1] > full[:, 1])
