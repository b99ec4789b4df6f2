"""GPU parity on the reference's own test suite (test/test.cpp).

Each of the 76 tests of the reference's table, restated as a command stream
(tests/reference_suite.py), rendered by the HIP library on the MI355X and
compared with the oracle to the parity bar (tests/parity.py), and held to the
reference suite's own pass criterion: the image hash within Hamming distance 5
of the golden table (test.cpp:2186-2261, 2614-2618).
"""
import pytest

import canvas_ity_amd as cia
import oracle
import reference_suite
import scenes
from parity import assert_parity

pytestmark = pytest.mark.gpu

HASHES = scenes.reference_hashes()


@pytest.mark.parametrize("name", sorted(HASHES))
def test_reference_test_on_gpu(device, name):
    stream = reference_suite.stream(name)
    r8, rl = oracle.render(stream, "port")
    g8, gl = cia.render(stream, device)
    assert_parity(r8, rl, g8, gl, name)
    assert scenes.hash_distance(scenes.image_hash(g8), HASHES[name]["hash"]) <= 5


def test_reference_suite_as_one_batch(device):
    # the whole suite as one batch: every canvas as rendered alone
    streams = [reference_suite.stream(n) for n in sorted(HASHES)]
    b = cia.Batch(streams, device)
    b.execute(to_host=True)
    for i, name in enumerate(sorted(HASHES)):
        assert scenes.hash_distance(scenes.image_hash(b.rgba8(i)), HASHES[name]["hash"]) <= 5, name
    b.close()
