"""Round joins on the GPU, bit for bit (collected late).

A round join's cubic comes from acosf and tanf of the turn (canvas_ity.hpp:
1994-1997).  The kernels evaluate both as glibc does, so a stroke's polygons
equal the reference's stroked lines bit for bit, and its fp32 pixels the
reference's render.  The CPU oracle's seams stand in where the reference
build is absent (tests/test_oracle.py pins them to it).
"""
import numpy as np
import pytest

import canvas_ity_amd as cia
import oracle
import scenes

pytestmark = pytest.mark.gpu

KIND = "reference" if oracle.available("reference") else "port"
STROKES = scenes.round_join_strokes()


@pytest.mark.parametrize("name", sorted(STROKES))
def test_round_join_strokes_bit_for_bit(device, name):
    one, before = STROKES[name]
    w, h = cia.stream_size(one)
    ref = oracle.Canvas(w, h, KIND)
    ref.replay(before)
    rp, rs = ref.path_lines(True)
    c = cia.Canvas(w, h, device)
    c.replay(one)
    c.flush()
    gp, gs = c.trace_polygons(0)
    c.close()
    assert np.array_equal(rs, gs.astype(np.int32))
    assert np.array_equal(rp.view(np.uint32), gp.view(np.uint32))
    _, rl = oracle.render(one, KIND)
    _, gl = cia.render(one, device)
    assert np.array_equal(rl.view(np.uint32), gl.view(np.uint32))
