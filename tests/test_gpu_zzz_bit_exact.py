"""Seeded scenes bit for bit on the GPU (collected last).

The kernels follow the reference's float operations with no contraction,
correctly rounded division and square root, round joins' acosf/tanf as
glibc computes them and read-back bytes as glibc's powf gives them, and they
sum equal-magnitude runs in the oracle's order.  Two seeded scenes of each
configuration, and two with even-odd fills (the oracle's extension), must
then render with fp32 pixels and sRGB8 bytes equal to the
CPU oracle's, as they do in the CPU emulation of the kernels
(tests/test_cpu_emulation.py::test_scenes_bit_for_bit).  Tighter than the
parity bar of the other GPU tests.
"""
import numpy as np
import pytest

import canvas_ity_amd as cia
import oracle
import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,seed,w,h,n,flags", scenes.BIT_EXACT_SCENES,
                         ids=["c%d-%x" % (c[0], c[1]) for c in scenes.BIT_EXACT_SCENES])
def test_seeded_scene_bit_for_bit(device, cfg, seed, w, h, n, flags):
    stream = cia.scene(cfg, seed, w, h, n, flags)
    want8, wantl = oracle.render(stream, "port")
    got8, gotl = cia.render(stream, device)
    differ = (gotl.view(np.uint32) != wantl.view(np.uint32)).any(axis=2)
    assert not differ.any(), (int(differ.sum()), float(np.abs(gotl - wantl).max()))
    assert np.array_equal(got8, want8)
