"""The batch's fine-stage launch shapes on the GPU (collected last).

A batch's fine stage has six configurations (cih_batch_configure): work items
of 256, 128 or 64 pixels of a canvas row, and one workgroup per item or a
persistent grid whose waves drain per-XCD item queues.  Each canvas's pixels
do not depend on which runs them: every configuration must give the same bits
as the default, and the tuner must pick one of them.
"""
import numpy as np
import pytest

import canvas_ity_amd as cia
import scenes

pytestmark = pytest.mark.gpu


def test_fine_configurations_are_bit_identical(device):
    streams = [cia.scene(4, cia.scene_seed(0xC4, i)) for i in range(10)]
    b = cia.Batch(streams, device)
    b.execute(to_host=True)
    ref = [b.linear(i).copy() for i in range(len(streams))]
    for config in range(cia.FINE_CONFIGS):
        b.configure(config)
        for _ in range(2):  # the checked run, then the captured graph
            b.execute(to_host=True)
            for i in range(len(streams)):
                got = b.linear(i)
                assert np.array_equal(ref[i].view(np.uint32), got.view(np.uint32)), (config, i)
    best, ms = b.tune(reps=2)
    assert 0 <= best < cia.FINE_CONFIGS and min(ms) > 0.0
    b.execute(to_host=True)
    for i in range(len(streams)):
        assert np.array_equal(ref[i].view(np.uint32), b.linear(i).view(np.uint32)), i
    b.close()


def test_masked_frames_keep_the_default_configuration(device):
    # a frame with draws under a clip mask runs 256-pixel items, one
    # workgroup each, whatever the batch is configured to
    clips = scenes.clip_scenes()
    streams = [clips["nested"], cia.scene(1, 0xC1, 72, 56, 10, 0), clips["op_source_in"]]
    b = cia.Batch(streams, device)
    b.execute(to_host=True)
    ref = [b.linear(i).copy() for i in range(len(streams))]
    for config in range(1, cia.FINE_CONFIGS):
        b.configure(config)
        b.execute(to_host=True)
        for i in range(len(streams)):
            assert np.array_equal(ref[i].view(np.uint32), b.linear(i).view(np.uint32)), (config, i)
    b.close()
