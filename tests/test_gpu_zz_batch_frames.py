"""Batches cut into many frames on the GPU (collected late).

A batch cuts its draws, in order, into frames: one over several canvases
holds at most 2048 draws and 2M canvas pixels, one over a single canvas at
most 8192 draws.  Each canvas of a batch of several frames, some across a
canvas, renders bit-identical to the plain per-canvas path, on the checked
run and on the graph replay.
"""
import numpy as np
import pytest

import canvas_ity_amd as cia

pytestmark = pytest.mark.gpu

FRAME_DRAWS, BATCH_DRAWS, BATCH_PIXELS = 8192, 2048, 2 << 20


def check_frames(b):
    stats, frames = b.stats(), b.frames()
    at = 0
    for f in frames:
        assert f["first_draw"] == at and f["draws"] > 0, (f, at)
        at += f["draws"]
        if f["canvases"] == 1:
            assert f["draws"] <= FRAME_DRAWS, f
        else:
            assert f["draws"] <= BATCH_DRAWS and f["pixels"] <= BATCH_PIXELS, f
    assert at == stats["draws"] and len(frames) == stats["frames"], (at, stats)
    return frames


BATCHES = {
    # three 64x64 canvases of 2000 paths each (about 6,000 draws)
    "dense": lambda: [cia.scene(4, cia.scene_seed(0x0C, i), 64, 64, 2000) for i in range(3)],
    # 24 canvases of the batch mix (sizes log-uniform 64..1024, 10..400 paths)
    "mix": lambda: [cia.scene(4, cia.scene_seed(0xC4, i)) for i in range(24)],
}


@pytest.mark.parametrize("name", sorted(BATCHES))
def test_batch_of_many_frames_matches_plain(device, name):
    streams = BATCHES[name]()
    alone = [cia.render(s, device) for s in streams]
    b = cia.Batch(streams, device)
    frames = check_frames(b)
    assert len(frames) > 1, frames
    for _ in range(2):  # the checked run, then the captured graph
        b.execute(to_host=True)
        for i in range(len(streams)):
            assert np.array_equal(alone[i][1].view(np.uint32), b.linear(i).view(np.uint32)), \
                (name, i)
            assert np.array_equal(alone[i][0], b.rgba8(i)), (name, i)
    b.close()
