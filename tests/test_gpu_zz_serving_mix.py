"""The server at the serving mix's own path counts on the GPU (collected late).

Sixteen requests of the serving mix as it is configured -- canvases
log-uniform 128..2048 per side, 50..2000 paths from the config-1 and effects
mixes, shadows (with draw_image) on about 10% -- submitted together.  The
server admits several requests per batch up to the multi-canvas frame bounds
and runs a larger request in a batch of its own; every result must equal the
request rendered alone on the plain path, bit for bit.
"""
import numpy as np
import pytest

import canvas_ity_amd as cia
from canvas_ity_amd.stream import records

pytestmark = pytest.mark.gpu

SHADOW_BLUR = 9  # opcode of set_shadow_blur in the command stream


def test_server_matches_alone_at_the_configured_path_counts(device):
    streams = [cia.scene(5, cia.scene_seed(0xC5, i)) for i in range(16)]
    shadowed = [i for i, s in enumerate(streams) if any(r[0] == SHADOW_BLUR for r in records(s))]
    assert shadowed, "the sample holds no shadowed request"
    alone = [cia.render(s, device)[0] for s in streams]
    server = cia.Server(hbm_budget=8 << 30, max_batch=64, device=device)
    ids = [server.submit(s) for s in streams]
    got = {}
    while len(got) < len(ids):
        rid = server.wait(timeout_ms=120000)
        assert rid is not None, "server stalled"
        got[rid] = server.result(rid)[0]
    stats = server.stats()
    server.stop()
    assert stats["served"] == len(streams), stats
    for i, rid in enumerate(ids):
        assert np.array_equal(alone[i], got[rid]), (i, i in shadowed)
