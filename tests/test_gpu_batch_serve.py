"""Batch and serving paths on the GPU.

* A batch of seeded canvases of the batch mix (sizes log-uniform 64..1024,
  10..400 paths) renders each canvas exactly as the plain per-canvas path does
  (bit-identical fp32 and sRGB8: both run the same kernels, and each draw's
  coverage comes from its own runs alone, so the grouping of canvases into
  frames cannot change a result), and within the parity bar of the CPU oracle.
* A batch above 2^30 canvas pixels is refused before anything runs on the
  device (batches of many draws are cut into bounded frames; their GPU test
  is collected last, in test_gpu_zz_batch_frames.py).
* The server returns, for requests submitted together and interleaved, the
  same pixels each request gets when rendered alone; a memory budget smaller
  than the burst forces admission to wait for memory and still completes
  everything with the same pixels; a request larger than the budget is
  refused.
"""
import numpy as np
import pytest

import canvas_ity_amd as cia
import oracle
from parity import assert_parity

pytestmark = pytest.mark.gpu


def batch_streams(n, seed=0xC4):
    return [cia.scene(4, cia.scene_seed(seed, i)) for i in range(n)]


def test_batch_matches_plain_and_oracle(device):
    streams = batch_streams(10)
    b = cia.Batch(streams, device)
    b.execute(to_host=True)
    for i, s in enumerate(streams):
        alone8, alonel = cia.render(s, device)
        got8, gotl = b.rgba8(i), b.linear(i)
        assert np.array_equal(alonel.view(np.uint32), gotl.view(np.uint32)), i
        assert np.array_equal(alone8, got8), i
        if i < 4:
            ref8, refl = oracle.render(s, "port")
            assert_parity(ref8, refl, got8, gotl, "batch canvas %d" % i)
    b.close()


def test_batch_refuses_oversized(device):
    # 2 canvases of 32768x32768 hold 2^31 pixels, above the 2^30-pixel bound
    from canvas_ity_amd.stream import Stream
    streams = [Stream(32768, 32768).tobytes() for _ in range(2)]
    with pytest.raises(cia.CihError, match="budget"):
        cia.Batch(streams, device)


def test_batch_graph_replay_is_stable(device):
    streams = batch_streams(6, seed=0x77)
    b = cia.Batch(streams, device)
    b.execute(to_host=True)
    first = [b.rgba8(i) for i in range(len(streams))]
    for _ in range(3):  # later executes replay the captured graph
        b.execute(to_host=True)
        for i in range(len(streams)):
            assert np.array_equal(first[i], b.rgba8(i)), i
    b.close()


def serve_streams(n):
    out = []
    for i in range(n):
        flags = cia.SCENE_SHADOWS if i % 4 == 3 else 0
        out.append(cia.scene(5, cia.scene_seed(0xC5, i), 0, 0, 60 + 20 * i, flags))
    return out


def run_server(streams, budget, max_batch=8):
    server = cia.Server(hbm_budget=budget, max_batch=max_batch)
    ids = [server.submit(s) for s in streams]
    results = {}
    while len(results) < len(ids):
        rid = server.wait(timeout_ms=60000)
        assert rid is not None, "server stalled"
        results[rid] = server.result(rid)[0]
    stats = server.stats()
    server.stop()
    return [results[i] for i in ids], stats


def test_server_alone_equals_mix(device):
    streams = serve_streams(8)
    alone = [cia.render(s, device)[0] for s in streams]
    mixed, stats = run_server(streams, budget=8 << 30)
    for i in range(len(streams)):
        assert np.array_equal(alone[i], mixed[i]), i
    assert stats["served"] == len(streams)


def test_server_admission_waits_for_memory(device):
    streams = serve_streams(5)
    largest = max(16 * w * h for w, h in map(cia.stream_size, streams))
    budget = max(1 << 20, int(largest * 1.2))  # room for about one request at a time
    alone = [cia.render(s, device)[0] for s in streams]
    mixed, stats = run_server(streams, budget=budget)
    for i in range(len(streams)):
        assert np.array_equal(alone[i], mixed[i]), i
    assert stats["batches"] >= 2


def test_server_refuses_oversized_request(device):
    server = cia.Server(hbm_budget=1 << 20, max_batch=4)
    with pytest.raises(cia.CihError):
        server.submit(cia.scene(1, 1))  # 512x512 needs 4 MiB of pixels
    server.stop()
