"""Renders scenes with the CPU emulation of the library (AddressSanitizer must be
preloaded) and compares each with the oracle.  TEST INFRASTRUCTURE ONLY.

Usage: run_scenes.py LIBRARY NAME...   prints "NAME ok" per scene, exits 1 on a
parity failure; an out-of-bounds access aborts with AddressSanitizer's report.
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import canvas_ity_amd as cia  # noqa: E402
import oracle  # noqa: E402
import scenes  # noqa: E402
from parity import assert_parity  # noqa: E402


def streams():
    out = {
        "plain_small": cia.scene(1, 0xC1, 72, 56, 10, 0),
        "large_small": cia.scene(2, 0xC2, 96, 80, 24, 0),
        "effects_small": cia.scene(3, 0xC3, 96, 96, 10, 0),
    }
    out.update({"clip_" + k: v for k, v in scenes.clip_scenes().items()})
    out.update({"reference_" + k: v for k, v in scenes.reference_clip_tests(64).items()})
    measure = lambda font, size, text: oracle.measure_text(font, size, text)
    out.update({"text_" + k: v
                for k, v in scenes.reference_text_tests(128, measure=measure).items()})
    out.update({"text_scene_%d" % k: scenes.text_scene(k, 96) for k in range(4)})
    out.update({"put_scene_%d" % k: scenes.put_scene(k) for k in range(2)})
    return out


def batch_and_server(device):
    """put_image_data recorded in a batch and in server requests: each canvas
    as rendered alone, where it is applied at once."""
    streams = [scenes.put_scene(1, 32), scenes.put_scene(2, 40)]
    alone = [cia.render(s, device) for s in streams]
    b = cia.Batch(streams, device)
    for _ in range(2):  # the checked run, then the captured graph
        b.execute(to_host=True)
        for i in range(len(streams)):
            assert np.array_equal(alone[i][1].view(np.uint32), b.linear(i).view(np.uint32)), i
    b.close()
    server = cia.Server(hbm_budget=1 << 28, max_batch=2, device=device)
    ids = [server.submit(s) for s in streams]
    got = {}
    while len(got) < len(ids):
        rid = server.wait(timeout_ms=600000)
        got[rid] = server.result(rid)[0]
    server.stop()
    for i, rid in enumerate(ids):
        assert np.array_equal(alone[i][0], got[rid]), i


def batch_small(device):
    """Small configurations, a clip scene and a text scene as one multi-canvas
    batch (one frame): each canvas as rendered alone."""
    all_streams = streams()
    names = ["plain_small", "large_small", "effects_small", "clip_nested", "text_scene_1"]
    alone = [cia.render(all_streams[n], device) for n in names]
    b = cia.Batch([all_streams[n] for n in names], device)
    b.execute(to_host=True)
    for i, n in enumerate(names):
        assert np.array_equal(alone[i][1].view(np.uint32), b.linear(i).view(np.uint32)), n
    b.close()


FRAME_DRAWS = 8192  # draws of a frame over one canvas


def frame_bounds():
    """The multi-canvas frame bounds the library was built with (the
    emulation lowers them through EMU_BATCH_DRAWS / EMU_BATCH_PIXELS)."""
    return (int(os.environ.get("EMU_BATCH_DRAWS", "2048")),
            int(os.environ.get("EMU_BATCH_PIXELS", str(2 << 20))))


def check_frames(b):
    """The batch's frames cover its draws in order, each within its bound;
    returns them."""
    draws_bound, pixels_bound = frame_bounds()
    stats, frames = b.stats(), b.frames()
    assert len(frames) == stats["frames"] and frames, stats
    at = 0
    for f in frames:
        assert f["first_draw"] == at and f["draws"] > 0 and f["canvases"] > 0, (f, at)
        at += f["draws"]
        if f["canvases"] == 1:
            assert f["draws"] <= FRAME_DRAWS, f
        else:
            assert f["draws"] <= draws_bound and f["pixels"] <= pixels_bound, f
    assert at == stats["draws"], (at, stats)
    # frames are filled greedily: fewer than twice the frames either bound asks
    pixels = sum(f["pixels"] for f in frames)
    need = -(-stats["draws"] // draws_bound) + -(-pixels // pixels_bound)
    assert len(frames) <= 2 * need + 2, (len(frames), need)
    return frames


def batch_frames(device):
    """A batch of any draw count is cut into frames within the multi-canvas
    bounds: the 128-canvas batch of the batch mix (14.4M pixels, 25,187
    draws; as one frame it faulted the device) and three 64x64 canvases of
    2000 paths each.  A batch above 2^30 canvas pixels is refused before
    anything is allocated."""
    big = [cia.scene(4, cia.scene_seed(0xC4, i)) for i in range(128)]
    many = [cia.scene(4, cia.scene_seed(0x0C, i), 64, 64, 2000) for i in range(3)]
    for streams in (big, many):
        b = cia.Batch(streams, device)
        frames = check_frames(b)
        assert len(frames) > 1, frames
        b.close()
    from canvas_ity_amd.stream import Stream
    huge = [Stream(32768, 32768).tobytes() for _ in range(2)]
    try:
        cia.Batch(huge, device)
    except cia.CihError as e:
        assert "budget" in str(e) and "canvas pixels" in str(e), e
    else:
        raise AssertionError("a batch above 2^30 pixels was accepted")


def batch_frames_render(device):
    """With the frame bounds lowered (EMU_BATCH_DRAWS=50,
    EMU_BATCH_PIXELS=10000) eight small canvases of the batch mix's paths are
    cut into many frames, some of them across a canvas: every canvas renders
    as it does alone, on the checked run and on the graph replay."""
    assert frame_bounds() == (50, 10000)
    sizes = [(48, 32, 20), (64, 48, 40), (32, 32, 12), (80, 60, 60), (40, 50, 24),
             (56, 36, 30), (32, 64, 45), (72, 48, 28)]
    streams = [cia.scene(4, cia.scene_seed(0x5F, i), w, h, n)
               for i, (w, h, n) in enumerate(sizes)]
    alone = [cia.render(s, device)[1] for s in streams]
    b = cia.Batch(streams, device)
    frames = check_frames(b)
    assert len(frames) >= 6, frames
    assert any(f["canvases"] > 1 for f in frames), frames
    assert sum(f["canvases"] for f in frames) > len(streams), frames  # a canvas split
    for _ in range(2):  # the checked run, then the captured graph
        b.execute(to_host=True)
        for i in range(len(streams)):
            assert np.array_equal(alone[i].view(np.uint32), b.linear(i).view(np.uint32)), i
    b.close()


def rectangles(n, seed):
    from canvas_ity_amd.stream import Stream
    rng = np.random.default_rng(seed)
    s = Stream(16, 16)
    for _ in range(n):
        s.set_color(0, *rng.uniform(0.2, 1.0, 4))
        x, y = rng.uniform(0, 12, 2)
        s.fill_rectangle(float(x), float(y), 3.5, 2.25)
    return s.tobytes()


def payload_limit(device):
    """Run with EMU_PAYLOAD_LIMIT=20000 (the emulation's stand-in for the 4 GiB
    of 32-bit payload offsets): a canvas holding more packet bytes is refused,
    a batch whose joined packets exceed it is refused, and the server serves
    such requests in separate batches with the pixels each gets alone."""
    assert os.environ.get("EMU_PAYLOAD_LIMIT") == "20000"
    try:
        cia.render(rectangles(100, 1), device)
    except cia.CihError as e:
        assert "budget" in str(e) and "bytes of draw packets" in str(e), e
    else:
        raise AssertionError("a canvas above the payload limit was accepted")
    pair = [rectangles(40, 2), rectangles(40, 3)]
    alone = [cia.render(s, device)[0] for s in pair]
    try:
        cia.Batch(pair, device)
    except cia.CihError as e:
        assert "budget" in str(e) and "bytes of draw packets" in str(e), e
    else:
        raise AssertionError("a batch above the payload limit was accepted")
    server = cia.Server(hbm_budget=1 << 24, max_batch=8, device=device)
    ids = [server.submit(s) for s in pair]
    got = {}
    while len(got) < len(ids):
        rid = server.wait(timeout_ms=600000)
        got[rid] = server.result(rid)[0]
    stats = server.stats()
    server.stop()
    for i, rid in enumerate(ids):
        assert np.array_equal(alone[i], got[rid]), i
    assert stats["batches"] == 2, stats


def fine_configs(device):
    """Every fine-stage configuration of a batch (item widths 256/128/64
    pixels, one workgroup per item or a persistent grid draining per-XCD
    queues) gives bit-identical pixels; the tuner picks one of them (timed
    on the smaller, masked batch) and the tuning cache round-trips it."""
    import tempfile
    from canvas_ity_amd import tuning
    all_streams = streams()
    # the second batch is one frame with masked draws (clip_nested): it keeps
    # the default configuration whatever the batch is configured to
    groups = (["plain_small", "large_small", "effects_small", "text_scene_1"],
              ["clip_nested", "plain_small"])
    for names in groups:
        batch_streams = [all_streams[n] for n in names]
        b = cia.Batch(batch_streams, device)
        b.execute(to_host=True)
        ref = [b.linear(i).copy() for i in range(len(names))]
        for config in range(1, cia.FINE_CONFIGS):
            b.configure(config)
            b.execute(to_host=True)
            for i, n in enumerate(names):
                got = b.linear(i)
                assert np.array_equal(ref[i].view(np.uint32), got.view(np.uint32)), (config, n)
        if names is not groups[-1]:
            b.close()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "tune.json")
        config, key, tuned = tuning.configure(b, batch_streams, path=path, reps=1)
        assert tuned and tuning.load(path)[key]["config"] == config
        assert 0 <= config < cia.FINE_CONFIGS
    b.execute(to_host=True)
    for i in range(len(names)):
        assert np.array_equal(ref[i].view(np.uint32), b.linear(i).view(np.uint32)), i
    b.close()


def canvas_execute(device):
    """The benchmark's path: a flushed canvas program run again from HBM
    (cih_canvas_execute, unchecked) gives the first render's pixels, and
    its statistics pass counts composited pixels."""
    stream = streams()["large_small"]
    w, h = cia.stream_size(stream)
    c = cia.Canvas(w, h, device)
    c.replay(stream)
    first8, firstl = c.rgba8().copy(), c.linear().copy()
    for _ in range(2):
        c.execute(clear=True, readback=True)
        device.synchronize()
        assert np.array_equal(firstl.view(np.uint32), c.linear().view(np.uint32))
        assert np.array_equal(first8, c.rgba8())
    stats = c.stats()
    assert stats["pixel_draws"] > 0 and stats["runs"] > 0 and stats["points"] > 0, stats
    assert np.array_equal(firstl.view(np.uint32), c.linear().view(np.uint32))
    c.close()


def reference_suite_batch(device):
    """The reference's 76 tests (test.cpp) as one batch: every canvas within
    the parity bar of the oracle and within the reference suite's own pass
    criterion, image hash within Hamming distance 5 of its golden table."""
    import reference_suite
    hashes = scenes.reference_hashes()
    names = sorted(hashes)
    batch_streams = [reference_suite.stream(n) for n in names]
    b = cia.Batch(batch_streams, device)
    b.execute(to_host=True)
    worst = 0
    for i, n in enumerate(names):
        g8, gl = b.rgba8(i), b.linear(i)
        r8, rl = oracle.render(batch_streams[i], "port")
        assert_parity(r8, rl, g8, gl, n)
        dist = scenes.hash_distance(scenes.image_hash(g8), hashes[n]["hash"])
        assert dist <= 5, (n, dist)
        worst = max(worst, dist)
    b.close()
    print("reference suite: %d canvases in one batch, worst hash distance %d"
          % (len(names), worst), flush=True)


def batch_mix(device, n):
    """n canvases of the batch mix (sizes log-uniform 64..1024, 10..400 paths)
    as one batch, cut into bounded frames: each canvas's fp32 pixels and sRGB8
    bytes bit for bit equal to the oracle's."""
    streams = [cia.scene(4, cia.scene_seed(0xC4, i)) for i in range(n)]
    b = cia.Batch(streams, device)
    print("batch-mix %d: %s" % (n, b.stats()), flush=True)
    b.execute(to_host=True)
    for i, s in enumerate(streams):
        r8, rl = oracle.render(s, "port")
        assert np.array_equal(rl.view(np.uint32), b.linear(i).view(np.uint32)), i
        assert np.array_equal(r8, b.rgba8(i)), i
    b.close()


def serve_burst(device):
    """A closed-loop burst of serving-mix requests: the config-1 mix with
    shadows on every third and the effects mix (composite operators, shadows,
    draw_image), some streams sent twice, all submitted at once.  Admitted up
    to 64 per batch within the frame bounds (which the run may lower with
    EMU_BATCH_DRAWS / EMU_BATCH_PIXELS), with at most 3 per batch, and three
    larger requests under an HBM budget that holds two of them, so admission
    waits for memory.  Every result equals the request rendered alone."""
    sizes = [(40, 32), (64, 48), (32, 56), (72, 40), (48, 48), (56, 64)]
    streams = []
    for i in range(8):
        w, h = sizes[i % len(sizes)]
        seed = cia.scene_seed(0xB5, i % 6)
        if i % 4 == 1:
            streams.append(cia.scene(3, seed, w, h, 5))  # one draw_image per 5 paths
        else:
            flags = cia.SCENE_SHADOWS if i % 3 == 0 else 0
            streams.append(cia.scene(5, seed, w, h, 4 + 2 * (i % 4), flags))
    large = [cia.scene(1, cia.scene_seed(0xB6, i), 256, 160, 4, 0) for i in range(3)]
    alone = {id(s): cia.render(s, device)[0] for s in streams + large}
    two = 2 * 16 * 256 * 160 + 4096
    for batch, budget, max_batch in ((streams, 1 << 26, 64), (streams, 1 << 26, 3),
                                     (large, two, 64)):
        server = cia.Server(hbm_budget=budget, max_batch=max_batch, device=device)
        ids = [server.submit(s) for s in batch]
        got = {}
        while len(got) < len(ids):
            rid = server.wait(timeout_ms=600000)
            assert rid is not None, "server stalled"
            got[rid] = server.result(rid)[0]
        stats = server.stats()
        server.stop()
        assert stats["served"] == len(batch), stats
        if batch is large:
            assert stats["batches"] >= 2 and stats["max_in_flight"] <= 2, stats
        for s, rid in zip(batch, ids):
            assert np.array_equal(alone[id(s)], got[rid]), (budget, max_batch, rid)


def round_joins(device):
    """Strokes with round joins (scenes.round_join_strokes): the kernels'
    polygons equal the oracle's stroked lines bit for bit, and their fp32
    pixels the oracle's render (the join's acos and tan as glibc computes
    them, which the oracle and the reference call)."""
    for name, (one, before) in scenes.round_join_strokes().items():
        w, h = cia.stream_size(one)
        ref = oracle.Canvas(w, h, "port")
        ref.replay(before)
        rp, rs = ref.path_lines(True)
        c = cia.Canvas(w, h, device)
        c.replay(one)
        c.flush()
        gp, gs = c.trace_polygons(0)
        c.close()
        assert np.array_equal(rs, gs.astype(np.int32)), name
        assert np.array_equal(rp.view(np.uint32), gp.view(np.uint32)), name
        _, pl = oracle.render(one, "port")
        _, gl = cia.render(one, device)
        assert np.array_equal(pl.view(np.uint32), gl.view(np.uint32)), name


def bit_exact_scenes(device):
    """Seeded scenes of every configuration render with fp32 pixels and
    sRGB8 bytes bit for bit equal to the oracle's -- tighter than the parity
    bar: the kernels follow the reference's float operations, its glibc
    acosf/tanf in round joins and its read-back bytes, and sum equal-magnitude
    runs in the oracle's order (where the reference's std::sort may pick
    another, the oracle and the reference differ in the last bit)."""
    for cfg, seed, w, h, n, flags in scenes.BIT_EXACT_SCENES:
        s = cia.scene(cfg, seed, w, h, n, flags)
        p8, pl = oracle.render(s, "port")
        g8, gl = cia.render(s, device)
        assert np.array_equal(pl.view(np.uint32), gl.view(np.uint32)), (cfg, hex(seed))
        assert np.array_equal(p8, g8), (cfg, hex(seed))


def sweep(device, n):
    """n seeded scenes (configurations 1-5 in turn; 48-160 x 48-128 pixels,
    8-40 shapes; the shadowed serving variant on 1 in 10 of configuration 5),
    each with fp32 pixels and sRGB8 bytes bit for bit equal to the oracle's;
    a line per scene also compares with the reference build where it exists."""
    rng = np.random.default_rng(7)
    kind = "reference" if oracle.available("reference") else None
    for i in range(n):
        cfg = 1 + i % 5
        seed = int(rng.integers(1, 1 << 40))
        w, h = int(rng.integers(48, 160)), int(rng.integers(48, 128))
        shapes = int(rng.integers(8, 40))
        flags = cia.SCENE_SHADOWS if (cfg == 5 and i % 10 == 4) else 0
        s = cia.scene(cfg, seed, w, h, shapes, flags)
        p8, pl = oracle.render(s, "port")
        g8, gl = cia.render(s, device)
        line = "scene %d: configuration %d seed %#x %dx%d, %d shapes" % (i, cfg, seed, w, h,
                                                                         shapes)
        if kind:
            r8, rl = oracle.render(s, kind)
            line += ", reference: %d pixels differ" % int(np.count_nonzero(
                (gl.view(np.uint32) != rl.view(np.uint32)).any(axis=2)))
        print(line, flush=True)
        assert np.array_equal(pl.view(np.uint32), gl.view(np.uint32)), line
        assert np.array_equal(p8, g8), line


def full_config(device, cfg):
    """A benchmark configuration at full size (seed 0xC0 + cfg): fp32 pixels
    and sRGB8 bytes bit for bit equal to the oracle's, and the reference
    build's where it exists (configuration 1 takes minutes, 3 hours)."""
    s = cia.scene(cfg, 0xC0 + cfg)
    p8, pl = oracle.render(s, "port")
    g8, gl = cia.render(s, device)
    differ = int(np.count_nonzero((gl.view(np.uint32) != pl.view(np.uint32)).any(axis=2)))
    print("configuration %d: %d pixels differ from the oracle" % (cfg, differ), flush=True)
    if oracle.available("reference"):
        r8, rl = oracle.render(s, "reference")
        print("configuration %d: %d pixels differ from the reference" % (cfg, int(np.count_nonzero(
            (gl.view(np.uint32) != rl.view(np.uint32)).any(axis=2)))), flush=True)
    assert differ == 0 and np.array_equal(p8, g8)


def trace_stages(device):
    """cih_trace_stage reports only the stages a draw has: an undashed
    stroke's dash stage is empty, a dashed one's is not, and a rectangle
    (a polyline draw) has its own corners as its flatten stage."""
    from canvas_ity_amd.stream import Stream
    s = Stream(40, 30)
    s.fill_rectangle(3.5, 4.0, 20.0, 10.0)
    s.begin_path().move_to(2, 20).line_to(30, 25).line_to(35, 5)
    s.stroke()
    s.set_line_dash([4.0, 3.0])
    s.stroke()
    c = cia.Canvas(40, 30, device)
    c.replay(s.tobytes())
    c.flush()
    pts, subs = c.trace_stage(0, 0)
    assert subs.shape[0] == 1 and int(subs[0, 0]) == len(pts), (pts, subs)
    corners = {(3.5, 4.0), (23.5, 4.0), (23.5, 14.0), (3.5, 14.0)}
    assert {tuple(map(float, p)) for p in pts} == corners, pts
    assert c.trace_stage(0, 1)[0].shape[0] == 0
    assert c.trace_stage(1, 1)[0].shape[0] == 0 and c.trace_stage(1, 1)[1].shape[0] == 0
    dashed, dsubs = c.trace_stage(2, 1)
    assert dashed.shape[0] > 0 and dsubs.shape[0] > 1, (dashed.shape, dsubs.shape)
    c.close()


MODES = {"batch_and_server": batch_and_server, "round_joins": round_joins,
         "trace_stages": trace_stages, "bit_exact_scenes": bit_exact_scenes, "batch_small": batch_small, "serve_burst": serve_burst,
         "batch_frames": batch_frames, "batch_frames_render": batch_frames_render,
         "payload_limit": payload_limit,
         "fine_configs": fine_configs, "canvas_execute": canvas_execute,
         "reference_suite_batch": reference_suite_batch}


def serve_mix(device, n):
    """n requests of the serving mix as configured (canvases log-uniform
    128..2048 per side, 50..2000 paths, shadows on about 10%) submitted to the
    server together: each result's sRGB8 bytes equal to the oracle's."""
    streams = [cia.scene(5, cia.scene_seed(0xC5, i)) for i in range(n)]
    server = cia.Server(hbm_budget=8 << 30, max_batch=64, device=device)
    ids = [server.submit(s) for s in streams]
    got = {}
    while len(got) < len(ids):
        rid = server.wait(timeout_ms=3600000)
        assert rid is not None, "server stalled"
        got[rid] = server.result(rid)[0]
    print("serve-mix %d: %s" % (n, server.stats()), flush=True)
    server.stop()
    for i, (s, rid) in enumerate(zip(streams, ids)):
        r8, _ = oracle.render(s, "port")
        assert np.array_equal(r8, got[rid]), i


def mode(name):
    if name.startswith("batch_mix_"):
        n = int(name[len("batch_mix_"):])
        return lambda device: batch_mix(device, n)
    if name.startswith("serve_mix_"):
        n = int(name[len("serve_mix_"):])
        return lambda device: serve_mix(device, n)
    if name.startswith("sweep_"):
        n = int(name[len("sweep_"):])
        return lambda device: sweep(device, n)
    if name.startswith("config_"):
        cfg = int(name[len("config_"):])
        return lambda device: full_config(device, cfg)
    return MODES.get(name)


def main():
    cia.LIBRARY = sys.argv[1]
    all_streams = streams()
    device = cia.default_device()
    failed = False
    for name in sys.argv[2:]:
        if mode(name):
            mode(name)(device)
            print(name, "ok", flush=True)
            continue
        s = all_streams[name]
        g8, gl = cia.render(s, device)
        r8, rl = oracle.render(s, "port")
        try:
            assert_parity(r8, rl, g8, gl, name)
            print(name, "ok", flush=True)
        except AssertionError as e:
            print(name, "PARITY FAILURE", e, flush=True)
            failed = True
    return 1 if failed else 0


if __name__ == "__main__":
    sys.exit(main())
