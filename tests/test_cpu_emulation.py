"""The kernel source run on the CPU under AddressSanitizer.

tests/cpu_emulation builds the unmodified kernels (cih_kernels.hip) and
runtime against an emulated HIP runtime -- each workgroup's lanes as
coroutines meeting at every wave intrinsic and barrier, device buffers as
exactly-sized host allocations, each checked frame run again with arenas
exactly as large as it asked for -- so that any kernel access outside an
allocation aborts with AddressSanitizer's report, and the emulated renders
must meet the parity bar against the oracle.  This is a CPU check of the GPU
code's memory safety and arithmetic; the GPU tests are the parity tests of
the product itself.  Skipped when g++ with AddressSanitizer is unavailable.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "cpu_emulation"))

import build as emulation  # noqa: E402

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not available")


@pytest.fixture(scope="module")
def library():
    try:
        emulation.asan_runtime()
    except (OSError, subprocess.CalledProcessError):
        pytest.skip("AddressSanitizer runtime not found")
    return emulation.build()


STRICT = {"EMU_POISON": "1", "EMU_REDZONES": "1", "EMU_STRICT_STREAMS": "1"}


def run(library, names, shuffle=None, strict=False, threads=None, extra=None,
        sanitizer="address"):
    env = dict(os.environ)
    env["LD_PRELOAD"] = emulation.sanitizer_runtime(sanitizer)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["TSAN_OPTIONS"] = "halt_on_error=1:report_signal_unsafe=0"
    for k in ["EMU_SHUFFLE", "EMU_THREADS"] + list(STRICT):
        env.pop(k, None)
    if threads is not None:
        env["EMU_THREADS"] = str(threads)
    if shuffle is not None:
        env["EMU_SHUFFLE"] = shuffle
    if strict:
        env.update(STRICT)
    for k in ("EMU_PAYLOAD_LIMIT", "EMU_SMALL_ARENAS", "EMU_TRACE_ARENAS", "EMU_ASYNC_STREAMS",
              "EMU_BATCH_DRAWS", "EMU_BATCH_PIXELS"):
        env.pop(k, None)
    env.update(extra or {})
    p = subprocess.run([sys.executable, os.path.join(HERE, "cpu_emulation", "run_scenes.py"),
                        library] + names, env=env, capture_output=True, text=True,
                       timeout=900)
    report = (p.stdout + p.stderr)[-6000:]
    assert p.returncode == 0, report
    for n in names:
        assert "%s ok" % n in p.stdout, report
    return p.stdout + p.stderr


# Every test's emulation runs start together when the module's first test
# asks for them, JOBS at a time, longest first; each test then checks its own.
JOBS = 4
RUNS = [
    # the batch's fine-stage launch shapes (64/128/256-pixel items, one
    # workgroup per item or a persistent grid draining per-XCD queues) give
    # the same bits, with workgroups on 8 host threads and the strict checks
    ("fine_configs", dict(names=["fine_configs"], strict=True, threads=8)),
    # every stream's work queued until the host waits for it (EMU_ASYNC_STREAMS:
    # a result read before its wait, or pinned staging rewritten before its
    # copy ran, gives wrong pixels), with the strict checks, on 4 threads: a
    # closed-loop serving burst under lowered frame bounds and a tight HBM
    # budget, a multi-canvas batch, puts in batches and server requests, and
    # the re-run path of a flushed canvas
    ("deferred", dict(names=["serve_burst", "batch_small", "batch_and_server", "canvas_execute"],
                      strict=True, threads=4,
                      extra={"EMU_ASYNC_STREAMS": "1", "EMU_BATCH_DRAWS": "40",
                             "EMU_BATCH_PIXELS": "6000"})),
    # with the frame bounds lowered, a batch of many frames, some across a
    # canvas, renders every canvas as it renders alone
    ("frames_render", dict(names=["batch_frames_render"], threads=4,
                           extra={"EMU_BATCH_DRAWS": "50", "EMU_BATCH_PIXELS": "10000"})),
    # workgroups one at a time, in index order; and the benchmark's path, a
    # flushed canvas run again from HBM
    ("configurations", dict(names=["plain_small", "large_small", "effects_small",
                                   "canvas_execute"])),
    # first arenas of 1024 entries: every frame overflows, and the host grows
    # the arenas it ran out of (scaled by the share of draws that got through)
    # and runs the frame again, for checked frames and for server batches
    ("arenas", dict(names=["large_small", "effects_small", "clip_nested", "batch_small",
                           "batch_and_server"], threads=8,
                    extra={"EMU_SMALL_ARENAS": "1", "EMU_TRACE_ARENAS": "1"})),
    # device buffers and LDS start as 0xA5 garbage, every arena sub-allocation
    # is followed by a poisoned red zone (a draw overrunning into the next
    # draw's region aborts), NULL-stream uploads are held back until the NULL
    # stream is synchronised (the library's streams do not wait for it), and a
    # kernel's workgroups run on 8 host threads at once (atomics are host
    # atomics, LDS is per thread).  (Full-size batch-mix batches:
    # run_scenes.py batch_mix_<n> with the same environment.)
    ("strict", dict(names=["large_small", "effects_small", "clip_nested", "text_scene_1",
                           "batch_small", "batch_and_server"], strict=True, threads=8)),
    # a GPU runs a kernel's workgroups concurrently, in no fixed order; here
    # they are handed out in a seeded random order (EMU_SHUFFLE) to 8 host
    # threads
    ("random_order", dict(names=["large_small", "effects_small", "clip_nested", "text_scene_1",
                                 "batch_small"], shuffle="0xC4", threads=8)),
    # the batch regime that faulted the device (tens of thousands of draws
    # over many canvases in one frame) is cut into frames within the
    # multi-canvas bounds; above 2^30 pixels a batch is refused
    ("frames", dict(names=["batch_frames"])),
    # strokes with round joins: polygons and fp32 pixels bit for bit equal
    # to the oracle's (the join's acos and tan evaluated as glibc does)
    ("round_joins", dict(names=["round_joins", "trace_stages"], threads=4)),
    # two seeded scenes of each configuration bit for bit equal to the oracle
    # in fp32 and sRGB8, tighter than the parity bar
    ("bit_exact", dict(names=["bit_exact_scenes"], threads=4)),
    # the host runtime's threads (server workers and scheduler, batch replay
    # threads, the caller) under ThreadSanitizer, the kernels and the emulated
    # HIP runtime uninstrumented: a closed-loop serving burst under lowered
    # frame bounds, a multi-canvas batch, puts in batches and server requests,
    # and the canvas re-run path report no data race
    ("host_races", dict(names=["serve_burst", "batch_small", "batch_and_server",
                               "canvas_execute"], threads=4, sanitizer="thread",
                        extra={"EMU_BATCH_DRAWS": "40", "EMU_BATCH_PIXELS": "6000"})),
    # puts recorded as packets in batch and server canvases, applied in
    # painter's order: the same pixels as the plain canvas, which applies them
    # at once
    ("put", dict(names=["batch_and_server"], threads=4)),
    # draw records address packets with 32-bit offsets; the emulation lowers
    # the 4 GiB limit to 20000 bytes to check the refusals and the server's
    # admission without gigabytes of packets
    ("payload", dict(names=["payload_limit"], extra={"EMU_PAYLOAD_LIMIT": "20000"})),
    ("text", dict(names=["text_font", "text_fill_text", "text_stroke_text", "text_scene_1"],
                  threads=4)),
    ("clip", dict(names=["clip_nested", "clip_op_destination_in", "clip_shadow_image",
                         "clip_star_empty", "reference_global_composite_operation"],
                  threads=4)),
]


@pytest.fixture(scope="module")
def runs(library):
    libraries = {"address": library}
    try:
        emulation.sanitizer_runtime("thread")
        libraries["thread"] = emulation.build(sanitizer="thread")
    except (OSError, subprocess.CalledProcessError, RuntimeError):
        pass  # the ThreadSanitizer run is skipped
    pool = ThreadPoolExecutor(max_workers=JOBS)
    futures = {key: pool.submit(run, libraries[kw.get("sanitizer", "address")], **kw)
               for key, kw in RUNS if kw.get("sanitizer", "address") in libraries}
    yield futures
    pool.shutdown(wait=False, cancel_futures=True)


def test_configurations_memory_safe_and_in_parity(runs):
    runs["configurations"].result()


def test_clip_masks_memory_safe_and_in_parity(runs):
    runs["clip"].result()


def test_text_memory_safe_and_in_parity(runs):
    runs["text"].result()


def test_put_image_data_in_batches_and_server(runs):
    runs["put"].result()


def test_workgroups_in_random_order(runs):
    runs["random_order"].result()


def test_payloads_above_the_offset_range_are_refused(runs):
    runs["payload"].result()


def test_batches_are_cut_into_bounded_frames(runs):
    runs["frames"].result()
    runs["frames_render"].result()


def test_strict_checks_with_concurrent_workgroups(runs):
    runs["strict"].result()


def test_fine_configurations_are_bit_identical(runs):
    runs["fine_configs"].result()


def test_arena_growth_after_overflow(runs):
    out = runs["arenas"].result()
    attempts = [line for line in out.splitlines() if line.startswith("[arenas]")]
    assert sum(" overflow 1" in line for line in attempts) >= 3, out[-3000:]


def test_deferred_streams_and_a_serving_burst(runs):
    runs["deferred"].result()


def test_round_joins_bit_for_bit(runs):
    runs["round_joins"].result()


def test_host_threads_free_of_data_races(runs):
    if "host_races" not in runs:
        pytest.skip("ThreadSanitizer runtime not found")
    out = runs["host_races"].result()
    assert "ThreadSanitizer" not in out, out[-3000:]


def test_scenes_bit_for_bit(runs):
    runs["bit_exact"].result()
