"""Benchmark of the rasterise-and-composite hot path on one MI355X.

Workload (one "step"): the large-2048 configuration -- a 2048x2048 canvas with
5000 seeded shapes (60% fills / 40% strokes; cubic and quadratic paths, arcs,
rectangles; solid, linear and radial brushes with 2-8 stops; dashes; all
joins and caps; global alpha) -- rendered from scratch by the plain
per-canvas GPU path: clear, flatten/dash/stroke, scan conversion, shading and
compositing of every draw, and the sRGB8 conversion of the whole canvas.  The
draw packets are resident in HBM before the timed region (the host-side
command replay and the PCIe transfers are outside it); value = canvas pixels
per second.

The JSON line also carries the roofline of the dominant kernel (its time
from HIP events recorded by the library on its own stream during the timed
steps), an end-to-end figure with the PCIe transfers, a CPU baseline (the
reference canvas_ity compiled from source, oracle/_ref, timed on this host)
and the batch mix: 10 canvases rendered as one batch against the same
canvases one at a time.
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "Mpix/s + canvases/s (batched vs plain GPU), p99 latency at fixed QPS, HBM GB/s"
SEED = 0xC2


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Threads for the one-instance-per-core figure: the CPUs this process may
    run on, at most 16 (the host share of one GPU on the measurement boxes)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _median_time(fn, runs):
    fn()  # warm-up
    times = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t)
    times.sort()
    return times[len(times) // 2]


def _concurrent_time(fn, threads):
    """Wall time of `threads` concurrent calls of fn (the renders run in C
    with the interpreter lock released)."""
    workers = [threading.Thread(target=fn) for _ in range(threads)]
    t = time.perf_counter()
    for w in workers:
        w.start()
    for w in workers:
        w.join()
    return time.perf_counter() - t


def cpu_baseline(cia, stream, width, height, sample_draws, threads, runs=5):
    """The CPU reference (oracle/_ref, canvas_ity compiled from its source; the
    C restatement `port` only where that build is absent, and then named so)
    on this host.  Large-2048 is sampled: one whole canvas takes 40-90 s on
    one core, so five runs plus the per-core run would not fit the bench's
    few minutes; the first `sample_draws` draws of the same seeded stream
    (identically distributed) are timed and scaled to the 5000 draws.
    Plain-512, the reference's own CPU-sized case, is timed on whole
    canvases."""
    import oracle
    from canvas_ity_amd.stream import count_draws, truncate_draws
    kind = "reference" if oracle.available("reference") else "port"
    model = cpu_model()
    total = count_draws(stream)
    sample = truncate_draws(stream, sample_draws)
    scale = total / float(sample_draws)
    single = _median_time(lambda: oracle.render(sample, kind), runs) * scale
    wall = _concurrent_time(lambda: oracle.render(sample, kind), threads) * scale
    result = {
        "value": round(width * height / 1e6 / single, 4), "unit": "Mpix/s", "cores": 1,
        "kind": kind, "cpu": model,
        "sample": "large-2048 stream %#x: first %d of %d draws, median of %d runs after a "
                  "warm-up, one thread, scaled x%.3g to the whole canvas (%.2f s per canvas); "
                  "sampled because five whole canvases and the per-core run would take "
                  "several minutes of the bench" % (SEED, sample_draws, total, runs, scale, single),
        "per_core": {"value": round(threads * width * height / 1e6 / wall, 4),
                     "unit": "Mpix/s", "cores": threads,
                     "sample": "%d concurrent instances of the same sample, one per thread, "
                               "wall time scaled x%.3g" % (threads, scale)}}
    c1 = cia.scene(1, 0xC1)
    w1, h1 = cia.stream_size(c1)
    t1 = _median_time(lambda: oracle.render(c1, kind), runs)
    tw = _concurrent_time(lambda: oracle.render(c1, kind), threads)
    result["plain_512_whole_canvas"] = {
        "canvases_per_s": round(1.0 / t1, 3), "mpix_per_s": round(w1 * h1 / 1e6 / t1, 4),
        "per_core_canvases_per_s": round(threads / tw, 3), "cores": threads,
        "sample": "plain-512 stream 0xc1, whole canvas through get_image_data, median of %d "
                  "runs after a warm-up; then %d concurrent instances" % (runs, threads)}
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=250,
                    help="draws of the large-2048 stream timed on the CPU reference (0: skip "
                         "the CPU baseline)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="instances for the per-core CPU figure (0: the CPUs this process "
                         "may use, at most 16)")
    ap.add_argument("--end-to-end", type=int, default=3, metavar="K",
                    help="canvases timed from command stream to sRGB8 in host memory "
                         "(replay, upload, render, copy back); 0: skip")
    ap.add_argument("--batch", type=int, default=10, metavar="N",
                    help="also measure the batch mix: N canvases rendered as one batch vs "
                         "one canvas at a time (canvases/s; 0: skip); the batch cuts its "
                         "draws into frames of at most 2048 draws and 2M pixels (about 10 "
                         "canvases of this mix each) and holds at most 2^30 pixels")
    ap.add_argument("--tune", action="store_true",
                    help="with --batch: pick the batch's fine-stage launch shape with the "
                         "autotuner (cached per size class in $CIH_TUNE_CACHE or "
                         "~/.cache/canvas_ity_hip/fine_tuning.json)")
    ap.add_argument("--serve", type=float, default=0.0, metavar="QPS",
                    help="also measure the serving loop: Poisson arrivals at QPS requests/s "
                         "of the serving mix, p50/p99 latency (0: skip)")
    ap.add_argument("--requests", type=int, default=10000,
                    help="requests sent in the serving measurement (p50/p99 over them)")
    ap.add_argument("--skip-main", action="store_true",
                    help="skip the large-2048 line (only with --batch / --serve)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device 3 times (calibrates the HBM "
                         "byte counters of a profiler against a copy of known size)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print("bench.py measures one GPU: canvases are independent and this build runs "
              "on one GPU; it has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import canvas_ity_amd as cia
    if args.skip_main:
        device = cia.default_device()
        line = {}
        if args.batch:
            optional_leg(line, "batch",
                         lambda: bench_batch(cia, device, args.batch, tune=args.tune))
        if args.serve:
            optional_leg(line, "serving",
                         lambda: bench_serve(cia, device, args.serve, args.requests))
        print(json.dumps(line))
        return 0
    stream = cia.scene(2, SEED)
    width, height = cia.stream_size(stream)
    device = cia.default_device()
    canvas = cia.Canvas(width, height, device)
    canvas.replay(stream)  # host state machine -> draw packets
    canvas.flush()  # packets and decoded draws now resident in HBM

    for _ in range(args.warmup):
        canvas.execute(clear=True, readback=True)
    device.synchronize()

    device.profile(True)
    device.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        canvas.execute(clear=True, readback=True)
    device.synchronize()
    t1 = time.perf_counter()
    prof = device.profile_read()
    device.profile(False)

    ms = (t1 - t0) * 1000.0 / args.steps
    mpix = width * height / 1e6 / (ms / 1000.0)
    frames = max(prof.pop("frames"), 1)
    kernels = {k: v / frames for k, v in prof.items()}  # ms per launch
    dominant = max(kernels, key=kernels.get)

    # Algorithmic bytes of the composite kernel: 32 B (fp32 RGBA read + write)
    # per pixel-draw, counted exactly by the library's statistics pass.
    stats = canvas.stats()
    pixel_draws = stats["pixel_draws"]
    line = {
        "metric": METRIC, "value": round(mpix, 3), "unit": "Mpix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded command stream %#x)" % SEED,
        "config": {"workload": "large-2048 plain path: 2048x2048 canvas, 5000 seeded shapes "
                               "(fills/strokes, beziers/arcs/rects, gradients, dashes, joins), "
                               "sRGB8 read-back", "canvas": [width, height], "draws": 5000,
                   "path": "plain per-canvas GPU",
                   "timed": "clear, every kernel of the pipeline and the sRGB8 conversion, "
                            "from draw packets resident in HBM"},
        "canvases_per_s": round(1000.0 / ms, 3),
        "kernel_ms": {k: round(v, 4) for k, v in kernels.items()},
    }
    fine_ms = kernels.get("k_fine", 0.0)
    achieved = 32.0 * pixel_draws / (fine_ms / 1000.0) / 1e9 if fine_ms > 0 else 0.0
    traffic, traffic_source = traffic_bytes()
    line["roofline"] = {"kernel": "k_fine", "dominant_kernel": dominant, "bound": "hbm",
                        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                        "traffic_source": traffic_source, "pixel_draws": pixel_draws}
    if args.end_to_end > 0:
        optional_leg(line, "end_to_end",
                     lambda: bench_end_to_end(cia, device, stream, args.end_to_end))
    if args.cpu_sample > 0:
        threads = args.cpu_threads or cpu_threads()
        optional_leg(line, "cpu_baseline",
                     lambda: cpu_baseline(cia, stream, width, height, args.cpu_sample, threads))
    canvas.close()
    if args.batch:
        optional_leg(line, "batch", lambda: bench_batch(cia, device, args.batch, tune=args.tune))
    if args.serve:
        optional_leg(line, "serving",
                     lambda: bench_serve(cia, device, args.serve, args.requests))
    print(json.dumps(line))
    return 0


def optional_leg(line, key, measure):
    """Runs one of the line's further measurements; an error is recorded in
    its place (and on stderr) so the headline line is still printed."""
    try:
        line[key] = measure()
    except Exception as e:  # noqa: BLE001 -- the line is printed either way
        print("bench.py: %s failed: %s: %s" % (key, type(e).__name__, e), file=sys.stderr)
        line[key] = {"error": "%s: %s" % (type(e).__name__, e)}


def bench_end_to_end(cia, device, stream, k):
    """Large-2048 from the command stream in host memory to sRGB8 in host
    memory, one canvas at a time: the host state machine's replay into draw
    packets, their upload and decoding, the whole pipeline and the copy back
    over PCIe (the `value` figure starts with packets resident in HBM)."""
    width, height = cia.stream_size(stream)
    times = []
    for _ in range(k + 1):
        t = time.perf_counter()
        c = cia.Canvas(width, height, device)
        c.replay(stream)
        c.rgba8()  # flushes, renders and reads back
        times.append(time.perf_counter() - t)
        c.close()
    times = sorted(times[1:])
    t = times[len(times) // 2]
    return {"ms_per_canvas": round(1000.0 * t, 3),
            "mpix_per_s": round(width * height / 1e6 / t, 3),
            "sample": "median of %d canvases after one warm-up: replay, upload, render, "
                      "sRGB8 to host" % k}


def bench_batch(cia, device, n, seed=0xC4, reps=3, tune=False):
    """Batch mix (canvases log-uniform 64..1024, 10..400 paths each): one
    batch of n canvases vs the same canvases one at a time; both from draws
    resident in HBM, sRGB8 written to device memory."""
    streams = [cia.scene(4, cia.scene_seed(seed, i)) for i in range(n)]
    pixels = sum(w * h for w, h in map(cia.stream_size, streams))
    try:
        t0 = time.perf_counter()
        b = cia.Batch(streams, device)  # replay, decode, frames, upload
        created = time.perf_counter() - t0
    except cia.CihError as e:
        return {"workload": "batch mix: %d canvases (seed %#x)" % (n, seed), "pixels": pixels,
                "refused": str(e)}
    frames = b.stats()["frames"]
    fine = {"config": cia.FINE_CONFIG_NAMES[0], "tuned": False}
    if tune:
        from canvas_ity_amd import tuning
        config, key, tuned_now = tuning.configure(b, streams)
        fine = {"config": cia.FINE_CONFIG_NAMES[config], "tuned": True, "size_class": key,
                "tuned_in_this_run": tuned_now,
                "ms_per_config": dict(zip(cia.FINE_CONFIG_NAMES,
                                          tuning.load().get(key, {}).get("ms", [])))}
    b.execute(to_host=False)  # sizes the arenas, captures the graph
    b.execute(to_host=False)
    device.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.execute(to_host=False)
    device.synchronize()
    batched = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    b.execute(to_host=True)
    with_host = time.perf_counter() - t0
    b.close()
    canvases = []
    for s in streams:
        w, h = cia.stream_size(s)
        c = cia.Canvas(w, h, device)
        c.replay(s)
        c.flush()
        canvases.append(c)
    for c in canvases:
        c.execute(clear=True, readback=True)
    device.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for c in canvases:
            c.execute(clear=True, readback=True)
    device.synchronize()
    plain = (time.perf_counter() - t0) / reps
    for c in canvases:
        c.close()
    return {"workload": "batch mix: %d canvases log-uniform 64..1024, 10..400 paths (seed %#x)"
                        % (n, seed),
            "pixels": pixels, "frames": frames,
            "batched_canvases_per_s": round(n / batched, 2),
            "plain_canvases_per_s": round(n / plain, 2), "ratio": round(plain / batched, 3),
            "batched_mpix_per_s": round(pixels / batched / 1e6, 2),
            "batched_with_srgb8_to_host_canvases_per_s": round(n / with_host, 2),
            "batched_with_transfers_canvases_per_s": round(n / (created + with_host), 2),
            "transfers": "with transfers: batch creation from the command streams in host "
                         "memory (replay, decode, upload) plus one execute with the sRGB8 "
                         "copy to host; the other figures start from draws resident in HBM",
            "fine_stage": fine}


def bench_serve(cia, device, qps, nreq, seed=0xC5, distinct=64):
    """Serving mix (canvases log-uniform 128..2048, 50..2000 paths, shadows on
    10% of requests): open-loop Poisson arrivals at `qps`; latency is submit
    to sRGB8 on the host.  `distinct` seeded requests are sent in turn."""
    import numpy as np
    streams = [cia.scene(5, cia.scene_seed(seed, i)) for i in range(distinct)]
    server = cia.Server(hbm_budget=64 << 30, max_batch=64, device=device)
    for s in streams[:8]:  # warm-up
        server.submit(s)
    for _ in range(8):
        server.result(server.wait(timeout_ms=120000))
    gaps = np.random.default_rng(seed).exponential(1.0 / qps, nreq)
    arrivals = np.cumsum(gaps)
    latencies = []
    failed = []
    finished = [0.0]

    def collect():
        for _ in range(nreq):
            rid = server.wait(timeout_ms=300000)
            if rid is None:
                break
            try:
                latencies.append(server.result(rid)[1])
            except cia.CihError as e:
                failed.append(str(e))
        finished[0] = time.perf_counter()
    collector = threading.Thread(target=collect)
    collector.start()
    t0 = time.perf_counter()
    for i in range(nreq):
        delay = t0 + arrivals[i] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        server.submit(streams[i % distinct])
    collector.join()
    stats = server.stats()
    server.stop()
    wall = finished[0] - t0
    lat = np.array(latencies)
    return {"workload": "serving mix: canvases log-uniform 128..2048, 50..2000 paths, "
                        "10%% shadowed (seed %#x, %d distinct)" % (seed, distinct),
            "offered_qps": qps, "requests": nreq, "completed": int(lat.size),
            "failed": len(failed), "first_failure": failed[0] if failed else None,
            "achieved_qps": round(lat.size / wall, 2) if wall > 0 else None,
            "p50_ms": round(float(np.percentile(lat, 50)), 3) if lat.size else None,
            "p99_ms": round(float(np.percentile(lat, 99)), 3) if lat.size else None,
            "batches": stats["batches"], "max_in_flight": stats["max_in_flight"]}


def copy_probe(mib):
    import torch
    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n, "launches": 3}))
    return 0


def traffic_bytes():
    """HBM bytes per k_fine launch and where the figure comes from: the
    committed rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE runs of
    this workload, scaled by a copy probe), which is not measured in this
    run; (None, reason) when the summary is absent."""
    rel = os.path.join("profiles", "k_fine_large2048_pmc.json")
    try:
        with open(os.path.join(ROOT, rel)) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None, "no counter summary"
    return data.get("hbm_bytes_per_launch"), (
        "%s: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this workload, recorded earlier, "
        "not in this run" % rel)


if __name__ == "__main__":
    sys.exit(main())
