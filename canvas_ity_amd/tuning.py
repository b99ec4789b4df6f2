"""The fine-stage autotuner's cache: one tuned configuration per batch size
class, kept in a JSON file.

A batch's fine stage has six launch shapes (cih_batch_configure in
include/cih_abi.h): work items of 256, 128 or 64 pixels of a canvas row, as
one workgroup per item or as a persistent grid draining per-XCD queues.  All
give bit-identical pixels; which is fastest depends on the canvases' widths
and draw counts.  `configure(batch, streams)` looks the batch's size class up
in the cache, and on a miss times every shape on the batch itself
(Batch.tune) and records the winner.
"""
import json
import os

import numpy as np

from canvas_ity_amd.stream import count_draws


def cache_path():
    """$CIH_TUNE_CACHE, else ~/.cache/canvas_ity_hip/fine_tuning.json."""
    path = os.environ.get("CIH_TUNE_CACHE")
    if path:
        return path
    return os.path.join(os.path.expanduser("~"), ".cache", "canvas_ity_hip",
                        "fine_tuning.json")


def _bucket(v):
    """The power of two nearest to v (at least 1)."""
    return 1 << max(0, int(round(np.log2(max(v, 1.0)))))


def size_class(sizes, draws):
    """Key of a batch: median canvas width and height and median draws per
    canvas, each rounded to a power of two, and the number of canvases."""
    w = float(np.median([s[0] for s in sizes]))
    h = float(np.median([s[1] for s in sizes]))
    d = float(np.median(draws))
    return "w%d_h%d_d%d_n%d" % (_bucket(w), _bucket(h), _bucket(d), _bucket(len(sizes)))


def load(path=None):
    path = path or cache_path()
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return {}
    return data if isinstance(data, dict) else {}


def save(entries, path=None):
    path = path or cache_path()
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(entries, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def configure(batch, streams, path=None, reps=3):
    """Applies the cached configuration of the batch's size class, or tunes
    the batch and caches the result.  Returns (config, key, tuned_now)."""
    from canvas_ity_amd import FINE_CONFIGS, stream_size
    key = size_class([stream_size(s) for s in streams], [count_draws(s) for s in streams])
    entries = load(path)
    entry = entries.get(key)
    if isinstance(entry, dict) and 0 <= entry.get("config", -1) < FINE_CONFIGS:
        batch.configure(entry["config"])
        return entry["config"], key, False
    best, ms = batch.tune(reps)
    entries[key] = {"config": best, "ms": [round(t, 4) for t in ms]}
    save(entries, path)
    return best, key, True
