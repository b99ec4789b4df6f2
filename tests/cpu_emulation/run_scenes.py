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
    return out


def main():
    cia.LIBRARY = sys.argv[1]
    all_streams = streams()
    device = cia.default_device()
    failed = False
    for name in sys.argv[2:]:
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
