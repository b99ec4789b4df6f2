"""Runs the GPU test files against an emulation build of the library (on the
CPU).  TEST INFRASTRUCTURE ONLY: the product loads the HIP library and fails
loudly without it; this points the package at an emulation build in this
process only, to run the GPU tests' logic and parity checks before a GPU
does.

Usage: run_gpu_tests.py LIBRARY [pytest arguments...]
       (LIBRARY: build.py --plain's output for speed, or the AddressSanitizer
       build with its runtime preloaded)
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import canvas_ity_amd as cia  # noqa: E402


def main():
    cia.LIBRARY = os.path.abspath(sys.argv[1])
    import pytest
    return pytest.main(["-q", "-p", "no:cacheprovider", "-m", "gpu"] + sys.argv[2:])


if __name__ == "__main__":
    sys.exit(main())
