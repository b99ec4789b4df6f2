"""Builds libcanvas_ity_hip.so in-tree for gfx950 (MI355X).

The kernels and the HIP runtime side are compiled with hipcc; host-only
translation units (the canvas header's state machine behind the replay entry
point, and the scene generator) with g++.  Every unit is built with
-ffp-contract=off and no fast-math: flattening and stroking decisions are
float threshold tests that a contracted multiply-add could flip.  Packed FP32
instructions are disabled for all device code.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIBRARY = os.path.join(HERE, "libcanvas_ity_hip.so")
ARCH = os.environ.get("CIH_OFFLOAD_ARCH", "gfx950")


def _hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build")
    return path


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError("build step failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    common = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
              "-fvisibility=hidden", "-I" + INCLUDE, "-I" + CSRC]
    device = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
              "-packed-fp32-ops"]
    objects = []
    for name in ("cih_kernels.hip", "cih_runtime.cpp"):
        obj = os.path.join(BUILD, name + ".o")
        lang = ["-x", "hip"] if name.endswith(".hip") else ["-D__HIP_PLATFORM_AMD__"]
        _run([hipcc] + lang + common + (device if name.endswith(".hip") else []) +
             ["-c", os.path.join(CSRC, name), "-o", obj])
        objects.append(obj)
    for name in ("cih_replay.cpp", "cih_scene.cpp", "cih_text.cpp"):
        obj = os.path.join(BUILD, name + ".o")
        _run(["g++", "-O2", "-fPIC", "-std=c++11", "-ffp-contract=off", "-fno-fast-math",
              "-fvisibility=hidden", "-fvisibility-inlines-hidden", "-Wall", "-Wextra",
              "-I" + INCLUDE, "-I" + CSRC, "-c", os.path.join(CSRC, name), "-o", obj])
        objects.append(obj)
    tmp = LIBRARY + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objects +
         ["-Wl,--no-undefined"])
    os.replace(tmp, LIBRARY)
    if verbose:
        print("built", LIBRARY)
    return LIBRARY


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
