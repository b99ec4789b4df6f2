"""The kernels' libm-dependent arithmetic against the host libm.

The reference computes a round join's cubic from acosf and tanf
(canvas_ity.hpp:1994-1997), so the join's points depend on their last bits.
canvas_ity_amd/csrc/cih_geometry.hpp evaluates both as glibc does (the float
versions of fdlibm's e_acos.c and k_tan.c), the same code on host and device.
Read-back encodes sRGB with powf(v, 1/2.4) (hpp:1282-1284);
canvas_ity_amd/csrc/cih_color.hpp evaluates it in double, rounded once, except
on the four inputs where glibc's rounding changes a dithered byte.

This compiles both headers on the host with the library's floating-point
flags and compares join_acos with acosf on [-1, 1] and join_tan with tanf on
[0, float(pi)/4] bit for bit, and the read-back byte of every dither offset
with the reference's expression on [0.0031308, 1]: every 101st input plus
every branch edge and the known special inputs, or every input with
CIH_EXHAUSTIVE_JOIN_MATH=1 (about a minute).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "canvas_ity_amd", "csrc")

PROGRAM = r"""
#include "cih_color.hpp"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

using cih::bits_float;
using cih::float_bits;

static long checked = 0, wrong = 0;

static void acos_at(uint32_t u) {
    float x = bits_float(u);
    float want = acosf(x), got = cih::join_acos(x);
    ++checked;
    if (float_bits(want) != float_bits(got) && !(isnan(want) && isnan(got))) {
        if (wrong++ < 5)
            printf("acos(%a): libm %a, join_acos %a\n", x, want, got);
    }
}

static void tan_at(uint32_t u) {
    float x = bits_float(u);
    float want = tanf(x), got = cih::join_tan(x);
    ++checked;
    if (float_bits(want) != float_bits(got)) {
        if (wrong++ < 5)
            printf("tan(%a): libm %a, join_tan %a\n", x, want, got);
    }
}

// the reference's read-back byte (hpp:1282-1284, 3373-3379) against the
// kernels' for each of the 16 ordered-dither offsets (hpp:3358-3362)
static void srgb_at(uint32_t u) {
    float x = bits_float(u);
    float want = 1.055f * powf(x, 1.0f / 2.4f) - 0.055f, got = cih::delinearized(x);
    ++checked;
    for (int k = 0; k < 16; ++k) {
        float t = ((float)k + 0.5f) / 16.0f;
        if ((int)(t + want * 255.0f) != (int)(t + got * 255.0f)) {
            if (wrong++ < 5)
                printf("srgb(%a) dither %d: libm %a, kernels %a\n", x, k, want, got);
            return;
        }
    }
}

int main(int argc, char **argv) {
    uint32_t stride = (uint32_t)strtoul(argv[1], 0, 0);
    for (uint64_t u = float_bits(0.0031308f); u <= 0x3f800000u; u += stride)
        srgb_at((uint32_t)u);
    const uint32_t srgb_special[] = {0x3c8979b7u, 0x3e17c58du, 0x3e48b85au, 0x3e7ec3d6u,
                                     0x3c8979b6u, 0x3c8979b8u};
    for (uint32_t u : srgb_special)
        srgb_at(u);
    for (uint64_t u = 0; u <= 0x3f800000u; u += stride) {
        acos_at((uint32_t)u);
        acos_at((uint32_t)u | 0x80000000u);
    }
    for (uint64_t u = 0; u <= 0x3f490fdbu; u += stride)
        tan_at((uint32_t)u);
    // branch edges of both, on either side
    const uint32_t acos_edges[] = {0x32800000u, 0x3f000000u, 0x3f800000u, 0x00000001u};
    for (uint32_t e : acos_edges)
        for (uint32_t d = 0; d < 3; ++d)
            for (uint32_t sign = 0; sign < 2; ++sign) {
                acos_at((e - 1 + d) | (sign << 31));
            }
    const uint32_t tan_edges[] = {0x39000000u, 0x3f2ca140u, 0x3f490fdau, 0x00000001u};
    for (uint32_t e : tan_edges)
        for (uint32_t d = 0; d < 3; ++d)
            tan_at(e - 1 + d);
    tan_at(0x3f490fdbu);
    acos_at(0x7fc00000u);
    // inputs where glibc's result is not the correctly rounded one (an acos
    // or tan evaluated in double and rounded once fails these)
    acos_at(0x3288859eu); // 0x1.110b3cp-26
    tan_at(0x3aca1f9au);  // 0x1.943f34p-10
    printf("checked %ld, wrong %ld\n", checked, wrong);
    return wrong != 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not available")
def test_join_math_and_srgb_encoding_match_the_host_libm(tmp_path):
    src = tmp_path / "join_math.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "join_math"
    # -fno-builtin: the libm's own acosf/tanf, never the compiler's folding
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math",
                    "-fno-builtin", "-I" + CSRC, str(src), "-o", str(exe), "-lm"], check=True)
    stride = "1" if os.environ.get("CIH_EXHAUSTIVE_JOIN_MATH") else "101"
    p = subprocess.run([str(exe), stride], capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stdout[-2000:]
    assert "wrong 0" in p.stdout
