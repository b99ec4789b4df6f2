"""Parity measures shared by the tests.

The bar (from the project's parity gate): sRGB8 read-back within 1 LSB per
channel on every pixel with at least 99.9% of pixels identical, and the fp32
linear premultiplied buffer within 1e-5 absolute of the oracle.
"""
import numpy as np

SRGB_MAX_LSB = 1
SRGB_IDENTICAL = 0.999
LINEAR_ATOL = 1e-5


def measure(ref8, refl, got8, gotl):
    d8 = np.abs(ref8.astype(np.int32) - got8.astype(np.int32))
    identical = float(np.all(d8 == 0, axis=-1).mean())
    dl = float(np.abs(refl.astype(np.float64) - gotl.astype(np.float64)).max()) if refl.size else 0.0
    return {"srgb_max": int(d8.max()) if d8.size else 0, "identical": identical,
            "linear_max": dl}


def assert_parity(ref8, refl, got8, gotl, what=""):
    m = measure(ref8, refl, got8, gotl)
    assert m["srgb_max"] <= SRGB_MAX_LSB, (what, m)
    assert m["identical"] >= SRGB_IDENTICAL, (what, m)
    assert m["linear_max"] <= LINEAR_ATOL, (what, m)
    return m
