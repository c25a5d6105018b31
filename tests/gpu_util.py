"""Shared helpers of the GPU parity tests."""
import glob
import os

import numpy as np

from oracle import oracle as O

GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")
GOLDEN = sorted(glob.glob(os.path.join(GOLDEN_DIR, "*.npz")))

# parity bar: relative power error, absolute T0 error (days)
POWER_RTOL = 1e-6
T0_ATOL = 1e-6


def power_rel_err(g, r):
    """max relative error over periods with a qualifying window; a zero
    oracle power must be matched within 1e-12 of the largest power."""
    fin = np.isfinite(r)
    assert np.array_equal(np.isfinite(g), fin), "qualifying-window mask"
    if not fin.any():
        return 0.0
    scale = np.maximum(np.abs(r[fin]), 1e-12 * np.max(np.abs(r[fin])) +
                       1e-300)
    return float(np.max(np.abs(g[fin] - r[fin]) / scale))


def assert_parity(gpu, ref, label=""):
    """gpu/ref: dicts of the 7 per-period fields."""
    rel = power_rel_err(gpu["power"], ref["power"])
    assert rel <= POWER_RTOL, "%s power rel err %.3e" % (label, rel)
    fin = np.isfinite(ref["power"])
    if fin.any():
        ib = int(np.argmax(ref["power"]))
        gb = int(np.argmax(gpu["power"]))
        assert gb == ib, "%s best period %d vs oracle %d" % (label, gb, ib)
        dt0 = abs(gpu["transit_time"][ib] - ref["transit_time"][ib])
        assert dt0 <= T0_ATOL, "%s T0 err %.3e" % (label, dt0)
        assert gpu["duration"][ib] == ref["duration"][ib]
    return rel


def window_agreement(gpu, ref):
    """fraction of periods whose chosen window (duration, transit time)
    equals the oracle's."""
    fin = np.isfinite(ref["power"])
    same = (gpu["duration"] == ref["duration"]) & (
        np.abs(gpu["transit_time"] - ref["transit_time"]) <= T0_ATOL)
    return float(np.mean(same[fin])) if fin.any() else 1.0
