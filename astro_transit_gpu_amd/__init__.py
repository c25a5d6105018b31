"""MI355X-native Box Least Squares transit search.

Drop-in for the reference's ``BoxLeastSquaresGPU(t, y, dy).power(periods,
durations, method="fast"|"parity")`` (src/astrotransit_gpu/__init__.py:1-5)
backed by hand-written gfx950 HIP kernels behind a C ABI
(include/atg_bls.h, libatg_bls.so).
"""
from .api import BLSResult, BoxLeastSquaresGPU, Candidate, get_top_k_candidates
from .engine import (Context, DeviceBatch, DeviceBest, DeviceResult,
                     default_context)

__version__ = "0.1.0"
__all__ = ["BoxLeastSquaresGPU", "BLSResult", "Candidate",
           "get_top_k_candidates", "Context", "DeviceBatch", "DeviceBest",
           "DeviceResult", "default_context"]
