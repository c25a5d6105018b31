"""The C-ABI library loads and exports every symbol include/atg_bls.h
declares; host-side validation (CPU only, no compute calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    from astro_transit_gpu_amd.build import LIB, build_library
    if not os.path.exists(LIB):
        build_library()
    return ctypes.CDLL(LIB)


def header_symbols():
    src = open(os.path.join(ROOT, "include", "atg_bls.h")).read()
    return sorted(set(re.findall(r"\b(atg_[a-z_0-9]+)\s*\(", src)))


def test_header_symbols_exported(lib):
    from astro_transit_gpu_amd import _abi
    syms = header_symbols()
    assert set(syms) == set(_abi.EXPORTS), syms
    for s in syms:
        assert hasattr(lib, s), s


def test_abi_version(lib):
    assert lib.atg_abi_version() == 1


def test_curve_status_codes_match_header():
    """Every per-curve status the header defines has its message in the
    Python table the API raises from, and no other code does."""
    from astro_transit_gpu_amd import _abi
    src = open(os.path.join(ROOT, "include", "atg_bls.h")).read()
    codes = {int(v) for v in re.findall(r"#define ATG_LC_[A-Z_]+ (\d+)", src)}
    assert codes == set(_abi.LC_STATUS), (codes, sorted(_abi.LC_STATUS))


def test_struct_layout_matches_header():
    """ctypes structs have the C sizes (8-byte aligned pointers)."""
    from astro_transit_gpu_amd import _abi
    assert ctypes.sizeof(_abi.Batch) == 4 * 8 + 3 * 8 + 2 * 4 + 5 * 8 + 4 * 4
    assert ctypes.sizeof(_abi.Out) == 9 * 8
    assert ctypes.sizeof(_abi.Best) == 10 * 8
    assert ctypes.sizeof(_abi.Tuning) == 16


def test_ctx_create_without_gpu_fails_cleanly(lib):
    """No GPU in this container: context creation returns an error code
    instead of crashing."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    h = ctypes.c_void_p()
    lib.atg_ctx_create.argtypes = [ctypes.c_int, ctypes.c_void_p,
                                   ctypes.c_uint64,
                                   ctypes.POINTER(ctypes.c_void_p)]
    rc = lib.atg_ctx_create(0, None, 0, ctypes.byref(h))
    assert rc != 0 and not h.value


def test_null_arguments_rejected(lib):
    lib.atg_bls_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_void_p]
    assert lib.atg_bls_run(None, None, None) == -1
    lib.atg_bls_run_best.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p]
    assert lib.atg_bls_run_best(None, None, None) == -1
    lib.atg_last_error.restype = ctypes.c_char_p
    lib.atg_last_error.argtypes = [ctypes.c_void_p]
    assert lib.atg_last_error(None) == b"null context"


@pytest.mark.parametrize("kw,msg", [
    (dict(periods=[-1.0, 2.0]), "All periods must be positive."),
    (dict(durations=[0.0]), "All durations must be positive."),
    (dict(y_nan=True), "Input time or flux contains NaNs."),
    (dict(dy=-1.0), "flux_err must be strictly positive for weighted BLS."),
    (dict(dy_len=True), "flux_err and flux must have the same length."),
    (dict(durations=[0.5, 3.0]), "The maximum transit duration must be "
     "shorter than the minimum period"),
])
def test_reference_validation_messages(kw, msg):
    """power() raises the reference's ValueError messages (api.py:54-70)
    before touching the GPU."""
    from astro_transit_gpu_amd import BoxLeastSquaresGPU
    t = np.linspace(0, 10, 100)
    y = np.ones(100)
    dy = np.full(100, 1e-3)
    if kw.get("y_nan"):
        y[3] = np.nan
    if "dy" in kw:
        dy[:] = kw["dy"]
    if kw.get("dy_len"):
        dy = dy[:50]
    periods = kw.get("periods", [1.0, 2.0])
    durations = kw.get("durations", [0.1])
    with pytest.raises(ValueError, match=re.escape(msg)):
        BoxLeastSquaresGPU(t, y, dy).power(periods, durations,
                                           method="parity")
