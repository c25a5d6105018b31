"""ctypes binding of include/atg_bls.h (libatg_bls.so).

The library is the only compute path: if it is missing or fails to load,
every entry point raises; there is no CPU fallback.
"""
import ctypes
import os

from .build import LIB

ATG_OK = 0
ATG_ERR_ARG = -1
ATG_ERR_BUDGET = -2
ATG_ERR_HIP = -3
ATG_ERR_LDS = -4
ATG_ERR_DEVICE = -5

LC_STATUS = {
    0: "ok", 1: "invalid period", 2: "invalid duration", 3: "NaN in t or y",
    4: "dy <= 0 or NaN", 5: "period needs more bins than LDS holds",
    6: "empty light curve or grid", 7: "more than 64 durations",
    8: "more than 2^28 points in one light curve",
}

ATG_F32 = 1
ATG_F64 = 2
# static device workspace of atg_bls_run_noalloc
ATG_NOALLOC_SCRATCH_BYTES = 64 << 20
OBJECTIVES = {"likelihood": 0, "snr": 1}
MODES = {"parity": 0, "fast": 1}

ATG_STATS_FIELDS = 16
STATS_STATUS = {0: "ok", 1: "no point inside the transit box",
                2: "more transits than max_transits"}

# symbols declared in include/atg_bls.h
EXPORTS = (
    "atg_abi_version", "atg_ctx_create", "atg_ctx_destroy",
    "atg_ctx_set_stream", "atg_last_error", "atg_ctx_set_tuning",
    "atg_ctx_get_tuning", "atg_bls_workspace_size", "atg_bls_run",
    "atg_bls_run_sync", "atg_timing_main_kernel", "atg_bls_top_k",
    "atg_bls_stats", "atg_bls_model", "atg_bls_run_best",
    "atg_bls_run_noalloc", "atg_ctx_pin_workspace",
    "atg_bls_workspace_size_best",
)

_vp = ctypes.c_void_p


class Batch(ctypes.Structure):
    _fields_ = [
        ("n_lc", ctypes.c_int64), ("n_points", ctypes.c_int64),
        ("n_periods", ctypes.c_int64), ("n_durations", ctypes.c_int64),
        ("t", _vp), ("y", _vp), ("dy", _vp),
        ("y_dtype", ctypes.c_int32), ("dy_dtype", ctypes.c_int32),
        ("lc_offsets", _vp), ("periods", _vp), ("period_offsets", _vp),
        ("durations", _vp), ("dur_offsets", _vp),
        ("oversample", ctypes.c_int32), ("objective", ctypes.c_int32),
        ("mode", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


class Out(ctypes.Structure):
    _fields_ = [
        ("power", _vp), ("depth", _vp), ("depth_err", _vp),
        ("duration", _vp), ("transit_time", _vp), ("depth_snr", _vp),
        ("log_likelihood", _vp), ("best_index", _vp), ("status", _vp),
    ]


class Best(ctypes.Structure):
    _fields_ = [
        ("power", _vp), ("depth", _vp), ("depth_err", _vp),
        ("duration", _vp), ("transit_time", _vp), ("depth_snr", _vp),
        ("log_likelihood", _vp), ("period", _vp), ("best_index", _vp),
        ("status", _vp),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("periods_per_item", ctypes.c_int32), ("lds_bins", ctypes.c_int32),
        ("wg_per_cu", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


_lib = None


def lib():
    """Load libatg_bls.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise RuntimeError(
            "libatg_bls.so is not built (%s); run "
            "`python -m astro_transit_gpu_amd.build`" % LIB)
    L = ctypes.CDLL(LIB)
    L.atg_abi_version.restype = ctypes.c_int
    L.atg_abi_version.argtypes = []
    L.atg_ctx_create.restype = ctypes.c_int
    L.atg_ctx_create.argtypes = [ctypes.c_int, _vp, ctypes.c_uint64,
                                 ctypes.POINTER(_vp)]
    L.atg_ctx_destroy.restype = None
    L.atg_ctx_destroy.argtypes = [_vp]
    L.atg_ctx_pin_workspace.restype = ctypes.c_int
    L.atg_ctx_pin_workspace.argtypes = [_vp, ctypes.c_int]
    L.atg_ctx_set_stream.restype = ctypes.c_int
    L.atg_ctx_set_stream.argtypes = [_vp, _vp]
    L.atg_last_error.restype = ctypes.c_char_p
    L.atg_last_error.argtypes = [_vp]
    L.atg_ctx_set_tuning.restype = ctypes.c_int
    L.atg_ctx_set_tuning.argtypes = [_vp, ctypes.POINTER(Tuning)]
    L.atg_ctx_get_tuning.restype = ctypes.c_int
    L.atg_ctx_get_tuning.argtypes = [_vp, ctypes.POINTER(Tuning)]
    L.atg_bls_workspace_size.restype = ctypes.c_int
    L.atg_bls_workspace_size.argtypes = [_vp, ctypes.POINTER(Batch),
                                         ctypes.POINTER(ctypes.c_uint64)]
    for fn in (L.atg_bls_run, L.atg_bls_run_sync, L.atg_bls_run_noalloc):
        fn.restype = ctypes.c_int
        fn.argtypes = [_vp, ctypes.POINTER(Batch), ctypes.POINTER(Out)]
    L.atg_bls_run_best.restype = ctypes.c_int
    L.atg_bls_run_best.argtypes = [_vp, ctypes.POINTER(Batch),
                                   ctypes.POINTER(Best)]
    L.atg_bls_workspace_size_best.restype = ctypes.c_int
    L.atg_bls_workspace_size_best.argtypes = [
        _vp, ctypes.POINTER(Batch), ctypes.POINTER(ctypes.c_uint64)]
    L.atg_bls_top_k.restype = ctypes.c_int
    L.atg_bls_top_k.argtypes = [_vp, ctypes.POINTER(Batch),
                                ctypes.POINTER(Out), ctypes.c_int,
                                ctypes.c_int, _vp, _vp]
    L.atg_bls_stats.restype = ctypes.c_int
    L.atg_bls_stats.argtypes = [_vp, ctypes.POINTER(Batch), ctypes.c_int64,
                                _vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp,
                                _vp, _vp]
    L.atg_bls_model.restype = ctypes.c_int
    L.atg_bls_model.argtypes = [_vp, ctypes.c_int64, _vp, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double, _vp, _vp]
    L.atg_timing_main_kernel.restype = ctypes.c_int
    L.atg_timing_main_kernel.argtypes = [_vp, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double)]
    if L.atg_abi_version() != 1:
        raise RuntimeError("libatg_bls.so ABI version mismatch")
    _lib = L
    return L


class AtgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("atg error %d: %s" % (code, msg))
        self.code = code


class MemoryBudgetError(AtgError, MemoryError):
    pass
