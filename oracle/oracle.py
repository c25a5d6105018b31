"""CPU oracle for the BLS hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only as the checker (or the timed CPU baseline).
The product (astro_transit_gpu_amd/) never imports it.

* :func:`power` restates astropy ``BoxLeastSquares.power`` (astropy
  timeseries/periodograms/bls/core.py:234-330 input preparation,
  702-720 result assembly) around the C restatement of the compiled
  ``bls_impl`` in ``bls_oracle.c`` (built by ``oracle/Makefile``).
* :func:`bls_slow` restates astropy's brute-force ``bls_slow``
  (bls/methods.py:98-143) in numpy for tiny inputs.  It is an independent
  definition of every output field, so fast == slow is an invariant check.
* :func:`compute_stats`, :func:`model`, :func:`transit_mask` restate
  astropy's vetting helpers (bls/core.py:375-603) in numpy.
* :func:`power_variant` evaluates single-precision alternatives of the
  algorithm for the precision study (see atg_oracle_bls_variant).

Pinning (every function here has at least one):
  - ``power``/``bls_impl``: bit-for-bit against the outputs of astropy's own
    bls_impl stored in tests/golden/*.npz, including astropy's
    test_32bit_bug known answer (test_bls.py:382-397).
  - ``bls_slow``: astropy's test_fast_method invariant (test_bls.py:414-428,
    fast == slow on all fields) on the astropy ``data`` fixture, plus the
    closed forms loglike = W_in*depth^2/2 and snr = depth/depth_err.
  - ``compute_stats`` / ``model`` / ``transit_mask``: astropy's
    test_compute_stats and test_model known answers (test_bls.py:546-658:
    per-transit counts [9, 7, 7, 7, 8], harmonic amplitude and delta log
    likelihood with and without dy, agreement with the oversample-1000
    periodogram, the model equal to a two-level linear regression).
  - ``power_variant``: equal to ``power`` bit for bit with both switches
    off (its FP32 arithmetic is what the study measures).
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libatg_oracle.so")
FIELDS = ("power", "depth", "depth_err", "duration", "transit_time",
          "depth_snr", "log_likelihood")

_lib = None


def build():
    """Compile the C restatement with the committed Makefile."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", HERE])


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run make -C oracle")
        lib = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.atg_oracle_bls.restype = ctypes.c_int
        lib.atg_oracle_bls.argtypes = [
            ctypes.c_int64, dp, dp, dp, ctypes.c_int64, dp, ctypes.c_int64,
            dp, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [dp] * 7
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def bls_impl(t_rel, y_centered, ivar, periods, durations, oversample=10,
             use_likelihood=True, n_threads=0):
    """The restated compiled core: returns the 7 output arrays with the
    transit *phase* (t_ref not yet added)."""
    lib = _load()
    t_rel = np.ascontiguousarray(t_rel, dtype=np.float64)
    y_centered = np.ascontiguousarray(y_centered, dtype=np.float64)
    ivar = np.ascontiguousarray(ivar, dtype=np.float64)
    periods = np.ascontiguousarray(np.atleast_1d(periods), dtype=np.float64)
    durations = np.ascontiguousarray(np.atleast_1d(durations),
                                     dtype=np.float64)
    outs = [np.empty(periods.size) for _ in FIELDS]
    rc = lib.atg_oracle_bls(t_rel.size, _p(t_rel), _p(y_centered), _p(ivar),
                            periods.size, _p(periods), durations.size,
                            _p(durations), int(oversample),
                            int(bool(use_likelihood)), int(n_threads),
                            *[_p(o) for o in outs])
    if rc == 1:
        raise ValueError("invalid period")
    if rc == 2:
        raise ValueError("invalid duration")
    if rc != 0:
        raise RuntimeError("oracle failed with code %d" % rc)
    return dict(zip(FIELDS, outs))


def power_variant(t, y, dy, periods, durations, objective="likelihood",
                  oversample=10, fold32=False, acc32=False, n_threads=0):
    """Precision-study variant of :func:`power`: FP32 phase fold and/or FP32
    bin accumulation (see atg_oracle_bls_variant).  Returns power, depth,
    duration and transit_time."""
    lib = _load()
    if not hasattr(lib, "_variant_ready"):
        dp = ctypes.POINTER(ctypes.c_double)
        lib.atg_oracle_bls_variant.restype = ctypes.c_int
        lib.atg_oracle_bls_variant.argtypes = [
            ctypes.c_int64, dp, dp, dp, ctypes.c_int64, dp, ctypes.c_int64,
            dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_int] + [dp] * 4
        lib._variant_ready = True
    t_rel, yc, ivar, t_ref = prepare(t, y, dy)
    periods = np.ascontiguousarray(np.atleast_1d(periods), dtype=np.float64)
    durations = np.ascontiguousarray(np.atleast_1d(durations),
                                     dtype=np.float64)
    outs = [np.empty(periods.size) for _ in range(4)]
    rc = lib.atg_oracle_bls_variant(
        t_rel.size, _p(t_rel), _p(yc), _p(ivar), periods.size, _p(periods),
        durations.size, _p(durations), int(oversample),
        int(objective == "likelihood"), int(bool(fold32)), int(bool(acc32)),
        int(n_threads), *[_p(o) for o in outs])
    if rc != 0:
        raise RuntimeError("oracle variant failed with code %d" % rc)
    out = dict(zip(("power", "depth", "duration", "transit_time"), outs))
    out["transit_time"] = out["transit_time"] + t_ref
    return out


def prepare(t, y, dy):
    """astropy power() input preparation (core.py:304-311, 327):
    FP64 t, t_ref = min(t), ivar = 1/dy^2 (ones if dy is None),
    y - median(y)."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    t_ref = np.min(t)
    y = np.ascontiguousarray(y, dtype=np.float64)
    if dy is None:
        ivar = np.ones_like(y)
    else:
        ivar = 1.0 / np.ascontiguousarray(dy, dtype=np.float64) ** 2
    return t - t_ref, y - np.median(y), ivar, t_ref


def power(t, y, dy, periods, durations, objective="likelihood",
          oversample=10, n_threads=0):
    """astropy BoxLeastSquares(t, y, dy).power(periods, durations,
    objective=objective, oversample=oversample) on plain float arrays."""
    periods = np.atleast_1d(np.abs(np.asarray(periods, dtype=np.float64)))
    durations = np.atleast_1d(np.asarray(durations, dtype=np.float64))
    if not np.min(periods) > np.max(durations):
        raise ValueError("The maximum transit duration must be shorter "
                         "than the minimum period")
    t_rel, yc, ivar, t_ref = prepare(t, y, dy)
    out = bls_impl(t_rel, yc, ivar, periods, durations, oversample,
                   objective == "likelihood", n_threads)
    out["transit_time"] = out["transit_time"] + t_ref
    return out


def bls_slow(t, y, dy, periods, durations, objective="likelihood",
             oversample=10):
    """Brute-force periodogram restated from astropy bls/methods.py:98-143
    (``_bls_slow_one``), with the power() preparation applied first."""
    t_rel, y, ivar, t_ref = prepare(t, y, dy)
    use_likelihood = objective == "likelihood"
    durations = np.atleast_1d(np.asarray(durations, dtype=np.float64))
    rows = []
    min_t = np.min(t_rel)
    for period in np.atleast_1d(periods):
        best = (-np.inf, None)
        hp = 0.5 * period
        for dur in durations:
            d_phase = dur / oversample
            for t0 in np.arange(0, period + d_phase, d_phase):
                m_in = np.abs((t_rel - min_t - t0 + hp) % period - hp) \
                    < 0.5 * dur
                m_out = ~m_in
                ivar_in = np.sum(ivar[m_in])
                ivar_out = np.sum(ivar[m_out])
                with np.errstate(divide="ignore", invalid="ignore"):
                    y_in = np.sum(y[m_in] * ivar[m_in]) / ivar_in
                    y_out = np.sum(y[m_out] * ivar[m_out]) / ivar_out
                    depth = y_out - y_in
                    depth_err = np.sqrt(1.0 / ivar_in + 1.0 / ivar_out)
                    snr = depth / depth_err
                loglike = -0.5 * np.sum((y_in - y[m_in]) ** 2 * ivar[m_in])
                loglike += 0.5 * np.sum((y_out - y[m_in]) ** 2 * ivar[m_in])
                objective_v = loglike if use_likelihood else snr
                if depth > 0 and objective_v > best[0]:
                    best = (objective_v,
                            (objective_v, depth, depth_err, dur,
                             (t0 + min_t) % period, snr, loglike))
        rows.append(best[1])
    out = dict(zip(FIELDS, map(np.array, zip(*rows))))
    out["transit_time"] = out["transit_time"] + t_ref
    return out


def _box_mask(t, period, duration, transit_time, phase_shift=0.5):
    hp = phase_shift * period
    return np.abs((t - transit_time + hp) % period - hp) < 0.5 * duration


def transit_mask(t, period, duration, transit_time):
    """astropy BoxLeastSquares.transit_mask (core.py:572-603) for float
    inputs: |(t - T0 + P/2) mod P - P/2| < duration/2."""
    return _box_mask(np.asarray(t, dtype=np.float64), float(period),
                     float(duration), float(transit_time))


def model(t, y, dy, t_model, period, duration, transit_time):
    """astropy BoxLeastSquares.model (core.py:375-425) for float inputs:
    the weighted means inside and outside the transit mask of the data,
    evaluated on t_model."""
    t = np.asarray(t, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    ivar = np.ones_like(y) if dy is None else \
        1.0 / np.asarray(dy, dtype=np.float64) ** 2
    m_in = _box_mask(t, period, duration, transit_time)
    y_in = np.sum(y[m_in] * ivar[m_in]) / np.sum(ivar[m_in])
    y_out = np.sum(y[~m_in] * ivar[~m_in]) / np.sum(ivar[~m_in])
    t_model = np.asarray(t_model, dtype=np.float64)
    out = np.full(t_model.shape, y_out)
    out[_box_mask(t_model, period, duration, transit_time)] = y_in
    return out


def compute_stats(t, y, dy, period, duration, transit_time):
    """astropy BoxLeastSquares.compute_stats (core.py:427-570) for float
    inputs (raw t and y; t is not shifted by min(t) there)."""
    t = np.asarray(t, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    ivar = np.ones_like(y) if dy is None else \
        1.0 / np.asarray(dy, dtype=np.float64) ** 2
    P, d, T0 = float(period), float(duration), float(transit_time)

    def level(m, ref=None):
        # weighted mean of the masked points, or the depth below ``ref``
        # (level, variance) with its uncertainty; (0, inf) when undefined
        if np.any(m) and (ref is None or np.isfinite(ref[1])):
            var = 1.0 / np.sum(ivar[m])
            mean = np.sum(y[m] * ivar[m]) * var
            if ref is None:
                return mean, var
            return ref[0] - mean, np.sqrt(var + ref[1])
        return 0.0, np.inf

    m_in = _box_mask(t, P, d, T0)
    m_out = ~m_in
    m_odd = np.abs((t - T0) % (2 * P) - P) < 0.5 * d
    m_even = np.abs((t - T0 + P) % (2 * P) - P) < 0.5 * d
    out = level(m_out)
    depth = level(m_in, out)
    depth_odd = level(m_odd, out)
    depth_even = level(m_even, out)
    y_out = out[0]
    y_in = y_out - depth[0]
    m_phase = np.abs((t - T0) % P - 0.5 * P) < 0.5 * d
    depth_phased = level(m_phase, level(~m_phase & m_out))
    m_half = np.abs((t - T0 + 0.25 * P) % (0.5 * P) - 0.25 * P) < 0.5 * d
    depth_half = level(m_half, level(~m_half))

    ids = np.round((t[m_in] - T0) / P).astype(np.int64)
    first = ids.min()
    n_tr = ids.max() - first + 1
    counts = np.bincount(ids - first, minlength=n_tr)
    ll = -0.5 * ivar[m_in] * ((y[m_in] - y_in) ** 2 - (y[m_in] - y_out) ** 2)
    lls = np.array([np.sum(ll[ids - first == i]) for i in range(n_tr)])
    full_ll = -0.5 * np.sum(ivar[m_in] * (y[m_in] - y_in) ** 2)
    full_ll -= 0.5 * np.sum(ivar[m_out] * (y[m_out] - y_out) ** 2)
    A = np.vstack((np.sin(2 * np.pi * t / P), np.cos(2 * np.pi * t / P),
                   np.ones_like(t))).T
    w = np.linalg.solve(A.T @ (A * ivar[:, None]), A.T @ (y * ivar))
    sin_ll = -0.5 * np.sum((y - A @ w) ** 2 * ivar)
    return dict(transit_times=P * np.arange(first, first + n_tr) + T0,
                per_transit_count=counts, per_transit_log_likelihood=lls,
                depth=depth, depth_phased=depth_phased, depth_half=depth_half,
                depth_odd=depth_odd, depth_even=depth_even,
                harmonic_amplitude=float(np.sqrt(np.sum(w[:2] ** 2))),
                harmonic_delta_log_likelihood=float(sin_ll - full_ll))


def load_golden(path):
    """Load a golden fixture (data only; no pickles)."""
    d = np.load(path, allow_pickle=False)
    case = dict(t=d["t"], y=d["y"], dy=d["dy"] if "dy" in d.files else None,
                periods=d["periods"], durations=d["durations"],
                oversample=int(d["oversample"]),
                objective=str(d["objective"]))
    case["out"] = {k: d["out_" + k] for k in FIELDS}
    return case
