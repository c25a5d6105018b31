"""Drop-in of the reference's public BLS API on the MI355X engine.

Mirrors src/astrotransit_gpu/search/api.py:
  * ``Candidate`` (api.py:7-13) and ``BLSResult`` (api.py:15-25), whose
    first nine fields keep the reference's names and order; astropy's
    per-period fields (astropy timeseries/periodograms/bls/core.py:760-801)
    follow as extra fields.
  * ``BoxLeastSquaresGPU(t, y, dy).power(periods, durations, n_bins=200,
    dtype=np.float32, method="fast", **kwargs)`` (api.py:27-140) with the
    same validation messages (api.py:51-70).

Both methods compute astropy's BoxLeastSquares objective (the reference's
oracle, search/cpu_reference_bls.py:11-16) on the GPU with an FP64 phase
fold and int64 fixed-point bins, so both are bit-reproducible run to run:
  * ``method="parity"``  the oracle's exact bin rule and an oracle-order
    recomputation of the near-best periods: within 1e-6 (relative power)
    of the FP64 oracle, the best period's fields equal to it;
  * ``method="fast"``    a fused-multiply-add fold without the bin-edge
    test and no recomputation: a point exactly on a bin edge may land in
    the neighbouring bin (a stated drift).
The top candidates come from the device top-k (``Context.top_k``), with
the rule of the reference's get_top_k_candidates.
"""
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional

import numpy as np

from .engine import FIELDS, DeviceBatch, default_context


@dataclass
class Candidate:
    period: float
    t0: float
    duration: float
    depth: float
    power: float


@dataclass
class BLSResult:
    best_period: float
    best_t0: float
    best_duration: float
    best_depth: float
    best_power: float
    periods: np.ndarray
    power: np.ndarray
    top_candidates: List[Candidate]
    metadata: Dict[str, Any]
    # astropy BoxLeastSquaresResults fields (per period)
    objective: str = "snr"
    depth: Optional[np.ndarray] = None
    depth_err: Optional[np.ndarray] = None
    duration: Optional[np.ndarray] = None
    transit_time: Optional[np.ndarray] = None
    depth_snr: Optional[np.ndarray] = None
    log_likelihood: Optional[np.ndarray] = None
    best_index: int = 0

    @property
    def period(self):
        return self.periods

    @property
    def snr(self):
        return self.depth_snr


def get_top_k_candidates(results, k=5, min_dist_bins=10):
    """Top-k peaks of the power spectrum, as the reference's
    get_top_k_candidates (search/gpu_bls.py:146-174): scipy find_peaks with
    a minimum separation, strongest first; argsort fallback when there is no
    peak."""
    from scipy.signal import find_peaks

    power = np.asarray(results["power"])
    finite = np.where(np.isfinite(power), power, -np.inf)
    peaks, _ = find_peaks(finite, distance=min_dist_bins)
    if len(peaks) == 0:
        peaks = np.argsort(finite)[-k:][::-1]
    order = np.argsort(finite[peaks])[::-1]
    top = peaks[order[:k]]
    return [dict(period=float(results["periods"][i]),
                 t0=float(results["all_t0s"][i]),
                 duration=float(results["all_durs"][i]),
                 depth=float(results["all_depths"][i]),
                 power=float(power[i])) for i in top]


def _host(a):
    if a is None:
        return None
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a)


class BoxLeastSquaresGPU:
    """Astropy-compatible Box Least Squares on an MI355X (gfx950) GPU."""

    def __init__(self, t, y, dy=None):
        self.t = t
        self.y = y
        self.dy = dy

    def power(self, periods, durations, n_bins: int = 200,
              dtype: Any = np.float32, method: str = "fast",
              objective: str = "snr", oversample: int = 10, device=None,
              **kwargs) -> BLSResult:
        """Compute the BLS power spectrum.

        Args:
            periods: periods to search.
            durations: transit durations to search.
            n_bins: accepted for compatibility; the phase grid follows the
                oracle's rule (bins of min(durations)/oversample).
            dtype: accepted for compatibility; the fold is FP64 and the bins
                int64 whatever the dtype.
            method: 'parity' (bit-reproducible) or 'fast'.
            objective: 'snr' (the reference's oracle, cpu_reference_bls.py:
                16) or 'likelihood' (astropy's default).
            oversample: bins per minimum duration (astropy default 10).
            **kwargs: ``oversample`` and ``max_bins`` as in the reference
                (max_bins is accepted; bins are never capped).
        """
        periods = np.atleast_1d(_host(periods))
        durations = np.atleast_1d(_host(durations))
        t = _host(self.t)
        y = _host(self.y)
        dy = _host(self.dy)
        oversample = int(kwargs.get("oversample", oversample))

        if np.any(periods <= 0):
            raise ValueError("All periods must be positive.")
        if np.any(durations <= 0):
            raise ValueError("All durations must be positive.")
        if np.any(np.isnan(t)) or np.any(np.isnan(y)):
            raise ValueError("Input time or flux contains NaNs.")
        if dy is not None:
            if np.any(dy <= 0):
                raise ValueError(
                    "flux_err must be strictly positive for weighted BLS.")
            if len(dy) != len(y):
                raise ValueError(
                    "flux_err and flux must have the same length.")
        if len(t) != len(y):
            raise ValueError("time and flux must have the same length")
        if not np.min(periods) > np.max(durations):
            raise ValueError("The maximum transit duration must be shorter "
                             "than the minimum period")
        if oversample < 1:
            raise ValueError("oversample must be greater than or equal to 1")
        if objective not in ("snr", "likelihood"):
            raise ValueError("objective must be 'snr' or 'likelihood'")
        mode = "parity" if method == "parity" else "fast"

        ctx = default_context(device)
        y32 = y.dtype == np.float32 and (dy is None or dy.dtype == np.float32)
        batch = DeviceBatch.from_host(
            [t], [y], None if dy is None else [dy], periods, durations,
            device=ctx.device, y_dtype=np.float32 if y32 else np.float64)
        dres = ctx.run(batch, objective=objective, mode=mode,
                       oversample=oversample, sync=False)
        tk_idx, tk_cnt = ctx.top_k(batch, dres, k=5, distance=10, sync=False)
        res = dres.to_host()
        status = int(res["status"][0])
        if status != 0:
            from ._abi import LC_STATUS
            raise ValueError("BLS failed: %s" % LC_STATUS.get(status, status))
        best = int(res["best_index"][0])
        f = {k: res[k] for k in FIELDS}
        top = [Candidate(float(periods[i]), float(f["transit_time"][i]),
                         float(f["duration"][i]), float(f["depth"][i]),
                         float(f["power"][i]))
               for i in tk_idx[0, :int(tk_cnt[0])].tolist()]
        return BLSResult(
            best_period=float(periods[best]),
            best_t0=float(f["transit_time"][best]),
            best_duration=float(f["duration"][best]),
            best_depth=float(f["depth"][best]),
            best_power=float(f["power"][best]),
            periods=periods,
            power=f["power"],
            top_candidates=top,
            metadata={"n_bins": "dynamic (min(durations)/oversample)",
                      "dtype": str(np.dtype(dtype)),
                      "n_data": len(t), "method": method,
                      "objective": objective, "oversample": oversample},
            objective=objective,
            depth=f["depth"], depth_err=f["depth_err"],
            duration=f["duration"], transit_time=f["transit_time"],
            depth_snr=f["depth_snr"], log_likelihood=f["log_likelihood"],
            best_index=best,
        )

    def compute_stats(self, period, duration, transit_time, device=None,
                       max_transits=4096):
        """astropy BoxLeastSquares.compute_stats (astropy timeseries/
        periodograms/bls/core.py:427-570) for float inputs, on the GPU
        (atg_bls_stats): depth, depth_phased, depth_half, depth_odd,
        depth_even as (value, error), harmonic_amplitude,
        harmonic_delta_log_likelihood, transit_times, per_transit_count,
        per_transit_log_likelihood."""
        period, duration = _check_candidate(period, duration)
        ctx = default_context(device)
        st = ctx.compute_stats(_vetting_batch(self, ctx), [0], period,
                               duration, float(transit_time), max_transits)[0]
        if st["status"] == 1:
            raise ValueError("no data point lies inside the transit box")
        if st["status"] == 2:
            raise ValueError("more than %d transits in the baseline"
                             % max_transits)
        st.pop("status")
        st.pop("model_levels")
        return st

    def model(self, t_model, period, duration, transit_time, device=None):
        """astropy BoxLeastSquares.model (core.py:375-425) for float
        inputs: the weighted mean flux inside and outside the box, evaluated
        at t_model (atg_bls_stats for the levels, atg_bls_model for the
        times)."""
        period, duration = _check_candidate(period, duration)
        ctx = default_context(device)
        st = ctx.compute_stats(_vetting_batch(self, ctx), [0], period,
                               duration, float(transit_time), 1)[0]
        if st["status"] == 1:
            raise ValueError("no data point lies inside the transit box")
        y_in, y_out = st["model_levels"]
        return ctx.model(_host(t_model), period, duration, transit_time, y_in,
                         y_out)[1]

    def transit_mask(self, t, period, duration, transit_time, device=None):
        """astropy BoxLeastSquares.transit_mask (core.py:572-603) for float
        inputs (atg_bls_model)."""
        period, duration = _check_candidate(period, duration)
        return default_context(device).model(_host(t), period, duration,
                                             transit_time)[0]


def _vetting_batch(obj, ctx):
    t = _host(obj.t)
    y = _host(obj.y)
    dy = _host(obj.dy)
    y32 = y.dtype == np.float32 and (dy is None or dy.dtype == np.float32)
    return DeviceBatch.from_host([t], [y], None if dy is None else [dy],
                                 np.array([1.0]), np.array([0.5]),
                                 device=ctx.device,
                                 y_dtype=np.float32 if y32 else np.float64)


def _check_candidate(period, duration):
    period = float(np.asarray(period))
    duration = float(np.asarray(duration))
    if period <= 0:
        raise ValueError("All periods must be positive.")
    if duration <= 0:
        raise ValueError("All durations must be positive.")
    if not period > duration:
        raise ValueError("The maximum transit duration must be shorter "
                         "than the minimum period")
    return period, duration

