"""The reference's host-wrapper functions on the MI355X engine.

Same names, arguments and result keys as the reference's
  * ``run_gpu_bls``            (src/astrotransit_gpu/search/gpu_bls.py:42-144)
  * ``run_vbls_exact_parity``  (src/astrotransit_gpu/search/v42_parity.py:27-107)
  * ``run_vbls_massive``       (src/astrotransit_gpu/search/vbls.py:62-188)
so that callers of those functions (screener, refiner, benchmarks) can switch
imports.  Every one of them runs the BLS objective of astropy (the
reference's oracle) through the C ABI; arrays come back as numpy arrays.

Differences forced by the oracle: ``power`` is the chosen objective
(``"snr"`` by default, the reference oracle's), ``n_bins``/``max_bins``/
``dtype``/batch sizes are accepted and ignored (bins follow astropy's rule
and one launch covers the whole batch), t0 is the absolute mid-transit time.
Zero weights (the screener's padding, search/screener.py:55) become points
with infinite dy: they keep their place in the time axis but carry no weight.
"""
import numpy as np

from .engine import FIELDS, DeviceBatch, default_context


def _host(a):
    if a is None:
        return None
    if hasattr(a, "get"):
        a = a.get()
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a)


def _matrix_batch(time_array, flux_matrix, weights_matrix, periods,
                  durations, ctx):
    t = np.asarray(_host(time_array), dtype=np.float64)
    flux = np.atleast_2d(_host(flux_matrix))
    n_t = flux.shape[0]
    if weights_matrix is None:
        dys = None
    else:
        w = np.atleast_2d(np.asarray(_host(weights_matrix), dtype=np.float64))
        with np.errstate(divide="ignore"):
            dys = [np.where(w[i] > 0, 1.0 / np.sqrt(w[i]), np.inf)
                   for i in range(n_t)]
    return DeviceBatch.from_host([t] * n_t, [flux[i] for i in range(n_t)],
                                 dys, np.asarray(_host(periods), np.float64),
                                 np.asarray(_host(durations), np.float64),
                                 device=ctx.device, y_dtype=np.float64)


def run_gpu_bls(time, flux, periods, durations, flux_err=None, n_bins=200,
                dtype=np.float32, objective="snr", method="fast"):
    """One curve; result keys of the reference's run_gpu_bls."""
    if len(time) != len(flux):
        raise ValueError("time and flux must have the same length")
    if flux_err is not None and len(flux_err) != len(flux):
        raise ValueError("flux_err must have the same length as flux")
    ctx = default_context()
    b = DeviceBatch.from_host(
        [np.asarray(_host(time), np.float64)], [_host(flux)],
        None if flux_err is None else [_host(flux_err)],
        np.asarray(_host(periods), np.float64),
        np.asarray(_host(durations), np.float64), device=ctx.device)
    h = ctx.run(b, objective=objective,
                mode="fast" if method == "fast" else "parity").to_host()
    i = int(h["best_index"][0])
    periods = np.asarray(_host(periods), np.float64)
    return {"best_period": float(periods[i]),
            "best_t0": float(h["transit_time"][i]),
            "best_duration": float(h["duration"][i]),
            "best_depth": float(h["depth"][i]),
            "power": h["power"], "snr": float(h["power"][i]),
            "all_t0s": h["transit_time"], "all_durs": h["duration"],
            "all_depths": h["depth"], "periods": periods}


def run_vbls_exact_parity(time_array, flux_matrix, periods, durations,
                          weights_matrix=None, oversample=10, max_bins=2000,
                          dtype=None, objective="snr"):
    """T curves on one time axis, per-period arrays (T, P); result keys of
    the reference's run_vbls_exact_parity."""
    ctx = default_context()
    b = _matrix_batch(time_array, flux_matrix, weights_matrix, periods,
                      durations, ctx)
    h = ctx.run(b, objective=objective, mode="parity",
                oversample=oversample).to_host()
    n_t = b.n_lc
    per = np.asarray(_host(periods), np.float64)
    shape = (n_t, per.size)
    best = h["best_index"]
    rows = np.arange(n_t)
    pw = h["power"].reshape(shape)
    t0 = h["transit_time"].reshape(shape)
    du = h["duration"].reshape(shape)
    de = h["depth"].reshape(shape)
    return {"snr": pw[rows, best], "best_period": per[best],
            "best_t0": t0[rows, best], "best_duration": du[rows, best],
            "best_depth": de[rows, best], "power_array": pw,
            "t0_array": t0, "dur_array": du, "depth_array": de,
            "status": h["status"]}


def run_vbls_massive(time_array, flux_matrix, periods, durations,
                     weights_matrix=None, n_bins=128, dtype=None,
                     target_batch_size=4000, period_batch_size=20000,
                     use_blackwell=None, objective="snr", mode="parity"):
    """Best period per curve only; result keys of the reference's
    run_vbls_massive.  Unlike the reference (per-target CAS race on the
    winner's metadata, vbls_blackwell_v39.cu:15-24; 128-bin scan,
    vbls_blackwell_v39.cu:111) the winner is the deterministic first
    maximum over the full periodogram."""
    out = run_vbls_exact_parity(time_array, flux_matrix, periods, durations,
                                weights_matrix, objective=objective) \
        if mode == "parity" else None
    if out is None:
        ctx = default_context()
        b = _matrix_batch(time_array, flux_matrix, weights_matrix, periods,
                          durations, ctx)
        h = ctx.run(b, objective=objective, mode=mode).to_host()
        per = np.asarray(_host(periods), np.float64)
        n_t = b.n_lc
        rows = np.arange(n_t)
        best = h["best_index"]
        sh = (n_t, per.size)
        out = {"snr": h["power"].reshape(sh)[rows, best],
               "best_period": per[best],
               "best_t0": h["transit_time"].reshape(sh)[rows, best],
               "best_duration": h["duration"].reshape(sh)[rows, best],
               "best_depth": h["depth"].reshape(sh)[rows, best]}
    return {k: out[k] for k in ("best_period", "best_t0", "best_depth",
                                "best_duration", "snr")}
