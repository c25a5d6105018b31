"""The reference's host-wrapper functions on the MI355X engine.

Same names, arguments and result keys as the reference's
  * ``run_gpu_bls``            (src/astrotransit_gpu/search/gpu_bls.py:42-144)
  * ``run_vbls_exact_parity``  (src/astrotransit_gpu/search/v42_parity.py:27-107)
  * ``run_vbls_massive``       (src/astrotransit_gpu/search/vbls.py:62-188)
and the batched form of the refiner's per-target re-search
  * ``refine_targets``  (src/astrotransit_gpu/search/refiner.py:102-140)
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
    run_vbls_massive.  One best-only launch (atg_bls_run_best): the
    per-period arrays stay on the device and only the winners come back.
    Unlike the reference (per-target CAS race on the winner's metadata,
    vbls_blackwell_v39.cu:15-24; 128-bin scan, vbls_blackwell_v39.cu:111)
    the winner is the deterministic first maximum over the full
    periodogram."""
    ctx = default_context()
    b = _matrix_batch(time_array, flux_matrix, weights_matrix, periods,
                      durations, ctx)
    h = ctx.run_best(b, objective=objective, mode=mode).to_host()
    return {"best_period": h["period"], "best_t0": h["transit_time"],
            "best_depth": h["depth"], "best_duration": h["duration"],
            "snr": h["power"]}


def refine_targets(targets, n_periods=10000, top_k=5, method="fast",
                   objective="snr", p_min=0.5, p_max=20.0,
                   durations=None):
    """The re-search of the reference's CandidateRefiner.refine_from_csv
    (search/refiner.py:102-140) for every selected target in one launch.

    ``targets`` maps a TIC id to its (time, flux, flux_err) arrays (the
    reference's cache.get_target_data; flux_err may be None).  The grid is
    the refiner's: ``np.linspace(p_min, p_max, n_periods)`` periods and
    ``np.linspace(0.01, 0.2, 5)`` durations, searched with
    BoxLeastSquaresGPU.power's defaults (fast method, snr objective).
    Returns the rows the refiner writes (tic_id, status, period, t0, depth,
    duration, power, n_points, error): the first ``top_k`` candidates of
    each target by the rule of get_top_k_candidates (gpu_bls.py:146-174),
    equal to a per-target BoxLeastSquaresGPU(...).power call's
    top_candidates.  A target the search rejects gets one row with status
    "error" and the reason, where the refiner prints the exception and
    skips it.
    """
    periods = np.linspace(p_min, p_max, int(n_periods))
    durs = np.linspace(0.01, 0.2, 5) if durations is None else \
        np.atleast_1d(np.asarray(durations, np.float64))
    ids = list(targets)
    if not ids:
        return []
    data = [tuple(_host(a) for a in targets[i]) for i in ids]
    ctx = default_context()
    ts = [np.asarray(d[0], np.float64) for d in data]
    ys = [np.asarray(d[1], np.float64) for d in data]
    dys = None
    if any(len(d) > 2 and d[2] is not None for d in data):
        dys = [np.asarray(d[2], np.float64) if len(d) > 2 and d[2] is not None
               else np.ones_like(ys[k]) for k, d in enumerate(data)]
    b = DeviceBatch.from_host(ts, ys, dys, periods, durs, device=ctx.device,
                              y_dtype=np.float64)
    res = ctx.run(b, objective=objective,
                  mode="parity" if method == "parity" else "fast")
    cands = ctx.candidates(b, res, k=5, distance=10)
    status = res.status.cpu().numpy()
    from ._abi import LC_STATUS
    rows = []
    for k, tid in enumerate(ids):
        if status[k] != 0:
            rows.append({"tic_id": int(tid), "status": "error",
                         "period": np.nan, "t0": np.nan, "depth": np.nan,
                         "duration": np.nan, "power": np.nan,
                         "n_points": int(len(ts[k])),
                         "error": "BLS failed: %s"
                                  % LC_STATUS.get(int(status[k]),
                                                  int(status[k]))})
            continue
        for c in cands[k][:top_k]:
            rows.append({"tic_id": int(tid), "status": "ok",
                         "period": float(c["period"]), "t0": float(c["t0"]),
                         "depth": float(c["depth"]),
                         "duration": float(c["duration"]),
                         "power": float(c["power"]),
                         "n_points": int(len(ts[k])), "error": ""})
    return rows
