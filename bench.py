"""Benchmark of the BLS hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (one "step"): the BLS periodogram of 1,000 seeded light curves x
15,000 points (FP64 t, FP32 y/dy, 27 d baseline, 2-min cadence, two random
gaps, half with an injected box) over 5,000 log-uniform periods
0.5-13.5 d x 5 durations 0.04-0.20 d, likelihood objective, parity mode,
inputs resident in HBM, one batched launch sequence per step.

Printed beside the headline: the plain one-curve-per-call GPU form (the
API, host arrays in and out, as the reference's per-curve loop), fast mode,
the continuous-batching server (LC/s with p50/p99 latency), the roofline of
the main kernel (VALU lane-ops/s and LDS bytes/s from its average duration,
timed with HIP events on its stream) and the FP64 CPU oracle on the host
cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("LC/s at 15k pts x 5k periods x 5 durs vs plain GPU; p50/p99 "
          "latency")

# MI355X public peaks: 256 CUs at 2.4 GHz, 64 lanes per CU per clock for
# FP64/FP32 non-packed VALU ops (FP64 vector 78.6 TFLOP/s = 39.3 T FMA/s),
# LDS 128 B/clk/CU.
PEAK_VALU = 256 * 64 * 2.4e9          # lane-ops/s
PEAK_LDS = 256 * 128 * 2.4e9          # bytes/s
PEAK_HBM = 8.0e12                     # bytes/s


def algorithmic_work(n_points, periods, durations, oversample=10):
    """Per-curve algorithmic work of the path (bin rule of the oracle):
    folds = N*P; box evaluations E = sum_p sum_k (n_bins(p) - m_k + 1);
    VALU lane-ops = 6*folds + 4*sum n_bins + 16*E; LDS bytes with 8-byte
    bins = folds*2*2*8 + sum_p n_bins(p)*(2*2*8 + (1+D)*2*8)."""
    bd = np.min(durations) / oversample
    nb = np.ceil(periods / bd).astype(np.int64) + oversample
    m = np.round(np.asarray(durations) / bd).astype(np.int64)
    E = int(np.sum(np.maximum(nb[:, None] - m[None, :] + 1, 0)))
    folds = int(n_points) * periods.size
    D = len(durations)
    valu = 6 * folds + 4 * int(nb.sum()) + 16 * E
    lds = folds * 2 * 2 * 8 + int(nb.sum()) * (2 * 2 * 8 + (1 + D) * 2 * 8)
    return dict(folds=folds, box_evals=E, valu_lane_ops=valu, lds_bytes=lds,
                sum_bins=int(nb.sum()))


def profile_block(name):
    """The ``key=value`` lines of a committed rocprofv3 summary under
    profiles/ (its machine-readable block), as floats; None if absent."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        kv = {}
        for line in open(path):
            k, sep, v = line.strip().partition("=")
            if sep and k and " " not in k:
                try:
                    kv[k] = float(v)
                except ValueError:
                    kv[k] = v
        return kv or None
    except OSError:
        return None


def measured_traffic(n_lc):
    """HBM bytes per launch of the periodogram stage (k_main, k_big and
    k_fields) from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (copy-kernel calibrated), scaled to n_lc curves."""
    kv = profile_block("periodogram_stage_hbm_traffic_1000x15k_5000p_5d.md")
    try:
        return int(kv["traffic_bytes"] * n_lc / kv["n_lc"])
    except (TypeError, KeyError):
        return None


def measured_rates(n_lc, kernel_ms, algo_valu):
    """Counter-measured VALU and LDS rates of k_main at the headline shape:
    the per-launch counts of the committed rocprofv3 summary (SQ_INSTS_VALU,
    SQ_ACTIVE_INST_VALU, the LDS _BANDWIDTH counters, SQ_LDS_BANK_CONFLICT),
    scaled to n_lc curves, over this run's live stage time ``kernel_ms``."""
    name = "k_main_counters_1000x15k_5000p_5d.md"
    kv = profile_block(name)
    try:
        f = n_lc / kv["n_lc"]
        lane_ops = kv["valu_insts"] * 64.0 * f
        lds = kv["lds_bytes"] * f
        sec = kernel_ms * 1e-3
        return {
            "source": "profiles/" + name,
            "valu_lane_ops_per_launch": lane_ops,
            "valu_achieved": round(lane_ops / sec / 1e12, 3),
            "valu_frac": round(lane_ops / sec / PEAK_VALU, 4),
            "valu_busy": round(kv["valu_busy"], 4),
            "algorithmic_share_of_issued": round(algo_valu / lane_ops, 4),
            "lds_bytes_per_launch": lds,
            "lds_achieved_GBps": round(lds / sec / 1e9, 1),
            "lds_frac": round(lds / sec / PEAK_LDS, 4),
            "lds_bank_conflict_frac": round(kv["lds_bank_conflict_frac"], 4),
            "profiled_kernel_ms": kv["duration_ms"]}
    except (TypeError, KeyError, ZeroDivisionError):
        return None


def host_threads():
    """Host threads this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS where that is set (a shared GPU host gives each GPU's
    process a share of its cores through it)."""
    n = len(os.sched_getaffinity(0))
    try:
        n = min(n, int(os.environ.get("OMP_NUM_THREADS", n)))
    except ValueError:
        pass
    return max(1, n)


def cpu_baseline(d, n_sample, objective, gpu=None, n_single=8):
    """The FP64 C oracle on the host cores over the first ``n_sample``
    curves (OpenMP over periods on every thread this process may use), and
    single-threaded over the first ``n_single`` curves (the like of the
    reference's single-process astropy figure).  With ``gpu`` (host copies
    of the timed run's outputs and period offsets) the same oracle outputs
    give the parity of those curves."""
    from oracle import oracle
    cores = host_threads()
    refs = []
    s = time.perf_counter()
    for i in range(n_sample):
        refs.append(oracle.power(d["t"][i], d["y"][i], d["dy"][i],
                                 d["periods"], d["durations"], objective,
                                 n_threads=cores))
    el = time.perf_counter() - s
    s = time.perf_counter()
    for i in range(n_single):
        oracle.power(d["t"][i], d["y"][i], d["dy"][i], d["periods"],
                     d["durations"], objective, n_threads=1)
    el1 = time.perf_counter() - s
    out = dict(value=n_sample / el, unit="LC/s", cores=cores, kind="port",
               sample="%d of the 1,000 curves, full 5,000x5 grid, FP64 C "
               "oracle (OpenMP over periods, %d threads), %.1f s" % (
                   n_sample, cores, el),
               single_core={"value": n_single / el1, "unit": "LC/s",
                            "cores": 1,
                            "sample": "%d of the 1,000 curves, one thread, "
                            "%.1f s" % (n_single, el1)})
    if gpu is None:
        return out, None
    h, po = gpu
    rel_max = dt0_max = 0.0
    same = exact = 0
    for i, r in enumerate(refs):
        g = {k: h[k][po[i]:po[i + 1]] for k in r}
        fin = np.isfinite(r["power"]) & (r["power"] != 0)
        rel = np.abs(g["power"][fin] - r["power"][fin]) / np.abs(
            r["power"][fin])
        rel_max = max(rel_max, float(rel.max()))
        ib = int(np.argmax(r["power"]))
        same += int(int(h["best_index"][i]) == ib)
        dt0_max = max(dt0_max, abs(g["transit_time"][ib] -
                                   r["transit_time"][ib]))
        exact += int(all(g[k][ib] == r[k][ib] for k in r))
    return out, {"curves": n_sample, "power_rel_err_max": float(
        "%.3g" % rel_max), "best_period_same": "%d/%d" % (same, n_sample),
        "best_fields_bit_exact": "%d/%d" % (exact, n_sample),
        "t0_err_at_best_max_d": float("%.3g" % dt0_max)}


def hetero_leg(d, args, objective, headline_s):
    """The headline workload with per-point dy (each point's dy scaled by a
    seeded U(0.9, 1.1)), parity mode, on a context tiled by the autotuner
    (64-curve sample), and with the headline's default tiling."""
    import torch
    from astro_transit_gpu_amd import Context, DeviceBatch
    from astro_transit_gpu_amd.autotune import autotune
    rng = np.random.default_rng(22)
    dyh = [(dd * rng.uniform(0.9, 1.1, dd.size)).astype(np.float32)
           for dd in d["dy"]]
    nh = max(2, args.steps // 2)
    res = {}
    for name in ("default", "autotuned"):
        ctx = Context(0)
        tuning = None
        if name == "autotuned":
            tuning, _ = autotune(ctx, d["t"], d["y"], dyh, d["periods"],
                                 d["durations"], objective=objective,
                                 mode="parity", use_cache=False)
        b = DeviceBatch.from_host(d["t"], d["y"], dyh, d["periods"],
                                  d["durations"], device=ctx.device)
        out = ctx.run(b, objective=objective, mode="parity")
        torch.cuda.synchronize()
        ctx.timing(True, True)
        s = time.perf_counter()
        for _ in range(nh):
            ctx.run(b, objective=objective, mode="parity", out=out,
                    sync=False)
        torch.cuda.synchronize()
        el = time.perf_counter() - s
        _, kms = ctx.timing(False, True)
        res[name] = {"value": round(args.n_lc * nh / el, 2), "unit": "LC/s",
                     "kernel_ms": round(kms, 3),
                     "over_headline_time": round((el / nh) / headline_s, 3),
                     "tiling": tuning or "default",
                     "statuses_ok": int((out.status.cpu() == 0).sum())}
        del b, out, ctx
    line = dict(res["autotuned"])
    line["mode"] = "parity"
    line["default_tiling"] = res["default"]
    return line


def ragged_leg(n_lc, objective):
    """Ragged CSR batch: N log-uniform 1k-20k, baselines 10-27 d, per-curve
    autoperiod grids clipped to 1k-20k periods, 5-8 durations.  Its own
    context, tiled by the autotuner on a 64-curve sample of the workload;
    the launch sequence is captured as a HIP graph and one replay is
    timed."""
    import torch
    from astro_transit_gpu_amd import Context, DeviceBatch, synth
    from astro_transit_gpu_amd.autotune import autotune
    t0 = time.perf_counter()
    d = synth.config3(n_lc=n_lc, seed=3)
    gen = time.perf_counter() - t0
    ctx = Context(0)
    tuning, _ = autotune(ctx, d["t"], d["y"], d["dy"], d["periods"],
                         d["durations"], objective=objective, mode="parity",
                         use_cache=False)
    b = DeviceBatch.from_host(d["t"], d["y"], d["dy"], d["periods"],
                              d["durations"], device=ctx.device)
    out = ctx.run(b, objective=objective)
    g = ctx.capture(b, out, objective=objective)
    g.replay()
    torch.cuda.synchronize()
    s = time.perf_counter()
    g.replay()
    torch.cuda.synchronize()
    el = time.perf_counter() - s
    folds = sum(np.size(t) * np.size(p) for t, p in zip(d["t"], d["periods"]))
    st = out.status.cpu().numpy()
    # the replay's results on the first n_check curves against the FP64
    # oracle (the tiling changes the time only)
    from oracle import oracle
    h = out.to_host()
    po = b.host_period_offsets
    n_check = min(32, n_lc)
    rel_max, same, exact = 0.0, 0, 0
    for i in range(n_check):
        r = oracle.power(d["t"][i], d["y"][i], d["dy"][i], d["periods"][i],
                         d["durations"][i], objective,
                         n_threads=host_threads())
        g = {k: h[k][po[i]:po[i + 1]] for k in r}
        fin = np.isfinite(r["power"]) & (r["power"] != 0)
        rel_max = max(rel_max, float(np.max(
            np.abs(g["power"][fin] / r["power"][fin] - 1.0))))
        ib = int(np.argmax(r["power"]))
        same += int(int(h["best_index"][i]) == ib)
        exact += int(all(g[k][ib] == r[k][ib] for k in r))
    return {"value": round(n_lc / el, 1), "unit": "LC/s", "curves": n_lc,
            "parity_vs_oracle": {
                "curves": n_check, "power_rel_err_max": float(
                    "%.3g" % rel_max),
                "best_period_same": "%d/%d" % (same, n_check),
                "best_fields_bit_exact": "%d/%d" % (exact, n_check)},
            "points": int(sum(np.size(t) for t in d["t"])),
            "periods": int(sum(np.size(p) for p in d["periods"])),
            "folds_per_s": float("%.4g" % (folds / el)),
            "graph_replay_ms": round(el * 1e3, 1), "tiling": tuning,
            "statuses_ok": int((st == 0).sum()),
            "generation_s": round(gen, 1)}


def _stream(srv, reqs, rate, seconds, seed, label):
    """Poisson arrivals at ``rate`` req/s for ``seconds`` (requests drawn
    cyclically from ``reqs``) into ``srv`` until all are served."""
    rng = np.random.default_rng(seed)
    n_req = int(rate * seconds)
    arr = np.cumsum(rng.exponential(1.0 / rate, n_req))
    pool = len(reqs)
    handles = []
    start = time.perf_counter()
    next_report = 10.0
    i = 0
    while i < n_req or srv.pending or srv.admitted or srv.inflight:
        now = time.perf_counter() - start
        while i < n_req and arr[i] <= now:
            handles.append(srv.submit(*reqs[i % pool],
                                      arrival=start + arr[i]))
            i += 1
        if now >= next_report:
            print("%s: %.0f s, %d submitted, %d served" % (
                label, now, i, srv.completed), file=sys.stderr, flush=True)
            next_report += 10.0
        if not srv.poll():
            time.sleep(0.0005)
    el = time.perf_counter() - start
    lat = np.array([h.latency for h in handles]) * 1e3
    pts = sum(np.size(h.t) for h in handles)
    return {"value": round(len(handles) / el, 1), "unit": "LC/s",
            "offered_rate": round(rate, 1), "requests": len(handles),
            "arrival_seconds": seconds, "seconds": round(el, 2),
            "mean_points": int(pts / len(handles)),
            "p50_ms": round(float(np.percentile(lat, 50)), 1),
            "p99_ms": round(float(np.percentile(lat, 99)), 1),
            "statuses_ok": int(sum(h.status == 0 for h in handles))}


def serve5_leg(seconds, objective, rate=2000.0, pool=1024, budget=200e9,
               sub_load=0.7, sub_seconds=20.0):
    """Serving stream: Poisson arrivals at ``rate`` req/s for ``seconds``,
    each request one light curve with N log-uniform 1k-200k, baseline
    10-360 d, a random gap, an optional injected box and an autoperiod grid
    capped at 20,000 periods (``pool`` distinct seeded requests, drawn
    cyclically); admission by an HBM budget of ``budget`` bytes.  The
    offered load exceeds the device, so the first stream reports sustained
    throughput and latency at saturation; a second stream at ``sub_load``
    times that sustained rate reports the latency below saturation."""
    from astro_transit_gpu_amd import synth
    from astro_transit_gpu_amd.serving import BLSServer
    t0 = time.perf_counter()
    reqs = synth.config5_pool(pool)
    gen = time.perf_counter() - t0
    srv = BLSServer(objective=objective, hbm_budget_bytes=budget,
                    typical_n=40000, typical_p=20000, max_batch=64,
                    keep_results=False)
    sat = _stream(srv, reqs, rate, seconds, 7, "served stream")
    sub = _stream(srv, reqs, sub_load * sat["value"], sub_seconds, 8,
                  "served stream below saturation")
    sat.update({"hbm_budget_GB": budget / 1e9, "period_cap": 20000,
                "distinct_requests": pool, "generation_s": round(gen, 1),
                "below_saturation": dict(sub, load=sub_load)})
    return sat


def _err_stats(got, ref):
    """Errors of one curve's per-period outputs against the FP64 oracle."""
    fin = np.isfinite(ref["power"]) & (ref["power"] != 0)
    g = got["power"][fin]
    r = ref["power"][fin]
    rel = np.abs(g - r) / np.abs(r)
    ib = int(np.argmax(np.where(np.isfinite(ref["power"]), ref["power"],
                                -np.inf)))
    gb = int(np.argmax(np.where(np.isfinite(got["power"]), got["power"],
                                -np.inf)))
    same = (got["duration"] == ref["duration"]) & (
        np.abs(got["transit_time"] - ref["transit_time"]) <= 1e-6)
    return dict(rel=rel, best_same=gb == ib,
                dt0_best=abs(got["transit_time"][ib] - ref["transit_time"][ib]),
                ddepth_best=abs(got["depth"][ib] - ref["depth"][ib]) /
                max(abs(ref["depth"][ib]), 1e-300),
                window_same=float(np.mean(same[np.isfinite(ref["power"])])))


def _summarise(stats):
    rel = np.concatenate([s["rel"] for s in stats])
    return {"power_rel_max": float("%.3g" % rel.max()),
            "power_rel_p50": float("%.3g" % np.median(rel)),
            "power_rel_p99": float("%.3g" % np.percentile(rel, 99)),
            "frac_periods_rel_gt_1e-6": float("%.4g" % np.mean(rel > 1e-6)),
            "best_period_same": "%d/%d" % (sum(s["best_same"] for s in stats),
                                           len(stats)),
            "t0_err_at_best_max_d": float("%.3g" % max(s["dt0_best"]
                                                       for s in stats)),
            "depth_rel_err_at_best_max": float("%.3g" % max(
                s["ddepth_best"] for s in stats)),
            "window_agreement_min": float("%.4g" % min(s["window_same"]
                                                       for s in stats))}


def precision_leg(n_lc, n_sample, objective, seconds_cap=None):
    """Precision study on the long-baseline workload: 512 stitched curves x
    200,000 points over 360 d, 50,000 log-uniform periods 1-100 d, six
    durations.  The GPU runs the whole batch (FP64 fold, int64 bins) in
    parity and fast mode; on ``n_sample`` curves every output is compared
    with the FP64 oracle, beside the single-precision alternatives (FP32
    fold, FP32 bin sums and box scan) evaluated by the oracle's variants on
    the host."""
    import torch
    from oracle import oracle
    from astro_transit_gpu_amd import Context, DeviceBatch, synth
    t0 = time.perf_counter()
    d = synth.config4(n_lc=n_lc, seed=4)
    gen = time.perf_counter() - t0
    print("precision: generated %d curves in %.1f s" % (n_lc, gen),
          file=sys.stderr, flush=True)
    ctx = Context(0)
    b = DeviceBatch.from_host(d["t"], d["y"], d["dy"], d["periods"],
                              d["durations"], device=ctx.device)
    out = {}
    gpu = {}
    for mode in ("parity", "fast"):
        torch.cuda.synchronize()
        ctx.timing(True, True)
        s = time.perf_counter()
        res = ctx.run(b, objective=objective, mode=mode)
        torch.cuda.synchronize()
        el = time.perf_counter() - s
        _, kms = ctx.timing(False, True)
        h = res.to_host()
        out[mode] = [{k: v.copy() for k, v in res.curve(i, h).items()}
                     for i in range(n_sample)]
        gpu[mode] = {"LC_per_s": round(n_lc / el, 2), "seconds": round(el, 2),
                     "stage_ms": round(kms, 1),
                     "statuses_ok": int((h["status"] == 0).sum())}
        # k_big's rocprofv3 counters at this shape (64 curves; rates are
        # per launch and do not depend on the curve count)
        kv = profile_block("k_big_counters_64x200k_50000p_6d_%s.md" % mode)
        if kv:
            gpu[mode]["k_big_counters"] = {
                "source": "profiles/k_big_counters_64x200k_50000p_6d_%s.md"
                % mode,
                "valu_issue_frac": round(kv["valu_insts"] * 64.0 /
                                         (kv["duration_ms"] * 1e-3) /
                                         PEAK_VALU, 4),
                "valu_busy": round(kv["valu_busy"], 4),
                "lds_frac": round(kv["lds_bytes"] /
                                  (kv["duration_ms"] * 1e-3) / PEAK_LDS, 4),
                "lds_bank_conflict_frac": round(
                    kv["lds_bank_conflict_frac"], 4)}
        print("precision: GPU %s %.1f s" % (mode, el), file=sys.stderr,
              flush=True)
        del res, h
    cores = host_threads()
    stats = {k: [] for k in ("gpu_parity_f64fold_i64bins",
                             "gpu_fast_f64fmafold_i64bins",
                             "fp32_fold", "fp32_bins_and_scan",
                             "fp32_fold_bins_scan")}
    cpu_s = 0.0
    for i in range(n_sample):
        s = time.perf_counter()
        ref = oracle.power(d["t"][i], d["y"][i], d["dy"][i], d["periods"],
                           d["durations"], objective, n_threads=cores)
        cpu_s += time.perf_counter() - s
        stats["gpu_parity_f64fold_i64bins"].append(
            _err_stats(out["parity"][i], ref))
        stats["gpu_fast_f64fmafold_i64bins"].append(
            _err_stats(out["fast"][i], ref))
        for name, f32, a32 in (("fp32_fold", 1, 0),
                               ("fp32_bins_and_scan", 0, 1),
                               ("fp32_fold_bins_scan", 1, 1)):
            v = oracle.power_variant(d["t"][i], d["y"][i], d["dy"][i],
                                     d["periods"], d["durations"], objective,
                                     fold32=f32, acc32=a32, n_threads=cores)
            stats[name].append(_err_stats(v, ref))
        print("precision: curve %d compared" % i, file=sys.stderr,
              flush=True)
    return {"metric": "precision study: error of each result vs the FP64 "
            "oracle", "workload": "%d curves x 200,000 pts over 360 d "
            "(2-min grid, 13 periodic 1 d gaps, FP32 y/dy) x 50,000 "
            "log-uniform periods 1-100 d x 6 durations 0.04-0.3 d, %s" % (
                n_lc, objective),
            "compared_curves": n_sample, "gpu": gpu,
            "errors": {k: _summarise(v) for k, v in stats.items()}
            if n_sample else None,
            "cpu_oracle_LC_per_s": round(n_sample / cpu_s, 4)
            if n_sample else None,
            "cpu_cores": cores, "generation_s": round(gen, 1)}


def parity_sweep(n_lc, objective):
    """Every curve of the first ``n_lc`` of the 1,000-curve workload and of
    the ragged workload, parity and fast mode, against the FP64 oracle on the
    host cores: worst relative power error, share of periods beyond 1e-6,
    best-period agreement (a different best period whose oracle power equals
    the maximum within 1e-12 is counted as a tie) and bit-exact best
    fields."""
    import torch
    from oracle import oracle
    from astro_transit_gpu_amd import Context, DeviceBatch, synth
    cores = host_threads()
    ctx = Context(0)
    out = {}
    for name, d in (("1000x15k_5000p_5d", synth.config2(n_lc=n_lc, seed=2)),
                    ("ragged", synth.config3(n_lc=n_lc, seed=3))):
        per = d["periods"] if isinstance(d["periods"], list) else \
            [d["periods"]] * n_lc
        dur = d["durations"] if isinstance(d["durations"], list) else \
            [d["durations"]] * n_lc
        b = DeviceBatch.from_host(d["t"], d["y"], d["dy"], d["periods"],
                                  d["durations"], device=ctx.device)
        res = {m: ctx.run(b, objective=objective, mode=m).to_host()
               for m in ("parity", "fast")}
        po = b.host_period_offsets
        st = {m: dict(rel=0.0, big=0, tot=0, same=0, tie=0, exact=0)
              for m in res}
        for i in range(n_lc):
            r = oracle.power(d["t"][i], d["y"][i], d["dy"][i], per[i],
                             dur[i], objective, n_threads=cores)
            ib = int(np.argmax(r["power"]))
            fin = np.isfinite(r["power"]) & (r["power"] != 0)
            for m, h in res.items():
                g = {k: h[k][po[i]:po[i + 1]] for k in r}
                rel = np.abs(g["power"][fin] / r["power"][fin] - 1.0)
                s_ = st[m]
                s_["rel"] = max(s_["rel"], float(rel.max()))
                s_["big"] += int((rel > 1e-6).sum())
                s_["tot"] += int(rel.size)
                gb = int(h["best_index"][i])
                s_["same"] += int(gb == ib)
                s_["tie"] += int(gb != ib and r["power"][gb] >=
                                 r["power"][ib] * (1 - 1e-12))
                s_["exact"] += int(all(g[k][ib] == r[k][ib] for k in r))
            if i % 100 == 99:
                print("parity sweep: %s curve %d" % (name, i),
                      file=sys.stderr, flush=True)
        out[name] = {m: {"power_rel_err_max": float("%.3g" % v["rel"]),
                         "frac_periods_rel_gt_1e-6":
                             float("%.3g" % (v["big"] / max(v["tot"], 1))),
                         "best_period_same": "%d/%d" % (v["same"], n_lc),
                         "other_best_with_equal_power": v["tie"],
                         "best_fields_bit_exact": "%d/%d" % (v["exact"],
                                                             n_lc)}
                     for m, v in st.items()}
        del b, res
        torch.cuda.empty_cache()
    return {"metric": "parity sweep against the FP64 oracle", "curves": n_lc,
            "objective": objective, "cpu_cores": cores, "results": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n-lc", type=int, default=1000)
    ap.add_argument("--plain", type=int, default=20,
                    help="curves timed in the one-curve-per-call form")
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--ragged-lc", type=int, default=16384,
                    help="curves of the ragged-batch leg")
    ap.add_argument("--serve5-seconds", type=float, default=60.0,
                    help="arrival time of the long-baseline request stream "
                    "(2,000 req/s, 200 GB budget)")
    ap.add_argument("--skip-extras", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="tile the headline leg with the autotuner (default: "
                    "the library's default tiling, the one the committed "
                    "rocprofv3 summaries profile; the ragged leg always "
                    "autotunes)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="accepted for older command lines (the default)")
    ap.add_argument("--precision", action="store_true",
                    help="run only the long-baseline precision study")
    ap.add_argument("--parity-sweep", type=int, default=0, metavar="N",
                    help="run only the oracle comparison of N curves of the "
                    "1,000-curve and ragged workloads, parity and fast mode")
    ap.add_argument("--precision-lc", type=int, default=512)
    ap.add_argument("--precision-sample", type=int, default=4)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB device-to-device copy first (a kernel "
                    "of known HBM traffic for counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this engine has no multi-GPU "
              "path (curves and periods shard independently; see the design "
              "notes).", file=sys.stderr)
        sys.exit(2)

    if args.parity_sweep:
        print(json.dumps(parity_sweep(args.parity_sweep, "likelihood")),
              flush=True)
        return
    if args.precision:
        print(json.dumps(precision_leg(args.precision_lc,
                                       args.precision_sample, "likelihood")),
              flush=True)
        return

    import torch
    from astro_transit_gpu_amd import (BoxLeastSquaresGPU, Context,
                                       DeviceBatch, synth)

    objective = "likelihood"
    d = synth.config2(n_lc=args.n_lc, seed=2)
    ctx = Context(0)
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device=ctx.device)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    from astro_transit_gpu_amd.autotune import autotune
    t0 = time.perf_counter()
    if not args.autotune:
        tuning = ctx.get_tuning()
        if tuning["periods_per_item"] == 0:  # chosen per batch
            tuning["periods_per_item"] = (
                "per batch: min(24, max(4, periods / (8 x workgroups)))")
    else:
        tuning, _ = autotune(ctx, d["t"], d["y"], d["dy"], d["periods"],
                             d["durations"], objective=objective,
                             mode="parity", use_cache=False)
    tune_s = time.perf_counter() - t0
    batch = DeviceBatch.from_host(d["t"], d["y"], d["dy"], d["periods"],
                                  d["durations"], device=ctx.device)
    out = ctx.run(batch, objective=objective, mode="parity")

    def timed(mode, steps, warmup, bt=None):
        bt = batch if bt is None else bt
        for _ in range(warmup):
            ctx.run(bt, objective=objective, mode=mode, out=out, sync=False)
        torch.cuda.synchronize()
        ctx.timing(True, True)
        s = time.perf_counter()
        for _ in range(steps):
            ctx.run(bt, objective=objective, mode=mode, out=out, sync=False)
        torch.cuda.synchronize()
        el = time.perf_counter() - s
        n, kms = ctx.timing(False, True)
        return el, kms

    el, kms = timed("parity", args.steps, args.warmup)
    value = args.n_lc * args.steps / el
    work = algorithmic_work(15000, d["periods"], d["durations"])
    per_launch_valu = work["valu_lane_ops"] * args.n_lc
    per_launch_lds = work["lds_bytes"] * args.n_lc
    ach_valu = per_launch_valu / (kms * 1e-3)
    ach_lds = per_launch_lds / (kms * 1e-3)
    hbm_in = args.n_lc * 15000 * (8 + 4 + 4)
    hbm_out = args.n_lc * 5000 * 7 * 8
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "LC/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+i64",
        "data": "synthetic (seeded, numpy default_rng seed 2)",
        "config": {"workload": "1,000 curves x 15,000 pts (FP64 t, FP32 "
                   "y/dy) x 5,000 log-uniform periods 0.5-13.5 d x 5 "
                   "durations 0.04-0.20 d, likelihood, parity mode",
                   "n_lc": args.n_lc, "n_points": 15000, "n_periods": 5000,
                   "n_durations": 5, "objective": objective,
                   "mode": "parity", "tiling": tuning,
                   "autotune_s": round(tune_s, 2)},
        "roofline": {
            "bound": "valu", "achieved": round(ach_valu / 1e12, 3),
            "peak": round(PEAK_VALU / 1e12, 3), "unit": "Tlane-op/s",
            "frac": round(ach_valu / PEAK_VALU, 4),
            "traffic": measured_traffic(args.n_lc),
            # the periodogram stage: k_main and the phase-tile kernel k_big
            # that follows it (timed together with HIP events)
            "kernel": "k_main+k_big+k_fields", "kernel_ms": round(kms, 3),
            "valu_lane_ops_per_launch": per_launch_valu,
            "lds": {"achieved_GBps": round(ach_lds / 1e9, 1),
                    "peak_GBps": PEAK_LDS / 1e9,
                    "frac": round(ach_lds / PEAK_LDS, 4),
                    "bytes_per_launch": per_launch_lds},
            # achieved/frac above count the algorithmic work; "measured"
            # counts what the hardware issued (rocprofv3 counters)
            "measured": measured_rates(args.n_lc, kms, per_launch_valu),
            # the algorithmic minimum (inputs 16 B/point, outputs) and what
            # the stage moves once by design: the packed lane-major points
            # (16 B/point; every period of this workload fits the LDS
            # histogram, so no point-order reads), each period's winner
            # written by k_main and read back with its period by k_fields
            # (24 + 32 B), outputs
            "hbm_algorithmic_bytes_per_launch": hbm_in + hbm_out,
            "hbm_design_bytes_per_launch": args.n_lc * 15000 * 16 + hbm_out +
            args.n_lc * 5000 * 56,
        },
    }
    if not args.skip_extras:
        el_f, kms_f = timed("fast", max(2, args.steps // 2), 1)
        line["fast"] = {"value": round(args.n_lc * max(2, args.steps // 2) /
                                       el_f, 2), "unit": "LC/s",
                        "kernel_ms": round(kms_f, 3)}
        # the same workload with per-point dy (each point's dy scaled by a
        # seeded U(0.9, 1.1)): unpacked bins, two int64 LDS adds per point
        # and period, so twice the LDS per bin; its own context, tiled by
        # the autotuner on a 64-curve sample (the headline's tiling is
        # timed too)
        line["hetero_dy"] = hetero_leg(d, args, objective, el / args.steps)
        # top-5 candidates of every curve: device top-k vs the reference's
        # host rule (scipy find_peaks over each periodogram)
        from astro_transit_gpu_amd import get_top_k_candidates
        ctx.run(batch, objective=objective, mode="parity", out=out)
        ctx.top_k(batch, out, 5, 10)
        torch.cuda.synchronize()
        s = time.perf_counter()
        idx, cnt = ctx.top_k(batch, out, 5, 10)
        dev_ms = (time.perf_counter() - s) * 1e3
        hp = out.to_host()
        po = batch.host_period_offsets
        s = time.perf_counter()
        host_idx = []
        for i in range(args.n_lc):
            a, z = po[i], po[i + 1]
            host_idx.append([int(c["period"]) for c in get_top_k_candidates(
                dict(power=hp["power"][a:z], periods=np.arange(z - a),
                     all_t0s=hp["transit_time"][a:z],
                     all_durs=hp["duration"][a:z],
                     all_depths=hp["depth"][a:z]), k=5)])
        host_ms = (time.perf_counter() - s) * 1e3
        hidx, hcnt = idx.cpu().numpy(), cnt.cpu().numpy()
        from scipy.signal import find_peaks
        same = tied = 0
        for i in range(args.n_lc):
            if host_idx[i] == list(hidx[i, :int(hcnt[i])]):
                same += 1
                continue
            # numpy's default argsort orders equal peak powers differently
            # from machine to machine; the device ranks the later first
            x = hp["power"][po[i]:po[i + 1]]
            pk = find_peaks(x)[0]
            tied += int(len(np.unique(x[pk])) < len(pk))
        line["top_k"] = {"k": 5, "distance": 10,
                         "device_ms": round(dev_ms, 3),
                         "host_scipy_ms": round(host_ms, 1),
                         "curves_identical_to_host_rule": same,
                         "other_curves_with_tied_peak_powers": tied,
                         "curves": args.n_lc}
        # plain form: one curve per call, host arrays in and out
        lat = []
        for i in range(args.plain + 1):
            s = time.perf_counter()
            BoxLeastSquaresGPU(d["t"][i], d["y"][i], d["dy"][i]).power(
                d["periods"], d["durations"], method="parity",
                objective=objective)
            if i:
                lat.append(time.perf_counter() - s)
        plain = len(lat) / sum(lat)
        line["plain_gpu"] = {"value": round(plain, 2), "unit": "LC/s",
                             "p50_ms": round(np.percentile(lat, 50) * 1e3, 3),
                             "p99_ms": round(np.percentile(lat, 99) * 1e3, 3),
                             "curves": len(lat)}
        line["batched_over_plain"] = round(value / plain, 2)
        try:
            from astro_transit_gpu_amd.serving import serve_benchmark
            line["served"] = serve_benchmark(d, seconds=args.serve_seconds,
                                             batched_rate=value,
                                             objective=objective)
        except ImportError:
            line["served"] = None
        from astro_transit_gpu_amd.engine import release_host_cache
        # the server's freed staging buffers stay in torch's pinned cache;
        # the next legs stage gigabytes of their own
        release_host_cache()
        line["ragged"] = ragged_leg(args.ragged_lc, objective)
        release_host_cache()
        line["served_stream"] = serve5_leg(args.serve5_seconds, objective)
        ctx.run(batch, objective=objective, mode="parity", out=out)
        line["cpu_baseline"], line["parity_vs_oracle"] = cpu_baseline(
            d, args.cpu_sample, objective,
            gpu=(out.to_host(), batch.host_period_offsets))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
