"""Autotuner for the main kernel's tiling.

Candidates (each a legal launch on gfx950):
  * lds_bins          bins of the LDS histogram per pass (16 B each): the
                      number of periods folded together per pass over the
                      points and the threshold of the phase-tiled path;
  * periods_per_item  periods per work item of the persistent queue;
  * wg_per_cu         persistent workgroups per CU, limited by 160 KiB of LDS
                      and by the kernel's register use (4 waves per SIMD).

Each candidate runs the main kernel on a sample of the workload and is timed
with HIP events on the context's stream; the fastest is applied and cached
per workload signature (curve count bucket, mean points, mean bins per
period, durations, mode, objective, and whether the weights are uniform:
uniform-weight curves take the packed 8-byte bins, the others twice the
LDS per bin) in a JSON file next to the library.  Every tiling gives the
same results bit for bit (integer bins; tested), so the choice only
changes the time.
"""
import json
import os
import time

import numpy as np
import torch

from .build import LIB_DIR
from .engine import Context, DeviceBatch

CACHE = os.path.join(LIB_DIR, "tuning.json")
LDS_PER_CU = 160 * 1024
STATIC_LDS = 256           # the main kernel's static LDS
MAX_WAVES_PER_SIMD = 4     # register-limited
# bumped whenever the kernel's speed profile changes, so that cached
# tilings of an older build are not reused
KERNEL_GENERATION = 7


def candidates(lds_options=(2048, 2432, 2540, 3072, 3392, 4096, 5120, 6144),
               ppi_options=(4, 8, 16, 24)):
    out = []
    for lds in lds_options:
        per_wg = lds * 16 + STATIC_LDS
        wg_max = min(LDS_PER_CU // per_wg, MAX_WAVES_PER_SIMD)
        for wg in range(max(1, wg_max - 1), wg_max + 1):
            for ppi in ppi_options:
                out.append(dict(lds_bins=lds, periods_per_item=ppi,
                                wg_per_cu=wg))
    return out


def signature(n_lc, n_points_mean, bins_mean, n_dur, mode,
              objective="likelihood", uniform_weights=True):
    """Workload key: buckets wide enough that nearby shapes share a
    tiling."""
    return "g%d_lc%d_n%d_b%d_d%d_%s_%s_%s" % (
        KERNEL_GENERATION,
        int(np.log2(max(n_lc, 1))), int(np.log2(max(n_points_mean, 1))),
        int(np.log2(max(bins_mean, 1))), n_dur, mode, objective,
        "uw" if uniform_weights else "vw")


def uniform_weights(dys):
    """True when every curve has one dy value (or none): the curves whose
    bins the kernel packs (barring extreme outliers)."""
    if dys is None:
        return True
    return all(d is None or np.size(d) == 0 or
               np.all(np.asarray(d) == np.asarray(d).ravel()[0])
               for d in dys)


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(cache):
    try:
        with open(CACHE + ".tmp", "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(CACHE + ".tmp", CACHE)
    except OSError:
        pass


def autotune(ctx: Context, ts, ys, dys, periods, durations,
             objective="likelihood", mode="parity", oversample=10,
             sample=64, reps=2, use_cache=True, verbose=False):
    """Choose and apply the fastest tiling for curves like ``ts`` with the
    given grid.  Returns (tuning dict, timings list)."""
    n = min(sample, len(ts))
    per = periods if not isinstance(periods, (list, tuple)) else periods[0]
    dur = durations if not isinstance(durations, (list, tuple)) else \
        durations[0]
    bd = np.min(dur) / oversample
    bins_mean = float(np.mean(np.ceil(np.asarray(per) / bd))) + oversample
    sig = signature(len(ts), np.mean([np.size(t) for t in ts]), bins_mean,
                    np.size(dur), mode, objective,
                    uniform_weights(None if dys is None else dys[:n]))
    cache = _load() if use_cache else {}
    if sig in cache:
        ctx.set_tuning(**cache[sig])
        return cache[sig], []
    pick = lambda a: a if not isinstance(a, (list, tuple)) else a[:n]
    b = DeviceBatch.from_host(ts[:n], ys[:n], None if dys is None else
                              dys[:n], pick(periods), pick(durations),
                              device=ctx.device)
    out = ctx.run(b, objective=objective, mode=mode, oversample=oversample)
    timings = []
    for cand in candidates():
        try:
            ctx.set_tuning(**cand)
            ctx.run(b, objective=objective, mode=mode, oversample=oversample,
                    out=out)
        except Exception:
            continue
        ctx.timing(True, True)
        for _ in range(reps):
            ctx.run(b, objective=objective, mode=mode, oversample=oversample,
                    out=out, sync=False)
        torch.cuda.synchronize(ctx.device)
        _, ms = ctx.timing(False, True)
        timings.append((ms, cand))
        if verbose:
            print("autotune %s %.3f ms" % (cand, ms), flush=True)
    timings.sort(key=lambda x: x[0])
    best = timings[0][1]
    ctx.set_tuning(**best)
    cache[sig] = best
    if use_cache:
        _store(cache)
    return best, timings
