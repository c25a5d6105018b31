"""GPU parity of the HIP path against the astropy golden vectors and the
FP64 oracle (runs through libatg_bls.so)."""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from gpu_util import (GOLDEN, T0_ATOL, assert_parity, power_rel_err,
                      window_agreement)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from astro_transit_gpu_amd import Context
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    return Context(0)


def run_single(ctx, c, mode="parity"):
    from astro_transit_gpu_amd import DeviceBatch
    b = DeviceBatch.from_host([c["t"]], [c["y"]],
                              None if c["dy"] is None else [c["dy"]],
                              c["periods"], c["durations"], device=ctx.device)
    res = ctx.run(b, objective=c["objective"], mode=mode,
                  oversample=c["oversample"])
    h = res.to_host()
    return h


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_parity_vs_astropy_golden(ctx, path):
    c = O.load_golden(path)
    print("case", os.path.basename(path), flush=True)
    h = run_single(ctx, c)
    assert int(h["status"][0]) == 0
    rel = assert_parity(h, c["out"], os.path.basename(path))
    agree = window_agreement(h, c["out"])
    print("power rel err %.3e, window agreement %.4f" % (rel, agree))
    if np.isfinite(c["out"]["power"]).any():
        ib = int(np.argmax(c["out"]["power"]))
        assert int(h["best_index"][0]) == ib
        # the best period is recomputed in the oracle's operation order:
        # every field equals astropy's bit for bit
        for k in c["out"]:
            assert h[k][ib] == c["out"][k][ib], k


def test_ragged_batch_matches_golden(ctx):
    """All ragged fixtures of one objective in a single CSR launch."""
    from astro_transit_gpu_amd import DeviceBatch
    for obj in ("likelihood", "snr"):
        cases = [O.load_golden(p) for p in GOLDEN
                 if "ragged_" in p]
        cases = [c for c in cases if c["objective"] == obj]
        b = DeviceBatch.from_host([c["t"] for c in cases],
                                  [c["y"] for c in cases],
                                  [c["dy"] for c in cases],
                                  [c["periods"] for c in cases],
                                  [c["durations"] for c in cases],
                                  device=ctx.device)
        res = ctx.run(b, objective=obj, mode="parity")
        h = res.to_host()
        for i, c in enumerate(cases):
            assert_parity(res.curve(i, h), c["out"], "ragged %d" % i)
            assert int(h["best_index"][i]) == int(np.argmax(c["out"]["power"]))


def test_parity_bit_reproducible(ctx):
    c = O.load_golden([p for p in GOLDEN if p.endswith("cfg1_likelihood.npz")][0])
    a = run_single(ctx, c)
    b = run_single(ctx, c)
    for k in ("power", "depth", "depth_err", "duration", "transit_time",
              "depth_snr", "log_likelihood"):
        assert np.array_equal(a[k], b[k], equal_nan=True), k


def test_batched_equals_single(ctx):
    """A curve gives bit-identical results alone and inside a batch."""
    from astro_transit_gpu_amd import DeviceBatch, synth
    d = synth.config2(n_lc=6, seed=21, n_points=4000, n_periods=400)
    b = DeviceBatch.from_host(d["t"], d["y"], d["dy"], d["periods"],
                              d["durations"], device=ctx.device)
    hb = ctx.run(b, objective="likelihood").to_host()
    res_b = ctx.run(b, objective="likelihood")
    for i in range(6):
        s = DeviceBatch.from_host([d["t"][i]], [d["y"][i]], [d["dy"][i]],
                                  d["periods"], d["durations"],
                                  device=ctx.device)
        hs = ctx.run(s, objective="likelihood").to_host()
        bi = res_b.curve(i, hb)
        for k in bi:
            assert np.array_equal(bi[k], hs[k]), (i, k)
        assert hb["best_index"][i] == hs["best_index"][0]


def test_against_oracle_fp32_batch(ctx):
    """FP32 y/dy inputs (the 1,000-curve workload's format), several
    curves, against the C oracle."""
    from astro_transit_gpu_amd import DeviceBatch, synth
    d = synth.config2(n_lc=4, seed=22, n_points=6000, n_periods=800)
    b = DeviceBatch.from_host(d["t"], d["y"], d["dy"], d["periods"],
                              d["durations"], device=ctx.device)
    res = ctx.run(b, objective="likelihood")
    h = res.to_host()
    for i in range(4):
        ref = O.power(d["t"][i], d["y"][i], d["dy"][i], d["periods"],
                      d["durations"], "likelihood")
        rel = assert_parity(res.curve(i, h), ref, "curve %d" % i)
        print("curve %d rel %.2e agree %.4f" % (
            i, rel, window_agreement(res.curve(i, h), ref)))


def test_status_codes(ctx):
    from astro_transit_gpu_amd import DeviceBatch
    t = np.linspace(0, 10, 500)
    y = 1 + 1e-3 * np.random.default_rng(0).standard_normal(500)
    y_nan = y.copy()
    y_nan[7] = np.nan
    dy_bad = np.full(500, 1e-3)
    dy_bad[3] = 0.0
    b = DeviceBatch.from_host([t, t, t, t], [y, y_nan, y, y],
                              [np.full(500, 1e-3), np.full(500, 1e-3),
                               dy_bad, np.full(500, 1e-3)],
                              [np.linspace(1, 3, 20)] * 3 +
                              [np.linspace(0.05, 3, 20)],
                              [0.1], device=ctx.device)
    h = ctx.run(b, objective="likelihood").to_host()
    assert list(h["status"]) == [0, 3, 4, 2]
    assert np.isnan(h["power"][20:]).all()
    assert np.isfinite(h["power"][:20]).all()


@pytest.mark.parametrize("mode", ["parity", "fast"])
def test_bin_count_limit(ctx, mode):
    """A grid whose bins would overflow 32-bit indices (a 1e-9 d shortest
    duration: 5e10 bins at P = 5 d) is rejected as too many bins, with NaN
    results, while the other curve of the batch runs normally."""
    from astro_transit_gpu_amd import DeviceBatch
    rng = np.random.default_rng(5)
    t = np.sort(rng.uniform(0, 20, 500))
    y = 1 + 1e-3 * rng.standard_normal(500)
    dy = np.full(500, 1e-3)
    per = np.linspace(1, 5, 20)
    b = DeviceBatch.from_host([t, t], [y, y], [dy, dy], [per, per],
                              [np.array([0.1, 0.2]), np.array([1e-9, 0.1])],
                              device=ctx.device)
    h = ctx.run(b, objective="likelihood", mode=mode).to_host()
    assert list(h["status"]) == [0, 5]
    assert np.isfinite(h["power"][:20]).all()
    assert np.isnan(h["power"][20:]).all()


@pytest.mark.parametrize("name", ["cfg1_small_likelihood", "cfg1_small_snr",
                                  "hetero_dy", "unsorted_t"])
def test_segmented_bins_bit_identical(name):
    """Periods that do not fit the LDS histogram go through the
    phase-tiled path; with a tiny LDS budget every period does, and the
    outputs must equal the in-LDS path bit for bit.  Time-sorted curves
    take the ranged tiles, unsorted_t the refolding ones."""
    from astro_transit_gpu_amd import Context
    c = O.load_golden([p for p in GOLDEN if p.endswith(name + ".npz")][0])
    big = Context(0)
    small = Context(0, tuning=dict(lds_bins=300))
    a = run_single(big, c)
    b = run_single(small, c)
    assert int(a["status"][0]) == 0 and int(b["status"][0]) == 0
    for k in ("power", "depth", "depth_err", "duration", "transit_time",
              "depth_snr", "log_likelihood", "best_index"):
        assert np.array_equal(a[k], b[k]), k
    ib = int(np.argmax(c["out"]["power"]))
    for k in c["out"]:
        assert b[k][ib] == c["out"][k][ib], k


@pytest.mark.parametrize("name", ["cfg1_small_likelihood", "hetero_dy",
                                  "unsorted_t"])
def test_fast_mode_tiling_independent(name):
    """Fast mode folds with its own rule in the phase tiles too, so its
    (drifting) results do not depend on which periods the LDS budget sends
    there."""
    from astro_transit_gpu_amd import Context
    c = O.load_golden([p for p in GOLDEN if p.endswith(name + ".npz")][0])
    a = run_single(Context(0), c, mode="fast")
    b = run_single(Context(0, tuning=dict(lds_bins=300)), c, mode="fast")
    for k in ("power", "depth", "depth_err", "duration", "transit_time",
              "depth_snr", "log_likelihood", "best_index"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("name", ["cfg1_likelihood", "cfg1_snr",
                                  "hetero_dy"])
def test_fast_mode_drift(ctx, name):
    """Fast mode (FMA fold without the bin-edge exactness check, no
    refinement) stays within 1e-6 relative power of astropy with the same
    best period, and is run-to-run bit-identical (int64 bins)."""
    c = O.load_golden([p for p in GOLDEN if p.endswith(name + ".npz")][0])
    a = run_single(ctx, c, mode="fast")
    b = run_single(ctx, c, mode="fast")
    for k in ("power", "depth", "transit_time", "duration"):
        assert np.array_equal(a[k], b[k]), k
    rel = assert_parity(a, c["out"], name + " fast")
    print("fast drift %.3e window agreement %.4f" % (
        rel, window_agreement(a, c["out"])))


def test_hipgraph_replays_match_eager():
    """The launch sequence captured as a HIP graph gives the eager result on
    every replay, with eager work and synchronisations in between."""
    from astro_transit_gpu_amd import (Context, DeviceBatch, DeviceResult,
                                       synth)
    ctx = Context(0)  # its own workspace, sized by this test's batches
    d = synth.config2(n_lc=8, seed=23, n_points=5000, n_periods=600)
    b = DeviceBatch.from_host(d["t"], d["y"], d["dy"], d["periods"],
                              d["durations"], device=ctx.device)
    eager = ctx.run(b, objective="likelihood").to_host()
    out = DeviceResult.empty(b.n_periods, b.n_lc, b.host_period_offsets,
                             ctx.device)
    g = ctx.capture(b, out, objective="likelihood")
    for rep in range(4):
        for k in ("power", "depth"):
            getattr(out, k).fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        h = out.to_host()
        for k in h:
            assert np.array_equal(h[k], eager[k]), (rep, k)
        ctx.run(b, objective="snr")  # unrelated eager work between replays
        torch.cuda.synchronize()
    # the graph pins the workspace: a batch needing more is refused until
    # the graph is released, then runs
    from astro_transit_gpu_amd._abi import MemoryBudgetError
    d2 = synth.config2(n_lc=16, seed=24, n_points=6000, n_periods=600)
    b2 = DeviceBatch.from_host(d2["t"], d2["y"], d2["dy"], d2["periods"],
                               d2["durations"], device=ctx.device)
    assert ctx.workspace_bytes(b2) > ctx.workspace_bytes(b)
    with pytest.raises(MemoryBudgetError, match="pinned"):
        ctx.run(b2, objective="likelihood")
    g.release()
    assert int(ctx.run(b2, objective="likelihood").status.cpu().max()) == 0


@pytest.mark.parametrize("mode", ["parity", "fast"])
def test_long_baseline_parity(ctx, mode):
    """Two 200,000-point curves over 360 d with 13 gaps (the precision-study
    workload) on every 100th of its 50,000 periods 1-100 d: the long periods
    need up to 25,010 bins and take the phase-tiled path.  Parity mode holds
    the parity bar with every best-period field equal to the oracle's; fast
    mode stays within its stated drift (points on bin edges of the 2-min grid
    may change bin): its best period has the oracle's maximum power within
    1e-6 and the oracle's T0 and duration there, fewer than 1% of periods
    move beyond 1e-6 and none beyond 10%."""
    from astro_transit_gpu_amd import DeviceBatch, synth
    rng = np.random.default_rng(404)
    curves = [synth.config4_curve(rng) for _ in range(2)]
    periods = np.geomspace(1.0, 100.0, 50000)[::100]
    durs = synth.DURATIONS_6
    print("long baseline", mode, flush=True)
    b = DeviceBatch.from_host([c[0] for c in curves], [c[1] for c in curves],
                              [c[2] for c in curves], periods, durs,
                              device=ctx.device)
    res = ctx.run(b, objective="likelihood", mode=mode)
    h = res.to_host()
    assert list(h["status"]) == [0, 0]
    for i, c in enumerate(curves):
        ref = O.power(c[0], c[1], c[2], periods, durs, "likelihood")
        g = res.curve(i, h)
        ib = int(np.argmax(ref["power"]))
        if mode == "parity":
            rel = assert_parity(g, ref, "curve %d" % i)
            for k in ref:
                assert g[k][ib] == ref[k][ib], (i, k)
        else:
            r = np.abs(g["power"] / ref["power"] - 1.0)
            rel = float(r.max())
            assert np.mean(r > 1e-6) < 0.01 and rel < 0.1, (i, rel)
            gb = int(h["best_index"][i])
            assert ref["power"][gb] >= ref["power"][ib] * (1 - 1e-6), i
            assert abs(g["transit_time"][ib] - ref["transit_time"][ib]) <= \
                T0_ATOL, i
            assert g["duration"][ib] == ref["duration"][ib], i
        print("%s curve %d rel %.2e agree %.4f" % (
            mode, i, rel, window_agreement(g, ref)))


def test_long_plateau_of_tied_periods(ctx):
    """Curve 45 of the 1,000-curve workload (seed 2): 59 neighbouring
    periods share the best window's points, so their integer powers are
    equal; the oracle's FP64 rounding puts its maximum at index 4944, deep
    inside the plateau.  Every plateau member must be recomputed in the
    oracle's order for the best period and its fields to be the oracle's."""
    from astro_transit_gpu_amd import DeviceBatch, synth
    d = synth.config2(n_lc=46, seed=2)
    t, y, dy = d["t"][45], d["y"][45], d["dy"][45]
    ref = O.power(t, y, dy, d["periods"], d["durations"], "likelihood")
    mx = ref["power"].max()
    assert np.sum(ref["power"] >= mx * (1 - 1e-9)) > 16
    print("plateau curve: oracle best", int(np.argmax(ref["power"])),
          flush=True)
    b = DeviceBatch.from_host([t], [y], [dy], d["periods"], d["durations"],
                              device=ctx.device)
    h = ctx.run(b, objective="likelihood").to_host()
    assert_parity(h, ref, "plateau")
    ib = int(np.argmax(ref["power"]))
    assert int(h["best_index"][0]) == ib
    for k in ref:
        assert h[k][ib] == ref[k][ib], k


def test_refinement_queue_beyond_one_workgroup(ctx):
    """Flat flux: every window has depth 0, so all 1,500 periods tie at
    power 0 and the refinement queues every one of them (more than one
    workgroup's threads).  The oracle's answer (first period,
    all fields) must come back, and a second curve in the same batch must
    be unaffected."""
    from astro_transit_gpu_amd import DeviceBatch, synth
    rng = np.random.default_rng(77)
    t = np.sort(rng.uniform(0.0, 20.0, 800))
    flat = np.ones_like(t)
    y2 = synth.inject_box(t, 1.0 + 1e-3 * rng.standard_normal(t.size), 2.3,
                          0.7, 0.1, 3e-3)
    dy = np.full(t.size, 1e-3)
    periods = np.geomspace(0.6, 9.0, 1500)
    durs = np.array([0.05, 0.1])
    print("flat curve: %d tied periods" % periods.size, flush=True)
    b = DeviceBatch.from_host([t, t], [flat, y2], [dy, dy], periods, durs,
                              device=ctx.device)
    res = ctx.run(b, objective="likelihood")
    h = res.to_host()
    for i, y in enumerate((flat, y2)):
        ref = O.power(t, y, dy, periods, durs, "likelihood")
        g = res.curve(i, h)
        assert_parity(g, ref, "curve %d" % i)
        ib = int(np.argmax(ref["power"]))
        assert int(h["best_index"][i]) == ib
        for k in ref:
            assert g[k][ib] == ref[k][ib], (i, k)


def plateau_beyond_1024():
    """A sparse curve (20 points over 27 d) whose best window holds only one
    deep point for nearly every period: the exact (integer) power is the
    same on ~17,800 of 20,000 periods, and the oracle's FP64 rounding gives
    three slightly different values.  The grid drops every period of the
    largest value among the first 1,500 plateau members, so the oracle's
    maximum lies thousands of plateau members in."""
    rng = np.random.default_rng(2024)
    n = 20
    t = np.sort(rng.uniform(0.0, 27.0, n))
    y = 1.0 + 1e-3 * rng.standard_normal(n)
    y[n // 2] = 0.9
    dy = np.full(n, 1e-3)
    dense = np.geomspace(0.5, 13.5, 20000)
    durs = np.array([0.04, 0.08, 0.12, 0.16, 0.20])
    p = O.power(t, y, dy, dense, durs, "likelihood")["power"]
    near = np.nonzero(p >= p.max() * (1 - 1e-9))[0]
    keep = np.ones(p.size, bool)
    keep[:near[1500]] &= p[:near[1500]] != p.max()
    return t, y, dy, dense[keep], durs


def test_refinement_beyond_1024_tied_periods(ctx):
    """Every near-tied period is recomputed in the oracle's order, however
    many there are: here the oracle's maximum is the ~3,900th member of a
    ~17,800-period plateau, so a refinement limited to the first 1,024
    members would return another period."""
    from astro_transit_gpu_amd import DeviceBatch
    t, y, dy, periods, durs = plateau_beyond_1024()
    ref = O.power(t, y, dy, periods, durs, "likelihood")
    mx = ref["power"].max()
    near = np.nonzero(ref["power"] >= mx * (1 - 1e-9))[0]
    ib = int(np.argmax(ref["power"]))
    rank = int(np.searchsorted(near, ib))
    print("plateau %d periods, oracle best %d (plateau member %d)" % (
        near.size, ib, rank), flush=True)
    assert rank > 1024
    b = DeviceBatch.from_host([t], [y], [dy], periods, durs,
                              device=ctx.device)
    h = ctx.run(b, objective="likelihood").to_host()
    assert int(h["status"][0]) == 0
    assert_parity(h, ref, "plateau beyond 1024")
    assert int(h["best_index"][0]) == ib
    for k in ref:
        assert h[k][ib] == ref[k][ib], k


OUTLIER_CASES = [(1e4, False, 1.0), (1e6, False, 1.0), (1e6, False, -1.0),
                 (1e8, False, 1.0), (1e4, True, 1.0), (1e6, True, 1.0)]


@pytest.mark.parametrize("sigmas,hetero,sign", OUTLIER_CASES,
                         ids=lambda v: str(v))
def test_flux_outlier_parity(ctx, sigmas, hetero, sign):
    """A 15,000-point curve with one flux outlier of 1e4 to 1e8 sigma (above
    or below the curve), uniform dy (packed bins unless the outlier would
    coarsen their flux quantum) or dy varied per point (unpacked bins):
    every period within the parity bar of the FP64 oracle, the best period
    and all its fields the oracle's."""
    from astro_transit_gpu_amd import DeviceBatch, synth
    rng = np.random.default_rng(int(np.log10(sigmas)) * 10 + hetero +
                                (2 if sign < 0 else 0))
    t, y, dy = synth.outlier_curve(rng, 15000, sigmas, hetero, sign)
    periods = np.geomspace(0.5, 13.5, 1000)
    durs = synth.DURATIONS_5
    print("outlier %g sigma, hetero %s, sign %+d" % (sigmas, hetero, sign),
          flush=True)
    ref = O.power(t, y, dy, periods, durs, "likelihood")
    b = DeviceBatch.from_host([t], [y], [dy], periods, durs,
                              device=ctx.device)
    h = ctx.run(b, objective="likelihood").to_host()
    assert int(h["status"][0]) == 0
    rel = assert_parity(h, ref, "outlier")
    print("max rel power err %.3e" % rel, flush=True)
    ib = int(np.argmax(ref["power"]))
    assert int(h["best_index"][0]) == ib
    for k in ref:
        assert h[k][ib] == ref[k][ib], k
