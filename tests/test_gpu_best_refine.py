"""Best-only screening (atg_bls_run_best) and the batched refiner
re-search, on the GPU, against the full periodogram run and the per-target
API call they replace."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIELDS = ("power", "depth", "depth_err", "duration", "transit_time",
          "depth_snr", "log_likelihood")


@pytest.fixture(scope="module")
def ctx():
    from astro_transit_gpu_amd import Context
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    return Context(0)


def _mixed_batch(ctx, n_ok, seed):
    """n_ok synthetic curves of the 1,000-curve workload's kind (periods up
    to 13.5 d, so phase tiles too) plus a NaN curve, a curve with a zero
    dy, a curve whose durations exceed its shortest period and a curve
    whose points all share one time (no window qualifies: power -inf)."""
    from astro_transit_gpu_amd import DeviceBatch, synth
    d = synth.config2(n_lc=n_ok, seed=seed, n_points=5000, n_periods=500)
    ts, ys, dys = list(d["t"]), list(d["y"]), list(d["dy"])
    pers = [d["periods"]] * n_ok
    durs = [d["durations"]] * n_ok
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0, 20, 800))
    y = (1 + 1e-3 * rng.standard_normal(800)).astype(np.float32)
    dy = np.full(800, 1e-3, np.float32)
    y_nan = y.copy()
    y_nan[5] = np.nan
    dy0 = dy.copy()
    dy0[9] = 0.0
    ts += [t, t, t, np.full(300, 4.25)]
    ys += [y_nan, y, y, y[:300]]
    dys += [dy, dy0, dy, dy[:300]]
    pers += [np.linspace(1, 5, 64), np.linspace(1, 5, 64),
             np.linspace(0.05, 5, 64), np.linspace(1, 5, 64)]
    durs += [np.array([0.1, 0.2])] * 4
    return DeviceBatch.from_host(ts, ys, dys, pers, durs, device=ctx.device)


def _check_best_equals_full(ctx, b, objective, mode):
    full = ctx.run(b, objective=objective, mode=mode).to_host()
    best = ctx.run_best(b, objective=objective, mode=mode).to_host()
    per = b.periods.cpu().numpy()
    off = b.host_period_offsets
    for c in range(b.n_lc):
        i = int(full["best_index"][c])
        assert best["best_index"][c] == i, c
        assert best["status"][c] == full["status"][c], c
        q = off[c] + i
        assert best["period"][c] == per[q], c
        for k in FIELDS:
            a, e = best[k][c], full[k][q]
            assert (a == e) or (np.isnan(a) and np.isnan(e)), (c, k, a, e)
    return full, best


@pytest.mark.parametrize("mode", ["parity", "fast"])
@pytest.mark.parametrize("objective", ["likelihood", "snr"])
def test_best_only_equals_full_run(ctx, mode, objective):
    """Every field of the best-only run equals the full run's at its
    best_index, bit for bit, for good and rejected curves; a second, smaller
    batch on the same context reuses the workspace."""
    b = _mixed_batch(ctx, 6, 31)
    full, best = _check_best_equals_full(ctx, b, objective, mode)
    assert list(full["status"][6:]) == [3, 4, 2, 0]
    assert best["power"][9] == -np.inf and best["depth"][9] == 0.0
    assert best["transit_time"][9] == 4.25
    assert np.isnan(best["power"][6:9]).all()
    small = _mixed_batch(ctx, 2, 32)
    _check_best_equals_full(ctx, small, objective, mode)


def test_best_only_parity_against_oracle(ctx):
    """The winners of a best-only parity run are the oracle's."""
    from astro_transit_gpu_amd import DeviceBatch, synth
    d = synth.config2(n_lc=3, seed=33, n_points=4000, n_periods=400)
    b = DeviceBatch.from_host(d["t"], d["y"], d["dy"], d["periods"],
                              d["durations"], device=ctx.device)
    best = ctx.run_best(b, objective="likelihood").to_host()
    for c in range(3):
        ref = O.power(d["t"][c], d["y"][c], d["dy"][c], d["periods"],
                      d["durations"], "likelihood")
        ib = int(np.argmax(ref["power"]))
        assert best["best_index"][c] == ib
        assert best["period"][c] == d["periods"][ib]
        assert abs(best["power"][c] - ref["power"][ib]) <= \
            1e-6 * abs(ref["power"][ib])
        assert abs(best["transit_time"][c] - ref["transit_time"][ib]) <= 1e-6
        assert best["duration"][c] == ref["duration"][ib]


def test_best_only_workspace_size(ctx):
    b = _mixed_batch(ctx, 2, 34)
    extra = ctx.workspace_bytes_best(b) - ctx.workspace_bytes(b)
    assert extra >= 7 * 8 * b.n_periods and extra < 7 * 8 * b.n_periods + 256


def test_refine_targets_equals_per_target_power():
    """The batched re-search returns, target by target, the rows the
    reference's refiner builds from BoxLeastSquaresGPU(...).power(...)
    .top_candidates (search/refiner.py:118-135); a rejected target gets an
    error row."""
    from astro_transit_gpu_amd import BoxLeastSquaresGPU, synth
    from astro_transit_gpu_amd.compat import refine_targets
    rng = np.random.default_rng(35)
    targets = {}
    for k, tid in enumerate((1001, 2002, 3003)):
        t = np.sort(rng.uniform(0, 27, 3000))
        y = 1 + 1e-3 * rng.standard_normal(t.size)
        y = synth.inject_box(t, y, rng.uniform(1, 9), rng.uniform(0, 1),
                             0.1, 3e-3)
        dy = None if k == 1 else np.full(t.size, 1e-3)
        targets[tid] = (t, y, dy)
    t_bad = np.sort(rng.uniform(0, 27, 500))
    y_bad = np.ones(500)
    y_bad[3] = np.nan
    targets[4004] = (t_bad, y_bad, None)
    rows = refine_targets(targets, n_periods=2000)
    periods = np.linspace(0.5, 20.0, 2000)
    durs = np.linspace(0.01, 0.2, 5)
    for tid in (1001, 2002, 3003):
        t, y, dy = targets[tid]
        res = BoxLeastSquaresGPU(t, y, dy=dy).power(periods, durs)
        want = [(float(c.period), float(c.t0), float(c.depth),
                 float(c.duration), float(c.power))
                for c in res.top_candidates[:5]]
        got = [(r["period"], r["t0"], r["depth"], r["duration"], r["power"])
               for r in rows if r["tic_id"] == tid]
        assert got == want, tid
        assert all(r["status"] == "ok" and r["n_points"] == t.size
                   for r in rows if r["tic_id"] == tid)
        # anchored on the FP64 oracle (snr objective, the API's default):
        # the first row is an oracle maximum within the fast mode's drift
        # and every row's power is the oracle's at its period within 1e-6
        ref = O.power(t, y, dy, periods, durs, "snr")["power"]
        for r in (x for x in rows if x["tic_id"] == tid):
            i = int(np.nonzero(periods == r["period"])[0][0])
            assert abs(r["power"] - ref[i]) <= 1e-6 * abs(ref[i]), (tid, i)
        assert got[0][4] >= ref.max() * (1 - 1e-6), tid
    bad = [r for r in rows if r["tic_id"] == 4004]
    assert len(bad) == 1 and bad[0]["status"] == "error" and bad[0]["error"]
