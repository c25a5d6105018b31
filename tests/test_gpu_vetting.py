"""Vetting statistics, model and transit mask on the GPU (atg_bls_stats,
atg_bls_model) against the oracle's restatement of astropy's helpers."""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
RTOL = 1e-9   # FP64 sums in a different order, device sin/cos


@pytest.fixture(scope="module")
def ctx():
    from astro_transit_gpu_amd import Context
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    return Context(0)


def _close(a, b, rtol=RTOL):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    both_inf = np.isinf(a) & np.isinf(b)
    assert np.array_equal(np.isinf(a), np.isinf(b))
    np.testing.assert_allclose(a[~both_inf], b[~both_inf], rtol=rtol,
                               atol=1e-300)


def _compare(got, ref):
    assert np.array_equal(got["transit_times"], ref["transit_times"])
    assert np.array_equal(got["per_transit_count"], ref["per_transit_count"])
    _close(got["per_transit_log_likelihood"],
           ref["per_transit_log_likelihood"])
    for k in ("depth", "depth_phased", "depth_half", "depth_odd",
              "depth_even"):
        _close(got[k], ref[k])
    _close(got["harmonic_amplitude"], ref["harmonic_amplitude"])
    _close(got["harmonic_delta_log_likelihood"],
           ref["harmonic_delta_log_likelihood"], rtol=1e-8)


@pytest.mark.parametrize("with_err", [True, False])
def test_compute_stats_astropy_fixture(ctx, with_err):
    from astro_transit_gpu_amd import BoxLeastSquaresGPU
    c = O.load_golden(os.path.join(GOLD, "adata_oversample1000.npz"))
    dy = c["dy"] if with_err else None
    print("vetting fixture dy", with_err, flush=True)
    got = BoxLeastSquaresGPU(c["t"], c["y"], dy).compute_stats(2.0, 0.16,
                                                               0.5)
    assert list(got["per_transit_count"]) == [9, 7, 7, 7, 8]
    _compare(got, O.compute_stats(c["t"], c["y"], dy, 2.0, 0.16, 0.5))


def test_compute_stats_batch(ctx):
    """Candidates of several curves in one call: each curve's BLS best
    period plus an off-peak candidate; bit-identical on a second call."""
    from astro_transit_gpu_amd import DeviceBatch, synth
    d = synth.config2(n_lc=4, seed=41, n_points=5000, n_periods=600)
    b = DeviceBatch.from_host(d["t"], d["y"], d["dy"], d["periods"],
                              d["durations"], device=ctx.device)
    res = ctx.run(b, objective="likelihood")
    h = res.to_host()
    lc, per, dur, t0 = [], [], [], []
    for i in range(4):
        cur = res.curve(i, h)
        k = int(h["best_index"][i])
        for j in (k, (k + 137) % 600):
            lc.append(i)
            per.append(d["periods"][j])
            dur.append(cur["duration"][j])
            t0.append(cur["transit_time"][j])
    print("vetting batch candidates", len(lc), flush=True)
    got = ctx.compute_stats(b, lc, per, dur, t0)
    again = ctx.compute_stats(b, lc, per, dur, t0)
    for i in range(len(lc)):
        assert got[i]["status"] == 0
        ref = O.compute_stats(d["t"][lc[i]], d["y"][lc[i]], d["dy"][lc[i]],
                              per[i], dur[i], t0[i])
        _compare(got[i], ref)
        for k in ("per_transit_log_likelihood", "depth", "depth_half",
                  "harmonic_delta_log_likelihood"):
            assert np.array_equal(np.asarray(got[i][k]),
                                  np.asarray(again[i][k])), k
        # the fiducial depth is the BLS depth of that period's window
        cur = res.curve(lc[i], h)
        j = int(np.flatnonzero(d["periods"] == per[i])[0])
        np.testing.assert_allclose(got[i]["depth"][0], cur["depth"][j],
                                   rtol=1e-6)


def test_model_and_mask(ctx):
    from astro_transit_gpu_amd import BoxLeastSquaresGPU
    c = O.load_golden(os.path.join(GOLD, "adata_oversample1000.npz"))
    m = BoxLeastSquaresGPU(c["t"], c["y"], c["dy"])
    tm = np.linspace(-1.0, 11.0, 4001)
    _close(m.model(tm, 2.0, 0.16, 0.5),
           O.model(c["t"], c["y"], c["dy"], tm, 2.0, 0.16, 0.5), rtol=1e-12)
    assert np.array_equal(m.transit_mask(tm, 2.0, 0.16, 0.5),
                          O.transit_mask(tm, 2.0, 0.16, 0.5))
    assert np.array_equal(m.transit_mask(c["t"], 2.0, 0.16, 0.5),
                          O.transit_mask(c["t"], 2.0, 0.16, 0.5))


def test_stats_status(ctx):
    from astro_transit_gpu_amd import BoxLeastSquaresGPU, DeviceBatch
    t = np.arange(0.0, 10.0, 0.5)   # no point within 0.05 d of 0.25 + k
    y = np.ones_like(t)
    b = DeviceBatch.from_host([t], [y], None, [1.0], [0.1],
                              device=ctx.device)
    st = ctx.compute_stats(b, [0, 0], [1.0, 1.0], [0.1, 0.6], [0.25, 0.25],
                           max_transits=4)
    assert st[0]["status"] == 1
    assert st[1]["status"] == 2       # ten transits, room for four
    with pytest.raises(ValueError):
        BoxLeastSquaresGPU(t, y).compute_stats(1.0, 0.1, 0.25)
