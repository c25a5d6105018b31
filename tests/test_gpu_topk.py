"""Device top-k (atg_bls_top_k) against scipy's find_peaks rule."""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from gpu_util import GOLDEN
from topk_model import (local_maxima, reference_top_k, rounds_top_k,
                        sample_arrays)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from astro_transit_gpu_amd import Context
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    return Context(0)


def _run_top_k(ctx, arrays, k, distance, status=None):
    """One ragged call over power arrays given directly."""
    from astro_transit_gpu_amd import DeviceBatch, DeviceResult
    dev = ctx.device
    n = len(arrays)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(a) for a in arrays])
    z64 = torch.zeros(n, dtype=torch.float64, device=dev)
    b = DeviceBatch(t=z64, y=z64.float(), dy=None,
                    lc_offsets=torch.arange(n + 1, device=dev),
                    periods=torch.arange(int(off[-1]), dtype=torch.float64,
                                         device=dev),
                    period_offsets=torch.as_tensor(off, device=dev),
                    durations=torch.ones(n, dtype=torch.float64, device=dev),
                    dur_offsets=torch.arange(n + 1, device=dev),
                    host_period_offsets=off)
    r = DeviceResult.empty(int(off[-1]), n, off, dev)
    r.power.copy_(torch.as_tensor(np.concatenate(arrays), device=dev))
    r.status.copy_(torch.as_tensor(np.zeros(n, np.int32) if status is None
                                   else status, device=dev))
    idx, cnt = ctx.top_k(b, r, k=k, distance=distance)
    return idx.cpu().numpy(), cnt.cpu().numpy()


@pytest.mark.parametrize("k,distance", [(5, 10), (3, 1), (8, 4)])
def test_top_k_matches_rule(ctx, k, distance):
    rng = np.random.default_rng(30 + k)
    arrays = sample_arrays(rng, 350)
    print("top-k arrays", len(arrays), "k", k, "distance", distance,
          flush=True)
    idx, cnt = _run_top_k(ctx, arrays, k, distance)
    n_scipy = 0
    for c, x in enumerate(arrays):
        got = idx[c, :cnt[c]]
        assert np.all(idx[c, cnt[c]:] == -1)
        assert np.array_equal(got, rounds_top_k(x, k, distance)), c
        pk = local_maxima(x)
        if len(np.unique(x[pk])) == len(pk):
            assert np.array_equal(got, reference_top_k(x, k, distance)), c
            n_scipy += 1
    assert n_scipy > 100


def test_top_k_status_and_periodograms(ctx):
    """Real periodograms (astropy's outputs in the golden fixtures) in one
    call, one curve flagged bad."""
    arrays = []
    for p in GOLDEN:
        c = O.load_golden(p)
        pw = c["out"]["power"]
        if pw.size >= 3 and np.isfinite(pw).all():
            arrays.append(pw)
    status = np.zeros(len(arrays), np.int32)
    status[1] = 3
    idx, cnt = _run_top_k(ctx, arrays, 5, 10, status)
    assert cnt[1] == 0 and np.all(idx[1] == -1)
    for c, x in enumerate(arrays):
        if c == 1:
            continue
        assert np.array_equal(idx[c, :cnt[c]], rounds_top_k(x, 5, 10)), c


def test_api_top_candidates(ctx):
    """BoxLeastSquaresGPU.power's top candidates (device top-k) equal the
    host get_top_k_candidates on the same periodogram."""
    from astro_transit_gpu_amd import BoxLeastSquaresGPU, get_top_k_candidates
    c = O.load_golden([p for p in GOLDEN if p.endswith("cfg1_snr.npz")][0])
    r = BoxLeastSquaresGPU(c["t"], c["y"], c["dy"]).power(
        c["periods"], c["durations"], method="parity", objective="snr")
    host = get_top_k_candidates(dict(power=r.power, periods=r.periods,
                                     all_t0s=r.transit_time,
                                     all_durs=r.duration,
                                     all_depths=r.depth), k=5)
    assert len(r.top_candidates) == len(host) == 5
    for a, b in zip(r.top_candidates, host):
        assert (a.period, a.t0, a.duration, a.depth, a.power) == (
            b["period"], b["t0"], b["duration"], b["depth"], b["power"])
    assert r.top_candidates[0].period == r.best_period
    # anchored on astropy's periodogram of the same curve: the host rule on
    # the golden power array picks the same five periods, and each
    # candidate's power is astropy's at that period within the parity bar
    ref = c["out"]
    gold = get_top_k_candidates(dict(power=ref["power"], periods=c["periods"],
                                     all_t0s=ref["transit_time"],
                                     all_durs=ref["duration"],
                                     all_depths=ref["depth"]), k=5)
    assert [g["period"] for g in gold] == [a.period for a in r.top_candidates]
    for a in r.top_candidates:
        i = int(np.nonzero(c["periods"] == a.period)[0][0])
        assert abs(a.power - ref["power"][i]) <= 1e-6 * abs(ref["power"][i])


def test_candidates_batch(ctx):
    """Context.candidates over a batch: the reference's dicts per curve."""
    from astro_transit_gpu_amd import DeviceBatch, synth
    d = synth.config2(n_lc=5, seed=31, n_points=4000, n_periods=700)
    b = DeviceBatch.from_host(d["t"], d["y"], d["dy"], d["periods"],
                              d["durations"], device=ctx.device)
    res = ctx.run(b, objective="likelihood")
    h = res.to_host()
    cands = ctx.candidates(b, res, k=5)
    for i in range(5):
        cur = res.curve(i, h)
        want = rounds_top_k(cur["power"], 5, 10)
        assert [c["period"] for c in cands[i]] == list(d["periods"][want])
        assert [c["power"] for c in cands[i]] == list(cur["power"][want])
        assert [c["t0"] for c in cands[i]] == list(cur["transit_time"][want])
