"""Reference wrapper functions and sector-cache ingest."""
import os

import numpy as np
import pytest


def _curves(seed, k, n_lo=600, n_hi=1312):
    from astro_transit_gpu_amd import synth
    rng = np.random.default_rng(seed)
    t = np.arange(1312) * (2.0 / 1440.0) * 14.0  # ~25.5 d common axis
    out = []
    for i in range(k):
        n = int(rng.integers(n_lo, n_hi + 1))
        y = 1.0 + 1e-3 * rng.standard_normal(1312)
        if i % 2 == 0:
            y = synth.inject_box(t, y, rng.uniform(1.0, 5.0),
                                 rng.uniform(0, 1), 0.15, 4e-3)
        out.append((n, y.astype(np.float32),
                    np.full(1312, 1e-3, np.float32)))
    return t, out


def test_sector_cache_roundtrip(tmp_path):
    """Vectorized cache (padding flux_err = 1.0) -> ragged curves with the
    padding removed (sector_cache.py:196-202)."""
    from astro_transit_gpu_amd.ingest import (load_sector_cache,
                                              write_sector_cache)
    t, cur = _curves(3, 7)
    path = os.path.join(tmp_path, "data.npz")
    write_sector_cache(path, t, [c[1][:c[0]] for c in cur],
                       [c[2][:c[0]] for c in cur], np.arange(7) + 100)
    d = load_sector_cache(path)
    assert list(d["tic_ids"]) == list(range(100, 107))
    for i, (n, y, dy) in enumerate(cur):
        assert d["t"][i].dtype == np.float64
        np.testing.assert_array_equal(d["t"][i],
                                      t[:n].astype(np.float32).astype(np.float64))
        np.testing.assert_array_equal(d["y"][i], y[:n])
        np.testing.assert_array_equal(d["dy"][i], dy[:n])


def test_sector_cache_flat(tmp_path):
    from astro_transit_gpu_amd.ingest import load_sector_cache
    rng = np.random.default_rng(4)
    lens = [5, 9, 3]
    t = np.concatenate([np.sort(rng.uniform(0, 10, n)) for n in lens])
    f = rng.standard_normal(t.size).astype(np.float32)
    e = np.full(t.size, 1e-3, np.float32)
    e[2] = 1.0  # a padded point
    offs = np.array([0, 5, 14])
    path = os.path.join(tmp_path, "flat.npz")
    np.savez(path, time=t, flux=f, flux_err=e, tic_ids=np.array([1, 2, 3]),
             offsets=offs, is_vectorized=False)
    d = load_sector_cache(path)
    assert [x.size for x in d["t"]] == [4, 9, 3]


@pytest.mark.gpu
def test_wrappers_match_engine_and_oracle():
    """run_vbls_exact_parity / run_vbls_massive / run_gpu_bls on a shared
    time axis with zero-weight padding: per-curve results equal the oracle
    run on the unpadded curve with the padded points given infinite dy."""
    from astro_transit_gpu_amd.compat import (run_gpu_bls, run_vbls_massive,
                                              run_vbls_exact_parity)
    from gpu_util import assert_parity
    from oracle import oracle as O
    t, cur = _curves(5, 6)
    flux = np.stack([c[1] for c in cur]).astype(np.float64)
    w = np.stack([1.0 / c[2].astype(np.float64) ** 2 for c in cur])
    for i, c in enumerate(cur):
        w[i, c[0]:] = 0.0
    periods = np.geomspace(0.6, 8.0, 300)
    durs = np.array([0.05, 0.1, 0.15])
    ep = run_vbls_exact_parity(t, flux, periods, durs, weights_matrix=w)
    ms = run_vbls_massive(t, flux, periods, durs, weights_matrix=w)
    for i in range(len(cur)):
        dy = np.where(w[i] > 0, 1.0 / np.sqrt(np.where(w[i] > 0, w[i], 1)),
                      np.inf)
        ref = O.power(t, flux[i], dy, periods, durs, "snr")
        got = {"power": ep["power_array"][i], "depth": ep["depth_array"][i],
               "duration": ep["dur_array"][i],
               "transit_time": ep["t0_array"][i]}
        ref_sub = {k: ref[k] for k in got}
        assert_parity(got, ref_sub, "curve %d" % i)
        ib = int(np.argmax(ref["power"]))
        assert ms["best_period"][i] == periods[ib]
        assert ms["snr"][i] == ep["snr"][i]
    one = run_gpu_bls(t, flux[0], periods, durs, flux_err=None)
    assert set(one) >= {"best_period", "best_t0", "best_duration",
                        "best_depth", "power", "snr", "all_t0s", "all_durs",
                        "all_depths", "periods"}
