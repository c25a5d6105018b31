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
    # flat targets keep every point, the flux_err = 1.0 one included
    # (data/sector_cache.py:201-210 applies no mask in this layout)
    assert [x.size for x in d["t"]] == [5, 9, 3]
    for i, (a, b) in enumerate(((0, 5), (5, 14), (14, 17))):
        np.testing.assert_array_equal(d["t"][i], t[a:b])
        np.testing.assert_array_equal(d["y"][i], f[a:b])
        np.testing.assert_array_equal(d["dy"][i], e[a:b])
    # offsets with the final end stored give the same targets
    path2 = os.path.join(tmp_path, "flat_end.npz")
    np.savez(path2, time=t, flux=f, flux_err=e, tic_ids=np.array([1, 2, 3]),
             offsets=np.array([0, 5, 14, 17]), is_vectorized=False)
    d2 = load_sector_cache(path2)
    for k in ("t", "y", "dy"):
        for a, b in zip(d[k], d2[k]):
            np.testing.assert_array_equal(a, b)


def test_target_data_lookup(tmp_path):
    """target_data follows get_target_data: the first target with the TIC
    id, None when absent."""
    from astro_transit_gpu_amd.ingest import load_sector_cache, target_data
    t = np.arange(6, dtype=np.float64)
    path = os.path.join(tmp_path, "flat.npz")
    np.savez(path, time=t, flux=t.astype(np.float32) + 1,
             flux_err=np.full(6, 1e-3, np.float32),
             tic_ids=np.array([7, 9, 7]), offsets=np.array([0, 2, 4]),
             is_vectorized=False)
    d = load_sector_cache(path)
    got = target_data(d, 7)
    np.testing.assert_array_equal(got["time"], t[0:2])
    np.testing.assert_array_equal(target_data(d, 9)["flux"],
                                  np.array([3.0, 4.0], np.float32))
    assert target_data(d, 8) is None


def reference_layout_npz(path, seed=21, n_targets=6, length=1312):
    """A sector cache written with numpy alone in the layout the reference's
    SectorCache.build saves (data/sector_cache.py:141-148): FP32 common
    time axis, (targets, 1312) FP32 flux and flux_err matrices, curves
    shorter than 1312 padded with flux 0 and flux_err 1.0, int64 tic_ids,
    is_vectorized=True.  Returns the unpadded curves."""
    from astro_transit_gpu_amd import synth
    rng = np.random.default_rng(seed)
    time = (1468.27 + np.arange(length) * (10.0 / 1440.0)).astype(np.float32)
    flux = np.zeros((n_targets, length), np.float32)
    err = np.zeros((n_targets, length), np.float32)
    curves = []
    for i in range(n_targets):
        n = int(rng.integers(700, length + 1))
        y = 1.0 + 8e-4 * rng.standard_normal(n)
        if i % 2 == 0:
            y = synth.inject_box(time[:n].astype(np.float64), y,
                                 rng.uniform(0.8, 3.0), rng.uniform(0, 0.8),
                                 0.1, 3e-3)
        dy = rng.uniform(7e-4, 9e-4, n)
        flux[i, :n] = y
        err[i, :n] = dy
        err[i, n:] = 1.0
        curves.append((time[:n].astype(np.float64), flux[i, :n].copy(),
                       err[i, :n].copy()))
    np.savez(path, time=time, flux=flux, flux_err=err,
             tic_ids=np.arange(n_targets, dtype=np.int64) + 1000,
             is_vectorized=True)
    return curves


def test_reference_layout_ingest(tmp_path):
    """A cache in the reference's exact vectorized layout loads to the
    unpadded curves, values unchanged."""
    from astro_transit_gpu_amd.ingest import load_sector_cache
    path = os.path.join(tmp_path, "data.npz")
    curves = reference_layout_npz(path)
    d = load_sector_cache(path)
    assert list(d["tic_ids"]) == [1000 + i for i in range(len(curves))]
    for i, (t, y, dy) in enumerate(curves):
        np.testing.assert_array_equal(d["t"][i], t)
        np.testing.assert_array_equal(d["y"][i], y)
        np.testing.assert_array_equal(d["dy"][i], dy)


@pytest.mark.gpu
def test_reference_layout_ingest_to_gpu(tmp_path):
    """Sector cache in the reference's layout -> ingest -> one batched HIP
    launch: every target within the parity bar of the FP64 oracle run on
    its unpadded arrays, the best period and its fields the oracle's."""
    from astro_transit_gpu_amd import Context, DeviceBatch
    from astro_transit_gpu_amd.ingest import load_sector_cache
    from gpu_util import assert_parity
    from oracle import oracle as O
    path = os.path.join(tmp_path, "data.npz")
    curves = reference_layout_npz(path)
    d = load_sector_cache(path)
    periods = np.geomspace(0.5, 4.0, 400)
    durs = np.array([0.04, 0.08, 0.12])
    ctx = Context(0)
    b = DeviceBatch.from_host(d["t"], d["y"], d["dy"], periods, durs,
                              device=ctx.device)
    res = ctx.run(b, objective="likelihood")
    h = res.to_host()
    for i, (t, y, dy) in enumerate(curves):
        ref = O.power(t, y, dy, periods, durs, "likelihood")
        g = res.curve(i, h)
        assert_parity(g, ref, "target %d" % i)
        ib = int(np.argmax(ref["power"]))
        assert int(h["best_index"][i]) == ib
        for k in ref:
            assert g[k][ib] == ref[k][ib], (i, k)


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
    # its per-period outputs against the oracle on the same curve (unit
    # weights), in parity mode
    onep = run_gpu_bls(t, flux[0], periods, durs, flux_err=None,
                       method="parity")
    ref = O.power(t, flux[0], None, periods, durs, "snr")
    got = {"power": np.asarray(onep["power"]),
           "duration": np.asarray(onep["all_durs"]),
           "transit_time": np.asarray(onep["all_t0s"]),
           "depth": np.asarray(onep["all_depths"])}
    assert_parity(got, {k: ref[k] for k in got}, "run_gpu_bls")
    ib = int(np.argmax(ref["power"]))
    assert onep["best_period"] == periods[ib]
    assert onep["best_t0"] == ref["transit_time"][ib]
