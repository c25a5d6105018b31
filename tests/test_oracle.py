"""The oracle pinned against astropy's own outputs and against the
mathematics (CPU only)."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden",
                                       "*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: os.path.basename(p))
def test_oracle_bit_exact_vs_astropy(oracle_lib, path):
    """The C restatement reproduces astropy's compiled bls_impl bit for bit
    (every field of every period with a qualifying window)."""
    c = O.load_golden(path)
    out = O.power(c["t"], c["y"], c["dy"], c["periods"], c["durations"],
                  c["objective"], c["oversample"])
    g = c["out"]
    fin = np.isfinite(g["power"])
    np.testing.assert_array_equal(np.isfinite(out["power"]), fin)
    for k in O.FIELDS:
        np.testing.assert_array_equal(out[k][fin], g[k][fin], err_msg=k)


def test_astropy_known_answer_32bit_bug(oracle_lib):
    """astropy test_bls.py:382-397, restated: inputs regenerated from the
    seed, expected values typed from the test."""
    rand = np.random.default_rng(42)
    t = rand.uniform(0, 10, 500)
    y = np.ones_like(t)
    y[np.abs((t + 1.0) % 2.0 - 1) < 0.08] = 1.0 - 0.1
    y += 0.01 * rand.standard_normal(len(t))
    out = O.power(t, y, None, np.linspace(1.9, 2.1, 5), 0.16)
    np.testing.assert_allclose(
        out["power"],
        [0.01723948, 0.0643028, 0.1338783, 0.09428816, 0.03577543],
        rtol=1.1e-7)
    from astro_transit_gpu_amd.synth import autoperiod
    grid = autoperiod(t, 0.16)
    auto = O.power(t, y, None, grid, 0.16)
    assert grid[np.argmax(auto["power"])] == pytest.approx(2.000412388152837,
                                                         abs=1e-12)


def _astropy_data():
    c = O.load_golden(os.path.join(os.path.dirname(__file__), "golden",
                                   "adata_fast_likelihood.npz"))
    return c["t"], c["y"], c["dy"]


@pytest.mark.parametrize("objective", ["likelihood", "snr"])
@pytest.mark.parametrize("offset", [False, True])
def test_fast_equals_slow(oracle_lib, objective, offset):
    """astropy test_fast_method (test_bls.py:414-428): the binned method
    equals the brute-force definition on every field."""
    t, y, dy = _astropy_data()
    if offset:
        t = t - 0.5 + 2.0
    periods = np.exp(np.linspace(np.log(2.0) - 1, np.log(2.0) + 1, 10))
    fast = O.power(t, y, dy, periods, 0.16, objective)
    slow = O.bls_slow(t, y, dy, periods, 0.16, objective)
    for k in O.FIELDS:
        np.testing.assert_allclose(fast[k], slow[k], rtol=1e-7, err_msg=k)


@pytest.mark.parametrize("objective", ["likelihood", "snr"])
def test_correct_model(oracle_lib, objective):
    """astropy test_correct_model (test_bls.py:400-411): the argmax recovers
    the injected box within 0.01 and |chi| < 1."""
    t, y, dy = _astropy_data()
    periods = np.exp(np.linspace(np.log(2.0) - 0.1, np.log(2.0) + 0.1, 1000))
    out = O.power(t, y, dy, periods, 0.16, objective)
    i = int(np.argmax(out["power"]))
    truth = dict(period=2.0, transit_time=0.5, duration=0.16, depth=0.2)
    got = dict(period=periods[i], transit_time=out["transit_time"][i],
               duration=out["duration"][i], depth=out["depth"][i])
    for k, v in truth.items():
        assert abs(got[k] - v) < 0.01, k
    assert abs((out["depth"][i] - 0.2) / out["depth_err"][i]) < 1


def test_closed_forms(oracle_lib):
    """loglike = W_in*depth^2/2, snr = depth/depth_err, depth_err =
    sqrt(1/W_in + 1/W_out) with W_in + W_out = sum(ivar)."""
    c = O.load_golden(os.path.join(os.path.dirname(__file__), "golden",
                                   "hetero_dy.npz"))
    out = O.power(c["t"], c["y"], c["dy"], c["periods"], c["durations"])
    w_tot = np.sum(1.0 / np.asarray(c["dy"], dtype=np.float64) ** 2)
    d, e = out["depth"], out["depth_err"]
    # solve W_in from depth_err and W_tot: 1/W_in + 1/(W-W_in) = e^2
    s = e ** 2
    disc = np.sqrt(np.maximum(w_tot ** 2 - 4 * w_tot / s, 0))
    w_in_a = (w_tot - disc) / 2
    w_in_b = (w_tot + disc) / 2
    ll_a = 0.5 * w_in_a * d ** 2
    ll_b = 0.5 * w_in_b * d ** 2
    ok = (np.isclose(ll_a, out["log_likelihood"], rtol=1e-6) |
          np.isclose(ll_b, out["log_likelihood"], rtol=1e-6))
    assert ok.all()
    np.testing.assert_allclose(out["depth_snr"], d / e, rtol=1e-14)
    assert (d >= 0).all()


def test_shift_invariance(oracle_lib):
    """Adding a constant to y changes nothing (median subtraction); adding a
    constant to t shifts transit_time by it and leaves the rest."""
    c = O.load_golden(os.path.join(os.path.dirname(__file__), "golden",
                                   "cfg1_small_likelihood.npz"))
    a = O.power(c["t"], c["y"], c["dy"], c["periods"], c["durations"])
    b = O.power(c["t"], c["y"] + 0.5, c["dy"], c["periods"], c["durations"])
    np.testing.assert_allclose(a["power"], b["power"], rtol=1e-9)
    s = O.power(c["t"] + 1000.0, c["y"], c["dy"], c["periods"],
                c["durations"])
    np.testing.assert_allclose(a["power"], s["power"], rtol=1e-9)
    np.testing.assert_allclose(a["transit_time"] + 1000.0, s["transit_time"],
                               atol=1e-9)


def test_dy_scaling(oracle_lib):
    """Constant dy = sigma vs dy = None: depth unchanged, likelihood power
    divided by sigma^2, snr power divided by sigma."""
    c = O.load_golden(os.path.join(os.path.dirname(__file__), "golden",
                                   "wrap_box.npz"))
    sig = 2e-3
    for obj, scale in (("likelihood", sig ** 2), ("snr", sig)):
        a = O.power(c["t"], c["y"], None, c["periods"], c["durations"], obj)
        b = O.power(c["t"], c["y"], np.full(c["t"].size, sig), c["periods"],
                    c["durations"], obj)
        np.testing.assert_allclose(a["depth"], b["depth"], rtol=1e-9)
        np.testing.assert_allclose(a["power"] / scale, b["power"], rtol=1e-9)


def test_planted_padding_mistake_is_caught(oracle_lib):
    """A plausible mistake (padding that does not overwrite the last partial
    bin) changes the output on the golden data, so the bit-exact check
    above would catch it."""
    c = O.load_golden(os.path.join(os.path.dirname(__file__), "golden",
                                   "wrap_box.npz"))
    t_rel, yc, ivar, t_ref = O.prepare(c["t"], c["y"], c["dy"])
    g = c["out"]["power"]
    bd = np.min(c["durations"]) / c["oversample"]
    bad = []
    for p in c["periods"]:
        nb = int(np.ceil(p / bd)) + c["oversample"]
        ind = (np.floor((t_rel - p * np.floor(t_rel / p)) / bd)).astype(int) + 1
        wy = np.bincount(ind, yc * ivar, nb + 2)[:nb + 1]
        w = np.bincount(ind, ivar, nb + 2)[:nb + 1]
        # mistake: append copies after the data instead of overwriting
        wy = np.concatenate([wy, wy[1:c["oversample"] + 1]])
        w = np.concatenate([w, w[1:c["oversample"] + 1]])
        cw, cy = np.cumsum(w), np.cumsum(wy)
        best = -np.inf
        for d in c["durations"]:
            m = int(round(d / bd))
            wi = cw[m:] - cw[:-m]
            yi = cy[m:] - cy[:-m]
            wo = ivar.sum() - wi
            ok = (wi > 0) & (wo > 0)
            dep = (yc @ ivar - yi[ok]) / wo[ok] - yi[ok] / wi[ok]
            ll = 0.5 * wi[ok] * dep ** 2
            ll[dep < 0] = -np.inf
            best = max(best, ll.max())
        bad.append(best)
    assert not np.allclose(np.array(bad), g, rtol=1e-12)
