"""Pinning of the oracle's vetting helpers (compute_stats, model,
transit_mask) with astropy's own known answers (astropy
timeseries/periodograms/bls/tests/test_bls.py:546-658) on its ``data``
fixture (period 2.0, transit time 0.5, duration 0.16, depth 0.2)."""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
PARAMS = dict(period=2.0, duration=0.16, transit_time=0.5)


def _data():
    c = O.load_golden(os.path.join(GOLD, "adata_oversample1000.npz"))
    return c["t"], c["y"], c["dy"], c["out"]


@pytest.mark.parametrize("with_err", [True, False])
def test_compute_stats_known_answers(with_err):
    t, y, dy, out1000 = _data()
    if not with_err:
        dy = None
    st = O.compute_stats(t, y, dy, **PARAMS)
    tt = 2.0 * np.arange(int(t.max() / 2.0) + 1) + 0.5
    np.testing.assert_allclose(st["transit_times"], tt)
    assert list(st["per_transit_count"]) == [9, 7, 7, 7, 8]
    if not with_err:
        np.testing.assert_allclose(st["harmonic_amplitude"],
                                   0.029945029964964204, rtol=1e-7)
        np.testing.assert_allclose(st["harmonic_delta_log_likelihood"],
                                   -0.5875918155223113, rtol=1e-7)
        return
    # consistency with the oversample-1000 periodogram at the same period
    # (astropy's outputs, golden fixture) and at half the period (oracle)
    np.testing.assert_allclose(np.sum(st["per_transit_log_likelihood"]),
                               out1000["log_likelihood"][0], rtol=1e-7)
    np.testing.assert_allclose(st["depth"][0], out1000["depth"][0],
                               rtol=1e-7)
    np.testing.assert_allclose(st["depth"][1], out1000["depth_err"][0],
                               rtol=1e-7)
    half = O.power(t, y, dy, [1.0], [0.16], oversample=1000)
    np.testing.assert_allclose(st["depth_half"][0], half["depth"][0],
                               rtol=1e-7)
    np.testing.assert_allclose(st["depth_half"][1], half["depth_err"][0],
                               rtol=1e-7)
    np.testing.assert_allclose(st["harmonic_amplitude"],
                               0.033027988742275853, rtol=1e-7)
    np.testing.assert_allclose(st["harmonic_delta_log_likelihood"],
                               -12407.505922833765, rtol=1e-7)
    for f, k in zip((1.0, 1.0, 1.0, 0.0),
                    ("depth", "depth_even", "depth_odd", "depth_phased")):
        assert abs((st[k][0] - f * 0.2) / st[k][1]) < 1, k


def test_model_is_two_level_regression():
    t, y, dy, _ = _data()
    p, d, t0 = 2.0, 0.16, 0.5
    m_in = np.abs((t - t0 + 0.5 * p) % p - 0.5 * p) < 0.5 * d
    A = np.zeros((len(t), 2))
    A[~m_in, 0] = 1.0
    A[m_in, 1] = 1.0
    w = np.linalg.solve(A.T @ (A / dy[:, None] ** 2), A.T @ (y / dy ** 2))
    m = O.model(t, y, dy, t, p, d, t0)
    np.testing.assert_allclose(m, A @ w)
    mask = O.transit_mask(t, p, d, t0)
    assert np.array_equal(mask, (m - m.max()) < 0.0)
