"""Seeded synthetic light curves for tests, fixtures and benchmarks.

Pure numpy (no torch import) so that the golden-fixture script can load this
file by path under another interpreter.  Every generator is deterministic for
a given seed (numpy ``default_rng`` / PCG64, whose streams for ``uniform``,
``standard_normal`` and ``permutation`` are stable across numpy releases).

The box-transit injector follows the reference's ``inject_box_transit``
(src/astrotransit_gpu/inject/box.py:21-27): phase centred on ``t0``, points
with ``|phase| < duration/2`` are multiplied by ``1 - depth``.
"""

import numpy as np

CADENCE_2MIN = 2.0 / (24.0 * 60.0)
DURATIONS_5 = np.array([0.04, 0.08, 0.12, 0.16, 0.20])


def inject_box(t, y, period, t0, duration, depth):
    """Multiplicative box transit (reference inject/box.py:21-27)."""
    y = np.array(y, copy=True)
    phase = (t - t0 + 0.5 * period) % period - 0.5 * period
    y[np.abs(phase) < 0.5 * duration] *= (1.0 - depth)
    return y


def _slots(baseline, cadence):
    n = int(np.floor(baseline / cadence + 1e-9))
    return np.arange(n, dtype=np.float64) * cadence


def sampled_grid(rng, n_points, baseline, gaps, cadence=CADENCE_2MIN):
    """Sorted subsample of ``n_points`` slots of a regular cadence grid over
    ``[0, baseline)`` with the half-open ``gaps`` ``[(start, stop), ...]``
    removed.  A 2-minute grid over 27 d has 19,440 slots, so 15,000 points
    cover the baseline sparsely rather than as a 20.8 d contiguous run."""
    s = _slots(baseline, cadence)
    keep = np.ones(s.shape, dtype=bool)
    for a, b in gaps:
        keep &= ~((s >= a) & (s < b))
    s = s[keep]
    if n_points > s.size:
        raise ValueError("grid has %d slots, %d requested" % (s.size, n_points))
    pick = np.sort(rng.permutation(s.size)[:n_points])
    return s[pick]


def config1(seed=1):
    """One light curve: 15,000 points on a 2-min grid over 27 d with a 1 d gap
    centred at 13.5 d; flux 1 + N(0, 1e-3) with a box transit at P 3.7 d,
    duration 0.10 d, t0 1.2 d, depth 2e-3; dy 1e-3; 1,000 log-uniform periods
    0.5-13.5 d; five durations 0.04-0.20 d."""
    rng = np.random.default_rng(seed)
    t = sampled_grid(rng, 15000, 27.0, [(13.0, 14.0)])
    y = 1.0 + 1e-3 * rng.standard_normal(t.size)
    y = inject_box(t, y, 3.7, 1.2, 0.10, 2e-3)
    dy = np.full(t.size, 1e-3)
    periods = np.geomspace(0.5, 13.5, 1000)
    return dict(t=t, y=y, dy=dy, periods=periods, durations=DURATIONS_5.copy())


def config2_curve(rng, n_points=15000, baseline=27.0, sigma=1e-3):
    """One light curve of the 1,000-curve batch: two random 0.5-1.5 d gaps,
    FP64 t, FP32 y/dy, 50% chance of one injected box."""
    gaps = []
    for _ in range(2):
        ln = rng.uniform(0.5, 1.5)
        a = rng.uniform(0.0, baseline - ln)
        gaps.append((a, a + ln))
    t = sampled_grid(rng, n_points, baseline, gaps)
    y = 1.0 + sigma * rng.standard_normal(t.size)
    truth = None
    if rng.uniform() < 0.5:
        depth = float(np.exp(rng.uniform(np.log(3e-4), np.log(1e-2))))
        period = float(rng.uniform(0.5, 13.0))
        duration = float(rng.uniform(0.03, 0.25))
        t0 = float(rng.uniform(0.0, period))
        y = inject_box(t, y, period, t0, duration, depth)
        truth = dict(period=period, t0=t0, duration=duration, depth=depth)
    return (t, y.astype(np.float32), np.full(t.size, sigma, dtype=np.float32),
            truth)


def outlier_curve(rng, n_points=15000, sigmas=1e6, hetero=False, sign=1.0,
                  sigma=1e-3):
    """A curve of the 1,000-curve batch (2-min grid over 27 d, two gaps, one
    injected box) with one flux outlier ``sigmas`` standard deviations from
    the mean at a random point, FP64 y/dy.  ``hetero``: dy varies by +-10%
    from point to point (otherwise uniform dy).  A single extreme value sets
    the magnitude of the flux sums, which is what a fixed-point binning has
    to survive."""
    gaps = []
    for _ in range(2):
        ln = rng.uniform(0.5, 1.5)
        a = rng.uniform(0.0, 27.0 - ln)
        gaps.append((a, a + ln))
    t = sampled_grid(rng, n_points, 27.0, gaps)
    y = 1.0 + sigma * rng.standard_normal(t.size)
    period = float(rng.uniform(0.7, 9.0))
    y = inject_box(t, y, period, float(rng.uniform(0.0, period)),
                   float(rng.uniform(0.05, 0.2)),
                   float(np.exp(rng.uniform(np.log(1e-3), np.log(5e-3)))))
    y[int(rng.integers(0, t.size))] += sign * sigmas * sigma
    dy = np.full(t.size, sigma)
    if hetero:
        dy = dy * rng.uniform(0.9, 1.1, t.size)
    return t, y, dy


def config2(n_lc=1000, seed=2, n_points=15000, n_periods=5000):
    """Batch of ``n_lc`` curves x ``n_points``; 5,000 log-uniform periods
    0.5-13.5 d shared by all curves; five durations 0.04-0.20 d."""
    rng = np.random.default_rng(seed)
    ts, ys, dys, truths = [], [], [], []
    for _ in range(n_lc):
        t, y, dy, truth = config2_curve(rng, n_points)
        ts.append(t)
        ys.append(y)
        dys.append(dy)
        truths.append(truth)
    return dict(t=ts, y=ys, dy=dys, truths=truths,
                periods=np.geomspace(0.5, 13.5, n_periods),
                durations=DURATIONS_5.copy())


def autoperiod(t, durations, minimum_period=None, maximum_period=None,
               minimum_n_transit=3, frequency_factor=1.0):
    """Period grid uniform in frequency, following astropy
    ``BoxLeastSquares.autoperiod`` (timeseries/periodograms/bls/core.py:
    175-214): df = ff*min(dur)/baseline^2, default minimum period 2*max(dur),
    default maximum period baseline/(n_transit-1), grid 1/(f_max - df*k)."""
    durations = np.atleast_1d(np.asarray(durations, dtype=np.float64))
    baseline = float(np.max(t) - np.min(t))
    df = frequency_factor * float(np.min(durations)) / baseline ** 2
    if minimum_period is None:
        minimum_period = 2.0 * float(np.max(durations))
    if maximum_period is None:
        if minimum_n_transit <= 1:
            raise ValueError("minimum_n_transit must be greater than 1")
        maximum_period = baseline / (minimum_n_transit - 1)
    if maximum_period < minimum_period:
        minimum_period, maximum_period = maximum_period, minimum_period
    if minimum_period <= 0.0:
        raise ValueError("minimum_period must be positive")
    fmin = 1.0 / maximum_period
    fmax = 1.0 / minimum_period
    nf = 1 + int(np.round((fmax - fmin) / df))
    return 1.0 / (fmax - df * np.arange(nf))


def capped_grid(t, durations, n_min, n_max):
    """``autoperiod`` grid whose length is clipped to ``[n_min, n_max]`` by
    rescaling the frequency factor (the grid stays uniform in frequency)."""
    p = autoperiod(t, durations)
    if p.size > n_max:
        p = autoperiod(t, durations, frequency_factor=p.size / float(n_max))
        p = p[:n_max]
    elif p.size < n_min:
        p = autoperiod(t, durations, frequency_factor=p.size / float(n_min))
        p = p[:n_min]
    return p


def ragged_curve(rng, n_lo, n_hi, base_lo, base_hi, sigma=1e-3,
                 transit_prob=0.5):
    """Curve with N log-uniform in [n_lo, n_hi], baseline uniform in
    [base_lo, base_hi], times sorted uniform on the baseline with one random
    gap, optional injected box."""
    n = int(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
    base = float(rng.uniform(base_lo, base_hi))
    t = np.sort(rng.uniform(0.0, base, n))
    ln = rng.uniform(0.0, 0.1 * base)
    a = rng.uniform(0.0, base - ln)
    t = t[(t < a) | (t >= a + ln)]
    y = 1.0 + sigma * rng.standard_normal(t.size)
    if rng.uniform() < transit_prob:
        period = float(rng.uniform(0.6, max(0.7, base / 3.0)))
        y = inject_box(t, y, period, float(rng.uniform(0, period)),
                       float(rng.uniform(0.03, 0.2)),
                       float(np.exp(rng.uniform(np.log(5e-4), np.log(1e-2)))))
    dy = np.full(t.size, sigma)
    return t, y, dy


def config3(n_lc=16384, seed=3, n_lo=1000, n_hi=20000, p_lo=1000,
            p_hi=20000):
    """Ragged batch: N log-uniform 1k-20k, baselines 10-27 d, per-curve
    autoperiod grids clipped to 1k-20k periods, 5-8 durations each."""
    rng = np.random.default_rng(seed)
    out = dict(t=[], y=[], dy=[], periods=[], durations=[])
    for _ in range(n_lc):
        t, y, dy = ragged_curve(rng, n_lo, n_hi, 10.0, 27.0)
        k = int(rng.integers(5, 9))
        durs = np.linspace(rng.uniform(0.03, 0.06), rng.uniform(0.15, 0.25), k)
        out["t"].append(t)
        out["y"].append(y)
        out["dy"].append(dy)
        out["durations"].append(durs)
        out["periods"].append(capped_grid(t, durs, p_lo, p_hi))
    return out


DURATIONS_6 = np.array([0.04, 0.06, 0.09, 0.135, 0.2, 0.3])


def config4_curve(rng, n_points=200000, baseline=360.0, n_gaps=13,
                  sigma=1e-3):
    """One stitched long-baseline curve: ``n_points`` slots of the 2-min grid
    over ``[0, baseline)`` with ``n_gaps`` periodic 1 d gaps centred at
    ``baseline*(j+1)/(n_gaps+1)`` (a 30-min grid over 360 d has only 17,280
    slots, so the point count is kept and the cadence is the 2-min one);
    flux 1 + N(0, sigma) with one box at period log-uniform 1-100 d, depth
    log-uniform 1e-4-5e-3, duration uniform 0.04-0.3 d; FP32 y/dy."""
    step = baseline / (n_gaps + 1)
    gaps = [(step * (j + 1) - 0.5, step * (j + 1) + 0.5)
            for j in range(n_gaps)]
    t = sampled_grid(rng, n_points, baseline, gaps)
    y = 1.0 + sigma * rng.standard_normal(t.size)
    period = float(np.exp(rng.uniform(np.log(1.0), np.log(100.0))))
    depth = float(np.exp(rng.uniform(np.log(1e-4), np.log(5e-3))))
    duration = float(rng.uniform(0.04, 0.3))
    t0 = float(rng.uniform(0.0, period))
    y = inject_box(t, y, period, t0, duration, depth)
    truth = dict(period=period, t0=t0, duration=duration, depth=depth)
    return (t, y.astype(np.float32), np.full(t.size, sigma, dtype=np.float32),
            truth)


def config4(n_lc=512, seed=4, n_points=200000, n_periods=50000):
    """Precision-study batch: ``n_lc`` stitched 360 d curves (see
    :func:`config4_curve`), 50,000 log-uniform periods 1-100 d, six
    durations 0.04-0.3 d.  Bins per period reach ceil(100/0.004)+10 =
    25,010, above the LDS histogram, so the long periods take the
    phase-tiled path."""
    rng = np.random.default_rng(seed)
    ts, ys, dys, truths = [], [], [], []
    for _ in range(n_lc):
        t, y, dy, truth = config4_curve(rng, n_points)
        ts.append(t)
        ys.append(y)
        dys.append(dy)
        truths.append(truth)
    return dict(t=ts, y=ys, dy=dys, truths=truths,
                periods=np.geomspace(1.0, 100.0, n_periods),
                durations=DURATIONS_6.copy())


def config5_pool(n=128, seed=5, n_lo=1000, n_hi=200000, base_lo=10.0,
                 base_hi=360.0, p_cap=20000):
    """Pool of serving requests: N log-uniform 1k-200k, baseline uniform
    10-360 d, sorted uniform times with a random gap, optional injected box,
    durations 0.04-0.20 d, autoperiod grid capped at ``p_cap`` periods (the
    uncapped grid reaches millions of periods at 360 d)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        t, y, dy = ragged_curve(rng, n_lo, n_hi, base_lo, base_hi)
        durs = DURATIONS_5.copy()
        per = capped_grid(t, durs, 1000, p_cap)
        out.append((t, y.astype(np.float32), dy.astype(np.float32), per,
                    durs))
    return out
