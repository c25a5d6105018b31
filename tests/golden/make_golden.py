"""Write the golden BLS fixtures in this directory.

Run ONCE in the build container with the conda interpreter that carries
astropy 4.3.1:

    /opt/conda/bin/python3.9 tests/golden/make_golden.py [NAME ...]

(with names, only those fixtures are rewritten).

What it runs.  The oracle of the hot path is astropy's
``BoxLeastSquares.power(method="fast")`` (the reference's CPU path,
src/astrotransit_gpu/search/cpu_reference_bls.py:11-16).  Its arithmetic
lives in the compiled ``bls_impl`` of
``astropy/timeseries/periodograms/bls/_impl*.so`` (astropy 4.3.1).  Importing
``astropy.timeseries`` fails in that environment (astropy 4.3.1 references
``numpy.asscalar``, removed from numpy 1.26), so the extension module is
loaded by file path and called directly; the Python wrapper around it
(astropy bls/core.py:304-330 for the input preparation, 702-720 for adding
t_ref back) is restated below line for line.  The library is only run to
produce outputs: nothing here reads or inspects the binary.

The fixtures hold inputs and outputs only (``np.savez_compressed``, no
pickles).  The tests in this directory's parent load them with
``allow_pickle=False``.
"""

import ast
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SITE = "/opt/conda/lib/python3.9/site-packages/astropy"
IMPL = SITE + "/timeseries/periodograms/bls/_impl.cpython-39-x86_64-linux-gnu.so"
ASTROPY_TEST = SITE + "/timeseries/periodograms/bls/tests/test_bls.py"

FIELDS = ("power", "depth", "depth_err", "duration", "transit_time",
          "depth_snr", "log_likelihood")


def load_impl():
    spec = importlib.util.spec_from_file_location("_impl", IMPL)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.bls_impl


def load_synth():
    path = os.path.join(HERE, "..", "..", "astro_transit_gpu_amd", "synth.py")
    spec = importlib.util.spec_from_file_location("synth", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


BLS_IMPL = load_impl()
SYNTH = load_synth()


def astropy_power(t, y, dy, period, duration, objective="likelihood",
                  oversample=10):
    """astropy BoxLeastSquares(t, y, dy).power(period, duration, objective,
    method="fast", oversample) for plain (unit-less) float arrays."""
    period = np.atleast_1d(np.abs(period))
    duration = np.atleast_1d(duration)
    if not np.min(period) > np.max(duration):
        raise ValueError("max duration must be shorter than min period")
    t = np.ascontiguousarray(t, dtype=np.float64)
    t_ref = np.min(t)
    y = np.ascontiguousarray(y, dtype=np.float64)
    if dy is None:
        ivar = np.ones_like(y)
    else:
        ivar = 1.0 / np.ascontiguousarray(dy, dtype=np.float64) ** 2
    period_fmt = np.ascontiguousarray(period, dtype=np.float64)
    duration = np.ascontiguousarray(duration, dtype=np.float64)
    res = BLS_IMPL(t - t_ref, y - np.median(y), ivar, period_fmt, duration,
                   int(oversample), objective == "likelihood")
    res = list(res)
    res[4] = res[4] + t_ref
    return dict(zip(FIELDS, res))


ONLY = set(sys.argv[1:])  # names to (re)write; empty: every fixture


def save(name, t, y, dy, periods, durations, objective="likelihood",
         oversample=10, **extra):
    if ONLY and name not in ONLY:
        return astropy_power(t, y, dy, periods, durations, objective,
                             oversample)
    out = astropy_power(t, y, dy, periods, durations, objective, oversample)
    arrs = dict(t=np.asarray(t), y=np.asarray(y),
                periods=np.atleast_1d(np.asarray(periods, dtype=np.float64)),
                durations=np.atleast_1d(np.asarray(durations,
                                                   dtype=np.float64)),
                oversample=np.int64(oversample),
                objective=np.array(objective))
    if dy is not None:
        arrs["dy"] = np.asarray(dy)
    for k, v in out.items():
        arrs["out_" + k] = np.asarray(v, dtype=np.float64)
    arrs.update(extra)
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrs)
    best = int(np.argmax(out["power"]))
    print("%-28s N=%-6d P=%-5d D=%d best=%d power=%.9g" % (
        name, np.size(t), np.size(periods), np.size(durations), best,
        out["power"][best]))
    return out


def astropy_test_data():
    """The hard-coded ``data`` fixture of astropy's test_bls.py (lines 40-379):
    500 times, 500 dy, 500 standard-normal draws; box P=2, t0=0.5,
    duration 0.16, depth 0.2.  The arrays are parsed from the file text."""
    src = open(ASTROPY_TEST).read()
    tree = ast.parse(src)
    arrays = {}
    for node in ast.walk(tree):
        if isinstance(node, ast.FunctionDef) and node.name == "data":
            for st in node.body:
                if (isinstance(st, ast.Assign) and
                        isinstance(st.value, ast.Call) and
                        getattr(st.value.func, "attr", "") == "array"):
                    arrays[st.targets[0].id] = np.array(
                        ast.literal_eval(st.value.args[0]))
    t = arrays["t"]
    dy = arrays["dy"]
    period, transit_time, duration, depth = 2.0, 0.5, 0.16, 0.2
    y = np.ones_like(t)
    m = np.abs((t - transit_time + 0.5 * period) % period
               - 0.5 * period) < 0.5 * duration
    y[m] = 1.0 - depth
    y += dy * arrays["randn_arr"]
    return t, y, dy


def main():
    # 1. astropy test_32bit_bug (test_bls.py:382-397) known-answer input.
    rand = np.random.default_rng(42)
    t = rand.uniform(0, 10, 500)
    y = np.ones_like(t)
    y[np.abs((t + 1.0) % 2.0 - 1) < 0.08] = 1.0 - 0.1
    y += 0.01 * rand.standard_normal(len(t))
    save("k32_linspace", t, y, None, np.linspace(1.9, 2.1, 5), 0.16)
    auto = SYNTH.autoperiod(t, 0.16)
    out = save("k32_autopower", t, y, None, auto, 0.16)
    assert abs(auto[np.argmax(out["power"])] - 2.000412388152837) < 1e-12

    # 2. astropy ``data`` fixture: test_fast_method (414-428) grids, both
    #    objectives, with and without the time offset; test_correct_model
    #    (400-411) fine grid.
    t, y, dy = astropy_test_data()
    for obj in ("likelihood", "snr"):
        grid = np.exp(np.linspace(np.log(2.0) - 1, np.log(2.0) + 1, 10))
        save("adata_fast_%s" % obj, t, y, dy, grid, 0.16, obj)
        save("adata_fast_offset_%s" % obj, t - 0.5 + 2.0, y, dy, grid, 0.16,
             obj)
        fine = np.exp(np.linspace(np.log(2.0) - 0.1, np.log(2.0) + 0.1, 1000))
        save("adata_model_%s" % obj, t, y, dy, fine, 0.16, obj)
    save("adata_nody", t, y, None, np.linspace(1.5, 2.5, 50),
         [0.08, 0.16, 0.24])
    save("adata_oversample1000", t, y, dy, np.array([2.0]), 0.16,
         oversample=1000)
    # test_transit_time_in_range (811-826): absolute-time offset 10230 d.
    save("adata_offset_10230", t + 10230.0, y, dy,
         SYNTH.autoperiod(t, 0.16), 0.16)

    # 3. Scaled-down and full light curve of the 15k x 1k x 5 workload.
    c1 = SYNTH.config1()
    for obj in ("likelihood", "snr"):
        save("cfg1_%s" % obj, c1["t"], c1["y"], c1["dy"], c1["periods"],
             c1["durations"], obj)
    r = np.random.default_rng(11)
    sub = np.sort(r.permutation(c1["t"].size)[:2000])
    for obj in ("likelihood", "snr"):
        save("cfg1_small_%s" % obj, c1["t"][sub], c1["y"][sub],
             c1["dy"][sub], np.geomspace(0.5, 13.5, 200), c1["durations"],
             obj)

    # 4. FP32 y/dy input as in the 1,000-curve batch (astropy widens to FP64).
    r = np.random.default_rng(12)
    tt, yy, dd, _ = SYNTH.config2_curve(r, 3000)
    save("fp32_inputs", tt, yy, dd, np.geomspace(0.5, 13.5, 300),
         SYNTH.DURATIONS_5)

    # 5. Edge and ambiguity cases.
    r = np.random.default_rng(13)
    t = np.sort(r.uniform(0, 20, 1500))
    base = 1.0 + 1e-3 * r.standard_normal(t.size)
    # transit straddling phase 0 of the fold (t0 on the wrap point).
    y = SYNTH.inject_box(t, base, 2.5, 0.0, 0.12, 4e-3)
    save("wrap_box", t, y, None, np.linspace(2.4, 2.6, 41), [0.06, 0.12])
    # periods whose P/bin_duration is (nearly) an integer: bd = 0.01.
    save("int_bins", t, y, None, np.array([1.0, 2.0, 2.5, 3.0, 4.0, 5.0]),
         [0.1, 0.2])
    # all-equal flux: every window has depth 0.
    save("flat", t, np.ones_like(t), None, np.linspace(1.0, 3.0, 7), [0.1])
    # brightening box only.
    yb = SYNTH.inject_box(t, np.ones_like(t), 2.0, 0.3, 0.1, -5e-3)
    save("brighten", t, yb, None, np.linspace(1.5, 2.5, 21), [0.05, 0.1])
    # brightening box plus noise.
    save("brighten_noisy", t, yb + 1e-4 * r.standard_normal(t.size), None,
         np.linspace(1.5, 2.5, 21), [0.05, 0.1])
    # unsorted durations, oversample 1 and 3.
    save("dur_unsorted", t, y, None, np.linspace(1.0, 4.0, 60),
         [0.12, 0.04, 0.2, 0.08])
    save("oversample1", t, y, None, np.linspace(1.0, 4.0, 60), [0.05, 0.1],
         oversample=1)
    save("oversample3", t, y, None, np.linspace(1.0, 4.0, 60), [0.05, 0.1],
         oversample=3)
    # unsorted time stamps (astropy does not require sorted t).
    perm = r.permutation(t.size)
    save("unsorted_t", t[perm], y[perm], None, np.linspace(2.4, 2.6, 41),
         [0.06, 0.12])
    # sparse data: long empty stretches give many windows with equal sums.
    ts = np.concatenate([np.sort(r.uniform(0, 0.5, 40)),
                         np.sort(r.uniform(7.0, 7.4, 40))])
    ys = 1.0 + 1e-3 * r.standard_normal(ts.size)
    save("sparse_ties", ts, ys, None, np.linspace(1.0, 3.0, 25), [0.1, 0.2])
    # heteroscedastic errors.
    dyh = np.exp(r.uniform(np.log(5e-4), np.log(5e-3), t.size))
    save("hetero_dy", t, y + dyh * r.standard_normal(t.size), dyh,
         np.geomspace(0.7, 6.0, 120), [0.05, 0.1, 0.15])
    # tiny inputs.
    save("n1", np.array([3.0]), np.array([1.0]), None, np.array([1.0, 2.0]),
         [0.1])
    save("n3", np.array([0.0, 0.5, 1.7]), np.array([1.0, 0.9, 1.1]), None,
         np.array([1.0, 2.0]), [0.1, 0.3])

    # 6. A small ragged batch (uneven N, grids and duration sets).
    r = np.random.default_rng(14)
    for i in range(8):
        t, y, dy = SYNTH.ragged_curve(r, 300, 4000, 5.0, 27.0)
        k = int(r.integers(1, 7))
        durs = np.sort(np.linspace(r.uniform(0.02, 0.06),
                                   r.uniform(0.1, 0.25), k))
        per = SYNTH.capped_grid(t, durs, 50, 400)
        obj = "likelihood" if i % 2 == 0 else "snr"
        save("ragged_%d" % i, t, y, dy, per, durs, obj)

    # 7. One extreme flux outlier (1e4 and 1e6 sigma, above and below the
    #    curve), with uniform dy and with dy varied +-10% per point.
    r = np.random.default_rng(15)
    for name, sig, het, sgn in (("outlier_1e4", 1e4, False, 1.0),
                                ("outlier_1e6", 1e6, False, 1.0),
                                ("outlier_1e6_low", 1e6, False, -1.0),
                                ("outlier_1e4_hetero", 1e4, True, 1.0),
                                ("outlier_1e6_hetero", 1e6, True, 1.0)):
        t, y, dy = SYNTH.outlier_curve(r, 4000, sig, het, sgn)
        save(name, t, y, dy, np.geomspace(0.5, 13.5, 300),
             SYNTH.DURATIONS_5)


if __name__ == "__main__":
    sys.exit(main())
