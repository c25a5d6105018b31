"""The autotuner's tilings change the time, never the results: every field
and the best index bit-identical to the default tiling, for a seeded sample
of the candidates on golden fixtures (packed and unpacked bins, unsorted
times, both objectives); the tuner itself returns one of its candidates."""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from gpu_util import GOLDEN

pytestmark = pytest.mark.gpu

CASES = ("cfg1_small_likelihood.npz", "cfg1_small_snr.npz", "hetero_dy.npz",
         "unsorted_t.npz", "outlier_1e6_hetero.npz")
FIELDS = ("power", "depth", "depth_err", "duration", "transit_time",
          "depth_snr", "log_likelihood", "best_index", "status")


def _run(ctx, c, mode):
    from astro_transit_gpu_amd import DeviceBatch
    b = DeviceBatch.from_host([c["t"]], [c["y"]],
                              None if c["dy"] is None else [c["dy"]],
                              c["periods"], c["durations"], device=ctx.device)
    return ctx.run(b, objective=c["objective"], mode=mode,
                   oversample=c["oversample"]).to_host()


@pytest.mark.parametrize("mode", ["parity", "fast"])
def test_tilings_bit_identical(mode):
    from astro_transit_gpu_amd import Context
    from astro_transit_gpu_amd.autotune import candidates
    assert torch.cuda.is_available(), "GPU test run without a GPU"
    cands = candidates()
    rng = np.random.default_rng(9)
    pick = [cands[i] for i in rng.choice(len(cands), 12, replace=False)]
    paths = [p for p in GOLDEN if os.path.basename(p) in CASES]
    assert len(paths) == len(CASES)
    for path in paths:
        c = O.load_golden(path)
        ref = _run(Context(0), c, mode)
        for cand in pick:
            print("%s %s %s" % (os.path.basename(path), mode, cand),
                  flush=True)
            got = _run(Context(0, tuning=cand), c, mode)
            for k in FIELDS:
                assert np.array_equal(got[k], ref[k], equal_nan=True), (
                    os.path.basename(path), cand, k)


def test_autotune_picks_a_candidate():
    from astro_transit_gpu_amd import Context, synth
    from astro_transit_gpu_amd.autotune import autotune, candidates
    d = synth.config2(n_lc=16, seed=12, n_points=3000, n_periods=400)
    ctx = Context(0)
    best, timings = autotune(ctx, d["t"], d["y"], d["dy"], d["periods"],
                             d["durations"], sample=8, reps=1,
                             use_cache=False)
    assert best in candidates()
    assert timings and timings[0][1] == best
    assert ctx.get_tuning() == best
