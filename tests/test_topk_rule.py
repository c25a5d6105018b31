"""The round-based peak thinning of the device top-k equals scipy's greedy
find_peaks(distance=...) selection (host only; scipy is the checker)."""
import numpy as np
from scipy.signal import find_peaks

from topk_model import (local_maxima, reference_top_k, rounds_top_k,
                        sample_arrays)


def test_local_maxima_equal_scipy():
    rng = np.random.default_rng(10)
    for x in sample_arrays(rng, 700):
        assert np.array_equal(local_maxima(x), find_peaks(x)[0])


def test_rounds_equal_greedy():
    rng = np.random.default_rng(11)
    checked = 0
    for x in sample_arrays(rng, 700):
        for d in (1, 3, 10):
            pk = local_maxima(x)
            # scipy's greedy order among equal peak values depends on its
            # sort; compare where the peak values are distinct
            if len(pk) and len(np.unique(x[pk])) != len(pk):
                continue
            assert np.array_equal(rounds_top_k(x, 5, d),
                                  reference_top_k(x, 5, d)), (x, d)
            checked += 1
    assert checked > 1000
